"""Oracle subspace linear algebra (test infrastructure; see oracle/__init__.py).

Restates math/_linalg.py:110-183: ``ortho_matrix`` with Cholesky,
U = (L^-1)^H via a triangular solve, and ``eighg``: E, V = eigh(U^H H U),
V <- U V, eigenvalues ascending, lower triangle used (torch/LAPACK default).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg


def dagger(x: np.ndarray) -> np.ndarray:
    return np.swapaxes(x.conj(), -2, -1)


def ortho_matrix(O: np.ndarray) -> np.ndarray:
    out = np.empty_like(O)
    for idx in np.ndindex(O.shape[:-2]):
        L = np.linalg.cholesky(O[idx])
        Linv = scipy.linalg.solve_triangular(L, np.eye(O.shape[-1]), lower=True)
        out[idx] = Linv.conj().T
    return out


def eigh_lower(M: np.ndarray):
    return np.linalg.eigh(M, UPLO="L")


def eighg(H: np.ndarray, O: np.ndarray):
    U = ortho_matrix(O)
    E, V = eigh_lower(dagger(U) @ (H @ U))
    return E, U @ V
