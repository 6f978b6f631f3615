"""Dense-H brute force (test infrastructure; see oracle/__init__.py).

Assembles H(G, G') = KE(G) delta + V~(G - G') + sum_p beta_p(G) D_p
conj(beta_p(G')) (+ 1/2 (G + k) . (G' + k) V~tau(G - G') with a meta-GGA
potential) on the in-sphere plane waves of each k, independently of
the FFT path: V~ = fftn(V)/N and the difference G - G' is indexed modulo
the box. Its LAPACK eigenvalues are the exact answer the iterative solvers
must reach.
"""
from __future__ import annotations

import numpy as np

from .basis import Basis
from .hamiltonian import projectors


def dense_hamiltonian(basis: Basis, ik: int, V=None, beta=None, D=None,
                      V_tau=None) -> np.ndarray:
    spec = basis.spec
    V = spec.V if V is None else V
    D = spec.D if D is None else D
    beta = projectors(basis) if beta is None else beta
    n = int(basis.n[ik])
    iG = basis.iG[ik, :n]  # (n, 3) signed
    Vt = np.fft.fftn(V) / basis.N
    diff = (iG[:, None, :] - iG[None, :, :]) % np.array(basis.shape)
    H = Vt[diff[..., 0], diff[..., 1], diff[..., 2]].astype(np.complex128)
    H[np.diag_indices(n)] += basis.KE[ik, :n]
    if V_tau is not None:  # -1/2 sum_d (i G_d)(i G'_d) V~tau(G - G') = 1/2 (G . G') V~tau
        Gk = (iG + basis.k[ik]) @ basis.Gbasis.T
        Tt = np.fft.fftn(V_tau) / basis.N
        H += 0.5 * (Gk @ Gk.T) * Tt[diff[..., 0], diff[..., 1], diff[..., 2]]
    b = beta[ik, :, :n]  # (P, n)
    H += (b.T * D[None, :]) @ b.conj()
    return H


def dense_eigenvalues(basis: Basis, n_bands: int, **kw) -> np.ndarray:
    beta = projectors(basis)
    out = []
    for ik in range(len(basis.k)):
        H = dense_hamiltonian(basis, ik, beta=beta, **kw)
        out.append(np.linalg.eigvalsh(H)[:n_bands])
    return np.array(out)
