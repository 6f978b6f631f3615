"""Oracle electron density (test infrastructure; see oracle/__init__.py).

Restates Basis.collect_density for one spin channel and scalar
wavefunctions (dft/electrons/_basis_ops.py:207-287 with
_CollectDensityKernel, :290-343): prefactor f w_sk / volume (:230); each
band scattered into a zero box at ``fft_index`` and inverse transformed
unscaled (``expand_ifft``, :144-163, i.e. ``ifftn(norm="forward")``); then
``accum_norm_`` (math/_linalg.py:31-76): rho += f |psi|^2, summed over k
and bands, skipping the zero fillings it walks past with ``argwhere``.

Pinned (tests/test_oracle.py): a single plane wave gives the uniform density
f w |c|^2 / volume; two plane waves give |a|^2 + |b|^2 + 2 Re(a b*
exp(i (G1 - G2) . r)) at every grid point; the integral over the cell equals
sum f w for normalised bands (Parseval); against a direct O(N^2) inverse DFT
on a tiny box.

collect_density_cross restates the spinor cross term of need_Mvec
(_basis_ops.py:330-336: accum_prod_(f, psi_1, conj(psi_0)), math/_linalg.py:
77-107): sum over k and bands of f w / volume psi_1 conj(psi_0). Pinned by
the polarization identities against collect_density, rho(C0 + C1) - rho(C0)
- rho(C1) = 2 Re X and rho(C0 + i C1) - rho(C0) - rho(C1) = -2 Im X, and by
the two-plane-wave closed form b conj(a) exp(i (G1 - G0) . r).
"""
from __future__ import annotations

import numpy as np
import scipy.fft

from .basis import Basis
from .hamiltonian import WORKERS


def collect_density(basis: Basis, C: np.ndarray, f: np.ndarray) -> np.ndarray:
    """rho (S0, S1, S2) float64 of bands C (K, B, G) with fillings f (K, B)."""
    K, B, G = C.shape
    prefac = f * (basis.w_sk[:, None] / basis.volume)
    rho = np.zeros(basis.shape)
    for ik in range(K):
        live = np.nonzero(prefac[ik])[0]
        if live.size == 0:
            continue
        box = np.zeros((live.size, basis.N), dtype=C.dtype)
        box[:, basis.fft_index[ik]] = C[ik, live]
        box = box.reshape((live.size,) + basis.shape)
        psi = scipy.fft.ifftn(box, axes=(1, 2, 3), norm="forward", workers=WORKERS)
        rho += np.einsum("b,bxyz->xyz", prefac[ik, live], psi.real**2 + psi.imag**2)
    return rho


def _psi(basis: Basis, ik: int, Ck: np.ndarray) -> np.ndarray:
    box = np.zeros((Ck.shape[0], basis.N), dtype=Ck.dtype)
    box[:, basis.fft_index[ik]] = Ck
    box = box.reshape((Ck.shape[0],) + basis.shape)
    return scipy.fft.ifftn(box, axes=(1, 2, 3), norm="forward", workers=WORKERS)


def collect_density_cross(basis: Basis, C0: np.ndarray, C1: np.ndarray,
                          f: np.ndarray) -> np.ndarray:
    """X (S0, S1, S2) complex128 = sum_{k,b} f w / V psi_1 conj(psi_0) of the
    spinor components C0, C1 (K, B, G) with fillings f (K, B)."""
    K, B, G = C0.shape
    prefac = f * (basis.w_sk[:, None] / basis.volume)
    X = np.zeros(basis.shape, dtype=np.complex128)
    for ik in range(K):
        live = np.nonzero(prefac[ik])[0]
        if live.size == 0:
            continue
        p0, p1 = _psi(basis, ik, C0[ik, live]), _psi(basis, ik, C1[ik, live])
        X += np.einsum("b,bxyz->xyz", prefac[ik, live], p1 * p0.conj())
    return X
