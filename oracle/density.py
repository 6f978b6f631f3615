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
