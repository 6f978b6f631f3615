"""Oracle H*psi (test infrastructure; see oracle/__init__.py).

Restates:
- ``apply_ke``: dft/electrons/_basis_ops.py:29-37;
- ``apply_potential``: _basis_ops.py:39-113 and :152-205 — scatter the
  sphere into a zero box at ``fft_index``, unscaled inverse FFT
  (``ifftn(norm="forward")``), multiply by V(r), forward FFT scaled by 1/N
  (``fftn(norm="forward")``), gather at ``fft_index``, zero the padding;
- ``projectors``: dft/ions/_ions_projectors.py:28-72 with analytic Gaussian
  radials: beta = f(|G+k|) * r^l Y_lm(G+k) * i^l * exp(-2 pi i (iG+k).tau)
  / sqrt(volume), real solid harmonics ordered l^2 + l + m with
  (y, z, x) for l = 1 (math/spherical_harmonics.py:74-87,111-137);
- ``apply_nonlocal``: dft/electrons/_hamiltonian.py:21-24, i.e.
  ``beta @ (D_all @ (beta ^ C))`` with the _dot/_matmul index conventions of
  _wavefunction_dot.py:66-84,168-185;
- ``hamiltonian``: _hamiltonian.py:7-29 (tau and +U branches off).
"""
from __future__ import annotations

import os

import numpy as np
import scipy.fft

from .basis import Basis

WORKERS = int(os.environ.get("ORACLE_THREADS", os.cpu_count() or 1))

Y00 = 0.5 / np.sqrt(np.pi)  # 1/sqrt(4 pi)
Y1 = np.sqrt(3.0 / (4.0 * np.pi))


def projectors(basis: Basis) -> np.ndarray:
    """beta (K, P, n_tot) complex128, projector order atom -> radial -> (l, m)."""
    spec = basis.spec
    iGk = basis.iG + basis.k[:, None, :]  # (K, G, 3)
    Gk = iGk @ basis.Gbasis.T
    Gmag2 = (Gk**2).sum(-1)
    # Solid harmonics times the i^l reciprocal phase:
    Ylm = {0: [Y00 * np.ones_like(Gmag2, dtype=complex)],
           1: [1j * Y1 * Gk[..., 1], 1j * Y1 * Gk[..., 2], 1j * Y1 * Gk[..., 0]]}
    phase = np.exp(-2j * np.pi * (iGk @ spec.positions.T))  # (K, G, atoms)
    phase /= np.sqrt(spec.volume)
    rows = []
    for i_atom in range(spec.n_atoms):
        for rc in spec.proj_rc:
            f = np.exp(-0.5 * rc**2 * Gmag2)
            for l in spec.proj_l:
                for Y in Ylm[l]:
                    rows.append(f * Y * phase[..., i_atom])
    beta = np.stack(rows, axis=1)  # (K, P, G)
    beta[np.broadcast_to(basis.pad_mask[:, None, :], beta.shape)] = 0.0
    return beta


def apply_ke(basis: Basis, C: np.ndarray) -> np.ndarray:
    return C * basis.KE[:, None, :]


def apply_potential(basis: Basis, V: np.ndarray, C: np.ndarray) -> np.ndarray:
    K, B, G = C.shape
    N = basis.N
    out = np.empty_like(C)
    for ik in range(K):
        box = np.zeros((B, N), dtype=C.dtype)
        box[:, basis.fft_index[ik]] = C[ik]
        box = box.reshape((B,) + basis.shape)
        r = scipy.fft.ifftn(box, axes=(1, 2, 3), norm="forward", workers=WORKERS)
        r *= V[None]
        t = scipy.fft.fftn(r, axes=(1, 2, 3), norm="forward", workers=WORKERS)
        out[ik] = t.reshape(B, N)[:, basis.fft_index[ik]]
    basis.constrain(out)
    return out


def dot(C1: np.ndarray, C2: np.ndarray) -> np.ndarray:
    """(K, B1, B2) = conj(C1) @ C2^T per k (_wavefunction_dot.py:66-84)."""
    return np.matmul(C1.conj(), C2.transpose(0, 2, 1))


def matmul(C: np.ndarray, M: np.ndarray) -> np.ndarray:
    """(K, B_out, G) = M^T @ C per k (_wavefunction_dot.py:168-185)."""
    return np.matmul(M.transpose(0, 2, 1), C)


def apply_nonlocal(beta: np.ndarray, D: np.ndarray, C: np.ndarray) -> np.ndarray:
    proj = dot(beta, C)  # (K, P, B)
    return matmul(beta, D[None, :, None] * proj)


class Hamiltonian:
    """H = T + V_loc + beta D beta^H on one cell, FP64."""

    def __init__(self, basis: Basis, V: np.ndarray | None = None):
        self.basis = basis
        spec = basis.spec
        self.V = spec.V if V is None else V
        self.beta = projectors(basis)
        self.D = spec.D.astype(np.float64)

    def __call__(self, C: np.ndarray) -> np.ndarray:
        HC = apply_ke(self.basis, C)
        HC += apply_potential(self.basis, self.V, C)
        HC += apply_nonlocal(self.beta, self.D, C)
        return HC
