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
- ``apply_gradient``: _basis_ops.py:15-26 with get_gradient = i (G + k)
  (_basis.py:240-254); ``apply_tau``: the need_tau branch of _hamiltonian.py:
  13-20, -1/2 sum_dir grad_dir apply_potential(V_tau, grad_dir C). Pinned
  (tests/test_oracle.py): a constant V_tau = c gives c KE C (the sign, the
  factor 1/2 and all three directions), Hermiticity for a real V_tau, and
  <C|H_tau|C> = 1/2 sum_dir <grad C|V_tau|grad C> >= 0 for V_tau >= 0;
- ``hamiltonian``: _hamiltonian.py:7-29 (+U branch off; tau when V_tau is set).
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


def apply_gradient(basis: Basis, C: np.ndarray, i_dir: int) -> np.ndarray:
    Gk = (basis.iG + basis.k[:, None, :]) @ basis.Gbasis.T  # (K, G, 3)
    return C * (1j * Gk[:, None, :, i_dir])  # C (K, B, G)


def apply_tau(basis: Basis, V_tau: np.ndarray, C: np.ndarray) -> np.ndarray:
    out = np.zeros_like(C)
    for i_dir in range(3):
        out -= 0.5 * apply_gradient(
            basis, apply_potential(basis, V_tau, apply_gradient(basis, C, i_dir)), i_dir)
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

    def __init__(self, basis: Basis, V: np.ndarray | None = None,
                 V_tau: np.ndarray | None = None):
        self.basis = basis
        spec = basis.spec
        self.V = spec.V if V is None else V
        self.V_tau = V_tau
        self.beta = projectors(basis)
        self.D = spec.D.astype(np.float64)

    def __call__(self, C: np.ndarray) -> np.ndarray:
        HC = apply_ke(self.basis, C)
        HC += apply_potential(self.basis, self.V, C)
        if self.V_tau is not None:
            HC += apply_tau(self.basis, self.V_tau, C)
        HC += apply_nonlocal(self.beta, self.D, C)
        return HC
