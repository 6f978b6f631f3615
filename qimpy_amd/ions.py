"""Nonlocal projectors and couplings consumed by H (reference: dft/ions).

The band solve consumes ``ions.beta`` (a Wavefunction (1, K, P, 1, G)),
``ions.D_all`` (P x P) and the ``beta_version`` cache key that invalidates
``Wavefunction.proj`` (_wavefunction.py:106-132). Projectors are built on the
device as get_projectors does (dft/ions/_ions_projectors.py:28-72):
beta = f(|G+k|) * r^l Y_lm(G+k) * i^l * exp(-2 pi i (iG+k).tau) / sqrt(V),
with the analytic Gaussian radials f(G) = exp(-rc^2 G^2 / 2) of the synthetic
cells instead of UPF splines, real solid harmonics ordered l^2 + l + m
((y, z, x) for l = 1) and projector order atom -> radial -> (l, m). D_all is
block diagonal per atom (_ions_update.py:229-243); the cells make it diagonal.
"""
from __future__ import annotations

import numpy as np
import torch

from . import rc

Y00 = 0.5 / np.sqrt(np.pi)
Y1 = np.sqrt(3.0 / (4.0 * np.pi))


def build_projectors(positions: torch.Tensor, proj_rc, proj_l, iG: torch.Tensor,
                     k: torch.Tensor, Gbasis: np.ndarray, volume: float,
                     n: torch.Tensor) -> torch.Tensor:
    """beta (K, P, G) complex128 for atoms at `positions` (fractional) on the
    plane waves iG (K, G, 3) + k (K, 3); entries g >= n[k] are zero
    (_ions_projectors.py:28-72). Projector order atom -> radial -> (l, m)."""
    dev = iG.device
    K, G = iG.shape[0], iG.shape[1]
    n_atoms = positions.shape[0]
    if n_atoms == 0 or not proj_rc or not proj_l:
        return torch.zeros((K, 0, G), dtype=torch.complex128, device=dev)
    iGk = iG.to(torch.float64) + k[:, None, :]
    Gb = torch.from_numpy(np.ascontiguousarray(np.asarray(Gbasis, dtype=np.float64).T))
    Gk = iGk @ (Gb.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else Gb)
    Gmag2 = Gk.square().sum(-1)
    Y = []
    for l in proj_l:
        if l == 0:
            Y.append(torch.full_like(Gmag2, Y00).to(torch.complex128))
        else:  # i^l r Y_1m, m order (y, z, x)
            Y.extend(1j * Y1 * Gk[..., a] for a in (1, 2, 0))
    Y = torch.stack(Y, dim=-1)  # (K, G, n_lm)
    rc_t = torch.tensor(proj_rc, dtype=torch.float64)
    rc_t = rc_t.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else rc_t
    f = torch.exp(-0.5 * rc_t ** 2
                  * Gmag2[..., None])  # (K, G, n_rc)
    # exp(-2 pi i (iG + k) . tau) / sqrt(V), atoms last (_ions.py:304-311):
    phase = torch.polar(torch.full((), 1.0 / np.sqrt(volume), device=dev, dtype=torch.float64),
                        (-2 * np.pi) * (iGk @ positions.T))  # (K, G, A)
    proj = phase[..., :, None, None] * f[..., None, :, None] * Y[..., None, None, :]
    proj = proj.reshape(K, G, -1).transpose(1, 2).contiguous()  # (K, P, G)
    pad = torch.arange(G, device=dev)[None, None, :] >= n.to(dev)[:, None, None]
    return proj.masked_fill_(pad, 0.0)


class Ions:
    beta_version: int

    def __init__(self, positions: np.ndarray, proj_rc, proj_l, D: np.ndarray):
        self.positions = torch.tensor(np.asarray(positions, dtype=np.float64), device=rc.device)
        self.proj_rc = tuple(float(r) for r in proj_rc)
        self.proj_l = tuple(int(l) for l in proj_l)
        if any(l > 1 for l in self.proj_l):
            raise NotImplementedError("projectors with l > 1 are not on this path")
        self.D_diag = torch.tensor(np.asarray(D, dtype=np.float64), device=rc.device)
        self.D_all = torch.diag(self.D_diag).to(torch.complex128)
        self.beta = None
        self.beta_full = None
        self.beta_version = 0
        self._beta_cast = {}

    @property
    def n_atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_projectors(self) -> int:
        return self.n_atoms * len(self.proj_rc) * sum(2 * l + 1 for l in self.proj_l)

    def translation_phase(self, iG: torch.Tensor) -> torch.Tensor:
        """exp(-2 pi i iG . tau) with atoms last (_ions.py:304-311)."""
        return torch.polar(torch.ones((), device=iG.device, dtype=torch.float64),
                           (-2 * np.pi) * (iG @ self.positions.T))

    def update(self, basis) -> None:
        """(Re)build beta on `basis` and bump beta_version."""
        from .wavefunction import Wavefunction

        proj = build_projectors(self.positions, self.proj_rc, self.proj_l, basis.iG, basis.k,
                                basis.Gbasis, basis.volume, basis.n)
        self.beta = Wavefunction(basis, coeff=proj[None, :, :, None, :].contiguous())
        self.beta_version += 1
        self._beta_cast = {}

    def beta_coeff(self, dtype: torch.dtype) -> torch.Tensor:
        """beta coefficients in the working precision (cached per version)."""
        if dtype == torch.complex128:
            return self.beta.coeff
        cached = self._beta_cast.get(dtype)
        if cached is None:
            cached = self.beta.coeff.to(dtype)
            self._beta_cast[dtype] = cached
        return cached

    @classmethod
    def from_spec(cls, spec) -> "Ions":
        return cls(spec.positions, spec.proj_rc, spec.proj_l, spec.D)
