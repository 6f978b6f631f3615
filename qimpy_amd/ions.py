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

        iGk = basis.iG.to(torch.float64) + basis.k[:, None, :]
        Gk = iGk @ torch.tensor(basis.Gbasis.T, device=rc.device)
        Gmag2 = Gk.square().sum(-1)
        Ylm = {
            0: [torch.full_like(Gmag2, Y00).to(torch.complex128)],
            1: [1j * Y1 * Gk[..., 1], 1j * Y1 * Gk[..., 2], 1j * Y1 * Gk[..., 0]],
        }
        phase = self.translation_phase(iGk) / np.sqrt(basis.volume)  # (K, G, atoms)
        rows = []
        for i_atom in range(self.n_atoms):
            ph = phase[..., i_atom]
            for rc_ in self.proj_rc:
                f = torch.exp(-0.5 * rc_**2 * Gmag2)
                for l in self.proj_l:
                    for Y in Ylm[l]:
                        rows.append(f * Y * ph)
        proj = torch.stack(rows, dim=1)[None, :, :, None, :]  # (1, K, P, 1, G)
        proj[basis.pad_index] = 0.0
        self.beta = Wavefunction(basis, coeff=proj.contiguous())
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
