"""Electronic system for a fixed-Hamiltonian band solve (reference:
dft/electrons/_electrons.py, _hamiltonian.py).

``Electrons.hamiltonian`` is H C = T C + V_loc C + beta (D_all (beta^H C))
(_hamiltonian.py:7-29, tau and +U branches off): the kinetic and local
terms run fused in one sparse-FFT pass (qmx_apply_local), the nonlocal term as
two band-space GEMMs. ``solve()`` runs the configured diagonalizer and, in the
mixed form, finishes with an FP64 Rayleigh-Ritz: complex64 iterations, then
C is promoted to complex128, H C is applied in FP64, and eighg(C^H HC, C^H C)
in FP64 gives the reported eigenvalues and residual norms.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from types import SimpleNamespace
from typing import Optional

import numpy as np
import torch

from . import gemm, rc
from .basis import Basis
from .chefsi import CheFSI
from .davidson import Davidson
from .ions import Ions
from .linalg import eighg
from .wavefunction import Wavefunction


def _n_bands_extra(n_bands: int) -> int:
    return max(1, int(np.ceil(n_bands * 0.1)))  # _fillings.py:238-243


@dataclass
class SolveReport:
    eig: torch.Tensor  #: float64 (S, K, n_bands)
    residual_norm: torch.Tensor  #: float64 (S, K, n_bands) of ||HC - eig C||
    n_iterations: int
    deig_max: float
    Eband: float
    history: list = field(default_factory=list)


class Electrons:
    def __init__(
        self,
        basis: Basis,
        V: torch.Tensor,
        *,
        n_bands: int,
        n_bands_extra: Optional[int] = None,
        dtype: torch.dtype = torch.complex128,
        diagonalizer: str = "davidson",
        n_iterations: int = 100,
        eig_threshold: float = 1e-8,
        filter_order: int = 10,
        init_threshold: float = 1e-1,
        residual_threshold: Optional[float] = None,
        V_tau: Optional[torch.Tensor] = None,
    ):
        self.basis = basis
        self.ions = basis.ions
        self.V_ks = torch.as_tensor(V, device=rc.device, dtype=torch.float64).contiguous()
        self.V_max = float(self.V_ks.max().item())
        #: meta-GGA potential dE/dtau on the box (tau_tilde.grad), or None
        self.V_tau = (None if V_tau is None else
                      torch.as_tensor(V_tau, device=rc.device, dtype=torch.float64).contiguous())
        self.dtype = dtype
        nbe = _n_bands_extra(n_bands) if n_bands_extra is None else n_bands_extra
        self.fillings = SimpleNamespace(n_bands=n_bands, n_bands_extra=nbe, n_bands_min=n_bands)
        self.C = Wavefunction(basis, n_bands=n_bands, dtype=dtype)
        self.eig: Optional[torch.Tensor] = None
        self.deig_max = np.inf
        if diagonalizer == "davidson":
            self.diagonalize = Davidson(electrons=self, n_iterations=n_iterations,
                                        eig_threshold=eig_threshold,
                                        residual_threshold=residual_threshold)
        elif diagonalizer == "chefsi":
            self.diagonalize = CheFSI(electrons=self, n_iterations=n_iterations,
                                      eig_threshold=eig_threshold, filter_order=filter_order,
                                      init_threshold=init_threshold,
                                      residual_threshold=residual_threshold)
        else:
            raise ValueError(f"unknown diagonalizer {diagonalizer!r}")

    # --------------------------------------------------------------- H * psi
    def hamiltonian(self, C: Wavefunction) -> Wavefunction:
        basis = C.basis
        HC = basis.apply_ke_potential(self.V_ks, C)
        if self.V_tau is not None:  # need_tau branch (_hamiltonian.py:13-20)
            basis.apply_tau(self.V_tau, C.coeff, HC.coeff)
        ions = basis.ions
        if ions is not None and ions.n_projectors:
            beta_C = C.proj  # (S, K, P, B), cached on C
            DP = ions.D_diag.to(beta_C.real.dtype)[:, None] * beta_C
            beta = ions.beta_coeff(C.coeff.dtype).flatten(-2)  # (1, K, P, G)
            gemm.rotate(DP, beta, out=HC.coeff.flatten(-2), beta=1.0)  # HC += beta (D P)
        return HC

    # --------------------------------------------------------------- set-up
    def initialize_wavefunctions(self, seed: int = 0) -> None:
        """Random bands then Cholesky orthonormalisation (_electrons.py:271-287).
        Both run in FP64 whatever the working precision: bandwidth-limited
        random bands span their space with a condition number beyond float32,
        so only the orthonormal result is cast."""
        C = self.C if self.dtype == torch.complex128 else self.C.to(torch.complex128)
        C.randomize(seed=seed)
        C = C.orthonormalize()
        self.C = C if self.dtype == torch.complex128 else C.to(self.dtype)
        self.deig_max = np.inf

    def initialize_lcao(self, orbitals, seed: int = 0) -> None:
        """LCAO start of a fixed-H solve (Electrons.initialize_wavefunctions
        with lcao, _electrons.py:255-287, and LCAO.update's subspace
        diagonalisation, _lcao.py:53-68): the first bands are the atomic
        orbitals of every ion (`orbitals`, a RadialSet of the orbital radials,
        built by qmx_projectors as get_atomic_orbitals does), the remaining
        bands random; then eighg(C^H H C, C^H O C) in FP64 and C = C V. With
        more orbitals than bands, the lowest eigenvectors of the full orbital
        subspace are kept (the reference keeps them all). The subspace
        minimisation with smearing belongs to SCF and is not on this path."""
        from .ions import spline_projectors

        if self.basis.ions is None:
            raise ValueError("LCAO needs ions")
        n_total = self.C.n_bands()
        psi = spline_projectors(self.basis, self.basis.ions.positions, orbitals)  # (K, A, G)
        n_atomic = psi.shape[1]
        S = self.C.coeff.shape[0]
        if n_atomic < n_total:
            C = Wavefunction(self.basis, n_bands=n_total, dtype=torch.complex128)
            C.coeff[:, :, :n_atomic, 0] = psi[None]
            C.randomize(seed=seed, b_start=n_atomic)
        else:
            C = Wavefunction(self.basis, coeff=psi[None, :, :, None, :].expand(
                S, -1, -1, -1, -1).contiguous())
        HC = self.hamiltonian(C)
        eig, V = eighg(C ^ HC, C.dot_O(C).wait())
        C = C @ V[..., :n_total]
        self.eig = eig[..., :n_total]
        self.C = C if self.dtype == torch.complex128 else C.to(self.dtype)
        self.deig_max = np.inf

    def get_Eband(self) -> float:
        n = self.fillings.n_bands
        return float((self.basis.w_sk * self.eig[..., :n]).sum().item())

    # --------------------------------------------------------------- solve
    def rayleigh_ritz_fp64(self) -> SolveReport:
        """FP64 Rayleigh-Ritz on the current subspace; updates C (complex128) and eig."""
        C = self.C.to(torch.complex128) if self.C.coeff.dtype != torch.complex128 else self.C
        HC = self.hamiltonian(C)
        eig, V = eighg(C ^ HC, C.dot_O(C).wait())
        C = C @ V
        HC = HC @ V
        self.C = C
        self.eig = eig
        resid = HC.coeff - C.coeff * eig[..., None, None].to(C.coeff.real.dtype)
        rnorm = torch.view_as_real(resid).square().sum(dim=(-3, -2, -1)).sqrt()
        return SolveReport(eig[..., : self.fillings.n_bands], rnorm[..., : self.fillings.n_bands],
                           0, self.deig_max, self.get_Eband())

    def solve(self, fp64_finish: Optional[bool] = None) -> SolveReport:
        """Run the diagonalizer; mixed form (complex64) ends with FP64 Rayleigh-Ritz."""
        self.diagonalize()
        diag = self.diagonalize
        if fp64_finish is None:
            fp64_finish = self.C.coeff.dtype != torch.complex128
        if fp64_finish:
            rep = self.rayleigh_ritz_fp64()
        else:
            HC = getattr(diag, "last_HC", None)
            if HC is None or HC.n_bands() != self.C.n_bands():
                HC = self.hamiltonian(self.C)
            resid = HC.coeff - self.C.coeff * self.eig[..., None, None].to(self.C.coeff.real.dtype)
            rnorm = torch.view_as_real(resid).square().sum(dim=(-3, -2, -1)).sqrt().to(torch.float64)
            n = self.fillings.n_bands
            rep = SolveReport(self.eig[..., :n], rnorm[..., :n], 0, self.deig_max, self.get_Eband())
        rep.n_iterations = diag._i_iter
        rep.history = list(diag.history)
        return rep


def electrons_from_spec(spec, dtype=torch.complex128, **kw) -> Electrons:
    """Basis, projectors and electrons of a synthetic cell, on the GPU."""
    rc.require_gpu()
    ions = Ions.from_spec(spec)
    basis = Basis.from_spec(spec, ions=ions)
    ions.update(basis)
    args = dict(
        n_bands=spec.n_bands, n_bands_extra=spec.n_bands_extra, dtype=dtype,
        n_iterations=spec.n_iterations, eig_threshold=spec.eig_threshold,
    )
    if spec.filter_order:
        args.update(diagonalizer="chefsi", filter_order=spec.filter_order)
    args.update(kw)
    return Electrons(basis, torch.as_tensor(spec.V), **args)
