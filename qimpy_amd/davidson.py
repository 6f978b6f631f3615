"""Davidson diagonalization (reference: dft/electrons/_davidson.py).

Same algorithm and side effects as the reference: it mutates
``electrons.C``, ``electrons.eig`` (float64 (S, K, n)) and
``electrons.deig_max``. The residual, inverse-kinetic preconditioner and
expansion norms run as one fused HIP pass (``qmx_precond_residual``), the
normalisation and padding constraint as another (``qmx_scale_constrain``);
H*psi, the overlaps and rotations run on the GPU in the wavefunction
precision, and eighg/eigh always in FP64.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib, rc
from .basis import DTYPE_CODE
from .linalg import dagger, eigh, eighg, eighg_robust
from .wavefunction import Wavefunction


class _PhaseTimer:
    """Accumulates GPU time per solver phase with events on the current stream
    (enabled only when a caller asks for a breakdown)."""

    def __init__(self):
        self.enabled = False
        self.times: dict = {}
        self._pending: list = []

    def mark(self, name: str):
        if not self.enabled:
            return None
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return (name, ev)

    def done(self, token) -> None:
        if token is None:
            return
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self._pending.append((token[0], token[1], ev))

    def collect(self) -> dict:
        torch.cuda.synchronize()
        for name, a, b in self._pending:
            self.times[name] = self.times.get(name, 0.0) + a.elapsed_time(b)
        self._pending = []
        return self.times


class Davidson:
    """Davidson diagonalization of the Hamiltonian in `electrons`."""

    def __init__(self, *, electrons, n_iterations: int = 100, eig_threshold: float = 1e-8,
                 canonical_overlap: Optional[bool] = None,
                 residual_threshold: Optional[float] = None):
        self.electrons = electrons
        #: When set, bands count as converged once ||H C - eig C|| (accumulated
        #: in double) is below this, instead of the eigenvalue change of
        #: _davidson.py:126-139: in float32 the eigenvalue change has a noise
        #: floor near 1e-6 Ha that an FP64-sized threshold never passes.
        self.residual_threshold = residual_threshold
        #: Solve the expanded subspace by canonical orthogonalisation instead
        #: of Cholesky; defaults to on for complex64 wavefunctions, whose
        #: float32-accumulated overlaps can lose positive definiteness.
        self.canonical_overlap = canonical_overlap
        self.n_iterations = n_iterations
        self.eig_threshold = eig_threshold
        self._line_prefix = "Davidson"
        self._norm_cut = np.sqrt(electrons.basis.n_tot * 1e-15)  # :56-58
        self._i_iter = 0
        self.history: list = []
        self.timer = _PhaseTimer()

    def __repr__(self) -> str:
        return f"Davidson(n_iterations: {self.n_iterations}, eig_threshold: {self.eig_threshold:g})"

    def precondition(self, Cerr: Wavefunction, KEref: torch.Tensor) -> Wavefunction:
        """Inverse-kinetic preconditioner (:92-99)."""
        basis = self.electrons.basis
        x = basis.get_ke()[None, :, None, None, :] / KEref[..., None, None]
        x += torch.exp(-x)
        return Cerr * (1.0 / x).to(Cerr.coeff.real.dtype)

    def _residual_expansion(
        self, C_sel: Wavefunction, HC_sel: Wavefunction, eig_sel: torch.Tensor,
        KEref: torch.Tensor,
    ) -> tuple[Wavefunction, torch.Tensor]:
        """precondition(HC_sel - C_sel eig_sel, KEref) and its band norms (:182-184),
        fused into one pass over C and HC."""
        c, h = C_sel.coeff, HC_sel.coeff
        S, K, B, X, G = c.shape
        out = torch.empty((S, K, B, X, G), dtype=c.dtype, device=c.device)
        norm2 = torch.empty((S, K, B), dtype=torch.float64, device=c.device)
        rnorm2 = torch.empty((S, K, B), dtype=torch.float64, device=c.device)
        ke = self.electrons.basis.get_ke()
        eig = eig_sel.contiguous()
        kref = KEref.contiguous()
        for s in range(S):
            cs, hs = c[s], h[s]
            if cs.stride() != hs.stride() or cs.stride(-1) != 1:
                hs = hs.contiguous()
                cs = cs.contiguous()
            _lib.call(
                "qmx_precond_residual", DTYPE_CODE[c.dtype], cs.data_ptr(), hs.data_ptr(),
                cs.stride(1), cs.stride(0), eig[s].data_ptr(), kref[s].data_ptr(),
                ke.data_ptr(), ke.stride(0), out[s].data_ptr(), out.stride(2), out.stride(1),
                K, B, G, norm2[s].data_ptr(), rnorm2[s].data_ptr(), rc.stream_ptr(),
            )
        self.residual_norm = rnorm2.sqrt()
        return Wavefunction(C_sel.basis, coeff=out), norm2.sqrt()

    def _orthogonalize(self, C: Wavefunction) -> bool:
        ortho = self.canonical_overlap
        return (C.coeff.dtype == torch.complex64) if ortho is None else ortho

    def _project_out(self, C: Wavefunction, Cexp: Wavefunction, drop_below: float = 1e-3) -> None:
        """Cexp <- (1 - C C^H) Cexp twice (block classical Gram-Schmidt with
        re-orthogonalisation), then renormalise. Directions that keep less
        than `drop_below` of their norm lie in span(C) up to float32 rounding;
        they are re-drawn at random (as _regularize does for null bands) and
        projected again. Used for float32 subspaces: the reference's expansion
        (_davidson.py:183-197) leaves the near-dependence between C and Cexp
        to its FP64 generalized eigensolve, which float32 overlaps cannot
        resolve. In exact arithmetic the span, and so the Ritz values, are
        unchanged."""
        for attempt in range(2):
            for _ in range(2):
                M = (C ^ Cexp).wait()
                Cexp -= C @ M
            norm = Cexp.band_norm()
            low = norm < drop_below
            i_spin, i_k, i_band = torch.where(low)
            if attempt or not len(i_spin):
                break
            Cexp.randomize_selected(i_spin, i_k, i_band, 1000 + self._i_iter)
            Cexp.scale_constrain_(1.0 / Cexp.band_norm())
        Cexp.scale_constrain_(1.0 / norm.clamp_min(1e-300))

    def _regularize(self, C: Wavefunction, norm: torch.Tensor, i_iter: int) -> None:
        """Randomize null bands, seeded by the iteration (:101-115)."""
        low_norm = norm < self._norm_cut
        i_spin, i_k, i_band = torch.where(low_norm)
        if not len(i_spin):
            return
        C.randomize_selected(i_spin, i_k, i_band, i_iter)
        norm[i_spin, i_k, i_band] = 1.0

    def get_Eband(self) -> float:
        el = self.electrons
        n_bands = el.fillings.n_bands
        return float((el.basis.w_sk * el.eig[..., :n_bands]).sum().item())

    def _check_deig(self, deig: torch.Tensor, eig_threshold: float) -> tuple[float, int]:
        n_bands = self.electrons.fillings.n_bands
        d = deig[..., :n_bands]
        deig_max = float(d.max().item())
        pending = torch.where((d > eig_threshold).flatten(0, 1).any(dim=0))[0]
        n_eigs_done = int(pending[0].item()) if len(pending) else n_bands
        return deig_max, n_eigs_done

    def _report(self, n_eigs_done: int = 0, converged: bool = False) -> None:
        self.history.append((self._i_iter, self.electrons.deig_max, n_eigs_done))

    def __call__(self, n_iterations: Optional[int] = None, eig_threshold: Optional[float] = None) -> None:
        el = self.electrons
        n_bands = el.fillings.n_bands
        n_bands_max = n_bands + el.fillings.n_bands_extra
        helper = type(self) != Davidson
        inner_loop = not (helper or ((n_iterations is None) and (eig_threshold is None)))
        n_iterations = n_iterations if n_iterations else self.n_iterations
        eig_threshold = eig_threshold if eig_threshold else self.eig_threshold
        n_eigs_target = el.fillings.n_bands_min if inner_loop else n_bands

        if 2 * n_bands_max >= el.basis.n_min:
            raise ValueError(
                f"n_bands + n_bands_extra = {n_bands_max} exceeds"
                f" min(n_basis)/2 = {el.basis.n_min // 2} in Davidson"
            )
        self.history = []
        HC = el.hamiltonian(el.C)
        el.eig, V = eigh((el.C ^ HC).wait().to(torch.complex128))
        el.deig_max = np.inf
        el.C = el.C @ V
        HC = HC @ V
        self._i_iter = 0
        self._report()
        n_eigs_done = 0

        for self._i_iter in range(self._i_iter + 1, n_iterations + 1):
            n_bands_cur = el.C.n_bands()
            eig_sel = el.eig[:, :, n_eigs_done:]
            C_sel = el.C[:, :, n_eigs_done:] if n_eigs_done else el.C
            HC_sel = HC[:, :, n_eigs_done:] if n_eigs_done else HC
            tk = self.timer.mark("residual")
            KEref = C_sel.band_ke()
            Cexp, norm_exp = self._residual_expansion(C_sel, HC_sel, eig_sel, KEref)
            self.timer.done(tk)
            res_thr = self.residual_threshold
            if res_thr is not None:
                # residuals of the bands not yet locked (index >= n_eigs_done):
                r = self.residual_norm[..., : max(n_bands - n_eigs_done, 0)]
                pending = torch.where((r > res_thr).flatten(0, 1).any(dim=0))[0]
                n_res_done = n_eigs_done + (
                    int(pending[0].item()) if len(pending) else n_bands - n_eigs_done)
                if n_res_done >= n_eigs_target:
                    self._report(n_res_done, True)
                    break
            self._regularize(Cexp, norm_exp, self._i_iter)
            Cexp.scale_constrain_(1.0 / norm_exp)
            if self._orthogonalize(el.C):
                tk = self.timer.mark("project")
                self._project_out(el.C, Cexp)
                self.timer.done(tk)
            n_bands_exp = Cexp.n_bands()
            n_bands_new = n_bands_cur + n_bands_exp

            Cnew = el.C.cat(Cexp, clear=True)
            Cexp = Cnew[:, :, n_bands_cur:]

            tk = self.timer.mark("overlap")
            C_OC = torch.eye(n_bands_cur, device=rc.device, dtype=torch.float64)[None, None]
            C_OC_new = TileExpansion(C_OC, Cnew.dot_O(Cexp))
            self.timer.done(tk)

            tk = self.timer.mark("hamiltonian")
            HCexp = el.hamiltonian(Cexp)
            self.timer.done(tk)
            del Cexp
            tk = self.timer.mark("overlap")
            C_HC = torch.diag_embed(el.eig)
            C_HC_new = TileExpansion(C_HC, Cnew ^ HCexp)
            self.timer.done(tk)
            tk = self.timer.mark("eighg")

            canonical = self.canonical_overlap
            if canonical is None:
                canonical = Cnew.coeff.dtype == torch.complex64
            if canonical:
                eig_new, V_new = eighg_robust(C_HC_new, C_OC_new.wait())
            else:
                eig_new, V_new = eighg(C_HC_new, C_OC_new.wait())
            self.timer.done(tk)
            n_bands_next = min(n_bands_new, n_bands_max)
            V_new = V_new[..., :n_bands_next]
            Vcur, Vexp = V_new.split((n_bands_cur, n_bands_exp), dim=2)

            tk = self.timer.mark("rotate")
            el.C = Cnew @ V_new
            del Cnew
            HC = HC @ Vcur
            HC += HCexp @ Vexp
            del HCexp
            self.timer.done(tk)
            deig = torch.abs(el.eig - eig_new[..., :n_bands_cur])
            el.eig = eig_new[..., :n_bands_next]

            tk = self.timer.mark("check")
            el.deig_max, n_eigs_done = self._check_deig(deig, eig_threshold)
            self.timer.done(tk)
            if res_thr is not None:
                n_eigs_done = n_res_done  # lock by residual; test at the next start
            converged = res_thr is None and n_eigs_done >= n_eigs_target
            self._report(n_eigs_done, converged)
            if converged:
                break
        self.last_HC = HC  # H C of the final subspace, for residual reports
        if helper:
            self._HC = HC

    diagonalize = __call__


@dataclass
class TileExpansion:
    """[[C_XC, C_XCexp], [C_XCexp^H, Cexp_XCexp]] from Cnew^X(Cexp) (:240-264),
    assembled in complex128 for the FP64 subspace solve."""

    C_XC: torch.Tensor
    Cnew_XCexp: object

    def wait(self) -> torch.Tensor:
        Cnew_XCexp = self.Cnew_XCexp.wait().to(torch.complex128)
        n_spins, nk, n_bands_new, n_bands_exp = Cnew_XCexp.shape
        n_bands_cur = n_bands_new - n_bands_exp
        C_XCexp, Cexp_XCexp = Cnew_XCexp.split((n_bands_cur, n_bands_exp), dim=2)
        result = torch.zeros(
            (n_spins, nk, n_bands_new, n_bands_new), device=C_XCexp.device, dtype=torch.complex128
        )
        result[:, :, :n_bands_cur, :n_bands_cur] += self.C_XC
        result[:, :, :n_bands_cur, n_bands_cur:] = C_XCexp
        result[:, :, n_bands_cur:, :n_bands_cur] = dagger(C_XCexp)
        result[:, :, n_bands_cur:, n_bands_cur:] = Cexp_XCexp
        return result
