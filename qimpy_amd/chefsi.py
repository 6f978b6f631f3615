"""Chebyshev-filtered subspace iteration (reference: dft/electrons/_chefsi.py).

Drop-in for QimPy's CheFSI: the same constructor (``filter_order``,
``init_threshold``), call signature and side effects on ``electrons.C``,
``eig`` and ``deig_max``, and the same mathematics per outer iteration
(_chefsi.py:72-176):

1. a Davidson warm start while ``deig_max > init_threshold``, else a Ritz
   rotation of the current bands;
2. the filter window of each spin and k-point: the spectrum above the
   largest current Ritz value, up to ``ke_cutoff + max V(r)``, is damped,
   with the filter scaled at the smallest Ritz value;
3. ``filter_order - 1`` applications of H in the scaled three-term
   Chebyshev recurrence to the bands not yet converged;
4. Rayleigh-Ritz on all bands in FP64 and a convergence count of the leading
   bands whose eigenvalue changed by less than the threshold.

Here the window is a small object of per-(spin, k) device tensors
(``_Window``), and every recurrence step after the first H application is
a single fused pass over the bands (``qmx_axpby``) that writes the new
iterate over the one two steps back, so a filter holds three band blocks
(the current and previous iterates and H applied to the current) however
high its order.
"""
from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import numpy as np
import torch

from . import _lib, rc
from .basis import DTYPE_CODE
from .davidson import Davidson
from .linalg import eighg, eighg_robust
from .wavefunction import Wavefunction


@dataclass
class _Window:
    """Chebyshev filter window per (spin, k): damp [lower, upper], scale the
    filter at `anchor` (below the window). Tensors are float64 (S, K)."""

    center: torch.Tensor  #: (lower + upper) / 2
    half: torch.Tensor  #: (upper - lower) / 2
    sigma0: torch.Tensor  #: half / (center - anchor)

    @classmethod
    def from_ritz(cls, eig: torch.Tensor, upper: float) -> "_Window":
        lower = eig.max(dim=2)[0]
        anchor = eig.min(dim=2)[0]
        center = 0.5 * (lower + upper)
        half = 0.5 * (upper - lower)
        return cls(center, half, half / (center - anchor))


def _combine(a: torch.Tensor, X: Wavefunction, b: Optional[torch.Tensor] = None,
             Y: Optional[Wavefunction] = None, c: Optional[torch.Tensor] = None,
             Z: Optional[Wavefunction] = None, out: Optional[Wavefunction] = None) -> Wavefunction:
    """out = a X + b Y + c Z with coefficients per (spin, k) (float64 (S, K)
    tensors) in one pass (qmx_axpby); out may be one of the operands."""
    x = X.coeff
    S, K, B, Xs, G = x.shape
    for W in (Y, Z):
        if W is not None and W.coeff.shape != x.shape:
            raise ValueError("combine: operands of different shapes")
    blocks = [W.coeff if W is not None else None for W in (X, Y, Z)]
    if out is None:
        out = Wavefunction(X.basis, coeff=torch.empty_like(x))
    blocks = [t if t is None or t.is_contiguous() else t.contiguous() for t in blocks]
    o = out.coeff
    if not o.is_contiguous():
        raise ValueError("combine: out must be contiguous")
    coef = [None if t is None else t.to(torch.float64).reshape(S * K).contiguous()
            for t in (a, b, c)]
    _lib.call("qmx_axpby", DTYPE_CODE[x.dtype], S * K, B, Xs * G,
              coef[0].data_ptr(), None if coef[1] is None else coef[1].data_ptr(),
              None if coef[2] is None else coef[2].data_ptr(), 1, 0,
              blocks[0].data_ptr(), None if blocks[1] is None else blocks[1].data_ptr(),
              None if blocks[2] is None else blocks[2].data_ptr(), o.data_ptr(), Xs * G,
              B * Xs * G, rc.stream_ptr())
    out._proj_invalidate()
    return out


class CheFSI(Davidson):
    def __init__(self, *, electrons, n_iterations: int = 100, eig_threshold: float = 1e-8,
                 filter_order: int = 10, init_threshold: float = 1e-1,
                 residual_threshold: Optional[float] = None):
        super().__init__(electrons=electrons, n_iterations=n_iterations,
                         eig_threshold=eig_threshold, residual_threshold=residual_threshold)
        self.filter_order = filter_order
        self.init_threshold = init_threshold
        self._line_prefix = "CheFSI"

    def __repr__(self) -> str:
        return (f"CheFSI(n_iterations: {self.n_iterations}, eig_threshold: {self.eig_threshold:g},"
                f" filter_order: {self.filter_order}, init_threshold: {self.init_threshold:g})")

    # ------------------------------------------------------------ pieces
    def _start(self, n_iterations: int) -> Wavefunction:
        """Bands and H C to filter from: Davidson down to init_threshold
        (its residual stop off), or the Ritz rotation of the current bands
        when they are already that good (_chefsi.py:86-101)."""
        el = self.electrons
        if el.deig_max > self.init_threshold:
            self._line_prefix = "ChefSI_init"
            res_thr, self.residual_threshold = self.residual_threshold, None
            try:
                super().__call__(n_iterations, self.init_threshold)
            finally:
                self.residual_threshold = res_thr
            self._line_prefix = "CheFSI"
            HC, self._HC = self._HC, None
            return HC
        HC = self._rotate_to_ritz(el.C)
        el.deig_max = np.inf
        self._i_iter = 0
        self.history = []
        self._report()
        return HC

    def _filter(self, C: Wavefunction, HC: Wavefunction, w: _Window) -> Wavefunction:
        """The scaled Chebyshev polynomial of H of degree filter_order - 1
        applied to C (_chefsi.py:117-140): Y_1 = (HC - c C) s_1 / e, then
        Y_{n+1} = (2 s_{n+1} / e)(H Y_n - c Y_n) - s_n s_{n+1} Y_{n-1} with
        s_{n+1} = 1 / (2 / s_1 - s_n); c, e the window's center and half
        width. Returns the last iterate."""
        el = self.electrons
        c, e, s1 = w.center, w.half, w.sigma0
        Y = _combine(s1 / e, HC, -c * s1 / e, C)
        prev, s = C, s1
        for _ in range(self.filter_order - 2):
            s_next = 1.0 / (2.0 / s1 - s)
            HY = el.hamiltonian(Y)
            a = 2.0 * s_next / e
            # The new iterate overwrites the one two steps back (not C itself).
            target = prev if prev is not C else None
            Ynew = _combine(a, HY, -a * c, Y, -s * s_next, prev, out=target)
            del HY
            prev, Y, s = Y, Ynew, s_next
        return Y

    def _residual_locked(self, HC: Wavefunction, n_done: int) -> tuple[int, bool]:
        """Mixed form: the eigenvalue change stalls at the complex64 rounding
        of the subspace, so bands lock (leave the filter) and the solve stops
        by residual ||HC - eig C|| below residual_threshold, counted from the
        lowest band; the FP64 Rayleigh-Ritz that follows sets the final
        eigenvalues."""
        el = self.electrons
        n = el.fillings.n_bands
        rdt = el.C.coeff.real.dtype
        R = HC.coeff[:, :, :n] - el.C.coeff[:, :, :n] * el.eig[:, :, :n, None, None].to(rdt)
        r = torch.linalg.vector_norm(R, dim=(-2, -1))
        bad = torch.nonzero((r > self.residual_threshold).flatten(0, 1).any(dim=0))
        n_res = int(bad[0, 0].item()) if len(bad) else n
        return max(n_done, n_res), n_res >= n

    # ------------------------------------------------------------ iteration
    def __call__(self, n_iterations: Optional[int] = None,
                 eig_threshold: Optional[float] = None) -> None:
        el = self.electrons
        n_bands = el.fillings.n_bands
        limit = n_iterations or self.n_iterations
        threshold = eig_threshold or self.eig_threshold
        HC = self._start(limit)
        upper = el.basis.ke_cutoff + el.V_max  # above every eigenvalue of H
        n_done = 0
        for self._i_iter in range(self._i_iter + 1, limit + 1):
            w = _Window.from_ritz(el.eig, upper)
            if n_done:  # converged bands stay out of the filter
                C_fixed, HC_fixed = el.C[:, :, :n_done], HC[:, :, :n_done]
                C_act, HC_act = el.C[:, :, n_done:], HC[:, :, n_done:]
            else:
                C_act, HC_act = el.C, HC
            Y = self._filter(C_act, HC_act, w)
            del C_act, HC_act
            HY = el.hamiltonian(Y)
            if n_done:
                C, HC = C_fixed.cat(Y, clear=True), HC_fixed.cat(HY, clear=True)
            else:
                C, HC = Y, HY
            del Y, HY
            eig_prev = el.eig
            solve = eighg_robust if C.coeff.dtype == torch.complex64 else eighg
            el.eig, V = solve(C ^ HC, C.dot_O(C).wait())
            el.C = C @ V
            HC = HC @ V
            del C
            el.deig_max, n_done = self._check_deig((el.eig - eig_prev).abs(), threshold)
            converged = n_done == n_bands
            if not converged and self.residual_threshold is not None:
                n_done, converged = self._residual_locked(HC, n_done)
            self._report(n_done, converged)
            if converged:
                break
        self.last_HC = HC
        if type(self) is not CheFSI:
            self._HC = HC

    diagonalize = __call__
