"""Chebyshev filter subspace iteration (reference: dft/electrons/_chefsi.py).

Davidson warm start while ``deig_max > init_threshold`` (:86-92); filter
bounds b_up = ke_cutoff + max V(r) (:103-111), b_lo / a_lo the largest and
smallest current eigenvalue (:112-113); the scaled three-term Chebyshev
recurrence with filter_order - 1 H applications per outer iteration
(:121-146); Rayleigh-Ritz eighg(C^H HC, C^H C) in FP64 (:155-158).
"""
from __future__ import annotations

from typing import Optional

import numpy as np
import torch

from .davidson import Davidson
from .linalg import eigh, eighg, eighg_robust


class CheFSI(Davidson):
    def __init__(self, *, electrons, n_iterations: int = 100, eig_threshold: float = 1e-8,
                 filter_order: int = 10, init_threshold: float = 1e-1,
                 residual_threshold: Optional[float] = None):
        super().__init__(electrons=electrons, n_iterations=n_iterations, eig_threshold=eig_threshold,
                         residual_threshold=residual_threshold)
        self.filter_order = filter_order
        self.init_threshold = init_threshold
        self._line_prefix = "CheFSI"

    def __repr__(self) -> str:
        return (
            f"CheFSI(n_iterations: {self.n_iterations}, eig_threshold: {self.eig_threshold:g},"
            f" filter_order: {self.filter_order}, init_threshold: {self.init_threshold:g})"
        )

    def __call__(self, n_iterations: Optional[int] = None, eig_threshold: Optional[float] = None) -> None:
        el = self.electrons
        n_bands = el.fillings.n_bands
        n_iterations = n_iterations if n_iterations else self.n_iterations
        eig_threshold = eig_threshold if eig_threshold else self.eig_threshold
        if el.deig_max > self.init_threshold:
            self._line_prefix = "ChefSI_init"
            # The warm start converges on |deig| < init_threshold only (a
            # residual threshold belongs to the filter iterations below).
            res_thr, self.residual_threshold = self.residual_threshold, None
            try:
                super().__call__(n_iterations, self.init_threshold)
            finally:
                self.residual_threshold = res_thr
            self._line_prefix = "CheFSI"
            HC = self._HC
            del self._HC
        else:
            HC = el.hamiltonian(el.C)
            el.eig, V = eigh((el.C ^ HC).wait().to(torch.complex128))
            el.deig_max = np.inf
            el.C = el.C @ V
            HC = HC @ V
            self._i_iter = 0
            self.history = []
            self._report()

        KEmax = el.basis.ke_cutoff
        PEmax = el.V_max
        n_eigs_done = 0
        rdtype = el.C.coeff.real.dtype
        for self._i_iter in range(self._i_iter + 1, n_iterations + 1):
            b_up = KEmax + PEmax
            b_lo = el.eig.max(dim=2)[0]
            a_lo = el.eig.min(dim=2)[0]
            b_shape = a_lo.shape + (1, 1, 1)
            if n_eigs_done:
                C0, el.C = el.C[:, :, :n_eigs_done], el.C[:, :, n_eigs_done:]
                HC0, HC = HC[:, :, :n_eigs_done], HC[:, :, n_eigs_done:]
            b_diff = (0.5 * (b_up - b_lo)).view(b_shape)
            b_mid = (0.5 * (b_lo + b_up)).view(b_shape)
            sigma = b_diff / (b_mid - a_lo.view(b_shape))
            tau = 2 / sigma
            Y = (HC - el.C * b_mid.to(rdtype)) * (sigma / b_diff).to(rdtype)
            del HC
            for i_filter in range(self.filter_order - 2):
                sigma_new = 1 / (tau - sigma)
                HY = el.hamiltonian(Y)
                Yt = HY - Y * b_mid.to(rdtype)
                Yt *= (2 * sigma_new / b_diff).to(rdtype)
                Yt -= el.C * (sigma * sigma_new).to(rdtype)
                el.C = Y
                Y = Yt
                sigma = sigma_new
                del HY
            HC = el.hamiltonian(Y)
            el.C = Y
            del Y
            if n_eigs_done:
                el.C = C0.cat(el.C, clear=True)
                HC = HC0.cat(HC, clear=True)
                del C0, HC0
            eig_prev = el.eig
            if el.C.coeff.dtype == torch.complex64:
                el.eig, V = eighg_robust(el.C ^ HC, el.C.dot_O(el.C).wait())
            else:
                el.eig, V = eighg(el.C ^ HC, el.C.dot_O(el.C).wait())
            el.C = el.C @ V
            HC = HC @ V
            deig = torch.abs(el.eig - eig_prev)
            el.deig_max, n_eigs_done = self._check_deig(deig, eig_threshold)
            converged = n_eigs_done == n_bands
            if not converged and self.residual_threshold is not None:
                # Mixed form: |deig| stalls at the complex64 rounding of the
                # subspace, so bands lock (leave the filter) and the solve
                # stops by residual |HC - eig C| < threshold, counted from the
                # lowest band like the deig locking (the FP64 Rayleigh-Ritz
                # that follows sets the final eigenvalues).
                R = HC.coeff[:, :, :n_bands] - el.C.coeff[:, :, :n_bands] * el.eig[
                    :, :, :n_bands, None, None].to(rdtype)
                r = torch.linalg.vector_norm(R, dim=(-2, -1))  # (S, K, n_bands)
                pending = torch.where((r > self.residual_threshold).flatten(0, 1).any(dim=0))[0]
                n_res_done = int(pending[0].item()) if len(pending) else n_bands
                n_eigs_done = max(n_eigs_done, n_res_done)
                converged = n_res_done >= n_bands
            self._report(n_eigs_done, converged)
            if converged:
                break
        self.last_HC = HC
        if type(self) != CheFSI:
            self._HC = HC

    diagonalize = __call__
