"""Oracle Davidson and CheFSI (test infrastructure; see oracle/__init__.py).

``davidson`` restates dft/electrons/_davidson.py:141-235 for a fixed
Hamiltonian (not an inner SCF loop, so the target is n_bands):
eigh of C^H HC and rotation (:164-168); per iteration the preconditioned
residual of the unconverged bands (:178-183, precondition :93-99),
band norms and regularisation of null bands (:184-185, :101-115), expansion
(:186-192), TileExpansion overlaps (:196-203, :240-264), eighg and
truncation to n_bands + n_bands_extra (:206-209), rotations (:212-215) and
the convergence test (:217-222, _check_deig :126-139).

``chefsi`` restates dft/electrons/_chefsi.py:72-176: Davidson warm start to
init_threshold, filter bounds b_up = ke_cutoff + max V, b_lo = max eig,
a_lo = min eig, the scaled three-term Chebyshev recurrence with
filter_order - 1 H applications, and Rayleigh-Ritz via eighg(C^H HC, C^H C).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Optional

import numpy as np

from .basis import Basis
from .hamiltonian import dot, matmul
from .init import randomize, randomize_selected
from .linalg import dagger, eigh_lower, eighg


@dataclass
class SolveResult:
    eig: np.ndarray  #: (K, n) eigenvalues of the retained subspace
    C: np.ndarray
    HC: np.ndarray
    deig_max: float
    n_iterations: int
    converged: bool
    n_eigs_done: int
    history: list = field(default_factory=list)


def band_ke(basis: Basis, C: np.ndarray) -> np.ndarray:
    return np.einsum("kbg,kg->kb", np.abs(C) ** 2, basis.KE)


def band_norm(C: np.ndarray) -> np.ndarray:
    return np.sqrt((np.abs(C) ** 2).sum(-1))


def precondition(basis: Basis, Cerr: np.ndarray, KEref: np.ndarray) -> np.ndarray:
    x = basis.KE[:, None, :] / KEref[..., None]
    x = x + np.exp(-x)
    return Cerr / x


def tile_expansion(C_XC: np.ndarray, Cnew_XCexp: np.ndarray) -> np.ndarray:
    K, n_new, n_exp = Cnew_XCexp.shape
    n_cur = n_new - n_exp
    out = np.zeros((K, n_new, n_new), dtype=Cnew_XCexp.dtype)
    out[:, :n_cur, :n_cur] += C_XC
    out[:, :n_cur, n_cur:] = Cnew_XCexp[:, :n_cur]
    out[:, n_cur:, :n_cur] = dagger(Cnew_XCexp[:, :n_cur])
    out[:, n_cur:, n_cur:] = Cnew_XCexp[:, n_cur:]
    return out


def check_deig(deig: np.ndarray, n_bands: int, threshold: float):
    d = deig[:, :n_bands]
    deig_max = float(d.max())
    pending = np.where((d > threshold).any(axis=0))[0]
    return deig_max, int(pending[0]) if len(pending) else n_bands


def initial_state(basis: Basis, n_bands: int, seed: int = 0) -> np.ndarray:
    """Random, bandwidth-limited, Cholesky-orthonormalized start
    (_electrons.py:279-287)."""
    C = np.zeros((len(basis.k), n_bands, basis.n_tot), dtype=np.complex128)
    randomize(basis, C, seed=seed)
    from .linalg import ortho_matrix

    return matmul(C, ortho_matrix(dot(C, C)))


def davidson(
    basis: Basis,
    H: Callable[[np.ndarray], np.ndarray],
    C: np.ndarray,
    n_bands: int,
    n_bands_extra: int,
    n_iterations: int = 100,
    eig_threshold: float = 1e-8,
    log: Optional[list] = None,
) -> SolveResult:
    n_bands_max = n_bands + n_bands_extra
    if 2 * n_bands_max >= basis.n_min:
        raise ValueError(
            f"n_bands + n_bands_extra = {n_bands_max} exceeds"
            f" min(n_basis)/2 = {basis.n_min // 2} in Davidson"
        )
    norm_cut = np.sqrt(basis.n_tot * 1e-15)
    HC = H(C)
    eig, V = eigh_lower(dot(C, HC))
    C = matmul(C, V)
    HC = matmul(HC, V)
    n_eigs_done, deig_max, converged, i_iter = 0, np.inf, False, 0
    history = []
    for i_iter in range(1, n_iterations + 1):
        n_cur = C.shape[1]
        eig_sel = eig[:, n_eigs_done:, None]
        C_sel = C[:, n_eigs_done:]
        HC_sel = HC[:, n_eigs_done:]
        KEref = band_ke(basis, C_sel)
        Cexp = precondition(basis, HC_sel - C_sel * eig_sel, KEref)
        norm = band_norm(Cexp)
        low_k, low_b = np.where(norm < norm_cut)
        if len(low_k):
            randomize_selected(basis, Cexp, low_k, low_b, i_iter)
            norm[low_k, low_b] = 1.0
        Cexp *= 1.0 / norm[..., None]
        basis.constrain(Cexp)
        n_exp = Cexp.shape[1]
        n_new = n_cur + n_exp
        Cnew = np.concatenate((C, Cexp), axis=1)
        O_new = tile_expansion(np.eye(n_cur)[None], dot(Cnew, Cexp))
        HCexp = H(Cexp)
        H_new = tile_expansion(
            np.stack([np.diag(e) for e in eig]), dot(Cnew, HCexp)
        )
        eig_new, V_new = eighg(H_new, O_new)
        n_next = min(n_new, n_bands_max)
        V_new = V_new[..., :n_next]
        Vcur, Vexp = V_new[:, :n_cur], V_new[:, n_cur:]
        C = matmul(Cnew, V_new)
        HC = matmul(HC, Vcur) + matmul(HCexp, Vexp)
        deig = np.abs(eig - eig_new[:, :n_cur])
        eig = eig_new[:, :n_next]
        deig_max, n_eigs_done = check_deig(deig, n_bands, eig_threshold)
        history.append((i_iter, deig_max, n_eigs_done))
        if log is not None:
            log.append(history[-1])
        converged = n_eigs_done >= n_bands
        if converged:
            break
    return SolveResult(eig, C, HC, deig_max, i_iter, converged, n_eigs_done, history)


def chefsi(
    basis: Basis,
    H: Callable[[np.ndarray], np.ndarray],
    C: np.ndarray,
    n_bands: int,
    n_bands_extra: int,
    V_max: float,
    filter_order: int = 10,
    n_iterations: int = 100,
    eig_threshold: float = 1e-8,
    init_threshold: float = 1e-1,
) -> SolveResult:
    warm = davidson(
        basis, H, C, n_bands, n_bands_extra, n_iterations, init_threshold
    )
    C, HC, eig = warm.C, warm.HC, warm.eig
    i_start = warm.n_iterations
    b_up = basis.spec.ke_cutoff + V_max
    n_eigs_done, deig_max, converged = 0, np.inf, False
    i_iter = i_start
    history = list(warm.history)
    for i_iter in range(i_start + 1, n_iterations + 1):
        b_lo = eig.max(axis=1)[:, None, None]
        a_lo = eig.min(axis=1)[:, None, None]
        if n_eigs_done:
            C0, C = C[:, :n_eigs_done], C[:, n_eigs_done:]
            HC0, HC = HC[:, :n_eigs_done], HC[:, n_eigs_done:]
        b_diff = 0.5 * (b_up - b_lo)
        b_mid = 0.5 * (b_lo + b_up)
        sigma = b_diff / (b_mid - a_lo)
        tau = 2.0 / sigma
        Y = (HC - b_mid * C) * (sigma / b_diff)
        for _ in range(filter_order - 2):
            sigma_new = 1.0 / (tau - sigma)
            HY = H(Y)
            Yt = (HY - b_mid * Y) * (2.0 * sigma_new / b_diff)
            Yt -= (sigma * sigma_new) * C
            C, Y, sigma = Y, Yt, sigma_new
        HC = H(Y)
        C = Y
        if n_eigs_done:
            C = np.concatenate((C0, C), axis=1)
            HC = np.concatenate((HC0, HC), axis=1)
        eig_prev = eig
        eig, V = eighg(dot(C, HC), dot(C, C))
        C = matmul(C, V)
        HC = matmul(HC, V)
        deig = np.abs(eig - eig_prev)
        deig_max, n_eigs_done = check_deig(deig, n_bands, eig_threshold)
        history.append((i_iter, deig_max, n_eigs_done))
        converged = n_eigs_done == n_bands
        if converged:
            break
    return SolveResult(eig, C, HC, deig_max, i_iter, converged, n_eigs_done, history)


def e_band(basis: Basis, eig: np.ndarray, n_bands: int) -> float:
    """sum_{k} w_sk sum_b eig (_davidson.py:117-124)."""
    return float((basis.w_sk[:, None] * eig[:, :n_bands]).sum())
