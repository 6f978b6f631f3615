"""Per-band eigenvalue error estimates from residual norms (host, numpy).

For a normalised Ritz vector x with Rayleigh quotient theta = x^H H x and
residual r = ||H x - theta x||, the Kato-Temple inequality bounds the error
of theta against the eigenvalue lambda it approximates,
    |theta - lambda| <= r^2 / delta,
where delta is the distance from theta to the rest of the spectrum. With
near-degenerate Ritz values the single-vector gap is tiny and the bound
useless; for a cluster of Ritz values the same form holds with the
cluster's block residual r_c^2 = sum of its r^2 and the cluster's distance
to the other eigenvalues (the block Kato-Temple / sin-theta argument). The
unknown spectrum is approximated by the neighbouring Ritz values, reduced
by their own residuals so that the gap is not overstated; where no
neighbour is known, or the reduced gap does not exceed r_c, only the
residual bound |theta - lambda| <= r_c is used.

The batched band solver stops a mixed-precision job when every wanted band
of every k-point has an estimate below its tolerance, and every result
carries the estimate of its final FP64 Rayleigh-Ritz. QimPy's Davidson stops
on the eigenvalue change between iterations instead (dft/electrons/
_davidson.py:126-139); in complex64 that change has a rounding floor near
1e-6 Ha, which the residual-based estimate does not share.
"""
from __future__ import annotations

import numpy as np

#: Ritz values closer than this (Ha) are treated as one cluster.
CLUSTER_TOL = 1e-3


def eig_error_estimate(theta: np.ndarray, r: np.ndarray, n_want: int,
                       cluster_tol: float = CLUSTER_TOL) -> np.ndarray:
    """Error estimates of the n_want lowest Ritz values of each row.

    theta: (..., m) ascending Ritz values; r: (..., m) residual norms of
    their normalised vectors (m >= n_want; values beyond n_want serve as
    the neighbours of the top wanted bands). Returns (..., n_want)."""
    theta = np.asarray(theta, dtype=np.float64)
    r = np.asarray(r, dtype=np.float64)
    lead = theta.shape[:-1]
    m = theta.shape[-1]
    T = theta.reshape(-1, m)
    Rr = r.reshape(-1, m)
    rows = T.shape[0]
    brk = np.diff(T, axis=1) > cluster_tol  # a cluster ends between c and c + 1
    first = np.concatenate([np.ones((rows, 1), bool), brk], axis=1)
    last = np.concatenate([brk, np.ones((rows, 1), bool)], axis=1)
    gid = np.cumsum(first.ravel()) - 1  # cluster id of every entry
    n_cl = int(gid[-1]) + 1 if gid.size else 0
    rc = np.sqrt(np.bincount(gid, weights=(Rr**2).ravel(), minlength=n_cl))
    # Below the lowest Ritz value no eigenvalue is missing; above the last one
    # the spectrum is unknown (gap 0: the residual bound only).
    lo = np.full((rows, m), np.inf)
    lo[:, 1:] = T[:, 1:] - T[:, :-1] - Rr[:, :-1]
    hi = np.zeros((rows, m))
    hi[:, :-1] = T[:, 1:] - T[:, :-1] - Rr[:, 1:]
    gap = np.empty(n_cl)
    gap[gid[first.ravel()]] = lo.ravel()[first.ravel()]
    gap_hi = np.empty(n_cl)
    gap_hi[gid[last.ravel()]] = hi.ravel()[last.ravel()]
    gap = np.minimum(gap, gap_hi)
    with np.errstate(divide="ignore", invalid="ignore"):
        est = np.where(gap > rc, rc * rc / np.where(gap > 0, gap, 1.0), rc)
    return est[gid].reshape(rows, m)[:, :n_want].reshape(lead + (n_want,))
