"""Oracle wavefunction initialisation (test infrastructure; see oracle/__init__.py).

Restates dft/electrons/_wavefunction_init.py for S = X = 1 on one process:
- ``xorshift_draw``: x ^= x << 13; x ^= x >> 7 (arithmetic); x ^= x << 17 on
  signed int64, u = 2^-64 x + 1/2 in float64 (:16-20);
- ``randn``: Box-Muller sqrt(-2 ln u0) exp(2 pi i u1) (:11-24);
- ``randomize``: state 1 + seed + g + n_max k (:75-93), discard
  b_start + 1 draws (:98-99), one draw per band, zero padding, bandwidth
  limit 1 / (1 + ((4/3) KE)^6) (:120-125);
- ``randomize_selected``: state 1 + seed + g + n_max (b + n_bands k), one
  discarded draw, padding of every band zeroed, bandwidth limit on the
  selected bands (:145-188).
"""
from __future__ import annotations

import numpy as np

from .basis import Basis


def xorshift_draw(x: np.ndarray) -> np.ndarray:
    """Advance int64 states in place and return the float64 uniforms."""
    x ^= x << np.int64(13)
    x ^= x >> np.int64(7)
    x ^= x << np.int64(17)
    return (0.5**64) * x.astype(np.float64) + 0.5


def randn(x: np.ndarray) -> np.ndarray:
    u0 = xorshift_draw(x)
    u1 = xorshift_draw(x)
    return np.sqrt(-2.0 * np.log(u0)) * np.exp((2j * np.pi) * u1)


def _bandwidth_limit(KE: np.ndarray) -> np.ndarray:
    return 1.0 / (1.0 + ((4.0 / 3) * KE) ** 6)


def randomize(
    basis: Basis, C: np.ndarray, seed: int = 0, b_start: int = 0, b_stop=None
) -> None:
    K, B, G = C.shape
    b_stop = B if b_stop is None else b_stop
    if b_start >= b_stop:
        return
    with np.errstate(over="ignore"):
        x = (1 + seed) + np.arange(G, dtype=np.int64)[None] + basis.n_max * np.arange(
            K, dtype=np.int64
        )[:, None]
        for _ in range(b_start + 1):
            randn(x)
        for b in range(b_start, b_stop):
            C[:, b] = randn(x)
    sel = C[:, b_start:b_stop]
    sel[np.broadcast_to(basis.pad_mask[:, None, :], sel.shape)] = 0.0
    sel *= _bandwidth_limit(basis.KE)[:, None, :]
    C[:, b_start:b_stop] = sel


def randomize_selected(
    basis: Basis, C: np.ndarray, i_k: np.ndarray, i_band: np.ndarray, seed: int
) -> None:
    K, B, G = C.shape
    with np.errstate(over="ignore"):
        x = (
            (1 + seed)
            + np.arange(G, dtype=np.int64)[None]
            + basis.n_max * (i_band.astype(np.int64) + B * i_k.astype(np.int64))[:, None]
        )
        randn(x)
        vals = randn(x)
    C[i_k, i_band] = vals
    # The padding of every band is zeroed, selected or not (:176-177):
    C[np.broadcast_to(basis.pad_mask[:, None, :], C.shape)] = 0.0
    C[i_k, i_band] *= _bandwidth_limit(basis.KE[i_k])
