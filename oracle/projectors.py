"""Oracle projector generation (test infrastructure; see oracle/__init__.py).

Restates Ions.get_projectors (dft/ions/_ions_projectors.py:14-72) with
spline radial functions:
    beta[k, atom * n_proj + p, G] = R_{i_rf[p]}(|G + k|) * Ytilde_{i_lm[p]}(G + k)
                                    * exp(-2 pi i (iG + k) . tau_atom) / sqrt(V)
zeroed on the padding (:67-70), where
- ``quintic_coeff`` restates quintic_spline.get_coeff (math/quintic_spline.py:
  12-53): blip coefficients of uniform samples, mirror-symmetric at the left
  end and natural (zero third and fourth derivative) at the right end, two
  padding coefficients each side;
- ``blip_eval`` evaluates them as Interpolator does (:56-128) but through
  scipy's B-spline (the blip is the cardinal quintic B-spline scaled to 1 at
  its centre, 120/66), not through the reference's polynomial table;
- ``real_solid_harmonics`` is the definition the reference tests its
  harmonics against (math/_spherical_harmonics_generate.py:10-27): scipy's
  complex Y_l^m times r^l and (-1)^m, real part at m >= 0 (times sqrt 2 for
  m > 0) and imaginary part times sqrt 2 at index l - m; ``harmonics_tilde``
  multiplies by i^l (math/spherical_harmonics.py:111-137).

``radial_transform`` restates RadialFunction.transform (math/
_radial_function.py:62-123; Simpson weights at :84-89) with scipy's
spherical_jn for j_l(x) / x^l (the reference's jl_by_xl, math/
spherical_bessel.py:8-49); pinned by the Gaussian closed form
f(r) = exp(-r^2 / 2 s^2) -> (2 pi)^(3/2) s^(2l + 3) exp(-s^2 G^2 / 2) for l <= 3.

Pinned (tests/test_oracle.py): the harmonics against closed forms for
l <= 2 and orthonormality on the sphere by exact quadrature for l <= 3; the
spline reproduces its samples at the nodes, meets the reference's own
accuracy test (math/test_quintic_spline.py: |f - interp| < dx^4 over a fine
grid), keeps the mirror symmetry and the zero third and fourth derivatives at
the right end; projectors of sampled Gaussian radials match the analytic
Gaussian projectors (oracle.hamiltonian.projectors).
"""
from __future__ import annotations

import numpy as np
import scipy.special
from scipy.interpolate import BSpline
from scipy.linalg import solve_banded

BLIP_SCALE = 120.0 / 66.0  # cardinal quintic B-spline at its centre is 66/120


def quintic_coeff(samples: np.ndarray) -> np.ndarray:
    """(N + 4, ...) blip coefficients of samples (..., N) on a uniform grid."""
    samples = np.asarray(samples, dtype=np.float64)
    N = samples.shape[-1]
    b1, b2 = 13.0 / 33.0, 1.0 / 66.0  # blip one and two nodes from its centre
    band = np.zeros((5, N))
    band[0], band[1], band[2], band[3], band[4] = b2, b1, 1.0, b1, b2
    # mirror at the left end: c[-1] = c[1], c[-2] = c[2]
    band[2, 1] += b2
    band[1, 1] += b1
    band[0, 2] += b2
    # natural right end: c[N] and c[N+1] extrapolated from the last three
    ext = np.array([[1.0, -3.0, 3.0], [3.0, -8.0, 6.0]])
    cols = (N - 3, N - 2, N - 1)
    for j, col in enumerate(cols):  # row N-2 gets b2 c[N]; row N-1 gets b1 c[N] + b2 c[N+1]
        band[2 + (N - 2) - col, col] += b2 * ext[0, j]
        band[2 + (N - 1) - col, col] += b1 * ext[0, j] + b2 * ext[1, j]
    flat = samples.reshape(-1, N).T
    c = solve_banded((2, 2), band, flat)
    c = np.vstack((c[2:0:-1], c, ext @ c[-3:]))
    return c.reshape((N + 4,) + samples.shape[:-1])


def blip_eval(coeff: np.ndarray, x: np.ndarray, dx: float) -> np.ndarray:
    """Spline of blip coefficients (N + 4, ...) at points x (any shape);
    result shape coeff.shape[1:] + x.shape."""
    n = coeff.shape[0]
    knots = (np.arange(n + 6) - 5.0) * dx
    flat = coeff.reshape(n, -1) * BLIP_SCALE
    out = BSpline(knots, flat, 5, extrapolate=False)(np.asarray(x).reshape(-1))  # (x, rest)
    return out.T.reshape(coeff.shape[1:] + np.asarray(x).shape)


def real_solid_harmonics(l_max: int, r: np.ndarray) -> np.ndarray:
    """((l_max + 1)^2, ...) r^l Y_lm(r), index l^2 + l + m."""
    r = np.asarray(r, dtype=np.float64)
    rmag = np.linalg.norm(r, axis=-1)
    safe = np.where(rmag > 0, rmag, 1.0)
    theta = np.arccos(np.clip(r[..., 2] / safe, -1.0, 1.0))
    phi = np.arctan2(r[..., 1], r[..., 0])
    out = []
    for l in range(l_max + 1):
        res = np.zeros((2 * l + 1,) + r.shape[:-1])
        for m in range(l + 1):
            y = ((-1) ** m) * rmag**l * scipy.special.sph_harm_y(l, m, theta, phi)
            if m == 0:
                res[l] = y.real
            else:
                res[l + m] = np.sqrt(2) * y.real
                res[l - m] = np.sqrt(2) * y.imag
        out.append(res)
    return np.concatenate(out, axis=0)


def harmonics_tilde(l_max: int, G: np.ndarray) -> np.ndarray:
    Y = real_solid_harmonics(l_max, G).astype(np.complex128)
    for l in range(l_max + 1):
        Y[l * l:(l + 1) ** 2] *= 1j ** l
    return Y


def projectors(basis, positions: np.ndarray, coeff: np.ndarray, dG: float,
               l_rf, volume: float) -> np.ndarray:
    """beta (K, n_atoms * n_proj, n_tot) complex128 for one species: radial
    functions with blip coefficients coeff (N + 4, n_rf) on spacing dG and
    angular momenta l_rf[i_rf]; per atom, radial function i_rf then m = -l..l
    (dft/ions/_pseudo_quantum_numbers.py:28-56)."""
    i_rf = [i for i, l in enumerate(l_rf) for _ in range(2 * l + 1)]
    i_lm = [l * (l + 1) + m for l in l_rf for m in range(-l, l + 1)]
    iGk = basis.iG + basis.k[:, None, :]
    Gk = iGk @ basis.Gbasis.T  # (K, G, 3)
    R = blip_eval(coeff, np.linalg.norm(Gk, axis=-1), dG)  # (n_rf, K, G)
    Y = harmonics_tilde(max(l_rf), Gk)  # (lm, K, G)
    atom = R[i_rf] * Y[i_lm]  # (n_proj, K, G)
    phase = np.exp(-2j * np.pi * (iGk @ np.asarray(positions).T)) / np.sqrt(volume)  # (K, G, A)
    beta = atom[None] * phase.transpose(2, 0, 1)[:, None]  # (A, n_proj, K, G)
    beta = beta.transpose(2, 0, 1, 3).reshape(iGk.shape[0], -1, iGk.shape[1])
    beta[np.broadcast_to(basis.pad_mask[:, None, :], beta.shape)] = 0.0
    return beta


def jl_by_xl(l: int, x: np.ndarray) -> np.ndarray:
    x = np.asarray(x, dtype=np.float64)
    out = np.empty_like(x)
    small = x < 1e-3
    dfact = np.prod(np.arange(1, 2 * l + 2, 2, dtype=np.float64))  # (2l + 1)!!
    xs = x[small]
    out[small] = (1.0 - xs**2 / (2 * (2 * l + 3))) / dfact  # two terms of the series
    xl = x[~small]
    out[~small] = scipy.special.spherical_jn(l, xl) / xl**l
    return out


def radial_transform(r: np.ndarray, dr: np.ndarray, f: np.ndarray, l_rf, G_max: float,
                     dG: float) -> np.ndarray:
    """f_tilde (n_f, nG), nG = ceil(G_max / dG) + 5, of f (n_f, n_r)."""
    w = (2.0 / 3.0) * 4.0 * np.pi * r**2 * dr
    w[1::2] *= 2.0
    w[0] *= 0.5
    w[-1] *= 0.5
    G = np.arange(int(np.ceil(G_max / dG)) + 5) * dG
    out = np.empty((len(l_rf), len(G)))
    for i, l in enumerate(l_rf):
        out[i] = (w * f[i] * r ** (2 * l)) @ jl_by_xl(l, np.outer(r, G))
    return out
