"""Nonlocal projectors and couplings consumed by H (reference: dft/ions).

Two producers: ``spline_projectors`` (qmx_projectors) for sampled radial
functions of any l <= 6, as get_projectors with pseudopotential radials, and
``build_projectors`` for the Gaussian radials of the synthetic cells.

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

from typing import Optional

import numpy as np
import torch

from . import _lib, rc

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
    if dev.type == "cuda":
        return _gaussian_projectors(positions, proj_rc, proj_l, iG, k, Gbasis, volume, n)
    iGk = iG.to(torch.float64) + k[:, None, :]
    Gb = torch.from_numpy(np.ascontiguousarray(np.asarray(Gbasis, dtype=np.float64).T))
    Gb = Gb.pin_memory().to(dev, non_blocking=True) if dev.type == "cuda" else Gb
    # Three-term sums rather than matmul: no library GEMM for a 3-long axis.
    Gk = iGk[..., 0:1] * Gb[0] + iGk[..., 1:2] * Gb[1] + iGk[..., 2:3] * Gb[2]
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
                        (-2 * np.pi) * (iGk[..., None, 0] * positions[:, 0]
                                        + iGk[..., None, 1] * positions[:, 1]
                                        + iGk[..., None, 2] * positions[:, 2]))  # (K, G, A)
    proj = phase[..., :, None, None] * f[..., None, :, None] * Y[..., None, None, :]
    proj = proj.reshape(K, G, -1).transpose(1, 2).contiguous()  # (K, P, G)
    pad = torch.arange(G, device=dev)[None, None, :] >= n.to(dev)[:, None, None]
    return proj.masked_fill_(pad, 0.0)


def _gaussian_projectors(positions, proj_rc, proj_l, iG, k, Gbasis, volume, n) -> torch.Tensor:
    """build_projectors on the device in one launch (qmx_gaussian_projectors):
    the same projectors, atom -> radial -> (l, m) with m = -1, 0, 1 the
    (y, z, x) harmonics of l = 1."""
    K, G = iG.shape[0], iG.shape[1]
    A = int(positions.shape[0])
    l_p, m_p, r_p = [], [], []
    for r in range(len(proj_rc)):
        for l in proj_l:
            if l > 1:
                raise NotImplementedError("Gaussian projectors with l > 1 are not on this path")
            for m in range(-l, l + 1):
                l_p.append(l)
                m_p.append(m)
                r_p.append(r)
    l_p, m_p, r_p = (np.array(x, dtype=np.int32) for x in (l_p, m_p, r_p))
    rc_h = np.array(proj_rc, dtype=np.float64)
    Gb = np.ascontiguousarray(np.asarray(Gbasis, dtype=np.float64))
    iG32 = iG if iG.dtype == torch.int32 and iG.is_contiguous() else iG.to(torch.int32).contiguous()
    n32 = n if n.dtype == torch.int32 else n.to(torch.int32)
    kk = k.to(torch.float64).contiguous()
    pos = positions.to(device=iG.device, dtype=torch.float64).contiguous()
    out = torch.empty((K, A * len(l_p), G), dtype=torch.complex128, device=iG.device)
    c_i32 = lambda a: a.ctypes.data_as(_lib.c_i32p)
    _lib.call("qmx_gaussian_projectors", K, G, iG32.data_ptr(), kk.data_ptr(), n32.data_ptr(),
              Gb.ctypes.data_as(_lib.c_f64p), A, pos.data_ptr(), len(l_p), c_i32(l_p), c_i32(m_p),
              c_i32(r_p), len(rc_h), rc_h.ctypes.data_as(_lib.c_f64p), 1.0 / np.sqrt(volume),
              out.data_ptr(), out.stride(1), out.stride(0), rc.stream_ptr())
    return out


DG_DEFAULT = 0.02  #: node spacing of the radial splines (math/_radial_function.py:22)


def quintic_coeff(samples: np.ndarray) -> np.ndarray:
    """Blip coefficients (N + 4, ...) of samples (..., N) on a uniform grid
    (quintic_spline.get_coeff, math/quintic_spline.py:12-53): the
    pentadiagonal system of blip values 1, 13/33, 1/66 at 0, 1 and 2 nodes,
    mirror-symmetric at the left end (c[-j] = c[j]) and natural at the right
    end (c[N], c[N+1] extrapolated so that the third and fourth derivatives
    vanish), then the two padding coefficients on each side. A host set-up
    step, once per radial function set, as in the reference."""
    from scipy.linalg import solve_banded

    samples = np.asarray(samples, dtype=np.float64)
    N = samples.shape[-1]
    if N < 4:
        raise ValueError("a radial spline needs at least 4 samples")
    near, far = 13.0 / 33.0, 1.0 / 66.0
    A = np.zeros((5, N))  # A[2 + i - j, j] = matrix element (i, j)
    A[0, :], A[1, :], A[2, :], A[3, :], A[4, :] = far, near, 1.0, near, far
    A[2, 1] += far  # row 1: c[-1] = c[1]
    A[1, 1] += near  # row 0: c[-1] = c[1]
    A[0, 2] += far  # row 0: c[-2] = c[2]
    ext = np.array([[1.0, -3.0, 3.0], [3.0, -8.0, 6.0]])  # c[N], c[N+1] from c[N-3:N]
    for j in range(3):
        col = N - 3 + j
        A[2 + (N - 2) - col, col] += far * ext[0, j]
        A[2 + (N - 1) - col, col] += near * ext[0, j] + far * ext[1, j]
    c = solve_banded((2, 2), A, samples.reshape(-1, N).T)
    c = np.concatenate([c[2:0:-1], c, ext @ c[N - 3:]], axis=0)
    return c.reshape((N + 4,) + samples.shape[:-1])


def simpson_weights(r: np.ndarray, dr: np.ndarray) -> np.ndarray:
    """4 pi r^2 dr with Simpson's 1/3 rule on the radial grid (weights 1, 4,
    2, 4, ..., 1 over 3; math/_radial_function.py:84-89)."""
    w = (2.0 / 3.0) * 4.0 * np.pi * np.asarray(r, dtype=np.float64) ** 2 * np.asarray(dr)
    w[1::2] *= 2.0
    w[0] *= 0.5
    w[-1] *= 0.5
    return w


def radial_transform(r: np.ndarray, dr: np.ndarray, f: np.ndarray, l_rf, G_max: float,
                     dG: float = DG_DEFAULT) -> np.ndarray:
    """f_tilde (n_f, nG) on G = iG dG, nG = ceil(G_max / dG) + 5, of radial
    functions f (n_f, n_r) with r^l removed (RadialFunction.transform,
    math/_radial_function.py:62-123, through qmx_radial_transform)."""
    f = np.atleast_2d(np.asarray(f, dtype=np.float64))
    l32 = np.ascontiguousarray(np.asarray(l_rf, dtype=np.int32).reshape(-1))
    if f.shape[0] != len(l32) or f.shape[1] != len(r):
        raise ValueError("f must be (len(l_rf), len(r))")
    nG = int(np.ceil(G_max / dG)) + 5
    dev = rc.device
    rt = torch.tensor(np.asarray(r, dtype=np.float64), device=dev)
    wt = torch.tensor(simpson_weights(r, dr), device=dev)
    ft = torch.tensor(f, device=dev)
    out = torch.empty((len(l32), nG), dtype=torch.float64, device=dev)
    _lib.call("qmx_radial_transform", len(l32), len(r), rt.data_ptr(), wt.data_ptr(),
              ft.data_ptr(), l32.ctypes.data_as(_lib.c_i32p), nG, float(dG), out.data_ptr(),
              rc.stream_ptr())
    return out.cpu().numpy()


class RadialSet:
    """Radial functions of one species on a uniform reciprocal grid (the
    f_tilde of RadialFunction, math/_radial_function.py:11-123, with G^l
    removed) as blip coefficients on the device, and the angular momentum
    of each; projectors follow the pseudo quantum numbers
    (dft/ions/_pseudo_quantum_numbers.py:28-56): per radial function,
    m = -l..l."""

    def __init__(self, l_rf, f_tilde: np.ndarray, dG: float = DG_DEFAULT):
        self.l_rf = tuple(int(l) for l in l_rf)
        f_tilde = np.atleast_2d(np.asarray(f_tilde, dtype=np.float64))
        if f_tilde.shape[0] != len(self.l_rf):
            raise ValueError("one sampled radial function per l")
        self.dG = float(dG)
        self.G_max = (f_tilde.shape[1] - 1) * self.dG  # last node
        coeff = np.ascontiguousarray(quintic_coeff(f_tilde))  # (N + 4, n_rf)
        self.coeff = torch.tensor(coeff, device=rc.device)
        self.l = np.array([l for l in self.l_rf for _ in range(2 * l + 1)], dtype=np.int32)
        self.m = np.array([m for l in self.l_rf for m in range(-l, l + 1)], dtype=np.int32)
        self.i_rf = np.array([i for i, l in enumerate(self.l_rf) for _ in range(2 * l + 1)],
                             dtype=np.int32)

    @property
    def n_proj(self) -> int:
        return len(self.l)

    @classmethod
    def from_real_space(cls, l_rf, r: np.ndarray, dr: np.ndarray, f: np.ndarray,
                        G_max: float, dG: float = DG_DEFAULT) -> "RadialSet":
        """Radial functions f (n_f, n_r) on the grid r (weights dr), r^l
        removed, transformed up to G_max (radial_transform) and fitted."""
        return cls(l_rf, radial_transform(r, dr, f, l_rf, G_max, dG), dG)


def spline_projectors(basis, positions: torch.Tensor, radial: RadialSet,
                      out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """beta (K, n_atoms * n_proj, n_tot) complex128 of one species on
    `basis` (qmx_projectors; _ions_projectors.py:28-72), padding zeroed;
    written into `out` (contiguous, that shape) when given."""
    K, G = basis.nk, basis.n_tot
    A = int(positions.shape[0])
    shape = (K, A * radial.n_proj, G)
    if out is None:
        out = torch.empty(shape, dtype=torch.complex128, device=rc.device)
    elif (tuple(out.shape) != shape or out.dtype != torch.complex128
          or not out.is_contiguous()):
        raise ValueError(f"out must be a contiguous complex128 tensor of shape {shape}")
    if A == 0 or radial.n_proj == 0 or K == 0:
        return out.zero_()
    iG = getattr(basis, "_iG_i32", None)
    if iG is None:
        iG = basis._iG_i32 = basis.iG.to(torch.int32).contiguous()
    pos = positions.to(device=rc.device, dtype=torch.float64).contiguous()
    Gb = np.ascontiguousarray(np.asarray(basis.Gbasis, dtype=np.float64))
    G_max = float(np.sqrt(2.0 * basis.ke_cutoff))  # the sphere: |G + k|^2 / 2 < ke_cutoff
    if G_max > radial.G_max - 4 * radial.dG:
        raise ValueError(f"radial functions sampled to G = {radial.G_max:g}, the basis needs "
                         f"{G_max:g} (plus four nodes)")
    c_i32 = lambda a: a.ctypes.data_as(_lib.c_i32p)
    _lib.call("qmx_projectors", K, G, iG.data_ptr(), basis.k.data_ptr(), basis.n_i32.data_ptr(),
              Gb.ctypes.data_as(_lib.c_f64p), A, pos.data_ptr(), radial.n_proj, c_i32(radial.l),
              c_i32(radial.m), c_i32(radial.i_rf), len(radial.l_rf), radial.coeff.shape[0],
              radial.coeff.data_ptr(), radial.dG, G_max, 1.0 / np.sqrt(basis.volume),
              out.data_ptr(), out.stride(1), out.stride(0), rc.stream_ptr())
    return out


class Ions:
    beta_version: int

    def __init__(self, positions: np.ndarray, proj_rc, proj_l, D: np.ndarray,
                 radial: "RadialSet | None" = None):
        """Gaussian radials exp(-rc^2 G^2 / 2) for every (rc, l) of proj_rc x
        proj_l (l <= 1), or, with `radial`, the sampled radial functions of a
        species (l <= 6) through qmx_projectors; D: the diagonal couplings."""
        self.positions = torch.tensor(np.asarray(positions, dtype=np.float64), device=rc.device)
        self.radial = radial
        self.proj_rc = tuple(float(r) for r in proj_rc)
        self.proj_l = tuple(int(l) for l in proj_l)
        if radial is not None and (self.proj_rc or self.proj_l):
            raise ValueError("give either Gaussian radials or a RadialSet, not both")
        if any(l > 1 for l in self.proj_l):
            raise NotImplementedError("Gaussian projectors with l > 1 are not on this path")
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
        if self.radial is not None:
            return self.n_atoms * self.radial.n_proj
        return self.n_atoms * len(self.proj_rc) * sum(2 * l + 1 for l in self.proj_l)

    def translation_phase(self, iG: torch.Tensor) -> torch.Tensor:
        """exp(-2 pi i iG . tau) with atoms last (_ions.py:304-311)."""
        return torch.polar(torch.ones((), device=iG.device, dtype=torch.float64),
                           (-2 * np.pi) * (iG @ self.positions.T))

    def update(self, basis) -> None:
        """(Re)build beta on `basis` and bump beta_version."""
        from .wavefunction import Wavefunction

        if self.radial is not None:
            proj = spline_projectors(basis, self.positions, self.radial)
        else:
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
