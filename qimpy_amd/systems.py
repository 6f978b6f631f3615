"""Seeded synthetic band-solve inputs (cells, k-points, potentials, projectors).

Every quantity is drawn from a counter-based splitmix64 stream keyed by
``(seed, stream)``, so the same arrays come out on any host, in any order of
generation, and the GPU path and the CPU checker see bit-identical inputs.

The cell recipes follow the five workloads this package is measured on:

0. 2-atom FCC diamond cell (a = 10.26 bohr), 36^3 box, 20 Ha, 2x2x2 MP k-mesh,
   16 bands, plain FP64 Davidson.
1. 64-atom cubic cell (a = 20.52 bohr), 96^3 box, Gamma only, ~16k plane waves,
   256 bands, mixed precision.
2. 64 uneven cells: boxes from {24,30,36,40,48,60}^3, 8-64 bands, 1-16 k,
   2-16 atoms, 20 Ha.
3. A Poisson request stream (200 req/s) of workload-2 sized cells.
4. 256-atom cubic cell (a = 32.58 bohr), 144^3 box, Gamma only, ~60k plane
   waves, 1024 bands, CheFSI degree 20.

Conventions follow the reference: lattice vectors are the columns of
``Rbasis`` and ``Gbasis = 2 pi inv(Rbasis^T)`` (lattice/_lattice.py:217-218);
Monkhorst-Pack k are wrapped to [-1/2, 1/2) (lattice/_kpoints.py:133-138);
``n_bands_extra = max(1, ceil(0.1 n_bands))`` (dft/electrons/_fillings.py:241);
the box is C-ordered with z fastest and signed frequencies
``i <= S/2 ? i : i - S`` (grid/_fft.py:67-71).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

__all__ = [
    "CellSpec",
    "splitmix64_uniform",
    "splitmix64_normal",
    "make_cell",
    "make_cfg2_cells",
    "make_request_stream",
    "signed_freqs",
]

_MASK64 = np.uint64(0xFFFFFFFFFFFFFFFF)
_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def _mix64(z: np.ndarray) -> np.ndarray:
    """splitmix64 finalizer (Steele, Lea, Flood 2014) on a uint64 array."""
    z = z.copy()
    z ^= z >> np.uint64(30)
    z *= _M1
    z ^= z >> np.uint64(27)
    z *= _M2
    z ^= z >> np.uint64(31)
    return z


def _stream_key(seed: int, stream: int) -> np.uint64:
    k = np.array([(seed * 1_000_003 + stream) & 0xFFFFFFFFFFFFFFFF], dtype=np.uint64)
    return _mix64(k)[0]


def splitmix64_uniform(seed: int, stream: int, n: int) -> np.ndarray:
    """n uniforms in [0, 1) from counters 1..n of stream `(seed, stream)`."""
    with np.errstate(over="ignore"):
        ctr = np.arange(1, n + 1, dtype=np.uint64) * _GOLDEN + _stream_key(seed, stream)
        z = _mix64(ctr)
    return (z >> np.uint64(11)).astype(np.float64) * (2.0**-53)


def splitmix64_normal(seed: int, stream: int, n: int) -> np.ndarray:
    """n standard normals (Box-Muller on pairs of stream uniforms)."""
    m = (n + 1) // 2
    u = splitmix64_uniform(seed, stream, 2 * m)
    u0 = 1.0 - u[0::2]  # (0, 1]
    r = np.sqrt(-2.0 * np.log(u0))
    phi = 2.0 * np.pi * u[1::2]
    return np.concatenate((r * np.cos(phi), r * np.sin(phi)))[:n]


def splitmix64_int(seed: int, stream: int, n: int, lo: int, hi: int) -> np.ndarray:
    """n integers uniform in [lo, hi] (inclusive)."""
    u = splitmix64_uniform(seed, stream, n)
    return lo + np.minimum((u * (hi - lo + 1)).astype(np.int64), hi - lo)


def signed_freqs(s: int) -> np.ndarray:
    """Signed integer frequencies of a length-s FFT axis (grid/_fft.py:67-71)."""
    iv = np.arange(s)
    return np.where(iv <= s // 2, iv, iv - s)


# Streams used for each drawn quantity:
_S_V, _S_D, _S_POS, _S_SIZE, _S_K, _S_ARRIVE = 1, 2, 3, 4, 5, 6


@dataclass
class CellSpec:
    """One band-solve request: geometry, basis, potential and projectors."""

    name: str
    seed: int
    Rbasis: np.ndarray  #: 3x3, lattice vectors in columns (bohr)
    positions: np.ndarray  #: (n_atoms, 3) fractional
    ke_cutoff: float  #: wavefunction cutoff (Ha)
    shape: tuple[int, int, int]  #: FFT box
    k: np.ndarray  #: (K, 3) fractional k-points
    wk: np.ndarray  #: (K,) weights (sum 1)
    n_bands: int
    n_bands_extra: int
    V: np.ndarray = field(repr=False)  #: real potential on the box (Ha), f64
    proj_rc: tuple[float, ...] = (0.8, 1.3)  #: Gaussian radial widths (bohr)
    proj_l: tuple[int, ...] = (0, 1)  #: angular momenta per radial
    D: np.ndarray = field(default=None, repr=False)  #: (P,) diagonal couplings (Ha)
    n_iterations: int = 100
    eig_threshold: float = 1e-8
    filter_order: int = 0  #: >0 selects CheFSI with this filter order

    @property
    def Gbasis(self) -> np.ndarray:
        return 2 * np.pi * np.linalg.inv(self.Rbasis.T)

    @property
    def volume(self) -> float:
        return float(abs(np.linalg.det(self.Rbasis)))

    @property
    def n_atoms(self) -> int:
        return int(self.positions.shape[0])

    @property
    def n_proj_atom(self) -> int:
        return len(self.proj_rc) * sum(2 * l + 1 for l in self.proj_l)

    @property
    def n_projectors(self) -> int:
        return self.n_atoms * self.n_proj_atom

    @property
    def n_bands_max(self) -> int:
        return self.n_bands + self.n_bands_extra


def _n_extra(n_bands: int) -> int:
    return max(1, int(np.ceil(n_bands * 0.1)))


def _make_potential(
    seed: int, shape: Sequence[int], Gbasis: np.ndarray, sigma: float = 1.0,
    amplitude: float = 0.5,
) -> np.ndarray:
    """Gaussian-smoothed white noise, zero mean, max|V| = amplitude: the
    noise times exp(-sigma^2 G^2 / 2) in reciprocal space (real transforms
    on the half box, G^2 from the reciprocal metric)."""
    import scipy.fft

    shape = tuple(int(s) for s in shape)
    n = int(np.prod(shape))
    noise = splitmix64_normal(seed, _S_V, n).reshape(shape)
    g = Gbasis.T @ Gbasis  # metric: G^2 = sum_ab g_ab iG_a iG_b
    f0 = signed_freqs(shape[0]).astype(np.float64)[:, None, None]
    f1 = signed_freqs(shape[1]).astype(np.float64)[None, :, None]
    f2 = np.arange(shape[2] // 2 + 1, dtype=np.float64)[None, None, :]
    G2 = (g[0, 0] * f0 * f0 + g[1, 1] * f1 * f1 + g[2, 2] * f2 * f2
          + 2.0 * (g[0, 1] * f0 * f1 + g[0, 2] * f0 * f2 + g[1, 2] * f1 * f2))
    Vt = scipy.fft.rfftn(noise)
    Vt *= np.exp(-0.5 * sigma**2 * G2)
    V = scipy.fft.irfftn(Vt, s=shape)
    V -= V.mean()
    V *= amplitude / np.abs(V).max()
    return np.ascontiguousarray(V)


def _mp_mesh(size: Sequence[int], offset: Sequence[float]) -> np.ndarray:
    grids = [(offset[i] + np.arange(size[i])) / size[i] for i in range(3)]
    mesh = np.stack(np.meshgrid(*grids, indexing="ij")).reshape(3, -1).T
    return mesh - np.floor(0.5 + mesh)


def _finish(
    name, seed, Rbasis, positions, ke_cutoff, shape, k, wk, n_bands, **kw
) -> CellSpec:
    spec = CellSpec(
        name=name, seed=seed, Rbasis=np.asarray(Rbasis, dtype=np.float64),
        positions=np.asarray(positions, dtype=np.float64), ke_cutoff=float(ke_cutoff),
        shape=tuple(int(s) for s in shape), k=np.asarray(k, dtype=np.float64),
        wk=np.asarray(wk, dtype=np.float64), n_bands=int(n_bands),
        n_bands_extra=int(kw.pop("n_bands_extra", _n_extra(n_bands))), V=None, **kw,
    )
    spec.V = _make_potential(seed, spec.shape, spec.Gbasis)
    spec.D = 2.0 * splitmix64_uniform(seed, _S_D, spec.n_projectors) - 1.0
    return spec


# Diamond conventional-cell basis (fractional, 8 atoms):
_DIAMOND8 = np.array(
    [[0, 0, 0], [0, 0.5, 0.5], [0.5, 0, 0.5], [0.5, 0.5, 0],
     [0.25, 0.25, 0.25], [0.25, 0.75, 0.75], [0.75, 0.25, 0.75], [0.75, 0.75, 0.25]]
)


def make_cell(cfg: int, cell_id: int = 0, **overrides) -> CellSpec:
    """Build the cell of workload `cfg` (0, 1 or 4; 2 and 3 via their helpers).

    `overrides` replace recipe parameters (e.g. ``shape``, ``ke_cutoff``,
    ``n_bands``, ``kmesh``) to make small cells of the same recipe for tests.
    """
    seed = cfg * 1000 + cell_id
    if cfg == 0:
        a = overrides.pop("a", 10.26)
        R = 0.5 * a * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0]], dtype=float)
        pos = np.array([[0, 0, 0], [0.25, 0.25, 0.25]])
        kmesh = overrides.pop("kmesh", (2, 2, 2))
        k = _mp_mesh(kmesh, (0.5, 0.5, 0.5))
        wk = np.full(len(k), 1.0 / len(k))
        args = dict(ke_cutoff=20.0, shape=(36, 36, 36), n_bands=16,
                    n_iterations=20, eig_threshold=1e-8)
    elif cfg == 1:
        a = overrides.pop("a", 20.52)
        R = a * np.eye(3)
        base = np.concatenate(
            [(_DIAMOND8 + np.array(o)) / 2 for o in np.ndindex(2, 2, 2)]
        )
        disp = 0.05 / a * splitmix64_normal(seed, _S_POS, base.size).reshape(base.shape)
        pos = (base + disp) % 1.0
        k = np.zeros((1, 3))
        wk = np.ones(1)
        args = dict(ke_cutoff=11.5, shape=(96, 96, 96), n_bands=256,
                    n_iterations=100, eig_threshold=1e-8)
    elif cfg == 4:
        a = overrides.pop("a", 32.58)
        R = a * np.eye(3)
        n_atoms = overrides.pop("n_atoms", 256)
        pos = splitmix64_uniform(seed, _S_POS, 3 * n_atoms).reshape(n_atoms, 3)
        k = np.zeros((1, 3))
        wk = np.ones(1)
        args = dict(ke_cutoff=11.0, shape=(144, 144, 144), n_bands=1024,
                    n_iterations=10, eig_threshold=1e-8, filter_order=20)
    else:
        raise ValueError(f"make_cell builds workloads 0, 1 and 4, not {cfg}")
    if "n_atoms" in overrides:
        n_atoms = overrides.pop("n_atoms")
        pos = splitmix64_uniform(seed, _S_POS, 3 * n_atoms).reshape(n_atoms, 3)
    args.update(overrides)
    return _finish(f"cfg{cfg}", seed, R, pos, k=k, wk=wk, **args)


CFG2_GRIDS = (24, 30, 36, 40, 48, 60)


def make_cfg2_cell(cell_id: int, cfg: int = 2, ke_cutoff: float = 20.0) -> CellSpec:
    """One uneven cell: cubic box g^3, a = g pi / sqrt(8 Ecut) so that the
    box exactly holds the 4 Ecut density sphere (grid/_grid.py:114-127)."""
    seed = cfg * 1000 + cell_id
    g, nb, nk, na = (
        int(CFG2_GRIDS[splitmix64_int(seed, _S_SIZE, 1, 0, 5)[0]]),
        int(splitmix64_int(seed, _S_SIZE + 10, 1, 8, 64)[0]),
        int(splitmix64_int(seed, _S_SIZE + 20, 1, 1, 16)[0]),
        int(splitmix64_int(seed, _S_SIZE + 30, 1, 2, 16)[0]),
    )
    a = g * np.pi / np.sqrt(8.0 * ke_cutoff)
    k = splitmix64_uniform(seed, _S_K, 3 * nk).reshape(nk, 3) - 0.5
    pos = splitmix64_uniform(seed, _S_POS, 3 * na).reshape(na, 3)
    return _finish(
        f"cfg{cfg}.{cell_id}", seed, a * np.eye(3), pos, ke_cutoff, (g, g, g), k,
        np.full(nk, 1.0 / nk), nb, n_iterations=100, eig_threshold=1e-8,
    )


def make_cfg2_cells(n_cells: int = 64, cfg: int = 2) -> list[CellSpec]:
    return [make_cfg2_cell(i, cfg) for i in range(n_cells)]


def make_request_stream(
    n_requests: int, rate: float = 200.0, seed: int = 3000
) -> tuple[np.ndarray, list[int]]:
    """Poisson arrival times (s) and cell ids of a workload-3 request stream."""
    u = splitmix64_uniform(seed, _S_ARRIVE, n_requests)
    gaps = -np.log(1.0 - u) / rate
    return np.cumsum(gaps), list(range(n_requests))


def small_cell(
    shape=(12, 15, 10), ke_cutoff: Optional[float] = None, n_bands: int = 4,
    kmesh=(2, 2, 1), n_atoms: int = 2, seed: int = 777, triclinic: bool = True,
) -> CellSpec:
    """Tiny non-cubic cell for brute-force checks (dense H is a few hundred wide)."""
    if triclinic:
        R = np.array([[4.1, 0.3, -0.2], [0.1, 4.6, 0.25], [-0.15, 0.2, 3.8]])
    else:
        R = 4.0 * np.eye(3)
    # Largest cutoff whose 4x sphere fits in the box:
    if ke_cutoff is None:
        shape_arr = np.array(shape, dtype=float)
        gmax = np.min(shape_arr * np.pi / np.linalg.norm(R, axis=0))
        ke_cutoff = 0.95 * gmax**2 / 8.0
    k = _mp_mesh(kmesh, (0.5, 0.5, 0.5)) if np.prod(kmesh) > 1 else np.zeros((1, 3))
    pos = splitmix64_uniform(seed, _S_POS, 3 * n_atoms).reshape(n_atoms, 3)
    return _finish(
        "small", seed, R, pos, ke_cutoff, shape, k, np.full(len(k), 1.0 / len(k)),
        n_bands, n_iterations=50, eig_threshold=1e-10,
    )
