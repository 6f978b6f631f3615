"""Batched band solves of many independent cells on one GPU.

QimPy solves one cell at a time: each request below is one fixed-Hamiltonian
band solve (all k-points and bands of a cell) with the Davidson iteration of
dft/electrons/_davidson.py:141-235 and H of dft/electrons/_hamiltonian.py:7-29.
This module runs many such solves together.

Layout. Jobs (one cell each) that share an FFT box and a band class form a
bucket, which stacks the k-points of all its jobs along one axis:
    X  = [C; R]     (Kb, 2 nb, Gb)  bands, then their preconditioned residuals
    HX = [HC; HR]   (Kb, 2 nb, Gb)  H applied to both
    eig             (Kb, nb)        float64 Ritz values
    beta, D         (Kb, Pb, Gb), (Kb, Pb)  projectors / couplings, zero-padded
    V               (J, N)          one potential box per job (plan slot)
with Gb = the largest basis of the bucket's k-points, Pb = the largest
projector count and nb the band class (>= n_bands + n_bands_extra of every
job). Every pass of an iteration is then one launch over all jobs of the
bucket: the sparse FFT through a plan whose k-points carry the potential
slot of their job, the overlaps, rotations and nonlocal products as batched
GEMMs over Kb, and the subspace eigensolves as one batched on-chip launch.

Iteration. The reference Davidson step with every band expanded in every
iteration (the reference's locking counter n_eigs_done stays 0), so shapes
stay fixed while jobs join and leave between iterations. The expansion is
made orthonormal before H is applied: projected out of span(C) once
(classical Gram-Schmidt; the rest enters the expanded overlap, which a
Loewdin factor removes) and orthonormalised by the FP64 eigenvectors of its
own overlap, dropping directions whose overlap
eigenvalue falls below a relative 1e-8 (they sit at +1e6 Ha in the projected
H and sort out of the kept bands). The subspace problem is then an ordinary
FP64 Hermitian eigenproblem of size 2 nb, with the same span and Ritz values
as the reference's generalized one (_davidson.py:199-206). No step needs the
host: one device-to-host read per iteration of the solver carries every
bucket's convergence data.

Each job converges on its own n_bands: the plain (complex128) form by the
reference's test, eigenvalue change below eig_threshold
(_davidson.py:126-139); the mixed (complex64) form when the Kato-Temple
estimate of every wanted band (qimpy_amd.accuracy, from the FP64-accumulated
residual norms ||HC - eig C|| and the Ritz gaps) is below half its
eigenvalue tolerance. Both finish with an FP64 Rayleigh-Ritz (complex128
overlap, projected H and eigenproblem; the mixed form applies H in
complex128 first), whose Ritz values are the results and whose residual
norms give each result's error estimate. A job starts as in the
reference: xorshift bands with its own seed and basis
(_wavefunction_init.py:27-125), orthonormalised in FP64, then rotated to
Ritz vectors, so its result does not depend on the other jobs of its bucket.
"""
from __future__ import annotations

import copy
import ctypes as ct
import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib, gemm, rc
from .basis import DTYPE_CODE, REAL_OF
from .ions import build_projectors
from .accuracy import eig_error_estimate
from .linalg import dagger, eigh_grouped, eigh_nosync
from .systems import signed_freqs

C64, C128 = torch.complex64, torch.complex128

#: Largest sparse-FFT workspace of one bucket; bands beyond it are blocked.
WORKSPACE_CAP = 4 << 30
#: A mixed-precision job stops when every wanted band's error estimate is
#: below this fraction of its tolerance.
KT_SAFETY = 0.5


# ---------------------------------------------------------------- geometry
def sphere_tables(Rbasis: np.ndarray, k: np.ndarray, ke_cutoff: float, shape):
    """Plane waves with KE(G + k) < ke_cutoff of each k (_basis.py:167-194):
    per k, the box linear indices in increasing order and the integer
    G-vectors. Only the bounding box of the sphere is enumerated:
    |iG_a + k_a| = |R_a . G| / 2 pi < |R_a| sqrt(2 ke_cutoff) / 2 pi."""
    Rbasis = np.asarray(Rbasis, dtype=np.float64)
    Gbasis = 2 * np.pi * np.linalg.inv(Rbasis.T)
    gmax = np.sqrt(2.0 * ke_cutoff)
    shape = tuple(int(s) for s in shape)
    out = []
    for kk in np.asarray(k, dtype=np.float64).reshape(-1, 3):
        axes = []
        for a in range(3):
            m = gmax * np.linalg.norm(Rbasis[:, a]) / (2 * np.pi)
            f = signed_freqs(shape[a])
            lo = max(int(np.ceil(-m - kk[a])), int(f.min()))
            hi = min(int(np.floor(m - kk[a])), int(f.max()))
            axes.append(np.arange(lo, hi + 1))
        mesh = np.stack(np.meshgrid(*axes, indexing="ij"), -1).reshape(-1, 3)
        Gk = (mesh + kk) @ Gbasis.T
        sel = mesh[0.5 * np.einsum("gi,gi->g", Gk, Gk) < ke_cutoff]
        lin = ((sel[:, 0] % shape[0]) * shape[1] + (sel[:, 1] % shape[1])) * shape[2] + (
            sel[:, 2] % shape[2])
        order = np.argsort(lin, kind="stable")
        out.append((lin[order], sel[order]))
    return out


def choose_band_classes(n_max_bands: Sequence[int], overhead_bands: float = 12.0,
                        granule: int = 8) -> list[int]:
    """Band classes (ascending) for jobs needing `n_max_bands` bands each.

    A job runs padded to its class; every class is a separate bucket whose
    iteration costs a fixed launch and synchronisation overhead, worth about
    `overhead_bands` bands of work. Classes are multiples of `granule`;
    dynamic programming over the sorted demands minimises
    sum(class size over its jobs) + overhead_bands * (number of classes)."""
    if not len(n_max_bands):
        return []
    cand = sorted({int(np.ceil(n / granule)) * granule for n in n_max_bands})
    need = sorted(int(n) for n in n_max_bands)
    # cost[i][j]: jobs with demand in (cand[i-1], cand[j]] padded to cand[j].
    m = len(cand)
    best = [0.0] + [np.inf] * m
    choice = [-1] * (m + 1)
    for j in range(1, m + 1):
        for i in range(j):
            lo = cand[i - 1] if i else 0
            cnt = sum(1 for n in need if lo < n <= cand[j - 1])
            c = best[i] + cnt * cand[j - 1] + overhead_bands
            if c < best[j]:
                best[j], choice[j] = c, i
    classes = []
    j = m
    while j > 0:
        classes.append(cand[j - 1])
        j = choice[j]
    return sorted(classes)


# ---------------------------------------------------------------- requests
@dataclass
class CellResult:
    job_id: int
    eig: np.ndarray  #: float64 (K, n_bands)
    residual_norm: np.ndarray  #: float64 (K, n_bands)
    #: largest Kato-Temple error estimate (Ha) of the wanted bands after the
    #: final FP64 Rayleigh-Ritz (qimpy_amd.accuracy)
    err_bound: float
    Eband: float
    iterations: int
    converged: bool
    arrival: float
    admitted: float
    finished: float

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


class CellJob:
    """One band-solve request: a cell, its k-points and its bands."""

    def __init__(self, spec, job_id: int = 0, arrival: float = 0.0, seed: int = 0):
        self.spec = spec
        self.job_id = job_id
        self.arrival = arrival
        self.seed = seed
        self.shape = tuple(int(s) for s in spec.shape)
        self.N = int(np.prod(self.shape))
        self.K = int(np.asarray(spec.k).reshape(-1, 3).shape[0])
        self.n_bands = int(spec.n_bands)
        self.n_bands_max = int(spec.n_bands + spec.n_bands_extra)
        tables = sphere_tables(spec.Rbasis, spec.k, spec.ke_cutoff, self.shape)
        self.n = np.array([len(t[0]) for t in tables], dtype=np.int64)
        self.n_tot = int(self.n.max())
        if 2 * self.n_bands_max >= int(self.n.min()):  # _davidson.py:159-163
            raise ValueError(f"n_bands + n_bands_extra = {self.n_bands_max} exceeds"
                             f" min(n_basis)/2 = {int(self.n.min()) // 2} in Davidson")
        self.fft_index = np.zeros((self.K, self.n_tot), dtype=np.int64)
        self.iG_np = np.zeros((self.K, self.n_tot, 3), dtype=np.int32)
        for i, (lin, sel) in enumerate(tables):
            self.fft_index[i, : len(lin)] = lin
            self.iG_np[i, : len(lin)] = sel
        self.P = int(spec.n_projectors)
        self.wk = np.asarray(spec.wk, dtype=np.float64)
        self.iterations = 0
        self.admitted = None
        self.result: Optional[CellResult] = None
        self._dev = None
        # Host copies of the device tables, pinned here (by the thread that
        # builds the job) so that admission copies them without waiting.
        host = dict(k=np.asarray(spec.k, dtype=np.float64).reshape(-1, 3), iG=self.iG_np,
                    GbT=np.asarray(spec.Gbasis, dtype=np.float64).T, n=self.n.astype(np.int32),
                    V=np.asarray(spec.V, dtype=np.float64).reshape(-1),
                    pos=np.asarray(spec.positions, dtype=np.float64),
                    D=np.asarray(spec.D, dtype=np.float64))
        self._host = {name: torch.from_numpy(np.ascontiguousarray(a)) for name, a in host.items()}
        if torch.cuda.is_available():
            self._host = {name: t.pin_memory() for name, t in self._host.items()}

    def device_tables(self) -> dict:
        """iG, k, KE and the potential on the device (built once)."""
        if self._dev is None:
            dev = rc.device
            h = {name: t.to(dev, non_blocking=True) for name, t in self._host.items()}
            x = h["iG"].to(torch.float64) + h["k"][:, None, :]
            GbT = h["GbT"]
            Gk = x[..., 0:1] * GbT[0] + x[..., 1:2] * GbT[1] + x[..., 2:3] * GbT[2]
            ke = 0.5 * Gk.square().sum(-1)
            ke[torch.arange(self.n_tot, device=dev)[None, :] >= h["n"][:, None]] = 0.0
            self._dev = dict(k=h["k"], iG=h["iG"], ke=ke, V=h["V"], n=h["n"], pos=h["pos"],
                             D=h["D"])
        return self._dev

    def fresh(self, arrival: Optional[float] = None) -> "CellJob":
        """A new request for the same cell: shares this job's tables (host
        and staged device copies), with its solve state reset."""
        job = copy.copy(self)
        job.arrival = self.arrival if arrival is None else float(arrival)
        job.iterations = 0
        job.admitted = None
        job.result = None
        job.converged = False
        job._beta = None
        return job

    def stage(self) -> None:
        """Copy the job's inputs to the device now and drop the pinned host
        copies (the copies are ordered on the current stream)."""
        self.device_tables()
        self._host = None

    def release(self) -> None:
        """Drop the job's device tables; their memory is reused only after
        the work queued so far on the current stream (they may have been
        made on another one)."""
        st = torch.cuda.current_stream() if torch.cuda.is_available() else None
        for t in list((self._dev or {}).values()) + [getattr(self, "_beta", None)]:
            if st is not None and t is not None and t.is_cuda:
                t.record_stream(st)
        self._dev = None
        self._beta = None

    def projectors(self, G: int) -> torch.Tensor:
        """beta (K, P, G) complex128, zero beyond each k's sphere; built once
        per job (on its own basis) and kept on the device while it is
        resident."""
        beta = getattr(self, "_beta", None)
        if beta is None:
            t = self.device_tables()
            beta = build_projectors(t["pos"], self.spec.proj_rc, self.spec.proj_l, t["iG"], t["k"],
                                    self.spec.Gbasis, self.spec.volume, t["n"])
            self._beta = beta
        if G > self.n_tot:
            return torch.nn.functional.pad(beta, (0, G - self.n_tot))
        return beta[..., :G]

    def hbm_bytes(self, nb: int, dtype: torch.dtype, G: Optional[int] = None) -> int:
        """Device bytes this job holds inside a bucket of band class nb whose
        rows are G plane waves long (the bucket pads every job to its
        largest basis): X and HX double-buffered, the copy of both a restack
        makes, GEMM outputs, projectors (complex128 and working copies) and
        the potential."""
        G = max(G or 0, self.n_tot)
        cs = 8 if dtype == C64 else 16
        state = 2 * 2 * (2 * nb) * G * cs * self.K  # X and HX, double-buffered
        restack = 2 * (2 * nb) * G * cs * self.K
        temps = 3 * nb * G * cs * self.K + 16 * self.P * G * self.K  # GEMM outputs, beta c128
        return int(state + restack + temps + self.P * G * cs * self.K + 12 * self.N)




# ---------------------------------------------------------------- raw kernels
def _c(x: torch.Tensor) -> int:
    return x.data_ptr()


def _to_device(a, dtype=None) -> torch.Tensor:
    """Host array -> device tensor through pinned memory without waiting for
    the work queued on the stream (a pageable copy would)."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.pin_memory().to(rc.device, non_blocking=True)


class _BlockCopies:
    """Row-block copies between device tensors, issued together as one
    qmx_copy_blocks launch (a restack moves many small and a few large
    blocks; one torch copy each costs more host time than the device
    needs)."""

    WG_BYTES = 32 << 10  # bytes written per workgroup
    _WORDS = {torch.float64: 1, C64: 1, C128: 2}
    _NARROW = {(C128, C64): 2, (torch.float64, torch.float32): 1}  # reals per element

    def __init__(self):
        self.rows: list = []

    def add(self, dst: torch.Tensor, src: Optional[torch.Tensor] = None) -> None:
        """dst <- src along the last axis, zero past src's length (zero
        throughout without src). 1-3 dimensional views with unit stride
        along the last axis; src broadcasts over dst's leading axes. Same
        dtype (float64, complex64, complex128), or complex128 into
        complex64 / float64 into float32 (rounded)."""
        if dst.numel() == 0:
            return
        if dst.dim() > 3 or dst.stride(-1) != 1:
            raise ValueError("copy destination needs 1-3 axes and unit stride along the last")
        lead = tuple(dst.shape[:-1])
        if src is not None:
            src = src.expand(*lead, src.shape[-1])
            if src.stride(-1) != 1 and src.shape[-1] > 1:
                raise ValueError("copy source needs unit stride along its last axis")
            if src.shape[-1] > dst.shape[-1]:
                raise ValueError("copy source longer than its destination")
        if src is None or src.dtype == dst.dtype:
            mode, us, ud = 0, self._WORDS[dst.dtype], self._WORDS[dst.dtype]
        else:
            mode = 1
            us = ud = self._NARROW[(src.dtype, dst.dtype)]

        def dims3(t, u):
            st = [x * u for x in t.stride()[:-1]]
            return [0] * (2 - len(st)) + st  # (the padded leading axes have length 1)

        shape = [1] * (3 - dst.dim()) + list(dst.shape)
        d_s = dims3(dst, ud)
        s_s = dims3(src, us) if src is not None else [0, 0]
        dst_cols = dst.shape[-1] * ud
        src_cols = src.shape[-1] * us if src is not None else 0
        unit_bytes = 4 if mode == 1 else 8
        rpw = max(1, self.WG_BYTES // (dst_cols * unit_bytes))
        self.rows.append([src.data_ptr() if src is not None else 0, dst.data_ptr(), shape[0],
                          shape[1], src_cols, dst_cols, s_s[0], s_s[1], d_s[0], d_s[1], rpw, 0,
                          mode, 0])

    def run(self) -> None:
        if not self.rows:
            return
        a = np.array(self.rows, dtype=np.int64)
        wg = -(-(a[:, 2] * a[:, 3]) // a[:, 10])
        a[:, 11] = np.cumsum(wg) - wg
        d = _to_device(a)
        _lib.call("qmx_copy_blocks", _c(d), len(a), int(wg.sum()), rc.stream_ptr())
        self.rows = []


#: Plans of restacked buckets, destroyed at the next host synchronisation
#: (freeing device memory waits for the whole device).
_retired_plans: list = []  # (stamp, plan handle)
_stamp = [0]  # advanced by every solver tick


def flush_retired_plans(upto: Optional[int] = None) -> None:
    """Destroy the plans retired at stamps <= `upto` (all if None); the
    caller guarantees that no queued device work still uses them."""
    keep = []
    for stamp, plan in _retired_plans:
        if upto is not None and stamp > upto:
            keep.append((stamp, plan))
            continue
        try:
            _lib.lib().qmx_plan_destroy(plan)
        except Exception:  # pragma: no cover - interpreter shutdown
            pass
    _retired_plans[:] = keep


def _band_norm(X: torch.Tensor) -> torch.Tensor:
    K, B, G = X.shape
    norm2 = torch.empty((K, B), dtype=torch.float64, device=X.device)
    _lib.call("qmx_band_reduce", DTYPE_CODE[X.dtype], _c(X), X.stride(1), X.stride(0), K, B, G,
              None, 0, _c(norm2), None, rc.stream_ptr())
    return norm2.sqrt()


def _scale_constrain(X: torch.Tensor, scale: Optional[torch.Tensor], n32: torch.Tensor) -> None:
    K, B, G = X.shape
    s = None if scale is None else scale.to(torch.float64).contiguous()
    _lib.call("qmx_scale_constrain", DTYPE_CODE[X.dtype], _c(X), X.stride(1), X.stride(0), K, B,
              G, None if s is None else _c(s), _c(n32), rc.stream_ptr())


def _dot(A: torch.Tensor, B: torch.Tensor) -> torch.Tensor:
    """conj(A) @ B^T over the basis axis, batched over k, complex128."""
    return gemm.dot128(A, B)


def _lowdin(O: torch.Tensor) -> torch.Tensor:
    """O^(-1/2) for Hermitian overlaps within ~1e-3 of the identity:
    I - E/2 + 3 E^2/8 with E = O - I (error O(|E|^3)); no factorisation,
    no host synchronisation."""
    E = 0.5 * (O + dagger(O))
    E.diagonal(dim1=-2, dim2=-1).sub_(1.0)
    S = E.clone()
    gemm.matmul(E, E, out=S, alpha=0.375, beta=-0.5)
    S.diagonal(dim1=-2, dim2=-1).add_(1.0)
    return S


# ---------------------------------------------------------------- buckets
class Bucket:
    """Jobs sharing an FFT box and a band class, iterated in lockstep."""

    #: Overlap eigenvalue (of unit-norm residuals) below which an expansion
    #: direction is dropped, in the plain and the mixed form.
    DROP_RTOL = 1e-10
    DROP_RTOL_MIXED = 1e-4
    #: Diagonal shift (Ha) that moves dropped directions above every kept band.
    DROP_SHIFT = 1e6

    def __init__(self, shape, nb: int, dtype: torch.dtype, *, residual_threshold: float,
                 eig_threshold: float, max_iterations: int,
                 eig_tolerance: Optional[float] = None):
        self.shape = tuple(int(s) for s in shape)
        self.eig_tolerance = eig_tolerance
        self.N = int(np.prod(self.shape))
        self.nb = int(nb)
        self.dtype = dtype
        self.rdtype = REAL_OF[dtype]
        self.mixed = dtype == C64
        self.residual_threshold = residual_threshold
        self.eig_threshold = eig_threshold
        self.max_iterations = max_iterations
        self.jobs: list[CellJob] = []
        self.k0: list[int] = []
        self._plan = None
        self.Kb = self.Gb = self.Pb = 0
        self.n_ticks = 0
        self._info: list[torch.Tensor] = []

    # ------------------------------------------------------------ structure
    def __del__(self):
        self._destroy_plan()

    def _destroy_plan(self):
        plan = getattr(self, "_plan", None)
        if plan is not None and plan.value:
            _retired_plans.append((_stamp[0], plan))  # destroyed once no work uses it
        self._plan = None

    #: Capacity step of the bucket buffers (k-points), and the factor of
    #: headroom they are allocated with.
    K_QUANTUM, K_GROW = 16, 1.25
    #: Block Gram-Schmidt passes projecting the residuals out of span(C)
    #: before their canonical orthonormalisation. One suffices: what a
    #: float32 pass leaves of C in the residual block (about 1e-5 relative)
    #: enters the expanded overlap E, whose Loewdin factor (error O(|E|^3))
    #: removes it. Measured on a 1300-request served stream against two
    #: passes: the same iteration counts (13.7-13.8 mixed, 14.8-14.9 plain),
    #: the same largest error against FP64 (1.5-1.7e-8 Ha), 4% more band
    #: solves/s.
    GS_PASSES = 1

    def _buffers(self, Kb: int, Gb: int, Pb: int) -> bool:
        """Make the bucket's buffers hold Kb k-points of Gb plane waves and Pb
        projectors: X and HX twice over (the iteration state and the buffers
        the next rotation writes), the projectors, couplings and kinetic
        energies. They are reused while the jobs fit, so a restack allocates
        nothing of the iteration state in the steady state of a server (and
        the caching allocator sees few sizes instead of a new one per
        restack); the basis length never shrinks and the k-point capacity
        shrinks only below a third. Returns True when new buffers were
        allocated (the previous ones, if any, stay readable until replaced)."""
        kcap, gcap, pcap = getattr(self, "_cap", (0, 0, 0))
        g_need = max(gcap, -(-Gb // 64) * 64)
        p_need = max(pcap, -(-Pb // 8) * 8)
        if Kb <= kcap and g_need == gcap and p_need == pcap and 3 * Kb >= kcap:
            return False
        kcap = max(self.K_QUANTUM, -(-int(np.ceil(Kb * self.K_GROW)) // self.K_QUANTUM)
                   * self.K_QUANTUM)
        dev, nb, dt = rc.device, self.nb, self.dtype
        self._cap = (kcap, g_need, p_need)
        # Not cleared: every row a step reads is written first (kept rows and
        # moved jobs' projector, coupling and kinetic rows by the restack copy
        # with their padding zeroed, new jobs' bands by init_steps, rotations
        # and the residual pass over the whole basis length).
        self._full = [torch.empty((kcap, 2 * nb, g_need), dtype=dt, device=dev) for _ in range(4)]
        self._beta_full = torch.empty((kcap, p_need, g_need), dtype=dt, device=dev)
        # FP64 couplings: they scale the complex128 projections directly.
        self._D_full = torch.zeros((kcap, p_need), dtype=torch.float64, device=dev)
        self._ke_full = torch.zeros((kcap, g_need), dtype=torch.float64, device=dev)
        return True

    def _restack(self, keep: list[CellJob], new: list[CellJob]) -> None:
        """Rebuild the stacked state for jobs `keep` (state carried over)
        followed by `new` (state to be initialised by init_steps) in the
        bucket's reusable buffers."""
        dev = rc.device
        nb, dt = self.nb, self.dtype
        jobs = keep + new
        if not jobs:
            self.jobs, self.k0, self.Kb, self.row0 = [], [], 0, 0
            for name in ("X", "HX", "X2", "HX2", "eig", "deig", "rnorm", "pnorm2", "beta", "D",
                         "ke", "ke_r", "n32", "V", "V64", "workspace", "_full", "_beta_full",
                         "_D_full", "_ke_full"):
                setattr(self, name, None)
            self._cap = (0, 0, 0)
            self._destroy_plan()
            return
        Kb = sum(j.K for j in jobs)
        old_X, old_HX = getattr(self, "X", None), getattr(self, "HX", None)
        # Kept jobs' old rows, absolute in the full buffers (the stack sits at
        # row0 of them).
        old_row0 = getattr(self, "row0", 0)
        old_at = {id(j): a for j, a in zip(getattr(self, "jobs", []), getattr(self, "k0", []))}
        realloc = self._buffers(Kb, max(j.n_tot for j in jobs), max(j.P for j in jobs))
        kcap, Gb, Pb = self._cap
        k0, acc = [], 0
        for j in jobs:
            k0.append(acc)
            acc += j.K
        # Where the new stack starts: the start that leaves the most kept rows
        # where they are (jobs keep their order; a retirement near the front
        # moves the few rows before it, one near the back the rows after it).
        row0 = 0
        if not realloc and keep:
            def moved(r0):
                return sum(j.K for j, a in zip(keep, k0) if old_row0 + old_at[id(j)] != r0 + a)
            cands = {0} | {old_row0 + old_at[id(j)] - a for j, a in zip(keep, k0)}
            row0 = min((r for r in cands if 0 <= r and r + Kb <= kcap), key=lambda r: (moved(r), r))
        Xn, HXn, X2n, HX2n = (b[row0 : row0 + Kb] for b in self._full)
        eig = torch.zeros((Kb, nb), dtype=torch.float64, device=dev)
        deig = torch.full((Kb, nb), np.inf, dtype=torch.float64, device=dev)
        cp, cp2 = _BlockCopies(), _BlockCopies()
        if keep:
            # Kept jobs in runs of consecutive old rows, one copy per run. With
            # new buffers every kept row is copied across; in the same buffers
            # only the rows that move, through the spare pair (first copy) and
            # back (second launch), since a move can overlap its own rows.
            runs = []  # (new row, old row, rows), relative to the new and old stacks
            for j, a in zip(keep, k0):
                o = old_at[id(j)]
                if not realloc and old_row0 + o == row0 + a:
                    continue
                if runs and runs[-1][1] + runs[-1][2] == o and runs[-1][0] + runs[-1][2] == a:
                    runs[-1][2] += j.K
                else:
                    runs.append([a, o, j.K])
            for a, o, n_rows in runs:  # (zero past the old basis length)
                if realloc:
                    cp.add(Xn[a : a + n_rows, :nb], old_X[o : o + n_rows, :nb])
                    cp.add(HXn[a : a + n_rows, :nb], old_HX[o : o + n_rows, :nb])
                else:
                    cp.add(X2n[a : a + n_rows, :nb], old_X[o : o + n_rows, :nb])
                    cp.add(HX2n[a : a + n_rows, :nb], old_HX[o : o + n_rows, :nb])
                    cp2.add(Xn[a : a + n_rows, :nb], X2n[a : a + n_rows, :nb])
                    cp2.add(HXn[a : a + n_rows, :nb], HX2n[a : a + n_rows, :nb])
            for j, a in zip(keep, k0):
                o = old_at[id(j)]
                cp.add(eig[a : a + j.K], self.eig[o : o + j.K])
                cp.add(deig[a : a + j.K], self.deig[o : o + j.K])
        # Rows of the projector, coupling and kinetic-energy buffers that a
        # job already holds at its new place (same buffers, same rows).
        old_abs = {i: old_row0 + a for i, a in old_at.items()}
        ke = self._ke_full[row0 : row0 + Kb]
        beta = self._beta_full[row0 : row0 + Kb]
        D = self._D_full[row0 : row0 + Kb]
        V64 = torch.empty((len(jobs), self.N), dtype=torch.float64, device=dev)
        v_slot = np.zeros(Kb, dtype=np.int32)
        for i, (j, a) in enumerate(zip(jobs, k0)):
            t = j.device_tables()
            if realloc or old_abs.get(id(j)) != row0 + a:
                # The job's rows with their padding zeroed (entries past its
                # basis and projectors), written only where they moved.
                rows = slice(a, a + j.K)
                cp.add(ke[rows], t["ke"])
                if j.P:
                    cp.add(beta[rows, : j.P], j.projectors(j.n_tot))
                    cp.add(D[rows], t["D"])
                else:
                    cp.add(D[rows])
                cp.add(beta[rows, j.P :])
            cp.add(V64[i], t["V"])
            v_slot[a : a + j.K] = i
        cp.run()
        cp2.run()
        n = np.concatenate([j.n for j in jobs])
        self._destroy_plan()
        geom, keepalive = _lib.make_geom_rows(self.shape, n, [j.fft_index for j in jobs], Gb,
                                              v_slot)
        self._plan = _lib.create_plan(geom)
        del keepalive
        self.jobs, self.k0, self.Kb, self.Gb, self.Pb, self.row0 = jobs, k0, Kb, Gb, Pb, row0
        self.X, self.HX, self.X2, self.HX2 = Xn, HXn, X2n, HX2n
        self.eig, self.deig = eig, deig
        self.ke = ke
        self.ke_r = ke.to(self.rdtype)
        self.n32 = _to_device(n.astype(np.int32))
        self.beta, self.D = beta, D
        self.V64 = V64
        self.V = V64 if dt == C128 else V64.to(self.rdtype)
        self.rnorm = self.pnorm2 = None
        need = int(_lib.lib().qmx_plan_workspace_bytes(self._plan, DTYPE_CODE[dt], 2 * nb))
        one = int(_lib.lib().qmx_plan_workspace_bytes(self._plan, DTYPE_CODE[C128], 1))
        size = max(one, min(need, WORKSPACE_CAP))
        ws = getattr(self, "workspace", None)
        if ws is None or ws.numel() < size or ws.numel() > 2 * size + (256 << 20):
            size = -(-size // (64 << 20)) * (64 << 20)  # 64 MB steps: few distinct sizes
            self.workspace = torch.empty(size, dtype=torch.uint8, device=dev)

    # ------------------------------------------------------------ operators
    def hamiltonian(self, Cin: torch.Tensor, out: torch.Tensor, k0: int, nk: int,
                    beta: Optional[torch.Tensor] = None, D: Optional[torch.Tensor] = None) -> None:
        """out = H Cin for the bucket k-points [k0, k0 + nk) (Cin, out:
        (nk, B, Gb) views with unit stride along G, either precision)."""
        B = Cin.shape[1]
        dt = Cin.dtype
        V = self.V64 if dt == C128 else self.V
        ke = self.ke if dt == C128 else self.ke_r
        if gemm.recorder is not None:
            key = ("plane", DTYPE_CODE[dt]) + tuple(int(x) for x in self.shape)
            gemm.recorder[key] = gemm.recorder.get(key, 0) + 1
        _lib.call("qmx_apply_local_k", self._plan, k0, nk, DTYPE_CODE[dt], _c(Cin), Cin.stride(1),
                  Cin.stride(0), _c(out), out.stride(1), out.stride(0), _c(V), _c(ke[k0:]),
                  ke.stride(0), B, 0, _c(self.workspace), self.workspace.numel(), rc.stream_ptr())
        if self.Pb:
            if beta is None:
                beta, D = self.beta[k0 : k0 + nk], self.D[k0 : k0 + nk]
            proj = _dot(beta, Cin)  # (nk, P, B), complex128
            proj *= D[:, :, None] if D.dtype == torch.float64 else D[:, :, None].double()
            gemm.rotate(proj, beta, out=out, beta=1.0)  # out += beta^T (D P)

    # ------------------------------------------------------------ job start
    def init_steps(self, new: list[CellJob]):
        """Start of the new jobs, whose rows close the stack: random bands in
        FP64 (_electrons.py:271-287; bandwidth-limited random bands have a
        condition number beyond float32), orthonormalised (canonical, on the
        FP64 eigenvectors of their overlap), then rotated to the Ritz vectors
        of their span (_davidson.py:164-169). A generator that yields its two
        Hermitian eigenproblems (see expand_steps)."""
        nb = self.nb
        a0 = self.k0[len(self.jobs) - len(new)]
        Kn = self.Kb - a0
        C = torch.zeros((Kn, nb, self.Gb), dtype=C128, device=rc.device)
        for j, a in zip(self.jobs[len(self.jobs) - len(new):], self.k0[len(self.jobs) - len(new):]):
            Cj = C[a - a0 : a - a0 + j.K]
            # Seeds follow the job's own basis (n_max = its n_tot):
            _lib.call("qmx_randomize", DTYPE_CODE[C128], _c(Cj), Cj.stride(1), Cj.stride(0), j.K,
                      self.Gb, j.n_tot, 0, _c(self.n32[a:]), _c(self.ke[a:]), self.ke.stride(0),
                      j.seed, 0, nb, rc.stream_ptr())
            j.admitted = time.perf_counter()
        lam, W = yield _dot(C, C)
        keep = lam > 1e-10 * lam[..., -1:].clamp_min(1e-300)
        U = W * torch.where(keep, lam.clamp_min(1e-300).rsqrt(), torch.zeros_like(lam))[..., None, :]
        C = gemm.rotate(U, C)
        # Bandwidth-limited random bands are nearly dependent: a direction
        # kept at overlap eigenvalue lambda is orthonormal only to about
        # eps lambda_max / lambda (1e-6 at the cut). A second pass (Loewdin on
        # the now near-identity overlap, error O(1e-18)) restores C^H C = I to
        # rounding, which every later step assumes.
        O1 = _dot(C, C)
        O1.diagonal(dim1=-2, dim2=-1).add_(torch.where(keep, 0.0, 1.0))  # dropped (zero) columns
        C = gemm.rotate(_lowdin(O1), C).to(self.dtype)
        HC = self.HX[a0:, :nb]
        self.hamiltonian(C, HC, a0, Kn)
        Hs = _dot(C, HC)
        Hs.diagonal(dim1=-2, dim2=-1).add_(torch.where(keep, 0.0, self.DROP_SHIFT))
        E, Vr = yield 0.5 * (Hs + dagger(Hs))
        gemm.rotate(Vr, C, out=self.X[a0:, :nb])
        self.HX[a0:, :nb] = gemm.rotate(Vr, HC)  # HC is these rows: rotate out of place
        self.eig[a0:] = E

    # ------------------------------------------------------------ iteration
    def residuals(self) -> None:
        """R = precondition(HC - C eig) into X[:, nb:], with the
        unpreconditioned residual norms for the convergence read
        (_davidson.py:182-184). Zero where C and HC are (the padding). Not
        normalised: the step folds the normalisation of _davidson.py:186-187
        into its small matrices."""
        nb = self.nb
        C, HC, R = self.X[:, :nb], self.HX[:, :nb], self.X[:, nb:]
        norm2 = torch.empty((self.Kb, nb), dtype=torch.float64, device=rc.device)
        rnorm2 = torch.empty_like(norm2)
        eig = self.eig.contiguous()
        # KEref (band_ke of C) is formed inside the launch (null pointer).
        _lib.call("qmx_precond_residual", DTYPE_CODE[self.dtype], _c(C), _c(HC), C.stride(1),
                  C.stride(0), _c(eig), None, _c(self.ke), self.ke.stride(0), _c(R),
                  R.stride(1), R.stride(0), self.Kb, nb, self.Gb, _c(norm2), _c(rnorm2),
                  rc.stream_ptr())
        self.rnorm = rnorm2.sqrt()
        self.pnorm2 = norm2

    def pack(self) -> torch.Tensor:
        """Convergence data for the host read: [rnorm, deig, eig, number of
        unconverged QL eigensolves since the last read] in one launch (the
        host checks the values for non-finite entries itself)."""
        if self._info:
            fail = torch.stack([(i != 0).sum() for i in self._info]).sum().reshape(1).to(
                torch.float64)
            self._info = []
        else:
            fail = getattr(self, "_zero1", None)
            if fail is None or fail.device != self.eig.device:
                fail = self._zero1 = torch.zeros(1, dtype=torch.float64, device=self.eig.device)
        return torch.cat([self.rnorm.flatten(), self.deig.flatten(), self.eig.flatten(), fail])

    def converged_jobs(self, packed: np.ndarray) -> list[CellJob]:
        nb = self.nb
        if packed[-1] != 0:
            raise RuntimeError(f"batched QL eigensolver did not converge ({int(packed[-1])} matrices)"
                               f" in bucket {self.shape}, {nb} bands")
        n = self.Kb * nb
        rnorm = packed[:n].reshape(self.Kb, nb)
        deig = packed[n : 2 * n].reshape(self.Kb, nb)
        eig = packed[2 * n : 3 * n].reshape(self.Kb, nb)
        bad = int((~np.isfinite(eig)).sum() + (~np.isfinite(rnorm)).sum())
        if bad:
            raise RuntimeError(f"non-finite eigenvalues or residuals ({bad}) in bucket"
                               f" {self.shape}, {nb} bands")
        done = []
        for j, a in zip(self.jobs, self.k0):
            rows = slice(a, a + j.K)
            if self.mixed and self.eig_tolerance is not None:
                # Kato-Temple estimate of every wanted band against half the
                # tolerance: the FP64 Rayleigh-Ritz that follows sets the
                # final eigenvalues and their own estimate.
                est = eig_error_estimate(eig[rows], rnorm[rows], j.n_bands)
                ok = est.max() < KT_SAFETY * self.eig_tolerance
            elif self.mixed:
                ok = rnorm[rows, : j.n_bands].max() < self.residual_threshold
            else:
                ok = deig[rows, : j.n_bands].max() < self.eig_threshold
            if ok or j.iterations >= self.max_iterations:
                j.converged = bool(ok)
                done.append(j)
        return done

    def expand_steps(self):
        """One Davidson step on every row (_davidson.py:180-233, no locking),
        as a generator that yields its two FP64 Hermitian eigenproblems
        (residual-block overlap, then the projected H) and receives their
        (eigenvalues, eigenvectors), so that a solver can run every bucket's
        eigenproblems in one launch."""
        nb, dt = self.nb, self.dtype
        C, R = self.X[:, :nb], self.X[:, nb:]
        HR = self.HX[:, nb:]
        for _ in range(self.GS_PASSES):  # R <- (1 - C C^H) R
            gemm.rotate(_dot(C, R), C, out=R, alpha=-1.0, beta=1.0)
        # Canonical orthonormalisation of the normalised residuals (drops
        # near-dependent directions). The normalisation (_davidson.py:186-187)
        # acts on the small matrices only: s = 1 / |R_i| before the
        # projection above (0 for a null row, below 1e-30) scales the Gram
        # matrix by s_i s_j and the canonical basis along its input axis, so
        # no pass over R is needed (row scaling commutes with the projection).
        d = self.pnorm2
        sc = torch.where(d > 1e-60, d.clamp_min(1e-300).rsqrt(), torch.zeros_like(d))
        lam, W = yield _dot(R, R) * (sc[:, :, None] * sc[:, None, :])
        tol = self.DROP_RTOL_MIXED if self.mixed else self.DROP_RTOL
        Ut = torch.empty((self.Kb, nb, nb), dtype=dt, device=rc.device)
        _lib.call("qmx_canonical_basis", DTYPE_CODE[dt], _c(W), _c(lam), self.Kb, nb, tol, _c(Ut),
                  rc.stream_ptr())
        Ut *= sc[:, None, :].to(Ut.real.dtype)  # Ut[j][i] s_i: R' = Ut diag(s) R
        if nb <= gemm.inplace_rows_max(dt):
            gemm.rotate(Ut.transpose(-2, -1), R, out=R)  # Ut @ R in place
        else:
            R.copy_(gemm.rotate(Ut.transpose(-2, -1), R))
        self.hamiltonian(R, HR, 0, self.Kb)
        XO = _dot(self.X, R)  # (Kb, 2 nb, nb): [C^H R; R^H R]
        XH = _dot(self.X, HR)
        n2 = 2 * nb
        E = torch.empty((self.Kb, n2, n2), dtype=C128, device=rc.device)
        H = torch.empty_like(E)
        _lib.call("qmx_subspace_assemble", _c(XO), _c(XH), _c(self.eig), _c(lam), self.Kb, nb, tol,
                  self.DROP_SHIFT, _c(E), _c(H), rc.stream_ptr())
        # Loewdin factor of O = I + E: I - E/2 + 3 E^2 / 8 (error O(|E|^3)).
        S = E.clone()
        gemm.matmul(E, E, out=S, alpha=0.375, beta=-0.5)
        S.diagonal(dim1=-2, dim2=-1).add_(1.0)
        # The eigensolver reads the lower triangle of S H S.
        E, W2 = yield gemm.matmul(gemm.matmul(S, H), S), nb  # only the nb lowest eigenvectors
        V = gemm.matmul(S, W2[:, :, :nb])  # (Kb, 2 nb, nb) complex128
        gemm.rotate(V, self.X, out=self.X2[:, :nb])
        gemm.rotate(V, self.HX, out=self.HX2[:, :nb])
        self.X, self.X2 = self.X2, self.X
        self.HX, self.HX2 = self.HX2, self.HX
        f = self._full
        self._full = [f[2], f[3], f[0], f[1]]
        E = E[:, :nb]
        self.deig = (E - self.eig).abs()
        self.eig = E.contiguous()
        for j in self.jobs:
            j.iterations += 1
        self.n_ticks += 1

    # ------------------------------------------------------------ results
    def finish_steps(self, job: CellJob):
        """Device results of a converged (or exhausted) job, as a generator
        that yields one Hermitian eigenproblem and returns a _Pending: an
        FP64 Rayleigh-Ritz on the job's nb bands (overlap, projected H and
        its eigenproblem in complex128) and the residual norms of the
        rotated bands, from which the host forms each band's error
        estimate. The mixed form first applies H in complex128 to its
        promoted bands; the plain form uses its H C, which every rotation
        has kept consistent with C. The packed result holds the nb
        eigenvalues and residual norms of every k-point, then a slot for the
        count of failed eigensolves (filled in by the solver)."""
        a = self.k0[self.jobs.index(job)]
        rows = slice(a, a + job.K)
        nb = self.nb
        if self.mixed:
            C = self.X[rows, :nb].to(C128)
            HC = torch.empty_like(C)
            beta = job.projectors(self.Gb) if job.P else None
            D = job.device_tables()["D"][None, :].expand(job.K, -1) if job.P else None
            self.hamiltonian(C, HC, a, job.K, beta, D)
        else:
            C, HC = self.X[rows, :nb], self.HX[rows, :nb]
        S = _lowdin(_dot(C, C))  # C is orthonormal to float32 rounding or better
        Hh = gemm.matmul(gemm.matmul(S, _dot(C, HC)), S)
        E, W = yield 0.5 * (Hh + dagger(Hh))
        Vs = gemm.matmul(S, W)
        C = gemm.rotate(Vs, C)
        HC = gemm.rotate(Vs, HC)
        rnorm = torch.linalg.vector_norm(HC - C * E[:, :, None].to(C128), dim=-1)
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return _Pending(job, torch.cat([E.flatten(), rnorm.flatten(),
                                        torch.zeros(1, dtype=torch.float64, device=rc.device)]), ev)


@dataclass
class _Pending:
    job: CellJob
    data: torch.Tensor
    event: "torch.cuda.Event"


@dataclass
class _GroupTail:
    """What the end of a group's step left for the host: its buckets with
    their convergence data, its finished jobs with their results, the
    number of failed eigensolves (host copies, valid once `event` has
    completed)."""
    event: Optional["torch.cuda.Event"] = None
    packs: list = field(default_factory=list)  # (bucket, host tensor)
    pending: list = field(default_factory=list)  # (_Pending, host tensor)
    fails: Optional[torch.Tensor] = None


# ---------------------------------------------------------------- solver
class Scheduler:
    """Host bookkeeping of the served solver, with no device calls (so that
    CPU tests drive it with simulated jobs and buckets): the FIFO queue, the
    HBM budget held by admitted jobs, the group of every bucket key and
    cancellations. Jobs need ``job_id``, ``arrival``, ``_key`` (shape, band
    class), ``K``, ``n_tot`` and ``hbm_bytes(nb, dtype, G)``."""

    def __init__(self, hbm_budget: float, max_k_per_bucket: int, max_bucket_bytes: float,
                 csize: int, dtype=None, n_groups: int = 1):
        self.hbm_budget = hbm_budget
        self.max_k_per_bucket = max_k_per_bucket
        self.max_bucket_bytes = max_bucket_bytes
        self.csize = csize
        self.dtype = dtype
        self.n_groups = max(1, int(n_groups))
        self.queue: list = []
        self.hbm_in_use = 0
        self.held: dict[int, int] = {}
        self.group_of: dict = {}  # bucket key -> group
        self.cancelled: set = set()  # job ids withdrawn (never give a result)
        self._cancel_running: set = set()  # running jobs to drop at their next tick

    def submit(self, jobs) -> None:
        self.queue.extend(jobs)

    def cancel(self, job_id: int, running_ids) -> bool:
        """Withdraw a job: from the queue at once, or (running) at its
        group's next tick. False when the job is neither queued nor
        running."""
        for i, j in enumerate(self.queue):
            if j.job_id == job_id:
                del self.queue[i]
                self.cancelled.add(job_id)
                return True
        if job_id in running_ids:
            self._cancel_running.add(job_id)
            return True
        return False

    def split_retiring(self, jobs, converged) -> tuple[list, list]:
        """(jobs to finish, jobs to drop) of one bucket at a tick: converged
        jobs finish unless cancelled; cancelled ones are dropped, result-less."""
        drop = [j for j in jobs if j.job_id in self._cancel_running]
        for j in drop:
            self._cancel_running.discard(j.job_id)
            self.cancelled.add(j.job_id)
        dropped = {j.job_id for j in drop}
        return [j for j in converged if j.job_id not in dropped], drop

    def release(self, job_id: int) -> None:
        """A job left its bucket: its HBM estimate is free for admission."""
        self.hbm_in_use -= self.held.pop(job_id, 0)

    def admit(self, now: Optional[float], group: int, buckets: dict, reserved: int,
              loads) -> dict:
        """Take queued jobs (arrived by `now`) that fit into buckets of
        `group` (a key seen for the first time joins the group with the
        least work, `loads()` giving each group's); returns key -> jobs.
        `buckets` maps key -> (k-points, basis size) of the bucket's current
        jobs. The budget holds against `reserved` (the bytes live device
        tensors hold) plus the estimates admitted in this call, and against
        the estimates of every admitted job; the first job is always taken.
        The queue is FIFO: once the budget is spent no later job is taken,
        and a full bucket (k-points or state bytes) takes no later job of
        its key."""
        taken: dict = {}
        rest = []
        load = None
        pending = 0
        full: set = set()
        for pos, job in enumerate(self.queue):
            if now is not None and job.arrival > now:
                rest.extend(self.queue[pos:])  # arrivals are in order
                break
            key = job._key
            if key in full:
                rest.append(job)
                continue
            if key not in self.group_of:
                if load is None:
                    load = loads()
                g = int(np.argmin(load))
                self.group_of[key] = g
                load[g] += float(job.K) * job.n_tot * key[1]
            if self.group_of[key] != group:
                rest.append(job)
                continue
            kb, gb = buckets.get(key, (0, 0))
            joining = taken.get(key, [])
            k_in = kb + sum(j.K for j in joining)
            G = max([gb] + [j.n_tot for j in joining] + [job.n_tot])
            need = job.hbm_bytes(key[1], self.dtype, G)
            if max(self.hbm_in_use, reserved + pending) + need > self.hbm_budget and (
                    self.hbm_in_use > 0):
                rest.extend(self.queue[pos:])
                break
            in_bucket = 4 * (2 * key[1]) * G * self.csize * (k_in + job.K)  # X, HX, buffers
            if k_in > 0 and (k_in + job.K > self.max_k_per_bucket
                             or in_bucket > self.max_bucket_bytes):
                full.add(key)
                rest.append(job)
                continue
            pending += need
            self.hbm_in_use += need
            self.held[job.job_id] = need
            taken.setdefault(key, []).append(job)
        self.queue = rest
        return taken


class BatchSolver:
    """Runs band-solve jobs in buckets, admitting new jobs as others finish.

    ``submit`` queues jobs (arrival times relative to the solver's start).
    Buckets are dealt to ``groups`` groups (balanced by basis size) that
    ``tick`` serves in turn: a tick of group g waits for g's convergence data
    and finished results (copied to the host at the end of g's previous
    step), retires g's converged jobs, admits arrived jobs into g's buckets
    (HBM budget and per-bucket k-point limit) and runs one Davidson step of
    each of g's buckets, ending with their residuals and the device-to-host
    copy of the next convergence data. With more than one group the device
    runs one group's step while the host works on another, which pays when
    the host, not the device, limits the rate; each job still takes one
    Davidson step per visit of its group and its results are those of a
    single-group solver. Two groups are the default: measured on the mixed
    served stream (1300 requests), two give 92 band solves/s against 84-86
    for one, and three no more than two. Each group's launches go to a
    stream of its own, so that one group's latency-bound phases (the grouped
    eigensolves) overlap the other group's kernels; cached workspaces are
    kept per stream. ``run`` ticks until every submitted job is done.
    """

    def __init__(self, dtype: torch.dtype = C64, *, eig_tolerance: Optional[float] = 1e-6,
                 residual_threshold: float = 3e-4,
                 eig_threshold: float = 1e-8, max_iterations: int = 200,
                 band_classes: Optional[Sequence[int]] = (32, 48, 72), hbm_budget: float = 180e9,
                 max_k_per_bucket: int = 2048, max_bucket_bytes: float = 16e9, groups: int = 2,
                 streams: bool = True):
        rc.require_gpu()
        from . import autotune

        #: tuned launch choices applied from this device's cache (once per process)
        self.tuned = autotune.load_default()
        self.dtype = dtype
        #: mixed form: stop a job when every wanted band's Kato-Temple error
        #: estimate is below KT_SAFETY * eig_tolerance (Ha); None stops on
        #: residual norms below residual_threshold instead. The plain form
        #: stops on the reference's eigenvalue change below eig_threshold.
        self.eig_tolerance = eig_tolerance
        self.residual_threshold = residual_threshold
        self.eig_threshold = eig_threshold
        self.max_iterations = max_iterations
        self.band_classes = sorted(band_classes) if band_classes else None
        self.hbm_budget = hbm_budget
        self.max_k_per_bucket = max_k_per_bucket
        #: Largest iteration state (X, HX, double-buffered) of one bucket:
        #: beyond it a larger batch adds no parallelism the device lacks.
        self.max_bucket_bytes = max_bucket_bytes
        self._csize = torch.empty((), dtype=dtype).element_size()
        self.buckets: dict = {}
        self.results: dict[int, CellResult] = {}
        self.pending: list[_Pending] = []
        self._info: list[torch.Tensor] = []
        self.n_groups = max(1, int(groups))
        self.sched = Scheduler(hbm_budget, max_k_per_bucket, max_bucket_bytes, self._csize,
                               dtype, self.n_groups)
        self._tail = [_GroupTail() for _ in range(self.n_groups)]
        self._turn = 0
        self.n_ticks = 0
        # One stream per bucket group (default priority): a bucket stays in
        # its group, so each group's work is ordered on its stream, and while
        # one group is in a latency-bound phase (its grouped eigensolves above
        # all, a few workgroups per matrix) the other group's kernels fill the
        # device.
        # streams=False keeps every launch on the caller's current stream.
        self._streams = ([torch.cuda.Stream() for _ in range(self.n_groups)]
                         if streams and self.n_groups > 1 else None)
        torch.cuda.synchronize()
        self._t0_event = torch.cuda.Event(enable_timing=True)
        self._t0_event.record()
        torch.cuda.synchronize()
        self.t0 = time.perf_counter()

    def clock(self) -> float:
        """Seconds since the solver started (the time base of arrivals)."""
        return time.perf_counter() - self.t0

    @property
    def queue(self) -> list:
        return self.sched.queue

    @property
    def hbm_in_use(self) -> int:
        return self.sched.hbm_in_use

    @property
    def cancelled(self) -> set:
        return self.sched.cancelled

    def cancel(self, job_id: int) -> bool:
        """Withdraw a request: a queued job leaves the queue; a running job
        leaves its bucket at its group's next tick (its HBM and bucket rows
        then free for admission). A cancelled job gives no result; the
        other jobs' results are those they would have had without it.
        False when the job is neither queued nor running."""
        running = {j.job_id for b in self.buckets.values() for j in b.jobs}
        return self.sched.cancel(job_id, running)

    def band_class(self, job: CellJob) -> int:
        """Smallest configured class that holds the job's n_bands +
        n_bands_extra and keeps the expanded subspace (twice the class)
        below its smallest basis (_davidson.py:159-163); else its own size
        rounded up to a multiple of 8 within that limit."""
        limit = (int(job.n.min()) - 1) // 2
        if self.band_classes:
            for c in self.band_classes:
                if job.n_bands_max <= c <= limit:
                    return c
        c = int(np.ceil(job.n_bands_max / 8) * 8)
        return c if c <= limit else job.n_bands_max

    def submit(self, jobs: Sequence[CellJob]) -> None:
        """Queue jobs (FIFO); their bucket key is fixed here, once."""
        for job in jobs:
            job._key = (job.shape, self.band_class(job))
        self.sched.submit(jobs)

    def _bucket(self, key) -> Bucket:
        b = self.buckets.get(key)
        if b is None:
            b = Bucket(key[0], key[1], self.dtype, residual_threshold=self.residual_threshold,
                       eig_threshold=self.eig_threshold, max_iterations=self.max_iterations,
                       eig_tolerance=self.eig_tolerance)
            self.buckets[key] = b
        return b

    def _group_loads(self) -> list[float]:
        load = [0.0] * self.n_groups
        for key, b in self.buckets.items():
            if b.jobs:
                load[self.sched.group_of[key]] += float(b.Kb) * b.Gb * b.nb
        return load

    def _admit(self, now: Optional[float], group: int = 0) -> dict:
        """Queued jobs (arrived by `now`) that fit into buckets of `group`
        (Scheduler.admit against the bytes live tensors hold: a host-side
        counter of the caching allocator, whose cached free blocks are
        reused and so do not count); returns key -> jobs."""
        state = {key: (b.Kb, b.Gb) for key, b in self.buckets.items() if b.jobs}
        return self.sched.admit(now, group, state, torch.cuda.memory_allocated(rc.device),
                                self._group_loads)

    def _collect(self, pending: list, host: list) -> None:
        for p, h in zip(pending, host):
            job = p.job
            K, B = job.K, job.n_bands
            if h[-1] != 0:
                raise RuntimeError(f"batched QL eigensolver did not converge (job {job.job_id})")
            nb = (len(h) - 1) // (2 * K)
            eig_all = h[: K * nb].reshape(K, nb)
            r_all = h[K * nb : 2 * K * nb].reshape(K, nb)
            res = CellResult(
                job_id=job.job_id, eig=eig_all[:, :B].copy(), residual_norm=r_all[:, :B].copy(),
                err_bound=float(eig_error_estimate(eig_all, r_all, B).max()),
                Eband=float(2.0 * (job.wk[:, None] * eig_all[:, :B]).sum()), iterations=job.iterations,
                converged=getattr(job, "converged", False), arrival=job.arrival,
                admitted=(job.admitted - self.t0) if job.admitted else job.arrival,
                finished=self._t0_event.elapsed_time(p.event) * 1e-3)
            job.result = res
            self.results[job.job_id] = res

    def tick(self, now: Optional[float] = None) -> list[CellResult]:
        """One iteration of the buckets of the next group, on that group's
        stream; returns the results that became available (that group's jobs
        finished in its previous tick)."""
        if self._streams is None:
            return self._tick(now)
        with torch.cuda.stream(self._streams[self._turn]):
            return self._tick(now)

    def _tick(self, now: Optional[float]) -> list[CellResult]:
        g = self._turn
        self._turn = (g + 1) % self.n_groups
        _stamp[0] += 1
        tail, self._tail[g] = self._tail[g], _GroupTail()
        # 1. the group's convergence data and finished results (copied at
        #    the end of its previous step; the other groups' work stays
        #    queued). Work queued n_groups ticks ago is complete.
        if tail.event is not None:
            tail.event.synchronize()
        flush_retired_plans(_stamp[0] - self.n_groups)
        if tail.fails is not None and int(tail.fails) != 0:
            raise RuntimeError(f"grouped QL eigensolver did not converge ({int(tail.fails)} matrices)")
        self._collect([p for p, _ in tail.pending], [h.numpy() for _, h in tail.pending])
        new_results = [self.results[p.job.job_id] for p, _ in tail.pending]
        active = [b for b, _ in tail.packs if b.jobs]
        host = {id(b): h.numpy() for b, h in tail.packs}
        # 2. converged jobs (their HBM is free for admission) and admissions
        #    (host decisions only).
        retire, drop = {}, {}
        for b in active:
            retire[id(b)], drop[id(b)] = self.sched.split_retiring(
                b.jobs, b.converged_jobs(host[id(b)]))
            for job in retire[id(b)] + drop[id(b)]:
                self.sched.release(job.job_id)
        newcomers = self._admit(now, g)
        # 3. the Davidson step of every bucket that keeps its jobs starts
        #    first (up to its first eigenproblem).
        steps, smats = [], []
        for b in active:
            if not retire[id(b)] and not drop[id(b)] and (b.shape, b.nb) not in newcomers:
                self._start(b.expand_steps(), steps, smats)
        # 4. retirement (each finish runs up to its eigenproblem before the
        #    restack), admission, restack and start of new jobs.
        gens, mats = [], []
        for b in active:
            for job in retire[id(b)]:
                self._start(b.finish_steps(job), gens, mats)
        keys = sorted(set(newcomers) | {(b.shape, b.nb) for b in active})
        changed = []
        for key in keys:
            b = self._bucket(key)
            new = newcomers.get(key, [])
            gone = retire.get(id(b), []) + drop.get(id(b), [])
            if new or gone:
                b._restack([j for j in b.jobs if j not in gone], new)
                for job in gone:
                    job.release()
                if not b.jobs:
                    continue
                changed.append(b)
                if new:
                    self._start(b.init_steps(new), gens, mats)
        self.pending.extend(self._drive(gens, mats))
        for b in changed:
            b.residuals()
        # 5. the Davidson step of the changed buckets; all steps' later
        #    eigenproblems run grouped.
        for b in changed:
            self._start(b.expand_steps(), steps, smats)
        self._drive(steps, smats)
        # 6. the group's tail: residuals of the new state and the copies
        #    the group's next tick reads.
        nxt = self._tail[g]
        for key, b in self.buckets.items():
            if b.jobs and self.sched.group_of.get(key) == g:
                b.residuals()
                nxt.packs.append((b, b.pack().to("cpu", non_blocking=True)))
        nxt.pending = [(p, p.data.to("cpu", non_blocking=True)) for p in self.pending]
        self.pending = []
        if self._info:
            nxt.fails = torch.stack([(i != 0).sum() for i in self._info]).sum().to(
                "cpu", non_blocking=True)
            self._info = []
        # Recorded every tick (also one that left nothing to read): the
        # group's next tick waits on it before it frees retired plans, whose
        # last launches may be on this group's stream.
        nxt.event = torch.cuda.Event()
        nxt.event.record()
        self.n_ticks += 1
        return new_results

    def _start(self, gen, gens: list, mats: list) -> None:
        """Run `gen` to its first eigenproblem (queued in gens/mats); a
        generator that returns before yielding delivers its value to the
        pending results."""
        try:
            mats.append(next(gen))
            gens.append(gen)
        except StopIteration as stop:
            if stop.value is not None:
                self.pending.append(stop.value)

    def _drive(self, gens: list, mats: list) -> list:
        """Run generators that yield Hermitian eigenproblems to completion,
        all pending eigenproblems of a round in one grouped launch; returns
        the generators' return values (those that are not None). A generator
        yields a matrix batch, or (matrix batch, n_want) when it needs only
        the eigenvectors of the n_want lowest eigenvalues."""
        out = []
        while gens:
            want = [m[1] if isinstance(m, tuple) else None for m in mats]
            mats = [m[0] if isinstance(m, tuple) else m for m in mats]
            results, info = eigh_grouped(mats, want)
            self._info.append(info)
            nxt_gens, nxt_mats = [], []
            for g, r in zip(gens, results):
                try:
                    nxt_mats.append(g.send(r))
                    nxt_gens.append(g)
                except StopIteration as stop:
                    if stop.value is not None:
                        out.append(stop.value)
            gens, mats = nxt_gens, nxt_mats
        return out

    @property
    def idle(self) -> bool:
        return (not self.queue and not self.pending and not any(t.pending for t in self._tail)
                and not any(b.jobs for b in self.buckets.values()))

    def run(self) -> dict[int, CellResult]:
        while not self.idle:
            self.tick()
        torch.cuda.synchronize()
        flush_retired_plans()
        return self.results
