"""ctypes binding of libqmx.so, the C ABI declared in include/qmx.h.

The library is loaded once, from ``qimpy_amd/_lib/libqmx.so`` (built in-tree
by ``qimpy_amd.build``). There is no fallback: if the library is missing or a
call fails, a ``RuntimeError`` carrying ``qmx_last_error()`` is raised.
"""
from __future__ import annotations

import ctypes as ct
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libqmx.so"

QMX_C64, QMX_C128 = 0, 1
QMX_KERNEL_PROJ, QMX_KERNEL_C128 = 10, 16

c_i32p = ct.POINTER(ct.c_int32)
c_i64p = ct.POINTER(ct.c_int64)
c_f64p = ct.POINTER(ct.c_double)
vp = ct.c_void_p
i64 = ct.c_int64
i32 = ct.c_int


class QmxGeom(ct.Structure):
    _fields_ = [
        ("shape", ct.c_int32 * 3),
        ("nk", ct.c_int32),
        ("n_tot", ct.c_int32),
        ("n", c_i32p),
        ("fft_index", c_i64p),
        ("v_slot", c_i32p),
        ("fft_rows", ct.POINTER(c_i64p)),
    ]


# name -> (restype, argtypes)
SIGNATURES = {
    "qmx_copy_blocks": (i32, [vp, i64, i64, vp]),
    "qmx_geom_tables": (i32, [ct.POINTER(QmxGeom), c_i64p, c_i32p, c_i32p, c_i32p, c_i32p, c_i32p]),
    "qmx_plan_create": (i32, [ct.POINTER(QmxGeom), ct.POINTER(vp)]),
    "qmx_plan_destroy": (i32, [vp]),
    "qmx_plan_workspace_bytes": (i64, [vp, i32, i64]),
    "qmx_plan_info": (i32, [vp, c_i64p]),
    "qmx_plan_set_local_path": (i32, [vp, i32]),
    "qmx_plan_local_path": (i32, [vp, i32]),
    "qmx_apply_local": (
        i32,
        [vp, i32, vp, i64, i64, vp, i64, i64, vp, vp, i64, i64, i32, vp, i64, vp],
    ),
    "qmx_apply_local_k": (
        i32,
        [vp, i64, i64, i32, vp, i64, i64, vp, i64, i64, vp, vp, i64, i64, i32, vp, i64, vp],
    ),
    "qmx_fft3d": (i32, [i32, c_i32p, vp, i64, i32, ct.c_double, vp]),
    "qmx_band_reduce": (i32, [i32, vp, i64, i64, i64, i64, i64, vp, i64, vp, vp, vp]),
    "qmx_precond_residual": (
        i32,
        [i32, vp, vp, i64, i64, vp, vp, vp, i64, vp, i64, i64, i64, i64, i64, vp, vp, vp],
    ),
    "qmx_precond_residual_normalize": (
        i32,
        [i32, vp, vp, i64, i64, vp, vp, vp, i64, vp, i64, i64, i64, i64, i64, vp, vp, vp, vp],
    ),
    "qmx_scale_constrain": (i32, [i32, vp, i64, i64, i64, i64, i64, vp, vp, vp]),
    "qmx_axpby": (i32, [i32, i64, i64, i64, vp, vp, vp, i64, i64, vp, vp, vp, vp, i64, i64, vp]),
    "qmx_apply_gradient": (i32, [i32, vp, i64, i64, i64, i64, i64, vp, i32, ct.c_double, vp, i64,
                                 i64, i32, vp]),
    "qmx_randomize": (
        i32,
        [i32, vp, i64, i64, i64, i64, i64, i64, vp, vp, i64, i64, i64, i64, vp],
    ),
    "qmx_heevj_workspace_bytes": (i64, [i64, i64]),
    "qmx_heevj_batched": (i32, [vp, i64, i64, i64, i64, vp, vp, i64, i64, vp, i64, i32, vp]),
    "qmx_heev_small_workspace_bytes": (i64, [i64, i64]),
    "qmx_heev_small_batched": (i32, [vp, i64, i64, i64, i64, vp, vp, i64, i64, vp, vp, i64, vp]),
    "qmx_dot_workspace_bytes": (i64, [i32, i64, i64, i64, i64]),
    "qmx_dot": (i32, [i32, vp, i64, i64, vp, i64, i64, i64, i64, i64, i64, vp, i64, i64, i32,
                      vp, i64, vp]),
    "qmx_density_workspace_bytes": (i64, [vp, i32, i64]),
    "qmx_collect_density": (i32, [vp, i64, i64, i32, vp, i64, i64, i64, vp, i64, vp, i32, vp, i64,
                                  vp]),
    "qmx_density_cross_workspace_bytes": (i64, [vp, i32, i64]),
    "qmx_projectors": (i32, [i64, i64, vp, vp, vp, vp, i64, vp, i64, vp, vp, vp, i64, i64, vp,
                             ct.c_double, ct.c_double, ct.c_double, vp, i64, i64, vp]),
    "qmx_gaussian_projectors": (i32, [i64, i64, vp, vp, vp, vp, i64, vp, i64, vp, vp, vp, i64, vp,
                                      ct.c_double, vp, i64, i64, vp]),
    "qmx_radial_transform": (i32, [i64, i64, vp, vp, vp, vp, i64, ct.c_double, vp, vp]),
    "qmx_collect_density_cross": (i32, [vp, i64, i64, i32, vp, vp, i64, i64, i64, vp, i64, vp, i32,
                                        vp, i64, vp]),
    "qmx_heev_small_work_bytes": (i64, [i64]),
    "qmx_heev_small_grouped": (i32, [vp, i64, i64, vp]),
    "qmx_heevx_work_bytes": (i64, [i64, i64]),
    "qmx_heevx_grouped": (i32, [vp, i64, i64, vp]),
    "qmx_heevx_large_work_bytes": (i64, [i64, i64]),
    "qmx_heevx_large": (i32, [vp, i64, i64, i64, vp, vp, i64, i64, vp, vp, i64, i32, vp]),
    "qmx_potrf_inv_work_bytes": (i64, [i64]),
    "qmx_potrf_inv": (i32, [vp, i64, i64, vp, vp, vp, i64, i32, vp]),
    "qmx_tune_bucket": (i64, [i64]),
    "qmx_tune_batch_bucket": (i64, [i64]),
    "qmx_tune_set": (i32, [i32, vp, i32, vp, i32]),
    "qmx_tune_clear": (i32, []),
    "qmx_tune_count": (i64, []),
    "qmx_normalize_constrain": (i32, [i32, vp, i64, i64, i64, i64, i64, vp, vp, vp]),
    "qmx_canonical_basis": (i32, [i32, vp, vp, i64, i64, ct.c_double, vp, vp]),
    "qmx_subspace_assemble": (i32, [vp, vp, vp, vp, i64, i64, ct.c_double, ct.c_double, vp, vp,
                                    vp]),
    "qmx_dot_noalloc_max_bytes": (i64, []),
    "qmx_dot_noalloc": (i32, [i32, vp, i64, i64, vp, i64, i64, i64, i64, i64, i64, vp, i64, i64,
                              i32, vp]),
    "qmx_gemm": (i32, [i32, i32, i32, i32, i64, i64, i64, i64, ct.c_double, vp, i64, i64, vp, i64,
                       i64, ct.c_double, vp, i64, i64, vp]),
    "qmx_gemm_c64": (i32, [i32, i32, i64, i64, i64, ct.c_double, vp, i64, i64, vp, i64, i64,
                           ct.c_double, vp, i64, i64, i64, vp]),
    "qmx_gemm_c64_bf16x3": (i32, [i32, i32, i64, i64, i64, ct.c_double, vp, i64, i64, vp, i64,
                                  i64, ct.c_double, vp, i64, i64, i64, vp]),
    "qmx_gemm_c128": (i32, [i32, i32, i64, i64, i64, ct.c_double, vp, i64, i64, vp, i64, i64,
                            ct.c_double, vp, i64, i64, i64, vp]),
    "qmx_profile_enable": (i32, [i32]),
    "qmx_profile_read": (i32, [vp, vp, vp, i32]),
    "qmx_profile_read_flops": (i32, [vp, vp, vp, vp, i32]),
    "qmx_last_error": (ct.c_char_p, []),
    "qmx_version": (ct.c_char_p, []),
}

_lib = None


def lib() -> ct.CDLL:
    """The loaded library (raises if it was never built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m qimpy_amd.build`"
            )
        handle = ct.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def check(status: int, what: str) -> None:
    if status != 0:
        msg = lib().qmx_last_error().decode(errors="replace")
        raise RuntimeError(f"{what} failed (status {status}): {msg}")


def call(name: str, *args) -> int:
    status = getattr(lib(), name)(*args)
    check(status, name)
    return status


def make_geom(shape, n: np.ndarray, fft_index: np.ndarray,
              v_slot: np.ndarray | None = None) -> tuple[QmxGeom, tuple]:
    """Geometry struct over host arrays (returned tuple keeps them alive)."""
    n32 = np.ascontiguousarray(n, dtype=np.int32)
    fi = np.ascontiguousarray(fft_index, dtype=np.int64)
    nk, n_tot = fi.shape
    g = QmxGeom()
    for i in range(3):
        g.shape[i] = int(shape[i])
    g.nk = nk
    g.n_tot = n_tot
    g.n = n32.ctypes.data_as(c_i32p)
    g.fft_index = fi.ctypes.data_as(c_i64p)
    vs = _set_v_slot(g, nk, v_slot)
    return g, (n32, fi, vs)


def _set_v_slot(g: QmxGeom, nk: int, v_slot) -> np.ndarray | None:
    if v_slot is None:
        return None
    vs = np.ascontiguousarray(v_slot, dtype=np.int32)
    if vs.shape != (nk,):
        raise ValueError("v_slot needs one entry per k-point")
    g.v_slot = vs.ctypes.data_as(c_i32p)
    return vs


def make_geom_rows(shape, n: np.ndarray, blocks: list, n_tot: int,
                   v_slot: np.ndarray | None = None) -> tuple[QmxGeom, tuple]:
    """Geometry struct whose k-points are the rows of `blocks` (int64 arrays
    (K_i, n_tot_i), C-contiguous, in order) through per-k row pointers, so
    that the index rows of many cells need no concatenated copy. The blocks
    must stay alive while the struct is used (the returned tuple holds
    them)."""
    n32 = np.ascontiguousarray(n, dtype=np.int32)
    rows = []
    for b in blocks:
        if b.dtype != np.int64 or not b.flags.c_contiguous or b.ndim != 2:
            raise ValueError("index blocks must be C-contiguous int64 matrices")
        if b.shape[1] > n_tot:
            raise ValueError("index block wider than n_tot")
        rows.append(b.ctypes.data + b.strides[0] * np.arange(b.shape[0], dtype=np.uint64))
    ptrs = np.ascontiguousarray(np.concatenate(rows) if rows else np.zeros(0, np.uint64))
    nk = len(ptrs)
    if n32.shape != (nk,):
        raise ValueError("n needs one entry per k-point")
    if np.any(n32 > np.concatenate([np.full(b.shape[0], b.shape[1]) for b in blocks])):
        raise ValueError("n[k] exceeds its index row")
    g = QmxGeom()
    for i in range(3):
        g.shape[i] = int(shape[i])
    g.nk = nk
    g.n_tot = int(n_tot)
    g.n = n32.ctypes.data_as(c_i32p)
    g.fft_rows = ptrs.ctypes.data_as(ct.POINTER(c_i64p))
    vs = _set_v_slot(g, nk, v_slot)
    return g, (n32, ptrs, vs, blocks)


def create_plan(geom: QmxGeom) -> ct.c_void_p:
    """qmx_plan_create, with the local-apply kernel sequence taken from
    QMX_LOCAL_PATH (0 automatic, 1 separate passes, 2 fused plane pass where
    it fits; default 0)."""
    handle = ct.c_void_p()
    call("qmx_plan_create", ct.byref(geom), ct.byref(handle))
    path = int(os.environ.get("QMX_LOCAL_PATH", "0"))
    if path:
        call("qmx_plan_set_local_path", handle, path)
    return handle


def geom_tables(shape, n: np.ndarray, fft_index: np.ndarray) -> dict:
    """Host-side sparse-FFT tables (qmx_geom_tables) as numpy arrays."""
    g, keep = make_geom(shape, n, fft_index)
    sizes = np.zeros(2, dtype=np.int64)
    null = ct.cast(None, c_i32p)
    call("qmx_geom_tables", ct.byref(g), sizes.ctypes.data_as(c_i64p), null, null, null, null, null)
    nxa, ncol = int(sizes[0]), int(sizes[1])
    nk, n_tot = keep[1].shape
    out = {
        "x_list": np.zeros(nxa, np.int32),
        "pcol_start": np.zeros(nxa + 1, np.int32),
        "col_iy": np.zeros(ncol, np.int32),
        "cstart": np.zeros((nk, ncol + 1), np.int32),
        "t1idx": np.zeros((nk, n_tot), np.int32),
    }
    p = lambda a: a.ctypes.data_as(c_i32p)
    call(
        "qmx_geom_tables", ct.byref(g), sizes.ctypes.data_as(c_i64p), p(out["x_list"]),
        p(out["pcol_start"]), p(out["col_iy"]), p(out["cstart"]), p(out["t1idx"]),
    )
    out["nxa"], out["ncol"] = nxa, ncol
    return out


_BASE_NAMES = {
    0: "k_sphere_to_cols", 1: "k_cols_to_planes", 2: "k_xpass_potential",
    3: "k_planes_to_cols", 4: "k_cols_to_sphere", 5: "k_fft", 6: "k_gemm",
    7: "k_plane_potential", 8: "k_potential_zmajor",
    9: "local_apply",  # the whole apply (a span of launches, not one kernel)
    10: "k_projectors", 11: "k_rotate", 12: "k_rotate_bf16x3",
}
#: Profile ids: the kernel id for complex64 launches, + 16 for complex128.
KERNEL_NAMES = {**{i: f"{n}<c64>" for i, n in _BASE_NAMES.items()},
                **{16 + i: f"{n}<c128>" for i, n in _BASE_NAMES.items()}}
N_KERNEL_IDS = 32


def profile_enable(on: bool) -> None:
    call("qmx_profile_enable", int(on))


def profile_read() -> dict:
    """{kernel id: (total ms, total compulsory bytes, launches, total
    algorithmic flops)} since the last read."""
    ms = np.zeros(N_KERNEL_IDS)
    by = np.zeros(N_KERNEL_IDS)
    fl = np.zeros(N_KERNEL_IDS)
    cnt = np.zeros(N_KERNEL_IDS, dtype=np.int64)
    call("qmx_profile_read_flops", ms.ctypes.data, by.ctypes.data, fl.ctypes.data,
         cnt.ctypes.data, N_KERNEL_IDS)
    return {i: (float(ms[i]), float(by[i]), int(cnt[i]), float(fl[i]))
            for i in range(N_KERNEL_IDS) if cnt[i]}
