"""Autotuner of the launch choices of qmx_dot, qmx_gemm and the fused plane
pass (no counterpart in the reference, which has no device kernels to
tune).

- ``record()``: a context in which the library's entry points note every
  shape they launch (gemm.recorder);
- ``tune(shapes)``: for each distinct tuning key of those shapes, times the
  candidate choices on random operands of the recorded shape with device
  events (median of REPS launches after a warm-up), and sets the fastest in
  the library's table (qmx_tune_set) when it beats the library's own choice
  by more than MARGIN; returns the entries;
- ``save(path, entries)`` and ``load(path)``: the JSON cache, with entries
  per device name; ``load_default()`` applies the cache shipped for the
  current device model (qimpy_amd/tuning/<arch>_<CUs>cu.json), once per
  process.

Keys follow include/qmx.h (QMX_TUNE_*): the overlap product by (dtype,
batch bucket, M, N, basis bucket), the rotation by (dtype, matrix dtype,
variant, m, K, basis bucket), the plane pass by (dtype, box).

Run ``python -m qimpy_amd.autotune`` on the device to tune the shapes of the
served workload (cells of the batched distribution through BatchSolver, both
precisions) and write the default cache.
"""
from __future__ import annotations

import contextlib
import ctypes as ct
import json
import math
import os
import sys
import time
from pathlib import Path
from typing import Optional

import numpy as np
import torch

from . import _lib, gemm, rc

TUNE_DOT, TUNE_ROTATE, TUNE_PLANE = 1, 2, 3
REPS = 5
MARGIN = 0.03  # a tuned choice must beat the default by 3 %
TUNING_DIR = Path(__file__).resolve().parent / "tuning"
_KB_ROWS16, _KB_TILES = 14, 48  # block-kernel limits of qmx_dot (bm16 + bn16, bm16 * bn16)
_loaded_default = False


def device_name() -> str:
    """The device model the cache is for: its architecture and CU count
    (e.g. gfx950_256cu)."""
    if not torch.cuda.is_available():
        return "cpu"
    prop = torch.cuda.get_device_properties(0)
    arch = str(getattr(prop, "gcnArchName", "gpu")).split(":")[0]
    return f"{arch}_{prop.multi_processor_count}cu"


@contextlib.contextmanager
def record():
    """Collect the shapes launched inside the block (shape tuple -> count)."""
    prev, gemm.recorder = gemm.recorder, {}
    try:
        yield gemm.recorder
    finally:
        shapes, gemm.recorder = gemm.recorder, prev
        if prev is not None:
            for k, v in shapes.items():
                prev[k] = prev.get(k, 0) + v


def _set(op: int, key, choice) -> None:
    k = (ct.c_int64 * len(key))(*[int(v) for v in key])
    c = (ct.c_int64 * len(choice))(*[int(v) for v in choice])
    _lib.call("qmx_tune_set", op, ct.addressof(k), len(key), ct.addressof(c), len(choice))


def clear() -> None:
    _lib.call("qmx_tune_clear")


def _timed(fn, reps: int = REPS) -> float:
    """Median device time (ms) of fn() after one warm-up call."""
    fn()
    ts = []
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        ts.append(a.elapsed_time(b))
    return float(np.median(ts))


def _dtype(code: int) -> torch.dtype:
    return torch.complex64 if code == 0 else torch.complex128


def _rand(shape, dtype, seed):
    g = torch.Generator(device="cuda").manual_seed(seed)
    return torch.randn(shape, dtype=dtype, device="cuda", generator=g)


# ---------------------------------------------------------------- qmx_dot
def dot_key(code, batch, M, N, G) -> tuple:
    lib = _lib.lib()
    return (code, int(lib.qmx_tune_batch_bucket(batch)), M, N, int(lib.qmx_tune_bucket(G)))


def dot_candidates(batch, M, N, G) -> list:
    """(kernel, bm16, bn16, chunks): block shapes giving 1-4 blocks along each
    axis within the kernel's limits, the 64 x 64 tile kernel, each with chunk
    counts that aim at 256..4096 workgroups (at least 256 basis points per
    chunk)."""
    M16, N16 = -(-M // 16), -(-N // 16)
    shapes = set()
    for mb in range(1, 5):
        for nb in range(1, 5):
            bm, bn = -(-M16 // mb), -(-N16 // nb)
            if bm + bn <= _KB_ROWS16 and bm * bn <= _KB_TILES:
                shapes.add((0, bm, bn))
    shapes.add((1, 0, 0))
    most = max(1, -(-G // 256))
    out = set()
    for kern, bm, bn in shapes:
        tiles = batch * (-(-M16 // bm) * -(-N16 // bn) if kern == 0 else -(-M // 64) * -(-N // 64))
        for target in (256, 512, 1024, 2048, 4096):
            out.add((kern, bm, bn, min(most, max(1, -(-target // tiles)))))
    return sorted(out)


def tune_dot(code, batch, M, N, G, seed=0) -> Optional[dict]:
    dt = _dtype(code)
    A = _rand((batch, M, G), dt, seed)
    B = _rand((batch, N, G), dt, seed + 1)
    out = torch.empty((batch, M, N), dtype=torch.complex128, device="cuda")
    key = dot_key(code, batch, M, N, G)
    run = lambda: gemm.dot128(A, B, out=out)  # noqa: E731
    base = _timed(run)
    best, best_t = None, base
    for cand in dot_candidates(batch, M, N, G):
        _set(TUNE_DOT, key, cand)
        t = _timed(run)
        if t < best_t:
            best, best_t = cand, t
    clear_key = best is None or best_t > base * (1 - MARGIN)
    entry = {"op": TUNE_DOT, "key": list(key), "choice": None if clear_key else list(best),
             "default_ms": base, "tuned_ms": base if clear_key else best_t,
             "shape": [batch, M, N, G]}
    _set(TUNE_DOT, key, best if not clear_key else (2, 0, 0, 0))  # kernel 2: not a choice
    return entry


# --------------------------------------------------------------- qmx_gemm
def rotate_key(code, mcode, variant, m, K, G) -> tuple:
    return (code, mcode, variant, m, K, int(_lib.lib().qmx_tune_bucket(G)))


def tune_rotate(code, mcode, variant, batch, m, K, G, seed=0) -> Optional[dict]:
    dt = _dtype(code)
    M = _rand((batch, K, m), _dtype(mcode), seed)
    C = _rand((batch, K, G), dt, seed + 1)
    out = torch.empty((batch, m, G), dtype=dt, device="cuda")
    key = rotate_key(code, mcode, variant, m, K, G)
    old = gemm.bf16x3_enabled()
    gemm.set_bf16x3(bool(variant))
    try:
        run = lambda: gemm.rotate(M, C, out=out)  # noqa: E731
        base = _timed(run)
        best, best_t = None, base
        mt_max = 6 if variant else (5 if code == 1 else 8)
        for mt in range(1, mt_max + 1):
            _set(TUNE_ROTATE, key, (mt,))
            t = _timed(run)
            if t < best_t:
                best, best_t = mt, t
    finally:
        gemm.set_bf16x3(old)
    keep = best is not None and best_t <= base * (1 - MARGIN)
    _set(TUNE_ROTATE, key, (best,) if keep else (0,))  # 0: not a choice
    return {"op": TUNE_ROTATE, "key": list(key), "choice": [best] if keep else None,
            "default_ms": base, "tuned_ms": best_t if keep else base,
            "shape": [batch, m, K, G]}


# -------------------------------------------------------------- plane pass
def tune_plane(code, shape, n_bands=32, seed=0) -> Optional[dict]:
    from . import systems
    from .basis import Basis
    from .wavefunction import Wavefunction

    spec = None
    for i in range(400):
        s = systems.make_cfg2_cell(i, cfg=3)
        if tuple(int(x) for x in s.shape) == tuple(shape):
            spec = s
            break
    if spec is None:
        return None
    basis = Basis.from_spec(spec)
    if basis.local_path(_dtype(code)) != "plane":
        return None
    dt = _dtype(code)
    coeff = _rand((1, basis.nk, n_bands, 1, basis.n_tot), dt, seed)
    C = Wavefunction(basis, coeff=coeff)
    C.constrain()
    V = torch.as_tensor(spec.V, device="cuda")
    key = (code,) + tuple(int(x) for x in shape)
    run = lambda: basis.apply_potential(V, C)  # noqa: E731
    base = _timed(run)
    best, best_t = None, base
    for choice in (0, 1, 2):
        _set(TUNE_PLANE, key, (choice,))
        t = _timed(run)
        if t < best_t:
            best, best_t = choice, t
    keep = best is not None and best_t <= base * (1 - MARGIN)
    _set(TUNE_PLANE, key, (best,) if keep else (0,))
    return {"op": TUNE_PLANE, "key": list(key), "choice": [best] if keep else None,
            "default_ms": base, "tuned_ms": best_t if keep else base, "shape": list(shape)}


# ------------------------------------------------------------------ driver
def tune(shapes: dict, progress=None) -> list:
    """Tune every distinct key of the recorded shapes (the most frequent
    shape of a key stands for it); returns the entries, including those
    whose default stays (choice None)."""
    groups: dict = {}
    for sh, count in shapes.items():
        if sh[0] == "dot":
            _, code, batch, M, N, G = sh
            k = ("dot",) + dot_key(code, batch, M, N, G)
        elif sh[0] == "rotate":
            _, code, mcode, variant, batch, m, K, G = sh
            k = ("rotate",) + rotate_key(code, mcode, variant, m, K, G)
        else:
            k = sh
        cur = groups.get(k)
        if cur is None or count > cur[1]:
            groups[k] = (sh, count)
    entries = []
    for i, (k, (sh, _)) in enumerate(sorted(groups.items(), key=lambda kv: str(kv[0]))):
        if sh[0] == "dot":
            e = tune_dot(*sh[1:], seed=i)
        elif sh[0] == "rotate":
            e = tune_rotate(*sh[1:], seed=i)
        else:
            e = tune_plane(sh[1], sh[2:], seed=i)
        if e is not None:
            entries.append(e)
            if progress:
                progress(f"{k}: default {e['default_ms']:.3f} ms, tuned {e['tuned_ms']:.3f} ms"
                         f" {e['choice']}")
    return entries


def apply(entries: list) -> int:
    """Set the entries that carry a choice; returns how many."""
    n = 0
    for e in entries:
        if e.get("choice"):
            _set(int(e["op"]), e["key"], e["choice"])
            n += 1
    return n


def save(path, entries: list) -> None:
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    data = {"device": device_name(), "entries": entries}
    path.write_text(json.dumps(data, indent=1))


def load(path) -> int:
    """Apply a cache file written for this device (entries of another device
    are ignored); returns the number of choices set."""
    path = Path(path)
    if not path.exists():
        return 0
    data = json.loads(path.read_text())
    if data.get("device") != device_name():
        return 0
    return apply(data.get("entries", []))


def default_path() -> Path:
    return TUNING_DIR / f"{device_name()}.json"


def load_default() -> int:
    """The shipped cache of this device, once per process (0 if none)."""
    global _loaded_default
    if _loaded_default:
        return int(_lib.lib().qmx_tune_count())
    _loaded_default = True
    if os.environ.get("QMX_NO_TUNING"):
        return 0
    return load(default_path())


def main(argv=None) -> None:
    import argparse

    from . import systems
    from .batched import BatchSolver, CellJob

    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--cells", type=int, default=96, help="cells of the served distribution")
    ap.add_argument("--out", default=None, help="cache file (default: the shipped one)")
    ap.add_argument("--band-classes", default=None,
                    help="comma-separated band classes of the solver (default: its own)")
    args = ap.parse_args(argv)
    classes = ({"band_classes": tuple(int(c) for c in args.band_classes.split(",") if c)}
               if args.band_classes else {})
    rc.init()
    rc.require_gpu()
    global _loaded_default
    _loaded_default = True  # measure against the library's own choices
    clear()
    t0 = time.perf_counter()
    specs = [systems.make_cfg2_cell(i, cfg=3) for i in range(args.cells)]
    with record() as shapes:
        for dt in (torch.complex64, torch.complex128):
            solver = BatchSolver(dt, max_iterations=60, **classes)
            solver.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
            solver.run()
            del solver
    print(f"[autotune] {len(shapes)} shapes recorded in {time.perf_counter() - t0:.1f} s",
          flush=True)
    entries = tune(shapes, progress=lambda m: print(f"[autotune] {m}", flush=True))
    path = Path(args.out) if args.out else default_path()
    save(path, entries)
    tuned = sum(1 for e in entries if e["choice"])
    print(f"[autotune] {tuned} of {len(entries)} keys tuned, written to {path}", flush=True)


if __name__ == "__main__":
    main(sys.argv[1:])
