"""Build libqmx.so (the gfx950 HIP kernels behind include/qmx.h) in-tree.

Run as ``python -m qimpy_amd.build``. hipcc cross-compiles for gfx950, so
this works on a host without a GPU; the library lands in
``qimpy_amd/_lib/libqmx.so`` next to the sources it is built from.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
LIB = LIBDIR / "libqmx.so"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["qmx_local.hip", "qmx_ops.hip", "qmx_eig.hip", "qmx_heevx.hip", "qmx_subspace.hip", "qmx_gemm.hip",
           "qmx_projectors.hip", "qmx_rotate.hip", "qmx_plane_fixed.hip", "qmx_cols_fixed.hip",
           "qmx_tune.hip"]

FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-result",
    # Keep packed-FP32 instructions out of the device code:
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


# Per-source additions. The rotations keep their MFMA accumulators in
# ordinary VGPRs: left to its heuristics the compiler parks the FP64 ones in
# AGPRs and copies them back and forth every k slab (and drains the MFMA
# pipeline to do so), and the complex64 kernel needs one more register than
# four waves per SIMD allow.
EXTRA_FLAGS = {"qmx_rotate.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}


def build(verbose: bool = False, jobs: int | None = None) -> Path:
    """Compile every source (in parallel, one hipcc per file) and link."""
    from concurrent.futures import ThreadPoolExecutor

    LIBDIR.mkdir(exist_ok=True)
    cmds, objs = [], []
    for src in SOURCES:
        path = CSRC / src
        if not path.exists():
            continue
        obj = LIBDIR / (path.stem + ".o")
        cmds.append([HIPCC, *FLAGS, *EXTRA_FLAGS.get(src, []), "-c", str(path), "-o", str(obj)])
        objs.append(str(obj))
    if verbose:
        for cmd in cmds:
            print(" ".join(cmd), flush=True)
    jobs = jobs or min(len(cmds), max(1, min(8, os.cpu_count() or 1)))
    with ThreadPoolExecutor(max_workers=jobs) as pool:
        done = list(pool.map(lambda c: subprocess.run(c, capture_output=not verbose), cmds))
    for cmd, r in zip(cmds, done):
        if r.returncode != 0:
            if r.stderr:
                sys.stderr.write(r.stderr.decode(errors="replace"))
            raise subprocess.CalledProcessError(r.returncode, cmd)
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", *objs, "-o", str(tmp)]
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    for o in objs:
        os.remove(o)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-q" not in sys.argv))
