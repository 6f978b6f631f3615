"""Host-side checks of libqmx.so: it loads without a GPU, exports every entry
point include/qmx.h declares, and its sparse-FFT index tables (built by the
same C++ code a plan uploads) reproduce the FP64 oracle's local-potential
apply when the five passes are emulated in numpy on them."""
import re
from pathlib import Path

import numpy as np
import pytest

from oracle.basis import Basis
from oracle.hamiltonian import apply_potential
from qimpy_amd import _lib, systems

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "qmx.h").read_text()
    return sorted(set(re.findall(r"\b(qmx_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_header():
    lib = _lib.lib()
    names = declared_symbols()
    assert len(names) >= 12
    for name in names:
        assert hasattr(lib, name), name
        assert name in _lib.SIGNATURES, f"{name} has no ctypes signature"
    assert lib.qmx_version().decode().startswith("qmx")


def emulate_local(basis, V, C, T):
    """Numpy model of passes A-E on the plan tables."""
    K, B, G = C.shape
    nx, ny, nz = basis.shape
    N = nx * ny * nz
    out = np.zeros_like(C)
    for k in range(K):
        n = basis.n[k]
        t = T["t1idx"][k, :n]
        cs = T["cstart"][k]
        assert cs[-1] == n and np.all(np.diff(cs) >= 0)
        cols = np.repeat(np.arange(T["ncol"]), np.diff(cs))
        assert np.array_equal(t // nz, cols)  # each column's points are contiguous
        for b in range(B):
            T1 = np.zeros((T["ncol"], nz), complex)
            T1[t // nz, t % nz] = C[k, b, :n]
            T1 = np.fft.ifft(T1, axis=1) * nz
            T2 = np.zeros((T["nxa"], ny, nz), complex)
            for xa in range(T["nxa"]):
                c0, c1 = T["pcol_start"][xa], T["pcol_start"][xa + 1]
                T2[xa, T["col_iy"][c0:c1]] = T1[c0:c1]
            T2 = np.fft.ifft(T2, axis=1) * ny
            box = np.zeros((nx, ny, nz), complex)
            box[T["x_list"]] = T2
            box = np.fft.fft(np.fft.ifft(box, axis=0) * nx * V, axis=0)
            T2 = np.fft.fft(box[T["x_list"]], axis=1)
            for xa in range(T["nxa"]):
                c0, c1 = T["pcol_start"][xa], T["pcol_start"][xa + 1]
                T1[c0:c1] = T2[xa, T["col_iy"][c0:c1]]
            T1 = np.fft.fft(T1, axis=1) / N
            out[k, b, :n] = T1[t // nz, t % nz]
    return out


@pytest.mark.parametrize("which", ["small", "cfg0"])
def test_pass_model_matches_oracle(which):
    spec = systems.small_cell(n_bands=3) if which == "small" else systems.make_cell(0, n_bands=2, kmesh=(1, 1, 2))
    b = Basis(spec)
    T = _lib.geom_tables(spec.shape, b.n, b.fft_index)
    assert T["nxa"] <= spec.shape[0] and T["ncol"] <= spec.shape[0] * spec.shape[1]
    assert np.all(np.diff(T["x_list"]) > 0)
    rng = np.random.default_rng(0)
    C = rng.standard_normal((len(b.k), 2, b.n_tot)) + 1j * rng.standard_normal((len(b.k), 2, b.n_tot))
    b.constrain(C)
    ref = apply_potential(b, spec.V, C)
    assert np.abs(emulate_local(b, spec.V, C, T) - ref).max() < 1e-13


def test_workload_sparsity():
    """Active planes and columns of the measured boxes (cfg0 36^3, cfg1 96^3)."""
    b0 = Basis(systems.make_cell(0))
    T0 = _lib.geom_tables(b0.shape, b0.n, b0.fft_index)
    b1 = Basis(systems.make_cell(1))
    T1 = _lib.geom_tables(b1.shape, b1.n, b1.fft_index)
    assert T0["nxa"] < 36 // 2 + 2 and T1["nxa"] < 96 // 2
    assert T1["ncol"] < 0.2 * 96 * 96


def test_geom_errors():
    fi = np.array([[5, 3, 7]], dtype=np.int64)  # not increasing
    with pytest.raises(RuntimeError, match="increasing"):
        _lib.geom_tables((4, 4, 4), np.array([3]), fi)
    with pytest.raises(RuntimeError, match="2, 3 and 5"):
        _lib.geom_tables((7, 4, 4), np.array([1]), np.array([[0]], dtype=np.int64))
    with pytest.raises(RuntimeError, match="n\\[k\\]"):
        _lib.geom_tables((4, 4, 4), np.array([5]), np.array([[0, 1, 2]], dtype=np.int64))


@pytest.mark.parametrize("n,n_want", [(1, 1), (2, 1), (37, 37), (144, 72), (144, 144), (256, 3)])
def test_heevx_work_contract(n, n_want):
    """Work area of the bisection eigensolver per matrix: the complex working
    copy (16 n^2), four n x n_want real inverse-iteration arrays and d, e, tau
    (32 n), 256-byte aligned; n_want <= 0 or > n means all n; n outside
    1..256 is refused (-1)."""
    lib = _lib.lib()
    w = int(lib.qmx_heevx_work_bytes(n, n_want))
    assert w >= 16 * n * n + 32 * n * n_want + 32 * n
    assert w % 256 == 0
    assert int(lib.qmx_heevx_work_bytes(n, 0)) == int(lib.qmx_heevx_work_bytes(n, n))
    assert int(lib.qmx_heevx_work_bytes(n, n + 5)) == int(lib.qmx_heevx_work_bytes(n, n))
    assert int(lib.qmx_heevx_work_bytes(0, 1)) == -1
    assert int(lib.qmx_heevx_work_bytes(257, 1)) == -1


@pytest.mark.parametrize("n", [1, 2, 37, 127, 128, 129, 143, 144, 200, 256])
def test_heev_small_workspace_contract(n):
    """Workspace of the batched QL eigensolver: per matrix at least the complex
    working copy (16 n^2 bytes) plus, above n = 128, the real QL accumulator
    (8 n^2 bytes); linear in the batch, 256-byte aligned per matrix (the
    kernel offsets matrix b by b times the one-matrix size)."""
    lib = _lib.lib()
    one = int(lib.qmx_heev_small_workspace_bytes(n, 1))
    assert one >= 16 * n * n + (8 * n * n if n > 128 else 0)
    assert one % 256 == 0
    for batch in (0, 3, 100):
        assert int(lib.qmx_heev_small_workspace_bytes(n, batch)) == batch * one


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus 2 exits with status 2 and says why,
    before touching a device."""
    import subprocess
    import sys
    from pathlib import Path

    root = Path(__file__).resolve().parents[1]
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_geom_tables_many_cells():
    """The index tables of a plan over the k-points of many cells of one box
    (large enough that the build splits the k-points over threads), given
    as per-k row pointers or as one padded matrix: identical, and each
    sphere point's column and z reproduce its box index."""
    from qimpy_amd.batched import CellJob

    jobs = []
    i = 0
    while len(jobs) < 12:
        spec = systems.make_cfg2_cell(i, cfg=3)
        i += 1
        if tuple(spec.shape) == (48, 48, 48):
            jobs.append(CellJob(spec, job_id=i))
    shape = jobs[0].shape
    Gb = max(j.n_tot for j in jobs)
    n = np.concatenate([j.n for j in jobs])
    fi = np.zeros((len(n), Gb), dtype=np.int64)
    a = 0
    for j in jobs:
        fi[a : a + j.K, : j.n_tot] = j.fft_index
        a += j.K
    assert len(n) * Gb > 2 << 18  # several threads
    T = _lib.geom_tables(shape, n, fi)
    g, keep = _lib.make_geom_rows(shape, n, [j.fft_index for j in jobs], Gb)
    sizes = np.zeros(2, dtype=np.int64)
    out = {k: np.zeros_like(T[k]) for k in ("x_list", "pcol_start", "col_iy", "cstart", "t1idx")}
    p = lambda x: x.ctypes.data_as(_lib.c_i32p)
    _lib.call("qmx_geom_tables", _lib.ct.byref(g), sizes.ctypes.data_as(_lib.c_i64p),
              p(out["x_list"]), p(out["pcol_start"]), p(out["col_iy"]), p(out["cstart"]),
              p(out["t1idx"]))
    assert (sizes[0], sizes[1]) == (T["nxa"], T["ncol"])
    for k in out:
        assert np.array_equal(out[k], T[k]), k
    nx, ny, nz = shape
    col_ix = np.repeat(T["x_list"], np.diff(T["pcol_start"]))
    for k in range(len(n)):
        t = T["t1idx"][k]
        col, z = t[: n[k]] // nz, t[: n[k]] % nz
        assert np.array_equal((col_ix[col] * ny + T["col_iy"][col]) * nz + z, fi[k, : n[k]])
        assert np.all(t[n[k] :] == 0)
        assert np.array_equal(np.diff(T["cstart"][k]), np.bincount(col, minlength=T["ncol"]))
