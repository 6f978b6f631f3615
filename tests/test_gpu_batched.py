"""Batched cell solver on the GPU against the FP64 dense oracle.

Several small cells of different boxes, band counts and k-meshes run
together; every cell's eigenvalues must match LAPACK on its dense H within
1e-8 Ha (plain, complex128) or 1e-6 Ha (mixed, complex64 + FP64
Rayleigh-Ritz), and a cell's result must not depend on the company it runs
in (alone vs in the mix, same tolerances), on a budget that defers it, or
on other requests cancelled while queued or running."""
import numpy as np
import pytest
import torch

from oracle.basis import Basis as OBasis
from oracle.dense import dense_eigenvalues
from qimpy_amd import systems
from qimpy_amd.batched import BatchSolver, CellJob

pytestmark = pytest.mark.gpu

C64, C128 = torch.complex64, torch.complex128
TOL = {C128: 1e-8, C64: 1e-6}


def _specs():
    return [
        systems.small_cell(n_bands=4, seed=11),
        systems.small_cell(n_bands=6, kmesh=(1, 1, 1), seed=12),
        systems.small_cell(shape=(10, 12, 15), n_bands=5, kmesh=(2, 1, 1), seed=13),
        systems.small_cell(n_bands=9, seed=14),
        systems.small_cell(shape=(16, 16, 16), n_bands=8, kmesh=(1, 2, 2), seed=15),
    ]


@pytest.fixture(scope="module")
def dense_refs():
    return [dense_eigenvalues(OBasis(s), s.n_bands) for s in _specs()]


def _solver(dtype, **kw):
    args = dict(residual_threshold=3e-4, eig_threshold=1e-10, max_iterations=150)
    args.update(kw)
    return BatchSolver(dtype, **args)


@pytest.mark.parametrize("dtype", [C128, C64], ids=["plain", "mixed"])
def test_group_streams_match_one_stream(gpu, dtype):
    """Two bucket groups on streams of their own give the results of the
    same groups on one stream: the same launches per group in the same
    order, only overlapped across groups (eigenvalues and iteration counts
    equal, bit for bit)."""
    specs = _specs()
    runs = []
    for streams in (True, False):
        solver = _solver(dtype, groups=2, streams=streams)
        solver.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
        runs.append(solver.run())
    a, b = runs
    assert sorted(a) == sorted(b) == list(range(len(specs)))
    for i in a:
        print(f"[group streams {dtype}] cell {i}: {a[i].iterations} / {b[i].iterations} iterations",
              flush=True)
        assert a[i].iterations == b[i].iterations
        assert np.array_equal(a[i].eig, b[i].eig)


@pytest.mark.parametrize("groups", [1, 2])
@pytest.mark.parametrize("dtype", [C128, C64], ids=["plain", "mixed"])
def test_batched_cells_match_dense(gpu, dense_refs, dtype, groups):
    specs = _specs()
    solver = _solver(dtype, groups=groups)
    solver.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
    results = solver.run()
    assert sorted(results) == list(range(len(specs)))
    for i, ref in enumerate(dense_refs):
        r = results[i]
        err = np.abs(r.eig - ref).max()
        print(f"[batched {dtype}] cell {i}: {r.iterations} iterations, max|eig - dense| = {err:.2e}",
              flush=True)
        assert r.converged
        assert err < TOL[dtype]
        assert abs(r.Eband - 2.0 * (specs[i].wk[:, None] * ref).sum()) < 2 * TOL[dtype] * ref.size


@pytest.mark.parametrize("dtype", [C128, C64], ids=["plain", "mixed"])
def test_alone_equals_mixed(gpu, dtype):
    specs = _specs()
    mix = _solver(dtype)
    mix.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
    in_mix = mix.run()
    for i in (0, 3):
        alone = _solver(dtype)
        alone.submit([CellJob(specs[i], job_id=i)])
        r = alone.run()[i]
        d = np.abs(r.eig - in_mix[i].eig).max()
        print(f"[alone vs mix {dtype}] cell {i}: {d:.2e}", flush=True)
        assert d < TOL[dtype]


def test_jobs_join_running_bucket(gpu, dense_refs):
    """Jobs arriving while others iterate join the same bucket between
    iterations and still converge to the dense answer."""
    specs = _specs()
    solver = _solver(C64)
    solver.submit([CellJob(s, job_id=i, arrival=float(i)) for i, s in enumerate(specs)])
    now, sizes = 0.0, []
    while not solver.idle:
        solver.tick(now=now)
        sizes.append(sum(b.Kb for b in solver.buckets.values()))
        now += 0.5
    assert max(sizes) > min(s for s in sizes)  # membership changed while running
    for i, ref in enumerate(dense_refs):
        assert np.abs(solver.results[i].eig - ref).max() < TOL[C64]


def test_bucket_groups_served_in_turn(gpu, dense_refs):
    """Buckets dealt to two groups that tick in turn (each group's step
    ends with the copies its next tick reads): every result still matches
    the dense answer, and jobs keep one Davidson step per visit."""
    specs = _specs()
    solver = _solver(C64, groups=2)
    solver.submit([CellJob(s, job_id=i, arrival=0.01 * i) for i, s in enumerate(specs)])
    now = 0.0
    while not solver.idle:
        solver.tick(now=now)
        now += 0.01
    assert len(set(solver.sched.group_of.values())) == 2  # both groups hold buckets
    for i, ref in enumerate(dense_refs):
        r = solver.results[i]
        assert r.converged and np.abs(r.eig - ref).max() < TOL[C64]


def test_served_stream_matches_dense_and_alone(gpu, dense_refs):
    """Requests arriving over time (continuous batching): every result equals
    the dense answer and its own result served alone, within tolerance."""
    from qimpy_amd.serve import serve
    specs = _specs()
    arrivals = [0.0, 0.01, 0.02, 0.05, 0.08]
    rep = serve(specs, arrivals, _solver(C64))
    assert rep.n_requests == len(specs) and rep.p99_s >= rep.p50_s > 0
    for i, ref in enumerate(dense_refs):
        assert np.abs(rep.results[i].eig - ref).max() < TOL[C64]
    alone = serve([specs[2]], [0.0], _solver(C64)).results[0]
    assert np.abs(alone.eig - rep.results[2].eig).max() < TOL[C64]


@pytest.mark.parametrize("dtype", [C128, C64], ids=["plain", "mixed"])
def test_budget_and_cancellation(gpu, dense_refs, dtype):
    """A budget that holds about two of the cells (the rest deferred until
    others retire), one request cancelled while queued and one while
    running: the cancelled ones give no result and every other cell's
    eigenvalues are those of an unconstrained run (to the tolerance) and of
    the dense reference."""
    specs = _specs()
    free_run = _solver(dtype)
    free_run.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
    ref_res = free_run.run()
    del free_run
    torch.cuda.empty_cache()
    solver = _solver(dtype)
    jobs = [CellJob(s, job_id=i) for i, s in enumerate(specs)]
    one = max(j.hbm_bytes(solver.band_class(j), dtype, j.n_tot) for j in jobs)
    solver.sched.hbm_budget = torch.cuda.memory_allocated() + 1.8 * one
    solver.submit(jobs)
    print(f"[budget/cancel {dtype}] budget {solver.sched.hbm_budget / 1e9:.3f} GB, largest job"
          f" {one / 1e6:.1f} MB", flush=True)
    solver.tick()
    queued = [j.job_id for j in solver.queue]
    assert queued, "the budget should defer some cells"
    assert solver.cancel(queued[-1])
    running = [j.job_id for b in solver.buckets.values() for j in b.jobs]
    assert solver.cancel(running[0])
    assert not solver.cancel(999)
    while not solver.idle:
        solver.tick()
    torch.cuda.synchronize()
    gone = {queued[-1], running[0]}
    assert solver.cancelled == gone
    assert set(solver.results) == set(range(len(specs))) - gone
    for i, r in solver.results.items():
        err = np.abs(r.eig - ref_res[i].eig).max()
        print(f"  cell {i}: |eig - unconstrained| {err:.2e}", flush=True)
        assert err < TOL[dtype]
        assert np.abs(r.eig - dense_refs[i]).max() < TOL[dtype]
    assert solver.hbm_in_use == 0


@pytest.mark.parametrize("seed,n_atoms", [(21, 3), (22, 1), (23, 5)])
def test_projectors_kernel_match_oracle(gpu, seed, n_atoms):
    """qmx_gaussian_projectors (build_projectors on device tensors) against
    the oracle's projectors, which pin the torch formula on the CPU."""
    from oracle.hamiltonian import projectors as o_projectors
    from qimpy_amd.ions import build_projectors

    spec = systems.small_cell(seed=seed, n_atoms=n_atoms)
    ob = OBasis(spec)
    dev = torch.device("cuda")
    got = build_projectors(torch.tensor(spec.positions, device=dev), spec.proj_rc, spec.proj_l,
                           torch.tensor(ob.iG, device=dev), torch.tensor(ob.k, device=dev),
                           spec.Gbasis, spec.volume, torch.tensor(ob.n, device=dev)).cpu().numpy()
    ref = o_projectors(ob)
    assert got.shape == ref.shape
    assert np.abs(got - ref).max() < 1e-14 * max(1.0, np.abs(ref).max())
