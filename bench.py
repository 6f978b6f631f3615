"""Band-solve benchmark on one MI355X: served band solves/s with p50/p99
latency, mixed precision against plain FP64, every result within 1e-6 Ha.

Headline workload. One continuous served stream per precision:
(warmup + steps) x window band-solve requests arriving as a seeded Poisson
process at 200 req/s, each a cell of the uneven distribution of the batched
workload (cubic boxes 24..60^3, 8-64 bands, 1-16 k-points, 2-16 atoms with 8
projectors each, 20 Ha). Every request's inputs (sphere tables, potential
box, k-points, positions, couplings) are staged in HBM before the stream
starts; a request enters the server at its arrival time, which admits it
against a 180 GB HBM budget and iterates every resident request in lockstep
batches (continuous batching) until its FP64 result is ready. A step is a
window of `window` consecutive completed requests (in completion order):
the first `warmup` windows are untimed, the next `steps` are timed, from
the completion of the last warm-up request to that of the last timed one
(device event times). value = timed requests / timed wall time; latency =
result-ready time - arrival time over the timed requests. At 200 req/s the
server is saturated (arrivals outpace completions), so this throughput is
the saturation rate and the latency is queueing-dominated; a second mixed
stream at 80% of the measured saturation rate gives the latency of a stable
queue.

The mixed form iterates in complex64 and stops each request when the
Kato-Temple estimate of every wanted band is below half of 1e-6 Ha, then
finishes with an FP64 Rayleigh-Ritz; each result carries the estimate of
that final FP64 step. The plain form iterates in complex128 to the
reference's |deig| < 1e-8. The line reports, over every request of the
stream, the largest mixed-vs-plain eigenvalue difference and the largest
error estimate.

Side workloads in the same line: the 2-atom diamond cell (36^3, 8 k, 16
bands) through the drop-in Davidson and the batched engine, the 64-atom cell
(96^3, 256 bands), 64 uneven cells in one batch, the 256-atom CheFSI solve
(144^3, 1024 bands), the electron density and the projectors; the roofline of
the dominant bandwidth kernel from live HIP-event timing of every library
launch on its stream; the FP64 CPU oracle on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
#: Dense MFMA peaks of MI355X (TFLOP/s): FP32 for complex64 data, FP64 for complex128.
MFMA_PEAK_TFLOPS = {"c64": 157.3, "c128": 78.6}
RATE = 200.0  # requests per second
EIG_TOLERANCE = 1e-6  # mixed form: per-band eigenvalue error target (Ha)
PLAIN_THRESHOLD = 1e-8  # plain form: the reference's |deig| stop
#: Band classes of the batched solver (jobs pad to the smallest that holds them).
BAND_CLASSES = (32, 48, 72)
#: Bucket groups the served solver deals its buckets to (served in turn).
GROUPS = 2
CHEFSI_RESIDUAL = 3e-5
#: Library kernels that are bandwidth (not latency) bound; the roofline
#: reports the one with the largest total time.
BANDWIDTH_KERNELS = {f"{k}<{t}>" for t in ("c64", "c128") for k in (
    "k_sphere_to_cols", "k_cols_to_planes", "k_xpass_potential", "k_planes_to_cols",
    "k_cols_to_sphere", "k_plane_potential", "k_gemm", "k_rotate", "k_rotate_bf16x3")}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed windows of completed requests")
    ap.add_argument("--warmup", type=int, default=5, help="untimed windows before them")
    ap.add_argument("--window", type=int, default=100, help="requests per window")
    ap.add_argument("--rate", type=float, default=RATE, help="arrival rate of the headline stream")
    ap.add_argument("--stable-requests", type=int, default=2200,
                    help="requests of the stable-rate latency stream (first tenth untimed); 0 skips it")
    ap.add_argument("--host-inputs", action="store_true",
                    help="requests enter as host arrays prepared during the stream (not staged in HBM)")
    ap.add_argument("--only", choices=("all", "mixed", "plain"), default="all",
                    help="headline streams to run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--bf16x3", action="store_true",
                    help="complex64 rotations as BF16 x 3 split-precision products")
    ap.add_argument("--no-extra", action="store_true", help="skip the side workloads")
    ap.add_argument("--band-classes", default=",".join(map(str, BAND_CLASSES)),
                    help="comma-separated band classes of the batched solver")
    ap.add_argument("--groups", type=int, default=GROUPS,
                    help="bucket groups of the served solver (served in turn)")
    ap.add_argument("--one-stream", action="store_true",
                    help="serve the headline stream with both bucket groups on one HIP stream "
                         "and take the roofline from it (no separate roofline pass)")
    ap.add_argument("--side-only", action="store_true",
                    help="only the density and projector measurements (no served streams)")
    return ap.parse_args()


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(specs, budget_s: float = 20.0) -> dict:
    """The FP64 oracle (numpy/scipy, this host's physical cores) solving the
    first requests of the stream in order, whole band solves to the plain
    form's |deig| < 1e-8, until about `budget_s` of run time: a request
    whose predicted time (the measured seconds per k-point x plane wave x
    band of the requests before it) would pass 1.5 budget_s ends the sample;
    then the 2-atom diamond cell (8 k, 16 bands, 20 iterations)."""
    import psutil
    from threadpoolctl import threadpool_limits

    from oracle.basis import Basis as OBasis
    from oracle.davidson import davidson, initial_state
    from oracle.hamiltonian import Hamiltonian
    from qimpy_amd import systems

    cores = psutil.cpu_count(logical=False) or os.cpu_count() or 1
    t_run, names, iters, work = 0.0, [], [], 0.0
    with threadpool_limits(limits=cores):
        for spec in specs:
            size = len(spec.k) * spec.n_bands_max * np.pi * spec.shape[0] ** 3 / 48
            if names and t_run + t_run / work * size > 1.5 * budget_s:
                break
            work += size
            ob = OBasis(spec)
            H = Hamiltonian(ob)
            C = initial_state(ob, spec.n_bands)
            t0 = time.perf_counter()
            r = davidson(ob, H, C, spec.n_bands, spec.n_bands_extra, 200, PLAIN_THRESHOLD)
            t_run += time.perf_counter() - t0
            names.append(f"{spec.shape[0]}^3/{len(spec.k)}k/{spec.n_bands}b")
            iters.append(r.n_iterations)
            if t_run > budget_s:
                break
        spec0 = systems.make_cell(0)
        ob = OBasis(spec0)
        H = Hamiltonian(ob)
        C = initial_state(ob, spec0.n_bands)
        t0 = time.perf_counter()
        davidson(ob, H, C, spec0.n_bands, spec0.n_bands_extra, spec0.n_iterations,
                 spec0.eig_threshold)
        t0_cfg0 = time.perf_counter() - t0
    return {
        "value": round(len(names) / t_run, 4), "unit": "band solves/s", "cores": int(cores),
        "kind": "port",
        "sample": (f"FP64 numpy/scipy oracle, whole band solves (|deig| < 1e-8) of the first "
                   f"{len(names)} requests of the stream ({', '.join(names)}; {sum(iters)} "
                   f"iterations in {t_run:.1f} s) on {cores} threads of {cpu_model()}"),
        "diamond_2_atom_s_per_solve": round(t0_cfg0, 3),
    }


def windows(rep, arrivals, warmup: int, steps: int, window: int) -> dict:
    """Throughput and latency of the timed windows of a served stream."""
    res = sorted(rep.results.values(), key=lambda r: r.finished)
    need = (warmup + steps) * window
    if len(res) < need:
        raise RuntimeError(f"stream completed {len(res)} requests, {need} needed")
    t_start = res[warmup * window - 1].finished if warmup else float(np.min(arrivals))
    t_end = res[need - 1].finished
    timed = res[warmup * window:need]
    lat = np.array([r.latency for r in timed])
    step_s = []
    prev = t_start
    for w in range(warmup, warmup + steps):
        t = res[(w + 1) * window - 1].finished
        step_s.append(t - prev)
        prev = t
    wall = t_end - t_start
    return {"value": steps * window / wall, "wall_s": wall, "ms_per_step": 1e3 * wall / steps,
            "step_ms_min_max": [round(1e3 * min(step_s), 1), round(1e3 * max(step_s), 1)],
            "p50": float(np.percentile(lat, 50)), "p99": float(np.percentile(lat, 99)),
            "n_latency": int(lat.size),
            "mean_iterations": float(np.mean([r.iterations for r in timed]))}


def served(specs, arrivals, dtype, staged=None, profile=False, streams=True):
    """One served stream; `staged` (jobs from stage_requests) holds the
    requests' inputs in HBM, None makes them enter as host arrays. streams
    False puts both bucket groups on one HIP stream (no two kernels of the
    solver overlap, so per-launch times are each kernel's own)."""
    import torch

    from qimpy_amd import _lib
    from qimpy_amd.batched import BatchSolver
    from qimpy_amd.serve import serve

    jobs = None if staged is None else [j.fresh(a) for j, a in zip(staged, arrivals)]
    solver = BatchSolver(dtype, eig_tolerance=EIG_TOLERANCE, eig_threshold=PLAIN_THRESHOLD,
                         max_iterations=200, hbm_budget=180e9, band_classes=BAND_CLASSES,
                         groups=GROUPS, streams=streams)
    if profile:
        _lib.profile_enable(True)
        _lib.profile_read()
    torch.cuda.synchronize()
    rep = serve(specs, arrivals, solver, jobs=jobs)
    prof = None
    if profile:
        _lib.profile_enable(False)
        prof = _lib.profile_read()
    del jobs
    torch.cuda.empty_cache()
    return rep, prof


def batched(specs, dtype):
    import torch

    from qimpy_amd.batched import BatchSolver, CellJob

    solver = BatchSolver(dtype, eig_tolerance=EIG_TOLERANCE, eig_threshold=PLAIN_THRESHOLD,
                         max_iterations=200, band_classes=BAND_CLASSES)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
    res = solver.run()
    return len(res) / (time.perf_counter() - t0), res


def single(spec, dtype, repeats=1, **kw):
    """Drop-in solve (Electrons + its diagonalizer): one untimed solve, then
    the median of `repeats` timed ones from the same start."""
    import torch

    from qimpy_amd.electrons import electrons_from_spec

    el = electrons_from_spec(spec, dtype=dtype, **kw)
    el.initialize_wavefunctions(seed=0)
    el.solve()
    times, rep = [], None
    for _ in range(repeats):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        el.initialize_wavefunctions(seed=0)
        rep = el.solve()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
    return float(np.median(times)), rep


def diamond_measurement(spec) -> dict:
    """The 2-atom diamond cell (36^3, 8 k, 16 bands): the drop-in Davidson in
    FP64 (the reference's algorithm, 20 iterations to |deig| < 1e-8), and the
    batched engine on the same cell in both forms."""
    import torch

    C64, C128 = torch.complex64, torch.complex128
    t_plain, rp = single(spec, C128, repeats=3, n_iterations=spec.n_iterations,
                         eig_threshold=spec.eig_threshold)
    vm, rm = batched([spec], C64)
    vb, rb = batched([spec], C128)
    return {"workload": "2-atom diamond cell a=10.26 bohr, 36^3 box, 8 k (2x2x2 MP), 16 bands (+2), "
                        "16 projectors, 20 Ha",
            "drop_in_davidson_plain_fp64": {
                "s_per_solve": round(t_plain, 4), "iterations": rp.n_iterations,
                "deig_max_after": float(rp.deig_max),
                "note": "the reference's configuration: 20 iterations, |deig| < 1e-8"},
            "batched_engine": {"mixed": round(vm, 3), "plain_fp64": round(vb, 3),
                               "unit": "band solves/s",
                               "mixed_eig_err_vs_fp64_Ha": float(np.abs(rm[0].eig - rb[0].eig).max()),
                               "mixed_err_bound_Ha": rm[0].err_bound,
                               "iterations": [rm[0].iterations, rb[0].iterations]}}


def density_measurement(spec) -> dict:
    """Basis.collect_density (qmx_collect_density) on the 64-atom cell: 282
    normalised random bands, complex64 and complex128, HIP events around
    ten calls. Algorithmic bytes per band follow the dense-transform
    convention of the local apply: a sphere read b G and one box transform
    read and write 2 b N, plus the FP64 density once per call (16 N)."""
    import torch

    from qimpy_amd.electrons import electrons_from_spec

    el = electrons_from_spec(spec)
    b = el.basis
    K, G, N = b.nk, b.n_tot, b.grid.N
    B = int(spec.n_bands + spec.n_bands_extra)
    g = torch.Generator(device="cuda").manual_seed(5)
    res = {"workload": f"64-atom cell, {b.grid.shape[0]}^3 box, {K} k, {B} bands, fillings 1",
           "unit": "bands/s"}
    for dt, name in ((torch.complex64, "c64"), (torch.complex128, "c128")):
        C = torch.randn((1, K, B, 1, G), dtype=dt, device="cuda", generator=g)
        C /= torch.linalg.vector_norm(C, dim=-1, keepdim=True)
        f = torch.ones((1, K, B), dtype=torch.float64)
        rho = b.collect_density(C, f)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(10):
            rho = b.collect_density(C, f)
        e1.record()
        torch.cuda.synchronize()
        t = e0.elapsed_time(e1) / 10 * 1e-3
        cs = C.element_size()
        by = K * B * (cs * G + 2 * cs * N) + 16 * N
        charge = float(rho[0].sum()) * b.volume / N
        res[name] = {"ms_per_call": round(t * 1e3, 3), "bands_per_s": round(K * B / t, 1),
                     "GBps_dense_convention": round(by / t / 1e9, 1),
                     "frac": round(by / t / 1e9 / HBM_PEAK_GBS, 4),
                     "charge_minus_sum_fw": charge - float((f[0] * b.w_sk[0].cpu()).sum())}
    return res


def projectors_measurement(spec) -> dict:
    """qmx_projectors on the 64-atom cell: the cell's Gaussian radials
    sampled at spacing 0.02 as spline radial functions, ten calls into the
    same output, timed by the library's HIP events around each launch.
    Algorithmic bytes: the complex128 output, 16 bytes per (k, atom,
    projector, plane wave) entry."""
    import torch

    from qimpy_amd import _lib
    from qimpy_amd.electrons import electrons_from_spec
    from qimpy_amd.ions import RadialSet, spline_projectors

    el = electrons_from_spec(spec)
    b = el.basis
    dG = 0.02
    G = np.arange(int(np.ceil(np.sqrt(2 * spec.ke_cutoff) / dG)) + 10) * dG
    l_rf = [l for rc in spec.proj_rc for l in spec.proj_l]
    radial = RadialSet(l_rf, np.stack([np.exp(-0.5 * rc**2 * G**2)
                                       for rc in spec.proj_rc for _ in spec.proj_l]), dG)
    pos = torch.tensor(spec.positions, device="cuda")
    beta = spline_projectors(b, pos, radial)
    err = float((beta - el.ions.beta.coeff[0, :, :, 0]).abs().max()
                / el.ions.beta.coeff.abs().max())
    torch.cuda.synchronize()
    _lib.profile_read()  # clear
    _lib.profile_enable(True)
    for _ in range(10):
        spline_projectors(b, pos, radial, out=beta)
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    ms, kby, cnt = prof[_lib.QMX_KERNEL_PROJ + _lib.QMX_KERNEL_C128][:3]
    tk = ms / cnt * 1e-3
    return {"workload": f"64-atom cell, {b.grid.shape[0]}^3 box, {b.nk} k, n_tot {b.n_tot}, "
                        f"{beta.shape[1]} projectors (l <= {max(l_rf)})",
            "kernel_us": round(tk * 1e6, 2), "GBps": round(kby / cnt / tk / 1e9, 1),
            "frac": round(kby / cnt / tk / 1e9 / HBM_PEAK_GBS, 4), "err_vs_analytic": err}


def chefsi_measurement(spec) -> dict:
    """One CheFSI solve of the 256-atom cell (filter order 20, ten outer
    iterations at most) in each precision, plain FP64 first. The mixed form
    locks bands and stops on ||HC - eig C|| < CHEFSI_RESIDUAL, then the FP64
    Rayleigh-Ritz; its band-energy error is taken against the plain solve."""
    import torch

    from qimpy_amd.electrons import electrons_from_spec

    res = {"workload": "256-atom cubic cell a=32.58 bohr, 144^3 box, Gamma, 1024 bands, "
                       "2048 projectors, CheFSI order 20", "unit": "s per solve",
           "mixed_residual_threshold": CHEFSI_RESIDUAL}
    out = {}
    for dt, name, kw in ((torch.complex128, "plain_fp64", {}),
                         (torch.complex64, "mixed", {"residual_threshold": CHEFSI_RESIDUAL})):
        el = electrons_from_spec(spec, dtype=dt, **kw)
        el.initialize_wavefunctions(seed=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = el.solve()
        torch.cuda.synchronize()
        res[name] = round(time.perf_counter() - t0, 3)
        res[name + "_outer_iterations"] = rep.n_iterations
        out[name] = (rep.eig.cpu().numpy(), float(rep.Eband))
        del el, rep
        torch.cuda.empty_cache()
    res["plain_over_mixed_time"] = round(res["plain_fp64"] / res["mixed"], 3)
    res["eig_err_vs_fp64_Ha"] = float(np.abs(out["mixed"][0] - out["plain_fp64"][0]).max())
    res["Eband_err_vs_fp64_Ha"] = abs(out["mixed"][1] - out["plain_fp64"][1])
    return res


def roofline(prof, wall_s: float) -> tuple[dict, dict]:
    """The dominant kernel of the profiled stream (largest total time among
    the library's kernels) against its roofline: MFMA-bound (TFLOP/s against
    the dense FP32 or FP64 MFMA peak) when its algorithmic flops per byte
    exceed the ridge peak_flops / 8 TB/s, else HBM-bound (GB/s of its
    compulsory bytes). The dominant bandwidth-bound kernel is reported
    beside it, and the whole local apply under the dense-transform
    convention."""
    from qimpy_amd import _lib

    names = {i: _lib.KERNEL_NAMES.get(i, str(i)) for i in prof}

    def entry(i):
        ms, by, cnt, fl = prof[i]
        avg_s = ms / cnt * 1e-3
        prec = "c128" if names[i].endswith("<c128>") else "c64"
        peak_tf = MFMA_PEAK_TFLOPS[prec]
        ridge = peak_tf * 1e12 / (HBM_PEAK_GBS * 1e9)
        gbs = by / cnt / avg_s / 1e9
        e = {"kernel": names[i], "avg_launch_us": round(avg_s * 1e6, 2),
             "bytes_per_launch": round(by / cnt), "launches": cnt}
        if fl and fl / by > ridge:
            tf = fl / cnt / avg_s / 1e12
            e.update({"bound": "mfma", "achieved": round(tf, 2), "peak": peak_tf,
                      "unit": "TFLOP/s", "frac": round(tf / peak_tf, 4),
                      "flops_per_launch": round(fl / cnt), "flop_per_byte": round(fl / by, 1),
                      "GBps": round(gbs, 1)})
        else:
            e.update({"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS,
                      "unit": "GB/s", "frac": round(gbs / HBM_PEAK_GBS, 4)})
        e["traffic"] = None
        traffic = measured_traffic(names[i])
        if traffic is not None:
            e["traffic"] = round(traffic["ratio"] * by / cnt)
            e["traffic_over_compulsory"] = traffic["ratio"]
            e["traffic_source"] = traffic["source"]
            if "valu_busy" in traffic:
                # HBM is not this kernel's limit: the counters put it on the
                # VALU (issue-cycle fraction, packed FP32 disabled) and on
                # LDS (waits and bank conflicts).
                e["limiter"] = {"unit": "VALU issue fraction", "achieved": traffic["valu_busy"],
                                "source": traffic["source"]}
                for key in ("lds_wait_fraction", "lds_conflict_share"):
                    if key in traffic:
                        e["limiter"][key] = traffic[key]
        return e

    lib_ids = [i for i in prof if not names[i].startswith("local_apply")]
    kid = max(lib_ids, key=lambda i: prof[i][0])
    out = entry(kid)
    bw = [i for i in lib_ids if names[i] in BANDWIDTH_KERNELS and entry(i)["bound"] == "hbm"]
    if bw:
        bid = max(bw, key=lambda i: prof[i][0])
        if bid != kid:
            out["dominant_bandwidth_kernel"] = entry(bid)
    kernels = {}
    for i in lib_ids:
        ms, by, cnt, fl = prof[i]
        k = {"launches": cnt, "avg_us": round(ms / cnt * 1e3, 2),
             "bytes_per_launch": round(by / cnt),
             "GBps": round(by / cnt / (ms / cnt * 1e-3) / 1e9, 1),
             "share_of_stream": round(ms * 1e-3 / wall_s, 3)}
        if fl:
            k["flops_per_launch"] = round(fl / cnt)
            k["TFLOPs"] = round(fl / (ms * 1e-3) / 1e12, 2)
        kernels[names[i]] = k
    apply = [prof[i] for i in prof if names[i].startswith("local_apply")]
    apply_ms = sum(a[0] for a in apply)
    if apply_ms:
        gbs = sum(a[1] for a in apply) / (apply_ms * 1e-3) / 1e9
        out["local_apply_dense_convention"] = {
            "bytes_per_band": "b(2G + 4N) + b_V N / B", "achieved": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4), "calls": sum(a[2] for a in apply),
            "avg_call_us": round(apply_ms / sum(a[2] for a in apply) * 1e3, 2)}
    return out, kernels


def measured_traffic(kernel: str):
    """HBM traffic of `kernel` relative to its compulsory bytes, from the two
    rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE, each scaled by the same
    counter on a 1 GiB copy) recorded in profiles/pmc_traffic.json; the ratio
    is applied to this run's bytes per launch. None without an entry."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for e in table.get("entries", []):
        if e.get("kernel") == kernel:
            out = {"ratio": e["traffic_over_compulsory"],
                   "source": f"{e['profile']} via profiles/pmc_traffic.json"}
            for key in ("valu_busy", "lds_wait_fraction", "lds_conflict_share"):
                if key in e:
                    out[key] = e[key]
            return out
    return None


def accuracy(mixed, plain) -> dict:
    common = sorted(set(mixed.results) & set(plain.results))
    diff = [float(np.abs(mixed.results[i].eig - plain.results[i].eig).max()) for i in common]
    eband = [abs(mixed.results[i].Eband - plain.results[i].Eband) for i in common]
    bound = [mixed.results[i].err_bound for i in mixed.results]
    return {"requests_compared": len(common), "eig_err_vs_fp64_Ha_max": max(diff),
            "eig_err_vs_fp64_Ha_p99": float(np.percentile(diff, 99)),
            "Eband_err_vs_fp64_Ha_max": max(eband),
            "mixed_err_bound_Ha_max": max(bound),
            "requests_within_1e-6": int(sum(d <= 1e-6 for d in diff)),
            "all_converged": bool(all(r.converged for r in mixed.results.values()))}


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requests, cells and k-points would shard across GPUs with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from qimpy_amd import autotune, gemm, rc, systems

    global BAND_CLASSES, GROUPS
    BAND_CLASSES = tuple(int(c) for c in args.band_classes.split(",") if c)
    GROUPS = args.groups
    rc.init()
    rc.require_gpu()
    if args.bf16x3:
        gemm.set_bf16x3(True)
    if args.side_only:
        spec1 = systems.make_cell(1)
        print(json.dumps({"collect_density": density_measurement(spec1),
                          "projectors": projectors_measurement(spec1)}), flush=True)
        return
    C64, C128 = torch.complex64, torch.complex128
    from concurrent.futures import ThreadPoolExecutor

    from qimpy_amd.serve import stage_requests

    n_req = (args.warmup + args.steps) * args.window
    arrivals, ids = systems.make_request_stream(n_req, rate=args.rate)
    n_cells = max(n_req, args.stable_requests)
    with ThreadPoolExecutor(max_workers=16) as pool:
        all_specs = list(pool.map(lambda i: systems.make_cfg2_cell(i, cfg=3), range(n_cells)))
    specs = [all_specs[i] for i in ids]
    t0 = time.perf_counter()
    staged = None if args.host_inputs else stage_requests(all_specs, np.zeros(n_cells))
    print(f"[bench] stream of {n_req} requests over {arrivals[-1]:.1f} s at {args.rate:.0f} req/s, "
          f"{args.warmup} + {args.steps} windows of {args.window}; inputs "
          f"{'host' if staged is None else f'staged in HBM in {time.perf_counter() - t0:.1f} s'}",
          flush=True)
    out = {}
    mixed = plain = None
    if args.only in ("all", "mixed"):
        mixed, prof = served(specs, arrivals, C64, staged, profile=args.one_stream,
                             streams=not args.one_stream)
        wm = windows(mixed, arrivals, args.warmup, args.steps, args.window)
        print(f"[bench] mixed: {wm['value']:.2f} solves/s, p50 {wm['p50']:.2f} s, "
              f"p99 {wm['p99']:.2f} s", flush=True)
        # The roofline pass: the same stream with both bucket groups on one
        # HIP stream. With a stream per group, two groups' kernels share the
        # device and a launch's event-to-event time includes the other's
        # work; on one stream each launch's time is its own.
        rl_run = mixed
        if not args.one_stream:
            rl_run, prof = served(specs, arrivals, C64, staged, profile=True, streams=False)
            w1 = windows(rl_run, arrivals, args.warmup, args.steps, args.window)
            print(f"[bench] mixed, one stream (roofline pass): {w1['value']:.2f} solves/s",
                  flush=True)
    if args.only in ("all", "plain"):
        plain, _ = served(specs, arrivals, C128, staged, streams=not args.one_stream)
        wp = windows(plain, arrivals, args.warmup, args.steps, args.window)
        print(f"[bench] plain: {wp['value']:.2f} solves/s", flush=True)
    head, wh = (mixed, wm) if mixed is not None else (plain, wp)
    out.update({
        "metric": "served band solves/s with p50/p99 latency (mixed c64/FP64 vs plain FP64, "
                  "<=1e-6 Ha vs FP64)",
        "value": round(wh["value"], 3), "unit": "band solves/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wh["ms_per_step"], 1),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64" if mixed is not None else "c128",
        "data": "synthetic (seeded cells, potentials, projectors, arrivals and start bands)",
        "config": {"workload": (f"served stream: {n_req} requests, Poisson {args.rate:.0f} req/s, "
                                "cells of the uneven batched distribution (boxes 24..60^3, 8-64 "
                                "bands, 1-16 k, 2-16 atoms x 8 projectors, 20 Ha), inputs staged "
                                "in HBM, HBM budget 180 GB; a step is a window of "
                                f"{args.window} completed requests"),
                   "requests": n_req, "rate_per_s": args.rate, "window": args.window,
                   "inputs": "host arrays" if args.host_inputs else "staged in HBM",
                   "band_classes": list(BAND_CLASSES), "groups": GROUPS,
                   "mixed_eig_tolerance_Ha": EIG_TOLERANCE,
                   "plain_eig_threshold": PLAIN_THRESHOLD},
        "latency_s": {"p50": round(wh["p50"], 3), "p99": round(wh["p99"], 3),
                      "n": wh["n_latency"], "note": "saturated server: queueing-dominated"},
        "step_ms_min_max": wh["step_ms_min_max"],
        "mean_iterations": round(wh["mean_iterations"], 1),
        "tuned_launch_choices": autotune.load_default(),
    })
    if mixed is not None and plain is not None:
        out["plain_fp64"] = {"value": round(wp["value"], 3), "unit": "band solves/s",
                             "latency_s": {"p50": round(wp["p50"], 3), "p99": round(wp["p99"], 3)},
                             "mean_iterations": round(wp["mean_iterations"], 1)}
        out["mixed_over_plain"] = round(wm["value"] / wp["value"], 3)
        out["accuracy"] = acc = accuracy(mixed, plain)
        print(f"[bench] accuracy: {acc}", flush=True)
    if mixed is not None:
        # The profile covers the whole mixed stream (warm-up windows too) of
        # the roofline pass.
        out["roofline"], out["kernels"] = roofline(prof, rl_run.wall_s + 1e-9)
        out["roofline"]["measured_on"] = (
            "the headline stream served with both bucket groups on one HIP stream (kernels not "
            "overlapping); HIP events around every library launch on that stream")
        if rl_run is not mixed:
            out["roofline"]["pass_value"] = round(w1["value"], 3)
        del mixed, plain, rl_run
        torch.cuda.empty_cache()
        if args.stable_requests:
            rate2 = round(0.8 * wm["value"], 1)
            arr2, ids2 = systems.make_request_stream(args.stable_requests, rate=rate2, seed=3001)
            specs2 = [all_specs[i] for i in ids2]
            st, _ = served(specs2, arr2, C64, None if staged is None else [staged[i] for i in ids2])
            w0 = max(1, args.stable_requests // 10)  # untimed start
            res = sorted(st.results.values(), key=lambda r: r.finished)[w0:]
            lat = np.array([r.latency for r in res])
            out["latency_stable_rate"] = {
                "rate_per_s": rate2, "requests": args.stable_requests, "timed": int(lat.size),
                "p50_s": round(float(np.percentile(lat, 50)), 4),
                "p99_s": round(float(np.percentile(lat, 99)), 4),
                "throughput": round(lat.size / (res[-1].finished - res[0].finished), 3),
                "note": "mixed form at 80% of the measured saturation rate"}
            print(f"[bench] stable rate {rate2}/s: {out['latency_stable_rate']}", flush=True)
            del st
    if not args.no_extra:
        out["diamond_2_atom"] = diamond_measurement(systems.make_cell(0))
        print(f"[bench] diamond: {out['diamond_2_atom']}", flush=True)
        cells = systems.make_cfg2_cells(64)
        vm, rm = batched(cells, C64)
        vp, rp = batched(cells, C128)
        out["batched_64_cells"] = {
            "mixed": round(vm, 3), "plain_fp64": round(vp, 3), "unit": "band solves/s",
            "mixed_over_plain": round(vm / vp, 3),
            "eig_err_vs_fp64_Ha": max(float(np.abs(rm[i].eig - rp[i].eig).max()) for i in rm),
            "mixed_err_bound_Ha_max": max(rm[i].err_bound for i in rm)}
        print(f"[bench] batched 64 cells: {out['batched_64_cells']}", flush=True)
        spec1 = systems.make_cell(1)
        t1m, r1m = single(spec1, C64, n_iterations=200, residual_threshold=3e-4)
        t1p, r1p = single(spec1, C128, n_iterations=200, eig_threshold=PLAIN_THRESHOLD)
        out["single_64_atom_cell"] = {
            "workload": "64-atom cubic cell a=20.52 bohr, 96^3 box, Gamma, 256 bands (+26), "
                        "512 projectors, drop-in Davidson", "unit": "s per solve",
            "mixed": round(t1m, 3), "plain_fp64": round(t1p, 3),
            "plain_over_mixed_time": round(t1p / t1m, 3),
            "iterations": [r1m.n_iterations, r1p.n_iterations],
            "mixed_subspace_gemms": "BF16x3" if gemm.bf16x3_enabled() else "FP32 MFMA",
            "eig_err_vs_fp64_Ha": float(np.abs(r1m.eig.cpu().numpy() - r1p.eig.cpu().numpy()).max())}
        print(f"[bench] single 64-atom cell: {out['single_64_atom_cell']}", flush=True)
        out["chefsi_256_atom_cell"] = chefsi_measurement(systems.make_cell(4))
        print(f"[bench] CheFSI 256-atom cell: {out['chefsi_256_atom_cell']}", flush=True)
        out["collect_density"] = density_measurement(spec1)
        out["projectors"] = projectors_measurement(spec1)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(specs[:64])
        print(f"[bench] cpu baseline: {out['cpu_baseline']}", flush=True)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
