"""Band-solve benchmark on one MI355X: band solves/s and p50/p99 latency of
a served request stream, mixed precision vs plain FP64.

Workload (one step = one served stream): band-solve requests arrive as a
seeded Poisson process at 200 req/s, each a cell drawn from the uneven
distribution of the batched workload (cubic boxes 24..60^3, 8-64 bands,
1-16 k-points, 2-16 atoms with 8 projectors each, 20 Ha). The server
(qimpy_amd.serve) builds each request's basis when it arrives, admits it
against a 200 GB HBM budget and iterates every resident request in lockstep
batches (continuous batching); a request's latency runs from its arrival to
its FP64 result. The mixed form iterates in complex64 (residual 3e-4) and
finishes each request with an FP64 Rayleigh-Ritz; the plain form iterates
in complex128 to the reference's |deig| < 1e-8. Inputs are synthetic and
resident on the host before timing; the stream is timed from its first
arrival to its last result.

Printed (one JSON line): served band solves/s of the mixed form with p50 and
p99 latency, the plain FP64 form on the same stream, mixed-vs-FP64
eigenvalue and band-energy errors, an alone-vs-in-the-mix identity check,
the 64-cell batched workload and the single 64-atom solve in both forms, the
roofline of the dominant bandwidth kernel (live HIP-event timing of every
library launch on its stream) and the FP64 CPU oracle timed on a bounded
sample of the same requests.
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
RATE = 200.0  # requests per second
MIXED_RESIDUAL = 5e-4  # residual stop of the mixed form; calibrated in the design notes
PLAIN_THRESHOLD = 1e-8
#: Library kernels that are bandwidth (not latency) bound; the roofline
#: reports the one with the largest total time.
BANDWIDTH_KERNELS = {f"{k}<{t}>" for t in ("c64", "c128") for k in (
    "k_sphere_to_cols", "k_cols_to_planes", "k_xpass_potential", "k_planes_to_cols",
    "k_cols_to_sphere", "k_plane_potential", "k_gemm")}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1, help="timed served streams")
    ap.add_argument("--warmup", type=int, default=1, help="untimed warm-up streams")
    ap.add_argument("--requests", type=int, default=128, help="requests per stream")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle sample")
    ap.add_argument("--no-extra", action="store_true",
                    help="skip the batched and single-solve side measurements")
    ap.add_argument("--side-only", action="store_true",
                    help="only the density and projector measurements (no served streams)")
    ap.add_argument("--chefsi-256", action="store_true",
                    help="also time the 256-atom CheFSI solve (144^3 box, 1024 bands), ~20 s")
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


def cpu_baseline(specs, gpu_iterations, budget_s: float = 12.0) -> dict:
    """FP64 oracle (numpy/scipy) on this host for the first requests of the
    stream, in order, until about `budget_s` of CPU run time: 2 Davidson
    iterations each, extrapolated to the iteration count the FP64 GPU form
    needed for that request."""
    import psutil
    from threadpoolctl import threadpool_limits

    from oracle.basis import Basis as OBasis
    from oracle.davidson import davidson, initial_state
    from oracle.hamiltonian import Hamiltonian

    cores = psutil.cpu_count(logical=False) or os.cpu_count() or 1
    t_est, t_run, names = 0.0, 0.0, []
    with threadpool_limits(limits=cores):
        for spec, its in zip(specs, gpu_iterations):
            ob = OBasis(spec)
            H = Hamiltonian(ob)
            C = initial_state(ob, spec.n_bands)
            t0 = time.perf_counter()
            davidson(ob, H, C, spec.n_bands, spec.n_bands_extra, 2, 0.0)
            t = time.perf_counter() - t0
            t_run += t
            t_est += t / 3.0 * (its + 1)  # the start counts as one more iteration of work
            names.append(f"{spec.shape[0]}^3/{len(spec.k)}k/{spec.n_bands}b")
            if t_run > budget_s:
                break
    return {
        "value": len(names) / t_est, "unit": "band solves/s", "cores": int(cores), "kind": "port",
        "sample": (f"FP64 numpy/scipy oracle on the first {len(names)} requests of the stream "
                   f"({', '.join(names)}): 2 Davidson iterations each ({t_run:.1f} s on {cores} "
                   f"threads of {cpu_model()}), extrapolated to the iterations of their FP64 GPU "
                   f"solves"),
    }


def measured_traffic(kernel: str, n_requests: int):
    """HBM traffic of `kernel` relative to its compulsory bytes on this served
    stream, from the two rocprofv3 PMC passes (FETCH_SIZE, WRITE_SIZE, each
    scaled by the same counter on a 1 GiB copy) recorded in
    profiles/pmc_traffic.json. The stream's batching depends on timing, so
    the ratio (measured over the profiled runs' dispatches) is applied to
    this run's bytes per launch. None when the file has no entry."""
    path = os.path.join(ROOT, "profiles", "pmc_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
    except (OSError, ValueError):
        return None
    for e in table.get("entries", []):
        if e.get("kernel") == kernel and e.get("requests") == n_requests:
            return {"ratio": e["traffic_over_compulsory"],
                    "source": f"{e['profile']} via profiles/pmc_traffic.json"}
    return None


def served(specs, arrivals, dtype, profile=False):
    import torch

    from qimpy_amd import _lib
    from qimpy_amd.batched import BatchSolver
    from qimpy_amd.serve import serve

    solver = BatchSolver(dtype, residual_threshold=MIXED_RESIDUAL, eig_threshold=PLAIN_THRESHOLD,
                         max_iterations=200, hbm_budget=200e9)
    if profile:
        _lib.profile_enable(True)
        _lib.profile_read()
    torch.cuda.synchronize()
    rep = serve(specs, arrivals, solver)
    prof = None
    if profile:
        _lib.profile_enable(False)
        prof = _lib.profile_read()
    return rep, prof


def batched(specs, dtype):
    import torch

    from qimpy_amd.batched import BatchSolver, CellJob

    solver = BatchSolver(dtype, residual_threshold=MIXED_RESIDUAL, eig_threshold=PLAIN_THRESHOLD,
                         max_iterations=200)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    solver.submit([CellJob(s, job_id=i) for i, s in enumerate(specs)])
    res = solver.run()
    return len(res) / (time.perf_counter() - t0), res


def single(spec, dtype, **kw):
    import torch

    from qimpy_amd.electrons import electrons_from_spec

    el = electrons_from_spec(spec, dtype=dtype, n_iterations=200, **kw)
    el.initialize_wavefunctions(seed=0)
    el.solve()  # warm-up
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    el.initialize_wavefunctions(seed=0)
    rep = el.solve()
    torch.cuda.synchronize()
    return 1.0 / (time.perf_counter() - t0), rep


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
    same output. The kernel is timed by the library's HIP events around each
    launch (qmx_profile, on the launch stream); the Python call, whose host
    work is longer than the 32 us kernel, is reported beside it. Algorithmic
    bytes: the complex128 output, 16 bytes per (k, atom, projector, plane
    wave) entry (the inputs are a few MB)."""
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
    stream = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    _lib.profile_read()  # clear
    _lib.profile_enable(True)
    e0.record(stream)
    for _ in range(10):
        spline_projectors(b, pos, radial, out=beta)
    e1.record(stream)
    torch.cuda.synchronize()
    prof = _lib.profile_read()
    _lib.profile_enable(False)
    t = e0.elapsed_time(e1) / 10 * 1e-3
    ms, kby, cnt = prof[_lib.QMX_KERNEL_PROJ + _lib.QMX_KERNEL_C128]
    tk = ms / cnt * 1e-3  # the kernel alone: library HIP events around each launch
    by = beta.numel() * 16
    return {"workload": f"64-atom cell, {b.grid.shape[0]}^3 box, {b.nk} k, n_tot {b.n_tot}, "
                        f"{beta.shape[1]} projectors (l <= {max(l_rf)})",
            "kernel_us": round(tk * 1e6, 2), "GBps": round(kby / cnt / tk / 1e9, 1),
            "frac": round(kby / cnt / tk / 1e9 / HBM_PEAK_GBS, 4),
            "ms_per_python_call": round(t * 1e3, 3), "err_vs_analytic": err}


def chefsi_measurement(spec, residual: float = 3e-5) -> dict:
    """One CheFSI solve of the 256-atom cell (filter order 20, ten outer
    iterations at most) in each precision, plain FP64 first. The mixed form
    locks bands and stops on ||HC - eig C|| < residual, then the FP64
    Rayleigh-Ritz; its error is taken against the plain solve."""
    import torch

    from qimpy_amd.electrons import electrons_from_spec

    res = {"workload": "256-atom cubic cell, 144^3 box, Gamma, 1024 bands, CheFSI order 20",
           "unit": "s", "mixed_residual_threshold": residual}
    out = {}
    for dt, name, kw in ((torch.complex128, "plain_fp64", {}),
                         (torch.complex64, "mixed", {"residual_threshold": residual})):
        el = electrons_from_spec(spec, dtype=dt, **kw)
        el.initialize_wavefunctions(seed=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rep = el.solve()
        torch.cuda.synchronize()
        res[name] = round(time.perf_counter() - t0, 3)
        res[name + "_iterations"] = rep.n_iterations
        out[name] = (rep.eig.cpu().numpy(), float(rep.Eband))
        del el, rep
        torch.cuda.empty_cache()
    res["eig_err_vs_fp64_Ha"] = float(np.abs(out["mixed"][0] - out["plain_fp64"][0]).max())
    res["Eband_err_vs_fp64_Ha"] = abs(out["mixed"][1] - out["plain_fp64"][1])
    return res


def main():
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(requests, cells and k-points would shard across GPUs with no collectives)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from qimpy_amd import _lib, rc, systems
    from qimpy_amd.batched import BatchSolver, CellJob

    rc.init()
    rc.require_gpu()
    if args.side_only:  # kernel evidence for the side measurements, no headline line
        spec1 = systems.make_cell(1)
        print(json.dumps({"collect_density": density_measurement(spec1),
                          "projectors": projectors_measurement(spec1)}), flush=True)
        return
    C64, C128 = torch.complex64, torch.complex128
    arrivals, ids = systems.make_request_stream(args.requests, rate=RATE)
    specs = [systems.make_cfg2_cell(i, cfg=3) for i in ids]
    print(f"[bench] stream of {len(specs)} requests over {arrivals[-1]:.2f} s", flush=True)

    # Warm-up streams (kernels, allocator, library handles):
    # Warm-up steps: the same stream, untimed, in both precisions (library
    # modules, allocator and plan-block pools reach their working sizes).
    for _ in range(args.warmup):
        for dt in (C64, C128):
            served(specs, arrivals, dt)
    print("[bench] warm-up done", flush=True)

    reps, prof = [], None
    for step in range(args.steps):
        rep, p = served(specs, arrivals, C64, profile=(step == args.steps - 1))
        reps.append(rep)
        prof = p or prof
        print(f"[bench] mixed stream {step}: {rep.throughput:.2f} solves/s, p50 {rep.p50_s:.2f} s,"
              f" p99 {rep.p99_s:.2f} s", flush=True)
    plain, _ = served(specs, arrivals, C128)
    print(f"[bench] plain stream: {plain.throughput:.2f} solves/s", flush=True)
    mixed = reps[-1]
    err = max(float(np.abs(mixed.results[i].eig - plain.results[i].eig).max()) for i in mixed.results)
    eband_err = max(abs(mixed.results[i].Eband - plain.results[i].Eband) for i in mixed.results)

    # The same request alone gives the same result as in the mix:
    ident = None
    if not args.no_extra:
        ident = 0.0
        for i in (0, len(specs) // 2):
            alone = BatchSolver(C64, residual_threshold=MIXED_RESIDUAL, max_iterations=200)
            alone.submit([CellJob(specs[i], job_id=i)])
            r = alone.run()[i]
            ident = max(ident, float(np.abs(r.eig - mixed.results[i].eig).max()))

    # Roofline of the dominant bandwidth kernel over the timed mixed stream:
    names = {i: _lib.KERNEL_NAMES.get(i, str(i)) for i in prof}
    bw = {i: v for i, v in prof.items() if names[i] in BANDWIDTH_KERNELS}
    kid = max(bw, key=lambda i: bw[i][0])
    ms, by, cnt = prof[kid]
    avg_s = ms / cnt * 1e-3
    achieved = by / cnt / avg_s / 1e9
    wall = mixed.wall_s
    kernels = {
        names[i]: {"launches": c, "avg_us": round(m / c * 1e3, 2),
                   "GBps": round(b / c / (m / c * 1e-3) / 1e9, 1),
                   "share_of_stream": round(m * 1e-3 / wall, 3)}
        for i, (m, b, c) in prof.items() if not names[i].startswith("local_apply")
    }
    # The whole local apply (all its launches, V reorder included) under the
    # dense-transform byte convention b(2G + 4N) + b_V N / B per band:
    apply = [prof[i] for i in prof if names[i].startswith("local_apply")]
    apply_ms = sum(a[0] for a in apply)
    apply_gbs = sum(a[1] for a in apply) / (apply_ms * 1e-3) / 1e9 if apply_ms else None
    out = {
        "metric": "served band solves/s with p50/p99 latency (mixed c64/FP64 vs plain FP64, "
                  "<=1e-6 Ha vs FP64)",
        "value": round(mixed.throughput, 3),
        "unit": "band solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * wall, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64+f64",
        "data": "synthetic (seeded cells, potentials, projectors, arrivals and start bands)",
        "config": {"workload": (f"served stream: {len(specs)} requests, Poisson {RATE:.0f} req/s, "
                                "cells of the uneven batched distribution (boxes 24..60^3, 8-64 "
                                "bands, 1-16 k, 2-16 atoms x 8 projectors, 20 Ha), HBM budget 200 GB"),
                   "requests": len(specs), "rate_per_s": RATE,
                   "mixed_residual_threshold": MIXED_RESIDUAL,
                   "plain_eig_threshold": PLAIN_THRESHOLD},
        "latency_s": {"p50": round(mixed.p50_s, 3), "p99": round(mixed.p99_s, 3)},
        "plain_fp64": {"value": round(plain.throughput, 3), "unit": "band solves/s",
                       "latency_s": {"p50": round(plain.p50_s, 3), "p99": round(plain.p99_s, 3)},
                       "mean_iterations": round(plain.mean_iterations, 1)},
        "mixed_over_plain": round(mixed.throughput / plain.throughput, 3),
        "mean_iterations": round(mixed.mean_iterations, 1),
        "eig_err_vs_fp64_Ha": err,
        "Eband_err_vs_fp64_Ha": eband_err,
        "alone_vs_mix_max_diff_Ha": ident,
        "roofline": {
            "kernel": names[kid], "bound": "hbm", "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None, "avg_launch_us": round(avg_s * 1e6, 2),
            "bytes_per_launch": round(by / cnt),
            "local_apply_dense_convention": None if apply_gbs is None else {
                "bytes_per_band": "b(2G + 4N) + b_V N / B",
                "achieved": round(apply_gbs, 1), "frac": round(apply_gbs / HBM_PEAK_GBS, 4),
                "calls": sum(a[2] for a in apply),
                "avg_call_us": round(apply_ms / sum(a[2] for a in apply) * 1e3, 2),
            },
        },
        "kernels": kernels,
    }
    traffic = measured_traffic(names[kid], len(specs))
    if traffic is not None:
        out["roofline"]["traffic"] = round(traffic["ratio"] * by / cnt)
        out["roofline"]["traffic_over_compulsory"] = traffic["ratio"]
        out["roofline"]["traffic_source"] = traffic["source"]
    if not args.no_extra:
        cells = systems.make_cfg2_cells(64)
        vm, rm = batched(cells, C64)
        vp, rp = batched(cells, C128)
        berr = max(float(np.abs(rm[i].eig - rp[i].eig).max()) for i in rm)
        out["batched_64_cells"] = {"mixed": round(vm, 3), "plain_fp64": round(vp, 3),
                                   "unit": "band solves/s", "mixed_over_plain": round(vm / vp, 3),
                                   "eig_err_vs_fp64_Ha": berr}
        print(f"[bench] batched 64 cells: mixed {vm:.2f}, plain {vp:.2f} solves/s", flush=True)
        spec1 = systems.make_cell(1)
        v1m, r1m = single(spec1, C64, residual_threshold=MIXED_RESIDUAL)
        v1p, r1p = single(spec1, C128, eig_threshold=PLAIN_THRESHOLD)
        out["single_64_atom_cell"] = {
            "workload": "64-atom cubic cell a=20.52 bohr, 96^3 box, Gamma, 256 bands (+26), "
                        "512 projectors, Davidson",
            "mixed": round(v1m, 3), "plain_fp64": round(v1p, 3), "unit": "band solves/s",
            "iterations": [r1m.n_iterations, r1p.n_iterations],
            "eig_err_vs_fp64_Ha": float(np.abs(r1m.eig.cpu().numpy() - r1p.eig.cpu().numpy()).max())}
        print(f"[bench] single 64-atom cell: mixed {v1m:.3f}, plain {v1p:.3f} solves/s", flush=True)
        out["collect_density"] = density_measurement(spec1)
        print(f"[bench] collect_density: {out['collect_density']}", flush=True)
        out["projectors"] = projectors_measurement(spec1)
        print(f"[bench] projectors: {out['projectors']}", flush=True)
    if args.chefsi_256:
        out["chefsi_256_atom_cell"] = chefsi_measurement(systems.make_cell(4))
        print(f"[bench] CheFSI 256-atom cell: {out['chefsi_256_atom_cell']}", flush=True)
    if not args.no_cpu:
        n_cpu = min(8, len(specs))
        out["cpu_baseline"] = cpu_baseline(
            specs[:n_cpu], [plain.results[i].iterations for i in range(n_cpu)])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
