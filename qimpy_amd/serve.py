"""Request loop for band solves: continuous batching on one GPU.

Requests (one cell each, with an arrival time) enter a BatchSolver as they
arrive. Every iteration of the solver retires converged jobs, admits arrived
requests into the freed capacity (an HBM budget and a per-bucket k-point
limit) and advances every resident job by one Davidson step, so requests that
arrive while others run share the device with them; state (bands, H applied
to them, potentials, projectors) stays resident from admission to result.
Latency is result-ready time minus arrival time; the report gives p50/p99
beside throughput. QimPy has no request loop: each request is the band solve
of dft/electrons/_davidson.py:141-235 for one cell.
"""
from __future__ import annotations

import time
from collections import deque
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from .batched import BatchSolver, CellJob, CellResult


@dataclass
class ServeReport:
    n_requests: int
    wall_s: float  #: first arrival to last result
    throughput: float  #: results per second over wall_s
    p50_s: float
    p99_s: float
    mean_iterations: float
    results: dict = field(repr=False, default_factory=dict)

    def as_dict(self) -> dict:
        return {k: v for k, v in self.__dict__.items() if k != "results"}


def stage_requests(specs: Sequence, arrivals: Sequence[float], workers: int = 16) -> list:
    """Jobs of the requests with their inputs resident in HBM: sphere
    tables, potential box, k-points, positions and couplings built on the
    host by `workers` threads and copied to the device (the pinned host
    copies are dropped). Served with serve(..., jobs=...), a request then
    enters the solver at its arrival time with nothing left to upload."""
    def build(i):
        job = CellJob(specs[i], job_id=i, arrival=float(arrivals[i]))
        job.stage()
        return job
    with ThreadPoolExecutor(max_workers=max(1, workers)) as pool:
        jobs = list(pool.map(build, range(len(specs))))
    torch.cuda.synchronize()
    return jobs


def serve(specs: Sequence, arrivals: Sequence[float], solver: BatchSolver, *,
          on_result: Optional[Callable[[CellResult], None]] = None,
          progress: Optional[Callable[[str], None]] = None, build_workers: int = 4,
          jobs: Optional[Sequence[CellJob]] = None) -> ServeReport:
    """Serve requests `specs[i]` arriving at `arrivals[i]` seconds after the
    solver's clock start (ascending). Without `jobs`, a request's job
    (sphere tables, on the host) is built by a pool of `build_workers`
    threads from its arrival on, overlapping the solver's device work, and
    its tables are copied to the device at admission; with `jobs` (from
    stage_requests) the staged job of request i enters at its arrival.
    Built jobs enter the solver in arrival order."""
    order = np.argsort(np.asarray(arrivals, dtype=np.float64), kind="stable")
    n = len(order)
    nxt = 0
    building: deque = deque()
    t_last_report = time.perf_counter()
    with ThreadPoolExecutor(max_workers=max(1, build_workers)) as pool:
        while nxt < n or building or not solver.idle:
            now = solver.clock()
            while nxt < n and arrivals[order[nxt]] <= now:
                i = int(order[nxt])
                if jobs is not None:
                    solver.submit([jobs[i]])
                else:
                    building.append(pool.submit(CellJob, specs[i], job_id=i,
                                                arrival=float(arrivals[i])))
                nxt += 1
            while building and building[0].done():
                solver.submit([building.popleft().result()])
            if solver.idle:
                if building:  # nothing resident: wait for the next job to be built
                    building[0].result()
                elif nxt < n:  # or for the next arrival
                    time.sleep(max(0.0, float(arrivals[order[nxt]]) - solver.clock()))
                continue
            for res in solver.tick(now=now):
                if on_result is not None:
                    on_result(res)
            if progress is not None and time.perf_counter() - t_last_report > 5.0:
                t_last_report = time.perf_counter()
                progress(f"t={solver.clock():.1f}s arrived {nxt}/{n} done {len(solver.results)}"
                         f" queued {len(solver.queue)}")
    torch.cuda.synchronize()
    res = solver.results
    lat = np.array([r.latency for r in res.values()])
    t_first = float(np.min(np.asarray(arrivals)))
    t_end = max(r.finished for r in res.values())
    wall = t_end - t_first
    return ServeReport(
        n_requests=len(res), wall_s=wall, throughput=len(res) / wall if wall > 0 else np.inf,
        p50_s=float(np.percentile(lat, 50)), p99_s=float(np.percentile(lat, 99)),
        mean_iterations=float(np.mean([r.iterations for r in res.values()])), results=res)
