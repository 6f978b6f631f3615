"""Admission of the served solver on the host (qimpy_amd.batched.Scheduler,
the bookkeeping BatchSolver.tick runs): a simulated HBM budget that defers
requests, requests cancelled while queued and while running, and bucket
slots freed mid-run. Simulated jobs take a fixed number of steps and
produce a result that depends on the job alone; every test checks that the
results of the other jobs are unchanged by the deferral or cancellation and
that the held estimates never exceed the budget once a job is resident.
"""
import pytest

from qimpy_amd.batched import Scheduler

UNIT = 10**9


class SimJob:
    def __init__(self, job_id, key=((24, 24, 24), 32), K=1, n_tot=500, units=1.0, steps=3,
                 arrival=0.0):
        self.job_id = job_id
        self._key = key
        self.K = K
        self.n_tot = n_tot
        self.nbytes = int(units * UNIT)
        self.steps = steps
        self.arrival = arrival
        self.done = 0

    def hbm_bytes(self, nb, dtype, G=None):
        return self.nbytes

    def result(self):
        return (self.job_id, self.steps, self.K * 1000 + self.n_tot)


def simulate(jobs, sched, cancel_at=None, max_ticks=100):
    """Tick loop in the order of BatchSolver.tick: retire (converged or
    cancelled) and release, admit, step every resident job. Returns
    (results, admitted tick, finished tick, peak held bytes with more than
    one job resident)."""
    cancel_at = cancel_at or {}
    sched.submit(sorted(jobs, key=lambda j: j.arrival))
    running: dict = {}
    results, admitted, finished = {}, {}, {}
    peak = 0
    for t in range(max_ticks):
        for jid in cancel_at.get(t, []):
            ids = {j.job_id for js in running.values() for j in js}
            sched.cancel(jid, ids)
        for key, js in list(running.items()):
            conv = [j for j in js if j.done >= j.steps]
            fin, drop = sched.split_retiring(js, conv)
            for j in fin:
                results[j.job_id] = j.result()
                finished[j.job_id] = t
            for j in fin + drop:
                sched.release(j.job_id)
                js.remove(j)
        state = {k: (sum(j.K for j in js), max(j.n_tot for j in js))
                 for k, js in running.items() if js}
        taken = sched.admit(float(t), 0, state, 0, lambda: [0.0] * sched.n_groups)
        for key, js in taken.items():
            running.setdefault(key, []).extend(js)
            for j in js:
                admitted[j.job_id] = t
        resident = sum(len(js) for js in running.values())
        if resident > 1:
            peak = max(peak, sched.hbm_in_use)
        for js in running.values():
            for j in js:
                j.done += 1
        if not sched.queue and not any(running.values()):
            break
    return results, admitted, finished, peak


def make(**kw):
    args = dict(hbm_budget=2.5 * UNIT, max_k_per_bucket=2048, max_bucket_bytes=1e18, csize=8)
    args.update(kw)
    return Scheduler(**args)


def test_budget_defers_then_admits():
    jobs = [SimJob(i) for i in range(4)]
    res, adm, fin, peak = simulate(jobs, make())
    assert adm == {0: 0, 1: 0, 2: 3, 3: 3}  # two fit; the others wait for the first to finish
    assert peak <= 2.5 * UNIT
    alone = {}
    for i in range(4):
        r, _, _, _ = simulate([SimJob(i)], make())
        alone.update(r)
    assert res == alone


def test_first_job_always_admitted_and_fifo():
    """A job larger than the budget runs alone; later small jobs do not
    overtake a deferred head of the queue."""
    jobs = [SimJob(0, units=4.0, steps=2), SimJob(1, units=1.0), SimJob(2, units=0.5)]
    res, adm, fin, peak = simulate(jobs, make())
    assert adm[0] == 0 and adm[1] == 2 and adm[2] == 2
    assert set(res) == {0, 1, 2}
    assert peak <= 2.5 * UNIT


def test_arrivals_respected():
    jobs = [SimJob(0, arrival=0.0), SimJob(1, arrival=2.0), SimJob(2, arrival=5.0)]
    _, adm, _, _ = simulate(jobs, make(hbm_budget=100 * UNIT))
    assert adm == {0: 0, 1: 2, 2: 5}


def test_cancel_queued():
    jobs = [SimJob(i) for i in range(4)]
    res, adm, _, peak = simulate(jobs, make(), cancel_at={1: [2]})
    assert 2 not in res and 2 not in adm
    assert adm == {0: 0, 1: 0, 3: 3}
    base, _, _, _ = simulate([SimJob(i) for i in (0, 1, 3)], make())
    assert res == base  # the other results are unchanged
    sched = make()
    assert not sched.cancel(99, set())  # unknown job


def test_cancel_running_frees_budget():
    """Cancelling a running job drops it at the next tick, releases its
    estimate and lets the deferred job in at once; it gives no result."""
    jobs = [SimJob(0, steps=10), SimJob(1, steps=10), SimJob(2, steps=3)]
    sched = make()
    res, adm, fin, peak = simulate(jobs, sched, cancel_at={2: [1]})
    assert 1 not in res and 1 in sched.cancelled
    assert adm[2] == 2  # admitted in the tick that dropped job 1
    assert set(res) == {0, 2}
    alone = {}
    for i in (0, 2):
        r, _, _, _ = simulate([SimJob(i, steps=jobs[i].steps)], make())
        alone.update(r)
    assert res == alone
    assert sched.hbm_in_use == 0 and not sched.held
    assert peak <= 2.5 * UNIT


def test_slot_freed_mid_run():
    """A bucket full by k-points takes the next job of its key when a slot
    frees; a job of another key is not held back by the full bucket."""
    key_a, key_b = ((24, 24, 24), 32), ((30, 30, 30), 32)
    jobs = [SimJob(0, key=key_a, K=2, steps=2), SimJob(1, key=key_a, K=2, steps=5),
            SimJob(2, key=key_a, K=2, steps=2), SimJob(3, key=key_b, K=2, steps=2)]
    res, adm, fin, _ = simulate(jobs, make(hbm_budget=100 * UNIT, max_k_per_bucket=4))
    assert adm[0] == 0 and adm[1] == 0 and adm[3] == 0
    assert adm[2] == fin[0] == 2  # the slot job 0 frees, in the same tick
    assert set(res) == {0, 1, 2, 3}


def test_bucket_byte_cap():
    """The per-bucket state cap defers like the k-point limit."""
    jobs = [SimJob(i, K=4, n_tot=1000, steps=2) for i in range(3)]
    one = 4 * (2 * 32) * 1000 * 8 * 4  # X, HX and buffers of one job
    res, adm, _, _ = simulate(jobs, make(hbm_budget=100 * UNIT, max_bucket_bytes=2 * one))
    assert adm == {0: 0, 1: 0, 2: 2}
    assert set(res) == {0, 1, 2}


def test_groups_balance_new_keys():
    sched = Scheduler(100 * UNIT, 2048, 1e18, 8, n_groups=2)
    keys = [((24, 24, 24), 32), ((30, 30, 30), 32), ((36, 36, 36), 32)]
    sched.submit([SimJob(i, key=k) for i, k in enumerate(keys)])
    loads = [0.0, 0.0]
    taken0 = sched.admit(0.0, 0, {}, 0, lambda: loads)
    taken1 = sched.admit(0.0, 1, {}, 0, lambda: loads)
    assert set(taken0) | set(taken1) == set(keys)
    assert set(taken0).isdisjoint(taken1) and taken0 and taken1
    assert not sched.queue


@pytest.mark.parametrize("reserved", [0, int(2.0 * UNIT)])
def test_live_bytes_count(reserved):
    """Bytes already held by live tensors count against the budget."""
    sched = make()
    sched.submit([SimJob(0), SimJob(1), SimJob(2)])
    sched.hbm_in_use = 1  # something resident
    taken = sched.admit(0.0, 0, {}, reserved, lambda: [0.0])
    n = sum(len(v) for v in taken.values())
    assert n == (2 if reserved == 0 else 0)
