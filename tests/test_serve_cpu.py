"""The serving loop's admission and completion logic on the CPU, without a
device: gbserve_run_custom runs the same dispatcher and worker threads as
gbserve_run, with the work of a job done by a Python function that sleeps.
Checks, under a request rate well above what the workers drain without
waiting, that every job completes, that no more jobs than slots and no
more predicted device bytes than the budget are resident at once, that
jobs are admitted in arrival order (head of line waits for memory), that
a job larger than the budget is rejected and that runner failures come back
as job statuses."""
import os
import threading
import time

import numpy as np
import pytest

import nabc_amd as na

NSTEPS = 10


def _parms(gen_dir):
    out = []
    for n in (300, 1200, 4000):
        top = os.path.join(gen_dir, "srv_%d.prmtop" % n)
        rst = os.path.join(gen_dir, "srv_%d.rst7" % n)
        if not os.path.exists(top):
            na.generate(n, 17 + n, top, rst)
        out.append(na.read_parm(top))
    return out


def _footprint(p, kind):
    return na.lib.gbhip_footprint_bytes(p.contents.Natom, NSTEPS if kind == na.GBSERVE_MD else 0)


def _jobs(parms, njobs, rate, seed):
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, njobs))
    sizes = rng.integers(0, len(parms), njobs)
    kinds = rng.integers(0, 2, njobs)
    return [dict(prm=parms[s], kind=int(kd), nsteps=NSTEPS, arrival_s=float(t))
            for s, kd, t in zip(sizes, kinds, arrivals)]


class _Runner:
    """Sleeps per job; records residency as the runner sees it."""

    def __init__(self, jobs, durations, fail_every=0):
        self.jobs, self.durations, self.fail_every = jobs, durations, fail_every
        self.lock = threading.Lock()
        self.active = self.mem = self.max_active = self.max_mem = 0
        self.fp = [_footprint(j["prm"], j["kind"]) for j in jobs]

    def __call__(self, k, slot):
        with self.lock:
            self.active += 1
            self.mem += self.fp[k]
            self.max_active = max(self.max_active, self.active)
            self.max_mem = max(self.max_mem, self.mem)
        time.sleep(self.durations[k])
        with self.lock:
            self.active -= 1
            self.mem -= self.fp[k]
        if self.fail_every and k % self.fail_every == 0:
            return -5, 0.0
        return 0, float(k)


@pytest.mark.timeout(120)
def test_dispatcher_soak_slots_memory_order(gen_dir):
    parms = _parms(gen_dir)
    slots = 8
    jobs = _jobs(parms, 2000, rate=4000.0, seed=3)
    big = max(_footprint(p, na.GBSERVE_MD) for p in parms)
    budget = int(big * 3.5)  # at most three of the largest jobs at once
    rng = np.random.default_rng(4)
    run = _Runner(jobs, rng.uniform(0.0, 0.004, len(jobs)))
    t0 = time.perf_counter()
    res = na.serve_custom(jobs, run, slots=slots, hbm_budget=budget)
    wall = time.perf_counter() - t0
    assert wall < 60.0
    assert run.max_active <= slots and run.max_mem <= budget
    assert run.max_active >= 2  # the load overlapped jobs
    for k, r in enumerate(res):
        assert r["status"] == 0 and r["energy"] == float(k)
        assert 0 <= r["slot"] < slots and 1 <= r["max_active"] <= slots
        assert r["footprint"] == run.fp[k]
        assert r["start_s"] >= jobs[k]["arrival_s"] - 1e-3
        assert r["end_s"] >= r["start_s"] and r["latency_s"] > 0.0
    starts = np.array([r["start_s"] for r in res])
    assert np.all(np.diff(starts) >= -1e-9), "admission out of arrival order"


@pytest.mark.timeout(60)
def test_dispatcher_rejects_oversized_and_reports_failures(gen_dir):
    parms = _parms(gen_dir)
    jobs = _jobs(parms, 200, rate=2000.0, seed=5)
    small = _footprint(parms[0], na.GBSERVE_MD)
    large = _footprint(parms[2], na.GBSERVE_CG)
    budget = (small + large) // 2  # the 4,000-atom jobs can never fit
    run = _Runner(jobs, np.full(len(jobs), 0.001), fail_every=7)
    res = na.serve_custom(jobs, run, slots=4, hbm_budget=budget)
    for k, (j, r) in enumerate(zip(jobs, res)):
        if run.fp[k] > budget:
            assert r["status"] == -2 and r["slot"] == -1
        elif k % 7 == 0:
            assert r["status"] == -5
        else:
            assert r["status"] == 0 and r["energy"] == float(k)
    assert any(r["status"] == -2 for r in res) and run.max_mem <= budget


@pytest.mark.timeout(60)
def test_pool_blocks_of_resident_jobs_never_overlap(gen_dir):
    """Each admitted job holds a contiguous block of its footprint in the
    device pool (first fit, 4 KiB granules); the blocks of jobs resident at
    the same time are disjoint and inside the pool."""
    parms = _parms(gen_dir)
    jobs = _jobs(parms, 400, rate=3000.0, seed=9)
    big = max(_footprint(p, na.GBSERVE_MD) for p in parms)
    budget = int(big * 2.5)
    rng = np.random.default_rng(10)
    run = _Runner(jobs, rng.uniform(0.0, 0.003, len(jobs)))
    res = na.serve_custom(jobs, run, slots=6, hbm_budget=budget)
    rnd = lambda b: (b + 4095) // 4096 * 4096
    iv = [(r["start_s"], r["end_s"], r["pool_offset"], r["pool_offset"] + rnd(r["footprint"])) for r in res]
    for a0, a1, o0, o1 in iv:
        assert 0 <= o0 < o1 <= budget
    for i in range(len(iv)):
        for j in range(i + 1, len(iv)):
            (s0, e0, p0, q0), (s1, e1, p1, q1) = iv[i], iv[j]
            if s0 < e1 and s1 < e0:  # resident together
                assert q0 <= p1 or q1 <= p0, (i, j, iv[i], iv[j])
