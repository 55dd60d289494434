"""Serving loop: concurrent CG / MD jobs, admission by slots and HBM footprint.

Each job's result must be the same whether it runs alone or in the mix
with other resident jobs (within the mixed-precision tolerance, 1e-6
relative on the final energy), and the jobs resident at any time must fit
the device-memory budget.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = "gb=5, cut=20.0, rgbmax=20.0, nsnb=10, dt=0.001, gamma_ln=5.0, temp0=300, tempi=300"


def _jobs(na, spec):
    jobs = []
    for k, (natom, kind, nsteps, t) in enumerate(spec):
        p, a, x = na.build_system(natom, 900 + k)
        jobs.append(dict(prm=p, x=x, kind=kind, nsteps=nsteps, seed=31 + k, arrival_s=t))
    return jobs


def test_alone_equals_in_the_mix(gpu):
    """MD jobs run 40 steps: Langevin dynamics amplifies the last-bit
    differences of FP64 atomic accumulation order (run-to-run spread of one
    job alone: ~1e-13 relative at 50 steps, ~2e-7 at 100), so a longer run
    would measure chaos rather than cross-talk between resident jobs."""
    na = gpu
    na.lib.mm_reset_options()
    spec = [(300, na.GBSERVE_CG, 30, 0.0), (1200, na.GBSERVE_MD, 40, 0.0), (800, na.GBSERVE_CG, 30, 0.01),
            (2000, na.GBSERVE_MD, 40, 0.02), (500, na.GBSERVE_MD, 40, 0.02), (1500, na.GBSERVE_CG, 30, 0.03)]
    jobs = _jobs(na, spec)
    mix = na.serve(jobs, slots=4, mm_opts=OPTS, check_fp64=True)
    assert all(r["status"] >= -3 for r in mix), [r["status"] for r in mix]
    assert max(r["max_active"] for r in mix) >= 2  # work really overlapped
    for k, j in enumerate(jobs):
        j2 = dict(j, arrival_s=0.0)
        alone = na.serve([j2], slots=1, mm_opts=OPTS)[0]
        assert alone["status"] == mix[k]["status"]
        assert abs(alone["energy"] - mix[k]["energy"]) <= 1e-6 * abs(mix[k]["energy"]), (k, alone, mix[k])
    for r in mix:
        # mixed vs FP64 at the same final coordinates
        assert abs(r["energy_final_eval"] - r["energy_fp64"]) <= 1e-6 * abs(r["energy_fp64"])


def test_admission_respects_hbm_budget(gpu):
    na = gpu
    na.lib.mm_reset_options()
    spec = [(3000, na.GBSERVE_MD, 60, 0.0), (2500, na.GBSERVE_MD, 60, 0.0), (2800, na.GBSERVE_MD, 60, 0.0),
            (400, na.GBSERVE_CG, 10, 0.0), (3000, na.GBSERVE_MD, 60, 0.0)]
    jobs = _jobs(na, spec)
    fp = [na.lib.gbhip_footprint_bytes(j["prm"].contents.Natom, j["nsteps"] if j["kind"] else 0) for j in jobs]
    budget = int(max(fp) * 1.6)  # at most one large job at a time (plus the small one)
    res = na.serve(jobs, slots=4, hbm_budget=budget, mm_opts=OPTS)
    assert all(r["status"] >= -3 for r in res)
    events = sorted([(r["start_s"], fp[k]) for k, r in enumerate(res)] +
                    [(r["end_s"], -fp[k]) for k, r in enumerate(res)], key=lambda e: (e[0], e[1]))
    used = 0
    for t, d in events:
        used += d
        assert used <= budget
    # a job that can never fit is rejected, not run
    res = na.serve(jobs[:1], slots=2, hbm_budget=fp[0] // 2, mm_opts=OPTS)
    assert res[0]["status"] == -2
