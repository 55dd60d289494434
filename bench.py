"""Benchmark: GB energy + gradient evaluations/s at 25,000 atoms, mixed vs FP64.

Workload (one "step" = one generalized-Born energy + gradient evaluation,
egb() of the reference, with inputs resident in HBM): a seeded synthetic
25,000-atom polymer (0.1 atoms/A^3), igb=8 (GB-neck), 25 A cutoff,
rgbmax 25 A, Debye-Hueckel kappa 0.127315 1/A (0.15 M salt), pair list
rebuilt every 20 evaluations (nsnb=20), as in the reference's mme34 cadence.

Reported: mixed-precision evals/s (value), the plain FP64 evals/s beside it,
per-pass kernel times from HIP events on the engine stream, the roofline
fraction of the dominant kernel, and the oracle (OpenMP FP64 CPU
restatement) timed on this host on a bounded sample.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import hashlib
import json
import os
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

NATOM = 25000
SEED = 2025
GB = 8
CUT = 25.0
RGBMAX = 25.0
KAPPA = 0.127315
NSNB = 20
# Algorithmic flops per listed (unordered) pair, from the reference's
# expressions on the most common (Taylor) branch with +,-,x = 1, FMA = 2 and
# each sqrt/1/x/exp/log = 1: pass 1 F1 = 24 per ordered pair (+10 neck),
# pass 2 F2 = 68 per unordered pair, pass 3 F3 = 37 per ordered pair (+15
# neck), i.e. W = 240 per pair for gb 7/8.  Split along this build's kernel
# boundaries: k_born does pass 1 and the datmp part of pass 3 (2 x (28 +
# 15)), k_chain_st the pass-3 force application (2 x 9), k_pairs pass 2.
FLOPS = {"born": 2 * (24 + 10) + 2 * (28 + 15), "pairs": 68.0, "chain": 2 * 9.0}
W_PER_PAIR = 240.0
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X vector FP64 / FP32
# HBM bytes per launch of the pair kernels at this workload, measured with
# rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE and scaled by the copy-kernel
# calibration (profiles/gb_25k_igb8_hbm_traffic.md)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "gb_25k_igb8_traffic.json")


def kernel_name(pass_, real, split):
    k = ("born", "pairs", "chain").index(pass_)
    sp = "true" if split[k] > 1 else "false"
    if pass_ == "born" and real == "double":
        # FP64 pass 1 is the queued variant
        return "k_born_q<double, %s>" % sp
    base = ("k_born", "k_pairs", "k_chain_st")[k]
    return "%s<%s, %s>" % (base, real, sp)


def profiled(kernel, key):
    """A figure of the committed profile summary for this workload."""
    try:
        with open(TRAFFIC_FILE) as fp:
            return json.load(fp)["kernels"].get(kernel, {}).get(key)
    except (OSError, ValueError, KeyError):
        return None


def measured_traffic(kernel):
    return profiled(kernel, "bytes_per_launch")


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as fp:
        h.update(fp.read())
    return h.hexdigest()


def run_gpu(na, arr, x, precision, steps, warmup, tune=True, fixed_launch=None, det=False):
    ctx = na.GBContext(arr, x, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA, precision=precision)
    if det:
        ctx.set_deterministic(1)
    ctx.nblist()
    npairs = ctx.npairs()
    ntiles, ntp = ctx.tile_stats()
    # launch-shape autotuning is set-up, outside the timed region
    if fixed_launch:
        ctx.set_launch(*fixed_launch)
        launch = ctx.launch()
    else:
        launch = ctx.autotune(5) if tune else ctx.launch()
    store_need, store_cap, store_used = ctx.store_stats()
    split = ctx.split()
    for k in range(warmup):
        if k % NSNB == 0:
            ctx.nblist_async()
        ctx.egb_device()
    ctx.sync()
    ctx.timing_start(steps)
    t0 = time.perf_counter()
    for k in range(steps):
        if k % NSNB == 0:
            ctx.nblist_async()
        ctx.egb_device()
    ctx.sync()
    t1 = time.perf_counter()
    ms, n = ctx.timing_stop()
    # separate timing of the list rebuild
    ctx.sync()
    tl0 = time.perf_counter()
    for _ in range(5):
        ctx.nblist_async()
    ctx.sync()
    tl = (time.perf_counter() - tl0) / 5
    epol, evdw, eel, g = ctx.egb()
    ctx.close()
    return dict(evals_per_s=steps / (t1 - t0), ms_per_eval=1e3 * (t1 - t0) / steps,
                pass_ms=dict(born=ms[0] / n, radii=ms[1] / n, pairs=ms[2] / n, chain=ms[3] / n),
                nblist_ms=1e3 * tl, npairs=npairs, ntiles=ntiles, tilepairs=ntp, launch=launch, split=split,
                store=dict(entries=store_need, allocated=store_cap, in_use=store_used),
                energies=dict(epol=epol, evdw=evdw, eelec=eel), grad=g)


def cpu_baseline(parm, natom, x, budget_s=15.0):
    import oracle
    # every CPU this process may run on, unless the environment allots fewer
    # OpenMP threads (a share of a shared host)
    visible = len(os.sched_getaffinity(0))
    threads = visible
    omp = os.environ.get("OMP_NUM_THREADS", "")
    if omp.isdigit() and 0 < int(omp) < visible:
        threads = int(omp)
    oracle.set_threads(threads)
    o = oracle.Oracle(parm, natom, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA)
    t0 = time.perf_counter()
    o.nblist(x)
    t_list = time.perf_counter() - t0
    n, t_eval = 0, 0.0
    ref = None
    while t_eval < budget_s and n < 200:
        t0 = time.perf_counter()
        ref = o.egb(x)
        t_eval += time.perf_counter() - t0
        n += 1
    per_eval = t_eval / n + t_list / NSNB
    o.close()
    cpu = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return dict(value=1.0 / per_eval, unit="evals/s", cores=threads, kind="port",
                host_cpus={"affinity": visible, "os_cpu_count": os.cpu_count(),
                           "omp_num_threads": omp or None, "model": cpu},
                sample="%d oracle egb() evaluations + 1 pair-list build (amortised over nsnb=%d) "
                       "of the same 25,000-atom igb=8 system; OpenMP, %d threads (%d CPUs in this "
                       "process's affinity, OMP_NUM_THREADS=%s), %s" %
                       (n, NSNB, threads, visible, omp or "unset", cpu or "cpu model unknown")), ref


def bench_batch(args):
    """Batched scoring: 1,024 molecules, N log-uniform on [200, 3000], seeded
    per instance, igb=5, no cutoff, one evaluation per molecule in one launch
    per pass; every instance checked against the oracle."""
    import numpy as np
    import nabc_amd as na
    import oracle
    nmol = args.nmol
    rng = np.random.default_rng(4242)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(200), np.log(3000), nmol))]
    t0 = time.perf_counter()
    systems = [na.build_system(n, 100000 + k) for k, n in enumerate(sizes)]
    t_gen = time.perf_counter() - t0
    res = {}
    for name, prec in (("fp64", na.GBHIP_FP64), ("mixed", na.GBHIP_MIXED)):
        b = na.Batch([(a, x) for _, a, x in systems], gb=5, precision=prec)
        npairs, natoms, ntp = b.stats()
        launch = b.autotune(3) if not args.no_tune else None
        store_used = b.store_stats()[2]
        for _ in range(args.warmup):
            b.eval_device()
        ene, grads = b.eval()
        b.timing_start(args.steps)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            b.eval_device()
        ene2, _ = b.eval()
        t1 = time.perf_counter()
        ms, n = b.timing_stop()
        dt = (t1 - t0) / (args.steps + 1)
        res[name] = dict(batch_ms=1e3 * dt, molecules_per_s=nmol / dt, pairs_per_s=npairs / dt,
                         pass_ms={k: round(v / n, 4) for k, v in zip(("born", "radii", "pairs", "chain"), ms)},
                         ene=ene, grads=grads, npairs=npairs, natoms=natoms, tilepairs=ntp, launch=launch,
                         store=store_used)
        b.close()
    # per-instance check against the oracle
    worst = dict(fp64_e=0.0, fp64_g=0.0, mixed_e=0.0, mixed_g=0.0)
    t0 = time.perf_counter()
    for k, (p, a, x) in enumerate(systems):
        o = oracle.Oracle(p, a["natom"], gb=5, cut=999.0, rgbmax=999.0)
        o.nblist(x)
        r = o.egb(x)
        for name in ("fp64", "mixed"):
            e = res[name]["ene"][k]
            g = res[name]["grads"][k]
            worst[name + "_e"] = max(worst[name + "_e"], max(abs(e[c] - r[c]) / abs(r[c]) for c in range(3)))
            worst[name + "_g"] = max(worst[name + "_g"],
                                     float(np.sqrt(np.mean((g - r[3]) ** 2) / np.mean(r[3] ** 2))))
        o.close()
    t_check = time.perf_counter() - t0
    out = {"metric": "batched GB energy+grad, molecules/s (1,024 molecules, N log-uniform 200-3000)",
           "value": round(res["mixed"]["molecules_per_s"], 1), "unit": "molecules/s",
           "fp64_molecules_per_s": round(res["fp64"]["molecules_per_s"], 1),
           "mixed_vs_fp64": round(res["mixed"]["molecules_per_s"] / res["fp64"]["molecules_per_s"], 3),
           "batch_ms": {k: round(res[k]["batch_ms"], 3) for k in res},
           "pair_interactions_per_s": {k: float("%.4g" % res[k]["pairs_per_s"]) for k in res},
           "pass_ms": {k: res[k]["pass_ms"] for k in res},
           "npairs": res["mixed"]["npairs"], "natoms": res["mixed"]["natoms"],
           "tile_pairs": res["mixed"]["tilepairs"], "nmol": nmol,
           "launch_blocks": {k: res[k]["launch"] for k in res},
           "derivative_store_in_use": {k: res[k]["store"] for k in res},
           "worst_instance_error_vs_oracle": worst, "generate_s": round(t_gen, 1),
           "oracle_check_s": round(t_check, 1)}
    print(json.dumps(out))


def bench_accuracy(args):
    """Per-term accuracy of the full mme() (bonded + 1-4 + GB) through the
    host API, FP64 and mixed kernels, against the oracle's mme34 on the same
    coordinates and list: C1 (600 atoms, igb=5, no cutoff), C2 (2,500,
    igb=5, no cutoff), C3 (25,000, igb=8, 25 A, 0.15 M salt)."""
    import numpy as np
    import nabc_amd as na
    import oracle
    oracle.set_threads(min(16, os.cpu_count() or 1))
    names = {0: "total", 1: "enb", 2: "eel", 3: "bond", 4: "angle", 5: "dihedral", 7: "enb14",
             8: "eel14", 10: "egb"}
    cfgs = {"c1": (600, 5, 999.0, 999.0, 0.0), "c2": (2500, 5, 999.0, 999.0, 0.0),
            "c3": (25000, 8, 25.0, 25.0, 0.127315)}
    out = {"metric": "mme() per-term relative error vs the oracle (FP64 mme34)", "unit": "relative",
           "configs": {}}
    for cfg, (natom, gb, cut, rgbmax, kappa) in cfgs.items():
        parm, arr, x = na.build_system(natom, SEED)
        x = np.ascontiguousarray(x)
        orc = oracle.Oracle(parm, natom, gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa)
        r_ene, r_frms, r_g = orc.mme(x)
        orc.close()
        res = {}
        for prec in ("fp64", "mixed"):
            na.lib.mm_reset_options()
            assert na.lib.mm_options(("gb=%d, cut=%g, rgbmax=%g, kappa=%g, diel=C, gbprec=%s, ntpr=100000000" % (
                gb, cut, rgbmax, kappa, prec)).encode()) == 0
            assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
            f = np.zeros(3 * natom)
            it = na.C.c_int(-2)
            na.lib.mme(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
            ene = np.zeros(13)
            frms = na.C.c_double()
            na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
            # relative to the term, with a 1 kcal/mol floor (the generated
            # bonds sit at Req, so the bond energy is ~1e-26)
            terms = {names[k]: abs(ene[k] - r_ene[k]) / max(abs(r_ene[k]), 1.0) for k in names}
            res[prec] = {"terms": terms,
                         "grad_rel_rms": float(np.sqrt(np.mean((f - r_g) ** 2) / np.mean(r_g ** 2))),
                         "grad_max_component_abs": float(np.max(np.abs(f - r_g))),
                         "grad_max_component_rel_to_max": float(np.max(np.abs(f - r_g)) / np.max(np.abs(r_g))),
                         "frms_rel": abs(frms.value - r_frms) / r_frms}
            it = na.C.c_int(-3)
            na.lib.mme(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
        out["configs"][cfg] = {"natom": natom, "gb": gb, "cut": cut, "kappa": kappa,
                               "oracle_terms": {names[k]: float(r_ene[k]) for k in names}, **res}
    worst = max(v for c in out["configs"].values() for p in ("fp64", "mixed") for v in c[p]["terms"].values())
    out["value"] = worst
    out["worst_mixed_grad_rel_rms"] = max(c["mixed"]["grad_rel_rms"] for c in out["configs"].values())
    print(json.dumps(out))


_MME_FN = None


def bench_md(args):
    """Dynamics through the host API: md() over mme() with every step a
    hipGraph launch (bonded + 1-4 + GB on the device)."""
    import numpy as np
    import nabc_amd as na
    global _MME_FN
    _MME_FN = na.MMEFUNC(lambda xp, fp, itp: na.lib.mme(xp, fp, itp))
    if args.md_steps <= 0:
        args.md_steps = {"c1": 1000, "c2": 10000, "c3": 5000}[args.md]
    natom, opts = {"c1": (600, "gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, nsnb=10"),
                   "c2": (2500, "gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, nsnb=10"),
                   "c3": (25000, "gb=8, cut=25.0, rgbmax=25.0, kappa=0.127315, nsnb=20")}[args.md]
    parm, arr, x0 = na.build_system(natom, SEED)
    # the generated chain is strained (about 9 kcal/mol/atom above its
    # relaxed energy), so dynamics starts from a conjugate-gradient minimum
    # found through the same library (conjgrad over the device mme)
    # the list is rebuilt at every call while minimising (pairs crossing the
    # cutoff would otherwise make the energy jump between line-search
    # points), and CG restarts after a failed line search
    na.lib.mm_reset_options()
    assert na.lib.mm_options(("%s, nsnb=1, ntpr=100000000" % opts).encode()) == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    x0 = np.ascontiguousarray(x0.copy())
    n3, fmin = na.C.c_int(3 * natom), na.C.c_double()
    t0 = time.perf_counter()
    left, cg_ret, runs = args.md_min, 0, 0
    while left > 0 and runs < 5:
        rms, dfp, mx = na.C.c_double(0.1), na.C.c_double(10.0), na.C.c_int(left)
        cg_ret = na.lib.conjgrad(x0.ctypes.data_as(na.dp), na.C.byref(n3), na.C.byref(fmin), _MME_FN,
                                 na.C.byref(rms), na.C.byref(dfp), na.C.byref(mx))
        runs += 1
        if cg_ret != -1:
            break
        left -= 50
    t_min = time.perf_counter() - t0
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
    out = {"metric": "MD steps/s (%s, %d atoms)" % (args.md, natom), "unit": "steps/s",
           "step_launch": "stream" if args.md_no_graph else "hipGraph",
           "start": {"cg_iterations_max": args.md_min, "cg_return": cg_ret, "cg_runs": runs,
                     "energy": fmin.value, "seconds": round(t_min, 2)}}
    for name, prec in (("fp64", "fp64"), ("mixed", "mixed")):
        na.lib.mm_reset_options()
        assert na.lib.mm_options(("%s, gbprec=%s, dt=0.001, temp0=300, tempi=300, gamma_ln=5.0, ntpr=1000000, "
                                  "ntpr_md=1000000" % (opts, prec)).encode()) == 0
        assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
        ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
        if not args.no_tune:
            # launch shapes tuned on the starting list, before the step graphs are captured
            assert na.lib.gbhip_set_coords(ctx, x0.ctypes.data_as(na.dp)) == 0
            assert na.lib.gbhip_nblist(ctx, None) == 0
            assert na.lib.gbhip_autotune(ctx, 3, None) == 0
        masses = arr["mass"]
        o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                         boltz2=9.93595e-4, nsnb=int(opts.split("nsnb=")[1]), zerov=0, seed=SEED)
        assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                    na.C.byref(o), args.md_steps) == 0
        use_graph = 0 if args.md_no_graph else 1
        assert na.lib.gbhip_md_run(ctx, 20, use_graph) == 0
        assert na.lib.gbhip_md_state(ctx, None, None, None) == 0
        t0 = time.perf_counter()
        assert na.lib.gbhip_md_run(ctx, args.md_steps, use_graph) == 0
        assert na.lib.gbhip_md_state(ctx, None, None, None) == 0
        t1 = time.perf_counter()
        rows = np.zeros((args.md_steps, 16))
        nr = na.C.c_int()
        na.lib.gbhip_md_log(ctx, rows.ctypes.data_as(na.dp), args.md_steps, na.C.byref(nr))
        out[name] = {"steps_per_s": round(args.md_steps / (t1 - t0), 1),
                     "ms_per_step": round(1e3 * (t1 - t0) / args.md_steps, 4),
                     "ns_per_day": round(args.md_steps / (t1 - t0) * 0.001 * 86400 / 1000, 2),
                     "final_temp": round(float(rows[-200:, 15].mean()), 1),
                     "final_epot": float(rows[-1, 0])}
        it = na.C.c_int(-3)
        na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
    out["value"] = out["mixed"]["steps_per_s"]
    print(json.dumps(out))


def bench_serve(args):
    """Serving: Poisson arrivals at a fixed rate of jobs on generated
    molecules of 300-8000 atoms (log-uniform), each 50 CG iterations or 500
    MD steps (Bernoulli 1/2), igb=5, 20 A cutoff; continuous admission by
    free slot and HBM footprint.  Reports throughput, p50/p99 latency and the
    mixed-vs-FP64 energy error at each job's final coordinates."""
    import numpy as np
    import nabc_amd as na
    rng = np.random.default_rng(55)
    n = args.jobs
    sizes = np.exp(rng.uniform(np.log(300), np.log(8000), n)).astype(int)
    kinds = rng.integers(0, 2, n)
    arrivals = np.cumsum(rng.exponential(1.0 / args.rate, n))
    t0 = time.perf_counter()
    jobs = []
    for k in range(n):
        p, a, x = na.build_system(int(sizes[k]), 700000 + k)
        jobs.append(dict(prm=p, x=x, kind=int(kinds[k]), nsteps=50 if kinds[k] == 0 else 500,
                         seed=k + 1, arrival_s=float(arrivals[k])))
    t_gen = time.perf_counter() - t0
    opts = ("gb=5, cut=20.0, rgbmax=20.0, nsnb=10, dt=0.001, gamma_ln=5.0, temp0=300, tempi=300, "
            "ntpr=100000000")
    res = na.serve(jobs, slots=args.slots, hbm_budget=int(args.hbm_gb * (1 << 30)), precision=na.GBHIP_MIXED,
                   check_fp64=True, mm_opts=opts, cg_rmsgrad=1e-4, cg_dfpred=10.0, progress=20)
    ok = [r for r in res if r["status"] >= -4]  # conjgrad's -1..-4 are ordinary endings
    lat = np.array([r["latency_s"] for r in ok])
    span = max(r["end_s"] for r in ok) - min(j["arrival_s"] for j in jobs)
    err = np.array([abs(r["energy_final_eval"] - r["energy_fp64"]) / abs(r["energy_fp64"]) for r in ok
                    if r["energy_fp64"] != 0.0])
    lat_cg = np.array([r["latency_s"] for r, j in zip(res, jobs) if j["kind"] == 0 and r["status"] >= -4])
    lat_md = np.array([r["latency_s"] for r, j in zip(res, jobs) if j["kind"] == 1 and r["status"] >= -4])
    out = {"metric": "serving p99 latency at a fixed request rate (CG/MD jobs, 300-8000 atoms)",
           "value": round(float(np.percentile(lat, 99)), 4), "unit": "s", "higher_is_better": False,
           "request_rate_per_s": args.rate, "jobs": n, "completed": len(ok), "slots": args.slots,
           "hbm_budget_gb": args.hbm_gb, "throughput_jobs_per_s": round(len(ok) / span, 2),
           "p50_latency_s": round(float(np.percentile(lat, 50)), 4),
           "p99_latency_s": round(float(np.percentile(lat, 99)), 4),
           "p50_cg_s": round(float(np.percentile(lat_cg, 50)), 4) if len(lat_cg) else None,
           "p50_md_s": round(float(np.percentile(lat_md, 50)), 4) if len(lat_md) else None,
           "max_resident_jobs": max(r["max_active"] for r in res),
           "p50_setup_ms": round(1e3 * float(np.median([r["setup_s"] for r in ok])), 3),
           "mixed_vs_fp64_final_energy_rel_err": {"median": float(np.median(err)), "max": float(err.max())},
           "atoms_mean": float(sizes.mean()), "generate_s": round(t_gen, 1),
           "job_mix": "Bernoulli(1/2) of CG maxiter=50 / MD 500 steps, seed 55"}
    print(json.dumps(out))


def bench_cg(args):
    """Minimiser: conjgrad() over mme() with its vectors on the device
    (cgdevice=1, gbhip_conjgrad) and on the host (cgdevice=0, the
    reference's arrangement: coordinates in and gradient out every
    evaluation) at 600, 2,500 and 8,000 atoms (gb=5, cut = rgbmax = 20 A,
    nsnb 10, mixed precision), maxiter iterations from the generated start;
    wall time per evaluation beside the device time of one mme() evaluation
    (captured graph, back to back)."""
    import numpy as np
    import nabc_amd as na
    maxiter = args.cg_iters
    rows = []
    for natom in (600, 2500, 8000):
        parm, arr, x0 = na.build_system(natom, 4242 + natom)
        row = {"natom": natom}
        for dev in (1, 0):
            na.lib.mm_reset_options()
            opts = "gb=5, cut=20.0, rgbmax=20.0, nsnb=10, ntpr=100000000, gbprec=mixed, cgdevice=%d" % dev
            assert na.lib.mm_options(opts.encode()) == 0
            assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
            best = None
            for rep in range(3):  # the first run captures graphs and stages memory
                x = np.ascontiguousarray(x0.copy())
                n = na.C.c_int(3 * natom)
                f = na.C.c_double()
                rms = na.C.c_double(1e-12)
                dfp = na.C.c_double(10.0)
                mx = na.C.c_int(maxiter)
                t0 = time.perf_counter()
                ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), na.MME_C,
                                      na.C.byref(rms), na.C.byref(dfp), na.C.byref(mx))
                dt = time.perf_counter() - t0
                if rep > 0 and (best is None or dt < best[0]):
                    best = (dt, ret, f.value)
            nev = maxiter + 1 if best[1] == -3 else abs(best[1])
            row["device" if dev else "host"] = {"us_per_eval": round(1e6 * best[0] / nev, 1), "return": best[1],
                                                "final_energy": best[2]}
            if dev:
                # device time of one evaluation: back-to-back captured graphs
                ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
                for _ in range(3):
                    na.lib.gbhip_mme(ctx, None, None)
                na.lib.gbhip_sync(ctx)
                t0 = time.perf_counter()
                for _ in range(50):
                    na.lib.gbhip_mme(ctx, None, None)
                row["mme_us_with_wait"] = round(1e6 * (time.perf_counter() - t0) / 50, 1)
            it = na.C.c_int(-3)
            na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
        row["host_over_device"] = round(row["host"]["us_per_eval"] / row["device"]["us_per_eval"], 3)
        rows.append(row)
        print("# %s" % json.dumps(row), file=sys.stderr, flush=True)
    print(json.dumps({"metric": "conjgrad wall time per evaluation, device vs host vectors",
                      "unit": "us/evaluation", "iterations": maxiter, "rows": rows}))


def bench_hessian(args):
    """Second derivatives: mme2() (the device Hessian of the full mme energy
    by central differences of the FP64 gradient, 6N device evaluations)
    and nmode() (mass weighting, the eigenproblem on the GPU and the
    thermochemistry) at 300, 1,000 and 2,500 atoms (gb=5, cut = rgbmax = 20
    A), after one untimed mme2 call that creates the context."""
    import ctypes
    import numpy as np
    import nabc_amd as na
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    na.lib.mm_set_output.argtypes = [ctypes.c_void_p]
    rows = []
    for natom in (300, 1000, 2500):
        parm, arr, x = na.build_system(natom, 5150 + natom)
        x = np.ascontiguousarray(x)
        na.lib.mm_reset_options()
        assert na.lib.mm_options(b"gb=5, cut=20.0, rgbmax=20.0, ntpr=100000000") == 0
        assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
        n3 = 3 * natom
        f, H, m = np.zeros(n3), np.zeros(n3 * n3), np.zeros(natom)
        name = ctypes.create_string_buffer(natom + 1)
        nat, it = na.C.c_int(), na.C.c_int(-2)
        dummy = [na.C.c_int(0) for _ in range(7)]

        def mme2():
            return na.lib.mme2(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), H.ctypes.data_as(na.dp),
                               m.ctypes.data_as(na.dp), None, *[na.C.byref(d) for d in dummy], na.C.byref(nat),
                               na.C.byref(it), name)
        mme2()
        t0 = time.perf_counter()
        e = mme2()
        t_h = time.perf_counter() - t0
        fp = libc.fopen(b"/dev/null", b"w")
        na.lib.mm_set_output(fp)
        t0 = time.perf_counter()
        ret = na.lib.nmode(x.ctypes.data_as(na.dp), n3, na.MME2_C, 0, 1, 0.0, 0.0, 0)
        t_nm = time.perf_counter() - t0
        na.lib.mm_set_output(None)
        libc.fclose(fp)
        itr = na.C.c_int(-3)
        na.lib.mme(x.ctypes.data_as(na.dp), x.copy().ctypes.data_as(na.dp), na.C.byref(itr))
        row = {"natom": natom, "hessian_s": round(t_h, 4), "evaluations": 2 * n3 + 1,
               "us_per_evaluation": round(1e6 * t_h / (2 * n3 + 1), 1), "nmode_s": round(t_nm, 4),
               "nmode_return": ret, "energy": e}
        rows.append(row)
        print("# %s" % json.dumps(row), file=sys.stderr, flush=True)
    print(json.dumps({"metric": "second derivatives: mme2 Hessian and nmode wall time", "unit": "s",
                      "rows": rows}))


# hierarchical charge partitioning workload: a generated polymer in 20-atom
# residues, strands of 50 residues, complexes of 10 strands, the sff.c
# default thresholds, gb=5, 0.1 M salt
HCP_NATOM, HCP_SEED = 100000, 7
HCP_H = (15.0, 50.0, 150.0)
HCP_KAPPA = 0.10395
# algorithmic flops per unit of the hierarchical walk, counted from the
# reference expressions (hcp_gb.c) as the pair path's are: a component
# distance check; per far charge and per near ordered atom pair in the
# Born pass, the energy pass (GB + Coulomb, + LJ for atom pairs) and the
# chain-rule pass
HCP_FLOPS = dict(check=8, born_far=23, born_near=24, energy_far=56, energy_near=68, deriv_far=34, deriv_near=37)


def hcp_layers(natom, res_atoms=20, strand_res=50, complex_strands=10):
    import numpy as np
    ipres = np.array(list(range(0, natom, res_atoms)) + [natom], np.int32)
    ips = np.arange(1, len(ipres), strand_res, dtype=np.int32)
    ipc = np.arange(1, len(ips) + 1, complex_strands, dtype=np.int32)
    return ipres, ips, ipc


def hcp_walk_counts(x, ipres, ips, ipc, h):
    """Units of one evaluation's hierarchical walk (as each atom walks it in
    hcp_gb.c): component distance checks, far components, near ordered
    atom pairs."""
    import numpy as np
    xyz = x.reshape(-1, 3)
    n, nres = len(xyz), len(ipres) - 1
    rfirst = np.append(ips - 1, nres)
    sfirst = np.append(ipc - 1, len(ips))
    natres = np.diff(ipres)
    res_of = np.repeat(np.arange(nres), natres)
    cres = np.add.reduceat(xyz, ipres[:-1], axis=0) / natres[:, None]
    cstr = np.array([xyz[ipres[rfirst[s]]:ipres[rfirst[s + 1]]].mean(0) for s in range(len(ips))])
    ccpx = np.array([xyz[ipres[rfirst[sfirst[c]]]:ipres[rfirst[sfirst[c + 1]]]].mean(0) for c in range(len(ipc))])
    s_of_r = np.repeat(np.arange(len(ips)), np.diff(rfirst))
    c_of_s = np.repeat(np.arange(len(ipc)), np.diff(sfirst))
    checks = far = near = 0
    for i0 in range(0, n, 2000):
        xi = xyz[i0:i0 + 2000]
        d3 = ((xi[:, None, :] - ccpx[None]) ** 2).sum(-1) > h[2] ** 2
        checks += d3.size
        far += int(d3.sum())
        in3 = ~d3[:, c_of_s]
        d2 = ((xi[:, None, :] - cstr[None]) ** 2).sum(-1) > h[1] ** 2
        checks += int(in3.sum())
        far += int((in3 & d2).sum())
        in2 = (in3 & ~d2)[:, s_of_r]
        d1 = ((xi[:, None, :] - cres[None]) ** 2).sum(-1) > h[0] ** 2
        checks += int(in2.sum())
        far += int((in2 & d1).sum())
        in1 = in2 & ~d1
        near += int((in1 * natres[None, :]).sum()) - int(in1[np.arange(len(xi)), res_of[i0:i0 + 2000]].sum())
    return checks, far, near


def bench_hcp(args):
    """Hierarchical charge partitioning (egb_hcp) at 100,000 atoms (the
    system size the reference offers it for): evaluations/s of hcp 1, 2 and
    4 in the mixed and the FP64 variant (HIP events around each pass, after
    warm-up), the roofline of the energy kernel (algorithmic flops of the
    walk per launch over its duration), errors against the oracle, and the
    oracle's egb_hcp on the host cores."""
    import numpy as np
    import nabc_amd as na
    parm, arr, x = na.build_system(HCP_NATOM, HCP_SEED)
    ipres, ips, ipc = hcp_layers(HCP_NATOM)
    steps = max(5, min(args.steps, 100))
    runs, grads = {}, {}
    for pname, prec in (("mixed", na.GBHIP_MIXED), ("fp64", na.GBHIP_FP64)):
        rows = {}
        for hcp in (1, 2, 4):
            ctx = na.GBContext(arr, x, gb=5, cut=999.0, rgbmax=999.0, kappa=HCP_KAPPA, precision=prec)
            ctx.set_hcp(hcp, ipres, ips, ipc, h1=HCP_H[0], h2=HCP_H[1], h3=HCP_H[2])
            ctx.nblist()
            for _ in range(3):
                ctx.egb_device()
            ctx.sync()
            ctx.timing_start(steps)
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.egb_device()
            ctx.sync()
            wall = (time.perf_counter() - t0) / steps
            ms, ne = ctx.timing_stop()
            e = ctx.egb()
            ctx.close()
            if hcp == 1:
                grads[pname] = e[3]
            rows[hcp] = dict(evals_per_s=round(1.0 / wall, 2), ms_per_eval=round(1e3 * wall, 4),
                             pass_ms=dict(zip(("charges_born", "component_radii", "energy", "deriv"),
                                              [round(v / ne, 4) for v in ms])),
                             energies=dict(epol=e[0], evdw=e[1], eelec=e[2]))
            print("# %s hcp %d: %s" % (pname, hcp, json.dumps(rows[hcp])), file=sys.stderr, flush=True)
        runs[pname] = rows
    checks, far, near = hcp_walk_counts(x, ipres, ips, ipc, HCP_H)
    F = HCP_FLOPS
    w_energy = checks * F["check"] + far * F["energy_far"] + near * F["energy_near"]
    w_eval = (3 * checks * F["check"] + far * (F["born_far"] + F["energy_far"] + F["deriv_far"]) +
              near * (F["born_near"] + F["energy_near"] + F["deriv_near"]))

    def roof(pname, kernel):
        t = runs[pname][1]["pass_ms"]["energy"] * 1e-3
        a = w_energy / t / 1e12
        return {"bound": "valu", "kernel": kernel, "kernel_ms": round(1e3 * t, 4), "achieved": round(a, 3),
                "peak": PEAK[pname], "unit": "TFLOP/s", "frac": round(a / PEAK[pname], 4),
                "eval_achieved": round(w_eval / (runs[pname][1]["ms_per_eval"] * 1e-3) / 1e12, 3)}

    m1, f1 = runs["mixed"][1], runs["fp64"][1]
    line = {"metric": "HCP GB energy+grad evals/s at 100k atoms", "value": m1["evals_per_s"],
            "unit": "evals/s", "n_gpus": 1, "steps": steps, "higher_is_better": True,
            "dtype": "f32-pair/f64-accumulate", "data": "synthetic (seeded 100,000-atom polymer, nabc_amd/csrc/gen.c)",
            "config": {"workload": "egb_hcp hcp 1, 100,000 atoms in 5,000 20-atom residues, 100 strands, "
                                   "10 complexes, thresholds 15/50/150 A, gb=5, kappa 0.10395",
                       "natom": HCP_NATOM, "seed": HCP_SEED},
            "mixed": {str(k): v for k, v in runs["mixed"].items()},
            "fp64": {str(k): v for k, v in runs["fp64"].items()},
            "mixed_vs_fp64": round(m1["evals_per_s"] / f1["evals_per_s"], 3),
            "walk": {"checks": checks, "far_components": far, "near_pairs": near,
                     "flops_per_eval_hcp1": w_eval, "flops_per_unit": F},
            "roofline": roof("mixed", "k_hcp_energy<float, true>"),
            "roofline_fp64": roof("fp64", "k_hcp_energy<double, true>")}
    if not args.no_cpu:
        import oracle
        visible = len(os.sched_getaffinity(0))
        threads = visible
        omp = os.environ.get("OMP_NUM_THREADS", "")
        if omp.isdigit() and 0 < int(omp) < visible:
            threads = int(omp)
        oracle.set_threads(threads)
        o = oracle.Oracle(parm, HCP_NATOM, gb=5, cut=999.0, rgbmax=999.0, kappa=HCP_KAPPA)
        o.set_hcp(1, ipres, ips, ipc, h1=HCP_H[0], h2=HCP_H[1], h3=HCP_H[2])
        n, t = 0, 0.0
        while t < 15.0 and n < 20:
            t0 = time.perf_counter()
            ref = o.egb(x)
            t += time.perf_counter() - t0
            n += 1
        o.close()
        line["cpu_baseline"] = {"value": round(n / t, 4), "unit": "evals/s", "cores": threads, "kind": "port",
                                "sample": "%d oracle egb_hcp (hcp 1) evaluations of the same system, OpenMP, "
                                          "%d threads" % (n, threads)}
        for pname in ("mixed", "fp64"):
            err = {k: abs(runs[pname][1]["energies"][k] - ref[j]) / abs(ref[j])
                   for j, k in enumerate(("epol", "evdw", "eelec"))}
            err["grad_rel_rms"] = float(np.sqrt(np.mean((grads[pname] - ref[3]) ** 2) / np.mean(ref[3] ** 2)))
            line["error_vs_oracle_hcp1_" + pname] = err
    print(json.dumps(line))


def bench_hcp_sweep(args):
    """Size sweep of the egb_hcp workload (hcp 1, 20-atom residues, strands
    of 50 residues, complexes of 10 strands, thresholds 15/50/150 A, gb=5,
    kappa 0.10395) at 25k-200k atoms: evaluations/s per precision, HIP-event
    pass times, the walk's units and the energy kernel's rate."""
    import numpy as np
    import nabc_amd as na
    F = HCP_FLOPS
    rows = []
    for natom in (25000, 50000, 100000, 200000):
        parm, arr, x = na.build_system(natom, HCP_SEED)
        ipres, ips, ipc = hcp_layers(natom)
        print("# %d atoms: counting the walk's units on the host" % natom, file=sys.stderr, flush=True)
        checks, far, near = hcp_walk_counts(x, ipres, ips, ipc, HCP_H)
        w_energy = checks * F["check"] + far * F["energy_far"] + near * F["energy_near"]
        row = {"natom": natom, "residues": len(ipres) - 1, "strands": len(ips), "complexes": len(ipc),
               "walk": {"checks": checks, "far_components": far, "near_pairs": near}}
        steps = max(10, min(100, int(4e6 / natom)))
        for pname, prec in (("mixed", na.GBHIP_MIXED), ("fp64", na.GBHIP_FP64)):
            ctx = na.GBContext(arr, x, gb=5, cut=999.0, rgbmax=999.0, kappa=HCP_KAPPA, precision=prec)
            ctx.set_hcp(1, ipres, ips, ipc, h1=HCP_H[0], h2=HCP_H[1], h3=HCP_H[2])
            ctx.nblist()
            for _ in range(3):
                ctx.egb_device()
            ctx.sync()
            ctx.timing_start(steps)
            t0 = time.perf_counter()
            for _ in range(steps):
                ctx.egb_device()
            ctx.sync()
            wall = (time.perf_counter() - t0) / steps
            ms, ne = ctx.timing_stop()
            ctx.close()
            pm = [v / ne for v in ms]
            row[pname] = {"evals_per_s": round(1.0 / wall, 2), "ms_per_eval": round(1e3 * wall, 4),
                          "pass_ms": [round(v, 4) for v in pm],
                          "energy_tflops": round(w_energy / (pm[2] * 1e-3) / 1e12, 3)}
        print("# %s" % json.dumps(row), file=sys.stderr, flush=True)
        rows.append(row)
    print(json.dumps({"metric": "HCP GB energy+grad evals/s by system size", "unit": "evals/s",
                      "pass_order": ["charges_born", "component_radii", "energy", "deriv"],
                      "data": "synthetic (seeded polymers, nabc_amd/csrc/gen.c)", "rows": rows}))


def bench_sweep(args):
    """Size sweep of the headline workload (igb=8, cut = rgbmax = 25 A,
    kappa 0.127315, list rebuilt every 20 evaluations): evaluations/s and
    pair interactions/s per precision at 5k-100k atoms, default launch
    shapes and step split (no autotuning), HIP-event pass times."""
    import numpy as np
    import nabc_amd as na
    rows = []
    for natom in (5000, 10000, 25000, 50000, 100000):
        parm, arr, x = na.build_system(natom, SEED)
        row = {"natom": natom}
        for name, prec in (("fp64", na.GBHIP_FP64), ("mixed", na.GBHIP_MIXED)):
            ctx = na.GBContext(arr, x, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA, precision=prec)
            ctx.nblist()
            npairs = ctx.npairs()
            steps = max(20, min(200, int(2e10 / max(1, npairs))))
            for k in range(5):
                ctx.egb_device()
            ctx.sync()
            ctx.timing_start(steps)
            t0 = time.perf_counter()
            for k in range(steps):
                if k % NSNB == 0:
                    ctx.nblist_async()
                ctx.egb_device()
            ctx.sync()
            t1 = time.perf_counter()
            ms, n = ctx.timing_stop()
            row[name] = {"evals_per_s": round(steps / (t1 - t0), 1),
                         "pairs_per_s": float("%.4g" % (npairs * steps / (t1 - t0))),
                         "pass_ms": [round(v / n, 4) for v in ms], "split": list(ctx.split())}
            row["npairs"] = npairs
            ctx.close()
        row["mixed_vs_fp64"] = round(row["mixed"]["evals_per_s"] / row["fp64"]["evals_per_s"], 3)
        rows.append(row)
        print("# %d atoms: %s" % (natom, json.dumps(row)), file=sys.stderr, flush=True)
    print(json.dumps({"metric": "GB energy+grad evals/s vs system size (igb=8, 25 A)", "unit": "evals/s",
                      "data": "synthetic (nabc_amd/csrc/gen.c, seed %d)" % SEED, "rows": rows}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes, no autotuning")
    ap.add_argument("--launch", default="", metavar="B1,B2,B3",
                    help="fixed workgroup counts of passes 1-3 (both precisions), no autotuning; "
                         "reproduces a tuned run's shapes under a profiler")
    ap.add_argument("--launch64", default="", metavar="B1,B2,B3",
                    help="fixed workgroup counts for the FP64 run only (else --launch applies to both)")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device 3 times (calibrates HBM byte counters)")
    ap.add_argument("--config", default="c3",
                    choices=["c3", "batch", "md", "serve", "accuracy", "sweep", "cg", "hessian", "hcp", "hcp-sweep"],
                    help="c3: 25k-atom evaluation rate (default); batch: 1,024-molecule scoring; "
                         "md: dynamics (see --md); serve: request loop; accuracy: per-term mme() "
                         "errors against the oracle at C1/C2/C3; sweep: evaluation rate at 5k-100k atoms")
    ap.add_argument("--jobs", type=int, default=200)
    ap.add_argument("--rate", type=float, default=10.0)
    ap.add_argument("--slots", type=int, default=8)
    ap.add_argument("--hbm-gb", type=float, default=64.0)
    ap.add_argument("--nmol", type=int, default=1024)
    ap.add_argument("--md", default="c3", choices=["c1", "c2", "c3"])
    ap.add_argument("--md-steps", type=int, default=0,
                    help="MD steps (default per config: c1 1000, c2 10000, c3 5000)")
    ap.add_argument("--md-min", type=int, default=500, help="CG iterations before dynamics")
    ap.add_argument("--cg-iters", type=int, default=100, help="conjgrad iterations of --config cg")
    ap.add_argument("--deterministic", action="store_true",
                    help="--config c3 with fixed-order reductions (gbhip_set_deterministic)")
    ap.add_argument("--md-no-graph", action="store_true",
                    help="launch each step's kernels on the stream instead of replaying the captured graphs")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.config == "batch":
        return bench_batch(args)
    if args.config == "md":
        return bench_md(args)
    if args.config == "serve":
        return bench_serve(args)
    if args.config == "accuracy":
        return bench_accuracy(args)
    if args.config == "sweep":
        return bench_sweep(args)
    if args.config == "hcp-sweep":
        return bench_hcp_sweep(args)
    if args.config == "cg":
        return bench_cg(args)
    if args.config == "hessian":
        return bench_hessian(args)
    if args.config == "hcp":
        return bench_hcp(args)

    import numpy as np
    import nabc_amd as na

    if args.copy_probe > 0:
        import torch
        a = torch.ones(args.copy_probe * 262144, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b

    with tempfile.TemporaryDirectory() as d:
        top, rst = os.path.join(d, "c3.prmtop"), os.path.join(d, "c3.rst7")
        na.generate(NATOM, SEED, top, rst)
        digests = dict(prmtop=sha256(top), rst7=sha256(rst))
        parm = na.read_parm(top)
        arr = na.parm_arrays(parm)
        x, _ = na.read_rst7(rst, arr["natom"])

    fixed = tuple(int(v) for v in args.launch.split(",")) if args.launch else None
    if fixed is not None and len(fixed) != 3:
        raise SystemExit("--launch takes three workgroup counts")
    fixed64 = tuple(int(v) for v in args.launch64.split(",")) if args.launch64 else fixed
    if fixed64 is not None and len(fixed64) != 3:
        raise SystemExit("--launch64 takes three workgroup counts")
    fp64 = run_gpu(na, arr, x, na.GBHIP_FP64, args.steps, args.warmup, not args.no_tune, fixed64, args.deterministic)
    mixed = run_gpu(na, arr, x, na.GBHIP_MIXED, args.steps, args.warmup, not args.no_tune, fixed, args.deterministic)

    # accuracy of the mixed kernels against the FP64 kernels (same list)
    acc = {k: abs(mixed["energies"][k] - fp64["energies"][k]) / abs(fp64["energies"][k])
           for k in fp64["energies"]}
    gerr = float(np.sqrt(np.mean((mixed["grad"] - fp64["grad"]) ** 2) / np.mean(fp64["grad"] ** 2)))

    P = mixed["npairs"]
    dom = max(("born", "pairs", "chain"), key=lambda k: mixed["pass_ms"][k])
    achieved = FLOPS[dom] * P / (mixed["pass_ms"][dom] * 1e-3) / 1e12
    # the kernels are instantiated with and without the step split (the
    # profiler names them k_born<float, false> etc.)
    kname = kernel_name(dom, "float", mixed["split"])
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK["mixed"], unit="TFLOP/s",
                    frac=round(achieved / PEAK["mixed"], 4), traffic=measured_traffic(kname), kernel=kname,
                    kernel_ms=round(mixed["pass_ms"][dom], 4),
                    kernel_ms_source="HIP events around the kernel on the engine stream, this run's timed region")
    # the same figure at the kernel's average duration in the committed
    # rocprofv3 kernel trace (the profiler serialises launches, so it is longer)
    prof_us = profiled(kname, "avg_us_kernel_trace")
    if prof_us:
        a_prof = FLOPS[dom] * P / (prof_us * 1e-6) / 1e12
        roofline["profile_kernel_ms"] = round(prof_us / 1e3, 4)
        roofline["achieved_at_profile_ms"] = round(a_prof, 3)
        roofline["frac_at_profile_ms"] = round(a_prof / PEAK["mixed"], 4)
    dom64 = max(("born", "pairs", "chain"), key=lambda k: fp64["pass_ms"][k])
    a64 = FLOPS[dom64] * P / (fp64["pass_ms"][dom64] * 1e-3) / 1e12
    # whole-evaluation figure: W = 240 P flops per evaluation at the measured rate
    roofline["eval_achieved"] = round(W_PER_PAIR * P * mixed["evals_per_s"] / 1e12, 3)
    roofline["eval_frac"] = round(W_PER_PAIR * P * mixed["evals_per_s"] / 1e12 / PEAK["mixed"], 4)
    # the ceiling this build can reach: independent FMA chains compiled like
    # the pair kernels (no packed FP32), measured on this device
    probe32, probe64 = na.valu_peak(na.GBHIP_MIXED), na.valu_peak(na.GBHIP_FP64)
    roofline["valu_probe_tflops"] = {"fp32": round(probe32, 2), "fp64": round(probe64, 2)}
    roofline["frac_of_probe"] = round(achieved / probe32, 4)

    cpu, vs_oracle = None, None
    gmax = float(np.max(np.abs(fp64["grad"])))
    acc["grad_max_component_abs"] = float(np.max(np.abs(mixed["grad"] - fp64["grad"])))
    acc["grad_max_component_rel_to_max"] = acc["grad_max_component_abs"] / gmax
    if not args.no_cpu:
        cpu, ref = cpu_baseline(parm, arr["natom"], x)
        # both GPU variants against the oracle's FP64 egb() on the same list
        vs_oracle = {}
        r_epol, r_evdw, r_eel, r_g = ref[0], ref[1], ref[2], ref[3]
        for name, res in (("fp64", fp64), ("mixed", mixed)):
            e = res["energies"]
            vs_oracle[name] = {
                "epol": abs(e["epol"] - r_epol) / abs(r_epol), "evdw": abs(e["evdw"] - r_evdw) / abs(r_evdw),
                "eelec": abs(e["eelec"] - r_eel) / abs(r_eel),
                "grad_rel_rms": float(np.sqrt(np.mean((res["grad"] - r_g) ** 2) / np.mean(r_g ** 2))),
                "grad_max_component_abs": float(np.max(np.abs(res["grad"] - r_g))),
                "grad_max_component_rel_to_max": float(np.max(np.abs(res["grad"] - r_g)) / np.max(np.abs(r_g)))}

    out = {
        "metric": "GB energy+grad evals/s, mixed vs FP64, at 25k atoms",
        "value": round(mixed["evals_per_s"], 2),
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_eval"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-pair/f64-accumulate",
        "data": "synthetic (seeded 25,000-atom polymer, generator in nabc_amd/csrc/gen.c)",
        "config": {"workload": "GB energy+gradient, 25,000 atoms, igb=8, cut 25 A, rgbmax 25 A, "
                               "kappa 0.127315 (0.15 M), nsnb 20",
                   "natom": NATOM, "seed": SEED, "npairs": P, "tiles": mixed["ntiles"],
                   "tile_pairs": mixed["tilepairs"], "sha256": digests,
                   "launch_blocks": {"mixed": mixed["launch"], "fp64": fp64["launch"]},
                   "step_split": {"mixed": list(mixed["split"]), "fp64": list(fp64["split"])},
                   "derivative_store": {"mixed": mixed["store"], "fp64": fp64["store"]},
                   "deterministic": bool(args.deterministic)},
        "fp64": {"value": round(fp64["evals_per_s"], 2), "ms_per_step": round(fp64["ms_per_eval"], 4),
                 "pass_ms": {k: round(v, 4) for k, v in fp64["pass_ms"].items()},
                 "dominant": kernel_name(dom64, "double", fp64["split"]), "achieved_tflops": round(a64, 3),
                 "frac": round(a64 / PEAK["fp64"], 4)},
        "mixed_vs_fp64": round(mixed["evals_per_s"] / fp64["evals_per_s"], 3),
        "mixed_pass_ms": {k: round(v, 4) for k, v in mixed["pass_ms"].items()},
        "nblist_ms": round(mixed["nblist_ms"], 4),
        "mixed_error_vs_fp64": {"epol": acc["epol"], "evdw": acc["evdw"], "eelec": acc["eelec"],
                                "grad_rel_rms": gerr, "grad_max_component_abs": acc["grad_max_component_abs"],
                                "grad_max_component_rel_to_max": acc["grad_max_component_rel_to_max"]},
        "error_vs_oracle": vs_oracle,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
