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
# algorithmic flops per unordered pair and pass for gb 7/8: pass 1 two
# directions x (24 + 10 neck), pass 2 68, pass 3 two directions x (37 + 15)
FLOPS = {"born": 68.0, "pairs": 68.0, "chain": 104.0}
PEAK = {"fp64": 78.6, "mixed": 157.3}  # TFLOP/s, MI355X vector FP64 / FP32


def sha256(path):
    h = hashlib.sha256()
    with open(path, "rb") as fp:
        h.update(fp.read())
    return h.hexdigest()


def run_gpu(na, arr, x, precision, steps, warmup):
    ctx = na.GBContext(arr, x, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA, precision=precision)
    ctx.nblist()
    npairs = ctx.npairs()
    ntiles, ntp = ctx.tile_stats()
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
                nblist_ms=1e3 * tl, npairs=npairs, ntiles=ntiles, tilepairs=ntp,
                energies=dict(epol=epol, evdw=evdw, eelec=eel), grad=g)


def cpu_baseline(parm, natom, x, budget_s=15.0):
    import oracle
    threads = min(16, os.cpu_count() or 1)
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
                sample="%d oracle egb() evaluations + 1 pair-list build (amortised over nsnb=%d) "
                       "of the same 25,000-atom igb=8 system; OpenMP, %d threads, %s" %
                       (n, NSNB, threads, cpu or "cpu model unknown")), ref


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path "
              "(run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import nabc_amd as na

    with tempfile.TemporaryDirectory() as d:
        top, rst = os.path.join(d, "c3.prmtop"), os.path.join(d, "c3.rst7")
        na.generate(NATOM, SEED, top, rst)
        digests = dict(prmtop=sha256(top), rst7=sha256(rst))
        parm = na.read_parm(top)
        arr = na.parm_arrays(parm)
        x, _ = na.read_rst7(rst, arr["natom"])

    fp64 = run_gpu(na, arr, x, na.GBHIP_FP64, args.steps, args.warmup)
    mixed = run_gpu(na, arr, x, na.GBHIP_MIXED, args.steps, args.warmup)

    # accuracy of the mixed kernels against the FP64 kernels (same list)
    acc = {k: abs(mixed["energies"][k] - fp64["energies"][k]) / abs(fp64["energies"][k])
           for k in fp64["energies"]}
    gerr = float(np.sqrt(np.mean((mixed["grad"] - fp64["grad"]) ** 2) / np.mean(fp64["grad"] ** 2)))

    P = mixed["npairs"]
    dom = max(("born", "pairs", "chain"), key=lambda k: mixed["pass_ms"][k])
    achieved = FLOPS[dom] * P / (mixed["pass_ms"][dom] * 1e-3) / 1e12
    roofline = dict(bound="valu", achieved=round(achieved, 3), peak=PEAK["mixed"], unit="TFLOP/s",
                    frac=round(achieved / PEAK["mixed"], 4), traffic=None, kernel="k_%s<float>" % dom,
                    kernel_ms=round(mixed["pass_ms"][dom], 4))
    dom64 = max(("born", "pairs", "chain"), key=lambda k: fp64["pass_ms"][k])
    a64 = FLOPS[dom64] * P / (fp64["pass_ms"][dom64] * 1e-3) / 1e12

    cpu = None
    if not args.no_cpu:
        cpu, _ = cpu_baseline(parm, arr["natom"], x)

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
                   "tile_pairs": mixed["tilepairs"], "sha256": digests},
        "fp64": {"value": round(fp64["evals_per_s"], 2), "ms_per_step": round(fp64["ms_per_eval"], 4),
                 "pass_ms": {k: round(v, 4) for k, v in fp64["pass_ms"].items()},
                 "dominant": "k_%s<double>" % dom64, "achieved_tflops": round(a64, 3),
                 "frac": round(a64 / PEAK["fp64"], 4)},
        "mixed_vs_fp64": round(mixed["evals_per_s"] / fp64["evals_per_s"], 3),
        "mixed_pass_ms": {k: round(v, 4) for k, v in mixed["pass_ms"].items()},
        "nblist_ms": round(mixed["nblist_ms"], 4),
        "mixed_error_vs_fp64": {"epol": acc["epol"], "evdw": acc["evdw"], "eelec": acc["eelec"],
                                "grad_rel_rms": gerr},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
