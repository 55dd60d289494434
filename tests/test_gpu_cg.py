"""Device-resident conjgrad (gbhip_conjgrad) and served minimisation jobs.

- conjgrad() over the library's own mme() runs with its vectors on the
  device; conjgrad() over any other function runs on the host.  Over the
  same device mme() the two take the same steps: the same number of
  evaluations, the same return code, every evaluation's energy to 1e-9
  relative and the final coordinates to 1e-7 A (FP64; the dot products are
  reduced in a different order, 1e-16 relative apart).
- The reference's gb5trx run (thioredoxin, igb=5 with salt) through the
  device path prints the ff: lines of gb5trx.out.check:11-21.
- Served CG jobs of the 300-8,000-atom serving workload (gb=5, cut =
  rgbmax = 20 A, 50 iterations, run concurrently in FP64) end where the
  host conjgrad over the oracle's mme34 ends from the same start: final
  energy to 1e-8 relative, coordinates to 1e-5 A RMS.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from common import GOLDEN, generated_system, golden, read_pdb_xyz

pytestmark = pytest.mark.gpu

CG_OPTS = "gb=5, cut=20.0, rgbmax=20.0, kappa=0.0, diel=C, nsnb=10, ntpr=100000"


def _init(na, parm, opts):
    na.lib.mm_reset_options()
    assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0, na.lib.mm_get_error()


def _cg(na, func, x0, maxiter, rmsgrad_dof=1e-4):
    n = na.C.c_int(len(x0))
    f = na.C.c_double()
    rms = na.C.c_double(len(x0) * rmsgrad_dof)
    dfp = na.C.c_double(10.0)
    mx = na.C.c_int(maxiter)
    x = np.ascontiguousarray(x0.copy())
    ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), func, na.C.byref(rms),
                          na.C.byref(dfp), na.C.byref(mx))
    return ret, f.value, x


def _release(na, x):
    it = na.C.c_int(-3)
    na.lib.mme(x.ctypes.data_as(na.dp), x.copy().ctypes.data_as(na.dp), na.C.byref(it))


@pytest.mark.parametrize("natom", [600, 2500])
def test_device_cg_takes_the_host_cg_steps(gpu, gen_dir, natom):
    na = gpu
    parm, arr, x0 = generated_system(natom, 51, gen_dir)
    runs = {}
    for path in ("device", "host"):
        _init(na, parm, CG_OPTS + ", gbprec=fp64")
        seen = []

        def cb(xp, fp, itp):
            e = na.lib.mme(xp, fp, itp)
            seen.append((itp[0], e))
            return e

        if path == "device":
            func = na.MME_C
        else:
            func = na.MMEFUNC(cb)
        ret, f, x = _cg(na, func, x0, 60)
        runs[path] = (ret, f, x, list(seen))
        _release(na, x)
    (rd, fd, xd, _), (rh, fh, xh, seen) = runs["device"], runs["host"]
    assert rd == rh, (rd, rh)
    assert abs(fd - fh) <= 1e-9 * abs(fh), (fd, fh)
    assert np.max(np.abs(xd - xh)) <= 1e-7, np.max(np.abs(xd - xh))
    assert fh < seen[0][1]  # it minimised


def test_device_cg_gb5trx_ff_lines(gpu, tmp_path):
    """conjgrad(x, 3N, f, mme, 3N*1e-4, 10, 10) of gb5trx.nab through the
    device-resident path: the printed ff: lines are the golden ones."""
    na = gpu
    g = golden()["gb5trx"]
    parm = na.read_parm(GOLDEN + "/prmtop.gb2")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/trxox.2.4ns.pdb")
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    na.lib.mm_set_output.argtypes = [ctypes.c_void_p]
    path = str(tmp_path / "ff.out")
    fp = libc.fopen(path.encode(), b"w")
    na.lib.mm_set_output(fp)
    try:
        _init(na, parm, "ntpr=1, gb=5, kappa=0.14701, rgbmax=99.0, cut=99.0, diel=C")
        ret, f, xf = _cg(na, na.MME_C, x, 10)
        _release(na, xf)
    finally:
        na.lib.mm_set_output(None)
        libc.fclose(fp)
    assert ret == -3
    ff = [line.split() for line in open(path) if line.startswith("ff:")]
    assert [int(v[1]) for v in ff] == [w["iter"] for w in g["ff"]]
    for v, w in zip(ff, g["ff"]):
        for k, col in (("total", 2), ("bad", 3), ("vdw", 4), ("elect", 5), ("genborn", 7)):
            assert abs(float(v[col]) - w[k]) <= 0.0051, (k, v, w)


def test_served_cg_jobs_match_oracle_cg(gpu):
    """Five CG jobs of the serving workload, served together in FP64,
    against conjgrad over the oracle's mme34 from the same starts."""
    na = gpu
    sizes = [300, 900, 2500, 5000, 8000]
    jobs = []
    for k, natom in enumerate(sizes):
        p, a, x = na.build_system(natom, 700000 + k)
        jobs.append(dict(prm=p, x=x, kind=na.GBSERVE_CG, nsteps=50, seed=k + 1, arrival_s=0.002 * k, arr=a))
    na.lib.mm_reset_options()
    res = na.serve(jobs, slots=4, hbm_budget=16 << 30, precision=na.GBHIP_FP64, mm_opts=CG_OPTS, want_x=True,
                   cg_rmsgrad=1e-4, cg_dfpred=10.0)
    assert max(r["max_active"] for r in res) >= 2
    for j, r in zip(jobs, res):
        assert r["status"] >= -4, r["status"]
        natom = j["arr"]["natom"]
        orc = oracle.Oracle(j["prm"], natom, gb=5, cut=20.0, rgbmax=20.0)
        oracle.lib.orc_set_current(orc.h, 10)
        n = na.C.c_int(3 * natom)
        f = na.C.c_double()
        rms = na.C.c_double(1e-4)
        dfp = na.C.c_double(10.0)
        mx = na.C.c_int(50)
        x = np.ascontiguousarray(j["x"].copy())
        ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f),
                              na.MMEFUNC(oracle.lib.orc_mme_cb), na.C.byref(rms), na.C.byref(dfp), na.C.byref(mx))
        oracle.lib.orc_set_current(None, 1)
        orc.close()
        rms_dx = float(np.sqrt(np.mean((r["x"] - x) ** 2)))
        print("served CG %d atoms: status %d / oracle %d, energy %.10f / %.10f, coordinate rms diff %.2e" %
              (natom, r["status"], ret, r["energy"], f.value, rms_dx))
        assert r["status"] == ret
        assert abs(r["energy"] - f.value) <= 1e-8 * abs(f.value)
        assert rms_dx <= 1e-5
