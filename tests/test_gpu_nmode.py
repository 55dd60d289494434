"""Second derivatives on the GPU: mme2, newton and nmode of the library.

- The device Hessian (gbhip_hessian behind mme2: central differences of the
  analytic FP64 gradient, two device evaluations per column) against the
  oracle's Hessian formed the same way from its own mme34, on a generated
  300-atom gb=5 system: 1e-7 of the largest element.  The gradient and
  energy mme2 returns against the oracle's to 1e-9.
- test/nab/asp_nm1 and asp_nm2 run through the library: conjgrad over mme
  (device vectors), newton over mme/mme2 (the library's dense Cholesky),
  then nmode (the library's symmetric eigensolver): the printed table matches
  asp_nm1.out.check (66 rows x 4 columns at their printed precision) and
  the lowest frequencies match asp_nm2.out.check (modes 7-20 to 1e-4; mode
  6 is a zero mode, see tests/test_oracle_nmode.py).
- The eigensolver's GPU path (Householder reduction on the device, used
  from order 1024 on) against numpy's LAPACK on a random symmetric matrix.
"""
import ctypes
import os
import re

import numpy as np
import pytest

import oracle
from oracle import nmode as onm
from common import GOLDEN, generated_system, golden, read_pdb_xyz

pytestmark = pytest.mark.gpu


def _mme2(na, x):
    n3 = len(x)
    f = np.zeros(n3)
    H = np.zeros(n3 * n3)
    m = np.zeros(n3 // 3)
    name = ctypes.create_string_buffer(n3 // 3 + 1)
    natom = na.C.c_int()
    it = na.C.c_int(1)
    dummy = [na.C.c_int(0) for _ in range(7)]
    e = na.lib.mme2(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), H.ctypes.data_as(na.dp),
                    m.ctypes.data_as(na.dp), None, *[na.C.byref(d) for d in dummy], na.C.byref(natom),
                    na.C.byref(it), name)
    return e, f, H.reshape(n3, n3), m, natom.value, name.raw[:natom.value]


def test_device_hessian_matches_oracle(gpu, gen_dir):
    na = gpu
    parm, arr, x = generated_system(300, 61, gen_dir)
    n = arr["natom"]
    orc = oracle.Oracle(parm, n, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1)
    mme = onm.MME(orc)
    r_e, r_g = mme(x)
    r_H = onm.hessian(mme, x)
    orc.close()
    na.lib.mm_reset_options()
    assert na.lib.mm_options(b"gb=5, cut=999.0, rgbmax=999.0, kappa=0.1, diel=C, ntpr=100000, gbprec=mixed") == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    try:
        e, f, H, m, natom, name = _mme2(na, np.ascontiguousarray(x))
        assert natom == n and np.allclose(m, arr["mass"])
        assert abs(e - r_e) <= 1e-9 * abs(r_e)
        assert np.max(np.abs(f - r_g)) <= 1e-9 * np.max(np.abs(r_g))
        assert np.array_equal(H, H.T)
        scale = np.max(np.abs(r_H))
        err = np.max(np.abs(H - r_H))
        print("device Hessian vs oracle: max |dH| %.2e of %.2e" % (err, scale))
        assert err <= 1e-7 * scale
    finally:
        it = na.C.c_int(-3)
        na.lib.mme2(None, None, None, None, None, None, None, None, None, None, None, None, None,
                    na.C.byref(it), None)
        na.lib.mme(np.zeros(3 * n).ctypes.data_as(na.dp), np.zeros(3 * n).ctypes.data_as(na.dp), na.C.byref(it))


def _asp_minimum(na):
    g = golden()["asp_nm1"]
    o = g["options"]
    parm = na.read_parm(GOLDEN + "/asp.top")
    arr = na.parm_arrays(parm)
    n = arr["natom"]
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    x = np.ascontiguousarray(x)
    na.lib.mm_reset_options()
    opts = "rgbmax=%g, cut=%g, ntpr=50, nsnb=%d, gb=%d, diel=C" % (o["rgbmax"], o["cut"], o["nsnb"], o["gb"])
    assert na.lib.mm_options(opts.encode()) == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    nn, f = na.C.c_int(3 * n), na.C.c_double()
    rms, dfp, mx = na.C.c_double(o["cg"]["rmsgrad"]), na.C.c_double(o["cg"]["dfpred"]), na.C.c_int(o["cg"]["maxiter"])
    ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(nn), na.C.byref(f), na.MME_C, na.C.byref(rms),
                          na.C.byref(dfp), na.C.byref(mx))
    assert ret == -3
    assert na.lib.mm_options(b"ntpr=1") == 0
    rms, nradd, mx = na.C.c_double(o["newton"]["rmsgrad"]), na.C.c_double(o["newton"]["nradd"]), na.C.c_int(o["newton"]["maxiter"])
    ret = na.lib.newton(x.ctypes.data_as(na.dp), na.C.byref(nn), na.C.byref(f), na.MME_C, na.MME2_C,
                        na.C.byref(rms), na.C.byref(nradd), na.C.byref(mx))
    assert ret > 0, ret
    return x, n


def _release(na, n):
    it = na.C.c_int(-3)
    na.lib.mme(np.zeros(3 * n).ctypes.data_as(na.dp), np.zeros(3 * n).ctypes.data_as(na.dp), na.C.byref(it))


def test_asp_nm1_through_the_library(gpu, tmp_path):
    na = gpu
    x, n = _asp_minimum(na)
    libc = ctypes.CDLL(None)
    libc.fopen.restype = ctypes.c_void_p
    libc.fopen.argtypes = [ctypes.c_char_p, ctypes.c_char_p]
    libc.fclose.argtypes = [ctypes.c_void_p]
    na.lib.mm_set_output.argtypes = [ctypes.c_void_p]
    path = str(tmp_path / "nmode.out")
    fp = libc.fopen(path.encode(), b"w")
    na.lib.mm_set_output(fp)
    try:
        ret = na.lib.nmode(x.ctypes.data_as(na.dp), 3 * n, na.MME2_C, 0, 1, 0.0, 0.0, 0)
    finally:
        na.lib.mm_set_output(None)
        libc.fclose(fp)
        _release(na, n)
    assert ret == 0
    rows = {}
    for line in open(path):
        f = line.split()
        if len(f) == 5 and f[0].isdigit():
            rows[int(f[0])] = [float(v) for v in f[1:]]
    gold = golden()["asp_nm1"]["modes"]
    assert sorted(rows) == [r["mode"] for r in gold]
    worst = 0.0
    for r in gold:
        got = rows[r["mode"]]
        for a, b in zip(got, (r["freq"], r["e"], r["cv"], r["s"])):
            worst = max(worst, abs(a - b))
    print("asp_nm1: worst printed difference %.4f" % worst)
    # the same 3 decimals, rounding of two computations allowed for
    assert worst <= 0.0011


def test_asp_nm2_through_the_library(gpu, tmp_path, capfd):
    na = gpu
    g = golden()["asp_nm2"]
    x, n = _asp_minimum(na)
    cwd = os.getcwd()
    os.chdir(tmp_path)  # nmode writes the modes file "vecs" (eigp > 0)
    try:
        capfd.readouterr()
        ret = na.lib.nmode(x.ctypes.data_as(na.dp), 3 * n, na.MME2_C, g["nmode"]["eigp"], g["nmode"]["ntrun"],
                           0.0, 0.0, 0)
        out = capfd.readouterr().out
        vecs = open("vecs").read()
    finally:
        os.chdir(cwd)
        _release(na, n)
    assert ret == 0
    got = {int(m.group(1)): float(m.group(2)) for m in re.finditer(r"^\s*(\d+)\s+(-?[\d.]+)\s*$", out, re.M)}
    for r in g["modes"]:
        if r["mode"] <= 6:
            continue
        assert abs(got[r["mode"]] - r["freq"]) <= g["tolerance"], (r, got[r["mode"]])
    assert vecs.startswith(" modes from NAB") and vecs.count("****") == g["nmode"]["eigp"]


@pytest.mark.parametrize("vectors", [0, 1])
def test_device_tridiagonalization_eigenvalues(gpu, vectors):
    """Matrices of order >= 1024 are reduced on the GPU (dense.hip):
    eigenvalues against numpy's LAPACK, eigenvectors by their residual."""
    na = gpu
    f = na.lib.nabc_dense_eigh
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, na.dp, na.dp, ctypes.c_int]
    n = 1200
    rng = np.random.default_rng(12)
    B = rng.normal(size=(n, n))
    A = np.ascontiguousarray(B + B.T)
    a = A.copy()
    w = np.zeros(n)
    assert f(n, a.ctypes.data_as(na.dp), w.ctypes.data_as(na.dp), vectors) == 0
    ref = np.linalg.eigvalsh(A)
    scale = np.abs(ref).max()
    assert np.max(np.abs(w - ref)) <= 1e-11 * scale
    if vectors:
        V = a.reshape(n, n)
        assert np.max(np.abs(V @ A - w[:, None] * V)) <= 1e-9 * scale
