"""The oracle's second-derivative path against the reference's normal modes.

test/nab/asp_nm1.nab and asp_nm2.nab minimise asp.top (gb=1, cut = rgbmax
= 100) by 500 CG and up to 15 Newton-Raphson iterations and print its
normal modes.  Here the same steps run over the oracle's mme34 (CG through
the library's host conjgrad, Newton-Raphson and the modes through
oracle/nmode.py, whose Hessian is central differences of the analytic
gradient), and every printed number must come out:
  - asp_nm1.out.check: modes 7-72, frequency and per-mode E, Cv, S, to
    their printed 3 decimals (half a unit of the last digit, plus 1e-4);
  - asp_nm2.out.check: modes 7-20 to 1e-4 (the reference script's
    tolerance).  Its mode 6 (0.00003) is one of the six zero modes of
    overall translation and rotation, whose computed values are numerical
    noise in both codes (here about 0.02 cm^-1, from the O(h^2) error of
    the differences); it is not compared.
No GPU is used.
"""
import numpy as np
import pytest

import oracle
from oracle import nmode as onm
from common import GOLDEN, golden, read_pdb_xyz


@pytest.fixture(scope="module")
def asp_minimum():
    import nabc_amd as na
    g = golden()["asp_nm1"]
    o = g["options"]
    parm = na.read_parm(GOLDEN + "/asp.top")
    arr = na.parm_arrays(parm)
    n = arr["natom"]
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    x = np.ascontiguousarray(x)
    orc = oracle.Oracle(parm, n, gb=o["gb"], cut=o["cut"], rgbmax=o["rgbmax"], kappa=o["kappa"])
    oracle.lib.orc_set_current(orc.h, o["nsnb"])
    nn, f = na.C.c_int(3 * n), na.C.c_double()
    rms, dfp, mx = na.C.c_double(o["cg"]["rmsgrad"]), na.C.c_double(o["cg"]["dfpred"]), na.C.c_int(o["cg"]["maxiter"])
    ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(nn), na.C.byref(f), oracle.MME_CB, na.C.byref(rms),
                          na.C.byref(dfp), na.C.byref(mx))
    oracle.lib.orc_set_current(None, 1)
    assert ret == -3  # the iteration limit, as in the reference run
    mme = onm.MME(orc)
    m = np.asarray(arr["mass"], dtype=np.float64)
    xm, code, e = onm.newton(mme, x, m, o["newton"]["rmsgrad"], o["newton"]["nradd"], o["newton"]["maxiter"])
    assert code > 0, code
    freq = onm.frequencies(onm.hessian(mme, xm), m)
    yield dict(x=xm, freq=freq, energy=e)
    orc.close()


def test_asp_nm1_modes_and_thermochemistry(asp_minimum):
    rows = golden()["asp_nm1"]["modes"]
    got = {r[0]: r for r in onm.thermo_modes(asp_minimum["freq"])}
    assert sorted(got) == [r["mode"] for r in rows]
    for r in rows:
        k, f, e, cv, s = got[r["mode"]]
        for a, b in ((f, r["freq"]), (e, r["e"]), (cv, r["cv"]), (s, r["s"])):
            assert abs(a - b) <= 0.0005 + 1e-4, (r, (f, e, cv, s))


def test_asp_nm2_lowest_frequencies(asp_minimum):
    g = golden()["asp_nm2"]
    for r in g["modes"]:
        if r["mode"] <= 6:
            continue
        assert abs(asp_minimum["freq"][r["mode"] - 1] - r["freq"]) <= g["tolerance"], r


def test_zero_modes_are_small(asp_minimum):
    """The six overall translations and rotations at the minimum."""
    assert np.all(np.abs(asp_minimum["freq"][:6]) < 0.1)
    assert asp_minimum["freq"][6] > 30.0


def test_hessian_translation_sum_and_symmetry():
    """Every row of the Hessian sums to zero over the x (y, z) coordinates
    of all atoms (a rigid translation changes no force), up to the
    O(h^2) difference error."""
    import nabc_amd as na
    p, a, x = na.build_system(40, 77)
    orc = oracle.Oracle(p, 40, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1)
    mme = onm.MME(orc)
    H = onm.hessian(mme, x)
    orc.close()
    assert np.array_equal(H, H.T)
    scale = np.abs(H).max()
    for k in range(3):
        assert np.abs(H[:, k::3].sum(axis=1)).max() <= 1e-5 * scale
