"""Hierarchical charge partitioning (egb_hcp) on the GPU against the oracle.

  - test/thcp1.c's thioredoxin (hcp=1, hcp_h1=15, hcp_h2=hcp_h3=999, gb=5,
    kappa 0.125) through the libsff host API (mm_options, mme_init_sff,
    mme): the first ff: line of test/2trx.log (bad, vdW, elect, genBorn to
    2 decimals) and every term against the oracle's mme34 to 1e-9;
  - the same molecule through the context API, gb 1/2/5 x hcp 1/2/4, and
    gb 7/8 on a generated 1,200-atom system (thioredoxin's radii are
    outside the neck table, which the reference refuses):
    energies and gradient (relative RMS) to 1e-9 for hcp 1 and 2 and 1e-7
    for hcp 4 (its point charges are single precision, and the device's
    FP64 acos/cos under them may differ from the host's by an ulp), Born
    radii to 1e-11 (1e-9);
  - a generated 3,000-atom system in 20-atom residues, strands of 5 and
    complexes of 3, thresholds 10 / 20 / 30 A so that every level is far
    for some atoms, every 9th atom frozen, gb 5 with salt: the same bars;
  - the mixed variant (descreening and GB terms in FP32, Coulomb, LJ and
    every sum in FP64) on thioredoxin and the layered system: energies to
    1e-6 relative, gradient to 1e-5 relative RMS, Born radii to 1e-5;
  - edge hierarchies (one residue for the whole molecule, one per atom),
    refused hierarchies and options (GBHIP_EINVAL), and the return to the
    pair-list path when HCP is switched off;
  - dynamics and minimisation through the host API with hcp on: 40 Langevin
    steps (one hipGraph each) against the oracle's md() (coordinates and
    velocities to 1e-8), and 15 conjgrad iterations with the device-resident
    minimiser against conjgrad over the oracle's mme34 (every evaluation's
    energy to 1e-9 relative, final coordinates to 1e-7 A).
"""
import numpy as np
import pytest

import oracle
from common import golden, layered_hierarchy, rel, rel_rms, trx_hcp_system

pytestmark = pytest.mark.gpu

TRX_OPTS = "gb=5, cut=15.0, rgbmax=15.0, kappa=0.125, diel=C, hcp=1, hcp_h1=15, hcp_h2=999, hcp_h3=999"


@pytest.fixture(scope="module")
def trx(tmp_path_factory):
    return trx_hcp_system(tmp_path_factory.mktemp("trx"))


def _init(na, parm, opts, frozen=None):
    na.lib.mm_reset_options()
    assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
    fz = None if frozen is None else np.ascontiguousarray(frozen, dtype=np.int32)
    assert na.lib.mme_init_sff(parm, fz.ctypes.data_as(na.ip) if fz is not None else None, None, None,
                               None) == 0, na.lib.mm_get_error()


def _release(na, x):
    it = na.C.c_int(-3)
    na.lib.mme(x.ctypes.data_as(na.dp), x.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_thcp1_first_line_through_mme(gpu, trx):
    na = gpu
    parm, arr, x, (ipres, ips, ipc) = trx
    want = golden()["thcp1"]["ff"][0]
    print("thcp1: 2trx, hcp=1 via mme()", flush=True)
    _init(na, parm, TRX_OPTS + ", ntpr=100000, gbprec=fp64")
    f = np.zeros_like(x)
    it = na.C.c_int(0)
    na.lib.mme(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
    ene = np.zeros(13)
    frms = na.C.c_double()
    na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
    _release(na, x)
    got = dict(bad=ene[3] + ene[4] + ene[5], vdw=ene[1] + ene[7], elect=ene[2] + ene[8], genborn=ene[10])
    for k, v in got.items():
        assert abs(v - want[k]) <= 0.0051, (k, v, want[k])
    o = oracle.Oracle(parm, arr["natom"], gb=5, cut=15.0, rgbmax=15.0, kappa=0.125)
    o.set_hcp(1, ipres, ips, ipc, h1=15.0, h2=999.0, h3=999.0)
    e0, frms0, g0 = o.mme(x)
    for t in (0, 1, 2, 3, 4, 5, 7, 8, 10):
        assert rel(ene[t], e0[t]) <= 1e-9, (t, ene[t], e0[t])
    assert rel_rms(f, g0) <= 1e-9
    assert abs(frms.value - frms0) <= 1e-9 * frms0


def _compare(na, arr, x, gb, hcp, hier, h, kappa=0.125, frozen=None, parm=None, tol=None, precision=0):
    ipres, ips, ipc = hier
    print("hcp parity: N=%d gb=%d hcp=%d h=%s precision=%d" % (arr["natom"], gb, hcp, h, precision), flush=True)
    ctx = na.GBContext(arr, x, gb=gb, cut=999.0, rgbmax=999.0, kappa=kappa, frozen=frozen, precision=precision)
    ctx.set_hcp(hcp, ipres, ips, ipc, h1=h[0], h2=h[1], h3=h[2])
    assert ctx.nblist() == 0
    epol, evdw, eel, g = ctx.egb()
    reff = ctx.reff()
    o = oracle.Oracle(parm, arr["natom"], gb=gb, cut=999.0, rgbmax=999.0, kappa=kappa, frozen=frozen)
    o.set_hcp(hcp, ipres, ips, ipc, h1=h[0], h2=h[1], h3=h[2])
    r_epol, r_evdw, r_eel, r_g, r_reff, _ = o.egb(x)
    print("  epol %.2e evdw %.2e eel %.2e grad %.2e reff %.2e" % (
        rel(epol, r_epol), rel(evdw, r_evdw), rel(eel, r_eel), rel_rms(g, r_g),
        np.max(np.abs(reff - r_reff) / r_reff)), flush=True)
    if precision:
        # the mixed variant's energy and force bounds; its Born radii (HCT's
        # 1 / (1/ri + sum) amplifies the sum's FP32 rounding) to 1e-5
        etol, gtol, rtol = 1e-6, 1e-5, 1e-5
    else:
        etol = gtol = tol or (1e-7 if hcp == 4 else 1e-9)
        rtol = 1e-9 if hcp == 4 else 1e-11
    for a, b in ((epol, r_epol), (evdw, r_evdw), (eel, r_eel)):
        assert rel(a, b) <= etol, (a, b)
    assert rel_rms(g, r_g) <= gtol, rel_rms(g, r_g)
    assert np.max(np.abs(reff - r_reff) / r_reff) <= rtol
    # a second evaluation at moved coordinates (the charges follow the atoms)
    x2 = x + 0.05 * np.sin(np.arange(len(x)))
    ctx.set_coords(x2)
    e2 = ctx.egb()
    r2 = o.egb(x2)
    assert rel(e2[0], r2[0]) <= etol and rel_rms(e2[3], r2[3]) <= gtol
    ctx.close()
    o.close()


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("gb", [1, 2, 5])
@pytest.mark.parametrize("hcp", [1, 2, 4])
def test_trx_parity(gpu, trx, gb, hcp, precision):
    parm, arr, x, hier = trx
    _compare(gpu, arr, x, gb, hcp, hier, (15.0, 999.0, 999.0), parm=parm, precision=precision)


@pytest.mark.parametrize("gb", [7, 8])
@pytest.mark.parametrize("hcp", [1, 4])
def test_neck_models_parity(gpu, gb, hcp):
    """gb 7/8 (thioredoxin's radii fall outside the neck table's 1-2 A,
    which the reference refuses, sff.c:1318): a generated system"""
    na = gpu
    p, a, x = na.build_system(1200, 43)
    _compare(na, a, x, gb, hcp, layered_hierarchy(1200), (12.0, 30.0, 999.0), parm=p)


@pytest.mark.parametrize("precision", [0, 1])
@pytest.mark.parametrize("hcp", [1, 2, 4])
def test_layered_parity_frozen(gpu, hcp, precision):
    na = gpu
    p, a, x = na.build_system(3000, 41)
    frozen = np.zeros(3000, np.int32)
    frozen[::9] = 1
    _compare(na, a, x, 5, hcp, layered_hierarchy(3000), (10.0, 20.0, 30.0), kappa=0.1, frozen=frozen, parm=p,
             precision=precision)


def test_md_follows_oracle(gpu, trx):
    na = gpu
    parm, arr, x0, (ipres, ips, ipc) = trx
    print("hcp md: 2trx 40 steps", flush=True)
    _init(na, parm, TRX_OPTS + ", ntpr=100000, ntpr_md=100000, gbprec=fp64")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx, na.lib.mm_get_error()
    o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                     boltz2=9.93595e-4, nsnb=10, zerov=0, seed=11)
    masses = arr["mass"]
    assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                na.C.byref(o), 40) == 0, na.lib.gbhip_last_error()
    assert na.lib.gbhip_md_run(ctx, 40, 1) == 0, na.lib.gbhip_last_error()
    xd = np.zeros_like(x0)
    vd = np.zeros_like(x0)
    assert na.lib.gbhip_md_state(ctx, xd.ctypes.data_as(na.dp), vd.ctypes.data_as(na.dp), None) == 0
    _release(na, x0)
    orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=15.0, rgbmax=15.0, kappa=0.125)
    orc.set_hcp(1, ipres, ips, ipc, h1=15.0, h2=999.0, h3=999.0)
    xr, vr, _ = orc.md(x0, None, masses, 40, dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, nsnb=10, seed=11)
    orc.close()
    assert np.max(np.abs(xd - x0)) > 1e-3
    assert np.max(np.abs(xd - xr)) < 1e-8, np.max(np.abs(xd - xr))
    assert np.max(np.abs(vd - vr)) < 1e-8, np.max(np.abs(vd - vr))


def test_cg_follows_oracle(gpu, trx):
    na = gpu
    parm, arr, x0, (ipres, ips, ipc) = trx
    print("hcp cg: 2trx 15 iterations", flush=True)
    n3 = len(x0)

    def run(func):
        n = na.C.c_int(n3)
        f = na.C.c_double()
        rms = na.C.c_double(n3 * 1e-4)
        dfp = na.C.c_double(10.0)
        mx = na.C.c_int(15)
        x = np.ascontiguousarray(x0.copy())
        ret = na.lib.conjgrad(x.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), func, na.C.byref(rms),
                              na.C.byref(dfp), na.C.byref(mx))
        return ret, f.value, x

    _init(na, parm, TRX_OPTS + ", ntpr=100000, nsnb=10, gbprec=fp64")
    rd, fd, xd = run(na.MME_C)
    _release(na, xd)
    orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=15.0, rgbmax=15.0, kappa=0.125)
    orc.set_hcp(1, ipres, ips, ipc, h1=15.0, h2=999.0, h3=999.0)
    oracle.lib.orc_set_current(orc.h, 10)
    oracle.lib.orc_log_clear()
    rr, fr, xr = run(na.MMEFUNC(oracle.lib.orc_mme_cb))
    oracle.lib.orc_set_current(None, 1)
    orc.close()
    assert rd == rr, (rd, rr)
    assert abs(fd - fr) <= 1e-9 * abs(fr), (fd, fr)
    assert np.max(np.abs(xd - xr)) <= 1e-7, np.max(np.abs(xd - xr))


def test_edge_hierarchies_and_errors(gpu):
    """One residue holding the whole molecule, one residue per atom, a
    single strand and complex; invalid hierarchies and options are refused
    with GBHIP_EINVAL; switching HCP off returns to the pair-list path."""
    na = gpu
    p, a, x = na.build_system(300, 61)
    n = a["natom"]
    cases = [(np.array([0, n], np.int32), (20.0, 50.0, 150.0)),
             (np.arange(n + 1, dtype=np.int32), (6.0, 50.0, 150.0))]
    for ipres, h in cases:
        for hcp in (1, 2):
            print("edge hierarchy: nres=%d hcp=%d" % (len(ipres) - 1, hcp), flush=True)
            _compare(na, a, x, 5, hcp, (ipres, np.array([1], np.int32), np.array([1], np.int32)), h, kappa=0.1,
                     parm=p)
    ctx = na.GBContext(a, x, gb=5, cut=999.0, rgbmax=999.0)
    good = layered_hierarchy(n)
    bad = [(3, good), (1, (np.array([1, n], np.int32), good[1], good[2])),           # hcp 3; not from atom 0
           (1, (np.array([0, n - 1], np.int32), good[1], good[2])),                    # not ending at natom
           (1, (np.array([0, 200, 100, n], np.int32), good[1], good[2])),              # decreasing
           (1, (good[0], np.array([2], np.int32), good[2])),                           # strand not from 1
           (1, (good[0], np.array([1, 1], np.int32), good[2])),                        # strands not increasing
           (1, (good[0], good[1], np.array([1, 99], np.int32)))]                       # complex beyond strands
    for hcp, hier in bad:
        with pytest.raises(na.GBHipError) as ei:
            ctx.set_hcp(hcp, *hier)
        assert "code -1" in str(ei.value), ei.value
    ctx.set_hcp(1, *good, h1=12.0)
    with pytest.raises(na.GBHipError):
        ctx.set_deterministic(1)
    with pytest.raises(na.GBHipError):
        ctx.autotune(1)
    e_hcp = ctx.egb()
    ctx.set_hcp(0, *good)
    ctx.nblist()
    e_pair = ctx.egb()
    o = oracle.Oracle(p, n, gb=5, cut=999.0, rgbmax=999.0)
    o.nblist(x)
    ref = o.egb(x)
    assert rel(e_pair[0], ref[0]) <= 1e-9 and rel_rms(e_pair[3], ref[3]) <= 1e-9
    assert abs(e_hcp[0] - e_pair[0]) > 1e-6 * abs(ref[0])  # the HCP result was a different one
    ctx.close()
