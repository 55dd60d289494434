"""The headline workload against the oracle at its full size.

25,000-atom generated polymer (seed 2025, the benchmark's system), igb=8
(GB-neck), cut = rgbmax = 25 A, Debye-Hueckel kappa 0.127315 1/A (0.15 M),
pair list rebuilt every nsnb = 20 calls.

- One evaluation: identical pair sets, FP64 energies and gradient to 1e-9
  relative (effective radii to 1e-11), mixed to 1e-6 per energy term and
  1e-5 relative RMS on the gradient; the largest gradient-component error
  relative to the largest component is printed beside the RMS.
- A 21-call mme() walk (iter 0..20) over drifting coordinates, against the
  oracle's mme34 with the same rebuild cadence: the list is built at calls 0
  and 20 only, so between them both sides evaluate the list frozen at the
  call-0 coordinates while pairs drift across the cutoff (eff.c:2564-2578).
  Every call's per-term energies and gradient are compared.
"""
import numpy as np
import pytest

import oracle
from common import oracle_pairs, rel, rel_rms

pytestmark = pytest.mark.gpu

NATOM, SEED = 25000, 2025
GB, CUT, RGBMAX, KAPPA, NSNB = 8, 25.0, 25.0, 0.127315, 20
TOL = {0: (1e-9, 1e-9), 1: (1e-6, 1e-5)}  # (energy per term, gradient rel RMS)

_sys = {}


def _system(na):
    if "s" not in _sys:
        _sys["s"] = na.build_system(NATOM, SEED)
    return _sys["s"]


def _oracle_eval(parm, x):
    if "orc" not in _sys:
        orc = oracle.Oracle(parm, NATOM, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA)
        nent = orc.nblist(x)
        pairs = oracle_pairs(orc, NATOM)
        _sys["orc"] = (nent, pairs, orc.egb(x))
        orc.close()
    return _sys["orc"]


def _max_comp(g, r):
    return float(np.max(np.abs(g - r)) / np.max(np.abs(r)))


@pytest.mark.parametrize("precision", [0, 1])
def test_headline_egb_matches_oracle(gpu, precision):
    na = gpu
    parm, arr, x = _system(na)
    nent_ref, pairs_ref, (r_epol, r_evdw, r_eel, r_g, r_reff, _) = _oracle_eval(parm, x)
    ctx = na.GBContext(arr, x, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA, precision=precision)
    assert ctx.nblist() == nent_ref
    got = ctx.pairs()
    assert got.shape == pairs_ref.shape and np.array_equal(got, pairs_ref)
    del got
    epol, evdw, eel, g = ctx.egb()
    e_tol, g_tol = TOL[precision]
    errs = dict(epol=rel(epol, r_epol), evdw=rel(evdw, r_evdw), eel=rel(eel, r_eel),
                grad_rms=rel_rms(g, r_g), grad_max_comp=_max_comp(g, r_g))
    print("headline 25k igb=8 precision=%d npairs=%d errors %s" % (precision, nent_ref // 2, errs))
    assert errs["epol"] <= e_tol and errs["evdw"] <= e_tol and errs["eel"] <= e_tol, errs
    assert errs["grad_rms"] <= g_tol, errs
    if precision == 0:
        assert np.max(np.abs(ctx.reff() - r_reff) / r_reff) <= 1e-11
    ctx.close()


def test_headline_mme_walk_nsnb20(gpu):
    """21 mme() calls with nsnb=20 through the host API, FP64 and mixed,
    against orc_mme_cb (the oracle's mme34 with the reference's cadence)."""
    na = gpu
    parm, arr, x0 = _system(na)
    rng = np.random.default_rng(20)
    drift = rng.normal(0.0, 0.03, x0.shape)  # 0.6 A per component by call 20
    walk = [np.ascontiguousarray(x0 + k * drift) for k in range(NSNB + 1)]
    orc = oracle.Oracle(parm, NATOM, gb=GB, cut=CUT, rgbmax=RGBMAX, kappa=KAPPA)
    oracle.lib.orc_set_current(orc.h, NSNB)
    oracle.lib.orc_log_clear()
    r_grads = []
    for k, xk in enumerate(walk):
        g = np.zeros(3 * NATOM)
        it = oracle.C.c_int(k)
        oracle.lib.orc_mme_cb(xk.ctypes.data_as(oracle.dp), g.ctypes.data_as(oracle.dp), oracle.C.byref(it))
        r_grads.append(g)
    log = oracle.log_entries()
    assert [e[0] for e in log] == list(range(NSNB + 1))
    # the call-20 rebuild changes the list: pairs did cross the cutoff
    n_frozen = orc.nblist(walk[0])
    n_new = orc.nblist(walk[NSNB])
    assert n_new != n_frozen
    orc.close()
    for precision in (0, 1):
        e_tol, g_tol = TOL[precision]
        na.lib.mm_reset_options()
        opts = ("gb=8, cut=25.0, rgbmax=25.0, kappa=0.127315, nsnb=20, diel=C, ntpr=100000, gbprec=%s"
                % ("fp64" if precision == 0 else "mixed"))
        assert na.lib.mm_options(opts.encode()) == 0
        assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
        worst = dict(term=0.0, grad=0.0)
        for k, xk in enumerate(walk):
            f = np.zeros(3 * NATOM)
            it = na.C.c_int(k)
            na.lib.mme(xk.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
            ene = np.zeros(13)
            frms = na.C.c_double()
            na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
            _, r_ene, r_frms = log[k]
            assert rel(ene[0], r_ene[0]) <= e_tol, (precision, k, ene[0], r_ene[0])
            for t in (1, 2, 3, 4, 5, 7, 8, 10):
                err = abs(ene[t] - r_ene[t]) / max(1.0, abs(r_ene[t]))
                worst["term"] = max(worst["term"], err)
                assert err <= e_tol, (precision, k, t, ene[t], r_ene[t])
            gerr = rel_rms(f, r_grads[k])
            worst["grad"] = max(worst["grad"], gerr)
            assert gerr <= g_tol, (precision, k, gerr)
            assert rel(frms.value, r_frms) <= (1e-9 if precision == 0 else 1e-5)
        print("headline mme walk precision=%d worst %s" % (precision, worst))
        it = na.C.c_int(-3)
        na.lib.mme(walk[0].ctypes.data_as(na.dp), np.zeros(3 * NATOM).ctypes.data_as(na.dp), na.C.byref(it))
