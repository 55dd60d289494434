"""gbhip_mme as one captured hipGraph per call (the minimizer's evaluation).

Replays must give what kernel-by-kernel launches give, over a sequence of
coordinates with list rebuilds in between (mme34's nsnb rule, eff.c:2564),
in both precisions; a change of launch shape or of the derivative store
must re-capture, and a list rebuild alone must not.
"""
import numpy as np
import pytest

from common import generated_system, rel_rms

pytestmark = pytest.mark.gpu


def _init(na, parm, opts):
    na.lib.mm_reset_options()
    assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0, na.lib.mm_get_error()


def _stats(na, ctx):
    cap, rep = na.C.c_longlong(), na.C.c_longlong()
    assert na.lib.gbhip_mme_graph_stats(ctx, na.C.byref(cap), na.C.byref(rep)) == 0
    return cap.value, rep.value


def _walk(na, x0, calls, seed):
    """mme() at calls coordinate sets (a random walk from x0): energies and
    gradients per call"""
    rng = np.random.default_rng(seed)
    x = x0.copy()
    out = []
    for it in range(1, calls + 1):
        f = np.zeros_like(x)
        itc = na.C.c_int(it)
        e = na.lib.mme(x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(itc))
        assert np.isfinite(e), na.lib.mm_get_error()
        ene = np.zeros(13)
        na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), None)
        out.append((ene, f))
        x = x + rng.normal(0.0, 0.02, x.shape)
    return out


@pytest.mark.parametrize("prec", ["fp64", "mixed"])
def test_graph_replays_match_plain_launches(gpu, gen_dir, prec):
    na = gpu
    parm, arr, x0 = generated_system(2500, 25, gen_dir)
    opts = "gb=8, cut=14.0, rgbmax=12.0, kappa=0.127315, nsnb=3, ntpr=100000, gbprec=" + prec
    runs = []
    for on in (1, 0):
        _init(na, parm, opts)
        ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
        assert ctx
        assert na.lib.gbhip_set_mme_graph(ctx, on) == 0
        runs.append(_walk(na, x0, 8, 11))
        cap, rep = _stats(na, ctx)
        if on:
            # one capture: the list rebuilds (calls 3 and 6) change only
            # device data; the store was sized at the first list build
            assert cap == 1 and rep == 8
        else:
            assert cap == 0 and rep == 0
        it = na.C.c_int(-3)
        na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
    for (eg, fg), (ep, fp) in zip(*runs):
        scale = np.abs(ep).max()
        # summation order of the device atomics differs between any two
        # evaluations: a few ulps of the largest term
        assert np.max(np.abs(eg - ep)) <= 1e-11 * scale
        assert rel_rms(fg, fp) < 1e-11
    # the walk moved: the calls are not one evaluation repeated
    assert abs(runs[0][0][0][0] - runs[0][-1][0][0]) > 1e-3


def test_graph_recaptures_on_launch_and_store_change(gpu, gen_dir):
    na = gpu
    parm, arr, x0 = generated_system(600, 26, gen_dir)
    _init(na, parm, "gb=5, cut=12.0, rgbmax=12.0, nsnb=1000, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    ref = _walk(na, x0, 2, 3)
    assert _stats(na, ctx) == (1, 2)
    # new launch shapes: captured again, same results
    b = (na.C.c_int * 3)(96, 80, 72)
    assert na.lib.gbhip_set_launch(ctx, b) == 0
    got = _walk(na, x0, 2, 3)
    assert _stats(na, ctx) == (2, 4)
    # the store freed (mode 0): its pointer changed, captured again
    assert na.lib.gbhip_set_derivative_store(ctx, 0) == 0
    got0 = _walk(na, x0, 2, 3)
    assert _stats(na, ctx) == (3, 6)
    # a timing session evaluates kernel by kernel (no replay counted)
    assert na.lib.gbhip_timing_start(ctx, 4) == 0
    tim = _walk(na, x0, 2, 3)
    ms = np.zeros(4)
    n = na.C.c_int()
    assert na.lib.gbhip_timing_stop(ctx, ms.ctypes.data_as(na.dp), na.C.byref(n)) == 0
    assert n.value == 2 and _stats(na, ctx) == (3, 6)
    for other in (got, got0, tim):
        for (e1, f1), (e2, f2) in zip(ref, other):
            assert np.max(np.abs(e1 - e2)) <= 1e-11 * np.abs(e2).max()
            assert rel_rms(f1, f2) < 1e-11
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_graph_after_async_rebuild_outgrows_store(gpu, gen_dir):
    """An asynchronous rebuild whose list no longer fits the derivative
    store leaves pass 3 to the recomputing kernel: the captured evaluation
    is re-captured with that kernel in it, and its gradient equals the one
    after a synchronous rebuild has grown the store again (FP64, same
    list)."""
    na = gpu
    parm, arr, x0 = generated_system(2500, 5, gen_dir)
    _init(na, parm, "gb=5, cut=10.0, rgbmax=10.0, kappa=0.1, nsnb=100000, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    f = np.zeros_like(x0)
    it = na.C.c_int(1)
    assert np.isfinite(na.lib.mme(x0.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it)))
    assert _stats(na, ctx)[0] == 1
    need, cap, used = na.C.c_longlong(), na.C.c_longlong(), na.C.c_int()
    xc = np.ascontiguousarray(x0 * 0.8)  # denser: about twice the pairs
    assert na.lib.gbhip_set_coords(ctx, xc.ctypes.data_as(na.dp)) == 0
    assert na.lib.gbhip_nblist_async(ctx) == 0
    assert na.lib.gbhip_store_stats(ctx, na.C.byref(need), na.C.byref(cap), na.C.byref(used)) == 0
    assert need.value > cap.value and not used.value
    g1, e1 = np.zeros_like(x0), np.zeros(14)
    assert na.lib.gbhip_mme(ctx, g1.ctypes.data_as(na.dp), e1.ctypes.data_as(na.dp)) == 0
    assert _stats(na, ctx) == (2, 2)
    np_ = na.C.c_longlong()
    assert na.lib.gbhip_nblist(ctx, na.C.byref(np_)) == 0
    assert na.lib.gbhip_store_stats(ctx, na.C.byref(need), na.C.byref(cap), na.C.byref(used)) == 0
    assert used.value
    g2, e2 = np.zeros_like(x0), np.zeros(14)
    assert na.lib.gbhip_mme(ctx, g2.ctypes.data_as(na.dp), e2.ctypes.data_as(na.dp)) == 0
    assert _stats(na, ctx) == (3, 3)
    assert np.max(np.abs(e1 - e2)) <= 1e-9 * np.abs(e2).max()
    assert rel_rms(g1, g2) < 1e-9
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_first_graph_call_sees_coordinates_copied_on_default_stream(gpu, gen_dir):
    """A context made without a stream copies coordinates on the default
    stream; its first captured evaluation runs on a stream of its own, which
    is ordered after that copy (the evaluation equals a plain one at the new
    coordinates, not one at the old)."""
    na = gpu
    parm, arr, x0 = generated_system(2500, 29, gen_dir)
    _init(na, parm, "gb=5, cut=12.0, rgbmax=12.0, nsnb=100000, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    np_ = na.C.c_longlong()
    assert na.lib.gbhip_nblist(ctx, na.C.byref(np_)) == 0
    x1 = np.ascontiguousarray(x0 + np.random.default_rng(6).normal(0, 0.05, x0.shape))
    g1, e1 = np.zeros_like(x0), np.zeros(14)
    assert na.lib.gbhip_set_coords(ctx, x1.ctypes.data_as(na.dp)) == 0
    assert na.lib.gbhip_mme(ctx, g1.ctypes.data_as(na.dp), e1.ctypes.data_as(na.dp)) == 0
    assert _stats(na, ctx) == (1, 1)
    assert na.lib.gbhip_set_mme_graph(ctx, 0) == 0
    g2, e2 = np.zeros_like(x0), np.zeros(14)
    assert na.lib.gbhip_set_coords(ctx, x1.ctypes.data_as(na.dp)) == 0
    assert na.lib.gbhip_mme(ctx, g2.ctypes.data_as(na.dp), e2.ctypes.data_as(na.dp)) == 0
    assert np.max(np.abs(e1 - e2)) <= 1e-11 * np.abs(e2).max()
    assert rel_rms(g1, g2) < 1e-11
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
