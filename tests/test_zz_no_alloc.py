"""gbhip_egb_static: the evaluation that allocates no device memory.

Every call takes its device arrays from the library's static buffer, so a
second call reads memory the first one left behind.  Two calls in a row,
the second with a smaller system, each against gbhip_ctx_create +
gbhip_nblist + gbhip_egb on the same input: the same kernels over the same
list, so energies agree to 1e-12 relative and gradients to 1e-10 of their
largest component (the only difference is the order of the atomic
accumulations).  A system too large for the buffer is refused with an
error that names the size.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

OPTS = dict(gb=5, cut=20.0, rgbmax=20.0, kappa=0.1)


def _ctx_eval(na, arr, x, **kw):
    c = na.GBContext(arr, x, **kw)
    try:
        c.nblist()
        return c.egb()
    finally:
        c.close()


def _compare(got, ref, label):
    for k, name in enumerate(("epol", "evdw", "eelec")):
        assert abs(got[k] - ref[k]) <= 1e-12 * max(1.0, abs(ref[k])), (label, name, got[k], ref[k])
    gmax = float(np.max(np.abs(ref[3])))
    dg = float(np.max(np.abs(got[3] - ref[3])))
    print("%s: epol %.12f / %.12f, max |dgrad| %.2e of %.2e" % (label, got[0], ref[0], dg, gmax), flush=True)
    assert dg <= 1e-10 * gmax, (label, dg, gmax)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_static_matches_context_two_calls(gpu, precision):
    na = gpu
    prec = na.GBHIP_FP64 if precision == "fp64" else na.GBHIP_MIXED
    for natom, seed in ((2500, 9101), (600, 9102)):
        _, arr, x = na.build_system(natom, seed)
        got = na.egb_static(arr, x, precision=prec, **OPTS)
        ref = _ctx_eval(na, arr, x, precision=prec, **OPTS)
        _compare(got, ref, "%s %d atoms" % (precision, natom))


def test_static_neck_model_and_repeat(gpu):
    """gb=8 (neck term) twice on the same input: both calls agree with the
    context and with each other."""
    na = gpu
    _, arr, x = na.build_system(1500, 9103)
    kw = dict(gb=8, cut=25.0, rgbmax=25.0, kappa=0.0)
    a = na.egb_static(arr, x, **kw)
    b = na.egb_static(arr, x, **kw)
    ref = _ctx_eval(na, arr, x, **kw)
    _compare(a, ref, "gb8 first")
    _compare(b, ref, "gb8 second")


def test_static_rejects_oversize(gpu):
    na = gpu
    _, arr, x = na.build_system(30000, 9104)
    with pytest.raises(na.GBHipError, match="static buffer"):
        na.egb_static(arr, x, **OPTS)
