"""Deterministic reductions (gbhip_set_deterministic).

With the mode on, every per-atom sum and every energy is added up in a
fixed order, so the same input gives bit-identical energies, gradient,
effective radii and sumdeijda on every run, whatever the launch shapes
(workgroup counts change which workgroup takes which tile pair and in
what order) and in a fresh context.  The atomic (default) and the
deterministic results are the same sums in another order: energies to
1e-12 relative, gradient to 1e-10 of its largest component.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _eval(c):
    ep, ev, ee, g = c.egb()
    return np.array([ep, ev, ee]), g, c.reff(), c.sumdeijda()


def _same_bits(a, b, label):
    for k, name in enumerate(("energies", "gradient", "reff", "sumdeijda")):
        x, y = np.asarray(a[k]), np.asarray(b[k])
        assert x.tobytes() == y.tobytes(), (label, name, float(np.max(np.abs(x - y))))


def _close(a, b, label):
    assert np.all(np.abs(a[0] - b[0]) <= 1e-12 * np.maximum(1.0, np.abs(b[0]))), (label, a[0], b[0])
    gmax = float(np.max(np.abs(b[1])))
    assert float(np.max(np.abs(a[1] - b[1]))) <= 1e-10 * gmax, label
    assert float(np.max(np.abs(a[2] - b[2]))) <= 1e-12 * float(np.max(np.abs(b[2]))), label


CASES = {
    "2500-gb5-fp64": (2500, 4401, dict(gb=5, cut=20.0, rgbmax=18.0, kappa=0.1)),
    "2500-gb5-mixed": (2500, 4401, dict(gb=5, cut=20.0, rgbmax=18.0, kappa=0.1)),
    "600-gb8-fp64": (600, 4402, dict(gb=8, cut=999.0, rgbmax=999.0, kappa=0.0)),
    "25000-gb8-fp64": (25000, 2025, dict(gb=8, cut=25.0, rgbmax=25.0, kappa=0.127315)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_deterministic_runs_are_bit_identical(gpu, case):
    na = gpu
    natom, seed, kw = CASES[case]
    prec = na.GBHIP_MIXED if case.endswith("mixed") else na.GBHIP_FP64
    _, arr, x = na.build_system(natom, seed)
    c = na.GBContext(arr, x, precision=prec, **kw)
    try:
        c.nblist()
        atomic = _eval(c)
        c.set_deterministic(1)
        first = _eval(c)
        _same_bits(_eval(c), first, case + " repeat")
        default = c.launch()
        for blocks in ((7, 13, 5), (4096, 4096, 4096), default):
            c.set_launch(*blocks)
            _same_bits(_eval(c), first, "%s grids %s" % (case, blocks))
        with pytest.raises(na.GBHipError, match="step split"):
            c.set_split(2, 2, 2)
        c.set_deterministic(0)
        _close(_eval(c), first, case + " atomic again")
    finally:
        c.close()
    _close(first, atomic, case + " atomic vs deterministic")
    c2 = na.GBContext(arr, x, precision=prec, **kw)
    try:
        c2.set_deterministic(1)
        c2.nblist()
        _same_bits(_eval(c2), first, case + " fresh context")
    finally:
        c2.close()
    print("%s: epol %.12f (atomic %.12f), bit-identical over repeats, grids and contexts"
          % (case, first[0][0], atomic[0][0]))


def test_deterministic_batch(gpu):
    na = gpu
    systems = [na.build_system(n, 4500 + k) for k, n in enumerate((300, 1200, 700))]
    b = na.Batch([(a, x) for _, a, x in systems], gb=5, cut=16.0, rgbmax=14.0)
    try:
        b.nblist()
        ea, ga = b.eval()
        b.set_deterministic(1)
        e1, g1 = b.eval()
        e2, g2 = b.eval()
        assert e1.tobytes() == e2.tobytes()
        assert np.all(np.abs(e1 - ea) <= 1e-12 * np.maximum(1.0, np.abs(ea)))
        for x1, x2, xa in zip(g1, g2, ga):
            assert x1.tobytes() == x2.tobytes()
            assert float(np.max(np.abs(x1 - xa))) <= 1e-10 * float(np.max(np.abs(xa)))
    finally:
        b.close()
