"""Edge cases of the GPU path against the oracle: molecules around the
64-atom tile size, empty pair lists, all-frozen systems, one-molecule and
tiny-molecule batches, gb 7 with a cutoff and salt.

Tolerances as in test_gpu_parity: FP64 1e-9, mixed 1e-6 (energies) and
1e-5 (relative RMS gradient); pair lists compared as sets.
"""
import numpy as np
import pytest

import oracle
from common import rel, rel_rms

pytestmark = pytest.mark.gpu

TOL = {0: (1e-9, 1e-9), 1: (1e-6, 1e-5)}


def _check(na, p, a, x, gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, precision=0, frozen=None):
    etol, gtol = TOL[precision]
    ctx = na.GBContext(a, x, gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa, precision=precision, frozen=frozen)
    nent = ctx.nblist()
    o = oracle.Oracle(p, a["natom"], gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa, frozen=frozen)
    assert o.nblist(x) == nent
    epol, evdw, eel, g = ctx.egb()
    r = o.egb(x)
    for got, ref in ((epol, r[0]), (evdw, r[1]), (eel, r[2])):
        if ref == 0.0:
            assert got == 0.0
        else:
            assert rel(got, ref) <= etol, (got, ref)
    if np.any(r[3] != 0.0):
        assert rel_rms(g, r[3]) <= gtol
    else:
        assert np.all(g == 0.0)
    ctx.close()
    o.close()
    return nent


@pytest.mark.parametrize("natom", [6, 17, 63, 64, 65, 127, 128, 129, 191])
@pytest.mark.parametrize("precision", [0, 1])
def test_sizes_around_the_tile(gpu, natom, precision):
    """Partially filled last tiles and single-tile molecules."""
    na = gpu
    p, a, x = na.build_system(natom, 300 + natom)
    _check(na, p, a, x, gb=5, precision=precision)
    _check(na, p, a, x, gb=8, cut=9.0, rgbmax=8.0, kappa=0.1, precision=precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_empty_pair_list(gpu, precision):
    """A cutoff below every distance: no pairs, only the GB self terms."""
    na = gpu
    p, a, x = na.build_system(300, 9)
    assert _check(na, p, a, x, gb=5, cut=1.0, rgbmax=1.0, precision=precision) == 0


def test_all_atoms_frozen(gpu):
    """Not-both-frozen rule empties the list; frozen atoms have no gradient."""
    na = gpu
    p, a, x = na.build_system(400, 10)
    frozen = np.ones(400, np.int32)
    assert _check(na, p, a, x, gb=5, cut=12.0, rgbmax=12.0, frozen=frozen) == 0


@pytest.mark.parametrize("precision", [0, 1])
def test_gb7_cutoff_salt(gpu, precision):
    na = gpu
    p, a, x = na.build_system(1500, 12)
    _check(na, p, a, x, gb=7, cut=14.0, rgbmax=12.0, kappa=0.14701, precision=precision)


@pytest.mark.parametrize("precision", [0, 1])
def test_batches_with_tiny_and_single_molecules(gpu, precision):
    na = gpu
    etol, gtol = TOL[precision]
    for sizes in ([700], [6, 2000, 17, 64, 65, 300, 9]):
        systems = [na.build_system(n, 500 + k) for k, n in enumerate(sizes)]
        b = na.Batch([(a, x) for _, a, x in systems], gb=5, precision=precision)
        ene, grads = b.eval()
        for k, (p, a, x) in enumerate(systems):
            o = oracle.Oracle(p, a["natom"], gb=5, cut=999.0, rgbmax=999.0)
            o.nblist(x)
            r = o.egb(x)
            for c in range(3):
                assert rel(ene[k, c], r[c]) <= etol, (sizes, k, c, ene[k, c], r[c])
            assert rel_rms(grads[k], r[3]) <= gtol
            o.close()
        b.close()


def test_valu_probe_in_range(gpu):
    """The VALU ceiling probe gives a plausible MI355X-class rate (the
    published vector peaks are 78.6 TFLOP/s FP64 and 157.3 FP32 with packed
    FP32, which this build does not emit) and rejects a bad precision."""
    na = gpu
    for prec in (na.GBHIP_FP64, na.GBHIP_MIXED):
        t = na.valu_peak(prec)
        assert 5.0 < t < 170.0, (prec, t)
    with pytest.raises(RuntimeError):
        na.valu_peak(7)
