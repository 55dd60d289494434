"""Batched scoring: many molecules of uneven size in one launch per pass.

Each molecule's energies and gradient must equal the oracle's for that
molecule alone (FP64 to 1e-9, mixed to 1e-6 / 1e-5), whatever it is
batched with -- including molecules with different topologies, LJ type
tables and Fsmax values.
"""
import numpy as np
import pytest

import oracle
from common import GOLDEN, golden, read_pdb_xyz, rel, rel_rms

pytestmark = pytest.mark.gpu


def _sizes(n, lo, hi, seed):
    rng = np.random.default_rng(seed)
    return [int(v) for v in np.exp(rng.uniform(np.log(lo), np.log(hi), n))]


@pytest.mark.parametrize("precision", [0, 1])
def test_batch_uneven_generated(gpu, precision):
    na = gpu
    systems = [na.build_system(n, 1000 + k) for k, n in enumerate(_sizes(20, 200, 3000, 5))]
    b = na.Batch([(a, x) for _, a, x in systems], gb=5, precision=precision)
    ene, grads = b.eval()
    etol, gtol = (1e-9, 1e-9) if precision == 0 else (1e-6, 1e-5)
    for k, (p, a, x) in enumerate(systems):
        o = oracle.Oracle(p, a["natom"], gb=5, cut=999.0, rgbmax=999.0)
        o.nblist(x)
        r = o.egb(x)
        for c in range(3):
            assert rel(ene[k, c], r[c]) <= etol, (k, c, ene[k, c], r[c])
        assert rel_rms(grads[k], r[3]) <= gtol


def test_batch_mixed_topologies_gb8_cutoff(gpu):
    """aspgb8 + thioredoxin + generated systems, igb=8 with a cutoff: per-molecule
    LJ tables, exclusions and the stale-Fsmax skip must stay per molecule."""
    na = gpu
    mols = []
    p = na.read_parm(GOLDEN + "/asp.top")
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    mols.append((p, na.parm_arrays(p), x))
    p = na.read_parm(GOLDEN + "/prmtop.gb2")
    x, _ = read_pdb_xyz(GOLDEN + "/trxox.2.4ns.pdb")
    mols.append((p, na.parm_arrays(p), x))
    for k, n in enumerate((300, 1200, 2500)):
        mols.append(na.build_system(n, 77 + k))
    b = na.Batch([(a, x) for _, a, x in mols], gb=8, cut=14.0, rgbmax=12.0, kappa=0.127315)
    ene, grads = b.eval()
    for k, (p, a, x) in enumerate(mols):
        o = oracle.Oracle(p, a["natom"], gb=8, cut=14.0, rgbmax=12.0, kappa=0.127315)
        o.nblist(x)
        r = o.egb(x)
        for c in range(3):
            assert rel(ene[k, c], r[c]) <= 1e-9, (k, c)
        assert rel_rms(grads[k], r[3]) <= 1e-9
    # and the batch agrees with single-molecule contexts
    a, x = mols[1][1], mols[1][2]
    ctx = na.GBContext(a, x, gb=8, cut=14.0, rgbmax=12.0, kappa=0.127315)
    ctx.nblist()
    e1 = ctx.egb()
    assert rel(e1[0], ene[1, 0]) <= 1e-12 * 1e3 and rel_rms(e1[3], grads[1]) <= 1e-10


def test_batch_rebuild_after_moves(gpu):
    na = gpu
    systems = [na.build_system(n, 500 + k) for k, n in enumerate((400, 900, 1600))]
    b = na.Batch([(a, x) for _, a, x in systems], gb=5, cut=10.0, rgbmax=10.0)
    rng = np.random.default_rng(3)
    moved = []
    for k, (p, a, x) in enumerate(systems):
        x1 = x + rng.normal(0, 0.4, x.shape)
        moved.append(x1)
        assert na.lib.gbhip_batch_set_coords(b.h, k, x1.ctypes.data_as(na.dp)) == 0
    b.nblist()
    ene, grads = b.eval()
    for k, (p, a, x) in enumerate(systems):
        o = oracle.Oracle(p, a["natom"], gb=5, cut=10.0, rgbmax=10.0)
        o.nblist(moved[k])
        r = o.egb(moved[k])
        assert rel(ene[k, 0], r[0]) <= 1e-9 and rel_rms(grads[k], r[3]) <= 1e-9
