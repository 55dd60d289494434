"""The oracle's hierarchical charge partitioning (egb_hcp) against the
reference's own output and what the mathematics fixes (no GPU).

  - test/2trx.log:21 (test/thcp1.c:34-46): the first ff: line of the
    reference's hcp=1 run on thioredoxin, 2 decimals (bad, vdW, elect,
    genBorn; its nonpolar column is the gbsa term, outside this path);
  - thresholds beyond every distance open every component down to its
    atoms, where egb_hcp is egb without cut-offs: for gb 1, 2 and 5 (gb 7/8
    differ by the neck term, which the HCP path lacks) and every hcp
    variant, energies and gradients to 1e-12 and the Born radii exactly;
  - vertex-transitive configurations (identical atoms on a regular
    polyhedron, each its own residue, strand and complex, no LJ or
    exclusions): with one threshold below every distance, every other atom
    acts through its component's point charge at its own position, and by
    symmetry the one-sided far-field chain rule (2 x atom i's own
    derivative, hcp_gb.c:1654) equals the exact two-sided one, so the
    energy and gradient equal the exact GB's at each level (hcp 1 and 2);
    a dropped factor 2 or a sign fails it;
  - component charges and radii at every level, from the atoms directly
    (not by the roll-up the oracle uses): hcp 1 and 2 charges are the
    total (or negative / positive) charge at its centre of charge; radii
    are cube roots of summed single-precision cubes; HCPO's two charges
    keep the total charge, keep the dipole wherever neither offset from the
    geometric centre was limited, and stay within 0.9 rmax of it; for a
    charged component q1 = Q d0 / ((1.5 rmax)^2 Q + d0), d0 the smallest
    eigenvalue of its quadrupole about the centre of charge, and each
    charge's offset points along its unlimited position (1.5 rmax along
    that eigenvector from the centre of charge, the other at -q1/q2 times
    it), with the length that limit gives.
"""
import numpy as np
import pytest

import nabc_amd as na
import oracle
from common import golden, layered_hierarchy, rel, rel_rms, trx_hcp_system

TRX_OPTS = dict(gb=5, cut=15.0, rgbmax=15.0, kappa=0.125)


@pytest.fixture(scope="module")
def trx(tmp_path_factory):
    return trx_hcp_system(tmp_path_factory.mktemp("trx"))


def test_thcp1_first_ff_line(trx):
    parm, arr, x, (ipres, ips, ipc) = trx
    want = golden()["thcp1"]["ff"][0]
    o = oracle.Oracle(parm, arr["natom"], **TRX_OPTS)
    o.set_hcp(1, ipres, ips, ipc, h1=15.0, h2=999.0, h3=999.0)
    ene, frms, _ = o.mme(x)
    got = dict(bad=ene[3] + ene[4] + ene[5], vdw=ene[1] + ene[7], elect=ene[2] + ene[8], genborn=ene[10])
    for k, v in got.items():
        assert abs(v - want[k]) <= 0.0051, (k, v, want[k])
    # frms carries the gbsa forces in the log (6.96e+01); without them 69.64
    assert abs(frms - want["frms"]) <= 0.06, frms


@pytest.mark.parametrize("gb", [1, 2, 5])
@pytest.mark.parametrize("hcp", [1, 2, 4])
def test_infinite_thresholds_are_exact_gb(trx, gb, hcp):
    parm, arr, x, (ipres, ips, ipc) = trx
    n = arr["natom"]
    ref = oracle.Oracle(parm, n, gb=gb, cut=999.0, rgbmax=999.0, kappa=0.125)
    ref.nblist(x)
    e0 = ref.egb(x)
    o = oracle.Oracle(parm, n, gb=gb, cut=999.0, rgbmax=999.0, kappa=0.125)
    o.set_hcp(hcp, ipres, ips, ipc, h1=1e4, h2=1e4, h3=1e4)
    e1 = o.egb(x)
    for k in range(3):
        assert rel(e1[k], e0[k]) <= 1e-12, (k, e1[k], e0[k])
    assert rel_rms(e1[3], e0[3]) <= 1e-12
    assert np.array_equal(e1[4], e0[4])


def test_infinite_thresholds_layered():
    """Several strands and complexes, gb 5 with salt.  (With frozen atoms
    the two differ by design: nblist drops pairs of two frozen atoms,
    nblist.c:748-765, and egb_hcp, which has no list, keeps them.)"""
    p, a, x = na.build_system(900, 3)
    ipres, ips, ipc = layered_hierarchy(900)
    ref = oracle.Oracle(p, 900, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1)
    ref.nblist(x)
    e0 = ref.egb(x)
    o = oracle.Oracle(p, 900, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1)
    for hcp in (1, 2, 4):
        o.set_hcp(hcp, ipres, ips, ipc, h1=1e4, h2=1e4, h3=1e4)
        e1 = o.egb(x)
        for k in range(3):
            assert rel(e1[k], e0[k]) <= 1e-12, (hcp, k)
        assert rel_rms(e1[3], e0[3]) <= 1e-12


def _icosahedron(radius):
    phi = (1 + 5 ** 0.5) / 2
    v = []
    for s1 in (-1, 1):
        for s2 in (-1, 1):
            v += [(0, s1, s2 * phi), (s1, s2 * phi, 0), (s2 * phi, 0, s1)]
    v = np.array(v, float)
    return (v / np.linalg.norm(v[0]) * radius).ravel()


def _identical_atoms(n, q):
    """n identical atoms (charge q, radius 1.5 A, screen 0.8), no LJ and no
    excluded pairs: a generated system overwritten in memory."""
    p, _, _ = na.build_system(n, 1)
    s = p.contents
    for k in range(n):
        s.Charges[k] = q
        s.Rborn[k] = 1.5
        s.Fs[k] = 0.8
    nt = s.Ntypes
    for k in range(nt * (nt + 1) // 2):
        s.Cn1[k] = 0.0
        s.Cn2[k] = 0.0
    for k in range(s.Nnb):
        s.ExclAt[k] = 0
    return p


@pytest.mark.parametrize("gb", [1, 5])
@pytest.mark.parametrize("hcp", [1, 2])
@pytest.mark.parametrize("level", [1, 2, 3])
def test_far_field_equals_exact_on_symmetric_cluster(gb, hcp, level):
    n = 12
    p = _identical_atoms(n, 0.7 * 18.2223)
    x = _icosahedron(6.0)  # edge 6.3 A > 4 sj: the Taylor case on both sides
    ref = oracle.Oracle(p, n, gb=gb, cut=999.0, rgbmax=999.0, kappa=0.1)
    ref.nblist(x)
    e0 = ref.egb(x)
    ipres = np.arange(n + 1, dtype=np.int32)
    each = np.arange(1, n + 1, dtype=np.int32)
    h = [1e4, 1e4, 1e4]
    h[level - 1] = 0.5  # every other atom's component is far at this level
    ips = each if level >= 2 else np.array([1], np.int32)
    ipc = each if level == 3 else np.array([1], np.int32)
    o = oracle.Oracle(p, n, gb=gb, cut=999.0, rgbmax=999.0, kappa=0.1)
    o.set_hcp(hcp, ipres, ips, ipc, h1=h[0], h2=h[1], h3=h[2])
    e1 = o.egb(x)
    # the component radii are single-precision cube roots (6e-8 relative)
    assert rel(e1[0], e0[0]) <= 1e-7 and rel(e1[2], e0[2]) <= 1e-12
    assert rel_rms(e1[3], e0[3]) <= 1e-6, rel_rms(e1[3], e0[3])
    assert np.max(np.abs(e1[4] - e0[4]) / e0[4]) <= 1e-7
    assert np.abs(e0[3]).max() > 0.1  # forces well above the tolerance


def _members(ipres, ips, ipc, level, k):
    """atoms of component k of a level"""
    nres = len(ipres) - 1
    rfirst = list(ips - 1) + [nres]
    sfirst = list(ipc - 1) + [len(ips)]
    if level == 1:
        return np.arange(ipres[k], ipres[k + 1])
    if level == 2:
        return np.arange(ipres[rfirst[k]], ipres[rfirst[k + 1]])
    return np.arange(ipres[rfirst[sfirst[k]]], ipres[rfirst[sfirst[k + 1]]])


@pytest.mark.parametrize("hcp", [1, 2])
def test_component_charges_and_radii(hcp):
    p, a, x = na.build_system(900, 5)
    ipres, ips, ipc = layered_hierarchy(900)
    o = oracle.Oracle(p, 900, gb=5, cut=999.0, rgbmax=999.0)
    o.set_hcp(hcp, ipres, ips, ipc, h1=12.0, h2=20.0, h3=30.0)
    o.egb(x)
    q = a["charge"]
    xyz = x.reshape(-1, 3)
    sj = a["fs"] * (a["rborn"] - 0.09)
    for level in (1, 2, 3):
        xc, qc, rc, _ = o.hcp_components(level)
        for k in range(len(xc)):
            m = _members(ipres, ips, ipc, level, k)
            assert np.allclose(xc[k], xyz[m].mean(axis=0), rtol=0, atol=1e-10)
            groups = [m] if hcp == 1 else [m[q[m] <= 0], m[q[m] > 0]]
            for g, grp in enumerate(groups):
                qt = q[grp].sum()
                assert abs(qc[k, g, 3] - qt) <= 1e-9 * max(1.0, np.abs(q[grp]).sum())
                if abs(qt) > 1e-5:
                    coq = (q[grp, None] * xyz[grp]).sum(axis=0) / qt
                    assert np.allclose(qc[k, g, :3], coq, rtol=0, atol=1e-8 * max(1.0, np.abs(coq).max()))
                cubes = (sj[grp].astype(np.float32) ** 3).astype(np.float64).sum()
                assert abs(rc[k, g] - np.cbrt(cubes)) <= 1e-6 * max(rc[k, g], 1e-6)


def test_hcpo_charges_keep_charge_and_dipole(trx):
    parm, arr, x, (ipres, ips, ipc) = trx
    o = oracle.Oracle(parm, arr["natom"], **TRX_OPTS)
    o.set_hcp(4, ipres, ips, ipc, h1=15.0, h2=999.0, h3=999.0)
    o.egb(x)
    q = arr["charge"]
    xyz = x.reshape(-1, 3)
    xc, qc, _, _ = o.hcp_components(1)
    unlimited = charged = 0
    for k in range(len(xc)):
        m = np.arange(ipres[k], ipres[k + 1])
        rcog = xyz[m].mean(axis=0)
        rmax = np.sqrt(((xyz[m] - rcog) ** 2).sum(axis=1).max())
        Q, pd = q[m].sum(), (q[m, None] * xyz[m]).sum(axis=0)
        scale = np.abs(q[m]).sum()
        assert abs(qc[k, 0, 3] + qc[k, 1, 3] - Q) <= 1e-5 * scale
        off = qc[k, :, :3] - rcog
        dist = np.linalg.norm(off, axis=1)
        assert dist.max() <= 0.9 * rmax * (1 + 1e-5)  # offsets limited to 0.9 rmax
        if dist.max() < 0.8 * rmax * (1 - 1e-4):
            unlimited += 1
            dip = qc[k, 0, 3] * qc[k, 0, :3] + qc[k, 1, 3] * qc[k, 1, :3]
            assert np.allclose(dip, pd, rtol=0, atol=2e-5 * scale * (np.abs(xyz[m]).max() + 1)), k
        if abs(Q) > 0.6 * 18.2223:
            charged += 1
            coq = pd / Q
            y = xyz[m] - coq
            Qm = 0.5 * (3 * np.einsum("a,ai,aj->ij", q[m], y, y) - np.eye(3) * (q[m] * (y ** 2).sum(axis=1)).sum())
            w, v = np.linalg.eigh(Qm)
            rbar = 1.5 * rmax
            q1 = Q * w[0] / (rbar * rbar * Q + w[0])
            assert abs(qc[k, 0, 3] - q1) <= 1e-4 * abs(Q), (k, qc[k, 0, 3], q1)
            # unlimited offsets from the geometric centre: +-rbar along the
            # axis and -q1/q2 times that, about the centre of charge
            best = None
            for sgn in (1.0, -1.0):
                u = [sgn * rbar * v[:, 0] + coq - rcog, -q1 / (Q - q1) * sgn * rbar * v[:, 0] + coq - rcog]
                err = max(1 - abs(off[j] @ u[j]) / (dist[j] * np.linalg.norm(u[j])) for j in range(2))
                if best is None or err < best[0]:
                    best = (err, u)
            assert best[0] <= 1e-4, (k, best[0])
            for j in range(2):
                d = np.linalg.norm(best[1][j])
                if d < 0.8 * rmax:
                    assert abs(dist[j] - d) <= 1e-4 * rmax
                elif d > 0.9 * rmax:
                    assert abs(dist[j] - 0.9 * rmax) <= 1e-4 * rmax
                else:
                    assert 0.8 * rmax * (1 - 1e-4) <= dist[j] <= 0.9 * rmax * (1 + 1e-4)
    assert unlimited >= 20 and charged >= 5, (unlimited, charged)
