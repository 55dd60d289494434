"""The oracle against what the mathematics fixes (no GPU).

These pin the oracle where the reference's fixtures leave it open:
  - pass-1 descreening, every analytic case (eff.c:1663-1697), against
    numerical quadrature of the Hawkins-Cramer-Truhlar integral
    (1/4pi) int_{sphere j outside r < ri} |r|^-4 dV, including the
    rgbmax-truncated case (integral cut at r = rgbmax);
  - HCT and OBC effective radii (eff.c:1712-1727) with the published OBC
    parameters (Onufriev, Bashford, Case 2004: gb=2 (0.8, 0, 2.909125),
    gb=5 (1.0, 0.8, 4.85));
  - the two-atom generalized-Born energy (Still et al. 1990 with the
    Debye-Hueckel factor of eff.c:1884-1893, :1962-1987);
  - the gradient against central finite differences of the total energy,
    for gb 1/2/5/7/8 with salt and with the rgbmax switch;
  - translation and rotation invariance, zero net force and torque;
  - the pair list against a brute-force count;
  - the 1-4 electrostatics of diel=R (eff.c:1032-1040: q_i q_j / (scee
    r^2)) and diel=C (q_i q_j / (scee r)) by how they scale when every
    distance doubles (1/4 and 1/2; the 1-4 pairs are topological), the
    diel=R gradient against finite differences, and dielc (sff.c:1153-1162:
    charges divided by sqrt(dielc)) by halving every charge product (eel,
    eel14 and the GB energy) at dielc = 2 and leaving the other terms alone.
A dropped term, a wrong sign or a swapped radius in any of these fails one
of them.  Inputs are generated systems whose radii and screening factors
are overwritten in memory to reach each case.
"""
import numpy as np
import pytest
from scipy import integrate

import oracle

GBOFF = 0.09  # gboffset of gb 1/2/5 (sff.c:1190)
NSYS = 6


def _two_atoms(na, rb, fs, q=(0.0, 0.0)):
    """A two-atom system with the given intrinsic radii, screens and charges
    (prmtop units): a small generated system (6 atoms) with its parameters
    overwritten in memory; atoms 3-6 are uncharged and placed beyond the
    999 A cutoff of everything (see _coords), so no pair of theirs is
    listed."""
    p, a, x = na.build_system(NSYS, 1)
    s = p.contents
    for k in range(NSYS):
        s.Rborn[k] = rb[k] if k < 2 else 1.5
        s.Fs[k] = fs[k] if k < 2 else 0.8
        s.Charges[k] = q[k] if k < 2 else 0.0
    s.Fsmax = max(fs[0] * (rb[0] - GBOFF), fs[1] * (rb[1] - GBOFF)) + 0.5
    return p


def _coords(d):
    x = np.zeros((NSYS, 3))
    x[1, 0] = d
    for k in range(2, NSYS):
        x[k, 0] = -3000.0 * (k - 1)
    return x.ravel()


def _hct_integral(ri, d, s, upper=np.inf):
    """(1/4pi) * volume integral of |r|^-4 over the sphere of radius s at
    distance d, outside the sphere r < ri around the origin, r < upper."""
    def frac(r):  # fraction of the shell |r| = r inside sphere j
        if r <= s - d:
            return 1.0
        c = (r * r + d * d - s * s) / (2.0 * r * d)
        return 0.5 * (1.0 - min(1.0, max(-1.0, c)))
    lo = max(ri, abs(d - s)) if d >= s else ri
    hi = min(d + s, upper)
    if hi <= lo:
        return 0.0
    pts = [v for v in (s - d,) if lo < v < hi]
    val, _ = integrate.quad(lambda r: frac(r) / (r * r), lo, hi, points=pts or None,
                            epsabs=1e-15, epsrel=1e-13, limit=200)
    return val


def _reff(na, p, d, gb, rgbmax=999.0):
    x = _coords(d)
    o = oracle.Oracle(p, NSYS, gb=gb, cut=999.0, rgbmax=rgbmax)
    assert o.nblist(x) == 2
    _, _, _, _, reff, _ = o.egb(x)
    gp = o.gbparams()
    o.close()
    return reff, gp


# (rb_i, rb_j, fs_i, fs_j, d): the case of i descreened by j, with
# ri = rb_i - 0.09, sj = fs_j (rb_j - 0.09)
CASES = {
    "taylor d>4sj": (1.5, 1.2, 0.8, 0.8, 6.0),
    "outside ri+sj<d<4sj": (1.5, 1.7, 0.8, 0.8, 3.5),
    "overlap |ri-sj|<d<ri+sj": (1.5, 1.7, 0.8, 0.8, 1.9),
    "engulfed ri<sj, d<sj-ri": (1.2, 3.0, 0.8, 1.0, 0.9),
    "j inside i": (3.0, 1.2, 0.8, 0.5, 0.8),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_hct_radius_matches_quadrature(case):
    import nabc_amd as na
    rbi, rbj, fsi, fsj, d = CASES[case]
    p = _two_atoms(na, (rbi, rbj), (fsi, fsj))
    reff, _ = _reff(na, p, d, gb=1)
    ri, sj = rbi - GBOFF, fsj * (rbj - GBOFF)
    integral = _hct_integral(ri, d, sj)
    expect = 1.0 / (1.0 / ri - integral)
    if expect < 0:  # eff.c:1715 clamp
        expect = 30.0
    tol = 2e-8 if case.startswith("taylor") else 1e-11  # series truncation (sj/d)^10
    assert abs(reff[0] - expect) <= tol * expect, (case, reff[0], expect)


def test_taylor_branch_continuity_at_4sj():
    """Both sides of dij = 4 sj agree to the series' truncation error."""
    import nabc_amd as na
    p = _two_atoms(na, (1.5, 1.2), (0.8, 0.8))
    sj = 0.8 * (1.2 - GBOFF)
    lo, _ = _reff(na, p, 4 * sj * (1 - 1e-9), gb=1)
    hi, _ = _reff(na, p, 4 * sj * (1 + 1e-9), gb=1)
    assert abs(lo[0] - hi[0]) <= 1e-7 * lo[0]


@pytest.mark.parametrize("d", [4.9, 5.3, 5.8])
def test_rgbmax_switch_matches_truncated_integral(d):
    """rgbmax - sj < d < rgbmax + sj: the integral is cut at r = rgbmax."""
    import nabc_amd as na
    rgbmax = 5.0
    p = _two_atoms(na, (1.5, 2.0), (0.8, 0.9))
    reff, _ = _reff(na, p, d, gb=1, rgbmax=rgbmax)
    ri, sj = 1.5 - GBOFF, 0.9 * (2.0 - GBOFF)
    expect = 1.0 / (1.0 / ri - _hct_integral(ri, d, sj, upper=rgbmax))
    assert abs(reff[0] - expect) <= 1e-11 * expect


@pytest.mark.parametrize("gb,abg", [(2, (0.8, 0.0, 2.909125)), (5, (1.0, 0.8, 4.85))])
def test_obc_radius_and_published_parameters(gb, abg):
    import nabc_amd as na
    p = _two_atoms(na, (1.5, 1.7), (0.8, 0.72))
    d = 2.6
    reff, gp = _reff(na, p, d, gb=gb)
    assert np.allclose(gp["alpha"], abg[0]) and np.allclose(gp["beta"], abg[1])
    assert np.allclose(gp["gamma"], abg[2])
    ri, sj = 1.5 - GBOFF, 0.72 * (1.7 - GBOFF)
    psi = _hct_integral(ri, d, sj) * ri
    a, b, g = abg
    expect = 1.0 / (1.0 / ri - np.tanh((a - b * psi + g * psi * psi) * psi) / 1.5)
    assert abs(reff[0] - expect) <= 1e-11 * expect


@pytest.mark.parametrize("kappa", [0.0, 0.127315])
def test_two_atom_gb_energy_closed_form(kappa):
    import nabc_amd as na
    q = (0.4 * 18.2223, -0.7 * 18.2223)
    p = _two_atoms(na, (1.5, 1.7), (0.8, 0.72), q)
    d, eps = 3.1, 78.5
    x = _coords(d)
    o = oracle.Oracle(p, NSYS, gb=5, cut=999.0, rgbmax=999.0, kappa=kappa, epsext=eps)
    assert o.nblist(x) == 2
    epol, _, _, _, R, _ = o.egb(x)
    o.close()

    def screen(r):
        return 1.0 - np.exp(-0.73 * kappa * r) / eps

    e = -0.5 * q[0] ** 2 * screen(R[0]) / R[0] - 0.5 * q[1] ** 2 * screen(R[1]) / R[1]
    f = np.sqrt(d * d + R[0] * R[1] * np.exp(-d * d / (4.0 * R[0] * R[1])))
    e -= q[0] * q[1] * screen(f) / f
    assert abs(epol - e) <= 1e-12 * abs(e)


def _energy(o, x):
    ene, _, g = o.mme(x, rebuild=False)
    return ene[0], g


@pytest.mark.parametrize("gb,kappa,rgbmax", [(1, 0.10395, 999.0), (2, 0.0, 999.0), (5, 0.14701, 999.0),
                                             (5, 0.0, 7.0), (7, 0.0, 999.0), (8, 0.127315, 8.0)])
def test_gradient_matches_finite_differences(gb, kappa, rgbmax):
    import nabc_amd as na
    p, a, x = na.build_system(60, 21)
    o = oracle.Oracle(p, 60, gb=gb, cut=999.0, rgbmax=rgbmax, kappa=kappa)
    o.nblist(x)
    e0, g = _energy(o, x)
    h = 1e-5
    fd = np.zeros_like(x)
    for k in range(len(x)):
        xp = x.copy(); xp[k] += h
        xm = x.copy(); xm[k] -= h
        fd[k] = (_energy(o, xp)[0] - _energy(o, xm)[0]) / (2 * h)
    o.close()
    err = np.sqrt(np.mean((fd - g) ** 2) / np.mean(g ** 2))
    assert err <= 1e-6, err


def test_translation_rotation_invariance():
    import nabc_amd as na
    p, a, x = na.build_system(200, 8)
    o = oracle.Oracle(p, 200, gb=8, cut=999.0, rgbmax=999.0, kappa=0.127315)
    o.nblist(x)
    e0, g = _energy(o, x)
    X = x.reshape(-1, 3)
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0.0], [np.sin(th), np.cos(th), 0.0], [0.0, 0.0, 1.0]])
    Rx = np.array([[1.0, 0.0, 0.0], [0.0, np.cos(0.3), -np.sin(0.3)], [0.0, np.sin(0.3), np.cos(0.3)]])
    xr = (X @ (Rz @ Rx).T + np.array([3.0, -7.0, 11.0])).ravel()
    e1, g1 = _energy(o, xr)
    o.close()
    assert abs(e1 - e0) <= 1e-10 * abs(e0)
    G = g.reshape(-1, 3)
    scale = np.abs(G).max()
    assert np.abs(G.sum(0)).max() <= 1e-9 * scale * len(G)
    torque = np.cross(X, G).sum(0)
    assert np.abs(torque).max() <= 1e-9 * scale * len(G) * np.abs(X).max()
    # gradients rotate with the coordinates
    np.testing.assert_allclose(g1.reshape(-1, 3), G @ (Rz @ Rx).T, atol=1e-9 * scale)


def test_pair_list_brute_force():
    import nabc_amd as na
    p, a, x = na.build_system(300, 4)
    frozen = np.zeros(300, np.int32)
    frozen[::4] = 1
    o = oracle.Oracle(p, 300, gb=5, cut=8.0, rgbmax=8.0, frozen=frozen)
    nent = o.nblist(x)
    lo, up, flat = o.lists()
    o.close()
    X = x.reshape(-1, 3)
    d2 = ((X[:, None, :] - X[None, :, :]) ** 2).sum(-1)
    keep = (d2 < 64.0) & ~(frozen[:, None].astype(bool) & frozen[None, :].astype(bool))
    np.fill_diagonal(keep, False)
    assert nent == keep.sum()
    off = 0
    for i in range(300):
        js = flat[off:off + lo[i] + up[i]]
        off += lo[i] + up[i]
        assert list(js[:lo[i]]) == sorted(np.nonzero(keep[i, :i])[0].tolist())
        assert list(js[lo[i]:]) == sorted((np.nonzero(keep[i, i + 1:])[0] + i + 1).tolist())
    # the vectorised extraction the GPU parity tests compare lists with
    from common import oracle_pairs
    o = oracle.Oracle(p, 300, gb=5, cut=8.0, rgbmax=8.0, frozen=frozen)
    o.nblist(x)
    got = oracle_pairs(o, 300)
    o.close()
    iu, ju = np.nonzero(np.triu(keep, 1))
    assert np.array_equal(got, np.stack([iu, ju], 1))


EEL14, ENB14, EEL, EGB = 8, 7, 2, 10


@pytest.mark.parametrize("dield,ratio", [(0, 4.0), (1, 2.0)])
def test_14_electrostatics_scale_with_dielectric_form(dield, ratio):
    import nabc_amd as na
    p, a, x = na.build_system(120, 31)
    o = oracle.Oracle(p, 120, gb=5, cut=999.0, rgbmax=999.0, dield=dield)
    e1, _, _ = o.mme(x)
    e2, _, _ = o.mme(2.0 * x)
    o.close()
    assert abs(e1[EEL14]) > 1.0
    assert abs(e1[EEL14] / e2[EEL14] - ratio) <= 1e-12 * ratio, (e1[EEL14], e2[EEL14])


def test_14_terms_depend_on_dielectric_only_in_electrostatics():
    import nabc_amd as na
    p, a, x = na.build_system(120, 32)
    oc = oracle.Oracle(p, 120, gb=5, cut=999.0, rgbmax=999.0, dield=1)
    orr = oracle.Oracle(p, 120, gb=5, cut=999.0, rgbmax=999.0, dield=0)
    ec, _, _ = oc.mme(x)
    er, _, _ = orr.mme(x)
    oc.close()
    orr.close()
    # (the oracle's OpenMP sums may differ in the last bit)
    for k in (1, 2, 3, 4, 5, ENB14, 10):
        assert abs(ec[k] - er[k]) <= 1e-13 * abs(ec[k]), k
    assert ec[EEL14] != er[EEL14]


def test_diel_r_gradient_matches_finite_differences():
    import nabc_amd as na
    p, a, x = na.build_system(60, 33)
    o = oracle.Oracle(p, 60, gb=2, cut=999.0, rgbmax=999.0, dield=0)
    o.nblist(x)
    e0, g = _energy(o, x)
    h = 1e-5
    fd = np.zeros_like(x)
    for k in range(len(x)):
        xp = x.copy(); xp[k] += h
        xm = x.copy(); xm[k] -= h
        fd[k] = (_energy(o, xp)[0] - _energy(o, xm)[0]) / (2 * h)
    o.close()
    err = np.sqrt(np.mean((fd - g) ** 2) / np.mean(g ** 2))
    assert err <= 1e-6, err


def test_dielc_halves_every_charge_product():
    import nabc_amd as na
    p, a, x = na.build_system(150, 34)
    o1 = oracle.Oracle(p, 150, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1, dielc=1.0)
    o2 = oracle.Oracle(p, 150, gb=5, cut=999.0, rgbmax=999.0, kappa=0.1, dielc=2.0)
    e1, _, _ = o1.mme(x)
    e2, _, _ = o2.mme(x)
    o1.close()
    o2.close()
    for k in (EEL, EEL14, EGB):
        assert abs(e2[k] - 0.5 * e1[k]) <= 1e-12 * abs(e1[k]), (k, e1[k], e2[k])
    for k in (1, 3, 4, 5, ENB14):
        assert abs(e2[k] - e1[k]) <= 1e-13 * abs(e1[k]), k
