"""GPU path (C-ABI) against the oracle and the reference's known answers.

Tolerances: the FP64 kernels must agree with the oracle's FP64 egb() to
1e-9 relative on energies and 1e-9 relative RMS on gradients (they differ
only in summation order and libm last-ulp effects); the mixed FP32-pair /
FP64-accumulate kernels to the project targets, 1e-6 relative per energy
term and 1e-5 relative RMS on gradients.  Pair lists must be identical.
"""
import numpy as np
import pytest

import oracle
from common import GOLDEN, generated_system, golden, oracle_pairs, read_pdb_xyz, rel, rel_rms

pytestmark = pytest.mark.gpu

FP64_E, FP64_G = 1e-9, 1e-9
MIX_E, MIX_G = 1e-6, 1e-5


def _compare(na, parm, arr, x, gb, cut, rgbmax, kappa=0.0, precision=0, frozen=None,
             e_tol=None, g_tol=None, check_list=True):
    e_tol = e_tol if e_tol is not None else (FP64_E if precision == 0 else MIX_E)
    g_tol = g_tol if g_tol is not None else (FP64_G if precision == 0 else MIX_G)
    ctx = na.GBContext(arr, x, gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa, precision=precision,
                       frozen=frozen)
    nent = ctx.nblist()
    orc = oracle.Oracle(parm, arr["natom"], gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa, frozen=frozen)
    r_nent = orc.nblist(x)
    assert nent == r_nent
    if check_list:
        ref = oracle_pairs(orc, arr["natom"])
        got = ctx.pairs()
        assert got.shape == ref.shape and np.array_equal(got, ref)
    epol, evdw, eel, g = ctx.egb()
    r_epol, r_evdw, r_eel, r_g, r_reff, r_sd = orc.egb(x)
    errs = dict(epol=rel(epol, r_epol), evdw=rel(evdw, r_evdw), eel=rel(eel, r_eel),
                grad=rel_rms(g, r_g))
    assert errs["epol"] <= e_tol and errs["evdw"] <= e_tol and errs["eel"] <= e_tol, errs
    assert errs["grad"] <= g_tol, errs
    if precision == 0:
        assert np.max(np.abs(ctx.reff() - r_reff) / r_reff) <= 1e-11
    return ctx, orc, errs


def test_aspgb8_known_answers(gpu):
    na = gpu
    g = golden()["aspgb8"]
    parm = na.read_parm(GOLDEN + "/asp.top")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    ctx, _, _ = _compare(na, parm, arr, x, gb=8, cut=100.0, rgbmax=100.0)
    reff = ctx.reff()
    for r in g["radii"]:
        assert abs(reff[r["atom"] - 1] - r["reff"]) <= 6e-8
    epol, evdw, eel, _ = ctx.egb()
    assert abs(epol - g["terms"]["egb"]["value"]) <= 1e-9
    assert abs(evdw - g["terms"]["enb"]["value"]) <= 1e-9
    assert abs(eel - g["terms"]["eel"]["value"]) <= 1e-9


@pytest.mark.parametrize("case", ["gb5trx", "gb8trx", "gb7trx", "gbtrx"])
@pytest.mark.parametrize("precision", [0, 1])
def test_trx_models(gpu, case, precision):
    na = gpu
    o = golden()[case]["options"]
    parm = na.read_parm(GOLDEN + "/prmtop.gb2")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/trxox.2.4ns.pdb")
    _compare(na, parm, arr, x, gb=o["gb"], cut=o["cut"], rgbmax=o["rgbmax"], kappa=o["kappa"],
             precision=precision)


@pytest.mark.parametrize("natom,gb,cut,rgbmax,kappa", [
    (600, 5, 999.0, 999.0, 0.0),
    (600, 1, 999.0, 999.0, 0.10395),
    (2500, 5, 999.0, 999.0, 0.0),
    (2500, 8, 12.0, 10.0, 0.127315),
    (2500, 2, 15.0, 15.0, 0.14701),
])
@pytest.mark.parametrize("precision", [0, 1])
def test_generated(gpu, gen_dir, natom, gb, cut, rgbmax, kappa, precision):
    na = gpu
    parm, arr, x = generated_system(natom, 11, gen_dir)
    _compare(na, parm, arr, x, gb=gb, cut=cut, rgbmax=rgbmax, kappa=kappa, precision=precision)


def test_frozen_atoms(gpu, gen_dir):
    """Not-both-frozen list rule, no self term and no chain rule for frozen atoms."""
    na = gpu
    parm, arr, x = generated_system(600, 3, gen_dir)
    frozen = np.zeros(arr["natom"], np.int32)
    frozen[::3] = 1
    frozen[100:160] = 1
    _compare(na, parm, arr, x, gb=5, cut=14.0, rgbmax=12.0, frozen=frozen)


def test_list_frozen_between_rebuilds(gpu, gen_dir):
    """Membership is fixed at build time: after the atoms move, both sides
    keep the old list until the next rebuild (eff.c:2564-2578)."""
    na = gpu
    parm, arr, x0 = generated_system(2500, 5, gen_dir)
    ctx = na.GBContext(arr, x0, gb=5, cut=10.0, rgbmax=10.0, kappa=0.1)
    orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=10.0, rgbmax=10.0, kappa=0.1)
    assert ctx.nblist() == orc.nblist(x0)
    rng = np.random.default_rng(0)
    x1 = x0 + rng.normal(0, 0.3, x0.shape)
    ctx.set_coords(x1)
    e = ctx.egb()
    r = orc.egb(x1)
    assert rel(e[0], r[0]) <= FP64_E and rel(e[2], r[2]) <= FP64_E and rel_rms(e[3], r[3]) <= FP64_G
    # rebuild at x1 on both sides
    assert ctx.nblist() == orc.nblist(x1)
    e = ctx.egb()
    r = orc.egb(x1)
    assert rel(e[0], r[0]) <= FP64_E and rel_rms(e[3], r[3]) <= FP64_G


def test_mme_full_path(gpu, gen_dir):
    """Host mme() (bonded + 1-4 + GB on the device) against the oracle's mme34."""
    na = gpu
    parm, arr, x = generated_system(2500, 7, gen_dir)
    na.lib.mm_reset_options()
    assert na.lib.mm_options(b"gb=8, cut=20.0, rgbmax=18.0, kappa=0.127315, diel=C, ntpr=100000") == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    f = np.zeros(3 * arr["natom"])
    it = na.C.c_int(-2)
    xc = np.ascontiguousarray(x)
    # first call builds the list (no list yet)
    e0 = na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
    ene = np.zeros(13)
    frms = na.C.c_double()
    na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
    # the oracle reads the same (gb-8-modified) topology
    orc = oracle.Oracle(parm, arr["natom"], gb=8, cut=20.0, rgbmax=18.0, kappa=0.127315)
    r_ene, r_frms, r_g = orc.mme(x)
    assert rel(e0, r_ene[0]) <= 1e-9
    for k in (1, 2, 3, 4, 5, 7, 8, 10):
        assert abs(ene[k] - r_ene[k]) <= 1e-9 * max(1.0, abs(r_ene[k])), (k, ene[k], r_ene[k])
    assert rel_rms(f, r_g) <= 1e-9
    assert rel(frms.value, r_frms) <= 1e-9
    it = na.C.c_int(-3)
    na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))


def test_cg_trajectory_on_gpu(gpu):
    """gb5trx.nab run entirely through the product: conjgrad over mme()."""
    na = gpu
    g = golden()["gb5trx"]
    parm = na.read_parm(GOLDEN + "/prmtop.gb2")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/trxox.2.4ns.pdb")
    na.lib.mm_reset_options()
    assert na.lib.mm_options(b"ntpr=1, gb=5, kappa=0.14701, rgbmax=99.0, cut=99.0, diel=C ") == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    seen = []

    def cb(xp, fp, itp):
        e = na.lib.mme(xp, fp, itp)
        ene = np.zeros(13)
        fr = na.C.c_double()
        na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(fr))
        seen.append((itp[0], ene.copy(), fr.value))
        return e

    fn = na.MMEFUNC(cb)
    n = na.C.c_int(3 * arr["natom"])
    f = na.C.c_double()
    rms = na.C.c_double(3 * arr["natom"] * 0.0001)
    dfp = na.C.c_double(10.0)
    mx = na.C.c_int(10)
    xc = np.ascontiguousarray(x)
    ret = na.lib.conjgrad(xc.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), fn, na.C.byref(rms),
                          na.C.byref(dfp), na.C.byref(mx))
    assert ret == -3
    assert [s[0] for s in seen] == [w["iter"] for w in g["ff"]]
    for (it, ene, frms), w in zip(seen, g["ff"]):
        assert abs(ene[0] - w["total"]) <= 0.0051
        assert abs(ene[10] - w["genborn"]) <= 0.0051
        assert abs(ene[2] + ene[8] - w["elect"]) <= 0.0051
    it = na.C.c_int(-3)
    na.lib.mme(xc.ctypes.data_as(na.dp), xc.ctypes.data_as(na.dp), na.C.byref(it))


@pytest.mark.parametrize("precision", [0, 1])
def test_derivative_store_off_matches_oracle(gpu, gen_dir, precision):
    """Pass 3 recomputing the derivatives (store disabled) gives the same
    answers as with the pass-1 store."""
    na = gpu
    parm, arr, x = generated_system(2500, 11, gen_dir)
    ctx, orc, _ = _compare(na, parm, arr, x, gb=8, cut=12.0, rgbmax=10.0, kappa=0.127315,
                           precision=precision)
    need, cap, used = ctx.store_stats()
    assert used and 0 < need <= cap
    e1 = ctx.egb()
    ctx.set_derivative_store(0)
    ctx.nblist()
    assert ctx.store_stats()[2] is False
    e0 = ctx.egb()
    r = orc.egb(x)
    tol_e = FP64_E if precision == 0 else MIX_E
    tol_g = FP64_G if precision == 0 else MIX_G
    assert rel(e0[0], r[0]) <= tol_e and rel_rms(e0[3], r[3]) <= tol_g
    # same list, same pair arithmetic: the two paths agree far below the bar
    assert rel(e0[0], e1[0]) <= 1e-12 and rel_rms(e0[3], e1[3]) <= (1e-12 if precision == 0 else 1e-6)


def _split_types(arr, copies):
    """The same topology with every LJ atom type split into `copies`
    identical types (atom a takes copy a mod copies): same coefficients for
    every pair, a table copies^2 times larger."""
    nt = int(arr["ntypes"])
    t = np.arange(nt * copies) % nt
    out = dict(arr)
    out["ntypes"] = nt * copies
    out["iac"] = (np.asarray(arr["iac"]) + nt * (np.arange(arr["natom"]) % copies)).astype(np.int32)
    out["cno"] = np.asarray(arr["cno"]).reshape(nt, nt)[t[:, None], t[None, :]].ravel().astype(np.int32)
    return out


@pytest.mark.parametrize("precision", [0, 1])
def test_lj_table_lds_and_global_agree(gpu, gen_dir, precision):
    """Pass 2 reads the LJ coefficients from LDS when the molecule has at
    most 16 types and from global memory otherwise: the 5-type generated
    system and the same system with 20 identical-coefficient types (400
    table entries) give the same energies and gradient (FP64 sums aside),
    and both match the oracle on the 5-type topology."""
    na = gpu
    parm, arr, x = generated_system(2500, 11, gen_dir)
    big = _split_types(arr, 4)
    assert big["ntypes"] ** 2 > 256 >= arr["ntypes"] ** 2
    out = []
    for a in (arr, big):
        ctx = na.GBContext(a, x, gb=5, cut=14.0, rgbmax=12.0, kappa=0.1, precision=precision)
        ctx.nblist()
        out.append(ctx.egb())
        ctx.close()
    for c in range(3):
        assert rel(out[1][c], out[0][c]) <= 1e-12, c
    assert rel_rms(out[1][3], out[0][3]) <= 1e-12
    orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=14.0, rgbmax=12.0, kappa=0.1)
    orc.nblist(x)
    r = orc.egb(x)
    tol_e = FP64_E if precision == 0 else MIX_E
    assert rel(out[1][1], r[1]) <= tol_e and rel(out[1][2], r[2]) <= tol_e
    orc.close()


def test_derivative_store_outgrown_by_async_rebuild(gpu, gen_dir):
    """An asynchronous rebuild whose list no longer fits the store falls back
    to recomputation (and stays correct); the next synchronous build grows
    the store and uses it again."""
    na = gpu
    parm, arr, x = generated_system(2500, 5, gen_dir)
    ctx = na.GBContext(arr, x, gb=5, cut=10.0, rgbmax=10.0, kappa=0.1)
    ctx.nblist()
    need0, cap0, used0 = ctx.store_stats()
    assert used0
    xc = x * 0.8  # denser: about twice the pairs
    ctx.set_coords(xc)
    ctx.nblist_async()
    need1, cap1, used1 = ctx.store_stats()
    assert need1 > cap0 and cap1 == cap0 and not used1
    orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=10.0, rgbmax=10.0, kappa=0.1)
    orc.nblist(xc)
    r = orc.egb(xc)
    e = ctx.egb()
    assert rel(e[0], r[0]) <= FP64_E and rel(e[2], r[2]) <= FP64_E and rel_rms(e[3], r[3]) <= FP64_G
    ctx.nblist()
    need2, cap2, used2 = ctx.store_stats()
    assert used2 and cap2 >= need2 == need1
    e = ctx.egb()
    assert rel(e[0], r[0]) <= FP64_E and rel_rms(e[3], r[3]) <= FP64_G


@pytest.mark.parametrize("precision", [0, 1])
def test_step_split_matches_oracle(gpu, gen_dir, precision):
    """Every step split (each tile pair's steps in 1, 2, 4, 8 or 16 work items),
    per pass and mixed across passes, with the derivative store and without
    it, gives the oracle's energies and gradient: the chunks' partial sums,
    and the store offsets each chunk starts from, add up to the whole
    sweep."""
    na = gpu
    parm, arr, x = generated_system(600, 12, gen_dir)
    ctx, orc, _ = _compare(na, parm, arr, x, gb=8, cut=12.0, rgbmax=10.0, kappa=0.127315,
                           precision=precision)
    assert ctx.split() == (8, 8, 8)  # 10 tiles, 55 tile pairs: the default splits
    r = orc.egb(x)
    tol_e = FP64_E if precision == 0 else MIX_E
    tol_g = FP64_G if precision == 0 else MIX_G
    res = {}
    combos = [(1, 1, 1), (2, 2, 2), (4, 4, 4), (8, 8, 8), (16, 16, 16), (8, 1, 2), (1, 4, 8), (16, 2, 1)]
    for store in (1, 0):
        ctx.set_derivative_store(store)
        ctx.nblist()
        assert ctx.store_stats()[2] is bool(store)
        for sp in combos:
            ctx.set_split(*sp)
            assert ctx.split() == sp
            e = ctx.egb()
            for c in range(3):
                assert rel(e[c], r[c]) <= tol_e, (store, sp, c)
            assert rel_rms(e[3], r[3]) <= tol_g, (store, sp)
            res[(store, sp)] = e
    # same arithmetic per pair, only the order of the sums differs
    base = res[(1, (1, 1, 1))]
    for key, e in res.items():
        assert rel(e[0], base[0]) <= 1e-12, key
        assert rel_rms(e[3], base[3]) <= (1e-12 if precision == 0 else 1e-6), key
    for bad in ((3, 1, 1), (1, 32, 1), (1, 1, -2)):
        with pytest.raises(na.GBHipError):
            ctx.set_split(*bad)
    ctx.set_split(0, 0, 0)
    assert ctx.split() == (8, 8, 8)
    ctx.close()
    orc.close()


@pytest.mark.parametrize("gb,cut,rgbmax", [(8, 14.0, 12.0), (5, 20.0, 20.0)])
def test_queued_pass1_matches_oracle(gpu, gen_dir, gb, cut, rgbmax):
    """Pass 1 of the mixed kernels evaluates work items whose listed pairs
    fill little of their lane-steps from a queue of those pairs, 64 at a
    time: never (fill 0), by default (0.7) and always (2), with the
    derivative store and without, whole and split, the energies and
    gradient match the oracle and each other (the same pair terms summed in
    another order)."""
    na = gpu
    parm, arr, x = generated_system(2500, 21, gen_dir)
    ctx, orc, _ = _compare(na, parm, arr, x, gb=gb, cut=cut, rgbmax=rgbmax, kappa=0.127315, precision=1)
    r = orc.egb(x)
    res = {}
    for store in (1, 0):
        ctx.set_derivative_store(store)
        ctx.nblist()
        for sp in ((1, 1, 1), (4, 2, 2)):
            ctx.set_split(*sp)
            for fill in (0.0, 0.7, 2.0):
                ctx.set_queue_fill(fill)
                e = ctx.egb()
                for c in range(3):
                    assert rel(e[c], r[c]) <= MIX_E, (store, sp, fill, c)
                assert rel_rms(e[3], r[3]) <= MIX_G, (store, sp, fill)
                res[(store, sp, fill)] = e
    # the queued and the swept form are separately compiled FP32 sequences
    # of the same terms (their last bits may differ), summed in FP64
    base = res[(1, (1, 1, 1), 0.0)]
    for key, e in res.items():
        assert rel(e[0], base[0]) <= 1e-7, key
        assert rel_rms(e[3], base[3]) <= 1e-6, key
    for bad in ((0.5, 1), (-1.0, 0), (3.0, 0)):
        with pytest.raises(na.GBHipError):
            ctx.set_queue_fill(bad[0], pass_=bad[1])
    ctx.close()
    orc.close()


def _two_clusters(n_each, gap, seed):
    """Two clusters of n_each atoms (random in spheres of 5 A, no two atoms
    closer than 1.6 A), the second shifted along x so that the closest
    pair across them is `gap` A apart."""
    rng = np.random.default_rng(seed)

    def cluster():
        pts = []
        while len(pts) < n_each:
            p = rng.uniform(-5.0, 5.0, 3)
            if np.linalg.norm(p) <= 5.0 and all(np.linalg.norm(p - q) >= 1.6 for q in pts):
                pts.append(p)
        return np.array(pts)

    a, b = cluster(), cluster()
    # shift b along x until the closest cross pair is at `gap`
    lo, hi = 0.0, 40.0
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        bb = b + np.array([mid, 0.0, 0.0])
        dm = np.sqrt(((a[:, None, :] - bb[None, :, :]) ** 2).sum(-1)).min()
        lo, hi = (mid, hi) if dm < gap else (lo, mid)
    return np.concatenate([a, b + np.array([hi, 0.0, 0.0])]).ravel()


@pytest.mark.parametrize("gb,rgbmax", [(8, 999.0), (8, 12.0), (5, 999.0)])
@pytest.mark.parametrize("precision", [0, 1])
def test_visit_kinds_at_their_thresholds(gpu, gen_dir, gb, rgbmax, precision):
    """Pass 1 compiles each tile pair's sweep without the near descreening
    cases (dij <= 4 sj) and/or the neck term (dij < rbi + rbj + 6.8 A) when
    a lower bound of its pair distances rules them out.  Two 64-atom
    clusters (one tile each) are placed with their closest pair at gaps
    from 3 to 12 A in quarter steps, across every threshold the
    classification uses; every placement matches the oracle."""
    na = gpu
    parm, arr, _ = generated_system(128, 77, gen_dir)
    for k, gap in enumerate(np.arange(3.0, 12.01, 0.25)):
        x = _two_clusters(64, gap, 1000 + k)
        ctx, orc, _ = _compare(na, parm, arr, x, gb=gb, cut=999.0, rgbmax=rgbmax, kappa=0.1,
                               precision=precision, check_list=False)
        if precision == 1:
            # the same kinds in the queued form of pass 1
            ctx.set_queue_fill(2.0)
            e = ctx.egb()
            r = orc.egb(x)
            for c in range(3):
                assert rel(e[c], r[c]) <= MIX_E, (gap, c)
            assert rel_rms(e[3], r[3]) <= MIX_G, gap
        ctx.close()
        orc.close()


def test_default_step_split_by_size(gpu, gen_dir):
    """The default step split per pass, from the tile-pair capacity: 8 in
    every pass below 128 tile pairs, pass 1 at most 4 above, passes 2 and 3
    split while they have fewer than 6 work items per CU, large systems
    unsplit (the headline 25,000-atom workload runs the unsplit kernels)."""
    na = gpu
    want = {600: (8, 8, 8), 2500: (4, 2, 2), 7000: (4, 1, 1), 25000: (1, 1, 1)}
    for natom, split in want.items():
        parm, arr, x = generated_system(natom, 30, gen_dir)
        ctx = na.GBContext(arr, x, gb=5, cut=20.0, rgbmax=20.0)
        ctx.nblist()
        nt, ntp = ctx.tile_stats()
        assert ctx.split() == split, (natom, nt, ntp, ctx.split())
        ctx.close()


TERM_INDEX = {"bond": 3, "angle": 4, "dihedral": 5, "enb14": 7, "eel14": 8, "enb": 1,
              "eel": 2, "egb": 10, "econs": 9, "esurf": 11, "Total": 0}


def test_aspgb8_mme_terms_on_gpu(gpu):
    """The product's whole mme() (bonded, 1-4 and GB terms on the device)
    reproduces every energy term and frms of test/nab/aspgb8.out at its
    printed 9 decimals."""
    na = gpu
    g = golden()["aspgb8"]
    parm = na.read_parm(GOLDEN + "/asp.top")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    na.lib.mm_reset_options()
    assert na.lib.mm_options(b"gb=8, cut=100.0, rgbmax=100.0, kappa=0.0, diel=C, ntpr=100000") == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
    f = np.zeros(3 * arr["natom"])
    xc = np.ascontiguousarray(x)
    it = na.C.c_int(-2)
    na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
    ene = np.zeros(13)
    frms = na.C.c_double()
    na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
    for name, t in g["terms"].items():
        got = frms.value if name == "frms" else ene[TERM_INDEX[name]]
        assert abs(got - t["value"]) <= 6e-10 + 1e-12 * abs(t["value"]), (name, got, t)
    it = na.C.c_int(-3)
    na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))


def test_general_torsion_phases_match_oracle(gpu, gen_dir):
    """Torsion terms whose phase is neither 0 nor pi take the general branch
    (E = Pk (1 + cos(Pn phi - phase)) with the signed Pn of a series,
    eff.c:591-599): the generated topology's four torsion types are given
    phases 0.7, 2.0, -1.1 and pi, and the device's mme() matches the
    oracle's mme34 per term and in the gradient to 1e-9."""
    na = gpu
    parm, arr, x = generated_system(2500, 9, gen_dir)
    s = parm.contents
    saved = [s.Phase[k] for k in range(4)]
    for k, ph in enumerate((0.7, 2.0, -1.1, 3.14159400)):
        s.Phase[k] = ph
    try:
        na.lib.mm_reset_options()
        assert na.lib.mm_options(b"gb=5, cut=16.0, rgbmax=16.0, kappa=0.1, diel=C, ntpr=100000") == 0
        assert na.lib.mme_init_sff(parm, None, None, None, None) == 0
        f = np.zeros(3 * arr["natom"])
        xc = np.ascontiguousarray(x)
        it = na.C.c_int(-2)
        e0 = na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
        ene = np.zeros(13)
        frms = na.C.c_double()
        na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(frms))
        orc = oracle.Oracle(parm, arr["natom"], gb=5, cut=16.0, rgbmax=16.0, kappa=0.1)
        r_ene, r_frms, r_g = orc.mme(x)
        orc.close()
        assert rel(e0, r_ene[0]) <= 1e-9
        for k in (3, 4, 5, 7, 8):
            assert abs(ene[k] - r_ene[k]) <= 1e-9 * max(1.0, abs(r_ene[k])), (k, ene[k], r_ene[k])
        assert rel_rms(f, r_g) <= 1e-9
        it = na.C.c_int(-3)
        na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
    finally:
        for k in range(4):
            s.Phase[k] = saved[k]


@pytest.mark.parametrize("precision", [0, 1])
def test_gbrna_old_format_gb1(gpu, precision):
    """gb=1 (HCT) with salt on the decamer of test/nab/gbrna.nab, read from
    its old-format prmtop: pair list, energies and gradient against the
    oracle (whose gb=1 CG trajectory reproduces gbrna.out.check)."""
    na = gpu
    parm = na.read_parm(GOLDEN + "/gbrna.prmtop")
    arr = na.parm_arrays(parm)
    x, _ = na.read_rst7(GOLDEN + "/gbrna.rst7", arr["natom"])
    _compare(na, parm, arr, x, gb=1, cut=99.0, rgbmax=99.0, kappa=0.10395, precision=precision)


@pytest.mark.parametrize("diel,dielc", [("R", 1.0), ("C", 2.0), ("R", 2.0)])
def test_dielectric_options_match_oracle(gpu, gen_dir, diel, dielc):
    """diel=R (1-4 electrostatics q_i q_j / (scee r^2)) and dielc != 1
    (charges scaled by 1/sqrt(dielc) at mme_init_sff): the device mme()
    per term and gradient against the oracle's mme34 to 1e-9, and for
    diel=R the device's 1-4 electrostatics scale by 1/4 when every distance
    doubles."""
    na = gpu
    parm, arr, x = generated_system(600, 12, gen_dir)
    n = arr["natom"]
    # the oracle copies the charges before mme_init_sff scales them in place
    orc = oracle.Oracle(parm, n, gb=5, cut=16.0, rgbmax=16.0, kappa=0.1, dielc=dielc, dield=0 if diel == "R" else 1)
    r_ene, r_frms, r_g = orc.mme(x)
    orc.close()
    na.lib.mm_reset_options()
    opts = "gb=5, cut=16.0, rgbmax=16.0, kappa=0.1, diel=%s, dielc=%.1f, ntpr=100000" % (diel, dielc)
    assert na.lib.mm_options(opts.encode()) == 0
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0

    def run(xc):
        xc = np.ascontiguousarray(xc)
        f = np.zeros(3 * n)
        it = na.C.c_int(-2)
        e = na.lib.mme(xc.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), na.C.byref(it))
        ene = np.zeros(13)
        fr = na.C.c_double()
        na.lib.mme_get_energies(ene.ctypes.data_as(na.dp), na.C.byref(fr))
        return e, ene, fr.value, f

    try:
        e0, ene, frms, f = run(x)
        assert rel(e0, r_ene[0]) <= 1e-9
        for k in (1, 2, 3, 4, 5, 7, 8, 10):
            assert abs(ene[k] - r_ene[k]) <= 1e-9 * max(1.0, abs(r_ene[k])), (k, ene[k], r_ene[k])
        assert rel_rms(f, r_g) <= 1e-9
        assert abs(frms - r_frms) <= 1e-9 * r_frms
        if diel == "R":
            _, ene2, _, _ = run(2.0 * x)
            assert abs(ene[8] / ene2[8] - 4.0) <= 1e-11, (ene[8], ene2[8])
    finally:
        it = na.C.c_int(-3)
        xc = np.ascontiguousarray(x)
        na.lib.mme(xc.ctypes.data_as(na.dp), np.zeros(3 * n).ctypes.data_as(na.dp), na.C.byref(it))
