"""The oracle against the reference's own known answers (CPU only).

Pins the CPU restatement in oracle/ before it is used as the checker:
  - test/nab/aspgb8.out: 24 igb=8 Born radii (7 decimals), every energy
    term and frms (9 decimals) of the current reference code;
  - test/nab/gb{5,8,7,}trx.out.check: the ff: lines of a 10-iteration CG
    minimisation of thioredoxin (1654 atoms) with igb 5 (salt), 8
    (rgbmax 15), 7 and 2, 2 decimals -- the whole trajectory, which also
    pins conjgrad() of the product's host library.
"""
import numpy as np
import pytest

import nabc_amd as na
import oracle
from common import GOLDEN, golden, read_pdb_xyz

TERM_INDEX = {"bond": 3, "angle": 4, "dihedral": 5, "enb14": 7, "eel14": 8, "enb": 1,
              "eel": 2, "egb": 10, "econs": 9, "esurf": 11, "Total": 0}


def _asp():
    parm = na.read_parm(GOLDEN + "/asp.top")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/asp.pdb")
    return parm, arr, x


def test_aspgb8_radii_and_terms():
    g = golden()["aspgb8"]
    parm, arr, x = _asp()
    o = oracle.Oracle(parm, arr["natom"], gb=8, cut=100.0, rgbmax=100.0)
    ene, frms, _ = o.mme(x)
    _, _, _, _, reff, _ = o.egb(x)
    for r in g["radii"]:
        assert abs(reff[r["atom"] - 1] - r["reff"]) <= 6e-8, r
    for name, t in g["terms"].items():
        got = frms if name == "frms" else ene[TERM_INDEX[name]]
        assert abs(got - t["value"]) <= 6e-10 + 1e-12 * abs(t["value"]), (name, got, t)


def test_aspgb8_check_file_differs_only_in_gb():
    """The older aspgb8.out.check agrees on every non-GB term."""
    g = golden()["aspgb8_check"]
    parm, arr, x = _asp()
    o = oracle.Oracle(parm, arr["natom"], gb=8, cut=100.0, rgbmax=100.0)
    ene, _, _ = o.mme(x)
    for name in ("bond", "angle", "dihedral", "enb14", "eel14", "enb", "eel"):
        assert abs(ene[TERM_INDEX[name]] - g["terms"][name]["value"]) <= 6e-10


def _ff_of(ene, frms):
    return dict(total=ene[0], bad=ene[3] + ene[4] + ene[5], vdw=ene[1] + ene[7],
                elect=ene[2] + ene[8], nonpolar=ene[9] + ene[11] + ene[12], genborn=ene[10],
                frms=frms)


def _match_ff(got, want):
    for k in ("total", "bad", "vdw", "elect", "genborn"):
        assert abs(got[k] - want[k]) <= 0.0051, (k, got[k], want)
    # %9.2e: three significant digits
    assert abs(got["frms"] - want["frms"]) <= 0.0051 * 10 ** np.floor(np.log10(want["frms"])), (got, want)


@pytest.mark.parametrize("case", ["gb5trx", "gb8trx", "gb7trx", "gbtrx"])
def test_trx_cg_trajectory(case):
    g = golden()[case]
    o = g["options"]
    parm = na.read_parm(GOLDEN + "/prmtop.gb2")
    arr = na.parm_arrays(parm)
    x, _ = read_pdb_xyz(GOLDEN + "/trxox.2.4ns.pdb")
    assert len(x) == 3 * arr["natom"]
    orc = oracle.Oracle(parm, arr["natom"], gb=o["gb"], cut=o["cut"], rgbmax=o["rgbmax"],
                        kappa=o["kappa"])
    # conjgrad(m_xyz, 3*natoms, fret, mme, dgrad, 10.0, 10) with
    # dgrad = 3*natoms*0.0001 (gb5trx.nab) -- the product's conjgrad over
    # the oracle's mme
    oracle.lib.orc_log_clear()
    oracle.lib.orc_set_current(orc.h, o["nsnb"])
    n = na.C.c_int(3 * arr["natom"])
    f = na.C.c_double()
    rms = na.C.c_double(3 * arr["natom"] * 0.0001)
    dfp = na.C.c_double(10.0)
    mx = na.C.c_int(10)
    xc = np.ascontiguousarray(x)
    cb = na.MMEFUNC(oracle.lib.orc_mme_cb)
    ret = na.lib.conjgrad(xc.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), cb,
                          na.C.byref(rms), na.C.byref(dfp), na.C.byref(mx))
    log = oracle.log_entries()
    assert ret == -3  # "conjgrad returns -3" (maxiter)
    assert [e[0] for e in log] == [w["iter"] for w in g["ff"]]
    for (it, ene, frms), want in zip(log, g["ff"]):
        _match_ff(_ff_of(ene, frms), want)


def test_gbrna_old_format_cg_trajectory():
    """gbrna.nab: the Dickerson decamer (640 atoms) from an old-format
    prmtop (no %FLAG sections; GB radii and screens from the first letter
    of each atom name, prm.c:1656-1741), gb=1 (HCT) with kappa 0.10395:
    the 11 ff: lines of its 10-iteration CG run (gbrna.out.check:12-22)
    from the oracle's mme through the product's conjgrad.  This pins the
    old-layout reader (every section, in the fixed order, or the energies
    would not come out) and the gb=1 totals."""
    g = golden()["gbrna"]
    o = g["options"]
    parm = na.read_parm(GOLDEN + "/gbrna.prmtop")
    arr = na.parm_arrays(parm)
    x, _ = na.read_rst7(GOLDEN + "/gbrna.rst7", arr["natom"])
    assert arr["natom"] == 640
    orc = oracle.Oracle(parm, arr["natom"], gb=o["gb"], cut=o["cut"], rgbmax=o["rgbmax"], kappa=o["kappa"])
    oracle.lib.orc_log_clear()
    oracle.lib.orc_set_current(orc.h, o["nsnb"])
    n = na.C.c_int(3 * arr["natom"])
    f = na.C.c_double()
    rms = na.C.c_double(3 * arr["natom"] * 0.0001)
    dfp = na.C.c_double(10.0)
    mx = na.C.c_int(10)
    xc = np.ascontiguousarray(x)
    cb = na.MMEFUNC(oracle.lib.orc_mme_cb)
    ret = na.lib.conjgrad(xc.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), cb,
                          na.C.byref(rms), na.C.byref(dfp), na.C.byref(mx))
    log = oracle.log_entries()
    oracle.lib.orc_set_current(None, 1)
    orc.close()
    assert ret == -3
    assert [e[0] for e in log] == [w["iter"] for w in g["ff"]]
    for (it, ene, frms), want in zip(log, g["ff"]):
        _match_ff(_ff_of(ene, frms), want)


def test_old_format_reader_fields():
    """The old-layout reader fills what rdparm fills: title, POINTERS, names
    (4 characters + blank), 0-based residue pointers with Ipres[Nres] =
    Natom, each atom's residue, the defaults a file without box or HCP
    sections gets, SCEE/SCNB defaults, and the Bondi radii of the name
    rule (e.g. H5T after nothing: oxygen's 0.8; H5'1 after C5': 1.3)."""
    p = na.read_parm(GOLDEN + "/gbrna.prmtop")
    s = p.contents
    assert s.ititl.startswith(b"PDB 5DNB, Dickerson decamer")
    assert (s.Natom, s.Ntypes, s.Nbonh, s.Mbona, s.Nres, s.Nat3, s.Nspm, s.Nstrand, s.Ncomplex) == \
        (640, 15, 220, 468, 20, 1920, 1, 1, 1)
    names = (na.C.c_char * 6 * s.Natom).from_address(s.AtomNames)
    assert bytes(names[0]) == b"H5T  \0" and bytes(names[3]) == b"H5'1 \0"
    assert [s.Ipres[i] for i in range(3)] == [0, 29, 60] and s.Ipres[s.Nres] == 640
    assert s.AtomRes[28] == 0 and s.AtomRes[29] == 1 and s.AtomRes[639] == 19
    assert s.Boundary[0] == 640 and s.Box[0] == 999999999.0
    assert all(abs(s.Scee[k] - 1 / 1.2) < 1e-15 and abs(s.Scnb[k] - 0.5) < 1e-15 for k in range(s.Nptra))
    assert (s.Rborn[0], s.Fs[0], s.Rborn[3], s.Rborn[1], s.Rborn[2]) == (0.8, 0.85, 1.3, 1.5, 1.7)
    fsmax = max(s.Fs[i] * (s.Rborn[i] - 0.09) for i in range(s.Natom))
    assert abs(s.Fsmax - fsmax) < 1e-15
    assert na.lib.free_prm(p) == 0
