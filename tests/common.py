"""Shared helpers for the test suite (fixture loading, generated systems)."""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")
REPO = os.path.dirname(HERE)


def golden():
    with open(os.path.join(GOLDEN, "golden.json")) as fp:
        return json.load(fp)


def read_pdb_xyz(path):
    """Coordinates of ATOM/HETATM records in file order (PDB columns 31-54)."""
    xyz, names = [], []
    for line in open(path):
        if line.startswith(("ATOM", "HETATM")):
            xyz.append((float(line[30:38]), float(line[38:46]), float(line[46:54])))
            names.append(line[12:16].strip())
    return np.array(xyz, dtype=np.float64).ravel(), names


_gen_cache = {}


def generated_system(natom, seed, tmpdir):
    """Write (once per session) and read back a synthetic system."""
    import nabc_amd as na
    key = (natom, seed)
    if key not in _gen_cache:
        top = os.path.join(tmpdir, "gen_%d_%d.prmtop" % key)
        rst = os.path.join(tmpdir, "gen_%d_%d.rst7" % key)
        na.generate(natom, seed, top, rst)
        _gen_cache[key] = (top, rst)
    top, rst = _gen_cache[key]
    parm = na.read_parm(top)
    arrs = na.parm_arrays(parm)
    x, _ = na.read_rst7(rst, arrs["natom"])
    return parm, arrs, x


def rel_rms(a, b):
    a = np.asarray(a); b = np.asarray(b)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def oracle_pairs(orc, natom):
    """The oracle's current list as a sorted (P, 2) array of (i, j), i < j:
    the upper part (j > i) of each atom's list (nblist.c:1146-1147)."""
    lo, up, flat = orc.lists()
    lo = lo.astype(np.int64)
    up = up.astype(np.int64)
    start = np.concatenate(([0], np.cumsum(lo + up)[:-1]))
    cum_up = np.concatenate(([0], np.cumsum(up)[:-1]))
    i = np.repeat(np.arange(natom, dtype=np.int64), up)
    pos = np.repeat(start + lo - cum_up, up) + np.arange(int(up.sum()), dtype=np.int64)
    j = flat[pos].astype(np.int64)
    assert np.all(j > i)
    key = i * natom + j
    key.sort()
    return np.stack([key // natom, key % natom], axis=1)


def gunzip_golden(name, tmpdir):
    """A gzipped fixture of tests/golden decompressed into tmpdir (once)."""
    import gzip
    import shutil
    dst = os.path.join(str(tmpdir), name[:-3] if name.endswith(".gz") else name)
    if not os.path.exists(dst):
        with gzip.open(os.path.join(GOLDEN, name), "rb") as fin, open(dst, "wb") as fout:
            shutil.copyfileobj(fin, fout)
    return dst


def trx_hcp_system(tmpdir):
    """test/thcp1.c's thioredoxin (2trx.parm7 / 2trx.rst7, 1653 atoms):
    (parm, arrays, x, hierarchy) with hierarchy = (ipres, ipstrand,
    ipcomplex) as the prmtop gives them."""
    import nabc_amd as na
    top = gunzip_golden("2trx.parm7.gz", tmpdir)
    rst = gunzip_golden("2trx.rst7.gz", tmpdir)
    parm = na.read_parm(top)
    arrs = na.parm_arrays(parm)
    x, _ = na.read_rst7(rst, arrs["natom"])
    return parm, arrs, x, parm_hierarchy(parm)


def parm_hierarchy(parm):
    """(ipres[nres + 1], ipstrand[nstrand], ipcomplex[ncomplex]) of a PARMSTRUCT_T."""
    p = parm.contents
    ipres = np.array([p.Ipres[i] for i in range(p.Nres + 1)], np.int32)
    ips = np.array([p.Ipstrand[i] for i in range(p.Nstrand)], np.int32)
    ipc = np.array([p.Ipcomplex[i] for i in range(p.Ncomplex)], np.int32)
    return ipres, ips, ipc


def layered_hierarchy(natom, res_atoms=20, strand_res=5, complex_strands=3):
    """Residues of res_atoms atoms, strands of strand_res residues,
    complexes of complex_strands strands (the last of each may be short)."""
    ipres = np.array(list(range(0, natom, res_atoms)) + [natom], np.int32)
    nres = len(ipres) - 1
    ips = np.arange(1, nres + 1, strand_res, dtype=np.int32)
    ipc = np.arange(1, len(ips) + 1, complex_strands, dtype=np.int32)
    return ipres, ips, ipc
