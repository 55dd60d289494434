"""Host-side checks that need no GPU: the C-ABI library's exports, prmtop and
rst7 round trips, option parsing, the synthetic-system generator.

These exercise the product's CPU code (prm.c, rst7.c, options.c, gen.c) and
the shared library's symbol table; no kernel is launched.
"""
import ctypes as C
import glob
import os
import re

import numpy as np
import pytest

from common import GOLDEN, REPO


def _declared_functions():
    """Every function declared in include/*.h, as (header, name)."""
    out = []
    for h in sorted(glob.glob(os.path.join(REPO, "include", "*.h"))):
        s = open(h).read()
        s = re.sub(r"/\*.*?\*/", "", s, flags=re.S)
        s = re.sub(r"//[^\n]*", "", s)
        s = re.sub(r"#[^\n]*", "", s)
        s = s.replace('extern "C" {', ";").replace('extern "C"', "")
        s = re.sub(r"\{[^{}]*\}", " ", s)  # struct bodies
        for st in s.split(";"):
            st = " ".join(st.replace("}", " ").split())
            if not st or st.startswith("typedef"):
                continue
            m = re.match(r"^[A-Za-z_][\w\s\*]*?\b([A-Za-z_]\w*)\s*\((.*)\)$", st)
            if m:
                out.append((os.path.basename(h), m.group(1)))
    return out


def test_library_exports_every_declared_function():
    import nabc_amd as na
    decl = _declared_functions()
    names = {n for _, n in decl}
    # the drop-in entry points of the reference's C API must be among them
    for n in ("rdparm", "wrparm", "getxv", "putxv", "mm_options", "mme_init_sff", "mme", "md",
              "conjgrad", "gbhip_egb", "gbhip_nblist", "gbhip_ctx_create", "gbhip_batch_eval",
              "gbserve_run"):
        assert n in names, n
    lib = C.CDLL(na.LIB_PATH)
    missing = [(h, n) for h, n in decl if not hasattr(lib, n)]
    assert not missing, missing
    assert len(decl) >= 50


def test_device_count_without_gpu_is_zero_or_more():
    import nabc_amd as na
    assert na.device_count() >= 0


def test_prmtop_write_read_round_trip(tmp_path):
    import nabc_amd as na
    p = na.read_parm(os.path.join(GOLDEN, "asp.top"))
    a = na.parm_arrays(p)
    out = str(tmp_path / "asp_copy.prmtop")
    assert na.lib.wrparm(p, out.encode()) == 0
    q = na.read_parm(out)
    b = na.parm_arrays(q)
    assert a.keys() == b.keys()
    for k in a:
        if isinstance(a[k], np.ndarray):
            assert a[k].shape == b[k].shape, k
            # %FORMAT(5E16.8) keeps 9 significant digits
            np.testing.assert_allclose(b[k], a[k], rtol=1e-8, atol=0, err_msg=k)
        else:
            assert a[k] == pytest.approx(b[k], rel=1e-8), k


def test_generated_prmtop_round_trip_is_exact_in_integers(tmp_path):
    import nabc_amd as na
    top, rst = str(tmp_path / "g.prmtop"), str(tmp_path / "g.rst7")
    na.generate(300, 7, top, rst)
    p = na.read_parm(top)
    a = na.parm_arrays(p)
    out = str(tmp_path / "g2.prmtop")
    assert na.lib.wrparm(p, out.encode()) == 0
    b = na.parm_arrays(na.read_parm(out))
    for k in a:
        if isinstance(a[k], np.ndarray) and a[k].dtype.kind == "i":
            np.testing.assert_array_equal(a[k], b[k], err_msg=k)


def test_rst7_round_trip(tmp_path):
    import nabc_amd as na
    rng = np.random.default_rng(3)
    n = 37
    x = rng.uniform(-500, 500, 3 * n)
    v = rng.normal(0, 1, 3 * n)
    path = str(tmp_path / "t.rst7")
    assert na.lib.putxv(path.encode(), b"round trip", n, 12.5, na._p(x), na._p(v)) == 0
    x2, v2 = na.read_rst7(path, n)
    # rst7 keeps 7 decimals (6F12.7)
    np.testing.assert_allclose(x2, x, atol=6e-8, rtol=0)
    np.testing.assert_allclose(v2, v, atol=6e-8, rtol=0)


def test_rst7_wrong_atom_count_is_an_error(tmp_path):
    import nabc_amd as na
    x = np.zeros(3 * 4)
    path = str(tmp_path / "t.rst7")
    assert na.lib.putxv(path.encode(), b"t", 4, 0.0, na._p(x), na._p(x)) == 0
    with pytest.raises(IOError):
        na.read_rst7(path, 5)


def test_missing_prmtop_is_an_error():
    import nabc_amd as na
    with pytest.raises(IOError):
        na.read_parm("/nonexistent/file.prmtop")


def test_mm_options_accepts_reference_keywords_and_rejects_others():
    import nabc_amd as na
    na.lib.mm_reset_options()
    assert na.lib.mm_options(b"cut=12.0, rgbmax=10.0, gb=8, kappa=0.1, nsnb=5, ntpr=10") == 0
    assert na.lib.mm_options(b"diel=C, dielc=1.0, gbprec=mixed, md_seed=4") == 0
    assert na.lib.mm_options(b"no_such_option=3") != 0
    assert b"no_such_option" in na.lib.mm_get_error()
    na.lib.mm_reset_options()


def test_generator_is_deterministic_and_respects_min_distance(tmp_path):
    import nabc_amd as na
    p1, a1, x1 = na.build_system(500, 11)
    p2, a2, x2 = na.build_system(500, 11)
    p3, a3, x3 = na.build_system(500, 12)
    np.testing.assert_array_equal(x1, x2)
    np.testing.assert_array_equal(a1["charge"], a2["charge"])
    assert not np.array_equal(x1, x3)
    X = x1.reshape(-1, 3)
    d = np.sqrt(((X[:, None, :] - X[None, :, :]) ** 2).sum(-1))
    # a chain: consecutive atoms 1.5 A apart, every other pair at least 1.75 A
    np.testing.assert_allclose(np.diagonal(d, 1), 1.5, atol=1e-9)
    far = np.triu(np.ones_like(d, dtype=bool), 2)
    assert d[far].min() >= 1.75 - 1e-9
    # inside the sphere of mean density 0.1 atoms/A^3
    R = (3 * 500 / (4 * np.pi * 0.1)) ** (1 / 3)
    assert np.sqrt(((X - X[0]) ** 2).sum(1)).max() <= 2 * R
    # net charge zero (prmtop unit e * 18.2223)
    assert abs(a1["charge"].sum()) < 1e-9 * 18.2223 * 500
    # written files read back to the same system
    top, rst = str(tmp_path / "a.prmtop"), str(tmp_path / "a.rst7")
    na.generate(500, 11, top, rst)
    b = na.parm_arrays(na.read_parm(top))
    xr, _ = na.read_rst7(rst, 500)
    np.testing.assert_allclose(xr, x1, atol=6e-8)
    np.testing.assert_allclose(b["charge"], a1["charge"], rtol=1e-8)


def test_kappa_from_salt_matches_reference_inputs():
    import nabc_amd as na
    # kappa values written in the reference's inputs for 0.1/0.15/0.2 M salt
    # (test/nab/txmin.nab:46, benchmarks/p450.nab:8, test/nab/gbtrx.nab:8)
    assert na.kappa_from_salt(0.1) == pytest.approx(0.10395, abs=1e-5)
    assert na.kappa_from_salt(0.15) == pytest.approx(0.127315, abs=1e-5)
    assert na.kappa_from_salt(0.2) == pytest.approx(0.14701, abs=1e-5)


def test_footprint_grows_with_atoms():
    import nabc_amd as na
    f = na.lib.gbhip_footprint_bytes
    f.restype = C.c_longlong
    f.argtypes = [C.c_int, C.c_int]
    a, b = f(1000, 0), f(10000, 0)
    assert 0 < a < b
    assert f(1000, 100) >= a
