"""The host boundary is a link-level drop-in for libsff (no GPU).

- PARMSTRUCT_T of include/sff_gb.h has the byte layout of the reference's
  sff.h:40-62 (sizeof and every field offset, recorded from the
  reference's own header by tests/golden/make_parm_layout.py), and so has
  the Python mirror of it;
- a libsff client (tests/c_driver/sff_driver.c: rdparm, getxv, mm_options,
  mme_init_sff, conjgrad, md, putxv, free_prm, nabout) written against the
  header name sff.h compiles against include/ and links with
  libnabc_amd.so; its parse-only mode runs here (no device work).
"""
import json
import os
import subprocess

import pytest

from common import GOLDEN, REPO

INC = os.path.join(REPO, "include")
LIBDIR = os.path.join(REPO, "nabc_amd")


def _fixture():
    with open(os.path.join(GOLDEN, "parm_layout.json")) as fp:
        return json.load(fp)


def test_parmstruct_layout_matches_reference(tmp_path):
    fx = _fixture()
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "sff_gb.h"', "int main(void) {",
             '  printf("sizeof %zu\\n", sizeof(PARMSTRUCT_T));']
    lines += ['  printf("%s %%zu\\n", offsetof(PARMSTRUCT_T, %s));' % (f, f) for f in fx["fields"]]
    lines += ["  return 0;", "}"]
    src, exe = tmp_path / "lay.c", tmp_path / "lay"
    src.write_text("\n".join(lines) + "\n")
    subprocess.run(["gcc", "-I", INC, str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    got = {k: int(v) for k, v in (line.split() for line in out.splitlines())}
    assert got == fx["layout"]


def test_python_mirror_layout():
    import nabc_amd as na
    fx = _fixture()
    assert na.C.sizeof(na.Parm) == fx["layout"]["sizeof"]
    for f in fx["fields"]:
        assert getattr(na.Parm, f).offset == fx["layout"][f], f


def test_libsff_client_compiles_links_and_reads(tmp_path):
    exe = tmp_path / "sff_driver"
    subprocess.run(["gcc", "-O2", "-Wall", "-Werror", "-std=gnu99", "-I", INC,
                    os.path.join(REPO, "tests", "c_driver", "sff_driver.c"), "-o", str(exe),
                    "-L", LIBDIR, "-lnabc_amd", "-Wl,-rpath," + LIBDIR, "-lm"], check=True)
    # every libsff entry point the client calls resolves in the product library
    nm = subprocess.run(["nm", "-D", "--defined-only", os.path.join(LIBDIR, "libnabc_amd.so")],
                        check=True, capture_output=True, text=True).stdout
    syms = {line.split()[-1] for line in nm.splitlines() if line.strip()}
    for s in ("rdparm", "wrparm", "free_prm", "getxv", "putxv", "mm_options", "mme_init_sff", "mme", "md",
              "conjgrad", "nabout"):
        assert s in syms, s
    out = subprocess.run([str(exe), GOLDEN + "/gbrna.prmtop", GOLDEN + "/gbrna.rst7", "--parse"],
                         check=True, capture_output=True, text=True).stdout
    f = out.split()
    got = dict(zip(f[0::2], f[1::2]))
    assert got["natom"] == "640" and got["nres"] == "20" and got["ipres_end"] == "640"
    assert got["nat3"] == "1920" and got["nspm"] == "1" and got["name0"] == "H5T"
    assert got["res0"] == "C5" and got["x0"] == "1.4800" and got["fsmax"] == "1.513600"


@pytest.mark.parametrize("top", ["asp.top", "prmtop.gb2", "gbrna.prmtop"])
def test_wrparm_rdparm_round_trip(tmp_path, top):
    """Written back as a %FLAG file and read again, a topology (new or old
    layout) keeps every array the energy path reads, 0-based residue
    pointers and names included."""
    import numpy as np
    import nabc_amd as na
    p = na.read_parm(os.path.join(GOLDEN, top))
    out = str(tmp_path / "w.prmtop")
    assert na.lib.wrparm(p, out.encode()) == 0
    q = na.read_parm(out)
    a, b = na.parm_arrays(p), na.parm_arrays(q)
    for k in a:
        assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
    s, t = p.contents, q.contents
    assert [s.Ipres[i] for i in range(s.Nres + 1)] == [t.Ipres[i] for i in range(t.Nres + 1)]
    rn = lambda z: bytes((na.C.c_char * 6 * z.Nres).from_address(z.ResNames))
    assert rn(s) == rn(t)
    na.lib.free_prm(p)
    na.lib.free_prm(q)
