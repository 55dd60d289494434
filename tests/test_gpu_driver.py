"""A libsff client program runs on the GPU path unchanged.

tests/c_driver/sff_driver (built by __graft_entry__.build() from
sff_driver.c against include/sff.h, linked with libnabc_amd.so) does what
test/nab/gbrna.nab does -- rdparm of an old-format prmtop, mm_options
"ntpr=1, gb=1, kappa=0.10395, rgbmax=99., cut=99.0, diel=C", mme_init_sff,
conjgrad(x, 3N, fret, mme, 3N*1e-4, 10, 10) -- then 20 steps of md() and
putxv.  Its ff: lines must be those of gbrna.out.check:12-22 at their
printed precision (2 decimals; frms 3 significant digits) and conjgrad must
return -3, as in the reference's run.
"""
import os
import subprocess

import numpy as np
import pytest

from common import GOLDEN, REPO, golden

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(REPO, "tests", "c_driver", "sff_driver")


def test_gbrna_through_libsff_client(gpu, tmp_path):
    assert os.path.exists(DRIVER), "tests/c_driver/sff_driver is not built (__graft_entry__.build())"
    g = golden()["gbrna"]
    out_rst = str(tmp_path / "final.rst7")
    r = subprocess.run([DRIVER, GOLDEN + "/gbrna.prmtop", GOLDEN + "/gbrna.rst7",
                        "ntpr=1, gb=1, kappa=0.10395, rgbmax=99., cut=99.0, diel=C, ntpr_md=5",
                        "10", "20", out_rst], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout[-2000:], r.stderr[-2000:])
    ff = [line.split() for line in r.stdout.splitlines() if line.startswith("ff:")]
    assert [int(f[1]) for f in ff] == [w["iter"] for w in g["ff"]]
    for f, w in zip(ff, g["ff"]):
        got = dict(total=float(f[2]), bad=float(f[3]), vdw=float(f[4]), elect=float(f[5]),
                   genborn=float(f[7]), frms=float(f[8]))
        for k in ("total", "bad", "vdw", "elect", "genborn"):
            assert abs(got[k] - w[k]) <= 0.0051, (k, got, w)
        assert abs(got["frms"] - w["frms"]) <= 0.0051 * 10 ** np.floor(np.log10(w["frms"])), (got, w)
    assert "conjgrad returns -3" in r.stdout
    md = [line.split() for line in r.stdout.splitlines() if line.startswith("md:")]
    assert len(md) >= 4 and all(np.isfinite(float(f[3])) for f in md)
    assert os.path.getsize(out_rst) > 0
