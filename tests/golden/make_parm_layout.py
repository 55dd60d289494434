"""Record the byte layout of the reference's PARMSTRUCT_T (sff.h:40-62).

Run in a container where dacase/nabc is mounted at /root/reference: it
compiles a few lines of C that include the reference's src/sff/sff.h and
print sizeof and the offset of every field, and writes them to
parm_layout.json.  tests/test_abi_cpu.py compares the product's
include/sff_gb.h (and the Python mirror of it) with these numbers, so a
program compiled against libsff's header reads the same fields at the same
offsets from a topology the product's rdparm returns.
"""
import json
import os
import subprocess
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
REF_INC = "/root/reference/src/sff"

FIELDS = ("ititl IfBox Nmxrs IfCap Natom Ntypes Nbonh Mbona Ntheth Mtheta Nphih Mphia Nhparm Nparm Nnb "
          "Nres Nbona Ntheta Nphia Numbnd Numang Nptra Natyp Nphb Nat3 Ntype2d Nttyp Nspm Iptres Nspsol "
          "Ipatm Natcap Numextra AtomNames ResNames AtomType AtomTree Charges Masses Rk Req Tk Teq Pk Pn "
          "Phase Solty Cn1 Cn2 HB12 HB10 Rborn Fs Box Cutcap Xcap Ycap Zcap Fsmax Iac Iblo Cno Ipres "
          "ExclAt TreeJoin AtomRes BondHAt1 BondHAt2 BondHNum BondAt1 BondAt2 BondNum AngleHAt1 "
          "AngleHAt2 AngleHAt3 AngleHNum AngleAt1 AngleAt2 AngleAt3 AngleNum DihHAt1 DihHAt2 DihHAt3 "
          "DihHAt4 DihHNum DihAt1 DihAt2 DihAt3 DihAt4 DihNum Boundary AtomicNum N14pairs N14pairlist "
          "Gvdw Scee Scnb N14sceelist N14scnblist Nstrand Ncomplex Ipstrand Ipcomplex").split()


def layout_program(header):
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "%s"' % header, "int main(void) {",
             '  printf("sizeof %zu\\n", sizeof(PARMSTRUCT_T));']
    lines += ['  printf("%s %%zu\\n", offsetof(PARMSTRUCT_T, %s));' % (f, f) for f in FIELDS]
    lines += ["  return 0;", "}"]
    return "\n".join(lines) + "\n"


def measure(header, incdir):
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        with open(src, "w") as fp:
            fp.write(layout_program(header))
        subprocess.run(["gcc", "-I", incdir, src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def main():
    lay = measure("sff.h", REF_INC)
    with open(os.path.join(HERE, "parm_layout.json"), "w") as fp:
        json.dump(dict(source="src/sff/sff.h:40-62 (gcc, x86-64)", fields=FIELDS, layout=lay), fp, indent=1)


if __name__ == "__main__":
    main()
