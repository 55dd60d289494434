"""Collect the reference's own known-answer data for the GB path as fixtures.

Run in a container where dacase/nabc is mounted at /root/reference.  It
copies the input files the reference's tests use (data files only) and
transcribes the expected numbers from their golden outputs into
golden.json, each with its source file:line.

  aspgb8   test/nab/aspgb8.nab:8-18, aspgb8.out: 24 igb=8 effective Born
           radii (7 decimals) and per-term energies and frms (9 decimals)
           of the reference's current code, cut = rgbmax = 100.
  aspgb8_check
           test/nab/aspgb8.out.check:12-47, the older golden file of the
           same input (its non-GB terms agree with aspgb8.out).
  gb5trx, gb8trx, gb7trx, gbtrx
           test/nab/gb*trx.nab, gb*trx.out.check: the "ff:" line of every
           mme() call of a 10-iteration conjgrad run on thioredoxin
           (1654 atoms), 2 decimals, frms to 3 significant digits.
  gbrna    test/nab/gbrna.nab, gbrna.out.check:12-22: the same for the
           Dickerson decamer (640 atoms) with gb=1 and salt, read from an
           old-format prmtop (GB radii from the atom names); gbrna.pdb's
           coordinates are also written as gbrna.rst7 (file order, the
           order of the prmtop's atoms).
  asp_nm1  test/nab/asp_nm1.nab, asp_nm1.out.check:1-66: normal modes of
           asp.top (gb=1, cut = rgbmax = 100) after 500 CG and 15
           Newton-Raphson iterations: modes 7-72, frequency (cm^-1) and
           the per-mode E, Cv, S of thermo(), 3 decimals.
  asp_nm2  test/nab/asp_nm2.nab, asp_nm2.out.check:1-15: the same minimum's
           lowest 20 frequencies (modes 6-20 kept by the test script), 5
           decimals, compared at 1e-4 (Run.nmode2: dacdif -a 1.e-4).
  thcp1    test/thcp1.c:20-46, test/2trx.log:21: hierarchical charge
           partitioning (hcp=1, hcp_h1=15, hcp_h2=hcp_h3=999) GB of
           thioredoxin (1653 atoms), gb=5, kappa 0.125: the ff: line of the
           first mme() call (iteration 0, at the input coordinates, where
           the wcons restraint energy is zero).  Its nonpolar column and
           frms include the gbsa=1 surface term; bad, vdW, elect and
           genBorn do not.  2trx.parm7 / 2trx.rst7 are stored gzipped.
"""
import gzip
import json
import os
import re
import shutil

REF = "/root/reference/test/nab"
REF_TEST = "/root/reference/test"
HERE = os.path.dirname(os.path.abspath(__file__))


def ff_lines(path, first=None):
    out = []
    for k, line in enumerate(open(path), 1):
        if first is not None and len(out) == first:
            break
        if line.startswith("ff:"):
            f = line.split()
            out.append(dict(line=k, iter=int(f[1]), total=float(f[2]), bad=float(f[3]),
                            vdw=float(f[4]), elect=float(f[5]), nonpolar=float(f[6]),
                            genborn=float(f[7]), frms=float(f[8])))
    return out


def main():
    for name in ("asp.top", "asp.pdb", "prmtop.gb2", "trxox.2.4ns.pdb"):
        shutil.copyfile(os.path.join(REF, name), os.path.join(HERE, name))
    g = {}
    # aspgb8: radii + per-term energies
    path = os.path.join(REF, "aspgb8.out.check")
    radii, terms = [], {}
    for k, line in enumerate(open(path), 1):
        m = re.match(r"^(\d+)\t\s*([\d.]+)\t\s*([\d.]+)\s*$", line)
        if m:
            radii.append(dict(line=k, atom=int(m.group(1)), rborn=float(m.group(2)),
                              reff=float(m.group(3))))
        m = re.match(r"^\s*(bond|angle|dihedral|enb14|eel14|enb|eel|egb|econs|esurf|Total):\s+(\S+)", line)
        if m:
            terms[m.group(1)] = dict(line=k, value=float(m.group(2)))
    g["aspgb8_check"] = dict(source="test/nab/aspgb8.out.check", prmtop="asp.top", pdb="asp.pdb",
                             options=dict(gb=8, cut=100.0, rgbmax=100.0, kappa=0.0, epsext=78.5,
                                          diel="C"),
                             radii=radii, terms=terms)
    # aspgb8.out: the same input run with the reference's current code (an MPI
    # run, so the radii lines are interleaved out of order); its igb=8 radii
    # and egb differ from the older .check file, whose other terms agree.
    path = os.path.join(REF, "aspgb8.out")
    radii, terms = [], {}
    for k, line in enumerate(open(path), 1):
        m = re.match(r"^(\d+)\t\s*([\d.]+)\t\s*([\d.]+)\s*$", line)
        if m:
            radii.append(dict(line=k, atom=int(m.group(1)), rborn=float(m.group(2)),
                              reff=float(m.group(3))))
        m = re.match(r"^\s*(bond|angle|dihedral|enb14|eel14|enb|eel|egb|econs|esurf|Total|frms):\s+(\S+)",
                     line)
        if m:
            terms[m.group(1)] = dict(line=k, value=float(m.group(2)))
    radii.sort(key=lambda r: r["atom"])
    g["aspgb8"] = dict(source="test/nab/aspgb8.out", prmtop="asp.top", pdb="asp.pdb",
                       options=dict(gb=8, cut=100.0, rgbmax=100.0, kappa=0.0, epsext=78.5,
                                    diel="C"),
                       radii=radii, terms=terms)
    for name, opt in (("gb5trx", dict(gb=5, kappa=0.14701, rgbmax=99.0, cut=99.0)),
                      ("gb8trx", dict(gb=8, kappa=0.0, rgbmax=15.0, cut=99.0)),
                      ("gb7trx", dict(gb=7, kappa=0.0, rgbmax=15.0, cut=99.0)),
                      ("gbtrx", dict(gb=2, kappa=0.14701, rgbmax=99.0, cut=99.0))):
        src = os.path.join(REF, name + ".out.check")
        opt.update(epsext=78.5, diel="C", nsnb=10, cg=dict(maxiter=10, dfpred=10.0,
                                                             rmsgrad_per_dof=0.0001))
        g[name] = dict(source="test/nab/%s.out.check" % name, prmtop="prmtop.gb2",
                       pdb="trxox.2.4ns.pdb", options=opt, ff=ff_lines(src))
    # gbrna: old-format prmtop, gb=1, kappa 0.10395 (gbrna.nab:8)
    shutil.copyfile(os.path.join(REF, "gbrna.prmtop"), os.path.join(HERE, "gbrna.prmtop"))
    xyz = []
    for line in open(os.path.join(REF, "gbrna.pdb")):
        if line.startswith(("ATOM", "HETATM")):
            xyz.extend((float(line[30:38]), float(line[38:46]), float(line[46:54])))
    with open(os.path.join(HERE, "gbrna.rst7"), "w") as fp:
        fp.write("gbrna.pdb coordinates\n%6d\n" % (len(xyz) // 3))
        for k in range(0, len(xyz), 6):
            fp.write("".join("%12.7f" % v for v in xyz[k:k + 6]) + "\n")
    g["gbrna"] = dict(source="test/nab/gbrna.out.check", prmtop="gbrna.prmtop", rst7="gbrna.rst7",
                      options=dict(gb=1, kappa=0.10395, rgbmax=99.0, cut=99.0, epsext=78.5, diel="C",
                                   nsnb=10, cg=dict(maxiter=10, dfpred=10.0, rmsgrad_per_dof=0.0001)),
                      ff=ff_lines(os.path.join(REF, "gbrna.out.check")))
    # asp_nm1 / asp_nm2: normal modes (asp_nm1.nab:14-30, asp_nm2.nab:14-30)
    nm = dict(gb=1, cut=100.0, rgbmax=100.0, kappa=0.0, diel="C", nsnb=999,
              cg=dict(rmsgrad=0.005, dfpred=1.0, maxiter=500), newton=dict(rmsgrad=1e-12, nradd=0.0, maxiter=15))
    rows = []
    for k, line in enumerate(open(os.path.join(REF, "asp_nm1.out.check")), 1):
        f = line.split()
        if len(f) == 5 and f[0].isdigit():
            rows.append(dict(line=k, mode=int(f[0]), freq=float(f[1]), e=float(f[2]), cv=float(f[3]),
                             s=float(f[4])))
    g["asp_nm1"] = dict(source="test/nab/asp_nm1.out.check", prmtop="asp.top", pdb="asp.pdb", options=nm,
                        nmode=dict(eigp=0, ntrun=1), modes=rows)
    rows = []
    for k, line in enumerate(open(os.path.join(REF, "asp_nm2.out.check")), 1):
        f = line.split()
        if len(f) == 2 and f[0].isdigit():
            rows.append(dict(line=k, mode=int(f[0]), freq=float(f[1])))
    g["asp_nm2"] = dict(source="test/nab/asp_nm2.out.check", prmtop="asp.top", pdb="asp.pdb", options=nm,
                        nmode=dict(eigp=20, ntrun=2), tolerance=1e-4, modes=rows)
    # thcp1: hcp GB (test/thcp1.c:34-46), first ff: line of test/2trx.log
    for name in ("2trx.parm7", "2trx.rst7"):
        with open(os.path.join(REF_TEST, name), "rb") as fin, open(os.path.join(HERE, name + ".gz"), "wb") as raw:
            with gzip.GzipFile(filename="", mode="wb", fileobj=raw, mtime=0) as fout:
                shutil.copyfileobj(fin, fout)
    g["thcp1"] = dict(source="test/2trx.log", prmtop="2trx.parm7.gz", rst7="2trx.rst7.gz",
                      options=dict(gb=5, cut=15.0, rgbmax=15.0, kappa=0.125, epsext=78.5, diel="C",
                                   hcp=1, hcp_h1=15.0, hcp_h2=999.0, hcp_h3=999.0, gbsa=1, wcons=5.0),
                      ff=ff_lines(os.path.join(REF_TEST, "2trx.log"), first=1))
    with open(os.path.join(HERE, "golden.json"), "w") as fp:
        json.dump(g, fp, indent=1)


if __name__ == "__main__":
    main()
