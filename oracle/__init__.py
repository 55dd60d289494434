"""The oracle: CPU FP64 restatement of the reference GB path (liboracle.so).

TEST INFRASTRUCTURE ONLY -- the checker for the GPU path.  Only tests/,
__graft_entry__.smoke() and bench.py's cpu_baseline leg may import it; the
product (nabc_amd) never does.  See oracle.h for the reference functions
each entry point restates.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
if not os.path.exists(LIB_PATH):
    raise ImportError("oracle/liboracle.so is not built (make -C oracle)")
lib = C.CDLL(LIB_PATH)

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


_sig("orc_create", vp, [vp, ip, C.c_int, C.c_double, C.c_double, C.c_double, C.c_double,
                        C.c_double, C.c_int])
_sig("orc_destroy", None, [vp])
_sig("orc_gbparams", None, [vp, dp, dp, dp, dp, dp, dp, ip, dp])
_sig("orc_nblist", C.c_longlong, [vp, dp])
_sig("orc_list_counts", None, [vp, ip, ip])
_sig("orc_list_copy", None, [vp, ip])
_sig("orc_egb", C.c_double, [vp, dp, dp, dp, dp, dp, dp])
_sig("orc_mme", C.c_double, [vp, dp, dp, C.c_int, dp, dp])
_sig("orc_set_current", None, [vp, C.c_int])
_sig("orc_mme_cb", C.c_double, [dp, dp, ip])
_sig("orc_log_count", C.c_int, [])
_sig("orc_log_get", None, [C.c_int, ip, dp, dp])
_sig("orc_log_clear", None, [])
_sig("orc_set_threads", None, [C.c_int])


class MDOpts(C.Structure):
    _fields_ = [("dt", C.c_double), ("temp0", C.c_double), ("tempi", C.c_double),
                ("gamma_ln", C.c_double), ("tautp", C.c_double), ("vlimit", C.c_double),
                ("boltz2", C.c_double), ("nsnb", C.c_int), ("zerov", C.c_int),
                ("seed", C.c_ulonglong)]


_sig("orc_gauss", C.c_double, [C.c_ulonglong, C.c_longlong, C.c_longlong])
_sig("orc_md", C.c_int, [vp, C.c_int, dp, dp, dp, C.POINTER(MDOpts), dp])

class HcpOpts(C.Structure):
    _fields_ = [("hcp", C.c_int), ("h1", C.c_double), ("h2", C.c_double), ("h3", C.c_double),
                ("dhcp", C.c_double), ("nres", C.c_int), ("nstrand", C.c_int), ("ncomplex", C.c_int),
                ("ipres", ip), ("ipstrand", ip), ("ipcomplex", ip)]


_sig("orc_set_hcp", C.c_int, [vp, C.POINTER(HcpOpts)])
_sig("orc_hcp_components", C.c_int, [vp, C.c_int, dp, dp, dp, dp])

MMEFUNC = C.CFUNCTYPE(C.c_double, dp, dp, ip)
MME_CB = MMEFUNC(("orc_mme_cb", lib))


def _p(a, t=dp):
    return a.ctypes.data_as(t)


class Oracle:
    """Reference-algorithm evaluation over a PARMSTRUCT_T read by rdparm."""

    def __init__(self, parm, natom, gb, cut, rgbmax, kappa=0.0, epsext=78.5, dielc=1.0,
                 dield=1, frozen=None):
        self.n = natom
        self._frozen = None if frozen is None else np.ascontiguousarray(frozen, dtype=np.int32)
        self.h = lib.orc_create(C.cast(parm, vp), _p(self._frozen, ip) if self._frozen is not None else None,
                                gb, cut, rgbmax, kappa, epsext, dielc, dield)

    def close(self):
        if getattr(self, "h", None):
            lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def nblist(self, x):
        return lib.orc_nblist(self.h, _p(np.ascontiguousarray(x, dtype=np.float64)))

    def lists(self):
        lo = np.zeros(self.n, np.int32)
        up = np.zeros(self.n, np.int32)
        lib.orc_list_counts(self.h, _p(lo, ip), _p(up, ip))
        flat = np.zeros(max(1, int(lo.sum() + up.sum())), np.int32)
        lib.orc_list_copy(self.h, _p(flat, ip))
        return lo, up, flat

    def egb(self, x):
        """egb() on the current list: (epol, evdw, eelec, grad, reff, sumdeijda)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(3 * self.n)
        reff = np.zeros(self.n)
        sd = np.zeros(self.n)
        enb, eel = C.c_double(), C.c_double()
        epol = lib.orc_egb(self.h, _p(x), _p(g), C.byref(enb), C.byref(eel), _p(reff), _p(sd))
        return epol, enb.value, eel.value, g, reff, sd

    def mme(self, x, rebuild=True):
        """mme34(): (ene[13], frms, grad)."""
        x = np.ascontiguousarray(x, dtype=np.float64)
        g = np.zeros(3 * self.n)
        ene = np.zeros(13)
        frms = C.c_double()
        lib.orc_mme(self.h, _p(x), _p(g), 1 if rebuild else 0, _p(ene), C.byref(frms))
        return ene, frms.value, g

    def md(self, x, v, masses, nsteps, dt=0.001, temp0=300.0, tempi=0.0, gamma_ln=0.0,
           tautp=999999.0, vlimit=10.0, nsnb=10, zerov=0, seed=71277):
        """md() (sff.c:2071-2366) for nsteps from (x, v): returns (x, v, rows),
        rows[k] = ene[0..12], frms, ekin, temperature of step k+1."""
        x = np.array(x, dtype=np.float64)
        v = np.zeros(3 * self.n) if v is None else np.array(v, dtype=np.float64)
        m = np.ascontiguousarray(masses, dtype=np.float64)
        rows = np.zeros((max(1, nsteps), 16))
        o = MDOpts(dt, temp0, tempi, gamma_ln, tautp, vlimit, 9.93595e-4, nsnb, zerov, seed)
        rc = lib.orc_md(self.h, nsteps, _p(x), _p(v), _p(m), C.byref(o), _p(rows))
        if rc != 0:
            raise MemoryError("orc_md")
        return x, v, rows[:nsteps]

    def set_hcp(self, hcp, ipres, ipstrand=(1,), ipcomplex=(1,), h1=15.0, h2=50.0, h3=150.0, dhcp=0.25):
        """egb_hcp in place of egb (hcp 1, 2 or 4; 0 switches back).  ipres:
        nres + 1 0-based residue starts ending at natom; ipstrand /
        ipcomplex: 1-based first residues / strands."""
        if hcp == 0:
            return lib.orc_set_hcp(self.h, None)
        self._hcp_arrays = [np.ascontiguousarray(a, dtype=np.int32) for a in (ipres, ipstrand, ipcomplex)]
        r, s, c = self._hcp_arrays
        o = HcpOpts(hcp, h1, h2, h3, dhcp, len(r) - 1, len(s), len(c), _p(r, ip), _p(s, ip), _p(c, ip))
        if lib.orc_set_hcp(self.h, C.byref(o)) != 0:
            raise ValueError("orc_set_hcp: bad hierarchy")
        self._hcp_nq = 1 if hcp == 1 else 2

    def hcp_components(self, level):
        """(centres[n,3], charges[n,nq,4], radii[n,nq], reff[n,nq]) of level 1-3."""
        m = lib.orc_hcp_components(self.h, level, None, None, None, None)
        if m < 0:
            raise ValueError("no HCP state")
        nq = self._hcp_nq
        xc, qc, rc, rf = np.zeros((m, 3)), np.zeros((m, nq, 4)), np.zeros((m, nq)), np.zeros((m, nq))
        lib.orc_hcp_components(self.h, level, _p(xc), _p(qc), _p(rc), _p(rf))
        return xc, qc, rc, rf

    def gbparams(self):
        n = self.n
        fs, a, b, g, q = (np.zeros(n) for _ in range(5))
        neck = np.zeros(n, np.int32)
        off, ns = C.c_double(), C.c_double()
        lib.orc_gbparams(self.h, C.byref(off), C.byref(ns), _p(fs), _p(a), _p(b), _p(g), _p(neck, ip), _p(q))
        return dict(gboffset=off.value, neckscale=ns.value, fs=fs, alpha=a, beta=b, gamma=g,
                    neckidx=neck, charge=q)


def set_threads(n):
    lib.orc_set_threads(int(n))


def log_entries():
    out = []
    for k in range(lib.orc_log_count()):
        it = C.c_int()
        e = np.zeros(13)
        fr = C.c_double()
        lib.orc_log_get(k, C.byref(it), _p(e), C.byref(fr))
        out.append((it.value, e, fr.value))
    return out
