"""nabc_amd -- MI355X generalized-Born energy path (Python bindings).

Thin ctypes bindings over libnabc_amd.so, the C-ABI declared in
include/gbhip.h (GPU engine) and include/sff_gb.h (libsff-compatible host
API: rdparm, mm_options, mme_init_sff, mme, md, conjgrad, getxv, putxv).
The library is built in-tree by ``make -C nabc_amd``; importing this module
raises if it is missing -- there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnabc_amd.so")

if not os.path.exists(LIB_PATH):
    raise ImportError(
        "libnabc_amd.so is not built (run `make -C nabc_amd` or __graft_entry__.build())")

lib = C.CDLL(LIB_PATH, mode=C.RTLD_GLOBAL)

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int)
vp = C.c_void_p

GBHIP_FP64 = 0
GBHIP_MIXED = 1

# mme34 energy-vector layout (eff.c:2617-2874); index 13 is the rms gradient
ENE_NAMES = {0: "total", 1: "enb", 2: "eel", 3: "bond", 4: "angle", 5: "dihedral",
             7: "enb14", 8: "eel14", 9: "econs", 10: "egb", 11: "esurf", 12: "evdwnp",
             13: "frms"}


class Topology(C.Structure):
    _fields_ = [("natom", C.c_int), ("ntypes", C.c_int), ("charge", dp), ("rborn", dp),
                ("fs", dp), ("iac", ip), ("cno", ip), ("cn1", dp), ("cn2", dp),
                ("iblo", ip), ("exclat", ip), ("frozen", ip), ("gbalpha", dp),
                ("gbbeta", dp), ("gbgamma", dp), ("neckidx", ip), ("fsmax", C.c_double)]


class Options(C.Structure):
    _fields_ = [("gb", C.c_int), ("precision", C.c_int), ("cut", C.c_double),
                ("rgbmax", C.c_double), ("gboffset", C.c_double),
                ("gbneckscale", C.c_double), ("kappa", C.c_double), ("epsext", C.c_double)]


class Parm(C.Structure):
    """PARMSTRUCT_T (include/sff_gb.h; the reference's sff.h:40-62 layout)."""
    _fields_ = ([("ititl", C.c_char * 81)] +
                [(f, C.c_int) for f in (
                    "IfBox Nmxrs IfCap Natom Ntypes Nbonh Mbona Ntheth Mtheta Nphih Mphia Nhparm "
                    "Nparm Nnb Nres Nbona Ntheta Nphia Numbnd Numang Nptra Natyp Nphb Nat3 Ntype2d "
                    "Nttyp Nspm Iptres Nspsol Ipatm Natcap Numextra").split()] +
                [(f, C.c_void_p) for f in "AtomNames ResNames AtomType AtomTree".split()] +
                [(f, dp) for f in "Charges Masses Rk Req Tk Teq Pk Pn Phase Solty Cn1 Cn2 HB12 HB10 Rborn Fs".split()] +
                [("Box", C.c_double * 4)] +
                [(f, C.c_double) for f in "Cutcap Xcap Ycap Zcap Fsmax".split()] +
                [(f, ip) for f in (
                    "Iac Iblo Cno Ipres ExclAt TreeJoin AtomRes BondHAt1 BondHAt2 BondHNum BondAt1 "
                    "BondAt2 BondNum AngleHAt1 AngleHAt2 AngleHAt3 AngleHNum AngleAt1 AngleAt2 "
                    "AngleAt3 AngleNum DihHAt1 DihHAt2 DihHAt3 DihHAt4 DihHNum DihAt1 DihAt2 DihAt3 "
                    "DihAt4 DihNum Boundary AtomicNum N14pairs N14pairlist").split()] +
                [(f, dp) for f in "Gvdw Scee Scnb N14sceelist N14scnblist".split()] +
                [("Nstrand", C.c_int), ("Ncomplex", C.c_int), ("Ipstrand", ip), ("Ipcomplex", ip)])


class MDOptions(C.Structure):
    _fields_ = [("dt", C.c_double), ("temp0", C.c_double), ("tempi", C.c_double),
                ("gamma_ln", C.c_double), ("tautp", C.c_double), ("vlimit", C.c_double),
                ("boltz2", C.c_double), ("nsnb", C.c_int), ("zerov", C.c_int),
                ("seed", C.c_ulonglong)]


class Bonded(C.Structure):
    _fields_ = [("nbond", C.c_int), ("bond_i", C.c_void_p), ("bond_j", C.c_void_p),
                ("bond_rk", C.c_void_p), ("bond_req", C.c_void_p), ("nangle", C.c_int),
                ("angle_i", C.c_void_p), ("angle_j", C.c_void_p), ("angle_k", C.c_void_p),
                ("angle_tk", C.c_void_p), ("angle_teq", C.c_void_p), ("ndihed", C.c_int),
                ("dihed_i", C.c_void_p), ("dihed_j", C.c_void_p), ("dihed_k", C.c_void_p),
                ("dihed_l", C.c_void_p), ("dihed_pk", C.c_void_p), ("dihed_pn", C.c_void_p),
                ("dihed_phase", C.c_void_p), ("npair14", C.c_int), ("p14_i", C.c_void_p),
                ("p14_j", C.c_void_p), ("p14_hb", C.c_void_p), ("p14_scee", C.c_void_p),
                ("p14_scnb", C.c_void_p), ("p14_a", C.c_void_p), ("p14_b", C.c_void_p),
                ("charge", C.c_void_p), ("dield", C.c_int)]


class ServeJob(C.Structure):
    _fields_ = [("prm", C.c_void_p), ("x", dp), ("kind", C.c_int), ("nsteps", C.c_int),
                ("seed", C.c_ulonglong), ("arrival_s", C.c_double), ("x_out", dp)]


class ServeConfig(C.Structure):
    _fields_ = [("slots", C.c_int), ("hbm_budget", C.c_size_t), ("precision", C.c_int),
                ("check_fp64", C.c_int), ("cg_rmsgrad", C.c_double), ("cg_dfpred", C.c_double),
                ("progress", C.c_int), ("watchdog_s", C.c_double)]


class ServeResult(C.Structure):
    _fields_ = [("start_s", C.c_double), ("end_s", C.c_double), ("latency_s", C.c_double),
                ("energy", C.c_double), ("energy_fp64", C.c_double),
                ("energy_final_eval", C.c_double), ("footprint", C.c_size_t),
                ("status", C.c_int), ("slot", C.c_int), ("max_active", C.c_int),
                ("setup_s", C.c_double), ("pool_offset", C.c_size_t)]


GBSERVE_CG = 0
GBSERVE_MD = 1

ParmP = C.POINTER(Parm)
MMEFUNC = C.CFUNCTYPE(C.c_double, dp, dp, ip)


def _sig(name, res, args):
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args
    return f


_sig("gbhip_last_error", C.c_char_p, [])
_sig("gbhip_device_count", C.c_int, [])
_sig("gbhip_ctx_create", C.c_int, [C.POINTER(Topology), C.POINTER(Options), dp, C.POINTER(vp)])
_sig("gbhip_ctx_destroy", C.c_int, [vp])
_sig("gbhip_set_coords", C.c_int, [vp, dp])
_sig("gbhip_nblist", C.c_int, [vp, C.POINTER(C.c_longlong)])
_sig("gbhip_nblist_async", C.c_int, [vp])
_sig("gbhip_egb", C.c_int, [vp, dp, dp, dp, dp])
_sig("gbhip_hessian", C.c_int, [vp, C.c_double, dp, dp, dp])
_sig("gbhip_egb_static", C.c_int, [C.POINTER(Topology), C.POINTER(Options), dp, dp, dp, dp, dp])
_sig("gbhip_get_reff", C.c_int, [vp, dp])
_sig("gbhip_get_sumdeijda", C.c_int, [vp, dp])
_sig("gbhip_set_stream", C.c_int, [vp, vp])
_sig("gbhip_set_coords_device", C.c_int, [vp, dp])
_sig("gbhip_egb_device", C.c_int, [vp, dp, dp])
_sig("gbhip_sync", C.c_int, [vp])
_sig("gbhip_npairs", C.c_int, [vp, C.POINTER(C.c_longlong)])
_sig("gbhip_tile_stats", C.c_int, [vp, ip, ip])
_sig("gbhip_get_pairs", C.c_int, [vp, ip, ip, C.c_longlong, C.POINTER(C.c_longlong)])
_sig("gbhip_timing_start", C.c_int, [vp, C.c_int])
_sig("gbhip_timing_stop", C.c_int, [vp, dp, ip])
_sig("gbhip_md_init", C.c_int, [vp, dp, dp, dp, C.POINTER(MDOptions), C.c_int])
_sig("gbhip_md_run", C.c_int, [vp, C.c_int, C.c_int])
_sig("gbhip_md_stats", C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)])
CG_REPORT = C.CFUNCTYPE(None, C.c_int, dp, C.c_void_p)
_sig("gbhip_conjgrad", C.c_int, [vp, dp, C.c_int, dp, C.c_double, C.c_double, C.c_int, C.c_int, CG_REPORT,
                                 C.c_void_p, ip])
_sig("gbhip_slot_create", C.c_int, [C.c_size_t, C.c_size_t, C.c_size_t, vp, C.POINTER(vp)])
_sig("gbhip_slot_set_region", C.c_int, [vp, vp, C.c_size_t])
_sig("gbhip_slot_destroy", C.c_int, [vp])
_sig("gbhip_ctx_create_in_slot", C.c_int, [C.POINTER(Topology), C.POINTER(Options), dp, vp, C.POINTER(vp)])
_sig("gbhip_md_state", C.c_int, [vp, dp, dp, dp])
_sig("gbhip_md_log", C.c_int, [vp, dp, C.c_int, ip])
_sig("gbhip_batch_timing_start", C.c_int, [vp, C.c_int])
_sig("gbhip_batch_timing_stop", C.c_int, [vp, dp, ip])
_sig("gbhip_set_launch", C.c_int, [vp, ip])
_sig("gbhip_valu_peak", C.c_int, [C.c_int, dp])
_sig("gbhip_set_derivative_store", C.c_int, [vp, C.c_int])
_sig("gbhip_batch_set_derivative_store", C.c_int, [vp, C.c_int])
_sig("gbhip_set_deterministic", C.c_int, [vp, C.c_int])


class Hcp(C.Structure):
    """gbhip_hcp (include/gbhip.h)."""
    _fields_ = [("hcp", C.c_int), ("h1", C.c_double), ("h2", C.c_double), ("h3", C.c_double),
                ("dhcp", C.c_double), ("nres", C.c_int), ("ipres", ip), ("nstrand", C.c_int),
                ("ipstrand", ip), ("ncomplex", C.c_int), ("ipcomplex", ip)]


_sig("gbhip_set_hcp", C.c_int, [vp, C.POINTER(Hcp)])
_sig("gbhip_batch_set_deterministic", C.c_int, [vp, C.c_int])
_sig("gbhip_store_stats", C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip])
_sig("gbhip_batch_store_stats", C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip])
_sig("gbhip_get_launch", C.c_int, [vp, ip])
_sig("gbhip_set_split", C.c_int, [vp, ip])
_sig("gbhip_get_split", C.c_int, [vp, ip])
_sig("gbhip_set_queue_fill", C.c_int, [vp, C.c_int, C.c_double])
_sig("gbhip_autotune", C.c_int, [vp, C.c_int, ip])
_sig("gbhip_batch_autotune", C.c_int, [vp, C.c_int, ip])
_sig("gbhip_set_bonded", C.c_int, [vp, vp])
_sig("gbhip_mme", C.c_int, [vp, dp, dp])
_sig("gbhip_mme_device", C.c_int, [vp, dp, dp])
_sig("gbhip_set_mme_graph", C.c_int, [vp, C.c_int])
_sig("gbhip_mme_graph_stats", C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)])
_sig("gbhip_batch_create", C.c_int, [C.c_int, C.POINTER(Topology), C.POINTER(Options),
                                     C.POINTER(dp), C.POINTER(vp)])
_sig("gbhip_batch_destroy", C.c_int, [vp])
_sig("gbhip_batch_set_coords", C.c_int, [vp, C.c_int, dp])
_sig("gbhip_batch_nblist", C.c_int, [vp])
_sig("gbhip_batch_eval", C.c_int, [vp, dp, C.POINTER(dp)])
_sig("gbhip_batch_eval_device", C.c_int, [vp])
_sig("gbhip_batch_stats", C.c_int, [vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), ip])
_sig("gbgen_build", C.c_int, [C.c_int, C.c_ulonglong, C.POINTER(ParmP), dp])
_sig("gbgen_write", C.c_int, [C.c_int, C.c_ulonglong, C.c_char_p, C.c_char_p])
_sig("rdparm", ParmP, [C.c_char_p])
_sig("wrparm", C.c_int, [ParmP, C.c_char_p])
_sig("free_prm", C.c_int, [ParmP])
_sig("getxv", C.c_int, [C.c_char_p, C.c_int, C.c_double, dp, dp])
_sig("putxv", C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_double, dp, dp])
_sig("mm_options", C.c_int, [C.c_char_p])
_sig("mm_reset_options", None, [])
_sig("mm_get_error", C.c_char_p, [])
_sig("mme_init_sff", C.c_int, [ParmP, ip, ip, dp, C.c_char_p])
_sig("mme", C.c_double, [dp, dp, ip])
_sig("mme_get_energies", None, [dp, dp])
_sig("mme_context", vp, [dp])
_sig("gbhip_footprint_bytes", C.c_size_t, [C.c_int, C.c_int])
_sig("gbserve_run", C.c_int, [C.POINTER(ServeJob), C.c_int, C.POINTER(ServeConfig), C.c_char_p,
                              C.POINTER(ServeResult)])
SERVE_RUNNER = C.CFUNCTYPE(C.c_int, C.POINTER(ServeJob), C.c_int, C.c_void_p, C.POINTER(C.c_double))
_sig("gbserve_run_custom", C.c_int, [C.POINTER(ServeJob), C.c_int, C.POINTER(ServeConfig), SERVE_RUNNER,
                                     C.c_void_p, C.POINTER(ServeResult)])


def serve(jobs, slots=8, hbm_budget=64 << 30, precision=GBHIP_MIXED, check_fp64=False,
          mm_opts="gb=5, cut=20.0, rgbmax=20.0", cg_rmsgrad=1e-4, cg_dfpred=10.0, progress=0,
          watchdog_s=0.0, want_x=False):
    """Run jobs through the serving loop.

    jobs: list of dicts with prm (PARMSTRUCT_T pointer), x, kind
    (GBSERVE_CG / GBSERVE_MD), nsteps, seed, arrival_s.
    Returns a list of result dicts (times in seconds from the run start);
    with want_x each also holds the job's final coordinates as "x"."""
    n = len(jobs)
    arr = (ServeJob * n)()
    keep, outs = [], []
    for k, j in enumerate(jobs):
        x = np.ascontiguousarray(j["x"], dtype=np.float64)
        xo = np.zeros_like(x) if want_x else None
        keep.append(x)
        outs.append(xo)
        arr[k] = ServeJob(C.cast(j["prm"], C.c_void_p), _p(x), j["kind"], j["nsteps"],
                          j.get("seed", k + 1), j["arrival_s"], _p(xo) if want_x else None)
    cfg = ServeConfig(slots, hbm_budget, precision, 1 if check_fp64 else 0, cg_rmsgrad, cg_dfpred, progress,
                      watchdog_s)
    res = (ServeResult * n)()
    rc = lib.gbserve_run(arr, n, C.byref(cfg), mm_opts.encode() if mm_opts else None, res)
    if rc != 0:
        raise GBHipError(lib.mm_get_error().decode(errors="replace"))
    out = [{f[0]: getattr(r, f[0]) for f in ServeResult._fields_} for r in res]
    if want_x:
        for r, xo in zip(out, outs):
            r["x"] = xo
    return out


def serve_custom(jobs, run, slots=8, hbm_budget=64 << 30):
    """The serving loop's admission and completion logic with run(k, slot)
    doing job k's work on the slot's worker thread (returns (status,
    energy)); no device work.  jobs as for serve(); results likewise."""
    n = len(jobs)
    arr = (ServeJob * n)()
    keep = []
    for k, j in enumerate(jobs):
        x = np.ascontiguousarray(j.get("x", np.zeros(3)), dtype=np.float64)
        keep.append(x)
        arr[k] = ServeJob(C.cast(j["prm"], C.c_void_p), _p(x), j["kind"], j["nsteps"],
                          j.get("seed", k + 1), j["arrival_s"], None)
    base = C.addressof(arr)
    size = C.sizeof(ServeJob)

    def tramp(job, slot, user, energy):
        k = (C.addressof(job.contents) - base) // size
        try:
            st, e = run(k, slot)
        except Exception:  # a Python error becomes a failed job, not a crash
            return -99
        energy[0] = e
        return st

    cb = SERVE_RUNNER(tramp)
    cfg = ServeConfig(slots, hbm_budget, GBHIP_MIXED, 0, 0.0, 0.0, 0, -1.0)
    res = (ServeResult * n)()
    rc = lib.gbserve_run_custom(arr, n, C.byref(cfg), cb, None, res)
    if rc != 0:
        raise GBHipError(lib.mm_get_error().decode(errors="replace"))
    return [{f[0]: getattr(r, f[0]) for f in ServeResult._fields_} for r in res]
_sig("md", C.c_int, [C.c_int, C.c_int, dp, dp, dp, MMEFUNC])
_sig("conjgrad", C.c_int, [dp, ip, dp, MMEFUNC, dp, dp, ip])
_sig("nabc_md_gauss", C.c_double, [C.c_ulonglong, C.c_longlong, C.c_longlong])

MME_C = MMEFUNC(("mme", lib))
MME2FUNC = C.CFUNCTYPE(C.c_double, dp, dp, dp, dp, dp, ip, ip, ip, ip, ip, ip, ip, ip, ip, C.c_char_p)
_sig("mme2", C.c_double, [dp, dp, dp, dp, dp, ip, ip, ip, ip, ip, ip, ip, ip, ip, C.c_char_p])
_sig("newton", C.c_int, [dp, ip, dp, MMEFUNC, MME2FUNC, dp, dp, ip])
_sig("nmode", C.c_int, [dp, C.c_int, MME2FUNC, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int])
MME2_C = MME2FUNC(("mme2", lib))


class GBHipError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise GBHipError(lib.gbhip_last_error().decode(errors="replace") + f" (code {rc})")


def _p(a, t=dp):
    return a.ctypes.data_as(t) if a is not None else None


def device_count():
    return lib.gbhip_device_count()


def valu_peak(precision):
    """Measured VALU TFLOP/s of independent FMA chains (GBHIP_FP64 or GBHIP_MIXED = FP32)."""
    t = C.c_double()
    _check(lib.gbhip_valu_peak(int(precision), C.byref(t)))
    return t.value


# ---------------------------------------------------------------- prmtop / rst7

def read_parm(path):
    p = lib.rdparm(path.encode())
    if not p:
        raise IOError(lib.mm_get_error().decode(errors="replace"))
    return p


def parm_arrays(p):
    """Copy the GB-relevant topology arrays of a PARMSTRUCT_T into numpy."""
    s = p.contents
    n, nt = s.Natom, s.Ntypes
    nttyp = nt * (nt + 1) // 2
    g = lambda ptr, k, dt: np.ctypeslib.as_array(ptr, shape=(k,)).astype(dt).copy()
    return dict(natom=n, ntypes=nt, charge=g(s.Charges, n, np.float64),
                rborn=g(s.Rborn, n, np.float64), fs=g(s.Fs, n, np.float64),
                mass=g(s.Masses, n, np.float64), iac=g(s.Iac, n, np.int32),
                cno=g(s.Cno, nt * nt, np.int32), cn1=g(s.Cn1, nttyp, np.float64),
                cn2=g(s.Cn2, nttyp, np.float64), iblo=g(s.Iblo, n, np.int32),
                exclat=g(s.ExclAt, max(1, s.Nnb), np.int32), fsmax=s.Fsmax,
                atomic_num=g(s.AtomicNum, n, np.int32))


def build_system(natom, seed):
    """Generate a synthetic system in memory: (PARMSTRUCT_T pointer, arrays, x)."""
    p = ParmP()
    x = np.zeros(3 * natom)
    rc = lib.gbgen_build(natom, seed, C.byref(p), _p(x))
    if rc != 0:
        raise RuntimeError(lib.mm_get_error().decode(errors="replace"))
    return p, parm_arrays(p), x


def read_rst7(path, natom):
    x = np.zeros(3 * natom)
    v = np.zeros(3 * natom)
    rc = lib.getxv(path.encode(), natom, 0.0, _p(x), _p(v))
    if rc != 0:
        raise IOError(lib.mm_get_error().decode(errors="replace"))
    return x, v


def generate(natom, seed, prmtop_path, rst7_path):
    rc = lib.gbgen_write(natom, seed, prmtop_path.encode(), rst7_path.encode())
    if rc != 0:
        raise RuntimeError(lib.mm_get_error().decode(errors="replace"))


# ------------------------------------------------------- GB parameter set-up

GB8 = {1: (0.788440, 0.798699, 0.437334, 1.425952), 6: (0.733756, 0.506378, 0.205844, 1.058554),
       7: (0.503364, 0.316828, 0.192915, 0.733599), 8: (0.867814, 0.876635, 0.387882, 1.061039),
       15: (1.0, 0.8, 4.85, 0.5), 16: (0.867814, 0.876635, 0.387882, -0.703469)}
GB7_FS = {1: 1.09085413633, 6: 0.484353823306, 7: 0.700147318409, 8: 1.06557401132,
          16: 0.602256336067}


def gb_params(gb, rborn, fs, atomic_num):
    """Per-atom GB parameters as mme_init_sff sets them (sff.c:1187-1333).

    Returns (fs, alpha, beta, gamma, neckidx, gboffset, neckscale)."""
    n = len(rborn)
    fs = np.array(fs, dtype=np.float64)
    a = np.zeros(n); b = np.zeros(n); g = np.zeros(n)
    gboffset, neckscale = 0.09, 0.0
    if gb == 2:
        a[:], b[:], g[:] = 0.8, 0.0, 2.909125
    elif gb == 5:
        a[:], b[:], g[:] = 1.0, 0.8, 4.85
    elif gb == 7:
        neckscale = 0.361825
        a[:], b[:], g[:] = 1.09511284, 1.90792938, 2.50798245
        for i, z in enumerate(atomic_num):
            fs[i] = GB7_FS.get(int(z), 0.5)
    elif gb == 8:
        gboffset, neckscale = 0.195141, 0.826836
        for i, z in enumerate(atomic_num):
            a[i], b[i], g[i], fs[i] = GB8.get(int(z), (1.0, 0.8, 4.85, 0.5))
    neck = ((np.asarray(rborn) - 1.0) * 20.0 + 0.5).astype(np.int32)
    return fs, a, b, g, neck, gboffset, neckscale


def kappa_from_salt(conc_molar):
    """Debye-Hueckel kappa (1/A) for a 1:1 salt at 300 K: sqrt(0.10806 c).

    Reproduces the values the reference's inputs use for 0.1, 0.15 and 0.2 M
    (txmin.c:45, benchmarks/p450.nab:8, benchmarks/mb.nab:10)."""
    return float(np.sqrt(0.10806 * conc_molar))


# --------------------------------------------------------------- GB context

class _Keep:
    """Holds numpy arrays alive while C holds pointers into them."""

    def __init__(self):
        self.refs = []

    def arr(self, a, dt):
        a = np.ascontiguousarray(a, dtype=dt)
        self.refs.append(a)
        return a


def make_topology(keep, top, gb, frozen=None, atomic_num=None):
    """Build a gbhip_topology from a dict of arrays (see parm_arrays)."""
    fs, a, b, g, neck, gboffset, neckscale = gb_params(
        gb, top["rborn"], top["fs"], atomic_num if atomic_num is not None else top.get("atomic_num", np.zeros(top["natom"])))
    t = Topology()
    t.natom = top["natom"]
    t.ntypes = top["ntypes"]
    t.charge = _p(keep.arr(top["charge"], np.float64))
    t.rborn = _p(keep.arr(top["rborn"], np.float64))
    t.fs = _p(keep.arr(fs, np.float64))
    t.iac = _p(keep.arr(top["iac"], np.int32), ip)
    t.cno = _p(keep.arr(top["cno"], np.int32), ip)
    t.cn1 = _p(keep.arr(top["cn1"], np.float64))
    t.cn2 = _p(keep.arr(top["cn2"], np.float64))
    t.iblo = _p(keep.arr(top["iblo"], np.int32), ip)
    t.exclat = _p(keep.arr(top["exclat"], np.int32), ip)
    t.frozen = _p(keep.arr(frozen, np.int32), ip) if frozen is not None else None
    t.gbalpha = _p(keep.arr(a, np.float64))
    t.gbbeta = _p(keep.arr(b, np.float64))
    t.gbgamma = _p(keep.arr(g, np.float64))
    t.neckidx = _p(keep.arr(neck, np.int32), ip)
    t.fsmax = top["fsmax"]
    return t, gboffset, neckscale


def make_options(gb, cut, rgbmax, kappa, epsext, gboffset, neckscale, precision=GBHIP_FP64):
    o = Options()
    o.gb, o.precision, o.cut, o.rgbmax = gb, precision, cut, rgbmax
    o.gboffset, o.gbneckscale = gboffset, neckscale
    o.kappa, o.epsext = kappa, epsext
    return o


def egb_static(top, x, gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, epsext=78.5, precision=GBHIP_FP64,
               frozen=None, atomic_num=None):
    """One evaluation through gbhip_egb_static (no device allocation).
    Returns (epol, evdw, eelec, grad[3N])."""
    keep = _Keep()
    t, gboffset, neckscale = make_topology(keep, top, gb, frozen, atomic_num)
    o = make_options(gb, cut, rgbmax, kappa, epsext, gboffset, neckscale, precision)
    x = np.ascontiguousarray(x, dtype=np.float64)
    g = np.zeros(3 * top["natom"])
    e = [C.c_double() for _ in range(3)]
    _check(lib.gbhip_egb_static(C.byref(t), C.byref(o), _p(x), _p(g), *[C.byref(v) for v in e]))
    return e[0].value, e[1].value, e[2].value, g


class GBContext:
    """One molecule on the GPU: pair list + egb() evaluation (gbhip.h)."""

    def __init__(self, top, x, gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, epsext=78.5,
                 precision=GBHIP_FP64, frozen=None, atomic_num=None):
        self._keep = _Keep()
        t, gboffset, neckscale = make_topology(self._keep, top, gb, frozen, atomic_num)
        o = make_options(gb, cut, rgbmax, kappa, epsext, gboffset, neckscale, precision)
        self.natom = top["natom"]
        x = np.ascontiguousarray(x, dtype=np.float64)
        h = vp()
        _check(lib.gbhip_ctx_create(C.byref(t), C.byref(o), _p(x), C.byref(h)))
        self.h = h
        self.set_coords(x)

    def close(self):
        if getattr(self, "h", None):
            lib.gbhip_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_coords(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        _check(lib.gbhip_set_coords(self.h, _p(x)))
        _check(lib.gbhip_sync(self.h))

    def nblist(self):
        n = C.c_longlong()
        _check(lib.gbhip_nblist(self.h, C.byref(n)))
        return n.value

    def egb(self):
        """Returns (epol, evdw, eelec, grad[3N])."""
        g = np.zeros(3 * self.natom)
        e = [C.c_double() for _ in range(3)]
        _check(lib.gbhip_egb(self.h, _p(g), *[C.byref(v) for v in e]))
        return e[0].value, e[1].value, e[2].value, g

    def reff(self):
        r = np.zeros(self.natom)
        _check(lib.gbhip_get_reff(self.h, _p(r)))
        return r

    def sumdeijda(self):
        r = np.zeros(self.natom)
        _check(lib.gbhip_get_sumdeijda(self.h, _p(r)))
        return r

    def tile_stats(self):
        a, b = C.c_int(), C.c_int()
        _check(lib.gbhip_tile_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def npairs(self):
        n = C.c_longlong()
        _check(lib.gbhip_npairs(self.h, C.byref(n)))
        return n.value

    def nblist_async(self):
        _check(lib.gbhip_nblist_async(self.h))

    def egb_device(self):
        """Asynchronous evaluation, results kept on the device."""
        _check(lib.gbhip_egb_device(self.h, None, None))

    def sync(self):
        _check(lib.gbhip_sync(self.h))

    def timing_start(self, max_evals):
        _check(lib.gbhip_timing_start(self.h, max_evals))

    def timing_stop(self):
        """Summed ms of (pass 1, radii, pass 2, pass 3) and the eval count."""
        ms = np.zeros(4)
        n = C.c_int()
        _check(lib.gbhip_timing_stop(self.h, _p(ms), C.byref(n)))
        return ms, n.value

    def set_launch(self, born=0, pairs=0, chain=0):
        """Workgroups of the three pair kernels (0 keeps a value)."""
        b = np.array([born, pairs, chain], np.int32)
        _check(lib.gbhip_set_launch(self.h, _p(b, ip)))

    def launch(self):
        b = np.zeros(3, np.int32)
        _check(lib.gbhip_get_launch(self.h, _p(b, ip)))
        return tuple(int(v) for v in b)

    def set_derivative_store(self, mode):
        """1: pass 1 stores the pass-3 derivatives when they fit; 0: never."""
        _check(lib.gbhip_set_derivative_store(self.h, mode))

    def set_deterministic(self, on):
        """Fixed-order reductions (bit-identical results run to run)."""
        _check(lib.gbhip_set_deterministic(self.h, int(on)))

    def set_hcp(self, hcp, ipres, ipstrand=(1,), ipcomplex=(1,), h1=15.0, h2=50.0, h3=150.0, dhcp=0.25):
        """Hierarchical charge partitioning (egb_hcp) for every later
        evaluation: hcp 1, 2 or 4 (0 returns to the pair-list path).  ipres:
        nres + 1 0-based residue starts ending at natom; ipstrand /
        ipcomplex: 1-based first residues / first strands."""
        if hcp == 0:
            _check(lib.gbhip_set_hcp(self.h, None))
            return
        r, s, c = (np.ascontiguousarray(a, dtype=np.int32) for a in (ipres, ipstrand, ipcomplex))
        o = Hcp(hcp, h1, h2, h3, dhcp, len(r) - 1, _p(r, ip), len(s), _p(s, ip), len(c), _p(c, ip))
        _check(lib.gbhip_set_hcp(self.h, C.byref(o)))

    def set_split(self, born=0, pairs=0, chain=0):
        """Step chunks per tile pair for the three pair passes (1, 2, 4, 8, 16;
        0 = the default for the system's size); resets the launch shapes to
        their defaults."""
        b = np.array([born, pairs, chain], np.int32)
        _check(lib.gbhip_set_split(self.h, _p(b, ip)))

    def split(self):
        b = np.zeros(3, np.int32)
        _check(lib.gbhip_get_split(self.h, _p(b, ip)))
        return tuple(int(v) for v in b)

    def set_queue_fill(self, fill, pass_=0):
        """Pass-1 visits filled below `fill` run from a queue of their listed
        pairs (0: never, above 1: always; mixed precision)."""
        _check(lib.gbhip_set_queue_fill(self.h, pass_, fill))

    def store_stats(self):
        """(entries needed by the current list, entries allocated, in use)."""
        need, cap, ok = C.c_longlong(), C.c_longlong(), C.c_int()
        _check(lib.gbhip_store_stats(self.h, C.byref(need), C.byref(cap), C.byref(ok)))
        return need.value, cap.value, bool(ok.value)

    def autotune(self, reps=5):
        """Pick the fastest grid size per pass (and derivative store on/off); returns the grid sizes."""
        b = np.zeros(3, np.int32)
        _check(lib.gbhip_autotune(self.h, reps, _p(b, ip)))
        return tuple(int(v) for v in b)

    def md_init(self, masses, x, v0=None, dt=0.001, temp0=300.0, tempi=0.0, gamma_ln=0.0,
                tautp=999999.0, vlimit=10.0, nsnb=10, zerov=0, seed=71277, log_cap=1024):
        o = MDOptions(dt=dt, temp0=temp0, tempi=tempi, gamma_ln=gamma_ln, tautp=tautp,
                      vlimit=vlimit, boltz2=9.93595e-4, nsnb=nsnb, zerov=zerov, seed=seed)
        m = np.ascontiguousarray(masses, dtype=np.float64)
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = None if v0 is None else np.ascontiguousarray(v0, dtype=np.float64)
        _check(lib.gbhip_md_init(self.h, _p(m), _p(x), _p(v) if v is not None else None,
                                 C.byref(o), log_cap))

    def pairs(self):
        """Current list as a sorted (P, 2) array of (min, max) atom indices."""
        cap = self.npairs()
        pi = np.zeros(max(cap, 1), np.int32)
        pj = np.zeros(max(cap, 1), np.int32)
        n = C.c_longlong()
        _check(lib.gbhip_get_pairs(self.h, _p(pi, ip), _p(pj, ip), cap, C.byref(n)))
        assert n.value == cap
        key = np.minimum(pi[:cap], pj[:cap]).astype(np.int64) * self.natom + np.maximum(pi[:cap], pj[:cap])
        key.sort()
        return np.stack([key // self.natom, key % self.natom], axis=1)


class Batch:
    """Many molecules scored in one launch per pass (gbhip_batch_*)."""

    def __init__(self, systems, gb=5, cut=999.0, rgbmax=999.0, kappa=0.0, epsext=78.5,
                 precision=GBHIP_FP64):
        """systems: list of (arrays, x) with arrays as returned by parm_arrays."""
        self._keep = _Keep()
        n = len(systems)
        tops = (Topology * n)()
        xs = (dp * n)()
        self.natoms = []
        gboffset = neckscale = None
        for k, (arr, x) in enumerate(systems):
            t, gboffset, neckscale = make_topology(self._keep, arr, gb)
            tops[k] = t
            xs[k] = _p(self._keep.arr(x, np.float64))
            self.natoms.append(arr["natom"])
        o = make_options(gb, cut, rgbmax, kappa, epsext, gboffset, neckscale, precision)
        h = vp()
        _check(lib.gbhip_batch_create(n, tops, C.byref(o), xs, C.byref(h)))
        self.h = h
        self.n = n

    def close(self):
        if getattr(self, "h", None):
            lib.gbhip_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def eval(self, grads=True):
        """Returns (ene[n,3] = epol, evdw, eelec per molecule, list of gradients)."""
        ene = np.zeros(3 * self.n)
        gl = [np.zeros(3 * m) for m in self.natoms] if grads else None
        gp = (dp * self.n)(*[_p(g) for g in gl]) if grads else None
        _check(lib.gbhip_batch_eval(self.h, _p(ene), gp))
        return ene.reshape(self.n, 3), gl

    def eval_device(self):
        _check(lib.gbhip_batch_eval_device(self.h))

    def nblist(self):
        _check(lib.gbhip_batch_nblist(self.h))

    def stats(self):
        p, a, t = C.c_longlong(), C.c_longlong(), C.c_int()
        _check(lib.gbhip_batch_stats(self.h, C.byref(p), C.byref(a), C.byref(t)))
        return p.value, a.value, t.value

    def timing_start(self, max_evals):
        _check(lib.gbhip_batch_timing_start(self.h, max_evals))

    def timing_stop(self):
        ms = np.zeros(4)
        n = C.c_int()
        _check(lib.gbhip_batch_timing_stop(self.h, _p(ms), C.byref(n)))
        return ms, n.value

    def set_derivative_store(self, mode):
        _check(lib.gbhip_batch_set_derivative_store(self.h, mode))

    def set_deterministic(self, on):
        _check(lib.gbhip_batch_set_deterministic(self.h, int(on)))

    def store_stats(self):
        need, cap, ok = C.c_longlong(), C.c_longlong(), C.c_int()
        _check(lib.gbhip_batch_store_stats(self.h, C.byref(need), C.byref(cap), C.byref(ok)))
        return need.value, cap.value, bool(ok.value)

    def autotune(self, reps=3):
        b = np.zeros(3, np.int32)
        _check(lib.gbhip_batch_autotune(self.h, reps, _p(b, ip)))
        return tuple(int(v) for v in b)
