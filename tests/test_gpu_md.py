"""Device-resident dynamics (md() on the GPU, hipGraph-captured steps).

Without noise (gamma_ln = 0) the device loop must follow the host loop of
md() (sff.c:2071-2366, one mme() per step) to within summation-order noise;
graph replay must match stream launches; with Langevin noise the
temperature must settle at temp0 (MD parity with the reference's serial
rand2.c stream is statistical, not bitwise).
"""
import numpy as np
import pytest

from common import generated_system, rel_rms

pytestmark = pytest.mark.gpu


def _init(na, parm, opts):
    na.lib.mm_reset_options()
    assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0, na.lib.mm_get_error()


def _md(na, n, steps, x, v, func=None):
    f = np.zeros(n)
    fn = func if func is not None else na.MME_C
    rc = na.lib.md(n, steps, x.ctypes.data_as(na.dp), f.ctypes.data_as(na.dp), v.ctypes.data_as(na.dp), fn)
    assert rc == 0, na.lib.mm_get_error()
    return f


def test_device_md_follows_host_loop(gpu, gen_dir):
    na = gpu
    parm, arr, x0 = generated_system(600, 21, gen_dir)
    n = 3 * arr["natom"]
    rng = np.random.default_rng(1)
    v0 = rng.normal(0, 0.05, n)
    opts = "gb=5, cut=12.0, rgbmax=12.0, nsnb=5, gamma_ln=0.0, tempi=0.0, dt=0.001, ntpr_md=1000, ntpr=100000"
    # device path: md() with this library's mme
    _init(na, parm, opts)
    xa, va = x0.copy(), v0.copy()
    _md(na, n, 25, xa, va)
    # host loop: the same md() driven through a Python wrapper of mme()
    _init(na, parm, opts)
    xb, vb = x0.copy(), v0.copy()
    wrap = na.MMEFUNC(lambda xp, fp, itp: na.lib.mme(xp, fp, itp))
    _md(na, n, 25, xb, vb, func=wrap)
    assert np.max(np.abs(xa - xb)) < 1e-8
    assert rel_rms(va, vb) < 1e-8
    assert np.max(np.abs(xa - x0)) > 1e-3  # it moved


def test_graph_replay_matches_stream(gpu, gen_dir):
    na = gpu
    parm, arr, x0 = generated_system(2500, 22, gen_dir)
    _init(na, parm, "gb=8, cut=14.0, rgbmax=12.0, kappa=0.127315, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                     boltz2=9.93595e-4, nsnb=10, zerov=0, seed=7)
    masses = na.parm_arrays(parm)["mass"]
    out = []
    for use_graph in (1, 0):
        assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                    na.C.byref(o), 64) == 0, na.lib.gbhip_last_error()
        assert na.lib.gbhip_md_run(ctx, 40, use_graph) == 0, na.lib.gbhip_last_error()
        x = np.zeros_like(x0)
        v = np.zeros_like(x0)
        assert na.lib.gbhip_md_state(ctx, x.ctypes.data_as(na.dp), v.ctypes.data_as(na.dp), None) == 0
        out.append((x, v))
    assert np.max(np.abs(out[0][0] - out[1][0])) < 1e-8
    assert rel_rms(out[0][1], out[1][1]) < 1e-7
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_langevin_temperature(gpu, gen_dir):
    na = gpu
    parm, arr, x0 = generated_system(600, 23, gen_dir)
    _init(na, parm, "gb=5, cut=15.0, rgbmax=15.0, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                     boltz2=9.93595e-4, nsnb=20, zerov=0, seed=11)
    masses = na.parm_arrays(parm)["mass"]
    assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                na.C.byref(o), 4000) == 0
    assert na.lib.gbhip_md_run(ctx, 4000, 1) == 0, na.lib.gbhip_last_error()
    rows = np.zeros((4000, 16))
    nr = na.C.c_int()
    assert na.lib.gbhip_md_log(ctx, rows.ctypes.data_as(na.dp), 4000, na.C.byref(nr)) == 0
    assert nr.value == 4000
    assert np.all(np.isfinite(rows))
    temp = rows[2000:, 15].mean()
    assert 270.0 < temp < 330.0, temp
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_md_store_holds_any_list(gpu, gen_dir):
    """The step graphs freeze the pass-3 derivative store, so gbhip_md_init
    sizes it for the largest list the molecule can have (every tile pair
    full): the store stays in use after the run, whatever the list grew to."""
    na = gpu
    parm, arr, x0 = generated_system(2500, 23, gen_dir)
    _init(na, parm, "gb=5, cut=12.0, rgbmax=12.0, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                     boltz2=9.93595e-4, nsnb=10, zerov=0, seed=3)
    masses = na.parm_arrays(parm)["mass"]
    assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                na.C.byref(o), 50) == 0, na.lib.gbhip_last_error()
    nt, ntp = na.C.c_int(), na.C.c_int()
    assert na.lib.gbhip_tile_stats(ctx, na.C.byref(nt), na.C.byref(ntp)) == 0
    need, cap, used = na.C.c_longlong(), na.C.c_longlong(), na.C.c_int()
    assert na.lib.gbhip_store_stats(ctx, na.C.byref(need), na.C.byref(cap), na.C.byref(used)) == 0
    full = nt.value * (nt.value + 1) // 2 * 64 * 64
    assert used.value == 1 and cap.value >= full > need.value
    assert na.lib.gbhip_md_run(ctx, 50, 1) == 0, na.lib.gbhip_last_error()
    assert na.lib.gbhip_store_stats(ctx, na.C.byref(need), na.C.byref(cap), na.C.byref(used)) == 0
    assert used.value == 1
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))


def test_set_stream_after_dynamics_keeps_callers_stream(gpu, gen_dir):
    """gbhip_md_init gives a context without a stream one of its own; a later
    gbhip_set_stream hands the context the caller's stream (dynamics then
    replay there) and the engine releases only the stream it created: the
    caller's stream is still usable after the context is freed."""
    import torch
    na = gpu
    parm, arr, x0 = generated_system(600, 24, gen_dir)
    _init(na, parm, "gb=5, cut=12.0, rgbmax=12.0, ntpr=100000")
    ctx = na.lib.mme_context(x0.ctypes.data_as(na.dp))
    assert ctx
    o = na.MDOptions(dt=0.001, temp0=300.0, tempi=300.0, gamma_ln=5.0, tautp=999999.0, vlimit=10.0,
                     boltz2=9.93595e-4, nsnb=10, zerov=0, seed=5)
    masses = na.parm_arrays(parm)["mass"]
    assert na.lib.gbhip_md_init(ctx, masses.ctypes.data_as(na.dp), x0.ctypes.data_as(na.dp), None,
                                na.C.byref(o), 40) == 0, na.lib.gbhip_last_error()
    assert na.lib.gbhip_md_run(ctx, 10, 1) == 0, na.lib.gbhip_last_error()
    s = torch.cuda.Stream(device=0)
    assert na.lib.gbhip_set_stream(ctx, na.C.c_void_p(s.cuda_stream)) == 0, na.lib.gbhip_last_error()
    assert na.lib.gbhip_md_run(ctx, 10, 1) == 0, na.lib.gbhip_last_error()
    x = np.zeros_like(x0)
    assert na.lib.gbhip_md_state(ctx, x.ctypes.data_as(na.dp), None, None) == 0
    assert np.all(np.isfinite(x)) and np.max(np.abs(x - x0)) > 1e-4
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
    with torch.cuda.stream(s):
        t = torch.ones(1024, device="cuda:0") * 2.0
    s.synchronize()
    assert float(t.sum()) == 2048.0


def test_device_md_frozen_langevin_follows_host_loop(gpu, gen_dir):
    """Frozen atoms under Langevin dynamics: the device step (drift in the
    evaluation's first kernel, kick in its gradient kernel) skips their kick
    and zeroes their gradient as the host loop of md() does, with the same
    counter-based noise, so the two trajectories agree to summation-order
    noise and the frozen atoms stay where they were."""
    na = gpu
    parm, arr, x0 = generated_system(600, 27, gen_dir)
    natom = arr["natom"]
    n = 3 * natom
    frozen = np.zeros(natom, dtype=np.int32)
    frozen[::7] = 1
    fz3 = np.repeat(frozen, 3).astype(bool)
    rng = np.random.default_rng(4)
    v0 = rng.normal(0, 0.05, n)
    v0[fz3] = 0.0
    opts = ("gb=5, cut=12.0, rgbmax=12.0, nsnb=5, gamma_ln=5.0, temp0=300.0, tempi=0.0, dt=0.001, "
            "md_seed=9, ntpr_md=1000, ntpr=100000")
    out = []
    for host in (False, True):
        na.lib.mm_reset_options()
        assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
        assert na.lib.mme_init_sff(parm, frozen.ctypes.data_as(na.ip), None, None, None) == 0, na.lib.mm_get_error()
        x, v = x0.copy(), v0.copy()
        wrap = na.MMEFUNC(lambda xp, fp, itp: na.lib.mme(xp, fp, itp)) if host else None
        _md(na, n, 25, x, v, func=wrap)
        out.append((x, v))
    (xa, va), (xb, vb) = out
    assert np.max(np.abs(xa - xb)) < 1e-8
    assert rel_rms(va, vb) < 1e-8
    assert np.array_equal(xa[fz3], x0[fz3]) and np.all(va[fz3] == 0.0)
    assert np.max(np.abs(xa[~fz3] - x0[~fz3])) > 1e-3  # the others moved


def test_md_writes_netcdf_trajectory(gpu, gen_dir, tmp_path):
    """mme_init_sff with a file name and ntwx > 0: md() appends a frame of
    single-precision coordinates every ntwx steps (sff.c:1164-1168,
    :2341-2342), from the device loop (which runs in chunks ending on frame
    steps) and from the host loop alike; the last frame is the final
    coordinates."""
    from scipy.io import netcdf_file
    na = gpu
    parm, arr, x0 = generated_system(600, 28, gen_dir)
    n = 3 * arr["natom"]
    rng = np.random.default_rng(2)
    v0 = rng.normal(0, 0.05, n)
    opts = "gb=5, cut=12.0, rgbmax=12.0, nsnb=5, gamma_ln=0.0, tempi=0.0, dt=0.001, ntwx=5, ntpr_md=1000, ntpr=100000"
    frames, finals = [], []
    for host in (False, True):
        path = str(tmp_path / ("host.nc" if host else "device.nc"))
        na.lib.mm_reset_options()
        assert na.lib.mm_options(opts.encode()) == 0, na.lib.mm_get_error()
        assert na.lib.mme_init_sff(parm, None, None, None, path.encode()) == 0, na.lib.mm_get_error()
        x, v = x0.copy(), v0.copy()
        wrap = na.MMEFUNC(lambda xp, fp, itp: na.lib.mme(xp, fp, itp)) if host else None
        _md(na, n, 20, x, v, func=wrap)
        f = netcdf_file(path, "r", mmap=False)
        c = f.variables["coordinates"][:].copy()
        tm = f.variables["time"][:].copy()
        f.close()
        assert c.shape == (4, arr["natom"], 3)  # steps 5, 10, 15, 20
        assert np.allclose(np.diff(tm), 0.005, atol=1e-6)
        assert np.array_equal(c[-1], x.astype(np.float32).reshape(-1, 3))
        frames.append(c)
        finals.append(x)
    # the two loops agree to summation-order noise, so their float frames
    # agree to a rounding step of single precision
    assert np.max(np.abs(frames[0] - frames[1])) < 1e-4
    assert np.max(np.abs(frames[0][0] - x0.reshape(-1, 3))) > 1e-4  # it moved
    it = na.C.c_int(-3)
    na.lib.mme(x0.ctypes.data_as(na.dp), x0.copy().ctypes.data_as(na.dp), na.C.byref(it))
    # a fresh initialisation closes the trajectory file
    assert na.lib.mme_init_sff(parm, None, None, None, None) == 0, na.lib.mm_get_error()


def test_device_md_without_cutoff_follows_host_loop(gpu, gen_dir):
    """Without a cutoff no pair can leave the list, so the device run skips
    the nsnb rebuilds; the host loop of md() rebuilds through mme() every
    nsnb steps, and the trajectories still agree (the rebuilt list is the
    same list)."""
    na = gpu
    parm, arr, x0 = generated_system(600, 31, gen_dir)
    n = 3 * arr["natom"]
    rng = np.random.default_rng(7)
    v0 = rng.normal(0, 0.05, n)
    opts = "gb=5, cut=999.0, rgbmax=999.0, nsnb=3, gamma_ln=0.0, tempi=0.0, dt=0.001, ntpr_md=1000, ntpr=100000"
    _init(na, parm, opts)
    xa, va = x0.copy(), v0.copy()
    _md(na, n, 25, xa, va)
    _init(na, parm, opts)
    xb, vb = x0.copy(), v0.copy()
    wrap = na.MMEFUNC(lambda xp, fp, itp: na.lib.mme(xp, fp, itp))
    _md(na, n, 25, xb, vb, func=wrap)
    assert np.max(np.abs(xa - xb)) < 1e-8
    assert rel_rms(va, vb) < 1e-8
    assert np.max(np.abs(xa - x0)) > 1e-3
