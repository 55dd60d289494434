"""The oracle's md() (orc_md, sff.c:2071-2366) against what the mechanics fixes (no GPU).

orc_md is the checker for the device dynamics, so it is pinned first:
  - the noise stream: orc_gauss is the same function of (seed, step, index)
    as the product's nabc_md_gauss (the documented stand-in for rand2.c),
    and its deviates have mean 0 and variance 1;
  - velocity Verlet without coupling (gamma_ln = 0, no Berendsen scaling:
    tautp huge) conserves the total energy (a wrong half-step factor, a
    force sign or a unit slip between dtx = 20.455 dt and the kcal/mol
    forces makes it drift by orders of magnitude more) and is time
    reversible: stepping forward, negating v and stepping back returns to
    the start;
  - the Langevin scheme holds the target temperature on average, Berendsen
    scaling (short tautp) pulls the temperature to temp0;
  - frozen atoms do not move; the per-step rows carry mme34's energies at
    the step's coordinates and the kinetic energy of the step's velocities.
"""
import numpy as np
import pytest

import oracle

N = 600
_min = {}


def _minimised(na):
    """A 600-atom generated system after 300 CG iterations of the oracle's
    mme (host conjgrad over orc_mme_cb): bonded terms near their minima, so
    the dynamics start from a physical configuration."""
    if "x" not in _min:
        p, a, x = na.build_system(N, 41)
        o = oracle.Oracle(p, N, gb=5, cut=999.0, rgbmax=999.0)
        oracle.lib.orc_set_current(o.h, 1)
        xc = np.ascontiguousarray(x)
        n = na.C.c_int(3 * N)
        f = na.C.c_double()
        rms = na.C.c_double(1e-3)
        dfp = na.C.c_double(10.0)
        mx = na.C.c_int(300)
        cb = na.MMEFUNC(oracle.lib.orc_mme_cb)
        na.lib.conjgrad(xc.ctypes.data_as(na.dp), na.C.byref(n), na.C.byref(f), cb, na.C.byref(rms),
                        na.C.byref(dfp), na.C.byref(mx))
        oracle.lib.orc_set_current(None, 1)
        o.close()
        _min["x"] = (p, a, xc.copy())
    return _min["x"]


def test_noise_stream_matches_product_and_is_standard_normal():
    import nabc_amd as na
    for seed, step, idx in ((71277, -1, 0), (1, 0, 5), (2 ** 63 + 5, 12345, 7777), (0, 999, 3)):
        assert oracle.lib.orc_gauss(seed, step, idx) == na.lib.nabc_md_gauss(seed, step, idx)
    z = np.array([oracle.lib.orc_gauss(9, s, i) for s in range(40) for i in range(2000)])
    assert abs(z.mean()) < 0.02 and abs(z.var() - 1.0) < 0.02
    assert abs(np.mean(z ** 4) - 3.0) < 0.1  # Gaussian kurtosis, not a uniform's 1.8


def test_nve_conserves_energy_and_is_reversible():
    """Over the same 0.1 ps at dt = 0.25 and 0.5 fs: the total energy's
    spread is a small fraction of the kinetic energy's (energy flows between
    the two, the sum stays), and it grows 4x when dt doubles, as a
    second-order integrator's error does."""
    import nabc_amd as na
    p, a, x0 = _minimised(na)
    o = oracle.Oracle(p, N, gb=5, cut=999.0, rgbmax=999.0)
    spread = {}
    for dt, nst in ((0.00025, 400), (0.0005, 200)):
        x1, v1, rows = o.md(x0, None, a["mass"], nst, dt=dt, tempi=300.0, tautp=1e300, seed=5)
        etot = rows[:, 0] + rows[:, 14]
        ekin = rows[:, 14]
        assert 200.0 < ekin.mean() < 600.0  # ~300 K shared with the potential energy
        assert etot.std() < 0.05 * ekin.std(), (dt, etot.std(), ekin.std())
        spread[dt] = etot.std()
    assert 3.0 < spread[0.0005] / spread[0.00025] < 5.0, spread
    # back along the same path: negate v and step again (no noise, no scaling)
    x2, v2, _ = o.md(x1, -v1, a["mass"], 200, dt=0.0005, tempi=0.0, tautp=1e300)
    assert np.max(np.abs(x2 - x0)) < 1e-7
    o.close()


def test_langevin_and_berendsen_temperature():
    import nabc_amd as na
    p, a, x0 = _minimised(na)
    o = oracle.Oracle(p, N, gb=5, cut=999.0, rgbmax=999.0)
    _, _, rows = o.md(x0, None, a["mass"], 600, dt=0.001, tempi=300.0, temp0=300.0, gamma_ln=5.0, seed=3)
    assert 270.0 < rows[200:, 15].mean() < 330.0
    _, _, rows = o.md(x0, None, a["mass"], 400, dt=0.001, tempi=100.0, temp0=300.0, tautp=0.02, seed=3)
    assert rows[:5, 15].mean() < 200.0 and 270.0 < rows[300:, 15].mean() < 330.0
    o.close()


def test_frozen_atoms_stay_and_rows_are_mme34():
    import nabc_amd as na
    p, a, x0 = _minimised(na)
    frozen = np.zeros(N, np.int32)
    frozen[::5] = 1
    o = oracle.Oracle(p, N, gb=5, cut=12.0, rgbmax=12.0, frozen=frozen)
    x1, v1, rows = o.md(x0, None, a["mass"], 30, dt=0.001, tempi=300.0, gamma_ln=5.0, nsnb=10, seed=8)
    fz = np.repeat(frozen.astype(bool), 3)
    assert np.array_equal(x1[fz], x0[fz]) and np.all(v1[fz] == 0.0)
    assert np.any(x1[~fz] != x0[~fz])
    # the last row: mme34 at the final coordinates with the list of step 30
    # (built there: 30 % nsnb == 0), and the kinetic energy of v (gamma > 0:
    # scaled by c_ave, sff.c:2309-2310)
    ene, frms, _ = o.mme(x1, rebuild=True)
    assert np.allclose(rows[-1, :13], ene, rtol=1e-12, atol=1e-9) and abs(rows[-1, 13] - frms) < 1e-12
    gammai = 5.0 / 20.455
    c_ave = 1.0 + gammai * 0.5 * 0.001 * 20.455
    m3 = np.repeat(a["mass"], 3)
    assert abs(rows[-1, 14] - 0.5 * np.sum(v1 * v1 * m3) * c_ave) < 1e-9 * rows[-1, 14]
    o.close()
