"""Host logic of nmode (nmode.c of the library), no GPU: the energy
function is a Python callback with the mme2 signature.

- A diatomic spring (force constant k along x, masses m1, m2) has one
  vibration, of eigenvalue k (1/m1 + 1/m2) of the mass-weighted Hessian,
  printed as 108.587 sqrt(k / mu) cm^-1 (nmode.c:1920-1945 of the
  reference); the other five modes are zero.  This pins the mass weighting
  and the conversion factor by a closed form.
- The error paths: Langevin modes (ntrun=3) are refused, and a callback
  whose atom count does not match n makes nmode fail with a message.
"""
import ctypes
import math
import re

import numpy as np
import pytest

K_SPRING = 570.0
MASSES = (12.01, 16.0)


def _diatomic(natom_report):
    import nabc_amd as na
    calls = {"n": 0}

    def func(x, f, g, m, grad, *rest):
        natom_p, iter_p = rest[7], rest[8]
        if iter_p[0] == -3:
            return 0.0
        calls["n"] += 1
        natom_p[0] = natom_report
        for i in range(6):
            f[i] = 0.0
        for i in range(36):
            g[i] = 0.0
        # d2E/dx1^2 = d2E/dx2^2 = k, d2E/dx1 dx2 = -k
        g[0] = g[3 * 6 + 3] = K_SPRING
        g[3] = g[3 * 6] = -K_SPRING
        m[0], m[1] = MASSES
        return -1.0

    return na, na.MME2FUNC(func), calls


def test_diatomic_frequency_closed_form(tmp_path, capfd, monkeypatch):
    na, cb, calls = _diatomic(2)
    x = np.array([0.0, 0.0, 0.0, 1.5, 0.0, 0.0])
    monkeypatch.chdir(tmp_path)  # eigp > 0 writes the modes file "vecs"
    capfd.readouterr()
    ret = na.lib.nmode(x.ctypes.data_as(na.dp), 6, cb, 6, 2, 0.0, 0.0, 0)
    out = capfd.readouterr().out
    assert ret == 0, na.lib.mm_get_error()
    assert calls["n"] == 1
    got = {int(a): float(b) for a, b in re.findall(r"^\s*(\d+)\s+(-?[\d.]+)\s*$", out, re.M)}
    mu = MASSES[0] * MASSES[1] / (MASSES[0] + MASSES[1])
    assert abs(got[6] - 108.587 * math.sqrt(K_SPRING / mu)) <= 1e-5
    assert all(abs(got[i]) < 1e-4 for i in range(1, 6))
    assert open("vecs").read().count("****") == 6


def test_langevin_modes_are_refused():
    na, cb, calls = _diatomic(2)
    x = np.zeros(6)
    assert na.lib.nmode(x.ctypes.data_as(na.dp), 6, cb, 0, 3, 0.0, 0.0, 0) == -1
    assert b"ntrun=3" in na.lib.mm_get_error()
    assert calls["n"] == 0


def test_atom_count_mismatch_is_an_error():
    na, cb, calls = _diatomic(3)
    x = np.zeros(6)
    assert na.lib.nmode(x.ctypes.data_as(na.dp), 6, cb, 0, 1, 0.0, 0.0, 0) == -1
    assert b"3 * natom" in na.lib.mm_get_error()
