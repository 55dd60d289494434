"""Amber netCDF trajectory writer (nctraj.c) on the host.

The file md() writes must have the layout of the reference's trajectories
(AmberNetcdf.c:595-735; pinned by the structure of the reference's own
test/trpcage.nc, tests/golden/nc_layout_trpcage.json): 64-bit-offset
classic netCDF, the same dimensions, variables, types, units and
conventions, frames of single-precision coordinates (AmberNetcdf.c:1000-
1011).  The file is read back with scipy's independent netCDF-3 reader.
"""
import ctypes as C
import json
import os

import numpy as np
from scipy.io import netcdf_file

from common import GOLDEN


class NCTraj(C.Structure):
    _fields_ = [("fp", C.c_void_p), ("natom", C.c_int), ("nframes", C.c_int),
                ("rec_begin", C.c_longlong), ("recsize", C.c_longlong)]


def _writer():
    import nabc_amd as na
    lib = na.lib
    lib.nabc_nctraj_create.argtypes = [C.POINTER(NCTraj), C.c_char_p, C.c_int]
    lib.nabc_nctraj_create.restype = C.c_int
    lib.nabc_nctraj_write.argtypes = [C.POINTER(NCTraj), C.POINTER(C.c_double), C.c_double]
    lib.nabc_nctraj_write.restype = C.c_int
    lib.nabc_nctraj_close.argtypes = [C.POINTER(NCTraj)]
    lib.nabc_nctraj_close.restype = None
    return lib


def _write(path, frames, times):
    lib = _writer()
    t = NCTraj()
    natom = frames[0].shape[0] // 3
    assert lib.nabc_nctraj_create(C.byref(t), path.encode(), natom) == 0
    for x, tm in zip(frames, times):
        x = np.ascontiguousarray(x, dtype=np.float64)
        assert lib.nabc_nctraj_write(C.byref(t), x.ctypes.data_as(C.POINTER(C.c_double)), tm) == 0
    return lib, t


def test_layout_matches_reference_trajectory(tmp_path):
    ref = json.load(open(os.path.join(GOLDEN, "nc_layout_trpcage.json")))
    rng = np.random.default_rng(3)
    natom = 37
    frames = [rng.normal(0, 20, 3 * natom) for _ in range(4)]
    times = [0.5, 1.0, 1.5, 2.0]
    path = str(tmp_path / "t.nc")
    lib, t = _write(path, frames, times)
    lib.nabc_nctraj_close(C.byref(t))
    f = netcdf_file(path, "r", mmap=False)
    assert f.version_byte == ref["version_byte"] == 2
    dims = dict(ref["dimensions"])
    dims["atom"] = natom
    assert f.dimensions == dims
    assert set(f.variables) == set(ref["variables"])
    for name, v in ref["variables"].items():
        got = f.variables[name]
        assert list(got.dimensions) == v["dimensions"], name
        assert got.typecode() == v["typecode"], name
        assert {a: b.decode() for a, b in got._attributes.items()} == v["attributes"], name
    g = {k: v.decode() for k, v in f._attributes.items()}
    for k in ("application", "Conventions", "ConventionVersion"):
        assert g[k] == ref["global_attributes"][k]
    assert set(g) == set(ref["global_attributes"])
    assert f.variables["spatial"][:].tobytes().decode() == ref["spatial"]
    assert f.variables["cell_spatial"][:].tobytes().decode() == ref["cell_spatial"]
    assert [r.tobytes().decode() for r in f.variables["cell_angular"][:]] == ref["cell_angular"]
    c = f.variables["coordinates"][:]
    assert c.shape == (4, natom, 3)
    for k in range(4):
        assert np.array_equal(c[k], frames[k].astype(np.float32).reshape(natom, 3))
    assert np.array_equal(f.variables["time"][:], np.array(times, dtype=np.float32))
    f.close()


def test_file_is_complete_after_every_frame(tmp_path):
    """The record count is rewritten and the file flushed per frame: a
    reader sees every frame written so far while the writer is open."""
    rng = np.random.default_rng(5)
    natom = 5
    path = str(tmp_path / "u.nc")
    lib = _writer()
    t = NCTraj()
    assert lib.nabc_nctraj_create(C.byref(t), path.encode(), natom) == 0
    f = netcdf_file(path, "r", mmap=False)
    assert f.variables["coordinates"].shape == (0, natom, 3)
    f.close()
    for k in range(1, 4):
        x = rng.normal(0, 5, 3 * natom)
        assert lib.nabc_nctraj_write(C.byref(t), x.ctypes.data_as(C.POINTER(C.c_double)), 0.1 * k) == 0
        f = netcdf_file(path, "r", mmap=False)
        assert f.variables["coordinates"].shape[0] == k
        assert np.array_equal(f.variables["coordinates"][k - 1], x.astype(np.float32).reshape(natom, 3))
        f.close()
    lib.nabc_nctraj_close(C.byref(t))


def test_bad_arguments_are_errors(tmp_path):
    lib = _writer()
    t = NCTraj()
    assert lib.nabc_nctraj_create(C.byref(t), str(tmp_path / "v.nc").encode(), 0) != 0
    assert lib.nabc_nctraj_create(C.byref(t), str(tmp_path / "no_such_dir" / "w.nc").encode(), 3) != 0
    x = np.zeros(9)
    assert lib.nabc_nctraj_write(C.byref(t), x.ctypes.data_as(C.POINTER(C.c_double)), 0.0) != 0


class AmberNetcdf(C.Structure):
    _fields_ = [("temp0", C.c_double), ("restartTime", C.c_double)] + [
        (n, C.c_int) for n in ("isNCrestart", "ncid", "frameDID", "ncframe", "currentFrame", "atomDID", "ncatom",
                               "ncatom3", "coordVID", "velocityVID", "cellAngleVID", "cellLengthVID", "spatialDID",
                               "labelDID", "cell_spatialDID", "cell_angularDID", "spatialVID", "timeVID",
                               "cell_spatialVID", "cell_angularVID", "TempVID")]


def _api():
    import nabc_amd as na
    lib = na.lib
    A, dp = C.POINTER(AmberNetcdf), C.POINTER(C.c_double)
    for name, args in (("netcdfLoad", [A, C.c_char_p]), ("netcdfClose", [A]), ("netcdfCreate", [A, C.c_char_p, C.c_int, C.c_int]),
                       ("netcdfGetFrame", [A, C.c_int, dp, dp]), ("netcdfGetNextFrame", [A, dp, dp]),
                       ("netcdfGetVelocity", [A, C.c_int, dp]), ("netcdfWriteFrame", [A, C.c_int, dp, dp]),
                       ("netcdfWriteNextFrame", [A, dp, dp]),
                       ("netcdfWriteRestart", [C.c_char_p, C.c_int, dp, dp, dp, C.c_double, C.c_double])):
        getattr(lib, name).argtypes = args
        getattr(lib, name).restype = C.c_int
    return lib


def _p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


def test_reads_reference_trajectory():
    """netcdfLoad / netcdfGetFrame on the reference's test/trpcage.nc (a
    trajectory sander wrote through the netCDF library) give what scipy
    reads from it."""
    lib = _api()
    path = os.path.join(GOLDEN, "trpcage.nc")
    f = netcdf_file(path, "r", mmap=False)
    ref = f.variables["coordinates"][:].astype(np.float64)
    f.close()
    A = AmberNetcdf()
    assert lib.netcdfLoad(C.byref(A), path.encode()) == 0
    assert (A.ncframe, A.ncatom, A.ncatom3, A.isNCrestart) == (101, 223, 669, 0)
    assert A.velocityVID == -1 and A.cellLengthVID == -1
    x = np.zeros(3 * 223)
    for k in (0, 50, 100):
        assert lib.netcdfGetFrame(C.byref(A), k, _p(x), None) == 0
        assert np.array_equal(x, ref[k].ravel())
    A.currentFrame = 99
    assert lib.netcdfGetNextFrame(C.byref(A), _p(x), None) == 0 and A.currentFrame == 100
    assert np.array_equal(x, ref[99].ravel())
    assert lib.netcdfGetFrame(C.byref(A), 101, _p(x), None) != 0
    assert lib.netcdfClose(C.byref(A)) == 0 and A.ncid == -1


def test_create_write_load_round_trip(tmp_path):
    lib = _api()
    rng = np.random.default_rng(8)
    natom = 11
    path = str(tmp_path / "c.nc").encode()
    A = AmberNetcdf()
    assert lib.netcdfCreate(C.byref(A), path, natom, 0) == 0
    frames = [rng.normal(0, 9, 3 * natom) for _ in range(3)]
    for fr in frames:
        assert lib.netcdfWriteNextFrame(C.byref(A), _p(fr), None) == 0
    assert A.currentFrame == 3 and A.ncframe == 3
    # a frame rewritten in place
    frames[1] = rng.normal(0, 9, 3 * natom)
    assert lib.netcdfWriteFrame(C.byref(A), 1, _p(frames[1]), None) == 0
    assert lib.netcdfClose(C.byref(A)) == 0
    assert lib.netcdfCreate(C.byref(A), path, natom, 1) != 0  # box output is not written
    B = AmberNetcdf()
    assert lib.netcdfLoad(C.byref(B), path) == 0
    assert (B.ncframe, B.ncatom, B.isNCrestart) == (3, natom, 0)
    x = np.zeros(3 * natom)
    for k in range(3):
        assert lib.netcdfGetFrame(C.byref(B), k, _p(x), None) == 0
        assert np.array_equal(x, frames[k].astype(np.float32).astype(np.float64))
    assert lib.netcdfClose(C.byref(B)) == 0


def test_restart_write_and_load(tmp_path):
    """netcdfWriteRestart: the AMBERRESTART layout of AmberNetcdf.c:366-593
    (FP64 coordinates, velocities with scale_factor 20.455, time, temp0),
    checked with scipy and read back with netcdfLoad."""
    lib = _api()
    rng = np.random.default_rng(9)
    natom = 13
    X = rng.normal(0, 10, 3 * natom)
    V = rng.normal(0, 0.1, 3 * natom)
    path = str(tmp_path / "r.ncrst")
    assert lib.netcdfWriteRestart(path.encode(), natom, _p(X), _p(V), None, 12.5, 300.0) == 0
    f = netcdf_file(path, "r", mmap=False)
    assert f.version_byte == 2
    assert f._attributes["Conventions"] == b"AMBERRESTART"
    assert "frame" not in f.dimensions and f.dimensions["atom"] == natom
    assert np.array_equal(f.variables["coordinates"][:].ravel(), X)
    assert np.array_equal(f.variables["velocities"][:].ravel(), V)
    assert f.variables["velocities"]._attributes["scale_factor"] == 20.455
    assert f.variables["velocities"]._attributes["units"] == b"angstrom/picosecond"
    assert float(f.variables["time"].getValue()) == 12.5
    assert float(f.variables["temp0"].getValue()) == 300.0
    f.close()
    A = AmberNetcdf()
    assert lib.netcdfLoad(C.byref(A), path.encode()) == 0
    assert (A.isNCrestart, A.ncframe, A.ncatom, A.restartTime, A.temp0) == (1, 1, natom, 12.5, 300.0)
    x, v = np.zeros(3 * natom), np.zeros(3 * natom)
    assert lib.netcdfGetFrame(C.byref(A), 0, _p(x), None) == 0 and np.array_equal(x, X)
    assert lib.netcdfGetVelocity(C.byref(A), 0, _p(v)) == 0 and np.array_equal(v, V)
    assert lib.netcdfClose(C.byref(A)) == 0
    # without velocities and temperature
    path2 = str(tmp_path / "r2.ncrst")
    assert lib.netcdfWriteRestart(path2.encode(), natom, _p(X), None, None, -1.0, -1.0) == 0
    assert lib.netcdfLoad(C.byref(A), path2.encode()) == 0
    assert A.velocityVID == -1 and A.TempVID == -1 and A.restartTime == 0.0
    assert lib.netcdfGetVelocity(C.byref(A), 0, _p(v)) != 0
    assert lib.netcdfClose(C.byref(A)) == 0
    assert lib.netcdfLoad(C.byref(A), str(tmp_path / "missing.nc").encode()) != 0
