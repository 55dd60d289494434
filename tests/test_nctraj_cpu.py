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
