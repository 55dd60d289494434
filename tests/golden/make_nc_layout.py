"""Writes nc_layout_trpcage.json: the structure of the reference's Amber
netCDF trajectory test/trpcage.nc (dimensions, variables with their
dimensions, types and attributes, global attributes), read with scipy's
netCDF-3 reader.  The trajectory writer is checked against it.

Run in a checkout that has the reference at /root/reference:
    python tests/golden/make_nc_layout.py
"""
import json
import os

from scipy.io import netcdf_file

SRC = "/root/reference/test/trpcage.nc"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "nc_layout_trpcage.json")


def main():
    f = netcdf_file(SRC, "r", mmap=False)
    lay = {
        "source": "test/trpcage.nc (written by sander)",
        "version_byte": int(f.version_byte),
        "dimensions": {k: v for k, v in f.dimensions.items()},
        "variables": {k: {"dimensions": list(v.dimensions), "typecode": v.typecode(),
                          "attributes": {a: b.decode() for a, b in v._attributes.items()}}
                      for k, v in f.variables.items()},
        "global_attributes": {k: v.decode() for k, v in f._attributes.items()},
        "spatial": f.variables["spatial"][:].tobytes().decode(),
        "cell_spatial": f.variables["cell_spatial"][:].tobytes().decode(),
        "cell_angular": [r.tobytes().decode() for r in f.variables["cell_angular"][:]],
        "frames": int(f.variables["coordinates"].shape[0]),
    }
    f.close()
    with open(OUT, "w") as fp:
        json.dump(lay, fp, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
