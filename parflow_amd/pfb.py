"""ParFlow binary (PFB) reader/writer.

Format (pfsimulator/parflow_lib/write_parflow_binary.c:122,
read_parflow_binary.c:80; pftools/python/parflow/tools/io.py:78-91,
235-250): big-endian header ``x y z`` (f64), ``nx ny nz`` (i32),
``dx dy dz`` (f64), ``num_subgrids`` (i32), then per subgrid
``ix iy iz nx ny nz rx ry rz`` (i32) followed by ``nx*ny*nz`` f64 values,
x fastest.
"""
from __future__ import annotations

import struct

import numpy as np


def read_pfb(path):
    """Returns (header dict, array of shape (nz, ny, nx))."""
    with open(path, "rb") as f:
        buf = f.read()
    x, y, z = struct.unpack(">ddd", buf[0:24])
    nx, ny, nz = struct.unpack(">iii", buf[24:36])
    dx, dy, dz = struct.unpack(">ddd", buf[36:60])
    ns = struct.unpack(">i", buf[60:64])[0]
    out = np.zeros((nz, ny, nx), dtype=np.float64)
    off = 64
    for _ in range(ns):
        ix, iy, iz, snx, sny, snz, _rx, _ry, _rz = struct.unpack(">9i", buf[off:off + 36])
        off += 36
        n = snx * sny * snz
        a = np.frombuffer(buf[off:off + 8 * n], dtype=">f8").reshape(snz, sny, snx)
        off += 8 * n
        out[iz:iz + snz, iy:iy + sny, ix:ix + snx] = a
    hdr = dict(x=x, y=y, z=z, nx=nx, ny=ny, nz=nz, dx=dx, dy=dy, dz=dz,
               num_subgrids=ns)
    return hdr, out


def write_pfb(path, data, x=0.0, y=0.0, z=0.0, dx=1.0, dy=1.0, dz=1.0):
    """Writes a single-subgrid PFB file from an (nz, ny, nx) array."""
    data = np.asarray(data, dtype=np.float64)
    nz, ny, nx = data.shape
    with open(path, "wb") as f:
        f.write(struct.pack(">ddd", x, y, z))
        f.write(struct.pack(">iii", nx, ny, nz))
        f.write(struct.pack(">ddd", dx, dy, dz))
        f.write(struct.pack(">i", 1))
        f.write(struct.pack(">9i", 0, 0, 0, nx, ny, nz, 1, 1, 1))
        f.write(data.astype(">f8").tobytes())
