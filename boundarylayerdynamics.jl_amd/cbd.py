"""Cartesian Binary Data (CBD) files, the snapshot format of the reference
(src/fileio.jl:1-115), read and written on the host.

Layout (little endian): 8 identifier bytes (magic 0xCB 0xDF, format version
1, zeros, and the bytes per value 8 or 4 as the last byte), the grid sizes
n1, n2, n3 as UInt64, xmin[3] and xmax[3] as Float64, the coordinates x1[n1],
x2[n2], x3[n3] as Float64, then the values (Float64 or Float32) in
column-major order (x1 fastest).
"""

import os

import numpy as np

MAGIC = bytes.fromhex("CBDF")
VERSION = 1


def write_cbd(dtype, filename, field, x1, x2, x3, xmin, xmax):
    """writecbd (fileio.jl:27-65) for one process."""
    dtype = np.dtype(dtype)
    if dtype not in (np.float64, np.float32):
        raise ValueError("Only 64-bit and 32-bit precision is supported.")
    field = np.asarray(field)
    n1, n2, n3 = field.shape
    if len(x1) != n1:
        raise ValueError("Length of x1-values incompatible with dimensions of field.")
    if len(x2) != n2:
        raise ValueError("Length of x2-values incompatible with dimensions of field.")
    if len(x3) != n3:
        raise ValueError("Length of x3-values incompatible with dimensions of field.")
    ident = bytearray(8)
    ident[0:2] = MAGIC
    ident[2] = VERSION
    ident[7] = dtype.itemsize
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    with open(filename, "wb") as f:
        f.write(bytes(ident))
        f.write(np.array([n1, n2, n3], dtype="<u8").tobytes())
        f.write(np.asarray(xmin, dtype="<f8").tobytes())
        f.write(np.asarray(xmax, dtype="<f8").tobytes())
        for x in (x1, x2, x3):
            f.write(np.asarray(x, dtype="<f8").tobytes())
        f.write(np.asarray(field, dtype=dtype.newbyteorder("<")).tobytes(order="F"))


def read_cbd(filename, x3range=(-np.inf, np.inf), tol=1e-9):
    """readcbd (fileio.jl:67-110) for one process: (xmin, xmax, x1, x2, x3, data)
    with the x3-levels inside x3range (within tol)."""
    with open(filename, "rb") as f:
        ident = f.read(8)
        if ident[0:2] != MAGIC:
            raise ValueError("Not a CBD file (magic number not matching).")
        if ident[2] != VERSION:
            raise ValueError("Unsupported version of CBD format.")
        if ident[7] not in (4, 8):
            raise ValueError("Unsupported data type: Only 32-bit and 64-bit precision is supported.")
        dtype = np.dtype("<f8") if ident[7] == 8 else np.dtype("<f4")
        n1, n2, n3 = (int(v) for v in np.frombuffer(f.read(24), dtype="<u8"))
        xmin = tuple(np.frombuffer(f.read(24), dtype="<f8"))
        xmax = tuple(np.frombuffer(f.read(24), dtype="<f8"))
        x1 = np.frombuffer(f.read(8 * n1), dtype="<f8").copy()
        x2 = np.frombuffer(f.read(8 * n2), dtype="<f8").copy()
        x3 = np.frombuffer(f.read(8 * n3), dtype="<f8").copy()
        lo, hi = max(x3[0], x3range[0]), min(x3[-1], x3range[-1])
        i3min = max(i for i in range(n3) if x3[i] < lo + tol)
        i3max = min(i for i in range(n3) if x3[i] > hi - tol)
        f.seek(n1 * n2 * i3min * dtype.itemsize, os.SEEK_CUR)
        count = n1 * n2 * (i3max - i3min + 1)
        data = np.frombuffer(f.read(count * dtype.itemsize), dtype=dtype)
        data = data.reshape((n1, n2, i3max - i3min + 1), order="F").astype(dtype.newbyteorder("="))
    return xmin, xmax, x1, x2, x3[i3min:i3max + 1], data


def hrange(length, n, centered):
    """x1/x2 sample positions (physical_space.jl:64-65, scaled by the length)."""
    pts = np.linspace(0.0, 1.0, 2 * n + 1)[(1 if centered else 0):-1:2]
    return length * pts
