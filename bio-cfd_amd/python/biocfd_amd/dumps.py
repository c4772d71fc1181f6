"""Reader for the record files written by the reference driver harness and by
this package's own dump routine.

Record layout (little endian): name (24 bytes, blank padded), kind (int64:
1 = float64, 2 = int64), ndim (int64), shape (ndim x int64, Fortran order),
then the data in Fortran (column-major) order.
"""
from __future__ import annotations

import struct
from typing import Dict

import numpy as np


def read_dump(path: str) -> Dict[str, np.ndarray]:
    out: Dict[str, np.ndarray] = {}
    with open(path, "rb") as f:
        buf = f.read()
    pos = 0
    while pos < len(buf):
        name = buf[pos:pos + 24].decode("ascii").strip()
        pos += 24
        kind, ndim = struct.unpack_from("<2q", buf, pos)
        pos += 16
        shape = struct.unpack_from(f"<{ndim}q", buf, pos) if ndim else ()
        pos += 8 * ndim
        count = int(np.prod(shape)) if ndim else 1
        dt = np.float64 if kind == 1 else np.int64
        arr = np.frombuffer(buf, dtype=dt, count=count, offset=pos)
        pos += 8 * count
        if ndim:
            arr = arr.reshape(shape, order="F")
        else:
            arr = arr[0]
        out[name] = arr.copy() if ndim else arr
    return out
