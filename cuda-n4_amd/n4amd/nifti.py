"""Minimal NIfTI-1 single-file (.nii / .nii.gz) read/write for float volumes.

Arrays are C-order [x][y][z] with z fastest (the library's layout); on disk
NIfTI stores i fastest, so an array of shape (fx, fy, fz) is the NIfTI volume
dim = (fz, fy, fx) -- the reference CLI's axis flip (main.cpp:153-155).
"""
from __future__ import annotations

import gzip
import struct

import numpy as np

_DTYPES = {2: np.uint8, 4: np.int16, 8: np.int32, 16: np.float32, 64: np.float64, 256: np.int8,
           512: np.uint16, 768: np.uint32}


def write(path: str, vol: np.ndarray, spacing=(1.0, 1.0, 1.0)) -> None:
    vol = np.ascontiguousarray(vol, dtype=np.float32)
    fx, fy, fz = vol.shape
    hdr = bytearray(348)
    struct.pack_into("<i", hdr, 0, 348)
    struct.pack_into("<8h", hdr, 40, 3, fz, fy, fx, 1, 1, 1, 1)
    struct.pack_into("<hh", hdr, 70, 16, 32)
    struct.pack_into("<8f", hdr, 76, 1.0, spacing[0], spacing[1], spacing[2], 0, 0, 0, 0)
    struct.pack_into("<fff", hdr, 108, 352.0, 1.0, 0.0)
    hdr[344:348] = b"n+1\0"
    blob = bytes(hdr) + b"\0\0\0\0" + vol.tobytes()
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "wb") as f:
        f.write(blob)


def read(path: str) -> np.ndarray:
    opener = gzip.open if path.endswith(".gz") else open
    with opener(path, "rb") as f:
        blob = f.read()
    if struct.unpack_from("<i", blob, 0)[0] != 348:
        raise ValueError("not a little-endian NIfTI-1 file")
    dim = struct.unpack_from("<8h", blob, 40)
    dt = struct.unpack_from("<h", blob, 70)[0]
    off = int(struct.unpack_from("<f", blob, 108)[0])
    slope, inter = struct.unpack_from("<ff", blob, 112)
    n = dim[1] * dim[2] * dim[3]
    a = np.frombuffer(blob, dtype=_DTYPES[dt], count=n, offset=off).astype(np.float64)
    if slope not in (0.0, 1.0) or inter != 0.0:
        a = a * slope + inter
    return a.astype(np.float32).reshape(dim[3], dim[2], dim[1])
