"""Minimal MRC2014 reader/writer (memory-mapped reads, float32 writes).

Supports the data modes used for maps and particle stacks: 0 (int8),
1 (int16), 2 (float32), 6 (uint16), 12 (float16).  Data are returned in
file order [sections][rows][columns] = [z][y][x].
"""
from __future__ import annotations

import os
import struct

import numpy as np

_MODES = {0: np.int8, 1: np.int16, 2: np.float32, 6: np.uint16, 12: np.float16}
HEADER_BYTES = 1024


def read_header(path):
    with open(path, "rb") as fh:
        raw = fh.read(HEADER_BYTES)
    if len(raw) < HEADER_BYTES:
        raise ValueError(f"{path}: truncated MRC header")
    stamp = raw[212:214]
    endian = "<" if stamp in (b"\x44\x44", b"\x44\x41", b"\x00\x00") else ">"
    nx, ny, nz, mode = struct.unpack(endian + "4i", raw[:16])
    nsymbt = struct.unpack(endian + "i", raw[92:96])[0]
    if mode not in _MODES:
        raise ValueError(f"{path}: unsupported MRC mode {mode}")
    return {"nx": nx, "ny": ny, "nz": nz, "mode": mode, "nsymbt": nsymbt, "endian": endian}


def mmap(path):
    """Memory-mapped read-only array [nz, ny, nx] (or [ny, nx] when nz == 1)."""
    h = read_header(path)
    dt = np.dtype(_MODES[h["mode"]]).newbyteorder(h["endian"])
    shape = (h["nz"], h["ny"], h["nx"])
    arr = np.memmap(path, dtype=dt, mode="r", offset=HEADER_BYTES + h["nsymbt"], shape=shape)
    return arr[0] if h["nz"] == 1 else arr


def read(path) -> np.ndarray:
    return np.array(mmap(path), dtype=np.float32, copy=True)


def write_mrc(path, data, voxel_size=1.0):
    """Write float32 data ([z][y][x] or [y][x]) with a standard MRC2014 header."""
    a = np.ascontiguousarray(np.asarray(data, np.float32))
    if a.ndim == 2:
        a = a[None]
    if a.ndim != 3:
        raise ValueError("MRC data must be 2-D or 3-D")
    nz, ny, nx = a.shape
    hdr = bytearray(HEADER_BYTES)
    struct.pack_into("<4i", hdr, 0, nx, ny, nz, 2)
    struct.pack_into("<3i", hdr, 28, nx, ny, nz)                       # mx, my, mz
    struct.pack_into("<3f", hdr, 40, nx * voxel_size, ny * voxel_size, nz * voxel_size)
    struct.pack_into("<3f", hdr, 52, 90.0, 90.0, 90.0)
    struct.pack_into("<3i", hdr, 64, 1, 2, 3)                          # mapc, mapr, maps
    struct.pack_into("<3f", hdr, 76, float(a.min()), float(a.max()), float(a.mean()))
    hdr[208:212] = b"MAP "
    hdr[212:216] = b"\x44\x44\x00\x00"
    struct.pack_into("<f", hdr, 216, float(a.std()))
    struct.pack_into("<i", hdr, 196, 20140)                            # nversion
    tmp = f"{os.fspath(path)}.tmp.{os.getpid()}"
    with open(tmp, "wb") as fh:
        fh.write(hdr)
        fh.write(a.astype("<f4").tobytes())
    os.replace(tmp, path)
