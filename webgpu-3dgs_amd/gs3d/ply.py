"""Native PLY ingest through libgs3d.so (gs3d_ply_parse / gs3d_load_ply).

The C++ parser restates the reference loader (src/workers/SceneWorker.ts:146-549):
the same header rules, error messages and f64 activations stored as f32.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from . import _lib as L
from .scene import Splats


class PlyError(ValueError):
    """A malformed or unsupported PLY (the reference loader's error message)."""


def probe(data: bytes | np.ndarray) -> L.PlyInfo:
    buf = np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray, memoryview)) else data
    info = L.PlyInfo()
    err = C.create_string_buffer(512)
    rc = L.lib().gs3d_ply_probe(L.ptr(buf), buf.nbytes, C.byref(info), err, len(err))
    if rc != L.GS3D_OK:
        raise PlyError(err.value.decode(errors="replace"))
    return info


def parse(data: bytes | np.ndarray, threads: int = 0) -> tuple[Splats, np.ndarray | None, tuple]:
    """Parse a whole PLY held in memory.

    Returns (splats, f_rest, bounds): splats in the loader's layouts (sh_rest
    holds the first (d+1)^2-1 coefficients of each channel when the file has
    more), f_rest the raw N x f_rest_count array in index order (or None) and
    bounds ((min x, y, z), (max x, y, z)) as float64.
    """
    buf = np.frombuffer(data, np.uint8) if isinstance(data, (bytes, bytearray, memoryview)) else data
    info = probe(buf)
    n, F = int(info.splat_count), int(info.f_rest_count)
    pos, sc, rot, col = (np.empty((n, 4), np.float32) for _ in range(4))
    dc = np.empty((n, 3), np.float32)
    rest = np.empty((n, F), np.float32) if F else None
    err = C.create_string_buffer(512)
    rc = L.lib().gs3d_ply_parse(L.ptr(buf), buf.nbytes, L.ptr(pos), L.ptr(sc), L.ptr(rot), L.ptr(col),
                                L.ptr(dc), L.ptr(rest) if rest is not None else None, threads,
                                C.byref(info), err, len(err))
    if rc != L.GS3D_OK:
        raise PlyError(err.value.decode(errors="replace"))
    deg = int(info.sh_degree)
    sh = None
    if deg > 0:
        per = (deg + 1) ** 2 - 1
        sh = np.ascontiguousarray(rest.reshape(n, 3, F // 3)[:, :, :per].reshape(n, 3 * per))
    bounds = (tuple(info.bounds_min), tuple(info.bounds_max))
    return Splats(pos, col, sc, rot, dc, sh, deg, bounds if n else ()), rest, bounds


def parse_file(path: str | Path, threads: int = 0):
    return parse(np.fromfile(str(path), np.uint8), threads)
