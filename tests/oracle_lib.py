"""ctypes access to the CPU oracle (oracle/liboracle_gs3d.so) for the tests.

The oracle is the checker: tests compare the HIP path against it.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent.parent
ORACLE = REPO / "oracle" / "liboracle_gs3d.so"

P, U32, I32, F32, SZ = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_size_t
F64 = C.c_double
_PROTOS = {
    "orc_set_threads": (I32, [I32]),
    "orc_f32_to_f16": (C.c_uint16, [F32]),
    "orc_f16_to_f32": (F32, [C.c_uint16]),
    "orc_f32_to_f16_array": (None, [P, P, SZ]),
    "orc_scan": (U32, [P, U32, P]),
    "orc_emit": (None, [P, P, U32, U32, P, P]),
    "orc_radix_sort": (I32, [P, P, SZ, I32]),
    "orc_identify": (None, [P, U32, U32, P]),
    "orc_preprocess_f32": (None, [P, U32, P, P, P, P, P, P, I32, U32, U32, F32, F32, P, P]),
    "orc_rasterize_f32": (None, [P, P, P, U32, U32, U32, U32, I32, P, P, P]),
    "orc_sh_basis_f32": (None, [F32, F32, F32, I32, P]),
    "orc_sh_basis_f64": (None, [F64, F64, F64, I32, P]),
    "orc_render_f32": (C.c_int64, [P, U32, P, P, P, P, P, P, I32, U32, U32, U32, U32, P, P, P]),
    "orc_render_f64": (C.c_int64, [P, U32, P, P, P, P, P, P, I32, U32, U32, U32, U32, P, P, P]),
    "orc_tile_overlay_f32": (None, [P, P, U32, U32, U32, U32, U32, P]),
    "orc_composite_f32": (None, [P, U32, U32, I32, P, P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        if not ORACLE.exists():
            raise RuntimeError("oracle not built; run __graft_entry__.build()")
        l = C.CDLL(str(ORACLE))
        for k, (r, a) in _PROTOS.items():
            f = getattr(l, k)
            f.restype = r
            f.argtypes = a
        _lib = l
    return _lib


def _p(a):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def set_threads(n: int) -> int:
    return lib().orc_set_threads(n)


def f16_round(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(x.shape, np.uint16)
    lib().orc_f32_to_f16_array(_p(x), _p(out), x.size)
    return out


def scan(counts: np.ndarray):
    counts = np.ascontiguousarray(counts, np.uint32)
    out = np.zeros(counts.size + 1, np.uint32)
    tot = lib().orc_scan(_p(counts), counts.size, _p(out))
    return out, int(tot)


def emit(projected: np.ndarray, offsets: np.ndarray, tiles_x: int):
    projected = np.ascontiguousarray(projected, np.float32)
    offsets = np.ascontiguousarray(offsets, np.uint32)
    n = offsets.size - 1
    k = int(offsets[n])
    keys = np.zeros(k, np.uint64)
    vals = np.zeros(k, np.uint32)
    lib().orc_emit(_p(projected), _p(offsets), n, tiles_x, _p(keys), _p(vals))
    return keys, vals


def radix_sort(keys: np.ndarray, values: np.ndarray, bits: int = 8):
    keys = np.array(keys, np.uint64, copy=True)
    values = np.array(values, np.uint32, copy=True)
    assert lib().orc_radix_sort(_p(keys), _p(values), keys.size, bits) == 0
    return keys, values


def identify(sorted_keys: np.ndarray, tile_count: int):
    sorted_keys = np.ascontiguousarray(sorted_keys, np.uint64)
    out = np.zeros(tile_count + 1, np.uint32)
    lib().orc_identify(_p(sorted_keys), sorted_keys.size, tile_count, _p(out))
    return out


def _splat_args(s):
    c = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    return [c(s.positions), c(s.scales), c(s.rotations), c(s.colors),
            c(s.f_dc) if s.sh_degree else None, c(s.sh_rest) if s.sh_degree else None]


def preprocess(cam: np.ndarray, s, tiles_x, tiles_y, inv_tx, inv_ty):
    cam = np.ascontiguousarray(cam, np.float32)
    args = _splat_args(s)
    proj = np.zeros((s.n, 16), np.float32)
    counts = np.zeros(s.n, np.uint32)
    lib().orc_preprocess_f32(_p(cam), s.n, *[_p(a) for a in args], s.sh_degree, tiles_x, tiles_y,
                             inv_tx, inv_ty, _p(proj), _p(counts))
    return proj, counts


def rasterize(projected, sorted_values, tile_offsets, tile, width, height, quantize=True):
    projected = np.ascontiguousarray(projected, np.float32)
    sorted_values = np.ascontiguousarray(sorted_values, np.uint32)
    tile_offsets = np.ascontiguousarray(tile_offsets, np.uint32)
    img = np.zeros((height, width, 4), np.float32)
    visited, blended = C.c_uint64(), C.c_uint64()
    lib().orc_rasterize_f32(_p(projected), _p(sorted_values), _p(tile_offsets), tile, tile, width,
                            height, 1 if quantize else 0, _p(img), C.byref(visited), C.byref(blended))
    return img, int(visited.value), int(blended.value)


def render(cam, s, width, height, tile=16, precision="f64"):
    """Whole pipeline on the CPU; returns (image HxWx4, K, pairs visited, pairs blended)."""
    cam = np.ascontiguousarray(cam, np.float32)
    args = _splat_args(s)
    dt = np.float64 if precision == "f64" else np.float32
    img = np.zeros((height, width, 4), dt)
    visited, blended = C.c_uint64(), C.c_uint64()
    fn = lib().orc_render_f64 if precision == "f64" else lib().orc_render_f32
    k = fn(_p(cam), s.n, *[_p(a) for a in args], s.sh_degree, width, height, tile, tile, _p(img),
           C.byref(visited), C.byref(blended))
    assert k >= 0
    return img, int(k), int(visited.value), int(blended.value)


def sh_basis(x, y, z, deg=3, precision="f64"):
    out = np.zeros(16, np.float64 if precision == "f64" else np.float32)
    fn = lib().orc_sh_basis_f64 if precision == "f64" else lib().orc_sh_basis_f32
    fn(x, y, z, deg, _p(out))
    return out


def pipeline_f32(cam, s, width, height, tile=16):
    """Stage-by-stage f32 mirror: every intermediate buffer of a rebin."""
    tiles_x, tiles_y = -(-width // tile), -(-height // tile)
    inv = float(np.float32(1.0 / tile))
    proj, counts = preprocess(cam, s, tiles_x, tiles_y, inv, inv)
    offsets, k = scan(counts)
    keys, vals = emit(proj, offsets, tiles_x)
    skeys, svals = radix_sort(keys, vals, 8)
    toff = identify(skeys, tiles_x * tiles_y)
    img, visited, blended = rasterize(proj, svals, toff, tile, width, height)
    return dict(projected=proj, counts=counts, offsets=offsets, k=k, keys=keys, values=vals,
                sorted_keys=skeys, sorted_values=svals, tile_offsets=toff, image=img,
                tiles=(tiles_x, tiles_y), visited=visited, blended=blended)


def tile_overlay(base, tile_offsets, tile, width, height, debug_ref=256):
    """Heat-map target (H x W x 4 f32, f16-representable) over `base` (same shape)."""
    base = np.ascontiguousarray(base, np.float32)
    tile_offsets = np.ascontiguousarray(tile_offsets, np.uint32)
    out = np.zeros((height, width, 4), np.float32)
    lib().orc_tile_overlay_f32(_p(base), _p(tile_offsets), tile, tile, width, height, debug_ref, _p(out))
    return out


def composite(src, width, height, clear=None, framebuffer=None, bgra=False):
    """Premultiplied "over" of src (H x W x 4 f32) onto an unorm8 framebuffer."""
    src = np.ascontiguousarray(src, np.float32)
    fb = (np.zeros((height, width, 4), np.uint8) if framebuffer is None
          else np.ascontiguousarray(framebuffer, np.uint8).copy())
    cl = None if clear is None else np.ascontiguousarray(clear, np.float32)
    lib().orc_composite_f32(_p(src), width, height, 1 if bgra else 0, _p(cl), _p(fb))
    return fb
