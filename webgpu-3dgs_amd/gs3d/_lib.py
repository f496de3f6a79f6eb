"""ctypes binding of libgs3d.so (the C ABI declared in include/gs3d.h).

Loading fails loudly when the built library is missing: there is no CPU
fallback anywhere in this package.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libgs3d.so"

GS3D_OK = 0
ERRORS = {
    -1: "GS3D_E_INVALID_ARG", -2: "GS3D_E_NO_DEVICE", -3: "GS3D_E_HIP", -4: "GS3D_E_OOM",
    -5: "GS3D_E_STATE", -6: "GS3D_E_UNKNOWN_NAME", -7: "GS3D_E_TOO_SMALL", -8: "GS3D_E_FORMAT",
    -9: "GS3D_E_IO",
}

BUF_PROJECTED_SPLATS = 0
BUF_SPLAT_REF_COUNTS = 1
BUF_SPLAT_REF_OFFSETS = 2
BUF_REF_COUNTER = 3
BUF_SORT_KEYS = 4
BUF_SORT_VALUES = 5
BUF_TILE_OFFSETS = 6
BUF_IMAGE_RGBA16F = 7
BUF_IMAGE_RGBA32F = 8
BUF_UNSORTED_KEYS = 9
BUF_UNSORTED_VALUES = 10
BUF_DEBUG_TARGET_RGBA16F = 11
BUF_FRAMEBUFFER = 12

FORMAT_RGBA8UNORM = 0
FORMAT_BGRA8UNORM = 1

PLY_ASCII, PLY_BINARY_LE, PLY_BINARY_BE = 0, 1, 2

BATCH_EAGER = 1
BATCH_LEARNED = 2
LEARNED_PARAMS = 5123
LEARNED_DIRS = 256
MAX_BATCH_VIEWS = 1024
# gs3d_view: 72-float camera uniform + width + height
VIEW_DTYPE = [("camera", "<f4", (72,)), ("width", "<u4"), ("height", "<u4")]

REBIN_PROFILE = 1
REBIN_KEEP_UNSORTED = 2

STAGES = ("preprocess-splats", "radix-sort-splats", "prefix-scan", "emit-tile-refs", "radix-sort",
          "identify-tile-ranges", "rasterize-tiles", "rebin-total", "composite-splats")
# the stages of one rebin (without the whole-rebin bracket and the display pass)
REBIN_STAGES = STAGES[:7]

# The pass labels of the reference's Profiler table, in its execution order
# (GaussianSplatRenderer.ts:1017-1260, Profiler.ts:135-228).
REFERENCE_PASSES = ("preprocess-splats", "prefix-scan-local", "prefix-scan-blocks", "prefix-scan-add",
                    "emit-tile-refs", "radix-histogram", "scan-local", "scan-add", "radix-scatter",
                    "identify-tile-ranges", "rasterize-tiles", "tile-overlay", "composite-splats")


def reference_pass_stages(label: str) -> tuple[str, ...]:
    """The stages (STAGES labels) doing the work of a reference pass
    (gs3d_reference_pass_stages)."""
    mask = lib().gs3d_reference_pass_stages(label.encode())
    return tuple(s for i, s in enumerate(STAGES) if mask >> i & 1)


def reference_pass_table(stage_us: dict) -> list[tuple[list[str], str, float]]:
    """The reference's pass labels grouped by the stage(s) doing their work, in
    execution order: [(reference labels, "stage[+stage]", microseconds)]."""
    rows: list[tuple[list[str], str, float]] = []
    for label in REFERENCE_PASSES:
        stages = reference_pass_stages(label)
        key = "+".join(stages)
        if rows and rows[-1][1] == key:
            rows[-1][0].append(label)
        else:
            rows.append(([label], key, sum(stage_us.get(s, 0.0) for s in stages)))
    return rows

P = C.c_void_p
U32 = C.c_uint32
F32 = C.c_float
I32 = C.c_int
SZ = C.c_size_t

_PROTOS = {
    "gs3d_abi_version": (I32, []),
    "gs3d_create": (I32, [I32, C.POINTER(P)]),
    "gs3d_destroy": (I32, [P]),
    "gs3d_last_error": (C.c_char_p, [P]),
    "gs3d_sync": (I32, [P]),
    "gs3d_device_limits": (I32, [P, P]),
    "gs3d_set_splats": (I32, [P, P, P, P, P, P, P, I32, U32]),
    "gs3d_splat_count": (U32, [P]),
    "gs3d_set_tile_size": (I32, [P, U32, U32, P]),
    "gs3d_get_tile_size": (I32, [P, P]),
    "gs3d_get_tile_count": (I32, [P, P]),
    "gs3d_set_kernel_variant": (I32, [P, C.c_char_p, C.c_char_p, I32]),
    "gs3d_get_kernel_variant": (I32, [P, C.c_char_p, C.c_char_p, C.POINTER(I32)]),
    "gs3d_rebin": (I32, [P, P, U32, U32, U32]),
    "gs3d_ref_count": (I32, [P, C.POINTER(C.c_uint64)]),
    "gs3d_read": (I32, [P, I32, P, SZ, C.POINTER(SZ)]),
    "gs3d_read_tile_counts": (I32, [P, P, SZ, C.POINTER(SZ)]),
    "gs3d_stage_times_us": (I32, [P, P, I32]),
    "gs3d_stage_label": (C.c_char_p, [I32]),
    "gs3d_profiler_reset": (I32, [P]),
    "gs3d_set_profile_mask": (I32, [P, U32]),
    "gs3d_stage_times_avg_us": (I32, [P, P, I32, C.POINTER(C.c_uint64)]),
    "gs3d_sort_passes": (I32, [P, C.POINTER(U32), P]),
    "gs3d_buffer_create": (I32, [P, C.c_char_p, SZ]),
    "gs3d_buffer_resize": (I32, [P, C.c_char_p, SZ]),
    "gs3d_buffer_size": (I32, [P, C.c_char_p, C.POINTER(SZ)]),
    "gs3d_buffer_write": (I32, [P, C.c_char_p, SZ, P, SZ, C.POINTER(SZ)]),
    "gs3d_buffer_read": (I32, [P, C.c_char_p, SZ, P, SZ]),
    "gs3d_buffer_destroy": (I32, [P, C.c_char_p]),
    "gs3d_buffer_list": (I32, [P, P, P, SZ, C.POINTER(SZ)]),
    "gs3d_stage_preprocess": (I32, [P, P, U32, U32, F32, F32, P, P]),
    "gs3d_stage_scan": (I32, [P, P, U32, P, C.POINTER(U32)]),
    "gs3d_stage_emit": (I32, [P, P, P, U32, U32, P, P]),
    "gs3d_stage_sort": (I32, [P, P, P, U32, U32, P, P]),
    "gs3d_sort_pairs_noalloc": (I32, [P, P, P, U32, U32, P, P]),
    "gs3d_launch_geometry": (I32, [P, C.c_char_p, P, P, P]),
    "gs3d_reference_pass_stages": (U32, [C.c_char_p]),
    "gs3d_stage_identify": (I32, [P, P, U32, U32, P]),
    "gs3d_stage_rasterize": (I32, [P, P, U32, P, U32, P, U32, U32, U32, U32, P]),
    "gs3d_set_debug_mode": (I32, [P, U32, U32]),
    "gs3d_get_debug_mode": (I32, [P, C.POINTER(U32), C.POINTER(U32)]),
    "gs3d_composite": (I32, [P, I32, P, P, SZ]),
    "gs3d_stage_composite": (I32, [P, P, P, U32, U32, U32, U32, U32, U32, I32, P, P, P]),
    "gs3d_rebin_batch": (I32, [P, P, U32, U32]),
    "gs3d_batch_ref_counts": (I32, [P, P, U32]),
    "gs3d_batch_read": (I32, [P, U32, I32, P, SZ, C.POINTER(SZ)]),
    "gs3d_learned_init": (I32, [P, C.c_uint64]),
    "gs3d_learned_train": (I32, [P, U32, F32, F32, P, U32]),
    "gs3d_learned_grad": (I32, [P, U32, P, P, C.POINTER(F32)]),
    "gs3d_learned_get": (I32, [P, P, P, P]),
    "gs3d_learned_set": (I32, [P, P, P]),
    "gs3d_learned_get_state": (I32, [P, P]),
    "gs3d_learned_set_state": (I32, [P, P]),
    "gs3d_learned_dirs": (I32, [P]),
    "gs3d_learned_colors": (I32, [P, P, P]),
    "gs3d_learned_validate": (I32, [P, P, U32, F32, C.POINTER(F32), C.POINTER(I32)]),
    "gs3d_learned_enable": (I32, [P, I32]),
    "gs3d_ply_probe": (I32, [P, SZ, P, C.c_char_p, SZ]),
    "gs3d_ply_parse": (I32, [P, SZ, P, P, P, P, P, P, I32, P, C.c_char_p, SZ]),
    "gs3d_load_ply": (I32, [P, C.c_char_p, I32, P]),
}


class PlyInfo(C.Structure):
    """gs3d_ply_info"""
    _fields_ = [("splat_count", C.c_uint64), ("f_rest_count", C.c_uint32), ("sh_degree", C.c_int32),
                ("format", C.c_int32), ("reserved", C.c_uint32), ("header_bytes", C.c_uint64),
                ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3)]

_lib = None


def lib() -> C.CDLL:
    """The loaded library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"libgs3d.so not built at {LIB_PATH}; run __graft_entry__.build()")
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _PROTOS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def exported_symbols() -> list[str]:
    return list(_PROTOS)


class Gs3dError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def ptr(a) -> int | None:
    """Address of a C-contiguous numpy array (None for None)."""
    if a is None:
        return None
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def launch_geometry(kernel: str, problem, handle=None) -> tuple[tuple[int, int, int], tuple[int, int, int]]:
    """(threads per workgroup, workgroups) of one launch of `kernel` over `problem`
    (gs3d_launch_geometry); handle None: the default variants."""
    import numpy as np

    p = np.ones(3, np.uint64)
    p[:len(problem)] = problem
    wg = np.zeros(3, np.uint32)
    grid = np.zeros(3, np.uint64)
    rc = lib().gs3d_launch_geometry(handle, kernel.encode(), ptr(p), ptr(wg), ptr(grid))
    if rc != 0:
        raise Gs3dError(rc, f"no pass named {kernel!r}")
    return tuple(int(x) for x in wg), tuple(int(x) for x in grid)


class LearnedState(C.Structure):
    """gs3d_learned_state: pointers to the six state arrays and the Adam step."""
    _fields_ = [("params", P), ("params_m", P), ("params_v", P), ("latents", P), ("latents_m", P),
                ("latents_v", P), ("step", U32)]
