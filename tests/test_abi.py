"""C-ABI checks that need no GPU: the library loads, every entry point that
include/gs3d.h declares is exported, and struct sizes match the contract."""
import ctypes as C
import re
from pathlib import Path

import pytest

from gs3d import _lib as L

REPO = Path(__file__).resolve().parent.parent


def declared_entry_points():
    text = (REPO / "include" / "gs3d.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|uint32_t)\s+(gs3d_\w+)\s*\(", text, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    lib = C.CDLL(str(L.LIB_PATH))
    names = declared_entry_points()
    assert len(names) >= 24
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) == set(L.exported_symbols())


def test_abi_version_and_labels():
    lib = L.lib()
    assert lib.gs3d_abi_version() == 1
    assert [lib.gs3d_stage_label(i).decode() for i in range(9)] == list(L.STAGES)
    assert lib.gs3d_stage_label(99) == b""


def test_null_arguments_rejected_without_device():
    lib = L.lib()
    assert lib.gs3d_create(0, None) == -1
    assert lib.gs3d_destroy(None) == -1
    assert lib.gs3d_sync(None) == -1
    assert lib.gs3d_rebin(None, None, 1, 1, 0) == -1
    assert lib.gs3d_set_debug_mode(None, 1, 256) == -1
    assert lib.gs3d_composite(None, 0, None, None, 0) == -1
    assert lib.gs3d_stage_composite(None, None, None, 16, 16, 1, 1, 0, 256, 0, None, None, None) == -1
    assert lib.gs3d_rebin_batch(None, None, 1, 0) == -1
    assert lib.gs3d_batch_read(None, 0, 7, None, 0, None) == -1
    assert lib.gs3d_learned_init(None, 0) == -1
    assert lib.gs3d_learned_train(None, 1, 1e-3, 1e-2, None, 0) == -1
    assert lib.gs3d_learned_validate(None, None, 1, 45.0, None, None) == -1
    assert lib.gs3d_load_ply(None, b"x.ply", 0, None) == -1
    assert lib.gs3d_ply_parse(None, 0, None, None, None, None, None, None, 0, None, None, 0) == -1
    assert lib.gs3d_last_error(None) == b"null context"


def test_create_without_gpu_reports_no_device():
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    h = C.c_void_p()
    assert L.lib().gs3d_create(0, C.byref(h)) == -2


def test_struct_sizes():
    text = (REPO / "include" / "gs3d.h").read_text()
    assert "float view[16];" in text and "uint32_t tile_max[2];" in text
    # 72 floats camera, 16 x 4-byte projected splat: checked against the oracle's static assert
    import oracle_lib

    oracle_lib.lib()  # built with _Static_assert(sizeof(projected) == 64)


def test_launch_geometry_without_device():
    """gs3d_launch_geometry (what WorkgroupManager reports) with no context: the default
    variants' launch shapes; an unknown pass is GS3D_E_UNKNOWN_NAME."""
    assert L.launch_geometry("preprocess-splats", [100_000]) == ((256, 1, 1), (391, 1, 1))
    assert L.launch_geometry("prefix-scan", [0]) == ((1024, 1, 1), (1, 1, 1))
    assert L.launch_geometry("radix-sort", [16_384]) == ((1024, 1, 1), (1, 1, 1))
    assert L.launch_geometry("radix-sort", [16_385]) == ((1024, 1, 1), (2, 1, 1))
    assert L.launch_geometry("rasterize-tiles", [1281, 720]) == ((256, 1, 1), (81 * 45, 1, 1))
    assert L.launch_geometry("identify-tile-ranges", [1280, 720]) == ((256, 1, 1), (15, 1, 1))
    assert L.launch_geometry("composite-splats", [1920, 1080])[1][0] == -(-1920 * 1080 // 1024)
    with pytest.raises(L.Gs3dError) as e:
        L.launch_geometry("rasterize", [64, 64])
    assert e.value.code == -6


def test_reference_pass_stages():
    """Every pass label of the reference's Profiler table maps to the fused stage(s)
    doing its work; an unknown label maps to none."""
    assert L.reference_pass_stages("prefix-scan-blocks") == ("prefix-scan",)
    assert L.reference_pass_stages("radix-scatter") == ("radix-sort-splats", "radix-sort")
    assert L.reference_pass_stages("tile-overlay") == ("composite-splats",)
    assert L.reference_pass_stages("grid") == ()
    table = L.reference_pass_table({s: float(i + 1) for i, s in enumerate(L.STAGES)})
    assert [lab for labels, _, _ in table for lab in labels] == list(L.REFERENCE_PASSES)
    assert [stage for _, stage, _ in table] == [
        "preprocess-splats", "prefix-scan", "emit-tile-refs", "radix-sort-splats+radix-sort",
        "identify-tile-ranges", "rasterize-tiles", "composite-splats"]
    assert table[3][2] == 2.0 + 5.0


def test_autotuner_candidates():
    """Each candidate keeps at least one wavefront per tile; list forms vary only at
    one pixel per lane (2 = picked per frame, 1 = quarter lists, 0 = block lists)."""
    from gs3d import autotune

    c = list(autotune.candidates(tiles=(8, 16), ppls=(1, 2, 4), radix_bits=(8,)))
    assert (8, 2, 8, 2) not in c and (8, 4, 8, 2) not in c
    assert [q for t, p, b, q in c if (t, p) == (8, 1)] == [2, 1, 0]
    assert [q for t, p, b, q in c if (t, p) == (16, 1)] == [2, 1, 0]
    assert [q for t, p, b, q in c if t == 16 and p > 1] == [2, 2]
    assert len(c) == 8
    assert autotune.Choice(16, 1, 8, 0.0, {}).quarter_lists == 2
