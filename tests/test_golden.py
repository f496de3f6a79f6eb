"""Golden fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py
from the CPU oracle): small rebins with every intermediate buffer. The oracle
as built now must reproduce them bit for bit (CPU), and so must the HIP path
through the C ABI (GPU; the image within 1/255 of the fixture's)."""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from gs3d import _lib as L
from gs3d import scene as S

GOLDEN = sorted((Path(__file__).resolve().parent / "golden").glob("*.npz"))


def load(path):
    f = np.load(path)
    deg = int(f["sh_degree"])
    s = S.Splats(f["positions"], f["colors"], f["scales"], f["rotations"], f["f_dc"] if deg else None,
                 f["sh_rest"] if deg else None, deg, None)
    W, H, tile = (int(v) for v in f["viewport"])
    return f, s, W, H, tile


def test_fixtures_present():
    assert len(GOLDEN) >= 2


@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_oracle_reproduces_golden(path):
    f, s, W, H, tile = load(path)
    ref = O.pipeline_f32(f["camera"], s, W, H, tile)
    for k in ("projected", "counts", "offsets", "keys", "values", "sorted_keys", "sorted_values", "tile_offsets"):
        assert np.array_equal(np.asarray(ref[k]).view(np.uint8), f[k].view(np.uint8)), k
    assert np.array_equal(O.f16_round(ref["image"]), f["image_f16"])


@pytest.mark.gpu
@pytest.mark.parametrize("path", GOLDEN, ids=[p.stem for p in GOLDEN])
def test_hip_path_reproduces_golden(renderer, path):
    f, s, W, H, tile = load(path)
    renderer.set_splat_data(s)
    renderer.set_tile_size(tile, tile)
    renderer.rebin(f["camera"], W, H, L.REBIN_KEEP_UNSORTED)
    valid = f["projected"].view(np.uint32)[:, 11] == 1
    gp = renderer.read(L.BUF_PROJECTED_SPLATS, np.uint32).reshape(-1, 16)
    assert np.array_equal(gp[valid], f["projected"].view(np.uint32)[valid])
    assert np.array_equal(renderer.read(L.BUF_SPLAT_REF_COUNTS, np.uint32), f["counts"])
    assert np.array_equal(renderer.read(L.BUF_SPLAT_REF_OFFSETS, np.uint32), f["offsets"])
    assert np.array_equal(renderer.read(L.BUF_UNSORTED_KEYS, np.uint64), f["keys"])
    assert np.array_equal(renderer.read(L.BUF_UNSORTED_VALUES, np.uint32), f["values"])
    assert np.array_equal(renderer.read(L.BUF_SORT_KEYS, np.uint64), f["sorted_keys"])
    assert np.array_equal(renderer.read(L.BUF_SORT_VALUES, np.uint32), f["sorted_values"])
    assert np.array_equal(renderer.read(L.BUF_TILE_OFFSETS, np.uint32), f["tile_offsets"])
    img = renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16).view(np.float16).astype(np.float32)
    want = f["image_f16"].view(np.float16).astype(np.float32).reshape(-1)
    assert float(np.abs(img - want).max()) <= 1.0 / 255.0
    renderer.set_tile_size(16, 16)
