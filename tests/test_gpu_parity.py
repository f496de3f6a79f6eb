"""HIP path vs the CPU oracle, through the C ABI (libgs3d.so).

Integer stages (scan, emit, sort, identify) are bit-exact; preprocess is
bit-exact against the fp32 oracle (same operation order, no contraction);
the image is within 1/255 max-abs per channel of the fp32 oracle, and its
error against the fp64 render is reported.
"""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from gs3d import _lib as L
from gs3d import camera as cam_mod
from gs3d import scene as sc

pytestmark = pytest.mark.gpu

TOL_IMAGE = 1.0 / 255.0


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# ------------------------------------------------------------------ scan ----
@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 65_537, 1_000_000])
def test_scan_bit_exact(renderer, n):
    say(f"scan n={n}")
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 64, n, dtype=np.uint32)
    counts[rng.random(n) < 0.25] = 0
    got, tot = renderer.stage_scan(counts)
    exp, etot = O.scan(counts)
    assert np.array_equal(got, exp) and tot == etot


def test_scan_wraps(renderer):
    counts = np.full(5000, 0x00F00000, np.uint32)
    counts[17] = 0xFFFFFFFF
    got, tot = renderer.stage_scan(counts)
    exp, etot = O.scan(counts)
    assert np.array_equal(got, exp) and tot == etot


# ------------------------------------------------------------------ sort ----
def make_keys(n, seed, tiles=8160, dup=0.3):
    rng = np.random.default_rng(seed)
    tile = rng.integers(0, tiles, n, dtype=np.uint64)
    depth = rng.uniform(0.2, 30.0, n).astype(np.float32).view(np.uint32).astype(np.uint64)
    keys = (tile << np.uint64(32)) | depth
    if n:
        idx = rng.random(n) < dup
        keys[idx] = keys[rng.integers(0, n, idx.sum())]
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


@pytest.mark.parametrize("bits", [8, 4])
@pytest.mark.parametrize("n", [0, 1, 63, 4095, 4096, 4097, 100_003, 3_400_000])
def test_sort_bit_exact(renderer, n, bits):
    say(f"sort n={n} bits={bits}")
    keys, vals = make_keys(n, n + bits)
    gk, gv = renderer.stage_sort(keys, vals, bits)
    order = np.lexsort((np.arange(n), keys))
    assert np.array_equal(gk, keys[order])
    assert np.array_equal(gv, vals[order])


def test_sort_full_64bit_keys(renderer):
    rng = np.random.default_rng(77)
    keys = rng.integers(0, 2 ** 63, 200_000, dtype=np.uint64) * np.uint64(2) + rng.integers(
        0, 2, 200_000, dtype=np.uint64)
    keys[::7] = keys[3]
    vals = np.arange(keys.size, dtype=np.uint32)
    gk, gv = renderer.stage_sort(keys, vals, 8)
    ok, ov = O.radix_sort(keys, vals, 8)
    assert np.array_equal(gk, ok) and np.array_equal(gv, ov)


# ------------------------------------------------------ emit / identify ----
def test_emit_bit_exact(renderer):
    from test_oracle import random_projected

    for n, seed in ((1, 1), (255, 2), (256, 3), (257, 4), (20_000, 5)):
        say(f"emit n={n}")
        proj, counts = random_projected(n, 120, 68, seed=seed)
        u = proj.view(np.uint32)
        big = np.random.default_rng(seed).random(n) < 0.01  # a few very wide splats
        u[big, 12], u[big, 13], u[big, 14], u[big, 15] = 0, 0, 119, 67
        u[big, 11] = 1
        counts = np.where(u[:, 11] == 1, (u[:, 14] - u[:, 12] + 1) * (u[:, 15] - u[:, 13] + 1),
                          0).astype(np.uint32)
        offs, _ = O.scan(counts)
        gk, gv = renderer.stage_emit(proj, offs, 120)
        ek, ev = O.emit(proj, offs, 120)
        assert np.array_equal(gk, ek) and np.array_equal(gv, ev)


@pytest.mark.parametrize("case", ["empty", "one", "dense", "sparse", "last_tile", "gap_start"])
def test_identify_bit_exact(renderer, case):
    T = 8160
    rng = np.random.default_rng(len(case))
    tiles = {"empty": np.array([], np.int64), "one": np.array([4000]),
             "dense": rng.integers(0, T, 300_000), "sparse": rng.choice([3, 900, 5000], 3000),
             "last_tile": rng.choice([0, T - 1], 50), "gap_start": rng.choice([8000, 8001], 9)}[case]
    keys = np.sort((tiles.astype(np.uint64) << np.uint64(32)) | rng.integers(
        0, 1 << 31, tiles.size, dtype=np.uint64))
    got = renderer.stage_identify(keys, T)
    assert np.array_equal(got, O.identify(keys, T))


# ------------------------------------------------------------ preprocess ----
@pytest.mark.parametrize("deg", [0, 1, 2, 3])
def test_preprocess_bit_exact(renderer, deg):
    say(f"preprocess deg={deg}")
    raw = sc.clustered(20 + deg, 40_000, 16, sh_degree=deg) if deg else sc.uniform_cube(20, 40_000)
    s = sc.activate(raw)
    renderer.set_splat_data(s)
    W, H = 1280, 720
    cu = cam_mod.view_camera(sc.orbit_eye(6.0, 30.0, 15.0) if deg else (0.0, 0.0, 3.0), W, H)
    tx, ty = 80, 45
    inv = float(np.float32(1 / 16))
    gp, gc = renderer.stage_preprocess(cu, tx, ty, inv, inv)
    ep, ec = O.preprocess(cu, s, tx, ty, inv, inv)
    assert np.array_equal(gc, ec)
    valid = ep.view(np.uint32)[:, 11] == 1
    assert np.array_equal(gp.view(np.uint32)[:, 11], ep.view(np.uint32)[:, 11])
    diff = gp.view(np.uint32)[valid] != ep.view(np.uint32)[valid]
    assert not diff.any(), f"{diff.any(axis=1).sum()} splats differ; fields {np.nonzero(diff.any(axis=0))}"


# --------------------------------------------------------------- raster ----
@pytest.mark.parametrize("tile", [16, 8, 32])
def test_rasterize_stage(renderer, tile):
    say(f"rasterize tile={tile}")
    raw = sc.uniform_cube(30, 60_000, log_scale_mean=-4.0)
    s = sc.activate(raw)
    W, H = 640, 360
    cu = cam_mod.view_camera((0.0, 0.0, 3.0), W, H)
    ref = O.pipeline_f32(cu, s, W, H, tile)
    img = renderer.stage_rasterize(ref["projected"], ref["sorted_values"], ref["tile_offsets"],
                                   tile, W, H)
    err = np.abs(img - ref["image"]).max()
    say(f"  max abs vs oracle f32: {err:.3g}")
    assert err <= TOL_IMAGE


# ------------------------------------------------------------- pipeline ----
def rebin_and_compare(renderer, s, eye, W, H, tile=16, fov=math.pi / 3):
    renderer.set_splat_data(s)
    renderer.set_tile_size(tile, tile)
    cu = cam_mod.view_camera(eye, W, H, fov)
    renderer.rebin(cu, W, H, L.REBIN_KEEP_UNSORTED)
    ref = O.pipeline_f32(cu, s, W, H, tile)
    k = renderer.ref_count()
    assert k == ref["k"]
    gp = renderer.read(L.BUF_PROJECTED_SPLATS, np.uint32).reshape(-1, 16)
    valid = ref["projected"].view(np.uint32)[:, 11] == 1
    assert np.array_equal(gp[valid], ref["projected"].view(np.uint32)[valid])
    assert np.array_equal(renderer.read(L.BUF_SPLAT_REF_COUNTS, np.uint32), ref["counts"])
    assert np.array_equal(renderer.read(L.BUF_SPLAT_REF_OFFSETS, np.uint32), ref["offsets"])
    assert int(renderer.read(L.BUF_REF_COUNTER, np.uint32)[0]) == ref["k"]
    assert np.array_equal(renderer.read(L.BUF_UNSORTED_KEYS, np.uint64), ref["keys"])
    assert np.array_equal(renderer.read(L.BUF_UNSORTED_VALUES, np.uint32), ref["values"])
    assert np.array_equal(renderer.read(L.BUF_SORT_KEYS, np.uint64), ref["sorted_keys"])
    assert np.array_equal(renderer.read(L.BUF_SORT_VALUES, np.uint32), ref["sorted_values"])
    assert np.array_equal(renderer.read(L.BUF_TILE_OFFSETS, np.uint32), ref["tile_offsets"])
    img = renderer.read(L.BUF_IMAGE_RGBA32F, np.float32).reshape(H, W, 4)
    err = float(np.abs(img - ref["image"]).max())
    return img, ref, err


def test_pipeline_uniform_100k_sh0(renderer):
    """The 100k SH0 workload at 1280x720: every buffer vs the oracle."""
    say("pipeline uniform-100k-sh0-720p")
    wl = sc.workload(0)
    s = sc.activate(wl.raw)
    (eye, W, H) = wl.views[0]
    img, ref, err = rebin_and_compare(renderer, s, eye, W, H)
    f64, _, _, _ = O.render(cam_mod.view_camera(eye, W, H), s, W, H, 16, "f64")
    err64 = float(np.abs(img - f64).max())
    say(f"  K={ref['k']} image max abs: vs oracle f32 {err:.3g}, vs fp64 render {err64:.3g}")
    assert err <= TOL_IMAGE
    assert err64 <= 2 * TOL_IMAGE


@pytest.mark.parametrize("tile", [8, 32])
def test_pipeline_tile_sizes(renderer, tile):
    say(f"pipeline tile={tile}")
    s = sc.activate(sc.clustered(41, 50_000, 16, sh_degree=3))
    _, _, err = rebin_and_compare(renderer, s, sc.orbit_eye(7.0, 45.0, 20.0), 960, 540, tile)
    assert err <= TOL_IMAGE


def test_pipeline_capacity_regrow(renderer):
    """Splats that cover many tiles make K exceed the initial 8N key budget:
    the frame is re-run with grown buffers and still matches the oracle."""
    say("pipeline regrow")
    s = sc.activate(sc.uniform_cube(42, 2_000, log_scale_mean=-1.5, log_scale_std=0.3))
    img, ref, err = rebin_and_compare(renderer, s, (0.0, 0.0, 3.0), 1280, 720)
    assert ref["k"] > 8 * 2_000
    assert err <= TOL_IMAGE


def test_pipeline_empty_and_all_culled(renderer):
    s = sc.activate(sc.uniform_cube(43, 100))
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    cu = cam_mod.view_camera((0.0, 0.0, -30.0), 64, 48, target=(0.0, 0.0, -60.0))  # facing away
    renderer.rebin(cu, 64, 48)
    assert renderer.ref_count() == 0
    assert not renderer.read(L.BUF_TILE_OFFSETS, np.uint32).any()
    assert not renderer.read(L.BUF_IMAGE_RGBA32F, np.float32).any()
    assert renderer.read_tile_counts().sum() == 0


def test_pipeline_no_splats(renderer):
    """An empty scene (SH3, N = 0): K = 0, zero tile ranges and a cleared image, for a
    single view and for a batch; a batch view's projected splats read back empty."""
    s = sc.activate(sc.uniform_cube(45, 0, sh_degree=3))
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    cam = cam_mod.view_camera((0.0, 0.0, 3.0), 96, 64)
    renderer.rebin(cam, 96, 64)
    assert renderer.ref_count() == 0
    assert not renderer.read(L.BUF_TILE_OFFSETS, np.uint32).any()
    assert not renderer.read(L.BUF_IMAGE_RGBA32F, np.float32).any()
    renderer.rebin_batch([(cam, 96, 64), (cam_mod.view_camera((0.0, 1.0, 3.0), 40, 30), 40, 30)])
    assert [int(k) for k in renderer.batch_ref_counts()] == [0, 0]
    for i in range(2):
        assert not renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16).any()
        assert not renderer.batch_read(i, L.BUF_TILE_OFFSETS, np.uint32).any()
        assert renderer.batch_read(i, L.BUF_PROJECTED_SPLATS, np.uint32).size == 0


@pytest.mark.parametrize("W,H,tile", [(1, 1, 16), (16, 16, 16), (17, 9, 8), (200, 100, 32), (33, 700, 16),
                                      (4096, 40, 16), (4200, 40, 16), (4096, 4096, 16), (4096, 4112, 16)])
def test_pipeline_tiny_and_ragged_viewports(renderer, W, H, tile):
    """Viewports of one tile (no tile-digit pass at all), a few tiles (one pass, so the
    result ends in the second buffer), partial tiles at the right and bottom edges, a
    tall, narrow one, wide ones at and past 256 tiles per row (tile rectangles
    packed 8 bits per coordinate up to 256, gathered at full width past it), and
    65,536 tiles (the last frame whose tile ids sort as 16 bits, top id 65,535) and
    65,792 (32-bit tile ids): every buffer bit-exact against the oracle."""
    say(f"pipeline viewport {W}x{H} tile={tile}")
    s = sc.activate(sc.uniform_cube(44, 3_000, log_scale_mean=-3.0))
    img, ref, err = rebin_and_compare(renderer, s, (0.0, 0.0, 2.5), W, H, tile)
    say(f"  K={ref['k']} passes={renderer.sort_passes()} image max abs {err:.3g}")
    assert ref["k"] > 0
    assert err <= TOL_IMAGE


def check_keys_inside_rects(keys, vals, projected, tiles_x):
    """Every sorted key's tile id lies in its splat's tile rectangle (independent of the
    oracle's emit), and every key belongs to a visible splat."""
    u = np.ascontiguousarray(projected.view(np.uint32).reshape(-1, 16)[:, 11:16])
    for a in range(0, keys.size, 1 << 24):
        tile = (keys[a:a + (1 << 24)] >> np.uint64(32)).astype(np.int64)
        tx, ty = tile % tiles_x, tile // tiles_x
        r = u[vals[a:a + (1 << 24)]]
        assert np.all(r[:, 0] == 1)
        assert np.all((tx >= r[:, 1]) & (tx <= r[:, 3]) & (ty >= r[:, 2]) & (ty <= r[:, 4]))


def full_size_every_buffer(renderer, index, tiles_x, tiles_y):
    """A whole workload frame with every buffer compared to the oracle: projected splats
    (visible rows), counts, id-order offsets, K, the emit-order and sorted keys and values,
    the tile ranges bit-exact; the image within 1/255 of the fp32 oracle, its error against
    the fp64 render printed; then properties that do not rest on the oracle, and
    determinism of a second rebin."""
    O.set_threads(16)
    wl = sc.workload(index)
    s = sc.activate(wl.raw)
    eye, W, H = wl.views[0]
    img, ref, err = rebin_and_compare(renderer, s, eye, W, H, 16, wl.fov_y)
    cu = cam_mod.view_camera(eye, W, H, wl.fov_y)
    f64, k64, _, _ = O.render(cu, s, W, H, 16, "f64")
    err64 = float(np.abs(img - f64).max())
    say(f"  K={ref['k']} image max abs: vs oracle f32 {err:.3g}, vs fp64 render {err64:.3g} "
        f"(fp64 pipeline K={k64})")
    assert err <= TOL_IMAGE
    assert err64 <= 2 * TOL_IMAGE
    keys, vals = ref["sorted_keys"], ref["sorted_values"]
    assert ref["k"] == int(ref["counts"].sum(dtype=np.uint64)) == keys.size
    check_keys_inside_rects(keys, vals, ref["projected"], tiles_x)
    same = keys[1:] == keys[:-1]
    assert np.all(keys[1:] >= keys[:-1]) and np.all(vals[1:][same] > vals[:-1][same])
    del same
    assert np.array_equal(np.bincount(vals, minlength=s.n), ref["counts"])
    img16 = renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16)
    renderer.rebin(cu, W, H)
    assert renderer.ref_count() == ref["k"]
    assert np.array_equal(renderer.read(L.BUF_SORT_VALUES, np.uint32), vals)
    assert np.array_equal(renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16), img16)
    return ref


@pytest.mark.timeout(600)
def test_full_size_1m_sh3_every_buffer(renderer):
    """The headline workload, 1M SH3 splats at 1920x1080: every buffer vs the oracle."""
    say("full-size clusters-1M-sh3-1080p, every buffer")
    ref = full_size_every_buffer(renderer, 1, 120, 68)
    assert ref["k"] > 1_000_000


@pytest.mark.timeout(900)
def test_full_size_6m_sh3_2160p_every_buffer(renderer):
    """The largest single-GPU workload: 6M SH3 splats at 3840x2160, ~119M keys. The key
    buffers start at 8N = 48M entries, so the frame regrows them and runs again; then
    every buffer vs the oracle, as at 1M."""
    say("full-size clusters-6M-sh3-2160p, every buffer")
    ref = full_size_every_buffer(renderer, 2, 240, 135)
    assert ref["k"] > 8 * 6_000_000  # the regrow path ran


def test_errors_through_abi(renderer):
    import gs3d.renderer as R

    with pytest.raises(L.Gs3dError) as e:
        renderer.set_tile_size(12, 12)
    assert e.value.code == -6
    fresh = R.GaussianSplatRenderer(0)
    with pytest.raises(L.Gs3dError) as e:
        fresh.rebin(np.zeros(72, np.float32), 64, 64)
    assert e.value.code == -5
    with pytest.raises(L.Gs3dError) as e:
        fresh.read(L.BUF_SORT_KEYS, np.uint64)
    assert e.value.code == -5
    assert fresh.read_tile_counts() is None  # readTileCounts: null on failure
    fresh.close()
    lim = renderer.device_limits()
    assert lim["maxTotalThreads"] >= 1024


@pytest.mark.timeout(300)
def test_raster_images_identical_across_blocks(renderer):
    """A splat that can reach a pixel lies in the list of every tile that holds the pixel
    (its tile rectangle bounds its radius circle), lists are ordered by (depth, id), and
    every variant does the same per-pixel arithmetic. The variants cull with different
    rectangles: 8 x 8 blocks and their 4 x 4 quarters at one pixel per lane (or the blocks
    alone, QUARTER_LISTS 0, or either chosen per frame, 2), 8 x 16 and 16 x 16 blocks at 2 and 4. A cull that dropped a splat
    some pixel blends would change that pixel in some variants and not others, so the f16
    images of the 1M SH3 1080p frame at tiles 8, 16, 32, 1, 2, 4 pixels per lane and both list
    forms must be bit-identical, at three camera positions."""
    wl = sc.workload(1)
    s = sc.activate(wl.raw)
    renderer.set_splat_data(s)
    _, W, H = wl.views[0]
    variants = [(16, 1, 1), (8, 1, 1), (32, 1, 1), (16, 1, 0), (8, 1, 0), (32, 1, 0), (16, 1, 2), (8, 1, 2),
                (16, 2, 1), (16, 4, 1), (32, 2, 1), (32, 4, 1)]
    try:
        for az in (0.0, 135.0, 270.0):
            eye = wl.views[0][0] if az == 0.0 else sc.orbit_eye(wl.orbit_radius, az, 20.0)
            cu = cam_mod.view_camera(eye, W, H, wl.fov_y)
            base = None
            for tile, ppl, quarters in variants:
                renderer.set_tile_size(tile, tile)
                renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", ppl)
                renderer.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", quarters)
                renderer.rebin(cu, W, H)
                img = renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16)
                if base is None:
                    base = img
                    assert np.count_nonzero(img) > img.size // 4  # the frame is not empty
                else:
                    assert np.array_equal(img, base), f"azimuth {az}: tile {tile}, {ppl} px/lane, quarters {quarters}"
            say(f"  azimuth {az}: {len(variants)} variants bit-identical")
    finally:
        renderer.set_tile_size(16, 16)
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", 1)
        renderer.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", 2)


@pytest.mark.parametrize("tile,ppl", [(16, 2), (16, 4), (32, 2), (32, 4)])
def test_pipeline_raster_variants(renderer, tile, ppl):
    say(f"pipeline tile={tile} pixels_per_lane={ppl}")
    s = sc.activate(sc.clustered(44, 40_000, 16, sh_degree=3))
    renderer.set_splat_data(s)
    renderer.set_tile_size(tile, tile)
    renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", ppl)
    try:
        _, _, err = rebin_and_compare(renderer, s, sc.orbit_eye(7.0, 200.0, 35.0), 800, 600, tile)
    finally:
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", 1)
    assert err <= TOL_IMAGE


def test_pipeline_radix_4bit(renderer):
    say("pipeline radix 4-bit digits")
    s = sc.activate(sc.clustered(45, 30_000, 16, sh_degree=3))
    renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 4)
    try:
        _, _, err = rebin_and_compare(renderer, s, sc.orbit_eye(6.0, 90.0, -20.0), 640, 480)
        assert renderer.sort_passes() == list(range(0, 32 + 11, 4))  # 1200 tiles -> 11 tile bits
    finally:
        renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 8)
    assert err <= TOL_IMAGE


def test_kernel_variant_errors(renderer):
    renderer.set_tile_size(8, 8)
    with pytest.raises(L.Gs3dError):
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", 2)  # 8x8 is one wavefront
    with pytest.raises(L.Gs3dError):
        renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 6)
    with pytest.raises(L.Gs3dError):
        renderer.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", 3)
    assert renderer.get_kernel_variant("rasterize-tiles", "QUARTER_LISTS") == 2
    with pytest.raises(L.Gs3dError):
        renderer.set_kernel_variant("nope", "X", 1)
    renderer.set_tile_size(16, 16)
    assert renderer.get_kernel_variant("rasterize-tiles", "TILE_X") == 16


# ------------------------------------------------- heat map + composite ----
def random_target_f16(rng, H, W):
    a = rng.uniform(0, 1, (H, W, 1))
    rgb = rng.uniform(0, 1, (H, W, 3)) * a
    return np.concatenate([rgb, a], axis=2).astype(np.float16)


@pytest.mark.parametrize("debug", [0, 1])
@pytest.mark.parametrize("fmt", [L.FORMAT_RGBA8UNORM, L.FORMAT_BGRA8UNORM])
@pytest.mark.parametrize("clear", [None, (0.12, 0.12, 0.13, 1.0)])
@pytest.mark.parametrize("W,H,tile", [(333, 77, 16), (1030, 9, 8)])
def test_composite_stage_bit_exact(renderer, debug, fmt, clear, W, H, tile):
    say(f"composite stage {W}x{H} tile={tile} debug={debug} fmt={fmt} clear={clear is not None}")
    rng = np.random.default_rng(W + H + debug)
    tgt = random_target_f16(rng, H, W)
    T = (-(-W // tile)) * (-(-H // tile))
    to = np.concatenate([[0], np.cumsum(rng.integers(0, 400, T))]).astype(np.uint32)
    to[-4:-1] = 0
    to[3] = to[4] + 5  # one tile whose end is below its start (the underflow guard)
    fb = rng.integers(0, 256, (H, W, 4), dtype=np.uint8)
    got, dbg = renderer.stage_composite(tgt.view(np.uint16), to if debug else None, tile, W, H,
                                        debug, 200, clear, fb, fmt)
    src = tgt.astype(np.float32)
    if debug:
        src = O.tile_overlay(src, to, tile, W, H, 200)
        assert np.array_equal(dbg, src.astype(np.float16).view(np.uint16))
    bgra = fmt == L.FORMAT_BGRA8UNORM
    want = O.composite(src, W, H, clear=clear, framebuffer=fb, bgra=bgra)
    assert np.array_equal(got, want)


def test_composite_after_rebin(renderer):
    """render()'s display step on a rebinned frame: heat map + composite vs the oracle on the
    frame's own target and tile ranges, then a plain composite onto a caller framebuffer."""
    say("composite after rebin (uniform-100k-sh0-720p)")
    wl = sc.workload(0)
    s = sc.activate(wl.raw)
    (eye, W, H) = wl.views[0]
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    renderer.rebin(cam_mod.view_camera(eye, W, H), W, H)
    tgt = renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16).reshape(H, W, 4)
    to = renderer.read(L.BUF_TILE_OFFSETS, np.uint32)
    src = tgt.view(np.float16).astype(np.float32)
    try:
        renderer.set_debug_mode(1, 64)
        assert renderer.get_debug_mode() == (1, 64)
        got = renderer.composite(W, H, clear=(0.12, 0.12, 0.13, 1.0))
        heat = O.tile_overlay(src, to, 16, W, H, 64)
        dbg = renderer.read(L.BUF_DEBUG_TARGET_RGBA16F, np.uint16).reshape(H, W, 4)
        assert np.array_equal(dbg, heat.astype(np.float16).view(np.uint16))
        assert np.array_equal(got, O.composite(heat, W, H, clear=(0.12, 0.12, 0.13, 1.0)))
        assert np.array_equal(renderer.read(L.BUF_FRAMEBUFFER, np.uint8).reshape(H, W, 4), got)
    finally:
        renderer.set_debug_mode(0, 256)
    fb = np.random.default_rng(9).integers(0, 256, (H, W, 4), dtype=np.uint8)
    got = renderer.composite(W, H, clear=None, framebuffer=fb, fmt=L.FORMAT_BGRA8UNORM)
    assert np.array_equal(got, O.composite(src, W, H, framebuffer=fb, bgra=True))
    # resident framebuffer: blending the same frame onto the previous result again
    again = renderer.composite(W, H, clear=None)
    assert np.array_equal(again, O.composite(src, W, H, framebuffer=got, bgra=False))


@pytest.mark.parametrize("key_tile", [8192, 16384])
@pytest.mark.parametrize("n", [1, 8191, 8193, 100_003, 3_400_000])
def test_sort_key_tile_shapes_bit_exact(renderer, n, key_tile):
    """Both tile shapes of the 64-bit key pass give the oracle's stable order."""
    say(f"sort key_tile={key_tile} n={n}")
    keys, vals = make_keys(n, n + key_tile)
    renderer.set_kernel_variant("radix-sort", "KEY_TILE", key_tile)
    try:
        ko, vo = renderer.stage_sort(keys, vals, 8)
    finally:
        renderer.set_kernel_variant("radix-sort", "KEY_TILE", 0)
    order = np.lexsort((np.arange(n), keys))
    assert np.array_equal(ko, keys[order]) and np.array_equal(vo, vals[order])


@pytest.mark.parametrize("depth_tile", [4096, 8192])
def test_depth_key_tile_shapes_bit_exact(renderer, depth_tile):
    """Both tile shapes of the 32-bit depth passes give the same frame as the oracle."""
    say(f"pipeline depth key tile {depth_tile}")
    renderer.set_kernel_variant("radix-sort-splats", "KEY_TILE", depth_tile)
    try:
        s = sc.activate(sc.clustered(43, 60_000, 16, sh_degree=1))
        _, _, err = rebin_and_compare(renderer, s, sc.orbit_eye(7.0, 10.0, 25.0), 640, 480)
        assert err <= TOL_IMAGE
    finally:
        renderer.set_kernel_variant("radix-sort-splats", "KEY_TILE", 0)
