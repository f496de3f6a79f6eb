"""Batched views (gs3d_rebin_batch): every view of a batch must equal a
single-view rebin of that camera and viewport bit for bit (the single-view
path is itself checked against the oracle in test_gpu_parity.py), whether the
batch runs from a captured graph (replayed with new cameras) or eagerly.
"""
import math
import sys

import numpy as np
import pytest

import oracle_lib as O
from gs3d import _lib as L
from gs3d import camera as cam_mod
from gs3d import scene as S

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def views_for(seed, count, shapes=((320, 180), (640, 360), (200, 120), (480, 272))):
    vs = S.orbit_views(seed, count)
    out = []
    for i, (eye, _, _) in enumerate(vs):
        w, h = shapes[i % len(shapes)]
        out.append((cam_mod.view_camera(eye, w, h), w, h))
    return out


def single_view(renderer, cam, w, h):
    renderer.rebin(cam, w, h)
    return {
        "k": renderer.ref_count(),
        "image": renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16),
        "tile_offsets": renderer.read(L.BUF_TILE_OFFSETS, np.uint32),
        "projected": renderer.read(L.BUF_PROJECTED_SPLATS, np.uint32),
        "values": renderer.read(L.BUF_SORT_VALUES, np.uint32),
        "keys": renderer.read(L.BUF_SORT_KEYS, np.uint64),
    }


def check_batch(renderer, views, eager=False):
    renderer.rebin_batch(views, eager=eager)
    ks = renderer.batch_ref_counts()
    got = []
    for i in range(len(views)):
        got.append({"k": int(ks[i]),
                    "image": renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16),
                    "tile_offsets": renderer.batch_read(i, L.BUF_TILE_OFFSETS, np.uint32),
                    "projected": renderer.batch_read(i, L.BUF_PROJECTED_SPLATS, np.uint32),
                    "values": renderer.batch_read(i, L.BUF_SORT_VALUES, np.uint32),
                    "keys": renderer.batch_read(i, L.BUF_SORT_KEYS, np.uint64)})
        assert int(renderer.batch_read(i, L.BUF_REF_COUNTER, np.uint32)[0]) == got[-1]["k"]
    for i, (cam, w, h) in enumerate(views):
        ref = single_view(renderer, cam, w, h)
        for key in ("k", "tile_offsets", "projected", "image", "values", "keys"):
            assert np.array_equal(np.asarray(got[i][key]), np.asarray(ref[key])), (i, key)
    return got


@pytest.fixture(scope="module")
def scene_sh3(renderer):
    s = S.activate(S.clustered(31, 40_000, 12, sh_degree=3))
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    return s


def test_batch_views_equal_single_view_rebins(renderer, scene_sh3):
    say("batch: 9 views of mixed resolution, graph")
    views = views_for(7, 9)
    got = check_batch(renderer, views)
    # one view straight against the oracle as well
    cam, w, h = views[1]
    ref = O.pipeline_f32(cam, scene_sh3, w, h, 16)
    assert got[1]["k"] == ref["k"]
    assert np.array_equal(got[1]["tile_offsets"], ref["tile_offsets"])


def test_batch_graph_replays_with_new_cameras(renderer, scene_sh3):
    """Same view shapes, new cameras: the captured graph is replayed; every
    replay's lookback epochs must differ from the previous replay's."""
    shapes = ((320, 180), (640, 360), (200, 120), (480, 272), (320, 180))
    for rep in range(3):
        say(f"batch: graph replay {rep}")
        check_batch(renderer, views_for(100 + rep, 5, shapes))


def test_batch_eager_equals_graph(renderer, scene_sh3):
    say("batch: eager vs graph")
    views = views_for(11, 6)
    a = check_batch(renderer, views, eager=True)
    renderer.rebin_batch(views)
    for i in range(len(views)):
        assert np.array_equal(renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16), a[i]["image"])


def test_batch_single_view_and_tile_sizes(renderer, scene_sh3):
    for tile in (8, 32):
        say(f"batch: tile {tile}")
        renderer.set_tile_size(tile, tile)
        try:
            check_batch(renderer, views_for(5, 3))
        finally:
            renderer.set_tile_size(16, 16)
    say("batch: one view")
    check_batch(renderer, views_for(6, 1))


def test_batch_four_bit_digits(renderer, scene_sh3):
    """4-bit digits: eight depth passes carry the packed tile rectangles through the
    batch's ping-pong buffers; the batch still equals single views."""
    say("batch: 4-bit digits")
    renderer.set_splat_data(scene_sh3)
    renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 4)
    try:
        check_batch(renderer, views_for(7, 3))
    finally:
        renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 8)


def test_batch_wide_view(renderer, scene_sh3):
    """A view more than 256 tiles wide keeps 8-B tile rectangles that the scan gathers
    (no packed rectangles carried by the depth passes), beside a packed view."""
    say("batch: wide view")
    renderer.set_splat_data(scene_sh3)
    cams = [(cam_mod.view_camera((5.0, 2.0, 6.0), 4200, 40), 4200, 40),
            (cam_mod.view_camera((4.0, -1.0, 7.0), 320, 180), 320, 180)]
    got = check_batch(renderer, cams)
    assert got[0]["k"] > 0 and got[1]["k"] > 0


def test_batch_records_follow_the_batch_splats(renderer, scene_sh3):
    """A batch view's projected records are made again on read from the splats the
    batch projected: after a new upload they are refused, the image still reads."""
    say("batch: records after a new upload")
    renderer.set_splat_data(scene_sh3)
    views = views_for(8, 2)
    got = check_batch(renderer, views)
    renderer.set_splat_data(scene_sh3)
    with pytest.raises(L.Gs3dError):
        renderer.batch_read(0, L.BUF_PROJECTED_SPLATS, np.uint32)
    assert np.array_equal(renderer.batch_read(1, L.BUF_IMAGE_RGBA16F, np.uint16), got[1]["image"])


def test_batch_capacity_regrow(renderer):
    """Large splats make K exceed the initial 8N key budget of a branch: the
    batch grows its buffers, captures again and still equals single views."""
    say("batch: capacity regrow")
    raw = S.uniform_cube(12, 3000, log_scale_mean=-1.2, log_scale_std=0.2)
    s = S.activate(raw)
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    views = [(cam_mod.view_camera((0.0, 0.0, 3.0), 640, 360), 640, 360),
             (cam_mod.view_camera((0.5, 0.2, 2.5), 480, 270), 480, 270)]
    got = check_batch(renderer, views)
    assert max(g["k"] for g in got) > 8 * s.n


def test_batch_errors(renderer, scene_sh3):
    lib = L.lib()
    assert lib.gs3d_rebin_batch(renderer._h, None, 0, 0) == -1
    bad = np.zeros(1, dtype=L.VIEW_DTYPE)
    assert lib.gs3d_rebin_batch(renderer._h, L.ptr(bad), 1, 0) == -1  # zero viewport
    renderer.rebin_batch(views_for(1, 2))
    size = __import__("ctypes").c_size_t()
    assert lib.gs3d_batch_read(renderer._h, 5, L.BUF_IMAGE_RGBA16F, None, 0, size) == -1
    assert lib.gs3d_batch_read(renderer._h, 0, L.BUF_UNSORTED_KEYS, None, 0, size) == -6


def test_batch_autotuner_picks_the_fastest_candidate(renderer, scene_sh3):
    """tune_batch times every (tile, pixels per lane, digit width) candidate on whole
    batches, applies the fastest, and the batch still equals single views there."""
    from gs3d import autotune

    say("batch: autotuner")
    renderer.set_splat_data(scene_sh3)
    views = views_for(3, 4)
    try:
        best, results = autotune.tune_batch(renderer, views, tiles=(8, 16), ppls=(1, 2), radix_bits=(8, 4),
                                            warmup=1, reps=2)
        assert len(results) == 12  # 8x8 tiles take only 1 pixel per lane; two raster modes each
        assert {r.raster_groups for r in results} == {0, 1}
        assert renderer.get_kernel_variant("rasterize-tiles-batch", "GROUPS") == best.raster_groups
        assert best.views_per_s == max(r.views_per_s for r in results)
        assert renderer.get_tile_size() == (best.tile, best.tile)
        assert renderer.get_kernel_variant("radix-sort", "RADIX_BITS") == best.radix_bits
        assert renderer.get_kernel_variant("rasterize-tiles", "QUARTER_LISTS") == 2
        check_batch(renderer, views)
    finally:
        renderer.set_tile_size(16, 16)
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", 1)
        renderer.set_kernel_variant("radix-sort", "RADIX_BITS", 8)
        renderer.set_kernel_variant("rasterize-tiles-batch", "GROUPS", 1)


@pytest.mark.parametrize("groups,wave_items", [(0, 0), (3, 0), (1, 1)])
def test_batch_raster_modes(renderer, scene_sh3, groups, wave_items):
    """The batch's other raster schedules give the same views: one raster launch per view
    on its binning stream (GROUPS 0), three persistent launches over view groups, and
    pixel-block work items (WAVE_ITEMS 1)."""
    say(f"batch: raster groups {groups}, wave items {wave_items}")
    renderer.set_splat_data(scene_sh3)
    renderer.set_kernel_variant("rasterize-tiles-batch", "GROUPS", groups)
    renderer.set_kernel_variant("rasterize-tiles-batch", "WAVE_ITEMS", wave_items)
    try:
        check_batch(renderer, views_for(12, 7))
    finally:
        renderer.set_kernel_variant("rasterize-tiles-batch", "GROUPS", 1)
        renderer.set_kernel_variant("rasterize-tiles-batch", "WAVE_ITEMS", 0)


def test_batch_all_culled_views(renderer, scene_sh3):
    """Views that see no splat (cameras facing away from the scene) next to one that does:
    K = 0, empty tile ranges and a cleared image, exactly as single views give them."""
    say("batch: all-culled views")
    renderer.set_splat_data(scene_sh3)
    cams = [(cam_mod.view_camera((0.0, 0.0, 30.0), 160, 90, target=(0.0, 0.0, 60.0)), 160, 90),
            (cam_mod.view_camera((5.0, 2.0, 6.0), 128, 96), 128, 96),
            (cam_mod.view_camera((0.0, 40.0, 0.5), 96, 96, target=(0.0, 80.0, 0.5)), 96, 96)]
    got = check_batch(renderer, cams)
    assert got[0]["k"] == 0 and got[2]["k"] == 0 and got[1]["k"] > 0
    assert not got[0]["image"].any() and not got[0]["tile_offsets"].any()


def test_batch_odd_splat_count(renderer):
    """An odd splat count (n % 4 != 0): every view's depth-key and rectangle slices of the
    batch arrays still start 16-B aligned, and every view equals its single-view rebin."""
    say("batch: odd splat count")
    s = S.activate(S.clustered(33, 40_001, 12, sh_degree=3))
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    check_batch(renderer, views_for(9, 5))
    s3 = S.activate(S.clustered(34, 3, 1, sh_degree=1))
    renderer.set_splat_data(s3)
    check_batch(renderer, views_for(10, 3))


@pytest.mark.timeout(900)
def test_workload_2m_sh3_64_views(renderer):
    """The batched workload itself: 2M SH3 splats, 64 seeded orbit views of uneven
    resolution (640x360 / 1280x720 / 1920x1080) in one call. Every view: K equals its
    counter and its splats' tile counts, the sorted keys are sorted and stable, each key's
    tile lies in its splat's rectangle, and the tile ranges are the closed form of the
    sorted keys. One view of each resolution against the oracle: keys, values and tile
    ranges bit-exact, image within 1/255."""
    say("batch: workload clusters-2M-sh3-64views")
    O.set_threads(16)
    wl = S.workload(3)
    s = S.activate(wl.raw)
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    views = [(cam_mod.view_camera(eye, w, h, wl.fov_y), w, h) for eye, w, h in wl.views]
    renderer.rebin_batch(views)
    ks = renderer.batch_ref_counts()
    say(f"  K total {int(ks.sum())}, largest view {int(ks.max())}")
    assert len(ks) == 64 and ks.min() > 0
    checked = set()
    for i, (cam, w, h) in enumerate(views):
        tx, ty = -(-w // 16), -(-h // 16)
        to = renderer.batch_read(i, L.BUF_TILE_OFFSETS, np.uint32)
        keys = renderer.batch_read(i, L.BUF_SORT_KEYS, np.uint64)
        vals = renderer.batch_read(i, L.BUF_SORT_VALUES, np.uint32)
        rec = renderer.batch_read(i, L.BUF_PROJECTED_SPLATS, np.uint32).reshape(-1, 16)
        assert int(renderer.batch_read(i, L.BUF_REF_COUNTER, np.uint32)[0]) == int(ks[i]) == keys.size == vals.size
        counts = np.where(rec[:, 11] == 1, (rec[:, 14] - rec[:, 12] + 1) * (rec[:, 15] - rec[:, 13] + 1), 0)
        assert np.array_equal(np.bincount(vals, minlength=s.n), counts)
        same = keys[1:] == keys[:-1]
        assert np.all(keys[1:] >= keys[:-1]) and np.all(vals[1:][same] > vals[:-1][same])
        tile = (keys >> np.uint64(32)).astype(np.int64)
        r = rec[vals]
        assert np.all((tile % tx >= r[:, 12]) & (tile % tx <= r[:, 14]) &
                      (tile // tx >= r[:, 13]) & (tile // tx <= r[:, 15]))
        assert np.array_equal(to, O.identify(keys, tx * ty))
        if (w, h) in checked:
            continue
        checked.add((w, h))
        ref = O.pipeline_f32(cam, s, w, h, 16)
        assert int(ks[i]) == ref["k"], (i, w, h)
        assert np.array_equal(keys, ref["sorted_keys"]) and np.array_equal(vals, ref["sorted_values"])
        assert np.array_equal(to, ref["tile_offsets"])
        img = renderer.batch_read(i, L.BUF_IMAGE_RGBA32F, np.float32).reshape(h, w, 4)
        err = float(np.abs(img - ref["image"]).max())
        say(f"  view {i} {w}x{h}: K={ref['k']} keys, values, ranges bit-exact; image max abs vs oracle f32 {err:.3g}")
        assert err <= TOL_IMAGE
    assert len(checked) == 3


TOL_IMAGE = 1.0 / 255.0
