"""Autotuner for the launch-shape knobs of the rebin path.

Knobs (each a compiled kernel variant): binning/raster tile size {8, 16, 32},
rasterizer pixels per lane {1, 2, 4} (a wavefront covers 64*ppl pixels of one
tile), at one pixel per lane survivor lists per 4x4 quarter of each 8x8 block
or per block (quarter_lists {1, 0}; the block lists win when most splats span
the whole block), and radix digit width {4, 8}; for batches also the raster mode: one
persistent rasterizer over every view's tiles after all views are binned
(raster_groups 1), or one raster launch per view on its binning stream
(raster_groups 0). `tune` times a candidate as the mean GPU
time of whole single-view rebins (hipEvent pairs inside the library) over the
given views after a warm-up; `tune_batch` times whole batched calls (one
replayed hipGraph over all the views, wall clock). The fastest wins. Tile size
changes the binning and so the image; callers that must keep a given tile size
pass `tiles=(t,)`.
"""
from __future__ import annotations

import itertools
import json
import time
from dataclasses import asdict, dataclass, field
from pathlib import Path

from . import camera as cam_mod


@dataclass
class Choice:
    tile: int
    pixels_per_lane: int
    radix_bits: int
    frame_us: float
    stages_us: dict
    quarter_lists: int = 2


def candidates(tiles=(8, 16, 32), ppls=(1, 2, 4), radix_bits=(4, 8), quarter_lists=(2, 1, 0)):
    """(tile, pixels per lane, digit bits, quarter lists); quarter lists only vary at one pixel per lane."""
    for t, p, b in itertools.product(tiles, ppls, radix_bits):
        if t * t // p >= 64:
            for q in (quarter_lists if p == 1 else (2,)):
                yield t, p, b, q


def tune(renderer, views, *, tiles=(8, 16, 32), ppls=(1, 2, 4), radix_bits=(8, 4), quarter_lists=(2, 1, 0),
         warmup=3, reps=10, fov_y=None) -> tuple[Choice, list[Choice]]:
    """Time every candidate over `views` [(eye, W, H), ...]; returns (best, all)."""
    import math

    fov = fov_y if fov_y is not None else math.pi / 3
    cams = [(cam_mod.view_camera(eye, w, h, fov), w, h) for eye, w, h in views]
    results = []
    for t, p, b, q in candidates(tiles, ppls, radix_bits, quarter_lists):
        renderer.set_tile_size(t, t)
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", p)
        renderer.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", q)
        renderer.set_kernel_variant("radix-sort", "RADIX_BITS", b)
        for _ in range(warmup):
            for cu, w, h in cams:
                renderer.rebin(cu, w, h)
        renderer.sync()
        renderer.profiler_reset()
        for _ in range(reps):
            for cu, w, h in cams:
                renderer.rebin(cu, w, h, 1)
        stages, frames = renderer.stage_times_avg_us()
        results.append(Choice(t, p, b, stages["rebin-total"], stages, q))
    best = min(results, key=lambda c: c.frame_us)
    apply(renderer, best)
    return best, results


@dataclass
class BatchChoice:
    tile: int
    pixels_per_lane: int
    radix_bits: int
    views_per_s: float
    ms_per_batch: float
    raster_groups: int = 1
    extra: dict = field(default_factory=dict)


def tune_batch(renderer, cams, *, tiles=(8, 16, 32), ppls=(1, 2), radix_bits=(8, 4), raster_groups=(1, 0),
               warmup=2, reps=4, learned=False) -> tuple[BatchChoice, list[BatchChoice]]:
    """Time every candidate on gs3d_rebin_batch over `cams` [(camera uniform, W, H), ...]
    (a re-capture per candidate, then `reps` timed replays); returns (best, all)."""
    results = []
    # the persistent rasterizer keeps the quarter lists; the batch tunes the rest, and single
    # frames keep the per-frame choice of list form
    for (t, p, b, _), g in itertools.product(candidates(tiles, ppls, radix_bits, quarter_lists=(2,)),
                                             raster_groups):
        apply(renderer, Choice(t, p, b, 0.0, {}))
        renderer.set_kernel_variant("rasterize-tiles-batch", "GROUPS", g)
        for _ in range(warmup):
            renderer.rebin_batch(cams, learned=learned)
        renderer.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            renderer.rebin_batch(cams, learned=learned)
        renderer.sync()
        per = (time.perf_counter() - t0) / reps
        results.append(BatchChoice(t, p, b, len(cams) / per, per * 1e3, g))
    best = max(results, key=lambda c: c.views_per_s)
    apply(renderer, Choice(best.tile, best.pixels_per_lane, best.radix_bits, 0.0, {}))
    renderer.set_kernel_variant("rasterize-tiles-batch", "GROUPS", best.raster_groups)
    return best, results


def apply(renderer, choice: Choice) -> None:
    renderer.set_tile_size(choice.tile, choice.tile)
    renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", choice.pixels_per_lane)
    renderer.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", choice.quarter_lists)
    renderer.set_kernel_variant("radix-sort", "RADIX_BITS", choice.radix_bits)


def save(path: str | Path, best: Choice, results: list[Choice]) -> None:
    Path(path).write_text(json.dumps({"best": asdict(best), "all": [asdict(r) for r in results]},
                                     indent=1))
