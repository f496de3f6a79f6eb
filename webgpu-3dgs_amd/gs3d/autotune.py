"""Autotuner for the launch-shape knobs of the rebin path.

Knobs (each a compiled kernel variant): binning/raster tile size {8, 16, 32},
rasterizer pixels per lane {1, 2, 4} (a wavefront covers 64*ppl pixels of one
tile) and radix digit width {4, 8}. A candidate is timed as the mean GPU time of
whole rebins (hipEvent pairs inside the library) over the given views after a
warm-up; the fastest wins. Tile size changes the binning and so the image;
callers that must keep a given tile size pass `tiles=(t,)`.
"""
from __future__ import annotations

import itertools
import json
from dataclasses import asdict, dataclass
from pathlib import Path

from . import camera as cam_mod


@dataclass
class Choice:
    tile: int
    pixels_per_lane: int
    radix_bits: int
    frame_us: float
    stages_us: dict


def candidates(tiles=(8, 16, 32), ppls=(1, 2, 4), radix_bits=(4, 8)):
    for t, p, b in itertools.product(tiles, ppls, radix_bits):
        if t * t // p >= 64:
            yield t, p, b


def tune(renderer, views, *, tiles=(8, 16, 32), ppls=(1, 2, 4), radix_bits=(8, 4), warmup=3,
         reps=10, fov_y=None) -> tuple[Choice, list[Choice]]:
    """Time every candidate over `views` [(eye, W, H), ...]; returns (best, all)."""
    import math

    fov = fov_y if fov_y is not None else math.pi / 3
    cams = [(cam_mod.view_camera(eye, w, h, fov), w, h) for eye, w, h in views]
    results = []
    for t, p, b in candidates(tiles, ppls, radix_bits):
        renderer.set_tile_size(t, t)
        renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", p)
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
        results.append(Choice(t, p, b, stages["rebin-total"], stages))
    best = min(results, key=lambda c: c.frame_us)
    apply(renderer, best)
    return best, results


def apply(renderer, choice: Choice) -> None:
    renderer.set_tile_size(choice.tile, choice.tile)
    renderer.set_kernel_variant("rasterize-tiles", "PIXELS_PER_LANE", choice.pixels_per_lane)
    renderer.set_kernel_variant("radix-sort", "RADIX_BITS", choice.radix_bits)


def save(path: str | Path, best: Choice, results: list[Choice]) -> None:
    Path(path).write_text(json.dumps({"best": asdict(best), "all": [asdict(r) for r in results]},
                                     indent=1))
