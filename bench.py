#!/usr/bin/env python3
"""Frames/s of the rebin-then-rasterize path on one MI355X.

A step is one rebin (preprocess -> scan -> emit -> radix sort -> identify ->
rasterize) of the 1M-splat SH3 scene at 1920x1080 with 16x16 tiles, splats
already resident in HBM. Prints ONE JSON line (rank 0).

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload 1]
                  [--no-cpu-baseline] [--cpu-seconds S] [--no-batched]

The same line carries "stress": the 6M-splat SH3 frame at 3840x2160 (about
119M references), with the sort's HBM fraction on its own bytes there;
"batched": the 2M-splat SH3 scene rendered for 64 seeded orbit views of uneven
resolution (640x360 / 1280x720 / 1920x1080) per gs3d_rebin_batch call, as
views/s from one replayed hipGraph, beside the same views launched eagerly and
one gs3d_rebin per view; and "learned": the learned-colour batch of the
3M-splat scene against exact SH.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

REPO = Path(__file__).resolve().parent
sys.path.insert(0, str(REPO / "webgpu-3dgs_amd"))

METRIC = "frames/s at 1M-splat SH3 1080p; batched views/s, learned vs plain @\u226545dB"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
RASTER_FLOP_PER_PAIR = 23  # ~22 flop + 1 exp per evaluated (pixel, splat) pair


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def algorithmic_bytes(n, k, tiles, pixels, depth_passes, tile_passes, deg, packed=True):
    """Bytes each stage must move at minimum, per launch group of one frame.

    With `packed` (both tile counts <= 256) the tile rectangle is 4 B and moves
    with the depth keys; otherwise it is 8 B and the scan gathers it.
    preprocess: vec4 inputs + SH coefficients in, 64-B projected splat + count
    + depth key + tile rectangle out; depth sort: per pass a u32 key + u32 id
    (+ the 4-B rectangle when packed) read and written (the first pass reads
    keys, and rectangles, only), plus one read of the keys to histogram;
    scan: the depth key and the rectangle (packed: read in order; else the id
    and the 8-B rectangle) of each splat in, a 16-B record per splat out; emit: the records and ids in,
    (tile id, 32-bit splat id) out; key sort: per key and pass the tile id and
    the splat id, read and written (tile ids are 16-bit up to 65,536 tiles, so
    12 B per key and pass, else 32-bit, 16 B); identify: one (T+1)-entry write plus, per tile, the ceil(log2 K)
    sorted keys its binary search reads; raster: the 48 B used of each listed
    splat, tile ranges, 8-B pixel out.
    """
    pre_in = 64 + (12 * (deg + 1) ** 2 if deg else 0)
    rect = 4 if packed else 8
    carry = 4 if packed else 0
    tid = 2 if tiles <= 65536 else 4  # bytes per tile id in emit's output and the tile-id passes
    return {
        "preprocess-splats": n * (pre_in + 64 + 4 + 4 + rect),
        "radix-sort-splats": n * (4 + 12 + 2 * carry + (16 + 2 * carry) * (depth_passes - 1)) if depth_passes else 0,
        "prefix-scan": (8 if packed else 16) * n + 16 * (n + 1),
        "emit-tile-refs": 16 * (n + 1) + 4 * n + (tid + 4) * k,
        "radix-sort": tile_passes * 2 * (tid + 4) * k,
        "identify-tile-ranges": 8 * tiles * max(1, int(k).bit_length()) + 4 * (tiles + 1),
        "rasterize-tiles": 48 * k + 8 * (tiles + 1) + 8 * pixels,
    }


def load_traffic(workload: str, kernel: str):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc FETCH_SIZE /
    WRITE_SIZE (calibrated against a plain copy), from profiles/traffic.json."""
    f = REPO / "profiles" / "traffic.json"
    if not f.exists():
        return None
    try:
        return json.loads(f.read_text()).get(workload, {}).get(kernel)
    except (ValueError, AttributeError):
        return None


def host_cpu() -> dict:
    """CPU model and core counts of the host this runs on."""
    model = "unknown"
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count() or 1
    return {"cpu_model": model, "host_cores_total": os.cpu_count() or 1, "host_cores_usable": usable}


def probe_cpu_webgpu() -> dict:
    """Look (install nothing) for a WebGPU implementation that could run the
    reference's WGSL on the host CPU: a Dawn or wgpu-native library, the
    `webgpu` Node package, or a CPU Vulkan driver (lavapipe / SwiftShader) for
    Dawn to sit on."""
    import glob
    import shutil
    import subprocess

    found = {}
    pats = {"wgpu_native": ["/usr/lib*/**/libwgpu_native*", "/usr/local/lib*/**/libwgpu_native*",
                            "/opt/**/libwgpu_native*"],
            "dawn": ["/usr/lib*/**/libdawn*", "/usr/local/lib*/**/libdawn*", "/opt/**/libwebgpu_dawn*"],
            "node_webgpu": ["/usr/lib/node_modules/webgpu", "/usr/local/lib/node_modules/webgpu",
                            "/root/node_modules/webgpu"],
            "vulkan_cpu_icd": ["/usr/share/vulkan/icd.d/*lvp*", "/usr/share/vulkan/icd.d/*swiftshader*",
                               "/etc/vulkan/icd.d/*lvp*", "/etc/vulkan/icd.d/*swiftshader*"]}
    for key, globs in pats.items():
        hits = []
        for g in globs:
            hits += glob.glob(g, recursive=True)[:4]
        found[key] = sorted(set(hits))[:4]
    libs = []
    for d in ("/usr/lib", "/usr/lib/x86_64-linux-gnu", "/usr/local/lib", "/lib/x86_64-linux-gnu"):
        for pat in ("libwgpu*", "libdawn*", "libwebgpu*", "libvulkan*"):
            libs += glob.glob(f"{d}/{pat}")
    found["system_libraries"] = sorted(set(libs))[:8]
    node = shutil.which("node")
    found["node"] = subprocess.run([node, "--version"], capture_output=True, text=True,
                                   timeout=10).stdout.strip() if node else None
    usable = bool(found["wgpu_native"] or found["dawn"] or found["node_webgpu"])
    found["usable_cpu_adapter"] = usable and bool(found["vulkan_cpu_icd"])
    found["note"] = "looked for, nothing installed; run before the GPU is initialised"
    return found


def cpu_baseline(cam, splats, W, H, tile, seconds, gpu_image=None):
    """The oracle's fp32 pipeline (a port of the reference WGSL passes) on the
    host cores, whole frames of the same workload, until `seconds` of work.
    With gpu_image (H x W x 4), also the GPU frame's max-abs error against the
    fp32 oracle frame and the fp64 render."""
    sys.path.insert(0, str(REPO / "tests"))
    import oracle_lib as O

    threads = max(1, min(os.cpu_count() or 1, 16))
    O.set_threads(threads)
    frames, t0, visited = 0, time.perf_counter(), 0
    while True:
        img, k, visited, blended = O.render(cam, splats, W, H, tile, "f32")
        frames += 1
        el = time.perf_counter() - t0
        if el >= seconds or frames >= 500:
            break
    out = {"value": frames / el, "unit": "frames/s", "cores": threads, "kind": "port",
           "sample": f"{frames} whole frames of the same workload (oracle fp32 pipeline, "
                     f"{threads} OpenMP threads, {el:.1f} s)", **host_cpu()}
    err = None
    if gpu_image is not None:
        f64, k64, _, _ = O.render(cam, splats, W, H, tile, "f64")
        err = {"vs_oracle_f32": float(np.abs(gpu_image - img).max()),
               "vs_render_f64": float(np.abs(gpu_image - f64).max()),
               "oracle_keys": int(k), "f64_keys": int(k64), "tolerance": 1.0 / 255.0}
    return out, visited, blended, err


def sort_roofline_of(algo: dict, stages: dict) -> dict | None:
    """The sort's HBM fraction on its own algorithmic bytes: the depth passes over
    the splats and the tile-id passes over the references, together and each."""
    sort_us = stages.get("radix-sort", 0.0) + stages.get("radix-sort-splats", 0.0)
    if sort_us <= 0:
        return None

    def frac(b, us):
        return round(b / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 4) if us > 0 else None

    own = algo["radix-sort-splats"] + algo["radix-sort"]
    own_gbs = own / (sort_us * 1e-6) / 1e9
    return {"bound": "hbm", "achieved": round(own_gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(own_gbs / HBM_PEAK_GBS, 4), "algo_bytes": int(own), "us": round(sort_us, 2),
            "depth_passes_frac": frac(algo["radix-sort-splats"], stages.get("radix-sort-splats", 0.0)),
            "tile_passes_frac": frac(algo["radix-sort"], stages.get("radix-sort", 0.0))}


def bench_stress(steps: int, warmup: int) -> dict:
    """The 6M-splat SH3 frame at 3840x2160 (about 119M references): frames/s and
    the sort's own-bytes HBM fraction, the size where the sort is bandwidth-bound
    rather than latency-bound."""
    from gs3d import camera, scene
    from gs3d._lib import STAGES
    from gs3d.renderer import GaussianSplatRenderer

    t = time.perf_counter()
    wl = scene.workload(2)
    splats = scene.activate(wl.raw)
    eye, W, H = wl.views[0]
    cu = camera.view_camera(eye, W, H, wl.fov_y)
    log(f"[bench] stress workload {wl.name}: {splats.n} splats, {W}x{H} (generated in "
        f"{time.perf_counter() - t:.1f} s)")
    r = GaussianSplatRenderer(0)
    r.set_splat_data(splats)
    r.set_tile_size(wl.tile, wl.tile)
    for _ in range(warmup):
        r.rebin(cu, W, H)  # the first frame grows the key buffers and runs again
    r.sync()
    k = r.ref_count()
    r.set_profile_mask(STAGES)
    r.profiler_reset()
    for _ in range(10):
        r.rebin(cu, W, H, 1)  # GS3D_REBIN_PROFILE
    stages, _ = r.stage_times_avg_us()
    r.sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        r.rebin(cu, W, H)
    r.sync()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    tiles_x, tiles_y = r.get_tile_count()
    shifts = r.sort_passes()
    depth_passes = sum(1 for b in shifts if b < 32)
    algo = algorithmic_bytes(splats.n, k, tiles_x * tiles_y, W * H, depth_passes, len(shifts) - depth_passes,
                             splats.sh_degree, packed=tiles_x <= 256 and tiles_y <= 256)
    r.close()
    out = {"workload": wl.name, "splats": splats.n, "width": W, "height": H, "keys": k,
           "value": round(1000.0 / ms, 2), "unit": "frames/s", "ms_per_frame": round(ms, 4), "steps": steps,
           "stages_us": {n: round(v, 2) for n, v in stages.items()},
           "sort_roofline": sort_roofline_of(algo, stages)}
    log(f"[bench] stress: {out['value']} frames/s ({ms:.3f} ms), K={k}; sort {out['sort_roofline']['us']} us, "
        f"{out['sort_roofline']['frac']:.3f} of HBM peak on its own bytes")
    return out


def bench_batched(steps: int, warmup: int, tune: bool = True) -> dict:
    """views/s of the 64-view batch (2M splats, SH3, uneven resolutions) at 16x16 tiles,
    1 pixel per lane and 8-bit digits; then the autotuner's pick and its views/s."""
    from gs3d import camera, scene
    from gs3d.renderer import GaussianSplatRenderer

    t = time.perf_counter()
    wl = scene.workload(3)
    splats = scene.activate(wl.raw)
    views = [(camera.view_camera(eye, w, h, wl.fov_y), w, h) for eye, w, h in wl.views]
    px = sum(w * h for _, w, h in views)
    log(f"[bench] batched workload {wl.name}: {splats.n} splats, {len(views)} views, "
        f"{px / 1e6:.1f} Mpx (generated in {time.perf_counter() - t:.1f} s)")
    r = GaussianSplatRenderer(0)
    r.set_splat_data(splats)
    r.set_tile_size(wl.tile, wl.tile)
    out = {"workload": wl.name, "views": len(views), "megapixels": round(px / 1e6, 2)}

    def timed(fn, reps):
        r.sync()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        r.sync()
        return (time.perf_counter() - t0) / reps

    for _ in range(warmup):
        r.rebin_batch(views)
    per_p = timed(lambda: r.rebin_batch(views), steps)
    ks = r.batch_ref_counts()
    # the same batch with one raster launch per view on its binning stream
    r.set_kernel_variant("rasterize-tiles-batch", "GROUPS", 0)
    for _ in range(warmup):
        r.rebin_batch(views)
    per_v = timed(lambda: r.rebin_batch(views), steps)
    r.set_kernel_variant("rasterize-tiles-batch", "GROUPS", 1)
    # value: the faster of the two raster schedules, both replayed from a captured graph
    per = min(per_p, per_v)
    out.update({"value": round(len(views) / per, 2), "unit": "views/s", "ms_per_batch": round(per * 1e3, 3),
                "mode": ("hipGraph replay, one persistent rasterizer over every view's tiles" if per_p <= per_v
                         else "hipGraph replay, a raster launch per view on its binning stream"),
                "persistent_raster_views_per_s": round(len(views) / per_p, 2),
                "per_view_raster_views_per_s": round(len(views) / per_v, 2),
                "keys_total": int(ks.sum()), "keys_max_view": int(ks.max())})
    r.rebin_batch(views)
    r.rebin_batch(views, eager=True)
    per_e = timed(lambda: r.rebin_batch(views, eager=True), max(1, steps // 2))
    out["eager_views_per_s"] = round(len(views) / per_e, 2)

    def one_by_one():
        for cam, w, h in views:
            r.rebin(cam, w, h)
    one_by_one()
    per_s = timed(one_by_one, 1)
    out["single_view_rebins_views_per_s"] = round(len(views) / per_s, 2)
    if tune:
        from gs3d import autotune

        best, results = autotune.tune_batch(r, views, tiles=(8, 16, 32), ppls=(1, 2), radix_bits=(8, 4),
                                            raster_groups=(1, 0), warmup=1, reps=3)
        out["autotune"] = {"best": {"tile": best.tile, "pixels_per_lane": best.pixels_per_lane,
                                    "radix_bits": best.radix_bits, "raster_groups": best.raster_groups,
                                    "views_per_s": round(best.views_per_s, 2)},
                           "candidates_note": "tile, pixels per lane, digit bits, raster groups (1: one persistent "
                                              "rasterizer, 0: a raster launch per view), views/s",
                           "candidates": [[c.tile, c.pixels_per_lane, c.radix_bits, c.raster_groups,
                                           round(c.views_per_s, 1)] for c in results]}
        per_t = timed(lambda: r.rebin_batch(views), steps)
        out["tuned_views_per_s"] = round(len(views) / per_t, 2)
        log(f"[bench] batch autotune: best tile {best.tile}, ppl {best.pixels_per_lane}, "
            f"{best.radix_bits}-bit digits, raster groups {best.raster_groups}: "
            f"{out['tuned_views_per_s']} views/s")
    log(f"[bench] batched: {out['value']} views/s; graph with the persistent rasterizer "
        f"{out['persistent_raster_views_per_s']}, with a raster launch per view "
        f"{out['per_view_raster_views_per_s']}, persistent form eager {out['eager_views_per_s']}, "
        f"{out['single_view_rebins_views_per_s']} one rebin per view; K total {out['keys_total']}")
    r.close()
    return out


def bench_learned(steps: int, warmup: int, epochs: int) -> dict:
    """The 32-view 1080p batch of the 3M-splat SH3 scene, coloured by exact SH and by the
    learned stage (trained here on the GPU); PSNR of the learned render against exact SH."""
    from gs3d import camera, scene
    from gs3d.renderer import GaussianSplatRenderer

    t = time.perf_counter()
    wl = scene.workload(4)
    splats = scene.activate(wl.raw)
    views = [(camera.view_camera(eye, w, h, wl.fov_y), w, h) for eye, w, h in wl.views]
    log(f"[bench] learned workload {wl.name}: {splats.n} splats, {len(views)} views "
        f"(generated in {time.perf_counter() - t:.1f} s)")
    r = GaussianSplatRenderer(0)
    r.set_splat_data(splats)
    r.set_tile_size(wl.tile, wl.tile)
    r.learned_init(seed=4)
    r.sync()
    t0 = time.perf_counter()
    losses = r.learned_train(epochs, lr=3e-3, lr_latent=3e-2)
    train_s = time.perf_counter() - t0
    psnr, enabled = r.learned_validate(views, 45.0)
    log(f"[bench] learned: {epochs} epochs in {train_s:.2f} s (loss {losses[0]:.5f} -> {losses[-1]:.5f}); "
        f"PSNR vs exact SH {psnr:.2f} dB -> {'learned colours' if enabled else 'falls back to exact SH'}")

    def rate(learned):
        # the best of 3 timed runs of `steps` batches, for both colourings alike:
        # one run is under 0.1 s, where a single host hiccup shows
        for _ in range(warmup):
            r.rebin_batch(views, learned=learned)
        r.sync()
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            for _ in range(steps):
                r.rebin_batch(views, learned=learned)
            r.sync()
            best = min(best, time.perf_counter() - t0)
        return len(views) * steps / best

    plain = rate(False)
    r.learned_enable(True)  # measured even when the 45 dB gate falls back
    learned = rate(True)
    r.learned_enable(enabled)
    out = {"workload": wl.name, "views": len(views), "plain_views_per_s": round(plain, 2),
           "learned_views_per_s": round(learned, 2), "learned_over_plain": round(learned / plain, 3),
           "psnr_db": round(psnr, 2), "psnr_gate_db": 45.0, "learned_enabled": enabled,
           "train_epochs": epochs, "train_seconds": round(train_s, 2),
           "train_samples_per_s": round(epochs * 256 * splats.n / train_s, 1),
           "loss_first_epoch": float(losses[0]), "loss_last_epoch": float(losses[-1])}
    log(f"[bench] learned batch {out['learned_views_per_s']} views/s vs exact SH {out['plain_views_per_s']} "
        f"({out['learned_over_plain']}x)")
    r.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", type=int, default=1)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--quarter-lists", type=int, default=2, choices=(0, 1, 2),
                    help="rasterizer survivor lists per 4x4 quarter (1), per 8x8 block (0), or chosen per "
                         "frame from the previous frame's references per splat (2)")
    ap.add_argument("--key-tile", type=int, default=0, help="keys per workgroup of the 64-bit sort passes "
                    "(0 = chosen by key count, 8192, 16384)")
    ap.add_argument("--depth-key-tile", type=int, default=0, help="keys per workgroup of the 32-bit depth "
                    "sort passes (0 = 8192, 4096)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-view batch measurement")
    ap.add_argument("--no-stress", action="store_true",
                    help="skip the 6M-splat 3840x2160 frame (frames/s and the sort's HBM fraction there)")
    ap.add_argument("--only-batched", action="store_true",
                    help="run only the 64-view batch measurement (for kernel traces of it)")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-colour measurement")
    ap.add_argument("--only-learned", action="store_true", help="run only the learned-colour measurement")
    ap.add_argument("--learned-epochs", type=int, default=16)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 256 MiB device to device 5 times (the reference reading for the "
                         "FETCH_SIZE / WRITE_SIZE counters) and exit")
    ap.add_argument("--batch-steps", type=int, default=5)
    ap.add_argument("--composite", type=int, default=0,
                    help="after the timed region, run this many composites per display mode "
                         "(plain, tile heat map) and report their wall time per call")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: this path runs on a single MI355X "
              "(a multi-GPU run would shard batches by view); use --gpus 1", file=sys.stderr)
        return 2

    webgpu_probe = probe_cpu_webgpu()  # before anything initialises the GPU (it starts node)
    if args.copy_calibration:
        import torch

        a = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        for _ in range(5):
            b.copy_(a)
        torch.cuda.synchronize()
        log(f"[bench] copied {a.numel() * 4} bytes x5")
        return 0
    if args.only_batched:
        print(json.dumps({"batched": bench_batched(args.batch_steps, 2)}), flush=True)
        return 0
    if args.only_learned:
        print(json.dumps({"learned": bench_learned(args.batch_steps, 2, args.learned_epochs)}), flush=True)
        return 0

    from gs3d import camera, scene
    from gs3d._lib import REBIN_STAGES, STAGES, reference_pass_table
    from gs3d.renderer import GaussianSplatRenderer

    t = time.perf_counter()
    wl = scene.workload(args.workload)
    splats = scene.activate(wl.raw)
    eye, W, H = wl.views[0]
    cu = camera.view_camera(eye, W, H, wl.fov_y)
    log(f"[bench] workload {wl.name}: {splats.n} splats SH{splats.sh_degree}, {W}x{H} "
        f"(generated in {time.perf_counter() - t:.1f} s)")

    r = GaussianSplatRenderer(0)
    r.set_splat_data(splats)
    r.set_tile_size(args.tile, args.tile)
    r.set_kernel_variant("radix-sort", "KEY_TILE", args.key_tile)
    r.set_kernel_variant("radix-sort-splats", "KEY_TILE", args.depth_key_tile)
    r.set_kernel_variant("rasterize-tiles", "QUARTER_LISTS", args.quarter_lists)
    for _ in range(args.warmup):
        r.rebin(cu, W, H)
    r.sync()
    k = r.ref_count()
    log(f"[bench] warmup done: K={k} keys, passes at bits {r.sort_passes()}")

    # Stage breakdown: a short run with a hipEvent pair around every stage.
    # Each event pair adds a dispatch gap, so these frames are not the timed ones.
    r.set_profile_mask(STAGES)
    r.profiler_reset()
    for _ in range(20):
        r.rebin(cu, W, H, 1)  # GS3D_REBIN_PROFILE
    stages, _ = r.stage_times_avg_us()
    dominant = max(REBIN_STAGES, key=lambda n: stages[n])

    # Timed region: K back-to-back rebins; every 8th one brackets the dominant
    # kernel with a hipEvent pair on the context's stream (the stream every
    # kernel is launched on), the rest queue no events at all.
    r.set_profile_mask([dominant])
    r.profiler_reset()
    r.sync()
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.rebin(cu, W, H, 1 if i % 8 == 0 else 0)
    r.sync()
    el = time.perf_counter() - t0
    timed, frames = r.stage_times_avg_us()
    dominant_us = timed[dominant]
    ms = el * 1000.0 / args.steps
    log(f"[bench] {args.steps} steps in {el:.3f} s: {ms:.3f} ms/frame; {dominant} {dominant_us:.1f} us "
        f"over {frames} sampled frames; profiled breakdown (us) "
        + ", ".join(f"{n}={v:.1f}" for n, v in stages.items()))

    tiles_x, tiles_y = r.get_tile_count()
    T = tiles_x * tiles_y
    shifts = r.sort_passes()
    depth_passes = sum(1 for b in shifts if b < 32)
    passes = len(shifts) - depth_passes
    algo = algorithmic_bytes(splats.n, k, T, W * H, depth_passes, passes, splats.sh_degree,
                             packed=tiles_x <= 256 and tiles_y <= 256)

    # Moving camera: the headline frame is one fixed camera, whose splat data
    # (64 MB of vec4 + 192 MB of SH at 1M SH3) could stay in the 256 MB
    # last-level cache between frames. The 8 orbit positions (azimuth k * 45
    # deg at the same radius and elevation) are timed each on its own, and
    # cycled frame by frame, so every frame renders a different view.
    orbit = []
    img_headline = r.read_image(W, H)
    if wl.orbit_radius:
        cams = [camera.view_camera(scene.orbit_eye(wl.orbit_radius, 45.0 * j, 0.0), W, H, wl.fov_y)
                for j in range(8)]
        reps = max(8, args.steps // 8)
        for j, cj in enumerate(cams):
            for _ in range(3):
                r.rebin(cj, W, H)
            r.sync()
            t0 = time.perf_counter()
            for _ in range(reps):
                r.rebin(cj, W, H)
            r.sync()
            orbit.append({"azimuth_deg": 45 * j, "ms": round((time.perf_counter() - t0) * 1e3 / reps, 4),
                          "keys": r.ref_count()})
        r.sync()
        t0 = time.perf_counter()
        cyc = 8 * reps
        for i in range(cyc):
            r.rebin(cams[i % 8], W, H)
        r.sync()
        cycle_ms = (time.perf_counter() - t0) * 1e3 / cyc
        log("[bench] orbit: " + ", ".join(f"{o['azimuth_deg']}deg {o['ms']:.3f} ms" for o in orbit)
            + f"; cycling all 8 every frame {cycle_ms:.3f} ms/frame")

    cpu = None
    visited = None
    image_error = None
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline (oracle port) ...")
        cpu, visited, _, image_error = cpu_baseline(cu, splats, W, H, args.tile, args.cpu_seconds,
                                                    gpu_image=img_headline)
        log(f"[bench] cpu baseline {cpu['value']:.3f} frames/s on {cpu['cores']} of {cpu['host_cores_total']} "
            f"cores ({cpu['cpu_model']}); image max abs vs oracle f32 {image_error['vs_oracle_f32']:.3g}, "
            f"vs fp64 {image_error['vs_render_f64']:.3g}")

    per_stage = {}
    for name, b in algo.items():
        us = stages.get(name, 0.0)
        gbs = b / (us * 1e-6) / 1e9 if us > 0 else 0.0
        per_stage[name] = {"us": round(us, 2), "algo_bytes": int(b), "GB/s": round(gbs, 1),
                           "hbm_frac": round(gbs / HBM_PEAK_GBS, 4)}
    if visited:
        us = dominant_us if dominant == "rasterize-tiles" else stages["rasterize-tiles"]
        tf = visited * RASTER_FLOP_PER_PAIR / (us * 1e-6) / 1e12
        per_stage["rasterize-tiles"].update({"pairs": visited, "TFLOP/s": round(tf, 2),
                                             "valu_frac": round(tf / FP32_PEAK_TFLOPS, 4)})
    sort_us = stages.get("radix-sort", 0.0) + stages.get("radix-sort-splats", 0.0)
    if sort_us > 0:
        # The whole sort priced on the reference's layout: one 64-bit key
        # (tile << 32 | depth) + 32-bit id per reference, sorted in
        # ceil((32 + ceil(log2 T)) / 8) 8-bit passes of 24 B per key each, plus
        # one 8-B histogram read (sort.wgsl). This pipeline reaches the same
        # order with four 32-bit passes over the splats and then passes over
        # 32-bit tile ids; the figure measures both against the bytes the
        # reference's sort has to move.
        ref_passes = -(-(32 + (max(1, T) - 1).bit_length()) // 8)
        ref_bytes = 8 * k + ref_passes * 24 * k
        gbs = ref_bytes / (sort_us * 1e-6) / 1e9
        per_stage["sort-reference-layout"] = {"us": round(sort_us, 2), "algo_bytes": int(ref_bytes),
                                              "passes": ref_passes, "GB/s": round(gbs, 1),
                                              "reference_layout_equivalent_frac": round(gbs / HBM_PEAK_GBS, 4),
                                              "note": "the reference's sort bytes over this sort's time: "
                                                      "not this sort's own traffic, not a roofline fraction"}
    d = per_stage[dominant]
    d["us_timed_region"] = round(dominant_us, 2)
    d["GB/s"] = round(algo[dominant] / (dominant_us * 1e-6) / 1e9, 1)
    d["hbm_frac"] = round(d["GB/s"] / HBM_PEAK_GBS, 4)
    traffic = load_traffic(wl.name, dominant)
    if dominant == "rasterize-tiles" and visited:
        # The rasterizer is bound by fp32 vector issue, not HBM: algorithmic
        # flops = oracle pixel-splat pairs x 23 over the kernel time, against
        # MI355X's fp32 peak.
        tf = visited * RASTER_FLOP_PER_PAIR / (dominant_us * 1e-6) / 1e12
        roofline = {"kernel": dominant, "bound": "valu", "achieved": round(tf, 2),
                    "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                    "frac": round(tf / FP32_PEAK_TFLOPS, 4), "traffic": traffic,
                    "hbm_frac": d["hbm_frac"]}
    else:
        roofline = {"kernel": dominant, "bound": "hbm", "achieved": d["GB/s"], "peak": HBM_PEAK_GBS,
                    "unit": "GB/s", "frac": d["hbm_frac"], "traffic": traffic}
    fps = 1000.0 / ms
    out = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": wl.name, "splats": splats.n, "sh_degree": splats.sh_degree,
                   "width": W, "height": H, "tile": args.tile, "keys": k, "tiles": T,
                   "radix_passes": {"splats": depth_passes, "keys": passes}, "digit_bits": 8,
                   "raster_lists": {0: "per 8x8 block", 1: "per 4x4 quarter",
                                    2: "per 4x4 quarter up to 8 references per splat, else per 8x8 block, "
                                       "chosen per frame"}[args.quarter_lists]},
        "mkeys_per_s": round(k / ((stages["radix-sort"] + stages["radix-sort-splats"]) * 1e-6) / 1e6, 1)
        if stages["radix-sort"] else None,
        "stages_us": {n: round(v, 2) for n, v in stages.items()},
        "reference_passes": [[labels, stage, round(us, 2)] for labels, stage, us
                             in reference_pass_table(stages) if stage != "composite-splats"],
        "stages_note": "stages_us from 20 frames with an event pair per stage (each pair adds a "
                       "dispatch gap); ms_per_step from frames without them",
        "dominant_sampled_frames": frames,
        "per_stage": per_stage,
        "roofline": roofline,
        "sort_roofline": sort_roofline_of(algo, stages),
        "cpu_baseline": cpu,
        "image_error": image_error,
        "orbit": {"positions": orbit,
                  "mean_ms": round(sum(o["ms"] for o in orbit) / len(orbit), 4) if orbit else None,
                  "cycling_ms_per_frame": round(cycle_ms, 4) if orbit else None,
                  "cycling_frames_per_s": round(1000.0 / cycle_ms, 2) if orbit else None},
        "cpu_webgpu_probe": webgpu_probe,
    }
    if args.composite > 0:
        # The display step (not part of frames/s): premultiplied over onto the
        # resident unorm8 framebuffer, without and with the tile heat map. Wall
        # time per call includes the host side of each call; the kernel time is
        # in the rocprof kernel trace of this same command.
        disp = {}
        for mode, label in ((0, "composite"), (1, "composite_heat_map")):
            r.set_debug_mode(mode, 256)
            r.composite(W, H, clear=None)  # allocate / warm
            r.sync()
            t0 = time.perf_counter()
            for _ in range(args.composite):
                r.composite_resident()
            r.sync()
            disp[label] = {"wall_us_per_call": round((time.perf_counter() - t0) * 1e6 / args.composite, 2),
                           "algo_bytes": (24 if mode else 16) * W * H}
        r.set_debug_mode(0, 256)
        out["display"] = disp
    r.close()
    if not args.no_stress and args.workload == 1:
        out["stress"] = bench_stress(30, 3)
    if not args.no_batched:
        out["batched"] = bench_batched(args.batch_steps, 2)
    if not args.no_learned:
        out["learned"] = bench_learned(args.batch_steps, 2, args.learned_epochs)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
