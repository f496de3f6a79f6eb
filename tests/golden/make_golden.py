"""Write the golden fixtures: small rebins computed by the CPU oracle.

Each fixture holds the inputs (activated splat arrays, the 288-B camera
uniform, viewport, tile size) and every output of the rebin: projected splats,
counts, offsets, emitted and sorted keys and values, tile ranges, the image as
binary16 bits. tests/test_golden.py re-runs the oracle on the inputs (CPU) and
the HIP path (GPU) and compares bit for bit, so a change of the oracle build or
of the kernels shows against data that neither produced in that run.

    python tests/golden/make_golden.py        # rewrites tests/golden/*.npz
"""
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parents[1] / "webgpu-3dgs_amd"))

import oracle_lib as O  # noqa: E402
from gs3d import camera, scene  # noqa: E402

# name: (splat generator, eye, width, height, tile)
CASES = {
    "sh0_64splats_64x48_tile8": (lambda: scene.uniform_cube(101, 64, log_scale_mean=-2.5), (0.0, 0.0, 3.0), 64, 48, 8),
    "sh3_300splats_96x64_tile16": (lambda: scene.clustered(102, 300, 3, log_scale_mean=-2.0, sh_degree=3),
                                   (7.0, 3.0, 10.0), 96, 64, 16),
}


def fixture(gen, eye, W, H, tile):
    s = scene.activate(gen())
    cam = camera.view_camera(eye, W, H)
    ref = O.pipeline_f32(cam, s, W, H, tile)
    out = dict(camera=cam, viewport=np.array([W, H, tile], np.uint32),
               positions=s.positions, scales=s.scales, rotations=s.rotations, colors=s.colors,
               sh_degree=np.array(s.sh_degree, np.int32))
    if s.sh_degree:
        out.update(f_dc=s.f_dc, sh_rest=s.sh_rest)
    for k in ("projected", "counts", "offsets", "keys", "values", "sorted_keys", "sorted_values", "tile_offsets"):
        out[k] = ref[k]
    out["image_f16"] = O.f16_round(ref["image"])
    return out


def main():
    O.set_threads(1)
    for name, (gen, eye, W, H, tile) in CASES.items():
        f = fixture(gen, eye, W, H, tile)
        np.savez_compressed(HERE / f"{name}.npz", **f)
        print(f"{name}: K={f['sorted_keys'].size}")


if __name__ == "__main__":
    main()
