"""Node host (webgpu-3dgs_amd/host) checks, driven through tests/js/host_check.js.

CPU tests: the PLY loader (ascii, binary little/big endian, extra elements,
f_rest ordering, SH degree, header errors) against the Python restatement of
the loader's activations (SceneWorker.ts:407-464); the Camera uniform against
the textbook look-at / perspective closed forms (Camera.ts:69-191) and the
Python mirror; WorkgroupManager layouts against the library's launch geometry;
PipelineManager variant validation. The GPU tests render through
GaussianSplatRenderer.render and renderBatch and compare with the oracle
pipeline, and read the Profiler's stage and reference-pass tables.
"""
import base64
import json
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from gs3d import camera as C
from gs3d import scene as S

ROOT = Path(__file__).resolve().parents[1]
DRIVER = ROOT / "tests" / "js" / "host_check.js"
ADDON = ROOT / "webgpu-3dgs_amd" / "addon" / "gs3d_napi.node"
NODE = shutil.which("node")

pytestmark = pytest.mark.skipif(NODE is None, reason="node is not installed")


def run_node(req: dict, tmp_path: Path, timeout: float = 120.0) -> dict:
    f = tmp_path / "req.json"
    f.write_text(json.dumps(req))
    p = subprocess.run([NODE, str(DRIVER), str(f)], capture_output=True, text=True, timeout=timeout)
    assert p.returncode == 0, p.stderr[-3000:]
    return json.loads(p.stdout)


def f32(s):
    return None if s is None else np.frombuffer(base64.b64decode(s), np.float32)


# ---------------------------------------------------------------- PLY writer

PROPS = ["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2", "opacity", "scale_0", "scale_1", "scale_2",
         "rot_0", "rot_1", "rot_2", "rot_3"]


def raw_columns(raw: S.RawSplats) -> dict:
    cols = {
        "x": raw.xyz[:, 0], "y": raw.xyz[:, 1], "z": raw.xyz[:, 2],
        "f_dc_0": raw.f_dc[:, 0], "f_dc_1": raw.f_dc[:, 1], "f_dc_2": raw.f_dc[:, 2],
        "opacity": raw.opacity_logit,
        "scale_0": raw.log_scale[:, 0], "scale_1": raw.log_scale[:, 1], "scale_2": raw.log_scale[:, 2],
        "rot_0": raw.rot[:, 0], "rot_1": raw.rot[:, 1], "rot_2": raw.rot[:, 2], "rot_3": raw.rot[:, 3],
    }
    if raw.f_rest is not None:
        for k in range(raw.f_rest.shape[1]):
            cols[f"f_rest_{k}"] = raw.f_rest[:, k]
    return {k: np.asarray(v, np.float32) for k, v in cols.items()}


def write_ply(path: Path, raw: S.RawSplats, fmt: str, order=None, extra_before=0, extra_after=0,
              drop=None):
    """A Gaussian-splat PLY with float properties in `order` (default: PROPS
    then f_rest_k ascending). Extra elements ('camera' with 2 floats) before or
    after the vertex element exercise the element walk."""
    cols = raw_columns(raw)
    names = order or (PROPS + [k for k in cols if k.startswith("f_rest_")])
    if drop:
        names = [n for n in names if n != drop]
    lines = ["ply", f"format {fmt} 1.0", "comment written by the host tests"]
    if extra_before:
        lines += [f"element camera {extra_before}", "property float cx", "property float cy"]
    lines += [f"element vertex {raw.n}"] + [f"property float {n}" for n in names]
    if extra_after:
        lines += [f"element extra {extra_after}", "property uchar flag", "property double weight"]
    lines.append("end_header")
    head = ("\n".join(lines) + "\n").encode()
    vert = np.stack([cols[n] for n in names], axis=1) if names else np.zeros((raw.n, 0), np.float32)
    before = np.arange(extra_before * 2, dtype=np.float32).reshape(extra_before, 2) + 0.25
    if fmt == "ascii":
        body = []
        body += [" ".join(repr(float(v)) for v in row) for row in before]
        # repr of the f64 widening: the loader parses ascii tokens as f64, so the
        # shortest f32 string would denote a different number than the f32 value
        body += [" ".join(repr(float(v)) for v in row) for row in vert]
        body += [f"{i % 7} {i * 0.5}" for i in range(extra_after)]
        path.write_bytes(head + ("\n".join(body) + "\n").encode())
        return
    end = "<" if fmt == "binary_little_endian" else ">"
    data = before.astype(end + "f4").tobytes() + vert.astype(end + "f4").tobytes()
    rec = np.zeros(extra_after, dtype=[("flag", "u1"), ("weight", end + "f8")])
    rec["flag"] = np.arange(extra_after) % 7
    rec["weight"] = np.arange(extra_after) * 0.5
    path.write_bytes(head + data + rec.tobytes())


def check_loaded(r: dict, raw: S.RawSplats):
    ref = S.activate(raw)
    assert r["ok"], r.get("error")
    assert r["n"] == raw.n
    assert r["degree"] == raw.sh_degree
    # f64 activations stored as f32: allow the last f32 bit for exp/sigmoid
    # evaluated by a different libm.
    for key, want in (("positions", ref.positions), ("colors", ref.colors), ("scales", ref.scales),
                      ("rotations", ref.rotations)):
        got = f32(r[key]).reshape(-1, 4)
        np.testing.assert_allclose(got, want, rtol=2.5e-7, atol=0, err_msg=key)
    np.testing.assert_array_equal(f32(r["fDc"]).reshape(-1, 3), raw.f_dc.astype(np.float32))
    if raw.f_rest is None:
        assert r["sh"] is None
    else:
        np.testing.assert_array_equal(f32(r["sh"]).reshape(raw.n, -1), raw.f_rest.astype(np.float32))
    lo, hi = raw.xyz.min(axis=0), raw.xyz.max(axis=0)
    np.testing.assert_array_equal(np.float32(r["bounds"][:3]), lo.astype(np.float32))
    np.testing.assert_array_equal(np.float32(r["bounds"][3:]), hi.astype(np.float32))


def small_raw(n=257, deg=0, seed=5) -> S.RawSplats:
    raw = S.uniform_cube(seed, n, sh_degree=deg)
    if n >= 2:
        raw.rot[0] = 0.0  # zero quaternion: the loader keeps it as is (inv = 1)
        raw.opacity_logit[1] = 80.0  # sigmoid saturates to 1
    return raw


@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("deg", [0, 3])
def test_ply_formats(tmp_path, fmt, deg):
    raw = small_raw(deg=deg)
    p = tmp_path / "s.ply"
    write_ply(p, raw, fmt, extra_before=2, extra_after=3)
    check_loaded(run_node({"cmd": "ply", "path": str(p)}, tmp_path), raw)


def test_ply_in_caller_thread_matches_worker(tmp_path):
    raw = small_raw(n=64, deg=1)
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian")
    a = run_node({"cmd": "ply", "path": str(p), "worker": True}, tmp_path)
    b = run_node({"cmd": "ply", "path": str(p), "worker": False}, tmp_path)
    assert a == b
    check_loaded(a, raw)


def test_ply_f_rest_ordered_by_index(tmp_path):
    """f_rest columns in shuffled header order are read back by their index
    (SceneWorker.ts:375-388), so f_rest_10 is not placed after f_rest_1."""
    raw = small_raw(n=50, deg=2)
    cols = list(raw_columns(raw))
    rest = [c for c in cols if c.startswith("f_rest_")]
    order = PROPS[:7] + rest[::-1] + PROPS[7:]
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian", order=order)
    r = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    check_loaded(r, raw)
    assert r["degree"] == 2


@pytest.mark.parametrize("n_rest,deg", [(0, 0), (8, 0), (9, 1), (23, 1), (24, 2), (44, 2), (45, 3)])
def test_ply_sh_degree_from_rest_count(tmp_path, n_rest, deg):
    raw = small_raw(n=8, deg=3)
    raw.f_rest = raw.f_rest[:, :n_rest] if n_rest else None
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian")
    r = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    assert r["degree"] == deg
    if n_rest:
        assert f32(r["sh"]).size == 8 * n_rest


def test_ply_empty_vertex_element(tmp_path):
    raw = small_raw(n=0)
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian")
    r = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    assert r["ok"] and r["n"] == 0


@pytest.mark.parametrize("case,msg", [
    ("missing", "missing property 'opacity'"),
    ("format", "Unsupported PLY format"),
    ("magic", "must start with `ply`"),
    ("list", "list properties"),
    ("truncated", "outside the bounds"),
    ("noheader", "missing end_header"),
])
def test_ply_errors(tmp_path, case, msg):
    raw = small_raw(n=4)
    p = tmp_path / "s.ply"
    if case == "missing":
        write_ply(p, raw, "binary_little_endian", drop="opacity")
    else:
        write_ply(p, raw, "binary_little_endian")
        b = p.read_bytes()
        if case == "format":
            b = b.replace(b"binary_little_endian", b"binary_middle_endian")
        elif case == "magic":
            b = b"plx" + b[3:]
        elif case == "list":
            b = b.replace(b"end_header", b"property list uchar int vertex_indices\nend_header")
        elif case == "truncated":
            b = b[:-5]
        elif case == "noheader":
            b = b.replace(b"end_header", b"end_headr")
        p.write_bytes(b)
    r = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    assert not r["ok"]
    assert msg in r["error"]


# ---------------------------------------------------------------- camera

def look_at_closed_form(eye, center, up):
    """Column-major right-handed look-at: rows of R are (s, u, -f), t = -R eye."""
    eye, center, up = (np.asarray(v, np.float64) for v in (eye, center, up))
    f = center - eye
    f /= np.linalg.norm(f)
    s = np.cross(f, up)
    s /= np.linalg.norm(s)
    u = np.cross(s, f)
    m = np.eye(4)
    m[0, :3], m[1, :3], m[2, :3] = s, u, -f
    m[:3, 3] = -m[:3, :3] @ eye
    return m.T.reshape(-1)  # column-major storage


def perspective_closed_form(fovy, aspect, near, far):
    """OpenGL projection, clip z in [-1, 1]."""
    t = 1.0 / math.tan(fovy / 2)
    m = np.zeros((4, 4))
    m[0, 0] = t / aspect
    m[1, 1] = t
    m[2, 2] = (far + near) / (near - far)
    m[2, 3] = 2 * far * near / (near - far)
    m[3, 2] = -1.0
    return m.T.reshape(-1)


CAMS = [
    dict(position=[10, 10, 10], target=[0, 0, 0], up=[0, 1, 0], fov=math.pi / 2, aspect=1.0, near=0.1,
         far=10000, width=800, height=800),
    dict(position=[0.5, -2.0, 7.25], target=[0.1, 0.2, -0.3], up=[0, 1, 0], fov=math.pi / 3,
         aspect=1920 / 1080, near=0.1, far=10000, width=1920, height=1080),
    dict(position=[-3, 1, -4], target=[1, 0.5, 2], up=[0, 1, 0], fov=1.0, aspect=0.75, near=0.5,
         far=200, width=600, height=800),
]


def test_camera_uniform_closed_forms(tmp_path):
    r = run_node({"cmd": "camera", "cameras": CAMS}, tmp_path)
    for c, out in zip(CAMS, r["cameras"]):
        u = f32(out["uniform"]).astype(np.float64)
        assert u.size == 72
        view = look_at_closed_form(c["position"], c["target"], c["up"])
        proj = perspective_closed_form(c["fov"], c["aspect"], c["near"], c["far"])
        np.testing.assert_allclose(u[0:16], view, rtol=0, atol=2e-6 * max(1, np.abs(view).max()))
        np.testing.assert_allclose(u[16:32], proj, rtol=3e-7, atol=1e-12)
        iv = np.linalg.inv(view.reshape(4, 4).T).T.reshape(-1)
        ip = np.linalg.inv(proj.reshape(4, 4).T).T.reshape(-1)
        np.testing.assert_allclose(u[32:48], iv, rtol=0, atol=2e-6 * max(1, np.abs(iv).max()))
        np.testing.assert_allclose(u[48:64], ip, rtol=1e-6, atol=1e-6)
        np.testing.assert_array_equal(u[64:68], np.float32(c["position"] + [1.0]))
        assert u[68] == c["width"] and u[69] == c["height"]
        assert u[70] == np.float32(c["width"] / c["height"]) and u[71] == 0


def test_camera_matches_python_mirror(tmp_path):
    """The JS Camera and gs3d.camera produce the same 72 floats, up to f32
    rounding of the inverses (the two hosts invert with different formulas)."""
    r = run_node({"cmd": "camera", "cameras": CAMS}, tmp_path)
    for c, out in zip(CAMS, r["cameras"]):
        py = C.Camera(c["position"], c["target"], c["up"], c["fov"], c["aspect"], c["near"], c["far"])
        u, v = f32(out["uniform"]), py.uniform_data(c["width"], c["height"])
        np.testing.assert_array_equal(u[:32], v[:32])
        np.testing.assert_array_equal(u[64:], v[64:])
        np.testing.assert_allclose(u[32:64], v[32:64], rtol=1e-6, atol=1e-6)


def test_camera_setters(tmp_path):
    """setOrbit puts the camera at the orbit position gs3d.scene.orbit_eye gives (the
    workloads' camera placement) and its uniform equals the Python mirror's; setPosition,
    setTarget and setAspect change the view or projection; each bumps the version."""
    base = dict(CAMS[1], target=[0.0, 0.0, 0.0])
    cams = [dict(base, ops=[["setOrbit", 8.0, 45.0, 20.0]]), dict(base, ops=[["setPosition", [1, 2, 3]]]),
            dict(base, ops=[["setTarget", [0.5, 0.5, -1.0]]]), dict(base, ops=[["setAspect", 2.0]])]
    r = run_node({"cmd": "camera", "cameras": cams}, tmp_path)["cameras"]
    eye = S.orbit_eye(8.0, 45.0, 20.0)
    np.testing.assert_allclose(r[0]["position"], eye, rtol=1e-6, atol=1e-6)
    py = C.Camera(r[0]["position"], [0, 0, 0], base["up"], base["fov"], base["aspect"], base["near"], base["far"])
    u, v = f32(r[0]["uniform"]), py.uniform_data(base["width"], base["height"])
    np.testing.assert_array_equal(u[:32], v[:32])
    np.testing.assert_allclose(u[32:64], v[32:64], rtol=1e-6, atol=1e-6)
    np.testing.assert_allclose(r[1]["position"], [1, 2, 3])
    view = look_at_closed_form([1, 2, 3], [0, 0, 0], base["up"])
    np.testing.assert_allclose(f32(r[1]["uniform"])[0:16], view, rtol=0, atol=2e-6 * max(1, np.abs(view).max()))
    np.testing.assert_allclose(r[2]["target"], [0.5, 0.5, -1.0])
    for out in r:
        assert out["versionDelta"] >= 1
    u = f32(r[3]["uniform"]).astype(np.float64)
    np.testing.assert_allclose(u[16:32], perspective_closed_form(base["fov"], 2.0, base["near"], base["far"]),
                               rtol=3e-7, atol=1e-12)


# ---------------------------------------------------------------- managers

@pytest.mark.skipif(not ADDON.exists(), reason="addon not built")
def test_workgroup_layouts_are_the_library_launches(tmp_path):
    """WorkgroupManager reports what libgs3d launches (gs3d_launch_geometry, the same
    shape functions the launchers size their grids with): the JS layouts equal the C
    ABI's for every pass, follow problem-size updates, and an unknown pass throws."""
    from gs3d import _lib as L

    cases = [
        dict(name="pre", kernel="preprocess-splats", problemSize=[1_000_000]),
        dict(name="hist", kernel="radix-histogram", problemSize=[1_000_000]),
        dict(name="depth", kernel="radix-sort-splats", problemSize=[1_000_000]),
        dict(name="scan", kernel="prefix-scan", problemSize=[1_000_000]),
        dict(name="emit", kernel="emit-tile-refs", problemSize=[1_000_000]),
        dict(name="sort", kernel="radix-sort", problemSize=[3_615_659]),
        dict(name="ident", kernel="identify-tile-ranges", problemSize=[1920, 1080]),
        dict(name="raster", kernel="rasterize-tiles", problemSize=[1920, 1080]),
        dict(name="comp", kernel="composite-splats", problemSize=[1920, 1080]),
        dict(name="raster", problemSize=[640, 360], update=True),
        dict(name="bad", kernel="no-such-pass", problemSize=[1]),
        dict(name="never", problemSize=[5], update=True),
    ]
    r = run_node({"cmd": "workgroups", "cases": cases}, tmp_path)
    out = r["layouts"]
    for c, lay in zip(cases[:9], out[:9]):
        wg, grid = L.launch_geometry(c["kernel"], c["problemSize"])
        assert lay == {"workgroupSize": list(wg), "dispatchSize": list(grid)}, c
    # the 1M-splat 1080p frame: thread per splat, 8,192-key depth tiles, 8,192 counts per
    # scan workgroup, 16,384-key tile-id tiles, one 16x16 workgroup per tile
    assert out[0] == {"workgroupSize": [256, 1, 1], "dispatchSize": [3907, 1, 1]}
    assert out[2]["dispatchSize"][0] == 123 and out[3]["dispatchSize"][0] == 123
    assert out[5] == {"workgroupSize": [1024, 1, 1], "dispatchSize": [221, 1, 1]}
    assert out[7] == {"workgroupSize": [256, 1, 1], "dispatchSize": [8160, 1, 1]}
    assert out[9]["dispatchSize"] == [40 * 23, 1, 1]
    assert "no pass named 'no-such-pass'" in out[10]["error"]
    assert "No workgroup layout registered for 'never'" in out[11]["error"]
    assert r["names"] == ["pre", "hist", "depth", "scan", "emit", "sort", "ident", "raster", "comp"]


def test_pipeline_variants(tmp_path):
    cases = [
        dict(name="radix-sort", codeConstants={"RADIX_BITS": 8}),
        dict(name="radix-sort", codeConstants={"RADIX_BITS": 4}, update=True),
        dict(name="radix-sort", codeConstants={"RADIX_BITS": 5}, update=True),
        dict(name="rasterize-tiles", codeConstants={"TILE_X": 16, "TILE_Y": 16, "TILE_AREA": 256}),
        dict(name="rasterize-tiles", codeConstants={"PIXELS_PER_LANE": 2}, update=True),
        dict(name="rasterize-tiles", codeConstants={"TILE_X": 8, "TILE_Y": 16}, update=True),
        dict(name="no-such-kernel", codeConstants={}),
        dict(name="prefix-scan", codeConstants={"FOO": 1}),
    ]
    r = run_node({"cmd": "pipelines", "cases": cases}, tmp_path)["results"]
    assert r[0] == {"ok": True, "constants": {"RADIX_BITS": 8}}
    assert r[1] == {"ok": True, "constants": {"RADIX_BITS": 4}}
    assert not r[2]["ok"] and "RADIX_BITS=5" in r[2]["error"]
    assert r[3]["ok"]
    assert r[4] == {"ok": True, "constants": {"TILE_X": 16, "TILE_Y": 16, "TILE_AREA": 256, "PIXELS_PER_LANE": 2}}
    assert not r[5]["ok"] and "square" in r[5]["error"]
    assert not r[6]["ok"] and "No compiled kernel" in r[6]["error"]
    assert not r[7]["ok"]


@pytest.mark.skipif(not ADDON.exists(), reason="addon not built")
def test_addon_exports(tmp_path):
    """The addon loads and exports one function per bound C entry point."""
    r = run_node({"cmd": "addonExports"}, tmp_path)
    want = {"create", "destroy", "lastError", "sync", "deviceLimits", "setSplats", "setTileSize",
            "getTileSize", "getTileCount", "rebin", "refCount", "read", "readTileCounts", "stageTimes",
            "stageTimesAvg", "profilerReset", "stageLabel", "bufferCreate", "bufferResize", "bufferSize",
            "bufferWrite", "bufferRead", "bufferDestroy", "bufferList", "setKernelVariant",
            "getKernelVariant", "setProfileMask", "sortPasses", "setDebugMode", "composite",
            "parsePly", "loadPly", "rebinBatch", "batchRefCounts", "batchRead", "learnedInit",
            "learnedTrain", "learnedValidate", "learnedEnable", "launchGeometry", "referencePassStages",
            "abiVersion"}
    assert set(r["names"]) == want
    assert r["abiVersion"] >= 1


@pytest.mark.skipif(not ADDON.exists(), reason="addon not built")
def test_addon_rejects_wrong_argument_types_cpu(tmp_path):
    """Entry points that need no device reject wrong-typed arguments with a TypeError."""
    r = run_node({"cmd": "argErrors", "noGpu": True}, tmp_path)["results"]
    for name in ("stageLabel", "parsePly"):
        assert r[name]["threw"] and r[name]["type"] == "TypeError", (name, r[name])


@pytest.mark.gpu
def test_addon_rejects_wrong_argument_types(tmp_path):
    """A wrong-typed argument throws a TypeError naming the call; nothing is silently
    converted to 0 (which would, e.g., read view 0 or create an empty buffer)."""
    r = run_node({"cmd": "argErrors"}, tmp_path, timeout=300)["results"]
    for name in ("stageLabel", "parsePly", "setTileSize", "rebin", "rebinFlags", "read", "batchRead",
                 "bufferCreate", "bufferCreateName", "bufferRead", "bufferWrite", "bufferResize",
                 "learnedEnable", "learnedTrain", "loadPly"):
        assert r[name]["threw"] and r[name]["type"] == "TypeError", (name, r[name])
    assert "batchRead(ctx, view, bufferId)" in r["batchRead"]["message"]
    assert "buffer name" in r["bufferCreateName"]["message"]
    assert r["bufferSizeOk"] == 64


# ---------------------------------------------------------------- GPU render through node

@pytest.mark.gpu
@pytest.mark.parametrize("tile,ppl", [(16, 1), (8, 1), (32, 2)])
def test_node_render_matches_oracle(tmp_path, tile, ppl):
    import oracle_lib as O

    raw = S.uniform_cube(11, 20_000, sh_degree=3)
    p = tmp_path / "scene.ply"
    write_ply(p, raw, "binary_little_endian")
    W, H = 320, 240
    eye = [3.0, 2.0, 4.0]
    r = run_node({"cmd": "render", "path": str(p), "tile": tile, "pixelsPerLane": ppl, "eye": eye,
                  "width": W, "height": H}, tmp_path, timeout=300)
    assert r["decisions"] == [True, False, True]  # the rebin gate: unchanged camera skips the rebin
    assert r["rebinCount"] == 2
    assert r["tileSize"] == [tile, tile]
    assert r["variant"] == ppl
    s = S.activate(raw)
    cam = f32(r["uniform"])
    ref = O.pipeline_f32(cam, s, W, H, tile)
    assert r["refCount"] == ref["k"]
    img = np.frombuffer(base64.b64decode(r["image"]), np.float16).astype(np.float32).reshape(H, W, 4)
    err = np.abs(img - ref["image"].reshape(H, W, 4)).max()
    assert err <= 1.0 / 255.0, err
    counts = np.frombuffer(base64.b64decode(r["tileCounts"]), np.uint32)
    np.testing.assert_array_equal(counts, np.diff(ref["tile_offsets"].astype(np.int64)).clip(0))
    # render() with a colour target: heat map (debug mode 1) + composite over the cleared target
    assert r["compositeRebinned"] is False
    fb = np.frombuffer(base64.b64decode(r["framebuffer"]), np.uint8).reshape(H, W, 4)
    heat = O.tile_overlay(img, ref["tile_offsets"], tile, W, H, r["debugRef"])
    np.testing.assert_array_equal(fb, O.composite(heat, W, H, clear=(0.12, 0.12, 0.13, 1.0)))


@pytest.mark.gpu
def test_node_render_batch_equals_single_views(tmp_path):
    """GaussianSplatRenderer.renderBatch (one hipGraph over all views) gives each view the
    image and key count of render() alone; the learned-colour calls run through the addon."""
    raw = S.clustered(13, 20_000, 6, sh_degree=3)
    p = tmp_path / "scene.ply"
    write_ply(p, raw, "binary_little_endian")
    views = [{"eye": [3.0, 2.0, 4.0], "width": 320, "height": 180},
             {"eye": [-4.0, 1.0, 3.0], "width": 256, "height": 256},
             {"eye": [0.5, -3.0, 4.5], "width": 200, "height": 120}]
    r = run_node({"cmd": "batch", "path": str(p), "views": views}, tmp_path, timeout=300)
    assert r["ok"]
    for i, sv in enumerate(r["single"]):
        assert r["ks"][i] == sv["k"]
        assert r["batchImages"][i] == sv["image"]
    assert len(r["losses"]) == 2 and r["losses"][1] < r["losses"][0]
    assert r["validation"]["enabled"] == (r["validation"]["psnr"] >= 45)


@pytest.mark.gpu
def test_node_profiler_table(tmp_path):
    """The Profiler over profiled frames with their display pass: every stage label in
    execution order with a time > 0, the stages adding up to the whole rebin (each event
    pair adds a dispatch gap, so their sum is a little below it), the reference's 13 pass
    labels each in one row of the reference table, and the launch layouts of the frame."""
    raw = S.clustered(14, 200_000, 16, sh_degree=3)
    p = tmp_path / "scene.ply"
    write_ply(p, raw, "binary_little_endian")
    r = run_node({"cmd": "profile", "path": str(p), "eye": [5.0, 2.0, 6.0], "width": 960, "height": 540,
                  "frames": 12}, tmp_path, timeout=300)
    t = r["timings"]
    from gs3d import _lib as L

    assert list(t) == list(L.STAGES)
    assert all(t[s] > 0 for s in L.STAGES), t
    parts = sum(t[s] for s in L.REBIN_STAGES)
    assert 0.5 * t["rebin-total"] <= parts <= 1.02 * t["rebin-total"], t
    ref = r["reference"]
    assert [lab for row in ref for lab in row["passes"]] == list(L.REFERENCE_PASSES)
    by_stage = {row["stages"]: row for row in ref}
    assert by_stage["prefix-scan"]["passes"] == ["prefix-scan-local", "prefix-scan-blocks", "prefix-scan-add"]
    assert by_stage["radix-sort-splats+radix-sort"]["passes"] == ["radix-histogram", "scan-local", "scan-add",
                                                                   "radix-scatter"]
    assert abs(by_stage["radix-sort-splats+radix-sort"]["us"] - t["radix-sort-splats"] - t["radix-sort"]) < 1e-3
    assert by_stage["composite-splats"]["passes"] == ["tile-overlay", "composite-splats"]
    lay = r["layouts"]
    assert lay["rasterize-tiles"] == {"workgroupSize": [256, 1, 1], "dispatchSize": [60 * 34, 1, 1]}
    assert lay["radix-sort"]["dispatchSize"][0] == -(-r["refCount"] // 16384)
