"""Native PLY ingest (libgs3d.so gs3d_ply_parse) against the node loader.

The node loader (webgpu-3dgs_amd/host/SceneLoader.js) restates the reference
loader src/workers/SceneWorker.ts and is itself checked against the Python
activations in test_host_js.py. Here the C++ parser must give the same bits
(positions, colours, scales, rotations, f_dc, f_rest, bounds) and the same error
messages on the same files. The upload entry point gs3d_load_ply is checked on
the GPU against gs3d_set_splats of the parsed arrays.
"""
import base64
import math
import time

import numpy as np
import pytest

from gs3d import _lib as L
from gs3d import camera as cam_mod
from gs3d import ply
from gs3d import scene as S
from test_host_js import NODE, PROPS, f32, raw_columns, run_node, small_raw, write_ply

needs_node = pytest.mark.skipif(NODE is None, reason="node is not installed")


def assert_same_as_node(path, tmp_path, threads=0):
    r = run_node({"cmd": "ply", "path": str(path)}, tmp_path, timeout=300)
    assert r["ok"], r.get("error")
    s, rest, bounds = ply.parse_file(path, threads)
    assert s.n == r["n"] and s.sh_degree == r["degree"]
    for key, got in (("positions", s.positions), ("colors", s.colors), ("scales", s.scales),
                     ("rotations", s.rotations), ("fDc", s.f_dc)):
        assert np.array_equal(got.reshape(-1).view(np.uint32), f32(r[key]).view(np.uint32)), key
    if r["sh"] is None:
        assert rest is None
    else:
        assert np.array_equal(rest.reshape(-1).view(np.uint32), f32(r["sh"]).view(np.uint32))
    if s.n:
        assert list(bounds[0]) + list(bounds[1]) == r["bounds"]
    return s, rest


@needs_node
@pytest.mark.parametrize("fmt", ["ascii", "binary_little_endian", "binary_big_endian"])
@pytest.mark.parametrize("deg", [0, 1, 3])
def test_native_matches_node_loader(tmp_path, fmt, deg):
    raw = small_raw(n=300, deg=deg, seed=11 + deg)
    p = tmp_path / "s.ply"
    write_ply(p, raw, fmt, extra_before=2, extra_after=3)
    assert_same_as_node(p, tmp_path)


@needs_node
def test_native_matches_node_on_many_rows(tmp_path):
    """100k SH3 rows: exp, sigmoid and the 4-norm of the rotation agree bit for bit
    with the JS Math functions over many values, and across thread counts."""
    raw = S.clustered(5, 100_000, 16, sh_degree=3)
    raw.rot[:7] = [[0, 0, 0, 0], [1e-30, 0, 0, 0], [3, 4, 0, 0], [-1, 1, -1, 1], [1e20, 1, 1, 1],
                   [0, 0, 0, -2], [1e-300, 1e-300, 0, 0]]
    raw.opacity_logit[:4] = [80.0, -80.0, 0.0, 745.0]
    p = tmp_path / "big.ply"
    write_ply(p, raw, "binary_little_endian")
    s1, _ = assert_same_as_node(p, tmp_path, threads=1)
    s8, _ = ply.parse_file(p, 8)[:2]
    for a, b in ((s1.positions, s8.positions), (s1.rotations, s8.rotations), (s1.colors, s8.colors),
                 (s1.scales, s8.scales), (s1.sh_rest, s8.sh_rest)):
        assert np.array_equal(a, b)


@needs_node
@pytest.mark.parametrize("fmt", ["ascii", "binary_big_endian"])
def test_node_native_option_matches_js_loader(tmp_path, fmt):
    """SceneLoader({native: true}) (the addon's parsePly) returns what the JS loader returns."""
    raw = small_raw(n=120, deg=3, seed=21)
    p = tmp_path / "s.ply"
    write_ply(p, raw, fmt, extra_before=1, extra_after=2)
    a = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    b = run_node({"cmd": "ply", "path": str(p), "native": True}, tmp_path)
    assert a == b and a["ok"]
    b = p.read_bytes()
    p.write_bytes(b.replace(b"element vertex 120", b"element vertex 120\nproperty list uchar int v"))
    a = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    b = run_node({"cmd": "ply", "path": str(p), "native": True}, tmp_path)
    assert a == b and not a["ok"]


@needs_node
def test_native_f_rest_shuffled_and_double_and_int_properties(tmp_path):
    """f_rest in shuffled header order, and non-float property types (double,
    int16, uchar) in the vertex element, read as the DataView getters would."""
    raw = small_raw(n=90, deg=2, seed=3)
    cols = raw_columns(raw)
    rest = [c for c in cols if c.startswith("f_rest_")]
    names = PROPS[:7] + rest[::-1] + PROPS[7:]
    types = {n: "float" for n in names}
    types["x"] = "double"
    lines = ["ply", "format binary_big_endian 1.0", f"element vertex {raw.n}"]
    lines += [f"property {types[n]} {n}" for n in names]
    lines += ["property short label", "property uchar flag", "end_header"]
    dt = [(n, ">f8" if types[n] == "double" else ">f4") for n in names] + [("label", ">i2"), ("flag", "u1")]
    rec = np.zeros(raw.n, dtype=dt)
    for n in names:
        rec[n] = raw.xyz[:, 0].astype(np.float64) * 1.000000001 if n == "x" else cols[n]
    rec["label"] = -np.arange(raw.n)
    rec["flag"] = np.arange(raw.n) % 250
    p = tmp_path / "mixed.ply"
    p.write_bytes(("\n".join(lines) + "\n").encode() + rec.tobytes())
    s, _ = assert_same_as_node(p, tmp_path)
    assert s.sh_degree == 2


@needs_node
def test_native_two_vertex_elements(tmp_path):
    """A second element named 'vertex': its rows are checked and counted in the
    bounds like the first one's, but only the first element's count is stored."""
    raw = small_raw(n=20, deg=0, seed=8)
    cols = raw_columns(raw)
    head = ["ply", "format binary_little_endian 1.0", f"element vertex {raw.n}"]
    head += [f"property float {n}" for n in PROPS]
    head += ["element vertex 3"] + [f"property float {n}" for n in PROPS] + ["end_header"]
    body = np.stack([cols[n] for n in PROPS], axis=1).astype("<f4")
    extra = body[:3] * 50.0  # far outside the first element's bounds
    p = tmp_path / "two.ply"
    p.write_bytes(("\n".join(head) + "\n").encode() + body.tobytes() + extra.tobytes())
    s, _ = assert_same_as_node(p, tmp_path)
    assert s.n == 20


ERROR_CASES = [
    ("missing", "missing property 'opacity'"),
    ("format", "Unsupported PLY format"),
    ("magic", "must start with `ply`"),
    ("list", "list properties"),
    ("truncated", "outside the bounds"),
    ("noheader", "missing end_header"),
    ("nan_row", "row: missing 'y'"),
    ("bad_type_after", "Unsupported PLY scalar type: float16"),
    ("ascii_token", "Invalid ASCII PLY value 'abc'"),
    ("ascii_short", "Unexpected end of ASCII PLY body"),
    ("count", "Invalid PLY element count for 'vertex'"),
    ("noformat", "missing format line"),
    ("novertex", "missing vertex element"),
]


@needs_node
@pytest.mark.parametrize("case,msg", ERROR_CASES)
def test_native_errors_match_node(tmp_path, case, msg):
    raw = small_raw(n=40)
    p = tmp_path / "s.ply"
    fmt = "ascii" if case.startswith("ascii") else "binary_little_endian"
    write_ply(p, raw, fmt, drop="opacity" if case == "missing" else None)
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
    elif case == "nan_row":
        hb = b.index(b"end_header") + len(b"end_header\n")
        arr = np.frombuffer(b[hb:], "<f4").copy().reshape(40, -1)
        arr[17, 1] = np.nan  # row 17's y
        arr[30, 0] = np.inf  # a later row's x: the earlier row is the one reported
        b = b[:hb] + arr.tobytes()
    elif case == "bad_type_after":
        b = b.replace(b"end_header", b"element extra 2\nproperty float16 h\nend_header") + b"\0" * 8
    elif case == "ascii_token":
        lines = b.split(b"\n")
        hl = lines.index(b"end_header")
        row = lines[hl + 5].split(b" ")
        row[3] = b"abc"
        lines[hl + 5] = b" ".join(row)
        b = b"\n".join(lines)
    elif case == "ascii_short":
        b = b.rstrip()[: -40]
    elif case == "count":
        b = b.replace(b"element vertex 40", b"element vertex -3")
    elif case == "noformat":
        b = b.replace(b"format binary_little_endian 1.0\n", b"")
    elif case == "novertex":
        b = b.replace(b"element vertex", b"element points")
    p.write_bytes(b)
    r = run_node({"cmd": "ply", "path": str(p)}, tmp_path)
    assert not r["ok"] and msg in r["error"], r
    with pytest.raises(ply.PlyError) as ei:
        ply.parse_file(p, 4)
    assert str(ei.value) == r["error"]


def test_native_header_probe_and_empty_body(tmp_path):
    raw = small_raw(n=0)
    p = tmp_path / "e.ply"
    write_ply(p, raw, "binary_little_endian")
    info = ply.probe(p.read_bytes())
    assert info.splat_count == 0 and info.format == L.PLY_BINARY_LE and info.sh_degree == 0
    s, rest, _ = ply.parse_file(p)
    assert s.n == 0 and rest is None


def test_native_activations_match_python_restatement(tmp_path):
    """Against gs3d.scene.activate (numpy f64): the 4-norm there is sqrt(sum of
    squares), not the normalised Kahan sum, so rotations may differ in the last bit."""
    raw = S.clustered(9, 5000, 8, sh_degree=3)
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian")
    s, rest, bounds = ply.parse_file(p)
    ref = S.activate(raw)
    for a, b in ((s.positions, ref.positions), (s.colors, ref.colors), (s.scales, ref.scales),
                 (s.f_dc, ref.f_dc), (s.sh_rest, ref.sh_rest)):
        np.testing.assert_allclose(a, b, rtol=2.5e-7, atol=0)
    np.testing.assert_allclose(s.rotations, ref.rotations, rtol=2.5e-7, atol=1e-30)
    assert np.allclose(bounds[0], raw.xyz.min(axis=0)) and np.allclose(bounds[1], raw.xyz.max(axis=0))


@needs_node
def test_native_parse_is_faster_than_node(tmp_path):
    """1M SH3 rows, binary little endian (248 MB): native parse time vs the node loader."""
    raw = S.clustered(1, 1_000_000, 64, sh_degree=3)
    p = tmp_path / "m.ply"
    write_ply(p, raw, "binary_little_endian")
    r = run_node({"cmd": "plyTime", "path": str(p)}, tmp_path, timeout=600)
    t_node = r["seconds"]
    rn = run_node({"cmd": "plyTime", "path": str(p), "native": True}, tmp_path, timeout=600)
    best = math.inf
    for _ in range(2):
        t = time.perf_counter()
        s, _, _ = ply.parse_file(p)
        best = min(best, time.perf_counter() - t)
    print(f"1M SH3 PLY parse: native {best:.3f} s (through node: {rn['seconds']:.3f} s incl. the file "
          f"read), node loader {t_node:.3f} s, ratio {t_node / best:.1f}x")
    assert s.n == 1_000_000 and r["n"] == 1_000_000 and rn["n"] == 1_000_000
    assert best < t_node


@pytest.mark.gpu
def test_load_ply_uploads_the_parsed_scene(renderer, tmp_path):
    """gs3d_load_ply (mmap, parse into pinned memory, upload) renders the same
    frame as gs3d_set_splats of the parsed arrays; 48 f_rest (16 per channel)
    keep the first 15 of each channel for SH degree 3."""
    raw = S.clustered(4, 30_000, 8, sh_degree=3)
    raw.f_rest = np.concatenate([raw.f_rest.reshape(-1, 3, 15),
                                 np.full((raw.n, 3, 1), 7.0, np.float32)], axis=2).reshape(raw.n, 48)
    p = tmp_path / "s.ply"
    write_ply(p, raw, "binary_little_endian")
    W, H = 640, 360
    cu = cam_mod.view_camera(S.orbit_eye(8.0, 30.0, 15.0), W, H)
    info = renderer.load_ply(p)
    assert info.splat_count == raw.n and info.sh_degree == 3 and info.f_rest_count == 48
    renderer.set_tile_size(16, 16)
    renderer.rebin(cu, W, H)
    img_a = renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16)
    proj_a = renderer.read(L.BUF_PROJECTED_SPLATS, np.uint32)
    s, _, _ = ply.parse_file(p)
    renderer.set_splat_data(s)
    renderer.rebin(cu, W, H)
    assert np.array_equal(renderer.read(L.BUF_PROJECTED_SPLATS, np.uint32), proj_a)
    assert np.array_equal(renderer.read(L.BUF_IMAGE_RGBA16F, np.uint16), img_a)
    raw.f_rest = raw.f_rest.reshape(-1, 3, 16)[:, :, :15].reshape(raw.n, 45)
    ref = S.activate(raw)
    assert np.array_equal(s.sh_rest, ref.sh_rest)
