"""Pin the CPU oracle against closed forms, library routines and brute force.

Every oracle function is checked by something that does not restate it:
numpy cumsum / lexsort / searchsorted / float16, scipy's spherical harmonics
and rotations, finite-difference Jacobians, analytic compositing values.
"""
import math

import numpy as np
import pytest
from scipy.spatial.transform import Rotation
from scipy.special import sph_harm_y

import oracle_lib as O
from gs3d import camera as cam_mod
from gs3d import scene as sc

TWO32 = 1 << 32


# ---------------------------------------------------------------- scan ----
@pytest.mark.parametrize("n", [0, 1, 7, 2048, 2049, 100_003])
def test_scan_matches_cumsum(n):
    rng = np.random.default_rng(n)
    counts = rng.integers(0, 40, n, dtype=np.uint32)
    counts[rng.random(n) < 0.3] = 0
    offs, tot = O.scan(counts)
    ref = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))]) % TWO32
    assert np.array_equal(offs, ref.astype(np.uint32))
    assert tot == int(ref[-1])


def test_scan_wraps_mod_2_32():
    counts = np.array([0xFFFFFFF0, 0x20, 5, 0xFFFFFFFF], np.uint32)
    offs, tot = O.scan(counts)
    ref = np.concatenate([[0], np.cumsum(counts.astype(np.uint64))]) % TWO32
    assert np.array_equal(offs, ref.astype(np.uint32)) and tot == int(ref[-1])


# ---------------------------------------------------------------- emit ----
def random_projected(n, tiles_x, tiles_y, seed=0, p_valid=0.8):
    rng = np.random.default_rng(seed)
    proj = np.zeros((n, 16), np.float32)
    u = proj.view(np.uint32)
    valid = rng.random(n) < p_valid
    x0 = rng.integers(0, tiles_x, n)
    y0 = rng.integers(0, tiles_y, n)
    x1 = np.minimum(x0 + rng.integers(0, 4, n), tiles_x - 1)
    y1 = np.minimum(y0 + rng.integers(0, 3, n), tiles_y - 1)
    proj[:, 2] = -rng.uniform(0.3, 30.0, n).astype(np.float32)  # view z < 0
    u[:, 11] = valid
    u[:, 12], u[:, 13], u[:, 14], u[:, 15] = x0, y0, x1, y1
    counts = np.where(valid, (x1 - x0 + 1) * (y1 - y0 + 1), 0).astype(np.uint32)
    return proj, counts


def test_emit_matches_enumeration():
    tx, ty = 9, 5
    proj, counts = random_projected(300, tx, ty, seed=3)
    offs, k = O.scan(counts)
    keys, vals = O.emit(proj, offs, tx)
    u = proj.view(np.uint32)
    exp_k, exp_v = [], []
    for i in range(300):
        if not u[i, 11]:
            continue
        lo = int(np.float32(-proj[i, 2]).view(np.uint32))
        for yy in range(u[i, 13], u[i, 15] + 1):  # row-major: y outer, x inner
            for xx in range(u[i, 12], u[i, 14] + 1):
                exp_k.append(((yy * tx + xx) << 32) | lo)
                exp_v.append(i)
    assert k == len(exp_k)
    assert np.array_equal(keys, np.array(exp_k, np.uint64))
    assert np.array_equal(vals, np.array(exp_v, np.uint32))


# ---------------------------------------------------------------- sort ----
@pytest.mark.parametrize("bits", [4, 8])
@pytest.mark.parametrize("n", [0, 1, 5, 1000, 70_000])
def test_sort_is_stable_lexsort(bits, n):
    rng = np.random.default_rng(n + bits)
    keys = (rng.integers(0, 50, n, dtype=np.uint64) << np.uint64(32)) | rng.integers(
        0, 1 << 32, n, dtype=np.uint64)
    keys[rng.random(n) < 0.4] = keys[0] if n else 0  # many duplicates
    vals = np.arange(n, dtype=np.uint32)
    sk, sv = O.radix_sort(keys, vals, bits)
    order = np.lexsort((vals, keys))
    assert np.array_equal(sk, keys[order]) and np.array_equal(sv, vals[order])


def test_sort_brute_force_tiny():
    rng = np.random.default_rng(9)
    keys = rng.integers(0, 6, 40, dtype=np.uint64) * np.uint64(0x1_0000_0001)
    vals = rng.integers(0, 1000, 40, dtype=np.uint32)
    sk, sv = O.radix_sort(keys, vals)
    pairs = sorted(range(40), key=lambda i: (int(keys[i]), i))
    assert [int(x) for x in sk] == [int(keys[i]) for i in pairs]
    assert [int(x) for x in sv] == [int(vals[i]) for i in pairs]


def test_sort_threads_deterministic():
    rng = np.random.default_rng(4)
    keys = rng.integers(0, 1 << 40, 200_000, dtype=np.uint64)
    vals = np.arange(keys.size, dtype=np.uint32)
    O.set_threads(1)
    a = O.radix_sort(keys, vals)
    O.set_threads(4)
    b = O.radix_sort(keys, vals)
    O.set_threads(1)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ------------------------------------------------------------ identify ----
def identify_closed_form(sorted_keys, tile_count):
    """offsets[t] = lower_bound(hi, t) for t <= last + 1, 0 after; all 0 if empty."""
    out = np.zeros(tile_count + 1, np.uint32)
    if sorted_keys.size == 0:
        return out
    hi = (sorted_keys >> np.uint64(32)).astype(np.int64)
    last = int(hi[-1])
    t = np.arange(0, min(last + 1, tile_count) + 1)
    out[t] = np.searchsorted(hi, t, side="left")
    return out


@pytest.mark.parametrize("case", ["empty", "one", "dense", "sparse", "last_tile", "gap_start"])
def test_identify_closed_form(case):
    T = 40
    rng = np.random.default_rng(len(case))
    if case == "empty":
        tiles = np.array([], np.int64)
    elif case == "one":
        tiles = np.array([7])
    elif case == "dense":
        tiles = rng.integers(0, T, 500)
    elif case == "sparse":
        tiles = rng.choice([3, 4, 17, 30], 50)
    elif case == "last_tile":
        tiles = rng.choice([0, T - 1], 20)
    else:
        tiles = rng.choice([25, 26], 10)
    keys = np.sort((tiles.astype(np.uint64) << np.uint64(32)) | rng.integers(
        0, 1 << 31, tiles.size, dtype=np.uint64))
    got = O.identify(keys, T)
    assert np.array_equal(got, identify_closed_form(keys, T))
    # per-tile counts up to the last populated tile
    if keys.size:
        last = int(keys[-1] >> np.uint64(32))
        cnt = np.bincount(tiles, minlength=T)
        assert np.array_equal(np.diff(got.astype(np.int64))[: last + 1], cnt[: last + 1])


# ---------------------------------------------------------------- fp16 ----
def test_f16_matches_numpy_rne():
    rng = np.random.default_rng(0)
    x = np.concatenate([
        rng.standard_normal(20000).astype(np.float32),
        rng.uniform(0, 1, 20000).astype(np.float32),
        np.array([0.0, -0.0, 65504.0, 65520.0, 70000.0, 1e-8, 6e-8, 5.96e-8, 2.98e-8,
                  6.1e-5, 6.09e-5, np.inf, -np.inf], np.float32),
        (np.arange(1, 2049, dtype=np.float32) / 2048.0 + np.float32(2 ** -12)),  # ties
    ])
    got = O.f16_round(x)
    with np.errstate(over="ignore"):
        exp = x.astype(np.float16).view(np.uint16)
    assert np.array_equal(got, exp)


# ----------------------------------------------------------- preprocess ----
def one_splat(pos, scale, quat, opacity=0.8, color=(0.2, 0.5, 0.9), sh_degree=0, f_dc=None,
              rest=None):
    s = sc.Splats(
        positions=np.array([[*pos, 1.0]], np.float32),
        colors=np.array([[*color, opacity]], np.float32),
        scales=np.array([[*scale, 0.0]], np.float32),
        rotations=np.array([quat], np.float32),
        f_dc=np.zeros((1, 3), np.float32) if f_dc is None else np.array([f_dc], np.float32),
        sh_rest=rest, sh_degree=sh_degree)
    return s


def prep(s, W=256, H=256, eye=(0.0, 0.0, 5.0), tile=16, fov=math.pi / 3):
    cu = cam_mod.view_camera(eye, W, H, fov)
    tx, ty = -(-W // tile), -(-H // tile)
    inv = float(np.float32(1.0 / tile))
    proj, counts = O.preprocess(cu, s, tx, ty, inv, inv)
    return cu, proj, counts


def cov2d_from_conic(p):
    con = np.array([[p[4], p[5]], [p[5], p[6]]], np.float64)
    return np.linalg.inv(con)


def test_centred_isotropic_kat():
    W = H = 256
    sigma, z0 = 0.05, 5.0
    s = one_splat((0, 0, 0), (sigma,) * 3, (1, 0, 0, 0))
    cu, proj, counts = prep(s, W, H, (0, 0, z0))
    p = proj[0]
    f = cu[16] * W / 2.0  # fx = fy (square viewport)
    a = sigma * sigma * f * f / (z0 * z0) + 0.3
    assert p.view(np.uint32)[11] == 1
    assert p[0] == pytest.approx(W / 2, abs=1e-4) and p[1] == pytest.approx(H / 2, abs=1e-4)
    assert p[2] == pytest.approx(-z0, rel=1e-6)
    assert p[4] == pytest.approx(1.0 / a, rel=1e-5) and abs(p[5]) < 1e-7
    assert p[6] == pytest.approx(1.0 / a, rel=1e-5)
    lam = a + math.sqrt(0.1)  # a == c: mid^2 - det = 0 -> max(0.1, 0)
    assert p[3] == max(2.0, math.ceil(3.0 * math.sqrt(lam)))
    r = p[3]
    t0, t1 = int((W / 2 - r) // 16), int((W / 2 + r) // 16)
    assert counts[0] == (t1 - t0 + 1) ** 2


@pytest.mark.parametrize("seed", range(5))
def test_centred_cov_matches_scipy_rotation(seed):
    """At the image centre with W = I, Sigma' = J2 Sigma_xy J2^T: checks the
    (w, x, y, z) quaternion convention and R S S^T R^T against scipy."""
    rng = np.random.default_rng(seed)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    scale = rng.uniform(0.02, 0.1, 3)
    W = H = 256
    z0 = 5.0
    s = one_splat((0, 0, 0), scale, q)
    cu, proj, _ = prep(s, W, H, (0, 0, z0))
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
    Sig = R @ np.diag(scale ** 2) @ R.T
    f = cu[16] * W / 2.0
    exp = np.array([[f * f / z0 ** 2 * Sig[0, 0], -f * f / z0 ** 2 * Sig[0, 1]],
                    [-f * f / z0 ** 2 * Sig[0, 1], f * f / z0 ** 2 * Sig[1, 1]]])
    got = cov2d_from_conic(proj[0]) - np.diag([0.3, 0.3])
    assert np.allclose(got, exp, rtol=2e-3, atol=2e-3)


@pytest.mark.parametrize("seed", range(4))
def test_offcentre_jacobian_finite_difference(seed):
    """Sigma' = J W Sigma W^T J^T with J the Jacobian of the screen mapping,
    taken here by central finite differences of the projection itself."""
    rng = np.random.default_rng(100 + seed)
    pos = rng.uniform(-1.0, 1.0, 3)
    q = rng.standard_normal(4)
    q /= np.linalg.norm(q)
    scale = rng.uniform(0.02, 0.08, 3)
    W, H = 320, 200
    eye = (1.5, 0.7, 6.0)
    s = one_splat(pos, scale, q)
    cu, proj, _ = prep(s, W, H, eye)
    V = cu[0:16].astype(np.float64).reshape(4, 4).T
    Pm = cu[16:32].astype(np.float64).reshape(4, 4).T

    def screen(pv):
        c = Pm @ np.append(pv, 1.0)
        ndc = c[:3] / c[3]
        return np.array([(ndc[0] * 0.5 + 0.5) * W, (1 - (ndc[1] * 0.5 + 0.5)) * H])

    pv = (V @ np.append(pos, 1.0))[:3]
    J = np.zeros((2, 3))
    h = 1e-5
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        J[:, k] = (screen(pv + e) - screen(pv - e)) / (2 * h)
    R = Rotation.from_quat([q[1], q[2], q[3], q[0]]).as_matrix()
    Sig = R @ np.diag(scale ** 2) @ R.T
    T = J @ V[:3, :3]
    exp = T @ Sig @ T.T
    got = cov2d_from_conic(proj[0]) - np.diag([0.3, 0.3])
    assert np.allclose(got, exp, rtol=5e-3, atol=5e-3)
    assert np.allclose(proj[0][:2], screen(pv), atol=2e-3)


def test_culling_and_floor_and_clamp():
    W = H = 128
    # behind the camera
    _, proj, counts = prep(one_splat((0, 0, 9), (0.05,) * 3, (1, 0, 0, 0)), W, H, (0, 0, 5))
    assert counts[0] == 0 and proj.view(np.uint32)[0, 11] == 0
    # outside the frustum (ndc.x > 1)
    _, proj, counts = prep(one_splat((50, 0, 0), (0.05,) * 3, (1, 0, 0, 0)), W, H, (0, 0, 5))
    assert counts[0] == 0
    # tiny splat: cov2d -> the 0.3 low-pass alone, lambda = 0.3 + sqrt(0.1), so the
    # radius is ceil(3 sqrt(0.616...)) = 3 and the reference's floor of 2 never binds
    _, proj, counts = prep(one_splat((0, 0, 0), (1e-6,) * 3, (1, 0, 0, 0)), W, H, (0, 0, 5))
    assert proj[0, 3] == math.ceil(3 * math.sqrt(0.3 + math.sqrt(0.1))) == 3 and counts[0] >= 1
    # near the right edge: tile rect clamped to the grid
    _, proj, counts = prep(one_splat((2.85, 0, 0), (0.3,) * 3, (1, 0, 0, 0)), W, H, (0, 0, 5))
    u = proj.view(np.uint32)[0]
    assert u[11] == 1 and u[14] == W // 16 - 1 and u[12] <= u[14]


# ------------------------------------------------------------------ SH ----
def real_sh_scipy(l, m, d):
    """Real SH from scipy's complex Y_l^m (Condon-Shortley phase included):
    m > 0: sqrt2 Re Y_l^m, m < 0: sqrt2 Im Y_l^|m|, m = 0: Y_l^0."""
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))
    phi = math.atan2(y, x)
    Y = sph_harm_y(l, abs(m), theta, phi)
    if m > 0:
        return math.sqrt(2) * Y.real
    if m < 0:
        return math.sqrt(2) * Y.imag
    return Y.real


def test_sh_basis_matches_scipy():
    rng = np.random.default_rng(1)
    for _ in range(50):
        d = rng.standard_normal(3)
        d /= np.linalg.norm(d)
        got = O.sh_basis(*d, deg=3)
        for l in range(4):
            for m in range(-l, l + 1):
                assert got[l * l + l + m] == pytest.approx(real_sh_scipy(l, m, d), abs=1e-12)
        got32 = O.sh_basis(*d, deg=3, precision="f32")
        assert np.allclose(got32, got, atol=2e-6)


def test_sh_zero_rest_reduces_to_dc():
    f_dc = (0.7, -1.2, 0.1)
    rest = np.zeros((1, 45), np.float32)
    s = one_splat((0.1, 0.2, 0.0), (0.05,) * 3, (1, 0, 0, 0), sh_degree=3, f_dc=f_dc, rest=rest)
    _, proj, _ = prep(s)
    ref = np.clip(0.5 + sc.SH_C0 * np.array(f_dc, np.float32).astype(np.float64), 0, 1)
    assert np.allclose(proj[0, 8:11], ref, atol=1e-7)


def test_sh_view_dependence_matches_basis():
    rng = np.random.default_rng(5)
    rest = rng.normal(0, 0.1, (1, 45)).astype(np.float32)
    f_dc = (0.2, -0.3, 0.4)
    pos = (0.3, -0.2, 0.1)
    eye = (0.0, 0.0, 5.0)
    s = one_splat(pos, (0.05,) * 3, (1, 0, 0, 0), sh_degree=3, f_dc=f_dc, rest=rest)
    _, proj, _ = prep(s, eye=eye)
    d = np.array(pos, np.float64) - np.array(eye)
    d /= np.linalg.norm(d)
    coef = np.zeros((3, 16))
    coef[:, 0] = f_dc
    coef[:, 1:] = rest.reshape(3, 15)  # f_rest: all R, then all G, then all B
    basis = np.array([real_sh_scipy(l, m, d) for l in range(4) for m in range(-l, l + 1)])
    ref = np.clip(0.5 + coef @ basis, 0, 1)
    assert np.allclose(proj[0, 8:11], ref, atol=2e-6)


# -------------------------------------------------------------- raster ----
def raster_scene(splats, W=32, H=32, tile=16):
    """Projected splats given directly: rows (mx, my, depth, r, cx, cy, cz, o, R, G, B)."""
    n = len(splats)
    tx, ty = -(-W // tile), -(-H // tile)
    proj = np.zeros((n, 16), np.float32)
    u = proj.view(np.uint32)
    for i, sp in enumerate(splats):
        proj[i, :11] = sp
        u[i, 11] = 1
        r = sp[3]
        u[i, 12] = min(int(max(0, sp[0] - r) // tile), tx - 1)
        u[i, 13] = min(int(max(0, sp[1] - r) // tile), ty - 1)
        u[i, 14] = min(int(min(W - 1, sp[0] + r) // tile), tx - 1)
        u[i, 15] = min(int(min(H - 1, sp[1] + r) // tile), ty - 1)
    counts = ((u[:, 14] - u[:, 12] + 1) * (u[:, 15] - u[:, 13] + 1)).astype(np.uint32)
    offs, _ = O.scan(counts)
    keys, vals = O.emit(proj, offs, tx)
    sk, sv = O.radix_sort(keys, vals)
    toff = O.identify(sk, tx * ty)
    img, _, _ = O.rasterize(proj, sv, toff, tile, W, H, quantize=False)
    return img


def test_raster_single_splat_centre():
    img = raster_scene([(8.5, 8.5, -2.0, 6.0, 0.5, 0.0, 0.5, 0.6, 0.2, 0.4, 1.0)])
    assert img[8, 8, 3] == pytest.approx(0.6, abs=1e-7)
    assert np.allclose(img[8, 8, :3], 0.6 * np.array([0.2, 0.4, 1.0]), atol=1e-7)
    # radius cull: |d| > r -> no contribution even though the Gaussian is > 0
    assert img[8, 15, 3] == 0.0
    # alpha at a nearby pixel = o * exp(-0.5 d^T conic d)
    a = 0.6 * math.exp(-0.5 * (0.5 * 2 * 2))
    assert img[8, 10, 3] == pytest.approx(a, rel=1e-6)


def test_raster_occlusion_and_termination_rule():
    front = (8.5, 8.5, -1.0, 6.0, 1e-3, 0.0, 1e-3, 1.0, 1.0, 0.0, 0.0)
    back = (8.5, 8.5, -3.0, 6.0, 1e-3, 0.0, 1e-3, 1.0, 0.0, 1.0, 0.0)
    third = (8.5, 8.5, -5.0, 6.0, 1e-3, 0.0, 1e-3, 1.0, 0.0, 0.0, 1.0)
    img = raster_scene([third, back, front])  # ids out of depth order on purpose
    p = img[8, 8]
    a = np.float32(0.99)
    # accumulate THEN stop: the second splat is blended, the third never is
    assert p[0] == pytest.approx(a, rel=1e-6)
    assert p[1] == pytest.approx((1 - a) * a, rel=1e-5)
    assert p[2] == 0.0
    assert p[3] == pytest.approx(1 - (1 - a) ** 2, rel=1e-6)


def test_raster_alpha_threshold():
    lo = raster_scene([(8.5, 8.5, -1.0, 6.0, 0.1, 0.0, 0.1, 1.0 / 255.0 - 1e-6, 1, 1, 1)])
    hi = raster_scene([(8.5, 8.5, -1.0, 6.0, 0.1, 0.0, 0.1, 1.0 / 255.0 + 1e-6, 1, 1, 1)])
    assert lo[8, 8, 3] == 0.0 and hi[8, 8, 3] > 0.0


def brute_force_image(proj, W, H, tile):
    """Per pixel, every valid splat whose tile rect holds the pixel's tile, in
    (depth key, id) order, composited in float64 with the reference rules."""
    u = proj.view(np.uint32)
    n = proj.shape[0]
    lo = np.array([np.float32(-proj[i, 2]).view(np.uint32) for i in range(n)], np.uint64)
    order = sorted([i for i in range(n) if u[i, 11]], key=lambda i: (int(lo[i]), i))
    img = np.zeros((H, W, 4))
    for y in range(H):
        for x in range(W):
            tx, ty = x // tile, y // tile
            C, T = np.zeros(3), 1.0
            for i in order:
                if not (u[i, 12] <= tx <= u[i, 14] and u[i, 13] <= ty <= u[i, 15]):
                    continue
                p = proj[i].astype(np.float64)
                dx, dy = x + 0.5 - p[0], y + 0.5 - p[1]
                if dx * dx + dy * dy > p[3] * p[3]:
                    continue
                pw = -0.5 * (p[4] * dx * dx + 2 * p[5] * dx * dy + p[6] * dy * dy)
                if pw > 0:
                    continue
                a = min(0.99, p[7] * math.exp(pw))
                if a < 1 / 255:
                    continue
                C += T * a * p[8:11]
                T *= 1 - a
                if T < 1e-4:
                    break
            img[y, x, :3], img[y, x, 3] = C, 1 - T
    return img


def test_raster_matches_brute_force():
    raw = sc.uniform_cube(7, 300, log_scale_mean=-3.0)
    s = sc.activate(raw)
    W, H, tile = 48, 40, 16
    cu = cam_mod.view_camera((0.0, 0.0, 3.0), W, H)
    out = O.pipeline_f32(cu, s, W, H, tile)
    bf = brute_force_image(out["projected"], W, H, tile)
    img, _, _ = O.rasterize(out["projected"], out["sorted_values"], out["tile_offsets"], tile, W, H,
                            quantize=False)
    assert np.abs(img - bf).max() < 2e-5


def test_render_f32_close_to_f64():
    raw = sc.clustered(11, 3000, 8, log_scale_mean=-3.5, sh_degree=3)
    s = sc.activate(raw)
    W, H = 160, 96
    cu = cam_mod.view_camera(sc.orbit_eye(8.0, 20.0, 10.0), W, H)
    i32, k32, _, _ = O.render(cu, s, W, H, 16, "f32")
    i64, k64, _, _ = O.render(cu, s, W, H, 16, "f64")
    assert abs(k32 - k64) <= max(4, k64 // 200)
    assert np.abs(i32 - i64).max() < 1.0 / 255.0


# ------------------------------------------------- heat map + composite ----
def f16_image(rng, H, W):
    """Random premultiplied f16-representable target: alpha in [0, 1], rgb <= alpha."""
    a = rng.uniform(0, 1, (H, W, 1))
    rgb = rng.uniform(0, 1, (H, W, 3)) * a
    img = np.concatenate([rgb, a], axis=2).astype(np.float16).astype(np.float32)
    img[0, 0] = 0.0          # fully transparent
    img[0, 1] = [0.25, 0.5, 0.75, 1.0]  # opaque
    return img


def test_composite_transparent_keeps_and_opaque_replaces():
    rng = np.random.default_rng(3)
    H, W = 16, 64
    fb = np.arange(H * W * 4, dtype=np.uint32).reshape(H, W, 4).astype(np.uint8)  # every byte value
    out = O.composite(np.zeros((H, W, 4), np.float32), W, H, framebuffer=fb)
    assert np.array_equal(out, fb)  # src alpha 0 leaves every stored byte as it was
    src = rng.uniform(0, 1, (H, W, 4)).astype(np.float32)
    src[..., 3] = 1.0
    out = O.composite(src, W, H, framebuffer=fb)
    assert np.array_equal(out, np.floor(src.astype(np.float64) * 255 + 0.5).astype(np.uint8))


def test_composite_clear_colour_and_over_formula():
    H, W = 8, 8
    out = O.composite(np.zeros((H, W, 4), np.float32), W, H, clear=(0.12, 0.12, 0.13, 1.0))
    assert np.all(out == np.array([31, 31, 33, 255], np.uint8))  # round(0.12*255)=31, round(0.13*255)=33
    rng = np.random.default_rng(4)
    src = f16_image(rng, 40, 50)
    fb = rng.integers(0, 256, (40, 50, 4), dtype=np.uint8)
    out = O.composite(src, 50, 40, framebuffer=fb)
    s64 = src.astype(np.float64)
    want = np.clip(s64 + fb / 255.0 * (1.0 - s64[..., 3:4]), 0, 1) * 255
    diff = np.abs(out.astype(np.int32) - np.floor(want + 0.5).astype(np.int32))
    assert diff.max() <= 1 and (diff == 0).mean() > 0.99
    # a blend weighted by dst alpha or by (1 - src colour) would miss by far more
    wrong = np.floor(np.clip(s64 + fb / 255.0 * (1.0 - fb[..., 3:4] / 255.0), 0, 1) * 255 + 0.5)
    assert np.abs(out - wrong).max() > 10


def test_composite_bgra_swaps_red_and_blue():
    rng = np.random.default_rng(5)
    src = f16_image(rng, 12, 20)
    fb = rng.integers(0, 256, (12, 20, 4), dtype=np.uint8)
    rgba = O.composite(src, 20, 12, framebuffer=fb)
    bgra = O.composite(src, 20, 12, framebuffer=fb[..., [2, 1, 0, 3]], bgra=True)
    assert np.array_equal(bgra, rgba[..., [2, 1, 0, 3]])
    c = O.composite(np.zeros((2, 2, 4), np.float32), 2, 2, clear=(1.0, 0.5, 0.0, 1.0), bgra=True)
    assert np.all(c == np.array([0, 128, 255, 255], np.uint8))


def overlay_f64(base, tile_offsets, tile, W, H, ref):
    tx = np.minimum(np.arange(W) // tile, -(-W // tile) - 1)
    ty = np.minimum(np.arange(H) // tile, -(-H // tile) - 1)
    tid = ty[:, None] * (-(-W // tile)) + tx[None, :]
    lo, hi = tile_offsets[tid].astype(np.int64), tile_offsets[tid + 1].astype(np.int64)
    count = np.where(hi >= lo, hi - lo, 0)
    t = np.clip(count / max(1.0, float(ref)), 0, 1)
    tint = np.stack([np.clip(2 * t, 0, 1), np.clip(2 - 2 * t, 0, 1), np.zeros_like(t)], axis=2)
    out = np.empty_like(base, dtype=np.float64)
    out[..., :3] = tint * 0.4 + base[..., :3] * 0.6
    out[..., 3] = 0.4 + base[..., 3] * 0.6
    return out


def test_overlay_colour_ramp_end_points():
    """count 0 -> green, count = ref/2 -> yellow, count >= ref -> red, all at opacity 0.4."""
    W, H, tile, ref = 64, 16, 16, 256
    tiles = 4
    to = np.zeros(tiles + 1, np.uint32)
    counts = [0, 128, 256, 1000]
    to[1:] = np.cumsum(counts)
    out = O.tile_overlay(np.zeros((H, W, 4), np.float32), to, tile, W, H, ref)
    f16 = lambda v: float(np.float16(v))
    want = [(0, f16(0.4), 0), (f16(0.4), f16(0.4), 0), (f16(0.4), 0, 0), (f16(0.4), 0, 0)]
    for t in range(tiles):
        blk = out[:, t * tile:(t + 1) * tile]
        assert np.all(blk[..., 0] == want[t][0]) and np.all(blk[..., 1] == want[t][1])
        assert np.all(blk[..., 2] == 0) and np.all(blk[..., 3] == f16(0.4))


def test_overlay_underflow_guard_and_ref_clamp():
    W, H, tile = 32, 16, 16
    # tile 0: offsets 5 -> 0 (hi < lo: reads as 0, green); tile 1: 0 -> 7
    to = np.array([5, 0, 7], np.uint32)
    out = O.tile_overlay(np.zeros((H, W, 4), np.float32), to, tile, W, H, 0)
    assert np.all(out[:, :16, 1] == np.float16(0.4)) and np.all(out[:, :16, 0] == 0)
    # debug_ref 0 is clamped to 1: any count >= 1 is full red
    assert np.all(out[:, 16:, 0] == np.float16(0.4)) and np.all(out[:, 16:, 1] == 0)


@pytest.mark.parametrize("W,H,tile", [(100, 37, 16), (33, 70, 8), (64, 64, 32)])
def test_overlay_matches_float64_on_ragged_viewports(W, H, tile):
    rng = np.random.default_rng(W * 1000 + H)
    T = (-(-W // tile)) * (-(-H // tile))
    to = np.concatenate([[0], np.cumsum(rng.integers(0, 600, T))]).astype(np.uint32)
    to[-3:-1] = 0  # trailing tiles past the last populated one stay 0 (identify's contract)
    base = f16_image(rng, H, W)
    out = O.tile_overlay(base, to, tile, W, H, 256)
    want = overlay_f64(base.astype(np.float64), to, tile, W, H, 256)
    # each channel within one binary16 step of the float64 value
    step = np.abs(want.astype(np.float16).astype(np.float64) - np.nextafter(
        want.astype(np.float16), np.float16(np.inf)).astype(np.float64))
    assert np.all(np.abs(out - want) <= step + 1e-12)
    assert np.array_equal(out, out.astype(np.float16).astype(np.float32))
