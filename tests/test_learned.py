"""Learned colour stage: the oracle (oracle/learned_oracle.py) pinned on the
CPU, then the HIP kernels (k_learned.hip) against it on the GPU: gradients of
one step, inference colours, and what training gains over the DC colour the
reference uses.
"""
import math
import sys
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as O
from gs3d import _lib as L
from gs3d import camera as cam_mod
from gs3d import scene as S

sys.path.insert(0, str(Path(__file__).resolve().parents[1] / "oracle"))
import learned_oracle as LO  # noqa: E402


def say(*a):
    print(*a, flush=True)


def random_params(rng, scale=1.0):
    W1 = rng.normal(0, math.sqrt(2 / 11) * scale, (64, 11))
    W2 = rng.normal(0, math.sqrt(2 / 64) * scale, (64, 64))
    W3 = rng.normal(0, math.sqrt(1 / 64) * scale, (3, 64))
    return LO.pack(W1, rng.normal(0, 0.1, 64), W2, rng.normal(0, 0.1, 64), W3, rng.normal(0, 0.1, 3))


# ---------------------------------------------------------------- oracle (CPU)
def test_oracle_backward_matches_finite_differences():
    """float64 oracle: analytic gradients of the mean squared error vs central differences."""
    rng = np.random.default_rng(1)
    S_ = 24
    p = random_params(rng)
    z = rng.normal(0, 0.5, (S_, 8))
    d = rng.normal(0, 1, (S_, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = rng.uniform(0, 1, (S_, 3))
    loss, gp, gz = LO.loss_and_grads(p, z, d, t, bf16=False)
    eps = 1e-6
    for idx in list(rng.choice(LO.N_PARAMS, 60, replace=False)) + [703, 704 + 3, 5120, 5122]:
        q = p.copy()
        q[idx] += eps
        lp = LO.loss_and_grads(q, z, d, t, bf16=False)[0]
        q[idx] -= 2 * eps
        lm = LO.loss_and_grads(q, z, d, t, bf16=False)[0]
        assert abs((lp - lm) / (2 * eps) - gp[idx]) <= 1e-6 + 1e-4 * abs(gp[idx]), idx
    for s_, k in [(0, 0), (5, 7), (23, 3)]:
        zz = z.copy()
        zz[s_, k] += eps
        lp = LO.loss_and_grads(p, zz, d, t, bf16=False)[0]
        zz[s_, k] -= 2 * eps
        lm = LO.loss_and_grads(p, zz, d, t, bf16=False)[0]
        assert abs((lp - lm) / (2 * eps) - gz[s_, k]) <= 1e-6 + 1e-4 * abs(gz[s_, k])


def test_oracle_bf16_close_to_float64_and_forward_by_hand():
    rng = np.random.default_rng(2)
    p = random_params(rng)
    z = rng.normal(0, 0.5, (40, 8))
    d = rng.normal(0, 1, (40, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    y64, _ = LO.forward(p, z, d, bf16=False)
    y16, _ = LO.forward(p, z, d, bf16=True)
    assert np.abs(y64 - y16).max() < 2e-2
    # one sample by explicit loops
    W1, b1, W2, b2, W3, b3 = LO.unpack(p)
    x = np.concatenate([z[0], d[0]])
    h1 = [max(0.0, sum(W1[i, k] * x[k] for k in range(11)) + b1[i]) for i in range(64)]
    h2 = [max(0.0, sum(W2[i, k] * h1[k] for k in range(64)) + b2[i]) for i in range(64)]
    y = [1 / (1 + math.exp(-(sum(W3[c, k] * h2[k] for k in range(64)) + b3[c]))) for c in range(3)]
    assert np.allclose(y, y64[0], atol=1e-12)


def test_oracle_sh_basis_and_targets():
    """The oracle's SH basis equals the C oracle's (itself pinned against scipy), and a
    zero f_rest gives the DC colour clamp01(0.5 + C0 f_dc) in every direction."""
    dirs = LO.fibonacci_dirs()
    assert np.allclose(np.linalg.norm(dirs.astype(np.float64), axis=1), 1.0, atol=1e-6)
    assert np.abs(dirs.astype(np.float64).mean(axis=0)).max() < 0.01  # evenly spread
    for dvec in dirs[::37]:
        want = O.sh_basis(float(dvec[0]), float(dvec[1]), float(dvec[2]), 3)
        assert np.allclose(LO.sh_basis(dvec.astype(np.float64)), want, atol=1e-12)
    f_dc = np.array([[0.3, -1.0, 4.0]])
    t = LO.sh_targets(f_dc, np.zeros((1, 45)), 3, dirs[:5])
    assert np.allclose(t, np.clip(0.5 + 0.28209479177387814 * f_dc, 0, 1)[:, None, :])


def test_oracle_subset_rows_equal_whole_scene_rows():
    """step_grads with n_total: the latent-gradient rows of a subset of splats equal the
    rows of the whole scene's step, bit for bit (every latent row depends only on its own
    splat's samples and the shared mean scale)."""
    rng = np.random.default_rng(6)
    n = 40
    p = random_params(rng)
    z = rng.normal(0, 0.3, (n, 8))
    f_dc = rng.normal(0, 0.5, (n, 3))
    rest = rng.normal(0, 0.1, (n, 45))
    _, _, full = LO.step_grads(p, z, f_dc, rest, 3, 32)
    pick = np.array([0, 7, 8, 39])
    _, _, sub = LO.step_grads(p, z[pick], f_dc[pick], rest[pick], 3, 32, n_total=n)
    assert np.array_equal(sub, full[pick])
    _, _, wrong = LO.step_grads(p, z[pick], f_dc[pick], rest[pick], 3, 32)
    assert not np.allclose(wrong, full[pick])


def test_library_directions_equal_oracle():
    from gs3d.renderer import GaussianSplatRenderer

    assert np.array_equal(GaussianSplatRenderer.learned_dirs(), LO.fibonacci_dirs())


# ---------------------------------------------------------------- GPU
@pytest.fixture(scope="module")
def small_scene(renderer):
    raw = S.clustered(21, 600, 6, sh_degree=3)
    s = S.activate(raw)
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    renderer.learned_init(seed=3)
    return s


def rel_err(a, b):
    return float(np.linalg.norm(np.asarray(a, np.float64) - b) / max(1e-30, np.linalg.norm(b)))


@pytest.mark.gpu
@pytest.mark.parametrize("first_dir", [0, 112, 240])
def test_step_gradients_match_oracle(renderer, small_scene, first_dir):
    say(f"learned: gradients of one step, directions {first_dir}..{first_dir + 15}")
    p, z, _ = renderer.learned_get()
    loss, gp, gl = renderer.learned_grad(first_dir)
    s = small_scene
    rl, rgp, rgl = LO.step_grads(p, z, s.f_dc, s.sh_rest, 3, first_dir)
    say(f"  loss gpu {loss:.6g} oracle {rl:.6g}; rel err params {rel_err(gp, rgp):.3g} latents {rel_err(gl, rgl):.3g}")
    # The kernels and the oracle round to bf16 at the same points and differ only in fp32
    # summation order: measured 1e-7 to 1e-6 (parameters) and up to 1.6e-5 (latents).
    # A missing bf16 rounding point or a wrong direction index is off by 1e-3 or more.
    assert abs(loss - rl) <= 1e-5 * rl + 1e-9
    assert rel_err(gp, rgp) < 1e-4
    assert rel_err(gl, rgl) < 1e-4
    # per block (W1, b1, W2, b2, W3, b3), so no block hides behind a larger one
    blocks = [(0, 704), (704, 768), (768, 4864), (4864, 4928), (4928, 5120), (5120, 5123)]
    per_block = [rel_err(gp[a:b], rgp[a:b]) for a, b in blocks]
    say("  per block: " + " ".join(f"{e:.2g}" for e in per_block))
    for (a, b), e in zip(blocks, per_block):
        assert e < 3e-4, (a, b, e)


@pytest.mark.gpu
def test_inference_colors_match_oracle(renderer, small_scene):
    say("learned: inference colours vs oracle forward")
    p, z, _ = renderer.learned_get()
    eye = np.array([3.0, 2.0, 6.0], np.float32)
    got = renderer.learned_colors(eye)
    pos = small_scene.positions[:, :3].astype(np.float32)
    d = pos - eye
    d = d / np.sqrt((d * d).sum(axis=1, keepdims=True))
    want, _ = LO.forward(p, z, d, bf16=True)
    assert np.abs(got - want).max() < 4e-3  # f16 output + fast exp


@pytest.mark.gpu
def test_training_gains_over_dc_colour(renderer):
    """Trained on the exact SH3 colours, the learned colours come closer to them than the
    DC colour the reference uses, per splat (mean squared error over the 256 training
    directions) and in rendered images (PSNR against the exact SH render).

    With random SH coefficients (45 per splat, N(0, 0.1)) the view-dependent part holds
    ~0.0119 of variance per channel; an 8-dim latent can carry only a few of those 45
    dimensions, so the gain over the DC colour is real but small."""
    say("learned: training and PSNR against exact SH")
    raw = S.clustered(22, 4000, 8, sh_degree=3)
    s = S.activate(raw)
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    renderer.learned_init(seed=5)
    losses = renderer.learned_train(40, lr=3e-3, lr_latent=3e-2)
    dirs = LO.fibonacci_dirs()
    t = LO.sh_targets(s.f_dc, s.sh_rest, 3, dirs)
    dc_loss = float(((np.clip(0.5 + 0.28209479177387814 * s.f_dc.astype(np.float64), 0, 1)[:, None, :]
                      - t) ** 2).mean())
    say(f"  loss epoch 0 {losses[0]:.5f} -> epoch {len(losses) - 1} {losses[-1]:.5f}; DC colour {dc_loss:.5f}")
    assert losses[-1] < 0.5 * losses[0]
    assert losses[-1] < 0.95 * dc_loss
    views = [(cam_mod.view_camera(eye, w, h), w, h) for eye, w, h in S.orbit_views(9, 4, fixed=(320, 180))]

    def images(learned):
        renderer.rebin_batch(views, learned=learned)
        return [renderer.batch_read(i, L.BUF_IMAGE_RGBA32F, np.float32).reshape(-1, 4)[:, :3]
                for i in range(len(views))]

    exact = np.concatenate(images(False))
    psnr, enabled = renderer.learned_validate(views, 0.0)
    assert enabled
    learned = np.concatenate(images(True))
    mse_l = float(((learned - exact) ** 2).mean())
    assert abs(10 * math.log10(1 / mse_l) - psnr) < 0.05  # the library's PSNR is the same quantity
    dc = S.Splats(s.positions, s.colors, s.scales, s.rotations, s.f_dc, None, 0, s.bounds)
    renderer.set_splat_data(dc)
    dc_img = np.concatenate(images(False))
    psnr_dc = 10 * math.log10(1 / float(((dc_img - exact) ** 2).mean()))
    say(f"  PSNR vs exact SH3: learned {psnr:.2f} dB, DC colour {psnr_dc:.2f} dB")
    assert psnr > psnr_dc + 0.2
    renderer.set_splat_data(s)


@pytest.mark.gpu
def test_fallback_to_exact_sh_below_threshold(renderer, small_scene):
    """Below the PSNR threshold a learned batch renders exact SH; at or above it, learned colours."""
    say("learned: fallback")
    renderer.set_splat_data(small_scene)
    renderer.learned_init(seed=3)
    views = [(cam_mod.view_camera(eye, w, h), w, h) for eye, w, h in S.orbit_views(4, 2, fixed=(200, 120))]
    renderer.rebin_batch(views)
    exact = [renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16) for i in range(2)]
    exact_rec = renderer.batch_read(1, L.BUF_PROJECTED_SPLATS, np.uint32).reshape(-1, 16)
    psnr, enabled = renderer.learned_validate(views, 200.0)
    assert not enabled and psnr < 200.0
    renderer.rebin_batch(views, learned=True)
    for i in range(2):
        assert np.array_equal(renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16), exact[i])
    psnr2, enabled = renderer.learned_validate(views, -100.0)
    assert enabled and abs(psnr2 - psnr) < 1e-3
    renderer.rebin_batch(views, learned=True)
    differs = any(not np.array_equal(renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16), exact[i])
                  for i in range(2))
    assert differs
    # the learned batch's records: the same geometry, the learned colours
    rec = renderer.batch_read(1, L.BUF_PROJECTED_SPLATS, np.uint32).reshape(-1, 16)
    geometry = [0, 1, 2, 3, 4, 5, 6, 7, 11, 12, 13, 14, 15]
    assert np.array_equal(rec[:, geometry], exact_rec[:, geometry])
    visible = exact_rec[:, 11] == 1
    assert visible.any() and not np.array_equal(rec[visible][:, 8:11], exact_rec[visible][:, 8:11])
    assert not rec[~visible].any()


@pytest.mark.gpu
def test_new_splats_reset_the_learned_state(renderer, small_scene):
    """Latents belong to the splats they were trained on: uploading splats again (even of
    the same count) drops them, and a learned batch falls back to exact SH until re-init."""
    say("learned: reset on upload")
    renderer.set_splat_data(small_scene)
    renderer.learned_init(seed=3)
    renderer.learned_enable(True)
    renderer.set_splat_data(small_scene)
    with pytest.raises(Exception):
        renderer.learned_get()
    views = [(cam_mod.view_camera(eye, w, h), w, h) for eye, w, h in S.orbit_views(4, 1, fixed=(200, 120))]
    renderer.rebin_batch(views)
    exact = renderer.batch_read(0, L.BUF_IMAGE_RGBA16F, np.uint16)
    renderer.rebin_batch(views, learned=True)
    assert np.array_equal(renderer.batch_read(0, L.BUF_IMAGE_RGBA16F, np.uint16), exact)


@pytest.mark.gpu
@pytest.mark.timeout(900)
def test_workload_3m_sh3_32_views_learned(renderer):
    """The learned-colour workload itself: 3M SH3 splats, 32 views at 1920x1080.
    Latent gradients of one step for a sample of splats against the oracle (rows of the
    whole scene's step), a short training that lowers the loss, inference colours of a
    sample against the oracle's forward, a gate decision consistent with the PSNR, and,
    below the gate, learned batches that render exactly the exact-SH images."""
    say("learned: workload clusters-3M-sh3-32views-learned")
    wl = S.workload(4)
    s = S.activate(wl.raw)
    renderer.set_splat_data(s)
    renderer.set_tile_size(16, 16)
    renderer.learned_init(seed=4)
    p, z, _ = renderer.learned_get()
    rng = np.random.default_rng(4)
    pick = np.sort(rng.choice(s.n, 3000, replace=False))
    for first_dir in (0, 128):
        _, _, gl = renderer.learned_grad(first_dir)
        _, _, rgl = LO.step_grads(p, z[pick], s.f_dc[pick], s.sh_rest[pick], 3, first_dir, n_total=s.n)
        e = rel_err(gl[pick], rgl)
        say(f"  latent gradients, directions {first_dir}..{first_dir + 15}, 3000 splats: rel err {e:.3g}")
        assert e < 1e-4
    losses = renderer.learned_train(2, lr=3e-3, lr_latent=3e-2)
    say(f"  2 epochs: loss {losses[0]:.5f} -> {losses[-1]:.5f}")
    assert losses[-1] < losses[0]
    p, z, _ = renderer.learned_get()
    eye = np.array(wl.views[5][0], np.float32)
    got = renderer.learned_colors(eye)
    d = s.positions[pick, :3].astype(np.float32) - eye
    d = d / np.sqrt((d * d).sum(axis=1, keepdims=True))
    want, _ = LO.forward(p, z[pick], d, bf16=True)
    assert np.abs(got[pick] - want).max() < 4e-3
    views = [(cam_mod.view_camera(e, w, h, wl.fov_y), w, h) for e, w, h in wl.views]
    renderer.rebin_batch(views)
    exact = [renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16) for i in range(len(views))]
    psnr, enabled = renderer.learned_validate(views, 45.0)
    say(f"  PSNR vs exact SH over 32 views: {psnr:.2f} dB -> {'learned' if enabled else 'exact SH fallback'}")
    assert enabled == (psnr >= 45.0)
    renderer.rebin_batch(views, learned=True)
    same = all(np.array_equal(renderer.batch_read(i, L.BUF_IMAGE_RGBA16F, np.uint16), exact[i])
               for i in range(len(views)))
    assert same == (not enabled)


@pytest.mark.gpu
def test_checkpoint_resume_is_exact(renderer):
    """gs3d_learned_get_state / set_state: training resumed from a checkpoint (parameters,
    latents, both Adam moments and the step) continues exactly as uninterrupted training
    (the training kernels are deterministic: no atomics); resuming without the moments
    does not. Setting a state leaves the moments as set (no decay)."""
    say("learned: checkpoint and resume")
    s = S.activate(S.clustered(23, 3000, 6, sh_degree=3))
    renderer.set_splat_data(s)
    renderer.learned_init(seed=7)
    renderer.learned_train(2, lr=3e-3, lr_latent=3e-2)
    ckpt = renderer.learned_get_state()
    assert ckpt["step"] == 2 * 16
    assert np.abs(ckpt["params_m"]).max() > 0 and np.abs(ckpt["latents_v"]).max() > 0
    renderer.learned_train(2, lr=3e-3, lr_latent=3e-2)
    straight = renderer.learned_get_state()
    renderer.learned_init(seed=99)  # a different state, then the checkpoint
    renderer.learned_set_state(ckpt)
    back = renderer.learned_get_state()
    for k in renderer.STATE_ARRAYS:
        assert np.array_equal(back[k], ckpt[k]), k
    assert back["step"] == ckpt["step"]
    _, _, z16 = renderer.learned_get()
    assert np.array_equal(z16, ckpt["latents"].astype(np.float16).view(np.uint16))
    renderer.learned_train(2, lr=3e-3, lr_latent=3e-2)
    resumed = renderer.learned_get_state()
    for k in renderer.STATE_ARRAYS:
        assert np.array_equal(resumed[k], straight[k]), k
    assert resumed["step"] == straight["step"] == 4 * 16
    # the same resume without the moments and the step diverges
    renderer.learned_init(seed=99)
    renderer.learned_set(ckpt["params"], ckpt["latents"])
    renderer.learned_train(2, lr=3e-3, lr_latent=3e-2)
    p, _, _ = renderer.learned_get()
    assert not np.array_equal(p, straight["params"])
