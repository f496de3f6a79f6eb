"""CPU oracle of the learned colour stage (numpy).

TEST INFRASTRUCTURE ONLY: tests/ use it to check the HIP kernels of
webgpu-3dgs_amd/csrc/k_learned.hip; the product never imports it.

The stage has no reference counterpart (the reference colours splats with the
DC term only, SceneWorker.ts:428-433). What it computes is fixed here:

  x   = (z0..z7, dx, dy, dz)            latent (through f16) + unit direction
  h1  = relu(W1 x + b1)                 64
  h2  = relu(W2 h1 + b2)                64
  y   = sigmoid(W3 h2 + b3)             3
  t   = clamp01(0.5 + sum_k b_k(d) c_k) exact SH colour per channel (the target)
  L   = mean over samples and channels of (y - t)^2

with parameters packed as W1[64][11], b1[64], W2[64][64], b2[64], W3[3][64], b3[3].

`bf16=True` rounds where the MFMA kernels round (inputs, weights, post-ReLU
activations and the back-propagated errors, all to bf16; sums in float64), so
the kernels can be compared within accumulation-order tolerance.
`bf16=False` is the plain float64 function: its analytic gradient is checked
against finite differences, which pins the backward formulas.
"""
from __future__ import annotations

import numpy as np

N_PARAMS = 5123
SH_C = (0.28209479177387814, 0.4886025119029199, 1.0925484305920792, -1.0925484305920792,
        0.31539156525252005, -1.0925484305920792, 0.5462742152960396, -0.5900435899266435,
        2.890611442640554, -0.4570457994644658, 0.3731763325901154, -0.4570457994644658,
        1.445305721320277, -0.5900435899266435)


def unpack(p):
    p = np.asarray(p, np.float64)
    o = 0
    W1 = p[o:o + 704].reshape(64, 11); o += 704
    b1 = p[o:o + 64]; o += 64
    W2 = p[o:o + 4096].reshape(64, 64); o += 4096
    b2 = p[o:o + 64]; o += 64
    W3 = p[o:o + 192].reshape(3, 64); o += 192
    b3 = p[o:o + 3]; o += 3
    assert o == N_PARAMS
    return W1, b1, W2, b2, W3, b3


def pack(W1, b1, W2, b2, W3, b3):
    return np.concatenate([W1.ravel(), b1, W2.ravel(), b2, W3.ravel(), b3])


def to_bf16(a):
    """Round to bfloat16 (nearest even), returned as float64."""
    f = np.asarray(a, np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def to_f16(a):
    return np.asarray(a, np.float32).astype(np.float16).astype(np.float64)


def fibonacci_dirs(n=256):
    """The training directions: z evenly spaced, azimuth stepped by the golden angle (as f32)."""
    k = np.arange(n, dtype=np.float64)
    z = 1.0 - (2.0 * k + 1.0) / n
    r = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    phi = np.pi * (3.0 - np.sqrt(5.0)) * k
    return np.stack([r * np.cos(phi), r * np.sin(phi), z], axis=1).astype(np.float32)


def sh_basis(d):
    """Real SH basis of degree 3 (16 values) at unit directions d (..., 3), float64."""
    d = np.asarray(d, np.float64)
    x, y, z = d[..., 0], d[..., 1], d[..., 2]
    xx, yy, zz, xy, yz, xz = x * x, y * y, z * z, x * y, y * z, x * z
    c = SH_C
    return np.stack([
        np.full_like(x, c[0]), -c[1] * y, c[1] * z, -c[1] * x,
        c[2] * xy, c[3] * yz, c[4] * (2 * zz - xx - yy), c[5] * xz, c[6] * (xx - yy),
        c[7] * y * (3 * xx - yy), c[8] * xy * z, c[9] * y * (4 * zz - xx - yy),
        c[10] * z * (2 * zz - 3 * xx - 3 * yy), c[11] * x * (4 * zz - xx - yy),
        c[12] * z * (xx - yy), c[13] * x * (xx - 3 * yy)], axis=-1)


def sh_targets(f_dc, sh_rest, deg, dirs):
    """clamp01(0.5 + sum_k b_k(d) c_k) for every splat (rows) and direction: (n, D, 3)."""
    n = f_dc.shape[0]
    R = (deg + 1) ** 2 - 1
    coef = np.zeros((n, 16, 3))
    coef[:, 0, :] = f_dc
    if deg:
        rest = np.asarray(sh_rest, np.float64).reshape(n, 3, R)
        coef[:, 1:R + 1, :] = rest.transpose(0, 2, 1)
    b = sh_basis(dirs)[:, :(deg + 1) ** 2]  # (D, ncoef)
    col = np.einsum("dk,nkc->ndc", b, coef[:, :(deg + 1) ** 2, :])
    return np.clip(col + 0.5, 0.0, 1.0)


def forward(p, z, d, bf16=True):
    """Colours for latents z (n, 8) and directions d (n, 3); returns (y, cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    zin = to_f16(z) if bf16 else np.asarray(z, np.float64)
    x = np.concatenate([zin, np.asarray(d, np.float64)], axis=1)
    xb = r(x)
    W1b, b1b, W2b, W3b = r(W1), r(b1), r(W2), r(W3)
    d1 = xb @ W1b.T + b1b
    h1 = r(np.maximum(d1, 0.0))
    d2 = h1 @ W2b.T + b2
    h2 = r(np.maximum(d2, 0.0))
    d3 = h2 @ W3b.T + b3
    y = 1.0 / (1.0 + np.exp(-d3))
    return y, dict(xb=xb, d1=d1, h1=h1, d2=d2, h2=h2, y=y, W1b=W1b, W2b=W2b, W3b=W3b)


def loss_and_grads(p, z, d, t, bf16=True, s_total=None):
    """Mean squared error over samples x channels and its gradients.

    z (S, 8) latents per sample, d (S, 3), t (S, 3) targets. Returns
    (loss, grad_params (5123,), grad_z (S, 8)). With s_total, the samples are
    a subset of s_total and the mean (and every gradient, rounded where the
    kernels round) is taken over s_total: grad_z rows then equal the full
    computation's rows for these samples.
    """
    S = z.shape[0] if s_total is None else s_total
    y, c = forward(p, z, d, bf16)
    r = to_bf16 if bf16 else (lambda a: np.asarray(a, np.float64))
    diff = y - t
    loss = float(np.sum(diff * diff)) / (3 * S)
    e3 = (2.0 / (3 * S)) * diff * y * (1.0 - y)
    e3b = r(e3)
    e2 = (e3b @ c["W3b"]) * (c["d2"] > 0)
    e2b = r(e2)
    e1 = (e2b @ c["W2b"]) * (c["d1"] > 0)
    e1b = r(e1)
    gW3 = e3b.T @ c["h2"]
    gb3 = e3.sum(axis=0)
    gW2 = e2b.T @ c["h1"]
    gb2 = e2.sum(axis=0)
    gW1 = e1b.T @ c["xb"][:, :11]
    gb1 = e1b.sum(axis=0)
    gz = (e1b @ c["W1b"])[:, :8]
    return loss, pack(gW1, gb1, gW2, gb2, gW3, gb3), gz


def step_grads(p, latents, f_dc, sh_rest, deg, first_dir, dirs=None, bf16=True, n_total=None):
    """The kernel's step: all splats x directions [first_dir, first_dir + 16).

    Returns (loss, grad_params, grad_latents (n, 8) summed over the 16 directions).
    With n_total, the splats given are a subset of a scene of n_total splats:
    grad_latents are those splats' rows of the whole scene's step (the loss and
    grad_params are then partial sums over the subset)."""
    dirs = fibonacci_dirs() if dirs is None else dirs
    n = latents.shape[0]
    dsel = dirs[first_dir:first_dir + 16].astype(np.float64)
    t = sh_targets(f_dc, sh_rest, deg, dsel)  # (n, 16, 3)
    z = np.repeat(np.asarray(latents, np.float64), 16, axis=0)
    d = np.tile(dsel, (n, 1))
    loss, gp, gz = loss_and_grads(p, z, d, t.reshape(-1, 3), bf16,
                                  None if n_total is None else 16 * n_total)
    return loss, gp, gz.reshape(n, 16, 8).sum(axis=1)
