'use strict';
// Column-major 4x4 matrix helpers for the host Camera.
// The reference builds its camera with gl-matrix 3.4.4 (src/Camera.ts:1); the
// functions here are written from the textbook definitions of the same
// operations (look-at view matrix, OpenGL-style perspective with NDC z in
// [-1, 1], cofactor inverse). Arithmetic is in
// JS doubles; results are stored in Float32Array like gl-matrix's defaults.

const EPSILON = 1e-6;

const mat4 = {
  create: () => { const m = new Float32Array(16); m[0] = m[5] = m[10] = m[15] = 1; return m; },
  identity: (o) => { o.fill(0); o[0] = o[5] = o[10] = o[15] = 1; return o; },
  // view matrix of an eye at `eye` looking at `center` with `up` (right-handed, -z forward)
  lookAt: (o, eye, center, up) => {
    const ex = eye[0], ey = eye[1], ez = eye[2];
    if (Math.abs(ex - center[0]) < EPSILON && Math.abs(ey - center[1]) < EPSILON &&
        Math.abs(ez - center[2]) < EPSILON) return mat4.identity(o);
    let z0 = ex - center[0], z1 = ey - center[1], z2 = ez - center[2];
    let l = 1 / Math.hypot(z0, z1, z2);
    z0 *= l; z1 *= l; z2 *= l;
    let x0 = up[1] * z2 - up[2] * z1, x1 = up[2] * z0 - up[0] * z2, x2 = up[0] * z1 - up[1] * z0;
    l = Math.hypot(x0, x1, x2);
    if (!l) { x0 = x1 = x2 = 0; } else { x0 /= l; x1 /= l; x2 /= l; }
    let y0 = z1 * x2 - z2 * x1, y1 = z2 * x0 - z0 * x2, y2 = z0 * x1 - z1 * x0;
    l = Math.hypot(y0, y1, y2);
    if (!l) { y0 = y1 = y2 = 0; } else { y0 /= l; y1 /= l; y2 /= l; }
    o[0] = x0; o[1] = y0; o[2] = z0; o[3] = 0;
    o[4] = x1; o[5] = y1; o[6] = z1; o[7] = 0;
    o[8] = x2; o[9] = y2; o[10] = z2; o[11] = 0;
    o[12] = -(x0 * ex + x1 * ey + x2 * ez);
    o[13] = -(y0 * ex + y1 * ey + y2 * ez);
    o[14] = -(z0 * ex + z1 * ey + z2 * ez);
    o[15] = 1;
    return o;
  },
  // OpenGL-style perspective: view z in [-near, -far] -> NDC z in [-1, 1]
  perspective: (o, fovy, aspect, near, far) => {
    const f = 1 / Math.tan(fovy / 2);
    o.fill(0);
    o[0] = f / aspect; o[5] = f; o[11] = -1;
    if (far != null && far !== Infinity) {
      const nf = 1 / (near - far);
      o[10] = (far + near) * nf; o[14] = 2 * far * near * nf;
    } else {
      o[10] = -1; o[14] = -2 * near;
    }
    return o;
  },
  // inverse by cofactor expansion; returns null for a singular matrix
  invert: (o, a) => {
    const m = Array.from(a);
    const inv = new Array(16);
    inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
    inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
    inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
    inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
    inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
    inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
    inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
    inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
    inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
    inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
    inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
    inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
    inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
    inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
    inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
    inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
    const det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
    if (!det) return null;
    for (let i = 0; i < 16; i++) o[i] = inv[i] / det;
    return o;
  },
};

module.exports = { mat4, EPSILON };
