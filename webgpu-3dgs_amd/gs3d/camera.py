"""Camera matrices and the 288-B CameraUniforms block.

Mirrors ``src/Camera.ts`` (constructor :33-67, updateView/setProjection :69-79,
getUniformData :172-191). The reference builds its matrices with gl-matrix
3.4.4 (``mat4.lookAt``, ``mat4.perspective`` = the OpenGL-style "NO" mapping of
view z in [-near, -far] to NDC z in [-1, 1], ``mat4.invert``), computing in
float64 (JS numbers) and storing float32 (gl-matrix's Float32Array). This
module does the same arithmetic in float64 and stores float32.
"""
from __future__ import annotations

import math

import numpy as np

EPSILON = 1e-6  # gl-matrix glMatrix.EPSILON


def _f32(v) -> np.ndarray:
    return np.asarray(v, dtype=np.float32)


def look_at(eye, center, up) -> np.ndarray:
    """Column-major view matrix (16 x f32); identity when eye == center."""
    ex, ey, ez = (float(v) for v in _f32(eye))
    cx, cy, cz = (float(v) for v in _f32(center))
    ux, uy, uz = (float(v) for v in _f32(up))
    if abs(ex - cx) < EPSILON and abs(ey - cy) < EPSILON and abs(ez - cz) < EPSILON:
        return np.eye(4, dtype=np.float32).reshape(16)
    z0, z1, z2 = ex - cx, ey - cy, ez - cz
    ln = 1.0 / math.hypot(z0, z1, z2)
    z0, z1, z2 = z0 * ln, z1 * ln, z2 * ln
    x0, x1, x2 = uy * z2 - uz * z1, uz * z0 - ux * z2, ux * z1 - uy * z0
    ln = math.hypot(x0, x1, x2)
    if not ln:
        x0 = x1 = x2 = 0.0
    else:
        x0, x1, x2 = x0 / ln, x1 / ln, x2 / ln
    y0, y1, y2 = z1 * x2 - z2 * x1, z2 * x0 - z0 * x2, z0 * x1 - z1 * x0
    ln = math.hypot(y0, y1, y2)
    if not ln:
        y0 = y1 = y2 = 0.0
    else:
        y0, y1, y2 = y0 / ln, y1 / ln, y2 / ln
    return _f32([
        x0, y0, z0, 0.0,
        x1, y1, z1, 0.0,
        x2, y2, z2, 0.0,
        -(x0 * ex + x1 * ey + x2 * ez), -(y0 * ex + y1 * ey + y2 * ez),
        -(z0 * ex + z1 * ey + z2 * ez), 1.0,
    ])


def perspective(fovy: float, aspect: float, near: float, far: float) -> np.ndarray:
    """Column-major OpenGL-style projection (NDC z in [-1, 1]); far=inf allowed."""
    f = 1.0 / math.tan(fovy / 2.0)
    m = [0.0] * 16
    m[0] = f / aspect
    m[5] = f
    m[11] = -1.0
    if far is not None and far != math.inf:
        nf = 1.0 / (near - far)
        m[10] = (far + near) * nf
        m[14] = 2.0 * far * near * nf
    else:
        m[10] = -1.0
        m[14] = -2.0 * near
    return _f32(m)


def invert(m: np.ndarray) -> np.ndarray:
    """float64 inverse of a column-major mat4, stored float32 (zeros if singular)."""
    a = np.asarray(m, np.float64).reshape(4, 4).T  # row-major view of the matrix
    try:
        inv = np.linalg.inv(a)
    except np.linalg.LinAlgError:
        return np.zeros(16, np.float32)
    return _f32(inv.T.reshape(16))


class Camera:
    """Position/target/up camera with the reference's uniform layout."""

    def __init__(self, position=(10.0, 10.0, 10.0), target=(0.0, 0.0, 0.0), up=(0.0, 1.0, 0.0),
                 fov: float = math.pi / 2.0, aspect: float = 1.0, near: float = 0.1,
                 far: float = 10000.0):
        self.position = _f32(position).copy()
        self.target = _f32(target).copy()
        self.up = _f32(up).copy()
        self.fov, self.aspect, self.near, self.far = fov, aspect, near, far
        self.version = 0
        self.update_view()
        self.set_projection()

    def update_view(self) -> None:
        self.view = look_at(self.position, self.target, self.up)
        self.inverse_view = invert(self.view)
        self.version += 1

    def set_projection(self) -> None:
        self.projection = perspective(self.fov, self.aspect, self.near, self.far)
        self.inverse_projection = invert(self.projection)
        self.version += 1

    def set_position(self, position) -> None:
        self.position = _f32(position).copy()
        self.update_view()

    def set_aspect(self, aspect: float) -> None:
        self.aspect = aspect
        self.set_projection()

    def uniform_data(self, width: int, height: int) -> np.ndarray:
        """72 x f32 CameraUniforms (Camera.getUniformData, Camera.ts:172-191)."""
        d = np.zeros(72, np.float32)
        d[0:16] = self.view
        d[16:32] = self.projection
        d[32:48] = self.inverse_view
        d[48:64] = self.inverse_projection
        d[64:67] = self.position
        d[67] = 1.0
        d[68] = width
        d[69] = height
        d[70] = np.float32(width / height)
        d[71] = 0.0
        return d


def view_camera(eye, width: int, height: int, fov_y: float = math.pi / 3.0,
                target=(0.0, 0.0, 0.0)) -> np.ndarray:
    """CameraUniforms for a camera at `eye` looking at `target`, aspect W/H,
    near/far 0.1/10000 (the reference's defaults, Viewport.ts:58-66)."""
    cam = Camera(eye, target, (0.0, 1.0, 0.0), fov_y, width / height, 0.1, 10000.0)
    return cam.uniform_data(width, height)
