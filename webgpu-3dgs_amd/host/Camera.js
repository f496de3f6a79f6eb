'use strict';
// The host camera: the view and projection one rebin needs, and the 288-byte
// CameraUniforms block that preprocess reads (the block's layout is the
// reference's Camera.getUniformData, src/Camera.ts:172-191, read by
// preprocess_splats.wgsl:1-8). The arithmetic is gs3d/camera.py's: a look-at
// view, an OpenGL-style perspective (NDC z in [-1, 1]) and both inverses,
// computed in doubles and stored as float32. `version` changes with every
// change of view or projection; GaussianSplatRenderer.render compares it to
// skip re-binning a still camera.
//
// Mouse-driven orbit / pan / zoom belongs to the application around the
// renderer; setOrbit places the camera on a sphere around its target, the
// camera placement the workloads and the bench use.

const { mat4 } = require('./glmath');

const vec = (v) => Float32Array.of(v[0], v[1], v[2]);

class Camera {
  constructor(position, target, up, fov, aspect, near, far) {
    this.position = vec(position);
    this.target = vec(target);
    this.up = vec(up);
    this.fov = fov;
    this.aspect = aspect;
    this.near = near;
    this.far = far;
    this.viewMatrix = new Float32Array(16);
    this.inverseViewMatrix = new Float32Array(16);
    this.projectionMatrix = new Float32Array(16);
    this.inverseProjectionMatrix = new Float32Array(16);
    this.version = 0;
    this.updateView();
    this.setProjection();
  }

  updateView() {
    mat4.lookAt(this.viewMatrix, this.position, this.target, this.up);
    if (!mat4.invert(this.inverseViewMatrix, this.viewMatrix)) this.inverseViewMatrix.fill(0);
    this.version++;
  }

  setProjection() {
    mat4.perspective(this.projectionMatrix, this.fov, this.aspect, this.near, this.far);
    if (!mat4.invert(this.inverseProjectionMatrix, this.projectionMatrix)) this.inverseProjectionMatrix.fill(0);
    this.version++;
  }

  setAspect(aspect) {
    this.aspect = aspect;
    this.setProjection();
  }

  setPosition(position) {
    this.position = vec(position);
    this.updateView();
  }

  setTarget(target) {
    this.target = vec(target);
    this.updateView();
  }

  // On the sphere of `radius` around the target: azimuth about +y measured
  // from +z towards +x, elevation above the xz-plane (degrees); the same
  // placement as gs3d.scene.orbit_eye around the origin.
  setOrbit(radius, azimuthDeg, elevationDeg) {
    const az = (azimuthDeg * Math.PI) / 180;
    const el = (elevationDeg * Math.PI) / 180;
    const t = this.target;
    this.position = Float32Array.of(t[0] + radius * Math.cos(el) * Math.sin(az), t[1] + radius * Math.sin(el),
      t[2] + radius * Math.cos(el) * Math.cos(az));
    this.updateView();
  }

  // CameraUniforms, 72 floats: view, projection, inverse view, inverse
  // projection (column-major), position.xyz1, viewport (w, h, w / h, 0).
  getUniformData(viewportWidth, viewportHeight) {
    const u = new Float32Array(72);
    u.set(this.viewMatrix, 0);
    u.set(this.projectionMatrix, 16);
    u.set(this.inverseViewMatrix, 32);
    u.set(this.inverseProjectionMatrix, 48);
    u.set(this.position, 64);
    u[67] = 1;
    u[68] = viewportWidth;
    u[69] = viewportHeight;
    u[70] = viewportWidth / viewportHeight;
    u[71] = 0;
    return u;
  }
}

module.exports = { Camera };
