'use strict';
// Camera with the reference's API and 288-byte uniform block
// (src/Camera.ts: constructor :33-67, updateView/setProjection :69-79,
// setPosition/setTarget :81-91, rotate/pan/zoom :93-170, getUniformData :172-191).

const { vec3, quat, mat4 } = require('./glmath');

class Camera {
  constructor(position, target, up, fov, aspect, near, far) {
    this.position = vec3.clone(position);
    this.target = vec3.clone(target);
    this.worldUp = vec3.clone(up);
    this.up = vec3.clone(up);
    this.right = vec3.create();
    this.fov = fov;
    this.aspect = aspect;
    this.near = near;
    this.far = far;
    this.projectionMatrix = mat4.create();
    this.viewMatrix = mat4.create();
    this.inverseProjectionMatrix = mat4.create();
    this.inverseViewMatrix = mat4.create();
    // bumped on every view/projection change; renderers skip work while it is unchanged
    this.version = 0;
    this.forward = vec3.create();
    vec3.sub(this.forward, this.target, this.position);
    vec3.normalize(this.forward, this.forward);
    this.orientation = quat.create();
    const dir = vec3.normalize(vec3.create(), vec3.sub(vec3.create(), this.target, this.position));
    quat.rotationTo(this.orientation, vec3.fromValues(0, 0, -1), dir);
    this.rotationSpeed = 0.01;
    this.panSpeed = 0.02;
    this.zoomSpeed = 0.01;
    this.updateView();
    this.setProjection();
  }

  updateView() {
    mat4.lookAt(this.viewMatrix, this.position, this.target, this.up);
    mat4.invert(this.inverseViewMatrix, this.viewMatrix);
    this.version++;
  }

  setProjection() {
    mat4.perspective(this.projectionMatrix, this.fov, this.aspect, this.near, this.far);
    mat4.invert(this.inverseProjectionMatrix, this.projectionMatrix);
    this.version++;
  }

  setAspect(aspect) {
    this.aspect = aspect;
    this.setProjection();
  }

  setPosition(position) {
    vec3.copy(this.position, position);
    this.updateVectors();
    this.updateView();
  }

  setTarget(target) {
    vec3.copy(this.target, target);
    this.updateVectors();
    this.updateView();
  }

  rotate(deltaX, deltaY) {
    const yaw = quat.setAxisAngle(quat.create(), vec3.fromValues(0, 1, 0), deltaX * this.rotationSpeed);
    const right = vec3.normalize(vec3.create(), vec3.cross(vec3.create(), this.up, this.forward));
    const pitch = quat.setAxisAngle(quat.create(), right, -deltaY * this.rotationSpeed);
    const rotation = quat.create();
    quat.multiply(rotation, rotation, pitch);
    quat.multiply(rotation, yaw, rotation);
    quat.normalize(this.orientation, rotation);
    const offset = vec3.sub(vec3.create(), this.position, this.target);
    const rotated = vec3.transformQuat(vec3.create(), offset, this.orientation);
    vec3.add(this.position, this.target, rotated);
    vec3.transformQuat(this.up, this.up, this.orientation);
    vec3.transformQuat(this.forward, this.forward, this.orientation);
    this.updateView();
  }

  pan(deltaX, deltaY) {
    const dir = vec3.normalize(vec3.create(), vec3.sub(vec3.create(), this.target, this.position));
    const right = vec3.normalize(vec3.create(), vec3.cross(vec3.create(), this.worldUp, dir));
    const up = vec3.normalize(vec3.create(), vec3.cross(vec3.create(), right, dir));
    const offset = vec3.create();
    vec3.scaleAndAdd(offset, offset, right, -deltaX * this.panSpeed);
    vec3.scaleAndAdd(offset, offset, up, deltaY * this.panSpeed);
    vec3.add(this.position, this.position, offset);
    vec3.add(this.target, this.target, offset);
    this.updateView();
  }

  zoom(deltaZoom) {
    const dir = vec3.normalize(vec3.create(), vec3.sub(vec3.create(), this.target, this.position));
    const distance = vec3.distance(this.position, this.target);
    let amount = deltaZoom * this.zoomSpeed * (distance / 10);
    if (distance - amount < this.near) amount = distance - this.near;
    vec3.scale(dir, dir, amount);
    vec3.add(this.position, this.position, dir);
    this.updateView();
  }

  // 72 floats: view, projection, inverse view, inverse projection, position.xyz1, viewport
  getUniformData(viewportWidth, viewportHeight) {
    const data = new Float32Array(72);
    data.set(this.viewMatrix, 0);
    data.set(this.projectionMatrix, 16);
    data.set(this.inverseViewMatrix, 32);
    data.set(this.inverseProjectionMatrix, 48);
    data[64] = this.position[0];
    data[65] = this.position[1];
    data[66] = this.position[2];
    data[67] = 1.0;
    data[68] = viewportWidth;
    data[69] = viewportHeight;
    data[70] = viewportWidth / viewportHeight;
    data[71] = 0.0;
    return data;
  }

  updateVectors() {
    vec3.sub(this.forward, this.target, this.position);
    vec3.normalize(this.forward, this.forward);
    vec3.cross(this.right, this.forward, this.worldUp);
    vec3.normalize(this.right, this.right);
    vec3.cross(this.up, this.right, this.forward);
    vec3.normalize(this.up, this.up);
  }
}

module.exports = { Camera };
