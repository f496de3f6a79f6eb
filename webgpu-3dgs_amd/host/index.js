'use strict';
// Node host of the gs3d path: the reference's scene / camera / renderer API
// over libgs3d.so through the N-API addon.

const { Camera } = require('./Camera');
const glmath = require('./glmath');
const managers = require('./managers');
const loader = require('./SceneLoader');
const { GaussianSplatRenderer, loadAddon, BUF, REBIN_PROFILE } = require('./GaussianSplatRenderer');

// Scene (src/Scene.ts): tile pixel size, splats, bounds.
class Scene {
  constructor() {
    this.tiles = [16, 16];
    this.splats = null;
    this.bounds = null;
  }

  clear() {
    this.splats = null;
    this.bounds = null;
  }
}

// Profiler (src/Profiler.ts): per-stage GPU times averaged over a window of
// profiled rebins (hipEvent pairs inside libgs3d.so) and the buffer-size panel.
class Profiler {
  constructor(renderer) {
    this.renderer = renderer;
  }

  static profile(label, fn) {
    const t = process.hrtime.bigint();
    const r = fn();
    console.log(`[Profile] ${label}: ${(Number(process.hrtime.bigint() - t) / 1e6).toFixed(2)} ms`);
    return r;
  }

  static async profileAsync(label, fn) {
    const t = process.hrtime.bigint();
    const r = await fn();
    console.log(`[Profile] ${label}: ${(Number(process.hrtime.bigint() - t) / 1e6).toFixed(2)} ms`);
    return r;
  }

  resetWindow() {
    this.renderer.addon.profilerReset(this.renderer.ctx);
  }

  // label -> averaged microseconds, in execution order
  timings() {
    const { addon, ctx } = this.renderer;
    const t = addon.stageTimesAvg(ctx);
    const m = new Map();
    for (let i = 0; i < t.length; i++) m.set(addon.stageLabel(i), t[i]);
    return m;
  }

  getBuffersSortedBySize() {
    return this.renderer.addon.bufferList(this.renderer.ctx);
  }

  getTotalBufferSize() {
    return this.getBuffersSortedBySize().reduce((s, b) => s + b.size, 0);
  }
}

module.exports = {
  Camera, Scene, Profiler, GaussianSplatRenderer, loadAddon, BUF, REBIN_PROFILE,
  ...managers, ...loader, glmath,
};
