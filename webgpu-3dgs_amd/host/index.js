'use strict';
// Node host of the gs3d path: the reference's scene / camera / renderer API
// over libgs3d.so through the N-API addon.

const { Camera } = require('./Camera');
const glmath = require('./glmath');
const managers = require('./managers');
const loader = require('./SceneLoader');
const { GaussianSplatRenderer, loadAddon, BUF, REBIN_PROFILE } = require('./GaussianSplatRenderer');

// The pass labels of the reference's Profiler table, in its execution order.
const REFERENCE_PASSES = ['preprocess-splats', 'prefix-scan-local', 'prefix-scan-blocks', 'prefix-scan-add',
  'emit-tile-refs', 'radix-histogram', 'scan-local', 'scan-add', 'radix-scatter', 'identify-tile-ranges',
  'rasterize-tiles', 'tile-overlay', 'composite-splats'];

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

  // The reference's table (its pass labels, GaussianSplatRenderer.ts:1017-1260):
  // rows of [reference labels, the stage(s) doing their work, microseconds],
  // in execution order; passes fused into one stage here share a row.
  referenceTimings() {
    const { addon } = this.renderer;
    const t = this.timings();
    const labels = [...t.keys()];
    const rows = [];
    for (const pass of REFERENCE_PASSES) {
      const mask = addon.referencePassStages(pass);
      const stages = labels.filter((_, i) => (mask >>> i) & 1);
      const key = stages.join('+');
      const last = rows[rows.length - 1];
      if (last && last.stages === key) last.passes.push(pass);
      else rows.push({ passes: [pass], stages: key, us: stages.reduce((s, l) => s + t.get(l), 0) });
    }
    return rows;
  }

  getBuffersSortedBySize() {
    return this.renderer.addon.bufferList(this.renderer.ctx);
  }

  getTotalBufferSize() {
    return this.getBuffersSortedBySize().reduce((s, b) => s + b.size, 0);
  }
}

module.exports = {
  Camera, Scene, Profiler, GaussianSplatRenderer, loadAddon, BUF, REBIN_PROFILE, REFERENCE_PASSES,
  ...managers, ...loader, glmath,
};
