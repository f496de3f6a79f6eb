'use strict';
// GaussianSplatRenderer with the reference's API (src/renderers/GaussianSplatRenderer.ts):
// init (:331-471), render with the rebin gate (:1033-1098), rebin (:1104-1266),
// setTileSize / clampTileSize (:301-317, :162-167), setDebugMode (:321-323),
// setAlwaysRebin (:327-329), getTileCount / getTileSize (:1268-1274),
// readTileCounts (:1278-1310) and the display step of render (heat map +
// composite, :1075-1097). The passes run in libgs3d.so on one HIP stream;
// this class keeps the host-side state machine and the resource registries.

const { WorkgroupManager, BufferManager, PipelineManager } = require('./managers');

const REBIN_PROFILE = 1;
const BATCH_EAGER = 1;
const BATCH_LEARNED = 2;
const BUF = {
  PROJECTED_SPLATS: 0, SPLAT_REF_COUNTS: 1, SPLAT_REF_OFFSETS: 2, REF_COUNTER: 3, SORT_KEYS: 4,
  SORT_VALUES: 5, TILE_OFFSETS: 6, IMAGE_RGBA16F: 7, IMAGE_RGBA32F: 8,
  UNSORTED_KEYS: 9, UNSORTED_VALUES: 10, DEBUG_TARGET_RGBA16F: 11, FRAMEBUFFER: 12,
};
// Framebuffer byte orders of composite (the swapchain format handed to init).
const FORMATS = { rgba8unorm: 0, bgra8unorm: 1 };
const MAX_TILES_PER_SPLAT = 16;  // Config.MAX_TILES_PER_SPLAT: the reference's key budget per splat
// passes sized by splat count, by key count and by viewport
const SPLAT_PASSES = ['preprocess-splats', 'radix-histogram', 'radix-sort-splats', 'prefix-scan', 'emit-tile-refs'];
const PIXEL_PASSES = ['identify-tile-ranges', 'rasterize-tiles', 'composite-splats'];

function loadAddon() {
  try {
    return require('../addon/gs3d_napi.node');
  } catch (e) {
    throw new Error(`gs3d N-API addon not built (webgpu-3dgs_amd/addon/gs3d_napi.node): ${e.message}`);
  }
}

class GaussianSplatRenderer {
  constructor(scene, { device = 0, addon = loadAddon(), profiler = null } = {}) {
    this.addon = addon;
    this.ctx = addon.create(device);
    this.scene = scene;
    this.bufferManager = new BufferManager(addon, this.ctx);
    this.pipelineManager = new PipelineManager(addon, this.ctx);
    this.workgroups = WorkgroupManager.fromDevice(addon, this.ctx);
    this.profiler = profiler;
    this.tileSizeX = 16;
    this.tileSizeY = 16;
    this.tileCountX = 1;
    this.tileCountY = 1;
    this.lastViewport = [-1, -1];
    this.lastCameraVersion = -1;
    this.alwaysRebin = false;
    this.debugMode = 0;
    this.debugRef = 256;
    this.lastSplatCount = -1;
    this.uploadedSplats = null;
    this.rebinCount = 0;
  }

  // The passes' launch layouts (registerBinningLayouts, :127-158): the
  // library's own launch geometry for the current problem sizes. The key
  // count starts at the reference's 16-per-splat budget and follows the last
  // rebin's K when launchLayouts() is read.
  registerBinningLayouts(splatCount) {
    for (const name of SPLAT_PASSES) this.workgroups.register({ name, problemSize: [splatCount] });
    this.workgroups.register({ name: 'radix-sort', problemSize: [Math.max(1, splatCount) * MAX_TILES_PER_SPLAT] });
    [this.tileSizeX, this.tileSizeY] = this.clampTileSize(this.scene.tiles[0], this.scene.tiles[1]);
    for (const name of PIXEL_PASSES) this.workgroups.register({ name, problemSize: [1, 1] });
  }

  // name -> {workgroupSize, dispatchSize} of every pass, as the library launches
  // them for the last frame (waits for the last rebin to read its key count).
  launchLayouts() {
    if (this.rebinCount > 0) this.workgroups.update('radix-sort', { problemSize: [this.refCount()] });
    return this.workgroups.layouts();
  }

  clampTileSize(x, y) {
    const lim = this.workgroups.limits;
    const tx = Math.max(1, Math.min(Math.floor(x) || 1, lim.maxSizeX, lim.maxTotalThreads));
    const ty = Math.max(1, Math.min(Math.floor(y) || 1, lim.maxSizeY, Math.floor(lim.maxTotalThreads / tx)));
    return [tx, ty];
  }

  init(format = 'rgba8unorm') {
    if (!(format in FORMATS)) throw new Error(`GaussianSplatRenderer.init: unsupported format ${format}`);
    this.format = format;
    const splatCount = Math.max(1, (this.scene.splats && this.scene.splats.splatCount) || 1);
    this.registerBinningLayouts(splatCount);
    this.lastSplatCount = -1;
    for (const name of ['preprocess-splats', 'prefix-scan', 'emit-tile-refs', 'identify-tile-ranges']) {
      this.pipelineManager.create({ name, type: 'compute' });
    }
    this.pipelineManager.create({ name: 'radix-sort', type: 'compute', codeConstants: { RADIX_BITS: 8 } });
    this.pipelineManager.create({
      name: 'rasterize-tiles', type: 'compute',
      codeConstants: { TILE_X: this.tileSizeX, TILE_Y: this.tileSizeY, TILE_AREA: this.tileSizeX * this.tileSizeY },
    });
    this.addon.setTileSize(this.ctx, this.tileSizeX, this.tileSizeY);
  }

  // SceneSyncer.setSplatData (src/SceneSyncer.ts:25-43): upload the loader's arrays.
  setSplatData(splats) {
    const deg = splats.sphericalHarmonicsDegree || 0;
    this.addon.setSplats(this.ctx, splats.positions, splats.scales, splats.rotations, splats.colors,
      deg > 0 ? splats.fDc : null, deg > 0 ? splats.sphericalHarmonics : null, deg, splats.splatCount);
    this.uploadedSplats = splats;
    this.lastSplatCount = -1;
  }

  setTileSize(x, y) {
    const [tx, ty] = this.clampTileSize(x, y);
    if (tx === this.tileSizeX && ty === this.tileSizeY) return [tx, ty];
    this.pipelineManager.update('rasterize-tiles', { codeConstants: { TILE_X: tx, TILE_Y: ty, TILE_AREA: tx * ty } });
    this.addon.setTileSize(this.ctx, tx, ty);
    this.tileSizeX = tx;
    this.tileSizeY = ty;
    this.scene.tiles = [tx, ty];
    this.lastViewport = [-1, -1];  // force the tile grid + target to be recomputed
    return [tx, ty];
  }

  // 0 = normal, 1 = splats-per-tile heat map (any other value draws normally,
  // as in the reference, which tests debugMode === 1).
  setDebugMode(mode) {
    this.debugMode = mode | 0;
    this.addon.setDebugMode(this.ctx, this.debugMode === 1 ? 1 : 0, this.debugRef);
  }

  setAlwaysRebin(on) {
    this.alwaysRebin = !!on;
  }

  // frame: {width, height, camera (Camera), cameraVersion?, colorTexture?, clearValue?}
  // colorTexture: a Uint8Array of width*height*4 bytes in the init format; when
  // given, the frame is composited onto it (premultiplied over, in place), after
  // clearing it to clearValue ({r, g, b, a}) if that is set.
  render(frame) {
    const splats = this.scene.splats;
    if (!splats) return false;
    if (this.uploadedSplats !== splats) this.setSplatData(splats);
    const splatCount = splats.splatCount;
    const width = Math.max(1, frame.width | 0);
    const height = Math.max(1, frame.height | 0);
    const tilesX = Math.ceil(width / this.tileSizeX);
    const tilesY = Math.ceil(height / this.tileSizeY);
    this.tileCountX = tilesX;
    this.tileCountY = tilesY;
    const splatCountChanged = splatCount !== this.lastSplatCount;
    if (splatCountChanged) {
      for (const name of SPLAT_PASSES) this.workgroups.update(name, { problemSize: [splatCount] });
      this.lastSplatCount = splatCount;
    }
    const viewportChanged = width !== this.lastViewport[0] || height !== this.lastViewport[1];
    if (viewportChanged) {
      this.lastViewport = [width, height];
      for (const name of PIXEL_PASSES) this.workgroups.update(name, { problemSize: [width, height] });
    }
    const version = frame.cameraVersion !== undefined ? frame.cameraVersion : frame.camera.version;
    const cameraMoved = version !== this.lastCameraVersion;
    this.lastCameraVersion = version;
    const needsRebin = this.alwaysRebin || cameraMoved || splatCountChanged || viewportChanged;
    if (needsRebin) this.rebin(frame.camera, splatCount, tilesX, tilesY, width, height);
    if (frame.colorTexture) this.composite(frame.colorTexture, frame.clearValue);
    return needsRebin;
  }

  // Heat map (debug mode 1) + composite of the current frame onto `target`.
  composite(target, clearValue = null) {
    const clear = clearValue ? new Float32Array([clearValue.r, clearValue.g, clearValue.b, clearValue.a]) : null;
    this.addon.composite(this.ctx, FORMATS[this.format || 'rgba8unorm'], clear, target);
  }

  rebin(camera, splatCount, tilesX, tilesY, width, height) {
    const flags = this.profiler ? REBIN_PROFILE : 0;
    this.addon.rebin(this.ctx, camera.getUniformData(width, height), width, height, flags);
    this.rebinCount++;
  }

  // Many views in one call (no reference counterpart): frames = [{width, height, camera}],
  // rendered by one replayed hipGraph; learned: colour with the learned stage when it
  // passed validation (exact SH otherwise). Returns the K of each view.
  renderBatch(frames, { learned = false, eager = false } = {}) {
    const splats = this.scene.splats;
    if (!splats) return null;
    if (this.uploadedSplats !== splats) this.setSplatData(splats);
    const views = frames.map((f) => ({ camera: f.camera.getUniformData(f.width, f.height), width: f.width, height: f.height }));
    this.addon.rebinBatch(this.ctx, views, (eager ? BATCH_EAGER : 0) | (learned ? BATCH_LEARNED : 0));
    this.batchViews = frames.length;
    return this.addon.batchRefCounts(this.ctx, frames.length);
  }

  readBatchImage(view, asFloat32 = true) {
    const ab = this.addon.batchRead(this.ctx, view, asFloat32 ? BUF.IMAGE_RGBA32F : BUF.IMAGE_RGBA16F);
    return asFloat32 ? new Float32Array(ab) : new Uint16Array(ab);
  }

  // Learned colour: train on the uploaded splats' exact SH colours, then check the
  // learned render against exact SH on `frames`; enabled iff PSNR >= minPsnrDb.
  trainLearnedColor({ epochs = 8, lr = 3e-3, lrLatent = 3e-2, seed = 0 } = {}) {
    if (this.uploadedSplats !== this.scene.splats) this.setSplatData(this.scene.splats);
    this.addon.learnedInit(this.ctx, seed);
    return this.addon.learnedTrain(this.ctx, epochs, lr, lrLatent);
  }

  validateLearnedColor(frames, minPsnrDb = 45) {
    const views = frames.map((f) => ({ camera: f.camera.getUniformData(f.width, f.height), width: f.width, height: f.height }));
    return this.addon.learnedValidate(this.ctx, views, minPsnrDb);
  }

  getTileCount() {
    return [this.tileCountX, this.tileCountY];
  }

  getTileSize() {
    return [this.tileSizeX, this.tileSizeY];
  }

  // Per-tile splat counts with the underflow guard; null on failure.
  async readTileCounts() {
    try {
      return this.addon.readTileCounts(this.ctx);
    } catch (e) {
      return null;
    }
  }

  readBuffer(id) {
    return this.addon.read(this.ctx, id);
  }

  // Premultiplied (C, 1 - T) target, rgba16float raw halves or widened to f32.
  readImage(asFloat32 = true) {
    const ab = this.addon.read(this.ctx, asFloat32 ? BUF.IMAGE_RGBA32F : BUF.IMAGE_RGBA16F);
    return asFloat32 ? new Float32Array(ab) : new Uint16Array(ab);
  }

  refCount() {
    return this.addon.refCount(this.ctx);
  }

  sync() {
    this.addon.sync(this.ctx);
  }

  destroy() {
    if (this.ctx) this.addon.destroy(this.ctx);
    this.ctx = null;
  }
}

module.exports = { GaussianSplatRenderer, loadAddon, BUF, FORMATS, REBIN_PROFILE };
