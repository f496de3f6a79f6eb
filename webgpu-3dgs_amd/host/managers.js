'use strict';
// The renderer's resource registries, over libgs3d.so through the N-API addon.
// They keep the names the reference's renderer uses for them
// (src/WorkgroupManager.ts, src/BufferManager.ts, src/PipelineManager.ts);
// what each one holds is this library's:
//   WorkgroupManager  launch geometry. The library owns its launches; a
//                     registered pass reports the workgroup and dispatch sizes
//                     the library launches for its problem size and the
//                     context's tile size and kernel variants.
//   BufferManager     the context's named device buffers.
//   PipelineManager   the choice of ahead-of-time compiled kernel variants (the
//                     reference compiles WGSL with __NAME__ constants
//                     substituted in; here each constant picks a variant).

// ---- WorkgroupManager -------------------------------------------------------

const LIMIT_NAMES = ['maxTotalThreads', 'maxSizeX', 'maxSizeY', 'maxSizeZ', 'maxDispatch', 'maxSharedMemory'];

class WorkgroupManager {
  // ctx null: the library's default variants (no device needed).
  constructor(addon, ctx = null) {
    this.addon = addon;
    this.ctx = ctx;
    this.passes = new Map();  // name -> {kernel, problemSize}
    this.limits = null;
    if (ctx) {
      const l = addon.deviceLimits(ctx);
      this.limits = Object.freeze(Object.fromEntries(LIMIT_NAMES.map((k, i) => [k, l[i]])));
    }
  }

  static fromDevice(addon, ctx) {
    return new WorkgroupManager(addon, ctx);
  }

  // A pass of the library by kernel name (gs3d_launch_geometry lists them) and
  // its problem size: [splats] or [keys] for the 1-D passes, [width, height]
  // for the per-pixel ones. An unknown kernel throws.
  register({ name, kernel = name, problemSize }) {
    const entry = { kernel, problemSize: Array.from(problemSize) };
    this.addon.launchGeometry(this.ctx, entry.kernel, entry.problemSize);
    this.passes.set(name, entry);
  }

  update(name, { problemSize }) {
    const entry = this.passes.get(name);
    if (!entry) throw new Error(`No workgroup layout registered for '${name}'`);
    entry.problemSize = Array.from(problemSize);
  }

  // {workgroupSize: [x, y, z], dispatchSize: [x, y, z]} as the library launches it now
  getLayout(name) {
    const entry = this.passes.get(name);
    if (!entry) throw new Error(`No workgroup layout registered for '${name}'`);
    return this.addon.launchGeometry(this.ctx, entry.kernel, entry.problemSize);
  }

  layouts() {
    const out = new Map();
    for (const name of this.passes.keys()) out.set(name, this.getLayout(name));
    return out;
  }
}

// ---- BufferManager ----------------------------------------------------------

class BufferManager {
  constructor(addon, ctx) {
    this.addon = addon;
    this.ctx = ctx;
    this.resizeListeners = [];
  }

  // a live view of the registry: name -> {name, size}
  get buffers() {
    const m = new Map();
    for (const { name, size } of this.addon.bufferList(this.ctx)) m.set(name, { name, size });
    return m;
  }

  initBuffers(configs) {
    for (const { name, size, data } of configs) {
      this.addon.bufferCreate(this.ctx, name, Math.max(size, 4));
      if (data) this.write(name, data);
    }
  }

  createBuffer(name, requestedSize) {
    this.addon.bufferCreate(this.ctx, name, requestedSize);
    return this.get(name);
  }

  // Clipped to the buffer; a write with nothing left warns and is skipped.
  write(name, data, offset = 0, dataOffset = 0, size) {
    if (this.addon.bufferSize(this.ctx, name) === undefined) return;
    const bytes = ArrayBuffer.isView(data)
      ? new Uint8Array(data.buffer, data.byteOffset + dataOffset * (data.BYTES_PER_ELEMENT || 1),
        size !== undefined ? size : data.byteLength - dataOffset * (data.BYTES_PER_ELEMENT || 1))
      : new Uint8Array(data, dataOffset, size !== undefined ? size : data.byteLength - dataOffset);
    const written = this.addon.bufferWrite(this.ctx, name, offset, bytes);
    if (written === 0) console.warn(`[BufferManager] Write to '${name}' skipped: size exceeds buffer limit.`);
  }

  get(name) {
    const size = this.addon.bufferSize(this.ctx, name);
    return size === undefined ? undefined : { name, size };
  }

  read(name, offset = 0, bytes) {
    const size = this.addon.bufferSize(this.ctx, name);
    if (size === undefined) return undefined;
    return this.addon.bufferRead(this.ctx, name, offset, bytes === undefined ? size - offset : bytes);
  }

  resize(name, newSize) {
    if (this.addon.bufferSize(this.ctx, name) === undefined) return undefined;
    const clamped = Math.max(newSize, 4);
    this.addon.bufferResize(this.ctx, name, clamped);
    this.emitResize(name, clamped);
    return this.get(name);
  }

  clear(name) {
    const size = this.addon.bufferSize(this.ctx, name);
    if (size === undefined) return;
    this.addon.bufferWrite(this.ctx, name, 0, new Uint8Array(size));
  }

  destroy(name) {
    this.addon.bufferDestroy(this.ctx, name);
  }

  onResize(listener) {
    this.resizeListeners.push(listener);
  }

  emitResize(name, newSize) {
    for (const l of this.resizeListeners) l(name, newSize);
  }
}

// ---- PipelineManager --------------------------------------------------------

// The compiled variants: per kernel, the values each constant may take.
const KERNEL_VARIANTS = {
  'preprocess-splats': {},
  'prefix-scan': {},
  'emit-tile-refs': {},
  'radix-sort': { RADIX_BITS: [4, 8], KEY_TILE: [0, 8192, 16384] },
  'radix-sort-splats': { KEY_TILE: [0, 4096, 8192] },
  'identify-tile-ranges': {},
  'rasterize-tiles': {
    TILE_X: [8, 16, 32], TILE_Y: [8, 16, 32], TILE_AREA: [64, 256, 1024], PIXELS_PER_LANE: [1, 2, 4], QUARTER_LISTS: [0, 1, 2],
  },
  'rasterize-tiles-batch': { GROUPS: [...Array(17).keys()], WAVE_ITEMS: [0, 1], WG_PER_CU: [...Array(17).keys()] },
};

// Constants that select a variant in the library (gs3d_set_kernel_variant);
// the tile constants follow the renderer's tile size (gs3d_set_tile_size).
const SELECTED_IN_LIBRARY = new Set([
  'RADIX_BITS', 'KEY_TILE', 'PIXELS_PER_LANE', 'QUARTER_LISTS', 'GROUPS', 'WAVE_ITEMS', 'WG_PER_CU',
]);

function checkConstants(name, kernel, constants) {
  const table = KERNEL_VARIANTS[kernel];
  if (!table) throw new Error(`No compiled kernel for pipeline "${name}"`);
  for (const [k, v] of Object.entries(constants)) {
    if (!table[k] || !table[k].includes(v)) throw new Error(`Pipeline "${name}": no compiled variant with ${k}=${v}`);
  }
  if ('TILE_X' in constants && constants.TILE_X !== constants.TILE_Y) {
    throw new Error(`Pipeline "${name}": tiles are square (8, 16 or 32)`);
  }
}

class PipelineManager {
  // With an addon and context, a pipeline's constants are applied to the
  // library when it is created or updated.
  constructor(addon = null, ctx = null) {
    this.addon = addon;
    this.ctx = ctx;
    this.selections = new Map();  // name -> {label, kernel, constants}
  }

  create({ name, kernel = name, codeConstants = {} }) {
    const constants = { ...codeConstants };
    checkConstants(name, kernel, constants);
    if (this.addon && this.ctx) {
      for (const [k, v] of Object.entries(constants)) {
        if (SELECTED_IN_LIBRARY.has(k)) this.addon.setKernelVariant(this.ctx, kernel, k, v);
      }
    }
    const selection = { label: `pipeline-${name}`, kernel, constants };
    this.selections.set(name, selection);
    return selection;
  }

  // New constants merged over the pipeline's current ones.
  update(name, { codeConstants = {} } = {}) {
    const current = this.selections.get(name);
    if (!current) throw new Error(`No pipeline config found for "${name}"`);
    return this.create({ name, kernel: current.kernel, codeConstants: { ...current.constants, ...codeConstants } });
  }

  get(name) {
    return this.selections.get(name);
  }

  has(name) {
    return this.selections.has(name);
  }

  delete(name) {
    return this.selections.delete(name);
  }

  clear() {
    this.selections.clear();
  }
}

module.exports = { WorkgroupManager, BufferManager, PipelineManager, KERNEL_VARIANTS };
