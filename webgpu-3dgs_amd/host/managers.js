'use strict';
// Host-side resource services with the reference's names and semantics:
//   WorkgroupManager + linear1D/tiled1D/tile2D  (src/WorkgroupManager.ts:10-147)
//   BufferManager                               (src/BufferManager.ts:13-180)
//   PipelineManager                             (src/PipelineManager.ts:27-117)
// Device memory and kernels live in libgs3d.so; these classes keep the
// reference's registry-by-name API over the N-API addon.

// ---- WorkgroupManager -------------------------------------------------------

const highestPowerOfTwoAtMost = (v) => 2 ** Math.floor(Math.log2(Math.max(1, v)));

const linear1D = (preferredSize = 256, elementsPerThread = 1) => ({ limits, problemSize }) => {
  const x = highestPowerOfTwoAtMost(Math.min(preferredSize, limits.maxTotalThreads, limits.maxSizeX));
  return { workgroupSize: [x, 1, 1], dispatchSize: [Math.max(1, Math.ceil(problemSize[0] / (elementsPerThread * x))), 1, 1] };
};

const tile2D = (tileX, tileY) => ({ problemSize }) => ({
  workgroupSize: [tileX, tileY, 1],
  dispatchSize: [Math.max(1, Math.ceil(problemSize[0] / tileX)), Math.max(1, Math.ceil(problemSize[1] / tileY)), 1],
});

const tiled1D = (elementsPerThread, bytesPerElement, preferredSize = Infinity) => ({ limits, problemSize }) => {
  let x = highestPowerOfTwoAtMost(Math.min(preferredSize, limits.maxTotalThreads, limits.maxSizeX));
  while (x > 1 && elementsPerThread * x * bytesPerElement > limits.maxSharedMemory) x /= 2;
  return { workgroupSize: [x, 1, 1], dispatchSize: [Math.max(1, Math.ceil(problemSize[0] / (elementsPerThread * x))), 1, 1] };
};

class WorkgroupManager {
  // limits: {maxTotalThreads, maxSizeX, maxSizeY, maxSizeZ, maxDispatch, maxSharedMemory}
  constructor(limits) {
    this.limits = Object.freeze({ ...limits });
    this.strategies = new Map();
    this.layouts = new Map();
  }

  static fromDevice(addon, ctx) {
    const l = addon.deviceLimits(ctx);
    return new WorkgroupManager({
      maxTotalThreads: l[0], maxSizeX: l[1], maxSizeY: l[2], maxSizeZ: l[3], maxDispatch: l[4], maxSharedMemory: l[5],
    });
  }

  register(config) {
    this.strategies.set(config.name, { ...config });
    this.layouts.set(config.name, this.computeLayout(config.problemSize, config.strategyFn(...config.strategyArgs)));
  }

  update(name, updates) {
    const config = this.strategies.get(name);
    if (!config) throw new Error(`No strategy registered for '${name}'`);
    if (updates.problemSize) config.problemSize = updates.problemSize;
    if (updates.strategyArgs) config.strategyArgs = updates.strategyArgs;
    this.layouts.set(name, this.computeLayout(config.problemSize, config.strategyFn(...config.strategyArgs)));
  }

  getLayout(name) {
    const layout = this.layouts.get(name);
    if (!layout) throw new Error(`No layout found for '${name}'`);
    return layout;
  }

  computeLayout(problemSize, strategy) {
    let result;
    try {
      result = strategy({ limits: this.limits, problemSize });
    } catch (err) {
      if (err instanceof Error) throw new Error(`Invalid workgroup layout: ${err.message}`);
      throw err;
    }
    const [x, y, z] = result.workgroupSize;
    if (x > this.limits.maxSizeX || y > this.limits.maxSizeY || z > this.limits.maxSizeZ ||
        x * y * z > this.limits.maxTotalThreads) {
      throw new Error('Workgroup layout exceeds hardware limits.');
    }
    const d = result.dispatchSize ||
      [Math.ceil(problemSize[0] / x), Math.ceil(problemSize[1] / y), Math.ceil(problemSize[2] / z)];
    return {
      workgroupSize: [x, y, z],
      dispatchSize: d.map((v) => Math.min(v, this.limits.maxDispatch)),
    };
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
// WGSL text substitution (__NAME__) becomes the choice of a kernel variant that
// libgs3d.so compiled ahead of time. The registry lists the variants.

const KERNEL_VARIANTS = {
  'preprocess-splats': { constants: {} },
  'prefix-scan': { constants: {} },
  'emit-tile-refs': { constants: {} },
  'radix-sort': { constants: { RADIX_BITS: [4, 8], KEY_TILE: [0, 8192, 16384] } },
  'identify-tile-ranges': { constants: {} },
  'rasterize-tiles': {
    constants: { TILE_X: [8, 16, 32], TILE_Y: [8, 16, 32], TILE_AREA: [64, 256, 1024], PIXELS_PER_LANE: [1, 2, 4] },
  },
};

// Constants the library selects by gs3d_set_kernel_variant; the tile constants
// follow gs3d_set_tile_size instead.
const LIBRARY_CONSTANTS = new Set(['RADIX_BITS', 'KEY_TILE', 'PIXELS_PER_LANE']);

class PipelineManager {
  // With an addon and context, creating or updating a pipeline also selects
  // the matching compiled variant in the library.
  constructor(addon = null, ctx = null) {
    this.addon = addon;
    this.ctx = ctx;
    this.pipelines = new Map();
    this.pipelineConfigs = new Map();
  }

  create(config) {
    const kernel = config.kernel || config.name;
    const variant = KERNEL_VARIANTS[kernel];
    if (!variant) throw new Error(`No compiled kernel for pipeline "${config.name}"`);
    const constants = { ...(config.codeConstants || {}) };
    for (const [k, v] of Object.entries(constants)) {
      const allowed = variant.constants[k];
      if (!allowed || !allowed.includes(v)) {
        throw new Error(`Pipeline "${config.name}": no compiled variant with ${k}=${v}`);
      }
    }
    if ('TILE_X' in constants && constants.TILE_X !== constants.TILE_Y) {
      throw new Error(`Pipeline "${config.name}": tiles are square (8, 16 or 32)`);
    }
    if (this.addon) {
      for (const [k, v] of Object.entries(constants)) {
        if (LIBRARY_CONSTANTS.has(k)) this.addon.setKernelVariant(this.ctx, kernel, k, v);
      }
    }
    const pipeline = { label: `pipeline-${config.name}`, kernel, constants };
    this.pipelines.set(config.name, pipeline);
    this.pipelineConfigs.set(config.name, { ...config, codeConstants: constants });
  }

  applyShaderConstants(src, constants) {
    let out = src;
    for (const [k, v] of Object.entries(constants)) out = out.replace(new RegExp(`__${k}__`, 'g'), String(v));
    return out;
  }

  update(name, updates) {
    const old = this.pipelineConfigs.get(name);
    if (!old) throw new Error(`No pipeline config found for "${name}"`);
    this.create({
      ...old, ...updates,
      constants: { ...old.constants, ...updates.constants },
      codeConstants: { ...old.codeConstants, ...updates.codeConstants },
    });
  }

  get(name) {
    return this.pipelines.get(name);
  }

  delete(name) {
    return this.pipelines.delete(name);
  }

  clear() {
    this.pipelines.clear();
  }
}

module.exports = {
  WorkgroupManager, linear1D, tiled1D, tile2D, BufferManager, PipelineManager, KERNEL_VARIANTS,
};
