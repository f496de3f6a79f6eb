'use strict';
// Gaussian-splat PLY loading with the reference loader's behaviour
// (src/workers/SceneWorker.ts): header parse and validation (:146-272),
// ascii / binary little- and big-endian bodies (:274-333, :508-549), the
// per-row activations (writeGaussianSplat :407-464), f_rest ordering by index
// (:375-388) and the SH degree from the f_rest count (:501-506). The output also
// keeps the raw f_dc, which view-dependent SH colour needs. Parsing runs in a
// worker thread (SceneLoader.ts:75-139) so the caller's thread stays free.

const fs = require('fs');
const path = require('path');

const SH_C0 = 0.28209479177387814;
const REQUIRED = ['x', 'y', 'z', 'opacity', 'scale_0', 'scale_1', 'scale_2',
  'rot_0', 'rot_1', 'rot_2', 'rot_3', 'f_dc_0', 'f_dc_1', 'f_dc_2'];

const SCALAR = {
  char: [1, 'getInt8'], int8: [1, 'getInt8'], uchar: [1, 'getUint8'], uint8: [1, 'getUint8'],
  short: [2, 'getInt16'], int16: [2, 'getInt16'], ushort: [2, 'getUint16'], uint16: [2, 'getUint16'],
  int: [4, 'getInt32'], int32: [4, 'getInt32'], uint: [4, 'getUint32'], uint32: [4, 'getUint32'],
  float: [4, 'getFloat32'], float32: [4, 'getFloat32'], double: [8, 'getFloat64'], float64: [8, 'getFloat64'],
};

function toArrayBuffer(buf) {
  if (buf instanceof ArrayBuffer) return buf;
  return buf.buffer.slice(buf.byteOffset, buf.byteOffset + buf.byteLength);
}

function parsePlyHeader(buffer) {
  const bytes = new Uint8Array(buffer);
  const prefix = Buffer.from(bytes.subarray(0, Math.min(bytes.byteLength, 1 << 20))).toString('utf8');
  const end = prefix.indexOf('end_header');
  if (end < 0) throw new Error('Invalid PLY: missing end_header');
  let headerByteLength = Buffer.byteLength(prefix.slice(0, end), 'utf8') + 'end_header'.length;
  while (headerByteLength < bytes.byteLength &&
         (bytes[headerByteLength] === 10 || bytes[headerByteLength] === 13)) headerByteLength++;
  const lines = Buffer.from(bytes.subarray(0, headerByteLength)).toString('utf8')
    .split(/\r?\n/).map((l) => l.trim()).filter(Boolean);
  if (lines[0] !== 'ply') throw new Error('Invalid PLY: file must start with `ply`');
  let format = null;
  let version = '1.0';
  const elements = [];
  let current = null;
  for (const line of lines.slice(1)) {
    if (line === 'end_header') break;
    if (line.startsWith('comment ') || line.startsWith('obj_info ')) continue;
    const t = line.split(/\s+/);
    if (t[0] === 'format') {
      format = t[1];
      version = t[2] || '1.0';
    } else if (t[0] === 'element') {
      current = { name: t[1], count: Number(t[2]), properties: [] };
      if (!Number.isFinite(current.count) || current.count < 0) {
        throw new Error(`Invalid PLY element count for '${current.name}'`);
      }
      elements.push(current);
    } else if (t[0] === 'property') {
      if (!current) throw new Error('Invalid PLY: property declared before element');
      if (t[1] === 'list') {
        throw new Error('Unsupported PLY: list properties are not expected in Gaussian Splat PLY files');
      }
      current.properties.push({ type: t[1], name: t[2] });
    }
  }
  if (!format) throw new Error('Invalid PLY: missing format line');
  if (!['ascii', 'binary_little_endian', 'binary_big_endian'].includes(format)) {
    throw new Error(`Unsupported PLY format: ${format}`);
  }
  const vertexElement = elements.find((e) => e.name === 'vertex');
  if (!vertexElement) throw new Error('Invalid Gaussian Splat PLY: missing vertex element');
  return { format, version, elements, vertexElement, headerByteLength };
}

function validateGaussianSplatHeader(header) {
  const names = new Set(header.vertexElement.properties.map((p) => p.name));
  for (const n of REQUIRED) {
    if (!names.has(n)) throw new Error(`Invalid Gaussian Splat PLY: missing property '${n}'`);
  }
}

function inferSphericalHarmonicsDegree(fRestCount) {
  if (fRestCount >= 45) return 3;
  if (fRestCount >= 24) return 2;
  if (fRestCount >= 9) return 1;
  return 0;
}

const clamp01 = (x) => Math.min(1, Math.max(0, x));
const sigmoid = (x) => 1 / (1 + Math.exp(-x));

function makeReader(buffer, header) {
  if (header.format === 'ascii') {
    const text = Buffer.from(new Uint8Array(buffer, header.headerByteLength)).toString('utf8');
    const tokens = text.trim().split(/\s+/);
    let i = 0;
    return () => {
      const tok = tokens[i++];
      if (tok === undefined) throw new Error('Unexpected end of ASCII PLY body');
      const v = Number(tok);
      if (!Number.isFinite(v)) throw new Error(`Invalid ASCII PLY value '${tok}'`);
      return v;
    };
  }
  const view = new DataView(buffer);
  const little = header.format === 'binary_little_endian';
  let offset = header.headerByteLength;
  return (type) => {
    const s = SCALAR[type];
    if (!s) throw new Error(`Unsupported PLY scalar type: ${type}`);
    if (offset + s[0] > view.byteLength) throw new RangeError('Offset is outside the bounds of the DataView');
    const v = view[s[1]](offset, little);
    offset += s[0];
    return v;
  };
}

function parseGaussianSplatPly(input) {
  const buffer = toArrayBuffer(input);
  const header = parsePlyHeader(buffer);
  validateGaussianSplatHeader(header);
  const next = makeReader(buffer, header);
  const vprops = header.vertexElement.properties;
  const fRestNames = vprops.map((p) => p.name).filter((n) => /^f_rest_\d+$/.test(n))
    .sort((a, b) => Number(a.slice(7)) - Number(b.slice(7)));
  const n = header.vertexElement.count;
  const fRestCount = fRestNames.length;
  const splats = {
    positions: new Float32Array(n * 4),
    colors: new Float32Array(n * 4),
    scales: new Float32Array(n * 4),
    rotations: new Float32Array(n * 4),
    fDc: new Float32Array(n * 3),
    sphericalHarmonics: fRestCount > 0 ? new Float32Array(n * fRestCount) : null,
    sphericalHarmonicsDegree: inferSphericalHarmonicsDegree(fRestCount),
    splatCount: n,
  };
  const bounds = { min: { x: Infinity, y: Infinity, z: Infinity }, max: { x: -Infinity, y: -Infinity, z: -Infinity } };
  const index = new Map(vprops.map((p, i) => [p.name, i]));
  const req = (row, name) => {
    const v = row[index.get(name)];
    if (!Number.isFinite(v)) throw new Error(`Invalid Gaussian Splat PLY row: missing '${name}'`);
    return v;
  };
  const restIdx = fRestNames.map((nm) => index.get(nm));
  let vi = 0;
  for (const el of header.elements) {
    const isVertex = el.name === 'vertex';
    const row = new Array(el.properties.length);
    for (let r = 0; r < el.count; r++) {
      for (let k = 0; k < el.properties.length; k++) row[k] = next(el.properties[k].type);
      if (!isVertex) continue;
      const o = vi * 4;
      const x = req(row, 'x'), y = req(row, 'y'), z = req(row, 'z');
      splats.positions[o] = x; splats.positions[o + 1] = y; splats.positions[o + 2] = z; splats.positions[o + 3] = 1;
      bounds.min.x = Math.min(bounds.min.x, x); bounds.min.y = Math.min(bounds.min.y, y); bounds.min.z = Math.min(bounds.min.z, z);
      bounds.max.x = Math.max(bounds.max.x, x); bounds.max.y = Math.max(bounds.max.y, y); bounds.max.z = Math.max(bounds.max.z, z);
      for (let c = 0; c < 3; c++) {
        const dc = req(row, `f_dc_${c}`);
        splats.fDc[vi * 3 + c] = dc;
        splats.colors[o + c] = clamp01(0.5 + SH_C0 * dc);
      }
      splats.colors[o + 3] = clamp01(sigmoid(req(row, 'opacity')));
      splats.scales[o] = Math.exp(req(row, 'scale_0'));
      splats.scales[o + 1] = Math.exp(req(row, 'scale_1'));
      splats.scales[o + 2] = Math.exp(req(row, 'scale_2'));
      splats.scales[o + 3] = 0;
      const r0 = req(row, 'rot_0'), r1 = req(row, 'rot_1'), r2 = req(row, 'rot_2'), r3 = req(row, 'rot_3');
      const len = Math.hypot(r0, r1, r2, r3);
      const inv = len > 0 ? 1 / len : 1;
      splats.rotations[o] = r0 * inv; splats.rotations[o + 1] = r1 * inv;
      splats.rotations[o + 2] = r2 * inv; splats.rotations[o + 3] = r3 * inv;
      if (splats.sphericalHarmonics) {
        const so = vi * fRestCount;
        for (let k = 0; k < fRestCount; k++) {
          const v = row[restIdx[k]];
          if (!Number.isFinite(v)) throw new Error(`Invalid Gaussian Splat PLY row: missing '${fRestNames[k]}'`);
          splats.sphericalHarmonics[so + k] = v;
        }
      }
      vi++;
    }
  }
  return { splats, bounds };
}

// SceneLoader: parse off the caller's thread, one worker per load
// (the reference restarts its worker on every load, SceneLoader.ts:103-106).
// With `native`, the addon's multi-threaded C++ parser (parsePly) does the
// work instead and gives the same arrays and error messages.
class SceneLoader {
  constructor({ useWorker = true, native = false, addon = null } = {}) {
    this.useWorker = useWorker;
    this.native = native;
    this.addon = addon;
  }

  loadArrayBuffer(buffer) {
    if (this.native) {
      const addon = this.addon || require('./GaussianSplatRenderer').loadAddon();
      return Promise.resolve().then(() => addon.parsePly(buffer));
    }
    if (!this.useWorker) return Promise.resolve().then(() => parseGaussianSplatPly(buffer));
    const { Worker } = require('worker_threads');
    return new Promise((resolve, reject) => {
      const w = new Worker(path.join(__dirname, 'SceneWorker.js'));
      const ab = toArrayBuffer(buffer);
      w.once('message', (msg) => {
        w.terminate();
        if (msg.type === 'loaded') resolve({ splats: msg.splats, bounds: msg.bounds });
        else reject(new Error(msg.error));
      });
      w.once('error', (e) => { w.terminate(); reject(e); });
      w.postMessage({ type: 'load-arraybuffer', buffer: ab }, [ab]);
    });
  }

  loadFile(file) {
    if (this.native) {
      const addon = this.addon || require('./GaussianSplatRenderer').loadAddon();
      return Promise.resolve().then(() => addon.parsePly(String(file)));
    }
    return fs.promises.readFile(file).then((b) => this.loadArrayBuffer(b));
  }
}

module.exports = {
  parsePlyHeader, validateGaussianSplatHeader, parseGaussianSplatPly, inferSphericalHarmonicsDegree,
  SceneLoader, SH_C0,
};
