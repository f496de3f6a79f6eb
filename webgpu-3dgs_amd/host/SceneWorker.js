'use strict';
// Worker side of SceneLoader: parse a PLY ArrayBuffer and transfer the arrays
// back (message protocol of src/workers/SceneWorker.ts:57-133).
const { parentPort } = require('worker_threads');
const { parseGaussianSplatPly } = require('./SceneLoader');

parentPort.on('message', (msg) => {
  try {
    if (msg.type !== 'load-arraybuffer') throw new Error(`Unknown worker message type: ${msg.type}`);
    const { splats, bounds } = parseGaussianSplatPly(msg.buffer);
    const transfer = [splats.positions.buffer, splats.colors.buffer, splats.scales.buffer,
      splats.rotations.buffer, splats.fDc.buffer];
    if (splats.sphericalHarmonics) transfer.push(splats.sphericalHarmonics.buffer);
    parentPort.postMessage({ type: 'loaded', splats, bounds }, transfer);
  } catch (err) {
    parentPort.postMessage({ type: 'error', error: err instanceof Error ? err.message : String(err) });
  }
});
