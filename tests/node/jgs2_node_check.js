'use strict';
// Drives the Node binding (webgpu-physics_amd/napi/jgs2hip.js) for
// tests/test_napi.py. Usage: node jgs2_node_check.js <mode> <scene dir> [out]
//   surface  print the binding's exported functions and class members
//   errors   print the JS error class and message of each invalid call
//   step     create, step, read, and write positions to <out> (needs a GPU)
// The scene directory holds scene.json (the JGS2GpuInput fields, arrays as
// {"file", "dtype"} references to raw little-endian .bin files) and the
// settings in camelCase.

const fs = require('fs');
const path = require('path');

const repo = path.resolve(__dirname, '..', '..');
const binding = require(path.join(repo, 'webgpu-physics_amd', 'napi', 'jgs2hip.js'));
const { JGS2GpuSolver, native } = binding;

function loadTree(dir, node) {
  if (node && typeof node === 'object' && !Array.isArray(node)) {
    if (node.file !== undefined && node.dtype !== undefined) {
      const buf = fs.readFileSync(path.join(dir, node.file));
      const copy = new ArrayBuffer(buf.length);
      new Uint8Array(copy).set(buf);
      return node.dtype === 'f32' ? new Float32Array(copy) : new Uint32Array(copy);
    }
    const out = {};
    for (const [k, v] of Object.entries(node)) out[k] = loadTree(dir, v);
    return out;
  }
  return node;
}

function load(dir) {
  const manifest = JSON.parse(fs.readFileSync(path.join(dir, 'scene.json'), 'utf8'));
  return {
    input: loadTree(dir, manifest.input),
    settings: manifest.settings,
    iterations: manifest.iterations,
    pinnedVertex: manifest.pinnedVertex,
  };
}

async function outcome(label, fn) {
  try {
    await fn();
    return { label, name: null, message: null };
  } catch (e) {
    return { label, name: e.constructor.name, message: String(e.message) };
  }
}

async function errors(scene) {
  const { input, settings } = scene;
  const results = [];
  results.push(await outcome('empty input', () => JGS2GpuSolver.create({})));
  results.push(await outcome('positions as Uint32Array', () =>
    JGS2GpuSolver.create({ ...input, positions: new Uint32Array(input.positions.length) }, settings)));
  results.push(await outcome('unknown schedule', () => JGS2GpuSolver.create(input, { ...settings, schedule: 'gauss' })));
  results.push(await outcome('zero timestep', () => JGS2GpuSolver.create(input, { ...settings, timestep: 0 })));
  const badTets = Uint32Array.from(input.tetIndices);
  badTets[0] = input.vertexCount + 5;
  results.push(await outcome('tet index out of range', () =>
    JGS2GpuSolver.create({ ...input, tetIndices: badTets }, settings)));
  const forces = new Float32Array(3 * input.vertexCount);
  forces[3 * scene.pinnedVertex] = 1;
  results.push(await outcome('force on a pinned vertex', () =>
    JGS2GpuSolver.create({ ...input, objectives: { externalForces: forces } }, settings)));
  results.push(await outcome('short positions', () =>
    JGS2GpuSolver.create({ ...input, positions: input.positions.subarray(4) }, settings)));
  results.push(await outcome('not a handle', () => native.step({}, {})));
  results.push(await outcome('valid input', async () => {
    const solver = await JGS2GpuSolver.create(input, settings);
    solver.destroy();
  }));
  return results;
}

async function step(scene, out) {
  const { input, settings, iterations } = scene;
  // The three-argument form of the reference: create(device, input, settings).
  const solver = await JGS2GpuSolver.create(null, input, settings);
  solver.stepExactIterations(iterations, settings);
  await solver.awaitIdle();
  const positions = await solver.readPositions();
  fs.writeFileSync(out, Buffer.from(positions.buffer, positions.byteOffset, positions.byteLength));
  const diagnostics = await solver.readGlobalizationDiagnostics();
  const oracle = await solver.readOracleDiagnostics(settings);
  const summary = {
    vertexCount: solver.vertexCount,
    tetCount: solver.tetCount,
    historyCount: diagnostics.historyCount,
    localRecords: diagnostics.local.length,
    acceptedLocal: diagnostics.local.filter((r) => r.accepted).length,
    lastHistory: diagnostics.history[diagnostics.historyCount - 1],
    oracleEnergy: oracle.energy,
    oracleVertices: oracle.vertices.length,
    oracleFinite: oracle.finite,
  };
  solver.destroy();
  solver.destroy(); // idempotent (jgs2-solver.ts:3029-3032)
  summary.afterDestroy = [
    await outcome('step after destroy', () => solver.step(settings)),
    await outcome('read after destroy', () => solver.readPositions()),
  ];
  return summary;
}

async function main() {
  const [mode, dir, out] = process.argv.slice(2);
  let result;
  if (mode === 'surface') {
    result = {
      native: Object.keys(native).sort(),
      members: Object.getOwnPropertyNames(JGS2GpuSolver.prototype).sort(),
      statics: Object.getOwnPropertyNames(JGS2GpuSolver).filter((k) => typeof JGS2GpuSolver[k] === 'function'),
      bound: native.boundEntryPoints(),
      deviceCount: native.deviceCount(),
    };
  } else {
    const scene = load(dir);
    result = mode === 'errors' ? await errors(scene) : await step(scene, out);
  }
  process.stdout.write(JSON.stringify(result) + '\n');
}

main().catch((e) => {
  process.stderr.write(`${e.constructor.name}: ${e.message}\n`);
  process.exit(1);
});
