'use strict';
// JGS2GpuSolver for Node over libjgs2hip (N-API addon ../lib/jgs2hip.node).
// Mirrors the reference class's public surface
// (src/simulation/gpu/jgs2-solver.ts:905-3051): async create, the four step
// entry points, the seven objective mutators, the async reads and the
// globalization / oracle diagnostics, awaitIdle and an idempotent destroy.
// A Node host has no GPUDevice: create(input, settings) takes the scene
// directly; create(device, input, settings) is accepted and ignores device.
//
// Errors follow the reference: RangeError for invalid sizes, values and
// infeasible poses, TypeError for malformed arguments, Error otherwise and
// for use after destroy ("JGS2GpuSolver has been destroyed.").

const path = require('path');

const native = require(path.join(__dirname, '..', 'lib', 'jgs2hip.node'));

const HISTORY_CAPACITY = 64; // jgs2-globalization.ts:2
const MAX_BACKTRACKS = 12; // jgs2-globalization.ts:23
const DETERMINANT_FLOOR = 1e-4; // jgs2-globalization.ts:20
const LOCAL_STATUS = ['accepted', 'pinned', 'zeroGradient', 'shiftLimit', 'nonDescent', 'lineSearchFailed',
  'nonfinite', 'choleskyFailed']; // jgs2-globalization.ts:39-81

function finite(v) {
  return Number.isFinite(v) ? [v, true] : [0, false];
}

function nonnegative(v) {
  return Number.isFinite(v) && v >= 0 ? [v, true] : [0, false];
}

function flag(v, label) {
  if (v !== 0 && v !== 1) throw new RangeError(`${label} must be encoded as exactly 0 or 1.`);
  return v === 1;
}

function count(v, label, maximum = Number.MAX_SAFE_INTEGER) {
  if (!Number.isSafeInteger(v) || v < 0 || v > maximum) {
    throw new RangeError(`${label} must be a nonnegative integer no greater than ${maximum}.`);
  }
  return v;
}

// One 16-float local record (decodeJGS2GlobalizationLocalDiagnostic,
// jgs2-globalization.ts:237-340).
function decodeLocal(p) {
  const dir = [finite(p[0]), finite(p[1]), finite(p[2])];
  const alpha = nonnegative(p[3]);
  const pd = [finite(p[4]), nonnegative(p[5]), nonnegative(p[6]), nonnegative(p[7])];
  const en = [finite(p[8]), finite(p[9]), finite(p[10]), nonnegative(p[11])];
  const minJ = finite(p[12]);
  const statusCode = p[15];
  if (!Number.isSafeInteger(statusCode) || statusCode < 0 || statusCode >= LOCAL_STATUS.length) {
    throw new RangeError(`Local globalization status must be an integer from 0 through ${LOCAL_STATUS.length - 1}.`);
  }
  const nonfiniteStatus = statusCode === 6;
  const geometryEvaluated = statusCode === 0 || statusCode === 5 || (nonfiniteStatus && minJ[0] > DETERMINANT_FLOOR);
  const directionValid = dir.every((e) => e[1]);
  const positiveDefiniteDiagnosticsValid = pd.every((e) => e[1]);
  const energyDiagnosticsValid = en.every((e) => e[1]);
  return {
    direction: [dir[0][0], dir[1][0], dir[2][0]],
    alpha: alpha[0],
    minimumEigenvalue: pd[0][0],
    scale: pd[1][0],
    diagonalShift: pd[2][0],
    normalizedShift: pd[3][0],
    acceptedEnergyDelta: en[0][0],
    armijoDeltaBound: en[1][0],
    gradientDotDirection: en[2][0],
    shiftedLinearResidual: en[3][0],
    minimumTrialDeformationDeterminant: minJ[0],
    backtrackCount: count(p[13], 'Local globalization backtrackCount', MAX_BACKTRACKS),
    energyEvaluationCount: count(p[14], 'Local globalization energyEvaluationCount'),
    statusCode,
    status: LOCAL_STATUS[statusCode],
    accepted: statusCode === 0,
    directionValid,
    alphaValid: alpha[1],
    positiveDefiniteDiagnosticsValid,
    energyDiagnosticsValid,
    minimumTrialDeformationDeterminantValid: minJ[1] && geometryEvaluated,
    finite: directionValid && alpha[1] && positiveDefiniteDiagnosticsValid && energyDiagnosticsValid && minJ[1] &&
      !nonfiniteStatus,
  };
}

// One 32-float history record (decodeJGS2GlobalizationHistoryRecord,
// jgs2-globalization.ts:342-509).
function decodeHistory(p) {
  const minJ = [finite(p[0]), finite(p[1]), finite(p[2])];
  const bits = count(p[3], 'geometryValidityBits', 15);
  const energies = [finite(p[8]), finite(p[9]), finite(p[10])];
  const norms = [4, 5, 6, 7, 8].map((i) => nonnegative(p[8 + i]));
  const residual = [nonnegative(p[17]), nonnegative(p[18]), nonnegative(p[19])];
  const update = [nonnegative(p[20]), nonnegative(p[21])];
  const shift = nonnegative(p[27]);
  const sourceGeometryValid = (bits & 1) !== 0 && minJ[0][1];
  const candidateGeometryValid = (bits & 2) !== 0 && minJ[1][1];
  const acceptedValid = (bits & 4) !== 0 && minJ[2][1];
  const localNumericsValid = (bits & 8) !== 0;
  const componentGradientNormsValid = norms.every((e) => e[1]);
  const residualDiagnosticsValid = residual.every((e) => e[1]);
  const updateDiagnosticsValid = update.every((e) => e[1]);
  const allNumericsValid = minJ.every((e) => e[1]) && energies.every((e) => e[1]) && componentGradientNormsValid &&
    residualDiagnosticsValid && updateDiagnosticsValid && shift[1];
  return {
    sourceMinimumDeformationDeterminant: minJ[0][0],
    candidateMinimumDeformationDeterminant: minJ[1][0],
    acceptedMinimumDeformationDeterminant: minJ[2][0],
    sourceGeometryValid,
    candidateGeometryValid,
    acceptedMinimumDeformationDeterminantValid: acceptedValid,
    localNumericsValid,
    geometryValid: sourceGeometryValid && candidateGeometryValid,
    assembledAccepted: flag(p[4], 'assembledAccepted'),
    assembledReverted: flag(p[5], 'assembledReverted'),
    localFailureCount: count(p[6], 'localFailureCount'),
    revertCount: count(p[7], 'revertCount', 1),
    sourceEnergy: energies[0][0],
    candidateEnergy: energies[1][0],
    acceptedEnergy: energies[2][0],
    energyValid: flag(p[11], 'energyValid') && energies.every((e) => e[1]),
    componentGradientNorms: Object.freeze({
      inertia: norms[0][0],
      material: norms[1][0],
      externalForce: norms[2][0],
      target: norms[3][0],
      contact: norms[4][0],
    }),
    componentGradientNormsValid,
    gradientNorm: residual[0][0],
    residualDenominator: residual[1][0],
    relativeResidual: residual[2][0],
    residualDiagnosticsValid,
    maximumUpdate: update[0][0],
    normalizedMaximumUpdate: update[1][0],
    updateDiagnosticsValid,
    residualSatisfied: flag(p[22], 'residualSatisfied'),
    updateSatisfied: flag(p[23], 'updateSatisfied'),
    converged: flag(p[24], 'converged'),
    finite: flag(p[25], 'finite') && allNumericsValid && sourceGeometryValid && candidateGeometryValid &&
      acceptedValid && localNumericsValid,
    iterationIndex: count(p[26], 'iterationIndex', HISTORY_CAPACITY - 1),
    maximumLocalNormalizedShift: shift[0],
    maximumLocalNormalizedShiftValid: shift[1],
    maximumBacktrackCount: count(p[28], 'maximumBacktrackCount', MAX_BACKTRACKS),
    infeasibleTrialCount: count(p[29], 'infeasibleTrialCount'),
    energyEvaluationCount: count(p[30], 'energyEvaluationCount'),
    active: flag(p[31], 'active'),
  };
}

// Per-vertex oracle record of 36 doubles (layout in include/jgs2hip.h) as
// the reference's JGS2GpuOracleVertexRecord (jgs2-diagnostics.ts:30-51).
function decodeOracleVertex(r, vertex) {
  return {
    vertex,
    active: r[3] !== 0,
    gradient: [r[0], r[1], r[2]],
    localHessian: Float32Array.from(r.subarray(4, 13)),
    inertiaEnergy: r[13],
    externalForceEnergy: r[28],
    targetEnergy: r[32],
    externalForceGradient: [r[25], r[26], r[27]],
    targetGradient: [r[29], r[30], r[31]],
    floorContactEnergy: r[14],
    updateMagnitude: r[15],
    updateMagnitudeValid: r[19] !== 0,
    gradientSquaredNorm: r[16],
    inertiaWeight: r[18],
    finite: r[17] !== 0,
  };
}

class JGS2GpuSolver {
  constructor(handle, info, input) {
    this._handle = handle;
    this._destroyed = false;
    this.vertexCount = info.vertexCount;
    this.tetCount = info.tetCount;
    this.cubatureK = info.cubatureK;
    this.bodyCount = info.bodyCount;
    this.scheduleColorCount = info.scheduleColorCount;
    this.globalizationEnabled = info.globalizationEnabled;
    this.sceneScale = info.sceneScale;
    const contact = input.contactCandidates;
    this.contactCandidateCount = contact ? contact.vertexTriangleCount + contact.edgeEdgeCount : 0;
    this.clothTriangleCount = input.cloth ? input.cloth.triangleIndices.length / 3 : 0;
    this._recordsAvailable = false;
  }

  // JGS2GpuSolver.create (jgs2-solver.ts:971-1502).
  static async create(...args) {
    if (args.length === 3 || (args.length === 2 && args[1] && args[1].vertexCount !== undefined)) args.shift();
    const [input, settings = {}] = args;
    const handle = native.create(input, settings);
    return new JGS2GpuSolver(handle, native.info(handle), input);
  }

  _assertUsable() {
    if (this._destroyed) throw new Error('JGS2GpuSolver has been destroyed.');
  }

  step(overrides = {}) {
    native.step(this._handle, overrides);
    this._recordsAvailable = true;
  }

  stepFrames(frameCount, overrides = {}) {
    native.stepFrames(this._handle, frameCount, overrides);
    this._recordsAvailable = true;
  }

  stepExactIterations(iterations, overrides = {}) {
    native.stepExactIterations(this._handle, iterations, overrides);
    this._recordsAvailable = true;
  }

  stepFramesExactIterations(frameCount, iterations, overrides = {}) {
    native.stepFramesExactIterations(this._handle, frameCount, iterations, overrides);
    this._recordsAvailable = true;
  }

  setExternalForce(vertex, force) {
    native.setExternalForce(this._handle, vertex, force);
  }

  replaceExternalForces(forces) {
    native.replaceExternalForces(this._handle, forces);
  }

  clearExternalForces() {
    native.clearExternalForces(this._handle);
  }

  setQuadraticTarget(vertex, target, stiffness) {
    native.setQuadraticTarget(this._handle, vertex, target, stiffness);
  }

  replaceQuadraticTargets(positions, stiffnesses) {
    native.replaceQuadraticTargets(this._handle, positions, stiffnesses);
  }

  releaseQuadraticTarget(vertex) {
    native.releaseQuadraticTarget(this._handle, vertex);
  }

  releaseAllQuadraticTargets() {
    native.releaseAllQuadraticTargets(this._handle);
  }

  async awaitIdle() {
    native.awaitIdle(this._handle);
  }

  async readPositions() {
    return native.readPositions(this._handle);
  }

  async readVelocities() {
    return native.readVelocities(this._handle);
  }

  async readPredictedPositions() {
    return native.readPredictedPositions(this._handle);
  }

  async readFinalIterationUpdates() {
    return native.readFinalIterationUpdates(this._handle);
  }

  async readVertexRotations() {
    return native.readVertexRotations(this._handle);
  }

  // readGlobalizationDiagnostics (jgs2-solver.ts:2548-2604).
  async readGlobalizationDiagnostics() {
    this._assertUsable();
    if (!this.globalizationEnabled) {
      throw new Error('Globalization diagnostics are available only for stable Neo-Hookean production solves.');
    }
    if (!this._recordsAvailable) {
      throw new Error('Globalization diagnostics are unavailable until a stable nonlinear iteration or ' +
        'diagnostic test kernel has produced records.');
    }
    const packedLocal = native.readLocalDiagnostics(this._handle);
    const control = native.readGlobalizationControl(this._handle);
    const packedHistory = native.readGlobalizationHistory(this._handle);
    const historyCount = control[7];
    if (!Number.isSafeInteger(historyCount) || historyCount < 0 || historyCount > HISTORY_CAPACITY) {
      throw new Error(`GPU returned invalid globalization history count ${historyCount}.`);
    }
    const local = [];
    for (let v = 0; v < this.vertexCount; v += 1) local.push(decodeLocal(packedLocal.subarray(16 * v, 16 * v + 16)));
    const history = [];
    for (let i = 0; i < historyCount; i += 1) history.push(decodeHistory(packedHistory.subarray(32 * i, 32 * i + 32)));
    return { local, history, historyCount };
  }

  // readOracleDiagnostics (jgs2-solver.ts:2882-2907; jgs2-diagnostics.ts:61-111).
  async readOracleDiagnostics(overrides = {}) {
    const d = native.readOracleDiagnostics(this._handle, overrides, null, null, true);
    const vertices = [];
    for (let v = 0; v < this.vertexCount; v += 1) {
      vertices.push(decodeOracleVertex(d.vertexRecords.subarray(36 * v, 36 * v + 36), v));
    }
    const tetrahedra = [];
    for (let t = 0; t < this.tetCount; t += 1) {
      const r = d.tetrahedronRecords.subarray(4 * t, 4 * t + 4);
      tetrahedra.push({ tetrahedron: t, elasticityEnergy: r[0], deformationDeterminant: r[1], finite: r[2] !== 0 });
    }
    const ok = d.finite !== 0;
    return {
      energy: d.inertia + d.elasticity + d.externalForce + d.quadraticTarget + d.floorContact,
      components: {
        inertia: d.inertia,
        elasticity: d.elasticity,
        externalForce: d.externalForce,
        quadraticTarget: d.quadraticTarget,
        floorContact: d.floorContact,
      },
      componentGradientNorms: {
        inertia: d.gradientNormInertia,
        material: d.gradientNormMaterial,
        externalForce: d.gradientNormExternalForce,
        target: d.gradientNormTarget,
        contact: d.gradientNormFloor,
      },
      gradientNorm: d.gradientNorm,
      residualNumerator: d.gradientNorm,
      relativeResidual: d.relativeResidual,
      relativeResidualValid: ok,
      residualDenominator: d.residualDenominator,
      gradientSquaredNorm: d.gradientNorm * d.gradientNorm,
      gradientValid: ok,
      maximumUpdate: d.maximumUpdate,
      maximumUpdateValid: d.maximumUpdateValid !== 0,
      minimumDeformationDeterminant: d.minimumDeformationDeterminant,
      minimumDeformationDeterminantValid: d.minimumDeterminantValid !== 0,
      minimumContactDistance: d.minimumFloorDistance,
      minimumContactDistanceValid: d.floorMetricsValid !== 0,
      activeContactCount: d.activeFloorContacts,
      activeContactCountValid: d.floorMetricsValid !== 0,
      candidateBufferOverflow: false,
      candidateBufferOverflowValid: false,
      activeVertexCount: d.activeVertices,
      finite: ok,
      vertices,
      tetrahedra,
      // Beyond the reference: the IPC barrier term and its closest pair.
      contactBarrierEnergy: d.contactBarrier,
      minimumCandidateDistance: d.minimumContactDistance,
      activeCandidatePairs: d.activeContactPairs,
    };
  }

  // destroy (jgs2-solver.ts:3029-3044): a second call is a no-op.
  destroy() {
    if (this._destroyed) return;
    this._destroyed = true;
    native.destroy(this._handle);
  }
}

module.exports = {
  JGS2GpuSolver,
  native,
  deviceCount: native.deviceCount,
  decodeLocal,
  decodeHistory,
};
