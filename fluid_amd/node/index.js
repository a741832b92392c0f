'use strict';
// Host-side drop-in of the MI355X pressure projection for the reference's
// TypeScript host (CommonJS so it runs without a TypeScript toolchain; types in
// index.d.ts).
//
//  * UniformPressureProjection mirrors WebGPUUniformPressureMultigrid
//    (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:228-607):
//    the constructor plans the hierarchy and throws RangeError on an
//    unplannable lattice (:251-256); project() runs divergence -> MGPCG ->
//    gradient subtract; diagnostics holds the decoded report; destroy() is
//    idempotent.  submit() queues the same projection on the context's
//    continuous-batching worker; projectBatch() solves several grids at once.
//  * TallCellPressureProjection is the same over a restricted tall-cell
//    layout (lib/core/tall-cell-grid.ts:309-436).
//  * mi355xProjectionMethod is a SimulationMethod
//    (lib/core/method-contract.ts:363-602) whose createSolverAsync resolves a
//    GPUSolverInstance (:209-361) built from a SceneDescription
//    (lib/core/model.ts:52-130): advanceTo() steps the scene's fixed step,
//    each step a gravity impulse and one projection; readStats() resolves a
//    GPUEulerianInfo (lib/core/webgpu-eulerian.ts:13-38).
const path = require('path');

const binding = require(path.join(__dirname, '..', '_lib', 'fluid_projection.node'));

const PRECISION = { fp32: 0, fp16: 1, bf16: 2 };
const SOLVER = { mgpcg: 0, cm11a: 1 };
const ADVECT_SCHEME = { 'semi-lagrangian': 0, maccormack: 1 };
const FLAGS = {
  NONFINITE: 0x1, NONPOSITIVE_CURVATURE: 0x2, NONPOSITIVE_PRECONDITIONER: 0x4, NONCONVERGENCE: 0x8,
  UNANCHORED_COMPONENT: 0x10, PRECISION_FALLBACK: 0x20, COARSE_SINGULAR: 0x40,
};

let sharedContext = null;
function defaultContext() {
  if (!sharedContext) sharedContext = binding.createContext(0, 0);
  return sharedContext;
}

function nativeOptions(options) {
  const opts = Object.assign({}, options);
  if (typeof opts.precision === 'string') {
    if (!(opts.precision in PRECISION)) throw new RangeError(`unknown precision ${opts.precision}`);
    opts.precision = PRECISION[opts.precision];
  }
  if (typeof opts.solver === 'string') {
    if (!(opts.solver in SOLVER)) throw new RangeError(`unknown solver ${opts.solver}`);
    opts.solver = SOLVER[opts.solver];
  }
  return opts;
}

// A grid handle that is freed only once no projection of it is in flight:
// destroy() while a project()/submit() promise is pending defers the native
// free until the promise settles.
class GridHandle {
  constructor(context, desc) {
    this.native = binding.createGrid(context, desc);
    this.pending = 0;
    this.destroyed = false;
  }

  track(promise) {
    this.pending += 1;
    const done = () => {
      this.pending -= 1;
      if (this.destroyed && this.pending === 0) this.free();
    };
    return promise.then((v) => { done(); return v; }, (e) => { done(); throw e; });
  }

  use() {
    if (this.destroyed) throw new Error('the projection grid was destroyed');
    return this.native;
  }

  free() {
    if (this.native) {
      binding.destroyGrid(this.native);
      this.native = null;
    }
  }

  destroy() {
    this.destroyed = true;
    if (this.pending === 0) this.free();
  }
}

class UniformPressureProjection {
  constructor(dims, spacing, context) {
    const [nx, ny, nz] = dims;
    const h = Array.isArray(spacing) ? spacing[0] : spacing;
    if (Array.isArray(spacing) && spacing.some((s) => Math.abs(s - h) > 1e-12 * h)) {
      throw new RangeError('MI355X projection needs cubic cells');
    }
    this.context = context || defaultContext();
    this.dims = [nx, ny, nz];
    this.h = h;
    this.handle = new GridHandle(this.context, { layout: 0, nx, ny, nz, h });
    this.diagnostics = undefined;
  }

  get grid() { return this.handle.native; }
  get levelCount() { return binding.gridLevels(this.handle.use()).length; }
  get levelDimensions() { return binding.gridLevels(this.handle.use()); }

  outputs() {
    const [nx, ny, nz] = this.dims;
    return {
      pressure: new Float32Array(nx * ny * nz),
      u: new Float32Array((nx + 1) * ny * nz),
      v: new Float32Array(nx * (ny + 1) * nz),
      w: new Float32Array(nx * ny * (nz + 1)),
    };
  }

  /** One projection; resolves {pressure, u, v, w, report}. */
  project(inputs, options = {}) {
    const { phi, u, v, w, targetDivergence, dt, rho } = inputs;
    const run = binding.project(this.handle.use(), { phi, u, v, w, targetDivergence, dt, rho }, this.outputs(),
      nativeOptions(options));
    return this.handle.track(run).then((res) => { this.diagnostics = res.report; return res; });
  }

  /** The same projection through the continuous-batching worker (fp_submit /
   * fp_poll): resolves when it completes, with its latency in report.latencyMs. */
  submit(inputs, options = {}) {
    const { phi, u, v, w, targetDivergence, dt, rho } = inputs;
    const run = binding.submit(this.handle.use(), { phi, u, v, w, targetDivergence, dt, rho }, this.outputs(),
      nativeOptions(options));
    return this.handle.track(run).then((res) => { this.diagnostics = res.report; return res; });
  }

  /** Velocity extrapolation into the air (fp_extrapolate_velocity; the
   * reference's WebGPUUniformVelocityExtrapolator stage,
   * lib/methods/uniform/webgpu-uniform-velocity-extrapolation.ts): u, v, w
   * are updated in place from the faces next to liquid (phi < 0, or
   * density > 1/2).  Synchronous; returns the report
   * {passes, activeLeft, levels, bandFaces, unknownFaces, deviceMs}. */
  extrapolate({ phi, density, u, v, w }, { openTop = false, maxPasses = 0 } = {}) {
    if (this.handle.pending > 0) throw new Error('extrapolate: a projection of this grid is in flight');
    const fields = phi ? { phi, u, v, w } : { density, u, v, w };
    return binding.extrapolateVelocity(this.handle.use(), fields, undefined, { openTop: openTop ? 1 : 0, maxPasses });
  }

  /** Velocity advection and body forces (fp_advect_velocity; the reference's
   * semiLagrangianAdvection / bounded MacCormack passes,
   * lib/methods/uniform/webgpu-uniform-reference.wgsl.ts:872-1004): u, v, w
   * are updated in place, traced through `transport` when given, else through
   * their own extrapolation.  Synchronous; returns {deviceMs}. */
  advect({ phi, density, u, v, w, transport }, { dt, scheme = 'semi-lagrangian', gravity = -9.80665,
    viscosity = 0, openTop = false, noSlip = false } = {}) {
    if (this.handle.pending > 0) throw new Error('advect: a projection of this grid is in flight');
    if (!(scheme in ADVECT_SCHEME)) throw new RangeError(`advect: unknown scheme ${scheme}`);
    const fields = phi ? { phi, u, v, w } : { density, u, v, w };
    if (transport) fields.transport = transport;
    return binding.advectVelocity(this.handle.use(), fields, null, {
      dt, scheme: ADVECT_SCHEME[scheme], gravity, viscosity, openTop: openTop ? 1 : 0, noSlip: noSlip ? 1 : 0 });
  }

  destroy() { this.handle.destroy(); }
}

/**
 * Restricted tall-cell grid: nx x nz columns of fineNy rows of cubic cells h,
 * each column a tall cell of columnBases[x * nz + z] rows under `band` cubic
 * cells (the reference's createTallCellLayout output, tall-cell-grid.ts:309-436,
 * with x and z named the other way round: this library's arrays are z-fastest).
 * project() takes phi on the samples (nx, band+2, nz) and the face slots
 * u (nx+1, S, nz), v (nx, band+2, nz), w (nx, S, nz+1), S = band+2+maxDelta.
 */
class TallCellPressureProjection {
  constructor({ nx, fineNy, nz, band, h, columnBases, maxDelta = 4 }, context) {
    this.context = context || defaultContext();
    this.nx = nx; this.fineNy = fineNy; this.nz = nz; this.band = band; this.maxDelta = maxDelta;
    this.h = h;
    this.columnBases = Int32Array.from(columnBases);
    this.handle = new GridHandle(this.context, {
      layout: 1, nx, ny: fineNy, nz, band, h, maxDelta, columnBases: this.columnBases,
    });
    this.diagnostics = undefined;
  }

  get grid() { return this.handle.native; }
  get levelCount() { return binding.gridLevels(this.handle.use()).length; }

  outputs() {
    const P = this.band + 2, S = this.band + 2 + this.maxDelta;
    const { nx, nz } = this;
    return {
      pressure: new Float32Array(nx * P * nz),
      u: new Float32Array((nx + 1) * S * nz),
      v: new Float32Array(nx * P * nz),
      w: new Float32Array(nx * S * (nz + 1)),
    };
  }

  project({ phi, u, v, w, dt, rho }, options = {}) {
    const run = binding.project(this.handle.use(), { phi, u, v, w, dt, rho }, this.outputs(), nativeOptions(options));
    return this.handle.track(run).then((res) => { this.diagnostics = res.report; return res; });
  }

  submit({ phi, u, v, w, dt, rho }, options = {}) {
    const run = binding.submit(this.handle.use(), { phi, u, v, w, dt, rho }, this.outputs(), nativeOptions(options));
    return this.handle.track(run).then((res) => { this.diagnostics = res.report; return res; });
  }

  /** Re-lay the grid out around the surface of phi and move phi, u, v, w to
   * the new layout in place (fp_tall_remesh: chooseTallCellBase and the
   * neighbour limit of tall-cell-grid.ts:61-106 on the device, the quantity
   * transfer of the tallCells paper's remeshing). */
  remesh({ phi, u, v, w }, { liquidHalo = 0, airHalo = 0 } = {}) {
    if (this.handle.pending > 0) throw new Error('remesh: a projection of this grid is in flight');
    const res = binding.remeshTallGrid(this.handle.use(), { phi, u, v, w }, undefined, { liquidHalo, airHalo });
    this.columnBases = res.columnBases;
    return res;
  }

  destroy() { this.handle.destroy(); }
}

/** Independent projections of several grids (one context) in one batched
 * solve (fp_project_batch); resolves one result per grid. */
function projectBatch(projections, inputs, options = {}) {
  const handles = projections.map((p) => p.handle);
  const run = binding.projectBatch(handles.map((h) => h.use()), inputs, projections.map((p) => p.outputs()),
    nativeOptions(options));
  let tracked = run;
  for (const h of handles) tracked = h.track(tracked);
  return tracked.then((results) => {
    results.forEach((r, i) => { projections[i].diagnostics = r.report; });
    return results;
  });
}

// ---- scene -> grid (SceneDescription, lib/core/model.ts:52-130) -------------
// tallCellSettings (lib/core/tall-cell-grid.ts:52-59)
const TALL_SETTINGS = {
  balanced: { regularLayers: 12, liquidHalo: 8, airHalo: 8, maximumNeighborDelta: 4 },
  high: { regularLayers: 32, liquidHalo: 16, airHalo: 8, maximumNeighborDelta: 4 },
  ultra: { regularLayers: 40, liquidHalo: 24, airHalo: 8, maximumNeighborDelta: 5 },
};

/** The scene's lattice: voxelDomain.finestCellSize_m over the container
 * (the sizing of createTallCellLayout, tall-cell-grid.ts:311-314). */
function sceneLattice(scene) {
  const c = scene.container;
  const size = scene.voxelDomain.finestCellSize_m;
  const nx = Math.max(8, Math.round(c.width_m / size));
  const ny = Math.max(8, Math.round(c.height_m / size));
  const nz = Math.max(8, Math.round(c.depth_m / size));
  const hx = c.width_m / nx, hy = c.height_m / ny, hz = c.depth_m / nz;
  if (Math.abs(hy - hx) > 1e-6 * hx || Math.abs(hz - hx) > 1e-6 * hx) {
    throw new RangeError(`MI355X projection needs cubic cells (container ${c.width_m} x ${c.height_m} x ${c.depth_m} m `
      + `at ${size} m gives ${hx}, ${hy}, ${hz} m)`);
  }
  return { nx, ny, nz, h: hx };
}

/** Normalised reservoir box of a dam-break scene (sceneDamBreakBox,
 * lib/core/initial-fluid.ts:16-63: the fill-derived square footprint of
 * height max(0.92, fill) unless authored dimensions / origin are given). */
function damBreakBox(scene) {
  const c = scene.container, f = scene.fluid;
  let size;
  if (f.initialDamBreakDimensions_m) {
    const d = f.initialDamBreakDimensions_m;
    size = { x: d.x / c.width_m, y: d.y / c.height_m, z: d.z / c.depth_m };
  } else {
    const fill = Math.max(0, Math.min(1, c.fillFraction));
    const height = fill === 0 ? 0 : Math.max(0.92, fill);
    const foot = fill === 0 ? 0 : Math.sqrt(fill / height);
    size = { x: foot, y: height, z: foot };
  }
  const o = f.initialDamBreakOrigin_m;
  const min = o ? { x: o.x / c.width_m, y: o.y / c.height_m, z: o.z / c.depth_m } : { x: 0, y: 0, z: 0 };
  return { min, max: { x: min.x + size.x, y: min.y + size.y, z: min.z + size.z } };
}

/** Signed distance (m) of a container-relative point to the reservoir box. */
function boxDistance(scene, box, px, py, pz) {
  const c = scene.container;
  const lo = [box.min.x * c.width_m, box.min.y * c.height_m, box.min.z * c.depth_m];
  const hi = [box.max.x * c.width_m, box.max.y * c.height_m, box.max.z * c.depth_m];
  const p = [px, py, pz];
  const q = p.map((x, a) => Math.abs(x - (lo[a] + hi[a]) / 2) - (hi[a] - lo[a]) / 2);
  const outside = Math.hypot(Math.max(q[0], 0), Math.max(q[1], 0), Math.max(q[2], 0));
  return outside + Math.min(Math.max(q[0], q[1], q[2]), 0);
}

function sceneCheck(scene) {
  if (!scene || !scene.container || !scene.voxelDomain || !scene.fluid || !scene.numerics) {
    throw new TypeError('createSolverAsync needs a SceneDescription (container, voxelDomain, fluid, numerics)');
  }
  if (scene.container.shape === 'sphere') {
    throw new RangeError('the MI355X projection supports box containers only');
  }
  const ic = scene.fluid.initialCondition;
  if (ic !== 'dam-break' && ic !== 'tank-fill') throw new RangeError(`unsupported initial condition ${ic}`);
}

/** Uniform grid level set: the reservoir box (dam-break) or the fill level
 * (tank-fill), cell centres at (i + 1/2) h. */
function uniformLevelSet(scene, lat) {
  const { nx, ny, nz, h } = lat;
  const phi = new Float32Array(nx * ny * nz);
  const box = scene.fluid.initialCondition === 'dam-break' ? damBreakBox(scene) : null;
  const level = scene.container.fillFraction * scene.container.height_m;
  for (let i = 0; i < nx; i += 1) for (let j = 0; j < ny; j += 1) for (let k = 0; k < nz; k += 1) {
    const x = (i + 0.5) * h, y = (j + 0.5) * h, z = (k + 0.5) * h;
    phi[(i * ny + j) * nz + k] = box ? boxDistance(scene, box, x, y, z) : y - level;
  }
  return phi;
}

// chooseTallCellBase / limitNeighboringTallCellBases
// (lib/core/tall-cell-grid.ts:61-106), restated for the column bases.
function chooseTallCellBase(lowest, highest, fineNy, s) {
  const maximumBase = Math.max(0, fineNy - s.regularLayers);
  const lower = highest + 1 + s.airHalo - s.regularLayers;
  const upper = lowest + 1 - s.liquidHalo;
  const desired = lower <= upper ? Math.round((lower + upper) / 2) : lower;
  const clamped = Math.max(0, Math.min(maximumBase, desired));
  return maximumBase >= 2 ? Math.max(2, clamped) : 0;
}

function limitBases(bases, nx, nz, delta) {
  let cur = Int32Array.from(bases);
  for (let pass = 0; pass < nx + nz; pass += 1) {
    const next = cur.slice();
    let changed = false;
    for (let x = 0; x < nx; x += 1) for (let z = 0; z < nz; z += 1) {
      const i = x * nz + z;
      let up = cur[i];
      if (x > 0) up = Math.min(up, cur[i - nz] + delta);
      if (x + 1 < nx) up = Math.min(up, cur[i + nz] + delta);
      if (z > 0) up = Math.min(up, cur[i - 1] + delta);
      if (z + 1 < nz) up = Math.min(up, cur[i + 1] + delta);
      if (up !== cur[i]) changed = true;
      next[i] = up;
    }
    cur = next;
    if (!changed) break;
  }
  return cur;
}

/** Tall-cell layout of a scene: bases from each column's liquid crossing,
 * samples (nx, B+2, nz) with the tall bottom at fine row 0, the tall top at
 * row base-1 and band row k at base+k (buildInitialPhi, tall-cell-grid.ts:171-190),
 * level set clamped to +-5h and air above the band. */
function tallLayout(scene, lat, quality) {
  const s = TALL_SETTINGS[quality] || TALL_SETTINGS.high;
  const { nx, ny: fineNy, nz, h } = lat;
  const band = Math.min(fineNy, s.regularLayers);
  if (fineNy - band < 2) throw new RangeError(`a ${fineNy}-row lattice leaves no tall cell under a ${band}-row band`);
  const box = scene.fluid.initialCondition === 'dam-break' ? damBreakBox(scene) : null;
  const level = scene.container.fillFraction * scene.container.height_m;
  const dist = (x, y, z) => (box ? boxDistance(scene, box, x, y, z) : y - level);
  const raw = new Int32Array(nx * nz);
  for (let x = 0; x < nx; x += 1) for (let z = 0; z < nz; z += 1) {
    let lowest = -1, highest = -1;
    for (let j = 0; j + 1 < fineNy; j += 1) {
      const a = dist((x + 0.5) * h, (j + 0.5) * h, (z + 0.5) * h) < 0;
      const b = dist((x + 0.5) * h, (j + 1.5) * h, (z + 0.5) * h) < 0;
      if (a !== b) { if (lowest < 0) lowest = j; highest = j; }
    }
    raw[x * nz + z] = lowest < 0 ? Math.max(2, fineNy - band) : chooseTallCellBase(lowest, highest, fineNy, {
      regularLayers: band, liquidHalo: Math.min(s.liquidHalo, band), airHalo: Math.min(s.airHalo, band) });
  }
  const bases = limitBases(raw, nx, nz, s.maximumNeighborDelta);
  const P = band + 2;
  const phi = new Float32Array(nx * P * nz);
  for (let x = 0; x < nx; x += 1) for (let z = 0; z < nz; z += 1) {
    const b = bases[x * nz + z];
    for (let py = 0; py < P; py += 1) {
      const row = py === 0 ? 0 : (py === 1 ? b - 1 : b + py - 2);
      const val = row < fineNy ? dist((x + 0.5) * h, (row + 0.5) * h, (z + 0.5) * h) : 5 * h;
      phi[(x * P + py) * nz + z] = Math.max(-5 * h, Math.min(5 * h, val));
    }
  }
  return { bases, band, maxDelta: s.maximumNeighborDelta, phi };
}

// ---- the method plugin --------------------------------------------------------
const PARAMS = [
  { kind: 'select', key: 'layout', label: 'Pressure layout', tier: 'coarse', update: 'solver', default: 'uniform',
    hint: 'Uniform MAC grid, or restricted tall cells under a cubic band (tallCells, Chentanez & Mueller 2011).',
    options: [{ value: 'uniform', label: 'Uniform' }, { value: 'tall-cell', label: 'Restricted tall cells' }] },
  { kind: 'select', key: 'precision', label: 'Multigrid precision', tier: 'coarse', update: 'runtime',
    default: 'fp16', hint: 'Storage of the V-cycle preconditioner; CG runs in FP32, refinement in FP64.',
    options: [{ value: 'fp16', label: 'FP16 (scaled)' }, { value: 'bf16', label: 'BF16' },
      { value: 'fp32', label: 'FP32' }] },
  { kind: 'number', key: 'smoothingSweeps', label: 'Smoothing sweeps', tier: 'fine', update: 'runtime', unit: '',
    min: 1, max: 4, step: 1, default: 2, hint: 'Red-black Gauss-Seidel sweeps before and after each coarse correction.' },
  { kind: 'select', key: 'pressureSolver', label: 'Pressure solver', tier: 'coarse', update: 'solver',
    default: 'mgpcg', hint: 'MGPCG to an FP64 relative residual with closed walls, or the reference CM11a '
      + 'multigrid (FP32, 3 full + 4 V-cycles) whose walls hold p >= 0 so liquid may separate from them; '
      + 'open-top containers always use CM11a.',
    options: [{ value: 'mgpcg', label: 'MGPCG (closed walls)' }, { value: 'cm11a', label: 'CM11a (separating walls)' }] },
  { kind: 'select', key: 'velocityTransport', label: 'Velocity advection', tier: 'coarse', update: 'runtime',
    default: 'semi-lagrangian', hint: 'Uniform layout: the velocity is advected through its extrapolation with '
      + 'gravity and viscosity on the GPU before each projection. Semi-Lagrangian traces once; bounded MacCormack '
      + 'adds a forward prediction, its extrapolation, a reverse trace and a limited correction.',
    options: [{ value: 'semi-lagrangian', label: 'Semi-Lagrangian · one pass' },
      { value: 'maccormack', label: 'Bounded MacCormack · three passes' }] },
];

const PRESETS = {
  balanced: { precision: 'fp16', smoothingSweeps: 1 },
  high: { precision: 'fp16', smoothingSweeps: 2 },
  ultra: { precision: 'fp32', smoothingSweeps: 2 },
};

const mi355xProjectionMethod = {
  id: 'mi355x-projection',
  label: 'MI355X MGPCG pressure projection',
  shortLabel: 'MI355X',
  badge: 'HIP',
  description: 'Free-surface MAC-grid pressure projection solved by multigrid-preconditioned CG on one AMD MI355X.',
  detail: 'Ghost-fluid Poisson assembly (theta clamped to [0.05, 1]), FP16/BF16/FP32 V-cycle preconditioner with '
    + 'red-black Gauss-Seidel smoothing, FP32 CG with FP64 alpha/beta, FP64 iterative refinement of the true '
    + 'residual, hand-written HIP kernels for gfx950 behind the fluid_projection C ABI; uniform or restricted '
    + 'tall-cell layouts; closed walls.',
  // The contract's only backend value.  The state lives in HIP device memory
  // behind the addon, not in WebGPU buffers.
  backend: 'webgpu',
  resource: {
    id: 'mi355x-projection', lane: 'fluid', label: 'MI355X pressure projection',
    provides: ['fluid-authority'], blocks: 'transport',
  },
  qualityLabels: {
    balanced: 'FP16 multigrid, one sweep per side, 12-row tall-cell band',
    high: 'FP16 multigrid, two sweeps per side, 32-row tall-cell band',
    ultra: 'FP32 multigrid, two sweeps per side, 40-row tall-cell band',
  },
  showQualityControl: true,
  supportedFieldModes: [],
  params: PARAMS,
  pressureMapping: 'scene.numerics.pressureRelativeTolerance is the FP64 relative residual ||b - Ap||/||b|| the '
    + 'projection reaches; scene.numerics.pressureMaxIterations caps the MGPCG iterations of each refinement pass.',
  presetFor(quality) { return Object.assign({}, PRESETS[quality] || PRESETS.high); },
  normalizeValues(values) {
    const v = Object.assign({}, values);
    v.smoothingSweeps = Math.max(1, Math.min(4, Math.round(Number(v.smoothingSweeps) || 2)));
    if (!(v.precision in PRECISION)) v.precision = 'fp16';
    if (v.layout !== 'tall-cell') v.layout = 'uniform';
    if (!(v.pressureSolver in SOLVER) || v.layout === 'tall-cell') v.pressureSolver = 'mgpcg';
    if (!(v.velocityTransport in ADVECT_SCHEME)) v.velocityTransport = 'semi-lagrangian';
    return v;
  },
  runtimeParamKeys: ['precision', 'smoothingSweeps', 'velocityTransport'],
  capabilities: { volumeRendering: false, adoptsRigidRosterShape: true },

  async createSolverAsync(_device, scene, quality = 'high', values = {}, _onRigidLoads, onProgress, signal) {
    sceneCheck(scene);
    const progress = (phase, label, completed, total) => {
      if (typeof onProgress === 'function') onProgress({ phase, label, completed, total });
      if (signal && signal.aborted) throw new Error('createSolverAsync aborted');
    };
    const vals = this.normalizeValues(Object.assign({}, this.presetFor(quality), values));
    progress('planning', 'Planning the pressure lattice', 0, 3);
    const lat = sceneLattice(scene);
    // an open top is air beyond the +y side: the CM11a solver's boundary
    // (params.boundary.w of the reference); it needs the uniform layout and a
    // lattice whose hierarchy halves exactly (planUniformCM11aHierarchy)
    const openTop = scene.container.top === 'open';
    if (openTop) {
      if (vals.layout === 'tall-cell') throw new RangeError('open-top containers need the uniform layout (CM11a)');
      vals.pressureSolver = 'cm11a';
    }
    if (vals.pressureSolver === 'cm11a' && !cm11aLattice(lat)) {
      throw new RangeError(`the CM11a solver needs a lattice that halves exactly down to 256 haloed cells `
        + `(got ${lat.nx} x ${lat.ny} x ${lat.nz})`);
    }
    const rho = scene.fluid.density_kg_m3;
    const gy = scene.fluid.gravity_m_s2.y;
    const dt = scene.numerics.fixedDt_s;
    let proj, phi, u, v, w, vmask, storedNy, cellCount, regularLayers = lat.ny, maxDelta = 0;
    // a band that leaves no tall cell is the ordinary grid (the reference's
    // ordinaryGridFallback, lib/core/tall-cell-grid.ts:326)
    const tallBand = Math.min(lat.ny, (TALL_SETTINGS[quality] || TALL_SETTINGS.high).regularLayers);
    const ordinaryGridFallback = vals.layout === 'tall-cell' && lat.ny - tallBand < 2;
    if (vals.layout === 'tall-cell' && !ordinaryGridFallback) {
      const t = tallLayout(scene, lat, quality);
      progress('allocation', 'Allocating the tall-cell grid', 1, 3);
      proj = new TallCellPressureProjection({ nx: lat.nx, fineNy: lat.ny, nz: lat.nz, band: t.band, h: lat.h,
        columnBases: t.bases, maxDelta: t.maxDelta });
      const out = proj.outputs();
      phi = t.phi;
      ({ u, v, w } = out);
      const P = t.band + 2;
      // gravity on the v slots that carry faces inside the domain: slot 0
      // the faces inside the tall cell (base >= 2), slot s >= 1 the face at
      // fine height base + s - 1, interior below the lid
      vmask = new Uint8Array(lat.nx * P * lat.nz);
      for (let x = 0; x < lat.nx; x += 1) for (let z = 0; z < lat.nz; z += 1) {
        const b = t.bases[x * lat.nz + z];
        for (let py = 0; py < P; py += 1) {
          vmask[(x * P + py) * lat.nz + z] = py === 0 ? 1 : (b + py - 1 < lat.ny ? 1 : 0);
        }
      }
      storedNy = P;
      cellCount = lat.nx * P * lat.nz;
      regularLayers = t.band;
      maxDelta = t.maxDelta;
    } else {
      progress('allocation', 'Allocating the uniform grid', 1, 3);
      proj = new UniformPressureProjection([lat.nx, lat.ny, lat.nz], lat.h);
      phi = uniformLevelSet(scene, lat);
      ({ u, v, w } = proj.outputs());
      storedNy = lat.ny;
      cellCount = lat.nx * lat.ny * lat.nz;
    }
    progress('upload', 'Level set ready', 2, 3);
    const equivalent = lat.nx * lat.ny * lat.nz;
    const info = {
      nx: lat.nx, ny: lat.ny, nz: lat.nz, storedNy, cellCount,
      equivalentUniformCells: equivalent, compressionRatio: equivalent / cellCount,
      regularLayers, maximumNeighborDelta: maxDelta, gridKind: 'uniform', quality,
      cellSize_m: lat.h, pressureIterations: 0,
      pressureSolver: vals.pressureSolver === 'cm11a'
        ? `MI355X HIP CM11a LCP multigrid (separating walls, FP32, 3 full + 4 V-cycles${openTop ? ', open top' : ''})`
        : `MI355X HIP MGPCG (${vals.precision} multigrid, FP32 CG, FP64 refinement, `
        + `${vals.layout === 'tall-cell' ? 'restricted tall cells' : 'uniform grid'})`,
      allocatedBytes: binding.bytesInUse(proj.context),
      time_s: 0, steps: 0, ordinaryGridFallback,
    };
    let time = 0;
    let pending = Promise.resolve();
    let failure;
    let destroyed = false;
    const opts = () => (vals.pressureSolver === 'cm11a' ? { solver: 'cm11a', openTop: openTop ? 1 : 0 } : {
      precision: vals.precision, preSweeps: vals.smoothingSweeps, postSweeps: vals.smoothingSweeps,
      relativeTolerance: scene.numerics.pressureRelativeTolerance,
      maxIterations: Math.max(1, Math.min(1000, scene.numerics.pressureMaxIterations)),
      warmStart: info.steps > 0 ? 1 : 0,
    });
    const nu = scene.fluid.dynamicViscosity_Pa_s / rho;
    const noSlip = scene.container.fluidWallMode === 'no-slip';
    const stepOnce = () => {
      if (vmask) {
        // tall cells: gravity impulse on the v slots inside the domain
        for (let i = 0; i < v.length; i += 1) if (vmask[i]) v[i] += gy * dt;
      } else {
        // uniform: advection through the extrapolated field, gravity and
        // viscosity on the GPU (the reference's step before its projection)
        const r = proj.advect({ phi, u, v, w }, { dt, scheme: vals.velocityTransport, gravity: gy, viscosity: nu,
          openTop, noSlip });
        info.advectionDeviceMs = r.deviceMs;
      }
      return proj.project({ phi, u, v, w, dt, rho }, opts()).then((res) => {
        ({ u, v, w } = res);
        const r = res.report;
        info.steps += 1;
        info.pressureIterations = r.iterations;
        info.pressureIterationsExecuted = r.iterations;
        info.pressureRelativeResidual = r.relativeResidual;
        info.pressureResidual = r.relativeResidual * r.rhsNorm;
        info.maxDivergenceBefore_s = r.maxDivergenceBefore;
        info.maxDivergenceAfter_s = r.maxDivergenceAfter;
        info.maxDivergence_s = r.maxDivergenceAfter;
        info.pressureSolveConverged = (r.flags & FLAGS.NONCONVERGENCE) === 0;
        info.pressureIterationCapReached = !info.pressureSolveConverged && (r.flags & 0x7) === 0;
        info.pressureConvergenceReason = info.pressureSolveConverged ? 'tolerance' : 'iteration-cap';
        info.pressureCurvatureBreakdown = (r.flags & FLAGS.NONPOSITIVE_CURVATURE) !== 0;
        info.pressureFlags = r.flags;
        info.pressureDeviceMs = r.deviceMs;
        if (vals.pressureSolver === 'cm11a') {
          // the reference's own CM11a readout: projected residual in 1/s, coarse-solve cap failures
          info.pressureResidual = r.lcpResidual;
          info.pressureIterationCapReached = r.coarseFailures > 0;
          info.pressureConvergenceReason = r.coarseFailures > 0 ? 'iteration-cap' : 'tolerance';
        }
        return res;
      });
    };
    progress('attach', 'Solver attached', 3, 3);
    return {
      info,
      // No WebGPU texture exists without a WebGPU device: the field lives in
      // HIP device memory behind the addon.
      volumeTexture: undefined,
      advanceTo(time_s, _bodies) {
        if (destroyed || failure) return false;
        let stepped = false;
        while (time + dt <= time_s + 1e-9 * dt) {
          time += dt;
          stepped = true;
          pending = pending.then(stepOnce).catch((e) => { failure = e; });
        }
        info.time_s = time;
        return stepped;
      },
      applyRuntimeValues(values) {
        const n = mi355xProjectionMethod.normalizeValues(Object.assign({}, vals, values));
        vals.precision = n.precision;
        vals.smoothingSweeps = n.smoothingSweeps;
        vals.velocityTransport = n.velocityTransport;
      },
      async readStats() {
        await pending;
        if (failure) throw failure;
        info.allocatedBytes = binding.bytesInUse(proj.context);
        return Object.assign({}, info);
      },
      destroy() {
        if (destroyed) return;
        destroyed = true;
        pending.then(() => proj.destroy(), () => proj.destroy());
      },
    };
  },
};

// True when the lattice's multigrid plan halves every axis exactly (or keeps
// it) down to a coarsest lattice of at most 256 cells with its halo: the
// hierarchies planUniformCM11aHierarchy accepts
// (lib/methods/uniform/webgpu-uniform-pressure-multigrid.ts:149-169).
function cm11aLattice(lat) {
  const d = [lat.nx, lat.ny, lat.nz];
  const haloed = (e) => (e[0] + 2) * (e[1] + 2) * (e[2] + 2);
  const m = Math.min(...d);
  if ((m & (m - 1)) === 0) {
    const f = m >= 2 ? m / 2 : 1;
    if (d.every((x) => x % f === 0) && haloed(d.map((x) => x / f)) <= 256) return true;
  }
  let e = d.slice();
  for (;;) {
    const n = e.map((x) => (x % 2 === 0 && x > 2 ? x / 2 : x));
    if (n.every((x, a) => x === e[a])) break;
    e = n;
  }
  return haloed(e) <= 256;
}

module.exports = {
  binding, UniformPressureProjection, TallCellPressureProjection, projectBatch, mi355xProjectionMethod,
  sceneLattice, damBreakBox, uniformLevelSet, tallLayout, cm11aLattice, PRECISION, SOLVER, ADVECT_SCHEME, FLAGS,
};
