'use strict';
// Drives the SimulationMethod plugin and the addon's entry points the way the
// reference's TypeScript host would (used by tests/test_node_addon.py).
// argv[2]: "contract" (no GPU) or "run" (projections on the GPU).
const path = require('path');
const m = require(path.join(__dirname, '..', '..', 'fluid_amd', 'node'));

function scene(initialCondition, extra = {}) {
  return Object.assign({
    schemaVersion: '1.0.0', sceneId: 'plugin-check', randomSeed: 7, duration_s: 1,
    container: { width_m: 1.2, height_m: 0.8, depth_m: 0.8, fillFraction: 0.22, top: 'closed', fluidWallMode: 'no-slip' },
    voxelDomain: { finestCellSize_m: 0.025, brickSize_cells: 8 },
    fluid: { density_kg_m3: 998.2, dynamicViscosity_Pa_s: 0.001002, surfaceTension_N_m: 0.072,
      gravity_m_s2: { x: 0, y: -9.80665, z: 0 }, initialCondition },
    nominalResolution: { length_m: 0.025 },
    numerics: { fixedDt_s: 0.016, maxDt_s: 0.016, pressureRelativeTolerance: 1e-6, pressureMaxIterations: 100 },
    rigidBodies: [],
  }, extra);
}

function contract() {
  const meth = m.mi355xProjectionMethod;
  const out = {};
  for (const k of ['id', 'label', 'shortLabel', 'badge', 'description', 'detail', 'backend', 'pressureMapping']) {
    out[k] = typeof meth[k];
  }
  out.backend_value = meth.backend;
  out.resource = meth.resource;
  out.qualityLabels = Object.keys(meth.qualityLabels).sort();
  out.params = meth.params.map((p) => ({ key: p.key, kind: p.kind, tier: p.tier, default: p.default,
    ok: p.kind === 'select' ? p.options.some((o) => o.value === p.default)
      : (p.min <= p.default && p.default <= p.max && typeof p.unit === 'string' && p.step > 0) }));
  out.presets = {};
  for (const q of ['balanced', 'high', 'ultra']) out.presets[q] = meth.presetFor(q);
  out.normalized = meth.normalizeValues({ precision: 'fp64', smoothingSweeps: 9, layout: 'x' });
  out.createSolverAsync = typeof meth.createSolverAsync;
  out.createSolverAsyncArity = meth.createSolverAsync.length;
  out.binding = Object.keys(m.binding).sort();
  out.lattice = m.sceneLattice(scene('dam-break'));
  out.openTop = (() => { try { m.sceneLattice(scene('dam-break')); return meth.createSolverAsync(null,
    scene('dam-break', { container: { width_m: 1, height_m: 1, depth_m: 1, fillFraction: 0.2, top: 'open',
      fluidWallMode: 'no-slip' } })).then(() => 'accepted', (e) => e.constructor.name); } catch (e) { return e.constructor.name; } })();
  return Promise.resolve(out.openTop).then((v) => { out.openTop = v; return out; });
}

async function run() {
  const out = {};
  const meth = m.mi355xProjectionMethod;
  for (const layout of ['uniform', 'tall-cell']) {
    const progress = [];
    // the tall-cell case in a 1.6 m tall tank: 64 rows under a 32-row band
    const sc = layout === 'uniform' ? scene('dam-break') : scene('dam-break', { container: { width_m: 1.2,
      height_m: 1.6, depth_m: 0.8, fillFraction: 0.22, top: 'closed', fluidWallMode: 'no-slip' } });
    const solver = await meth.createSolverAsync(null, sc, 'high', { layout }, undefined,
      (p) => progress.push(p.phase));
    const stepped = [solver.advanceTo(0.016, []), solver.advanceTo(0.048, []), solver.advanceTo(0.05, [])];
    const info = await solver.readStats();
    solver.destroy();
    solver.destroy();
    out[layout] = { stepped, info, progress };
  }
  // the CM11a solver: an open-top tank (32^3 lattice), and the closed tank by choice
  const openTank = scene('dam-break', { container: { width_m: 0.8, height_m: 0.8, depth_m: 0.8, fillFraction: 0.22,
    top: 'open', fluidWallMode: 'no-slip' } });
  for (const [name, sc, values] of [['cm11aOpenTop', openTank, {}],
    ['cm11aClosed', scene('dam-break'), { pressureSolver: 'cm11a' }],
    ['maccormack', scene('dam-break'), { velocityTransport: 'maccormack' }]]) {
    const solver = await meth.createSolverAsync(null, sc, 'high', values);
    const stepped = [solver.advanceTo(0.016, []), solver.advanceTo(0.048, [])];
    const info = await solver.readStats();
    solver.destroy();
    out[name] = { stepped, info };
  }
  // the advection scheme switched at run time (applyRuntimeValues)
  {
    const solver = await meth.createSolverAsync(null, scene('dam-break'), 'high', {});
    solver.advanceTo(0.016, []);
    solver.applyRuntimeValues({ velocityTransport: 'maccormack' });
    solver.advanceTo(0.032, []);
    const info = await solver.readStats();
    solver.destroy();
    out.runtimeSwitch = { info };
  }
  // addon entry points: project, submit and projectBatch agree
  const dims = [24, 20, 16];
  const h = 1 / 24;
  const phi = new Float32Array(dims[0] * dims[1] * dims[2]);
  for (let i = 0; i < dims[0]; i += 1) for (let j = 0; j < dims[1]; j += 1) for (let k = 0; k < dims[2]; k += 1) {
    phi[(i * dims[1] + j) * dims[2] + k] = (j + 0.5) * h - 0.4 + 0.05 * Math.sin(i * 0.7 + k * 0.3);
  }
  const a = new m.UniformPressureProjection(dims, h);
  const b = new m.UniformPressureProjection(dims, h);
  const vel = () => ({ u: new Float32Array(25 * 20 * 16).fill(0.1), v: new Float32Array(24 * 21 * 16).fill(-0.2),
    w: new Float32Array(24 * 20 * 17).fill(0.05) });
  const inputs = Object.assign({ phi, dt: 1 / 120, rho: 998.2 }, vel());
  const p1 = await a.project(inputs, { precision: 'fp16' });
  const p2 = await a.submit(inputs, { precision: 'fp16' });
  const [q1, q2] = await m.projectBatch([a, b], [inputs, inputs], { precision: 'fp16' });
  const same = (x, y) => x.length === y.length && x.every((v, i) => v === y[i]);
  out.entryPoints = {
    submitEqualsProject: same(p1.pressure, p2.pressure) && same(p1.v, p2.v),
    batchEqualsProject: same(p1.pressure, q1.pressure) && same(p1.pressure, q2.pressure),
    latencyMs: p2.report.latencyMs, flags: [p1.report.flags, p2.report.flags, q1.report.flags, q2.report.flags],
  };
  // destroy with a projection in flight: deferred until it settles
  const pending = a.project(inputs, { precision: 'fp16' });
  a.destroy();
  const late = await pending;
  out.destroyWhilePending = { flags: late.report.flags, freed: a.grid === null };
  b.destroy();
  return out;
}

const mode = process.argv[2] || 'contract';
(mode === 'run' ? run() : contract()).then((o) => console.log(JSON.stringify(o)))
  .catch((e) => { console.error(e.code, e.message, e.stack); process.exit(1); });
