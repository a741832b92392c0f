'use strict';
// Advects the velocities of a uniform grid through the N-API addon on inputs
// stored as raw float32 files and writes the results next to them (used by
// tests/test_node_addon.py).  argv: dir nx ny nz h dt scheme viscosity noSlip
const fs = require('fs');
const path = require('path');
const { UniformPressureProjection } = require(path.join(__dirname, '..', '..', 'fluid_amd', 'node'));

function readF32(file) {
  const b = fs.readFileSync(file);
  return new Float32Array(b.buffer.slice(b.byteOffset, b.byteOffset + b.byteLength));
}

const dir = process.argv[2];
const [nx, ny, nz, h, dt] = process.argv.slice(3, 8).map(Number);
const scheme = process.argv[8];
const viscosity = Number(process.argv[9]);
const noSlip = process.argv[10] === '1';
const proj = new UniformPressureProjection([nx, ny, nz], h);
const f = {};
for (const k of ['phi', 'u', 'v', 'w']) f[k] = readF32(path.join(dir, `${k}.f32`));
try {
  let refusedScheme = false;
  try {
    proj.advect(f, { dt, scheme: 'upwind' });
  } catch (e) {
    refusedScheme = e instanceof RangeError;
  }
  const rep = proj.advect(f, { dt, scheme, gravity: -9.80665, viscosity, noSlip });
  for (const k of ['u', 'v', 'w']) fs.writeFileSync(path.join(dir, `${k}.out.f32`), Buffer.from(f[k].buffer));
  console.log(JSON.stringify({ refusedScheme, timed: rep.deviceMs > 0 }));
} catch (e) {
  console.error(e.code, e.message);
  process.exitCode = 1;
}
proj.destroy();
