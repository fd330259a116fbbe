'use strict';
// Drives the reference "Cube Drop" scene flow (src/controller.ts:203-208,
// 247-248) through the TypeScript-facing class and prints a JSON summary.
// Usage: node run_scene.js [steps] [out.bin]
const fs = require('fs');
const { PBF } = require('./PBF.js');

const steps = parseInt(process.argv[2] || '3', 10);
const out = process.argv[3];
const sim = new PBF();
sim.reset();
const n = sim.voxelizeCube({ x: 0.15, y: 0.35, z: 0.15 }, { x: 0.65, y: 0.85, z: 0.65 });
sim.setConfig({ iteration: 5, XSPH: 0.01, vorticity: 0.1, surfaceTension: 0.5, gravity: 9.8 });
sim.setParticlePosition();
sim.start();
for (let s = 0; s < steps; s++) sim.run();
sim.sync();
const p = sim.readPositions();
let err = null;
try { sim.setParticlePosition(new Float32Array(10), 3); } catch (e) { err = e.message; }
if (out) fs.writeFileSync(out, Buffer.from(p.buffer));
console.log(JSON.stringify({ version: PBF.version(), particles: n, steps, y0: p[1], rho0: p[3], error: err }));
