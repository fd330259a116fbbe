'use strict';
// PBF.js -- CommonJS twin of PBF.ts (the same class, written for Node without
// a TypeScript compiler).  Member names follow the reference simulator class
// (src/simulator/SPH.ts, PBF/PBFConfig.ts, PBF/PBF.ts of
// LinzhouLi/WebGPU-Fluid-Simulation).
const addon = require('./pbf_addon.node');

const FIELD = { position: 0, velocity: 1, acceleration: 2, lambda: 3, omega: 4, normal: 5 };

class PBF {
  constructor(params) {
    this.pause = true;
    this.stepCount = 1;
    this.handle = addon.create(Object.assign({ capacity: PBF.MAX_PARTICLE_NUM }, params || {}));
  }
  reset() { addon.reset(this.handle); }
  setConfig(conf) { addon.setConfig(this.handle, conf); }
  optionsChange(e) { this.setConfig(e.object); }
  voxelizeCube(min, max, jitter, seed) { return addon.voxelizeCube(this.handle, min, max, jitter || 0, seed || 0); }
  voxelizeSphere(center, radius, jitter, seed) {
    return addon.voxelizeSphere(this.handle, center, radius, jitter || 0, seed || 0);
  }
  voxelizeMesh(positions, index, matrixWorld, jitter, seed) {
    return addon.voxelizeMesh(this.handle, positions, index, matrixWorld || null, jitter || 0, seed || 0);
  }
  setWarmStart(blob, mode, relax, fallbackThreshold, fallbackIterations) {
    addon.setWarmStart(this.handle, blob || null, mode === undefined ? 1 : mode, relax === undefined ? 0.5 : relax,
                       fallbackThreshold || 0, fallbackIterations || 0);
  }
  autotune(dt, iterations, repeats) { return addon.autotune(this.handle, dt, iterations || 0, repeats || 3); }
  setTuning(walkBlock, listVariant) { addon.setTuning(this.handle, walkBlock, listVariant); }
  setParticlePosition(positions, count) {
    if (positions) addon.setParticlePosition(this.handle, positions, count === undefined ? positions.length / 4 : count);
    else addon.setParticlePosition(this.handle);
  }
  setBoundaryData(data) { addon.setBoundaryData(this.handle, data); }
  setVolumeMap(scene, res, lo, hi, sdfVol) { addon.setVolumeMap(this.handle, scene, res, lo, hi, sdfVol); }
  get particleCount() { return addon.particleCount(this.handle); }
  switch() { this.pause = !this.pause; }
  stop() { this.pause = true; }
  start() { this.pause = false; }
  run() { if (!this.pause) addon.run(this.handle); }
  step(dt, substeps, iterations) { addon.step(this.handle, dt, substeps || 1, iterations || 0); }
  sync() { addon.sync(this.handle); }
  readPositions() { return addon.readField(this.handle, FIELD.position); }
  readVelocities() { return addon.readField(this.handle, FIELD.velocity); }
  readField(field) { return addon.readField(this.handle, field); }
  densityError(scene) { return addon.densityError(this.handle, scene); }
  static version() { return addon.version(); }
}
PBF.MAX_PARTICLE_NUM = 1 << 17;

module.exports = { PBF, FIELD };
