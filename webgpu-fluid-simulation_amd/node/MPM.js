// MPM.js -- runnable twin of MPM.ts (CommonJS, no TypeScript compiler needed).
'use strict';
const addon = require('./pbf_addon.node');

const MPM_FIELD = { position: 0, velocity: 1, affine: 2, grid: 3 };

class MPM {
  constructor(params) {
    params = params || {};
    this.particleCount = params.particleCount !== undefined ? params.particleCount : 9000;
    this.stepCount = 15;
    this.pause = true;
    this.nGrid = params.nGrid !== undefined ? params.nGrid : 32;
    this.handle = addon.mpmCreate(Object.assign({ capacity: this.particleCount }, params));
  }

  initResource(random) {
    random = random || Math.random;
    const p = new Float32Array(4 * this.particleCount);
    for (let i = 0; i < this.particleCount; i++)
      for (let a = 0; a < 3; a++) p[4 * i + a] = random() * 0.4 + 0.15;
    this.setParticlePosition(p);
    return p;
  }

  setParticlePosition(xyzw) {
    addon.mpmSetPositions(this.handle, xyzw);
    this.particleCount = xyzw.length / 4;
  }

  run(steps) { addon.mpmStep(this.handle, steps === undefined ? 1 : steps); }
  start() { this.pause = false; }
  stop() { this.pause = true; }
  sync() { addon.mpmSync(this.handle); }

  readPositions() { return addon.mpmRead(this.handle, MPM_FIELD.position); }
  readVelocities() { return addon.mpmRead(this.handle, MPM_FIELD.velocity); }
  readAffine() { return addon.mpmRead(this.handle, MPM_FIELD.affine); }
  readGrid() { return addon.mpmRead(this.handle, MPM_FIELD.grid, this.nGrid); }
}

module.exports = { MPM, MPM_FIELD };
