// MPM.ts -- the reference's MLS-MPM simulator class (src/simulator/MPM/MPM.ts,
// commented out in the reference) over the Node-API addon.  Same member names:
// the constructor's parameters (MPM.ts:60-83), initResource (MPM.ts:94-121)
// and run (MPM.ts:204-238).  MPM.js is the runnable twin for Node without a
// TypeScript compiler.
// eslint-disable-next-line @typescript-eslint/no-var-requires
const addon = require('./pbf_addon.node');

export interface MPMParams {
  particleCount?: number; nGrid?: number; dt?: number; bound?: number;
  rho?: number; E?: number; gravity?: number; flags?: number;
}

export const MPM_FIELD = { position: 0, velocity: 1, affine: 2, grid: 3 } as const;

export class MPM {
  public particleCount: number;
  public stepCount = 15;  // steps per frame (MPM.ts:64)
  public pause = true;
  private readonly nGrid: number;
  private readonly handle: unknown;

  constructor(params: MPMParams = {}) {
    this.particleCount = params.particleCount ?? 9000;
    this.nGrid = params.nGrid ?? 32;
    this.handle = addon.mpmCreate({ capacity: this.particleCount, ...params });
  }

  /** MPM.initResource: particles uniform in [0.15, 0.55]^3 (J = 1, v = 0, C = 0). */
  public initResource(random: () => number = Math.random): Float32Array {
    const p = new Float32Array(4 * this.particleCount);
    for (let i = 0; i < this.particleCount; i++)
      for (let a = 0; a < 3; a++) p[4 * i + a] = random() * 0.4 + 0.15;
    this.setParticlePosition(p);
    return p;
  }

  public setParticlePosition(xyzw: Float32Array): void {
    addon.mpmSetPositions(this.handle, xyzw);
    this.particleCount = xyzw.length / 4;
  }

  /** MPM.run: one P2G, grid and G2P pass (`steps` of them). */
  public run(steps = 1): void { addon.mpmStep(this.handle, steps); }
  public start(): void { this.pause = false; }
  public stop(): void { this.pause = true; }
  public sync(): void { addon.mpmSync(this.handle); }

  public readPositions(): Float32Array { return addon.mpmRead(this.handle, MPM_FIELD.position); }
  public readVelocities(): Float32Array { return addon.mpmRead(this.handle, MPM_FIELD.velocity); }
  public readAffine(): Float32Array { return addon.mpmRead(this.handle, MPM_FIELD.affine); }
  public readGrid(): Float32Array { return addon.mpmRead(this.handle, MPM_FIELD.grid, this.nGrid); }
}
