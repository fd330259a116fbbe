// PBF.ts -- TypeScript face of the MI355X engine, member for member the
// reference simulator class (LinzhouLi/WebGPU-Fluid-Simulation:
// src/simulator/SPH.ts, src/simulator/PBF/PBFConfig.ts, src/simulator/PBF/PBF.ts)
// over the Node-API addon (pbf_addon.node -> libpbf.so).  PBF.js is the
// CommonJS twin that Node runs directly (no TypeScript compiler is needed).
/* eslint-disable @typescript-eslint/no-var-requires */
const addon = require('./pbf_addon.node');

export interface Vec3 { x: number; y: number; z: number; }
export interface SimulationOptions {
  iteration: number; XSPH: number; vorticity: number; surfaceTension: number; gravity: number;
}
export interface PBFParams {
  capacity?: number; maxNeighbors?: number; particleRadius?: number; kernelRadius?: number;
  restDensity?: number; timeStep?: number; domainLo?: Vec3 | number[]; domainHi?: Vec3 | number[];
  gridMode?: number; hashLog2?: number; flags?: number; cellSlabs?: number;
}

export const FIELD = { position: 0, velocity: 1, acceleration: 2, lambda: 3, omega: 4, normal: 5 } as const;

export class PBF {
  // SPH.ts:24-27 as instance parameters
  static MAX_PARTICLE_NUM = 1 << 17;
  public pause = true;
  public stepCount = 1;
  private handle: unknown;

  constructor(params: PBFParams = {}) {
    this.handle = addon.create({ capacity: PBF.MAX_PARTICLE_NUM, ...params });
  }

  /** PBF.reset (PBF.ts:71-89). */
  public reset(): void { addon.reset(this.handle); }
  /** PBFConfig.setConfig (PBFConfig.ts:31-46). */
  public setConfig(conf: SimulationOptions): void { addon.setConfig(this.handle, conf); }
  public optionsChange(e: { object: SimulationOptions }): void { this.setConfig(e.object); }
  /** SPH.voxelizeCube (SPH.ts:99-121); returns the number of particles added. */
  public voxelizeCube(min: Vec3, max: Vec3, jitter = 0, seed = 0): number {
    return addon.voxelizeCube(this.handle, min, max, jitter, seed);
  }
  /** SPH.voxelizeSphere (SPH.ts:123-149). */
  public voxelizeSphere(center: Vec3, radius: number, jitter = 0, seed = 0): number {
    return addon.voxelizeSphere(this.handle, center, radius, jitter, seed);
  }
  /** SPH.voxelizeMesh (SPH.ts:151-187): pass mesh.geometry.attributes.position.array,
   *  mesh.geometry.index.array and mesh.matrixWorld.elements. */
  public voxelizeMesh(positions: Float32Array, index: Uint32Array | Uint16Array, matrixWorld?: ArrayLike<number>,
                      jitter = 0, seed = 0): number {
    return addon.voxelizeMesh(this.handle, positions, index, matrixWorld ?? null, jitter, seed);
  }
  /** Learned warm start (no reference counterpart): mode 0 plain, 1 network, 2 previous lambda sum,
   *  3 network factor on the previous lambda sum. */
  public setWarmStart(blob: Float32Array | null, mode = 1, relax = 0.5, fallbackThreshold = 0, fallbackIterations = 0): void {
    addon.setWarmStart(this.handle, blob, mode, relax, fallbackThreshold, fallbackIterations);
  }
  /** Times the launch settings on the current scene and keeps the fastest. */
  public autotune(dt: number, iterations = 0, repeats = 3): { walkBlock: number[]; listVariant: number } {
    return addon.autotune(this.handle, dt, iterations, repeats);
  }
  public setTuning(walkBlock: number[], listVariant: number): void { addon.setTuning(this.handle, walkBlock, listVariant); }
  /** PBFConfig.setParticlePosition (PBFConfig.ts:48-62): staged particles, or an explicit float4 array. */
  public setParticlePosition(positions?: Float32Array, count?: number): void {
    if (positions) addon.setParticlePosition(this.handle, positions, count ?? positions.length / 4);
    else addon.setParticlePosition(this.handle);
  }
  /** PBF.setBoundaryData (PBF.ts:91-94): .cdm text. */
  public setBoundaryData(data: string): void { addon.setBoundaryData(this.handle, data); }
  public setVolumeMap(scene: number, res: number[], lo: number[], hi: number[], sdfVol: Float32Array): void {
    addon.setVolumeMap(this.handle, scene, res, lo, hi, sdfVol);
  }
  public get particleCount(): number { return addon.particleCount(this.handle); }
  public switch(): void { this.pause = !this.pause; }
  public stop(): void { this.pause = true; }
  public start(): void { this.pause = false; }
  /** PBF.run (PBF.ts:398-448): one substep; the encoder argument of the WebGPU version is not needed. */
  public run(_encoder?: unknown): void { if (!this.pause) addon.run(this.handle); }
  public step(dt: number, substeps = 1, iterations = 0): void { addon.step(this.handle, dt, substeps, iterations); }
  public sync(): void { addon.sync(this.handle); }
  /** The `position` buffer (SPH.ts:39): float4 (x, y, z, density) per particle. */
  public readPositions(): Float32Array { return addon.readField(this.handle, FIELD.position); }
  public readVelocities(): Float32Array { return addon.readField(this.handle, FIELD.velocity); }
  public readField(field: number): Float32Array { return addon.readField(this.handle, field); }
  public densityError(scene?: number): [number, number] { return addon.densityError(this.handle, scene); }
  public static version(): string { return addon.version(); }
}
