"""Synthetic workloads of the PBF engine: dam breaks, obstacle drops, batches.

Every scene is generated on the host from a seed (lattice of spacing 2r via
libpbf's GPU-free voxelizers + seeded jitter), so the GPU engine and the CPU
oracle receive the same float32 inputs.  Scene geometry keeps r, h, g and dt
physical and scales the box to the particle block: a block of extent E sits in
the low corner of a cube of side 2 max(E), whose walls are both a trilinear
volume map and the wall clamp.

Reference scenes (r = 0.006, h = 0.025, unit box, dt = 1/300) follow
src/controller.ts:196-246 of LinzhouLi/WebGPU-Fluid-Simulation.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field as dc_field
from typing import List, Optional

import numpy as np

from pbf_host import (PBF, PBF_GRID_DENSE, PBF_GRID_HASH, VOLUME_REFERENCE, build_cubic_field, build_volume_map,
                      load_library)

SEED = 0x5EED


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


def lattice_cube(r: float, mn, mx, jitter: float = 0.0, seed: int = 0, serial0: int = 0) -> np.ndarray:
    """SPH.voxelizeCube (SPH.ts:99-121) through pbf_voxelize_cube_host."""
    lib = load_library()
    cnt = C.c_uint32()
    rc = lib.pbf_voxelize_cube_host(float(r), _d3(mn), _d3(mx), float(jitter), int(seed), int(serial0), None, 0,
                                    C.byref(cnt))
    assert rc == 0, rc
    out = np.zeros((cnt.value, 4), np.float32)
    rc = lib.pbf_voxelize_cube_host(float(r), _d3(mn), _d3(mx), float(jitter), int(seed), int(serial0),
                                    out.ctypes.data_as(C.POINTER(C.c_float)), out.shape[0], C.byref(cnt))
    assert rc == 0, rc
    return out


def lattice_sphere(r: float, center, radius: float, jitter: float = 0.0, seed: int = 0,
                   serial0: int = 0) -> np.ndarray:
    """SPH.voxelizeSphere (SPH.ts:123-149) through pbf_voxelize_sphere_host."""
    lib = load_library()
    cnt = C.c_uint32()
    rc = lib.pbf_voxelize_sphere_host(float(r), _d3(center), float(radius), float(jitter), int(seed), int(serial0),
                                      None, 0, C.byref(cnt))
    assert rc == 0, rc
    out = np.zeros((cnt.value, 4), np.float32)
    rc = lib.pbf_voxelize_sphere_host(float(r), _d3(center), float(radius), float(jitter), int(seed), int(serial0),
                                      out.ctypes.data_as(C.POINTER(C.c_float)), out.shape[0], C.byref(cnt))
    assert rc == 0, rc
    return out


@dataclass
class Field:
    kind: str                  # "trilinear" or "cubic"
    res: tuple
    lo: tuple
    hi: tuple
    sdf_vol: Optional[np.ndarray] = None   # trilinear (nodes, 2)
    sdf: Optional[np.ndarray] = None       # cubic
    vol: Optional[np.ndarray] = None


@dataclass
class Scene:
    name: str
    particle_radius: float
    kernel_radius: float
    domain_lo: tuple
    domain_hi: tuple
    positions: np.ndarray
    field: Optional[Field] = None
    config: dict = dc_field(default_factory=dict)
    dt: float = 1.0 / 60.0
    substeps: int = 2
    iterations: int = 4
    steps: int = 200
    grid_mode: int = PBF_GRID_DENSE
    hash_log2: int = 22

    @property
    def n(self) -> int:
        return int(self.positions.shape[0])

    def substep_dt(self) -> float:
        # pbf_step receives dt as float32 and divides in double
        return float(np.float32(float(np.float32(self.dt)) / self.substeps))


def _block(nx, ny, nz, r, lo, jitter, seed):
    d = 2.0 * r
    mn = (lo[0] + r, lo[1] + r, lo[2] + r)
    mx = (mn[0] + (nx + 0.5) * d, mn[1] + (ny + 0.5) * d, mn[2] + (nz + 0.5) * d)
    p = lattice_cube(r, mn, mx, jitter, seed)
    assert p.shape[0] == nx * ny * nz, (p.shape, nx, ny, nz)
    return p


def dam_break(nx: int, ny: int, nz: int, r: float = 0.025, map_res: int = 64, xsph: float = 0.01,
              vorticity: float = 0.0, tension: float = 0.0, iterations: int = 4, steps: int = 200,
              seed: int = SEED, name: str = "dam_break", with_map: bool = True) -> Scene:
    """Block of nx x ny x nz particles in the low corner of a cube of side 2 max(extent)."""
    h = 4.0 * r
    ext = 2.0 * r * max(nx, ny, nz)
    L = 2.0 * ext
    lo, hi = (0.0, 0.0, 0.0), (L, L, L)
    pos = _block(nx, ny, nz, r, lo, 0.01 * r, seed)
    fld = None
    if with_map:
        res = (map_res, map_res, map_res)
        fld = Field("trilinear", res, lo, hi, sdf_vol=build_volume_map(0, res, lo, hi, h, None, r))
    cfg = {"iteration": iterations, "XSPH": xsph, "vorticity": vorticity, "surfaceTension": tension, "gravity": 9.8}
    return Scene(name, r, h, lo, hi, pos, fld, cfg, 1.0 / 60.0, 2, iterations, steps)


def dam_break_16k(**kw) -> Scene:
    """16,384 particles (32x32x16), 64^3 box map, XSPH only, 200 steps."""
    return dam_break(32, 32, 16, map_res=64, name="dam_break_16k", seed=SEED + 0, **kw)


def dam_break_1m(**kw) -> Scene:
    """1,048,576 particles (128x128x64), 128^3 box map, vorticity 0.01, tension 0.1, 500 steps."""
    return dam_break(128, 128, 64, map_res=128, vorticity=0.01, tension=0.1, steps=500,
                     name="dam_break_1m", seed=SEED + 1, **kw)


def dam_break_physics(nx=32, ny=32, nz=16, map_res=32) -> Scene:
    """The 1M dam break's physics (vorticity + tension) at a size the oracle runs in seconds."""
    return dam_break(nx, ny, nz, map_res=map_res, vorticity=0.01, tension=0.1, steps=500,
                     name=f"dam_break_physics_{nx*ny*nz}", seed=SEED + 1)


def obstacle_drop(nx: int, ny: int, nz: int, r: float = 0.025, map_res: int = 256, hash_log2: int = 22,
                  seed: int = SEED + 2, name: str = "obstacle_drop") -> Scene:
    """Column of nx x ny x nz particles over a sphere (R = 0.2 L) and a torus
    (R = 0.3 L, r = 0.08 L, axis y) at the centre of a box of side L, hash grid."""
    h = 4.0 * r
    d = 2.0 * r
    L = 4.0 * d * max(nx, nz)
    Ly = L * 0.5 + 0.2 * L + ny * d + 4 * d
    lo, hi = (0.0, 0.0, 0.0), (L, max(L, Ly), L)
    c = (0.5 * L, 0.5 * L, 0.5 * L)
    shape = [c[0], c[1], c[2], 0.2 * L, c[0], c[1], c[2], 0.3 * L, 0.08 * L]
    y0 = c[1] + 0.2 * L + 2 * d
    mn = (c[0] - 0.5 * nx * d, y0, c[2] - 0.5 * nz * d)
    mx = (mn[0] + (nx + 0.5) * d, mn[1] + (ny + 0.5) * d, mn[2] + (nz + 0.5) * d)
    pos = lattice_cube(r, mn, mx, 0.01 * r, seed)
    res = (map_res, map_res, map_res)
    fld = Field("trilinear", res, lo, hi, sdf_vol=build_volume_map(1, res, lo, hi, h, shape, r))
    cfg = {"iteration": 4, "XSPH": 0.01, "vorticity": 0.0, "surfaceTension": 0.0, "gravity": 9.8}
    return Scene(name, r, h, lo, hi, pos, fld, cfg, 1.0 / 60.0, 2, 4, 300, PBF_GRID_HASH, hash_log2)


def obstacle_drop_4m(**kw) -> Scene:
    """4,194,304 particles (128x256x128) dropped on sphere + torus, 256^3 map, 2^22 hash table."""
    return obstacle_drop(128, 256, 128, map_res=256, name="obstacle_drop_4m", **kw)


def batch_specs(nscenes: int = 256, nmin: int = 2048, nmax: int = 65536, seed: int = SEED + 3):
    """Per-scene lattice dims: N log-uniform in [nmin, nmax], x/z aspect ratios U(0.5, 2)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(nscenes):
        n = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
        ax, az = rng.uniform(0.5, 2.0), rng.uniform(0.5, 2.0)
        c = (n / (ax * az)) ** (1.0 / 3.0)
        out.append((max(2, round(c * ax)), max(2, round(c)), max(2, round(c * az))))
    return out


def batch_scenes(nscenes: int = 256, map_res: int = 64, nmin: int = 2048, nmax: int = 65536) -> List[Scene]:
    """Independent dam breaks of uneven size, each with its own 64^3 box map."""
    scenes = []
    for k, (nx, ny, nz) in enumerate(batch_specs(nscenes, nmin, nmax)):
        scenes.append(dam_break(nx, ny, nz, map_res=map_res, seed=SEED + 3 + 1000 * (k + 1),
                                name=f"batch_scene{k}"))
    return scenes


# --- reference scenes (src/controller.ts:196-246) --------------------------
REF_R, REF_H, REF_DT = 0.006, 0.025, 1.0 / 300.0
REF_CONFIG = {"iteration": 5, "XSPH": 0.01, "vorticity": 0.1, "surfaceTension": 0.5, "gravity": 9.8}


def _ref_scene(name, pos, fld=None):
    return Scene(name, REF_R, REF_H, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), pos, fld, dict(REF_CONFIG), REF_DT, 1,
                 5, 300)


def ref_cube_drop() -> Scene:
    return _ref_scene("ref_cube_drop", lattice_cube(REF_R, (0.15, 0.35, 0.15), (0.65, 0.85, 0.65)))


def ref_water_droplet() -> Scene:
    a = lattice_cube(REF_R, (0.005, 0.005, 0.005), (0.995, 0.08, 0.995))
    b = lattice_sphere(REF_R, (0.5, 0.7, 0.5), 0.12)
    return _ref_scene("ref_water_droplet", np.concatenate([a, b]))


def ref_double_dam_break() -> Scene:
    a = lattice_cube(REF_R, (0.005, 0.005, 0.005), (0.3, 0.6, 0.3))
    b = lattice_cube(REF_R, (0.7, 0.005, 0.7), (0.995, 0.6, 0.995))
    return _ref_scene("ref_double_dam_break", np.concatenate([a, b]))


def ref_boundary(field: Field) -> Scene:
    return _ref_scene("ref_boundary", lattice_cube(REF_R, (0.15, 0.35, 0.15), (0.65, 0.85, 0.65)), field)


# --- cubic (.cdm layout) fields ------------------------------------------------
def cubic_nodes(res):
    """Node coordinates of the four .cdm node sets over the unit box."""
    R0, R1, R2 = res
    g = [np.arange(R + 1) / R for R in res]
    e = [(np.arange(2 * R) // 2 + (1 + np.arange(2 * R) % 2) / 3.0) / R for R in res]
    K, J, I = np.meshgrid(g[2], g[1], g[0], indexing="ij")
    corner = (I, J, K)
    K, J, I2 = np.meshgrid(g[2], g[1], e[0], indexing="ij")
    ex = (I2, J, K)
    I, K, J2 = np.meshgrid(g[0], g[2], e[1], indexing="ij")
    ey = (I, J2, K)
    J, I, K2 = np.meshgrid(g[1], g[0], e[2], indexing="ij")
    ez = (I, J, K2)
    return [corner, ex, ey, ez]


def cubic_field_from(fn_sdf, fn_vol, res=(20, 20, 20), lo=(0.0, 0.0, 0.0), hi=(1.0, 1.0, 1.0)) -> Field:
    """Sample analytic functions at the 32-node element's nodes (.cdm order)."""
    sdf, vol = [], []
    for (x, y, z) in cubic_nodes(res):
        X = lo[0] + (hi[0] - lo[0]) * x
        Y = lo[1] + (hi[1] - lo[1]) * y
        Z = lo[2] + (hi[2] - lo[2]) * z
        sdf.append(np.asarray(fn_sdf(X, Y, Z), np.float64).ravel())
        vol.append(np.asarray(fn_vol(X, Y, Z), np.float64).ravel())
    return Field("cubic", tuple(res), tuple(lo), tuple(hi), sdf=np.concatenate(sdf).astype(np.float32),
                 vol=np.concatenate(vol).astype(np.float32))


def unit_box_cdm_field(res=(20, 20, 20), h=REF_H) -> Field:
    """Unit-box walls in the .cdm layout, built by pbf_build_cubic_field with
    the reference's own volume definition (PBF_VOLUME_REFERENCE): the SDF is
    the distance to the nearest face and the volume reproduces
    public/model/domain_boundary.cdm's (tests/test_volume_definition.py)."""
    sdf, vol = build_cubic_field(VOLUME_REFERENCE, res, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), h)
    return Field("cubic", tuple(res), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), sdf=sdf, vol=vol)


def cdm_text(fld: Field) -> str:
    """Serialise a cubic field in the .cdm text layout (volumeMap.ts:23-39)."""
    head = f"{fld.res[0]} {fld.res[1]} {fld.res[2]}"
    s = " ".join(repr(float(v)) for v in fld.sdf) + " "
    v = " ".join(repr(float(v)) for v in fld.vol) + " "
    return f"{head}\r\n{fld.sdf.size}\r\n{s}\r\n{fld.vol.size}\r\n{v}\r\n"


# --- running a scene ----------------------------------------------------------
def make_engine(scene: Scene, capacity: Optional[int] = None, flags: int = 0, max_neighbors: int = 96,
                cell_slabs: int = 0) -> PBF:
    sim = PBF(particle_radius=scene.particle_radius, kernel_radius=scene.kernel_radius,
              capacity=capacity or max(scene.n, 1), domain_lo=scene.domain_lo, domain_hi=scene.domain_hi,
              grid_mode=scene.grid_mode, hash_log2=scene.hash_log2, flags=flags, max_neighbors=max_neighbors,
              time_step=scene.substep_dt(), cell_slabs=cell_slabs)
    sim.setConfig(scene.config)
    load_field(sim, scene.field)
    sim.setParticlePosition(scene.positions)
    return sim


def load_field(sim: PBF, fld: Optional[Field], scene_index: int = 0):
    if fld is None:
        return
    if fld.kind == "trilinear":
        sim.setVolumeMap(fld.res, fld.lo, fld.hi, fld.sdf_vol, scene_index)
    else:
        sim.setCubicField(fld.res, fld.lo, fld.hi, fld.sdf, fld.vol, scene_index)


def make_batch_engine(scenes: List[Scene], flags: int = 0, max_neighbors: int = 96, cell_slabs: int = 0) -> PBF:
    s0 = scenes[0]
    total = sum(s.n for s in scenes)
    sim = PBF(particle_radius=s0.particle_radius, kernel_radius=s0.kernel_radius, capacity=total, flags=flags,
              max_neighbors=max_neighbors, time_step=s0.substep_dt(), grid_mode=s0.grid_mode,
              hash_log2=s0.hash_log2, cell_slabs=cell_slabs,
              scenes=[{"domain_lo": s.domain_lo, "domain_hi": s.domain_hi, "count": s.n} for s in scenes])
    sim.setConfig(s0.config)
    for k, s in enumerate(scenes):
        load_field(sim, s.field, k)
    sim.setParticlePosition(np.concatenate([s.positions for s in scenes]))
    return sim
