"""Python host mirror of the reference simulator API over libpbf's C ABI.

The class ``PBF`` keeps the member names of the reference's ``SPH`` /
``PBFConfig`` / ``PBF`` classes (LinzhouLi/WebGPU-Fluid-Simulation,
src/simulator/SPH.ts, src/simulator/PBF/PBFConfig.ts, src/simulator/PBF/PBF.ts)
and raises ``PBFError`` with the reference's messages where it has them
('To Many Particles!', 'Illegal length of Particle Position Buffer!',
'Invalid Boundary Model File!').  All physics runs in libpbf.so (HIP, gfx950);
there is no CPU fallback: constructing a context without a GPU fails.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libpbf.so")

PBF_OK = 0
PBF_E_INVALID, PBF_E_CAPACITY, PBF_E_LENGTH, PBF_E_FORMAT, PBF_E_DEVICE, PBF_E_OVERFLOW, PBF_E_STATE = range(-1, -8, -1)
PBF_GRID_DENSE, PBF_GRID_HASH = 0, 1
PBF_FLAG_NO_GRAPH, PBF_FLAG_CUBIC_FIXED, PBF_FLAG_TIMING = 1, 2, 4
FIELDS = {"position": 0, "velocity": 1, "accel": 2, "lambda": 3, "omega": 4, "normal": 5,
          "boundary": 6, "predicted": 7, "lambda_sum": 8}
# the state a substep carries into the next (positions aside)
CHECKPOINT_FIELDS = ("velocity", "accel", "lambda", "lambda_sum")


class PBFError(RuntimeError):
    """Error raised for a non-zero status of the C ABI (message = pbf_last_error)."""

    def __init__(self, code: int, message: str):
        super().__init__(message)
        self.code = code


class Params(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("max_neighbors", C.c_uint32),
                ("particle_radius", C.c_double), ("kernel_radius", C.c_double),
                ("rest_density", C.c_double), ("time_step", C.c_double),
                ("domain_lo", C.c_double * 3), ("domain_hi", C.c_double * 3),
                ("grid_mode", C.c_uint32), ("hash_log2", C.c_uint32), ("flags", C.c_uint32),
                ("cell_slabs", C.c_uint32), ("reserved", C.c_uint32 * 4)]


class MpmParams(C.Structure):
    _fields_ = [("capacity", C.c_uint32), ("n_grid", C.c_uint32), ("bound", C.c_uint32), ("flags", C.c_uint32),
                ("dt", C.c_double), ("rho", C.c_double), ("E", C.c_double), ("gravity", C.c_double),
                ("reserved", C.c_uint32 * 4)]


MPM_POSITION, MPM_VELOCITY, MPM_AFFINE, MPM_GRID = range(4)


class Config(C.Structure):
    _fields_ = [("iteration", C.c_uint32), ("xsph", C.c_float), ("vorticity", C.c_float),
                ("surface_tension", C.c_float), ("gravity", C.c_float)]


class Tuning(C.Structure):
    _fields_ = [("walk_block", C.c_uint32 * 8), ("list_variant", C.c_uint32), ("reserved", C.c_uint32 * 7)]


class SceneDesc(C.Structure):
    _fields_ = [("domain_lo", C.c_double * 3), ("domain_hi", C.c_double * 3),
                ("particle_count", C.c_uint32), ("reserved", C.c_uint32 * 3)]


_lib: Optional[C.CDLL] = None


def load_library(path: str = LIB_PATH) -> C.CDLL:
    """Load libpbf.so (raises OSError when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"libpbf.so not built at {path}; run __graft_entry__.build()")
    lib = C.CDLL(path)
    P, I, U32, U64, D, F, S = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_double, C.c_float, C.c_size_t
    pd3 = C.POINTER(C.c_double)
    pf = C.POINTER(C.c_float)
    pu = C.POINTER(C.c_uint32)
    sig = {
        "pbf_create": (I, [C.POINTER(Params), C.POINTER(P)]),
        "pbf_batch_create": (I, [C.POINTER(Params), U32, C.POINTER(SceneDesc), C.POINTER(P)]),
        "pbf_destroy": (I, [P]), "pbf_reset": (I, [P]),
        "pbf_set_config": (I, [P, C.POINTER(Config)]),
        "pbf_voxelize_cube": (I, [P, pd3, pd3, D, U64, pu]),
        "pbf_voxelize_sphere": (I, [P, pd3, D, D, U64, pu]),
        "pbf_voxelize_cube_host": (I, [D, pd3, pd3, D, U64, U64, pf, U64, pu]),
        "pbf_voxelize_sphere_host": (I, [D, pd3, D, D, U64, U64, pf, U64, pu]),
        "pbf_voxelize_mesh": (I, [P, pf, U32, pu, U32, C.POINTER(C.c_double), D, U64, pu]),
        "pbf_voxelize_mesh_host": (I, [D, pf, U32, pu, U32, C.POINTER(C.c_double), D, U64, U64, pf, U64, pu]),
        "pbf_staged_count": (U32, [P]),
        "pbf_set_particle_position": (I, [P]),
        "pbf_set_positions": (I, [P, pf, S, U32]),
        "pbf_set_boundary_data": (I, [P, C.c_char_p, S]),
        "pbf_parse_cdm": (I, [C.c_char_p, S, C.POINTER(C.c_int32), pf, pf, S, C.POINTER(S)]),
        "pbf_set_cubic_field": (I, [P, U32, C.POINTER(C.c_int32), pd3, pd3, pf, pf, S]),
        "pbf_set_volume_map": (I, [P, U32, C.POINTER(C.c_int32), pd3, pd3, pf]),
        "pbf_clear_boundary": (I, [P]),
        "pbf_build_volume_map": (I, [U32, pd3, C.POINTER(C.c_int32), pd3, pd3, D, D, pf]),
        "pbf_build_volume_map_gpu": (I, [U32, pd3, C.POINTER(C.c_int32), pd3, pd3, D, D, pf]),
        "pbf_build_cubic_field": (I, [U32, pd3, C.POINTER(C.c_int32), pd3, pd3, D, D, pf, pf, S, C.POINTER(S)]),
        "pbf_write_cdm": (I, [C.POINTER(C.c_int32), pf, pf, S, C.c_char_p, S, C.POINTER(S)]),
        "pbf_step": (I, [P, F, U32, U32]), "pbf_run": (I, [P]), "pbf_sync": (I, [P]),
        "pbf_read_positions": (I, [P, pf, U32]), "pbf_read_velocities": (I, [P, pf, U32]),
        "pbf_read_field": (I, [P, U32, pf, U32]),
        "pbf_write_field": (I, [P, U32, pf, U32]),
        "pbf_num_cells": (U32, [P]),
        "pbf_hash_buckets_distinct": (I, [C.POINTER(C.c_int32), U32]),
        "pbf_read_cells": (I, [P, pu, pu]),
        "pbf_read_neighbors": (I, [P, pu, pu, U64, C.POINTER(U64)]),
        "pbf_search": (I, [P, F]),
        "pbf_density_error": (I, [P, U32, C.POINTER(D), C.POINTER(D)]),
        "pbf_record_density_error": (I, [P, U32]),
        "pbf_read_density_history": (I, [P, C.POINTER(D), U32, pu]),
        "pbf_particle_count": (U32, [P]), "pbf_scene_count": (U32, [P]),
        "pbf_last_error": (C.c_char_p, [P]), "pbf_version": (C.c_char_p, []),
        "pbf_pass_times": (I, [P, C.POINTER(D), pu, U32, pu, C.POINTER(C.c_char_p)]),
        "pbf_copy_bandwidth": (I, [S, U32, C.POINTER(D)]),
        "pbf_check_device_sqrt": (I, [C.POINTER(U64), C.POINTER(U64)]),
        "pbf_mlp_blob_size": (U32, []),
        "pbf_mlp_init": (I, [U64, pf]),
        "pbf_mlp_forward_backward": (I, [pf, pf, pf, U32, pf, pf, pf]),
        "pbf_mlp_train": (I, [pf, pf, U64, U32, U32, F, U64, pf, pf]),
        "pbf_mlp_train_weighted": (I, [pf, pf, U64, U32, U32, F, U64, F, pf, pf]),
        "pbf_mlp_forward_backward_noalloc": (I, [P, P, P, U32, P, P, P, P]),
        "pbf_set_warm_start": (I, [P, pf, S, U32, F, D, U32]),
        "pbf_warm_start_stats": (I, [P, C.POINTER(U64), C.POINTER(U64)]),
        "pbf_collect_training": (I, [P, F, U32, U32, pf, pf, U32, pu]),
        "pbf_get_tuning": (I, [P, C.POINTER(Tuning)]),
        "pbf_set_tuning": (I, [P, C.POINTER(Tuning)]),
        "pbf_autotune": (I, [P, F, U32, U32, C.POINTER(Tuning)]),
        "pbf_mpm_create": (I, [C.POINTER(MpmParams), C.POINTER(P)]), "pbf_mpm_destroy": (I, [P]),
        "pbf_mpm_set_positions": (I, [P, pf, U32]), "pbf_mpm_step": (I, [P, U32]), "pbf_mpm_sync": (I, [P]),
        "pbf_mpm_read": (I, [P, I, pf, S]), "pbf_mpm_write": (I, [P, I, pf, S]), "pbf_mpm_g2p": (I, [P]),
        "pbf_mpm_particle_count": (U32, [P]), "pbf_mpm_last_error": (C.c_char_p, [P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _d3(v: Sequence[float]):
    return (C.c_double * 3)(*[float(x) for x in v])


def _i3(v: Sequence[int]):
    return (C.c_int32 * 3)(*[int(x) for x in v])


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _uptr(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def hash_buckets_distinct(dims, hash_log2: int) -> bool:
    """True when the 27 cells around every cell of a dims grid hashed into
    2^hash_log2 buckets land in 27 different buckets (GPU-free)."""
    return bool(load_library().pbf_hash_buckets_distinct(_i3(dims), int(hash_log2)))


_map_on_gpu = None


def build_volume_map(kind: int, res, lo, hi, kernel_radius: float, shape=None,
                     particle_radius: float = 0.0, where: str = "auto") -> np.ndarray:
    """Analytic volume map (res+1)^3 x (sdf, vol); kind 0 box, 1 box + sphere + torus.
    where: "gpu", "host", or "auto" (the GPU when one answers, else the host;
    both give the same map)."""
    global _map_on_gpu
    lib = load_library()
    n = (res[0] + 1) * (res[1] + 1) * (res[2] + 1)
    out = np.zeros((n, 2), np.float32)
    sh = None
    if shape is not None:
        sh = (C.c_double * len(shape))(*[float(x) for x in shape])
    args = (kind, sh, _i3(res), _d3(lo), _d3(hi), float(kernel_radius), float(particle_radius), _fptr(out))
    if where in ("gpu", "auto") and _map_on_gpu is not False:
        rc = lib.pbf_build_volume_map_gpu(*args)
        if rc == PBF_OK:
            _map_on_gpu = True
            return out
        if where == "gpu" or rc != PBF_E_DEVICE:
            raise PBFError(rc, "pbf_build_volume_map_gpu failed")
        _map_on_gpu = False
    rc = lib.pbf_build_volume_map(*args)
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_build_volume_map failed")
    return out


def build_cubic_field(kind: int, res, lo, hi, kernel_radius: float, shape=None, particle_radius: float = 0.0):
    """(sdf, vol) at the 32-node element's nodes in the .cdm layout."""
    lib = load_library()
    sh = None if shape is None else (C.c_double * len(shape))(*[float(x) for x in shape])
    fl = C.c_size_t()
    rc = lib.pbf_build_cubic_field(kind, sh, _i3(res), _d3(lo), _d3(hi), float(kernel_radius), float(particle_radius),
                                   None, None, 0, C.byref(fl))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_build_cubic_field failed")
    sdf = np.zeros(fl.value, np.float32)
    vol = np.zeros(fl.value, np.float32)
    rc = lib.pbf_build_cubic_field(kind, sh, _i3(res), _d3(lo), _d3(hi), float(kernel_radius), float(particle_radius),
                                   _fptr(sdf), _fptr(vol), fl.value, C.byref(fl))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_build_cubic_field failed")
    return sdf, vol


def write_cdm(res, sdf, vol) -> str:
    """.cdm text of a cubic field."""
    lib = load_library()
    s = np.ascontiguousarray(sdf, np.float32)
    v = np.ascontiguousarray(vol, np.float32)
    n = C.c_size_t()
    rc = lib.pbf_write_cdm(_i3(res), _fptr(s), _fptr(v), s.size, None, 0, C.byref(n))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_write_cdm failed")
    buf = C.create_string_buffer(n.value + 1)
    rc = lib.pbf_write_cdm(_i3(res), _fptr(s), _fptr(v), s.size, buf, n.value + 1, C.byref(n))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_write_cdm failed")
    return buf.raw[: n.value].decode()


def parse_cdm(text) -> tuple:
    """Parse .cdm text without a GPU: (res, sdf, vol); raises PBFError on a bad file."""
    lib = load_library()
    b = text.encode() if isinstance(text, str) else bytes(text)
    res = (C.c_int32 * 3)()
    fl = C.c_size_t()
    rc = lib.pbf_parse_cdm(b, len(b), res, None, None, 0, C.byref(fl))
    if rc != PBF_OK:
        raise PBFError(rc, "Invalid Boundary Model File!")
    sdf = np.zeros(fl.value, np.float32)
    vol = np.zeros(fl.value, np.float32)
    rc = lib.pbf_parse_cdm(b, len(b), res, _fptr(sdf), _fptr(vol), fl.value, C.byref(fl))
    if rc != PBF_OK:
        raise PBFError(rc, "Invalid Boundary Model File!")
    return tuple(res), sdf, vol


def voxelize_mesh(particle_radius: float, vertices, indices, world=None, jitter: float = 0.0,
                  seed: int = 0) -> np.ndarray:
    """GPU-free SPH.voxelizeMesh: float4 rows (w = 0)."""
    lib = load_library()
    v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
    t = np.ascontiguousarray(indices, np.uint32).reshape(-1)
    w = None if world is None else (C.c_double * 16)(*[float(x) for x in np.asarray(world).reshape(-1)])
    cnt = C.c_uint32()
    rc = lib.pbf_voxelize_mesh_host(float(particle_radius), _fptr(v), v.size // 3, _uptr(t), t.size // 3, w,
                                    float(jitter), int(seed), 0, None, 0, C.byref(cnt))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_voxelize_mesh_host failed")
    out = np.zeros((max(cnt.value, 1), 4), np.float32)
    rc = lib.pbf_voxelize_mesh_host(float(particle_radius), _fptr(v), v.size // 3, _uptr(t), t.size // 3, w,
                                    float(jitter), int(seed), 0, _fptr(out), cnt.value, C.byref(cnt))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_voxelize_mesh_host failed")
    return out[: cnt.value]


def copy_bandwidth(nbytes: int = 1 << 30, iterations: int = 20) -> float:
    """Device-to-device streaming copy rate (read + write GB/s)."""
    lib = load_library()
    g = C.c_double()
    rc = lib.pbf_copy_bandwidth(int(nbytes), int(iterations), C.byref(g))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_copy_bandwidth failed")
    return g.value


def check_device_sqrt() -> tuple:
    """(failures, checked): the device square root of the pair terms against
    the correctly rounded value over every float in [2^-100, FLT_MAX]."""
    lib = load_library()
    bad, n = C.c_uint64(), C.c_uint64()
    rc = lib.pbf_check_device_sqrt(C.byref(bad), C.byref(n))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_check_device_sqrt failed")
    return bad.value, n.value


MLP_PARAMS = 5313
MLP_FEATURES = 16
WARM_PLAIN, WARM_LEARNED, WARM_PREVIOUS, WARM_RATIO = 0, 1, 2, 3
WARM_PROJECT = 0x100  # OR-ed into a mode: projected accumulation of lambda
WARM_SMOOTH = 0x200   # OR-ed into a mode: the warm lambda field smoothed over neighbours
VOLUME_REFERENCE = 0x100  # OR-ed into a map-builder kind: the reference's own boundary volume
WARM_RATIO_MAX = 2.0  # clamp of the ratio-mode factor, in units of relax


def mlp_blob_size() -> int:
    return int(load_library().pbf_mlp_blob_size())


def mlp_init(seed: int = 0) -> np.ndarray:
    """Fresh network blob (Glorot-uniform weights, identity normalisation)."""
    lib = load_library()
    blob = np.zeros(lib.pbf_mlp_blob_size(), np.float32)
    rc = lib.pbf_mlp_init(int(seed), _fptr(blob))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_mlp_init failed")
    return blob


def mlp_forward_backward(blob: np.ndarray, features: np.ndarray, targets: np.ndarray):
    """GPU forward + backward: (y normalised, gradient[MLP_PARAMS], loss)."""
    lib = load_library()
    blob = np.ascontiguousarray(blob, np.float32)
    f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1)
    if f.shape[0] != t.shape[0]:
        raise ValueError("features and targets disagree in length")
    n = f.shape[0]
    y = np.zeros(n, np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = np.zeros(1, np.float32)
    rc = lib.pbf_mlp_forward_backward(_fptr(blob), _fptr(f), _fptr(t), n, _fptr(y), _fptr(g), _fptr(loss))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_mlp_forward_backward failed")
    return y, g, float(loss[0])


def mlp_train(features: np.ndarray, targets: np.ndarray, blob: np.ndarray, epochs: int = 10, batch: int = 4096,
              lr: float = 1e-3, seed: int = 0, over_weight: float = 1.0):
    """Adam on the GPU; returns (trained blob, mean loss per epoch).
    over_weight > 1 weighs over-predictions more (pbf_mlp_train_weighted)."""
    lib = load_library()
    blob = np.array(blob, np.float32, copy=True)
    f = np.ascontiguousarray(features, np.float32).reshape(-1, MLP_FEATURES)
    t = np.ascontiguousarray(targets, np.float32).reshape(-1)
    losses = np.zeros(max(int(epochs), 1), np.float32)
    rc = lib.pbf_mlp_train_weighted(_fptr(f), _fptr(t), f.shape[0], int(epochs), int(batch), float(lr), int(seed),
                                    float(over_weight), _fptr(blob), _fptr(losses))
    if rc != PBF_OK:
        raise PBFError(rc, "pbf_mlp_train failed")
    return blob, losses[: int(epochs)]


class PBF:
    """Mirror of the reference ``PBF`` simulator (PBF.ts:27-575) on libpbf."""

    def __init__(self, particle_radius: float = 0.006, kernel_radius: float = 0.025, capacity: int = 1 << 17,
                 rest_density: float = 1000.0, time_step: float = 1.0 / 300.0, domain_lo=(0.0, 0.0, 0.0),
                 domain_hi=(1.0, 1.0, 1.0), max_neighbors: int = 96, grid_mode: int = PBF_GRID_DENSE,
                 hash_log2: int = 22, flags: int = 0, scenes: Optional[Sequence[dict]] = None, cell_slabs: int = 0):
        self.lib = load_library()
        p = Params()
        p.capacity = capacity
        p.max_neighbors = max_neighbors
        p.particle_radius = particle_radius
        p.kernel_radius = kernel_radius
        p.rest_density = rest_density
        p.time_step = time_step
        p.domain_lo = _d3(domain_lo)
        p.domain_hi = _d3(domain_hi)
        p.grid_mode = grid_mode
        p.hash_log2 = hash_log2
        p.flags = flags
        p.cell_slabs = cell_slabs  # z-slabs ordering a cell's particles: 1, 2 or 4 (0: 4)
        self.params = p
        self.ctx = C.c_void_p()
        if scenes:
            arr = (SceneDesc * len(scenes))()
            for k, s in enumerate(scenes):
                arr[k].domain_lo = _d3(s["domain_lo"])
                arr[k].domain_hi = _d3(s["domain_hi"])
                arr[k].particle_count = int(s["count"])
            rc = self.lib.pbf_batch_create(C.byref(p), len(scenes), arr, C.byref(self.ctx))
        else:
            rc = self.lib.pbf_create(C.byref(p), C.byref(self.ctx))
        if rc != PBF_OK:
            raise PBFError(rc, self.lib.pbf_last_error(None).decode())
        self.pause = True
        self.stepCount = 1
        self.particlePositionArray: list = []

    # -- lifetime -----------------------------------------------------------
    def close(self):
        if self.ctx:
            self.lib.pbf_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int):
        if rc != PBF_OK:
            raise PBFError(rc, self.lib.pbf_last_error(self.ctx).decode())

    # -- reference API ------------------------------------------------------
    def reset(self):
        """PBF.reset (PBF.ts:71-89)."""
        self.particlePositionArray = []
        self._check(self.lib.pbf_reset(self.ctx))

    @property
    def particleCount(self) -> int:
        return int(self.lib.pbf_particle_count(self.ctx)) or int(self.lib.pbf_staged_count(self.ctx))

    def voxelizeCube(self, mn, mx, jitter: float = 0.0, seed: int = 0) -> int:
        """SPH.voxelizeCube (SPH.ts:99-121)."""
        added = C.c_uint32()
        self._check(self.lib.pbf_voxelize_cube(self.ctx, _d3(mn), _d3(mx), float(jitter), int(seed), C.byref(added)))
        return int(added.value)

    def voxelizeSphere(self, center, radius: float, jitter: float = 0.0, seed: int = 0) -> int:
        """SPH.voxelizeSphere (SPH.ts:123-149)."""
        added = C.c_uint32()
        self._check(self.lib.pbf_voxelize_sphere(self.ctx, _d3(center), float(radius), float(jitter), int(seed),
                                                 C.byref(added)))
        return int(added.value)


    def voxelizeMesh(self, vertices, indices, world=None, jitter: float = 0.0, seed: int = 0) -> int:
        """SPH.voxelizeMesh: vertices (n, 3) mesh-local, indices (m, 3), world 4x4
        column-major (three.js Matrix4.elements) or None."""
        v = np.ascontiguousarray(vertices, np.float32).reshape(-1)
        t = np.ascontiguousarray(indices, np.uint32).reshape(-1)
        w = None if world is None else (C.c_double * 16)(*[float(x) for x in np.asarray(world).reshape(-1)])
        added = C.c_uint32()
        self._check(self.lib.pbf_voxelize_mesh(self.ctx, _fptr(v), v.size // 3, _uptr(t), t.size // 3, w,
                                               float(jitter), int(seed), C.byref(added)))
        return added.value
    def setParticlePosition(self, positions: Optional[np.ndarray] = None, count: Optional[int] = None):
        """PBFConfig.setParticlePosition (PBFConfig.ts:48-62); uploads staged or given float4 rows."""
        if positions is None:
            self._check(self.lib.pbf_set_particle_position(self.ctx))
            return
        a = np.ascontiguousarray(np.asarray(positions, np.float32).reshape(-1))
        n = a.size // 4 if count is None else int(count)
        self._check(self.lib.pbf_set_positions(self.ctx, _fptr(a), a.size, n))

    def setBoundaryData(self, data: str):
        """PBF.setBoundaryData (PBF.ts:91-94): .cdm text."""
        b = data.encode() if isinstance(data, str) else bytes(data)
        self._check(self.lib.pbf_set_boundary_data(self.ctx, b, len(b)))

    def setCubicField(self, res, lo, hi, sdf: np.ndarray, vol: np.ndarray, scene: int = 0):
        s = np.ascontiguousarray(sdf, np.float32)
        v = np.ascontiguousarray(vol, np.float32)
        self._check(self.lib.pbf_set_cubic_field(self.ctx, scene, _i3(res), _d3(lo), _d3(hi), _fptr(s), _fptr(v),
                                                 s.size))

    def setVolumeMap(self, res, lo, hi, sdf_vol: np.ndarray, scene: int = 0):
        a = np.ascontiguousarray(sdf_vol, np.float32)
        self._check(self.lib.pbf_set_volume_map(self.ctx, scene, _i3(res), _d3(lo), _d3(hi), _fptr(a)))

    def setConfig(self, conf: dict):
        """PBFConfig.setConfig (PBFConfig.ts:31-46)."""
        c = Config()
        c.iteration = int(conf.get("iteration", 5))
        c.xsph = float(conf.get("XSPH", 0.01))
        c.vorticity = float(conf.get("vorticity", 0.1))
        c.surface_tension = float(conf.get("surfaceTension", 0.5))
        c.gravity = float(conf.get("gravity", 9.8))
        self._check(self.lib.pbf_set_config(self.ctx, C.byref(c)))

    def start(self):
        self.pause = False

    def stop(self):
        self.pause = True

    def switch(self):
        self.pause = not self.pause

    def run(self):
        """PBF.run (PBF.ts:398-448): one substep of time_step; no-op when paused."""
        if self.pause:
            return
        self._check(self.lib.pbf_run(self.ctx))

    # -- extended API: diagnostics, warm starts, batching, tuning ----------
    def step(self, dt: float, substeps: int = 1, iterations: int = 0):
        self._check(self.lib.pbf_step(self.ctx, float(dt), int(substeps), int(iterations)))

    def sync(self):
        self._check(self.lib.pbf_sync(self.ctx))

    def search(self, dt: float):
        self._check(self.lib.pbf_search(self.ctx, float(dt)))

    def read_positions(self) -> np.ndarray:
        n = int(self.lib.pbf_particle_count(self.ctx))
        out = np.zeros((n, 4), np.float32)
        self._check(self.lib.pbf_read_positions(self.ctx, _fptr(out), n))
        return out

    def read_velocities(self) -> np.ndarray:
        n = int(self.lib.pbf_particle_count(self.ctx))
        out = np.zeros((n, 4), np.float32)
        self._check(self.lib.pbf_read_velocities(self.ctx, _fptr(out), n))
        return out

    def read_field(self, name: str) -> np.ndarray:
        n = int(self.lib.pbf_particle_count(self.ctx))
        out = np.zeros((n, 4), np.float32)
        self._check(self.lib.pbf_read_field(self.ctx, FIELDS[name], _fptr(out), n))
        return out

    def write_field(self, name: str, data: np.ndarray):
        """Checkpoint upload (pbf_write_field): velocity / accel rows (x, y, z),
        lambda / lambda_sum values, in the original particle order."""
        n = int(self.lib.pbf_particle_count(self.ctx))
        a = np.asarray(data, np.float32)
        buf = np.zeros((n, 4), np.float32)
        if name in ("lambda", "lambda_sum"):
            buf[:, 0] = a[:, 0] if a.ndim == 2 else a.reshape(n)
        else:
            buf[:, :3] = a.reshape(n, -1)[:, :3]
        self._check(self.lib.pbf_write_field(self.ctx, FIELDS[name], _fptr(buf), n))

    def checkpoint(self) -> dict:
        """Everything a later context needs to resume this run bit for bit:
        positions plus the carried per-particle state, in the original order."""
        snap = {"positions": self.read_positions()}
        for name in CHECKPOINT_FIELDS:
            snap[name] = self.read_field(name)
        return snap

    def resume(self, snap: dict):
        """Load a checkpoint() taken from a context with the same configuration."""
        self.setParticlePosition(snap["positions"])
        for name in CHECKPOINT_FIELDS:
            self.write_field(name, snap[name])

    def read_cells(self):
        nc = int(self.lib.pbf_num_cells(self.ctx))
        counts = np.zeros(nc, np.uint32)
        offsets = np.zeros(nc + 1, np.uint32)
        self._check(self.lib.pbf_read_cells(self.ctx, _uptr(counts), _uptr(offsets)))
        return counts, offsets

    def read_neighbors(self):
        n = int(self.lib.pbf_particle_count(self.ctx))
        off = np.zeros(n + 1, np.uint32)
        pairs = C.c_uint64()
        self._check(self.lib.pbf_read_neighbors(self.ctx, _uptr(off), None, 0, C.byref(pairs)))
        ids = np.zeros(max(int(pairs.value), 1), np.uint32)
        self._check(self.lib.pbf_read_neighbors(self.ctx, _uptr(off), _uptr(ids), ids.size, C.byref(pairs)))
        return off, ids[: int(pairs.value)]

    def density_error(self, scene: int = 0xFFFFFFFF):
        m, x = C.c_double(), C.c_double()
        self._check(self.lib.pbf_density_error(self.ctx, scene, C.byref(m), C.byref(x)))
        return m.value, x.value

    def record_density_error(self, enable: bool = True):
        self._check(self.lib.pbf_record_density_error(self.ctx, 1 if enable else 0))

    def density_history(self, cap: int = 4096) -> np.ndarray:
        out = np.zeros((cap, 2), np.float64)
        cnt = C.c_uint32()
        self._check(self.lib.pbf_read_density_history(self.ctx, out.ctypes.data_as(C.POINTER(C.c_double)), cap,
                                                      C.byref(cnt)))
        return out[: cnt.value]

    def pass_times(self) -> dict:
        ms = (C.c_double * 32)()
        cnt = (C.c_uint32 * 32)()
        k = C.c_uint32()
        names = C.c_char_p()
        self._check(self.lib.pbf_pass_times(self.ctx, ms, cnt, 32, C.byref(k), C.byref(names)))
        nm = names.value.decode().split("\n")
        return {nm[q]: (ms[q], cnt[q]) for q in range(min(k.value, len(nm)))}

    # -- learned warm start (no reference counterpart) ----------------------
    def set_warm_start(self, blob: Optional[np.ndarray] = None, mode: int = WARM_LEARNED, relax: float = 1.0,
                       fallback_threshold: float = 0.0, fallback_iterations: int = 0):
        ptr, length = None, 0
        if blob is not None:
            blob = np.ascontiguousarray(blob, np.float32)
            ptr, length = _fptr(blob), blob.size
        self._check(self.lib.pbf_set_warm_start(self.ctx, ptr, length, int(mode), float(relax),
                                                float(fallback_threshold), int(fallback_iterations)))

    def warm_start_stats(self):
        a, b = C.c_uint64(), C.c_uint64()
        self._check(self.lib.pbf_warm_start_stats(self.ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def collect_training(self, dt: float, long_iterations: int, iterations: int = 0):
        """Training pair of the current state (features (n, 16), lambda sums of a
        `long_iterations` solve (n,)), then one substep of the selected solver."""
        n = int(self.lib.pbf_particle_count(self.ctx))
        f = np.zeros((max(n, 1), MLP_FEATURES), np.float32)
        t = np.zeros(max(n, 1), np.float32)
        cnt = C.c_uint32()
        self._check(self.lib.pbf_collect_training(self.ctx, float(dt), int(long_iterations), int(iterations),
                                                  _fptr(f), _fptr(t), max(n, 1), C.byref(cnt)))
        return f[: cnt.value], t[: cnt.value]

    # -- launch tuning (no reference counterpart) ---------------------------
    WALKS = ("lambda", "dp", "density", "vorticity", "xsph", "diag")

    def get_tuning(self) -> dict:
        t = Tuning()
        self._check(self.lib.pbf_get_tuning(self.ctx, C.byref(t)))
        return {"walk_block": dict(zip(self.WALKS, list(t.walk_block)[:6])), "list_variant": int(t.list_variant)}

    def set_tuning(self, walk_block=None, list_variant: int = 0):
        t = Tuning()
        wb = walk_block if walk_block is not None else {}
        for k, name in enumerate(self.WALKS):
            t.walk_block[k] = int(wb.get(name, 256)) if isinstance(wb, dict) else int(wb)
        t.walk_block[6] = t.walk_block[7] = 256
        t.list_variant = int(list_variant)
        self._check(self.lib.pbf_set_tuning(self.ctx, C.byref(t)))

    def autotune(self, dt: float, iterations: int = 0, repeats: int = 3) -> dict:
        t = Tuning()
        self._check(self.lib.pbf_autotune(self.ctx, float(dt), int(iterations), int(repeats), C.byref(t)))
        return {"walk_block": dict(zip(self.WALKS, list(t.walk_block)[:6])), "list_variant": int(t.list_variant)}


class MPM:
    """Mirror of the reference's MLS-MPM simulator (src/simulator/MPM/MPM.ts,
    commented out in the reference) on libpbf: the constructor's parameters
    (MPM.ts:60-83), initResource (MPM.ts:94-121: particles uniform in
    [0.15, 0.55]^3, v = 0, C = 0, J = 1) and run (MPM.ts:204-238: one P2G,
    grid and G2P pass)."""

    def __init__(self, n_particle: int = 9000, n_grid: int = 32, dt: float = 8e-4, bound: int = 3,
                 rho: float = 1.0, E: float = 400.0, gravity: float = 9.8, flags: int = 0,
                 capacity: Optional[int] = None):
        self.lib = load_library()
        p = MpmParams()
        p.capacity = capacity or n_particle
        p.n_grid, p.bound, p.flags = n_grid, bound, flags
        p.dt, p.rho, p.E, p.gravity = dt, rho, E, gravity
        self.params = p
        self.particleCount = n_particle
        self.n_grid = n_grid
        self.m = C.c_void_p()
        rc = self.lib.pbf_mpm_create(C.byref(p), C.byref(self.m))
        if rc != PBF_OK:
            raise PBFError(rc, self.lib.pbf_mpm_last_error(None).decode())

    def _check(self, rc):
        if rc != PBF_OK:
            raise PBFError(rc, self.lib.pbf_mpm_last_error(self.m).decode())

    def initResource(self, seed: int = 0):
        """Random particle positions in [0.15, 0.55]^3 (MPM.ts:104-112, seeded)."""
        rng = np.random.default_rng(seed)
        x = np.zeros((self.particleCount, 4), np.float32)
        x[:, :3] = (rng.random((self.particleCount, 3)) * 0.4 + 0.15).astype(np.float32)
        self.setParticlePosition(x)
        return x

    def setParticlePosition(self, xyzw):
        a = np.ascontiguousarray(np.asarray(xyzw, np.float32).reshape(-1, 4))
        self._check(self.lib.pbf_mpm_set_positions(self.m, _fptr(a), a.shape[0]))
        self.particleCount = a.shape[0]

    def run(self, steps: int = 1):
        self._check(self.lib.pbf_mpm_step(self.m, int(steps)))

    def sync(self):
        self._check(self.lib.pbf_mpm_sync(self.m))

    def g2p(self):
        self._check(self.lib.pbf_mpm_g2p(self.m))

    def _shape(self, field):
        n = self.particleCount
        return {MPM_POSITION: (n, 4), MPM_VELOCITY: (n, 4), MPM_AFFINE: (n, 9),
                MPM_GRID: (self.n_grid ** 3, 4)}[field]

    def read(self, field: int) -> np.ndarray:
        out = np.zeros(self._shape(field), np.float32)
        self._check(self.lib.pbf_mpm_read(self.m, field, _fptr(out), out.shape[0]))
        return out

    def write(self, field: int, arr):
        a = np.ascontiguousarray(np.asarray(arr, np.float32).reshape(self._shape(field)))
        self._check(self.lib.pbf_mpm_write(self.m, field, _fptr(a), a.shape[0]))

    def close(self):
        if self.m:
            self.lib.pbf_mpm_destroy(self.m)
            self.m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
