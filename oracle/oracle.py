"""ctypes wrapper of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (to check
the GPU result) and bench.py's cpu_baseline leg.  The shipped library never loads it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

POSITION, VELOCITY, ACCEL, LAMBDA, OMEGA, NORMAL, BOUNDARY, PREDICTED = range(8)


class OrcParams(C.Structure):
    _fields_ = [("particle_radius", C.c_double), ("kernel_radius", C.c_double), ("rest_density", C.c_double),
                ("domain_lo", C.c_double * 3), ("domain_hi", C.c_double * 3), ("grid_mode", C.c_int),
                ("hash_log2", C.c_int), ("cell_slabs", C.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise OSError(f"oracle not built at {LIB}; run `make -C oracle`")
        L = C.CDLL(LIB)
        P, I, F, D = C.c_void_p, C.c_int, C.c_float, C.c_double
        pf = C.POINTER(C.c_float)
        pu = C.POINTER(C.c_uint32)
        pi = C.POINTER(C.c_int)
        pd = C.POINTER(C.c_double)
        sig = {
            "orc_create": (P, [C.POINTER(OrcParams), I, pf]), "orc_destroy": (None, [P]),
            "orc_set_config": (None, [P, F, F, F, F]),
            "orc_set_cubic_field": (I, [P, pi, pd, pd, pf, pf, I]),
            "orc_set_trilinear_field": (I, [P, pi, pd, pd, pf]),
            "orc_clear_field": (None, [P]),
            "orc_substep": (None, [P, F, I]), "orc_search_only": (None, [P, F]),
            "orc_get": (None, [P, I, pf]), "orc_count": (I, [P]), "orc_num_cells": (C.c_uint32, [P]),
            "orc_set_velocity": (None, [P, pf]),
            "orc_grid_dims": (None, [P, pi]),
            "orc_get_cells": (None, [P, pu, pu]), "orc_num_pairs": (C.c_uint64, [P]),
            "orc_get_neighbors": (None, [P, pu, pu]),
            "orc_density_error": (None, [P, pd, pd]),
            "orc_boundary_sample": (None, [P, I, pf, pf]), "orc_field_eval": (None, [P, I, pf, pf]),
            "orc_kernels": (None, [P, I, pf, pf, pf, pf]), "orc_sinf": (F, [F]),
            "orc_clamp": (None, [P, I, pf, pf]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _f(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _u(a):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _d3(v):
    return (C.c_double * 3)(*[float(x) for x in v])


class Oracle:
    """CPU restatement of one PBF scene (particles kept in id order)."""

    def __init__(self, xyzw, particle_radius=0.006, kernel_radius=0.025, rest_density=1000.0,
                 domain_lo=(0, 0, 0), domain_hi=(1, 1, 1), grid_mode=0, hash_log2=22, cell_slabs=0):
        self.L = lib()
        p = OrcParams()
        p.particle_radius = particle_radius
        p.kernel_radius = kernel_radius
        p.rest_density = rest_density
        p.domain_lo = _d3(domain_lo)
        p.domain_hi = _d3(domain_hi)
        p.grid_mode = grid_mode
        p.hash_log2 = hash_log2
        p.cell_slabs = cell_slabs  # z-slabs ordering a cell's particles (0: the library default, 4)
        a = np.ascontiguousarray(np.asarray(xyzw, np.float32).reshape(-1, 4))
        self.n = a.shape[0]
        self.s = self.L.orc_create(C.byref(p), self.n, _f(a))

    def __del__(self):
        try:
            if self.s:
                self.L.orc_destroy(self.s)
                self.s = None
        except Exception:
            pass

    def set_config(self, xsph=0.01, vorticity=0.1, tension=0.5, gravity=9.8):
        self.L.orc_set_config(self.s, xsph, vorticity, tension, gravity)

    def set_cubic_field(self, res, lo, hi, sdf, vol, quirk=1):
        s = np.ascontiguousarray(sdf, np.float32)
        v = np.ascontiguousarray(vol, np.float32)
        r = (C.c_int * 3)(*res)
        self.L.orc_set_cubic_field(self.s, r, _d3(lo), _d3(hi), _f(s), _f(v), int(quirk))

    def set_trilinear_field(self, res, lo, hi, sdf_vol):
        a = np.ascontiguousarray(sdf_vol, np.float32)
        r = (C.c_int * 3)(*res)
        self.L.orc_set_trilinear_field(self.s, r, _d3(lo), _d3(hi), _f(a))

    def substep(self, dt, iterations):
        self.L.orc_substep(self.s, float(dt), int(iterations))

    def step(self, dt, substeps, iterations):
        # the C ABI takes dt as float32 and divides in double
        sdt = np.float32(float(np.float32(dt)) / float(substeps))
        for _ in range(substeps):
            self.substep(float(sdt), iterations)

    def search(self, dt):
        self.L.orc_search_only(self.s, float(dt))

    def get(self, field):
        out = np.zeros((self.n, 4), np.float32)
        self.L.orc_get(self.s, field, _f(out))
        return out

    def set_velocity(self, v):
        a = np.zeros((self.n, 4), np.float32)
        a[:, :3] = np.asarray(v, np.float32).reshape(-1, 3)
        self.L.orc_set_velocity(self.s, _f(a))

    def cells(self):
        nc = int(self.L.orc_num_cells(self.s))
        c = np.zeros(nc, np.uint32)
        o = np.zeros(nc + 1, np.uint32)
        self.L.orc_get_cells(self.s, _u(c), _u(o))
        return c, o

    def neighbors(self):
        npairs = int(self.L.orc_num_pairs(self.s))
        off = np.zeros(self.n + 1, np.uint32)
        ids = np.zeros(max(npairs, 1), np.uint32)
        self.L.orc_get_neighbors(self.s, _u(off), _u(ids))
        return off, ids[:npairs]

    def density_error(self):
        m, x = C.c_double(), C.c_double()
        self.L.orc_density_error(self.s, C.byref(m), C.byref(x))
        return m.value, x.value

    def boundary_sample(self, x3):
        a = np.ascontiguousarray(np.asarray(x3, np.float32).reshape(-1, 3))
        out = np.zeros((a.shape[0], 4), np.float32)
        self.L.orc_boundary_sample(self.s, a.shape[0], _f(a), _f(out))
        return out

    def field_eval(self, x3):
        a = np.ascontiguousarray(np.asarray(x3, np.float32).reshape(-1, 3))
        out = np.zeros((a.shape[0], 5), np.float32)
        self.L.orc_field_eval(self.s, a.shape[0], _f(a), _f(out))
        return out

    def kernels(self, d2):
        a = np.ascontiguousarray(np.asarray(d2, np.float32).reshape(-1))
        p6, sp, co = (np.zeros_like(a) for _ in range(3))
        self.L.orc_kernels(self.s, a.size, _f(a), _f(p6), _f(sp), _f(co))
        return p6, sp, co

    def clamp(self, x3):
        a = np.ascontiguousarray(np.asarray(x3, np.float32).reshape(-1, 3))
        out = np.zeros_like(a)
        self.L.orc_clamp(self.s, a.shape[0], _f(a), _f(out))
        return out


def sinf(x):
    return float(lib().orc_sinf(float(x)))


# --- MLS-MPM (mpm_oracle.c) --------------------------------------------------
MPM_POSITION, MPM_VELOCITY, MPM_AFFINE, MPM_GRID = range(4)


class MpmParams(C.Structure):
    _fields_ = [("n_grid", C.c_int), ("bound", C.c_uint), ("dt", C.c_double), ("rho", C.c_double),
                ("E", C.c_double), ("gravity", C.c_double)]


def _mpm_lib():
    L = lib()
    if not getattr(L, "_mpm_sigs", False):
        P, I = C.c_void_p, C.c_int
        pf = C.POINTER(C.c_float)
        for name, (res, args) in {
            "orc_mpm_create": (P, [C.POINTER(MpmParams), I, pf]), "orc_mpm_destroy": (None, [P]),
            "orc_mpm_step": (None, [P]), "orc_mpm_get": (None, [P, I, pf]), "orc_mpm_set": (None, [P, I, pf]),
            "orc_mpm_p2g": (None, [P, pf]), "orc_mpm_g2p_only": (None, [P]),
        }.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        L._mpm_sigs = True
    return L


class MpmOracle:
    """CPU restatement of the MLS-MPM step (MPMShader.ts) on n particles."""

    def __init__(self, xyzw, n_grid=32, bound=3, dt=8e-4, rho=1.0, E=400.0, gravity=9.8):
        self.L = _mpm_lib()
        p = MpmParams(int(n_grid), int(bound), float(dt), float(rho), float(E), float(gravity))
        a = np.ascontiguousarray(np.asarray(xyzw, np.float32).reshape(-1, 4))
        self.n, self.ng = a.shape[0], int(n_grid)
        self.m = self.L.orc_mpm_create(C.byref(p), self.n, _f(a))

    def __del__(self):
        try:
            if self.m:
                self.L.orc_mpm_destroy(self.m)
                self.m = None
        except Exception:
            pass

    def _shape(self, field):
        return {MPM_POSITION: (self.n, 4), MPM_VELOCITY: (self.n, 4), MPM_AFFINE: (self.n, 9),
                MPM_GRID: (self.ng ** 3, 4)}[field]

    def step(self, steps=1):
        for _ in range(steps):
            self.L.orc_mpm_step(self.m)

    def get(self, field):
        out = np.zeros(self._shape(field), np.float32)
        self.L.orc_mpm_get(self.m, field, _f(out))
        return out

    def set(self, field, arr):
        a = np.ascontiguousarray(np.asarray(arr, np.float32).reshape(self._shape(field)))
        self.L.orc_mpm_set(self.m, field, _f(a))

    def p2g(self):
        """P2G + grid pass; returns the node sums (momentum xyz, mass) before the grid pass."""
        raw = np.zeros((self.ng ** 3, 4), np.float32)
        self.L.orc_mpm_p2g(self.m, _f(raw))
        return raw

    def g2p(self):
        self.L.orc_mpm_g2p_only(self.m)
