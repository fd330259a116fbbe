"""Helpers that run one scene on libpbf (GPU) and on the oracle (CPU)."""
import numpy as np

import oracle
import scenes


def make_oracle(sc: scenes.Scene, cell_slabs: int = 0) -> oracle.Oracle:
    o = oracle.Oracle(sc.positions, sc.particle_radius, sc.kernel_radius, 1000.0, sc.domain_lo, sc.domain_hi,
                      sc.grid_mode, sc.hash_log2, cell_slabs)
    c = sc.config
    o.set_config(c["XSPH"], c["vorticity"], c["surfaceTension"], c["gravity"])
    f = sc.field
    if f is not None:
        if f.kind == "trilinear":
            o.set_trilinear_field(f.res, f.lo, f.hi, f.sdf_vol)
        else:
            o.set_cubic_field(f.res, f.lo, f.hi, f.sdf, f.vol, quirk=1)
    return o


def gpu_fields(sim):
    return {k: sim.read_field(k) for k in ("position", "velocity", "accel", "lambda", "omega", "normal")}


ORACLE_FIELDS = {"position": oracle.POSITION, "velocity": oracle.VELOCITY, "accel": oracle.ACCEL,
                 "lambda": oracle.LAMBDA, "omega": oracle.OMEGA, "normal": oracle.NORMAL}


def oracle_fields(o):
    return {k: o.get(v) for k, v in ORACLE_FIELDS.items()}


def rel_err(a, b):
    """Norm-relative error max|a - b| / max|b| over the xyz (or x) components."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.abs(b).max()
    return float(np.abs(a - b).max() / den) if den > 0 else float(np.abs(a - b).max())


def ulp_diff(a, b):
    """Largest distance in float32 ulps between two float32 arrays."""
    ai = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    ai = np.where(ai < 0, -(ai & 0x7FFFFFFF), ai)
    bi = np.where(bi < 0, -(bi & 0x7FFFFFFF), bi)
    return int(np.abs(ai - bi).max()) if ai.size else 0
