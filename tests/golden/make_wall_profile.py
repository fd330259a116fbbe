"""Writes domain_boundary_wall_profile.json from the reference's unit-box
field (public/model/domain_boundary.cdm: 20^3 cells, 32-node cubic element).

The fixture holds
  * the points (0.5, 0.5, d) on the line normal to the z = 0 wall,
  * the 32 node values (SDF and volume) of each cubic element that line
    crosses for d < 0.1, with their flat indices in the .cdm arrays -- data
    copied from the reference's file, nothing computed,
  * SDF and volume at the points, evaluated over the FULL field with the
    oracle's cubic sampler in both index conventions: the reference's volume
    path (discreteFieldShader.ts:146-147,152-153 reads the sibling edge node)
    and the corrected one.

    python tests/golden/make_wall_profile.py   (reads /root/reference)
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "webgpu-fluid-simulation_amd"), os.path.join(ROOT, "oracle")]

import numpy as np  # noqa: E402

import oracle  # noqa: E402
import pbf_host  # noqa: E402

CDM = "/root/reference/public/model/domain_boundary.cdm"
DEPTHS = [0.0, 1.0 / 60.0, 1.0 / 30.0, 0.05, 0.07]
POINTS = [[0.5, 0.5, d] for d in DEPTHS]


def element_nodes(res, cell):
    """Flat indices of the 8 corner and 24 edge nodes of one element, in the
    .cdm layout (corners x-fastest, then x-, y-, z-edge blocks)."""
    R = np.array(res)
    nx, ny, nz = R + 1
    nc = nx * ny * nz
    nex, ney = nz * ny * 2 * R[0], nx * nz * 2 * R[1]
    i, j, k = cell
    out = []
    for ez in (0, 1):
        for ey in (0, 1):
            for ex in (0, 1):
                out.append(((k + ez) * ny + (j + ey)) * nx + (i + ex))
    for e1 in (0, 1):
        for e2 in (0, 1):
            for t in (0, 1):
                out.append(nc + ((k + e2) * ny + (j + e1)) * 2 * R[0] + 2 * i + t)
                out.append(nc + nex + ((i + e1) * nz + (k + e2)) * 2 * R[1] + 2 * j + t)
                out.append(nc + nex + ney + ((j + e2) * nx + (i + e1)) * 2 * R[2] + 2 * k + t)
    return sorted(int(v) for v in out)


def evaluate(res, sdf, vol):
    o = oracle.Oracle(np.zeros((1, 4), np.float32), 0.006, 0.025, 1000.0, (0, 0, 0), (1, 1, 1))
    pts = np.array(POINTS, np.float32)
    out = {}
    for name, quirk in (("reference_sampler", 1), ("corrected_sampler", 0)):
        o.set_cubic_field(res, (0, 0, 0), (1, 1, 1), sdf, vol, quirk=quirk)
        e = o.field_eval(pts)
        out[name] = {"sdf": [float(v) for v in e[:, 3]], "vol": [float(v) for v in e[:, 4]]}
    return out


def main():
    res, sdf, vol = pbf_host.parse_cdm(open(CDM, "rb").read())
    idx = sorted(set(element_nodes(res, (10, 10, 0)) + element_nodes(res, (10, 10, 1))))
    doc = {"source": "public/model/domain_boundary.cdm (LinzhouLi/WebGPU-Fluid-Simulation)",
           "res": [int(v) for v in res], "length": int(sdf.size), "points": POINTS,
           "nodes": {"index": idx, "sdf": [float(sdf[i]) for i in idx], "vol": [float(vol[i]) for i in idx]},
           **evaluate(res, sdf, vol)}
    with open(os.path.join(HERE, "domain_boundary_wall_profile.json"), "w") as f:
        json.dump(doc, f, indent=0)


if __name__ == "__main__":
    main()
