"""Writes domain_boundary_wall_profile.json from the reference's unit-box
field (public/model/domain_boundary.cdm: 20^3 cells, 32-node cubic element).

The fixture holds the points (0.5, 0.5, d) on the line normal to the z = 0
wall and the SDF and volume there, evaluated over the reference's field with
the oracle's cubic sampler in both index conventions: the reference's volume
path (discreteFieldShader.ts:146-147,152-153 reads the sibling edge node) and
the corrected one.  Only these evaluated values are kept; the field itself
stays in the reference checkout.

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
    doc = {"source": "public/model/domain_boundary.cdm (LinzhouLi/WebGPU-Fluid-Simulation), evaluated",
           "res": [int(v) for v in res], "length": int(sdf.size), "points": POINTS, **evaluate(res, sdf, vol)}
    with open(os.path.join(HERE, "domain_boundary_wall_profile.json"), "w") as f:
        json.dump(doc, f, indent=0)


if __name__ == "__main__":
    main()
