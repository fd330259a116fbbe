"""The reference's boundary-volume definition, pinned on its own data.

The reference ships two boundary fields (public/model/domain_boundary.cdm,
torus.cdm; read by src/simulator/boundary/volumeMap.ts:23-39) but not the
tool that made them.  Their volume values are
    V(x) = 0.8 * sum_{Gauss-Legendre 16^3 points xi in [-h, h]^3, |xi| <= h}
           w(xi) * gamma(sdf(x + xi)),
    gamma = 1 in the solid (sdf <= 0), W(sdf) / W(0) of the cubic spline of
    support h for 0 < sdf < h, 0 beyond and outside the field's box,
with h = 0.025 and sdf the field's own SDF (32-node cubic interpolation):
the first test recomputes that sum for nodes of both files (where the
reference checkout is mounted) and matches every stored value to its printed
precision -- a 15- or 17-point rule, a linear gamma, a factor other than 0.8
or a support other than h each miss one of the stored values by more than
1 %.  The library's builder with
PBF_VOLUME_REFERENCE evaluates the same definition on analytic SDFs; the
other tests pin it against the golden wall profile of the reference's unit
box and against a numpy restatement of the sum.
"""
import os

import numpy as np
import pytest

import oracle
import scenes
from pbf_host import VOLUME_REFERENCE, build_cubic_field

REF = "/root/reference/public/model"
H = 0.025
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "domain_boundary_wall_profile.json")


def _gl_ball(h, n=16):
    x, w = np.polynomial.legendre.leggauss(n)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    W = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel()
    P = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    keep = (P * P).sum(1) <= 1.0
    return P[keep] * h, 0.8 * W[keep] * h ** 3


def _gamma(phi, h):
    q = phi / h
    g = np.where(q <= 0.5, (6 * q ** 3 - 6 * q ** 2) + 1, 2 * (1 - q) ** 3)
    return np.where(phi <= 0, 1.0, np.where(phi < h, g, 0.0))


def _parse(path):
    lines = open(path).read().split("\n")
    res = tuple(int(v) for v in lines[0].split())
    return res, np.array(lines[2].split(), np.float32), np.array(lines[4].split(), np.float32)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference checkout not mounted")
@pytest.mark.parametrize("name", ["domain_boundary", "torus"])
def test_definition_reproduces_the_reference_fields(name):
    res, sdf, vol = _parse(os.path.join(REF, f"{name}.cdm"))
    o = oracle.Oracle(np.zeros((1, 4), np.float32), 0.006, H)
    o.set_cubic_field(res, (0, 0, 0), (1, 1, 1), sdf, vol, quirk=0)
    nodes = scenes.cubic_nodes(res)
    X = np.concatenate([np.stack([n[0].ravel(), n[1].ravel(), n[2].ravel()], 1) for n in nodes])
    rng = np.random.default_rng(3)
    pick = np.concatenate([rng.choice(np.nonzero(vol > 0)[0], 120, replace=False),
                           rng.choice(np.nonzero((vol == 0) & (sdf < 0.07))[0], 20, replace=False)])
    P, W = _gl_ball(H)
    for i in pick:
        y = X[i] + P
        inside = np.all((y >= 0.0) & (y <= 1.0), 1)
        phi = o.field_eval(np.clip(y, 0, 1).astype(np.float32))[:, 3].astype(np.float64)
        v = float((W * np.where(inside, _gamma(phi, H), 0.0)).sum())
        assert v == pytest.approx(float(vol[i]), rel=3e-5, abs=1e-11), (name, i, X[i], v, vol[i])


def test_generated_unit_box_reproduces_the_reference_wall_profile():
    import json
    gold = json.load(open(GOLDEN))
    res = tuple(gold["res"])
    sdf, vol = build_cubic_field(VOLUME_REFERENCE, res, (0, 0, 0), (1, 1, 1), H)
    o = oracle.Oracle(np.zeros((1, 4), np.float32), 0.006, H)
    pts = np.array(gold["points"], np.float32)
    for quirk, key in ((1, "reference_sampler"), (0, "corrected_sampler")):
        o.set_cubic_field(res, (0, 0, 0), (1, 1, 1), sdf, vol, quirk=quirk)
        ev = o.field_eval(pts)
        np.testing.assert_allclose(ev[:, 3], gold[key]["sdf"], atol=2e-7)
        np.testing.assert_allclose(ev[:, 4], gold[key]["vol"], rtol=2e-5, atol=1e-12)
    # the facts themselves: V(0) = 1.3763e-5, V(1/60) = 1.2389e-5, V(>= 0.05) = 0
    assert gold["corrected_sampler"]["vol"][0] == pytest.approx(1.3763e-5, rel=1e-4)
    assert gold["corrected_sampler"]["vol"][1] == pytest.approx(1.2389e-5, rel=1e-4)
    assert gold["corrected_sampler"]["vol"][3] == 0.0


def test_builder_matches_a_numpy_restatement():
    # sphere + torus in the unit box, at nodes of every kind
    shape = [0.5, 0.45, 0.5, 0.12, 0.5, 0.45, 0.5, 0.25, 0.05]
    res = (12, 12, 12)
    sdf, vol = build_cubic_field(VOLUME_REFERENCE | 1, res, (0, 0, 0), (1, 1, 1), H, shape)
    nodes = scenes.cubic_nodes(res)
    X = np.concatenate([np.stack([n[0].ravel(), n[1].ravel(), n[2].ravel()], 1) for n in nodes])

    def sdf_of(p):
        box = np.minimum.reduce([p[:, 0], 1 - p[:, 0], p[:, 1], 1 - p[:, 1], p[:, 2], 1 - p[:, 2]])
        sph = np.linalg.norm(p - shape[0:3], axis=1) - shape[3]
        e = p - shape[4:7]
        tor = np.hypot(np.hypot(e[:, 0], e[:, 2]) - shape[7], e[:, 1]) - shape[8]
        return np.minimum(np.minimum(box, sph), tor)

    P, W = _gl_ball(H)
    d = sdf_of(X)
    assert np.allclose(sdf, d, atol=1e-6)
    rng = np.random.default_rng(7)
    pick = np.concatenate([rng.choice(np.nonzero(vol > 0)[0], 200, replace=False),
                           rng.choice(np.nonzero(vol == 0)[0], 50, replace=False)])
    for i in pick:
        y = X[i] + P
        inside = np.all((y >= 0.0) & (y <= 1.0), 1)
        v = float((W * np.where(inside, _gamma(sdf_of(y), H), 0.0)).sum()) if d[i] < 2 * H else 0.0
        assert v == pytest.approx(float(vol[i]), rel=2e-6, abs=1e-13), (i, X[i], v, vol[i])
    # deep inside the sphere every point counts: 0.8 x the rule's ball volume
    assert vol.max() == pytest.approx(float(W.sum()), rel=1e-6)
