"""Pins of the CPU oracle (oracle/pbf_oracle.c) by closed forms, brute force,
invariants and published formulas -- no GPU needed.

Each test names the reference code the oracle restates and the independent
fact that pins it.
"""
import math
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle
import scenes

F32 = np.float32


def make(points, r=0.025, h=None, lo=(0, 0, 0), hi=(1, 1, 1), grid_mode=0, hash_log2=22, cell_slabs=0):
    h = 4 * r if h is None else h
    p = np.zeros((len(points), 4), np.float32)
    p[:, :3] = points
    return oracle.Oracle(p, r, h, 1000.0, lo, hi, grid_mode, hash_log2, cell_slabs)


# --------------------------------------------------------------------------
# SPH kernels (src/simulator/PBF/shader/common.ts:1-40)
def test_poly6_closed_forms():
    h = 0.1
    o = make(np.zeros((1, 3)), r=0.025, h=h)
    p6, sp, co = o.kernels(np.array([0.0], np.float32))
    assert p6[0] == pytest.approx(315 / (64 * math.pi * h ** 3), rel=2e-7)
    # normalisation: 4 pi int_0^h r^2 W(r) dr = 1, W evaluated by the oracle
    r = (np.arange(200000) + 0.5) / 200000 * h
    W, _, _ = o.kernels((r * r).astype(np.float32))
    integral = float(np.sum(4 * math.pi * r ** 2 * W.astype(np.float64)) * (h / 200000))
    assert integral == pytest.approx(1.0, abs=2e-5)
    # zero at and beyond h
    W, _, _ = o.kernels(np.array([h * h * 1.0001, 4 * h * h], np.float32))
    assert np.all(W == 0)


def test_spiky_gradient_matches_derivative_of_spiky_kernel():
    # W_spiky(r) = 15/(pi h^6) (h-r)^3 (Mueller et al. 2003); grad = W'(r) r_hat.
    h = 0.025
    o = make(np.zeros((1, 3)), r=0.006, h=h)
    r = np.linspace(0.05, 0.999, 400) * h
    _, f, _ = o.kernels((r * r).astype(np.float32))
    Ws = lambda x: 15 / (math.pi * h ** 6) * (h - x) ** 3
    eps = 1e-7 * h
    dW = (Ws(r + eps) - Ws(r - eps)) / (2 * eps)
    np.testing.assert_allclose(f.astype(np.float64) * r, dW, rtol=2e-4)
    _, f0, _ = o.kernels(np.array([0.0, h * h, 2 * h * h], np.float32))
    assert f0[0] == 0.0 and abs(f0[1]) < 1e-3 * abs(dW[0]) and f0[2] == 0.0


def test_cohesion_akinci_2013_and_continuity():
    # Akinci et al. 2013, eq. (2): C(r) = 32/(pi h^9) {(h-r)^3 r^3 if 2r > h; 2(h-r)^3 r^3 - h^6/64 if 0 < 2r <= h}
    h = 0.1
    o = make(np.zeros((1, 3)), r=0.025, h=h)
    r = np.linspace(0.01, 0.99, 300) * h
    _, _, f = o.kernels((r * r).astype(np.float32))
    C = np.where(2 * r > h, (h - r) ** 3 * r ** 3, 2 * (h - r) ** 3 * r ** 3 - h ** 6 / 64) * 32 / (math.pi * h ** 9)
    np.testing.assert_allclose(f.astype(np.float64) * r, C, rtol=2e-4, atol=1e-6 * np.abs(C).max())
    lo, hi = np.float32(0.5 * h * (1 - 1e-5)), np.float32(0.5 * h * (1 + 1e-5))
    _, _, fc = o.kernels(np.array([lo * lo, hi * hi], np.float32))
    assert fc[0] * lo == pytest.approx(fc[1] * hi, rel=1e-3)


# --------------------------------------------------------------------------
# wall clamp (common.ts:43-68) and its sin
def test_cephes_sinf_accuracy():
    xs = np.concatenate([np.linspace(-8000, 8000, 40001), np.linspace(-3.5, 3.5, 7001)]).astype(np.float32)
    err = max(abs(oracle.sinf(float(x)) - math.sin(float(x))) for x in xs)
    assert err < 3e-7


def test_clamp_box_and_inset():
    o = make(np.zeros((1, 3)), lo=(0, 0, 0), hi=(2, 1, 1))
    pts = np.array([[0.5, 0.5, 0.5], [-1.0, 0.3, 0.3], [2.5, 0.3, 1.5], [1.0, -0.1, 0.99995]], np.float32)
    out = o.clamp(pts)
    assert np.array_equal(out[0], pts[0])
    assert 0.0 <= out[1, 0] <= 1e-4 and out[1, 1] == pts[1, 1]
    assert 2.0 - 1e-4 <= out[2, 0] <= 2.0 and 1.0 - 1e-4 <= out[2, 2] <= 1.0
    assert 0.0 <= out[3, 1] <= 1e-4 and out[3, 2] <= 1.0
    assert np.array_equal(o.clamp(pts), out)


# --------------------------------------------------------------------------
# neighbour search (neighborShader.ts, neighborSearch.ts:270-301)
def brute_force_lists(x, h2):
    n = x.shape[0]
    lists = []
    for i in range(n):
        d = x - x[i]  # x_j - x_i, float32
        d2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        lists.append(set(np.nonzero(d2 <= h2)[0].tolist()))
    return lists


@pytest.mark.parametrize("grid_mode,hash_log2", [(0, 22), (1, 22), (1, 6)])
def test_neighbor_sets_vs_brute_force(grid_mode, hash_log2):
    rng = np.random.default_rng(7)
    r, h = 0.025, 0.1
    x = rng.uniform(0.0, 0.8, (1500, 3)).astype(np.float32)
    # pairs at exactly h and just beyond along an axis
    x[:10] = np.array([[0.3, 0.3, 0.3]] * 10, np.float32)
    x[1, 0] = np.float32(0.3) + np.float32(h)
    x[2, 0] = np.nextafter(x[1, 0], np.float32(1))
    x[3, 1] = np.float32(0.3) - np.float32(h)
    o = make(x, r=r, h=h, hi=(1, 1, 1), grid_mode=grid_mode, hash_log2=hash_log2)
    o.search(1e-9)
    xs = o.get(oracle.PREDICTED)[:, :3]
    h2 = np.float32(h * h)
    want = brute_force_lists(xs, h2)
    off, ids = o.neighbors()
    for i in range(x.shape[0]):
        got = ids[off[i]:off[i + 1]].tolist()
        assert len(got) == len(set(got)), "duplicate neighbour (hash slot visited twice)"
        assert set(got) == want[i], i


@pytest.mark.parametrize("slabs", [1, 2, 4, 0])
def test_neighbor_order_and_cells(slabs):
    # order of a list: cells x-major, then the z-slab inside the cell
    # (floor(S * (z / h - cz)), S = 4 by default), then id; the sets do not
    # depend on S
    rng = np.random.default_rng(3)
    r, h = 0.025, 0.1
    x = rng.uniform(0.0, 1.0, (3000, 3)).astype(np.float32)
    o = make(x, r=r, h=h, cell_slabs=slabs)
    o.search(1e-9)
    xs = o.get(oracle.PREDICTED)[:, :3]
    D = 10
    t = xs * np.float32(1 / h)
    c = np.clip(np.floor(t).astype(np.int64), 0, D - 1)
    key = (c[:, 0] * D + c[:, 1]) * D + c[:, 2]
    S = slabs or 4
    slab = np.clip(np.floor((t[:, 2] - c[:, 2].astype(np.float32)) * np.float32(S)).astype(np.int64), 0, S - 1)
    counts, offsets = o.cells()
    assert np.array_equal(counts, np.bincount(key, minlength=D ** 3).astype(np.uint32))
    assert np.array_equal(offsets, np.concatenate([[0], np.cumsum(counts)]).astype(np.uint32))
    off, ids = o.neighbors()
    ref = make(x, r=r, h=h, cell_slabs=1)
    ref.search(1e-9)
    roff, rids = ref.neighbors()
    assert np.array_equal(off, roff)
    for i in range(0, 3000, 37):
        lst = ids[off[i]:off[i + 1]]
        order = sorted(lst.tolist(), key=lambda j: (key[j], slab[j], j))
        assert lst.tolist() == order
        assert set(lst.tolist()) == set(rids[roff[i]:roff[i + 1]].tolist())
    if S > 1:  # the slab order is not the id order
        assert not np.array_equal(ids, rids)


# --------------------------------------------------------------------------
# boundary samplers (discreteFieldShader.ts:15-191, boundaryVolume.ts:26-57)
def cubic_poly(coef):
    mons = [(a, b, c) for a in range(4) for b in range(4) for c in range(4) if a + b + c <= 3]

    def f(x, y, z):
        return sum(k * x ** a * y ** b * z ** c for k, (a, b, c) in zip(coef, mons))

    def g(x, y, z):
        gx = sum(k * a * x ** max(a - 1, 0) * y ** b * z ** c for k, (a, b, c) in zip(coef, mons))
        gy = sum(k * b * x ** a * y ** max(b - 1, 0) * z ** c for k, (a, b, c) in zip(coef, mons))
        gz = sum(k * c * x ** a * y ** b * z ** max(c - 1, 0) for k, (a, b, c) in zip(coef, mons))
        return np.stack([gx, gy, gz], -1)
    return f, g, len(mons)


def test_cubic_field_reproduces_cubic_polynomials():
    rng = np.random.default_rng(11)
    f, g, nm = cubic_poly([0] * 20)
    coef = rng.uniform(-1, 1, nm)
    f, g, _ = cubic_poly(coef)
    fld = scenes.cubic_field_from(f, f, (6, 5, 7))
    o = make(np.zeros((1, 3)))
    o.set_cubic_field(fld.res, fld.lo, fld.hi, fld.sdf, fld.vol, quirk=0)
    pts = rng.uniform(0.0, 0.999, (400, 3)).astype(np.float32)
    out = o.field_eval(pts)
    P = pts.astype(np.float64)
    want = f(P[:, 0], P[:, 1], P[:, 2])
    np.testing.assert_allclose(out[:, 3], want, atol=2e-5)
    np.testing.assert_allclose(out[:, 4], want, atol=2e-5)
    np.testing.assert_allclose(out[:, :3], g(P[:, 0], P[:, 1], P[:, 2]), atol=5e-4)


def serendipity_numpy(res, sdf, pts, quirk):
    """Independent float64 evaluation of the 32-node element with explicit
    node coordinates; quirk swaps the edge nodes the reference reads twice."""
    R = np.array(res)
    nx, ny, nz = R + 1
    nc = nx * ny * nz
    nex = nz * ny * 2 * R[0]
    ney = nx * nz * 2 * R[1]
    out = []
    for p in pts.astype(np.float64):
        u = p * R
        c = np.minimum(np.floor(u).astype(int), R - 1)
        s = 2 * (u - c) - 1
        i, j, k = c
        val = 0.0
        for ex in (0, 1):
            for ey in (0, 1):
                for ez in (0, 1):
                    sx, sy, sz = 2 * ex - 1, 2 * ey - 1, 2 * ez - 1
                    N = (1 + sx * s[0]) * (1 + sy * s[1]) * (1 + sz * s[2]) * (9 * (s @ s) - 19) / 64
                    val += N * sdf[((k + ez) * ny + (j + ey)) * nx + (i + ex)]
        for ax in range(3):
            o1, o2 = [a for a in range(3) if a != ax]
            for t, tsgn in ((0, -1), (1, 1)):
                for e1 in (0, 1):
                    for e2 in (0, 1):
                        N = 9 / 64 * (1 - s[ax] ** 2) * (1 + 3 * tsgn * s[ax]) * (1 + (2 * e1 - 1) * s[o1]) * \
                            (1 + (2 * e2 - 1) * s[o2])
                        a2 = 2 * c[ax] + t
                        if ax == 0:
                            a2q = 2 * c[0] if (quirk and e2 == 1 and t == 1) else a2
                            idx = nc + ((k + e2) * ny + (j + e1)) * 2 * R[0] + a2q
                        elif ax == 1:
                            idx = nc + nex + ((i + e1) * nz + (k + e2)) * 2 * R[1] + a2
                        else:
                            a2q = 2 * c[2] if (quirk and t == 1) else a2
                            idx = nc + nex + ney + ((j + e2) * nx + (i + e1)) * 2 * R[2] + a2q
                        val += N * sdf[idx]
        out.append(val)
    return np.array(out)


@pytest.mark.parametrize("quirk", [0, 1])
def test_cubic_volume_path_index_quirk(quirk):
    # A field linear in x and z exposes the reference's repeated edge reads
    # (discreteFieldShader.ts:146-147,152-153) on the volume path only.
    fld = scenes.cubic_field_from(lambda x, y, z: 0.3 + x - 2 * z + 0.5 * y,
                                  lambda x, y, z: 0.3 + x - 2 * z + 0.5 * y, (4, 4, 4))
    o = make(np.zeros((1, 3)))
    o.set_cubic_field(fld.res, fld.lo, fld.hi, fld.sdf, fld.vol, quirk=quirk)
    rng = np.random.default_rng(5)
    pts = rng.uniform(0, 0.999, (200, 3)).astype(np.float32)
    out = o.field_eval(pts)
    np.testing.assert_allclose(out[:, 3], serendipity_numpy(fld.res, fld.sdf.astype(np.float64), pts, 0), atol=2e-6)
    want_vol = serendipity_numpy(fld.res, fld.vol.astype(np.float64), pts, quirk)
    np.testing.assert_allclose(out[:, 4], want_vol, atol=2e-6)
    exact = 0.3 + pts[:, 0] - 2 * pts[:, 2] + 0.5 * pts[:, 1]
    dev = np.abs(want_vol - exact).max()
    assert (dev > 1e-3) if quirk else (dev < 1e-5)


GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
REF_CDM = "/root/reference/public/model/domain_boundary.cdm"


def _wall_profile():
    import json
    with open(os.path.join(GOLDEN, "domain_boundary_wall_profile.json")) as f:
        return json.load(f)


def test_wall_profile_fixture_holds_the_reference_facts():
    # The planar-wall profile of the reference's unit-box field, as evaluated
    # by the oracle's sampler: V(0) = 1.3763e-5, V(1/60) = 1.2389e-5,
    # V(1/30) = 9.841e-7 on the corrected element (the reference's volume
    # path repeats the 1/60 edge node there), V = 0 from d = 0.05; the SDF
    # along the normal is the depth.
    doc = _wall_profile()
    d = np.array([p[2] for p in doc["points"]])
    for key in ("reference_sampler", "corrected_sampler"):
        np.testing.assert_allclose(doc[key]["sdf"], d, atol=1e-6)
    fix, ref = np.array(doc["corrected_sampler"]["vol"]), np.array(doc["reference_sampler"]["vol"])
    np.testing.assert_allclose(fix[:3], [1.3763e-5, 1.2389e-5, 9.841e-7], rtol=1e-4)
    assert np.all(fix[3:] == 0.0) and np.all(ref[3:] == 0.0)
    np.testing.assert_allclose(ref[:2], fix[:2], rtol=1e-6)
    assert ref[2] == pytest.approx(ref[1], rel=1e-6)


@pytest.mark.skipif(not os.path.exists(REF_CDM), reason="reference checkout not mounted")
def test_wall_profile_fixture_matches_reference_file():
    import pbf_host
    doc = _wall_profile()
    with open(REF_CDM, "rb") as f:
        res, sdf, vol = pbf_host.parse_cdm(f.read())
    assert list(res) == doc["res"] and sdf.size == doc["length"]
    pts = np.array(doc["points"], np.float32)
    o = make(np.zeros((1, 3)), r=0.006, h=0.025)
    for quirk, key in ((1, "reference_sampler"), (0, "corrected_sampler")):
        o.set_cubic_field(res, (0, 0, 0), (1, 1, 1), sdf, vol, quirk=quirk)
        out = o.field_eval(pts)
        assert np.array_equal(out[:, 3], np.array(doc[key]["sdf"], np.float32))
        assert np.array_equal(out[:, 4], np.array(doc[key]["vol"], np.float32))
        # an independent float64 evaluation of the same 32-node element
        np.testing.assert_allclose(out[:, 4], serendipity_numpy(res, vol.astype(np.float64), pts, quirk),
                                   rtol=1e-5, atol=1e-12)


def test_trilinear_field_linear_exact():
    res, lo, hi = (8, 6, 5), (0.0, -1.0, 0.5), (2.0, 1.0, 1.5)
    nodes = np.stack(np.meshgrid(*[np.linspace(lo[a], hi[a], res[a] + 1) for a in range(3)], indexing="ij"), -1)
    nodes = nodes.transpose(2, 1, 0, 3).reshape(-1, 3)  # x fastest
    sv = np.zeros((nodes.shape[0], 2), np.float32)
    sv[:, 0] = 0.25 + 0.5 * nodes[:, 0] - 0.25 * nodes[:, 1] + 0.75 * nodes[:, 2]
    sv[:, 1] = 1.0 + nodes[:, 2]
    o = make(np.zeros((1, 3)), lo=lo, hi=hi)
    o.set_trilinear_field(res, lo, hi, sv)
    rng = np.random.default_rng(2)
    pts = np.stack([rng.uniform(lo[a], hi[a], 300) for a in range(3)], -1).astype(np.float32)
    out = o.field_eval(pts)
    P = pts.astype(np.float64)
    np.testing.assert_allclose(out[:, 3], 0.25 + 0.5 * P[:, 0] - 0.25 * P[:, 1] + 0.75 * P[:, 2], atol=3e-6)
    np.testing.assert_allclose(out[:, 4], 1.0 + P[:, 2], atol=3e-6)
    np.testing.assert_allclose(out[:, :3], np.tile([0.5, -0.25, 0.75], (300, 1)), atol=3e-5)


def test_boundary_debug_query_known_answer():
    # DebugShader (discreteFieldShader.ts:194-237): point (0.5, 0.5, 0.02),
    # ParticleRadius 0.007353, unit-box field: d = 0.02, n = +z,
    # x_b = x - max(d + r/2, 2r) n.
    fld = scenes.unit_box_cdm_field()
    o = make(np.zeros((1, 3)), r=0.007353, h=0.025)
    o.set_cubic_field(fld.res, fld.lo, fld.hi, fld.sdf, fld.vol, quirk=1)
    out = o.boundary_sample(np.array([[0.5, 0.5, 0.02]], np.float32))[0]
    zb = 0.02 - max(0.02 + 0.5 * 0.007353, 2 * 0.007353)
    np.testing.assert_allclose(out[:3], [0.5, 0.5, zb], atol=2e-6)
    assert out[3] > 0
    # beyond h: no boundary sample
    assert np.all(o.boundary_sample(np.array([[0.5, 0.5, 0.03]], np.float32)) == 0)


# --------------------------------------------------------------------------
# physics invariants and closed forms (pressureConstrain.ts, attributeUpdate.ts)
def lattice(n, r, origin=0.5):
    g = np.arange(n) * 2 * r
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1) + origin


@pytest.mark.parametrize("r,h,expect", [(0.025, 0.1, 0.908798), (0.006, 0.025, 0.912940)])
def test_rest_lattice_density(r, h, expect):
    # V sum_j W_ij over a perfect 2r lattice, computed here in float64.
    m = int(math.ceil(h / (2 * r)))
    g = np.arange(-m, m + 1) * 2 * r
    d2 = (g[:, None, None] ** 2 + g[None, :, None] ** 2 + g[None, None, :] ** 2).ravel()
    W = np.where(d2 <= h * h, 315 / (64 * math.pi * h ** 9) * (h * h - d2) ** 3, 0)
    rho = 0.9 * (2 * r) ** 3 * W.sum()
    assert rho == pytest.approx(expect, abs=5e-7)
    # the oracle's K12 density of an interior particle (gravity off)
    x = lattice(11, r, 0.3)
    o = make(x, r=r, h=h, hi=(2, 2, 2))
    o.set_config(0.0, 0.0, 0.0, 0.0)
    o.substep(1e-4, 3)
    p = o.get(oracle.POSITION)
    centre = np.argmin(np.linalg.norm(x - x.mean(0), axis=1))
    assert p[centre, 3] == pytest.approx(1000 * rho, rel=1e-5)  # f32 sum of ~40 poly6 terms


def test_lattice_block_is_stationary_without_gravity():
    # rho~ < 1 everywhere -> C = 0 -> lambda = 0, dp = 0, v = 0; with tension
    # off nothing moves, bit for bit.
    x = lattice(8, 0.025, 0.4)
    o = make(x, hi=(2, 2, 2))
    o.set_config(0.01, 0.01, 0.0, 0.0)
    for _ in range(5):
        o.substep(1 / 120, 4)
    p = o.get(oracle.POSITION)
    assert np.array_equal(p[:, :3], x.astype(np.float32))
    assert np.all(o.get(oracle.VELOCITY) == 0) and np.all(o.get(oracle.LAMBDA)[:, 0] == 0)


def test_two_particle_lambda_and_correction_closed_form():
    # h = 2r makes V W(0) > 1 so a pair is over-dense; lambda and dp in closed form.
    r, h, d = 0.025, 0.05, 0.03
    x = np.array([[0.5, 0.5, 0.5], [0.5 + d, 0.5, 0.5]])
    o = make(x, r=r, h=h)
    o.set_config(0.0, 0.0, 0.0, 0.0)
    o.substep(1e-3, 1)
    V = 0.9 * (2 * r) ** 3
    W = lambda q: 315 / (64 * math.pi * h ** 9) * (h * h - q * q) ** 3
    dW = -45 / (math.pi * h ** 6) * (h - d) ** 2
    rho = V * (W(0) + W(d))
    lam = -(rho - 1) / (2 * V * V * dW * dW + 1e-6)
    assert o.get(oracle.LAMBDA)[0, 0] == pytest.approx(lam, rel=1e-5)
    # grad W(x_0 - x_1) = dW * (-x_hat); dp_0 = 0.5 V (lam + lam) grad W
    dp0 = -V * lam * dW
    p = o.get(oracle.POSITION)
    assert dp0 < 0  # the over-dense pair pushes apart
    assert p[0, 0] == pytest.approx(0.5 + dp0, rel=1e-6)
    assert p[1, 0] == pytest.approx(0.5 + d - dp0, rel=1e-6)


def test_free_fall_recursion():
    x = np.array([[0.5, 0.9, 0.5]])
    o = make(x, hi=(1, 1, 1))
    o.set_config(0.0, 0.0, 0.0, 9.8)
    dt = np.float32(1 / 120)
    y, v = 0.9, 0.0
    for _ in range(20):
        o.substep(float(dt), 2)
        vn = v + float(dt) * -9.8
        yn = y + float(dt) * vn
        v, y = (yn - y) / float(dt), yn
    p = o.get(oracle.POSITION)
    assert p[0, 1] == pytest.approx(y, rel=1e-5)
    assert o.get(oracle.VELOCITY)[0, 1] == pytest.approx(v, rel=1e-4)


def test_vorticity_normal_xsph_tension_f64():
    # K13 (vorticityConfinement.ts:46-62) and K14 (XSPH.ts:56-88) evaluated in
    # float64 from the published kernels (spiky gradient, poly6, Akinci
    # cohesion) on the oracle's own state and neighbour list after two
    # substeps of a random blob: a swapped cross product, a lost 1/rho_j, a
    # wrong Kij or a missing term moves a result by far more than the f32
    # rounding allowed here.
    rng = np.random.default_rng(11)
    r, h = 0.025, 0.1
    x = rng.uniform(0.35, 0.65, (400, 3))
    o = make(x, r=r, h=h)
    xsph, vort, tens = 0.05, 0.2, 0.5
    o.set_config(xsph, vort, tens, 9.8)
    dt = float(np.float32(1 / 120))
    o.substep(dt, 3)
    o.substep(dt, 3)
    p = o.get(oracle.POSITION).astype(np.float64)
    v = o.get(oracle.VELOCITY)[:, :3].astype(np.float64)
    om, nrm, acc = o.get(oracle.OMEGA), o.get(oracle.NORMAL), o.get(oracle.ACCEL)
    off, ids = o.neighbors()
    m = 0.9 * (2 * r) ** 3 * 1000.0
    ks, kp, kc = -45 / (math.pi * h ** 6), 315 / (64 * math.pi * h ** 9), 32 / (math.pi * h ** 9)
    x3, rho = p[:, :3], p[:, 3]

    def pairs(i):
        j = ids[off[i]:off[i + 1]].astype(np.int64)
        d = x3[i] - x3[j]
        q = np.sqrt((d * d).sum(1))
        inr, pos = q <= h, q > 0
        unit = np.where(pos[:, None], d / np.where(pos, q, 1)[:, None], 0)
        grad = np.where(inr, ks * (h - q) ** 2, 0)[:, None] * unit
        w = np.where(inr, kp * (h * h - q * q) ** 3, 0)
        t = (h - q) ** 3 * q ** 3
        c = np.where(q > h / 2, t, 2 * t - h ** 6 / 64)
        coh = (kc * np.where(inr, c, 0))[:, None] * unit
        return j, grad, w, coh

    n = len(x3)
    w64, n64 = np.zeros((n, 3)), np.zeros((n, 3))
    for i in range(n):
        j, grad, _, _ = pairs(i)
        vol = (m / rho[j])[:, None]
        w64[i] = (vol * np.cross(v[j] - v[i], grad)).sum(0)
        n64[i] = h * (vol * grad).sum(0)
    wabs = np.linalg.norm(w64, axis=1)
    a64, eta = np.zeros((n, 3)), np.zeros(n)
    for i in range(n):
        j, grad, w, coh = pairs(i)
        vel = ((v[j] - v[i]) * (w / rho[j])[:, None]).sum(0)
        e = ((wabs[j] / rho[j])[:, None] * grad).sum(0)
        K = 2000.0 / (rho[i] + rho[j])
        ten = -(K[:, None] * (m * coh + (n64[i] - n64[j]))).sum(0)
        eta[i] = np.linalg.norm(e)
        eh = e / eta[i] if eta[i] > 0 else 0 * e
        a64[i] = xsph / dt * m * vel + vort * np.cross(eh, w64[i]) + tens * ten
    scale = lambda a: np.abs(a).max()
    assert scale(w64) > 0.1 and scale(n64) > 0.1
    np.testing.assert_allclose(om[:, :3], w64, atol=2e-5 * scale(w64))
    np.testing.assert_allclose(om[:, 3], wabs, atol=2e-5 * scale(w64))
    np.testing.assert_allclose(nrm[:, :3], n64, atol=2e-5 * scale(n64))
    # eta-hat is ill-conditioned where eta nearly cancels
    ok = eta > 1e-3 * np.median(eta)
    assert ok.mean() > 0.95
    np.testing.assert_allclose(acc[ok, :3], a64[ok], atol=1e-4 * scale(a64))


def test_density_error_fp64_brute_force():
    sc = scenes.dam_break(6, 6, 6, map_res=16, name="tiny")
    o = oracle.Oracle(sc.positions, sc.particle_radius, sc.kernel_radius, 1000.0, sc.domain_lo, sc.domain_hi)
    o.set_trilinear_field(sc.field.res, sc.field.lo, sc.field.hi, sc.field.sdf_vol)
    o.set_config(0.01, 0, 0, 9.8)
    for _ in range(3):
        o.substep(sc.substep_dt(), 4)
    mean, mx = o.density_error()
    p = o.get(oracle.POSITION)[:, :3].astype(np.float64)
    off, ids = o.neighbors()
    h, r = sc.kernel_radius, sc.particle_radius
    V = 0.9 * (2 * r) ** 3
    kp = 315 / (64 * math.pi * h ** 9)
    b = o.boundary_sample(o.get(oracle.POSITION)[:, :3])
    C = []
    for i in range(p.shape[0]):
        d2 = ((p[ids[off[i]:off[i + 1]]] - p[i]) ** 2).sum(1)
        rho = V * np.sum(np.where(d2 <= h * h, kp * (h * h - d2) ** 3, 0))
        if b[i, 3] > 0:
            q = ((p[i] - b[i, :3].astype(np.float64)) ** 2).sum()
            rho += b[i, 3] * (kp * (h * h - q) ** 3 if q <= h * h else 0)
        C.append(max(rho - 1, 0))
    assert mean == pytest.approx(np.mean(C), rel=1e-9, abs=1e-15)
    assert mx == pytest.approx(np.max(C), rel=1e-9, abs=1e-15)


def test_oracle_is_deterministic_across_thread_counts(tmp_path):
    code = ("import sys; sys.path[:0]=%r\n"
            "import numpy as np, oracle, scenes\n"
            "sc=scenes.dam_break(12,12,8,map_res=16)\n"
            "o=oracle.Oracle(sc.positions,sc.particle_radius,sc.kernel_radius,1000.0,sc.domain_lo,sc.domain_hi)\n"
            "o.set_trilinear_field(sc.field.res,sc.field.lo,sc.field.hi,sc.field.sdf_vol)\n"
            "o.set_config(0.01,0.01,0.1,9.8)\n"
            "[o.substep(sc.substep_dt(),4) for _ in range(4)]\n"
            "np.save(sys.argv[1], o.get(oracle.POSITION))\n") % (sys.path[:3],)
    outs = []
    for t in ("1", "8"):
        f = tmp_path / f"p{t}.npy"
        env = dict(os.environ, OMP_NUM_THREADS=t)
        subprocess.run([sys.executable, "-c", code, str(f)], check=True, env=env)
        outs.append(np.load(f))
    assert np.array_equal(outs[0], outs[1])


# --------------------------------------------------------------------------
# K13 / K14 invariants (vorticityConfinement.ts:46-62, XSPH.ts:56-88): facts
# that hold whatever the summation order, each chosen so that a plausible
# misreading of the shaders (v_j for v_j - v_i, a swapped cross product, a
# lost m / rho_j or K_ij, a flipped normal difference) fails it.
def _lattice(n, spacing, origin):
    g = origin + spacing * np.arange(n)
    X, Y, Z = np.meshgrid(g, g, g, indexing="ij")
    return np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)


def test_uniform_velocity_gives_zero_vorticity_and_xsph_exactly():
    # Dyadic lattice (r = 2^-7, spacing 2^-6), dt = 1/64 and g = 4 make
    # every predicted position exact: x* = x - 2^-10 in y.  The rest lattice
    # has rho~ = 0.9088 < 1, so lambda = 0 and every particle gets exactly
    # v = (0, -1/16, 0): then omega = sum m/rho_j (v_j - v_i) x grad W = 0,
    # the XSPH sum of (v_j - v_i) W is 0 and the vorticity force
    # eps (eta^ x omega) is 0, bit for bit.
    r = 2.0 ** -7
    x = _lattice(10, 2 * r, 0.25)
    o = make(x, r=r)
    o.set_config(0.05, 0.2, 0.0, 4.0)
    o.substep(1.0 / 64.0, 3)
    v = o.get(oracle.VELOCITY)[:, :3]
    assert np.all(v[:, 1] == -1.0 / 16.0) and np.all(v[:, 0] == 0) and np.all(v[:, 2] == 0)
    assert np.all(o.get(oracle.POSITION)[:, :3] == (x + [0, -2.0 ** -10, 0]).astype(np.float32))
    assert np.all(o.get(oracle.OMEGA) == 0.0)
    assert np.all(o.get(oracle.ACCEL)[:, :3] == 0.0)
    n = o.get(oracle.NORMAL)[:, :3]
    assert np.abs(n).max() > 0.1          # the block's faces do have normals


def test_rigid_rotation_gives_omega_along_the_rotation_axis():
    # v = Omega x (x - c) on a 12^3 lattice (spacing 2r, h = 4r).  For an
    # interior particle, with x_ij = x_i - x_j and grad W_ij = f(r) x_ij
    # (f = -45/(pi h^6) (h - r)^2 / r < 0):
    #   (v_j - v_i) x grad W_ij = f (Omega r^2 - x_ij (x_ij . Omega)),
    # and on a cubic lattice sum_j f x_ij x_ij^T = (1/3) sum_j f r^2 I, so
    #   omega_i = S Omega,  S = (2/3) sum_j (m / rho_j) f(r_ij) r_ij^2 < 0.
    # Swapping the cross product flips the sign; dropping m / rho_j changes S.
    r, h = 0.025, 0.1
    x = _lattice(12, 2 * r, 0.3)
    c = x.mean(0)
    Om = np.array([0.3, 1.0, -0.5])
    o = make(x, r=r, hi=(2, 2, 2))
    o.set_config(0.0, 0.0, 0.0, 0.0)
    o.set_velocity(np.cross(Om, x - c))
    o.substep(1e-3, 2)
    p = o.get(oracle.POSITION).astype(np.float64)
    om = o.get(oracle.OMEGA)[:, :3].astype(np.float64)
    m = 0.9 * (2 * r) ** 3 * 1000.0
    ks = -45.0 / (math.pi * h ** 6)
    idx = np.arange(12)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    inner = ((I >= 2) & (I <= 9) & (J >= 2) & (J <= 9) & (K >= 2) & (K <= 9)).ravel()
    checked = 0
    for i in np.nonzero(inner)[0]:
        d = p[i, :3] - p[:, :3]
        q = np.sqrt((d * d).sum(1))
        nb = (q <= h) & (q > 0)
        S = (2.0 / 3.0) * np.sum(m / p[nb, 3] * ks * (h - q[nb]) ** 2 / q[nb] * q[nb] ** 2)
        assert S < 0
        assert np.linalg.norm(om[i] - S * Om) <= 2e-3 * np.linalg.norm(S * Om), (i, om[i], S * Om)
        checked += 1
    assert checked == 8 ** 3


def test_normals_vanish_inside_and_point_into_the_fluid_at_a_free_surface():
    # n_i = h sum m / rho_j grad W_ij: zero by symmetry inside a lattice; on
    # the top face every neighbour lies below (x_ij . y > 0, f < 0), so n
    # points down, into the fluid.
    r = 0.025
    x = _lattice(10, 2 * r, 0.3)
    o = make(x, r=r, hi=(2, 2, 2))
    o.set_config(0.0, 0.0, 0.0, 0.0)
    o.substep(1e-3, 2)
    n = o.get(oracle.NORMAL)[:, :3].astype(np.float64)
    idx = np.arange(10)
    I, J, K = np.meshgrid(idx, idx, idx, indexing="ij")
    I, J, K = I.ravel(), J.ravel(), K.ravel()
    top = (J == 9) & (I >= 2) & (I <= 7) & (K >= 2) & (K <= 7)
    inner = (I >= 2) & (I <= 7) & (J >= 2) & (J <= 7) & (K >= 2) & (K <= 7)
    ny = n[top, 1]
    # (the lattice coordinates 0.3 + 0.05 k are rounded to float32: 1e-5)
    assert np.all(ny < 0) and np.all(np.abs(n[top][:, [0, 2]]).max(1) < 1e-5 * np.abs(ny))
    assert np.abs(n[inner]).max() < 1e-5 * np.abs(ny).min()


def test_surface_tension_forces_sum_to_zero_over_an_isolated_cloud():
    # tension_i = -sum_j K_ij (m C(x_ij) + (n_i - n_j)): K_ij is symmetric,
    # the cohesion vector and n_i - n_j are antisymmetric, so the forces of
    # an isolated cloud (equal masses) cancel (momentum is conserved).
    rng = np.random.default_rng(5)
    r = 0.025
    x = rng.uniform(0.4, 0.6, (300, 3))
    o = make(x, r=r, hi=(2, 2, 2))
    o.set_config(0.0, 0.0, 0.5, 9.8)
    o.substep(1.0 / 240.0, 3)
    o.substep(1.0 / 240.0, 3)
    a = o.get(oracle.ACCEL)[:, :3].astype(np.float64)
    assert np.abs(a).sum() > 1.0
    assert np.abs(a.sum(0)).max() <= 1e-5 * np.abs(a).sum(0).max()


@pytest.mark.parametrize("d", [0.07, 0.03])
def test_two_particle_tension_closed_form(d):
    # Two particles at distance d (both branches of the Akinci cohesion
    # spline, Akinci et al. 2013 eq. 2): with x^ = (x_i - x_j) / d,
    #   rho = m (W(0) + W(d)),  K = 2 rho0 / (2 rho),
    #   n_i - n_j = 2 h (m / rho) spiky(d) x^,
    #   a_i = -gamma K (m C(d) + 2 h (m / rho) spiky(d)) x^.
    r, h, gamma = 0.025, 0.1, 1.0
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    xi = np.array([0.5, 0.5, 0.5])
    pts = np.array([xi, xi - d * u], np.float32)
    o = make(pts.astype(np.float64), r=r, hi=(2, 2, 2))
    o.set_config(0.0, 0.0, gamma, 0.0)
    o.substep(1e-3, 1)
    p = o.get(oracle.POSITION).astype(np.float64)
    a = o.get(oracle.ACCEL)[:, :3].astype(np.float64)
    dv = p[0, :3] - p[1, :3]
    dd = np.linalg.norm(dv)
    xh = dv / dd
    m = 0.9 * (2 * r) ** 3 * 1000.0
    W = lambda q: 315.0 / (64 * math.pi * h ** 9) * (h * h - q * q) ** 3
    rho = m * (W(0.0) + W(dd))
    assert p[0, 3] == pytest.approx(rho, rel=1e-5)
    t = (h - dd) ** 3 * dd ** 3
    Cd = 32.0 / (math.pi * h ** 9) * (t if 2 * dd > h else 2 * t - h ** 6 / 64)
    spiky = -45.0 / (math.pi * h ** 6) * (h - dd) ** 2
    K = 2000.0 / (2 * rho)
    want = -gamma * K * (m * Cd + 2 * h * (m / rho) * spiky) * xh
    np.testing.assert_allclose(a[0], want, rtol=2e-5, atol=1e-6 * np.abs(want).max())
    np.testing.assert_allclose(a[1], -want, rtol=2e-5, atol=1e-6 * np.abs(want).max())
