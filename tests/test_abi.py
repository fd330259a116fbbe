"""CPU checks of the drop-in boundary: libpbf.so loads, exports every entry
point declared in include/pbf.h, and its GPU-free host code (voxelizers,
.cdm parser, volume-map builder) matches independent restatements."""
import math
import os
import re
import subprocess
import sys

import numpy as np
import pytest

import pbf_host
import scenes

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "pbf.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pbf_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = pbf_host.load_library()
    out = subprocess.run(["nm", "-D", "--defined-only", pbf_host.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\sT\s(pbf_\w+)", out))
    missing = [s for s in declared_symbols() if s not in exported]
    assert not missing, missing
    assert len(declared_symbols()) >= 30
    assert lib.pbf_version().decode().startswith("pbf-mi355x")


def test_voxelize_cube_matches_reference_loop():
    # SPH.voxelizeCube (SPH.ts:99-121): floor((max-min)/2r) per axis, i outer.
    r, mn, mx = 0.006, (0.15, 0.35, 0.15), (0.65, 0.85, 0.65)
    p = scenes.lattice_cube(r, mn, mx)
    d = 2 * r
    n = [math.floor((mx[a] - mn[a]) / d) for a in range(3)]
    assert p.shape == (n[0] * n[1] * n[2], 4) == (41 ** 3, 4)
    I, J, K = np.meshgrid(*[np.arange(k) for k in n], indexing="ij")
    want = np.stack([I.ravel() * d + mn[0], J.ravel() * d + mn[1], K.ravel() * d + mn[2]], -1).astype(np.float32)
    assert np.array_equal(p[:, :3], want) and np.all(p[:, 3] == 0)


def test_voxelize_sphere_matches_reference_loop():
    r, c, R = 0.006, (0.5, 0.7, 0.5), 0.12
    p = scenes.lattice_sphere(r, c, R)
    n = math.floor(R / r / 2)
    g = np.arange(-n, n + 1) * 2 * r
    I, J, K = np.meshgrid(g, g, g, indexing="ij")
    q = np.stack([I.ravel(), J.ravel(), K.ravel()], -1)
    keep = (q[:, 0] ** 2 + q[:, 1] ** 2) + q[:, 2] ** 2 <= R * R
    assert np.array_equal(p[:, :3], (q[keep] + np.array(c)).astype(np.float32))


def test_jitter_is_seeded_and_bounded():
    r = 0.025
    a = scenes.lattice_cube(r, (0, 0, 0), (0.5, 0.5, 0.5), 0.01 * r, 7)
    b = scenes.lattice_cube(r, (0, 0, 0), (0.5, 0.5, 0.5), 0.01 * r, 7)
    c = scenes.lattice_cube(r, (0, 0, 0), (0.5, 0.5, 0.5), 0.01 * r, 8)
    z = scenes.lattice_cube(r, (0, 0, 0), (0.5, 0.5, 0.5))
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    dev = np.abs(a[:, :3].astype(np.float64) - z[:, :3])
    assert dev.max() <= 0.01 * r * (1 + 1e-5) and dev.max() > 0.005 * r
    assert abs(np.mean(a[:, :3] - z[:, :3])) < 0.001 * r


def test_cdm_round_trip_and_errors():
    fld = scenes.unit_box_cdm_field()
    res, sdf, vol = pbf_host.parse_cdm(scenes.cdm_text(fld))
    assert res == (20, 20, 20) and sdf.size == 62181
    assert np.array_equal(sdf, fld.sdf) and np.array_equal(vol, fld.vol)
    for bad in ["", "20 20 20\n1\n0.5 \n1\n0.5 \n", "x y z\n" + "\n".join(scenes.cdm_text(fld).split("\n")[1:])]:
        with pytest.raises(pbf_host.PBFError, match="Invalid Boundary Model File!"):
            pbf_host.parse_cdm(bad)


CDM = "/root/reference/public/model/domain_boundary.cdm"


@pytest.mark.skipif(not os.path.exists(CDM), reason="reference data not mounted")
def test_parser_on_reference_domain_boundary_file():
    # The reference's own unit-box field: SDF = min distance to the faces.
    res, sdf, vol = pbf_host.parse_cdm(open(CDM, "rb").read())
    assert res == (20, 20, 20) and sdf.size == 62181 == vol.size
    pts = scenes.cubic_nodes(res)
    box = np.concatenate([np.minimum.reduce([x, 1 - x, y, 1 - y, z, 1 - z]).ravel() for (x, y, z) in pts])
    assert np.abs(sdf - box).max() < 4e-7


def test_volume_map_builder_box():
    h, r = 0.1, 0.025
    res, lo, hi = (32, 32, 32), (0.0, 0.0, 0.0), (1.6, 1.6, 1.6)
    sv = pbf_host.build_volume_map(0, res, lo, hi, h, None, r)
    nodes = np.stack(np.meshgrid(*[np.linspace(lo[a], hi[a], res[a] + 1) for a in range(3)], indexing="ij"), -1)
    nodes = nodes.transpose(2, 1, 0, 3).reshape(-1, 3)
    box = np.minimum.reduce([nodes[:, 0], 1.6 - nodes[:, 0], nodes[:, 1], 1.6 - nodes[:, 1], nodes[:, 2],
                             1.6 - nodes[:, 2]])
    np.testing.assert_allclose(sv[:, 0], box, atol=1e-6)
    assert np.all(sv[box >= h, 1] == 0)
    # planar wall away from edges: V_b = 0.9 int_{z > d} W / W(max(d + r/2, 2r)),
    # with the poly6 mass of a slab: int over the disc = pi kp (h^2 - z^2)^4 / 4.
    kp = 315 / (64 * math.pi * h ** 9)
    for d in (0.0, 0.05):
        z = np.linspace(d, h, 20001)
        slab = np.trapezoid(math.pi * kp * (h * h - z * z) ** 4 / 4, z)
        s = max(d + r / 2, 2 * r)
        want = 0.9 * slab / (kp * (h * h - s * s) ** 3)
        sel = (np.abs(box - d) < 1e-9) & (np.abs(nodes[:, 0] - 0.8) < 1e-9) & (np.abs(nodes[:, 2] - 0.8) < 1e-9)
        got = sv[sel & (nodes[:, 1] < 0.8), 1]
        assert got.size == 1
        assert got[0] == pytest.approx(want, rel=0.06)  # 16^3 midpoint quadrature


def test_volume_map_builder_bounded_near_kernel_edge():
    """V_b = 0.9 int W / W(d + r/2) has a pole at d = h - r/2; the builder caps
    it at the deep-solid value 0.9 / W(2r).  Nodes 1.25 mm apart straddle the
    pole on a column far from the side walls."""
    h, r = 0.1, 0.025
    res, lo, hi = (2, 1280, 2), (0.0, 0.0, 0.0), (1.6, 1.6, 1.6)
    sv = pbf_host.build_volume_map(0, res, lo, hi, h, None, r)
    kp = 315 / (64 * math.pi * h ** 9)
    vcap = 0.9 / (kp * (h * h - 4 * r * r) ** 3)
    assert np.all(np.isfinite(sv))
    assert sv[:, 1].max() <= vcap * (1 + 1e-6)
    col = sv.reshape(3, 1281, 3, 2)[1, :, 1]          # x = z = 0.8, y = 0 .. 1.6
    d = col[:, 0]
    near = (d > h - r / 2 - 0.004) & (d < h - r / 2)
    assert near.sum() >= 2 and np.all(col[near, 1] <= vcap * (1 + 1e-6))
    assert np.all(col[d >= h, 1] == 0.0) and np.all(col[d <= 0.05, 1] > 0.0)


def test_volume_map_builder_obstacle_sdf():
    res, lo, hi = (8, 8, 8), (0.0, 0.0, 0.0), (1.0, 1.0, 1.0)
    shape = [0.5, 0.5, 0.5, 0.2, 0.5, 0.5, 0.5, 0.3, 0.08]
    sv = pbf_host.build_volume_map(1, res, lo, hi, 0.05, shape, 0.0125)
    nodes = np.stack(np.meshgrid(*[np.linspace(0, 1, 9)] * 3, indexing="ij"), -1).transpose(2, 1, 0, 3).reshape(-1, 3)
    box = np.minimum.reduce([nodes[:, 0], 1 - nodes[:, 0], nodes[:, 1], 1 - nodes[:, 1], nodes[:, 2], 1 - nodes[:, 2]])
    sph = np.linalg.norm(nodes - 0.5, axis=1) - 0.2
    q = np.hypot(nodes[:, 0] - 0.5, nodes[:, 2] - 0.5) - 0.3
    tor = np.hypot(q, nodes[:, 1] - 0.5) - 0.08
    np.testing.assert_allclose(sv[:, 0], np.minimum.reduce([box, sph, tor]), atol=1e-6)


def test_cubic_field_generator_and_cdm_writer():
    """The 32-node generator samples the volume-map definition at the .cdm
    node sets: its corner nodes are the trilinear map's nodes (same
    evaluator, bit-identical), every SDF node is the closed-form box
    distance, and .cdm text written and parsed again is bit-identical."""
    res, lo, hi, h, r = (10, 8, 6), (0.0, 0.0, 0.0), (1.0, 0.8, 0.6), 0.1, 0.025
    sdf, vol = pbf_host.build_cubic_field(0, res, lo, hi, h, None, r)
    n0 = 11 * 9 * 7
    assert sdf.size == n0 + 2 * 10 * 9 * 7 + 2 * 8 * 11 * 7 + 2 * 6 * 11 * 9
    tri = pbf_host.build_volume_map(0, res, lo, hi, h, None, r, where="host")
    assert np.array_equal(sdf[:n0], tri[:, 0]) and np.array_equal(vol[:n0], tri[:, 1])
    pts = scenes.cubic_nodes(res)
    box = np.concatenate([np.minimum.reduce([x * 1.0, 1.0 - x * 1.0, y * 0.8, 0.8 - y * 0.8, z * 0.6,
                                             0.6 - z * 0.6]).ravel() for (x, y, z) in pts])
    np.testing.assert_allclose(sdf, box, atol=2e-7)
    text = pbf_host.write_cdm(res, sdf, vol)
    assert text.startswith("10 8 6\r\n%d\r\n" % sdf.size)
    res2, s2, v2 = pbf_host.parse_cdm(text)
    assert res2 == res and np.array_equal(s2, sdf) and np.array_equal(v2, vol)
    with pytest.raises(pbf_host.PBFError):
        pbf_host.write_cdm(res, sdf[:-1], vol[:-1])


def test_host_entry_points_reject_non_finite_and_oversized_input():
    # GPU-free entry points validate what the reference's JS loops would turn
    # into endless or undefined loops: NaN / inf extents, spacings whose
    # lattice exceeds any 32-bit count, and non-finite map domains.
    import ctypes as C
    lib = pbf_host.load_library()
    d3 = lambda v: (C.c_double * 3)(*v)  # noqa: E731
    cnt = C.c_uint32()
    nan, inf = float("nan"), float("inf")
    cube = lambda r, mn, mx: lib.pbf_voxelize_cube_host(r, d3(mn), d3(mx), 0.0, 1, 0, None, 0, C.byref(cnt))  # noqa
    assert cube(0.01, (0, 0, 0), (nan, 1, 1)) == pbf_host.PBF_E_INVALID
    assert cube(0.01, (0, 0, 0), (inf, 1, 1)) == pbf_host.PBF_E_INVALID
    assert cube(inf, (0, 0, 0), (1, 1, 1)) == pbf_host.PBF_E_INVALID
    assert cube(1e-9, (0, 0, 0), (1, 1, 1)) == pbf_host.PBF_E_CAPACITY
    assert cube(0.01, (0, 0, 0), (-1, 1, 1)) == pbf_host.PBF_OK and cnt.value == 0
    sph = lambda r, c, R: lib.pbf_voxelize_sphere_host(r, d3(c), R, 0.0, 1, 0, None, 0, C.byref(cnt))  # noqa
    assert sph(0.01, (0, 0, 0), inf) == pbf_host.PBF_E_INVALID
    assert sph(0.01, (nan, 0, 0), 0.1) == pbf_host.PBF_E_INVALID
    assert sph(1e-6, (0, 0, 0), 1.0) == pbf_host.PBF_E_CAPACITY
    assert sph(0.01, (0, 0, 0), -0.5) == pbf_host.PBF_OK and cnt.value == 0
    with pytest.raises(pbf_host.PBFError):
        pbf_host.build_volume_map(0, (4, 4, 4), (0, 0, 0), (nan, 1, 1), 0.1, None, 0.025, where="host")
    with pytest.raises(pbf_host.PBFError):
        pbf_host.build_volume_map(0, (4, 4, 4), (0, 0, 0), (1, 1, 1), inf, None, 0.025, where="host")
    with pytest.raises(pbf_host.PBFError):
        pbf_host.build_volume_map(1, (4, 4, 4), (0, 0, 0), (1, 1, 1), 0.1, (0.5,) * 8 + (nan,), 0.025, where="host")
    # resolutions beyond 65536 per axis (the builders' and samplers' index range)
    big = (C.c_int32 * 3)(2147483647, 1, 1)
    length = C.c_size_t()
    assert lib.pbf_build_cubic_field(0, None, big, d3((0, 0, 0)), d3((1, 1, 1)), 0.1, 0.025, None, None, 0,
                                     C.byref(length)) == pbf_host.PBF_E_INVALID
    one = (C.c_float * 1)()
    assert lib.pbf_write_cdm((C.c_int32 * 3)(65537, 1, 1), one, one, 1, None, 0, None) == pbf_host.PBF_E_INVALID
    assert lib.pbf_build_volume_map(0, None, big, d3((0, 0, 0)), d3((1, 1, 1)), 0.1, 0.025, one) == pbf_host.PBF_E_INVALID


def test_create_rejects_bad_parameters_with_a_message():
    # Validation runs before any device call, so it is checked here; the
    # message of a failed create is pbf_last_error(NULL).  A dense grid whose
    # cell count passes 2^31 (4096^3 cells here, which a 32-bit product would
    # wrap to 0) and non-finite sizes are refused instead of indexed.
    nan, inf = float("nan"), float("inf")
    cases = [
        (dict(particle_radius=nan, kernel_radius=0.1), "particle_radius must be finite"),
        (dict(particle_radius=0.025, kernel_radius=inf), "must be finite"),
        (dict(particle_radius=0.025, kernel_radius=0.1, domain_hi=(nan, 1, 1)), "scene domain must be finite"),
        (dict(particle_radius=0.025, kernel_radius=0.1, domain_hi=(0, 1, 1)), "hi > lo"),
        (dict(particle_radius=0.025, kernel_radius=0.1, domain_hi=(409.6, 409.6, 409.6)), "too many grid cells"),
        (dict(particle_radius=0.025, kernel_radius=1e-12, domain_hi=(1, 1, 1), grid_mode=1, hash_log2=20),
         "too many grid cells"),
        (dict(particle_radius=0.025, kernel_radius=0.1, max_neighbors=70000), "max_neighbors"),
        (dict(particle_radius=0.025, kernel_radius=0.1, capacity=1 << 28), "capacity above"),
    ]
    for kw, msg in cases:
        kw.setdefault("capacity", 1024)
        with pytest.raises(pbf_host.PBFError, match=msg):
            pbf_host.PBF(**kw)


def test_bench_refuses_more_than_one_gpu():
    # the engine is single-GPU: bench.py --gpus N > 1 exits with a message
    # before touching a device
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr and r.stdout == ""


def _brute_distinct(dims, hash_log2):
    # every cell's 27-cell neighbourhood (inside the grid), bucket = row-major
    # index mod 2^k; distinct iff no neighbourhood repeats a bucket
    D0, D1, D2 = dims
    x, y, z = np.meshgrid(np.arange(D0), np.arange(D1), np.arange(D2), indexing="ij")
    x, y, z = x.ravel(), y.ravel(), z.ravel()
    buckets = []
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                nx, ny, nz = x + dx, y + dy, z + dz
                ok = (nx >= 0) & (ny >= 0) & (nz >= 0) & (nx < D0) & (ny < D1) & (nz < D2)
                b = ((nx * D1 + ny) * D2 + nz) & ((1 << hash_log2) - 1)
                buckets.append(np.where(ok, b, -1 - np.arange(27)[len(buckets)]))
    B = np.sort(np.stack(buckets, 1), 1)
    return not np.any((B[:, 1:] == B[:, :-1]) & (B[:, 1:] >= 0))


@pytest.mark.parametrize("dims,hash_log2", [
    ((32, 32, 32), 11), ((32, 32, 32), 12), ((32, 32, 32), 10), ((20, 20, 20), 9), ((20, 20, 20), 6),
    ((16, 16, 16), 9), ((16, 16, 16), 10), ((1, 8, 64), 9), ((3, 5, 7), 4), ((40, 24, 16), 10),
    ((12, 12, 12), 8), ((2, 2, 2), 4), ((64, 64, 64), 13)])
def test_hash_bucket_distinctness_matches_enumeration(dims, hash_log2):
    # The pruned list walk is taken only when this predicate holds; a wrong
    # predicate visits a repeated bucket twice and lists neighbours twice.
    # 32^3 / 2^11: cells x-1 and x+1 are 2 * 1024 apart in the row-major
    # index, one table length.
    assert pbf_host.hash_buckets_distinct(dims, hash_log2) == _brute_distinct(dims, hash_log2), (dims, hash_log2)


def test_hash_bucket_collision_cases_named():
    assert not pbf_host.hash_buckets_distinct((32, 32, 32), 11)
    assert not pbf_host.hash_buckets_distinct((64, 64, 64), 13)
    assert pbf_host.hash_buckets_distinct((20, 20, 20), 9)
    # 20^3 in 2^6 buckets: 400 and 20 are 16 and 20 mod 64, and no
    # 16 a + 20 b + c with a, b, c in [-2, 2] but zero is a multiple of 64
    assert pbf_host.hash_buckets_distinct((20, 20, 20), 6)
    assert not pbf_host.hash_buckets_distinct((20, 20, 20), 4)   # 400 = 25 * 16
