"""SPH.voxelizeMesh (SPH.ts:151-187) restated: the 2r lattice over the mesh's
world bounding box, kept where the ray along local +z first hits a triangle
facing along the ray.  Closed forms pin it: a box mesh keeps exactly the
voxelizeCube lattice of its box; a world matrix moves the lattice with it;
reversed winding keeps nothing; a fine icosphere keeps the voxelizeSphere
lattice up to the facets.  (The reference's ray caster, three-mesh-bvh, is not
in this image: on-face origins are a convention, documented in the header.)
"""
import numpy as np
import pytest

import pbf_host
from pbf_host import PBFError


def box_mesh(lo, hi):
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    # outward-facing (counter-clockwise seen from outside) triangles
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4],
                  [2, 3, 7], [2, 7, 6], [1, 2, 6], [1, 6, 5], [0, 4, 7], [0, 7, 3]], np.uint32)
    return v, t


def icosphere(level):
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    v = [np.array(p, float) / np.linalg.norm(p) for p in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(level):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]
        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    return np.array(v, np.float32), np.array(f, np.uint32)


def cube_lattice(r, lo, hi):
    out = np.zeros((100000, 4), np.float32)
    import ctypes as C
    cnt = C.c_uint32()
    lib = pbf_host.load_library()
    rc = lib.pbf_voxelize_cube_host(r, pbf_host._d3(lo), pbf_host._d3(hi), 0.0, 0, 0, pbf_host._fptr(out),
                                    out.shape[0], C.byref(cnt))
    assert rc == 0
    return out[: cnt.value]


def test_box_mesh_keeps_the_cube_lattice():
    r = 0.025
    lo, hi = (0.1, 0.2, 0.3), (0.63, 0.71, 0.58)
    v, t = box_mesh(lo, hi)
    got = pbf_host.voxelize_mesh(r, v, t)
    # the world box is the float32 vertex box
    want = cube_lattice(r, tuple(float(x) for x in v.min(0)), tuple(float(x) for x in v.max(0)))
    assert got.shape == want.shape and np.array_equal(got, want)


def test_world_matrix_moves_the_lattice():
    r = 0.05
    v, t = box_mesh((0.0, 0.0, 0.0), (1.0, 0.6, 0.4))
    base = pbf_host.voxelize_mesh(r, v, t)
    M = np.eye(4)
    M[:3, 3] = (2.0, -1.0, 0.5)                 # translation
    got = pbf_host.voxelize_mesh(r, v, t, M.T.reshape(-1))  # column-major
    assert got.shape == base.shape
    np.testing.assert_allclose(got[:, :3], base[:, :3] + np.array([2.0, -1.0, 0.5], np.float32), atol=1e-6)
    # a quarter turn about y: the local +z ray becomes world +x; the lattice of
    # the rotated box (extent 0.4 x 0.6 x 1.0) is kept whole
    R = np.array([[0, 0, 1, 0], [0, 1, 0, 0], [-1, 0, 0, 0], [0, 0, 0, 1]], float)
    rot = pbf_host.voxelize_mesh(r, v, t, R.T.reshape(-1))
    wv = (R[:3, :3] @ v.T.astype(float)).T
    want = cube_lattice(r, tuple(wv.min(0)), tuple(wv.max(0)))
    assert rot.shape[0] == want.shape[0]


def test_reversed_winding_keeps_nothing():
    v, t = box_mesh((0.0, 0.0, 0.0), (0.5, 0.5, 0.5))
    assert pbf_host.voxelize_mesh(0.025, v, t[:, ::-1].copy()).shape[0] == 0


def test_icosphere_matches_the_sphere_lattice():
    r, R = 0.02, 0.4
    v, t = icosphere(4)
    got = pbf_host.voxelize_mesh(r, v * R, t)
    p = got[:, :3].astype(np.float64)
    assert len(p) > 0 and np.all(np.linalg.norm(p, axis=1) <= R + 1e-6)
    ideal = 4.0 / 3.0 * np.pi * R ** 3 / (2 * r) ** 3
    assert abs(len(p) - ideal) / ideal < 0.03   # facets of a level-4 icosphere lose < 3 % of the volume


def test_bad_meshes_are_rejected():
    v, t = box_mesh((0, 0, 0), (1, 1, 1))
    bad = t.copy()
    bad[0, 0] = 99
    with pytest.raises(PBFError):
        pbf_host.voxelize_mesh(0.05, v, bad)
    with pytest.raises(PBFError):
        pbf_host.voxelize_mesh(0.05, v, t, np.zeros(16))   # singular matrix


@pytest.mark.gpu
def test_context_voxelize_mesh_stages_the_same_particles():
    from pbf_host import PBF
    v, t = box_mesh((0.1, 0.1, 0.1), (0.5, 0.4, 0.3))
    sim = PBF(particle_radius=0.01, kernel_radius=0.04, capacity=1 << 16)
    n = sim.voxelizeMesh(v, t, jitter=0.001, seed=3)
    host = pbf_host.voxelize_mesh(0.01, v, t, jitter=0.001, seed=3)
    assert n == host.shape[0] == sim.particleCount
    sim.setParticlePosition()
    np.testing.assert_array_equal(sim.read_positions()[:, :3], host[:, :3])
