"""Edge cases of the GPU path against the oracle: empty and single-particle
systems, ragged counts around wavefront / work-group / scan-block sizes,
particles on cell faces and domain walls, crowded cells and cells too full
for the 8-bit z-slab words, a three-level scan (16.7 M cells), and the
capacity limit.  All comparisons are bit-exact."""
import numpy as np
import pytest

import scenes
from pairs import gpu_fields, make_oracle, oracle_fields, ulp_diff
from pbf_host import PBF, PBFError, hash_buckets_distinct

pytestmark = pytest.mark.gpu


def _scene(pos, L=1.6, name="edge", iters=3, field=True, r=0.025):
    h = 4 * r
    lo, hi = (0.0, 0.0, 0.0), (L, L, L)
    fld = None
    if field:
        res = (16, 16, 16)
        fld = scenes.Field("trilinear", res, lo, hi,
                           sdf_vol=scenes.build_volume_map(0, res, lo, hi, h, None, r))
    cfg = {"iteration": iters, "XSPH": 0.01, "vorticity": 0.01, "surfaceTension": 0.1, "gravity": 9.8}
    return scenes.Scene(name, r, h, lo, hi, np.asarray(pos, np.float32), fld, cfg, 1.0 / 60.0, 2, iters, 10)


def _bit_exact(sc, steps=2):
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    for _ in range(steps):
        sim.step(sc.dt, sc.substeps, sc.iterations)
        o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    for k in g:
        assert ulp_diff(g[k], c[k]) == 0, (sc.name, k, ulp_diff(g[k], c[k]))
    sim.close()


def test_empty_context():
    sim = PBF(particle_radius=0.025, kernel_radius=0.1, capacity=1024)
    sim.setParticlePosition(np.zeros((0, 4), np.float32), 0)
    sim.step(1.0 / 60.0, 2, 4)
    sim.sync()
    assert sim.particleCount == 0
    with pytest.raises(PBFError):
        sim.density_error()
    sim.close()


def test_single_particle_free_fall():
    sc = _scene([[0.8, 1.2, 0.8, 0.0]], field=False, name="single")
    _bit_exact(sc, steps=3)
    # 6 substeps of free fall: y = y0 - g dt^2 (1 + 2 + ... + 6)
    sim = scenes.make_engine(sc)
    for _ in range(3):
        sim.step(sc.dt, sc.substeps, sc.iterations)
    dt = sc.dt / sc.substeps
    assert sim.read_positions()[0, 1] == pytest.approx(1.2 - 9.8 * dt * dt * 21, abs=2e-6)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 257, 2047, 2049])
def test_ragged_counts(n):
    rng = np.random.default_rng(n)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.3, 0.3 + 0.05 * max(2.0, n ** (1 / 3)), (n, 3))
    print(f"[ragged] n={n}", flush=True)
    _bit_exact(_scene(pos, name=f"ragged{n}"))


def test_cell_faces_and_walls():
    h = 0.1
    g = np.arange(0.0, 1.6 + 1e-9, h / 2)            # every other point on a cell face
    X, Y, Z = np.meshgrid(g[:10], g[:12], g[:10], indexing="ij")
    pos = np.zeros((X.size, 4), np.float32)
    pos[:, 0], pos[:, 1], pos[:, 2] = X.ravel(), Y.ravel() + 0.2, Z.ravel()
    pos[0, :3] = (0.0, 0.0, 0.0)                      # on the domain corner
    pos[1, :3] = (1.6, 1.6, 1.6)
    _bit_exact(_scene(pos, name="faces"))


def test_wide_slabs_mix_table_formats():
    # Two x-slabs of 64 x 64 cells with 8 particles each: a neighbour one
    # slab over sits ~32768 slots away, so some wavefronts keep 32-bit
    # neighbour entries while a small block elsewhere uses 16-bit offsets.
    r = 0.025
    g = (np.arange(128) + 0.5) * 2 * r
    X, Y, Z = np.meshgrid(g[:4] + 0.4, g, g, indexing="ij")
    wide = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    X, Y, Z = np.meshgrid(g[:8] + 4.0, g[:8] + 4.0, g[:8] + 4.0, indexing="ij")
    small = np.stack([X.ravel(), Y.ravel(), Z.ravel()], -1)
    pts = np.concatenate([wide, small])
    pos = np.zeros((len(pts), 4), np.float32)
    pos[:, :3] = pts + np.random.default_rng(4).uniform(-0.002, 0.002, pts.shape)
    print(f"[wide slabs] n={len(pos)}", flush=True)
    sc = _scene(pos, L=6.4, name="slabs", iters=2, field=False, r=r)
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    sim.search(sc.substep_dt())
    o.search(sc.substep_dt())
    assert np.array_equal(sim.read_neighbors()[1], o.neighbors()[1])
    sim.close()
    _bit_exact(sc, steps=1)


@pytest.mark.parametrize("L,hash_log2,distinct", [(2.0, 9, True), (2.0, 4, False), (3.2, 11, False),
                                                   (3.2, 12, True)])
def test_hashed_grid_wraps_and_collisions(L, hash_log2, distinct):
    # 20^3 cells (L = 2) in 2^9 buckets: the 27 cells around a cell hash to
    # distinct buckets, but the index wraps the table inside some z-columns
    # (visited cell by cell) and far cells share buckets.  2^4 buckets: cells
    # one x-column apart (400 = 25 * 16) share a bucket, so the walk visits
    # each bucket at its first occurrence.  32^3 cells (L = 3.2) in 2^11
    # buckets: cells x - 1 and x + 1 are 2 * 1024 apart in the row-major index,
    # one table length -- the same bucket, visited once.
    assert hash_buckets_distinct((round(L / 0.1),) * 3, hash_log2) == distinct
    rng = np.random.default_rng(hash_log2)
    n = 6000 if L == 2.0 else 16000
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = rng.uniform(0.05, L - 0.05, (n, 3))
    sc = _scene(pos, L=L, name=f"hash{hash_log2}", field=False)
    sc.grid_mode, sc.hash_log2 = 1, hash_log2
    print(f"[hash wrap] {round(L / 0.1)}^3 cells, 2^{hash_log2} buckets", flush=True)
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    sim.search(sc.substep_dt())
    o.search(sc.substep_dt())
    gc, go = sim.read_cells()
    oc, oo = o.cells()
    assert np.array_equal(gc, oc) and np.array_equal(go, oo)
    goff, gids = sim.read_neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    # no particle lists a neighbour twice
    for q in range(0, n, 97):
        row = gids[goff[q]:goff[q + 1]]
        assert len(np.unique(row)) == len(row)
    sim.close()
    _bit_exact(sc, steps=1)


def test_crowded_cells():
    rng = np.random.default_rng(3)
    pos = np.zeros((600, 4), np.float32)
    pos[:, :3] = 0.8 + rng.uniform(-0.15, 0.15, (600, 3))  # ~90 neighbours in the middle
    sc = _scene(pos, name="crowded", iters=2)
    sim = scenes.make_engine(sc, max_neighbors=160)
    o = make_oracle(sc)
    sim.step(sc.dt, sc.substeps, sc.iterations)
    o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    for k in g:
        assert ulp_diff(g[k], c[k]) == 0, k


def test_three_level_scan():
    # 256^3 = 16.7 M cells: the scan recurses 16.7 M -> 8192 -> 4 block sums
    L = 25.6
    rng = np.random.default_rng(8)
    pos = np.zeros((5000, 4), np.float32)
    pos[:, :3] = rng.uniform(0.0, L, (5000, 3))
    pos[:100, :3] = L - rng.uniform(0.0, 0.3, (100, 3))  # the last cells too
    sc = _scene(pos, L=L, field=False, name="scan3")
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    sim.search(sc.substep_dt())
    o.search(sc.substep_dt())
    gc, go = sim.read_cells()
    oc, oo = o.cells()
    assert gc.size == 256 ** 3 and int(gc.sum()) == 5000
    assert np.array_equal(gc, oc) and np.array_equal(go, oo)
    assert np.array_equal(sim.read_neighbors()[1], o.neighbors()[1])


def test_capacity_limit():
    pos = np.zeros((1000, 4), np.float32)
    pos[:, :3] = np.random.default_rng(1).uniform(0.2, 0.6, (1000, 3))
    sim = PBF(particle_radius=0.025, kernel_radius=0.1, capacity=1000, domain_hi=(1.6, 1.6, 1.6))
    sim.setParticlePosition(pos)                      # exactly the capacity
    sim.step(1.0 / 60.0, 1, 2)
    sim.sync()
    with pytest.raises(PBFError, match="To Many Particles!"):
        sim.setParticlePosition(np.zeros((1001, 4), np.float32))
    sim.close()


def test_argument_limits():
    sim = PBF(particle_radius=0.025, kernel_radius=0.1, capacity=256)
    sim.setParticlePosition(np.zeros((4, 4), np.float32) + 0.5)
    with pytest.raises(PBFError, match="iterations must be <= 4096"):
        sim.step(1.0 / 60.0, 1, 5000)
    with pytest.raises(PBFError, match="dt must be > 0"):
        sim.step(0.0, 1, 2)
    with pytest.raises(PBFError, match="iteration must be <= 4096"):
        sim.setConfig({"iteration": 5000, "XSPH": 0.0, "vorticity": 0.0, "surfaceTension": 0.0, "gravity": 9.8})
    sim.step(1.0 / 60.0, 1, 2)      # the context is still usable
    sim.sync()
    sim.close()
    with pytest.raises(PBFError):
        PBF(particle_radius=0.025, kernel_radius=0.1, capacity=1 << 28)


def test_boundary_field_limits():
    # field resolutions beyond the samplers' 32-bit index range and
    # non-finite or empty field boxes are refused; the context keeps working
    sim = PBF(particle_radius=0.025, kernel_radius=0.1, capacity=256)
    sim.setParticlePosition(np.zeros((4, 4), np.float32) + 0.5)
    one = np.zeros(2 * 2 * 2 * 2, np.float32)
    with pytest.raises(PBFError, match="resolution"):
        sim.setVolumeMap((70000, 1, 1), (0, 0, 0), (1, 1, 1), one)
    with pytest.raises(PBFError, match="field box"):
        sim.setVolumeMap((1, 1, 1), (0, 0, 0), (float("nan"), 1, 1), one)
    with pytest.raises(PBFError, match="field box"):
        sim.setVolumeMap((1, 1, 1), (0, 0, 0), (0, 1, 1), one)
    sim.setVolumeMap((1, 1, 1), (0, 0, 0), (1, 1, 1), one)
    sim.step(1.0 / 60.0, 1, 2)
    sim.sync()
    sim.close()


@pytest.mark.parametrize("slabs", [4, 2])
def test_cells_past_the_slab_word_range(slabs):
    """Cells of 254, 255, 256 and 400 particles: the z-slab word counts to
    254 and marks fuller cells as untrimmed; their neighbours' lists (order
    included) and the counts still equal the oracle's, with crowded cells as
    the first, middle and last cell of a column range."""
    rng = np.random.default_rng(21)
    parts = []

    def cell(c, k):  # k particles strictly inside cell c (h = 0.1), clear of its faces
        parts.append(np.asarray(c, np.float32) * 0.1 + rng.uniform(0.003, 0.097, (k, 3)).astype(np.float32))

    cell((5, 5, 5), 400)
    cell((5, 5, 6), 256)   # the next cell of the same column
    cell((5, 5, 4), 60)
    cell((8, 5, 5), 255)
    cell((5, 8, 5), 254)
    cell((8, 8, 8), 256)
    cell((8, 8, 9), 40)
    parts.append(rng.uniform(0.1, 1.5, (3000, 3)).astype(np.float32))
    pts = np.concatenate(parts)
    pos = np.zeros((len(pts), 4), np.float32)
    pos[:, :3] = pts
    sc = _scene(pos, name="saturated", field=False)
    print(f"[saturated cells] n={sc.n} slabs={slabs}", flush=True)
    sim = scenes.make_engine(sc, max_neighbors=2000, cell_slabs=slabs)
    o = make_oracle(sc, cell_slabs=slabs)
    dt = sc.substep_dt()
    sim.search(dt)
    o.search(dt)
    gc, go = sim.read_cells()
    oc, oo = o.cells()
    assert gc.max() >= 400 and np.array_equal(gc, oc) and np.array_equal(go, oo)
    goff, gids = sim.read_neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    sim.close()


@pytest.mark.parametrize("seed", range(10))
def test_random_scenes_bit_exact(seed):
    """Seeded random scenes: particle count, box, clustering, z-slabs per
    cell (1, 2, 4) and dense or hashed grid (any table size that can hold
    the cells' rows), each stepped once on both sides, bit for bit."""
    rng = np.random.default_rng(1000 + seed)
    L = float(rng.choice([1.2, 1.6, 2.4]))
    n = int(rng.integers(1, 5000))
    pts = rng.uniform(0.02, L - 0.02, (n, 3))
    k = int(rng.integers(0, 4))  # a few dense clumps
    for c in range(k):
        m = min(n, int(rng.integers(50, 400)))
        centre = rng.uniform(0.2, L - 0.2, 3)
        pts[rng.choice(n, m, replace=False)] = centre + rng.normal(0.0, 0.03, (m, 3))
    pts = np.clip(pts, 0.001, L - 0.001)
    pos = np.zeros((n, 4), np.float32)
    pos[:, :3] = pts
    sc = _scene(pos, L=L, name=f"random{seed}", iters=int(rng.integers(1, 5)), field=bool(rng.integers(0, 2)))
    slabs = int(rng.choice([1, 2, 4]))
    if rng.integers(0, 2):
        sc.grid_mode, sc.hash_log2 = 1, int(rng.integers(6, 14))
    print(f"[random {seed}] n={n} L={L} slabs={slabs} grid={sc.grid_mode}/{sc.hash_log2}", flush=True)
    sim = scenes.make_engine(sc, max_neighbors=512, cell_slabs=slabs)
    o = make_oracle(sc, cell_slabs=slabs)
    sim.step(sc.dt, sc.substeps, sc.iterations)
    o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    for key in g:
        assert ulp_diff(g[key], c[key]) == 0, (key, ulp_diff(g[key], c[key]))
    goff, gids = sim.read_neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    sim.close()
