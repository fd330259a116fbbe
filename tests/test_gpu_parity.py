"""GPU parity: libpbf's HIP path (through the C ABI) against the CPU oracle.

Bars: cell counts, scan offsets and neighbour sets bit-exact; one-step
positions and velocities within 1e-5 (norm-relative, float32); the density
error statistics of a 200-step run within 2 %.  The engine and the oracle
share one numerics specification (same summation order, no contraction), so
most comparisons below are in fact exact; they are asserted as such.
"""
import numpy as np
import pytest

import oracle
import scenes
from pairs import gpu_fields, make_oracle, oracle_fields, rel_err, ulp_diff
from pbf_host import PBF_FLAG_NO_GRAPH, PBFError, check_device_sqrt

pytestmark = pytest.mark.gpu


def test_device_pair_sqrt_is_correctly_rounded():
    """The pair terms' square root (hardware reciprocal square root plus one
    Newton step) equals the correctly rounded one -- the oracle's sqrtf --
    on every float in [2^-100, FLT_MAX]; the kernels' bit-exact results
    rest on it."""
    print("[sqrt] exhaustive device check", flush=True)
    bad, n = check_device_sqrt()
    assert n == 0x7F7FFFFF - 0x0D800000 + 1
    assert bad == 0


def small_scenes():
    return {
        "dam_break_small": scenes.dam_break(16, 16, 8, map_res=32, name="dam_break_small"),
        "dam_break_physics": scenes.dam_break_physics(16, 16, 8, map_res=32),
        "ref_cube_drop": scenes.ref_cube_drop(),
        "ref_boundary_cubic": scenes.ref_boundary(scenes.unit_box_cdm_field()),
        "hash_obstacle": scenes.obstacle_drop(12, 16, 12, map_res=48, hash_log2=12, name="hash_obstacle"),
    }


SCENES = small_scenes()


@pytest.mark.parametrize("name", list(SCENES))
def test_search_bit_exact(name):
    sc = SCENES[name]
    print(f"[search] {name} n={sc.n}", flush=True)
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    dt = sc.substep_dt()
    sim.search(dt)
    o.search(dt)
    gc, go = sim.read_cells()
    oc, oo = o.cells()
    assert gc.sum() == sc.n
    assert np.array_equal(gc, oc), "cell counts differ"
    assert np.array_equal(go, oo), "scan offsets differ"
    assert np.array_equal(sim.read_field("predicted")[:, :3], o.get(oracle.PREDICTED)[:, :3])
    goff, gids = sim.read_neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff, ooff), "neighbour counts differ"
    assert np.array_equal(gids, oids), "neighbour lists differ (set or walk order)"
    sim.close()


@pytest.mark.parametrize("name", list(SCENES))
def test_one_step_parity(name):
    sc = SCENES[name]
    print(f"[one step] {name} n={sc.n}", flush=True)
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    sim.step(sc.dt, sc.substeps, sc.iterations)
    o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    assert rel_err(g["position"][:, :3], c["position"][:, :3]) <= 1e-5
    assert rel_err(g["velocity"][:, :3], c["velocity"][:, :3]) <= 1e-5
    assert rel_err(g["position"][:, 3], c["position"][:, 3]) <= 1e-5
    for k in ("accel", "lambda", "omega", "normal"):
        assert rel_err(g[k], c[k]) <= 1e-5, k
    sim.close()


@pytest.mark.parametrize("name", ["dam_break_small", "ref_cube_drop", "ref_boundary_cubic", "hash_obstacle"])
def test_one_step_bit_exact(name):
    sc = SCENES[name]
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    sim.step(sc.dt, sc.substeps, sc.iterations)
    o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    for k in g:
        assert ulp_diff(g[k], c[k]) == 0, (k, ulp_diff(g[k], c[k]))
    sim.close()


@pytest.mark.parametrize("slabs", [1, 2, 4])
@pytest.mark.parametrize("name", ["dam_break_physics", "hash_obstacle"])
def test_cell_slabs_exact_and_sets_independent(name, slabs):
    """Each z-slab setting (a cell's particles ordered by slab, then id):
    lists, order included, and one step bit-exact against the oracle with
    the same setting; the neighbour sets equal those of one slab per cell
    (the build skips only candidates out of reach), and one step agrees with
    one slab per cell to rounding: positions within 1e-5; velocities, which
    are position differences over the substep dt, within 5e-5 (measured:
    1.3e-5 on the 2,048-particle scene between one and two slabs -- the
    spread any change of the order of a cell's terms gives, as the
    reference's atomic ranks inside a cell do from run to run)."""
    sc = SCENES[name]
    print(f"[slabs] {name} S={slabs} n={sc.n}", flush=True)
    sim = scenes.make_engine(sc, cell_slabs=slabs)
    one = scenes.make_engine(sc, cell_slabs=1)
    o = make_oracle(sc, cell_slabs=slabs)
    dt = sc.substep_dt()
    sim.search(dt)
    one.search(dt)
    o.search(dt)
    goff, gids = sim.read_neighbors()
    ooff, oids = o.neighbors()
    assert np.array_equal(goff, ooff) and np.array_equal(gids, oids)
    uoff, uids = one.read_neighbors()
    assert np.array_equal(goff, uoff)
    for i in range(sc.n):
        assert set(gids[goff[i]:goff[i + 1]].tolist()) == set(uids[uoff[i]:uoff[i + 1]].tolist()), i
    sim.close()
    one.close()
    sim = scenes.make_engine(sc, cell_slabs=slabs)
    one = scenes.make_engine(sc, cell_slabs=1)
    o = make_oracle(sc, cell_slabs=slabs)
    sim.step(sc.dt, sc.substeps, sc.iterations)
    one.step(sc.dt, sc.substeps, sc.iterations)
    o.step(sc.dt, sc.substeps, sc.iterations)
    g, c, u = gpu_fields(sim), oracle_fields(o), gpu_fields(one)
    for k in g:
        assert ulp_diff(g[k], c[k]) == 0, (k, ulp_diff(g[k], c[k]))
    ep, ev = rel_err(g["position"][:, :3], u["position"][:, :3]), rel_err(g["velocity"][:, :3], u["velocity"][:, :3])
    print(f"  vs one slab per cell: position {ep:.3g}, velocity {ev:.3g}")
    assert ep <= 1e-5
    assert ev <= 5e-5
    sim.close()
    one.close()


def test_many_steps_stay_exact():
    sc = scenes.dam_break_physics(16, 16, 8, map_res=32)
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    for s in range(30):
        sim.step(sc.dt, sc.substeps, sc.iterations)
        o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    assert rel_err(g["position"][:, :3], c["position"][:, :3]) <= 1e-5
    assert rel_err(g["velocity"][:, :3], c["velocity"][:, :3]) <= 1e-5
    print("ulp after 30 steps:", {k: ulp_diff(g[k], c[k]) for k in g})
    sim.close()


def test_density_error_fp64_matches_oracle():
    sc = scenes.dam_break(16, 16, 8, map_res=32, name="diag")
    sim = scenes.make_engine(sc)
    o = make_oracle(sc)
    for _ in range(5):
        sim.step(sc.dt, sc.substeps, sc.iterations)
        o.step(sc.dt, sc.substeps, sc.iterations)
    gm, gx = sim.density_error()
    om, ox = o.density_error()
    assert gm == pytest.approx(om, rel=1e-12, abs=1e-18)
    assert gx == pytest.approx(ox, rel=1e-12, abs=1e-18)


def test_200_step_density_statistics_dam_break_16k():
    sc = scenes.dam_break_16k()
    print(f"[200 steps] n={sc.n}", flush=True)
    sim = scenes.make_engine(sc)
    sim.record_density_error(True)
    o = make_oracle(sc)
    ohist = []
    for s in range(sc.steps):
        sim.step(sc.dt, sc.substeps, sc.iterations)
        for _ in range(sc.substeps):
            o.substep(sc.substep_dt(), sc.iterations)
            ohist.append(o.density_error())
    ghist = sim.density_history()
    ohist = np.array(ohist)
    assert ghist.shape == ohist.shape
    gmean, omean = ghist[:, 0].mean(), ohist[:, 0].mean()
    gmax, omax = ghist[:, 1].max(), ohist[:, 1].max()
    print(f"mean of mean error gpu {gmean:.6g} cpu {omean:.6g}; max gpu {gmax:.6g} cpu {omax:.6g}")
    assert gmean == pytest.approx(omean, rel=0.02)
    assert gmax == pytest.approx(omax, rel=0.02)


def test_batched_scenes_match_independent_runs():
    specs = scenes.batch_specs(6, 300, 3000, seed=99)
    scs = [scenes.dam_break(*d, map_res=24, seed=1234 + k, name=f"b{k}") for k, d in enumerate(specs)]
    print("[batch] sizes", [s.n for s in scs], flush=True)
    sim = scenes.make_batch_engine(scs)
    for _ in range(3):
        sim.step(scs[0].dt, scs[0].substeps, scs[0].iterations)
    gp, gv = sim.read_positions(), sim.read_velocities()
    base = 0
    for s in scs:
        o = make_oracle(s)
        for _ in range(3):
            o.step(s.dt, s.substeps, s.iterations)
        assert np.array_equal(gp[base:base + s.n], o.get(oracle.POSITION)), s.name
        assert np.array_equal(gv[base:base + s.n, :3], o.get(oracle.VELOCITY)[:, :3]), s.name
        base += s.n
    per = [sim.density_error(k) for k in range(len(scs))]
    assert all(np.isfinite(m) for m, _ in per)
    sim.close()


def test_graph_replay_equals_direct_launches():
    sc = scenes.dam_break_physics(16, 16, 8, map_res=32)
    a = scenes.make_engine(sc)
    b = scenes.make_engine(sc, flags=PBF_FLAG_NO_GRAPH)
    for s in range(6):
        a.step(sc.dt, sc.substeps, sc.iterations)
        b.step(sc.dt, sc.substeps, sc.iterations)
        # eager work and a synchronisation between graph replays
        pa, pb = a.read_positions(), b.read_positions()
        assert np.array_equal(pa, pb), s
    assert np.array_equal(a.read_velocities(), b.read_velocities())
    a.close()
    b.close()


def test_reference_error_messages():
    sc = SCENES["dam_break_small"]
    sim = scenes.make_engine(sc)
    with pytest.raises(PBFError, match="Illegal length of Particle Position Buffer!"):
        sim.setParticlePosition(np.zeros(10, np.float32), count=3)
    with pytest.raises(PBFError, match="To Many Particles!"):
        sim.setParticlePosition(np.zeros((sc.n + 1, 4), np.float32))
    with pytest.raises(PBFError, match="Invalid Boundary Model File!"):
        sim.setBoundaryData("20 20 20\n1\n0.5 \n1\n0.5 \n")
    sim.close()


def test_cell_slabs_setting_is_checked():
    from pbf_host import PBF
    for bad in (3, 8):
        with pytest.raises(PBFError, match="cell_slabs must be 1, 2 or 4"):
            PBF(cell_slabs=bad)


def test_reference_api_flow_and_cdm_text():
    # Controller.setSceneConfig flow (controller.ts:194-248): reset, voxelize,
    # boundary text, setParticlePosition, run.
    from pbf_host import PBF
    sim = PBF()
    sim.reset()
    n = sim.voxelizeCube((0.15, 0.35, 0.15), (0.65, 0.85, 0.65))
    assert n == 41 ** 3
    sim.setBoundaryData(scenes.cdm_text(scenes.unit_box_cdm_field()))
    sim.setConfig(scenes.REF_CONFIG)
    sim.setParticlePosition()
    sim.start()
    sim.run()
    sim.sync()
    p = sim.read_positions()
    assert p.shape == (n, 4) and np.isfinite(p).all() and (p[:, 3] > 0).all()
    sim.close()


def test_neighbour_overflow_is_reported():
    sc = scenes.dam_break(8, 8, 8, map_res=16, name="overflow")
    sim = scenes.make_engine(sc, max_neighbors=8)
    sim.step(sc.dt, 1, 1)
    with pytest.raises(PBFError, match="neighbour overflow"):
        sim.sync()
    sim.close()


def test_diverged_particle_is_reported():
    """A non-finite predicted position (a NaN velocity here) is reported at
    the next sync; the particle is binned by its slot rather than into one
    shared cell, and the other particles' steps stay finite."""
    sc = scenes.dam_break(8, 8, 8, map_res=16, name="diverged")
    sim = scenes.make_engine(sc)
    v = np.zeros((sc.n, 3), np.float32)
    v[[3, 100, 511]] = np.nan
    sim.write_field("velocity", v)
    sim.step(sc.dt, 1, 1)
    with pytest.raises(PBFError, match="non-finite particle position"):
        sim.sync()
    sim.sync()  # the flag was cleared when it was reported
    sim.close()


def test_generated_obstacle_cdm_runs_at_parity():
    """A reference-format field built by pbf_build_cubic_field (sphere + torus
    in the unit box, 24^3 elements, the reference's own volume definition),
    passed as .cdm text through setBoundaryData, steps bit-identically to the
    oracle's cubic sampler."""
    from pbf_host import VOLUME_REFERENCE, build_cubic_field, parse_cdm, write_cdm
    res = (24, 24, 24)
    shape = [0.5, 0.45, 0.5, 0.12, 0.5, 0.45, 0.5, 0.25, 0.05]
    sdf, vol = build_cubic_field(VOLUME_REFERENCE | 1, res, (0, 0, 0), (1, 1, 1), scenes.REF_H, shape)
    text = write_cdm(res, sdf, vol)
    _, s2, v2 = parse_cdm(text)
    fld = scenes.Field("cubic", res, (0.0, 0.0, 0.0), (1.0, 1.0, 1.0), sdf=s2, vol=v2)
    sc = scenes._ref_scene("ref_obstacle_cdm", scenes.lattice_cube(scenes.REF_R, (0.3, 0.62, 0.3), (0.7, 0.9, 0.7)),
                           fld)
    print(f"[generated cdm] n={sc.n}", flush=True)
    sim = scenes.make_engine(sc)
    sim.setBoundaryData(text)
    o = make_oracle(sc)
    for _ in range(3):
        sim.step(sc.dt, sc.substeps, sc.iterations)
        o.step(sc.dt, sc.substeps, sc.iterations)
    g, c = gpu_fields(sim), oracle_fields(o)
    for k in g:
        assert ulp_diff(g[k], c[k]) == 0, (k, ulp_diff(g[k], c[k]))
    sim.close()


def test_graph_cache_eviction_with_many_substep_shapes():
    """More substep shapes than the context caches executable graphs for (16),
    launched back to back without a host synchronisation: every eviction
    waits for the stream, and the results equal direct launches."""
    sc = scenes.dam_break(12, 12, 8, map_res=24, name="graph_cycle")
    a = scenes.make_engine(sc)
    b = scenes.make_engine(sc, flags=PBF_FLAG_NO_GRAPH)
    dt = sc.substep_dt()
    shapes = [(dt * (1.0 + 0.01 * (k % 5)), 1 + k % 7) for k in range(24)] * 2
    print(f"[graph cycle] {len(set(shapes))} shapes, {len(shapes)} substeps", flush=True)
    for sdt, it in shapes:
        a.step(sdt, 1, it)
        b.step(sdt, 1, it)
    assert np.array_equal(a.read_positions(), b.read_positions())
    assert np.array_equal(a.read_velocities(), b.read_velocities())
    a.close()
    b.close()
