"""Launch tuning: the autotuner and explicit settings change time only.

Every work-group size of the neighbour walks and every neighbour-list
variant must give bit-identical results, the autotuner must leave the
simulation state as it found it, and bad settings are rejected.
"""
import numpy as np
import pytest

import scenes
from pairs import gpu_fields
from pbf_host import PBFError

pytestmark = pytest.mark.gpu

SC = scenes.dam_break_physics(16, 16, 8, map_res=32)
MULTI = [scenes.dam_break(10, 12, 8, map_res=24, seed=11, name="batch_a"),
         scenes.dam_break(16, 8, 12, map_res=24, seed=12, name="batch_b")]
HASH = scenes.obstacle_drop(12, 16, 12, map_res=48, hash_log2=12, name="hash_obstacle")


@pytest.mark.parametrize("sc", [SC, HASH], ids=["dense", "hash"])
@pytest.mark.parametrize("block,variant", [(64, 1), (128, 2), (192, 3), (256, 0), (64, 3)])
def test_every_setting_gives_identical_results(sc, block, variant):
    print(f"[tuning] {sc.name} block={block} variant={variant}", flush=True)
    ref = scenes.make_engine(sc)
    sim = scenes.make_engine(sc)
    sim.set_tuning(block, variant)
    for _ in range(2):
        ref.step(sc.dt, sc.substeps, sc.iterations)
        sim.step(sc.dt, sc.substeps, sc.iterations)
    a, b = gpu_fields(ref), gpu_fields(sim)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    ra, rb = ref.read_neighbors(), sim.read_neighbors()
    assert np.array_equal(ra[0], rb[0]) and np.array_equal(ra[1], rb[1])


@pytest.mark.parametrize("variant", [0, 1, 3])
def test_batched_list_variants_identical(variant):
    print(f"[tuning batched] variant={variant}", flush=True)
    ref = scenes.make_batch_engine(MULTI)
    sim = scenes.make_batch_engine(MULTI)
    sim.set_tuning(128, variant)
    for s in (ref, sim):
        s.step(MULTI[0].dt, 2, 4)
    a, b = gpu_fields(ref), gpu_fields(sim)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_autotune_keeps_state_and_results():
    print("[autotune]", flush=True)
    ref = scenes.make_engine(SC)
    sim = scenes.make_engine(SC)
    for s in (ref, sim):
        s.step(SC.dt, SC.substeps, SC.iterations)
    before = gpu_fields(sim)
    t = sim.autotune(SC.substep_dt(), SC.iterations, repeats=2)
    assert set(t["walk_block"].values()) <= {64, 128, 192, 256} and 0 <= t["list_variant"] < 4
    assert sim.get_tuning() == t
    after = gpu_fields(sim)
    for k in before:
        assert np.array_equal(before[k], after[k]), k
    for s in (ref, sim):
        s.step(SC.dt, SC.substeps, SC.iterations)
    a, b = gpu_fields(ref), gpu_fields(sim)
    for k in a:
        assert np.array_equal(a[k], b[k]), k


def test_bad_settings_are_rejected():
    sim = scenes.make_engine(SC)
    for block in (0, 32, 96, 320):
        with pytest.raises(PBFError):
            sim.set_tuning(block, 0)
    with pytest.raises(PBFError):
        sim.set_tuning(256, 4)


@pytest.mark.timeout(300)
@pytest.mark.parametrize("name,steps", [("dam_break_1m", 60), ("obstacle_drop_4m", 8)])
def test_list_variants_full_size(name, steps):
    """Every neighbour-list variant at the full sizes, on evolved states where
    lists spill past the LDS rows and (in the 4M hashed scene) buckets wrap
    the table and offsets exceed 16 bits: every field and the ordered
    neighbour lists stay bit-identical to variant 3 after each step."""
    sc = getattr(scenes, name)()
    print(f"[list variants] {name} n={sc.n} steps={steps}", flush=True)
    sims = [scenes.make_engine(sc) for _ in range(4)]
    for s, v in zip(sims, (3, 0, 1, 2)):
        s.set_tuning(128, v)
    for k in range(steps):
        for s in sims:
            s.step(sc.dt, sc.substeps, sc.iterations)
        if k in (0, steps // 2, steps - 1):
            ref = sims[0].read_neighbors()
            for s in sims[1:]:
                got = s.read_neighbors()
                assert np.array_equal(ref[0], got[0]) and np.array_equal(ref[1], got[1]), k
    a = gpu_fields(sims[0])
    for s in sims[1:]:
        b = gpu_fields(s)
        for key in a:
            assert np.array_equal(a[key], b[key]), key
    for s in sims:
        s.close()
