"""GPU parity of the MLS-MPM step (libpbf's pbf_mpm_* entry points, HIP
kernels in csrc/pbf_mpm.hip) against the CPU oracle (oracle/mpm_oracle.c).

The engine and the oracle sum every P2G node in the same (base cell, id)
order and evaluate each expression in the shader's order without
contraction, so positions, velocities, J, the affine matrices and the grid
are compared bit for bit; the 1M-particle case is compared after two steps
and checked for mass-weighted momentum and finiteness.
"""
import numpy as np
import pytest

import oracle
from pairs import ulp_diff
from pbf_host import MPM, MPM_AFFINE, MPM_GRID, MPM_POSITION, MPM_VELOCITY, PBF_FLAG_NO_GRAPH, PBFError

pytestmark = pytest.mark.gpu

FIELDS = {"position": (MPM_POSITION, oracle.MPM_POSITION), "velocity": (MPM_VELOCITY, oracle.MPM_VELOCITY),
          "affine": (MPM_AFFINE, oracle.MPM_AFFINE), "grid": (MPM_GRID, oracle.MPM_GRID)}


def compare(sim, o, fields=FIELDS):
    for name, (gf, of) in fields.items():
        g, c = sim.read(gf), o.get(of)
        assert g.shape == c.shape, name
        assert ulp_diff(g, c) == 0, (name, ulp_diff(g, c))


def test_reference_configuration_bit_exact():
    """MPM.ts's own setup: 9,000 particles uniform in [0.15, 0.55]^3, 32^3
    grid, dt 8e-4, bound 3, E 400; two frames of 15 steps (MPM.ts:64)."""
    sim = MPM()
    x = sim.initResource(seed=11)
    o = oracle.MpmOracle(x)
    print(f"[mpm] reference configuration n={sim.particleCount}", flush=True)
    for frame in range(2):
        sim.run(15)
        o.step(15)
        compare(sim, o)
    sim.close()


@pytest.mark.parametrize("flags", [0, PBF_FLAG_NO_GRAPH])
def test_stressed_state_bit_exact(flags):
    """Random velocities, non-symmetric C and J in [0.8, 1.2] (stress and the
    affine term both active), a 48^3 grid without gravity."""
    rng = np.random.default_rng(5)
    n = 20000
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(0.2, 0.8, (n, 3))
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0, 0.5, (n, 3))
    v[:, 3] = rng.uniform(0.8, 1.2, n)
    C = rng.normal(0, 4.0, (n, 9)).astype(np.float32)
    sim = MPM(n_particle=n, n_grid=48, gravity=0.0, flags=flags)
    sim.setParticlePosition(x)
    sim.write(MPM_VELOCITY, v)
    sim.write(MPM_AFFINE, C)
    o = oracle.MpmOracle(x, n_grid=48, gravity=0.0)
    o.set(oracle.MPM_VELOCITY, v)
    o.set(oracle.MPM_AFFINE, C)
    for _ in range(4):
        sim.run(1)
        o.step(1)
        compare(sim, o)
    sim.close()


def test_state_round_trip_after_steps_keeps_caller_order():
    """The engine keeps particles in the cell order of its last step; reading
    every field after some steps and writing it back (through the slot -> id
    map) must leave the run on the oracle's track, which never reorders."""
    rng = np.random.default_rng(8)
    n = 12000
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(0.2, 0.7, (n, 3))
    x[:, 3] = rng.uniform(-1.0, 1.0, n)  # w is carried untouched
    sim = MPM(n_particle=n, n_grid=48)
    sim.setParticlePosition(x)
    sim.write(MPM_POSITION, x)
    o = oracle.MpmOracle(x, n_grid=48)
    o.set(oracle.MPM_POSITION, x)
    sim.run(6)
    o.step(6)
    compare(sim, o)
    state = {f: sim.read(f) for f in (MPM_POSITION, MPM_VELOCITY, MPM_AFFINE)}
    for f, a in state.items():
        sim.write(f, a)
    compare(sim, o)
    sim.run(6)
    o.step(6)
    compare(sim, o)
    sim.close()


def test_g2p_reproduces_a_linear_field():
    n, ng = 3000, 32
    rng = np.random.default_rng(6)
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(0.3, 0.7, (n, 3))
    A = np.array([[0.3, -1.2, 0.5], [0.7, 0.1, -0.4], [-0.9, 0.6, 0.2]])
    b = np.array([0.05, -0.2, 0.1])
    idx = np.stack(np.meshgrid(*[np.arange(ng)] * 3, indexing="ij"), -1).reshape(-1, 3)
    g = np.zeros((ng ** 3, 4), np.float32)
    g[:, :3] = (idx / ng) @ A.T + b
    sim = MPM(n_particle=n, gravity=0.0)
    sim.setParticlePosition(x)
    sim.write(MPM_GRID, g)
    sim.g2p()
    v = sim.read(MPM_VELOCITY)[:, :3].astype(np.float64)
    C = sim.read(MPM_AFFINE).astype(np.float64).reshape(n, 3, 3)
    assert np.abs(v - (x[:, :3] @ A.T + b)).max() < 2e-6
    assert np.abs(C - A[None]).max() < 2e-4
    sim.close()


def test_one_million_particles_two_steps_bit_exact():
    n, ng = 1 << 20, 128
    rng = np.random.default_rng(7)
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = (rng.random((n, 3)) * 0.4 + 0.15).astype(np.float32)
    print(f"[mpm] n={n} grid {ng}^3", flush=True)
    # dt scaled with the spacing (the explicit step's elastic-wave limit
    # dx / sqrt(E / rho) is 3.9e-4 on this grid)
    sim = MPM(n_particle=n, n_grid=ng, dt=2e-4)
    sim.setParticlePosition(x)
    o = oracle.MpmOracle(x, n_grid=ng, dt=2e-4)
    for _ in range(2):
        sim.run(1)
        o.step(1)
    compare(sim, o)
    v = sim.read(MPM_VELOCITY)
    assert np.isfinite(v).all()
    sim.close()


def test_errors():
    with pytest.raises(PBFError, match="n_grid"):
        MPM(n_grid=2)
    sim = MPM(n_particle=10)
    with pytest.raises(PBFError, match="To Many Particles!"):
        sim.setParticlePosition(np.zeros((11, 4), np.float32))
    sim.close()
