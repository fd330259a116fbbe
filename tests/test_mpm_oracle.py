"""CPU checks of the MLS-MPM oracle (oracle/mpm_oracle.c) against what the
mathematics fixes: the quadratic B-spline's partition of unity and vanishing
first moment (mass and momentum conservation of P2G), exact reproduction of a
linear grid velocity field and of its gradient by G2P (in the reference's
column-major C layout, MPMShader.ts:258-262), the sign of the pressure term,
the J update, the boundary layer, and an independent float64 scatter of P2G
(the reference's own per-particle form, MPMShader.ts:4-62) against the
oracle's per-node gather.
"""
import numpy as np
import pytest

import oracle

NG = 32
DX = 1.0 / NG


def cloud(n, lo=0.3, hi=0.7, seed=0):
    rng = np.random.default_rng(seed)
    x = np.zeros((n, 4), np.float32)
    x[:, :3] = rng.uniform(lo, hi, (n, 3))
    return x


def weights(fx):
    return np.stack([0.5 * (1.5 - fx) ** 2, 0.75 - (fx - 1.0) ** 2, 0.5 * (fx - 0.5) ** 2])  # [k][axis]


def scatter_p2g(x, v, C, J, dt=8e-4, E=400.0, rho=1.0):
    """P2G in float64, particle by particle (the shader's own loop)."""
    p_vol = (DX * 0.5) ** 2
    p_mass = rho * p_vol
    grid = np.zeros((NG, NG, NG, 4))
    for p in range(len(x)):
        Xp = x[p, :3].astype(np.float64) / DX
        base = np.floor(Xp - 0.5).astype(int)
        fx = Xp - base
        w = weights(fx)
        Cp = C[p].reshape(3, 3).astype(np.float64)  # row c = column c of the WGSL mat3x3
        Jp = J[p] * (1.0 + dt * np.trace(Cp))
        stress = -dt * 4.0 * E * p_vol * (Jp - 1.0) / DX ** 2
        affine_cols = stress * np.eye(3) + p_mass * Cp  # affine_cols[c] = column c
        for i in range(3):
            for j in range(3):
                for k in range(3):
                    dpos = (np.array([i, j, k]) - fx) * DX
                    wt = w[i, 0] * w[j, 1] * w[k, 2]
                    mv = wt * (p_mass * v[p, :3] + affine_cols.T @ dpos)
                    grid[base[0] + i, base[1] + j, base[2] + k] += [*mv, wt * p_mass]
    return grid.reshape(-1, 4)


def test_p2g_matches_float64_scatter():
    n = 400
    x = cloud(n)
    rng = np.random.default_rng(1)
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0, 0.3, (n, 3))
    v[:, 3] = rng.uniform(0.9, 1.1, n)  # J
    C = rng.normal(0, 2.0, (n, 9)).astype(np.float32)
    o = oracle.MpmOracle(x)
    o.set(oracle.MPM_VELOCITY, v)
    o.set(oracle.MPM_AFFINE, C)
    raw = o.p2g()
    ref = scatter_p2g(x, v, C, v[:, 3])
    scale = np.abs(ref).max(axis=0)
    assert np.abs(raw - ref).max(axis=0).max() / scale.max() < 2e-6
    assert np.abs(raw[:, 3] - ref[:, 3]).max() / scale[3] < 2e-6


def test_p2g_conserves_mass_and_momentum():
    n = 2000
    x = cloud(n, seed=2)
    rng = np.random.default_rng(3)
    v = np.zeros((n, 4), np.float32)
    v[:, :3] = rng.normal(0, 0.5, (n, 3))
    v[:, 3] = 1.0
    C = rng.normal(0, 3.0, (n, 9)).astype(np.float32)
    o = oracle.MpmOracle(x)
    o.set(oracle.MPM_VELOCITY, v)
    o.set(oracle.MPM_AFFINE, C)
    raw = o.p2g().astype(np.float64)
    p_mass = (DX * 0.5) ** 2
    # partition of unity: total node mass = total particle mass
    assert raw[:, 3].sum() == pytest.approx(n * p_mass, rel=1e-5)
    # first moment of the weights vanishes: the affine and stress terms add
    # no net momentum; J = 1 + dt tr C != 1 still adds none
    mom = (p_mass * v[:, :3].astype(np.float64)).sum(axis=0)
    assert np.abs(raw[:, :3].sum(axis=0) - mom).max() <= 1e-5 * np.abs(p_mass * v[:, :3]).sum()


def test_g2p_reproduces_a_linear_field_and_its_gradient():
    n = 500
    x = cloud(n, seed=4)
    A = np.array([[0.3, -1.2, 0.5], [0.7, 0.1, -0.4], [-0.9, 0.6, 0.2]])
    b = np.array([0.05, -0.2, 0.1])
    idx = np.stack(np.meshgrid(np.arange(NG), np.arange(NG), np.arange(NG), indexing="ij"), -1).reshape(-1, 3)
    g = np.zeros((NG ** 3, 4), np.float32)
    g[:, :3] = (idx * DX) @ A.T + b
    o = oracle.MpmOracle(x, gravity=0.0)
    o.set(oracle.MPM_GRID, g)
    o.g2p()
    v = o.get(oracle.MPM_VELOCITY)[:, :3].astype(np.float64)
    C = o.get(oracle.MPM_AFFINE).astype(np.float64).reshape(n, 3, 3)
    want_v = x[:, :3].astype(np.float64) @ A.T + b
    assert np.abs(v - want_v).max() < 2e-6
    # sum_i w (x_i - x_p)(x_i - x_p)^T = dx^2 / 4 I, so 4/dx^2 sum_i w v_i
    # dpos^T = A; the shader stores column c = v_c * dpos, i.e. C[c][r] = A[c][r]
    assert np.abs(C - A[None]).max() < 2e-4
    # positions advance by dt * v
    xp = o.get(oracle.MPM_POSITION)[:, :3].astype(np.float64)
    assert np.abs(xp - (x[:, :3] + 8e-4 * v)).max() < 1e-6


def test_pressure_pushes_a_compressed_particle_outward_and_j_update():
    x = np.array([[0.5 + 0.3 * DX, 0.5, 0.5 - 0.2 * DX, 0]], np.float32)
    o = oracle.MpmOracle(x, gravity=0.0)
    v = np.array([[0, 0, 0, 0.8]], np.float32)  # J = 0.8: compressed
    C = np.zeros((1, 9), np.float32)
    C[0, 0], C[0, 4], C[0, 8] = 1.0, 2.0, -0.5
    o.set(oracle.MPM_VELOCITY, v)
    o.set(oracle.MPM_AFFINE, C)
    raw = o.p2g()
    J = o.get(oracle.MPM_VELOCITY)[0, 3]
    assert J == pytest.approx(0.8 * (1 + 8e-4 * 2.5), rel=1e-6)
    idx = np.stack(np.meshgrid(np.arange(NG), np.arange(NG), np.arange(NG), indexing="ij"), -1).reshape(-1, 3)
    d = idx * DX - x[0, :3]
    m = raw[:, 3] > 0
    # without velocity, the node momentum is w (stress I + m C) dpos; the
    # stress term dominates and points away from the particle (J < 1)
    stress_dir = (raw[m, :3] * d[m]).sum(axis=1)
    assert (stress_dir > 0).sum() > 0.9 * m.sum()


def test_boundary_layer_stops_inflow():
    # a particle moving into the x = 0 wall, inside the 3-node layer
    x = np.array([[1.6 * DX, 0.5, 0.5, 0]], np.float32)
    o = oracle.MpmOracle(x, gravity=0.0)
    o.set(oracle.MPM_VELOCITY, np.array([[-1.0, 0.0, 0.0, 1.0]], np.float32))
    o.p2g()
    g = o.get(oracle.MPM_GRID).reshape(NG, NG, NG, 4)
    assert (g[:3, :, :, 0] == 0).all()        # nodes i < 3: no velocity toward -x
    assert (g[3:, :, :, 0] <= 0).all() and (g[3:, :, :, 0] < 0).any()


def test_free_fall_of_a_block():
    x = cloud(3000, seed=5)
    o = oracle.MpmOracle(x)
    o.step(15)
    v = o.get(oracle.MPM_VELOCITY)
    assert np.abs(v[:, 1] + 9.8 * 8e-4 * 15).max() < 1e-5
    assert np.abs(v[:, [0, 2]]).max() < 1e-6
    assert np.abs(v[:, 3] - 1.0).max() < 1e-6
