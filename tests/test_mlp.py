"""The learned warm start: network kernels, feature pass, learned solve.

CPU: the network oracle (oracle/mlp_oracle.py) is pinned by central finite
differences of its float64 loss and by bfloat16 rounding cases; the blob
initialiser is checked on the host.  GPU: libpbf's MFMA forward/backward
against the oracle with the same bfloat16 rounding points, training on a
known function, the collection mode against the plain solve (bit-exact),
the features against their definitions, the learned lambda of iteration 0
against the oracle's prediction, and the fallback to the plain solve.
"""
import os

import numpy as np
import pytest

import mlp_oracle as mo
import scenes
from pbf_host import MLP_PARAMS, PBFError, WARM_LEARNED, WARM_PLAIN, mlp_blob_size, mlp_forward_backward, \
    mlp_init, mlp_train


def _rng_blob(seed=3, scale=1.0):
    rng = np.random.default_rng(seed)
    b = np.zeros(mo.BLOB, np.float32)
    b[mo.OFF_W1:mo.OFF_B1] = rng.normal(0, 0.35, 1024) * scale
    b[mo.OFF_B1:mo.OFF_W2] = rng.normal(0, 0.1, 64)
    b[mo.OFF_W2:mo.OFF_B2] = rng.normal(0, 0.18, 4096) * scale
    b[mo.OFF_B2:mo.OFF_W3] = rng.normal(0, 0.1, 64)
    b[mo.OFF_W3:mo.OFF_B3] = rng.normal(0, 0.2, 64)
    b[mo.OFF_B3] = 0.05
    b[mo.OFF_FMU:mo.OFF_FSIG] = rng.normal(0, 0.5, 16)
    b[mo.OFF_FSIG:mo.OFF_TMU] = rng.uniform(0.5, 2.0, 16)
    b[mo.OFF_TMU] = -0.3
    b[mo.OFF_TSIG] = 1.7
    b[mo.OFF_LMIN] = -1e30
    return b


def _data(n, seed=5):
    rng = np.random.default_rng(seed)
    f = rng.normal(0, 1.0, (n, 16)).astype(np.float32)
    t = (np.tanh(f[:, 0] - 0.5 * f[:, 3]) + 0.3 * f[:, 7] * f[:, 2]).astype(np.float32)
    return f, t


# ---------------------------------------------------------------- CPU -----
def test_bf16_rounding_cases():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, 1.0 + 2.0 ** -8 + 2.0 ** -20, -2.5, 0.0, 3.0e38],
                 np.float32)
    # ties go to even; just above the tie rounds up; exactly representable values stay
    want = np.array([1.0, 1.0, 1.0 + 4 * 2.0 ** -8, 1.0 + 2.0 ** -7, -2.5, 0.0, np.float32(3.0e38)], np.float64)
    got = mo.bf16(x).astype(np.float64)
    assert got[:6].tolist() == want[:6].tolist()
    assert abs(got[6] - 3.0e38) / 3.0e38 < 2 ** -8
    # every output has its low 16 bits clear
    assert np.all(mo.bf16(np.random.default_rng(0).normal(size=1000).astype(np.float32)).view(np.uint32) & 0xFFFF
                  == 0)


def test_oracle_gradient_matches_finite_differences():
    b = _rng_blob()
    f, t = _data(48)
    _, g, loss = mo.forward_backward(b, f, t, quant=False)
    assert abs(loss - mo.loss_only(b.astype(np.float64), f, t)) < 1e-12
    rng = np.random.default_rng(11)
    # every bias and w3 entry, and a sample of both weight matrices
    idx = np.concatenate([np.arange(mo.OFF_B1, mo.OFF_W2), np.arange(mo.OFF_B2, mo.NPARAM),
                          rng.choice(1024, 40, replace=False) + mo.OFF_W1,
                          rng.choice(4096, 60, replace=False) + mo.OFF_W2])
    b64 = b.astype(np.float64)
    eps = 1e-6
    for q in idx:
        p, m = b64.copy(), b64.copy()
        p[q] += eps
        m[q] -= eps
        fd = (mo.loss_only(p, f, t) - mo.loss_only(m, f, t)) / (2 * eps)
        assert abs(fd - g[q]) <= 1e-6 * max(1.0, abs(fd)), (q, fd, g[q])


def test_oracle_quantised_close_to_float64():
    b = _rng_blob()
    f, t = _data(256)
    y64, g64, l64 = mo.forward_backward(b, f, t, quant=False)
    yq, gq, lq = mo.forward_backward(b, f, t, quant=True)
    assert np.abs(yq - y64).max() < 0.05 * np.abs(y64).max()
    assert abs(lq - l64) < 0.05 * l64
    assert np.linalg.norm(gq - g64) < 0.05 * np.linalg.norm(g64)


def test_blob_init_on_host():
    n = mlp_blob_size()
    assert n == mo.BLOB and MLP_PARAMS == mo.NPARAM
    a, b = mlp_init(1), mlp_init(1)
    assert np.array_equal(a, b) and not np.array_equal(a, mlp_init(2))
    P = mo.unpack(a)
    assert np.abs(P["W1"]).max() <= np.sqrt(6.0 / 80) and np.abs(P["W2"]).max() <= np.sqrt(6.0 / 128)
    assert np.abs(P["W1"]).std() > 0.05
    assert not P["b1"].any() and not P["b2"].any() and P["b3"] == 0.0
    assert (P["fsig"] == 1.0).all() and not P["fmu"].any() and P["tsig"] == 1.0 and P["tmu"] == 0.0
    assert P["lmin"] <= -1e38


def test_normalisation_from_data():
    import learned
    f, t = _data(1000)
    f[:, 9] = 0.0  # a constant feature keeps sigma = 1
    b = learned.set_normalisation(mlp_init(0), f, t)
    P = mo.unpack(b)
    assert np.allclose(P["fmu"], f.mean(0), atol=1e-6) and P["fsig"][9] == 1.0
    assert np.allclose(P["fsig"][:9], f.std(0)[:9], rtol=1e-5)
    assert abs(P["tmu"] - t.mean()) < 1e-6 and abs(P["tsig"] - t.std()) < 1e-5
    assert P["lmin"] == pytest.approx(1.5 * t.min(), rel=1e-6)


# ---------------------------------------------------------------- GPU -----
@pytest.mark.gpu
@pytest.mark.parametrize("n", [1, 31, 32, 77, 4096 + 13])
def test_gpu_forward_backward_matches_oracle(n):
    print(f"[mlp fwd/bwd] n={n}", flush=True)
    b = _rng_blob()
    f, t = _data(n)
    y, g, loss = mlp_forward_backward(b, f, t)
    yq, gq, lq = mo.forward_backward(b, f, t, quant=True)
    # same bf16 rounding points: what differs is float32 summation order and
    # the last bit of exp, which can move a bf16 rounding by one step
    assert np.abs(y - yq).max() <= 2e-3 * max(1.0, np.abs(yq).max()), np.abs(y - yq).max()
    assert abs(loss - lq) <= 2e-3 * lq
    for lo, hi, name in [(mo.OFF_W1, mo.OFF_B1, "W1"), (mo.OFF_B1, mo.OFF_W2, "b1"), (mo.OFF_W2, mo.OFF_B2, "W2"),
                         (mo.OFF_B2, mo.OFF_W3, "b2"), (mo.OFF_W3, mo.OFF_B3, "w3"), (mo.OFF_B3, mo.NPARAM, "b3")]:
        d = np.linalg.norm(g[lo:hi] - gq[lo:hi])
        assert d <= 1e-2 * np.linalg.norm(gq[lo:hi]) + 1e-7, (name, d, np.linalg.norm(gq[lo:hi]))


@pytest.mark.gpu
def test_gpu_training_fits_a_known_function():
    print("[mlp train]", flush=True)
    f, t = _data(32768, seed=9)
    import learned
    blob0 = learned.set_normalisation(mlp_init(4), f, t)
    _, _, l0 = mlp_forward_backward(blob0, f, t)
    blob, losses = mlp_train(f, t, blob0, epochs=15, batch=1024, lr=3e-3, seed=4)
    _, _, l1 = mlp_forward_backward(blob, f, t)
    assert losses.shape == (15,) and np.all(np.isfinite(losses))
    assert l1 < 0.05 * l0, (l0, l1, losses)
    assert losses[-1] < losses[0]
    # training is deterministic up to float atomics; the normalisation is kept
    assert np.array_equal(blob[mo.OFF_FMU:], blob0[mo.OFF_FMU:])


def _small_scene(seed=scenes.SEED):
    return scenes.dam_break(12, 12, 8, map_res=32, vorticity=0.01, tension=0.1, seed=seed, name="learn_small")


@pytest.mark.gpu
def test_collection_mode_is_the_plain_solve():
    """Collecting pairs (long solve undone, then the plain substep with the
    lambda sums kept) moves the particles exactly as the plain solve does."""
    sc = _small_scene()
    print(f"[collect == plain] n={sc.n}", flush=True)
    a, b = scenes.make_engine(sc), scenes.make_engine(sc)
    dt = sc.substep_dt()
    for _ in range(3):
        a.step(dt, 1, 3)
        f, t = b.collect_training(dt, 12, 3)
        assert f.shape == (sc.n, 16) and t.shape == (sc.n,)
    for k in ("position", "velocity", "accel"):
        assert np.array_equal(a.read_field(k), b.read_field(k)), k


@pytest.mark.gpu
def test_collected_features_and_targets():
    sc = _small_scene()
    print(f"[features] n={sc.n}", flush=True)
    dt = sc.substep_dt()
    a = scenes.make_engine(sc)
    f, t = a.collect_training(dt, 1, 2)               # target of a 1-iteration solve = plain lambda
    assert np.array_equal(t, f[:, 2])
    assert np.all(f[:, 3] == 0.0) and np.all(f[:, 4] == 0.0)  # no previous substep yet
    pairs = []
    for _ in range(12):
        # long solve = the deployed 2-iteration solve, so each target is the
        # lambda sum that the next pair reads back as feature 3
        pairs.append(a.collect_training(dt, 2, 2))
    compressed = 0
    for (f0, t0), (f1, _) in zip(pairs, pairs[1:]):
        assert np.array_equal(f1[:, 3], t0)
        assert np.array_equal(f0[:, 1], np.maximum(f0[:, 0], 0.0))          # C = max(rho~ - 1, 0)
        assert np.allclose(f0[:, 0], f0[:, 14] + f0[:, 15], atol=2e-6)      # fluid + boundary parts
        nz = f0[:, 1] > 0
        assert np.all(f0[nz, 2] < 0.0) and np.all(f0[~nz, 2] == 0.0)        # lambda = -C / (den + eps)
        assert np.all(t0 <= 0.0) and np.all(f0[:, 4] <= 0.0)
        assert np.all(f0[:, 8] >= 1) and np.all(f0[:, 8] <= 96)             # neighbour count (self included)
        assert np.all(f0[:, 9] >= 0.0) and np.all(f0[:, 5] >= f0[:, 6])     # den = |grad_i|^2 + sum |grad_j|^2
        compressed += int(nz.sum())
    assert compressed > 0 and np.any(pairs[-1][0][:, 3] < 0.0)


@pytest.mark.gpu
def test_learned_iteration0_lambda_matches_oracle():
    sc = _small_scene()
    print(f"[learned lambda] n={sc.n}", flush=True)
    dt = sc.substep_dt()
    c = scenes.make_engine(sc)
    f, _ = c.collect_training(dt, 1, 1)
    import learned
    blob = learned.set_normalisation(_rng_blob(7, 0.5), f, f[:, 2])
    s = scenes.make_engine(sc)
    s.set_warm_start(blob, WARM_LEARNED)
    s.step(dt, 1, 1)                                 # one iteration: lambda of iteration 0 is the network's
    got = s.read_field("lambda")[:, 0]
    want = mo.predict_lambda(blob, f, f[:, 2])
    tsig = float(blob[mo.OFF_TSIG])
    assert np.abs(got - want).max() <= 4e-3 * tsig, np.abs(got - want).max()
    assert np.all(got <= 0.0)
    assert s.warm_start_stats() == (1, 0)
    r = scenes.make_engine(sc)
    r.set_warm_start(blob, WARM_LEARNED, relax=0.5)
    r.step(dt, 1, 1)
    got = r.read_field("lambda")[:, 0]
    want = mo.predict_lambda(blob, f, f[:, 2], relax=0.5)
    assert np.abs(got - want).max() <= 2e-3 * tsig, np.abs(got - want).max()
    assert np.all(got >= 0.5 * blob[mo.OFF_LMIN])
    # a tighter lower bound clamps part of the predictions
    tight = blob.copy()
    tight[mo.OFF_LMIN] = np.float32(np.median(mo.predict_lambda(blob, f, f[:, 2])))
    q = scenes.make_engine(sc)
    q.set_warm_start(tight, WARM_LEARNED)
    q.step(dt, 1, 1)
    got = q.read_field("lambda")[:, 0]
    want = mo.predict_lambda(tight, f, f[:, 2])
    assert np.any(want == tight[mo.OFF_LMIN]) and np.all(got >= tight[mo.OFF_LMIN])
    assert np.abs(got - want).max() <= 4e-3 * tsig


@pytest.mark.gpu
def test_fallback_replaces_a_rejected_substep_with_the_plain_solve():
    sc = _small_scene()
    print(f"[fallback] n={sc.n}", flush=True)
    dt = sc.substep_dt()
    blob = _rng_blob(8)
    blob[mo.OFF_TMU] = -50.0                         # absurd lambdas: the substep must be rejected
    a = scenes.make_engine(sc)
    b = scenes.make_engine(sc)
    b.set_warm_start(blob, WARM_LEARNED, fallback_threshold=1e-12, fallback_iterations=3)
    for _ in range(3):
        a.step(dt, 1, 3)
        b.step(dt, 1, 2)
    assert b.warm_start_stats() == (3, 3)
    for k in ("position", "velocity", "accel"):
        assert np.array_equal(a.read_field(k), b.read_field(k)), k
    # a loose density threshold: the particles' speed rejects the substep
    c = scenes.make_engine(sc)
    c.set_warm_start(blob, WARM_LEARNED, fallback_threshold=10.0, fallback_iterations=3)
    c.step(dt, 1, 2)
    assert c.warm_start_stats() == (1, 1)
    # back to the plain solve
    b.set_warm_start(None, WARM_PLAIN)
    a.step(dt, 1, 3)
    b.step(dt, 1, 3)
    assert np.array_equal(a.read_field("position"), b.read_field("position"))


@pytest.mark.gpu
def test_device_side_fallback_graph_replay_equals_direct_launches():
    """The learned substep, its device-side check and the gated restore +
    plain re-solve are one graph: replayed with eager reads between the
    replays, with a threshold that keeps some substeps and rejects others,
    it gives the same states and the same rejection count as direct
    launches."""
    from pbf_host import PBF_FLAG_NO_GRAPH, WARM_PREVIOUS
    sc = _small_scene()
    dt = sc.substep_dt()
    probe = scenes.make_engine(sc)
    errs = []
    for _ in range(10):
        probe.step(dt, 1, 2)
        errs.append(probe.density_error()[0])
    thr = float(np.median(errs))           # some substeps above it, some below
    print(f"[device fallback] n={sc.n} threshold {thr:.4g}", flush=True)
    a = scenes.make_engine(sc)
    b = scenes.make_engine(sc, flags=PBF_FLAG_NO_GRAPH)
    for e in (a, b):
        e.set_warm_start(None, WARM_PREVIOUS, 1.0, fallback_threshold=thr, fallback_iterations=4)
    for k in range(10):
        a.step(dt, 1, 2)
        b.step(dt, 1, 2)
        assert np.array_equal(a.read_positions(), b.read_positions()), k
    assert np.array_equal(a.read_velocities(), b.read_velocities())
    sa, sb = a.warm_start_stats(), b.warm_start_stats()
    print(f"[device fallback] stats graph {sa} direct {sb}", flush=True)
    assert sa == sb and sa[0] == 10 and 0 < sa[1] < 10


@pytest.mark.gpu
def test_warm_start_argument_errors():
    sc = _small_scene()
    s = scenes.make_engine(sc)
    with pytest.raises(PBFError):
        s.set_warm_start(None, WARM_LEARNED)          # no network yet
    with pytest.raises(PBFError):
        s.set_warm_start(np.zeros(10, np.float32), WARM_LEARNED)
    bad = _rng_blob()
    bad[100] = np.nan
    with pytest.raises(PBFError):
        s.set_warm_start(bad, WARM_LEARNED)
    with pytest.raises(PBFError):
        s.set_warm_start(_rng_blob(), 7)
    with pytest.raises(PBFError):
        s.set_warm_start(_rng_blob(), WARM_LEARNED, relax=0.0)
    with pytest.raises(PBFError):
        s.set_warm_start(_rng_blob(), WARM_LEARNED, relax=1.5)


WARM_BLOB = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "webgpu-fluid-simulation_amd",
                         "warm_start", "dam_break.f32")


@pytest.mark.gpu
def test_learned_warm_start_gain_on_held_out_dam_break():
    """The committed network (trained on other block shapes and seeds) at 3
    iterations reaches a mean density error the plain solve needs 5 for,
    with a trajectory (mean height, kinetic energy) at least as close to a
    40-iteration reference run as the plain solve at 3 iterations."""
    import learned
    import train_warm_start as tw
    blob = np.fromfile(WARM_BLOB, "<f4")
    assert blob.size == mlp_blob_size()
    ev = tw.family(tw.EVAL_SHAPE, tw.EVAL_SEED)
    print(f"[warm-start gain] n={ev.n}", flush=True)
    sub = 150
    ref = learned.run(ev, sub, 40)
    plain3, plain5 = learned.run(ev, sub, 3), learned.run(ev, sub, 5)
    warm3 = learned.run(ev, sub, 3, blob, 0.5)
    e_p3, e_p5, e_w3 = (float(t.density_error.mean()) for t in (plain3, plain5, warm3))
    print(f"density error plain I=3 {e_p3:.5f}  plain I=5 {e_p5:.5f}  learned I=3 {e_w3:.5f}", flush=True)
    assert e_w3 <= 0.005 < e_p3
    assert e_w3 <= e_p5
    dh_p, dk_p = learned.fidelity(plain3, ref)
    dh_w, dk_w = learned.fidelity(warm3, ref)
    print(f"fidelity plain I=3 ({dh_p:.4f}, {dk_p:.3f})  learned I=3 ({dh_w:.4f}, {dk_w:.3f})", flush=True)
    assert dh_w <= dh_p and dk_w <= dk_p


@pytest.mark.gpu
def test_temporal_warm_start_runs_and_lowers_the_error():
    """PBF_WARM_PREVIOUS (no network): iteration 0 applies relax x the previous
    substep's lambda sum; on the held-out dam break it lowers the mean
    density error of the 2-iteration solve."""
    import learned
    import train_warm_start as tw
    from pbf_host import WARM_PREVIOUS
    ev = tw.family(tw.EVAL_SHAPE, tw.EVAL_SEED)
    print(f"[temporal warm start] n={ev.n}", flush=True)
    plain = learned.run(ev, 60, 2, every=30)
    prev = learned.run(ev, 60, 2, None, 0.8, every=30, mode=WARM_PREVIOUS)
    assert np.all(np.isfinite(prev.height)) and np.all(np.isfinite(prev.kinetic))
    assert prev.density_error.mean() < 0.8 * plain.density_error.mean()


@pytest.mark.gpu
def test_ratio_mode_with_a_constant_network_is_the_temporal_warm_start():
    """PBF_WARM_RATIO applies clamp(relax s, 0, 2 relax) x Lambda_prev with s the
    network output; zero weights and tmu = 1 make s = 1, so every substep must
    be bit-identical to PBF_WARM_PREVIOUS at the same relax."""
    import train_warm_start as tw
    from pbf_host import WARM_PREVIOUS, WARM_RATIO
    blob = np.zeros(mo.BLOB, np.float32)
    blob[mo.OFF_FSIG:mo.OFF_TMU] = 1.0
    blob[mo.OFF_TMU] = 1.0
    blob[mo.OFF_TSIG] = 1.0
    blob[mo.OFF_LMIN] = -1e30
    sc = tw.family(tw.EVAL_SHAPE, tw.EVAL_SEED)
    print(f"[ratio = temporal] n={sc.n}", flush=True)
    sims = [scenes.make_engine(sc, max_neighbors=128) for _ in range(2)]
    sims[0].set_warm_start(None, WARM_PREVIOUS, 0.6)
    sims[1].set_warm_start(blob, WARM_RATIO, 0.6)
    for _ in range(12):
        for s in sims:
            s.step(sc.substep_dt(), 1, 3)
    a, b = sims[0].read_positions(), sims[1].read_positions()
    assert np.array_equal(a, b)
    assert np.array_equal(sims[0].read_velocities(), sims[1].read_velocities())
    for s in sims:
        s.close()


RATIO_BLOB = os.path.join(os.path.dirname(WARM_BLOB), "dam_break_ratio.f32")


@pytest.mark.gpu
def test_ratio_warm_start_gain_on_held_out_dam_break():
    """The committed ratio network at 3 iterations (relax 0.6) reaches a mean
    density error the plain solve needs 5 for, with a trajectory at least as
    close to a 40-iteration reference run as the plain solve at 5."""
    import learned
    import train_warm_start as tw
    from pbf_host import WARM_RATIO
    blob = np.fromfile(RATIO_BLOB, "<f4")
    assert blob.size == mlp_blob_size()
    ev = tw.family(tw.EVAL_SHAPE, tw.EVAL_SEED)
    print(f"[ratio gain] n={ev.n}", flush=True)
    sub = 150
    ref = learned.run(ev, sub, 40)
    plain3, plain5 = learned.run(ev, sub, 3), learned.run(ev, sub, 5)
    warm3 = learned.run(ev, sub, 3, blob, 0.6, mode=WARM_RATIO)
    e_p3, e_p5, e_w3 = (float(t.density_error.mean()) for t in (plain3, plain5, warm3))
    print(f"density error plain I=3 {e_p3:.5f}  plain I=5 {e_p5:.5f}  ratio I=3 {e_w3:.5f}", flush=True)
    assert e_w3 <= 0.005 < e_p3
    assert e_w3 <= e_p5
    dh_p, dk_p = learned.fidelity(plain5, ref)
    dh_w, dk_w = learned.fidelity(warm3, ref)
    print(f"fidelity plain I=5 ({dh_p:.4f}, {dk_p:.3f})  ratio I=3 ({dh_w:.4f}, {dk_w:.3f})", flush=True)
    assert dh_w <= dh_p and dk_w <= dk_p
