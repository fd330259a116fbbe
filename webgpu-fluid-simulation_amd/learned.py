"""Learned warm start of the PBF lambda solve: data, training, evaluation.

The network (16 -> 64 -> 64 -> 1, SiLU, bf16 weights on MFMA) lives in
libpbf (csrc/pbf_mlp.hip); this module is the host-side workflow around it:

  collect   run a scene with the deployed solver and take, every substep,
            the iteration-0 features and the lambda summed over a long
            (converged) Jacobi solve from the same state (pbf_collect_training);
  train     set the feature/target normalisation from the data and fit the
            network with Adam on the GPU (pbf_mlp_train);
  evaluate  run plain and learned solves side by side and report the mean
            density error per substep, and the iterations each needs to
            reach a target error.

No reference counterpart: the reference solves lambda from zero with a fixed
iteration count every substep (PBF.ts:419-431).
"""
from __future__ import annotations

import time
from dataclasses import dataclass
from typing import Callable, List, Optional, Sequence

import numpy as np

from pbf_host import (MLP_FEATURES, WARM_LEARNED, WARM_PLAIN, WARM_RATIO_MAX, PBFError, mlp_init, mlp_train)
from scenes import Scene, make_engine

OFF_FMU, OFF_FSIG, OFF_TMU, OFF_TSIG, OFF_LMIN = 5313, 5329, 5345, 5346, 5347
LMIN_MARGIN = 1.5


def set_normalisation(blob: np.ndarray, features: np.ndarray, targets: np.ndarray) -> np.ndarray:
    """Feature mean / std and target mean / std into the blob (std 0 -> 1);
    the prediction is bounded below by 1.5 x the smallest target seen."""
    b = np.array(blob, np.float32, copy=True)
    f = np.asarray(features, np.float64).reshape(-1, MLP_FEATURES)
    t = np.asarray(targets, np.float64).reshape(-1)
    mu, sd = f.mean(0), f.std(0)
    sd[sd < 1e-12] = 1.0
    b[OFF_FMU:OFF_FSIG] = mu
    b[OFF_FSIG:OFF_TMU] = sd
    b[OFF_TMU] = t.mean()
    b[OFF_TSIG] = t.std() if t.std() > 1e-12 else 1.0
    b[OFF_LMIN] = LMIN_MARGIN * min(float(t.min()), 0.0) if t.size else 0.0
    return b


def collect(scene: Scene, substeps: int, long_iterations: int = 40, iterations: Optional[int] = None,
            blob: Optional[np.ndarray] = None, skip: int = 0, stride: int = 1, relax: float = 1.0,
            keep: float = 1.0, seed: int = 0, max_neighbors: int = 128, fallback_threshold: float = 0.0,
            mode: Optional[int] = None):
    """Training pairs along `substeps` substeps of `scene` with the deployed
    solver (plain, learned when `blob` is given, or the warm-start `mode`);
    `keep` < 1 keeps a seeded random fraction of the particles of every pair."""
    rng = np.random.default_rng(seed)
    sim = make_engine(scene, max_neighbors=max_neighbors)
    try:
        if mode is None:
            mode = WARM_LEARNED if blob is not None else WARM_PLAIN
        if mode != WARM_PLAIN:
            sim.set_warm_start(blob, mode, relax, fallback_threshold)
        it = iterations if iterations is not None else scene.iterations
        dt = scene.substep_dt()
        F: List[np.ndarray] = []
        T: List[np.ndarray] = []
        dropped = 0
        collect.last_error = None
        for k in range(substeps):
            if k >= skip and (k - skip) % stride == 0:
                try:
                    f, t = sim.collect_training(dt, long_iterations, it)
                except PBFError as e:  # an unstable learned run: keep what was gathered
                    collect.last_error = f"substep {k}: {e}"
                    break
                # a long Jacobi solve can blow up where the short one does not;
                # such pairs carry no target
                ok = np.isfinite(f).all(1) & np.isfinite(t)
                dropped += int((~ok).sum())
                if keep < 1.0:
                    ok &= rng.random(len(t)) < keep
                F.append(f[ok])
                T.append(t[ok])
            else:
                sim.step(dt, 1, it)
        collect.last_dropped = dropped
        return np.concatenate(F), np.concatenate(T)
    finally:
        sim.close()


RATIO_MIN_PREV = 1e-3  # |Lambda_prev| below which a pair carries no ratio


def ratio_pairs(features: np.ndarray, targets: np.ndarray):
    """Ratio-mode training pairs: the factor s = Lambda / Lambda_prev (feature
    3) that turns the previous substep's lambda sum into the long solve's,
    clipped to [0, WARM_RATIO_MAX]; pairs without a previous sum are dropped."""
    f = np.asarray(features, np.float32).reshape(-1, MLP_FEATURES)
    t = np.asarray(targets, np.float64).reshape(-1)
    prev = f[:, 3].astype(np.float64)
    ok = prev < -RATIO_MIN_PREV
    s = np.clip(t[ok] / prev[ok], 0.0, WARM_RATIO_MAX)
    return f[ok], s.astype(np.float32)


@dataclass
class TrainResult:
    blob: np.ndarray
    losses: np.ndarray


def train(features: np.ndarray, targets: np.ndarray, epochs: int = 20, batch: int = 4096, lr: float = 2e-3,
          seed: int = 7, blob: Optional[np.ndarray] = None, over_weight: float = 1.0) -> TrainResult:
    """Fit the network; a given blob is fine-tuned (its normalisation kept).
    over_weight > 1: over-predictions cost more (a low-side expectile)."""
    if blob is None:
        blob = set_normalisation(mlp_init(seed), features, targets)
    out, losses = mlp_train(features, targets, blob, epochs=epochs, batch=batch, lr=lr, seed=seed,
                            over_weight=over_weight)
    return TrainResult(out, losses)


@dataclass
class Trajectory:
    density_error: np.ndarray   # mean density error after every substep
    height: np.ndarray          # mean particle height at the checkpoints
    kinetic: np.ndarray         # mean kinetic energy per unit mass at the checkpoints
    seconds: float = 0.0        # wall time of the substeps (no checkpoints: every = 0)
    fallbacks: int = 0          # learned substeps replaced by the plain solve


def run(scene: Scene, substeps: int, iterations: int, blob: Optional[np.ndarray] = None, relax: float = 1.0,
        fallback_threshold: float = 0.0, every: int = 20, max_neighbors: int = 128,
        mode: Optional[int] = None, fallback_iterations: int = 0) -> Trajectory:
    """Density error per substep (FP64 diagnostics on the GPU) and, every
    `every` substeps, the mean height and kinetic energy of the fluid
    (every = 0: no checkpoints, and the wall time of the substeps).  mode:
    WARM_LEARNED when a blob is given, else plain; WARM_PREVIOUS needs none."""
    sim = make_engine(scene, max_neighbors=max_neighbors)
    try:
        if mode is None:
            mode = WARM_LEARNED if blob is not None else WARM_PLAIN
        if mode != WARM_PLAIN:
            sim.set_warm_start(blob, mode, relax, fallback_threshold, fallback_iterations)
        sim.record_density_error(True)
        dt = scene.substep_dt()
        hs, ks = [], []
        sim.sync()
        t0 = time.perf_counter()
        for k in range(substeps):
            sim.step(dt, 1, iterations)
            if every and (k + 1) % every == 0:
                p, v = sim.read_positions(), sim.read_velocities()
                hs.append(float(p[:, 1].mean()))
                ks.append(float(0.5 * (v[:, :3].astype(np.float64) ** 2).sum(1).mean()))
        sim.sync()
        secs = time.perf_counter() - t0
        fb = sim.warm_start_stats()[1]
        return Trajectory(sim.density_history()[:, 0].copy(), np.array(hs), np.array(ks), secs, fb)
    finally:
        sim.close()


def fidelity(t: Trajectory, ref: Trajectory):
    """(max |mean height - ref| / kernel-free units, max |KE - ref KE| / max ref KE)."""
    return (float(np.abs(t.height - ref.height).max()),
            float(np.abs(t.kinetic - ref.kinetic).max() / max(ref.kinetic.max(), 1e-30)))


def iterations_to_reach(scene: Scene, substeps: int, target: float, blob: Optional[np.ndarray] = None,
                        relax: float = 1.0,
                        candidates: Sequence[int] = (1, 2, 3, 4, 5, 6, 8, 10, 12, 16, 20, 24, 32)):
    """Smallest iteration count whose mean (over substeps) density error is <= target,
    with its trajectory (None, None when no candidate reaches it)."""
    for it in candidates:
        t = run(scene, substeps, it, blob, relax)
        if float(np.mean(t.density_error)) <= target:
            return it, t
    return None, None


def dagger(scene_fn: Callable[[int], Scene], seeds: Sequence[int], substeps: int, long_iterations: int,
           iterations: int, rounds: int = 2, epochs: int = 20, batch: int = 4096, lr: float = 2e-3, seed: int = 7,
           skip: int = 0, stride: int = 1, relax: float = 1.0):
    """Collect along the plain trajectory, train, then collect again along the
    learned trajectory (so the data covers the states the learned solver
    visits) and retrain on everything."""
    F, T = [], []
    for s in seeds:
        f, t = collect(scene_fn(s), substeps, long_iterations, iterations, None, skip, stride)
        F.append(f)
        T.append(t)
    res = train(np.concatenate(F), np.concatenate(T), epochs, batch, lr, seed)
    for _ in range(rounds - 1):
        for s in seeds:
            f, t = collect(scene_fn(s), substeps, long_iterations, iterations, res.blob, skip, stride, relax)
            F.append(f)
            T.append(t)
        res = train(np.concatenate(F), np.concatenate(T), epochs, batch, lr, seed)
    return res, np.concatenate(F), np.concatenate(T)
