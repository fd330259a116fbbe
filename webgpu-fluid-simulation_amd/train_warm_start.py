"""Train the learned warm start on the dam-break family and evaluate it.

    python webgpu-fluid-simulation_amd/train_warm_start.py [--out DIR] [--quick]

    python webgpu-fluid-simulation_amd/train_warm_start.py --ratio --iterations 12 --eval-relax 0.5,0.6,0.7

    python webgpu-fluid-simulation_amd/train_warm_start.py --deep

Training data: dam breaks of several block shapes and seeds (r = 0.025,
h = 4r, 1/60 s frames in 2 substeps, XSPH 0.01, vorticity 0.01, tension
0.1 -- the physics of the 1M-particle dam break), collected along the
deployed solver's trajectory (plain, then learned: DAgger), targets = the
lambda summed over a 40-iteration Jacobi solve.  Evaluation on a held-out
block shape and seed: the mean density error per iteration count, plain
vs learned, and the trajectory fidelity (mean height, kinetic energy)
against a 40-iteration reference run.  Writes `<out>/dam_break.f32` (the
network blob, raw little-endian float32) and `<out>/dam_break.json` (how it
was made and what it measured).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import numpy as np  # noqa: E402

import learned  # noqa: E402
import scenes  # noqa: E402
from pbf_host import (WARM_PLAIN, WARM_PREVIOUS, WARM_PROJECT, WARM_RATIO, WARM_SMOOTH,  # noqa: E402
                      PBFError)

TRAIN_SHAPES = [(16, 16, 8), (20, 20, 10), (24, 16, 12), (28, 24, 14)]
EVAL_SHAPE = (24, 18, 12)
EVAL_SEED = 21


def family(shape, seed):
    nx, ny, nz = shape
    return scenes.dam_break(nx, ny, nz, map_res=32, vorticity=0.01, tension=0.1, seed=seed, name="warm_start")


def evaluate(ev, sub, blob, long_iterations, plain_its, learned_its, relaxes, log=print):
    """Rows of (solver, iterations, relax, density error, height / KE deviation
    from a long-solve reference run)."""
    ref = learned.run(ev, sub, long_iterations)
    rows = []
    for it in plain_its:
        tr = learned.run(ev, sub, it)
        dh, dk = learned.fidelity(tr, ref)
        rows.append({"solver": "plain", "iterations": it, "relax": None,
                     "density_error": float(tr.density_error.mean()), "height_dev": dh, "kinetic_dev": dk})
        log(f"plain   I={it:2d}            err {rows[-1]['density_error']:.5f}  dh {dh:.4f}  dKE {dk:.3f}")
    for it in learned_its:
        for rx in relaxes:
            try:
                tr = learned.run(ev, sub, it, blob, rx)
            except PBFError as e:  # e.g. a neighbour overflow of an unstable run
                rows.append({"solver": "learned", "iterations": it, "relax": rx, "error": str(e)})
                log(f"learned I={it:2d} relax {rx:.2f} failed: {e}")
                continue
            dh, dk = learned.fidelity(tr, ref)
            rows.append({"solver": "learned", "iterations": it, "relax": rx,
                         "density_error": float(tr.density_error.mean()), "height_dev": dh, "kinetic_dev": dk})
            log(f"learned I={it:2d} relax {rx:.2f} err {rows[-1]['density_error']:.5f}  dh {dh:.4f}  dKE {dk:.3f}")
    return float(ref.density_error.mean()), rows


def evaluate_modes(ev, sub, long_iterations, plain_its, runs, log=print):
    """Mean density error, ms/substep and fidelity of the plain solve at each
    count and of each (name, mode, blob, relax, iterations) run."""
    ref = learned.run(ev, sub, long_iterations)
    rows = []

    def row(name, it, rx, tr):
        dh, dk = learned.fidelity(tr, ref)
        rows.append({"solver": name, "iterations": it, "relax": rx, "density_error": float(tr.density_error.mean()),
                     "height_dev": dh, "kinetic_dev": dk})
        log(f"{name:9s} I={it:2d} relax {rx if rx is not None else '-'}: err {rows[-1]['density_error']:.5f}"
            f"  dh {dh:.4f}  dKE {dk:.3f}")

    for it in plain_its:
        row("plain", it, None, learned.run(ev, sub, it))
    for name, mode, blob, rx, its in runs:
        for it in its:
            try:
                row(name, it, rx, learned.run(ev, sub, it, blob, rx, mode=mode))
            except PBFError as e:
                rows.append({"solver": name, "iterations": it, "relax": rx, "error": str(e)})
                log(f"{name} I={it} relax {rx} failed: {e}")
    return float(ref.density_error.mean()), rows


def main_ratio(a, log):
    """Ratio mode: the network predicts the factor between the long solve's
    lambda sum and the previous substep's (PBF_WARM_RATIO).  Pairs are
    collected along the temporal warm start's trajectory, then (DAgger) along
    the ratio mode's own."""
    shapes = TRAIN_SHAPES[:2] if a.quick else TRAIN_SHAPES
    sub = 60 if a.quick else a.substeps
    t0 = time.time()
    jobs = [(s, 100 + 7 * k + j, 1.0) for k, s in enumerate(shapes) for j in range(2)]
    if a.medium and not a.quick:
        jobs.append((tuple(int(x) for x in a.medium.split(",")), 200, a.medium_keep))
    F, T = [], []

    def gather(mode, blob, tag):
        for shape, seed, keep in jobs:
            f, t = learned.collect(family(shape, seed), sub, a.long, a.iterations, blob, relax=a.relax, keep=keep,
                                   seed=seed + (1 if blob is not None else 0), mode=mode)
            fr, sr = learned.ratio_pairs(f, t)
            log(f"collect ({tag}) {shape} seed {seed}: {len(sr)} ratio pairs of {len(t)}, "
                f"s mean {sr.mean():.3f} std {sr.std():.3f}")
            F.append(fr)
            T.append(sr)

    gather(WARM_PREVIOUS, None, "temporal")
    res = learned.train(np.concatenate(F), np.concatenate(T), a.epochs)
    gather(WARM_RATIO, res.blob, "ratio")
    Fa, Ta = np.concatenate(F), np.concatenate(T)
    res = learned.train(Fa, Ta, a.epochs)
    t_train = time.time() - t0
    log(f"samples {len(Ta)}  loss {res.losses[0]:.4f} -> {res.losses[-1]:.4f}  ({t_train:.1f} s)")
    relaxes = [float(x) for x in a.eval_relax.split(",")]
    runs = [("previous", WARM_PREVIOUS, None, rx, (2, 3, 4)) for rx in relaxes] + \
           [("ratio", WARM_RATIO, res.blob, rx, (2, 3, 4)) for rx in relaxes]
    ref_err, rows = evaluate_modes(family(EVAL_SHAPE, EVAL_SEED), sub, a.long, (2, 3, 4, 5, 6, 8), runs, log)
    large = None
    if a.eval_large and not a.quick:
        shape = tuple(int(x) for x in a.eval_large.split(","))
        log(f"-- large evaluation {shape}, {a.large_substeps} substeps")
        its = tuple(int(x) for x in a.large_iterations.split(","))
        lruns = [("previous", WARM_PREVIOUS, None, rx, its) for rx in relaxes] + \
                [("ratio", WARM_RATIO, res.blob, rx, its) for rx in relaxes]
        lref, lrows = evaluate_modes(family(shape, EVAL_SEED), a.large_substeps, a.long, (16, 32, 48), lruns, log)
        large = {"shape": shape, "seed": EVAL_SEED, "substeps": a.large_substeps, "reference_density_error": lref,
                 "rows": lrows}
    os.makedirs(a.out, exist_ok=True)
    res.blob.astype("<f4").tofile(os.path.join(a.out, "dam_break_ratio.f32"))
    meta = {
        "network": "16-64-64-1 SiLU, bf16 weights, fp32 accumulate; output = factor on the previous lambda sum",
        "mode": "PBF_WARM_RATIO",
        "train": [{"shape": s, "seed": sd, "keep": k} for s, sd, k in jobs], "substeps": sub,
        "long_iterations": a.long, "deployed_iterations": a.iterations, "collect_relax": a.relax,
        "epochs": a.epochs, "samples": int(len(Ta)), "loss_first_last": [float(res.losses[0]), float(res.losses[-1])],
        "physics": {"particle_radius": 0.025, "kernel_radius": 0.1, "substep_dt": "1/120 s", "xsph": 0.01,
                    "vorticity": 0.01, "surface_tension": 0.1},
        "eval": {"shape": EVAL_SHAPE, "seed": EVAL_SEED, "substeps": sub, "reference_iterations": a.long,
                 "reference_density_error": ref_err, "rows": rows},
        "eval_large": large,
        "train_seconds": t_train,
    }
    with open(os.path.join(a.out, "dam_break_ratio.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    return 0


DEEP_SHAPES = [((32, 128, 16), 300, 1.0), ((64, 128, 32), 301, 0.25), ((48, 96, 24), 302, 0.5),
               ((24, 64, 24), 303, 1.0)]


def deep_family(shape, seed):
    """Tall dam-break columns (up to 128 layers, the depth of the 1M block) in
    boxes of side 2 x their largest extent, 64^3 maps."""
    return scenes.dam_break(*shape, map_res=64, vorticity=0.01, tension=0.1, seed=seed, name="deep")


def evaluate_against(ev, sub, refs, runs, log=print):
    """Rows of (solver, iterations, relax, mean density error, ms/substep,
    fallbacks, largest mean-height / kinetic-energy deviation from each
    reference run).  runs: (name, mode, blob, relax, iteration counts,
    fallback threshold, fallback iterations)."""
    rows = []
    for name, mode, blob, rx, its, thr, fbi in runs:
        for it in its:
            t = learned.run(ev, sub, it, blob, rx, every=0, mode=mode, fallback_threshold=thr,
                            fallback_iterations=fbi)
            f = learned.run(ev, sub, it, blob, rx, every=20, mode=mode, fallback_threshold=thr,
                            fallback_iterations=fbi)
            row = {"solver": name, "iterations": it, "relax": rx, "density_error": float(t.density_error.mean()),
                   "max_substep_error": float(t.density_error.max()), "ms_per_substep": 1e3 * t.seconds / sub,
                   "fallbacks": int(t.fallbacks)}
            for key, ref in refs.items():
                dh, dk = learned.fidelity(f, ref)
                row[f"height_dev_vs_{key}"], row[f"kinetic_dev_vs_{key}"] = dh, dk
            rows.append(row)
            log(f"{name:18s} I={it:2d} relax {rx if rx is not None else '-'}: err {row['density_error']:.4%} "
                f"{row['ms_per_substep']:.3f} ms fb {row['fallbacks']} | " +
                " | ".join(f"vs {k}: h {row[f'height_dev_vs_{k}']:.4f} KE {row[f'kinetic_dev_vs_{k}']:.3f}"
                           for k in refs))
    return rows


def main_deep(a, log):
    """Ratio network for deep fluid (the 1M dam break): targets are the lambda
    sums of a long (`--long`, 200) Jacobi solve, i.e. close to converged;
    over-predictions weigh `--over-weight` (9) in the loss, so the factor errs
    towards pushing too little; pairs come from tall columns along the
    temporal warm start's trajectory (projected accumulation, relax 0.3),
    then (DAgger) along the network's own solve pushed hard (ratio with
    projected accumulation at relax `--dagger-relax`, 1.0): that trajectory
    visits the over-pushed states where the factor must fall below 1, which
    the deployed solve must never reach.  Deployed with PBF_WARM_RATIO |
    PBF_WARM_SMOOTH at relax `--relax` (0.5)."""
    t0 = time.time()
    F, T = [], []

    def gather(mode, blob, relax, tag):
        for shape, seed, keep in DEEP_SHAPES:
            f, t = learned.collect(deep_family(shape, seed), a.substeps, a.long, a.iterations, blob, relax=relax,
                                   keep=keep, seed=seed, mode=mode)
            fr, sr = learned.ratio_pairs(f, t)
            log(f"collect ({tag}) {shape} seed {seed}: {len(sr)} ratio pairs of {len(t)}, "
                f"s mean {sr.mean():.3f} std {sr.std():.3f}")
            F.append(fr)
            T.append(sr)

    gather(WARM_PREVIOUS | WARM_PROJECT, None, 0.3, "temporal")
    res = learned.train(np.concatenate(F), np.concatenate(T), a.epochs, over_weight=a.over_weight)
    gather(WARM_RATIO | WARM_PROJECT, res.blob, a.dagger_relax, "ratio")
    Fa, Ta = np.concatenate(F), np.concatenate(T)
    res = learned.train(Fa, Ta, a.epochs, over_weight=a.over_weight)
    t_train = time.time() - t0
    log(f"samples {len(Ta)}  loss {res.losses[0]:.4f} -> {res.losses[-1]:.4f}  ({t_train:.1f} s)")
    os.makedirs(a.out, exist_ok=True)
    res.blob.astype("<f4").tofile(os.path.join(a.out, "deep_ratio.f32"))
    ev = scenes.dam_break_1m()
    sub = 100
    log(f"-- evaluation: the 1M dam break ({ev.n} particles, a block shape and seed not trained on), {sub} substeps")
    refs = {"converged_200": learned.run(ev, sub, 200), "plain_40": learned.run(ev, sub, 40)}
    plain = evaluate_against(ev, sub, refs, [("plain", WARM_PLAIN, None, None, (16, 20, 32, 40, 44), 0.0, 0)], log)
    thr = next(r["max_substep_error"] for r in plain if r["iterations"] == 44)
    runs = [("deep ratio+smooth", WARM_RATIO | WARM_SMOOTH, res.blob, a.relax, (16, 18, 20), thr, 44),
            ("previous+smooth", WARM_PREVIOUS | WARM_SMOOTH, None, a.relax, (16, 18, 20), thr, 44),
            ("previous+smooth", WARM_PREVIOUS | WARM_SMOOTH, None, 0.4, (24, 28), thr, 44)]
    rows = plain + evaluate_against(ev, sub, refs, runs, log)
    meta = {
        "network": "16-64-64-1 SiLU, bf16 weights, fp32 accumulate; output = factor on the previous lambda sum",
        "mode": "PBF_WARM_RATIO | PBF_WARM_SMOOTH", "relax": a.relax,
        "train": [{"shape": sh, "seed": sd, "keep": k} for sh, sd, k in DEEP_SHAPES], "substeps": a.substeps,
        "long_iterations": a.long, "deployed_iterations": a.iterations, "over_weight": a.over_weight,
        "dagger": {"mode": "PBF_WARM_RATIO | PBF_WARM_PROJECT", "relax": a.dagger_relax},
        "epochs": a.epochs, "samples": int(len(Ta)), "loss_first_last": [float(res.losses[0]), float(res.losses[-1])],
        "physics": {"particle_radius": 0.025, "kernel_radius": 0.1, "substep_dt": "1/120 s", "xsph": 0.01,
                    "vorticity": 0.01, "surface_tension": 0.1},
        "eval": {"scene": "dam_break_1m", "substeps": sub, "fallback_threshold": thr, "fallback_iterations": 44,
                 "rows": rows},
        "train_seconds": t_train,
    }
    with open(os.path.join(a.out, "deep_ratio.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    return 0


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(HERE, "warm_start"))
    ap.add_argument("--substeps", type=int, default=150)
    ap.add_argument("--long", type=int, default=40)
    ap.add_argument("--iterations", type=int, default=3)
    ap.add_argument("--relax", type=float, default=0.5)
    ap.add_argument("--epochs", type=int, default=20)
    ap.add_argument("--quick", action="store_true", help="fewer shapes and substeps")
    ap.add_argument("--medium", default="64,64,32", help="a larger block, subsampled ('' to skip)")
    ap.add_argument("--medium-keep", type=float, default=0.25)
    ap.add_argument("--large", default="", help="the large block in training too, subsampled ('' to skip)")
    ap.add_argument("--large-keep", type=float, default=0.03)
    ap.add_argument("--eval-relax", default="0.25,0.5,0.75")
    ap.add_argument("--eval-iterations", default="1,2,3,4,6")
    ap.add_argument("--eval-large", default="128,128,64", help="block of the large evaluation ('' to skip)")
    ap.add_argument("--large-substeps", type=int, default=200)
    ap.add_argument("--ratio", action="store_true", help="train the ratio mode (PBF_WARM_RATIO)")
    ap.add_argument("--large-iterations", default="12,16,20,24", help="ratio mode: iteration counts at the large block")
    ap.add_argument("--deep", action="store_true",
                    help="the deep-fluid ratio network (near-converged targets, asymmetric loss, smoothing); "
                         "defaults: --long 200 --iterations 20 --relax 0.5 --substeps 120")
    ap.add_argument("--over-weight", type=float, default=9.0, help="--deep: loss weight of over-predictions")
    ap.add_argument("--dagger-relax", type=float, default=1.0, help="--deep: relax of the DAgger collection")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)  # noqa: E731
    if a.deep:
        for k, v in (("long", 200), ("iterations", 20), ("relax", 0.5), ("substeps", 120)):
            if f"--{k}" not in (argv if argv is not None else sys.argv):
                setattr(a, k, v)
        return main_deep(a, log)
    if a.ratio:
        return main_ratio(a, log)
    shapes = TRAIN_SHAPES[:2] if a.quick else TRAIN_SHAPES
    sub = 60 if a.quick else a.substeps
    t0 = time.time()
    jobs = [(s, 100 + 7 * k + j, 1.0) for k, s in enumerate(shapes) for j in range(2)]
    if a.medium and not a.quick:
        jobs.append((tuple(int(x) for x in a.medium.split(",")), 200, a.medium_keep))
    if a.large and not a.quick:
        jobs.append((tuple(int(x) for x in a.large.split(",")), 300, a.large_keep))
    F, T = [], []
    for shape, seed, keep in jobs:
        f, t = learned.collect(family(shape, seed), sub, a.long, a.iterations, keep=keep, seed=seed)
        log(f"collect {shape} seed {seed}: {len(t)} pairs, dropped {learned.collect.last_dropped}, "
            f"target range [{t.min():.3g}, {t.max():.3g}]")
        F.append(f)
        T.append(t)
    res = learned.train(np.concatenate(F), np.concatenate(T), a.epochs)
    for shape, seed, keep in jobs:  # DAgger: states the learned solver visits
        f, t = learned.collect(family(shape, seed), sub, a.long, a.iterations, res.blob, relax=a.relax, keep=keep,
                               seed=seed + 1, fallback_threshold=0.1)
        log(f"collect (learned) {shape} seed {seed}: {len(t)} pairs, dropped {learned.collect.last_dropped}"
            + (f", stopped at {learned.collect.last_error}" if learned.collect.last_error else ""))
        F.append(f)
        T.append(t)
    F, T = np.concatenate(F), np.concatenate(T)
    res = learned.train(F, T, a.epochs)
    t_train = time.time() - t0
    log(f"samples {len(T)}  loss {res.losses[0]:.4f} -> {res.losses[-1]:.4f}  ({t_train:.1f} s)")
    relaxes = [float(x) for x in a.eval_relax.split(",")]
    ref_err, rows = evaluate(family(EVAL_SHAPE, EVAL_SEED), sub, res.blob, a.long, (1, 2, 3, 4, 5, 6, 8, 10, 12),
                             [int(x) for x in a.eval_iterations.split(",")], relaxes, log)
    large = None
    if a.eval_large and not a.quick:
        shape = tuple(int(x) for x in a.eval_large.split(","))
        log(f"-- large evaluation {shape}, {a.large_substeps} substeps")
        lref, lrows = evaluate(family(shape, EVAL_SEED), a.large_substeps, res.blob, a.long, (2, 3, 4, 5, 6, 8, 12),
                               (2, 3, 4), relaxes, log)
        large = {"shape": shape, "seed": EVAL_SEED, "substeps": a.large_substeps, "reference_density_error": lref,
                 "rows": lrows}
    os.makedirs(a.out, exist_ok=True)
    res.blob.astype("<f4").tofile(os.path.join(a.out, "dam_break.f32"))
    meta = {
        "network": "16-64-64-1 SiLU, bf16 weights, fp32 accumulate",
        "train": [{"shape": s, "seed": sd, "keep": k} for s, sd, k in jobs], "substeps": sub,
        "long_iterations": a.long, "deployed_iterations": a.iterations, "collect_relax": a.relax,
        "epochs": a.epochs, "samples": int(len(T)), "loss_first_last": [float(res.losses[0]), float(res.losses[-1])],
        "physics": {"particle_radius": 0.025, "kernel_radius": 0.1, "substep_dt": "1/120 s", "xsph": 0.01,
                    "vorticity": 0.01, "surface_tension": 0.1},
        "eval": {"shape": EVAL_SHAPE, "seed": EVAL_SEED, "substeps": sub, "reference_iterations": a.long,
                 "reference_density_error": ref_err, "rows": rows},
        "eval_large": large,
        "train_seconds": t_train,
    }
    with open(os.path.join(a.out, "dam_break.json"), "w") as fh:
        json.dump(meta, fh, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
