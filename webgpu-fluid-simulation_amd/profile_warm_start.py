"""One-context driver of the warm-start network's kernels, for kernel traces.

1. The 1M dam break (128 x 128 x 64) for 20 substeps in ratio mode (relax 0.45,
   32 iterations): every substep runs iteration 0's feature pass and
   k_mlp_infer on all 1,048,576 particles.
2. Two epochs of training on 262,144 synthetic (feature, target) pairs in
   mini-batches of 4,096 (k_mlp_train, k_mlp_adam, k_mlp_pack); the kernels'
   work does not depend on the values.

Usage: python3 profile_warm_start.py   (under rocprofv3 --kernel-trace --stats)
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
if HERE not in sys.path:
    sys.path.insert(0, HERE)

import learned  # noqa: E402
import scenes  # noqa: E402
from pbf_host import WARM_RATIO  # noqa: E402


def main():
    sc = scenes.dam_break_1m()
    blob = np.fromfile(os.path.join(HERE, "warm_start", "dam_break_ratio.f32"), "<f4")
    sim = scenes.make_engine(sc, max_neighbors=128)
    try:
        sim.set_warm_start(blob, WARM_RATIO, 0.45, 0.0)
        dt = sc.substep_dt()
        for _ in range(20):
            sim.step(dt, 1, 32)
        sim.sync()
        learned_substeps, fallbacks = sim.warm_start_stats()
        print(f"ratio mode: {learned_substeps} learned substeps, {fallbacks} fallbacks, n={sc.n}", flush=True)
    finally:
        sim.close()
    rng = np.random.default_rng(5)
    n = 262144
    features = rng.standard_normal((n, 16)).astype(np.float32)
    targets = rng.standard_normal(n).astype(np.float32) * 0.01
    res = learned.train(features, targets, epochs=2, batch=4096)
    print(f"training: 2 epochs x {n // 4096} mini-batches, final loss {res.losses[-1]:.4g}", flush=True)


if __name__ == "__main__":
    main()
