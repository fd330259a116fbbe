"""CPU restatement of the warm-start network (16 -> 64 -> 64 -> 1, SiLU).

TEST INFRASTRUCTURE ONLY: imported by tests/, which check the GPU network
kernels (libpbf's k_mlp_infer / k_mlp_train) against it.  The shipped library
never loads it.  The network has no counterpart in the reference (which
solves lambda from zero every substep, PBF.ts:419-431).

Blob layout (float32):
  W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3 | fmu[16] |
  fsig[16] | tmu | tsig | lmin
Forward:  x = (f - fmu) / fsig;  z1 = x W1 + b1;  a1 = silu(z1);
          z2 = a1 W2 + b2;  a2 = silu(z2);  y = a2 . w3 + b3
Loss:     mean over samples of (y - (t - tmu) / tsig)^2.

With ``quant=True`` the MFMA operands are rounded to bfloat16 where the GPU
rounds them (x, W1, a1, W2 for the forward pass; dz2, W2 for da1; a1 and dz2
for dW2; x and dz1 for dW1); every product of two bfloat16 values is exact
in float32, so what is left between this and the GPU is the order of the
float32 sums and the last bit of exp.  ``quant=False`` evaluates everything
in float64: that path is pinned against central finite differences of the
loss.
"""
from __future__ import annotations

import numpy as np

OFF_W1, OFF_B1, OFF_W2, OFF_B2, OFF_W3, OFF_B3 = 0, 1024, 1088, 5184, 5248, 5312
OFF_FMU, OFF_FSIG, OFF_TMU, OFF_TSIG, OFF_LMIN, BLOB, NPARAM = 5313, 5329, 5345, 5346, 5347, 5348, 5313


def bf16(x: np.ndarray) -> np.ndarray:
    """float32 -> bfloat16 -> float32, round to nearest even (finite inputs)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32)


def unpack(blob: np.ndarray, dtype=np.float64):
    b = np.asarray(blob, np.float32).astype(dtype)
    return dict(W1=b[OFF_W1:OFF_B1].reshape(16, 64), b1=b[OFF_B1:OFF_W2], W2=b[OFF_W2:OFF_B2].reshape(64, 64),
                b2=b[OFF_B2:OFF_W3], w3=b[OFF_W3:OFF_B3], b3=b[OFF_B3], fmu=b[OFF_FMU:OFF_FSIG],
                fsig=b[OFF_FSIG:OFF_TMU], tmu=b[OFF_TMU], tsig=b[OFF_TSIG], lmin=b[OFF_LMIN])


def _sig(z):
    return 1.0 / (1.0 + np.exp(-z))


def forward_backward(blob: np.ndarray, feats: np.ndarray, targ: np.ndarray, quant: bool = True):
    """Returns (y (n,), grad (NPARAM,), loss) like pbf_mlp_forward_backward."""
    dt = np.float32 if quant else np.float64
    q = bf16 if quant else (lambda a: a)
    P = unpack(blob, dt)
    f = np.asarray(feats, np.float32).reshape(-1, 16).astype(dt)
    t = np.asarray(targ, np.float32).reshape(-1).astype(dt)
    n = f.shape[0]
    x = q(((f - P["fmu"]) / P["fsig"]).astype(dt))
    W1, W2 = q(P["W1"]), q(P["W2"])
    z1 = (x.astype(np.float64) @ W1.astype(np.float64)).astype(dt) + P["b1"]
    s1 = _sig(z1)
    a1f = z1 * s1
    a1 = q(a1f.astype(dt))
    z2 = (a1.astype(np.float64) @ W2.astype(np.float64)).astype(dt) + P["b2"]
    s2 = _sig(z2)
    a2 = z2 * s2
    y = a2 @ P["w3"] + P["b3"]
    tn = (t - P["tmu"]) / P["tsig"]
    e = y - tn
    loss = float(np.mean(e.astype(np.float64) ** 2))
    dy = 2.0 * e / n
    g = np.zeros(NPARAM, np.float64)
    g[OFF_B3] = dy.sum()
    g[OFF_W3:OFF_B3] = (a2 * dy[:, None]).sum(0)
    dz2 = P["w3"][None, :] * dy[:, None] * (s2 * (1.0 + z2 * (1.0 - s2)))
    g[OFF_B2:OFF_W3] = dz2.sum(0)
    dz2q = q(dz2.astype(dt))
    da1 = (dz2q.astype(np.float64) @ W2.astype(np.float64).T)
    dz1 = da1 * (s1 * (1.0 + z1 * (1.0 - s1)))
    g[OFF_B1:OFF_W2] = dz1.sum(0)
    dz1q = q(dz1.astype(dt))
    g[OFF_W2:OFF_B2] = (a1.astype(np.float64).T @ dz2q.astype(np.float64)).reshape(-1)
    g[OFF_W1:OFF_B1] = (x.astype(np.float64).T @ dz1q.astype(np.float64)).reshape(-1)
    return y.astype(np.float64), g, loss


def loss_only(blob64: np.ndarray, feats: np.ndarray, targ: np.ndarray) -> float:
    """float64 loss for a float64 parameter vector (finite differences)."""
    b = np.asarray(blob64, np.float64)
    W1 = b[OFF_W1:OFF_B1].reshape(16, 64)
    W2 = b[OFF_W2:OFF_B2].reshape(64, 64)
    x = (np.asarray(feats, np.float64) - b[OFF_FMU:OFF_FSIG]) / b[OFF_FSIG:OFF_TMU]
    z1 = x @ W1 + b[OFF_B1:OFF_W2]
    a1 = z1 * _sig(z1)
    z2 = a1 @ W2 + b[OFF_B2:OFF_W3]
    a2 = z2 * _sig(z2)
    y = a2 @ b[OFF_W3:OFF_B3] + b[OFF_B3]
    tn = (np.asarray(targ, np.float64) - b[OFF_TMU]) / b[OFF_TSIG]
    return float(np.mean((y - tn) ** 2))


def predict_lambda(blob: np.ndarray, feats: np.ndarray, lam_plain: np.ndarray, quant: bool = True,
                   relax: float = 1.0) -> np.ndarray:
    """lambda applied in iteration 0 by the learned solve (k_mlp_infer):
    relax (y tsig + tmu), replaced by the plain lambda where it is positive
    or not finite (lambda = -C / (den + eps) is never positive), then
    bounded below by relax lmin."""
    P = unpack(blob, np.float64)
    y, _, _ = forward_backward(blob, feats, np.zeros(len(feats), np.float32), quant)
    lam = (relax * (y * P["tsig"] + P["tmu"])).astype(np.float32)
    bad = ~(lam <= 0.0) | ~np.isfinite(lam)
    lam = np.where(bad, np.asarray(lam_plain, np.float32), lam)
    return np.maximum(lam, np.float32(relax * P["lmin"]))
