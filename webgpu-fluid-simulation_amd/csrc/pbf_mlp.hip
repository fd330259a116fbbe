// pbf_mlp.hip -- learned warm start for lambda: a 16 -> 64 -> 64 -> 1 SiLU MLP
// with BF16 weights on MFMA (v_mfma_f32_32x32x8_bf16, FP32 accumulate),
// inference and training kernels.
//
// Matrices are used transposed: hidden units are the M dimension and
// particles the N dimension of every MFMA (32 x 32 tiles, K = 8 per
// instruction).  Lane l of a wavefront then owns particle l % 32 of the tile
// and the accumulator rows 8(i/4) + 4(l/32) + i%4; those rows are exactly the
// B-operand rows of the next layer's K steps, so activations pass from layer
// to layer in registers.  Weight gradients reduce over particles and go
// through LDS.
//
// Packed parameter blob (float32, 5348 floats):
//   W1[16][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3 | fmu[16] |
//   fsig[16] | tmu | tsig | lmin
// The first PBF_MLP_PARAMS (5313) entries are trained; the rest normalise
// features (x = (f - fmu) / fsig) and the target (lambda = y tsig + tmu),
// and bound the prediction from below (lambda >= lmin).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_internal.h"

namespace pbf {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f16v __attribute__((ext_vector_type(16)));

enum : uint32_t {
  OFF_W1 = 0, OFF_B1 = 1024, OFF_W2 = 1088, OFF_B2 = 5184, OFF_W3 = 5248, OFF_B3 = 5312,
  OFF_FMU = 5313, OFF_FSIG = 5329, OFF_TMU = 5345, OFF_TSIG = 5346, OFF_LMIN = 5347, BLOB = 5348, NPARAM = 5313
};

// float -> bfloat16, round to nearest even (gfx950's v_cvt_pk_bf16_f32; the
// same bits as adding 0x7fff + lsb and truncating, for every finite value)
__device__ __forceinline__ short f2bf(float f) { return __builtin_bit_cast(short, (__bf16)f); }
// Logistic function with the hardware exp2 and reciprocal (each about 1 ulp):
// the network is checked against its CPU model to 2e-3, far above this, and the
// library-wide correctly rounded division would make the 64 activations per
// lane and tile, not the matrix products, the cost of a tile.  Large |z|
// saturates to 0 or 1 (exp2 -> inf or 0).
__device__ __forceinline__ float sigm(float z) {
  return __builtin_amdgcn_rcpf(1.0f + __builtin_amdgcn_exp2f(z * -1.44269504088896341f));
}
__device__ __forceinline__ int crow(int i, int h) { return 8 * (i >> 2) + 4 * h + (i & 3); }

// Forward pass of one 32-particle tile.  Lane l: particle l % 32, half h = l / 32.
struct TileAct {
  f16v z1[2], z2[2];  // pre-activations (accumulator layout)
  s4 a1b[8];          // bf16 SiLU(z1) as the K steps of layer 2
  s4 xb[2];           // bf16 normalised features as the K steps of layer 1
  float y;            // network output (normalised units), both halves
};

__device__ __forceinline__ void tile_forward(const MlpDev &M, const float *__restrict__ feats, uint32_t p, bool valid,
                                             TileAct &T) {
  const int l = __lane_id(), h = l >> 5, c = l & 31;
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = ks * 8 + h * 4 + r;
      float v = valid ? (feats[(size_t)p * 16 + f] - M.misc[OFF_FMU - OFF_B3 + f]) *
                            __builtin_amdgcn_rcpf(M.misc[OFF_FSIG - OFF_B3 + f])
                      : 0.0f;
      T.xb[ks][r] = f2bf(v);
    }
  float part = 0.0f;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f16v acc = {0};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const s4 a = *reinterpret_cast<const s4 *>(M.W1T + (mt * 32 + c) * 16 + ks * 8 + h * 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, T.xb[ks], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int hid = mt * 32 + crow(i, h);
      const float z = acc[i] + M.b1[hid];
      T.z1[mt][i] = z;
      T.a1b[mt * 4 + (i >> 2)][i & 3] = f2bf(z * sigm(z));
    }
  }
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f16v acc = {0};
#pragma unroll
    for (int s = 0; s < 8; ++s) {
      const s4 a = *reinterpret_cast<const s4 *>(M.W2T + (mt * 32 + c) * 64 + s * 8 + h * 4);
      acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, T.a1b[s], acc, 0, 0, 0);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int hid = mt * 32 + crow(i, h);
      const float z = acc[i] + M.b2[hid];
      T.z2[mt][i] = z;
      part = fmaf(M.w3[hid], z * sigm(z), part);
    }
  }
  part += __shfl_xor(part, 32, 64);
  T.y = part + M.misc[0];
}

// Inference: lambda_hat = y * tsig + tmu for every slot; written to X[i].w
// (the lambda of Jacobi iteration 0) and to the accumulated-lambda array.
// Ratio mode (smax > 0): the output is a factor s on the previous substep's
// lambda sum (feature 3), clamped to [0, smax]: lambda = s * Lambda_prev; a
// network that outputs the constant relax is the temporal warm start.
// The parameters one lane of an inference wavefront uses, loaded once before
// its tile loop (they are the same for every tile): the A operands of both
// layers and the biases / output weights of its accumulator rows.
struct InferConst {
  s4 a1[2][2], a2[2][8];
  float b1[2][16], b2[2][16], w3[2][16];
  float fmu[8], fisig[8];
};

__device__ __forceinline__ void load_infer_const(const MlpDev &M, InferConst &K) {
  const int l = __lane_id(), h = l >> 5, c = l & 31;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) K.a1[mt][ks] = *reinterpret_cast<const s4 *>(M.W1T + (mt * 32 + c) * 16 + ks * 8 + h * 4);
#pragma unroll
    for (int q = 0; q < 8; ++q) K.a2[mt][q] = *reinterpret_cast<const s4 *>(M.W2T + (mt * 32 + c) * 64 + q * 8 + h * 4);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const int hid = mt * 32 + crow(i, h);
      K.b1[mt][i] = M.b1[hid];
      K.b2[mt][i] = M.b2[hid];
      K.w3[mt][i] = M.w3[hid];
    }
  }
#pragma unroll
  for (int ks = 0; ks < 2; ++ks)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int f = ks * 8 + h * 4 + r;
      K.fmu[ks * 4 + r] = M.misc[OFF_FMU - OFF_B3 + f];
      K.fisig[ks * 4 + r] = __builtin_amdgcn_rcpf(M.misc[OFF_FSIG - OFF_B3 + f]);
    }
}

// tile_forward's arithmetic with the parameters already in registers; returns
// the network output (normalised units) of this lane's particle.
__device__ __forceinline__ float tile_forward_infer(const InferConst &K, float b3, const float *__restrict__ feats,
                                                    uint32_t p, bool valid) {
  const int l = __lane_id(), h = l >> 5;
  s4 xb[2];
#pragma unroll
  for (int ks = 0; ks < 2; ++ks) {
    float4 v = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (valid) v = *reinterpret_cast<const float4 *>(feats + (size_t)p * 16 + ks * 8 + h * 4);
    xb[ks][0] = f2bf(valid ? (v.x - K.fmu[ks * 4 + 0]) * K.fisig[ks * 4 + 0] : 0.0f);
    xb[ks][1] = f2bf(valid ? (v.y - K.fmu[ks * 4 + 1]) * K.fisig[ks * 4 + 1] : 0.0f);
    xb[ks][2] = f2bf(valid ? (v.z - K.fmu[ks * 4 + 2]) * K.fisig[ks * 4 + 2] : 0.0f);
    xb[ks][3] = f2bf(valid ? (v.w - K.fmu[ks * 4 + 3]) * K.fisig[ks * 4 + 3] : 0.0f);
  }
  s4 a1b[8];
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f16v acc = {0};
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(K.a1[mt][ks], xb[ks], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float z = acc[i] + K.b1[mt][i];
      a1b[mt * 4 + (i >> 2)][i & 3] = f2bf(z * sigm(z));
    }
  }
  float part = 0.0f;
#pragma unroll
  for (int mt = 0; mt < 2; ++mt) {
    f16v acc = {0};
#pragma unroll
    for (int q = 0; q < 8; ++q) acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(K.a2[mt][q], a1b[q], acc, 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      const float z = acc[i] + K.b2[mt][i];
      part = fmaf(K.w3[mt][i], z * sigm(z), part);
    }
  }
  part += __shfl_xor(part, 32, 64);
  return part + b3;
}

__global__ void __launch_bounds__(256) k_mlp_infer(uint32_t n, const float *__restrict__ feats, MlpDev M,
                                                   float4 *X, float *__restrict__ lam_acc, float smax) {
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int l = __lane_id();
  const uint32_t tiles = (n + 31) / 32;
  InferConst K;
  load_infer_const(M, K);
  const float b3 = M.misc[0], tsig = M.misc[OFF_TSIG - OFF_B3], tmu = M.misc[OFF_TMU - OFF_B3];
  const float lmin = M.misc[OFF_LMIN - OFF_B3];
  for (uint32_t t = wave; t < tiles; t += waves) {
    const uint32_t p = t * 32 + (l & 31);
    const bool valid = p < n;
    const float y = tile_forward_infer(K, b3, feats, p, valid);
    if (valid && l < 32) {
      // lambda = -C / (den + eps) is never positive; a non-finite or positive
      // prediction keeps the plain lambda already in X[p].w
      float *wp = reinterpret_cast<float *>(X) + 4 * (size_t)p + 3;
      float lam = fmaf(y, tsig, tmu);
      if (smax > 0.0f) {
        lam = fminf(fmaxf(lam, 0.0f), smax) * feats[(size_t)p * 16 + 3];  // fmaxf(NaN, 0) = 0
      } else {
        if (!(lam <= 0.0f) || !isfinite(lam)) lam = *wp;
        lam = fmaxf(lam, lmin);
      }
      *wp = lam;
      if (lam_acc) lam_acc[p] = lam;
    }
  }
}

// Training: forward + backward over a batch (`idx` selects the samples, or
// NULL for samples 0 .. nb-1),
// MSE in normalised target units; gradients summed into `grad` (NPARAM
// floats, scaled by inv_batch) and the loss into loss[0].
constexpr int TW = 4;  // waves per block
struct WaveLds {
  short a1[64 * 32], dz2[64 * 32], dz1[64 * 32], x[32 * 32];
};

__global__ void __launch_bounds__(256) k_mlp_train(uint32_t nb, const uint32_t *__restrict__ idx,
                                                   const float *__restrict__ feats, const float *__restrict__ targ,
                                                   MlpDev M, float inv_batch, float over_w,
                                                   float *__restrict__ grad,
                                                   float *__restrict__ loss, float *__restrict__ yout) {
  __shared__ float gacc[NPARAM];
  __shared__ WaveLds wl[TW];
  for (uint32_t q = threadIdx.x; q < NPARAM; q += blockDim.x) gacc[q] = 0.0f;
  __syncthreads();
  const int l = __lane_id(), h = l >> 5, c = l & 31;
  const int w = threadIdx.x >> 6;
  WaveLds &L = wl[w];
  const uint32_t waves = (gridDim.x * blockDim.x) >> 6;
  const uint32_t wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const uint32_t tiles = (nb + 31) / 32;
  float lsum = 0.0f;
  for (uint32_t t = wave; t < tiles; t += waves) {
    const uint32_t q = t * 32 + c;
    const bool valid = q < nb;
    const uint32_t p = valid ? (idx ? idx[q] : q) : 0u;  // no index array: sample q
    TileAct T;
    tile_forward(M, feats, p, valid, T);
    const float tn = valid ? (targ[p] - M.misc[OFF_TMU - OFF_B3]) / M.misc[OFF_TSIG - OFF_B3] : 0.0f;
    const float e = valid ? T.y - tn : 0.0f;
    if (yout && valid && h == 0) yout[q] = T.y;
    // over-predictions (e > 0) weigh over_w in the loss
    const float wgt = e > 0.0f ? over_w : 1.0f;
    if (h == 0) lsum = fmaf(wgt * e, e, lsum);
    const float dy = 2.0f * (wgt * e) * inv_batch;
    // b3 and w3
    float db3 = h == 0 ? dy : 0.0f;
    // dz2 = w3 dy silu'(z2); dw3 = a2 dy; db2 = dz2
    s4 dz2b[8];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int hid = mt * 32 + crow(i, h);
        const float z = T.z2[mt][i];
        const float s = sigm(z);
        const float a2 = z * s;
        const float dz = M.w3[hid] * dy * (s * (1.0f + z * (1.0f - s)));
        dz2b[mt * 4 + (i >> 2)][i & 3] = f2bf(dz);
        float gw3 = a2 * dy, gb2 = dz;
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) {
          gw3 += __shfl_xor(gw3, o, 64);
          gb2 += __shfl_xor(gb2, o, 64);
        }
        if (c == 0) {
          atomicAdd(&gacc[OFF_W3 + hid], gw3);
          atomicAdd(&gacc[OFF_B2 + hid], gb2);
        }
      }
#pragma unroll
    for (int o = 16; o > 0; o >>= 1) db3 += __shfl_xor(db3, o, 64);
    if (l == 0) atomicAdd(&gacc[OFF_B3], db3);
    // da1 = W2 dz2 (A = W2[i][o], B = dz2 by K steps of o); dz1 = da1 silu'(z1)
    s4 dz1b[8];
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f16v acc = {0};
#pragma unroll
      for (int s = 0; s < 8; ++s) {
        const s4 a = *reinterpret_cast<const s4 *>(M.W2N + (mt * 32 + c) * 64 + s * 8 + h * 4);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, dz2b[s], acc, 0, 0, 0);
      }
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int hid = mt * 32 + crow(i, h);
        const float z = T.z1[mt][i];
        const float s = sigm(z);
        const float dz = acc[i] * (s * (1.0f + z * (1.0f - s)));
        dz1b[mt * 4 + (i >> 2)][i & 3] = f2bf(dz);
        float gb1 = dz;
#pragma unroll
        for (int o = 16; o > 0; o >>= 1) gb1 += __shfl_xor(gb1, o, 64);
        if (c == 0) atomicAdd(&gacc[OFF_B1 + hid], gb1);
      }
    }
    // stage a1, dz2, dz1 as [hidden][particle] and x as [feature][particle] (bf16)
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int hid = s * 8 + h * 4 + r;
        L.a1[hid * 32 + c] = T.a1b[s][r];
        L.dz2[hid * 32 + c] = dz2b[s][r];
        L.dz1[hid * 32 + c] = dz1b[s][r];
      }
#pragma unroll
    for (int ks = 0; ks < 2; ++ks)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        L.x[(ks * 8 + h * 4 + r) * 32 + c] = T.xb[ks][r];
        L.x[(16 + ks * 8 + h * 4 + r) * 32 + c] = 0;
      }
    // one wavefront writes and reads its own LDS slice: LDS operations of a
    // wave complete in order, the wait only keeps the compiler from hoisting
    __builtin_amdgcn_s_waitcnt(0);
    __builtin_amdgcn_wave_barrier();
    // dW2[i][o] = sum_p a1[i][p] dz2[o][p]: M = i, N = o, K = particles
#pragma unroll
    for (int mt = 0; mt < 2; ++mt)
#pragma unroll
      for (int nt = 0; nt < 2; ++nt) {
        f16v acc = {0};
#pragma unroll
        for (int ks = 0; ks < 4; ++ks) {
          const s4 a = *reinterpret_cast<const s4 *>(&L.a1[(mt * 32 + c) * 32 + ks * 8 + h * 4]);
          const s4 b = *reinterpret_cast<const s4 *>(&L.dz2[(nt * 32 + c) * 32 + ks * 8 + h * 4]);
          acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 16; ++i) atomicAdd(&gacc[OFF_W2 + (mt * 32 + crow(i, h)) * 64 + nt * 32 + c], acc[i]);
      }
    // dW1[f][hid] = sum_p x[f][p] dz1[hid][p]: M = hid, N = f (padded to 32)
#pragma unroll
    for (int mt = 0; mt < 2; ++mt) {
      f16v acc = {0};
#pragma unroll
      for (int ks = 0; ks < 4; ++ks) {
        const s4 a = *reinterpret_cast<const s4 *>(&L.dz1[(mt * 32 + c) * 32 + ks * 8 + h * 4]);
        const s4 b = *reinterpret_cast<const s4 *>(&L.x[c * 32 + ks * 8 + h * 4]);
        acc = __builtin_amdgcn_mfma_f32_32x32x8bf16_1k(a, b, acc, 0, 0, 0);
      }
      if (c < 16) {
#pragma unroll
        for (int i = 0; i < 16; ++i) atomicAdd(&gacc[OFF_W1 + c * 64 + mt * 32 + crow(i, h)], acc[i]);
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_xor(lsum, o, 64);
  __syncthreads();
  for (uint32_t q = threadIdx.x; q < NPARAM; q += blockDim.x) atomicAdd(&grad[q], gacc[q]);
  if (l == 0) atomicAdd(loss, lsum * inv_batch);
}

// Adam on the master weights, then the bf16 / f32 device copies.
__global__ void k_mlp_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                           float *__restrict__ grad, float lr, float b1c, float b2c) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= NPARAM) return;
  const float g = grad[q];
  const float mq = 0.9f * m[q] + 0.1f * g;
  const float vq = 0.999f * v[q] + 0.001f * g * g;
  m[q] = mq;
  v[q] = vq;
  w[q] -= lr * (mq / b1c) / (sqrtf(vq / b2c) + 1e-8f);
  grad[q] = 0.0f;
}

__global__ void k_mlp_pack(const float *__restrict__ w, short *__restrict__ W1T, short *__restrict__ W2T,
                           short *__restrict__ W2N, float *__restrict__ f32) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < 1024) {  // W1[f][h] -> W1T[h][f]
    const uint32_t f = q / 64, hh = q % 64;
    W1T[hh * 16 + f] = f2bf(w[OFF_W1 + q]);
  }
  if (q < 4096) {  // W2[i][o] -> W2T[o][i], W2N[i][o]
    const uint32_t i = q / 64, o = q % 64;
    const short b = f2bf(w[OFF_W2 + q]);
    W2T[o * 64 + i] = b;
    W2N[q] = b;
  }
  if (q < BLOB - OFF_B1) {  // b1 .. lmin as f32 (W2 region copied too, unused)
    f32[q] = w[OFF_B1 + q];
  }
}

__global__ void k_mlp_zero(float *__restrict__ p, uint32_t n) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < n) p[q] = 0.0f;
}

__global__ void k_mlp_copy_out(const float *__restrict__ gacc, const float *__restrict__ lacc,
                               float *__restrict__ grad, float *__restrict__ loss) {
  const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
  if (grad && q < NPARAM) grad[q] = gacc[q];
  if (loss && q == 0) loss[0] = lacc[0];
}

// Scratch of the allocation-free forward + backward: the bf16 packings of W1
// (transposed) and W2 (both layouts), the float32 vectors b1 .. lmin, the
// gradient accumulator and the loss (byte offsets below; 57,344 bytes per
// GPU, independent of the sample count).  Host code reaches it only through
// the per-device address from hipGetSymbolAddress.
constexpr size_t kScrW1T = 0;
constexpr size_t kScrW2T = kScrW1T + 1024 * 2;
constexpr size_t kScrW2N = kScrW2T + 4096 * 2;
constexpr size_t kScrF32 = kScrW2N + 4096 * 2;
constexpr size_t kScrGrad = (kScrF32 + (size_t)(BLOB - OFF_B1) * 4 + 255) & ~(size_t)255;
constexpr size_t kScrLoss = kScrGrad + (size_t)NPARAM * 4;
constexpr size_t kScrBytes = (kScrLoss + 4 + 255) & ~(size_t)255;
static_assert(kScrBytes == 57344, "scratch size documented in include/pbf.h");
static __device__ __align__(256) unsigned char scratch[kScrBytes];

}  // namespace pbf

// ===========================================================================
// Host side.
using namespace pbf;

namespace {

struct MlpBuffers {
  short *W1T = nullptr, *W2T = nullptr, *W2N = nullptr;
  float *f32 = nullptr;  // b1 .. tsig
  float *w = nullptr;    // master blob
  bool ok = false;
};

hipError_t mlp_alloc(MlpBuffers &B) {
  hipError_t e;
  if ((e = hipMalloc((void **)&B.W1T, 1024 * 2)) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&B.W2T, 4096 * 2)) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&B.W2N, 4096 * 2)) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&B.f32, (BLOB - OFF_B1) * 4)) != hipSuccess) return e;
  if ((e = hipMalloc((void **)&B.w, BLOB * 4)) != hipSuccess) return e;
  B.ok = true;
  return hipSuccess;
}

void mlp_free(MlpBuffers &B) {
  void *p[] = {B.W1T, B.W2T, B.W2N, B.f32, B.w};
  for (void *q : p)
    if (q) (void)hipFree(q);
  B = MlpBuffers();
}

MlpDev mlp_view(const MlpBuffers &B) {
  MlpDev M;
  M.W1T = B.W1T;
  M.W2T = B.W2T;
  M.W2N = B.W2N;
  M.b1 = B.f32;
  M.b2 = B.f32 + (OFF_B2 - OFF_B1);
  M.w3 = B.f32 + (OFF_W3 - OFF_B1);
  M.misc = B.f32 + (OFF_B3 - OFF_B1);
  return M;
}

hipError_t mlp_upload(MlpBuffers &B, const float *blob, hipStream_t st) {
  hipError_t e = hipMemcpyAsync(B.w, blob, BLOB * 4, hipMemcpyHostToDevice, st);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_mlp_pack, dim3(17), dim3(256), 0, st, (const float *)B.w, B.W1T, B.W2T, B.W2N, B.f32);
  return hipGetLastError();
}

}  // namespace

namespace pbf {

// Engine hooks (pbf_engine.hip).
int mlp_engine_set(void **handle, const float *blob, hipStream_t st) {
  MlpBuffers *B = static_cast<MlpBuffers *>(*handle);
  if (!B) {
    B = new MlpBuffers();
    if (mlp_alloc(*B) != hipSuccess) {
      mlp_free(*B);
      delete B;
      return PBF_E_DEVICE;
    }
    *handle = B;
  }
  return mlp_upload(*B, blob, st) == hipSuccess ? PBF_OK : PBF_E_DEVICE;
}

// The applied lambda is relax * (y tsig + tmu): tmu and tsig of the device
// copy are rescaled from the master blob.
__global__ void k_mlp_relax(const float *__restrict__ w, float *__restrict__ f32, float relax) {
  if (threadIdx.x == 0) {
    f32[OFF_TMU - OFF_B1] = relax * w[OFF_TMU];
    f32[OFF_TSIG - OFF_B1] = relax * w[OFF_TSIG];
    f32[OFF_LMIN - OFF_B1] = relax * w[OFF_LMIN];
  }
}

int mlp_engine_relax(void *handle, float relax, hipStream_t st) {
  MlpBuffers *B = static_cast<MlpBuffers *>(handle);
  if (!B) return PBF_E_STATE;
  hipLaunchKernelGGL(k_mlp_relax, dim3(1), dim3(64), 0, st, (const float *)B->w, B->f32, relax);
  return hipGetLastError() == hipSuccess ? PBF_OK : PBF_E_DEVICE;
}

void mlp_engine_free(void **handle) {
  MlpBuffers *B = static_cast<MlpBuffers *>(*handle);
  if (B) {
    mlp_free(*B);
    delete B;
  }
  *handle = nullptr;
}

void launch_mlp_infer(hipStream_t st, void *handle, uint32_t n, const float *feats, float4 *X, float *lam_acc,
                      float smax, hipEvent_t e0, hipEvent_t e1) {
  MlpBuffers *B = static_cast<MlpBuffers *>(handle);
  if (!B || !n) return;
  const uint32_t tiles = (n + 31) / 32;
  // one resident round (2 wavefronts per SIMD at this kernel's register use,
  // 256 CUs): each wavefront loads the parameters once for all its tiles
  const uint32_t blocks = std::min<uint32_t>((tiles + 3) / 4, 512u);
  if (e0 || e1)
    hipExtLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, e0, e1, 0, n, feats, mlp_view(*B), X, lam_acc,
                          smax);
  else
    hipLaunchKernelGGL(k_mlp_infer, dim3(blocks), dim3(256), 0, st, n, feats, mlp_view(*B), X, lam_acc, smax);
}

}  // namespace pbf

namespace {
void init_blob(float *w, uint64_t seed) {
  // Glorot-uniform weights from the seeded SplitMix64 stream, zero biases,
  // identity normalisation, no lower bound (callers overwrite fmu, fsig,
  // tmu, tsig and lmin).
  uint64_t s = seed;
  auto uni = [&](double a) {
    s = splitmix64(s);
    return (float)((((double)(s >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0) * a);
  };
  std::memset(w, 0, BLOB * 4);
  const double a1 = std::sqrt(6.0 / (16 + 64)), a2 = std::sqrt(6.0 / (64 + 64)), a3 = std::sqrt(6.0 / (64 + 1));
  for (int q = 0; q < 1024; ++q) w[OFF_W1 + q] = uni(a1);
  for (int q = 0; q < 4096; ++q) w[OFF_W2 + q] = uni(a2);
  for (int q = 0; q < 64; ++q) w[OFF_W3 + q] = uni(a3);
  for (int f = 0; f < 16; ++f) w[OFF_FSIG + f] = 1.0f;
  w[OFF_TSIG] = 1.0f;
  w[OFF_LMIN] = -3.0e38f;
}

struct TrainState {
  MlpBuffers B;
  float *m = nullptr, *v = nullptr, *grad = nullptr, *loss = nullptr, *feats = nullptr, *targ = nullptr,
        *yout = nullptr;
  uint32_t *idx = nullptr;
  hipStream_t st = nullptr;
  ~TrainState() {
    mlp_free(B);
    void *p[] = {m, v, grad, loss, feats, targ, yout, idx};
    for (void *q : p)
      if (q) (void)hipFree(q);
    if (st) (void)hipStreamDestroy(st);
  }
};

#define TCHK(x)                          \
  do {                                   \
    if ((x) != hipSuccess) return PBF_E_DEVICE; \
  } while (0)

int train_setup(TrainState &S, const float *features, const float *targets, uint64_t n, const float *blob) {
  TCHK(hipStreamCreateWithFlags(&S.st, hipStreamNonBlocking));
  TCHK(mlp_alloc(S.B));
  TCHK(hipMalloc((void **)&S.m, NPARAM * 4));
  TCHK(hipMalloc((void **)&S.v, NPARAM * 4));
  TCHK(hipMalloc((void **)&S.grad, NPARAM * 4));
  TCHK(hipMalloc((void **)&S.loss, 4));
  TCHK(hipMalloc((void **)&S.feats, n * 16 * 4));
  TCHK(hipMalloc((void **)&S.targ, n * 4));
  TCHK(hipMalloc((void **)&S.yout, n * 4));
  TCHK(hipMalloc((void **)&S.idx, n * 4));
  TCHK(hipMemsetAsync(S.m, 0, NPARAM * 4, S.st));
  TCHK(hipMemsetAsync(S.v, 0, NPARAM * 4, S.st));
  TCHK(hipMemsetAsync(S.grad, 0, NPARAM * 4, S.st));
  TCHK(hipMemcpyAsync(S.feats, features, n * 16 * 4, hipMemcpyHostToDevice, S.st));
  TCHK(hipMemcpyAsync(S.targ, targets, n * 4, hipMemcpyHostToDevice, S.st));
  TCHK(mlp_upload(S.B, blob, S.st));
  return PBF_OK;
}

void launch_train(TrainState &S, uint32_t nb, const uint32_t *idx, float inv_batch, float *yout,
                  float over_w = 1.0f) {
  const uint32_t tiles = (nb + 31) / 32;
  const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((tiles + TW - 1) / TW, 512u));
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(64 * TW), 0, S.st, nb, idx, (const float *)S.feats,
                     (const float *)S.targ, mlp_view(S.B), inv_batch, over_w, S.grad, S.loss, yout);
}

// Address of the scratch array on the current device, looked up once per
// device (each GPU holds its own copy of a __device__ array).
std::mutex g_scratch_mu;
constexpr int kMaxDevices = 64;
unsigned char *g_scratch_dev[kMaxDevices] = {};

int mlp_fail(int code, const std::string &msg) {
  set_thread_error(msg);
  return code;
}

int scratch_on_device(unsigned char **out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return mlp_fail(PBF_E_DEVICE, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return mlp_fail(PBF_E_DEVICE, "device index above 63");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return mlp_fail(PBF_E_DEVICE, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_dev[dev] = static_cast<unsigned char *>(p);
  }
  *out = g_scratch_dev[dev];
  return PBF_OK;
}

}  // namespace

extern "C" {

uint32_t pbf_mlp_blob_size(void) { return BLOB; }

int pbf_mlp_init(uint64_t seed, float *blob) {
  if (!blob) return PBF_E_INVALID;
  init_blob(blob, seed);
  return PBF_OK;
}

int pbf_mlp_forward_backward(const float *blob, const float *features, const float *targets, uint32_t n, float *y,
                             float *grad, float *loss) {
  if (!blob || !features || !targets || n == 0) return PBF_E_INVALID;
  TrainState S;
  int rc = train_setup(S, features, targets, n, blob);
  if (rc != PBF_OK) return rc;
  std::vector<uint32_t> id(n);
  for (uint32_t q = 0; q < n; ++q) id[q] = q;
  TCHK(hipMemcpyAsync(S.idx, id.data(), n * 4, hipMemcpyHostToDevice, S.st));
  TCHK(hipMemsetAsync(S.loss, 0, 4, S.st));
  launch_train(S, n, S.idx, 1.0f / (float)n, S.yout);
  TCHK(hipGetLastError());
  if (y) TCHK(hipMemcpyAsync(y, S.yout, n * 4, hipMemcpyDeviceToHost, S.st));
  if (grad) TCHK(hipMemcpyAsync(grad, S.grad, NPARAM * 4, hipMemcpyDeviceToHost, S.st));
  float l = 0.0f;
  TCHK(hipMemcpyAsync(&l, S.loss, 4, hipMemcpyDeviceToHost, S.st));
  TCHK(hipStreamSynchronize(S.st));
  if (loss) *loss = l;
  return PBF_OK;
}

int pbf_mlp_forward_backward_noalloc(const float *blob, const float *features, const float *targets, uint32_t n,
                                     float *y, float *grad, float *loss, void *stream) {
  if (!blob || !features || !targets || n == 0)
    return mlp_fail(PBF_E_INVALID, "blob, features and targets are required and n must be >= 1");
  if (n > PBF_MLP_NOALLOC_MAX_SAMPLES)
    return mlp_fail(PBF_E_CAPACITY, "n above PBF_MLP_NOALLOC_MAX_SAMPLES (2^31 samples)");
  unsigned char *scr = nullptr;
  int rc = scratch_on_device(&scr);
  if (rc != PBF_OK) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream);
  MlpBuffers B;  // views into the scratch, nothing allocated
  B.W1T = reinterpret_cast<short *>(scr + kScrW1T);
  B.W2T = reinterpret_cast<short *>(scr + kScrW2T);
  B.W2N = reinterpret_cast<short *>(scr + kScrW2N);
  B.f32 = reinterpret_cast<float *>(scr + kScrF32);
  float *gacc = reinterpret_cast<float *>(scr + kScrGrad), *lacc = reinterpret_cast<float *>(scr + kScrLoss);
  // the scratch keeps the previous call's contents: repack, then clear the
  // accumulators (gradient and loss are adjacent) with a kernel, which a
  // captured graph replays correctly
  hipLaunchKernelGGL(k_mlp_pack, dim3(17), dim3(256), 0, st, blob, B.W1T, B.W2T, B.W2N, B.f32);
  hipLaunchKernelGGL(k_mlp_zero, dim3((NPARAM + 1 + 255) / 256), dim3(256), 0, st, gacc, (uint32_t)(NPARAM + 1));
  const uint32_t tiles = (n + 31) / 32;
  const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>((tiles + TW - 1) / TW, 512u));
  hipLaunchKernelGGL(k_mlp_train, dim3(blocks), dim3(64 * TW), 0, st, n, (const uint32_t *)nullptr, features,
                     targets, mlp_view(B), 1.0f / (float)n, 1.0f, gacc, lacc, y);
  if (grad || loss)
    hipLaunchKernelGGL(k_mlp_copy_out, dim3((NPARAM + 255) / 256), dim3(256), 0, st, (const float *)gacc,
                       (const float *)lacc, grad, loss);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return mlp_fail(PBF_E_DEVICE, std::string("launch: ") + hipGetErrorString(e));
  return PBF_OK;
}

int pbf_mlp_train(const float *features, const float *targets, uint64_t n, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, float *blob, float *loss_per_epoch) {
  return pbf_mlp_train_weighted(features, targets, n, epochs, batch, lr, seed, 1.0f, blob, loss_per_epoch);
}

int pbf_mlp_train_weighted(const float *features, const float *targets, uint64_t n, uint32_t epochs, uint32_t batch,
                           float lr, uint64_t seed, float over_weight, float *blob, float *loss_per_epoch) {
  if (!features || !targets || !blob || n == 0 || batch == 0 || n > 0xFFFFFFFFull) return PBF_E_INVALID;
  if (!(over_weight > 0.0f) || !std::isfinite(over_weight)) return PBF_E_INVALID;
  TrainState S;
  int rc = train_setup(S, features, targets, n, blob);
  if (rc != PBF_OK) return rc;
  std::vector<uint32_t> perm(n);
  uint64_t rs = seed ^ 0xA5A5A5A5ull;
  uint64_t t = 0;
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    for (uint64_t q = 0; q < n; ++q) perm[q] = (uint32_t)q;
    for (uint64_t q = n - 1; q > 0; --q) {  // Fisher-Yates with the seeded stream
      rs = splitmix64(rs);
      uint64_t r = rs % (q + 1);
      std::swap(perm[q], perm[r]);
    }
    TCHK(hipMemcpyAsync(S.idx, perm.data(), n * 4, hipMemcpyHostToDevice, S.st));
    TCHK(hipMemsetAsync(S.loss, 0, 4, S.st));
    const uint64_t nbatches = (n + batch - 1) / batch;
    for (uint64_t b = 0; b < nbatches; ++b) {
      const uint32_t start = (uint32_t)(b * batch);
      const uint32_t nb = (uint32_t)std::min<uint64_t>(batch, n - start);
      launch_train(S, nb, S.idx + start, 1.0f / (float)nb, nullptr, over_weight);
      ++t;
      const float b1c = 1.0f - std::pow(0.9f, (float)t), b2c = 1.0f - std::pow(0.999f, (float)t);
      hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, S.st, S.B.w, S.m, S.v, S.grad, lr,
                         b1c, b2c);
      hipLaunchKernelGGL(k_mlp_pack, dim3(17), dim3(256), 0, S.st, (const float *)S.B.w, S.B.W1T, S.B.W2T, S.B.W2N,
                         S.B.f32);
    }
    TCHK(hipGetLastError());
    if (loss_per_epoch) {
      float l = 0.0f;
      TCHK(hipMemcpyAsync(&l, S.loss, 4, hipMemcpyDeviceToHost, S.st));
      TCHK(hipStreamSynchronize(S.st));
      loss_per_epoch[ep] = l / (float)nbatches;
    }
  }
  TCHK(hipMemcpyAsync(blob, S.B.w, NPARAM * 4, hipMemcpyDeviceToHost, S.st));
  TCHK(hipStreamSynchronize(S.st));
  return PBF_OK;
}

}  // extern "C"
