// pbf_kernels.hip -- the PBF substep as hand-written CDNA4 kernels.
//
// Data lives in structure-of-float4 arrays that are physically re-sorted into
// cell order every substep, so every neighbour gather hits a short run of
// nearby slots (L2-resident per XCD, see xcd_block).  Pass order of one
// substep (reference order: src/simulator/PBF/PBF.ts:398-448):
//
//   predict_bin   x* = clamp(x + dt (v + dt (g + a))), cell key, wave-
//                 aggregated atomic rank                     (K1 + K2)
//   scan          counts -> offsets, LDS Blelloch, 3 kernels (K3)
//   scatter       (id, source slot) to the cell-sorted slot  (K4)
//   settle        stable order inside each cell (ascending id), move data
//   build_list    27-cell walk on frozen x*, slot-major list (K5 + K6 + K7)
//   lambda        boundary sample + density constraint + lambda (K8 + K9)
//   dp            position correction + clamp (K10 + K11)    x I
//   density       density + velocity (K12)
//   vorticity     omega and surface normal (K13)
//   xsph          XSPH + vorticity force + surface tension (K14)
//   diag          FP64 density-constraint error (optional)
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>

#include "pbf_device.h"
#include "pbf_internal.h"

namespace pbf {

constexpr int BS = 256;

template <bool MULTI>
__device__ __forceinline__ const Scene &scene_at(const Scene &s0, const Scene *__restrict__ scenes,
                                                 const uint32_t *__restrict__ sid, uint32_t i) {
  if constexpr (MULTI) return scenes[sid[i]];
  else return s0;
}

__device__ __forceinline__ uint32_t slot_index() {
  return xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
}

// The neighbour table is slot-major: entry k of particle i sits at
// nbr[k * ld + i], so one load instruction of a wavefront reads 256
// contiguous bytes.  A wavefront whose entries all lie within +-32767 slots
// of their particles is stored instead as 16-bit offsets j - i + 32768
// (wfmt[wave] = 1): half the bytes for every walk.  Entries 8c..8c+7 of slot
// i form one 16-byte chunk at chunk c * ld + i, so a lane reads eight
// entries with one load and a wavefront's load covers 1 KB of consecutive
// chunks.  Dense grids list neighbours in ascending slot order, so the first
// and last entries bound the offsets; hashed grids take the range over every
// entry.
__device__ __forceinline__ bool wave16(const Phys &P, uint32_t i) {
  return __builtin_amdgcn_readfirstlane(P.wfmt[i >> 6]) != 0u;
}
// Word holding 16-bit entries k and k + 1 (k even) of slot i.
__device__ __forceinline__ size_t chunk_word(uint32_t ld, uint32_t i, uint32_t k) {
  return ((size_t)(k >> 3) * ld + i) * 4u + ((k & 7u) >> 1);
}

// Per-lane walk (learned-feature and diagnostic passes): visits the
// neighbours in walk order.
template <int UNROLL = 4, typename F>
__device__ __forceinline__ void for_each_neighbor(const uint32_t *__restrict__ nbr, uint32_t i, const Phys &P,
                                                  uint32_t cnt, F &&f) {
  const uint32_t ld = P.ld;
  if (wave16(P, i)) {
    const uint32_t base = i - 32768u;
    uint32_t k = 0;
    for (; k + 2 <= cnt; k += 2) {
      const uint32_t w = P.nbr16[chunk_word(ld, i, k)];
      f(base + (w & 0xffffu));
      f(base + (w >> 16));
    }
    if (k < cnt) f(base + (P.nbr16[chunk_word(ld, i, k)] & 0xffffu));
    return;
  }
  const uint32_t *row = nbr + i;
  uint32_t k = 0;
  if constexpr (UNROLL >= 8) {
    for (; k + 8 <= cnt; k += 8) {  // eight index loads in flight, then eight bodies
      uint32_t j[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) j[q] = row[(size_t)(k + q) * ld];
#pragma unroll
      for (int q = 0; q < 8; ++q) f(j[q]);
    }
  }
  if constexpr (UNROLL >= 4) {
    for (; k + 4 <= cnt; k += 4) {  // four index loads in flight, then four bodies
      const uint32_t j0 = row[(size_t)k * ld], j1 = row[(size_t)(k + 1) * ld];
      const uint32_t j2 = row[(size_t)(k + 2) * ld], j3 = row[(size_t)(k + 3) * ld];
      f(j0); f(j1); f(j2); f(j3);
    }
  } else {
    for (; k + 2 <= cnt; k += 2) {
      const uint32_t j0 = row[(size_t)k * ld], j1 = row[(size_t)(k + 1) * ld];
      f(j0); f(j1);
    }
  }
  for (; k < cnt; ++k) f(row[(size_t)k * ld]);
}

// Largest value among the active lanes (values below 2^16), bit by bit from
// ballots, as a scalar: lanes that left the kernel early take no part.
__device__ __forceinline__ uint32_t active_max(uint32_t v) {
  uint32_t m = 0;
#pragma unroll
  for (int b = 15; b >= 0; --b) {
    const uint32_t t = m | (1u << b);
    if (__ballot(v >= t)) m = t;
  }
  return __builtin_amdgcn_readfirstlane(m);
}

// Raw buffer gathers (gfx9 descriptor word 3 = 0x00020000): a neighbour's
// record is addressed by a 32-bit lane offset j * 16 on a scalar base
// instead of 64-bit per-lane address arithmetic; rows of the neighbour table
// by a scalar row offset on a per-lane column offset.  Capacities stay below
// 2^27 slots, so every offset fits 32 bits.
constexpr int kBufWord3 = 0x00020000;
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x3 __attribute__((ext_vector_type(3)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));
struct Gather {
  __amdgpu_buffer_rsrc_t r;
  __device__ Gather(const float4 *p, uint32_t n)
      : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<float4 *>(p), 0, (int)(n * 16u), kBufWord3)) {}
  __device__ float4 operator[](uint32_t j) const {
    const f32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, j * 16u, 0, 0);
    return make_float4(v.x, v.y, v.z, v.w);
  }
  __device__ float3 xyz(uint32_t j, uint32_t q = 0) const {  // record j + q, q uniform
    const f32x3 v = __builtin_amdgcn_raw_buffer_load_b96(r, j * 16u, q * 16u, 0);
    return make_float3(v.x, v.y, v.z);
  }
};

// Wavefront-uniform walk: every lane runs as many bodies as the largest count
// among the active lanes of its wavefront -- batches of UNROLL, then one of 4,
// then singles, each loop test a ballot -- instead of each lane's own
// 8 / 4 / 1 split, which costs the wavefront the maximum of every phase
// (39 bodies for counts 29..33).  The list build fills rows [count, largest
// count of the wavefront) with the lane's own slot, so entries past a lane's
// count are visited as (i, false) and the body makes them contribute exactly
// nothing (d = 0: every gradient and difference term is 0; a poly6 sum takes
// them with weight 0).  Valid entries are visited in walk order, so every sum
// is the same as the per-lane loop's.
template <int UNROLL, int UNROLL32 = UNROLL, typename Fetch, typename Use>
__device__ __forceinline__ void for_each_neighbor_w(const uint32_t *__restrict__ nbr, uint32_t i, const Phys &P,
                                                    uint32_t cnt, Fetch &&fetch, Use &&use) {
  const uint32_t kmax = active_max(cnt);
  const uint32_t ld = P.ld, col = i * 4u, rstride = ld * 4u;
  using T = decltype(fetch(0u));
  uint32_t k = 0;
  if (wave16(P, i)) {
    const uint32_t base = i - 32768u, col16 = i * 16u;
    auto chunks = [&](uint32_t k) {  // chunk k/8 of the 16-bit table as one buffer
      return __builtin_amdgcn_make_buffer_rsrc(P.nbr16 + (size_t)(k >> 3) * ld * 4u, 0, -1, kBufWord3);
    };
    // each batch: one table load, then every gather, then the bodies
    if constexpr (UNROLL >= 8) {
      for (; k + 8 <= kmax; k += 8) {
        const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(chunks(k), col16, 0, 0);
        T d[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          d[2 * q] = fetch(base + (w[q] & 0xffffu));
          d[2 * q + 1] = fetch(base + (w[q] >> 16));
        }
#pragma unroll
        for (int q = 0; q < 8; ++q) use(d[q], k + q < cnt);
      }
    }
    for (; k + 4 <= kmax; k += 4) {  // k % 8 is 0 or 4
      const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(chunks(k), col16, (k & 7u) * 2u, 0);
      T d[4];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        d[2 * q] = fetch(base + (w[q] & 0xffffu));
        d[2 * q + 1] = fetch(base + (w[q] >> 16));
      }
#pragma unroll
      for (int q = 0; q < 4; ++q) use(d[q], k + q < cnt);
    }
    for (; k + 2 <= kmax; k += 2) {
      const uint32_t w = __builtin_amdgcn_raw_buffer_load_b32(chunks(k), col16, (k & 7u) * 2u, 0);
      const T d0 = fetch(base + (w & 0xffffu)), d1 = fetch(base + (w >> 16));
      use(d0, k < cnt);
      use(d1, k + 1 < cnt);
    }
    if (k < kmax)
      use(fetch(base + (__builtin_amdgcn_raw_buffer_load_b32(chunks(k), col16, (k & 7u) * 2u, 0) & 0xffffu)), k < cnt);
    return;
  }
  auto rows = [&](uint32_t k) {  // rows k.. of the table as one buffer
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<uint32_t *>(nbr + (size_t)k * ld), 0, -1, kBufWord3);
  };
  if constexpr (UNROLL32 >= 8) {
    for (; k + 8 <= kmax; k += 8) {
      const auto r = rows(k);
      uint32_t j[8];
      T d[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) j[q] = __builtin_amdgcn_raw_buffer_load_b32(r, col, q * rstride, 0);
#pragma unroll
      for (int q = 0; q < 8; ++q) d[q] = fetch(j[q]);
#pragma unroll
      for (int q = 0; q < 8; ++q) use(d[q], k + q < cnt);
    }
  }
  for (; k + 4 <= kmax; k += 4) {
    const auto r = rows(k);
    uint32_t j[4];
    T d[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) j[q] = __builtin_amdgcn_raw_buffer_load_b32(r, col, q * rstride, 0);
#pragma unroll
    for (int q = 0; q < 4; ++q) d[q] = fetch(j[q]);
#pragma unroll
    for (int q = 0; q < 4; ++q) use(d[q], k + q < cnt);
  }
  for (; k < kmax; ++k) use(fetch(__builtin_amdgcn_raw_buffer_load_b32(rows(k), col, 0, 0)), k < cnt);
}

// The same walk with 64-bit global loads for the table and a single body
// that gathers for itself: faster than the split form for the passes that
// gather two or three records per neighbour (vorticity, XSPH).
template <int UNROLL, typename F>
__device__ __forceinline__ void for_each_neighbor_g(const uint32_t *__restrict__ nbr, uint32_t i, const Phys &P,
                                                    uint32_t cnt, F &&f) {
  const uint32_t kmax = active_max(cnt);
  const uint32_t ld = P.ld;
  uint32_t k = 0;
  if (wave16(P, i)) {
    const uint32_t base = i - 32768u, col16 = i * 16u;
    auto chunks = [&](uint32_t k) {  // chunk k/8 of the 16-bit table as one buffer
      return __builtin_amdgcn_make_buffer_rsrc(P.nbr16 + (size_t)(k >> 3) * ld * 4u, 0, -1, kBufWord3);
    };
    if constexpr (UNROLL >= 8) {
      for (; k + 8 <= kmax; k += 8) {
        const u32x4 w = __builtin_amdgcn_raw_buffer_load_b128(chunks(k), col16, 0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          f(base + (w[q] & 0xffffu), k + 2 * q < cnt);
          f(base + (w[q] >> 16), k + 2 * q + 1 < cnt);
        }
      }
    }
    for (; k + 4 <= kmax; k += 4) {  // k % 8 is 0 or 4
      const u32x2 w = __builtin_amdgcn_raw_buffer_load_b64(chunks(k), col16, (k & 7u) * 2u, 0);
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        f(base + (w[q] & 0xffffu), k + 2 * q < cnt);
        f(base + (w[q] >> 16), k + 2 * q + 1 < cnt);
      }
    }
    for (; k + 2 <= kmax; k += 2) {
      const uint32_t w = P.nbr16[chunk_word(ld, i, k)];
      f(base + (w & 0xffffu), k < cnt);
      f(base + (w >> 16), k + 1 < cnt);
    }
    if (k < kmax) f(base + (P.nbr16[chunk_word(ld, i, k)] & 0xffffu), k < cnt);
    return;
  }
  const uint32_t *row = nbr + i;
  if constexpr (UNROLL >= 8) {
    for (; k + 8 <= kmax; k += 8) {
      uint32_t j[8];
#pragma unroll
      for (int q = 0; q < 8; ++q) j[q] = row[(size_t)(k + q) * ld];
#pragma unroll
      for (int q = 0; q < 8; ++q) f(j[q], k + q < cnt);
    }
  }
  for (; k + 4 <= kmax; k += 4) {
    uint32_t j[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) j[q] = row[(size_t)(k + q) * ld];
#pragma unroll
    for (int q = 0; q < 4; ++q) f(j[q], k + q < cnt);
  }
  for (; k < kmax; ++k) f(row[(size_t)k * ld], k < cnt);
}

// The gather walk with the split (fetch, use) interface of the buffer walk.
template <int UNROLL, typename Fetch, typename Use>
__device__ __forceinline__ void for_each_neighbor_gw(const uint32_t *__restrict__ nbr, uint32_t i, const Phys &P,
                                                     uint32_t cnt, Fetch &&fetch, Use &&use) {
  for_each_neighbor_g<UNROLL>(nbr, i, P, cnt, [&](uint32_t j, bool v) { use(fetch(j), v); });
}

__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    uint32_t w = (uint32_t)__shfl_xor((int)v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// ---------------------------------------------------------------------------
// K1 + K2: predict and bin.
template <bool MULTI>
__global__ void __launch_bounds__(BS) k_predict_bin(uint32_t n, const float4 *__restrict__ pos,
                                                    const float4 *__restrict__ vel, const float4 *__restrict__ acc,
                                                    float4 *__restrict__ xs, uint32_t *__restrict__ cellkey,
                                                    uint32_t *__restrict__ rank, uint32_t *__restrict__ counts,
                                                    Scene s0, const Scene *__restrict__ scenes,
                                                    const uint32_t *__restrict__ sid, Phys P,
                                                    uint32_t *__restrict__ flags) {
  const uint32_t i = slot_index();
  const bool active = i < n;
  uint32_t key = 0, slab = 0;
  if (active) {
    const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
    float4 x = pos[i], v = vel[i], a = acc[i];
    float vx = fmaf(P.dt, P.gx + a.x, v.x);
    float vy = fmaf(P.dt, P.gy + a.y, v.y);
    float vz = fmaf(P.dt, P.gz + a.z, v.z);
    float px = fmaf(P.dt, vx, x.x), py = fmaf(P.dt, vy, x.y), pz = fmaf(P.dt, vz, x.z);
    clamp_rand(sc, px, py, pz);
    xs[i] = make_float4(px, py, pz, 0.0f);
    if (__builtin_isfinite(px) && __builtin_isfinite(py) && __builtin_isfinite(pz)) {
      const int cz = cell_axis(sc, pz, 2);
      key = sc.cell_base + cell_local(sc, cell_axis(sc, px, 0), cell_axis(sc, py, 1), cz);
      slab = P.slab_log2 ? z_slab(sc, P.slab_log2, pz, cz) : 0u;
    } else {
      // a diverged particle: binned by its slot, so such particles spread
      // over the grid instead of all falling into one cell (whose sort and
      // walks are quadratic in its population); its distances are NaN, so
      // it has no neighbours and is nobody's.  Reported at the next sync.
      const uint32_t D0 = (uint32_t)sc.dims[0], D1 = (uint32_t)sc.dims[1], D2 = (uint32_t)sc.dims[2];
      key = sc.cell_base + cell_local(sc, (int)(i % D0), (int)((i / D0) % D1), (int)((i / (D0 * D1)) % D2));
      atomicOr(&flags[0], 2u);
    }
    cellkey[i] = key;
  }
  uint32_t r = wave_atomic_rank(counts, key, active);
  if (active) rank[i] = r | (slab << kRankSlabShift);
}

// ---------------------------------------------------------------------------
// K3: exclusive scan, reduce-then-scan with a Blelloch tree in LDS.
constexpr int SCAN_V = 8;
constexpr uint32_t SCAN_E = BS * SCAN_V;  // 2048 elements per block
__device__ __forceinline__ int pad32(int i) { return i + (i >> 5); }

__global__ void __launch_bounds__(BS) k_scan_reduce(const uint32_t *__restrict__ in, uint32_t n,
                                                    uint32_t *__restrict__ sums) {
  __shared__ uint32_t sh[BS / 64];
  const uint32_t base = blockIdx.x * SCAN_E + threadIdx.x * SCAN_V;
  uint32_t s = 0;
  if (base + SCAN_V <= n) {
    const uint4 *p = reinterpret_cast<const uint4 *>(in + base);
    uint4 a = p[0], b = p[1];
    s = ((a.x + a.y) + (a.z + a.w)) + ((b.x + b.y) + (b.z + b.w));
  } else {
    for (int k = 0; k < SCAN_V; ++k)
      if (base + k < n) s += in[base + k];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += (uint32_t)__shfl_xor((int)s, o, 64);
  if (__lane_id() == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < BS / 64; ++w) t += sh[w];
    sums[blockIdx.x] = t;
  }
}

// out[0..n] = exclusive scan of in[0..n) (+ prefix[block]); out[n] = total.
// zero_in: clear `in` (the cell counts, for the next binning) and, when not
// null, `zero2` (the per-cell slab words, rewritten by the reorder for the
// cells that hold particles) over the same range.
__global__ void __launch_bounds__(BS) k_scan_down(uint32_t *__restrict__ in, uint32_t n,
                                                  const uint32_t *__restrict__ prefix, uint32_t *__restrict__ out,
                                                  int zero_in, uint32_t *__restrict__ zero2) {
  __shared__ uint32_t sh[BS + BS / 32];
  const int t = threadIdx.x;
  const uint32_t base = blockIdx.x * SCAN_E + (uint32_t)t * SCAN_V;
  uint32_t v[SCAN_V];
  if (base + SCAN_V <= n) {
    uint4 *p = reinterpret_cast<uint4 *>(in + base);
    uint4 a = p[0], b = p[1];
    v[0] = a.x; v[1] = a.y; v[2] = a.z; v[3] = a.w; v[4] = b.x; v[5] = b.y; v[6] = b.z; v[7] = b.w;
    if (zero_in) { p[0] = make_uint4(0, 0, 0, 0); p[1] = make_uint4(0, 0, 0, 0); }
    if (zero2) {
      uint4 *z = reinterpret_cast<uint4 *>(zero2 + base);
      z[0] = make_uint4(0, 0, 0, 0);
      z[1] = make_uint4(0, 0, 0, 0);
    }
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k) {
      v[k] = (base + k < n) ? in[base + k] : 0u;
      if (zero_in && base + k < n) in[base + k] = 0u;
      if (zero2 && base + k < n) zero2[base + k] = 0u;
    }
  }
  uint32_t e[SCAN_V];
  e[0] = 0;
#pragma unroll
  for (int k = 1; k < SCAN_V; ++k) e[k] = e[k - 1] + v[k - 1];
  const uint32_t tot = e[SCAN_V - 1] + v[SCAN_V - 1];
  sh[pad32(t)] = tot;
  // Blelloch up-sweep
  int offset = 1;
  for (int d = BS >> 1; d > 0; d >>= 1) {
    __syncthreads();
    if (t < d) {
      int ai = offset * (2 * t + 1) - 1, bi = offset * (2 * t + 2) - 1;
      sh[pad32(bi)] += sh[pad32(ai)];
    }
    offset <<= 1;
  }
  if (t == 0) sh[pad32(BS - 1)] = 0;
  // down-sweep
  for (int d = 1; d < BS; d <<= 1) {
    offset >>= 1;
    __syncthreads();
    if (t < d) {
      int ai = offset * (2 * t + 1) - 1, bi = offset * (2 * t + 2) - 1;
      uint32_t tmp = sh[pad32(ai)];
      sh[pad32(ai)] = sh[pad32(bi)];
      sh[pad32(bi)] += tmp;
    }
  }
  __syncthreads();
  const uint32_t tp = sh[pad32(t)] + (prefix ? prefix[blockIdx.x] : 0u);
  if (base + SCAN_V <= n) {
    uint4 *o = reinterpret_cast<uint4 *>(out + base);
    o[0] = make_uint4(tp + e[0], tp + e[1], tp + e[2], tp + e[3]);
    o[1] = make_uint4(tp + e[4], tp + e[5], tp + e[6], tp + e[7]);
  } else {
#pragma unroll
    for (int k = 0; k < SCAN_V; ++k)
      if (base + k < n) out[base + k] = tp + e[k];
  }
  if (n > 0 && base <= n - 1 && n - 1 < base + SCAN_V) {
    uint32_t k = n - 1 - base;
    uint32_t last = 0;
#pragma unroll
    for (int q = 0; q < SCAN_V; ++q)
      if ((uint32_t)q == k) last = tp + e[q] + v[q];
    out[n] = last;
  }
}

// ---------------------------------------------------------------------------
// K4: scatter ids into cell order (rank inside a cell is still arbitrary);
// the sort key inside a cell is (z-slab, id).
__global__ void __launch_bounds__(BS) k_scatter(uint32_t n, const uint32_t *__restrict__ cellkey,
                                                const uint32_t *__restrict__ rank, const uint32_t *__restrict__ offsets,
                                                const uint32_t *__restrict__ ids, uint32_t *__restrict__ tmp_key,
                                                uint32_t *__restrict__ tmp_src) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const uint32_t rk = rank[i];
  uint32_t s = offsets[cellkey[i]] + (rk & kRankMask);
  tmp_key[s] = ids[i] | ((rk >> kRankSlabShift) << kIdBits);  // (z-slab, id)
  tmp_src[s] = i;
}

// Number of keys in [o, e) below `key`, four per load (aligned uint4) between
// a scalar head and tail: the ranges are short and L1-resident, so the loop
// is bound by load issue, and crowded buckets (hashed grids) make it long.
__device__ __forceinline__ uint32_t rank_below(const uint32_t *__restrict__ keys, uint32_t o, uint32_t e,
                                               uint32_t key) {
  uint32_t r = 0, t = o;
  for (; t < e && (t & 3u); ++t) r += keys[t] < key ? 1u : 0u;
  const uint4 *q = reinterpret_cast<const uint4 *>(keys + t);
  const uint32_t nv = (e - t) >> 2;
  for (uint32_t v = 0; v < nv; ++v) {
    const uint4 w = q[v];
    r += (w.x < key ? 1u : 0u) + (w.y < key ? 1u : 0u) + (w.z < key ? 1u : 0u) + (w.w < key ? 1u : 0u);
  }
  for (t += 4u * nv; t < e; ++t) r += keys[t] < key ? 1u : 0u;
  return r;
}

// The same count, plus the keys below z-slabs 1, 2 and 3 (the cell's slab
// boundaries in its sorted order).
__device__ __forceinline__ uint32_t rank_below_slabs(const uint32_t *__restrict__ keys, uint32_t o, uint32_t e,
                                                     uint32_t key, uint32_t &b1, uint32_t &b2, uint32_t &b3) {
  constexpr uint32_t s1 = 1u << kIdBits, s2 = 2u << kIdBits, s3 = 3u << kIdBits;
  uint32_t r = 0, c1 = 0, c2 = 0, c3 = 0, t = o;
  auto one = [&](uint32_t w) {
    r += w < key ? 1u : 0u;
    c1 += w < s1 ? 1u : 0u;
    c2 += w < s2 ? 1u : 0u;
    c3 += w < s3 ? 1u : 0u;
  };
  for (; t < e && (t & 3u); ++t) one(keys[t]);
  const uint4 *q = reinterpret_cast<const uint4 *>(keys + t);
  const uint32_t nv = (e - t) >> 2;
  for (uint32_t v = 0; v < nv; ++v) {
    const uint4 w = q[v];
    one(w.x); one(w.y); one(w.z); one(w.w);
  }
  for (t += 4u * nv; t < e; ++t) one(keys[t]);
  b1 = c1;
  b2 = c2;
  b3 = c3;
  return r;
}

// Stable order inside each cell (ascending (z-slab, id) key), then move the
// particle data into its final slot; dest[s] = that slot, for k_settle_extra
// (lam_out and dest may be null: the plain solve needs neither).  With
// z-slabs (cellsub not null) the cell's first particle also writes the
// cell's slab word: its particles below slabs 1, 2, 3 and its count, 8 bits
// each, or kSubSaturated for a cell of 255 or more.
__global__ void __launch_bounds__(BS) k_settle(uint32_t n, const uint32_t *__restrict__ tmp_key,
                                               const uint32_t *__restrict__ tmp_src,
                                               const uint32_t *__restrict__ cellkey,
                                               const uint32_t *__restrict__ offsets, const float4 *__restrict__ xs,
                                               const float4 *__restrict__ pos, const float *__restrict__ lam_in,
                                               float4 *__restrict__ X0, float4 *__restrict__ Xold,
                                               uint32_t *__restrict__ ids, float *__restrict__ lam_out,
                                               uint32_t *__restrict__ dest, uint32_t *__restrict__ cellsub) {
  const uint32_t s = slot_index();
  if (s >= n) return;
  const uint32_t p = tmp_src[s], key = tmp_key[s];
  const uint32_t c = cellkey[p];
  const uint32_t o = offsets[c], e = offsets[c + 1];
  uint32_t r;
  if (cellsub) {
    uint32_t b1, b2, b3;
    r = rank_below_slabs(tmp_key, o, e, key, b1, b2, b3);
    if (r == 0) {
      const uint32_t cnt = e - o;
      cellsub[c] = cnt >= 255u ? kSubSaturated : (b1 | (b2 << 8) | (b3 << 16) | (cnt << 24));
    }
  } else {
    r = rank_below(tmp_key, o, e, key);
  }
  const uint32_t f = o + r;
  X0[f] = xs[p];
  Xold[f] = pos[p];
  ids[f] = key & kIdMask;
  if (lam_out) lam_out[f] = lam_in[p];
  if (dest) dest[s] = f;
}

// Learned / data-generation modes: carry the previous substep's lambda sum and
// velocity along with the reorder (the slots k_settle chose).
__global__ void __launch_bounds__(BS) k_settle_extra(uint32_t n, const uint32_t *__restrict__ tmp_src,
                                                     const uint32_t *__restrict__ dest,
                                                     const float *__restrict__ acc_in, float *__restrict__ acc_out,
                                                     const float4 *__restrict__ vel_in, float4 *__restrict__ vel_out) {
  const uint32_t s = slot_index();
  if (s >= n) return;
  const uint32_t p = tmp_src[s], f = dest[s];
  acc_out[f] = acc_in[p];
  vel_out[f] = vel_in[p];
}

// Temporal warm start: iteration 0 applies relax x the previous substep's
// lambda sum (kept when it is not a finite non-positive number).
__global__ void __launch_bounds__(BS) k_warm_previous(uint32_t n, const float *__restrict__ acc_prev, float4 *X,
                                                      float *__restrict__ lam_acc, float relax) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  float *wp = reinterpret_cast<float *>(X) + 4 * (size_t)i + 3;
  float lam = relax * acc_prev[i];
  if (!(lam <= 0.0f) || !isfinite(lam)) lam = *wp;
  *wp = lam;
  lam_acc[i] = lam;
}

// Spatially smoothed warm start: the warm lambda in X.w replaced by its
// Shepard average over the particle's neighbours (poly6 weights, self
// included), so the first Jacobi displacement -- the gradient of that field --
// carries no particle-scale noise from a lambda field of the previous
// arrangement.  Two kernels: the average into tmp, then tmp into X.w and the
// substep's sum.
__global__ void __launch_bounds__(BS) k_smooth_warm(uint32_t n, const float4 *__restrict__ X,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ ncount, Phys P,
                                                    float *__restrict__ tmp) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = X[i];
  float sw = 0.0f, sl = 0.0f;
  // wavefront-uniform walk: padding entries (valid false) leave both sums
  // as they are, so they equal the per-lane sums over the list
  const Gather GX(X, n);
  for_each_neighbor_w<8>(nbr, i, P, ncount[i], [&](uint32_t j) { return GX[j]; }, [&](const float4 &xj, bool valid) {
    const float w = poly6(P, dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z));
    sw = valid ? sw + w : sw;
    sl = valid ? fmaf(w, xj.w, sl) : sl;
  });
  tmp[i] = sw > 0.0f ? sl / sw : xi.w;
}
__global__ void __launch_bounds__(BS) k_apply_warm(uint32_t n, const float *__restrict__ tmp, float4 *X,
                                                   float *__restrict__ lam_acc) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float lam = tmp[i];
  reinterpret_cast<float *>(X)[4 * (size_t)i + 3] = lam;
  lam_acc[i] = lam;
}

__global__ void __launch_bounds__(BS) k_copy_words(uint32_t *__restrict__ dst, const uint32_t *__restrict__ src,
                                                   size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// K5-K7: neighbour list over the 27 cells of the frozen predicted position.
// Dense grid: the three cells (x, y, z-1..z+1) are one contiguous slot range.
// The first R entries of each particle are staged in LDS and written as
// whole table rows (coalesced) once the block of B slots is done; later
// entries are stored directly.  (R, B) is chosen at run time among
// kListVariants instantiations (the autotuner's list knob).

template <bool MULTI, bool HASH, int LIST_R, int LIST_BS>
__global__ void __launch_bounds__(LIST_BS) k_build_list(uint32_t n, const float4 *__restrict__ X0,
                                                        const uint32_t *__restrict__ offsets,
                                                        uint32_t *__restrict__ nbr, uint32_t *__restrict__ ncount,
                                                        uint32_t *__restrict__ flags, Scene s0,
                                                        const Scene *__restrict__ scenes,
                                                        const uint32_t *__restrict__ sid, Phys P) {
  __shared__ uint32_t tile[(LIST_R + 1) * LIST_BS];  // row LIST_R: spare
  const uint32_t tid = threadIdx.x;
  const uint32_t base_slot = xcd_block(blockIdx.x, gridDim.x) * LIST_BS;
  const uint32_t i = base_slot + tid;
  static_assert(LIST_R % 8 == 0, "the LDS rows fill whole 8-entry chunks");
  auto store16 = [&](uint32_t k, uint32_t j) {  // entry k as a 16-bit offset
    reinterpret_cast<uint16_t *>(P.nbr16)[((size_t)(k >> 3) * P.ld + i) * 8u + (k & 7u)] =
        (uint16_t)(j - (i - 32768u));
  };
  uint32_t cnt = 0, last = i, smin = i, smax = i;
  bool fits16 = true;  // every entry within +-32767 slots of i
  if (i < n) {
    const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
    const float4 xi = X0[i];
    const int cx = cell_axis(sc, xi.x, 0), cy = cell_axis(sc, xi.y, 1), cz = cell_axis(sc, xi.z, 2);
    const int D0 = sc.dims[0], D1 = sc.dims[1], D2 = sc.dims[2];
    // Cells out of reach are skipped: the squared gap (in units of the cell
    // size h) from the particle to a neighbour cell's box is a^2 + b^2 + c^2
    // with a = fx (cell below), 0 (same), 1 - fx (cell above) per axis.  The
    // 1e-3 slack covers the float rounding of the position, of the cell
    // assignment and of the d2 <= h^2 test, so no candidate that could pass
    // the test is skipped and the list (and its order) is unchanged.
    const float fx = (xi.x - sc.lo[0]) * sc.inv_cell - (float)cx;
    const float fy = (xi.y - sc.lo[1]) * sc.inv_cell - (float)cy;
    const float fz = (xi.z - sc.lo[2]) * sc.inv_cell - (float)cz;
    constexpr float kReach = 1.001f;
    auto gap = [](int d, float f) { return d < 0 ? f * f : d > 0 ? (1.0f - f) * (1.0f - f) : 0.0f; };
    const uint32_t cbase = sc.cell_base;
    // Branch-free: every candidate is written to the particle's next free
    // LDS row (clamped to the spare row R) and the row is taken only when the
    // candidate is accepted, so a rejected one is overwritten by the next.
    // Entries past R are rare; a batch that produced any is tested again
    // (uniform branch) and those entries go straight to the table.
    const Gather GX(X0, n);
    auto test = [&](const float3 &xj, uint32_t j, bool live) {
      const float d2 = dist2(xj.x - xi.x, xj.y - xi.y, xj.z - xi.z);
      tile[(cnt < (uint32_t)LIST_R ? cnt : (uint32_t)LIST_R) * LIST_BS + tid] = j;
      cnt += (live && d2 <= P.h2) ? 1u : 0u;
    };
    // spilled entries go to both tables (the wavefront's format is decided
    // at the end); `last` keeps the latest one, smin / smax their range
    auto spill = [&](uint32_t j0, uint32_t m, uint32_t k) {  // k: count before the batch
      for (uint32_t q = 0; q < m; ++q) {
        const float4 xj = X0[j0 + q];
        const bool in = dist2(xj.x - xi.x, xj.y - xi.y, xj.z - xi.z) <= P.h2;
        if (in && k >= (uint32_t)LIST_R && k < P.maxn) {
          nbr[(size_t)k * P.ld + i] = j0 + q;
          store16(k, j0 + q);
          last = j0 + q;
          smin = smin < j0 + q ? smin : j0 + q;
          smax = smax > j0 + q ? smax : j0 + q;
        }
        k += in ? 1u : 0u;
      }
    };
    // Batches of eight candidate loads in flight.  A range's last 1..7
    // candidates are one masked batch instead of smaller batches: the eight
    // records loaded end at the range's end (they start before the range
    // when it is shorter; record 0 on when it ends below 8, always inside
    // the array's capacity of at least 256 slots) and only those inside the
    // range are accepted, in order.  So a range costs the wavefront
    // ceil(longest / 8) rounds of load latency.
    auto visit = [&](uint32_t t0, uint32_t t1) {
      uint32_t j = t0;
      for (; j + 8 <= t1; j += 8) {
        float3 c[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = GX.xyz(j, q);
        const uint32_t k0 = cnt;
#pragma unroll
        for (int q = 0; q < 8; ++q) test(c[q], j + q, true);
        if (__ballot(cnt > (uint32_t)LIST_R)) spill(j, 8, k0);
      }
      if (j < t1) {
        const uint32_t b = t1 >= 8u ? t1 - 8u : 0u;  // b <= j
        const uint32_t qlo = j - b, qhi = t1 - b;
        float3 c[8];
#pragma unroll
        for (int q = 0; q < 8; ++q) c[q] = GX.xyz(b, q);
        const uint32_t k0 = cnt;
#pragma unroll
        for (int q = 0; q < 8; ++q) test(c[q], b + q, (uint32_t)q >= qlo && (uint32_t)q < qhi);
        if (__ballot(cnt > (uint32_t)LIST_R)) spill(j, t1 - j, k0);
      }
    };
    if (!HASH || sc.distinct27) {
      // Each of the 9 columns (x, y) is visited as one slot range over its
      // cells z-1..z+1 -- contiguous cells, and in a hashed grid contiguous
      // buckets unless the row-major index wraps the table (then cell by
      // cell).  With every bucket of the 27 cells distinct (always on a dense
      // grid), skipping out-of-reach candidates leaves the list unchanged.
      // The range is cut to the z-slabs (height h / S) that the particle can
      // reach: [fz - w, fz + w] in cell units with w^2 = 1.001 - gxy; a
      // cell's particles are sorted by slab, so the reachable ones of the
      // first and last cell are a prefix / suffix of the cell's slab words.
      // The bounds are loaded up front so their latencies overlap.
      const uint32_t sl2 = P.slab_log2;
      const int S = 1 << sl2;
      const float Sf = (float)S;
      uint32_t t0[9], t1[9], u0[9], u1[9];
      int zl[9], zh[9], sl[9], sh[9];
      uint32_t wrap = 0;
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        const int dx = r / 3 - 1, dy = r % 3 - 1;
        const int nx = cx + dx, ny = cy + dy;
        const float gxy = gap(dx, fx) + gap(dy, fy);
        t0[r] = t1[r] = u0[r] = u1[r] = 0;
        zl[r] = zh[r] = sl[r] = sh[r] = 0;
        if (nx < 0 || nx >= D0 || ny < 0 || ny >= D1 || gxy > kReach) continue;
        const float w = __builtin_amdgcn_sqrtf(kReach - gxy);
        int klo = (int)floorf((fz - w) * Sf), khi = (int)floorf((fz + w) * Sf);
        const int kmin = cz > 0 ? -S : 0, kmax = cz + 1 < D2 ? 2 * S - 1 : S - 1;
        klo = klo < kmin ? kmin : klo > S - 1 ? S - 1 : klo;
        khi = khi > kmax ? kmax : khi < 0 ? 0 : khi;
        const int zlo = klo < 0 ? cz - 1 : cz, zhi = khi >= S ? cz + 1 : cz;
        sl[r] = klo & (S - 1);
        sh[r] = khi & (S - 1);
        const uint32_t blo = cell_local(sc, nx, ny, zlo), bhi = cell_local(sc, nx, ny, zhi);
        if constexpr (HASH) {
          if (bhi - blo != (uint32_t)(zhi - zlo)) wrap |= 1u << r;
        }
        zl[r] = zlo;
        zh[r] = zhi;
        t0[r] = offsets[cbase + blo];
        t1[r] = offsets[cbase + bhi + 1];
        if (sl2) {
          u0[r] = P.cellsub[cbase + blo];
          u1[r] = P.cellsub[cbase + bhi];
        }
      }
      if (sl2) {
        // first cell: skip its particles below slab sl; last cell: those
        // above slab sh (none to skip in a cell too full for 8-bit counts)
#pragma unroll
        for (int r = 0; r < 9; ++r) {
          const uint32_t a = u0[r], b = u1[r];
          const uint32_t below = (sl[r] == 0 || a >= kSubSaturated) ? 0u : (a >> (8 * (sl[r] - 1))) & 0xffu;
          uint32_t above = (sh[r] == S - 1 || b >= kSubSaturated) ? 0u : (b >> 24) - ((b >> (8 * sh[r])) & 0xffu);
          // a cut never leaves the column's range (the words of a sorted
          // cell always satisfy this; the clamps keep it unconditional)
          const uint32_t len = t1[r] - t0[r];
          above = above < len ? above : len;
          t0[r] += below < len - above ? below : len - above;
          t1[r] -= above;
        }
      }
#pragma unroll
      for (int r = 0; r < 9; ++r) {
        if (HASH && (wrap >> r & 1u)) {
          const int nx = cx + r / 3 - 1, ny = cy + r % 3 - 1;
          for (int z = zl[r]; z <= zh[r]; ++z) {
            const uint32_t k = cbase + cell_local(sc, nx, ny, z);
            visit(offsets[k], offsets[k + 1]);
          }
        } else {
          visit(t0[r], t1[r]);
        }
      }
    } else {
      // buckets may repeat among the 27 cells: each is visited at its first
      // occurrence; skipping a cell could move that occurrence, so no pruning
      uint32_t seen[27];
      int nseen = 0;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            const int nx = cx + dx, ny = cy + dy, nz = cz + dz;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= D0 || ny >= D1 || nz >= D2) continue;
            const uint32_t k = cbase + cell_local(sc, nx, ny, nz);
            bool dup = false;
            for (int q = 0; q < nseen; ++q) dup |= (seen[q] == k);
            if (dup) continue;
            seen[nseen++] = k;
            visit(offsets[k], offsets[k + 1]);
          }
    }
    ncount[i] = cnt < P.maxn ? cnt : P.maxn;
    const uint32_t c = cnt < P.maxn ? cnt : P.maxn;
    uint32_t lo, hi;
    if constexpr (HASH) {
      // bucket order is not slot order: the range over every entry
      lo = smin;
      hi = smax;
      const uint32_t rows = c < (uint32_t)LIST_R ? c : (uint32_t)LIST_R;
      for (uint32_t k = 0; k < rows; ++k) {
        const uint32_t v = tile[k * LIST_BS + tid];
        lo = lo < v ? lo : v;
        hi = hi > v ? hi : v;
      }
    } else {
      // ascending slot order: the first and last entries bound the rest
      if (c > 0 && c - 1 < (uint32_t)LIST_R) last = tile[(c - 1) * LIST_BS + tid];
      lo = c > 0 ? tile[tid] : i;
      hi = last;
    }
    fits16 = i - lo <= 32768u && hi - i <= 32767u;
  }
  const bool w16 = !__ballot(!fits16);  // lanes past n count as fitting
  if (__lane_id() == 0) P.wfmt[i >> 6] = w16 ? 1u : 0u;
  const uint32_t wm = wave_max_u32(cnt);  // lanes past n hold cnt = 0
  if (__lane_id() == 0 && wm > 0) {
    atomicMax(&flags[1], wm);
    if (wm > P.maxn) atomicOr(&flags[0], 1u);
  }
  // Rows [count, largest count of the wavefront) hold the lane's own slot:
  // the wavefront-uniform walks read them as padding.
  const uint32_t wcap = wm < P.maxn ? wm : P.maxn;
  const uint32_t lrows = wcap < (uint32_t)LIST_R ? wcap : (uint32_t)LIST_R;
  if (i < n) {
    const uint32_t c = cnt < P.maxn ? cnt : P.maxn;
    const uint32_t k1 = c > (uint32_t)LIST_R ? c : (uint32_t)LIST_R;
    if (w16) {
      // whole 8-entry chunks for the LDS rows (R is a multiple of 8; entries
      // of the last chunk past wcap are never read), then half-words; the
      // walks read whole pairs, so padding runs to wcap rounded up to even
      const uint32_t base = i - 32768u;
      for (uint32_t k = 0; k < lrows; k += 8) {
        uint32_t w[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const uint32_t ka = k + 2 * q, kb = ka + 1;
          const uint32_t a = ka < c ? tile[ka * LIST_BS + tid] : i;
          const uint32_t b = kb < c ? tile[kb * LIST_BS + tid] : i;
          w[q] = (a - base) | ((b - base) << 16);
        }
        reinterpret_cast<uint4 *>(P.nbr16)[(size_t)(k >> 3) * P.ld + i] = make_uint4(w[0], w[1], w[2], w[3]);
      }
      for (uint32_t k = k1; k < ((wcap + 1u) & ~1u); ++k) store16(k, i);
    } else {
      for (uint32_t k = 0; k < lrows; ++k) nbr[(size_t)k * P.ld + i] = k < c ? tile[k * LIST_BS + tid] : i;
      for (uint32_t k = k1; k < wcap; ++k) nbr[(size_t)k * P.ld + i] = i;
    }
  }
}


// ---------------------------------------------------------------------------
// K8: boundary sample of every particle at the current iterate (fields of
// any kind, the cubic 32-node element included).
template <bool MULTI>
__global__ void __launch_bounds__(BS) k_boundary(uint32_t n, const float4 *__restrict__ X, float4 *__restrict__ B,
                                                 const float *__restrict__ pool, Scene s0,
                                                 const Scene *__restrict__ scenes, const uint32_t *__restrict__ sid,
                                                 Phys P) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
  const float4 xi = X[i];
  B[i] = boundary_sample<0>(P, sc, pool, xi.x, xi.y, xi.z);
}

// K9: density constraint and lambda, lambda to X[i].w.  FK: 0 no boundary,
// 1 boundary samples already in B (k_boundary), 2 trilinear maps sampled here.
// ACC: 0 plain; 1 also add lambda to the per-substep sum (learned / data-
// generation modes); 2 projected accumulation: the increment comes from the
// unclamped constraint rho~ - 1, the substep's sum stays <= 0, and X[i].w
// receives the change of the sum (an over-pushed warm start is pulled back,
// never past a zero sum).  GATHER: wavefronts on the 32-bit table walk with
// 64-bit global gathers (hashed grids, where most wavefronts take that
// table: 292 -> 267 us per launch at the 4M obstacle drop; compiled into the
// dense-grid kernel it costs 1 us there); the pairs and their order are the
// same either way.
template <bool MULTI, int FK, int ACC = 0, bool GATHER = false>
__global__ void __launch_bounds__(BS) k_lambda(uint32_t n, float4 *X, float4 *__restrict__ B,
                                               const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ ncount,
                                               const float *__restrict__ pool, Scene s0,
                                               const Scene *__restrict__ scenes, const uint32_t *__restrict__ sid,
                                               Phys P, float *__restrict__ lam_acc) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = X[i];
  float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if constexpr (FK == 2) {
    const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
    b = boundary_sample<2>(P, sc, pool, xi.x, xi.y, xi.z);
    B[i] = b;
  } else if constexpr (FK == 1) {
    b = B[i];
  }
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, s2 = 0.0f, rho = 0.0f;
  // poly6 and the spiky gradient with their range tests shared (the values
  // of poly6() and spiky_f()); padding entries (j = i) have a zero gradient
  // and their poly6 term is not added
  const Gather GX(X, n);
  auto pair = [&](const float3 &xj, bool valid) {
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float d2 = dist2(dx, dy, dz);
    const bool in = d2 <= P.h2;
    const float r = sqrt_pair(d2);
    const float q = P.h - r;
    const float fs = div_n(P.Ks * (q * q), r);
    const float f = (in && d2 > 0.0f) ? fs : 0.0f;
    const float t = P.h2 - d2;
    const float w = P.Kp * ((t * t) * t);
    float ax = f * dx, ay = f * dy, az = f * dz;
    gx += ax; gy += ay; gz += az;
    s2 = fmaf(az, az, fmaf(ay, ay, fmaf(ax, ax, s2)));
    rho += (in && valid) ? w : 0.0f;
  };
  auto fetch = [&](uint32_t j) { return GX.xyz(j); };
  if constexpr (GATHER) {
    if (wave16(P, i))
      for_each_neighbor_w<8>(nbr, i, P, ncount[i], fetch, pair);
    else
      for_each_neighbor_gw<8>(nbr, i, P, ncount[i], [&](uint32_t j) {
        const float4 *q = X + j;
        return make_float3(q->x, q->y, q->z);
      }, pair);
  } else {
    for_each_neighbor_w<8, 4>(nbr, i, P, ncount[i], fetch, pair);
  }
  rho *= P.V;
  gx *= P.V; gy *= P.V; gz *= P.V;
  s2 *= P.V2;
  if (FK && b.w > 0.0f) {
    float dx = xi.x - b.x, dy = xi.y - b.y, dz = xi.z - b.z;
    float d2 = dist2(dx, dy, dz);
    float fb = b.w * spiky_f(P, d2);
    gx = fmaf(fb, dx, gx); gy = fmaf(fb, dy, gy); gz = fmaf(fb, dz, gz);
    rho = fmaf(b.w, poly6(P, d2), rho);
  }
  float c = rho - 1.0f;
  float den = fmaf(gz, gz, fmaf(gy, gy, fmaf(gx, gx, s2)));
  if constexpr (ACC == 2) {
    const float l0 = lam_acc[i];
    const float l1 = fminf(l0 + (-c) / (den + P.eps_lambda), 0.0f);
    reinterpret_cast<float *>(X)[4 * (size_t)i + 3] = l1 - l0;
    lam_acc[i] = l1;
  } else {
    float C = c > 0.0f ? c : 0.0f;
    const float lam = (-C) / (den + P.eps_lambda);
    reinterpret_cast<float *>(X)[4 * (size_t)i + 3] = lam;
    if constexpr (ACC == 1) lam_acc[i] += lam;
  }
}

struct FeatRec {
  float3 x;  // neighbour position
  float4 v;  // neighbour velocity of the previous substep
};

// Iteration 0 of the learned and data-generation modes: the lambda pass plus
// the 16 features of the warm-start network (previous-substep lambdas,
// density, constraint gradient terms, boundary, velocity divergence).
template <bool MULTI, int FK>
__global__ void __launch_bounds__(BS) k_lambda_feat(uint32_t n, float4 *X, float4 *__restrict__ B,
                                                    const uint32_t *__restrict__ nbr,
                                                    const uint32_t *__restrict__ ncount,
                                                    const float *__restrict__ pool, Scene s0,
                                                    const Scene *__restrict__ scenes,
                                                    const uint32_t *__restrict__ sid, Phys P, LearnBufs lb) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = X[i];
  float4 b = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if constexpr (FK == 2) {
    const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
    b = boundary_sample<2>(P, sc, pool, xi.x, xi.y, xi.z);
    B[i] = b;
  } else if constexpr (FK == 1) {
    b = B[i];
  }
  const float4 vi = lb.vel_prev[i];
  const uint32_t cnt = ncount[i];
  float gx = 0.0f, gy = 0.0f, gz = 0.0f, s2 = 0.0f, rho = 0.0f, dv = 0.0f;
  // the lambda pass's walk (same order, so the same lambda bit for bit) with
  // the previous velocity gathered too; padding entries (j = i) have a zero
  // gradient and velocity difference, and their poly6 term is not added
  const Gather GX(X, n), GV(lb.vel_prev, n);
  for_each_neighbor_w<8, 4>(nbr, i, P, cnt, [&](uint32_t j) { return FeatRec{GX.xyz(j), GV[j]}; },
                         [&](const FeatRec &r, bool valid) {
    float dx = xi.x - r.x.x, dy = xi.y - r.x.y, dz = xi.z - r.x.z;
    float d2 = dist2(dx, dy, dz);
    float f = spiky_f(P, d2);
    float ax = f * dx, ay = f * dy, az = f * dz;
    gx += ax; gy += ay; gz += az;
    s2 = fmaf(az, az, fmaf(ay, ay, fmaf(ax, ax, s2)));
    rho += valid ? poly6(P, d2) : 0.0f;
    dv = fmaf(r.v.x - vi.x, ax, fmaf(r.v.y - vi.y, ay, fmaf(r.v.z - vi.z, az, dv)));
  });
  rho *= P.V;
  gx *= P.V; gy *= P.V; gz *= P.V;
  s2 *= P.V2;
  const float rho_f = rho;
  float rho_b = 0.0f, bdist = 1.0f;
  if (FK && b.w > 0.0f) {
    float dx = xi.x - b.x, dy = xi.y - b.y, dz = xi.z - b.z;
    float d2 = dist2(dx, dy, dz);
    float fb = b.w * spiky_f(P, d2);
    gx = fmaf(fb, dx, gx); gy = fmaf(fb, dy, gy); gz = fmaf(fb, dz, gz);
    rho_b = b.w * poly6(P, d2);
    rho = fmaf(b.w, poly6(P, d2), rho);
    bdist = sqrtf(d2) / P.h;
  }
  float c = rho - 1.0f;
  float C = c > 0.0f ? c : 0.0f;
  float g2 = fmaf(gz, gz, fmaf(gy, gy, gx * gx));
  float den = fmaf(gz, gz, fmaf(gy, gy, fmaf(gx, gx, s2)));
  const float lam = (-C) / (den + P.eps_lambda);
  reinterpret_cast<float *>(X)[4 * (size_t)i + 3] = lam;
  lb.lam_acc[i] = lam;
  const float vs = P.dt / P.h;
  float4 *F = reinterpret_cast<float4 *>(lb.feats + 16 * (size_t)i);
  F[0] = make_float4(c, C, lam, lb.acc_prev[i]);
  F[1] = make_float4(lb.lam_prev[i], den, s2, g2);
  F[2] = make_float4((float)cnt, b.w / P.V, bdist, dv * P.V * P.dt);
  F[3] = make_float4(sqrtf(dist2(vi.x, vi.y, vi.z)) * vs, vi.y * vs, rho_f - 1.0f, rho_b);
}

// K10 + K11: position correction and wall clamp.
template <bool MULTI, int FK>
__global__ void __launch_bounds__(BS) k_dp(uint32_t n, const float4 *__restrict__ X, const float4 *__restrict__ B,
                                           const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ ncount,
                                           float4 *__restrict__ Xn, Scene s0, const Scene *__restrict__ scenes,
                                           const uint32_t *__restrict__ sid, Phys P) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = X[i];
  const float li = xi.w;
  float ax = 0.0f, ay = 0.0f, az = 0.0f;
  // padding entries (j = i): the gradient at d = 0 is 0 and dx = 0
  const Gather GX(X, n);
  for_each_neighbor_w<8>(nbr, i, P, ncount[i], [&](uint32_t j) { return GX[j]; }, [&](const float4 &xj, bool) {
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float fl = (li + xj.w) * spiky_f(P, dist2(dx, dy, dz));
    ax = fmaf(fl, dx, ax); ay = fmaf(fl, dy, ay); az = fmaf(fl, dz, az);
  });
  ax *= P.V; ay *= P.V; az *= P.V;
  if constexpr (FK != 0) {
    const float4 b = B[i];
    if (b.w > 0.0f) {
      float dx = xi.x - b.x, dy = xi.y - b.y, dz = xi.z - b.z;
      float fb = (b.w * li) * spiky_f(P, dist2(dx, dy, dz));
      ax = fmaf(fb, dx, ax); ay = fmaf(fb, dy, ay); az = fmaf(fb, dz, az);
    }
  }
  const Scene &sc = scene_at<MULTI>(s0, scenes, sid, i);
  float px = fmaf(0.5f, ax, xi.x), py = fmaf(0.5f, ay, xi.y), pz = fmaf(0.5f, az, xi.z);
  clamp_rand(sc, px, py, pz);
  Xn[i] = make_float4(px, py, pz, li);
}

// ---------------------------------------------------------------------------
// K12: density and velocity; the state arrays now hold the new order.
// FP64 density-constraint error of the diagnostic: the pair term and the
// boundary term with the clamp to C >= 0 (shared by k_diag and the density
// pass's fused form, so both give the same bits).
__device__ __forceinline__ double diag_pair(const Phys &P, float ax, float ay, float az, float bx, float by,
                                            float bz) {
  const double dx = (double)ax - (double)bx, dy = (double)ay - (double)by, dz = (double)az - (double)bz;
  const double d2 = (dx * dx + dy * dy) + dz * dz;
  if (d2 > P.h2_64) return 0.0;
  const double t = P.h2_64 - d2;
  return P.Kp64 * ((t * t) * t);
}
template <int FK>
__device__ __forceinline__ double diag_finish(const Phys &P, const Scene &sc, const float *__restrict__ pool,
                                              const float4 &xi, double rho) {
  rho *= P.V64;
  if constexpr (FK != 0) {
    const float4 b = boundary_sample<FK == 2 ? 2 : 0>(P, sc, pool, xi.x, xi.y, xi.z);
    if (b.w > 0.0f) rho += (double)b.w * diag_pair(P, xi.x, xi.y, xi.z, b.x, b.y, b.z);
  }
  const double c = rho - 1.0;
  return c > 0.0 ? c : 0.0;
}

// K12: density and velocity; the state arrays now hold the new order.  DIAG:
// also the FP64 density-constraint error of every particle into ci (the
// diagnostic's walk, fused: the same neighbours at the same positions, in
// the same order), for a substep whose error is recorded or checked.
template <bool DIAG, bool MULTI, int FK>
__global__ void __launch_bounds__(BS) k_density(uint32_t n, const float4 *__restrict__ X,
                                                const float4 *__restrict__ Xold, const uint32_t *__restrict__ nbr,
                                                const uint32_t *__restrict__ ncount, float4 *__restrict__ pos,
                                                float4 *__restrict__ vel, float *__restrict__ lam, Phys P,
                                                const float *__restrict__ pool, Scene s0,
                                                const Scene *__restrict__ scenes, const uint32_t *__restrict__ sid,
                                                double *__restrict__ ci) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = X[i];
  float rho = 0.0f;
  double rho64 = 0.0;
  const Gather GX(X, n);
  for_each_neighbor_w<8>(nbr, i, P, ncount[i], [&](uint32_t j) { return GX.xyz(j); },
                         [&](const float3 &xj, bool valid) {
    rho = fmaf(poly6(P, dist2(xi.x - xj.x, xi.y - xj.y, xi.z - xj.z)), valid ? 1.0f : 0.0f, rho);
    if constexpr (DIAG) rho64 += valid ? diag_pair(P, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z) : 0.0;
  });
  if constexpr (DIAG) ci[i] = diag_finish<FK>(P, scene_at<MULTI>(s0, scenes, sid, i), pool, xi, rho64);
  rho *= P.m;
  const float4 xo = Xold[i];
  // w carries 1/rho for the vorticity and XSPH passes
  vel[i] = make_float4((xi.x - xo.x) * P.inv_dt, (xi.y - xo.y) * P.inv_dt, (xi.z - xo.z) * P.inv_dt, 1.0f / rho);
  pos[i] = make_float4(xi.x, xi.y, xi.z, rho);
  lam[i] = xi.w;
}

// K13: vorticity, the Akinci surface normal and the XSPH velocity sum (the
// XSPH sum needs the same relative velocity and distance, so K14 gathers no
// velocities).
__global__ void __launch_bounds__(BS) k_vorticity(uint32_t n, const float4 *__restrict__ pos,
                                                  const float4 *__restrict__ vel, const uint32_t *__restrict__ nbr,
                                                  const uint32_t *__restrict__ ncount, float4 *__restrict__ om,
                                                  float4 *__restrict__ nrm, float4 *__restrict__ visc, Phys P) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = pos[i], vi = vel[i];
  float wx = 0.0f, wy = 0.0f, wz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f, sx = 0.0f, sy = 0.0f, sz = 0.0f;
  // padding entries (j = i): zero relative velocity and gradient
  for_each_neighbor_g<8>(nbr, i, P, ncount[i], [&](uint32_t j, bool) {
    const float4 xj = pos[j], vj = vel[j];
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float ux = vj.x - vi.x, uy = vj.y - vi.y, uz = vj.z - vi.z;
    float d2 = dist2(dx, dy, dz);
    float vol = P.m * vj.w;
    float f = spiky_f(P, d2);
    float gx = f * dx, gy = f * dy, gz = f * dz;
    float cx = fmaf(uy, gz, -(uz * gy)), cy = fmaf(uz, gx, -(ux * gz)), cz = fmaf(ux, gy, -(uy * gx));
    wx = fmaf(vol, cx, wx); wy = fmaf(vol, cy, wy); wz = fmaf(vol, cz, wz);
    nx = fmaf(vol, gx, nx); ny = fmaf(vol, gy, ny); nz = fmaf(vol, gz, nz);
    float sw = poly6(P, d2) * vj.w;
    sx = fmaf(ux, sw, sx); sy = fmaf(uy, sw, sy); sz = fmaf(uz, sw, sz);
  });
  const float wl = sqrtf(dist2(wx, wy, wz));
  om[i] = make_float4(wx, wy, wz, wl);
  // w carries |omega| / rho, the factor K14 applies to each neighbour's
  // gradient
  nrm[i] = make_float4(P.h * nx, P.h * ny, P.h * nz, wl * vi.w);
  visc[i] = make_float4(sx, sy, sz, 0.0f);
}

// K14: XSPH viscosity, vorticity confinement force and surface tension.
__global__ void __launch_bounds__(BS) k_xsph(uint32_t n, const float4 *__restrict__ pos,
                                             const float4 *__restrict__ om, const float4 *__restrict__ nrm,
                                             const float4 *__restrict__ visc, const uint32_t *__restrict__ nbr,
                                             const uint32_t *__restrict__ ncount, float4 *__restrict__ acc, Phys P) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = pos[i], ni = nrm[i], wi = om[i], ui = visc[i];
  const float rhoi = xi.w;
  float ex = 0.0f, ey = 0.0f, ez = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
  // padding entries (j = i): zero gradient, cohesion and normal difference
  for_each_neighbor_g<8>(nbr, i, P, ncount[i], [&](uint32_t j, bool) {
    const float4 xj = pos[j], nj = nrm[j];
    const float rhoj = xj.w;
    float dx = xi.x - xj.x, dy = xi.y - xj.y, dz = xi.z - xj.z;
    float d2 = dist2(dx, dy, dz);
    float K = div_n(P.two_rho0, rhoi + rhoj);
    float se = spiky_f(P, d2) * nj.w;
    ex = fmaf(se, dx, ex); ey = fmaf(se, dy, ey); ez = fmaf(se, dz, ez);
    float mc = P.m * cohesion_f(P, d2);
    tx = fmaf(-K, fmaf(mc, dx, ni.x - nj.x), tx);
    ty = fmaf(-K, fmaf(mc, dy, ni.y - nj.y), ty);
    tz = fmaf(-K, fmaf(mc, dz, ni.z - nj.z), tz);
  });
  const float ux = ui.x, uy = ui.y, uz = ui.z;
  float el = sqrtf(dist2(ex, ey, ez));
  float hx = 0.0f, hy = 0.0f, hz = 0.0f;
  if (el > 0.0f) {
    float ie = 1.0f / el;
    hx = ex * ie; hy = ey * ie; hz = ez * ie;
  }
  float cx = fmaf(hy, wi.z, -(hz * wi.y)), cy = fmaf(hz, wi.x, -(hx * wi.z)), cz = fmaf(hx, wi.y, -(hy * wi.x));
  acc[i] = make_float4(fmaf(P.tension, tx, fmaf(P.vort, cx, P.xsph_k * ux)),
                       fmaf(P.tension, ty, fmaf(P.vort, cy, P.xsph_k * uy)),
                       fmaf(P.tension, tz, fmaf(P.vort, cz, P.xsph_k * uz)), 0.0f);
}

// ---------------------------------------------------------------------------
// FP64 density-constraint error per particle, then one block per scene.
template <bool MULTI, int FK>
__global__ void __launch_bounds__(BS) k_diag(uint32_t n, const float4 *__restrict__ pos,
                                             const uint32_t *__restrict__ nbr, const uint32_t *__restrict__ ncount,
                                             const float *__restrict__ pool, Scene s0,
                                             const Scene *__restrict__ scenes, const uint32_t *__restrict__ sid,
                                             Phys P, double *__restrict__ ci) {
  const uint32_t i = slot_index();
  if (i >= n) return;
  const float4 xi = pos[i];
  double rho = 0.0;
  for_each_neighbor(nbr, i, P, ncount[i], [&](uint32_t j) {
    const float4 xj = pos[j];
    rho += diag_pair(P, xi.x, xi.y, xi.z, xj.x, xj.y, xj.z);
  });
  ci[i] = diag_finish<FK>(P, scene_at<MULTI>(s0, scenes, sid, i), pool, xi, rho);
}

// Per-scene reductions in two deterministic stages.  Stage 1: block (k, s)
// of a (K, nscenes) grid reduces the slots b + k 1024 + t, stride K 1024, of
// scene s (slots [b, b + count) with b = ranges[2s], count = ranges[2s+1])
// into partial[(s K + k) 2 + {0, 1}]; stage 2 combines the K partials of a
// scene in order.  KIND 0: (sum, max) of ci; KIND 1: (sum 0.5 |v|^2,
// sum g y) from the state (gravity is (0, -g, 0)).
template <int KIND>
__global__ void __launch_bounds__(1024) k_reduce_stage1(const double *__restrict__ ci, const float4 *__restrict__ pos,
                                                        const float4 *__restrict__ vel, float g,
                                                        const uint32_t *__restrict__ ranges,
                                                        double *__restrict__ partial) {
  __shared__ double sa[16], sb[16];
  const uint32_t k = blockIdx.x, K = gridDim.x, s = blockIdx.y;
  const uint32_t b = ranges[2 * s], cnt = ranges[2 * s + 1];
  double a = 0.0, m = 0.0;
  for (uint32_t q = k * 1024u + threadIdx.x; q < cnt; q += K * 1024u) {
    if constexpr (KIND == 0) {
      const double v = ci[b + q];
      a += v;
      m = v > m ? v : m;
    } else {
      const float4 v = vel[b + q];
      a += 0.5 * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
      m += (double)g * (double)pos[b + q].y;
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    const double w = __shfl_xor(m, o, 64);
    if constexpr (KIND == 0) m = w > m ? w : m;
    else m += w;
  }
  if (__lane_id() == 0) { sa[threadIdx.x >> 6] = a; sb[threadIdx.x >> 6] = m; }
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0, u = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) {
      t += sa[w];
      if constexpr (KIND == 0) u = sb[w] > u ? sb[w] : u;
      else u += sb[w];
    }
    partial[2 * ((size_t)s * K + k)] = t;
    partial[2 * ((size_t)s * K + k) + 1] = u;
  }
}

// Stage 2 (64 threads, K <= kMaxReduceBlocks = 64): the K partials of a
// scene are loaded at once, then combined in order by one thread.
template <int KIND>
__global__ void __launch_bounds__(64) k_reduce_stage2(const double *__restrict__ partial, uint32_t K,
                                                      double *__restrict__ out) {
  static_assert(kMaxReduceBlocks <= 64, "one partial per thread");
  __shared__ double pa[64], pb[64];
  const uint32_t s = blockIdx.x, k = threadIdx.x;
  if (k < K) {
    pa[k] = partial[2 * ((size_t)s * K + k)];
    pb[k] = partial[2 * ((size_t)s * K + k) + 1];
  }
  __syncthreads();
  if (k != 0) return;
  double t = 0.0, u = 0.0;
  for (uint32_t q = 0; q < K; ++q) {
    t += pa[q];
    const double v = pb[q];
    if constexpr (KIND == 0) u = v > u ? v : u;
    else u += v;
  }
  out[2 * s] = t;
  out[2 * s + 1] = u;
}

// The figures of a learned substep's check in one pass, each reduced exactly
// as its own two-stage reduction would (same slots per thread, same
// butterfly, same combine order, so the same bits): (sum, max) of the
// density error ci, (sum, max) of |v| dt / h (scale = dt / h), and the
// (kinetic, g y) energy sums.  Partials: six doubles per block.
__global__ void __launch_bounds__(1024) k_check_stage1(const double *__restrict__ ci, const float4 *__restrict__ pos,
                                                       const float4 *__restrict__ vel, float g, float scale,
                                                       const uint32_t *__restrict__ ranges,
                                                       double *__restrict__ partial) {
  __shared__ double sv[6][16];
  const uint32_t k = blockIdx.x, K = gridDim.x, s = blockIdx.y;
  const uint32_t b = ranges[2 * s], cnt = ranges[2 * s + 1];
  double r[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (uint32_t q = k * 1024u + threadIdx.x; q < cnt; q += K * 1024u) {
    const double e = ci[b + q];
    r[0] += e;
    r[1] = e > r[1] ? e : r[1];
    const float4 v = vel[b + q];
    const double sp = (double)(sqrtf(dist2(v.x, v.y, v.z)) * scale);
    r[2] += sp;
    r[3] = sp > r[3] ? sp : r[3];
    r[4] += 0.5 * ((double)v.x * v.x + (double)v.y * v.y + (double)v.z * v.z);
    r[5] += (double)g * (double)pos[b + q].y;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
#pragma unroll
    for (int j = 0; j < 6; ++j) {
      const double w = __shfl_xor(r[j], o, 64);
      r[j] = (j == 1 || j == 3) ? (w > r[j] ? w : r[j]) : r[j] + w;
    }
  }
  if (__lane_id() == 0)
#pragma unroll
    for (int j = 0; j < 6; ++j) sv[j][threadIdx.x >> 6] = r[j];
  __syncthreads();
  if (threadIdx.x < 6) {
    const int j = threadIdx.x;
    double t = 0.0;
    for (uint32_t w = 0; w < blockDim.x / 64; ++w) t = (j == 1 || j == 3) ? (sv[j][w] > t ? sv[j][w] : t) : t + sv[j][w];
    partial[6 * ((size_t)s * K + k) + j] = t;
  }
}
// Stage 2 (64 threads): the K partials of each scene loaded at once, then
// combined in order, into the three per-scene (x, y) pairs the gate reads.
__global__ void __launch_bounds__(64) k_check_stage2(const double *__restrict__ partial, uint32_t K,
                                                     double *__restrict__ err, double *__restrict__ speed,
                                                     double *__restrict__ energy) {
  __shared__ double pv[6][64];
  const uint32_t s = blockIdx.x, k = threadIdx.x;
  if (k < K)
#pragma unroll
    for (int j = 0; j < 6; ++j) pv[j][k] = partial[6 * ((size_t)s * K + k) + j];
  __syncthreads();
  const uint32_t j = threadIdx.x;
  if (j >= 6) return;
  double t = 0.0;
  for (uint32_t q = 0; q < K; ++q) {
    const double v = pv[j][q];
    t = (j == 1 || j == 3) ? (v > t ? v : t) : t + v;
  }
  double *out = j < 2 ? err : j < 4 ? speed : energy;
  out[2 * s + (j & 1u)] = t;
}

// The learned substep's check, on the device: *decision = 1 (redo the
// substep plainly) when
//   - the mean density error over all scenes is above the threshold,
//   - the mechanical energy (kinetic + potential, all scenes) rose by more
//     than 0.1 % of the kinetic energy it started from: the plain solve
//     dissipates energy every substep, while an over-pushed warm start lowers
//     the density error by turning the push into motion,
//   - a particle moved more than 2 h (|v| dt / h > 2: an explosion), or
//   - any of these figures is not finite;
// the rejections are counted.  `decision` is pinned host memory the host
// reads after the stream has reached this kernel.
__global__ void k_fallback_gate(const double *__restrict__ err, const double *__restrict__ speed,
                                const double *__restrict__ e0, const double *__restrict__ e1, uint32_t nscenes,
                                uint32_t n, const double *__restrict__ threshold, uint32_t *__restrict__ decision,
                                uint32_t *__restrict__ count) {
  // one wavefront: lane l takes scenes l, l + 64, ..., then a fixed butterfly
  double t = 0.0, v = 0.0, k0 = 0.0, E0 = 0.0, E1 = 0.0;
  bool finite = true;
  for (uint32_t s = threadIdx.x; s < nscenes; s += 64u) {
    t += err[2 * s];
    // a NaN speed drops out of the max but not out of the sum
    finite = finite && speed[2 * s] == speed[2 * s];
    const double m = speed[2 * s + 1];
    v = m > v ? m : v;
    k0 += e0[2 * s];
    E0 += e0[2 * s] + e0[2 * s + 1];
    E1 += e1[2 * s] + e1[2 * s + 1];
  }
  finite = !__any(!finite);
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    t += __shfl_xor(t, o, 64);
    const double w = __shfl_xor(v, o, 64);
    v = w > v ? w : v;
    k0 += __shfl_xor(k0, o, 64);
    E0 += __shfl_xor(E0, o, 64);
    E1 += __shfl_xor(E1, o, 64);
  }
  if (threadIdx.x != 0) return;
  const double mean = n ? t / (double)n : 0.0;
  finite = finite && E1 - E1 == 0.0;
  const bool bad = !finite || !(mean <= *threshold) || !(v <= 2.0) || E1 - E0 > 1e-3 * k0;
  *decision = bad ? 1u : 0u;
  if (bad) *count += 1u;
}

// Sum the per-scene totals into history[ctr % cap] = (mean, max) over all.
// One wavefront: lane l sums scenes l, l + 64, ..., then a fixed butterfly.
__global__ void __launch_bounds__(64) k_diag_history(const double *__restrict__ per_scene, uint32_t nscenes,
                                                     uint32_t n, double *__restrict__ hist, uint32_t cap,
                                                     uint32_t *__restrict__ ctr) {
  double t = 0.0, m = 0.0;
  for (uint32_t s = threadIdx.x; s < nscenes; s += 64u) {
    t += per_scene[2 * s];
    m = per_scene[2 * s + 1] > m ? per_scene[2 * s + 1] : m;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    t += __shfl_xor(t, o, 64);
    const double w = __shfl_xor(m, o, 64);
    m = w > m ? w : m;
  }
  if (threadIdx.x != 0) return;
  uint32_t c = *ctr;
  hist[2 * (c % cap)] = n ? t / (double)n : 0.0;
  hist[2 * (c % cap) + 1] = m;
  *ctr = c + 1;
}

// ---------------------------------------------------------------------------
// Utilities.
__global__ void __launch_bounds__(BS) k_unsort4(uint32_t n, const float4 *__restrict__ src,
                                                const uint32_t *__restrict__ ids, float4 *__restrict__ dst,
                                                int clear_w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float4 v = src[i];
    if (clear_w) v.w = 0.0f;
    dst[ids[i]] = v;
  }
}
__global__ void __launch_bounds__(BS) k_unsort1(uint32_t n, const float *__restrict__ src,
                                                const uint32_t *__restrict__ ids, float4 *__restrict__ dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[ids[i]] = make_float4(src[i], 0.0f, 0.0f, 0.0f);
}
// The inverse gathers (checkpoint upload): slot i takes the caller-order row
// of its particle id; a float4 row, or the .x of one for a scalar field.
__global__ void __launch_bounds__(BS) k_sort4(uint32_t n, const float4 *__restrict__ src,
                                              const uint32_t *__restrict__ ids, float4 *__restrict__ dst,
                                              int clear_w) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    float4 v = src[ids[i]];
    if (clear_w) v.w = 0.0f;
    dst[i] = v;
  }
}
__global__ void __launch_bounds__(BS) k_sort1(uint32_t n, const float4 *__restrict__ src,
                                              const uint32_t *__restrict__ ids, float *__restrict__ dst) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[ids[i]].x;
}
__global__ void __launch_bounds__(BS) k_fill_u32(uint32_t *p, uint32_t n, uint32_t v) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// Plain streaming copy (16 B per lane, grid-stride): the bandwidth probe and
// the scale reference for hardware byte counters.
__global__ void __launch_bounds__(BS) k_copy_probe(const uint4 *__restrict__ src, uint4 *__restrict__ dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}
// Exhaustive check of sqrt_pair over the float bit patterns [lo, hi]: counts
// the x whose result is not the correctly rounded square root.  Exact test
// in double: s is correctly rounded iff m_lo^2 < x < m_hi^2 for the
// midpoints m to s's float neighbours (25-bit values, so the squares are
// exact; x is a float, so it never equals one).
__global__ void __launch_bounds__(BS) k_check_sqrt(uint32_t lo, uint32_t hi, uint32_t per,
                                                   unsigned long long *bad) {
  const uint64_t first = (uint64_t)lo + ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * per;
  uint32_t nbad = 0;
  for (uint32_t k = 0; k < per; ++k) {
    const uint64_t bits = first + k;
    if (bits > hi) break;
    const float x = __uint_as_float((uint32_t)bits);
    const float s = sqrt_pair(x);
    const uint32_t sb = __float_as_uint(s);
    bool ok = s > 0.0f && sb < 0x7f800000u;
    if (ok) {
      const double d = (double)s, m0 = (d + (double)__uint_as_float(sb - 1)) * 0.5,
                   m1 = (d + (double)__uint_as_float(sb + 1)) * 0.5, X = (double)x;
      ok = m0 * m0 < X && X < m1 * m1;
    }
    nbad += ok ? 0u : 1u;
  }
  if (nbad) atomicAdd(bad, (unsigned long long)nbad);
}
__global__ void __launch_bounds__(BS) k_iota(uint32_t *p, uint32_t n) {
  uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = i;
}

// ---------------------------------------------------------------------------
// Launchers.  With timing events set, the first kernel of a pass carries the
// start event and the last one the stop event on its own dispatch packet
// (hipExtLaunchKernelGGL), so the interval is the kernels' execution time.
static inline uint32_t nblocks(uint32_t n, uint32_t bs) { return (n + bs - 1) / bs; }

template <typename K, typename... Args>
static void go(hipStream_t st, hipEvent_t e0, hipEvent_t e1, K kernel, dim3 g, dim3 b, Args... args) {
  if (e0 || e1) hipExtLaunchKernelGGL(kernel, g, b, 0, st, e0, e1, 0, args...);
  else hipLaunchKernelGGL(kernel, g, b, 0, st, args...);
}

void launch_predict_bin(const LaunchCtx &L, const float4 *pos, const float4 *vel, const float4 *acc, float4 *xs,
                        uint32_t *cellkey, uint32_t *rank, uint32_t *counts, uint32_t *flags) {
  if (!L.n) return;
  const dim3 g(nblocks(L.n, BS)), b(BS);
  if (L.multi) go(L.stream, L.ev0, L.ev1, k_predict_bin<true>, g, b, L.n, pos, vel, acc, xs, cellkey, rank, counts, L.s0, L.scenes, L.sid, L.P, flags);
  else go(L.stream, L.ev0, L.ev1, k_predict_bin<false>, g, b, L.n, pos, vel, acc, xs, cellkey, rank, counts, L.s0, L.scenes, L.sid, L.P, flags);
}

void launch_scan(hipStream_t st, uint32_t *in, uint32_t n, uint32_t *out, int zero_in, uint32_t *const *sums,
                 uint32_t *const *prefix, int level, hipEvent_t e0, hipEvent_t e1, uint32_t *zero2) {
  const dim3 b(BS);
  if (n <= SCAN_E) {
    go(st, e0, e1, k_scan_down, dim3(1), b, in, n, (const uint32_t *)nullptr, out, zero_in, zero2);
    return;
  }
  uint32_t nb = (n + SCAN_E - 1) / SCAN_E;
  go(st, e0, nullptr, k_scan_reduce, dim3(nb), b, (const uint32_t *)in, n, sums[level]);
  launch_scan(st, sums[level], nb, prefix[level], 0, sums, prefix, level + 1, nullptr, nullptr, nullptr);
  go(st, nullptr, e1, k_scan_down, dim3(nb), b, in, n, (const uint32_t *)prefix[level], out, zero_in, zero2);
}

void launch_scatter(const LaunchCtx &L, const uint32_t *cellkey, const uint32_t *rank, const uint32_t *offsets,
                    const uint32_t *ids, uint32_t *tmp_key, uint32_t *tmp_src) {
  if (!L.n) return;
  go(L.stream, L.ev0, L.ev1, k_scatter, dim3(nblocks(L.n, BS)), dim3(BS), L.n, cellkey, rank, offsets, ids, tmp_key,
     tmp_src);
}

void launch_settle(const LaunchCtx &L, const uint32_t *tmp_key, const uint32_t *tmp_src, const uint32_t *cellkey,
                   const uint32_t *offsets, const float4 *xs, const float4 *pos, const float *lam_in, float4 *X0,
                   float4 *Xold, uint32_t *ids, float *lam_out, uint32_t *dest) {
  if (!L.n) return;
  go(L.stream, L.ev0, L.ev1, k_settle, dim3(nblocks(L.n, BS)), dim3(BS), L.n, tmp_key, tmp_src, cellkey, offsets, xs,
     pos, lam_in, X0, Xold, ids, lam_out, dest, L.P.slab_log2 ? const_cast<uint32_t *>(L.P.cellsub) : nullptr);
}

template <int R, int B>
static void build_list_rb(const LaunchCtx &L, const float4 *X0, const uint32_t *offsets, uint32_t *nbr,
                          uint32_t *ncount, uint32_t *flags) {
  const dim3 g(nblocks(L.n, B)), b(B);
  if (L.multi) {
    if (L.hash) go(L.stream, L.ev0, L.ev1, k_build_list<true, true, R, B>, g, b, L.n, X0, offsets, nbr, ncount, flags, L.s0, L.scenes, L.sid, L.P);
    else go(L.stream, L.ev0, L.ev1, k_build_list<true, false, R, B>, g, b, L.n, X0, offsets, nbr, ncount, flags, L.s0, L.scenes, L.sid, L.P);
  } else {
    if (L.hash) go(L.stream, L.ev0, L.ev1, k_build_list<false, true, R, B>, g, b, L.n, X0, offsets, nbr, ncount, flags, L.s0, L.scenes, L.sid, L.P);
    else go(L.stream, L.ev0, L.ev1, k_build_list<false, false, R, B>, g, b, L.n, X0, offsets, nbr, ncount, flags, L.s0, L.scenes, L.sid, L.P);
  }
}

// List variants (LDS rows R, work-group size B): 0 (32, 128), 1 (16, 128),
// 2 (32, 64), 3 (56, 64): 16.9, 8.7, 8.4 and 14.6 KB of LDS (R + 1 rows).  (Staging each
// wavefront's nine candidate column ranges in LDS was measured too: 30-60 %
// slower, the candidate loads are L1 hits and not what bounds this pass.)
void launch_build_list(const LaunchCtx &L, const float4 *X0, const uint32_t *offsets, uint32_t *nbr, uint32_t *ncount,
                       uint32_t *flags) {
  if (!L.n) return;
  switch (L.lvar) {
    case 1: build_list_rb<16, 128>(L, X0, offsets, nbr, ncount, flags); break;
    case 2: build_list_rb<32, 64>(L, X0, offsets, nbr, ncount, flags); break;
    case 3: build_list_rb<56, 64>(L, X0, offsets, nbr, ncount, flags); break;
    default: build_list_rb<32, 128>(L, X0, offsets, nbr, ncount, flags); break;
  }
}

template <bool MULTI, template <bool, int> class Sel, typename... Args>
static void fk_dispatch(const LaunchCtx &L, hipEvent_t e0, hipEvent_t e1, dim3 g, Args... args) {
  if (L.fmode == 2) go(L.stream, e0, e1, Sel<MULTI, 2>::fn(), g, dim3(L.bs), args...);
  else if (L.fmode == 1) go(L.stream, e0, e1, Sel<MULTI, 1>::fn(), g, dim3(L.bs), args...);
  else go(L.stream, e0, e1, Sel<MULTI, 0>::fn(), g, dim3(L.bs), args...);
}
template <bool M, int F> struct SelLambda { static auto fn() { return k_lambda<M, F, 0>; } };
template <bool M, int F> struct SelLambdaGather { static auto fn() { return k_lambda<M, F, 0, true>; } };
template <bool M, int F> struct SelLambdaAcc { static auto fn() { return k_lambda<M, F, 1>; } };
template <bool M, int F> struct SelLambdaProj { static auto fn() { return k_lambda<M, F, 2>; } };
template <bool M, int F> struct SelLambdaFeat { static auto fn() { return k_lambda_feat<M, F>; } };
template <bool M, int F> struct SelDp { static auto fn() { return k_dp<M, F>; } };
template <bool M, int F> struct SelDiag { static auto fn() { return k_diag<M, F>; } };

static hipEvent_t boundary_pass(const LaunchCtx &L, float4 *X, float4 *B, dim3 g) {
  if (L.fmode != 1) return L.ev0;
  if (L.multi) go(L.stream, L.ev0, nullptr, k_boundary<true>, g, dim3(BS), L.n, (const float4 *)X, B, L.pool, L.s0, L.scenes, L.sid, L.P);
  else go(L.stream, L.ev0, nullptr, k_boundary<false>, g, dim3(BS), L.n, (const float4 *)X, B, L.pool, L.s0, L.scenes, L.sid, L.P);
  return nullptr;
}

void launch_lambda(const LaunchCtx &L, float4 *X, float4 *B, const uint32_t *nbr, const uint32_t *ncount,
                   float *lam_acc, bool project) {
  if (!L.n) return;
  hipEvent_t e0 = boundary_pass(L, X, B, dim3(nblocks(L.n, BS)));
  const dim3 g(nblocks(L.n, L.bs));
  if (lam_acc && project) {
    if (L.multi) fk_dispatch<true, SelLambdaProj>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
    else fk_dispatch<false, SelLambdaProj>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
  } else if (lam_acc) {
    if (L.multi) fk_dispatch<true, SelLambdaAcc>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
    else fk_dispatch<false, SelLambdaAcc>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
  } else if (L.hash) {
    if (L.multi) fk_dispatch<true, SelLambdaGather>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
    else fk_dispatch<false, SelLambdaGather>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
  } else {
    if (L.multi) fk_dispatch<true, SelLambda>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
    else fk_dispatch<false, SelLambda>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lam_acc);
  }
}

void launch_lambda_features(const LaunchCtx &L, float4 *X, float4 *B, const uint32_t *nbr, const uint32_t *ncount,
                            const LearnBufs &lb) {
  if (!L.n) return;
  hipEvent_t e0 = boundary_pass(L, X, B, dim3(nblocks(L.n, BS)));
  const dim3 g(nblocks(L.n, L.bs));
  if (L.multi) fk_dispatch<true, SelLambdaFeat>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lb);
  else fk_dispatch<false, SelLambdaFeat>(L, e0, L.ev1, g, L.n, X, B, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, lb);
}

void launch_settle_extra(const LaunchCtx &L, const uint32_t *tmp_src, const uint32_t *dest, const float *acc_in,
                         float *acc_out, const float4 *vel_in, float4 *vel_out) {
  if (!L.n) return;
  go(L.stream, nullptr, nullptr, k_settle_extra, dim3(nblocks(L.n, BS)), dim3(BS), L.n, tmp_src, dest, acc_in, acc_out,
     vel_in, vel_out);
}

void launch_warm_previous(const LaunchCtx &L, const float *acc_prev, float4 *X, float *lam_acc, float relax) {
  if (!L.n) return;
  go(L.stream, nullptr, nullptr, k_warm_previous, dim3(nblocks(L.n, BS)), dim3(BS), L.n, acc_prev, X, lam_acc, relax);
}

void launch_copy(hipStream_t st, void *dst, const void *src, size_t bytes) {
  size_t n = bytes / 4;
  uint32_t grid = (uint32_t)std::min<size_t>((n + BS - 1) / BS, 4096u);
  if (n) hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(BS), 0, st, (uint32_t *)dst, (const uint32_t *)src, n);
}

void launch_smooth_warm(const LaunchCtx &L, float4 *X, const uint32_t *nbr, const uint32_t *ncount, float *tmp,
                        float *lam_acc) {
  if (!L.n) return;
  go(L.stream, nullptr, nullptr, k_smooth_warm, dim3(nblocks(L.n, BS)), dim3(BS), L.n, (const float4 *)X, nbr, ncount,
     L.P, tmp);
  go(L.stream, nullptr, nullptr, k_apply_warm, dim3(nblocks(L.n, BS)), dim3(BS), L.n, (const float *)tmp, X, lam_acc);
}

void launch_fallback_gate(hipStream_t st, const double *err, const double *speed, const double *e0, const double *e1,
                          uint32_t nscenes, uint32_t n, const double *threshold, uint32_t *decision,
                          uint32_t *count) {
  hipLaunchKernelGGL(k_fallback_gate, dim3(1), dim3(64), 0, st, err, speed, e0, e1, nscenes, n, threshold, decision,
                     count);
}

void launch_dp(const LaunchCtx &L, const float4 *X, const float4 *B, const uint32_t *nbr, const uint32_t *ncount,
               float4 *Xn) {
  if (!L.n) return;
  const dim3 g(nblocks(L.n, L.bs));
  if (L.multi) fk_dispatch<true, SelDp>(L, L.ev0, L.ev1, g, L.n, X, B, nbr, ncount, Xn, L.s0, L.scenes, L.sid, L.P);
  else fk_dispatch<false, SelDp>(L, L.ev0, L.ev1, g, L.n, X, B, nbr, ncount, Xn, L.s0, L.scenes, L.sid, L.P);
}

void launch_density(const LaunchCtx &L, const float4 *X, const float4 *Xold, const uint32_t *nbr,
                    const uint32_t *ncount, float4 *pos, float4 *vel, float *lam, double *ci) {
  if (!L.n) return;
  const dim3 g(nblocks(L.n, L.bs));
  if (!ci)
    go(L.stream, L.ev0, L.ev1, k_density<false, false, 0>, g, dim3(L.bs), L.n, X, Xold, nbr, ncount, pos, vel, lam,
       L.P, L.pool, L.s0, L.scenes, L.sid, ci);
  else if (L.multi)
    go(L.stream, L.ev0, L.ev1, L.fmode ? k_density<true, true, 2> : k_density<true, true, 0>, g, dim3(L.bs), L.n, X,
       Xold, nbr, ncount, pos, vel, lam, L.P, L.pool, L.s0, L.scenes, L.sid, ci);
  else
    go(L.stream, L.ev0, L.ev1, L.fmode ? k_density<true, false, 2> : k_density<true, false, 0>, g, dim3(L.bs), L.n, X,
       Xold, nbr, ncount, pos, vel, lam, L.P, L.pool, L.s0, L.scenes, L.sid, ci);
}

void launch_vorticity(const LaunchCtx &L, const float4 *pos, const float4 *vel, const uint32_t *nbr,
                      const uint32_t *ncount, float4 *om, float4 *nrm, float4 *visc) {
  if (!L.n) return;
  go(L.stream, L.ev0, L.ev1, k_vorticity, dim3(nblocks(L.n, L.bs)), dim3(L.bs), L.n, pos, vel, nbr, ncount, om, nrm,
     visc, L.P);
}

void launch_xsph(const LaunchCtx &L, const float4 *pos, const float4 *om, const float4 *nrm, const float4 *visc,
                 const uint32_t *nbr, const uint32_t *ncount, float4 *acc) {
  if (!L.n) return;
  go(L.stream, L.ev0, L.ev1, k_xsph, dim3(nblocks(L.n, L.bs)), dim3(L.bs), L.n, pos, om, nrm, visc, nbr, ncount, acc,
     L.P);
}

template <int KIND>
static void reduce_scenes(const LaunchCtx &L, const double *ci, const float4 *pos, const float4 *vel, float g,
                          const uint32_t *ranges, uint32_t nscenes, double *out, hipEvent_t e1 = nullptr,
                          hipEvent_t e0 = nullptr) {
  go(L.stream, e0, nullptr, k_reduce_stage1<KIND>, dim3(L.red_k, nscenes), dim3(1024), ci, pos, vel, g, ranges,
     L.partial);
  go(L.stream, nullptr, e1, k_reduce_stage2<KIND>, dim3(nscenes), dim3(64), (const double *)L.partial, L.red_k, out);
}

void launch_diag(const LaunchCtx &L, const float4 *pos, const uint32_t *nbr, const uint32_t *ncount, double *ci,
                 const uint32_t *ranges, uint32_t nscenes, double *per_scene, double *hist, uint32_t hist_cap,
                 uint32_t *hist_ctr) {
  if (!L.n) return;
  const dim3 g(nblocks(L.n, L.bs));
  if (L.multi) fk_dispatch<true, SelDiag>(L, L.ev0, nullptr, g, L.n, pos, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, ci);
  else fk_dispatch<false, SelDiag>(L, L.ev0, nullptr, g, L.n, pos, nbr, ncount, L.pool, L.s0, L.scenes, L.sid, L.P, ci);
  LaunchCtx R = L;
  R.ev0 = nullptr;
  launch_diag_reduce(R, ci, ranges, nscenes, per_scene, hist, hist_cap, hist_ctr);
}

void launch_diag_reduce(const LaunchCtx &L, const double *ci, const uint32_t *ranges, uint32_t nscenes,
                        double *per_scene, double *hist, uint32_t hist_cap, uint32_t *hist_ctr) {
  if (!L.n) return;
  reduce_scenes<0>(L, ci, nullptr, nullptr, 0.0f, ranges, nscenes, per_scene, hist ? nullptr : L.ev1, L.ev0);
  if (hist)
    go(L.stream, nullptr, L.ev1, k_diag_history, dim3(1), dim3(64), (const double *)per_scene, nscenes, L.n, hist,
       hist_cap, hist_ctr);
}

void launch_diag_history(const LaunchCtx &L, const double *per_scene, uint32_t nscenes, double *hist,
                         uint32_t hist_cap, uint32_t *hist_ctr) {
  go(L.stream, L.ev0, L.ev1, k_diag_history, dim3(1), dim3(64), per_scene, nscenes, L.n, hist, hist_cap, hist_ctr);
}

void launch_check_reduce(const LaunchCtx &L, const double *ci, const float4 *pos, const float4 *vel,
                         const uint32_t *ranges, uint32_t nscenes, double *err, double *speed, double *energy) {
  if (!L.n) return;
  go(L.stream, nullptr, nullptr, k_check_stage1, dim3(L.red_k, nscenes), dim3(1024), ci, pos, vel, -L.P.gy,
     L.P.dt / L.P.h, ranges, L.partial);
  go(L.stream, nullptr, nullptr, k_check_stage2, dim3(nscenes), dim3(64), (const double *)L.partial, L.red_k, err,
     speed, energy);
}

void launch_energy(const LaunchCtx &L, const float4 *pos, const float4 *vel, const uint32_t *ranges,
                   uint32_t nscenes, double *out) {
  if (!L.n) return;
  reduce_scenes<1>(L, nullptr, pos, vel, -L.P.gy, ranges, nscenes, out);
}

void launch_unsort4(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float4 *dst, int clear_w) {
  if (n) hipLaunchKernelGGL(k_unsort4, dim3(nblocks(n, BS)), dim3(BS), 0, st, n, src, ids, dst, clear_w);
}
void launch_sort4(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float4 *dst, int clear_w) {
  if (n) hipLaunchKernelGGL(k_sort4, dim3(nblocks(n, BS)), dim3(BS), 0, st, n, src, ids, dst, clear_w);
}
void launch_sort1(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float *dst) {
  if (n) hipLaunchKernelGGL(k_sort1, dim3(nblocks(n, BS)), dim3(BS), 0, st, n, src, ids, dst);
}
void launch_unsort1(hipStream_t st, uint32_t n, const float *src, const uint32_t *ids, float4 *dst) {
  if (n) hipLaunchKernelGGL(k_unsort1, dim3(nblocks(n, BS)), dim3(BS), 0, st, n, src, ids, dst);
}
void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v) {
  if (n) hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(n, BS)), dim3(BS), 0, st, p, n, v);
}
void launch_copy_probe(hipStream_t st, const void *src, void *dst, size_t bytes) {
  size_t n = bytes / 16;
  uint32_t grid = (uint32_t)std::min<size_t>((n + BS - 1) / BS, 256u * 64u);
  if (n) hipLaunchKernelGGL(k_copy_probe, dim3(grid), dim3(BS), 0, st, (const uint4 *)src, (uint4 *)dst, n);
}
void launch_check_sqrt(hipStream_t st, uint32_t lo, uint32_t hi, unsigned long long *bad) {
  if (hi < lo) return;
  const uint32_t per = 256;
  const uint64_t threads = ((uint64_t)hi - lo + per) / per;
  hipLaunchKernelGGL(k_check_sqrt, dim3((uint32_t)((threads + BS - 1) / BS)), dim3(BS), 0, st, lo, hi, per, bad);
}
void launch_iota(hipStream_t st, uint32_t *p, uint32_t n) {
  if (n) hipLaunchKernelGGL(k_iota, dim3(nblocks(n, BS)), dim3(BS), 0, st, p, n);
}

}  // namespace pbf
