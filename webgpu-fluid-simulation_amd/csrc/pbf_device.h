// pbf_device.h -- device-side math of the PBF substep (gfx950).
//
// Numerics follow one fixed specification so that every pass is reproducible
// bit for bit: constants formed in double and rounded once, no implicit
// contraction (the library is built with -ffp-contract=off) and fused
// multiply-adds only where written as fmaf(), sums in evaluation order,
// correctly rounded division and square root, and the wall-clamp jitter's
// sin() evaluated with the Cephes single-precision algorithm (S. L. Moshier)
// in plain float arithmetic.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace pbf {

// ---------------------------------------------------------------------------
// Per-launch constants.
struct Phys {
  float h, h2, half_h, half_r, two_r;
  float V, V2, m, Kp, Ks, Kc, bias, two_rho0, eps_lambda;
  float dt, inv_dt;
  float gx, gy, gz;
  float xsph_k;  // (xsph * inv_dt) * m
  float vort, tension;
  uint32_t ld;   // leading dimension of the slot-major neighbour table
  uint32_t maxn; // neighbour capacity per particle
  uint32_t *nbr16;  // 16-bit neighbour table, 8-entry chunks, slot-major
  uint32_t *wfmt;   // per wavefront: 1 = 16-bit table, 0 = 32-bit table
  const uint32_t *cellsub;  // per cell: its particles below z-slabs 1, 2, 3 and its count (8 bits each)
  uint32_t slab_log2;       // log2 of the z-slabs per cell (0: one slab)
  double Kp64, h2_64, V64;
};

// Sort keys inside a cell: the particle id in bits 0..26 (capacities stay
// below 2^27) and its z-slab above it; the binning rank carries the slab in
// its top two bits until the scatter forms the key.
constexpr uint32_t kIdBits = 27;
constexpr uint32_t kIdMask = (1u << kIdBits) - 1u;
constexpr uint32_t kRankSlabShift = 30;
constexpr uint32_t kRankMask = (1u << kRankSlabShift) - 1u;
// cellsub word of a cell holding more particles than 8 bits count: no trimming
constexpr uint32_t kSubSaturated = 0xff000000u;

// One scene: clamp box, grid, particle range and boundary field.
struct Scene {
  float lo[3], hi[3];
  float inv_cell;
  int32_t dims[3];
  uint32_t cell_base;
  uint32_t ncells;
  uint32_t hash_mask;  // 0: dense grid
  uint32_t distinct27;  // the 27 cells around any cell map to distinct buckets (always, dense)
  uint32_t pbase, pcount;
  int32_t field_kind;  // 0 none, 1 cubic 32-node, 2 trilinear
  int32_t fres[3];
  float flo[3], finv[3], fscale[3];
  uint32_t quirk;
  uint64_t field_off;  // float offset of this scene's field in the pool
};

// ---------------------------------------------------------------------------
// Cephes sinf (Moshier): reduction modulo pi/4 with DP1+DP2+DP3, then the
// degree-7 sine / degree-8 cosine polynomials.
__device__ __forceinline__ float cephes_sinf(float xx) {
  float x = xx;
  int neg = 0;
  if (!(x == x)) return x;
  if (x < 0.0f) { neg = 1; x = -x; }
  if (x > 8192.0f) return 0.0f;
  uint32_t j = (uint32_t)(1.27323954473516f * x);
  float y = (float)j;
  if (j & 1u) { j += 1u; y += 1.0f; }
  j &= 7u;
  if (j > 3u) { neg = !neg; j -= 4u; }
  x = ((x - y * 0.78515625f) - y * 2.4187564849853515625e-4f) - y * 3.77489497744594108e-8f;
  float z = x * x;
  float res;
  if (j == 1u || j == 2u) {
    res = ((2.443315711809948e-5f * z - 1.388731625493765e-3f) * z + 4.166664568298827e-2f) * z * z;
    res = res - 0.5f * z;
    res = res + 1.0f;
  } else {
    res = ((-1.9515295891e-4f * z + 8.3321608736e-3f) * z - 1.6666654611e-1f) * z * x;
    res = res + x;
  }
  return neg ? -res : res;
}

__device__ __forceinline__ float fract_(float v) { return v - floorf(v); }

// Wall clamp with per-axis pseudo-random inset of at most 1e-4.
__device__ __forceinline__ void clamp_rand(const Scene &sc, float &px, float &py, float &pz) {
  const float A = 12.9898f, B = 78.233f;
  float b0 = fmaf(pz, B, py * A);
  float b1 = fmaf(pz, B, px * A);
  float b2 = fmaf(py, B, px * A);
  float r0 = fract_(cephes_sinf(b0) * 43758.5453f);
  float r1 = fract_(cephes_sinf(b1) * 43758.5453f);
  float r2 = fract_(cephes_sinf(b2) * 43758.5453f);
  float p[3] = {px, py, pz};
  float rr[3] = {r0, r1, r2};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    float e = 1e-4f * rr[c];
    float lo = sc.lo[c] + e;
    float hi = sc.hi[c] - e;
    float mm = p[c] < hi ? p[c] : hi;
    p[c] = mm > lo ? mm : lo;
  }
  px = p[0]; py = p[1]; pz = p[2];
}

// ---------------------------------------------------------------------------
// Correctly rounded square root and division for the operand ranges of the
// pair terms, equal to sqrtf() and '/' with
// -fhip-fp32-correctly-rounded-divide-sqrt there (the parity tests check
// them bit for bit against the CPU restatement).
//
// Square root of a pair distance d2: the hardware reciprocal square root y
// and one Newton step on s = d2 y, s + (d2 - s^2) y / 2 with the residual
// exact in an fma -- five instructions where testing the neighbours
// s -/+ 1 ulp of v_sqrt_f32 takes nine (v_sqrt_f32 is off by one ulp for
// about 15 % of inputs, in both directions).  The result is the correctly
// rounded square root for every float in [2^-100, FLT_MAX]: checked
// exhaustively on gfx950 (k_check_sqrt, pbf_check_device_sqrt).  d2 = 0
// gives NaN and the callers select their term to 0 there; a d2 in
// (0, 2^-100) needs two particles closer than 1e-15 in every coordinate,
// which positions of any useful scale cannot be.
__device__ __forceinline__ float sqrt_pair(float x) {
  const float y = __builtin_amdgcn_rsqf(x);
  const float s = x * y;
  const float r = fmaf(-s, s, x);
  return fmaf(r, 0.5f * y, s);
}
// Division: one Newton step on the hardware reciprocal (the correctly
// rounded 1/b for every b in [2^-100, 2^126), checked exhaustively on gfx950),
// then two fma remainder corrections of the quotient.  The general form adds
// pre- and post-scaling for denormal, huge or special operands; the pair
// terms never produce those (r lies in [2^-50, h], densities are ~rho0).
__device__ __forceinline__ float div_n(float a, float b) {
  float y = __builtin_amdgcn_rcpf(b);
  const float e = fmaf(-b, y, 1.0f);
  y = fmaf(e, y, y);
  float q = a * y;
  float r = fmaf(-b, q, a);
  q = fmaf(r, y, q);
  r = fmaf(-b, q, a);
  return fmaf(r, y, q);
}

// ---------------------------------------------------------------------------
// SPH kernels as functions of d2 = |x_ij|^2.  Written without branches
// (select after computing) so a neighbour loop stays one basic block and its
// gathers can be issued ahead of the arithmetic; the selected values are the
// same as those of the guarded forms.
__device__ __forceinline__ float poly6(const Phys &P, float d2) {
  float t = P.h2 - d2;
  float w = P.Kp * ((t * t) * t);
  return d2 <= P.h2 ? w : 0.0f;
}
__device__ __forceinline__ float spiky_f(const Phys &P, float d2) {
  float r = sqrt_pair(d2);
  float q = P.h - r;
  float f = div_n(P.Ks * (q * q), r);
  return (d2 > 0.0f && d2 <= P.h2) ? f : 0.0f;
}
__device__ __forceinline__ float cohesion_f(const Phys &P, float d2) {
  float r = sqrt_pair(d2);
  float a = P.h - r;
  float t = ((a * a) * a) * ((r * r) * r);
  float res = (r > P.half_h) ? t : fmaf(2.0f, t, P.bias);
  float f = div_n(P.Kc * res, r);
  return (d2 > 0.0f && d2 <= P.h2) ? f : 0.0f;
}

__device__ __forceinline__ float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

// ---------------------------------------------------------------------------
// Grid.
__device__ __forceinline__ int cell_axis(const Scene &sc, float x, int a) {
  int v = (int)floorf((x - sc.lo[a]) * sc.inv_cell);
  v = v < 0 ? 0 : v;
  v = v > sc.dims[a] - 1 ? sc.dims[a] - 1 : v;
  return v;
}
// z-slab of a position inside its (clamped) cell: floor(S * fz) with
// fz = (z - lo) / h - cz in float, clamped to [0, S), S = 2^slab_log2.
__device__ __forceinline__ uint32_t z_slab(const Scene &sc, uint32_t slab_log2, float z, int cz) {
  const float fz = (z - sc.lo[2]) * sc.inv_cell - (float)cz;
  const int top = (1 << slab_log2) - 1;
  int k = (int)floorf(fz * (float)(1u << slab_log2));
  k = k < 0 ? 0 : k;
  return (uint32_t)(k > top ? top : k);
}
__device__ __forceinline__ uint32_t cell_local(const Scene &sc, int cx, int cy, int cz) {
  // hashed grid: the row-major index modulo the table size (2^k), so nearby
  // cells land in nearby buckets and the bucket-sorted particles keep their
  // spatial locality; cells 2^k apart in that index share a bucket
  const uint32_t lin = ((uint32_t)cx * (uint32_t)sc.dims[1] + (uint32_t)cy) * (uint32_t)sc.dims[2] + (uint32_t)cz;
  return sc.hash_mask ? (lin & sc.hash_mask) : lin;
}

// ---------------------------------------------------------------------------
// Binning.
// Rank of this lane's key among equal keys, one global atomic per distinct
// key per wavefront (keys of cell-sorted particles repeat within a wave).
// The groups are found first with ballots only; then every group's leader
// issues its atomic in one instruction, so the wavefront waits for one
// atomic round trip instead of one per group.
__device__ __forceinline__ uint32_t wave_atomic_rank(uint32_t *counts, uint32_t key, bool active) {
  const uint32_t lane = __lane_id();
  unsigned long long pending = __ballot(active);
  const unsigned long long lt = (lane == 0) ? 0ull : ((~0ull) >> (64 - lane));
  unsigned long long my_same = 0;
  int my_leader = 0;
  while (pending) {
    const int leader = __ffsll((long long)pending) - 1;
    const uint32_t lkey = (uint32_t)__shfl((int)key, leader, 64);
    const bool mine = active && key == lkey;
    const unsigned long long same = __ballot(mine) & pending;
    if (mine) {
      my_same = same;
      my_leader = leader;
    }
    pending &= ~same;
  }
  uint32_t base = 0;
  if (active && (int)lane == my_leader) base = atomicAdd(&counts[key], (uint32_t)__popcll(my_same));
  base = (uint32_t)__shfl((int)base, my_leader, 64);
  return base + (uint32_t)__popcll(my_same & lt);
}

// ---------------------------------------------------------------------------
// Boundary fields.
__device__ __forceinline__ float dot4(const float *N, const float *v) {
  return ((N[0] * v[0] + N[1] * v[1]) + N[2] * v[2]) + N[3] * v[3];
}

// 32-node cubic serendipity element: values N[32] and derivatives dN[32][3]
// with respect to the local coordinate s in [-1, 1]^3.
__device__ __forceinline__ void cubic_shape(float x, float y, float z, float N[32], float dN[32][3]) {
  const float C1 = 0.015625f, C9 = 0.140625f, C19 = 0.296875f;
  float xx = x * x, yy = y * y, zz = z * z;
  float ssum = (xx + yy) + zz;
  float mx = 1.0f - x, px = 1.0f + x, my = 1.0f - y, py = 1.0f + y, mz = 1.0f - z, pz = 1.0f + z;
  float qx = 1.0f - xx, qy = 1.0f - yy, qz = 1.0f - zz;
  float ax = 1.0f - 3.0f * x, bx = 1.0f + 3.0f * x;
  float ay = 1.0f - 3.0f * y, by = 1.0f + 3.0f * y;
  float az = 1.0f - 3.0f * z, bz = 1.0f + 3.0f * z;
  float mxmy = mx * my, pxmy = px * my, mxpy = mx * py, pxpy = px * py;
  float mxmz = mx * mz, pxmz = px * mz, mxpz = mx * pz, pxpz = px * pz;
  float mymz = my * mz, pymz = py * mz, mypz = my * pz, pypz = py * pz;
  float fac = C9 * ssum - C19;
  float f0 = fac * mz, f1 = fac * pz;
  N[0] = f0 * mxmy; N[1] = f0 * pxmy; N[2] = f0 * mxpy; N[3] = f0 * pxpy;
  N[4] = f1 * mxmy; N[5] = f1 * pxmy; N[6] = f1 * mxpy; N[7] = f1 * pxpy;
  float e, ea, eb;
  e = C9 * qx; ea = e * ax; eb = e * bx;
  N[8] = ea * mymz;  N[9] = eb * mymz;  N[10] = ea * mypz; N[11] = eb * mypz;
  N[12] = ea * pymz; N[13] = eb * pymz; N[14] = ea * pypz; N[15] = eb * pypz;
  e = C9 * qy; ea = e * ay; eb = e * by;
  N[16] = ea * mxmz; N[17] = eb * mxmz; N[18] = ea * pxmz; N[19] = eb * pxmz;
  N[20] = ea * mxpz; N[21] = eb * mxpz; N[22] = ea * pxpz; N[23] = eb * pxpz;
  e = C9 * qz; ea = e * az; eb = e * bz;
  N[24] = ea * mxmy; N[25] = eb * mxmy; N[26] = ea * mxpy; N[27] = eb * mxpy;
  N[28] = ea * pxmy; N[29] = eb * pxmy; N[30] = ea * pxpy; N[31] = eb * pxpy;

  float wx = 9.0f * (2.0f * xx + ssum) - 19.0f;
  float wy = 9.0f * (2.0f * yy + ssum) - 19.0f;
  float wz = 9.0f * (2.0f * zz + ssum) - 19.0f;
  float DX[2] = {18.0f * x - wx, 18.0f * x + wx};
  float DY[2] = {18.0f * y - wy, 18.0f * y + wy};
  float DZ[2] = {18.0f * z - wz, 18.0f * z + wz};
  float Xs[2] = {mx, px}, Ys[2] = {my, py}, Zs[2] = {mz, pz};
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    int ex = c & 1, ey = (c >> 1) & 1, ez = (c >> 2) & 1;
    float X = Xs[ex], Y = Ys[ey], Z = Zs[ez];
    dN[c][0] = C1 * (DX[ex] * (Y * Z));
    dN[c][1] = C1 * (DY[ey] * (X * Z));
    dN[c][2] = C1 * (DZ[ez] * (X * Y));
  }
  float CX[2] = {(-(3.0f - 9.0f * xx)) - 2.0f * x, (3.0f - 9.0f * xx) - 2.0f * x};
  float CY[2] = {(-(3.0f - 9.0f * yy)) - 2.0f * y, (3.0f - 9.0f * yy) - 2.0f * y};
  float CZ[2] = {(-(3.0f - 9.0f * zz)) - 2.0f * z, (3.0f - 9.0f * zz) - 2.0f * z};
  float QX[2] = {qx * ax, qx * bx}, QY[2] = {qy * ay, qy * by}, QZ[2] = {qz * az, qz * bz};
#pragma unroll
  for (int ey = 0; ey < 2; ++ey)
#pragma unroll
    for (int ez = 0; ez < 2; ++ez)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        int k = 8 + 4 * ey + 2 * ez + t;
        float qa = QX[t];
        dN[k][0] = C9 * (CX[t] * (Ys[ey] * Zs[ez]));
        dN[k][1] = C9 * ((ey ? qa : -qa) * Zs[ez]);
        dN[k][2] = C9 * ((ez ? qa : -qa) * Ys[ey]);
      }
#pragma unroll
  for (int ez = 0; ez < 2; ++ez)
#pragma unroll
    for (int ex = 0; ex < 2; ++ex)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        int k = 16 + 4 * ez + 2 * ex + t;
        float qa = QY[t];
        dN[k][0] = C9 * ((ex ? qa : -qa) * Zs[ez]);
        dN[k][1] = C9 * (CY[t] * (Xs[ex] * Zs[ez]));
        dN[k][2] = C9 * ((ez ? qa : -qa) * Xs[ex]);
      }
#pragma unroll
  for (int ex = 0; ex < 2; ++ex)
#pragma unroll
    for (int ey = 0; ey < 2; ++ey)
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        int k = 24 + 4 * ex + 2 * ey + t;
        float qa = QZ[t];
        dN[k][0] = C9 * ((ex ? qa : -qa) * Ys[ey]);
        dN[k][1] = C9 * ((ey ? qa : -qa) * Xs[ex]);
        dN[k][2] = C9 * (CZ[t] * (Xs[ex] * Ys[ey]));
      }
}

// Node value indices of cell (i, j, k) in the .cdm layout: corner[k][j][i],
// edgex[k][j][i2], edgey[i][k][j2], edgez[j][i][k2] concatenated.
__device__ __forceinline__ void cubic_nodes(const int32_t *R, uint32_t i, uint32_t j, uint32_t k, uint32_t quirk,
                                            uint32_t idx[32]) {
  uint32_t nx = (uint32_t)R[0] + 1, ny = (uint32_t)R[1] + 1, nz = (uint32_t)R[2] + 1;
  uint32_t wx = 2u * (uint32_t)R[0], wy = 2u * (uint32_t)R[1], wz = 2u * (uint32_t)R[2];
  uint32_t ncorner = nx * ny * nz;
  uint32_t bex = ncorner, bey = ncorner + nz * ny * wx, bez = bey + nx * nz * wy;
  auto CO = [&](uint32_t a, uint32_t b, uint32_t c) { return (c * ny + b) * nx + a; };
  auto EX = [&](uint32_t a, uint32_t b, uint32_t c) { return bex + (c * ny + b) * wx + a; };
  auto EY = [&](uint32_t a, uint32_t b, uint32_t c) { return bey + (b * nz + c) * wy + a; };
  auto EZ = [&](uint32_t a, uint32_t b, uint32_t c) { return bez + (c * nx + b) * wz + a; };
  uint32_t i2 = 2u * i, j2 = 2u * j, k2 = 2u * k;
  idx[0] = CO(i, j, k); idx[1] = CO(i + 1, j, k); idx[2] = CO(i, j + 1, k); idx[3] = CO(i + 1, j + 1, k);
  idx[4] = CO(i, j, k + 1); idx[5] = CO(i + 1, j, k + 1); idx[6] = CO(i, j + 1, k + 1); idx[7] = CO(i + 1, j + 1, k + 1);
  uint32_t i2b = quirk ? i2 : i2 + 1;
  idx[8] = EX(i2, j, k); idx[9] = EX(i2 + 1, j, k); idx[10] = EX(i2, j, k + 1); idx[11] = EX(i2b, j, k + 1);
  idx[12] = EX(i2, j + 1, k); idx[13] = EX(i2 + 1, j + 1, k); idx[14] = EX(i2, j + 1, k + 1); idx[15] = EX(i2b, j + 1, k + 1);
  idx[16] = EY(j2, i, k); idx[17] = EY(j2 + 1, i, k); idx[18] = EY(j2, i + 1, k); idx[19] = EY(j2 + 1, i + 1, k);
  idx[20] = EY(j2, i, k + 1); idx[21] = EY(j2 + 1, i, k + 1); idx[22] = EY(j2, i + 1, k + 1); idx[23] = EY(j2 + 1, i + 1, k + 1);
  uint32_t k2b = quirk ? k2 : k2 + 1;
  idx[24] = EZ(k2, i, j); idx[25] = EZ(k2b, i, j); idx[26] = EZ(k2, i, j + 1); idx[27] = EZ(k2b, i, j + 1);
  idx[28] = EZ(k2, i + 1, j); idx[29] = EZ(k2b, i + 1, j); idx[30] = EZ(k2, i + 1, j + 1); idx[31] = EZ(k2b, i + 1, j + 1);
}

__device__ __forceinline__ float sum32(const float *N, const float *v) {
  float r[8];
#pragma unroll
  for (int g = 0; g < 8; ++g) r[g] = dot4(N + 4 * g, v + 4 * g);
  float q0 = r[0] + r[4], q1 = r[1] + r[5], q2 = r[2] + r[6], q3 = r[3] + r[7];
  return ((q0 + q1) + q2) + q3;
}

// sdf value, gradient and volume of a scene's field.  KIND 1: cubic 32-node
// element (.cdm layout), 2: trilinear node grid, 0: decided at run time.
template <int KIND>
__device__ __forceinline__ void field_eval(const Scene &sc, const float *__restrict__ pool, float x0, float x1, float x2,
                                           float g[3], float &sdf, float &vol, bool want_vol) {
  const float *F = pool + sc.field_off;
  float xv[3] = {x0, x1, x2};
  if (KIND == 1 || (KIND == 0 && sc.field_kind == 1)) {
    uint32_t cell[3];
    float sv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float u = (xv[a] - sc.flo[a]) * sc.finv[a];
      float fl = floorf(u);
      int c = (int)fl;
      if (c < 0) { c = 0; fl = 0.0f; }
      if (c > sc.fres[a] - 1) { c = sc.fres[a] - 1; fl = (float)c; }
      cell[a] = (uint32_t)c;
      sv[a] = 2.0f * (u - fl) - 1.0f;
    }
    float N[32], dN[32][3];
    cubic_shape(sv[0], sv[1], sv[2], N, dN);
    uint32_t idx[32];
    cubic_nodes(sc.fres, cell[0], cell[1], cell[2], 0u, idx);
    float v[32];
#pragma unroll
    for (int k = 0; k < 32; ++k) v[k] = F[idx[k]];
    sdf = sum32(N, v);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float G[8];
#pragma unroll
      for (int q = 0; q < 8; ++q)
        G[q] = ((dN[4 * q][a] * v[4 * q] + dN[4 * q + 1][a] * v[4 * q + 1]) + dN[4 * q + 2][a] * v[4 * q + 2]) +
               dN[4 * q + 3][a] * v[4 * q + 3];
      float r3 = ((G[0] + G[1]) + G[2]) + G[3];
      float r4 = ((G[4] + G[5]) + G[6]) + G[7];
      g[a] = sc.fscale[a] * (r3 + r4);
    }
    if (want_vol) {
      uint32_t nx = (uint32_t)sc.fres[0] + 1, ny = (uint32_t)sc.fres[1] + 1, nz = (uint32_t)sc.fres[2] + 1;
      uint32_t flen = nx * ny * nz + nz * ny * 2u * (uint32_t)sc.fres[0] + nx * nz * 2u * (uint32_t)sc.fres[1] +
                      nx * ny * 2u * (uint32_t)sc.fres[2];
      cubic_nodes(sc.fres, cell[0], cell[1], cell[2], sc.quirk, idx);
#pragma unroll
      for (int k = 0; k < 32; ++k) v[k] = F[flen + idx[k]];
      vol = sum32(N, v);
    }
  } else {
    int c[3];
    float f[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
      float u = (xv[a] - sc.flo[a]) * sc.finv[a];
      float R = (float)sc.fres[a];
      u = u > 0.0f ? u : 0.0f;
      u = u < R ? u : R;
      float fl = floorf(u);
      int ci = (int)fl;
      if (ci > sc.fres[a] - 1) { ci = sc.fres[a] - 1; fl = (float)ci; }
      c[a] = ci;
      f[a] = u - fl;
    }
    uint32_t nx = (uint32_t)sc.fres[0] + 1, ny = (uint32_t)sc.fres[1] + 1;
    uint32_t b000 = ((uint32_t)c[2] * ny + (uint32_t)c[1]) * nx + (uint32_t)c[0];
    const float2 *SV = reinterpret_cast<const float2 *>(F);
    uint32_t o[8] = {b000, b000 + 1, b000 + nx, b000 + nx + 1,
                     b000 + nx * ny, b000 + nx * ny + 1, b000 + nx * ny + nx, b000 + nx * ny + nx + 1};
    float cs[8], cv[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      float2 t = SV[o[k]];
      cs[k] = t.x;
      cv[k] = t.y;
    }
    float fx = f[0], fy = f[1], fz = f[2];
    float c00 = fmaf(fx, cs[1] - cs[0], cs[0]);
    float c10 = fmaf(fx, cs[3] - cs[2], cs[2]);
    float c01 = fmaf(fx, cs[5] - cs[4], cs[4]);
    float c11 = fmaf(fx, cs[7] - cs[6], cs[6]);
    float c0 = fmaf(fy, c10 - c00, c00);
    float c1 = fmaf(fy, c11 - c01, c01);
    sdf = fmaf(fz, c1 - c0, c0);
    float ex00 = cs[1] - cs[0], ex10 = cs[3] - cs[2], ex01 = cs[5] - cs[4], ex11 = cs[7] - cs[6];
    float ex0 = fmaf(fy, ex10 - ex00, ex00);
    float ex1 = fmaf(fy, ex11 - ex01, ex01);
    g[0] = fmaf(fz, ex1 - ex0, ex0) * sc.finv[0];
    float ey0 = c10 - c00, ey1 = c11 - c01;
    g[1] = fmaf(fz, ey1 - ey0, ey0) * sc.finv[1];
    g[2] = (c1 - c0) * sc.finv[2];
    float v00 = fmaf(fx, cv[1] - cv[0], cv[0]);
    float v10 = fmaf(fx, cv[3] - cv[2], cv[2]);
    float v01 = fmaf(fx, cv[5] - cv[4], cv[4]);
    float v11 = fmaf(fx, cv[7] - cv[6], cv[6]);
    float v0 = fmaf(fy, v10 - v00, v00);
    float v1 = fmaf(fy, v11 - v01, v01);
    vol = fmaf(fz, v1 - v0, v0);
  }
}

// Mirrored boundary particle (x_b, V_b) of a fluid particle at x: the field
// point max(d + r/2, 2r) below x along the SDF normal, when 0 < d < h.
template <int KIND>
__device__ __forceinline__ float4 boundary_sample(const Phys &P, const Scene &sc, const float *__restrict__ pool,
                                                  float x0, float x1, float x2) {
  float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (!sc.field_kind) return out;
  float g[3], d, vol = 0.0f;
  field_eval<KIND>(sc, pool, x0, x1, x2, g, d, vol, false);
  float len = sqrtf(dist2(g[0], g[1], g[2]));
  if (d > 0.0f && d < P.h && len > 1e-9f) {
    if (KIND == 1 || (KIND == 0 && sc.field_kind == 1)) field_eval<KIND>(sc, pool, x0, x1, x2, g, d, vol, true);
    if (vol > 0.0f) {
      float nx = g[0] / len, ny = g[1] / len, nz = g[2] / len;
      float a = d + P.half_r;
      float dd = a > P.two_r ? a : P.two_r;
      out.x = fmaf(-dd, nx, x0);
      out.y = fmaf(-dd, ny, x1);
      out.z = fmaf(-dd, nz, x2);
      out.w = vol;
    }
  }
  return out;
}

// ---------------------------------------------------------------------------
// Launch helpers.

// Workgroups are dispatched round-robin over the 8 XCDs; remap so that each
// XCD processes one contiguous run of particle blocks (its own L2 then holds
// that spatially compact slice of the cell-sorted arrays).
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
  const uint32_t NX = 8;
  uint32_t xcd = b % NX, local = b / NX;
  uint32_t per = nb / NX, rem = nb % NX;
  return xcd < rem ? xcd * (per + 1) + local : rem * (per + 1) + (xcd - rem) * per + local;
}

}  // namespace pbf
