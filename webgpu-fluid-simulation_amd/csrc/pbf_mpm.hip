// pbf_mpm.hip -- MLS-MPM step behind the pbf_mpm_* entry points of
// include/pbf.h (reference: src/simulator/MPM/MPMShader.ts, the P2G / grid /
// G2P shaders, and MPM.ts, their commented-out driver).
//
// The reference scatters every particle into its 27 grid nodes with float
// compare-and-swap loops (MPMShader.ts:8-56), so a node's sum depends on the
// order the atomics land in.  Here P2G is a gather: particles are binned by
// their base cell and stably sorted by id (the binning and scan of the PBF
// neighbour search), then one thread per grid node sums the particles of
// the 27 base cells that reach it in a fixed order (nine column partials,
// each over its cells and their particles in ascending (cell, id) order,
// added in column order), and applies the grid pass (MPMShader.ts:163-191).
// No atomics touch the grid, every step is deterministic, and the result is
// bit-identical to the CPU restatement that sums in the same order.
//
// Per step: prepare (J update, stress, affine matrix, base cell and binning
// rank; the per-particle P2G record written once), scan, scatter, settle
// (records moved into cell order), node gather + grid pass, G2P.
//
// The particle state (x, v, C, and ids: slot -> particle id) is kept in the
// cell order of the previous step, as the PBF engine keeps its particles:
// prepare then reads it in order and most of a wavefront's lanes share a
// cell (one atomic per cell and wavefront for the binning rank), and G2P
// writes the new state to its own slot.  Readback and upload go through ids.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <exception>
#include <new>
#include <string>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_device.h"
#include "pbf_internal.h"

using namespace pbf;

namespace {

constexpr uint32_t MBS = 256;
constexpr uint32_t kRec = 4;  // float4s per P2G record

struct MpmK {
  float dx, dt, p_vol, p_mass, E, inv4_dx2;
  float g[4];
  uint32_t ng, bound, G;
};

__device__ __forceinline__ uint32_t f2u_sat(float f) { return f > 0.0f ? (uint32_t)f : 0u; }

// quadratic B-spline weights of the shader (MPMShader.ts:106-110)
__device__ __forceinline__ void bspline(float f, float &w0, float &w1, float &w2) {
  const float t0 = 1.5f - f, t1 = f - 1.0f, t2 = f - 0.5f;
  w0 = (0.5f * t0) * t0;
  w1 = 0.75f - t1 * t1;
  w2 = (0.5f * t2) * t2;
}

__device__ __forceinline__ float pick3(int k, float a, float b, float c) { return k == 0 ? a : k == 1 ? b : c; }

// Prepare: J (MPMShader.ts:101), base cell and fractional offset
// (:103-105), stress and affine matrix (:112-113), p_mass v; the record
// (fx, p_mass v, affine columns, and J for G2P) of slot p and its binning
// rank, one atomic per distinct cell of the wavefront.
__global__ void __launch_bounds__(MBS) k_mpm_prepare(uint32_t n, const float4 *__restrict__ x,
                                                     const float4 *__restrict__ v, const float4 *__restrict__ C0,
                                                     const float4 *__restrict__ C1, const float4 *__restrict__ C2,
                                                     float4 *__restrict__ rec, uint32_t *__restrict__ cellkey,
                                                     uint32_t *__restrict__ rank, uint32_t *__restrict__ counts,
                                                     MpmK K) {
  const uint32_t p = blockIdx.x * MBS + threadIdx.x;
  uint32_t key = 0;
  if (p < n) {
    const float4 xp = x[p], vp = v[p], c0 = C0[p], c1 = C1[p], c2 = C2[p];
    const float J = vp.w * (1.0f + K.dt * ((c0.x + c1.y) + c2.z));
    const float X0 = xp.x / K.dx, X1 = xp.y / K.dx, X2 = xp.z / K.dx;
    const uint32_t b0 = f2u_sat(X0 - 0.5f), b1 = f2u_sat(X1 - 0.5f), b2 = f2u_sat(X2 - 0.5f);
    const float f0 = X0 - (float)b0, f1 = X1 - (float)b1, f2 = X2 - (float)b2;
    const float stress = (((((-K.dt) * 4.0f) * K.E) * K.p_vol) * (J - 1.0f)) / (K.dx * K.dx);
    const float z = 0.0f * stress;  // off-diagonal of stress * identity
    const float m = K.p_mass;
    // record (64 bytes): fx, p_mass v, the affine matrix's three columns
    const size_t s = (size_t)p * kRec;
    const float a00 = stress + m * c0.x, a01 = z + m * c0.y, a02 = z + m * c0.z;
    const float a10 = z + m * c1.x, a11 = stress + m * c1.y, a12 = z + m * c1.z;
    const float a20 = z + m * c2.x, a21 = z + m * c2.y, a22 = stress + m * c2.z;
    rec[s] = make_float4(f0, f1, f2, m * vp.x);
    rec[s + 1] = make_float4(m * vp.y, m * vp.z, a00, a01);
    rec[s + 2] = make_float4(a02, a10, a11, a12);
    rec[s + 3] = make_float4(a20, a21, a22, J);  // w: J, which P2G does not read
    key = (b0 < K.ng && b1 < K.ng && b2 < K.ng) ? (b0 * K.ng + b1) * K.ng + b2 : K.G;  // K.G: outside
    cellkey[p] = key;
  }
  // the settle orders each cell by id, so the order of the ranks is free
  const uint32_t r = wave_atomic_rank(counts, key, p < n);
  if (p < n) rank[p] = r;
}

// Slot p's particle id and slot to its binned position.
__global__ void __launch_bounds__(MBS) k_mpm_scatter(uint32_t n, const uint32_t *__restrict__ cellkey,
                                                     const uint32_t *__restrict__ rank,
                                                     const uint32_t *__restrict__ offsets,
                                                     const uint32_t *__restrict__ ids, uint32_t *__restrict__ tmp_id,
                                                     uint32_t *__restrict__ tmp_slot) {
  const uint32_t p = blockIdx.x * MBS + threadIdx.x;
  if (p >= n) return;
  const uint32_t t = offsets[cellkey[p]] + rank[p];
  tmp_id[t] = ids[p];
  tmp_slot[t] = p;
}

// Records into cell order, ascending id inside a cell (the bin of
// particles outside the grid is never read: left in scatter order), with
// each particle's position (xs) and id (ids_new) at its new slot.
__global__ void __launch_bounds__(MBS) k_mpm_settle(uint32_t n, const uint32_t *__restrict__ tmp_id,
                                                    const uint32_t *__restrict__ tmp_slot,
                                                    const uint32_t *__restrict__ cellkey,
                                                    const uint32_t *__restrict__ offsets,
                                                    const float4 *__restrict__ rec, const float4 *__restrict__ x,
                                                    float4 *__restrict__ srec, float4 *__restrict__ xs,
                                                    uint32_t *__restrict__ ids_new, uint32_t outside) {
  const uint32_t s = blockIdx.x * MBS + threadIdx.x;
  if (s >= n) return;
  const uint32_t id = tmp_id[s], q = tmp_slot[s], c = cellkey[q];
  uint32_t slot = s;
  if (c != outside) {
    const uint32_t o = offsets[c], e = offsets[c + 1];
    uint32_t r = 0;
    for (uint32_t t = o; t < e; ++t) r += tmp_id[t] < id ? 1u : 0u;
    slot = o + r;
  }
  const size_t f = (size_t)slot * kRec, g = (size_t)q * kRec;
#pragma unroll
  for (int k = 0; k < (int)kRec; ++k) srec[f + k] = rec[g + k];
  xs[slot] = x[q];
  ids_new[slot] = id;
}

// P2G as a gather (MPMShader.ts:4-62 summed per node) and the grid pass
// (MPMShader.ts:163-191).  A work-group of nine wavefronts takes 64 RUN
// consecutive nodes (z fastest); wavefront c sums base-cell column
// c = (bx, by) for all of them, lane l for the run of RUN nodes from
// K0 = RUN l (mod n_grid) in one node column.  The lane walks the cells
// bz = K0-2 .. K0+RUN-1 of its column once, each cell's records in slot
// order (four loads in flight), and adds every record into the partials of
// the nodes it reaches (bz .. bz+2 inside the run) -- a record is loaded
// once for up to three nodes and a node's partial still runs over its cells
// K-2..K in slot order.  Then each node's nine partials are added in column
// order and the grid pass applied.
// RUN nodes per lane (consecutive z): 4 on large grids (a record loaded once
// for up to three nodes; at 1M particles on 128^3 nodes runs of 2, 4 and 8
// take 395, 311 and 513 us), 1 on small ones (more lanes for few particles:
// 17 vs 29 us at the reference's 9,000 particles on 32^3 nodes).
constexpr uint32_t kP2GBS = 9 * 64;  // one wavefront per column
constexpr uint32_t kSmallGrid = 1u << 16;

template <int RUN>
__device__ __forceinline__ void p2g_run(const uint32_t *__restrict__ offsets, const float4 *__restrict__ srec,
                                        const MpmK &K, uint32_t c0, int oi, int oj, int K0, float4 acc[RUN]) {
#pragma unroll
  for (int c = 0; c < RUN + 2; ++c) {  // cell bz = K0 - 2 + c
    const int bz = K0 - 2 + c;
    if (bz < 0 || bz >= (int)K.ng) continue;
    const uint32_t t0 = offsets[c0 + (uint32_t)bz], t1 = offsets[c0 + (uint32_t)bz + 1];
    auto add = [&](const float4 &r0, const float4 &r1, const float4 &r2, const float4 &r3) {
      float wx0, wx1, wx2, wy0, wy1, wy2, wz[3];
      bspline(r0.x, wx0, wx1, wx2);
      bspline(r0.y, wy0, wy1, wy2);
      bspline(r0.z, wz[0], wz[1], wz[2]);
      const float wxy = pick3(oi, wx0, wx1, wx2) * pick3(oj, wy0, wy1, wy2);
      const float dpx = ((float)oi - r0.x) * K.dx, dpy = ((float)oj - r0.y) * K.dx;
#pragma unroll
      for (int ok = 0; ok < 3; ++ok) {
        const int node = c - 2 + ok;  // index of the target node in the run
        if (node < 0 || node >= RUN) continue;
        const float dpz = ((float)ok - r0.z) * K.dx;
        const float weight = wxy * wz[ok];
        // columns A0 = (r1.z, r1.w, r2.x), A1 = (r2.y, r2.z, r2.w), A2 = (r3.x, r3.y, r3.z)
        const float ad0 = (r1.z * dpx + r2.y * dpy) + r3.x * dpz;
        const float ad1 = (r1.w * dpx + r2.z * dpy) + r3.y * dpz;
        const float ad2 = (r2.x * dpx + r2.w * dpy) + r3.z * dpz;
        acc[node].x += weight * (r0.w + ad0);
        acc[node].y += weight * (r1.x + ad1);
        acc[node].z += weight * (r1.y + ad2);
        acc[node].w += weight * K.p_mass;
      }
    };
    uint32_t t = t0;
    for (; t + 4 <= t1; t += 4) {
      float4 r[4][kRec];
#pragma unroll
      for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int k = 0; k < (int)kRec; ++k) r[q][k] = srec[(size_t)(t + q) * kRec + k];
#pragma unroll
      for (int q = 0; q < 4; ++q) add(r[q][0], r[q][1], r[q][2], r[q][3]);
    }
    for (; t < t1; ++t) {
      const size_t b = (size_t)t * kRec;
      add(srec[b], srec[b + 1], srec[b + 2], srec[b + 3]);
    }
  }
}

template <int RUN>
__global__ void __launch_bounds__(kP2GBS) k_mpm_p2g_grid(const uint32_t *__restrict__ offsets,
                                                         const float4 *__restrict__ srec, float4 *__restrict__ grid,
                                                         MpmK K) {
  constexpr uint32_t kP2GNodes = 64u * RUN;  // nodes per work-group
  __shared__ float4 part[9 * kP2GNodes];
  const uint32_t tid = threadIdx.x, lane = tid % 64u, col = tid / 64u;
  // plain order: the work-groups in flight cover one contiguous stretch of
  // nodes (an XCD-contiguous mapping, whole or in runs of 16 work-groups,
  // measured 1.7-1.8x slower: the particles fill only part of the grid)
  const uint32_t base = blockIdx.x * kP2GNodes;
  // Work-group-uniform: does any particle lie in the cells its nodes gather
  // from?  With all nodes in one x-layer I, those are the cells (bx, by, *)
  // with bx in I-2..I and by in J_first-2..J_last, three contiguous cell
  // ranges.  Where none holds a particle every partial is +0, so the walks
  // and the LDS exchange are skipped and the nodes get the grid pass of a
  // zero sum (most work-groups: the particles fill a small part of the grid).
  bool any = true;
  {
    const uint32_t first = base, last = (base + kP2GNodes < K.G ? base + kP2GNodes : K.G) - 1u;
    const uint32_t ng2 = K.ng * K.ng, I0 = first / ng2;
    if (I0 == last / ng2) {
      const uint32_t J0 = (first / K.ng) % K.ng, J1 = (last / K.ng) % K.ng, jlo = J0 >= 2u ? J0 - 2u : 0u;
      any = false;
      for (uint32_t bx = I0 >= 2u ? I0 - 2u : 0u; bx <= I0; ++bx)
        any = any || offsets[(bx * K.ng + J1) * K.ng + K.ng] != offsets[(bx * K.ng + jlo) * K.ng];
    }
  }
  if (any) {
    // this lane's run: nodes base + RUN * lane .. + RUN - 1 (one node
    // column: ng is a multiple of RUN)
    const uint32_t n0 = base + RUN * lane;
    float4 acc[RUN];
#pragma unroll
    for (int k = 0; k < RUN; ++k) acc[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
    if (n0 < K.G) {
      const int I = (int)(n0 / (K.ng * K.ng)), J = (int)((n0 / K.ng) % K.ng), K0 = (int)(n0 % K.ng);
      // column col: bx = I - 2 + col / 3, by = J - 2 + col % 3 (ascending linear index)
      const int oi = 2 - (int)(col / 3u), oj = 2 - (int)(col % 3u);
      const int bx = I - oi, by = J - oj;
      if (bx >= 0 && by >= 0)
        p2g_run<RUN>(offsets, srec, K, ((uint32_t)bx * K.ng + (uint32_t)by) * K.ng, oi, oj, K0, acc);
    }
#pragma unroll
    for (int k = 0; k < RUN; ++k) part[col * kP2GNodes + RUN * lane + k] = acc[k];
    __syncthreads();
  }
  if (tid >= kP2GNodes) return;
  const uint32_t node = base + tid;
  if (node >= K.G) return;
  const int I = (int)(node / (K.ng * K.ng)), J = (int)((node / K.ng) % K.ng), Kz = (int)(node % K.ng);
  float4 a = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  if (any) {
    a = part[tid];
#pragma unroll
    for (int c = 1; c < 9; ++c) {
      const float4 q = part[c * kP2GNodes + tid];
      a.x = a.x + q.x;
      a.y = a.y + q.y;
      a.z = a.z + q.z;
      a.w = a.w + q.w;
    }
  }
  float v0 = 0.0f, v1 = 0.0f, v2 = 0.0f, v3 = 0.0f;
  if (a.w > 0.0f) {
    v0 = a.x / a.w;
    v1 = a.y / a.w;
    v2 = a.z / a.w;
    v3 = a.w / a.w;
  }
  v0 = v0 + K.dt * K.g[0];
  v1 = v1 + K.dt * K.g[1];
  v2 = v2 + K.dt * K.g[2];
  v3 = v3 + K.dt * K.g[3];
  const uint32_t hi = K.ng - K.bound;
  if (((uint32_t)I < K.bound && v0 < 0.0f) || ((uint32_t)I >= hi && v0 > 0.0f)) v0 = 0.0f;
  if (((uint32_t)J < K.bound && v1 < 0.0f) || ((uint32_t)J >= hi && v1 > 0.0f)) v1 = 0.0f;
  if (((uint32_t)Kz < K.bound && v2 < 0.0f) || ((uint32_t)Kz >= hi && v2 > 0.0f)) v2 = 0.0f;
  grid[node] = make_float4(v0, v1, v2, v3);
}

// G2P (MPMShader.ts:240-280): 27 nodes in the shader's order (i outer, k
// inner); nodes outside the grid read as zero.  STEP: slot s of the step's
// cell order takes its position from xs and J from its record (both moved
// there by the settle) and writes the new state, and its id, to slot s;
// otherwise (a G2P on its own) slot s reads and writes its own state.  In
// cell order a wavefront's node gathers fall on neighbouring nodes.
template <bool STEP>
__global__ void __launch_bounds__(MBS) k_mpm_g2p(uint32_t n, float4 *__restrict__ x, float4 *__restrict__ v,
                                                 float4 *__restrict__ C0, float4 *__restrict__ C1,
                                                 float4 *__restrict__ C2, const float4 *__restrict__ grid,
                                                 const float4 *__restrict__ xs, const float4 *__restrict__ srec,
                                                 const uint32_t *__restrict__ ids_new, uint32_t *__restrict__ ids,
                                                 MpmK K) {
  const uint32_t s = blockIdx.x * MBS + threadIdx.x;
  if (s >= n) return;
  const float4 xp = STEP ? xs[s] : x[s];
  const float X[3] = {xp.x / K.dx, xp.y / K.dx, xp.z / K.dx};
  uint32_t b[3];
  float f[3], w[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    b[a] = f2u_sat(X[a] - 0.5f);
    f[a] = X[a] - (float)b[a];
    bspline(f[a], w[0][a], w[1][a], w[2][a]);
  }
  float nv[3] = {0.0f, 0.0f, 0.0f}, nC[9] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int k = 0; k < 3; ++k) {
        const float dpos[3] = {((float)i - f[0]) * K.dx, ((float)j - f[1]) * K.dx, ((float)k - f[2]) * K.dx};
        const float weight = (w[i][0] * w[j][1]) * w[k][2];
        const uint32_t gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
        float vel[3] = {0.0f, 0.0f, 0.0f};
        if (gi < K.ng && gj < K.ng && gk < K.ng) {
          const float4 g = grid[(gi * K.ng + gj) * K.ng + gk];
          vel[0] = g.x;
          vel[1] = g.y;
          vel[2] = g.z;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) nv[a] = nv[a] + weight * vel[a];
        const float s = K.inv4_dx2 * weight;
#pragma unroll
        for (int c = 0; c < 3; ++c)
#pragma unroll
          for (int r = 0; r < 3; ++r) nC[3 * c + r] = nC[3 * c + r] + s * (vel[c] * dpos[r]);
      }
  const float J = STEP ? srec[(size_t)s * kRec + 3].w : v[s].w;
  v[s] = make_float4(nv[0], nv[1], nv[2], J);
  C0[s] = make_float4(nC[0], nC[1], nC[2], 0.0f);
  C1[s] = make_float4(nC[3], nC[4], nC[5], 0.0f);
  C2[s] = make_float4(nC[6], nC[7], nC[8], 0.0f);
  x[s] = make_float4(xp.x + K.dt * nv[0], xp.y + K.dt * nv[1], xp.z + K.dt * nv[2], xp.w);
  if (STEP) ids[s] = ids_new[s];
}

__global__ void __launch_bounds__(MBS) k_mpm_init(uint32_t n, const float4 *__restrict__ in, float4 *__restrict__ x,
                                                  float4 *__restrict__ v, float4 *__restrict__ C0,
                                                  float4 *__restrict__ C1, float4 *__restrict__ C2,
                                                  uint32_t *__restrict__ ids) {
  const uint32_t p = blockIdx.x * MBS + threadIdx.x;
  if (p >= n) return;
  const float4 a = in[p];
  x[p] = make_float4(a.x, a.y, a.z, 0.0f);
  v[p] = make_float4(0.0f, 0.0f, 0.0f, 1.0f);  // J = 1 (MPM.ts:115-119)
  C0[p] = C1[p] = C2[p] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  ids[p] = p;  // caller order
}

inline uint32_t nblk(uint64_t n) { return (uint32_t)((n + MBS - 1) / MBS); }

template <typename T>
hipError_t dalloc(T **p, size_t count) {
  return hipMalloc(reinterpret_cast<void **>(p), (count ? count : 1) * sizeof(T));
}

}  // namespace

struct pbf_mpm {
  pbf_mpm_params params;
  std::string err;
  int device = 0;
  hipStream_t stream = nullptr;
  uint32_t n = 0, cap = 0, ng = 0, G = 0;
  MpmK K;
  float4 *x = nullptr, *v = nullptr, *C0 = nullptr, *C1 = nullptr, *C2 = nullptr, *grid = nullptr;
  float4 *rec = nullptr, *srec = nullptr, *stage = nullptr;
  uint32_t *cellkey = nullptr, *rank = nullptr, *counts = nullptr, *offsets = nullptr, *tmp_id = nullptr;
  uint32_t *tmp_slot = nullptr;
  float4 *xs = nullptr;                           // positions moved into the step's cell order
  uint32_t *ids = nullptr, *ids_new = nullptr;  // slot -> particle id (state order; the step's new order)
  std::vector<uint32_t *> scan_sums, scan_prefix;
  hipGraphExec_t graph = nullptr;
  uint32_t graph_n = 0;
};

namespace {

int mfail(pbf_mpm *m, int code, const std::string &msg) {
  if (m) m->err = msg;
  set_thread_error(msg);
  return code;
}

#define MCHK(expr)                                                                     \
  do {                                                                                 \
    hipError_t e_ = (expr);                                                            \
    if (e_ != hipSuccess) return mfail(m, PBF_E_DEVICE, std::string(#expr ": ") + hipGetErrorString(e_)); \
  } while (0)

void enqueue_step(pbf_mpm *m) {
  const uint32_t n = m->n;
  hipStream_t st = m->stream;
  hipLaunchKernelGGL(k_mpm_prepare, dim3(nblk(n)), dim3(MBS), 0, st, n, m->x, m->v, m->C0, m->C1, m->C2, m->rec,
                     m->cellkey, m->rank, m->counts, m->K);
  // counts over G + 1 bins (the last: particles whose base cell is outside
  // the grid, which the reference's bounds test skips); zeroed by the scan
  launch_scan(st, m->counts, m->G + 1, m->offsets, 1, m->scan_sums.data(), m->scan_prefix.data(), 0);
  hipLaunchKernelGGL(k_mpm_scatter, dim3(nblk(n)), dim3(MBS), 0, st, n, m->cellkey, m->rank, m->offsets, m->ids,
                     m->tmp_id, m->tmp_slot);
  hipLaunchKernelGGL(k_mpm_settle, dim3(nblk(n)), dim3(MBS), 0, st, n, m->tmp_id, m->tmp_slot, m->cellkey,
                     m->offsets, m->rec, m->x, m->srec, m->xs, m->ids_new, m->G);
  if (m->G <= kSmallGrid)
    hipLaunchKernelGGL(k_mpm_p2g_grid<1>, dim3((m->G + 63u) / 64u), dim3(kP2GBS), 0, st, m->offsets, m->srec,
                       m->grid, m->K);
  else
    hipLaunchKernelGGL(k_mpm_p2g_grid<4>, dim3((m->G + 255u) / 256u), dim3(kP2GBS), 0, st, m->offsets, m->srec,
                       m->grid, m->K);
  hipLaunchKernelGGL(k_mpm_g2p<true>, dim3(nblk(n)), dim3(MBS), 0, st, n, m->x, m->v, m->C0, m->C1, m->C2, m->grid,
                     (const float4 *)m->xs, (const float4 *)m->srec, (const uint32_t *)m->ids_new, m->ids, m->K);
}

void drop_graph(pbf_mpm *m) {
  if (m->graph) {
    (void)hipStreamSynchronize(m->stream);
    (void)hipGraphExecDestroy(m->graph);
    m->graph = nullptr;
  }
}

}  // namespace

#define MPM_CATCH(m)                                                         \
  catch (const std::bad_alloc &) {                                           \
    return mfail(m, PBF_E_CAPACITY, "out of host memory");                   \
  }                                                                          \
  catch (const std::exception &ex) {                                         \
    return mfail(m, PBF_E_INVALID, ex.what());                               \
  }

extern "C" {

int pbf_mpm_create(const pbf_mpm_params *params, pbf_mpm **out) try {
  if (!params || !out) return mfail(nullptr, PBF_E_INVALID, "params and out are required");
  *out = nullptr;
  pbf_mpm_params p = *params;
  if (p.n_grid == 0) p.n_grid = 32;
  if (p.bound == 0) p.bound = 3;
  if (!(p.dt > 0.0)) p.dt = 8e-4;
  if (!(p.rho > 0.0)) p.rho = 1.0;
  if (!(p.E > 0.0)) p.E = 400.0;
  if (p.capacity == 0) p.capacity = 9000;
  if (!std::isfinite(p.dt) || !std::isfinite(p.rho) || !std::isfinite(p.E) || !std::isfinite(p.gravity))
    return mfail(nullptr, PBF_E_INVALID, "dt, rho, E and gravity must be finite");
  if (p.n_grid < 4 || p.n_grid > 1024 || p.n_grid % 4 != 0)
    return mfail(nullptr, PBF_E_INVALID, "n_grid must be a multiple of 4 in [4, 1024]");
  if (2 * p.bound > p.n_grid) return mfail(nullptr, PBF_E_INVALID, "bound must be at most n_grid / 2");
  if (p.capacity > (1u << 27)) return mfail(nullptr, PBF_E_CAPACITY, "capacity above 2^27 particles");
  pbf_mpm *m = new pbf_mpm();
  m->params = p;
  m->cap = p.capacity;
  m->ng = p.n_grid;
  m->G = p.n_grid * p.n_grid * p.n_grid;
  const double dx = 1.0 / (double)p.n_grid;
  MpmK &K = m->K;
  K.dx = (float)dx;
  K.dt = (float)p.dt;
  K.p_vol = (float)((0.5 * dx) * (0.5 * dx));  // MPM.ts:77-78
  K.p_mass = (float)(p.rho * ((0.5 * dx) * (0.5 * dx)));
  K.E = (float)p.E;
  K.inv4_dx2 = 4.0f / (K.dx * K.dx);
  K.g[0] = 0.0f;
  K.g[1] = -(float)p.gravity;
  K.g[2] = 0.0f;
  K.g[3] = 0.0f;
  K.ng = p.n_grid;
  K.bound = p.bound;
  K.G = m->G;
  int rc = PBF_OK;
  auto step = [&](hipError_t e, const char *what) {
    if (rc == PBF_OK && e != hipSuccess) rc = mfail(m, PBF_E_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
  };
  step(hipGetDevice(&m->device), "hipGetDevice");
  step(hipStreamCreateWithFlags(&m->stream, hipStreamNonBlocking), "hipStreamCreate");
  const size_t N = m->cap;
  step(dalloc(&m->x, N), "alloc"); step(dalloc(&m->v, N), "alloc");
  step(dalloc(&m->C0, N), "alloc"); step(dalloc(&m->C1, N), "alloc"); step(dalloc(&m->C2, N), "alloc");
  step(dalloc(&m->rec, kRec * N), "alloc"); step(dalloc(&m->srec, kRec * N), "alloc");
  step(dalloc(&m->stage, N), "alloc");
  step(dalloc(&m->cellkey, N), "alloc"); step(dalloc(&m->rank, N), "alloc"); step(dalloc(&m->tmp_id, N), "alloc");
  step(dalloc(&m->tmp_slot, N), "alloc"); step(dalloc(&m->xs, N), "alloc");
  step(dalloc(&m->ids, N), "alloc"); step(dalloc(&m->ids_new, N), "alloc");
  step(dalloc(&m->grid, (size_t)m->G), "alloc");
  step(dalloc(&m->counts, (size_t)m->G + 1), "alloc");
  step(dalloc(&m->offsets, (size_t)m->G + 2), "alloc");
  {
    uint32_t len = m->G + 1;
    while (rc == PBF_OK && len > 2048u) {
      const uint32_t nb = (len + 2047u) / 2048u;
      uint32_t *s = nullptr, *q = nullptr;
      step(dalloc(&s, nb), "alloc");
      step(dalloc(&q, (size_t)nb + 1), "alloc");
      m->scan_sums.push_back(s);
      m->scan_prefix.push_back(q);
      len = nb;
    }
    m->scan_sums.push_back(nullptr);
    m->scan_prefix.push_back(nullptr);
  }
  if (rc == PBF_OK) step(hipMemset(m->counts, 0, ((size_t)m->G + 1) * 4), "hipMemset");
  if (rc == PBF_OK) step(hipMemset(m->grid, 0, (size_t)m->G * 16), "hipMemset");
  if (rc != PBF_OK) {
    const std::string msg = m->err;
    pbf_mpm_destroy(m);
    set_thread_error(msg);
    return rc;
  }
  *out = m;
  return PBF_OK;
}
MPM_CATCH(nullptr)

int pbf_mpm_destroy(pbf_mpm *m) {
  if (!m) return PBF_OK;
  (void)hipSetDevice(m->device);
  if (m->stream) (void)hipStreamSynchronize(m->stream);
  drop_graph(m);
  void *bufs[] = {m->x, m->v, m->C0, m->C1, m->C2, m->grid, m->rec, m->srec, m->stage, m->cellkey, m->rank,
                  m->counts, m->offsets, m->tmp_id, m->tmp_slot, m->xs, m->ids, m->ids_new};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  for (auto q : m->scan_sums)
    if (q) (void)hipFree(q);
  for (auto q : m->scan_prefix)
    if (q) (void)hipFree(q);
  if (m->stream) (void)hipStreamDestroy(m->stream);
  delete m;
  return PBF_OK;
}

int pbf_mpm_set_positions(pbf_mpm *m, const float *xyzw, uint32_t count) try {
  if (!m || (!xyzw && count)) return mfail(m, PBF_E_INVALID, "context and positions are required");
  if (count > m->cap) return mfail(m, PBF_E_CAPACITY, "To Many Particles!");
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  drop_graph(m);
  m->n = count;
  if (count) {
    MCHK(hipMemcpy(m->stage, xyzw, (size_t)count * 16, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_mpm_init, dim3(nblk(count)), dim3(MBS), 0, m->stream, count, m->stage, m->x, m->v, m->C0,
                       m->C1, m->C2, m->ids);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(m->stream));
  }
  return PBF_OK;
}
MPM_CATCH(m)

int pbf_mpm_step(pbf_mpm *m, uint32_t steps) try {
  if (!m) return mfail(m, PBF_E_INVALID, "null context");
  if (m->n == 0 || steps == 0) return PBF_OK;
  MCHK(hipSetDevice(m->device));
  if (m->params.flags & PBF_FLAG_NO_GRAPH) {
    for (uint32_t s = 0; s < steps; ++s) enqueue_step(m);
    MCHK(hipGetLastError());
    return PBF_OK;
  }
  if (!m->graph || m->graph_n != m->n) {
    drop_graph(m);
    hipGraph_t g = nullptr;
    MCHK(hipStreamBeginCapture(m->stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(m);
    const hipError_t ec = hipStreamEndCapture(m->stream, &g);
    if (ec != hipSuccess) return mfail(m, PBF_E_DEVICE, std::string("hipStreamEndCapture: ") + hipGetErrorString(ec));
    const hipError_t ei = hipGraphInstantiate(&m->graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
      m->graph = nullptr;
      return mfail(m, PBF_E_DEVICE, std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
    }
    m->graph_n = m->n;
  }
  for (uint32_t s = 0; s < steps; ++s) MCHK(hipGraphLaunch(m->graph, m->stream));
  return PBF_OK;
}
MPM_CATCH(m)

int pbf_mpm_sync(pbf_mpm *m) try {
  if (!m) return mfail(m, PBF_E_INVALID, "null context");
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  return PBF_OK;
}
MPM_CATCH(m)

int pbf_mpm_read(pbf_mpm *m, int field, float *out, size_t count) try {
  if (!m || !out) return mfail(m, PBF_E_INVALID, "context and output are required");
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  const size_t n = m->n;
  if (field == PBF_MPM_GRID) {
    if (count != m->G) return mfail(m, PBF_E_LENGTH, "grid readback count must be n_grid^3");
    MCHK(hipMemcpy(out, m->grid, (size_t)m->G * 16, hipMemcpyDeviceToHost));
    return PBF_OK;
  }
  if (count != n) return mfail(m, PBF_E_LENGTH, "readback count must equal the particle count");
  if (!n) return PBF_OK;
  if (field != PBF_MPM_POSITION && field != PBF_MPM_VELOCITY && field != PBF_MPM_AFFINE)
    return mfail(m, PBF_E_INVALID, "unknown MPM field");
  // the state is in slot order: slot s holds particle ids[s]
  std::vector<uint32_t> ids(n);
  MCHK(hipMemcpy(ids.data(), m->ids, n * 4, hipMemcpyDeviceToHost));
  std::vector<float> c(n * 4);
  if (field == PBF_MPM_AFFINE) {
    float4 *cols[3] = {m->C0, m->C1, m->C2};
    for (int k = 0; k < 3; ++k) {
      MCHK(hipMemcpy(c.data(), cols[k], n * 16, hipMemcpyDeviceToHost));
      for (size_t s = 0; s < n; ++s)
        for (int r = 0; r < 3; ++r) out[9 * (size_t)ids[s] + 3 * k + r] = c[4 * s + r];
    }
  } else {
    MCHK(hipMemcpy(c.data(), field == PBF_MPM_POSITION ? m->x : m->v, n * 16, hipMemcpyDeviceToHost));
    for (size_t s = 0; s < n; ++s) std::memcpy(out + 4 * (size_t)ids[s], c.data() + 4 * s, 16);
  }
  return PBF_OK;
}
MPM_CATCH(m)

int pbf_mpm_write(pbf_mpm *m, int field, const float *in, size_t count) try {
  if (!m || !in) return mfail(m, PBF_E_INVALID, "context and input are required");
  MCHK(hipSetDevice(m->device));
  MCHK(hipStreamSynchronize(m->stream));
  const size_t n = m->n;
  if (field == PBF_MPM_GRID) {
    if (count != m->G) return mfail(m, PBF_E_LENGTH, "grid count must be n_grid^3");
    MCHK(hipMemcpy(m->grid, in, (size_t)m->G * 16, hipMemcpyHostToDevice));
    return PBF_OK;
  }
  if (count != n) return mfail(m, PBF_E_LENGTH, "count must equal the particle count");
  if (!n) return PBF_OK;
  if (field != PBF_MPM_POSITION && field != PBF_MPM_VELOCITY && field != PBF_MPM_AFFINE)
    return mfail(m, PBF_E_INVALID, "unknown MPM field");
  // into slot order: slot s takes particle ids[s]
  std::vector<uint32_t> ids(n);
  MCHK(hipMemcpy(ids.data(), m->ids, n * 4, hipMemcpyDeviceToHost));
  std::vector<float> c(n * 4, 0.0f);
  if (field == PBF_MPM_AFFINE) {
    float4 *cols[3] = {m->C0, m->C1, m->C2};
    for (int k = 0; k < 3; ++k) {
      for (size_t s = 0; s < n; ++s)
        for (int r = 0; r < 3; ++r) c[4 * s + r] = in[9 * (size_t)ids[s] + 3 * k + r];
      MCHK(hipMemcpy(cols[k], c.data(), n * 16, hipMemcpyHostToDevice));
    }
  } else {
    for (size_t s = 0; s < n; ++s) std::memcpy(c.data() + 4 * s, in + 4 * (size_t)ids[s], 16);
    MCHK(hipMemcpy(field == PBF_MPM_POSITION ? m->x : m->v, c.data(), n * 16, hipMemcpyHostToDevice));
  }
  return PBF_OK;
}
MPM_CATCH(m)

int pbf_mpm_g2p(pbf_mpm *m) try {
  if (!m) return mfail(m, PBF_E_INVALID, "null context");
  if (!m->n) return PBF_OK;
  MCHK(hipSetDevice(m->device));
  hipLaunchKernelGGL(k_mpm_g2p<false>, dim3(nblk(m->n)), dim3(MBS), 0, m->stream, m->n, m->x, m->v, m->C0, m->C1,
                     m->C2, m->grid, (const float4 *)nullptr, (const float4 *)nullptr, (const uint32_t *)nullptr,
                     (uint32_t *)nullptr, m->K);
  MCHK(hipGetLastError());
  return PBF_OK;
}
MPM_CATCH(m)

uint32_t pbf_mpm_particle_count(const pbf_mpm *m) { return m ? m->n : 0u; }

const char *pbf_mpm_last_error(const pbf_mpm *m) { return m ? m->err.c_str() : pbf_last_error(nullptr); }

}  // extern "C"
