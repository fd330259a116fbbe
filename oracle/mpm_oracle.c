/*
 * mpm_oracle.c -- CPU oracle for the MLS-MPM step (TEST INFRASTRUCTURE ONLY:
 * used by tests/ and bench.py's cpu_baseline leg, never by the library).
 *
 * Restates the reference's MLS-MPM shaders (LinzhouLi/WebGPU-Fluid-Simulation):
 *   P2G        src/simulator/MPM/MPMShader.ts:3-148 (P2GComputeShader)
 *   grid       src/simulator/MPM/MPMShader.ts:151-193 (GridComputeShader)
 *   G2P        src/simulator/MPM/MPMShader.ts:196-282 (G2PComputeShader)
 *   parameters src/simulator/MPM/MPM.ts:62-83 (n_grid 32, dt 8e-4, bound 3,
 *              rho 1, p_vol (dx/2)^2, E 400), pass order MPM.ts:204-238
 *
 * Numerics (one fixed reading of the WGSL; the GPU kernels follow it):
 *   - every expression is evaluated left to right as written in the shader,
 *     no contraction (built with -ffp-contract=off); x / dx is a division;
 *   - base = u32(X - 0.5) saturating at 0 (WGSL float -> u32 conversion);
 *   - mat3x3 values are column-major like WGSL: C[c][r] is column c, row r;
 *     affine * dpos = (col0 * dpos.x + col1 * dpos.y) + col2 * dpos.z;
 *   - P2G sums: the reference adds particles into a node with float atomics
 *     in whatever order they arrive.  Here the order is fixed: for each of
 *     the node's nine base-cell columns (bx, by) in ascending order, a
 *     partial sum over its cells bz = K-2..K, each cell's particles by
 *     ascending id; the node's sum then adds the nine partials in column
 *     order (s0 + s1) + ... + s8.
 *   - G2P sums the 27 nodes in the shader's order (i outer, k inner).
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_mpm {
  int n, ng;
  float dx, dt, p_vol, p_mass, E, inv4_dx2;
  float g[4];
  uint32_t bound;
  float *x;   /* n x 4: position (w unused) */
  float *v;   /* n x 4: velocity, J in w */
  float *C;   /* n x 9: C[c*3 + r] column-major */
  float *grid;/* ng^3 x 4: after a step, the grid velocity (w = 1 where mass) */
  /* per-step scratch */
  uint32_t *base, *cnt, *off, *sorted;
  float *fx, *aff, *pmv;
};

orc_mpm *orc_mpm_create(const orc_mpm_params *p, int n, const float *xyzw) {
  orc_mpm *m = (orc_mpm *)calloc(1, sizeof(orc_mpm));
  int ng = p->n_grid;
  m->n = n;
  m->ng = ng;
  m->dx = (float)(1.0 / (double)ng);
  m->dt = (float)p->dt;
  m->p_vol = (float)((0.5 / ng) * (0.5 / ng));
  m->p_mass = (float)(p->rho * ((0.5 / ng) * (0.5 / ng)));
  m->E = (float)p->E;
  m->inv4_dx2 = 4.0f / (m->dx * m->dx);
  m->bound = p->bound;
  m->g[0] = 0.0f; m->g[1] = -(float)p->gravity; m->g[2] = 0.0f; m->g[3] = 0.0f;
  size_t nn = (size_t)(n > 0 ? n : 1);
  size_t G = (size_t)ng * ng * ng;
  m->x = (float *)calloc(nn * 4, 4);
  m->v = (float *)calloc(nn * 4, 4);
  m->C = (float *)calloc(nn * 9, 4);
  m->grid = (float *)calloc(G * 4, 4);
  m->base = (uint32_t *)calloc(nn * 3, 4);
  m->cnt = (uint32_t *)calloc(G, 4);
  m->off = (uint32_t *)calloc(G + 1, 4);
  m->sorted = (uint32_t *)calloc(nn, 4);
  m->fx = (float *)calloc(nn * 3, 4);
  m->aff = (float *)calloc(nn * 9, 4);
  m->pmv = (float *)calloc(nn * 3, 4);
  for (int i = 0; i < n; ++i) {
    for (int a = 0; a < 3; ++a) m->x[4 * i + a] = xyzw[4 * i + a];
    m->v[4 * i + 3] = 1.0f; /* J = 1 (MPM.ts:115-119) */
  }
  return m;
}

void orc_mpm_destroy(orc_mpm *m) {
  if (!m) return;
  free(m->x); free(m->v); free(m->C); free(m->grid); free(m->base); free(m->cnt); free(m->off);
  free(m->sorted); free(m->fx); free(m->aff); free(m->pmv);
  free(m);
}

static inline uint32_t f2u_sat(float f) { return f > 0.0f ? (uint32_t)f : 0u; }

/* quadratic B-spline weights of the shader: w[k][a] for node offset k, axis a */
static inline void bspline(const float fx[3], float w[3][3]) {
  for (int a = 0; a < 3; ++a) {
    float t0 = 1.5f - fx[a], t1 = fx[a] - 1.0f, t2 = fx[a] - 0.5f;
    w[0][a] = (0.5f * t0) * t0;
    w[1][a] = 0.75f - t1 * t1;
    w[2][a] = (0.5f * t2) * t2;
  }
}

/* G2P (MPMShader.ts:240-280): velocity, affine matrix and position from the
   27 grid nodes (nodes outside the grid read as zero velocity). */
static void g2p(orc_mpm *m) {
  const int n = m->n, ng = m->ng;
  const float dx = m->dx, dt = m->dt;
#pragma omp parallel for schedule(static)
  for (int p = 0; p < n; ++p) {
    float X[3], fx[3];
    uint32_t b[3];
    for (int a = 0; a < 3; ++a) {
      X[a] = m->x[4 * p + a] / dx;
      b[a] = f2u_sat(X[a] - 0.5f);
      fx[a] = X[a] - (float)b[a];
    }
    float w[3][3];
    bspline(fx, w);
    float nv[3] = {0.0f, 0.0f, 0.0f}, nC[9] = {0};
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k < 3; ++k) {
          const float dpos[3] = {((float)i - fx[0]) * dx, ((float)j - fx[1]) * dx, ((float)k - fx[2]) * dx};
          const float weight = (w[i][0] * w[j][1]) * w[k][2];
          const uint32_t gi = b[0] + i, gj = b[1] + j, gk = b[2] + k;
          float vel[3] = {0.0f, 0.0f, 0.0f};
          if (gi < (uint32_t)ng && gj < (uint32_t)ng && gk < (uint32_t)ng) {
            const float *g = m->grid + 4 * (((size_t)gi * ng + gj) * ng + gk);
            vel[0] = g[0]; vel[1] = g[1]; vel[2] = g[2];
          }
          for (int a = 0; a < 3; ++a) nv[a] = nv[a] + weight * vel[a];
          const float s = m->inv4_dx2 * weight;
          for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 3; ++r) nC[3 * c + r] = nC[3 * c + r] + s * (vel[c] * dpos[r]);
        }
    for (int a = 0; a < 3; ++a) {
      m->v[4 * p + a] = nv[a];
      m->x[4 * p + a] = m->x[4 * p + a] + dt * nv[a];
    }
    memcpy(m->C + 9 * p, nC, sizeof nC);
  }
}

/* P2G (particle part, then the node sums) and the grid pass; raw (nullable)
   receives each node's sums (momentum xyz, mass) before the grid pass. */
static void p2g_grid(orc_mpm *m, float *raw) {
  const int n = m->n, ng = m->ng;
  const size_t G = (size_t)ng * ng * ng;
  const float dx = m->dx, dt = m->dt;
  /* P2G, particle part (MPMShader.ts:100-115): J, stress, affine, base, fx */
  for (int p = 0; p < n; ++p) {
    const float *C = m->C + 9 * p;
    float J = m->v[4 * p + 3] * (1.0f + dt * ((C[0] + C[4]) + C[8]));
    m->v[4 * p + 3] = J;
    float X[3];
    for (int a = 0; a < 3; ++a) {
      X[a] = m->x[4 * p + a] / dx;
      uint32_t b = f2u_sat(X[a] - 0.5f);
      m->base[3 * p + a] = b;
      m->fx[3 * p + a] = X[a] - (float)b;
    }
    float stress = (((((-dt) * 4.0f) * m->E) * m->p_vol) * (J - 1.0f)) / (dx * dx);
    for (int c = 0; c < 3; ++c)
      for (int r = 0; r < 3; ++r) {
        float idm = (c == r) ? stress : 0.0f * stress;
        m->aff[9 * p + 3 * c + r] = idm + m->p_mass * C[3 * c + r];
      }
    for (int a = 0; a < 3; ++a) m->pmv[3 * p + a] = m->p_mass * m->v[4 * p + a];
  }
  /* particles grouped by base cell (ascending id inside a cell) */
  memset(m->cnt, 0, G * 4);
  for (int p = 0; p < n; ++p) {
    const uint32_t *b = m->base + 3 * p;
    if (b[0] < (uint32_t)ng && b[1] < (uint32_t)ng && b[2] < (uint32_t)ng)
      m->cnt[((size_t)b[0] * ng + b[1]) * ng + b[2]]++;
  }
  m->off[0] = 0;
  for (size_t c = 0; c < G; ++c) m->off[c + 1] = m->off[c] + m->cnt[c];
  {
    uint32_t *fill = (uint32_t *)malloc(G * 4);
    memcpy(fill, m->off, G * 4);
    for (int p = 0; p < n; ++p) {
      const uint32_t *b = m->base + 3 * p;
      if (b[0] < (uint32_t)ng && b[1] < (uint32_t)ng && b[2] < (uint32_t)ng)
        m->sorted[fill[((size_t)b[0] * ng + b[1]) * ng + b[2]]++] = (uint32_t)p;
    }
    free(fill);
  }
  /* P2G sums per node, then the grid pass (MPMShader.ts:163-191) */
#pragma omp parallel for schedule(static)
  for (int node = 0; node < (int)G; ++node) {
    const int I = node / (ng * ng), Jn = (node / ng) % ng, K = node % ng;
    /* one partial sum per base-cell column (bx, by), over its cells
       bz = K-2..K and their particles in order; the node's sum adds the nine
       partials in column order (a column outside the grid adds +0) */
    float acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    int col = 0;
    for (int bx = I - 2; bx <= I; ++bx)
      for (int by = Jn - 2; by <= Jn; ++by, ++col) {
        float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int bz = K - 2; bz <= K; ++bz) {
          if (bx < 0 || by < 0 || bz < 0) continue;
          const size_t cell = ((size_t)bx * ng + by) * ng + bz;
          const int oi = I - bx, oj = Jn - by, ok = K - bz;
          for (uint32_t t = m->off[cell]; t < m->off[cell + 1]; ++t) {
            const int p = (int)m->sorted[t];
            const float *f = m->fx + 3 * p;
            float w[3][3];
            bspline(f, w);
            const float dpos[3] = {((float)oi - f[0]) * dx, ((float)oj - f[1]) * dx, ((float)ok - f[2]) * dx};
            const float weight = (w[oi][0] * w[oj][1]) * w[ok][2];
            const float *A = m->aff + 9 * p;
            for (int r = 0; r < 3; ++r) {
              float ad = (A[r] * dpos[0] + A[3 + r] * dpos[1]) + A[6 + r] * dpos[2];
              part[r] += weight * (m->pmv[3 * p + r] + ad);
            }
            part[3] += weight * m->p_mass;
          }
        }
        for (int a = 0; a < 4; ++a) acc[a] = col == 0 ? part[a] : acc[a] + part[a];
      }
    if (raw)
      for (int a = 0; a < 4; ++a) raw[4 * (size_t)node + a] = acc[a];
    /* grid pass */
    float vel[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    if (acc[3] > 0.0f)
      for (int a = 0; a < 4; ++a) vel[a] = acc[a] / acc[3];
    for (int a = 0; a < 4; ++a) vel[a] = vel[a] + dt * m->g[a];
    const uint32_t hi = (uint32_t)ng - m->bound;
    const uint32_t ix[3] = {(uint32_t)I, (uint32_t)Jn, (uint32_t)K};
    for (int a = 0; a < 3; ++a)
      if ((ix[a] < m->bound && vel[a] < 0.0f) || (ix[a] >= hi && vel[a] > 0.0f)) vel[a] = 0.0f;
    for (int a = 0; a < 4; ++a) m->grid[4 * (size_t)node + a] = vel[a];
  }
}

void orc_mpm_step(orc_mpm *m) {
  p2g_grid(m, NULL);
  g2p(m);
}

void orc_mpm_p2g(orc_mpm *m, float *raw) { p2g_grid(m, raw); }

void orc_mpm_get(const orc_mpm *m, int field, float *out) {
  const int n = m->n;
  if (field == ORC_MPM_POSITION) {
    memcpy(out, m->x, (size_t)n * 16);
  } else if (field == ORC_MPM_VELOCITY) {
    memcpy(out, m->v, (size_t)n * 16);
  } else if (field == ORC_MPM_AFFINE) {
    memcpy(out, m->C, (size_t)n * 36);
  } else if (field == ORC_MPM_GRID) {
    memcpy(out, m->grid, (size_t)m->ng * m->ng * m->ng * 16);
  }
}

void orc_mpm_set(orc_mpm *m, int field, const float *in) {
  const int n = m->n;
  if (field == ORC_MPM_VELOCITY) memcpy(m->v, in, (size_t)n * 16);
  else if (field == ORC_MPM_AFFINE) memcpy(m->C, in, (size_t)n * 36);
  else if (field == ORC_MPM_POSITION) memcpy(m->x, in, (size_t)n * 16);
  else if (field == ORC_MPM_GRID) memcpy(m->grid, in, (size_t)m->ng * m->ng * m->ng * 16);
}

/* G2P alone on the grid velocity field set with orc_mpm_set (the linear-
   field tests). */
void orc_mpm_g2p_only(orc_mpm *m) { g2p(m); }
