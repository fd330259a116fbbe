/*
 * pbf_oracle.c -- CPU oracle for the Position Based Fluids substep.
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() to check
 * the GPU result and by bench.py's cpu_baseline leg; never linked into the shipped
 * library.  See pbf_oracle.h for the reference files each part restates.
 *
 * Numerics (fixed here because the reference leaves them open):
 *   - physical constants are formed in double and rounded once to float;
 *   - no implicit contraction (built with -ffp-contract=off); fused
 *     multiply-adds only where fmaf() is written; every sum is written in
 *     the order it is evaluated;
 *   - |x_ij|^2 = fma(dz, dz, fma(dy, dy, dx*dx)); poly6 uses it directly;
 *   - spiky / cohesion gradients are f(r) * x_ij with one division by r;
 *   - sin() in the wall clamp is the Cephes single-precision sinf algorithm
 *     (S. L. Moshier), evaluated with plain float operations;
 *   - neighbour order: 27 cells x-major, then by z-slab inside the cell
 *     (cell_slabs slabs of height h / cell_slabs), then ascending particle id.
 */
#include "pbf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc_sim {
  int n;
  /* parameters */
  double r64, h64, rho064, V64;
  float r, h, h2, half_h, V, V2, m, Kp, Ks, Kc, bias, two_rho0, eps_lambda;
  float half_r, two_r;
  float lo[3], hi[3], inv_cell;
  int dims[3];
  int grid_mode, hash_log2;
  int slabs;      /* z-slabs per cell (order inside a cell) */
  uint32_t ncells;
  double Kp64, h2_64;
  /* config (src/common/shader.ts:50-57) */
  float xsph, vort, tension, gravity;
  /* boundary field */
  int field_kind; /* 0 none, 1 cubic, 2 trilinear */
  int fres[3];
  float flo[3], finv[3], fgrad_scale[3];
  float *fsdf, *fvol; /* cubic */
  float *fsv;         /* trilinear interleaved */
  int quirk;
  /* state, id order */
  float *x, *v, *a, *lam, *om, *nrm, *bnd, *xs, *xi, *xn;
  /* search */
  uint32_t *pcell, *cnt, *off, *sorted;
  uint8_t *pslab;
  uint32_t *noff, *nbr;
  uint64_t nbr_cap, npairs;
};

/* ------------------------------------------------------------------ */
/* Cephes sinf (Moshier), plain float operations.                      */
float orc_sinf(float xx) {
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

static inline float fractf_(float v) { return v - floorf(v); }

/* boundary_rand (src/simulator/PBF/shader/common.ts:55-67) generalised to the
 * box [lo, hi]: bounds lo + 1e-4*rand, hi - 1e-4*rand. */
static inline void clamp_rand(const orc_sim *s, float *p) {
  const float A = 12.9898f, B = 78.233f;
  float b0 = fmaf(p[2], B, p[1] * A);
  float b1 = fmaf(p[2], B, p[0] * A);
  float b2 = fmaf(p[1], B, p[0] * A);
  float rnd[3];
  rnd[0] = fractf_(orc_sinf(b0) * 43758.5453f);
  rnd[1] = fractf_(orc_sinf(b1) * 43758.5453f);
  rnd[2] = fractf_(orc_sinf(b2) * 43758.5453f);
  for (int c = 0; c < 3; ++c) {
    float e = 1e-4f * rnd[c];
    float lo = s->lo[c] + e;
    float hi = s->hi[c] - e;
    float m = p[c] < hi ? p[c] : hi;
    p[c] = m > lo ? m : lo;
  }
}

void orc_clamp(const orc_sim *s, int n, const float *x3, float *out3) {
  for (int i = 0; i < n; ++i) {
    float p[3] = {x3[3 * i], x3[3 * i + 1], x3[3 * i + 2]};
    clamp_rand(s, p);
    out3[3 * i] = p[0]; out3[3 * i + 1] = p[1]; out3[3 * i + 2] = p[2];
  }
}

static inline float dist2(float dx, float dy, float dz) { return fmaf(dz, dz, fmaf(dy, dy, dx * dx)); }

/* ------------------------------------------------------------------ */
/* SPH kernels (src/simulator/PBF/shader/common.ts:1-40).              */
static inline float k_poly6(const orc_sim *s, float d2) {
  if (!(d2 <= s->h2)) return 0.0f;
  float t = s->h2 - d2;
  return s->Kp * ((t * t) * t);
}
/* spiky gradient = f * x_ij */
static inline float k_spiky_f(const orc_sim *s, float d2) {
  if (!(d2 > 0.0f && d2 <= s->h2)) return 0.0f;
  float r = sqrtf(d2);
  float q = s->h - r;
  return (s->Ks * (q * q)) / r;
}
/* Akinci cohesion = f * x_ij */
static inline float k_cohesion_f(const orc_sim *s, float d2) {
  if (!(d2 > 0.0f && d2 <= s->h2)) return 0.0f;
  float r = sqrtf(d2);
  float a = s->h - r;
  float t = ((a * a) * a) * ((r * r) * r);
  float res = (r > s->half_h) ? t : fmaf(2.0f, t, s->bias);
  return (s->Kc * res) / r;
}

void orc_kernels(const orc_sim *s, int n, const float *d2, float *poly6, float *spiky_f, float *cohesion_f) {
  for (int i = 0; i < n; ++i) {
    poly6[i] = k_poly6(s, d2[i]);
    spiky_f[i] = k_spiky_f(s, d2[i]);
    cohesion_f[i] = k_cohesion_f(s, d2[i]);
  }
}

/* ------------------------------------------------------------------ */
/* Boundary fields.                                                    */

/* 32-node cubic serendipity element (discreteFieldShader.ts:15-130).
 * Node order: groups of four, see the value gather in cubic_eval. */
static void cubic_shape(const float sv[3], float N[32], float dN[32][3]) {
  const float C1 = 0.015625f, C9 = 0.140625f, C19 = 0.296875f;
  float x = sv[0], y = sv[1], z = sv[2];
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

  /* corner derivatives: d/ds_a of (1 +- s_a)(9|s|^2 - 19) = 18 s_a -+ w_a,
     w_a = 9 (2 s_a^2 + |s|^2) - 19 */
  float wx = 9.0f * (2.0f * xx + ssum) - 19.0f;
  float wy = 9.0f * (2.0f * yy + ssum) - 19.0f;
  float wz = 9.0f * (2.0f * zz + ssum) - 19.0f;
  float dmx = 18.0f * x - wx, dpx = 18.0f * x + wx;
  float dmy = 18.0f * y - wy, dpy = 18.0f * y + wy;
  float dmz = 18.0f * z - wz, dpz = 18.0f * z + wz;
  /* corners z-: (x-,y-) (x+,y-) (x-,y+) (x+,y+); then z+ */
  const float *Xs[2] = {&mx, &px};
  const float *Ys[2] = {&my, &py};
  const float *Zs[2] = {&mz, &pz};
  const float *DX[2] = {&dmx, &dpx};
  const float *DY[2] = {&dmy, &dpy};
  const float *DZ[2] = {&dmz, &dpz};
  for (int c = 0; c < 8; ++c) {
    int ex = c & 1, ey = (c >> 1) & 1, ez = (c >> 2) & 1;
    float X = *Xs[ex], Y = *Ys[ey], Z = *Zs[ez];
    dN[c][0] = C1 * (*DX[ex] * (Y * Z));
    dN[c][1] = C1 * (*DY[ey] * (X * Z));
    dN[c][2] = C1 * (*DZ[ez] * (X * Y));
  }
  /* edge derivative along the edge axis: d/ds[(1 - s^2)(1 -+ 3s)] */
  float cxm = (-(3.0f - 9.0f * xx)) - 2.0f * x, cxp = (3.0f - 9.0f * xx) - 2.0f * x;
  float cym = (-(3.0f - 9.0f * yy)) - 2.0f * y, cyp = (3.0f - 9.0f * yy) - 2.0f * y;
  float czm = (-(3.0f - 9.0f * zz)) - 2.0f * z, czp = (3.0f - 9.0f * zz) - 2.0f * z;
  float qax = qx * ax, qbx = qx * bx, qay = qy * ay, qby = qy * by, qaz = qz * az, qbz = qz * bz;
  /* x edges: node 8 + 4*ey + 2*ez + t ; factors Y (ey), Z (ez) */
  for (int ey = 0; ey < 2; ++ey)
    for (int ez = 0; ez < 2; ++ez)
      for (int t = 0; t < 2; ++t) {
        int k = 8 + 4 * ey + 2 * ez + t;
        float qa = t ? qbx : qax;
        float Y = *Ys[ey], Z = *Zs[ez];
        dN[k][0] = C9 * ((t ? cxp : cxm) * (Y * Z));
        dN[k][1] = C9 * ((ey ? qa : -qa) * Z);
        dN[k][2] = C9 * ((ez ? qa : -qa) * Y);
      }
  /* y edges: node 16 + 4*ez + 2*ex + t */
  for (int ez = 0; ez < 2; ++ez)
    for (int ex = 0; ex < 2; ++ex)
      for (int t = 0; t < 2; ++t) {
        int k = 16 + 4 * ez + 2 * ex + t;
        float qa = t ? qby : qay;
        float X = *Xs[ex], Z = *Zs[ez];
        dN[k][0] = C9 * ((ex ? qa : -qa) * Z);
        dN[k][1] = C9 * ((t ? cyp : cym) * (X * Z));
        dN[k][2] = C9 * ((ez ? qa : -qa) * X);
      }
  /* z edges: node 24 + 4*ex + 2*ey + t */
  for (int ex = 0; ex < 2; ++ex)
    for (int ey = 0; ey < 2; ++ey)
      for (int t = 0; t < 2; ++t) {
        int k = 24 + 4 * ex + 2 * ey + t;
        float qa = t ? qbz : qaz;
        float X = *Xs[ex], Y = *Ys[ey];
        dN[k][0] = C9 * ((ex ? qa : -qa) * Y);
        dN[k][1] = C9 * ((ey ? qa : -qa) * X);
        dN[k][2] = C9 * ((t ? czp : czm) * (X * Y));
      }
}

/* indices of the 32 node values of cell (i, j, k) in the .cdm layout */
static void cubic_nodes(const int R[3], int i, int j, int k, int quirk, uint32_t idx[32]) {
  uint32_t nx = (uint32_t)R[0] + 1, ny = (uint32_t)R[1] + 1, nz = (uint32_t)R[2] + 1;
  uint32_t ncorner = nx * ny * nz;
  uint32_t nex = nz * ny * (2u * R[0]);
  uint32_t ney = nx * nz * (2u * R[1]);
#define CO(a, b, c) ((uint32_t)(((c) * ny + (b)) * nx + (a)))
#define EX(a, b, c) (ncorner + (uint32_t)(((c) * ny + (b)) * (2u * R[0]) + (a)))
#define EY(a, b, c) (ncorner + nex + (uint32_t)(((b) * nz + (c)) * (2u * R[1]) + (a)))
#define EZ(a, b, c) (ncorner + nex + ney + (uint32_t)(((c) * nx + (b)) * (2u * R[2]) + (a)))
  uint32_t i2 = 2u * i, j2 = 2u * j, k2 = 2u * k;
  uint32_t ii = i, jj = j, kk = k;
  idx[0] = CO(ii, jj, kk); idx[1] = CO(ii + 1, jj, kk); idx[2] = CO(ii, jj + 1, kk); idx[3] = CO(ii + 1, jj + 1, kk);
  idx[4] = CO(ii, jj, kk + 1); idx[5] = CO(ii + 1, jj, kk + 1); idx[6] = CO(ii, jj + 1, kk + 1); idx[7] = CO(ii + 1, jj + 1, kk + 1);
  /* edgex[k][j][i2]: EX(i2, j, k) */
  idx[8] = EX(i2, jj, kk); idx[9] = EX(i2 + 1, jj, kk);
  idx[10] = EX(i2, jj, kk + 1); idx[11] = EX(quirk ? i2 : i2 + 1, jj, kk + 1);
  idx[12] = EX(i2, jj + 1, kk); idx[13] = EX(i2 + 1, jj + 1, kk);
  idx[14] = EX(i2, jj + 1, kk + 1); idx[15] = EX(quirk ? i2 : i2 + 1, jj + 1, kk + 1);
  /* edgey[i][k][j2]: EY(j2, i, k) */
  idx[16] = EY(j2, ii, kk); idx[17] = EY(j2 + 1, ii, kk); idx[18] = EY(j2, ii + 1, kk); idx[19] = EY(j2 + 1, ii + 1, kk);
  idx[20] = EY(j2, ii, kk + 1); idx[21] = EY(j2 + 1, ii, kk + 1); idx[22] = EY(j2, ii + 1, kk + 1); idx[23] = EY(j2 + 1, ii + 1, kk + 1);
  /* edgez[j][i][k2]: EZ(k2, i, j) */
  uint32_t k2b = quirk ? k2 : k2 + 1;
  idx[24] = EZ(k2, ii, jj); idx[25] = EZ(k2b, ii, jj); idx[26] = EZ(k2, ii, jj + 1); idx[27] = EZ(k2b, ii, jj + 1);
  idx[28] = EZ(k2, ii + 1, jj); idx[29] = EZ(k2b, ii + 1, jj); idx[30] = EZ(k2, ii + 1, jj + 1); idx[31] = EZ(k2b, ii + 1, jj + 1);
#undef CO
#undef EX
#undef EY
#undef EZ
}

static inline float dot32(const float *N, const float *v) {
  float r[8];
  for (int g = 0; g < 8; ++g)
    r[g] = ((N[4 * g] * v[4 * g] + N[4 * g + 1] * v[4 * g + 1]) + N[4 * g + 2] * v[4 * g + 2]) + N[4 * g + 3] * v[4 * g + 3];
  float q0 = r[0] + r[4], q1 = r[1] + r[5], q2 = r[2] + r[6], q3 = r[3] + r[7];
  return ((q0 + q1) + q2) + q3;
}

/* local coordinate of x in field cell: returns cell index, writes s in [-1,1] */
static inline void cubic_locate(const orc_sim *s, const float *x, int cell[3], float sv[3]) {
  for (int a = 0; a < 3; ++a) {
    float u = (x[a] - s->flo[a]) * s->finv[a];
    float fl = floorf(u);
    int c = (int)fl;
    if (c < 0) { c = 0; fl = 0.0f; }
    if (c > s->fres[a] - 1) { c = s->fres[a] - 1; fl = (float)c; }
    cell[a] = c;
    sv[a] = 2.0f * (u - fl) - 1.0f;
  }
}

/* sdf, gradient (and volume when want_vol) at x */
static void field_eval(const orc_sim *s, const float *x, float grad[3], float *sdf, float *vol, int want_vol) {
  if (s->field_kind == 1) {
    int cell[3];
    float sv[3], N[32], dN[32][3];
    cubic_locate(s, x, cell, sv);
    cubic_shape(sv, N, dN);
    uint32_t idx[32];
    cubic_nodes(s->fres, cell[0], cell[1], cell[2], 0, idx);
    float v[32];
    for (int k = 0; k < 32; ++k) v[k] = s->fsdf[idx[k]];
    *sdf = dot32(N, v);
    float G[8][3];
    for (int g = 0; g < 8; ++g)
      for (int a = 0; a < 3; ++a)
        G[g][a] = ((dN[4 * g][a] * v[4 * g] + dN[4 * g + 1][a] * v[4 * g + 1]) + dN[4 * g + 2][a] * v[4 * g + 2]) +
                  dN[4 * g + 3][a] * v[4 * g + 3];
    for (int a = 0; a < 3; ++a) {
      float r3 = ((G[0][a] + G[1][a]) + G[2][a]) + G[3][a];
      float r4 = ((G[4][a] + G[5][a]) + G[6][a]) + G[7][a];
      grad[a] = s->fgrad_scale[a] * (r3 + r4);
    }
    if (want_vol) {
      uint32_t vidx[32];
      cubic_nodes(s->fres, cell[0], cell[1], cell[2], s->quirk, vidx);
      for (int k = 0; k < 32; ++k) v[k] = s->fvol[vidx[k]];
      *vol = dot32(N, v);
    }
  } else {
    /* trilinear over (R+1)^3 nodes, x fastest, (sdf, vol) interleaved */
    int c[3];
    float f[3];
    for (int a = 0; a < 3; ++a) {
      float u = (x[a] - s->flo[a]) * s->finv[a];
      float R = (float)s->fres[a];
      u = u > 0.0f ? u : 0.0f;
      u = u < R ? u : R;
      float fl = floorf(u);
      int ci = (int)fl;
      if (ci > s->fres[a] - 1) { ci = s->fres[a] - 1; fl = (float)ci; }
      c[a] = ci;
      f[a] = u - fl;
    }
    uint32_t nx = (uint32_t)s->fres[0] + 1, ny = (uint32_t)s->fres[1] + 1;
    uint32_t b000 = ((uint32_t)c[2] * ny + (uint32_t)c[1]) * nx + (uint32_t)c[0];
    uint32_t o[8] = {b000, b000 + 1, b000 + nx, b000 + nx + 1,
                     b000 + nx * ny, b000 + nx * ny + 1, b000 + nx * ny + nx, b000 + nx * ny + nx + 1};
    float cs[8], cv[8];
    for (int k = 0; k < 8; ++k) { cs[k] = s->fsv[2 * o[k]]; cv[k] = s->fsv[2 * o[k] + 1]; }
    float fx = f[0], fy = f[1], fz = f[2];
    /* corners: 0 (000) 1 (100) 2 (010) 3 (110) 4 (001) 5 (101) 6 (011) 7 (111) */
    float c00 = fmaf(fx, cs[1] - cs[0], cs[0]);
    float c10 = fmaf(fx, cs[3] - cs[2], cs[2]);
    float c01 = fmaf(fx, cs[5] - cs[4], cs[4]);
    float c11 = fmaf(fx, cs[7] - cs[6], cs[6]);
    float c0 = fmaf(fy, c10 - c00, c00);
    float c1 = fmaf(fy, c11 - c01, c01);
    *sdf = fmaf(fz, c1 - c0, c0);
    float ex00 = cs[1] - cs[0], ex10 = cs[3] - cs[2], ex01 = cs[5] - cs[4], ex11 = cs[7] - cs[6];
    float ex0 = fmaf(fy, ex10 - ex00, ex00);
    float ex1 = fmaf(fy, ex11 - ex01, ex01);
    grad[0] = fmaf(fz, ex1 - ex0, ex0) * s->finv[0];
    float ey0 = c10 - c00, ey1 = c11 - c01;
    grad[1] = fmaf(fz, ey1 - ey0, ey0) * s->finv[1];
    grad[2] = (c1 - c0) * s->finv[2];
    if (want_vol) {
      float v00 = fmaf(fx, cv[1] - cv[0], cv[0]);
      float v10 = fmaf(fx, cv[3] - cv[2], cv[2]);
      float v01 = fmaf(fx, cv[5] - cv[4], cv[4]);
      float v11 = fmaf(fx, cv[7] - cv[6], cv[6]);
      float v0 = fmaf(fy, v10 - v00, v00);
      float v1 = fmaf(fy, v11 - v01, v01);
      *vol = fmaf(fz, v1 - v0, v0);
    }
  }
}

/* BoundaryVolumeShader (src/simulator/PBF/shader/boundaryVolume.ts:31-55) */
static void boundary_sample(const orc_sim *s, const float *x, float out[4]) {
  out[0] = out[1] = out[2] = out[3] = 0.0f;
  if (!s->field_kind) return;
  float g[3], d, vol = 0.0f;
  field_eval(s, x, g, &d, &vol, 0);
  float len = sqrtf(dist2(g[0], g[1], g[2]));
  if (d > 0.0f && d < s->h && len > 1e-9f) {
    field_eval(s, x, g, &d, &vol, 1);
    if (vol > 0.0f) {
      float nx = g[0] / len, ny = g[1] / len, nz = g[2] / len;
      float a = d + s->half_r;
      float dd = a > s->two_r ? a : s->two_r;
      out[0] = fmaf(-dd, nx, x[0]);
      out[1] = fmaf(-dd, ny, x[1]);
      out[2] = fmaf(-dd, nz, x[2]);
      out[3] = vol;
    }
  }
}

void orc_boundary_sample(const orc_sim *s, int n, const float *x3, float *out4) {
  for (int i = 0; i < n; ++i) boundary_sample(s, x3 + 3 * i, out4 + 4 * i);
}

void orc_field_eval(const orc_sim *s, int n, const float *x3, float *out5) {
  for (int i = 0; i < n; ++i) {
    float g[3] = {0, 0, 0}, d = 0.0f, vol = 0.0f;
    if (s->field_kind) field_eval(s, x3 + 3 * i, g, &d, &vol, 1);
    out5[5 * i] = g[0]; out5[5 * i + 1] = g[1]; out5[5 * i + 2] = g[2];
    out5[5 * i + 3] = d; out5[5 * i + 4] = vol;
  }
}

/* ------------------------------------------------------------------ */
/* Construction.                                                       */
orc_sim *orc_create(const orc_params *p, int n, const float *xyzw) {
  orc_sim *s = (orc_sim *)calloc(1, sizeof(orc_sim));
  s->n = n;
  s->r64 = p->particle_radius;
  s->h64 = p->kernel_radius;
  s->rho064 = p->rest_density;
  double d = 2.0 * s->r64;
  s->V64 = ((0.9 * d) * d) * d; /* SPH.ts:48-50 */
  const double PI = 3.141592653589793;
  s->r = (float)s->r64;
  s->h = (float)s->h64;
  s->h2 = (float)(s->h64 * s->h64);
  s->half_h = (float)(0.5 * s->h64);
  s->V = (float)s->V64;
  s->V2 = (float)(s->V64 * s->V64);
  s->m = (float)(s->V64 * s->rho064);
  s->Kp = (float)(315.0 / (64.0 * PI * pow(s->h64, 9)));
  s->Ks = (float)(-45.0 / (PI * pow(s->h64, 6)));
  s->Kc = (float)(32.0 / (PI * pow(s->h64, 9)));
  s->bias = (float)(-pow(s->h64, 6) / 64.0);
  s->two_rho0 = (float)(2.0 * s->rho064);
  s->eps_lambda = 1e-6f;
  s->half_r = 0.5f * s->r;
  s->two_r = 2.0f * s->r;
  s->Kp64 = 315.0 / (64.0 * PI * pow(s->h64, 9));
  s->h2_64 = s->h64 * s->h64;
  for (int a = 0; a < 3; ++a) {
    s->lo[a] = (float)p->domain_lo[a];
    s->hi[a] = (float)p->domain_hi[a];
    double L = p->domain_hi[a] - p->domain_lo[a];
    int D = (int)ceil(L / s->h64 - 1e-7);
    s->dims[a] = D < 1 ? 1 : D;
  }
  s->inv_cell = (float)(1.0 / s->h64);
  s->grid_mode = p->grid_mode;
  s->hash_log2 = p->hash_log2;
  s->slabs = (p->cell_slabs == 1 || p->cell_slabs == 2) ? p->cell_slabs : 4;
  if (s->grid_mode == 1)
    s->ncells = 1u << p->hash_log2;
  else
    s->ncells = (uint32_t)s->dims[0] * (uint32_t)s->dims[1] * (uint32_t)s->dims[2];
  s->xsph = 0.01f; s->vort = 0.1f; s->tension = 0.5f; s->gravity = 9.8f; /* config.ts:22-28 */
  size_t n4 = (size_t)(n > 0 ? n : 1) * 4;
  s->x = (float *)calloc(n4, 4); s->v = (float *)calloc(n4, 4); s->a = (float *)calloc(n4, 4);
  s->lam = (float *)calloc(n4, 4); s->om = (float *)calloc(n4, 4); s->nrm = (float *)calloc(n4, 4);
  s->bnd = (float *)calloc(n4, 4); s->xs = (float *)calloc(n4, 4); s->xi = (float *)calloc(n4, 4);
  s->xn = (float *)calloc(n4, 4);
  for (int i = 0; i < n; ++i) {
    s->x[4 * i] = xyzw[4 * i]; s->x[4 * i + 1] = xyzw[4 * i + 1]; s->x[4 * i + 2] = xyzw[4 * i + 2];
    s->x[4 * i + 3] = 0.0f;
  }
  s->pcell = (uint32_t *)calloc((size_t)(n > 0 ? n : 1), 4);
  s->pslab = (uint8_t *)calloc((size_t)(n > 0 ? n : 1), 1);
  s->sorted = (uint32_t *)calloc((size_t)(n > 0 ? n : 1), 4);
  s->cnt = (uint32_t *)calloc(s->ncells, 4);
  s->off = (uint32_t *)calloc((size_t)s->ncells + 1, 4);
  s->noff = (uint32_t *)calloc((size_t)n + 1, 4);
  s->nbr_cap = (uint64_t)(n > 0 ? n : 1) * 64;
  s->nbr = (uint32_t *)malloc(s->nbr_cap * 4);
  return s;
}

void orc_destroy(orc_sim *s) {
  if (!s) return;
  free(s->x); free(s->v); free(s->a); free(s->lam); free(s->om); free(s->nrm); free(s->bnd);
  free(s->xs); free(s->xi); free(s->xn); free(s->pcell); free(s->pslab); free(s->sorted); free(s->cnt); free(s->off);
  free(s->noff); free(s->nbr); free(s->fsdf); free(s->fvol); free(s->fsv);
  free(s);
}

void orc_set_config(orc_sim *s, float xsph, float vorticity, float tension, float gravity) {
  s->xsph = xsph; s->vort = vorticity; s->tension = tension; s->gravity = gravity;
}

void orc_clear_field(orc_sim *s) {
  free(s->fsdf); free(s->fvol); free(s->fsv);
  s->fsdf = s->fvol = s->fsv = NULL;
  s->field_kind = 0;
}

int orc_set_cubic_field(orc_sim *s, const int res[3], const double lo[3], const double hi[3],
                        const float *sdf, const float *vol, int quirk) {
  orc_clear_field(s);
  size_t n = (size_t)(res[0] + 1) * (res[1] + 1) * (res[2] + 1) + (size_t)(2 * res[0]) * (res[1] + 1) * (res[2] + 1) +
             (size_t)(res[0] + 1) * (2 * res[1]) * (res[2] + 1) + (size_t)(res[0] + 1) * (res[1] + 1) * (2 * res[2]);
  s->fsdf = (float *)malloc(n * 4);
  s->fvol = (float *)malloc(n * 4);
  memcpy(s->fsdf, sdf, n * 4);
  memcpy(s->fvol, vol, n * 4);
  for (int a = 0; a < 3; ++a) {
    s->fres[a] = res[a];
    s->flo[a] = (float)lo[a];
    double L = hi[a] - lo[a];
    s->finv[a] = (float)(res[a] / L);
    s->fgrad_scale[a] = (float)(2.0 * res[a] / L);
  }
  s->quirk = quirk;
  s->field_kind = 1;
  return 0;
}

int orc_set_trilinear_field(orc_sim *s, const int res[3], const double lo[3], const double hi[3],
                            const float *sdf_vol) {
  orc_clear_field(s);
  size_t n = (size_t)(res[0] + 1) * (res[1] + 1) * (res[2] + 1);
  s->fsv = (float *)malloc(n * 2 * 4);
  memcpy(s->fsv, sdf_vol, n * 2 * 4);
  for (int a = 0; a < 3; ++a) {
    s->fres[a] = res[a];
    s->flo[a] = (float)lo[a];
    double L = hi[a] - lo[a];
    s->finv[a] = (float)(res[a] / L);
    s->fgrad_scale[a] = s->finv[a];
  }
  s->field_kind = 2;
  return 0;
}

/* ------------------------------------------------------------------ */
/* Neighbour search (neighborShader.ts, neighborSearch.ts:270-301).    */
static inline void cell_coord(const orc_sim *s, const float *x, int c[3]) {
  for (int a = 0; a < 3; ++a) {
    int v = (int)floorf((x[a] - s->lo[a]) * s->inv_cell);
    if (v < 0) v = 0;
    if (v > s->dims[a] - 1) v = s->dims[a] - 1;
    c[a] = v;
  }
}

static inline uint32_t cell_key(const orc_sim *s, int cx, int cy, int cz) {
  if (s->grid_mode == 1) {
    /* the row-major cell index modulo the table size (2^k): nearby cells
       land in nearby buckets, so the bucket-sorted particles keep their
       spatial locality; cells 2^k apart in that index share a bucket */
    uint32_t lin = ((uint32_t)cx * (uint32_t)s->dims[1] + (uint32_t)cy) * (uint32_t)s->dims[2] + (uint32_t)cz;
    return lin & (s->ncells - 1u);
  }
  return ((uint32_t)cx * (uint32_t)s->dims[1] + (uint32_t)cy) * (uint32_t)s->dims[2] + (uint32_t)cz;
}

static void predict(orc_sim *s, float dt) {
  int n = s->n;
  float g[3] = {0.0f, -s->gravity, 0.0f};
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    float p[3];
    for (int c = 0; c < 3; ++c) {
      float vel = fmaf(dt, g[c] + s->a[4 * i + c], s->v[4 * i + c]);
      p[c] = fmaf(dt, vel, s->x[4 * i + c]);
    }
    clamp_rand(s, p);
    s->xs[4 * i] = p[0]; s->xs[4 * i + 1] = p[1]; s->xs[4 * i + 2] = p[2]; s->xs[4 * i + 3] = 0.0f;
  }
}

/* z-slab of a position inside its cell: floor(S * fz), fz = (z - lo) / h -
   cz in float (the cell's own coordinate, clamped like the cell), clamped to
   [0, S). */
static inline int z_slab(const orc_sim *s, const float *x, int cz) {
  float fz = (x[2] - s->lo[2]) * s->inv_cell - (float)cz;
  int k = (int)floorf(fz * (float)s->slabs);
  if (k < 0) k = 0;
  if (k > s->slabs - 1) k = s->slabs - 1;
  return k;
}

static void search(orc_sim *s) {
  int n = s->n;
  uint32_t C = s->ncells;
  memset(s->cnt, 0, (size_t)C * 4);
  for (int i = 0; i < n; ++i) {
    int c[3];
    cell_coord(s, s->xs + 4 * i, c);
    uint32_t k = cell_key(s, c[0], c[1], c[2]);
    s->pcell[i] = k;
    s->pslab[i] = (uint8_t)z_slab(s, s->xs + 4 * i, c[2]);
    s->cnt[k]++;
  }
  s->off[0] = 0;
  for (uint32_t c = 0; c < C; ++c) s->off[c + 1] = s->off[c] + s->cnt[c];
  {
    /* inside a cell (bucket): by z-slab, then ascending id */
    uint32_t *fill = (uint32_t *)malloc((size_t)C * 4);
    memcpy(fill, s->off, (size_t)C * 4);
    for (int z = 0; z < s->slabs; ++z)
      for (int i = 0; i < n; ++i)
        if (s->pslab[i] == z) s->sorted[fill[s->pcell[i]]++] = (uint32_t)i;
    free(fill);
  }
  /* neighbour counts then lists (27 cells x-major; inside a cell the order
     above) */
  uint32_t *ncount = (uint32_t *)calloc((size_t)n + 1, 4);
  float h2 = s->h2;
  for (int pass = 0; pass < 2; ++pass) {
#pragma omp parallel for schedule(dynamic, 256)
    for (int i = 0; i < n; ++i) {
      const float *xi = s->xs + 4 * i;
      int c[3];
      cell_coord(s, xi, c);
      uint32_t seen[27];
      int nseen = 0;
      uint32_t cntn = 0;
      uint32_t w = pass ? s->noff[i] : 0;
      for (int dx = -1; dx <= 1; ++dx)
        for (int dy = -1; dy <= 1; ++dy)
          for (int dz = -1; dz <= 1; ++dz) {
            int nx = c[0] + dx, ny = c[1] + dy, nz = c[2] + dz;
            if (nx < 0 || ny < 0 || nz < 0 || nx >= s->dims[0] || ny >= s->dims[1] || nz >= s->dims[2]) continue;
            uint32_t k = cell_key(s, nx, ny, nz);
            if (s->grid_mode == 1) {
              int dup = 0;
              for (int q = 0; q < nseen; ++q) dup |= (seen[q] == k);
              if (dup) continue;
              seen[nseen++] = k;
            }
            for (uint32_t t = s->off[k]; t < s->off[k + 1]; ++t) {
              uint32_t j = s->sorted[t];
              const float *xj = s->xs + 4 * j;
              float d2 = dist2(xj[0] - xi[0], xj[1] - xi[1], xj[2] - xi[2]);
              if (d2 <= h2) {
                if (pass) s->nbr[w++] = j;
                else cntn++;
              }
            }
          }
      if (!pass) ncount[i] = cntn;
    }
    if (!pass) {
      uint64_t tot = 0;
      for (int i = 0; i < n; ++i) { s->noff[i] = (uint32_t)tot; tot += ncount[i]; }
      s->noff[n] = (uint32_t)tot;
      s->npairs = tot;
      if (tot > s->nbr_cap) {
        free(s->nbr);
        s->nbr_cap = tot + 1024;
        s->nbr = (uint32_t *)malloc(s->nbr_cap * 4);
      }
    }
  }
  free(ncount);
}

void orc_search_only(orc_sim *s, float dt) {
  predict(s, dt);
  search(s);
}

/* ------------------------------------------------------------------ */
/* Substep (PBF.ts:398-448).                                           */
void orc_substep(orc_sim *s, float dt, int iterations) {
  int n = s->n;
  predict(s, dt);
  search(s);
  memcpy(s->xi, s->xs, (size_t)n * 16);
  for (int it = 0; it < iterations; ++it) {
    /* K8 boundary sample at the current iterate */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) boundary_sample(s, s->xi + 4 * i, s->bnd + 4 * i);
    /* K9 lambda (pressureConstrain.ts:31-85) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const float *xi = s->xi + 4 * i;
      float gx = 0.0f, gy = 0.0f, gz = 0.0f, s2 = 0.0f, rho = 0.0f;
      for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
        const float *xj = s->xi + 4 * s->nbr[q];
        float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        float d2 = dist2(dx, dy, dz);
        float f = k_spiky_f(s, d2);
        float ax = f * dx, ay = f * dy, az = f * dz;
        gx += ax; gy += ay; gz += az;
        s2 = fmaf(az, az, fmaf(ay, ay, fmaf(ax, ax, s2)));
        rho += k_poly6(s, d2);
      }
      rho *= s->V;
      gx *= s->V; gy *= s->V; gz *= s->V;
      s2 *= s->V2;
      const float *b = s->bnd + 4 * i;
      if (b[3] > 0.0f) {
        float dx = xi[0] - b[0], dy = xi[1] - b[1], dz = xi[2] - b[2];
        float d2 = dist2(dx, dy, dz);
        float fb = b[3] * k_spiky_f(s, d2);
        gx = fmaf(fb, dx, gx); gy = fmaf(fb, dy, gy); gz = fmaf(fb, dz, gz);
        rho = fmaf(b[3], k_poly6(s, d2), rho);
      }
      float c = rho - 1.0f;
      float C = c > 0.0f ? c : 0.0f;
      float den = fmaf(gz, gz, fmaf(gy, gy, fmaf(gx, gx, s2)));
      s->lam[4 * i] = (-C) / (den + s->eps_lambda);
    }
    /* K10 + K11 delta p and apply (pressureConstrain.ts:109-179) */
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; ++i) {
      const float *xi = s->xi + 4 * i;
      float li = s->lam[4 * i];
      float ax = 0.0f, ay = 0.0f, az = 0.0f;
      for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
        uint32_t j = s->nbr[q];
        const float *xj = s->xi + 4 * j;
        float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
        float fl = (li + s->lam[4 * j]) * k_spiky_f(s, dist2(dx, dy, dz));
        ax = fmaf(fl, dx, ax); ay = fmaf(fl, dy, ay); az = fmaf(fl, dz, az);
      }
      ax *= s->V; ay *= s->V; az *= s->V;
      const float *b = s->bnd + 4 * i;
      if (b[3] > 0.0f) {
        float dx = xi[0] - b[0], dy = xi[1] - b[1], dz = xi[2] - b[2];
        float fb = (b[3] * li) * k_spiky_f(s, dist2(dx, dy, dz));
        ax = fmaf(fb, dx, ax); ay = fmaf(fb, dy, ay); az = fmaf(fb, dz, az);
      }
      float p[3] = {fmaf(0.5f, ax, xi[0]), fmaf(0.5f, ay, xi[1]), fmaf(0.5f, az, xi[2])};
      clamp_rand(s, p);
      float *o = s->xn + 4 * i;
      o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; o[3] = 0.0f;
    }
    float *t = s->xi; s->xi = s->xn; s->xn = t;
  }
  /* K12 density + velocity (attributeUpdate.ts:25-58) */
  float inv_dt = (float)(1.0 / (double)dt);
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = s->xi + 4 * i;
    float rho = 0.0f;
    for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
      const float *xj = s->xi + 4 * s->nbr[q];
      rho += k_poly6(s, dist2(xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]));
    }
    rho *= s->m;
    for (int c = 0; c < 3; ++c) s->v[4 * i + c] = (xi[c] - s->x[4 * i + c]) * inv_dt;
    s->v[4 * i + 3] = 0.0f;
    s->xn[4 * i + 3] = rho; /* stash density; positions are written below */
  }
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    s->x[4 * i] = s->xi[4 * i]; s->x[4 * i + 1] = s->xi[4 * i + 1]; s->x[4 * i + 2] = s->xi[4 * i + 2];
    s->x[4 * i + 3] = s->xn[4 * i + 3];
  }
  /* K13 vorticity + normal (vorticityConfinement.ts:25-66) */
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = s->x + 4 * i, *vi = s->v + 4 * i;
    float wx = 0.0f, wy = 0.0f, wz = 0.0f, nx = 0.0f, ny = 0.0f, nz = 0.0f;
    for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
      uint32_t j = s->nbr[q];
      const float *xj = s->x + 4 * j, *vj = s->v + 4 * j;
      float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
      float ux = vj[0] - vi[0], uy = vj[1] - vi[1], uz = vj[2] - vi[2];
      float vol = s->m * (1.0f / xj[3]);
      float f = k_spiky_f(s, dist2(dx, dy, dz));
      float gx = f * dx, gy = f * dy, gz = f * dz;
      float cx = fmaf(uy, gz, -(uz * gy)), cy = fmaf(uz, gx, -(ux * gz)), cz = fmaf(ux, gy, -(uy * gx));
      wx = fmaf(vol, cx, wx); wy = fmaf(vol, cy, wy); wz = fmaf(vol, cz, wz);
      nx = fmaf(vol, gx, nx); ny = fmaf(vol, gy, ny); nz = fmaf(vol, gz, nz);
    }
    s->om[4 * i] = wx; s->om[4 * i + 1] = wy; s->om[4 * i + 2] = wz;
    s->om[4 * i + 3] = sqrtf(dist2(wx, wy, wz));
    s->nrm[4 * i] = s->h * nx; s->nrm[4 * i + 1] = s->h * ny; s->nrm[4 * i + 2] = s->h * nz;
    s->nrm[4 * i + 3] = 0.0f;
  }
  /* K14 XSPH + vorticity force + surface tension (XSPH.ts:31-91) */
  float kx = (s->xsph * inv_dt) * s->m;
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = s->x + 4 * i, *vi = s->v + 4 * i, *ni = s->nrm + 4 * i, *wi = s->om + 4 * i;
    float rhoi = xi[3];
    float ux = 0.0f, uy = 0.0f, uz = 0.0f, ex = 0.0f, ey = 0.0f, ez = 0.0f, tx = 0.0f, ty = 0.0f, tz = 0.0f;
    for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
      uint32_t j = s->nbr[q];
      const float *xj = s->x + 4 * j, *vj = s->v + 4 * j, *nj = s->nrm + 4 * j;
      float wj = s->om[4 * j + 3];
      float rhoj = xj[3];
      float dx = xi[0] - xj[0], dy = xi[1] - xj[1], dz = xi[2] - xj[2];
      float d2 = dist2(dx, dy, dz);
      float vx = vj[0] - vi[0], vy = vj[1] - vi[1], vz = vj[2] - vi[2];
      float irj = 1.0f / rhoj;
      float K = s->two_rho0 / (rhoi + rhoj);
      float sw = k_poly6(s, d2) * irj;
      ux = fmaf(vx, sw, ux); uy = fmaf(vy, sw, uy); uz = fmaf(vz, sw, uz);
      float se = k_spiky_f(s, d2) * (wj * irj); /* |omega_j| / rho_j formed per particle */
      ex = fmaf(se, dx, ex); ey = fmaf(se, dy, ey); ez = fmaf(se, dz, ez);
      float mc = s->m * k_cohesion_f(s, d2);
      tx = fmaf(-K, fmaf(mc, dx, ni[0] - nj[0]), tx);
      ty = fmaf(-K, fmaf(mc, dy, ni[1] - nj[1]), ty);
      tz = fmaf(-K, fmaf(mc, dz, ni[2] - nj[2]), tz);
    }
    float el = sqrtf(dist2(ex, ey, ez));
    float hx = 0.0f, hy = 0.0f, hz = 0.0f;
    if (el > 0.0f) {
      float ie = 1.0f / el;
      hx = ex * ie; hy = ey * ie; hz = ez * ie;
    }
    float cx = fmaf(hy, wi[2], -(hz * wi[1])), cy = fmaf(hz, wi[0], -(hx * wi[2])), cz = fmaf(hx, wi[1], -(hy * wi[0]));
    s->a[4 * i] = fmaf(s->tension, tx, fmaf(s->vort, cx, kx * ux));
    s->a[4 * i + 1] = fmaf(s->tension, ty, fmaf(s->vort, cy, kx * uy));
    s->a[4 * i + 2] = fmaf(s->tension, tz, fmaf(s->vort, cz, kx * uz));
    s->a[4 * i + 3] = 0.0f;
  }
}

/* ------------------------------------------------------------------ */
/* Accessors.                                                          */
void orc_set_velocity(orc_sim *s, const float *v4) {
  for (int i = 0; i < s->n; ++i) {
    for (int c = 0; c < 3; ++c) s->v[4 * i + c] = v4[4 * i + c];
    s->v[4 * i + 3] = 0.0f;
  }
}

void orc_get(const orc_sim *s, int field, float *out4) {
  const float *src = NULL;
  size_t n4 = (size_t)s->n * 4;
  switch (field) {
    case ORC_POSITION: src = s->x; break;
    case ORC_VELOCITY: src = s->v; break;
    case ORC_ACCEL: src = s->a; break;
    case ORC_LAMBDA: src = s->lam; break;
    case ORC_OMEGA: src = s->om; break;
    case ORC_NORMAL: src = s->nrm; break;
    case ORC_BOUNDARY: src = s->bnd; break;
    case ORC_PREDICTED: src = s->xs; break;
    default: memset(out4, 0, n4 * 4); return;
  }
  memcpy(out4, src, n4 * 4);
}

int orc_count(const orc_sim *s) { return s->n; }
uint32_t orc_num_cells(const orc_sim *s) { return s->ncells; }
void orc_grid_dims(const orc_sim *s, int dims[3]) { dims[0] = s->dims[0]; dims[1] = s->dims[1]; dims[2] = s->dims[2]; }
void orc_get_cells(const orc_sim *s, uint32_t *counts, uint32_t *offsets) {
  memcpy(counts, s->cnt, (size_t)s->ncells * 4);
  memcpy(offsets, s->off, ((size_t)s->ncells + 1) * 4);
}
uint64_t orc_num_pairs(const orc_sim *s) { return s->npairs; }
void orc_get_neighbors(const orc_sim *s, uint32_t *nbr_off, uint32_t *nbr_ids) {
  memcpy(nbr_off, s->noff, ((size_t)s->n + 1) * 4);
  memcpy(nbr_ids, s->nbr, s->npairs * 4);
}

/* ------------------------------------------------------------------ */
/* FP64 density-constraint error.                                      */
void orc_density_error(orc_sim *s, double *mean, double *maxv) {
  int n = s->n;
  double *ci = (double *)malloc((size_t)(n > 0 ? n : 1) * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; ++i) {
    const float *xi = s->x + 4 * i;
    double rho = 0.0;
    for (uint32_t q = s->noff[i]; q < s->noff[i + 1]; ++q) {
      const float *xj = s->x + 4 * s->nbr[q];
      double dx = (double)xi[0] - (double)xj[0], dy = (double)xi[1] - (double)xj[1], dz = (double)xi[2] - (double)xj[2];
      double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 <= s->h2_64) {
        double t = s->h2_64 - d2;
        rho += s->Kp64 * ((t * t) * t);
      }
    }
    rho *= s->V64;
    float b[4];
    boundary_sample(s, xi, b);
    if (b[3] > 0.0f) {
      double dx = (double)xi[0] - (double)b[0], dy = (double)xi[1] - (double)b[1], dz = (double)xi[2] - (double)b[2];
      double d2 = (dx * dx + dy * dy) + dz * dz;
      if (d2 <= s->h2_64) {
        double t = s->h2_64 - d2;
        rho += (double)b[3] * (s->Kp64 * ((t * t) * t));
      }
    }
    double c = rho - 1.0;
    ci[i] = c > 0.0 ? c : 0.0;
  }
  double sum = 0.0, mx = 0.0;
  for (int i = 0; i < n; ++i) {
    sum += ci[i];
    if (ci[i] > mx) mx = ci[i];
  }
  *mean = n > 0 ? sum / (double)n : 0.0;
  *maxv = mx;
  free(ci);
}
