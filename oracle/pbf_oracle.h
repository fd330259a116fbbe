/*
 * pbf_oracle.h -- CPU oracle for the Position Based Fluids substep.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, includes
 * or calls this code.  It is used by tests/, by __graft_entry__.smoke() to check
 * the GPU result, and by bench.py's cpu_baseline leg.  It is a plain-C restatement of
 * the reference step (LinzhouLi/WebGPU-Fluid-Simulation):
 *
 *   orchestration     src/simulator/PBF/PBF.ts:398-448
 *   predict           src/simulator/PBF/shader/timeIntegration.ts:20-31
 *   wall clamp        src/simulator/PBF/shader/common.ts:43-68
 *   kernels           src/simulator/PBF/shader/common.ts:1-40
 *   binning / sort    src/simulator/neighbor/neighborShader.ts:12-61
 *   scan              src/simulator/neighbor/exclusiveScan.ts:134-166
 *   neighbor lists    src/simulator/neighbor/neighborShader.ts:64-247
 *   boundary sample   src/simulator/PBF/shader/boundaryVolume.ts:26-57,
 *                     src/simulator/boundary/discreteFieldShader.ts:15-191
 *   lambda / dp       src/simulator/PBF/shader/pressureConstrain.ts:31-179
 *   density+velocity  src/simulator/PBF/shader/attributeUpdate.ts:25-59
 *   vorticity         src/simulator/PBF/shader/vorticityConfinement.ts:25-67
 *   XSPH + tension    src/simulator/PBF/shader/XSPH.ts:31-92
 *
 * Particles stay in input (id) order; the neighbour list of particle i walks
 * the 27 cells x-major (dx outer, dz inner) and, inside a cell, by z-slab
 * (floor(S * frac(z / h)), S = cell_slabs), then ascending id.  The reference
 * leaves the order inside a cell to its atomic ranks (neighborShader.ts:30),
 * so any fixed order is as faithful; this one lets the GPU list build skip
 * the far slabs of the cells above and below.
 * Every floating-point expression is written out operation by operation and
 * the file is compiled with -ffp-contract=off, so results are reproducible.
 */
#ifndef PBF_ORACLE_H
#define PBF_ORACLE_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_sim orc_sim;

typedef struct {
  double particle_radius;   /* r */
  double kernel_radius;     /* h */
  double rest_density;      /* rho0 */
  double domain_lo[3];
  double domain_hi[3];
  int grid_mode;            /* 0 dense, 1 spatial hash */
  int hash_log2;            /* table size 2^hash_log2 when grid_mode == 1 */
  int cell_slabs;           /* z-slabs per cell ordering a cell's particles: 1, 2 or 4 (0 -> 4) */
} orc_params;

/* field ids for orc_get */
enum {
  ORC_POSITION = 0,   /* x, y, z, density (kg/m^3) */
  ORC_VELOCITY = 1,
  ORC_ACCEL = 2,
  ORC_LAMBDA = 3,     /* lambda of the last Jacobi iteration, in .x */
  ORC_OMEGA = 4,      /* omega xyz, |omega| */
  ORC_NORMAL = 5,
  ORC_BOUNDARY = 6,   /* x_b xyz, V_b (last sample of the substep) */
  ORC_PREDICTED = 7   /* frozen predicted position x*_0 of the last substep */
};

orc_sim *orc_create(const orc_params *p, int n, const float *xyzw);
void orc_destroy(orc_sim *s);
void orc_set_config(orc_sim *s, float xsph, float vorticity, float tension, float gravity);
/* Cubic 32-node serendipity field over the box [lo, hi] (the .cdm layout of
 * src/simulator/boundary/discreteFieldShader.ts:1-12).  quirk=1 reproduces the
 * volume-edge indexing of discreteFieldShader.ts:146-153. */
int orc_set_cubic_field(orc_sim *s, const int res[3], const double lo[3], const double hi[3],
                        const float *sdf, const float *vol, int quirk);
/* Trilinear node grid (res+1)^3, x fastest, interleaved (sdf, vol). */
int orc_set_trilinear_field(orc_sim *s, const int res[3], const double lo[3], const double hi[3],
                            const float *sdf_vol);
void orc_clear_field(orc_sim *s);

/* One substep of length dt with `iterations` Jacobi iterations. */
void orc_substep(orc_sim *s, float dt, int iterations);
/* Predict + neighbour search only (state is not advanced). */
void orc_search_only(orc_sim *s, float dt);

void orc_get(const orc_sim *s, int field, float *out4);
/* Sets every particle's velocity (xyz of n float4 rows, id order): the
 * state the next substep's predict (timeIntegration.ts:20-31) starts from. */
void orc_set_velocity(orc_sim *s, const float *v4);
int orc_count(const orc_sim *s);
uint32_t orc_num_cells(const orc_sim *s);
void orc_grid_dims(const orc_sim *s, int dims[3]);
/* Search results of the last substep: counts[C], offsets[C+1], nbr_off[n+1]. */
void orc_get_cells(const orc_sim *s, uint32_t *counts, uint32_t *offsets);
uint64_t orc_num_pairs(const orc_sim *s);
void orc_get_neighbors(const orc_sim *s, uint32_t *nbr_off, uint32_t *nbr_ids);

/* FP64 density-constraint error: mean_i max(0, rho~_i - 1) and its max, on the
 * current positions, the last substep's neighbour sets and the boundary term. */
void orc_density_error(orc_sim *s, double *mean, double *maxv);

/* Boundary sampler alone: x (n x 3) -> (x_b, V_b) (n x 4), and raw field values. */
void orc_boundary_sample(const orc_sim *s, int n, const float *x3, float *out4);
void orc_field_eval(const orc_sim *s, int n, const float *x3, float *out5); /* grad xyz, sdf, vol */

/* Kernel functions for closed-form tests (d2 = |x_ij|^2). */
void orc_kernels(const orc_sim *s, int n, const float *d2, float *poly6, float *spiky_f, float *cohesion_f);
float orc_sinf(float x);
void orc_clamp(const orc_sim *s, int n, const float *x3, float *out3);

/* ------------------------------------------------------------------ */
/* MLS-MPM (mpm_oracle.c; src/simulator/MPM/MPMShader.ts, MPM.ts).      */
typedef struct orc_mpm orc_mpm;
typedef struct {
  int n_grid;       /* grid nodes per axis over the unit box (reference 32) */
  unsigned bound;   /* boundary layer in nodes (reference 3) */
  double dt;        /* reference 8e-4 */
  double rho;       /* particle density (reference 1) */
  double E;         /* Young's modulus (reference 400) */
  double gravity;   /* magnitude, applied as (0, -gravity, 0) */
} orc_mpm_params;
enum { ORC_MPM_POSITION = 0, ORC_MPM_VELOCITY = 1 /* v xyz, J in w */, ORC_MPM_AFFINE = 2 /* 9 per particle */,
       ORC_MPM_GRID = 3 /* n_grid^3 x 4 grid velocity */ };
orc_mpm *orc_mpm_create(const orc_mpm_params *p, int n, const float *xyzw);
void orc_mpm_destroy(orc_mpm *m);
void orc_mpm_step(orc_mpm *m);
void orc_mpm_get(const orc_mpm *m, int field, float *out);
void orc_mpm_set(orc_mpm *m, int field, const float *in);
/* P2G and the grid pass only (J advanced, grid velocity written); raw gets
 * the node sums (momentum xyz, mass) before the grid pass. */
void orc_mpm_p2g(orc_mpm *m, float *raw);
/* G2P only, on the grid velocity set with orc_mpm_set(ORC_MPM_GRID). */
void orc_mpm_g2p_only(orc_mpm *m);

#ifdef __cplusplus
}
#endif
#endif
