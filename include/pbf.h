/*
 * pbf.h -- C ABI of the MI355X Position Based Fluids engine (libpbf.so).
 *
 * Plain C: pointers, sizes and status codes only.  Every entry point stands
 * in for a member of the reference simulator class (LinzhouLi/WebGPU-Fluid-
 * Simulation, src/simulator/SPH.ts, PBF/PBFConfig.ts, PBF/PBF.ts); the
 * citation above each declaration names the member it replaces.  The
 * Node-API addon (node/pbf_addon.cc) and the Python host mirror (pbf_host.py)
 * bind exactly these symbols.
 *
 * Conventions
 *   - return value: PBF_OK (0) or a negative PBF_E_* code; pbf_last_error()
 *     returns the message (the reference's own text where it has one).
 *   - host pointers are borrowed for the duration of the call only.
 *   - a context is single-threaded and owns one HIP stream; pbf_step is
 *     asynchronous on that stream, every pbf_read_* synchronizes.
 *   - particle arrays cross the boundary in the caller's (original) order as
 *     float4 rows (x, y, z, w) like the reference's 16-byte-stride buffers.
 */
#ifndef PBF_H
#define PBF_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define PBF_ABI_VERSION 1

enum {
  PBF_OK = 0,
  PBF_E_INVALID = -1,   /* bad argument */
  PBF_E_CAPACITY = -2,  /* "To Many Particles!" and other capacity limits */
  PBF_E_LENGTH = -3,    /* "Illegal length of Particle Position Buffer!" */
  PBF_E_FORMAT = -4,    /* "Invalid Boundary Model File!" */
  PBF_E_DEVICE = -5,    /* HIP runtime error */
  PBF_E_OVERFLOW = -6,  /* a particle had more neighbours than max_neighbors, or a non-finite position */
  PBF_E_STATE = -7      /* call out of order (e.g. no particles yet) */
};

enum { PBF_GRID_DENSE = 0, PBF_GRID_HASH = 1 };
enum { PBF_FIELD_NONE = 0, PBF_FIELD_CUBIC = 1, PBF_FIELD_TRILINEAR = 2 };

typedef struct pbf_ctx pbf_ctx;

/* Construction parameters.  Reference: the constants of SPH.ts:24-27,33 and
 * PBFConfig.ts:21 (r = 0.006, h = 0.025, rho0 = 1000, dt = 1/300,
 * MAX_PARTICLE_NUM, MAX_NEIGHBOR_COUNT = 60) and the fixed unit-box domain of
 * PBF/shader/common.ts:43-68, all made runtime values. */
typedef struct {
  uint32_t capacity;        /* max particles over all scenes (SPH.MAX_PARTICLE_NUM), <= 2^27 - 256 */
  uint32_t max_neighbors;   /* per-particle neighbour capacity (SPH.MAX_NEIGHBOR_COUNT); 0 -> 96, < 65536 */
  double particle_radius;   /* r */
  double kernel_radius;     /* h; 0 -> 4 r */
  double rest_density;      /* rho0; 0 -> 1000 */
  double time_step;         /* dt used by pbf_run (SPH.TIME_STEP); 0 -> 1/300 */
  double domain_lo[3];      /* wall-clamp box; reference: (0,0,0) */
  double domain_hi[3];      /* reference: (1,1,1) */
  uint32_t grid_mode;       /* PBF_GRID_DENSE (cell = h) or PBF_GRID_HASH */
  uint32_t hash_log2;       /* hash table size 2^hash_log2 (PBF_GRID_HASH) */
  uint32_t flags;           /* PBF_FLAG_* */
  uint32_t cell_slabs;      /* 1, 2 or 4 (0 -> 4): a cell's particles are kept ordered by their
                               z-slab of height h / cell_slabs, then by id; the neighbour list
                               follows that order (the reference's order inside a cell is its
                               atomic ranks, neighborShader.ts:30) and its build skips the slabs
                               of the cells above and below that lie out of reach */
  uint32_t reserved[4];
} pbf_params;

enum {
  PBF_FLAG_NO_GRAPH = 1u << 0,      /* launch passes directly instead of replaying a hipGraph */
  PBF_FLAG_CUBIC_FIXED = 1u << 1,   /* cubic volume map: read the corrected edge nodes */
  PBF_FLAG_TIMING = 1u << 2         /* record per-pass HIP events (pbf_pass_times) */
};

/* Simulation options.  Reference: PBFConfig.setConfig (PBFConfig.ts:31-46)
 * writing the SimulationOptions uniform (src/common/shader.ts:50-57). */
typedef struct {
  uint32_t iteration;       /* Jacobi iterations per substep, <= 4096 */
  float xsph;               /* XSPH coefficient */
  float vorticity;          /* vorticity confinement epsilon */
  float surface_tension;    /* Akinci surface-tension gamma */
  float gravity;            /* magnitude; applied as (0, -gravity, 0) */
} pbf_config;

/* One scene of a batched context (no reference counterpart: the reference
 * runs a single scene in a fixed unit box). */
typedef struct {
  double domain_lo[3];
  double domain_hi[3];
  uint32_t particle_count;  /* particles of this scene, in scene order */
  uint32_t reserved[3];
} pbf_scene_desc;

/* Reference: new PBF() + PBF.initResource (PBF.ts:67-69,262-279). */
int pbf_create(const pbf_params *params, pbf_ctx **out);
/* Batched variant: `nscenes` independent scenes share every launch. */
int pbf_batch_create(const pbf_params *params, uint32_t nscenes, const pbf_scene_desc *scenes, pbf_ctx **out);
int pbf_destroy(pbf_ctx *ctx);
/* Reference: PBF.reset (PBF.ts:71-89) -> SPH.baseReset (SPH.ts:57-66). */
int pbf_reset(pbf_ctx *ctx);
/* Reference: PBFConfig.setConfig (PBFConfig.ts:31-46). */
int pbf_set_config(pbf_ctx *ctx, const pbf_config *cfg);

/* Reference: SPH.voxelizeCube (SPH.ts:99-121): appends a lattice of spacing
 * 2r to the context's staged particle array (f64 arithmetic, then f32).
 * jitter adds a seeded uniform offset in [-jitter, +jitter] per axis (0 = off). */
int pbf_voxelize_cube(pbf_ctx *ctx, const double min[3], const double max[3], double jitter, uint64_t seed,
                      uint32_t *added);
/* Reference: SPH.voxelizeSphere (SPH.ts:123-149). */
int pbf_voxelize_sphere(pbf_ctx *ctx, const double center[3], double radius, double jitter, uint64_t seed,
                        uint32_t *added);
/* GPU-free forms of the two voxelizers: write float4 rows (w = 0) for the
 * particles serial0, serial0+1, ... (the jitter stream index) into `out`
 * (NULL: count only). */
int pbf_voxelize_cube_host(double particle_radius, const double min[3], const double max[3], double jitter,
                           uint64_t seed, uint64_t serial0, float *out, uint64_t capacity, uint32_t *count);
int pbf_voxelize_sphere_host(double particle_radius, const double center[3], double radius, double jitter,
                             uint64_t seed, uint64_t serial0, float *out, uint64_t capacity, uint32_t *count);
/* Reference: SPH.voxelizeMesh (SPH.ts:151-187): the 2r lattice over the mesh's
 * world-space bounding box, keeping points whose ray along the mesh's local
 * +z axis first hits a triangle facing along the ray.  vertices: nverts xyz
 * (mesh-local), indices: 3 per triangle, world: 4x4 column-major
 * (three.js Matrix4.elements) or NULL for the identity. */
int pbf_voxelize_mesh(pbf_ctx *ctx, const float *vertices, uint32_t nverts, const uint32_t *indices, uint32_t ntris,
                      const double world[16], double jitter, uint64_t seed, uint32_t *added);
int pbf_voxelize_mesh_host(double particle_radius, const float *vertices, uint32_t nverts, const uint32_t *indices,
                           uint32_t ntris, const double world[16], double jitter, uint64_t seed, uint64_t serial0,
                           float *out, uint64_t capacity, uint32_t *count);
/* Number of staged particles (SPH.particleCount before upload). */
uint32_t pbf_staged_count(const pbf_ctx *ctx);
/* Reference: PBFConfig.setParticlePosition (PBFConfig.ts:48-62): uploads the
 * staged particles.  Errors: 'To Many Particles!'. */
int pbf_set_particle_position(pbf_ctx *ctx);
/* Upload an explicit float4 array (xyzw, w ignored) of `count` particles;
 * `length` is the number of floats.  Errors: 'Illegal length of Particle
 * Position Buffer!' (length != 4*count), 'To Many Particles!'. */
int pbf_set_positions(pbf_ctx *ctx, const float *xyzw, size_t length, uint32_t count);

/* Reference: PBF.setBoundaryData (PBF.ts:91-94) -> BoundaryModel.setData
 * (volumeMap.ts:23-39): parse .cdm text (SDF line, volume line) of a cubic
 * 32-node field over the unit box.  Error: 'Invalid Boundary Model File!'. */
int pbf_set_boundary_data(pbf_ctx *ctx, const char *text, size_t length);
/* GPU-free parse of .cdm text: resolution and field length first (sdf/vol
 * may be NULL), then the values when capacity >= length. */
int pbf_parse_cdm(const char *text, size_t length, int32_t res[3], float *sdf, float *vol, size_t capacity,
                  size_t *field_length);
/* Cubic 32-node field over [lo, hi] from arrays in the .cdm layout. */
int pbf_set_cubic_field(pbf_ctx *ctx, uint32_t scene, const int32_t res[3], const double lo[3], const double hi[3],
                        const float *sdf, const float *vol, size_t length);
/* Trilinear volume map: (res+1)^3 nodes over [lo, hi], x fastest, (sdf, vol)
 * interleaved.  No reference counterpart (the reference ships 20^3 cubic maps). */
int pbf_set_volume_map(pbf_ctx *ctx, uint32_t scene, const int32_t res[3], const double lo[3], const double hi[3],
                       const float *sdf_vol);
int pbf_clear_boundary(pbf_ctx *ctx);

/* OR'ed into the `kind` of the map builders below: store the reference's
 * own boundary volume instead of this library's definition -- the one its
 * .cdm files under public/model/ hold (read by volumeMap.ts:23-39, sampled by
 * boundaryVolume.ts:40-51): V(x) = 0.8 sum over the 16^3 Gauss-Legendre
 * points xi of [-h, h]^3 with |xi| <= h of w(xi) gamma(sdf(x + xi)), gamma = 1
 * in the solid, the cubic-spline ratio W(d) / W(0) of support h for
 * 0 < d < h, 0 beyond and for points outside the field's box [lo, hi]
 * (Bender et al. 2019 volume maps).  particle_radius is not used. */
#define PBF_VOLUME_REFERENCE 0x100u
/* Host-side map builder for analytic solids: kind 0 = inside of the box
 * [lo, hi] is fluid (walls outside; shape, if given, = box lo[3], hi[3]),
 * kind 1 = that box plus a sphere and a torus (shape[0..2] sphere centre,
 * shape[3] radius, shape[4..6] torus centre, shape[7] major R, shape[8]
 * minor r; torus axis = y).  Writes (res+1)^3 (sdf, vol) pairs, sdf > 0 in
 * the fluid.  vol: with particle_radius > 0 the boundary volume whose mirrored
 * sample at max(d + r/2, 2r) reproduces 0.9 x the poly6-weighted solid part
 * of the kernel ball; with particle_radius == 0 the solid volume in the ball.
 * No reference counterpart: the reference ships precomputed .cdm maps. */
int pbf_build_volume_map(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                         const double hi[3], double kernel_radius, double particle_radius, float *sdf_vol);
/* The same definitions sampled at the nodes of the 32-node cubic element in
 * the .cdm layout (corner, x-, y-, z-edge node sets) over [lo, hi]: a
 * reference-format field for any analytic solid and resolution (with
 * PBF_VOLUME_REFERENCE in kind, the reference's own volume values).
 * sdf / vol NULL: field_length only. */
int pbf_build_cubic_field(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                          const double hi[3], double kernel_radius, double particle_radius, float *sdf, float *vol,
                          size_t capacity, size_t *field_length);
/* .cdm text of a cubic field, the format volumeMap.ts:23-39 parses (9
 * significant digits: exact float round trip).  out NULL: size only. */
int pbf_write_cdm(const int32_t res[3], const float *sdf, const float *vol, size_t length, char *out,
                  size_t capacity, size_t *written);
/* The same map built on the GPU (same definition, quadrature and float64
 * arithmetic order, one thread per node); host arrays in and out.
 * PBF_E_DEVICE without a GPU. */
int pbf_build_volume_map_gpu(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                             const double hi[3], double kernel_radius, double particle_radius, float *sdf_vol);

/* Reference: PBF.run (PBF.ts:398-448) encodes one substep with the
 * compile-time dt and the configured iteration count.  pbf_step runs
 * `substeps` substeps of dt/substeps with `iterations` Jacobi iterations each
 * (0 = the configured count, at most 4096), asynchronously on the context
 * stream. */
int pbf_step(pbf_ctx *ctx, float dt, uint32_t substeps, uint32_t iterations);
/* Exactly PBF.run: one substep of params.time_step, configured iterations. */
int pbf_run(pbf_ctx *ctx);
int pbf_sync(pbf_ctx *ctx);

/* Reference: the public `position` buffer (SPH.ts:39) read by the renderer:
 * float4 (x, y, z, density) in the original particle order. */
int pbf_read_positions(pbf_ctx *ctx, float *xyzw, uint32_t count);
/* Reference: the protected `velocity` buffer (SPH.ts:40); float4, w = 0. */
int pbf_read_velocities(pbf_ctx *ctx, float *xyzw, uint32_t count);

enum {
  PBF_FIELD_POSITION = 0, PBF_FIELD_VELOCITY = 1, PBF_FIELD_ACCEL = 2, PBF_FIELD_LAMBDA = 3,
  PBF_FIELD_OMEGA = 4, PBF_FIELD_NORMAL = 5, PBF_FIELD_BOUNDARY = 6, PBF_FIELD_PREDICTED = 7,
  PBF_FIELD_LAMBDA_SUM = 8
};
/* Any per-particle float4 state in original order (debug readback of
 * PBF.debug, PBF.ts:514-573). LAMBDA (the last iteration's lambda) and
 * LAMBDA_SUM (the substep's lambda sum that the warm-start modes carry; zero
 * before any warm-start substep) are in .x; NORMAL and VELOCITY have w = 0;
 * PREDICTED is defined after pbf_search only (a step reuses its buffer). */
int pbf_read_field(pbf_ctx *ctx, uint32_t field, float *out4, uint32_t count);
/* Checkpoint upload, the inverse of pbf_read_field for the state a substep
 * carries into the next: VELOCITY and ACCEL (w ignored), LAMBDA and
 * LAMBDA_SUM (.x).  in4: `count` float4 rows in the original particle order,
 * count = the particle count.  A snapshot {positions, VELOCITY, ACCEL, LAMBDA,
 * LAMBDA_SUM} read after a step and written back after pbf_set_positions
 * resumes the run bit for bit.  The reference keeps this state GPU-resident
 * with no upload path (PBF.ts:71-89 resets it); no reference counterpart. */
int pbf_write_field(pbf_ctx *ctx, uint32_t field, const float *in4, uint32_t count);

/* Neighbour search results of the last substep (NeighborSearch.debug,
 * neighborSearch.ts:303-378).  Cells are indexed globally over all scenes.
 * counts/offsets may be NULL.  Neighbour lists are in original particle ids,
 * ordered as walked (27 cells x-major, ascending id inside a cell). */
uint32_t pbf_num_cells(const pbf_ctx *ctx);
/* GPU-free: 1 when, on a hashed grid of dims[0] x dims[1] x dims[2] cells
 * with 2^hash_log2 buckets (bucket = row-major cell index mod 2^hash_log2),
 * the 27 cells around every cell lie in 27 different buckets; 0 otherwise.
 * The neighbour search takes its pruned per-column walk only then, and
 * otherwise visits each repeated bucket once, at its first occurrence in the
 * x-major 27-cell order (the reference's dense grid, neighborShader.ts:64-154,
 * never repeats a cell).  No reference counterpart. */
int pbf_hash_buckets_distinct(const int32_t dims[3], uint32_t hash_log2);
int pbf_read_cells(pbf_ctx *ctx, uint32_t *counts, uint32_t *offsets);
int pbf_read_neighbors(pbf_ctx *ctx, uint32_t *nbr_off, uint32_t *nbr_ids, uint64_t capacity, uint64_t *pairs);
/* Predict + neighbour search only; the state is not advanced. */
int pbf_search(pbf_ctx *ctx, float dt);

/* FP64 density-constraint error of the last substep: mean and max over
 * particles of max(0, rho~ - 1), rho~ with the boundary term; computed and
 * reduced in FP64 on the GPU (no reference counterpart; nearest is K9's C,
 * pressureConstrain.ts:80). Per-scene values when scene < nscenes, else all. */
int pbf_density_error(pbf_ctx *ctx, uint32_t scene, double *mean, double *maxv);
/* Record the error after every substep into a ring of `capacity` entries. */
int pbf_record_density_error(pbf_ctx *ctx, uint32_t enable);
int pbf_read_density_history(pbf_ctx *ctx, double *mean_max, uint32_t capacity, uint32_t *count);

uint32_t pbf_particle_count(const pbf_ctx *ctx);
uint32_t pbf_scene_count(const pbf_ctx *ctx);
/* Message of the context's last failed call; with NULL, the message of the
 * last failed pbf_create / pbf_batch_create / pbf_mlp_forward_backward_noalloc
 * on the calling thread. */
const char *pbf_last_error(const pbf_ctx *ctx);
const char *pbf_version(void);

/* Per-pass GPU time in ms accumulated since the last call (PBF_FLAG_TIMING):
 * names are returned as one '\n'-separated string. */
int pbf_pass_times(pbf_ctx *ctx, double *ms, uint32_t *launches, uint32_t capacity, uint32_t *count,
                   const char **names);

/* Streaming-copy probe: copies `bytes` device-to-device `iterations` times
 * with a plain 16-byte-per-lane kernel and returns read + write GB/s.  The
 * measured HBM ceiling next to the 8 TB/s specification, and the reference
 * kernel for scaling hardware byte counters. No reference counterpart. */
int pbf_copy_bandwidth(size_t bytes, uint32_t iterations, double *gbps);

/* Numerics self-check, no reference counterpart: the pair terms take the
 * square root of a squared distance from the hardware reciprocal square
 * root and one Newton step, which is the correctly rounded square root (as
 * sqrtf) for every float in [2^-100, FLT_MAX] on gfx950.  This runs that
 * sequence on the device over all 1,912,602,624 of those floats and returns
 * how many differ from the correctly rounded value (0 on a device the
 * kernels' bit-exact results hold on) and how many were checked. */
int pbf_check_device_sqrt(uint64_t *failures, uint64_t *checked);

/* ---- Learned warm start of the lambda solve ----------------------------
 * No reference counterpart: the reference solves every substep from
 * lambda = 0 with a fixed iteration count (PBF.ts:419-431).  A 16 -> 64 ->
 * 64 -> 1 SiLU network (bf16 weights on MFMA, fp32 accumulation) predicts,
 * from 16 per-particle features computed in iteration 0, the lambda summed
 * over a long solve; that prediction is applied in iteration 0 and the
 * remaining iterations are the plain Jacobi passes.
 *
 * Network blob: pbf_mlp_blob_size() floats,
 *   W1[16][64] b1[64] W2[64][64] b2[64] w3[64] b3 fmu[16] fsig[16] tmu tsig
 *   lmin
 * (the first PBF_MLP_PARAMS are trained; features are normalised as
 * (f - fmu) / fsig and the output is lambda = max(y * tsig + tmu, lmin);
 * a positive or non-finite output keeps the plain lambda).
 * Features (iteration 0, per particle): 0 rho~-1, 1 C, 2 plain lambda,
 * 3 previous substep's lambda sum, 4 previous substep's last lambda,
 * 5 lambda denominator, 6 sum |grad_j C|^2, 7 |grad_i C|^2, 8 neighbour
 * count, 9 boundary volume / V, 10 boundary distance / h, 11 velocity
 * divergence * dt, 12 |v| dt / h, 13 v_y dt / h, 14 fluid part of rho~-1,
 * 15 boundary part of rho~. */
enum { PBF_MLP_PARAMS = 5313, PBF_MLP_FEATURES = 16 };
/* PBF_WARM_PREVIOUS: iteration 0 applies relax x the previous substep's
 * lambda sum, without a network (the classic temporal warm start).
 * PBF_WARM_RATIO: the network's output (y * tsig + tmu) is a per-particle
 * factor s on that sum (feature 3): iteration 0 applies
 * clamp(relax s, 0, PBF_WARM_RATIO_MAX relax) x Lambda_prev, and lmin is not
 * used.  A network with zero weights and tmu = 1 is exactly
 * PBF_WARM_PREVIOUS. */
enum { PBF_WARM_PLAIN = 0, PBF_WARM_LEARNED = 1, PBF_WARM_PREVIOUS = 2, PBF_WARM_RATIO = 3 };
/* OR'ed into a mode: projected accumulation.  Iterations 1.. take their
 * increment from the unclamped constraint, -(rho~ - 1) / (den + eps), add it
 * to the substep's lambda sum, keep that sum <= 0, and apply the change of
 * the sum (accumulated multipliers clamped as a whole, as in Catto's
 * warm-started sequential impulses).  A warm start that pushed particles
 * apart is pulled back by the later iterations, never past a zero sum.  With
 * PBF_WARM_PLAIN it is the cold-started solve in that form. */
#define PBF_WARM_PROJECT 0x100u
/* OR'ed into a warm mode: before iteration 0 applies it, the warm lambda
 * field is replaced by its poly6-weighted (Shepard) average over each
 * particle's neighbours, so its gradient -- the first displacement -- is
 * free of particle-scale noise. */
#define PBF_WARM_SMOOTH 0x200u
#define PBF_WARM_RATIO_MAX 2.0f

uint32_t pbf_mlp_blob_size(void);
/* Glorot-uniform weights from a SplitMix64 stream, zero biases, identity
 * normalisation, lmin = -3e38 (no bound). */
int pbf_mlp_init(uint64_t seed, float *blob);
/* One forward + backward pass on the GPU over n samples: y (normalised
 * output, n), grad (PBF_MLP_PARAMS, d loss / d parameter) and the loss
 * mean((y - (target - tmu) / tsig)^2).  Any output pointer may be NULL. */
int pbf_mlp_forward_backward(const float *blob, const float *features, const float *targets, uint32_t n, float *y,
                             float *grad, float *loss);
/* pbf_mlp_forward_backward without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight training
 * loops).  No counterpart in the reference.  Every pointer is DEVICE memory
 * of the current GPU: blob (pbf_mlp_blob_size() floats), features (16 n),
 * targets (n); outputs y (n), grad (PBF_MLP_PARAMS), loss (1), each may be
 * NULL.  Enqueued on `stream` (a hipStream_t; NULL = the default stream) and
 * asynchronous: the outputs are ready when the stream reaches that point.
 * The packed bf16 weights and the gradient / loss accumulators live in a
 * fixed 57,344-byte array per GPU that the library holds until the process
 * ends; calls that overlap on one GPU share it and race, so serialise them
 * (one stream, or an event between streams).  n must be in
 * [1, PBF_MLP_NOALLOC_MAX_SAMPLES]; errors are returned as codes with the
 * message in pbf_last_error(NULL). */
#define PBF_MLP_NOALLOC_MAX_SAMPLES 2147483648u
int pbf_mlp_forward_backward_noalloc(const float *blob, const float *features, const float *targets, uint32_t n,
                                     float *y, float *grad, float *loss, void *stream);
/* Adam (beta 0.9 / 0.999, eps 1e-8) over shuffled mini-batches on the GPU;
 * blob holds the initial network (and its normalisation) and receives the
 * trained weights; loss_per_epoch (epochs floats) may be NULL. */
int pbf_mlp_train(const float *features, const float *targets, uint64_t n, uint32_t epochs, uint32_t batch,
                  float lr, uint64_t seed, float *blob, float *loss_per_epoch);
/* pbf_mlp_train with an asymmetric loss: a sample whose prediction is above
 * its target weighs over_weight (> 0) times its squared error, one below it
 * weighs 1.  With over_weight > 1 the network learns a value on the low side
 * of each target's spread (an expectile), which for a factor on a negative
 * lambda sum means a warm start that errs towards pushing too little. */
int pbf_mlp_train_weighted(const float *features, const float *targets, uint64_t n, uint32_t epochs, uint32_t batch,
                           float lr, uint64_t seed, float over_weight, float *blob, float *loss_per_epoch);
/* Selects the solve: PBF_WARM_PLAIN, PBF_WARM_PREVIOUS, or PBF_WARM_LEARNED /
 * PBF_WARM_RATIO with the network in blob (NULL keeps the weights already
 * set).  relax in (0, 1] scales the
 * predicted lambda: the later iterations only remove compression, so an
 * over-predicted lambda pushes particles apart for good (kinetic energy);
 * relax < 1 keeps the warm start on the safe side.  fallback_threshold > 0
 * checks each learned substep on the device, in the substep's own launch
 * sequence, and the host reads the one-word decision after it: when its
 * mean density error is above the
 * threshold, its mechanical energy (kinetic + potential, summed in FP64) rose
 * by more than 0.1 % of the kinetic energy it started from, a particle moved
 * more than 2 h in it, or any of these is not finite, the substep is undone
 * and solved plainly with fallback_iterations (0: the same count).  The
 * rejections are counted on the device (pbf_warm_start_stats). */
int pbf_set_warm_start(pbf_ctx *ctx, const float *blob, size_t length, uint32_t mode, float relax,
                       double fallback_threshold, uint32_t fallback_iterations);
int pbf_warm_start_stats(const pbf_ctx *ctx, uint64_t *learned_substeps, uint64_t *fallbacks);
/* Training pairs along the trajectory of the selected solver: from the
 * current state, a solve of `long_iterations` plain Jacobi iterations gives,
 * per particle (in the caller's order), the 16 features of iteration 0 and
 * the lambda summed over all its iterations (the target); that solve is then
 * undone and the state advances by one substep of the selected solver with
 * `iterations` (0: the configured count), keeping the lambda sums that the
 * next pair's features read. */
int pbf_collect_training(pbf_ctx *ctx, float dt, uint32_t long_iterations, uint32_t iterations, float *features,
                         float *targets, uint32_t capacity, uint32_t *count);

/* ---- Launch tuning -------------------------------------------------------
 * No reference counterpart (WebGPU fixes workgroup sizes in the shaders).
 * The work-group size of each neighbour walk and the neighbour-list variant
 * change only the time, never a result: every particle's arithmetic is the
 * same under any setting. */
typedef struct {
  uint32_t walk_block[8];  /* lambda, dp, density, vorticity, xsph, diag, -, -: 64, 128, 192 or 256 */
  uint32_t list_variant;   /* 0..3: (LDS rows, work-group) = (32,128), (16,128), (32,64), (56,64) */
  uint32_t reserved[7];
} pbf_tuning;
int pbf_get_tuning(const pbf_ctx *ctx, pbf_tuning *out);
int pbf_set_tuning(pbf_ctx *ctx, const pbf_tuning *tuning);
/* Times `repeats` substeps of the current scene (dt, iterations) under each
 * candidate setting, keeps per pass the fastest, and leaves the simulation
 * state as it found it.  out (may be NULL) receives the chosen setting. */
int pbf_autotune(pbf_ctx *ctx, float dt, uint32_t iterations, uint32_t repeats, pbf_tuning *out);

/* ------------------------------------------------------------------------
 * MLS-MPM.  Reference: the MPM simulator class (src/simulator/MPM/MPM.ts,
 * commented out in the reference and imported by nothing) and its shaders
 * (src/simulator/MPM/MPMShader.ts: P2GComputeShader :3-148,
 * GridComputeShader :151-193, G2PComputeShader :196-282).  Particles live in
 * the unit box on an n_grid^3 node grid; one step = P2G, grid pass, G2P.
 * The reference's P2G adds with float atomics (arrival order); here a node
 * sums its particles in ascending (base cell, particle id) order, so every
 * step is deterministic.
 * ------------------------------------------------------------------------ */
typedef struct pbf_mpm pbf_mpm;

typedef struct {
  uint32_t capacity;  /* max particles (the reference's 9000, MPM.ts:64); 0 -> 9000, <= 2^27 */
  uint32_t n_grid;    /* nodes per axis (MPM.ts:62, 32); 0 -> 32, a multiple of 4 in [4, 1024] */
  uint32_t bound;     /* boundary layer in nodes (MPM.ts:69, 3); 0 -> 3, <= n_grid / 2 */
  uint32_t flags;     /* PBF_FLAG_NO_GRAPH: launch the passes directly */
  double dt;          /* MPM.ts:66, 8e-4; 0 -> 8e-4 */
  double rho;         /* particle density (MPM.ts:71, 1); 0 -> 1 */
  double E;           /* Young's modulus (MPM.ts:76, 400); 0 -> 400 */
  double gravity;     /* applied as (0, -gravity, 0); the reference never writes its gravity buffer */
  uint32_t reserved[4];
} pbf_mpm_params;

enum {
  PBF_MPM_POSITION = 0, /* n float4: x, y, z, 0 */
  PBF_MPM_VELOCITY = 1, /* n float4: v, J (the Particle struct's v: vec4, MPMShader.ts:72-75) */
  PBF_MPM_AFFINE = 2,   /* n x 9 floats: C column-major (WGSL mat3x3: column c at 3c) */
  PBF_MPM_GRID = 3      /* n_grid^3 float4, x-major: grid velocity after the last grid pass */
};

/* Reference: new MPM() + MPM.initResource's buffers (MPM.ts:60-83,94-103). */
int pbf_mpm_create(const pbf_mpm_params *params, pbf_mpm **out);
int pbf_mpm_destroy(pbf_mpm *m);
/* Reference: MPM.initResource's position write (MPM.ts:104-121): positions
 * (float4 rows, w ignored), v = 0, C = 0, J = 1.  Errors: 'To Many Particles!'. */
int pbf_mpm_set_positions(pbf_mpm *m, const float *xyzw, uint32_t count);
/* Reference: MPM.run (MPM.ts:204-238), `steps` times; asynchronous. */
int pbf_mpm_step(pbf_mpm *m, uint32_t steps);
int pbf_mpm_sync(pbf_mpm *m);
/* Readback / upload of a PBF_MPM_* field (synchronous); count = particles,
 * or n_grid^3 for the grid. */
int pbf_mpm_read(pbf_mpm *m, int field, float *out, size_t count);
int pbf_mpm_write(pbf_mpm *m, int field, const float *in, size_t count);
/* G2P alone (MPMShader.ts:240-280) on the current grid velocity. */
int pbf_mpm_g2p(pbf_mpm *m);
uint32_t pbf_mpm_particle_count(const pbf_mpm *m);
/* Message of the last failed call on m (m = NULL: of a failed create). */
const char *pbf_mpm_last_error(const pbf_mpm *m);

#ifdef __cplusplus
}
#endif
#endif
