// pbf_internal.h -- launchers shared by the engine and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>

#include "pbf_device.h"

namespace pbf {

// Device view of the warm-start MLP (bf16 MFMA operands + f32 vectors).
struct MlpDev {
  const short *W1T, *W2T, *W2N;
  const float *b1, *b2, *w3, *misc;  // misc: b3, fmu[16], fsig[16], tmu, tsig
};

// Extra per-slot arrays of the learned / data-generation modes.
struct LearnBufs {
  float *lam_acc;        // sum of the lambdas applied this substep (current order)
  const float *acc_prev; // previous substep's sum, reordered
  const float *lam_prev; // previous substep's last lambda, reordered
  const float4 *vel_prev;// previous substep's velocity, reordered
  float *feats;          // 16 features per slot
};

// Everything a pass launch needs besides its buffers.
struct LaunchCtx {
  hipStream_t stream;
  uint32_t n;
  bool multi;   // more than one scene: per-slot scene lookup
  bool hash;    // spatial-hash grid
  bool field;   // at least one scene has a boundary field
  int fmode;    // 0 no field, 1 sample in a separate pass (cubic present), 2 all trilinear
  Scene s0;     // the scene when !multi
  const Scene *scenes;
  const uint32_t *sid;
  const float *pool;
  Phys P;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;  // pass timing: start / stop on the dispatch packets
  uint32_t bs = 256;   // work-group size of the neighbour walks (64..256, a multiple of 64)
  uint32_t lvar = 0;   // neighbour-list kernel variant (LDS rows x work-group size)
  uint32_t red_k = 1;  // blocks per scene of the per-scene reductions
  double *partial = nullptr;  // their stage-1 partials: up to 6 x scenes x red_k doubles
};

constexpr uint32_t kListVariants = 4;
constexpr uint32_t kMaxReduceBlocks = 64;  // blocks per scene of a per-scene reduction (at most one wavefront)

// flags[0] bit 1: a predicted position was not finite
void launch_predict_bin(const LaunchCtx &L, const float4 *pos, const float4 *vel, const float4 *acc, float4 *xs,
                        uint32_t *cellkey, uint32_t *rank, uint32_t *counts, uint32_t *flags);
// zero_in: clear `in` (and `zero2`, when not null, over the same n words)
void launch_scan(hipStream_t st, uint32_t *in, uint32_t n, uint32_t *out, int zero_in, uint32_t *const *sums,
                 uint32_t *const *prefix, int level, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr,
                 uint32_t *zero2 = nullptr);
void launch_scatter(const LaunchCtx &L, const uint32_t *cellkey, const uint32_t *rank, const uint32_t *offsets,
                    const uint32_t *ids, uint32_t *tmp_key, uint32_t *tmp_src);
// dest[s]: the final slot of sorted slot s (read by launch_settle_extra)
void launch_settle(const LaunchCtx &L, const uint32_t *tmp_key, const uint32_t *tmp_src, const uint32_t *cellkey,
                   const uint32_t *offsets, const float4 *xs, const float4 *pos, const float *lam_in, float4 *X0,
                   float4 *Xold, uint32_t *ids, float *lam_out, uint32_t *dest);
void launch_build_list(const LaunchCtx &L, const float4 *X0, const uint32_t *offsets, uint32_t *nbr, uint32_t *ncount,
                       uint32_t *flags);
void launch_lambda(const LaunchCtx &L, float4 *X, float4 *B, const uint32_t *nbr, const uint32_t *ncount,
                   float *lam_acc = nullptr, bool project = false);
void launch_lambda_features(const LaunchCtx &L, float4 *X, float4 *B, const uint32_t *nbr, const uint32_t *ncount,
                            const LearnBufs &lb);
void launch_settle_extra(const LaunchCtx &L, const uint32_t *tmp_src, const uint32_t *dest, const float *acc_in,
                         float *acc_out, const float4 *vel_in, float4 *vel_out);
int mlp_engine_set(void **handle, const float *blob, hipStream_t st);
int mlp_engine_relax(void *handle, float relax, hipStream_t st);
void mlp_engine_free(void **handle);
void launch_mlp_infer(hipStream_t st, void *handle, uint32_t n, const float *feats, float4 *X, float *lam_acc,
                      float smax, hipEvent_t e0, hipEvent_t e1);
void launch_copy(hipStream_t st, void *dst, const void *src, size_t bytes);
void launch_energy(const LaunchCtx &L, const float4 *pos, const float4 *vel, const uint32_t *ranges,
                   uint32_t nscenes, double *out);
void launch_fallback_gate(hipStream_t st, const double *err, const double *speed, const double *e0, const double *e1,
                          uint32_t nscenes, uint32_t n, const double *threshold, uint32_t *decision,
                          uint32_t *count);
void launch_smooth_warm(const LaunchCtx &L, float4 *X, const uint32_t *nbr, const uint32_t *ncount, float *tmp,
                        float *lam_acc);
void launch_warm_previous(const LaunchCtx &L, const float *acc_prev, float4 *X, float *lam_acc, float relax);
// A learned substep's check figures in one pass per scene: (sum, max) of ci
// into err, (sum, max) of |v| dt / h into speed, the energy sums into energy.
void launch_check_reduce(const LaunchCtx &L, const double *ci, const float4 *pos, const float4 *vel,
                         const uint32_t *ranges, uint32_t nscenes, double *err, double *speed, double *energy);
void launch_dp(const LaunchCtx &L, const float4 *X, const float4 *B, const uint32_t *nbr, const uint32_t *ncount,
               float4 *Xn);
// ci: when not null, also the FP64 density-constraint error of every particle
// (the diagnostic's per-particle figure; launch_diag_reduce then reduces it).
// Not with cubic fields (fmode 1): the 32-node element's registers would
// cost the walk its occupancy, so those keep launch_diag's own walk.
void launch_density(const LaunchCtx &L, const float4 *X, const float4 *Xold, const uint32_t *nbr,
                    const uint32_t *ncount, float4 *pos, float4 *vel, float *lam, double *ci = nullptr);
// visc: the XSPH velocity sum, written by the vorticity pass, read by XSPH
void launch_vorticity(const LaunchCtx &L, const float4 *pos, const float4 *vel, const uint32_t *nbr,
                      const uint32_t *ncount, float4 *om, float4 *nrm, float4 *visc);
void launch_xsph(const LaunchCtx &L, const float4 *pos, const float4 *om, const float4 *nrm, const float4 *visc,
                 const uint32_t *nbr, const uint32_t *ncount, float4 *acc);
void launch_diag(const LaunchCtx &L, const float4 *pos, const uint32_t *nbr, const uint32_t *ncount, double *ci,
                 const uint32_t *ranges, uint32_t nscenes, double *per_scene, double *hist, uint32_t hist_cap,
                 uint32_t *hist_ctr);
// The reduction half of launch_diag, over a ci the density pass already
// wrote (launch_density with ci), with the same history entry.
void launch_diag_reduce(const LaunchCtx &L, const double *ci, const uint32_t *ranges, uint32_t nscenes,
                        double *per_scene, double *hist, uint32_t hist_cap, uint32_t *hist_ctr);
// Appends the mean / max over the per-scene totals of an earlier launch_diag
// to the history (the same entry launch_diag with a history writes).
void launch_diag_history(const LaunchCtx &L, const double *per_scene, uint32_t nscenes, double *hist,
                         uint32_t hist_cap, uint32_t *hist_ctr);
void launch_unsort4(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float4 *dst, int clear_w);
void launch_unsort1(hipStream_t st, uint32_t n, const float *src, const uint32_t *ids, float4 *dst);
void launch_sort4(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float4 *dst, int clear_w);
void launch_sort1(hipStream_t st, uint32_t n, const float4 *src, const uint32_t *ids, float *dst);
void launch_fill_u32(hipStream_t st, uint32_t *p, uint32_t n, uint32_t v);
void launch_iota(hipStream_t st, uint32_t *p, uint32_t n);
void launch_copy_probe(hipStream_t st, const void *src, void *dst, size_t bytes);
void launch_check_sqrt(hipStream_t st, uint32_t lo, uint32_t hi, unsigned long long *bad);

// Message of the last failed context-free call on this thread
// (pbf_last_error(NULL)).
void set_thread_error(const std::string &msg);

// Host utilities (pbf_host.cpp).
uint64_t splitmix64(uint64_t x);
int parse_cdm(const char *text, size_t len, int res[3], float **sdf, float **vol, size_t *flen);
size_t cubic_field_length(const int res[3]);

}  // namespace pbf
