// pbf_addon.cc -- Node-API binding of libpbf's C ABI (include/pbf.h).
//
// A thin layer: every function takes the context handle (a napi external)
// first, forwards to one pbf_* entry point and turns a non-zero status into
// a thrown JS Error carrying pbf_last_error(), the way the reference class
// throws `new Error('...')` (SPH.ts:70-75, PBFConfig.ts:50-55,
// volumeMap.ts:31-33).  Typed arrays are copied in and out; no JS memory is
// retained by the library.
#include <node_api.h>

#include <cstdint>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/pbf.h"

namespace {

#define NAPI_CALL(env, call)                                 \
  do {                                                       \
    if ((call) != napi_ok) {                                 \
      napi_throw_error((env), nullptr, "napi call failed");  \
      return nullptr;                                        \
    }                                                        \
  } while (0)

napi_value undefined(napi_env env) {
  napi_value u;
  napi_get_undefined(env, &u);
  return u;
}

napi_value throw_status(napi_env env, pbf_ctx *c, int rc) {
  std::string msg = c ? pbf_last_error(c) : "pbf error";
  if (msg.empty()) msg = "pbf error " + std::to_string(rc);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

struct Args {
  napi_value v[8];
  size_t n = 8;
};

bool get_args(napi_env env, napi_callback_info info, Args &a) {
  a.n = 8;
  return napi_get_cb_info(env, info, &a.n, a.v, nullptr, nullptr) == napi_ok;
}

pbf_ctx *ctx_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected a PBF context handle");
    return nullptr;
  }
  return static_cast<pbf_ctx *>(p);
}

double num(napi_env env, napi_value v, double dflt = 0.0) {
  double d = dflt;
  napi_valuetype t;
  if (napi_typeof(env, v, &t) == napi_ok && t == napi_number) napi_get_value_double(env, v, &d);
  return d;
}

double prop(napi_env env, napi_value obj, const char *name, double dflt) {
  bool has = false;
  if (napi_has_named_property(env, obj, name, &has) != napi_ok || !has) return dflt;
  napi_value v;
  napi_get_named_property(env, obj, name, &v);
  return num(env, v, dflt);
}

void vec3(napi_env env, napi_value v, double out[3]) {
  // accepts [x, y, z] or {x, y, z} (THREE.Vector3)
  bool is_arr = false;
  napi_is_array(env, v, &is_arr);
  if (is_arr) {
    for (uint32_t k = 0; k < 3; ++k) {
      napi_value e;
      napi_get_element(env, v, k, &e);
      out[k] = num(env, e);
    }
  } else {
    out[0] = prop(env, v, "x", 0.0);
    out[1] = prop(env, v, "y", 0.0);
    out[2] = prop(env, v, "z", 0.0);
  }
}

bool float_array(napi_env env, napi_value v, float **data, size_t *len) {
  napi_typedarray_type t;
  size_t n;
  void *p;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, &n, &p, &ab, &off) != napi_ok || t != napi_float32_array) {
    napi_throw_type_error(env, nullptr, "expected a Float32Array");
    return false;
  }
  *data = static_cast<float *>(p);
  *len = n;
  return true;
}

napi_value new_float_array(napi_env env, size_t n, float **data) {
  napi_value ab, arr;
  void *p = nullptr;
  if (napi_create_arraybuffer(env, n * 4, &p, &ab) != napi_ok) return nullptr;
  if (napi_create_typedarray(env, napi_float32_array, n, ab, 0, &arr) != napi_ok) return nullptr;
  *data = static_cast<float *>(p);
  return arr;
}

// A JS number as uint32; NaN, negative or too large values become 0xffffffff,
// which every uint32 parameter of pbf_params rejects or ignores.
uint32_t as_u32(double v) { return (v >= 0.0 && v <= 4294967295.0) ? (uint32_t)v : 0xffffffffu; }
// A JS number as a uint64 seed: NaN and negative values are 0, huge ones saturate.
uint64_t as_u64(double v) { return v >= 18446744073709549568.0 ? ~0ull : (v >= 0.0 ? (uint64_t)v : 0ull); }

void read_params(napi_env env, napi_value o, pbf_params &p) {
  std::memset(&p, 0, sizeof(p));
  p.capacity = as_u32(prop(env, o, "capacity", 1 << 17));
  p.max_neighbors = as_u32(prop(env, o, "maxNeighbors", 0));
  p.particle_radius = prop(env, o, "particleRadius", 0.006);
  p.kernel_radius = prop(env, o, "kernelRadius", 0.025);
  p.rest_density = prop(env, o, "restDensity", 1000.0);
  p.time_step = prop(env, o, "timeStep", 1.0 / 300.0);
  bool has = false;
  double lo[3] = {0, 0, 0}, hi[3] = {1, 1, 1};
  napi_has_named_property(env, o, "domainLo", &has);
  if (has) { napi_value v; napi_get_named_property(env, o, "domainLo", &v); vec3(env, v, lo); }
  napi_has_named_property(env, o, "domainHi", &has);
  if (has) { napi_value v; napi_get_named_property(env, o, "domainHi", &v); vec3(env, v, hi); }
  for (int a = 0; a < 3; ++a) { p.domain_lo[a] = lo[a]; p.domain_hi[a] = hi[a]; }
  p.grid_mode = as_u32(prop(env, o, "gridMode", PBF_GRID_DENSE));
  p.hash_log2 = as_u32(prop(env, o, "hashLog2", 22));
  p.flags = as_u32(prop(env, o, "flags", 0));
  p.cell_slabs = as_u32(prop(env, o, "cellSlabs", 0));
}

void finalize_ctx(napi_env, void *data, void *) { pbf_destroy(static_cast<pbf_ctx *>(data)); }

napi_value wrap_ctx(napi_env env, pbf_ctx *c) {
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, c, finalize_ctx, nullptr, &ext));
  return ext;
}

// create(params) -> handle
napi_value Create(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_status(env, nullptr, PBF_E_INVALID);
  pbf_params p;
  read_params(env, a.v[0], p);
  pbf_ctx *c = nullptr;
  int rc = pbf_create(&p, &c);
  if (rc != PBF_OK) {
    napi_throw_error(env, nullptr, pbf_last_error(nullptr));
    return nullptr;
  }
  return wrap_ctx(env, c);
}

// batchCreate(params, [{domainLo, domainHi, count}, ...]) -> handle
napi_value BatchCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 2) return throw_status(env, nullptr, PBF_E_INVALID);
  pbf_params p;
  read_params(env, a.v[0], p);
  uint32_t ns = 0;
  if (napi_get_array_length(env, a.v[1], &ns) != napi_ok || ns == 0 || ns > (1u << 20)) {
    napi_throw_range_error(env, nullptr, "scenes must be an array of 1 to 2^20 scene descriptions");
    return nullptr;
  }
  std::vector<pbf_scene_desc> d(ns);
  for (uint32_t k = 0; k < ns; ++k) {
    napi_value s, v;
    napi_get_element(env, a.v[1], k, &s);
    std::memset(&d[k], 0, sizeof(d[k]));
    napi_get_named_property(env, s, "domainLo", &v);
    vec3(env, v, d[k].domain_lo);
    napi_get_named_property(env, s, "domainHi", &v);
    vec3(env, v, d[k].domain_hi);
    d[k].particle_count = as_u32(prop(env, s, "count", 0));
  }
  pbf_ctx *c = nullptr;
  int rc = pbf_batch_create(&p, ns, d.data(), &c);
  if (rc != PBF_OK) {
    napi_throw_error(env, nullptr, pbf_last_error(nullptr));
    return nullptr;
  }
  return wrap_ctx(env, c);
}

#define CTX_OR_RETURN(idx)                  \
  Args a;                                   \
  if (!get_args(env, info, a)) return nullptr; \
  pbf_ctx *c = ctx_of(env, a.v[idx]);       \
  if (!c) return nullptr;

#define CHECK(rc)                           \
  do {                                      \
    int r_ = (rc);                          \
    if (r_ != PBF_OK) return throw_status(env, c, r_); \
  } while (0)

napi_value Reset(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  CHECK(pbf_reset(c));
  return undefined(env);
}

napi_value SetConfig(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  pbf_config cfg;
  cfg.iteration = as_u32(prop(env, a.v[1], "iteration", 5));
  cfg.xsph = (float)prop(env, a.v[1], "XSPH", 0.01);
  cfg.vorticity = (float)prop(env, a.v[1], "vorticity", 0.1);
  cfg.surface_tension = (float)prop(env, a.v[1], "surfaceTension", 0.5);
  cfg.gravity = (float)prop(env, a.v[1], "gravity", 9.8);
  CHECK(pbf_set_config(c, &cfg));
  return undefined(env);
}

napi_value VoxelizeCube(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  double mn[3], mx[3];
  vec3(env, a.v[1], mn);
  vec3(env, a.v[2], mx);
  double jitter = a.n > 3 ? num(env, a.v[3]) : 0.0;
  uint64_t seed = a.n > 4 ? as_u64(num(env, a.v[4])) : 0;
  uint32_t added = 0;
  CHECK(pbf_voxelize_cube(c, mn, mx, jitter, seed, &added));
  napi_value r;
  napi_create_uint32(env, added, &r);
  return r;
}

napi_value VoxelizeSphere(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  double ctr[3];
  vec3(env, a.v[1], ctr);
  double radius = num(env, a.v[2]);
  double jitter = a.n > 3 ? num(env, a.v[3]) : 0.0;
  uint64_t seed = a.n > 4 ? as_u64(num(env, a.v[4])) : 0;
  uint32_t added = 0;
  CHECK(pbf_voxelize_sphere(c, ctr, radius, jitter, seed, &added));
  napi_value r;
  napi_create_uint32(env, added, &r);
  return r;
}

// Index buffer of a THREE.BufferGeometry: Uint32Array or Uint16Array.
bool index_array(napi_env env, napi_value v, std::vector<uint32_t> &out) {
  napi_typedarray_type t;
  size_t n;
  void *p;
  napi_value ab;
  size_t off;
  if (napi_get_typedarray_info(env, v, &t, &n, &p, &ab, &off) != napi_ok ||
      (t != napi_uint32_array && t != napi_uint16_array)) {
    napi_throw_type_error(env, nullptr, "expected a Uint32Array or Uint16Array index");
    return false;
  }
  out.resize(n);
  for (size_t k = 0; k < n; ++k)
    out[k] = t == napi_uint32_array ? static_cast<uint32_t *>(p)[k] : static_cast<uint16_t *>(p)[k];
  return true;
}

// 16 numbers (Matrix4.elements, column-major) from an array or typed array.
bool matrix16(napi_env env, napi_value v, double m[16]) {
  for (uint32_t k = 0; k < 16; ++k) {
    napi_value e;
    if (napi_get_element(env, v, k, &e) != napi_ok) {
      napi_throw_type_error(env, nullptr, "expected 16 matrix elements");
      return false;
    }
    m[k] = num(env, e);
  }
  return true;
}

napi_value VoxelizeMesh(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  float *vp;
  size_t vlen;
  if (!float_array(env, a.v[1], &vp, &vlen)) return nullptr;
  std::vector<uint32_t> idx;
  if (!index_array(env, a.v[2], idx)) return nullptr;
  double m[16];
  const double *mp = nullptr;
  napi_valuetype t = napi_undefined;
  if (a.n > 3) napi_typeof(env, a.v[3], &t);
  if (t == napi_object) {
    if (!matrix16(env, a.v[3], m)) return nullptr;
    mp = m;
  }
  double jitter = a.n > 4 ? num(env, a.v[4]) : 0.0;
  uint64_t seed = a.n > 5 ? as_u64(num(env, a.v[5])) : 0;
  uint32_t added = 0;
  CHECK(pbf_voxelize_mesh(c, vp, (uint32_t)(vlen / 3), idx.data(), (uint32_t)(idx.size() / 3), mp, jitter, seed,
                          &added));
  napi_value r;
  napi_create_uint32(env, added, &r);
  return r;
}

// setWarmStart(ctx, blob | null, mode, relax, fallbackThreshold, fallbackIterations)
napi_value SetWarmStart(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  float *bp = nullptr;
  size_t blen = 0;
  napi_valuetype t = napi_undefined;
  napi_typeof(env, a.v[1], &t);
  if (t == napi_object && !float_array(env, a.v[1], &bp, &blen)) return nullptr;
  uint32_t mode = as_u32(num(env, a.v[2], 1));
  float relax = (float)num(env, a.v[3], 1.0);
  double thr = num(env, a.v[4], 0.0);
  uint32_t fb = as_u32(num(env, a.v[5], 0));
  CHECK(pbf_set_warm_start(c, bp, blen, mode, relax, thr, fb));
  return undefined(env);
}

napi_value tuning_object(napi_env env, const pbf_tuning &t) {
  napi_value o, arr, v;
  napi_create_object(env, &o);
  napi_create_array_with_length(env, 6, &arr);
  for (uint32_t k = 0; k < 6; ++k) {
    napi_create_uint32(env, t.walk_block[k], &v);
    napi_set_element(env, arr, k, v);
  }
  napi_set_named_property(env, o, "walkBlock", arr);
  napi_create_uint32(env, t.list_variant, &v);
  napi_set_named_property(env, o, "listVariant", v);
  return o;
}

// autotune(ctx, dt, iterations, repeats) -> {walkBlock, listVariant}
napi_value Autotune(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  pbf_tuning t;
  CHECK(pbf_autotune(c, (float)num(env, a.v[1], 1.0 / 300.0), as_u32(num(env, a.v[2], 0)),
                     as_u32(num(env, a.v[3], 3)), &t));
  return tuning_object(env, t);
}

// setTuning(ctx, walkBlock[6], listVariant)
napi_value SetTuning(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  pbf_tuning t;
  CHECK(pbf_get_tuning(c, &t));
  for (uint32_t k = 0; k < 6; ++k) {
    napi_value e;
    if (napi_get_element(env, a.v[1], k, &e) == napi_ok) t.walk_block[k] = as_u32(num(env, e, 256));
  }
  t.list_variant = as_u32(num(env, a.v[2], 0));
  CHECK(pbf_set_tuning(c, &t));
  return undefined(env);
}

napi_value SetParticlePosition(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  if (a.n >= 3) {
    float *d;
    size_t len;
    if (!float_array(env, a.v[1], &d, &len)) return nullptr;
    uint32_t count = as_u32(num(env, a.v[2]));
    CHECK(pbf_set_positions(c, d, len, count));
  } else {
    CHECK(pbf_set_particle_position(c));
  }
  return undefined(env);
}

napi_value SetBoundaryData(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  size_t len = 0;
  if (napi_get_value_string_utf8(env, a.v[1], nullptr, 0, &len) != napi_ok) {
    napi_throw_type_error(env, nullptr, "expected boundary text");
    return nullptr;
  }
  std::string s(len + 1, '\0');
  napi_get_value_string_utf8(env, a.v[1], &s[0], len + 1, &len);
  CHECK(pbf_set_boundary_data(c, s.data(), len));
  return undefined(env);
}

napi_value SetVolumeMap(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  uint32_t scene = as_u32(num(env, a.v[1]));
  double r[3], lo[3], hi[3];
  vec3(env, a.v[2], r);
  vec3(env, a.v[3], lo);
  vec3(env, a.v[4], hi);
  float *d;
  size_t len;
  if (!float_array(env, a.v[5], &d, &len)) return nullptr;
  for (int k = 0; k < 3; ++k)
    if (!(r[k] >= 1.0 && r[k] <= 65536.0) || r[k] != std::floor(r[k])) {
      napi_throw_range_error(env, nullptr, "volume map resolution must be integers in [1, 65536]");
      return nullptr;
    }
  int32_t res[3] = {(int32_t)r[0], (int32_t)r[1], (int32_t)r[2]};
  size_t need = 2 * ((size_t)res[0] + 1) * ((size_t)res[1] + 1) * ((size_t)res[2] + 1);
  if (len != need) {
    napi_throw_range_error(env, nullptr, "volume map length must be 2 (res+1)^3");
    return nullptr;
  }
  CHECK(pbf_set_volume_map(c, scene, res, lo, hi, d));
  return undefined(env);
}

napi_value Step(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  float dt = (float)num(env, a.v[1], 1.0 / 300.0);
  uint32_t sub = a.n > 2 ? as_u32(num(env, a.v[2], 1)) : 1;
  uint32_t it = a.n > 3 ? as_u32(num(env, a.v[3], 0)) : 0;
  CHECK(pbf_step(c, dt, sub, it));
  return undefined(env);
}

napi_value Run(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  CHECK(pbf_run(c));
  return undefined(env);
}

napi_value Sync(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  CHECK(pbf_sync(c));
  return undefined(env);
}

napi_value ReadField(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  uint32_t field = as_u32(num(env, a.v[1]));
  uint32_t n = pbf_particle_count(c);
  float *d = nullptr;
  napi_value arr = new_float_array(env, 4 * (size_t)n, &d);
  if (!arr) return nullptr;
  CHECK(pbf_read_field(c, field, d, n));
  return arr;
}

napi_value DensityError(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  uint32_t scene = a.n > 1 ? as_u32(num(env, a.v[1], 4294967295.0)) : 0xFFFFFFFFu;
  double m = 0, x = 0;
  CHECK(pbf_density_error(c, scene, &m, &x));
  napi_value arr, vm, vx;
  napi_create_array_with_length(env, 2, &arr);
  napi_create_double(env, m, &vm);
  napi_create_double(env, x, &vx);
  napi_set_element(env, arr, 0, vm);
  napi_set_element(env, arr, 1, vx);
  return arr;
}

napi_value ParticleCount(napi_env env, napi_callback_info info) {
  CTX_OR_RETURN(0);
  napi_value r;
  uint32_t n = pbf_particle_count(c);
  if (!n) n = pbf_staged_count(c);
  napi_create_uint32(env, n, &r);
  return r;
}

napi_value Version(napi_env env, napi_callback_info) {
  napi_value r;
  napi_create_string_utf8(env, pbf_version(), NAPI_AUTO_LENGTH, &r);
  return r;
}

// --- MLS-MPM (pbf_mpm_*): handles are externals of pbf_mpm*.
void finalize_mpm(napi_env, void *data, void *) { pbf_mpm_destroy(static_cast<pbf_mpm *>(data)); }

pbf_mpm *mpm_of(napi_env env, napi_value v) {
  void *p = nullptr;
  if (napi_get_value_external(env, v, &p) != napi_ok || !p) {
    napi_throw_type_error(env, nullptr, "expected an MPM context handle");
    return nullptr;
  }
  return static_cast<pbf_mpm *>(p);
}

napi_value throw_mpm(napi_env env, pbf_mpm *m, int rc) {
  std::string msg = pbf_mpm_last_error(m);
  if (msg.empty()) msg = "pbf_mpm error " + std::to_string(rc);
  napi_throw_error(env, nullptr, msg.c_str());
  return nullptr;
}

#define MPM_OR_RETURN(idx)                     \
  Args a;                                      \
  if (!get_args(env, info, a)) return nullptr; \
  pbf_mpm *m = mpm_of(env, a.v[idx]);          \
  if (!m) return nullptr;

#define MPM_CHECK(rc)                                  \
  do {                                                 \
    int r_ = (rc);                                     \
    if (r_ != PBF_OK) return throw_mpm(env, m, r_);    \
  } while (0)

// mpmCreate({capacity, nGrid, bound, dt, rho, E, gravity, flags}) -> handle
// (the reference's MPM constructor, MPM.ts:60-83)
napi_value MpmCreate(napi_env env, napi_callback_info info) {
  Args a;
  if (!get_args(env, info, a) || a.n < 1) return throw_mpm(env, nullptr, PBF_E_INVALID);
  pbf_mpm_params p;
  std::memset(&p, 0, sizeof(p));
  p.capacity = as_u32(prop(env, a.v[0], "capacity", 9000));
  p.n_grid = as_u32(prop(env, a.v[0], "nGrid", 32));
  p.bound = as_u32(prop(env, a.v[0], "bound", 3));
  p.flags = as_u32(prop(env, a.v[0], "flags", 0));
  p.dt = prop(env, a.v[0], "dt", 8e-4);
  p.rho = prop(env, a.v[0], "rho", 1.0);
  p.E = prop(env, a.v[0], "E", 400.0);
  p.gravity = prop(env, a.v[0], "gravity", 9.8);
  pbf_mpm *m = nullptr;
  const int rc = pbf_mpm_create(&p, &m);
  if (rc != PBF_OK) return throw_mpm(env, nullptr, rc);
  napi_value ext;
  NAPI_CALL(env, napi_create_external(env, m, finalize_mpm, nullptr, &ext));
  return ext;
}

// mpmSetPositions(handle, Float32Array xyzw) (MPM.initResource, MPM.ts:104-121)
napi_value MpmSetPositions(napi_env env, napi_callback_info info) {
  MPM_OR_RETURN(0);
  float *d = nullptr;
  size_t len = 0;
  if (a.n < 2 || !float_array(env, a.v[1], &d, &len)) return nullptr;
  if (len % 4 != 0 || len / 4 > 0xffffffffu) return throw_mpm(env, m, PBF_E_LENGTH);
  MPM_CHECK(pbf_mpm_set_positions(m, d, (uint32_t)(len / 4)));
  return undefined(env);
}

// mpmStep(handle, steps) (MPM.run, MPM.ts:204-238, `steps` times)
napi_value MpmStep(napi_env env, napi_callback_info info) {
  MPM_OR_RETURN(0);
  MPM_CHECK(pbf_mpm_step(m, a.n > 1 ? as_u32(num(env, a.v[1], 1)) : 1u));
  return undefined(env);
}

// mpmRead(handle, field[, nGrid]) -> Float32Array (PBF_MPM_POSITION /
// VELOCITY: 4 per particle, AFFINE: 9 per particle, GRID: 4 per node, of
// nGrid^3 nodes)
napi_value MpmRead(napi_env env, napi_callback_info info) {
  MPM_OR_RETURN(0);
  const int field = (int)as_u32(num(env, a.v[1]));
  size_t count = pbf_mpm_particle_count(m), per = 4;
  if (field == PBF_MPM_AFFINE) per = 9;
  if (field == PBF_MPM_GRID) {
    const uint32_t g = a.n > 2 ? as_u32(num(env, a.v[2])) : 0u;
    if (g == 0 || g > 1024) return throw_mpm(env, m, PBF_E_INVALID);
    count = (size_t)g * g * g;
  }
  float *d = nullptr;
  napi_value arr = new_float_array(env, per * count, &d);
  if (!arr) return nullptr;
  MPM_CHECK(pbf_mpm_read(m, field, d, count));
  return arr;
}

napi_value MpmSync(napi_env env, napi_callback_info info) {
  MPM_OR_RETURN(0);
  MPM_CHECK(pbf_mpm_sync(m));
  return undefined(env);
}

napi_value Init(napi_env env, napi_value exports) {
  struct {
    const char *name;
    napi_callback fn;
  } fns[] = {{"create", Create},
             {"batchCreate", BatchCreate},
             {"reset", Reset},
             {"setConfig", SetConfig},
             {"voxelizeCube", VoxelizeCube},
             {"voxelizeSphere", VoxelizeSphere},
             {"voxelizeMesh", VoxelizeMesh},
             {"setWarmStart", SetWarmStart},
             {"autotune", Autotune},
             {"setTuning", SetTuning},
             {"setParticlePosition", SetParticlePosition},
             {"setBoundaryData", SetBoundaryData},
             {"setVolumeMap", SetVolumeMap},
             {"step", Step},
             {"run", Run},
             {"sync", Sync},
             {"readField", ReadField},
             {"densityError", DensityError},
             {"particleCount", ParticleCount},
             {"version", Version},
             {"mpmCreate", MpmCreate},
             {"mpmSetPositions", MpmSetPositions},
             {"mpmStep", MpmStep},
             {"mpmRead", MpmRead},
             {"mpmSync", MpmSync}};
  for (auto &f : fns) {
    napi_value v;
    napi_create_function(env, f.name, NAPI_AUTO_LENGTH, f.fn, nullptr, &v);
    napi_set_named_property(env, exports, f.name, v);
  }
  return exports;
}

}  // namespace

NAPI_MODULE(NODE_GYP_MODULE_NAME, Init)
