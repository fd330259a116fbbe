// pbf_host.cpp -- host-side helpers of libpbf: the seeded jitter PRNG, the
// .cdm boundary-file parser and the analytic volume-map builder.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_quadrature.h"

namespace pbf {

// SplitMix64 finaliser (Steele, Lea, Flood 2014).
uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// Values of one 32-node cubic field: corner nodes (R+1)^3, then the three
// edge-node sets (src/simulator/boundary/volumeMap.ts:6-11).
// Field resolutions accepted per axis (the device samplers index fields with
// 32-bit arithmetic; the builders below stay in range with this bound).
static bool res_ok(const int32_t *r) {
  return r && r[0] >= 1 && r[1] >= 1 && r[2] >= 1 && r[0] <= 65536 && r[1] <= 65536 && r[2] <= 65536;
}

size_t cubic_field_length(const int r[3]) {  // r >= 1 (callers check)
  const size_t x = (size_t)r[0], y = (size_t)r[1], z = (size_t)r[2];
  return (x + 1) * (y + 1) * (z + 1) + 2 * x * (y + 1) * (z + 1) + (x + 1) * 2 * y * (z + 1) + (x + 1) * (y + 1) * 2 * z;
}

namespace {
struct Line {
  const char *b;
  const char *e;
};

std::vector<Line> split_lines(const char *t, size_t n) {
  std::vector<Line> out;
  const char *s = t, *end = t + n;
  while (s <= end) {
    const char *q = s;
    while (q < end && *q != '\n') ++q;
    out.push_back({s, q});
    if (q >= end) break;
    s = q + 1;
  }
  return out;
}

// Parse every whitespace-separated number on a line (the reference splits on
// ' ' and drops the trailing token, volumeMap.ts:28-29).
bool parse_numbers(const Line &l, std::vector<float> &out) {
  std::string s(l.b, l.e);
  const char *p = s.c_str();
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\r') ++p;
    if (!*p) break;
    char *q = nullptr;
    double v = std::strtod(p, &q);
    if (q == p) return false;
    out.push_back((float)v);
    p = q;
  }
  return true;
}
}  // namespace

// .cdm text: line 0 "Rx Ry Rz", line 1 count, line 2 SDF values, line 3
// count, line 4 volume values (src/simulator/boundary/volumeMap.ts:23-39).
int parse_cdm(const char *text, size_t len, int res[3], float **sdf, float **vol, size_t *flen) {
  *sdf = *vol = nullptr;
  try {
    auto lines = split_lines(text, len);
    if (lines.size() < 5) return -1;
    std::vector<float> head, a, b;
    if (!parse_numbers(lines[0], head) || head.size() != 3) return -1;
    for (int k = 0; k < 3; ++k) {
      if (!(head[k] >= 1.0f && head[k] <= 4096.0f) || head[k] != std::floor(head[k])) return -1;
      res[k] = (int)head[k];
    }
    if (!parse_numbers(lines[2], a) || !parse_numbers(lines[4], b)) return -1;
    size_t fl = cubic_field_length(res);
    if (a.size() != fl || b.size() != fl) return -1;
    *sdf = (float *)std::malloc(fl * 4);
    *vol = (float *)std::malloc(fl * 4);
    if (!*sdf || !*vol) {
      std::free(*sdf);
      std::free(*vol);
      *sdf = *vol = nullptr;
      return -1;
    }
    std::memcpy(*sdf, a.data(), fl * 4);
    std::memcpy(*vol, b.data(), fl * 4);
    *flen = fl;
    return 0;
  } catch (const std::bad_alloc &) {  // text too large for host memory: not a valid file here
    return -1;
  }
}

// Seeded jitter of particle `serial`, axis `axis`: uniform in [-amp, amp].
double jitter_value(uint64_t seed, uint64_t serial, int axis, double amp) {
  if (amp == 0.0) return 0.0;
  uint64_t z = splitmix64(seed ^ (0x9E3779B97F4A7C15ull * (3ull * serial + (uint64_t)axis + 1ull)));
  double u = (double)(z >> 11) * (1.0 / 9007199254740992.0);
  return (2.0 * u - 1.0) * amp;
}

namespace {
// Signed distance, positive in the fluid.
struct Solid {
  uint32_t kind;
  double box_lo[3], box_hi[3];
  double sc[3], sr;            // sphere
  double tc[3], tR, tr;        // torus (axis y)
  double sdf(const double *p) const {
    // inside of the box: distance to the nearest wall; outside: -distance
    double din = 1e300, q2 = 0.0;
    for (int a = 0; a < 3; ++a) {
      double lo = p[a] - box_lo[a], hi = box_hi[a] - p[a];
      din = std::fmin(din, std::fmin(lo, hi));
      double o = std::fmax(std::fmax(-lo, -hi), 0.0);
      q2 += o * o;
    }
    double d = din >= 0.0 ? din : -std::sqrt(q2 > 0.0 ? q2 : din * din);
    if (kind == 1) {
      double dx = p[0] - sc[0], dy = p[1] - sc[1], dz = p[2] - sc[2];
      d = std::fmin(d, std::sqrt(dx * dx + dy * dy + dz * dz) - sr);
      double ex = p[0] - tc[0], ey = p[1] - tc[1], ez = p[2] - tc[2];
      double qx = std::sqrt(ex * ex + ez * ez) - tR;
      d = std::fmin(d, std::sqrt(qx * qx + ey * ey) - tr);
    }
    return d;
  }
};
}  // namespace

}  // namespace pbf

// Volume map of an analytic solid.  With particle_radius > 0 the stored
// volume is the boundary volume V_b whose single mirrored sample reproduces
// the density deficit of the solid part of the kernel ball:
//   V_b(x) = 0.9 * int_{solid} W(x - y) dy / W(max(d + r/2, 2r)),
// 0.9 being the fluid's packing (V = 0.9 (2r)^3, SPH.ts:49) and
// max(d + r/2, 2r) the mirror distance of the sampler.  With
// particle_radius == 0 it is the plain solid volume inside the ball.
// The ratio has a pole where d + r/2 reaches h while a thin cap of solid is
// still inside the ball; V_b is therefore capped at its deep-solid value
// 0.9 / W(2r) (a sample never stands for more than a ball full of solid),
// which keeps interpolated values bounded near that shell.
namespace {
// The map definition above at one point: solid SDF, then the quadrature of
// the solid part of the kernel ball (stratified 16^3 midpoint cells).
struct VolumeEval {
  pbf::Solid S;
  double h, r, kp, cellv, ballv, vcap;
  std::vector<double> qp, qw;
  bool ref = false;                  // PBF_VOLUME_REFERENCE
  double map_lo[3], map_hi[3];       // the field's box
  std::vector<double> gp, gw;        // its quadrature

  int init(uint32_t kind_flags, const double *shape, const double lo[3], const double hi[3], double kernel_radius,
           double particle_radius) {
    ref = (kind_flags & PBF_VOLUME_REFERENCE) != 0;
    const uint32_t kind = kind_flags & ~(uint32_t)PBF_VOLUME_REFERENCE;
    if (!lo || !hi || kind > 1 || !(kernel_radius > 0.0) || !std::isfinite(kernel_radius)) return PBF_E_INVALID;
    if (!(particle_radius >= 0.0) || !std::isfinite(particle_radius)) return PBF_E_INVALID;
    for (int a = 0; a < 3; ++a)
      if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] > lo[a])) return PBF_E_INVALID;
    if (shape)
      for (int a = 0; a < (kind == 1 ? 9 : 6); ++a)
        if (!std::isfinite(shape[a])) return PBF_E_INVALID;
    std::memset(&S, 0, sizeof(S));
    S.kind = kind;
    for (int a = 0; a < 3; ++a) { S.box_lo[a] = lo[a]; S.box_hi[a] = hi[a]; }
    if (kind == 1) {
      if (!shape) return PBF_E_INVALID;
      for (int a = 0; a < 3; ++a) { S.sc[a] = shape[a]; S.tc[a] = shape[4 + a]; }
      S.sr = shape[3];
      S.tR = shape[7];
      S.tr = shape[8];
    } else if (shape) {
      for (int a = 0; a < 3; ++a) { S.box_lo[a] = shape[a]; S.box_hi[a] = shape[3 + a]; }
    }
    h = kernel_radius;
    r = particle_radius;
    for (int a = 0; a < 3; ++a) { map_lo[a] = lo[a]; map_hi[a] = hi[a]; }
    if (ref) pbf::reference_volume_quadrature(h, gp, gw);
    const int M = 16;
    const double PI = 3.141592653589793;
    kp = 315.0 / (64.0 * PI * std::pow(h, 9));
    for (int i = 0; i < M; ++i)
      for (int j = 0; j < M; ++j)
        for (int k = 0; k < M; ++k) {
          double o[3] = {(-1.0 + (2.0 * i + 1.0) / M) * h, (-1.0 + (2.0 * j + 1.0) / M) * h,
                         (-1.0 + (2.0 * k + 1.0) / M) * h};
          double rr = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
          if (rr <= h * h) {
            qp.push_back(o[0]); qp.push_back(o[1]); qp.push_back(o[2]);
            double t = h * h - rr;
            qw.push_back(kp * t * t * t);
          }
        }
    cellv = std::pow(2.0 * h / M, 3);
    const double t2r = h * h - 4.0 * r * r;
    vcap = t2r > 0.0 ? 0.9 / (kp * t2r * t2r * t2r) : 0.0;
    ballv = 4.0 / 3.0 * 3.141592653589793 * h * h * h;
    return PBF_OK;
  }

  // The reference's definition (pbf_quadrature.h): gamma of the SDF at the
  // quadrature points inside the field's box, in the points' order.
  double reference_volume(const double p[3], double d) const {
    if (d >= 2.0 * h) return 0.0;
    double v = 0.0;
    for (size_t t = 0, u = 0; t < gp.size(); t += 3, ++u) {
      const double y[3] = {p[0] + gp[t], p[1] + gp[t + 1], p[2] + gp[t + 2]};
      if (y[0] < map_lo[0] || y[0] > map_hi[0] || y[1] < map_lo[1] || y[1] > map_hi[1] || y[2] < map_lo[2] ||
          y[2] > map_hi[2])
        continue;
      const double phi = S.sdf(y);
      double g;
      if (phi <= 0.0) g = 1.0;
      else if (phi < h) {
        const double q = phi / h;
        g = q <= 0.5 ? (6.0 * q * q * q - 6.0 * q * q) + 1.0 : 2.0 * ((1.0 - q) * (1.0 - q) * (1.0 - q));
      } else g = 0.0;
      v += gw[u] * g;
    }
    return v;
  }

  void eval(const double p[3], float &sdf, float &vol) const {
    double d = S.sdf(p);
    double v;
    if (ref) v = reference_volume(p, d);
    else if (d >= h) v = 0.0;
    else {
      double vin = 0.0, win = 0.0;
      if (d <= -h) { vin = ballv; win = 1.0; }
      else {
        for (size_t t = 0, u = 0; t < qp.size(); t += 3, ++u) {
          double y[3] = {p[0] + qp[t], p[1] + qp[t + 1], p[2] + qp[t + 2]};
          if (S.sdf(y) < 0.0) { vin += cellv; win += qw[u] * cellv; }
        }
      }
      if (r > 0.0) {
        double s = std::fmax(d + 0.5 * r, 2.0 * r);
        double t = h * h - s * s;
        double ws = t > 0.0 ? kp * t * t * t : 0.0;
        v = ws > 0.0 ? std::fmin(0.9 * win / ws, vcap) : 0.0;
      } else {
        v = vin;
      }
    }
    sdf = (float)d;
    vol = (float)v;
  }
};

// Runs f(q) for q in [0, total) on up to 16 threads.
template <typename F>
void parallel_for(size_t total, F f) {
  unsigned nt = std::thread::hardware_concurrency();
  if (nt == 0) nt = 1;
  if (nt > 16) nt = 16;
  std::vector<std::thread> th;
  size_t chunk = (total + nt - 1) / nt;
  for (unsigned t = 0; t < nt; ++t) {
    size_t b = t * chunk, e = std::min(total, b + chunk);
    if (b < e) th.emplace_back([=] { for (size_t q = b; q < e; ++q) f(q); });
  }
  for (auto &x : th) x.join();
}
}  // namespace

extern "C" int pbf_build_volume_map(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                                    const double hi[3], double kernel_radius, double particle_radius,
                                    float *sdf_vol) {
  if (!pbf::res_ok(res) || !sdf_vol) return PBF_E_INVALID;
  VolumeEval E;
  int rc = E.init(kind, shape, lo, hi, kernel_radius, particle_radius);
  if (rc != PBF_OK) return rc;
  const size_t nx = (size_t)res[0] + 1, ny = (size_t)res[1] + 1, nz = (size_t)res[2] + 1;
  parallel_for(nx * ny * nz, [&](size_t q) {
    size_t i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
    double p[3] = {lo[0] + (hi[0] - lo[0]) * (double)i / res[0], lo[1] + (hi[1] - lo[1]) * (double)j / res[1],
                   lo[2] + (hi[2] - lo[2]) * (double)k / res[2]};
    E.eval(p, sdf_vol[2 * q], sdf_vol[2 * q + 1]);
  });
  return PBF_OK;
}

// The same definition sampled at the nodes of the 32-node cubic element in
// the .cdm layout (volumeMap.ts:23-39, discreteFieldShader.ts:1-12): corner
// nodes [k][j][i] ((R+1)^3), then x-edge nodes [k][j][i2], y-edge nodes
// [i][k][j2], z-edge nodes [j][i][k2] (2R(R+1)^2 each), an edge node i2 at
// lattice coordinate floor(i2 / 2) + (1 + i2 % 2) / 3.  field_length = the
// node count; sdf / vol receive it when capacity allows.
extern "C" int pbf_build_cubic_field(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                                     const double hi[3], double kernel_radius, double particle_radius, float *sdf,
                                     float *vol, size_t capacity, size_t *field_length) {
  using namespace pbf;
  if (!res_ok(res)) return PBF_E_INVALID;
  const int R[3] = {res[0], res[1], res[2]};
  const size_t fl = cubic_field_length(R);
  if (field_length) *field_length = fl;
  if (!sdf && !vol) return PBF_OK;
  if (!sdf || !vol) return PBF_E_INVALID;
  if (capacity < fl) return PBF_E_CAPACITY;
  VolumeEval E;
  int rc = E.init(kind, shape, lo, hi, kernel_radius, particle_radius);
  if (rc != PBF_OK) return rc;
  const size_t n0 = (size_t)(R[0] + 1) * (R[1] + 1) * (R[2] + 1);
  const size_t ne[3] = {(size_t)2 * R[0] * (R[1] + 1) * (R[2] + 1), (size_t)2 * R[1] * (R[0] + 1) * (R[2] + 1),
                        (size_t)2 * R[2] * (R[0] + 1) * (R[1] + 1)};
  auto lat = [&](int a, double u) { return lo[a] + (hi[a] - lo[a]) * u / R[a]; };
  auto edge = [](size_t i2) { return (double)(i2 / 2) + (1.0 + (double)(i2 % 2)) / 3.0; };
  parallel_for(fl, [&](size_t q) {
    double p[3];
    if (q < n0) {  // corner[k][j][i]
      size_t i = q % (R[0] + 1), j = (q / (R[0] + 1)) % (R[1] + 1), k = q / ((size_t)(R[0] + 1) * (R[1] + 1));
      p[0] = lat(0, (double)i); p[1] = lat(1, (double)j); p[2] = lat(2, (double)k);
    } else if (q < n0 + ne[0]) {  // edgex[k][j][i2]
      size_t t = q - n0, i2 = t % (2 * R[0]), j = (t / (2 * R[0])) % (R[1] + 1), k = t / (2 * (size_t)R[0] * (R[1] + 1));
      p[0] = lat(0, edge(i2)); p[1] = lat(1, (double)j); p[2] = lat(2, (double)k);
    } else if (q < n0 + ne[0] + ne[1]) {  // edgey[i][k][j2]
      size_t t = q - n0 - ne[0], j2 = t % (2 * R[1]), k = (t / (2 * R[1])) % (R[2] + 1), i = t / (2 * (size_t)R[1] * (R[2] + 1));
      p[0] = lat(0, (double)i); p[1] = lat(1, edge(j2)); p[2] = lat(2, (double)k);
    } else {  // edgez[j][i][k2]
      size_t t = q - n0 - ne[0] - ne[1], k2 = t % (2 * R[2]), i = (t / (2 * R[2])) % (R[0] + 1), j = t / (2 * (size_t)R[2] * (R[0] + 1));
      p[0] = lat(0, (double)i); p[1] = lat(1, (double)j); p[2] = lat(2, edge(k2));
    }
    E.eval(p, sdf[q], vol[q]);
  });
  return PBF_OK;
}

// .cdm text of a cubic field (the layout volumeMap.ts:23-39 reads): "Rx Ry Rz",
// the SDF count, the SDF values, the volume count, the volume values, each
// value followed by a space, CRLF line ends; values with 9 significant
// digits (exact float round trip).  out == NULL: size only (written gets the
// byte count without a terminator).
extern "C" int pbf_write_cdm(const int32_t res[3], const float *sdf, const float *vol, size_t length, char *out,
                             size_t capacity, size_t *written) {
  using namespace pbf;
  if (!res_ok(res) || !sdf || !vol) return PBF_E_INVALID;
  const int R[3] = {res[0], res[1], res[2]};
  if (length != cubic_field_length(R)) return PBF_E_LENGTH;
  try {  // the text is built in host memory: an allocation failure is a status, not an exception
    std::string t;
    t.reserve(length * 32 + 64);
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%d %d %d\r\n%zu\r\n", R[0], R[1], R[2], length);
    t += buf;
    for (size_t q = 0; q < length; ++q) { std::snprintf(buf, sizeof(buf), "%.9g ", (double)sdf[q]); t += buf; }
    std::snprintf(buf, sizeof(buf), "\r\n%zu\r\n", length);
    t += buf;
    for (size_t q = 0; q < length; ++q) { std::snprintf(buf, sizeof(buf), "%.9g ", (double)vol[q]); t += buf; }
    t += "\r\n";
    if (written) *written = t.size();
    if (!out) return PBF_OK;
    if (capacity < t.size()) return PBF_E_CAPACITY;
    std::memcpy(out, t.data(), t.size());
    return PBF_OK;
  } catch (const std::bad_alloc &) {
    return PBF_E_CAPACITY;
  }
}

namespace {
// Lattice points along one axis of a box of this extent (floor(extent /
// spacing), 0 for an empty box), or -1 when the extent or the spacing is not
// finite or the count is out of any 32-bit range.
long lattice_count(double extent, double diam) {
  if (!std::isfinite(extent) || !std::isfinite(diam) || !(diam > 0.0)) return -1;
  const double n = std::floor(extent / diam);
  if (!(n > 0.0)) return 0;
  if (n > 2147483647.0) return -1;
  return (long)n;
}

bool all_finite(const double *v, int n) {
  for (int a = 0; a < n; ++a)
    if (!std::isfinite(v[a])) return false;
  return true;
}
}  // namespace

// SPH.voxelizeCube (SPH.ts:99-121): lattice of spacing 2r from `mn`, i (x)
// outermost, k (z) innermost, f64 arithmetic rounded once to f32; plus the
// seeded jitter of particle serial0 + q.  out == NULL: count only.
extern "C" int pbf_voxelize_cube_host(double r, const double mn[3], const double mx[3], double jitter, uint64_t seed,
                                      uint64_t serial0, float *out, uint64_t capacity, uint32_t *count) {
  using namespace pbf;
  if (!mn || !mx || !(r > 0.0) || !all_finite(mn, 3) || !all_finite(mx, 3)) return PBF_E_INVALID;
  const double diam = 2.0 * r;
  const long nx = lattice_count(mx[0] - mn[0], diam);
  const long ny = lattice_count(mx[1] - mn[1], diam);
  const long nz = lattice_count(mx[2] - mn[2], diam);
  if (nx < 0 || ny < 0 || nz < 0) return PBF_E_INVALID;
  if ((double)nx * (double)ny * (double)nz > 4294967295.0) return PBF_E_CAPACITY;
  const uint64_t total = (uint64_t)nx * (uint64_t)ny * (uint64_t)nz;
  if (count) *count = (uint32_t)total;
  if (!out) return PBF_OK;
  if (total > capacity) return PBF_E_CAPACITY;
  uint64_t q = 0;
  for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
      for (long k = 0; k < nz; ++k, ++q) {
        double p[3] = {(double)i * diam + mn[0], (double)j * diam + mn[1], (double)k * diam + mn[2]};
        for (int a = 0; a < 3; ++a) out[4 * q + a] = (float)(p[a] + jitter_value(seed, serial0 + q, a, jitter));
        out[4 * q + 3] = 0.0f;
      }
  return PBF_OK;
}

// SPH.voxelizeSphere (SPH.ts:123-149): lattice points of spacing 2r with
// |p|^2 <= R^2 around `center`.
extern "C" int pbf_voxelize_sphere_host(double r, const double center[3], double radius, double jitter, uint64_t seed,
                                        uint64_t serial0, float *out, uint64_t capacity, uint32_t *count) {
  using namespace pbf;
  if (!center || !(r > 0.0) || !all_finite(center, 3) || !std::isfinite(radius)) return PBF_E_INVALID;
  const double diam = 2.0 * r;
  if (radius < 0.0) {  // the reference's loops from -n to n run empty
    if (count) *count = 0;
    return PBF_OK;
  }
  const double nd = std::floor(radius / r / 2.0);  // the reference's expression, rounding included
  if (!std::isfinite(nd)) return PBF_E_INVALID;
  if (2.0 * nd + 1.0 > 1625.0) return PBF_E_CAPACITY;  // (2n+1)^3 candidates < 2^32
  const long n = (long)nd;
  const double r2 = radius * radius;
  uint64_t q = 0;
  for (long i = -n; i <= n; ++i)
    for (long j = -n; j <= n; ++j)
      for (long k = -n; k <= n; ++k) {
        double p[3] = {(double)i * diam, (double)j * diam, (double)k * diam};
        if ((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2] <= r2) {
          if (out) {
            if (q >= capacity) return PBF_E_CAPACITY;
            for (int a = 0; a < 3; ++a)
              out[4 * q + a] = (float)(p[a] + center[a] + jitter_value(seed, serial0 + q, a, jitter));
            out[4 * q + 3] = 0.0f;
          }
          ++q;
        }
      }
  if (count) *count = (uint32_t)q;
  return PBF_OK;
}

// SPH.voxelizeMesh (SPH.ts:151-187): the lattice of spacing 2r over the
// mesh's world-space bounding box (i outermost, k innermost), keeping a point
// when the first triangle hit by the ray from it along the mesh's local +z
// axis faces along the ray (the ray leaves the solid first).  The reference
// casts the ray with three-mesh-bvh (absent here); this is a plain
// double-precision ray / triangle test (Moller-Trumbore) over a 2-D grid of
// the triangles' local-x/y bounds.  Hits at t <= 1e-12 (the origin on a
// face) and rays in a face's plane do not count -- the reference leaves
// those degenerate cases to its ray caster.  `world` is the mesh's 4x4
// world matrix in column-major order (three.js Matrix4.elements), NULL for
// the identity.
namespace {
struct MeshRays {
  std::vector<double> v;          // local-space vertices, xyz
  std::vector<uint32_t> tri;      // 3 per triangle
  double lo[2], cell[2];
  int nx = 1, ny = 1;
  std::vector<std::vector<uint32_t>> bins;

  void build() {
    const size_t nt = tri.size() / 3;
    double hi[2];
    lo[0] = lo[1] = 1e300;
    hi[0] = hi[1] = -1e300;
    for (size_t q = 0; q < v.size() / 3; ++q)
      for (int a = 0; a < 2; ++a) {
        lo[a] = std::fmin(lo[a], v[3 * q + a]);
        hi[a] = std::fmax(hi[a], v[3 * q + a]);
      }
    const int side = (int)std::max(1.0, std::min(512.0, std::sqrt((double)nt)));
    nx = ny = side;
    for (int a = 0; a < 2; ++a) cell[a] = std::fmax(hi[a] - lo[a], 1e-300) / side;
    bins.assign((size_t)nx * ny, {});
    for (size_t t = 0; t < nt; ++t) {
      double bl[2] = {1e300, 1e300}, bh[2] = {-1e300, -1e300};
      for (int e = 0; e < 3; ++e)
        for (int a = 0; a < 2; ++a) {
          bl[a] = std::fmin(bl[a], v[3 * tri[3 * t + e] + a]);
          bh[a] = std::fmax(bh[a], v[3 * tri[3 * t + e] + a]);
        }
      int i0 = bin(bl[0], 0), i1 = bin(bh[0], 0), j0 = bin(bl[1], 1), j1 = bin(bh[1], 1);
      for (int i = i0; i <= i1; ++i)
        for (int j = j0; j <= j1; ++j) bins[(size_t)j * nx + i].push_back((uint32_t)t);
    }
  }
  int bin(double x, int a) const {
    int k = (int)std::floor((x - lo[a]) / cell[a]);
    int n = a == 0 ? nx : ny;
    return k < 0 ? 0 : k >= n ? n - 1 : k;
  }
  // +1 the first hit faces along +z, -1 against it, 0 no hit
  int first_hit(const double o[3]) const {
    if (o[0] < lo[0] - cell[0] || o[1] < lo[1] - cell[1]) return 0;
    const std::vector<uint32_t> &b = bins[(size_t)bin(o[1], 1) * nx + bin(o[0], 0)];
    double best = 1e300;
    int face = 0;
    for (uint32_t t : b) {
      const double *A = &v[3 * tri[3 * t]], *B = &v[3 * tri[3 * t + 1]], *C = &v[3 * tri[3 * t + 2]];
      const double e1[3] = {B[0] - A[0], B[1] - A[1], B[2] - A[2]};
      const double e2[3] = {C[0] - A[0], C[1] - A[1], C[2] - A[2]};
      // direction d = (0, 0, 1): p = d x e2, det = e1 . p
      const double p[3] = {-e2[1], e2[0], 0.0};
      const double det = e1[0] * p[0] + e1[1] * p[1];
      if (std::fabs(det) < 1e-300) continue;
      const double inv = 1.0 / det;
      const double s[3] = {o[0] - A[0], o[1] - A[1], o[2] - A[2]};
      const double u = (s[0] * p[0] + s[1] * p[1]) * inv;
      if (u < 0.0 || u > 1.0) continue;
      const double qv[3] = {s[1] * e1[2] - s[2] * e1[1], s[2] * e1[0] - s[0] * e1[2], s[0] * e1[1] - s[1] * e1[0]};
      const double w = qv[2] * inv;
      if (w < 0.0 || u + w > 1.0) continue;
      const double tt = (e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2]) * inv;
      if (tt <= 1e-12 || tt >= best) continue;
      best = tt;
      // face normal (C - B) x (A - B) as three.js Triangle.getNormal; its z
      const double cb[2] = {C[0] - B[0], C[1] - B[1]}, ab[2] = {A[0] - B[0], A[1] - B[1]};
      const double nz = cb[0] * ab[1] - cb[1] * ab[0];
      face = nz > 0.0 ? 1 : -1;
    }
    return face;
  }
};

bool invert4(const double m[16], double r[16]) {  // column-major 4x4 inverse (cofactors)
  double inv[16];
  inv[0] = m[5] * m[10] * m[15] - m[5] * m[11] * m[14] - m[9] * m[6] * m[15] + m[9] * m[7] * m[14] + m[13] * m[6] * m[11] - m[13] * m[7] * m[10];
  inv[4] = -m[4] * m[10] * m[15] + m[4] * m[11] * m[14] + m[8] * m[6] * m[15] - m[8] * m[7] * m[14] - m[12] * m[6] * m[11] + m[12] * m[7] * m[10];
  inv[8] = m[4] * m[9] * m[15] - m[4] * m[11] * m[13] - m[8] * m[5] * m[15] + m[8] * m[7] * m[13] + m[12] * m[5] * m[11] - m[12] * m[7] * m[9];
  inv[12] = -m[4] * m[9] * m[14] + m[4] * m[10] * m[13] + m[8] * m[5] * m[14] - m[8] * m[6] * m[13] - m[12] * m[5] * m[10] + m[12] * m[6] * m[9];
  inv[1] = -m[1] * m[10] * m[15] + m[1] * m[11] * m[14] + m[9] * m[2] * m[15] - m[9] * m[3] * m[14] - m[13] * m[2] * m[11] + m[13] * m[3] * m[10];
  inv[5] = m[0] * m[10] * m[15] - m[0] * m[11] * m[14] - m[8] * m[2] * m[15] + m[8] * m[3] * m[14] + m[12] * m[2] * m[11] - m[12] * m[3] * m[10];
  inv[9] = -m[0] * m[9] * m[15] + m[0] * m[11] * m[13] + m[8] * m[1] * m[15] - m[8] * m[3] * m[13] - m[12] * m[1] * m[11] + m[12] * m[3] * m[9];
  inv[13] = m[0] * m[9] * m[14] - m[0] * m[10] * m[13] - m[8] * m[1] * m[14] + m[8] * m[2] * m[13] + m[12] * m[1] * m[10] - m[12] * m[2] * m[9];
  inv[2] = m[1] * m[6] * m[15] - m[1] * m[7] * m[14] - m[5] * m[2] * m[15] + m[5] * m[3] * m[14] + m[13] * m[2] * m[7] - m[13] * m[3] * m[6];
  inv[6] = -m[0] * m[6] * m[15] + m[0] * m[7] * m[14] + m[4] * m[2] * m[15] - m[4] * m[3] * m[14] - m[12] * m[2] * m[7] + m[12] * m[3] * m[6];
  inv[10] = m[0] * m[5] * m[15] - m[0] * m[7] * m[13] - m[4] * m[1] * m[15] + m[4] * m[3] * m[13] + m[12] * m[1] * m[7] - m[12] * m[3] * m[5];
  inv[14] = -m[0] * m[5] * m[14] + m[0] * m[6] * m[13] + m[4] * m[1] * m[14] - m[4] * m[2] * m[13] - m[12] * m[1] * m[6] + m[12] * m[2] * m[5];
  inv[3] = -m[1] * m[6] * m[11] + m[1] * m[7] * m[10] + m[5] * m[2] * m[11] - m[5] * m[3] * m[10] - m[9] * m[2] * m[7] + m[9] * m[3] * m[6];
  inv[7] = m[0] * m[6] * m[11] - m[0] * m[7] * m[10] - m[4] * m[2] * m[11] + m[4] * m[3] * m[10] + m[8] * m[2] * m[7] - m[8] * m[3] * m[6];
  inv[11] = -m[0] * m[5] * m[11] + m[0] * m[7] * m[9] + m[4] * m[1] * m[11] - m[4] * m[3] * m[9] - m[8] * m[1] * m[7] + m[8] * m[3] * m[5];
  inv[15] = m[0] * m[5] * m[10] - m[0] * m[6] * m[9] - m[4] * m[1] * m[10] + m[4] * m[2] * m[9] + m[8] * m[1] * m[6] - m[8] * m[2] * m[5];
  double det = m[0] * inv[0] + m[1] * inv[4] + m[2] * inv[8] + m[3] * inv[12];
  if (!(std::fabs(det) > 1e-300)) return false;
  for (int k = 0; k < 16; ++k) r[k] = inv[k] / det;
  return true;
}

void apply4(const double m[16], const double p[3], double out[3]) {  // column-major, w = 1
  for (int a = 0; a < 3; ++a) out[a] = m[a] * p[0] + m[4 + a] * p[1] + m[8 + a] * p[2] + m[12 + a];
}
}  // namespace

extern "C" int pbf_voxelize_mesh_host(double r, const float *vertices, uint32_t nverts, const uint32_t *indices,
                                      uint32_t ntris, const double world[16], double jitter, uint64_t seed,
                                      uint64_t serial0, float *out, uint64_t capacity, uint32_t *count) {
  using namespace pbf;
  if (!vertices || !indices || nverts == 0 || ntris == 0 || !(r > 0.0)) return PBF_E_INVALID;
  for (uint64_t q = 0; q < 3ull * ntris; ++q)
    if (indices[q] >= nverts) return PBF_E_INVALID;
  double M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1}, Minv[16];
  if (world) std::memcpy(M, world, sizeof(M));
  if (!invert4(M, Minv)) return PBF_E_INVALID;
  MeshRays R;
  R.v.resize(3 * (size_t)nverts);
  R.tri.assign(indices, indices + 3 * (size_t)ntris);
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  for (uint32_t q = 0; q < nverts; ++q) {
    double p[3] = {vertices[3 * q], vertices[3 * q + 1], vertices[3 * q + 2]}, w[3];
    for (int a = 0; a < 3; ++a) R.v[3 * q + a] = p[a];
    apply4(M, p, w);
    for (int a = 0; a < 3; ++a) {
      lo[a] = std::fmin(lo[a], w[a]);
      hi[a] = std::fmax(hi[a], w[a]);
    }
  }
  const double diam = 2.0 * r;
  const long nx = lattice_count(hi[0] - lo[0], diam);
  const long ny = lattice_count(hi[1] - lo[1], diam);
  const long nz = lattice_count(hi[2] - lo[2], diam);
  if (nx < 0 || ny < 0 || nz < 0) return PBF_E_INVALID;  // non-finite vertices or matrix
  if ((double)nx * (double)ny * (double)nz > 4294967295.0) return PBF_E_CAPACITY;
  R.build();
  uint64_t q = 0;
  for (long i = 0; i < nx; ++i)
    for (long j = 0; j < ny; ++j)
      for (long k = 0; k < nz; ++k) {
        double p[3] = {(double)i * diam + lo[0], (double)j * diam + lo[1], (double)k * diam + lo[2]}, o[3];
        apply4(Minv, p, o);
        if (R.first_hit(o) <= 0) continue;
        if (out) {
          if (q >= capacity) return PBF_E_CAPACITY;
          for (int a = 0; a < 3; ++a) out[4 * q + a] = (float)(p[a] + jitter_value(seed, serial0 + q, a, jitter));
          out[4 * q + 3] = 0.0f;
        }
        ++q;
      }
  if (count) *count = (uint32_t)q;
  return PBF_OK;
}

extern "C" int pbf_parse_cdm(const char *text, size_t length, int32_t res[3], float *sdf, float *vol,
                             size_t capacity, size_t *field_length) {
  using namespace pbf;
  if (!text || !res) return PBF_E_INVALID;
  int r3[3];
  float *a = nullptr, *b = nullptr;
  size_t fl = 0;
  if (parse_cdm(text, length, r3, &a, &b, &fl) != 0) return PBF_E_FORMAT;
  for (int k = 0; k < 3; ++k) res[k] = r3[k];
  if (field_length) *field_length = fl;
  int rc = PBF_OK;
  if (sdf || vol) {
    if (capacity < fl) rc = PBF_E_CAPACITY;
    else {
      if (sdf) std::memcpy(sdf, a, fl * 4);
      if (vol) std::memcpy(vol, b, fl * 4);
    }
  }
  std::free(a);
  std::free(b);
  return rc;
}

// A hashed grid stores cell (x, y, z) in bucket ((x D1 + y) D2 + z) mod 2^k.
// Two of the 27 cells around a cell differ in that row-major index by
// a D1 D2 + b D2 + c with a, b, c in [-2, 2] (|a| < D0, |b| < D1, |c| < D2,
// since both cells lie in the grid), so the 27 buckets are distinct for every
// cell exactly when no such nonzero difference is a multiple of 2^k.
extern "C" int pbf_hash_buckets_distinct(const int32_t dims[3], uint32_t hash_log2) {
  if (!dims || dims[0] < 1 || dims[1] < 1 || dims[2] < 1 || hash_log2 > 62) return 0;
  const uint64_t mask = (1ull << hash_log2) - 1ull;
  const int64_t D1 = dims[1], D2 = dims[2];
  for (int a = -2; a <= 2; ++a)
    for (int b = -2; b <= 2; ++b)
      for (int c = -2; c <= 2; ++c) {
        if (a == 0 && b == 0 && c == 0) continue;
        if (std::abs(a) >= dims[0] || std::abs(b) >= dims[1] || std::abs(c) >= dims[2]) continue;
        const int64_t d = (int64_t)a * D1 * D2 + (int64_t)b * D2 + c;
        if (((uint64_t)d & mask) == 0) return 0;
      }
  return 1;
}
