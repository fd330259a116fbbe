// pbf_mapgen.hip -- the volume-map generator on the GPU.
//
// Same definition, quadrature and arithmetic order as the host builder
// (pbf_host.cpp, pbf_build_volume_map): one thread per map node evaluates the
// analytic SDF, and for nodes within h of the solid sums the poly6 weight of
// the solid quadrature cells of the kernel ball in float64, in the host's
// order.  A 256^3 map with a few million boundary nodes takes milliseconds
// here against minutes on the host.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <exception>
#include <vector>

#include "../../include/pbf.h"
#include "pbf_quadrature.h"

using pbf::reference_volume_quadrature;

namespace {

struct SolidDev {
  uint32_t kind;
  double box_lo[3], box_hi[3];
  double sc[3], sr;
  double tc[3], tR, tr;
};

// Signed distance, positive in the fluid (the host's Solid::sdf, term by term).
__device__ double solid_sdf(const SolidDev &S, const double *p) {
  double din = 1e300, q2 = 0.0;
  for (int a = 0; a < 3; ++a) {
    double lo = p[a] - S.box_lo[a], hi = S.box_hi[a] - p[a];
    din = fmin(din, fmin(lo, hi));
    double o = fmax(fmax(-lo, -hi), 0.0);
    q2 += o * o;
  }
  double d = din >= 0.0 ? din : -sqrt(q2 > 0.0 ? q2 : din * din);
  if (S.kind == 1) {
    double dx = p[0] - S.sc[0], dy = p[1] - S.sc[1], dz = p[2] - S.sc[2];
    d = fmin(d, sqrt(dx * dx + dy * dy + dz * dz) - S.sr);
    double ex = p[0] - S.tc[0], ey = p[1] - S.tc[1], ez = p[2] - S.tc[2];
    double qx = sqrt(ex * ex + ez * ez) - S.tR;
    d = fmin(d, sqrt(qx * qx + ey * ey) - S.tr);
  }
  return d;
}

struct MapArgs {
  SolidDev S;
  double lo[3], hi[3];
  int res[3];
  double h, r, kp, cellv, ballv, vcap;
  uint32_t nq;
  int refv;  // PBF_VOLUME_REFERENCE: qp / qw hold the reference quadrature
};

// The reference's boundary volume (pbf_quadrature.h), in the host's order.
__device__ double reference_volume(const MapArgs &A, const double *__restrict__ qp, const double *__restrict__ qw,
                                   const double p[3], double d) {
  if (d >= 2.0 * A.h) return 0.0;
  double v = 0.0;
  for (uint32_t u = 0; u < A.nq; ++u) {
    const double y[3] = {p[0] + qp[3 * u], p[1] + qp[3 * u + 1], p[2] + qp[3 * u + 2]};
    if (y[0] < A.lo[0] || y[0] > A.hi[0] || y[1] < A.lo[1] || y[1] > A.hi[1] || y[2] < A.lo[2] || y[2] > A.hi[2])
      continue;
    const double phi = solid_sdf(A.S, y);
    double g;
    if (phi <= 0.0) g = 1.0;
    else if (phi < A.h) {
      const double q = phi / A.h;
      g = q <= 0.5 ? (6.0 * q * q * q - 6.0 * q * q) + 1.0 : 2.0 * ((1.0 - q) * (1.0 - q) * (1.0 - q));
    } else g = 0.0;
    v += qw[u] * g;
  }
  return v;
}

__global__ void __launch_bounds__(256) k_volume_map(MapArgs A, const double *__restrict__ qp,
                                                    const double *__restrict__ qw, float *__restrict__ out,
                                                    size_t total) {
  const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= total) return;
  const size_t nx = (size_t)A.res[0] + 1, ny = (size_t)A.res[1] + 1;
  const size_t i = q % nx, j = (q / nx) % ny, k = q / (nx * ny);
  double p[3] = {A.lo[0] + (A.hi[0] - A.lo[0]) * (double)i / A.res[0],
                 A.lo[1] + (A.hi[1] - A.lo[1]) * (double)j / A.res[1],
                 A.lo[2] + (A.hi[2] - A.lo[2]) * (double)k / A.res[2]};
  const double d = solid_sdf(A.S, p);
  double v;
  if (A.refv) {
    v = reference_volume(A, qp, qw, p, d);
  } else if (d >= A.h) {
    v = 0.0;
  } else {
    double vin = 0.0, win = 0.0;
    if (d <= -A.h) {
      vin = A.ballv;
      win = 1.0;
    } else {
      for (uint32_t u = 0; u < A.nq; ++u) {
        double y[3] = {p[0] + qp[3 * u], p[1] + qp[3 * u + 1], p[2] + qp[3 * u + 2]};
        if (solid_sdf(A.S, y) < 0.0) {
          vin += A.cellv;
          win += qw[u] * A.cellv;
        }
      }
    }
    if (A.r > 0.0) {
      double s = fmax(d + 0.5 * A.r, 2.0 * A.r);
      double t = A.h * A.h - s * s;
      double ws = t > 0.0 ? A.kp * t * t * t : 0.0;
      v = ws > 0.0 ? fmin(0.9 * win / ws, A.vcap) : 0.0;
    } else {
      v = vin;
    }
  }
  out[2 * q] = (float)d;
  out[2 * q + 1] = (float)v;
}

}  // namespace

extern "C" int pbf_build_volume_map_gpu(uint32_t kind, const double *shape, const int32_t res[3], const double lo[3],
                                        const double hi[3], double kernel_radius, double particle_radius,
                                        float *sdf_vol) try {
  const int refv = (kind & PBF_VOLUME_REFERENCE) ? 1 : 0;
  kind &= ~(uint32_t)PBF_VOLUME_REFERENCE;
  if (!res || !lo || !hi || !sdf_vol || kind > 1 || !(kernel_radius > 0.0)) return PBF_E_INVALID;
  for (int a = 0; a < 3; ++a)
    if (res[a] < 1 || res[a] > 65536) return PBF_E_INVALID;
  if (!std::isfinite(kernel_radius) || !(particle_radius >= 0.0) || !std::isfinite(particle_radius))
    return PBF_E_INVALID;
  for (int a = 0; a < 3; ++a)
    if (!std::isfinite(lo[a]) || !std::isfinite(hi[a]) || !(hi[a] > lo[a])) return PBF_E_INVALID;
  if (shape)
    for (int a = 0; a < (kind == 1 ? 9 : 6); ++a)
      if (!std::isfinite(shape[a])) return PBF_E_INVALID;
  MapArgs A = {};
  A.S.kind = kind;
  for (int a = 0; a < 3; ++a) {
    A.S.box_lo[a] = lo[a];
    A.S.box_hi[a] = hi[a];
    A.lo[a] = lo[a];
    A.hi[a] = hi[a];
    A.res[a] = res[a];
  }
  if (kind == 1) {
    if (!shape) return PBF_E_INVALID;
    for (int a = 0; a < 3; ++a) {
      A.S.sc[a] = shape[a];
      A.S.tc[a] = shape[4 + a];
    }
    A.S.sr = shape[3];
    A.S.tR = shape[7];
    A.S.tr = shape[8];
  } else if (shape) {
    for (int a = 0; a < 3; ++a) {
      A.S.box_lo[a] = shape[a];
      A.S.box_hi[a] = shape[3 + a];
    }
  }
  const double h = kernel_radius;
  const int M = 16;
  const double PI = 3.141592653589793;
  A.h = h;
  A.r = particle_radius;
  A.kp = 315.0 / (64.0 * PI * std::pow(h, 9));
  std::vector<double> qp, qw;
  for (int i = 0; i < M; ++i)
    for (int j = 0; j < M; ++j)
      for (int k = 0; k < M; ++k) {
        double o[3] = {(-1.0 + (2.0 * i + 1.0) / M) * h, (-1.0 + (2.0 * j + 1.0) / M) * h,
                       (-1.0 + (2.0 * k + 1.0) / M) * h};
        double rr = o[0] * o[0] + o[1] * o[1] + o[2] * o[2];
        if (rr <= h * h) {
          qp.push_back(o[0]);
          qp.push_back(o[1]);
          qp.push_back(o[2]);
          double t = h * h - rr;
          qw.push_back(A.kp * t * t * t);
        }
      }
  A.refv = refv;
  if (refv) reference_volume_quadrature(h, qp, qw);
  A.nq = (uint32_t)qw.size();
  A.cellv = std::pow(2.0 * h / M, 3);
  A.ballv = 4.0 / 3.0 * 3.141592653589793 * h * h * h;
  const double t2r = h * h - 4.0 * particle_radius * particle_radius;
  A.vcap = t2r > 0.0 ? 0.9 / (A.kp * t2r * t2r * t2r) : 0.0;
  const size_t total = ((size_t)res[0] + 1) * ((size_t)res[1] + 1) * ((size_t)res[2] + 1);
  if ((total + 255) / 256 > 0x7fffffffu) return PBF_E_CAPACITY;  // one launch, 1-D grid
  double *dqp = nullptr, *dqw = nullptr;
  float *dout = nullptr;
  hipStream_t st = nullptr;
  int rc = PBF_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == PBF_OK) rc = PBF_E_DEVICE;
    return rc == PBF_OK;
  };
  if (ok(hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) && ok(hipMalloc((void **)&dqp, qp.size() * 8)) &&
      ok(hipMalloc((void **)&dqw, qw.size() * 8)) && ok(hipMalloc((void **)&dout, total * 8)) &&
      ok(hipMemcpyAsync(dqp, qp.data(), qp.size() * 8, hipMemcpyHostToDevice, st)) &&
      ok(hipMemcpyAsync(dqw, qw.data(), qw.size() * 8, hipMemcpyHostToDevice, st))) {
    const size_t blocks = (total + 255) / 256;
    hipLaunchKernelGGL(k_volume_map, dim3((uint32_t)blocks), dim3(256), 0, st, A, (const double *)dqp,
                       (const double *)dqw, dout, total);
    if (ok(hipGetLastError()) && ok(hipMemcpyAsync(sdf_vol, dout, total * 8, hipMemcpyDeviceToHost, st)))
      ok(hipStreamSynchronize(st));
  }
  if (dqp) (void)hipFree(dqp);
  if (dqw) (void)hipFree(dqw);
  if (dout) (void)hipFree(dout);
  if (st) (void)hipStreamDestroy(st);
  return rc;
} catch (const std::exception &) {  // host vectors of the quadrature: a status, not an exception
  return PBF_E_CAPACITY;
}
