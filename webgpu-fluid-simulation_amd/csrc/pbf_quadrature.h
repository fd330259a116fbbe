// pbf_quadrature.h -- quadrature of the reference-format boundary volume,
// shared by the host and GPU map builders (header-only, plain C++).
//
// The reference's boundary fields (public/model/*.cdm, read by
// src/simulator/boundary/volumeMap.ts:23-39) store, at every node x,
//   V(x) = 0.8 * sum over the 16^3 Gauss-Legendre points xi of [-h, h]^3
//          with |xi| <= h of  w(xi) * gamma(phi(x + xi)),
//   gamma(phi) = 1 (phi <= 0, solid), W(phi) / W(0) of the cubic spline
//   kernel of support h (0 < phi < h), 0 beyond, and 0 for points outside
//   the field's box,
// the boundary volume of Bender et al. 2019 ("Volume Maps", MIG 2019) as
// evaluated by that quadrature: it reproduces both reference files to their
// printed precision.
#pragma once
#include <cmath>
#include <vector>

namespace pbf {

// n-point Gauss-Legendre nodes and weights on [-1, 1] (Newton iteration on
// P_n from the Chebyshev-like initial guesses, Golub & Welsch's classical
// alternative being an eigenproblem).
inline void gauss_legendre(int n, std::vector<double> &x, std::vector<double> &w) {
  const double PI = 3.141592653589793;
  x.assign(n, 0.0);
  w.assign(n, 0.0);
  for (int i = 0; i < n; ++i) {
    double z = std::cos(PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
    for (int it = 0; it < 100; ++it) {
      double p0 = 1.0, p1 = z;
      for (int k = 2; k <= n; ++k) {
        const double p2 = ((2.0 * k - 1.0) * z * p1 - (k - 1.0) * p0) / k;
        p0 = p1;
        p1 = p2;
      }
      dp = n * (z * p1 - p0) / (z * z - 1.0);
      const double dz = p1 / dp;
      z -= dz;
      if (std::fabs(dz) < 1e-16) break;
    }
    x[i] = z;
    w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
  }
}

// Points (3 per entry) and weights of the reference volume's quadrature for
// support h: the 16^3 tensor rule on [-h, h]^3 restricted to the ball, each
// weight times the 0.8 factor.
inline void reference_volume_quadrature(double h, std::vector<double> &pts, std::vector<double> &wts) {
  std::vector<double> x, w;
  gauss_legendre(16, x, w);
  pts.clear();
  wts.clear();
  for (int i = 0; i < 16; ++i)
    for (int j = 0; j < 16; ++j)
      for (int k = 0; k < 16; ++k) {
        const double o[3] = {x[i] * h, x[j] * h, x[k] * h};
        if (o[0] * o[0] + o[1] * o[1] + o[2] * o[2] > h * h) continue;
        pts.push_back(o[0]);
        pts.push_back(o[1]);
        pts.push_back(o[2]);
        wts.push_back(0.8 * (w[i] * w[j] * w[k]) * (h * h * h));
      }
}

}  // namespace pbf
