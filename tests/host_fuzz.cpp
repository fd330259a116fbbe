// Randomised and edge-case inputs for libpbf's GPU-free entry points (.cdm
// parser and writer, voxelizers, analytic map builders), built with
// AddressSanitizer and UndefinedBehaviorSanitizer by tests/test_host_sanitizers.py.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "pbf.h"

static std::mt19937_64 rng(12345);
static double U(double a, double b) { return std::uniform_real_distribution<double>(a, b)(rng); }
static int I(int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); }

static std::string make_cdm(int R, bool corrupt) {
  size_t n = (size_t)(R + 1) * (R + 1) * (R + 1) + 3 * (size_t)(R + 1) * (R + 1) * 2 * R;
  std::string s = std::to_string(R) + " " + std::to_string(R) + " " + std::to_string(R) + "\r\n" + std::to_string(n) + "\r\n";
  for (size_t k = 0; k < n; ++k) s += std::to_string(U(-1, 1)) + " ";
  s += "\r\n" + std::to_string(n) + "\r\n";
  for (size_t k = 0; k < n; ++k) s += std::to_string(U(0, 1e-5)) + " ";
  s += "\r\n";
  if (corrupt) {
    int m = I(1, 8);
    for (int q = 0; q < m; ++q) {
      size_t p = (size_t)I(0, (int)s.size() - 1);
      switch (I(0, 3)) {
        case 0: s[p] = (char)I(0, 255); break;
        case 1: s.erase(p, (size_t)I(1, 20)); break;
        case 2: s.insert(p, std::string((size_t)I(1, 20), (char)I(32, 126))); break;
        default: s.resize(p); break;
      }
      if (s.empty()) break;
    }
  }
  return s;
}

int main() {
  int ok = 0, fmt = 0;
  for (int it = 0; it < 400; ++it) {
    int R = I(1, 4);
    std::string t = make_cdm(R, it % 4 != 0);
    int32_t res[3];
    size_t fl = 0;
    int rc = pbf_parse_cdm(t.data(), t.size(), res, nullptr, nullptr, 0, &fl);
    if (rc == PBF_OK) {
      std::vector<float> a(fl), b(fl);
      rc = pbf_parse_cdm(t.data(), t.size(), res, a.data(), b.data(), fl, &fl);
      ++ok;
      std::vector<char> out(64 * fl + 256);
      size_t w = 0;
      pbf_write_cdm(res, a.data(), b.data(), fl, out.data(), out.size(), &w);
      pbf_write_cdm(res, a.data(), b.data(), fl, out.data(), 10, &w);  // too small
    } else {
      ++fmt;
    }
    // random bytes
    std::string junk((size_t)I(0, 200), '\0');
    for (auto &c : junk) c = (char)I(0, 255);
    pbf_parse_cdm(junk.data(), junk.size(), res, nullptr, nullptr, 0, &fl);
  }
  for (int it = 0; it < 200; ++it) {
    double mn[3], mx[3];
    for (int a = 0; a < 3; ++a) { mn[a] = U(-1, 1); mx[a] = mn[a] + U(-0.2, 0.6); }
    double r = U(0.005, 0.05);
    uint32_t cnt = 0;
    uint64_t cap = (uint64_t)I(0, 4000);
    std::vector<float> out(4 * (cap + 1));
    pbf_voxelize_cube_host(r, mn, mx, U(0, 0.01), 7, 0, out.data(), cap, &cnt);
    double c[3] = {U(-1, 1), U(-1, 1), U(-1, 1)};
    pbf_voxelize_sphere_host(r, c, U(-0.1, 0.3), U(0, 0.01), 7, 0, out.data(), cap, &cnt);
    // meshes: random triangle soups
    uint32_t nv = (uint32_t)I(0, 30), nt = (uint32_t)I(0, 20);
    std::vector<float> v(3 * (nv + 1));
    for (auto &x : v) x = (float)U(-0.5, 0.5);
    std::vector<uint32_t> idx(3 * (nt + 1));
    for (auto &x : idx) x = (uint32_t)I(0, (int)nv + 2);  // some out of range
    double M[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    pbf_voxelize_mesh_host(r * 4, v.data(), nv, idx.data(), nt, M, 0.0, 1, 0, out.data(), cap, &cnt);
  }
  for (int it = 0; it < 30; ++it) {
    int32_t res[3] = {I(1, 8), I(1, 8), I(1, 8)};
    double lo[3] = {0, 0, 0}, hi[3] = {U(0.1, 2), U(0.1, 2), U(0.1, 2)};
    double shape[9] = {0.5, 0.5, 0.5, 0.2, 0.5, 0.5, 0.5, 0.3, 0.08};
    size_t n = (size_t)(res[0] + 1) * (res[1] + 1) * (res[2] + 1);
    std::vector<float> sv(2 * n);
    pbf_build_volume_map(I(0, 1), shape, res, lo, hi, U(0.02, 0.2), U(0, 0.03), sv.data());
    size_t fl = 0;
    pbf_build_cubic_field(I(0, 1), shape, res, lo, hi, U(0.02, 0.2), U(0, 0.03), nullptr, nullptr, 0, &fl);
    std::vector<float> a(fl), b(fl);
    pbf_build_cubic_field(I(0, 1), shape, res, lo, hi, U(0.02, 0.2), U(0, 0.03), a.data(), b.data(), fl, &fl);
  }
  std::printf("fuzz done: %d parsed, %d rejected\n", ok, fmt);
  return 0;
}
// edge cases: NaN / inf / empty inputs
int edge_cases() {
  const double nan = std::nan(""), inf = 1.0 / 0.0;
  double mn[3] = {0, 0, 0}, mx[3] = {nan, 1, 1}, mx2[3] = {inf, 1, 1};
  uint32_t cnt = 0;
  std::vector<float> out(4 * 100);
  int r1 = pbf_voxelize_cube_host(0.01, mn, mx, 0, 1, 0, out.data(), 100, &cnt);
  int r2 = pbf_voxelize_cube_host(0.01, mn, mx2, 0, 1, 0, out.data(), 100, &cnt);
  int r3 = pbf_voxelize_cube_host(nan, mn, mn, 0, 1, 0, out.data(), 100, &cnt);
  int r4 = pbf_voxelize_cube_host(0.0, mn, mx, 0, 1, 0, out.data(), 100, &cnt);
  double c[3] = {0, 0, 0};
  int r5 = pbf_voxelize_sphere_host(0.01, c, inf, 0, 1, 0, out.data(), 100, &cnt);
  int r6 = pbf_voxelize_sphere_host(1e-30, c, 1.0, 0, 1, 0, nullptr, 0, &cnt);
  int32_t res[3];
  size_t fl;
  int r7 = pbf_parse_cdm("", 0, res, nullptr, nullptr, 0, &fl);
  int r8 = pbf_parse_cdm(nullptr, 5, res, nullptr, nullptr, 0, &fl);
  const char *big = "4096 4096 4096\r\n1\r\n0 \r\n1\r\n0 \r\n";
  int r9 = pbf_parse_cdm(big, std::strlen(big), res, nullptr, nullptr, 0, &fl);
  int32_t rr[3] = {1, 1, 1};
  double lo[3] = {0, 0, 0}, hi[3] = {nan, 1, 1};
  std::vector<float> sv(16);
  int r10 = pbf_build_volume_map(0, nullptr, rr, lo, hi, 0.1, 0.025, sv.data());
  std::printf("edge rcs: %d %d %d %d %d %d(count %u) %d %d %d %d\n", r1, r2, r3, r4, r5, r6, cnt, r7, r8, r9, r10);
  return 0;
}
static int dummy = edge_cases();
