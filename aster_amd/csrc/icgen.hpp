// Per-body draws of the seeded initial-condition generators, shared by the
// host generators (ic.cpp) and the device cluster loader (ic.hip). Body i of
// a generator depends only on (seed, i), and both builds evaluate the same
// operations in the same order, so host and device give identical bits.
#pragma once

#include <algorithm>
#include <cstdint>
#include <vector>

#include "icmath.hpp"

namespace aster {
namespace icg {

// Isotropic vector of length r: cos(polar) uniform in [-1, 1], azimuth
// uniform in [0, 2 pi). Two draws.
ICM_HD void iso(icm::Stream& s, double r, double* out) {
  const double ct = 1.0 - 2.0 * s.u();
  const double w = 1.0 - ct * ct;
  const double st = sqrt(w > 0.0 ? w : 0.0);
  double sp, cp;
  icm::sincos(2.0 * icm::kPi * s.u(), &sp, &cp);
  out[0] = r * st * cp;
  out[1] = r * st * sp;
  out[2] = r * ct;
}

// Plummer sphere body (Aarseth, Henon & Wielen 1974), scale radius b, radius
// truncated at r <= rmax (in units of b), velocity scale sqrt(G M / b):
//   r = (X^(-2/3) - 1)^(-1/2) by inversion of M(<r) = r^3 (1 + r^2)^(-3/2);
//   speed q v_esc with g(q) = q^2 (1 - q^2)^(7/2) sampled by von Neumann
//   rejection under the bound 0.1, v_esc = sqrt(2) (1 + r^2)^(-1/4);
//   mass_mode 1: raw mass 10^(2 U) (log-uniform over two decades), else 1.
ICM_HD void plummer_body(uint64_t seed, uint64_t i, double b, double rmax, double vscale, int mass_mode, double* p,
                         double* v, double* m) {
  icm::Stream s(seed, i);
  double r = rmax;
  for (int tries = 0; tries < 1000; ++tries) {
    const double x1 = s.u();
    if (x1 <= 0.0) continue;
    const double rt = 1.0 / sqrt(icm::pow(x1, -2.0 / 3.0) - 1.0);
    if (rt <= rmax) {
      r = rt;
      break;
    }
  }
  iso(s, r * b, p);
  double q = 0.0;
  for (int tries = 0; tries < 1000; ++tries) {
    const double x4 = s.u(), x5 = s.u();
    const double w = 1.0 - x4 * x4;
    const double g = x4 * x4 * (w * w * w * sqrt(w));
    q = x4;
    if (0.1 * x5 < g) break;
  }
  const double ve = sqrt(2.0) / sqrt(sqrt(1.0 + r * r)) * vscale;
  iso(s, q * ve, v);
  *m = mass_mode == 1 ? icm::exp(2.0 * s.u() * icm::kLn10) : 1.0;
}

// Uniform cube [-h, h]^3 at rest. Three draws.
ICM_HD void cube_body(uint64_t seed, uint64_t i, double h, double* p, double* v) {
  icm::Stream s(seed, i);
  for (int c = 0; c < 3; ++c) p[c] = -h + 2.0 * h * s.u();
  v[0] = v[1] = v[2] = 0.0;
}

// Disk preset (reference add_disk, src/clusters/presets_impl.cpp:207-255).
struct Disk {
  double radius, inner, thickness, scale;  // rng_point_in_cylinder(radius, inner, thickness) * scale
  double gm;                               // G * central mass
  double avr_mass;                         // body mass = U(0, 1] * avr_mass
  double center[3], vel[3];                // added before the rotation
  double cphi, sphi, cth, sth;             // rotate_point angles phi = rotation.x, theta = rotation.z
};

// rotate_point (presets_impl.cpp:47-64): a rotation by phi in the z-x plane,
// then by theta in the y-z plane; its intermediates are single precision, so
// every component is rounded to FP32.
ICM_HD void rotate_point(const double* v, const Disk& d, double* out) {
  const float x1 = (float)(v[0] * d.cphi + v[2] * d.sphi);
  const float z1 = (float)(-v[0] * d.sphi + v[2] * d.cphi);
  const float y1 = (float)(v[1] * d.cth + (double)z1 * d.sth);
  const float z2 = (float)(-v[1] * d.sth + (double)z1 * d.cth);
  out[0] = (double)x1;
  out[1] = (double)y1;
  out[2] = (double)z2;
}

// One disk body: rng_point_in_cylinder (presets_impl.cpp:130-147: angle
// uniform in [0, 2 pi), radius uniform in [inner, radius), height
// (U - 1/2) * (-thickness) as its sign flip leaves it), scaled; circular
// speed sqrt(G M / |p|) about the z axis; then translated and rotated.
ICM_HD void disk_body(uint64_t seed, uint64_t i, const Disk& d, double* p, double* v, double* m) {
  icm::Stream s(seed, i);
  const double theta = 2.0 * icm::kPi * s.u();
  const double rr = d.inner + (d.radius - d.inner) * s.u();
  const double z = (s.u() - 0.5) * -d.thickness;
  double st, ct;
  icm::sincos(theta, &st, &ct);
  const double pos[3] = {rr * ct * d.scale, rr * st * d.scale, z * d.scale};
  const double r = sqrt(pos[0] * pos[0] + pos[1] * pos[1] + pos[2] * pos[2]);
  const double mv = sqrt(d.gm / r);
  const double vel[3] = {pos[1] / r * mv, -pos[0] / r * mv, 0.0};
  *m = s.u_open0() * d.avr_mass;
  const double pc[3] = {pos[0] + d.center[0], pos[1] + d.center[1], pos[2] + d.center[2]};
  const double vc[3] = {vel[0] + d.vel[0], vel[1] + d.vel[1], vel[2] + d.vel[2]};
  rotate_point(pc, d, p);
  rotate_point(vc, d, v);
}

// Reduction shape shared with the device sums (ic.hip: k_cluster_sums).
constexpr int kSumThreads = 256;
constexpr int kSumBlocksMax = 512;

inline int sum_blocks(int64_t n) {
  return (int)std::min<int64_t>((n + kSumThreads - 1) / kSumThreads, kSumBlocksMax);
}

// Host evaluation of sums over i < n of get(i, vals) (nq values per body) in
// the device reduction's order: thread g of B * 256 adds i = g, g + B * 256,
// ... in turn; the 256 partials of a block are added pairwise with strides
// 128, 64, ..., 1; the block totals, padded with zeros to 512, are added
// pairwise with strides 256, 128, ..., 1.
template <class Get>
void ordered_sums(int64_t n, int nq, Get get, double* out) {
  const int B = sum_blocks(n);
  const int64_t stride = (int64_t)B * kSumThreads;
  std::vector<double> red((size_t)kSumThreads * nq);
  std::vector<double> tot((size_t)kSumBlocksMax * nq, 0.0);
  double vals[8];
  for (int b = 0; b < B; ++b) {
    for (int t = 0; t < kSumThreads; ++t) {
      double* s = &red[(size_t)t * nq];
      for (int q = 0; q < nq; ++q) s[q] = 0.0;
      for (int64_t i = (int64_t)b * kSumThreads + t; i < n; i += stride) {
        get(i, vals);
        for (int q = 0; q < nq; ++q) s[q] += vals[q];
      }
    }
    for (int w = kSumThreads / 2; w > 0; w >>= 1)
      for (int t = 0; t < w; ++t)
        for (int q = 0; q < nq; ++q) red[(size_t)t * nq + q] += red[(size_t)(t + w) * nq + q];
    for (int q = 0; q < nq; ++q) tot[(size_t)b * nq + q] = red[q];
  }
  for (int w = kSumBlocksMax / 2; w > 0; w >>= 1)
    for (int t = 0; t < w; ++t)
      for (int q = 0; q < nq; ++q) tot[(size_t)t * nq + q] += tot[(size_t)(t + w) * nq + q];
  for (int q = 0; q < nq; ++q) out[q] = tot[q];
}

}  // namespace icg
}  // namespace aster
