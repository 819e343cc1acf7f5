// Device helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace aster {

// One coordinate of a body's kick/drift (update_leapfrog,
// src/physics/tool_chain_impl.cpp:169-185): the previous stage's second
// half-kick (dt_prev, when not 0), the drift and this stage's first
// half-kick, all with the same acceleration, in the reference's operation
// order and without contraction.
__device__ __forceinline__ void kick_drift_1(double& x, double& v, double a, double dt_prev, double dt) {
#pragma clang fp contract(off)
  double vc = v;
  if (dt_prev != 0.0) vc = vc + 0.5 * a * dt_prev;
  x = x + (vc * dt + 0.5 * a * dt * dt);
  v = vc + 0.5 * a * dt;
}

// get_pn1 (src/physics/tool_chain_impl.cpp:110-118) evaluated in the
// reference's operation order; `n` is the caller's -norm (:133).
__device__ __forceinline__ void pn1_term64(double m1, double m2, double v1x, double v1y, double v1z, double v2x,
                                           double v2y, double v2z, double inv_r, double nx, double ny, double nz,
                                           double G, double& ax, double& ay, double& az) {
#pragma clang fp contract(off)
  const double v2_sq = v2x * v2x + v2y * v2y + v2z * v2z;
  const double v1_sq = v1x * v1x + v1y * v1y + v1z * v1z;
  const double dot4 = (4.0 * v1x) * v2x + (4.0 * v1y) * v2y + (4.0 * v1z) * v2z;
  const double nv2 = nx * v2x + ny * v2y + nz * v2z;
  double s = -v1_sq - 2.0 * v2_sq;
  s = s + dot4;
  s = s + 1.5 * (nv2 * nv2);
  s = s + 5.0 * G * m1 * inv_r;
  s = s + 4.0 * G * m2 * inv_r;
  ax = nx * s;
  ay = ny * s;
  az = nz * s;
  const double n4v1 = (nx * 4.0) * v1x + (ny * 4.0) * v1y + (nz * 4.0) * v1z;
  const double n3v2 = (nx * 3.0) * v2x + (ny * 3.0) * v2y + (nz * 3.0) * v2z;
  const double t = n4v1 - n3v2;
  ax = ax + (v1x - v2x) * t;
  ay = ay + (v1y - v2y) * t;
  az = az + (v1z - v2z) * t;
  ax = ax * G * m2 * inv_r * inv_r;
  ay = ay * G * m2 * inv_r * inv_r;
  az = az * G * m2 * inv_r * inv_r;
}

// get_pn2 (tool_chain_impl.cpp:87-107), the two-body 2PN acceleration of
// body 1, in the reference's operation order: every product left to right,
// every sum left to right, the fraction constants folded as written (3.0/2.0,
// 52.0/3.0, ...). `n` is the caller's -norm.
__device__ __forceinline__ void pn2_term64(double m1, double m2, double v1x, double v1y, double v1z, double v2x,
                                           double v2y, double v2z, double inv_r, double nx, double ny, double nz,
                                           double G, double& ax, double& ay, double& az) {
#pragma clang fp contract(off)
  const double v2_sq = v2x * v2x + v2y * v2y + v2z * v2z;
  const double v1_sq = v1x * v1x + v1y * v1y + v1z * v1z;
  const double v_prod = v1x * v2x + v1y * v2y + v1z * v2z;
  const double nv1 = nx * v1x + ny * v1y + nz * v1z;
  const double nv2 = nx * v2x + ny * v2y + nz * v2z;
  const double nv22 = nv2 * nv2;
  const double pn_1 = G * m1 * inv_r;
  const double pn_2 = G * m2 * inv_r;
  const double nv12 = nv1 * nv1;
  double A = -2.0 * v2_sq * v2_sq + 4.0 * v2_sq * v_prod;
  A = A - 2.0 * v_prod * v_prod;
  A = A + 3.0 / 2.0 * v1_sq * nv22;
  A = A + 9.0 / 2.0 * v2_sq * nv22;
  A = A - 6.0 * v_prod * nv22;
  A = A - 15.0 / 8.0 * nv22 * nv22;
  double g1 = -15.0 / 4.0 * v1_sq + 5.0 / 4.0 * v2_sq;
  g1 = g1 - 5.0 / 2.0 * v_prod;
  g1 = g1 + 39.0 / 2.0 * nv12;
  g1 = g1 - 39.0 * nv1 * nv2;
  g1 = g1 + 17.0 / 2.0 * nv22;
  A = A + pn_1 * g1;
  double g2 = 4.0 * v2_sq - 8.0 * v_prod;
  g2 = g2 + 2.0 * nv12;
  g2 = g2 - 4.0 * nv1 * nv2;
  g2 = g2 - 6.0 * nv22;
  A = A + pn_2 * g2;
  double B = v1_sq * nv2 + 4.0 * v2_sq * nv1;
  B = B - 5.0 * v2_sq * nv2;
  B = B - 4.0 * v_prod * nv1;
  B = B + 4.0 * v_prod * nv2;
  B = B - 6.0 * nv1 * nv22;
  B = B + 9.0 / 2.0 * nv22 * nv2;
  B = B + pn_1 * (-63.0 / 4.0 * nv1 + 55.0 / 4.0 * nv2);
  B = B + pn_2 * (-2.0 * nv1 - 2.0 * nv2);
  const double s1 = G * m2 * inv_r * inv_r;
  const double s2 = G * G * G * m2 * inv_r * inv_r * inv_r * inv_r;
  double C = -57.0 / 4.0 * m1 * m1 - 9.0 * m2 * m2;
  C = C - 69.0 / 2.0 * m1 * m2;
  ax = (nx * A + (v1x - v2x) * B) * s1 + (nx * s2) * C;
  ay = (ny * A + (v1y - v2y) * B) * s1 + (ny * s2) * C;
  az = (nz * A + (v1z - v2z) * B) * s1 + (nz * s2) * C;
}

// get_pn25 (tool_chain_impl.cpp:73-84), the two-body 2.5PN radiation-reaction
// acceleration of body 1, in the reference's operation order. `n` is the
// caller's -norm.
__device__ __forceinline__ void pn25_term64(double m1, double m2, double v1x, double v1y, double v1z, double v2x,
                                            double v2y, double v2z, double inv_r, double nx, double ny, double nz,
                                            double G, double& ax, double& ay, double& az) {
#pragma clang fp contract(off)
  const double pn_1 = G * m1 * inv_r;
  const double pn_2 = G * m2 * inv_r;
  const double dvx = v1x - v2x, dvy = v1y - v2y, dvz = v1z - v2z;
  const double sq = dvx * dvx + dvy * dvy + dvz * dvz;
  const double k1 = -sq + 2.0 * pn_1 - 8.0 * pn_2;
  const double ndv = (nx * v1x + ny * v1y + nz * v1z) - (nx * v2x + ny * v2y + nz * v2z);
  const double k2 = 3.0 * sq - 6.0 * pn_1 + 52.0 / 3.0 * pn_2;
  double w[3] = {dvx * k1 + (nx * ndv) * k2, dvy * k1 + (ny * ndv) * k2, dvz * k1 + (nz * ndv) * k2};
#pragma unroll
  for (int c = 0; c < 3; ++c) w[c] = w[c] * 4.0 / 5.0 * G * G * m1 * m2 * inv_r * inv_r * inv_r;
  ax = w[0];
  ay = w[1];
  az = w[2];
}

// The post-Newtonian terms of compute_gravitational (tool_chain_impl.cpp:
// 129-136) that `pnf` selects (ASTER_PN1 0x10, ASTER_PN2 0x100, ASTER_PN25
// 0x1000), each scaled by 1/c^k (products of ic = 1/c taken left to right)
// and by m1, added to the force f in the reference's order.
__device__ __forceinline__ void pn_add64(int pnf, double m1, double m2, double v1x, double v1y, double v1z,
                                         double v2x, double v2y, double v2z, double inv_r, double nx, double ny,
                                         double nz, double G, double ic, double& fx, double& fy, double& fz) {
#pragma clang fp contract(off)
  double px, py, pz;
  if (pnf & 0x10) {
    const double ic2 = ic * ic;
    pn1_term64(m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, nx, ny, nz, G, px, py, pz);
    fx = fx + (px * ic2) * m1;
    fy = fy + (py * ic2) * m1;
    fz = fz + (pz * ic2) * m1;
  }
  if (pnf & 0x100) {
    const double ic4 = ic * ic * ic * ic;
    pn2_term64(m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, nx, ny, nz, G, px, py, pz);
    fx = fx + (px * ic4) * m1;
    fy = fy + (py * ic4) * m1;
    fz = fz + (pz * ic4) * m1;
  }
  if (pnf & 0x1000) {
    const double ic5 = ic * ic * ic * ic * ic;
    pn25_term64(m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, nx, ny, nz, G, px, py, pz);
    fx = fx + (px * ic5) * m1;
    fy = fy + (py * ic5) * m1;
    fz = fz + (pz * ic5) * m1;
  }
}

// compute_gravitational (tool_chain_impl.cpp:120-137): force on body 1 from
// body 2 in the reference's operation order (IEEE division and square root,
// no contraction), with the post-Newtonian terms `pnf` selects. ic = 1/c.
// The caller divides by m1 (builder_endpoint.cpp:747-752) or multiplies by
// 1/m1 (barnes_hut.cpp:238-239).
__device__ __forceinline__ void grav64(double m1, double m2, double v1x, double v1y, double v1z, double v2x,
                                       double v2y, double v2z, double p1x, double p1y, double p1z, double p2x,
                                       double p2y, double p2z, double G, double ic, double soft, int pnf,
                                       double& fx, double& fy, double& fz) {
#pragma clang fp contract(off)
  const double dx = p2x - p1x, dy = p2y - p1y, dz = p2z - p1z;
  const double r2 = dx * dx + dy * dy + dz * dz + soft;
  const double inv_r = 1.0 / sqrt(r2);
  const double inv_r2 = inv_r * inv_r;
  const double nx = dx * inv_r, ny = dy * inv_r, nz = dz * inv_r;
  const double scal = G * m2 * m1 * inv_r2;
  fx = nx * scal;
  fy = ny * scal;
  fz = nz * scal;
  if (pnf) pn_add64(pnf, m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, -nx, -ny, -nz, G, ic, fx, fy, fz);
}

// 1/sqrt(x) in FP64: hardware estimate (v_rsq_f64, relative error up to
// 5.2e-8) plus Newton steps y += y/2 (1 - x y^2). Measured over log-uniform x
// in [1e-30, 1e30] against a long double reference: one step 4.1e-15, two
// steps 1.4e-16 (under one ulp). Not bit-identical to 1.0/sqrt(x).
template <int STEPS>
__device__ __forceinline__ double rsqrt64_n(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < STEPS; ++it) {
    const double p = x * y;
    const double e = fma(-p, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}
__device__ __forceinline__ double rsqrt64(double x) { return rsqrt64_n<2>(x); }

// Kahan-compensated FP32 accumulation: s + (-c) carries the running sum.
__device__ __forceinline__ void kahan_add32(float& s, float& c, float t) {
  const float y = t - c;
  const float u = s + y;
  c = (u - s) - y;
  s = u;
}

// Compiler and memory ordering point for LDS traffic between the lanes of one
// wavefront (no cross-wave synchronisation).
// Stumpff functions c0..c3 of z: quarter z until |z| <= 0.1, 8-term nested
// series (multiplying by reciprocal constants folded at compile time, no
// division), then c3 = (c2 + c0 c3)/4, c2 = c1^2/2, c1 = c0 c1,
// c0 = 2 c0^2 - 1 once per quartering. Only +, -, *, / in a fixed order, so the result is
// the same bits on any IEEE machine (contraction off).
__device__ __forceinline__ void stumpff64(double z, double& c0, double& c1, double& c2, double& c3) {
#pragma clang fp contract(off)
  int nq = 0;
  while (fabs(z) > 0.1) {
    z = z * 0.25;
    ++nq;
  }
  c2 = 0.5 * (1.0 - z * (1.0 / 12.0) * (1.0 - z * (1.0 / 30.0) * (1.0 - z * (1.0 / 56.0) *
       (1.0 - z * (1.0 / 90.0) * (1.0 - z * (1.0 / 132.0) * (1.0 - z * (1.0 / 182.0) *
       (1.0 - z * (1.0 / 240.0) * (1.0 - z * (1.0 / 306.0)))))))));
  c3 = (1.0 / 6.0) * (1.0 - z * (1.0 / 20.0) * (1.0 - z * (1.0 / 42.0) * (1.0 - z * (1.0 / 72.0) *
       (1.0 - z * (1.0 / 110.0) * (1.0 - z * (1.0 / 156.0) * (1.0 - z * (1.0 / 210.0) *
       (1.0 - z * (1.0 / 272.0) * (1.0 - z * (1.0 / 342.0)))))))));
  c1 = 1.0 - z * c3;
  c0 = 1.0 - z * c2;
  for (int k = 0; k < nq; ++k) {
    c3 = (c2 + c0 * c3) * 0.25;
    c2 = c1 * c1 * 0.5;
    c1 = c0 * c1;
    c0 = 2.0 * c0 * c0 - 1.0;
  }
}

// Universal Kepler equation F(s) = r0 G1 + eta0 G2 + mu G3 - dt with its
// first two derivatives (G_k = s^k c_k(beta s^2)).
__device__ __forceinline__ void kepler_eval64(double r0, double eta0, double zeta0, double beta, double mu, double s,
                                              double dt, double& F, double& Fp, double& Fpp) {
#pragma clang fp contract(off)
  double c0, c1, c2, c3;
  stumpff64(beta * s * s, c0, c1, c2, c3);
  const double g1 = s * c1, g2 = s * s * c2, g3 = s * s * s * c3;
  F = r0 * g1 + eta0 * g2 + mu * g3 - dt;
  Fp = r0 * c0 + eta0 * g1 + mu * g2;
  Fpp = eta0 * c0 + zeta0 * g1;
}

// Kepler drift of relative position q and velocity u about mu for dt
// (universal variables). F is increasing (F' = r > 0): on a bound orbit dt is
// reduced modulo the period (fmod is exact) and the root lies within one
// revolution, s in [0, 2 pi / sqrt(beta)]; otherwise the bracket is grown by
// doubling. Halley steps from s = dt / r0, replaced by bisection whenever they
// leave the bracket, at most 100, until |ds| <= 1e-15 |s|; then the f and g
// functions. Every lane runs its own iteration count.
__device__ inline void kepler_drift64(double* q, double* u, double mu, double dt) {
#pragma clang fp contract(off)
  const double r0 = sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  const double v2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  const double eta0 = q[0] * u[0] + q[1] * u[1] + q[2] * u[2];
  const double beta = 2.0 * mu / r0 - v2;
  const double zeta0 = mu - beta * r0;
  const double sgn = dt < 0.0 ? -1.0 : 1.0;
  double F, Fp, Fpp;
  double lo = 0.0, hi;
  if (beta > 0.0) {
    const double sq = sqrt(beta);
    dt = fmod(dt, 6.283185307179586 * mu / (beta * sq));
    hi = sgn * (6.283185307179586 / sq);
  } else {
    hi = dt / r0;
    for (int k = 0; k < 200; ++k) {
      kepler_eval64(r0, eta0, zeta0, beta, mu, hi, dt, F, Fp, Fpp);
      if (sgn * F >= 0.0) break;
      lo = hi;
      hi = 2.0 * hi;
    }
  }
  double s = dt / r0;
  if (!(sgn * (s - lo) > 0.0 && sgn * (hi - s) > 0.0)) s = 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    kepler_eval64(r0, eta0, zeta0, beta, mu, s, dt, F, Fp, Fpp);
    if (sgn * F < 0.0)
      lo = s;
    else
      hi = s;
    double sn = s - F / (Fp - 0.5 * F * Fpp / Fp);
    if (!(sgn * (sn - lo) > 0.0 && sgn * (hi - sn) > 0.0)) sn = 0.5 * (lo + hi);
    const double ds = sn - s;
    s = sn;
    if (fabs(ds) <= 1e-15 * fabs(s)) break;
  }
  double c0, c1, c2, c3;
  stumpff64(beta * s * s, c0, c1, c2, c3);
  const double g1 = s * c1, g2 = s * s * c2;
  const double r = r0 * c0 + eta0 * g1 + mu * g2;
  const double f = 1.0 - mu * g2 / r0;
  const double g = r0 * g1 + eta0 * g2;
  const double fd = -mu * g1 / (r * r0);
  const double gd = 1.0 - mu * g2 / r;
  double qn[3], un[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    qn[k] = f * q[k] + g * u[k];
    un[k] = fd * q[k] + gd * u[k];
  }
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    q[k] = qn[k];
    u[k] = un[k];
  }
}

// A value known to be equal in every active lane, moved to a scalar register.
__device__ __forceinline__ float uniform_f32(float x) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x)));
}


__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

}  // namespace aster
