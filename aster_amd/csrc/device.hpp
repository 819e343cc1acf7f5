// Device helpers shared by the HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

namespace aster {

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

// compute_gravitational (tool_chain_impl.cpp:120-137): force on body 1 from
// body 2 in the reference's operation order (IEEE division and square root,
// no contraction). The caller divides by m1 (builder_endpoint.cpp:747-752) or
// multiplies by 1/m1 (barnes_hut.cpp:238-239).
__device__ __forceinline__ void grav64(double m1, double m2, double v1x, double v1y, double v1z, double v2x,
                                       double v2y, double v2z, double p1x, double p1y, double p1z, double p2x,
                                       double p2y, double p2z, double G, double ic2, double soft, bool pn,
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
  if (pn) {
    double px, py, pz;
    pn1_term64(m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, -nx, -ny, -nz, G, px, py, pz);
    fx = fx + (px * ic2) * m1;
    fy = fy + (py * ic2) * m1;
    fz = fz + (pz * ic2) * m1;
  }
}

// 1/sqrt(x) in FP64: hardware estimate plus two Newton steps (relative error
// a few ulps). Not bit-identical to 1.0/sqrt(x).
__device__ __forceinline__ double rsqrt64(double x) {
  double y = __builtin_amdgcn_rsq(x);
#pragma unroll
  for (int it = 0; it < 2; ++it) {
    const double p = x * y;
    const double e = fma(-p, y, 1.0);
    y = fma(0.5 * y, e, y);
  }
  return y;
}

// Kahan-compensated FP32 accumulation: s + (-c) carries the running sum.
__device__ __forceinline__ void kahan_add32(float& s, float& c, float t) {
  const float y = t - c;
  const float u = s + y;
  c = (u - s) - y;
  s = u;
}

// Compiler and memory ordering point for LDS traffic between the lanes of one
// wavefront (no cross-wave synchronisation).
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
