// All-pairs gravitational accelerations on CDNA4.
//
// FP32 paths: one target body per lane; source bodies are read with
// wave-uniform addresses, so they arrive through the scalar cache into SGPRs
// and feed VALU instructions directly (no LDS staging, no barriers). The
// source range is split into slices (grid.y) so that a 64K-body problem puts
// several waves on every SIMD; slice partials are combined in FP64 in a fixed
// order, so results are deterministic. Per interaction the plain FP32 loop
// issues 3 sub + 3 fma (r^2) + 1 rsq + 3 mul + 3 fma = 12 VALU + 1
// transcendental; the mixed method forms the separation in FP64 and rounds it
// (3 add_f64 + 3 cvt instead of 3 sub: 15 VALU + 1 transcendental) and
// accumulates with Kahan compensation across 64-source tiles.
//
// FP64 path: the reference's operation order (Parallelized::compute_forces,
// src/utils/builder_endpoint.cpp:729-757, with compute_gravitational,
// src/physics/tool_chain_impl.cpp:120-137), each target's sources summed in j
// order, IEEE division and square root, no contraction: bit-identical to the
// CPU path. The pair contributions of a chunk of sources are formed by a whole
// block and summed per target afterwards.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <mutex>
#include <string>

#include "common.hpp"
#include "device.hpp"

namespace aster {

// ---------------------------------------------------------------- FP32 ----

enum { ACC_PLAIN = 0, ACC_TILE_KAHAN = 1, ACC_FULL_KAHAN = 2 };

__device__ __forceinline__ void kahan_add(float& s, float& c, float t) { kahan_add32(s, c, t); }

// Sources, padded to npad with zero-mass bodies at the origin. Bodies outside
// the source range [lower, upper) (Solver::set_bounds) get zero mass, which
// adds exactly zero force. Masses are scaled by msc (Fp32Scales).
//   k_src32: FP32 {x, y, z, m} (positions rounded to FP32 once).
//   k_src64: FP64 {x, y, z, m}; the kernel forms x_j - x_i in FP64 and rounds
//            the separation to FP32, so the pair force is FP32 but a close pair
//            does not lose ~|x|/r ulps to position rounding.
__global__ __launch_bounds__(256) void k_src32(int n, int npad, int lower, int upper, float msc,
                                               const double* __restrict__ x, const double* __restrict__ m,
                                               float4* __restrict__ src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const bool in = i >= lower && i < upper;
  src[i] = i < n ? make_float4((float)x[i], (float)x[n + i], (float)x[2 * n + i], in ? (float)m[i] * msc : 0.f)
                 : make_float4(0.f, 0.f, 0.f, 0.f);
}

// The w slot carries the FP32 mass in its low 32 bits, so the kernel reads it
// straight from the low SGPR of the pair without a conversion.
__device__ __forceinline__ double pack_mass(float m) {
  return __longlong_as_double((long long)(unsigned long long)__float_as_uint(m));
}
__device__ __forceinline__ float unpack_mass(double w) {
  return __uint_as_float((unsigned)(unsigned long long)__double_as_longlong(w));
}

__global__ __launch_bounds__(256) void k_src64(int n, int npad, int lower, int upper, float msc,
                                               const double* __restrict__ x, const double* __restrict__ m,
                                               double4* __restrict__ src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const bool in = i >= lower && i < upper;
  src[i] = i < n ? make_double4(x[i], x[n + i], x[2 * n + i], pack_mass(in ? (float)m[i] * msc : 0.f))
                 : make_double4(0.0, 0.0, 0.0, pack_mass(0.f));
}

__device__ __forceinline__ float src_mass(const float4& q) { return q.w; }
__device__ __forceinline__ float src_mass(const double4& q) { return unpack_mass(q.w); }

template <bool SRC64>
struct SrcT;
template <>
struct SrcT<false> {
  using type = float4;
};
template <>
struct SrcT<true> {
  using type = double4;
};

// One target per lane (natural order), sources [s*slice_len, ...) of the
// padded source array; out: [S][3][n] doubles.
template <int ACC, bool MASK, bool SRC64>
__global__ __launch_bounds__(256) void k_allpairs_f32(const typename SrcT<SRC64>::type* __restrict__ src,
                                                      const double* __restrict__ x, int n, int npad, int slice_len,
                                                      float eps2, double* __restrict__ out) {
  const int i = blockIdx.x * kApBlock + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int s = blockIdx.y;
  const int jb = s * slice_len;
  const int je = min(jb + slice_len, npad);
  const double px = x[ii], py = x[n + ii], pz = x[2 * n + ii];
  const float fx = (float)px, fy = (float)py, fz = (float)pz;
  float sx = 0.f, sy = 0.f, sz = 0.f;
  float cx = 0.f, cy = 0.f, cz = 0.f;
  for (int t = jb; t < je; t += kApTile) {
    float tx = 0.f, ty = 0.f, tz = 0.f;
#pragma unroll 16
    for (int k = 0; k < kApTile; ++k) {
      const auto q = src[t + k];
      float dx, dy, dz;
      if (SRC64) {
        dx = (float)(q.x - px);
        dy = (float)(q.y - py);
        dz = (float)(q.z - pz);
      } else {
        dx = q.x - fx;
        dy = q.y - fy;
        dz = q.z - fz;
      }
      float r2 = fmaf(dx, dx, eps2);
      r2 = fmaf(dy, dy, r2);
      r2 = fmaf(dz, dz, r2);
      float inv = __builtin_amdgcn_rsqf(r2);
      if (MASK) inv = (r2 > 0.f) ? inv : 0.f;
      const float inv2 = inv * inv;
      const float f = (src_mass(q) * inv) * inv2;
      if (ACC == ACC_FULL_KAHAN) {
        kahan_add(sx, cx, f * dx);
        kahan_add(sy, cy, f * dy);
        kahan_add(sz, cz, f * dz);
      } else if (ACC == ACC_TILE_KAHAN) {
        tx = fmaf(f, dx, tx);
        ty = fmaf(f, dy, ty);
        tz = fmaf(f, dz, tz);
      } else {
        sx = fmaf(f, dx, sx);
        sy = fmaf(f, dy, sy);
        sz = fmaf(f, dz, sz);
      }
    }
    if (ACC == ACC_TILE_KAHAN) {
      kahan_add(sx, cx, tx);
      kahan_add(sy, cy, ty);
      kahan_add(sz, cz, tz);
    }
  }
  if (i < n) {
    double* o = out + (size_t)s * 3 * n;
    o[i] = (double)sx - (double)cx;
    o[n + i] = (double)sy - (double)cy;
    o[2 * n + i] = (double)sz - (double)cz;
  }
}

// Sum S slice partials (and the optional FP64 1PN term) in slice order, then
// scale by G (times the mass scale's 2^k): a[c][i] = G * sum_s part[s][c][i]
// + pn[c][i].
__global__ void k_combine(int n, int S, const double* __restrict__ part, const double* __restrict__ pn, double G,
                          double* __restrict__ a) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // 8 slices per round, their 24 loads issued before the first add (clamped
  // slice index); the adds stay in slice order
  double s[3] = {0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < S; k0 += 8) {
    double v[8][3];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int k = min(k0 + u, S - 1);
#pragma unroll
      for (int c = 0; c < 3; ++c) v[u][c] = part[((size_t)k * 3 + c) * n + i];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (k0 + u < S)
#pragma unroll
        for (int c = 0; c < 3; ++c) s[c] += v[u][c];
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double r = G * s[c];
    if (pn) r += pn[(size_t)c * n + i];
    a[(size_t)c * n + i] = r;
  }
}

// ---------------------------------------------------------------- FP64 ----

// FP64 accelerations in the reference's arithmetic and order: for target i,
// sum over sources j in [lower, upper), j != i, in j order, of f/m_i with f =
// compute_gravitational (Parallelized::compute_forces,
// src/utils/builder_endpoint.cpp:729-757; tool_chain_impl.cpp:120-137), IEEE
// division and square root, no contraction. MODE 0: Newton (+ the
// post-Newtonian terms pnf selects if PN); MODE 1: only the post-Newtonian
// terms (added to an FP32 Newtonian result). x, v: SoA planes [3][n]; ic = 1/c.
// A block owns kF64Targets targets and walks the sources in chunks of
// kF64Chunk: all 256 threads form the chunk's pair contributions f/m_i
// (4 each) into LDS, then each target's thread adds its row in j order (the
// self pair and sources past upper are skipped by a select). Every target sums
// the same doubles in the same order as a sequential loop would, so the
// result is bit-identical; a 4096-body problem runs 256 blocks instead of 16
// wavefronts of long dependent chains.
constexpr int kF64Targets = 16;
constexpr int kF64Chunk = 64;

template <bool PN, int MODE>
__global__ __launch_bounds__(256) void k_allpairs_f64_tiled(int n, int lower, int upper,
                                                            const double* __restrict__ x,
                                                            const double* __restrict__ v,
                                                            const double* __restrict__ m, double G, double ic,
                                                            int pnf, double soft, double* __restrict__ a) {
#pragma clang fp contract(off)
  __shared__ double cf[3][kF64Targets][kF64Chunk + 1];
  const int t = threadIdx.x;
  const int ti = t % kF64Targets, sj = t / kF64Targets;  // this thread's target and first source slot
  const int i = blockIdx.x * kF64Targets + ti;
  const int ii = i < n ? i : n - 1;
  const double p1x = x[ii], p1y = x[n + ii], p1z = x[2 * n + ii];
  const double m1 = m[ii];
  double v1x = 0.0, v1y = 0.0, v1z = 0.0;
  if (PN) {
    v1x = v[ii];
    v1y = v[n + ii];
    v1z = v[2 * n + ii];
  }
  double ax = 0.0, ay = 0.0, az = 0.0;  // the sums (threads t < kF64Targets)
  for (int jb = lower; jb < upper; jb += kF64Chunk) {
#pragma unroll
    for (int q = 0; q < kF64Chunk * kF64Targets / 256; ++q) {
      const int s = sj + q * (256 / kF64Targets);
      const int j = min(jb + s, upper - 1);
      const double p2x = x[j], p2y = x[n + j], p2z = x[2 * n + j];
      const double m2 = m[j];
      const double dx = p2x - p1x, dy = p2y - p1y, dz = p2z - p1z;
      const double r2 = dx * dx + dy * dy + dz * dz + soft;
      const double inv_r = 1.0 / sqrt(r2);
      const double inv_r2 = inv_r * inv_r;
      const double nx = dx * inv_r, ny = dy * inv_r, nz = dz * inv_r;
      double fx = 0.0, fy = 0.0, fz = 0.0;
      if (MODE == 0) {
        const double scal = G * m2 * m1 * inv_r2;
        fx = nx * scal;
        fy = ny * scal;
        fz = nz * scal;
      }
      if (PN) {
        const double v2x = v[j], v2y = v[n + j], v2z = v[2 * n + j];
        pn_add64(pnf, m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, -nx, -ny, -nz, G, ic, fx, fy, fz);
      }
      cf[0][ti][s] = fx / m1;
      cf[1][ti][s] = fy / m1;
      cf[2][ti][s] = fz / m1;
    }
    __syncthreads();
    if (t < kF64Targets) {
      const int cnt = min(kF64Chunk, upper - jb);
      for (int s = 0; s < kF64Chunk; ++s) {
        const bool take = s < cnt && jb + s != ii;
        const double nax = ax + cf[0][t][s], nay = ay + cf[1][t][s], naz = az + cf[2][t][s];
        ax = take ? nax : ax;
        ay = take ? nay : ay;
        az = take ? naz : az;
      }
    }
    __syncthreads();  // the next chunk rewrites the contributions
  }
  if (t < kF64Targets && i < n) {
    a[i] = ax;
    a[n + i] = ay;
    a[2 * n + i] = az;
  }
}

// One target per thread, the sources in j order (the same sums as
// k_allpairs_f64_tiled, which is faster when there are few targets): four
// sources per iteration, whose pair forces are independent, added one by one.
template <bool PN, int MODE>
__global__ __launch_bounds__(256) void k_allpairs_f64(int n, int lower, int upper, const double* __restrict__ x,
                                                      const double* __restrict__ v, const double* __restrict__ m,
                                                      double G, double ic, int pnf, double soft,
                                                      double* __restrict__ a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const double p1x = x[ii], p1y = x[n + ii], p1z = x[2 * n + ii];
  const double m1 = m[ii];
  double v1x = 0.0, v1y = 0.0, v1z = 0.0;
  if (PN) {
    v1x = v[ii];
    v1y = v[n + ii];
    v1z = v[2 * n + ii];
  }
  double ax = 0.0, ay = 0.0, az = 0.0;
#pragma unroll 4
  for (int j = lower; j < upper; ++j) {
    const double p2x = x[j], p2y = x[n + j], p2z = x[2 * n + j];
    const double m2 = m[j];
    const double dx = p2x - p1x, dy = p2y - p1y, dz = p2z - p1z;
    const double r2 = dx * dx + dy * dy + dz * dz + soft;
    const double inv_r = 1.0 / sqrt(r2);
    const double inv_r2 = inv_r * inv_r;
    const double nx = dx * inv_r, ny = dy * inv_r, nz = dz * inv_r;
    double fx = 0.0, fy = 0.0, fz = 0.0;
    if (MODE == 0) {
      const double scal = G * m2 * m1 * inv_r2;
      fx = nx * scal;
      fy = ny * scal;
      fz = nz * scal;
    }
    if (PN) {
      const double v2x = v[j], v2y = v[n + j], v2z = v[2 * n + j];
      pn_add64(pnf, m1, m2, v1x, v1y, v1z, v2x, v2y, v2z, inv_r, -nx, -ny, -nz, G, ic, fx, fy, fz);
    }
    const double nax = ax + fx / m1, nay = ay + fy / m1, naz = az + fz / m1;
    ax = j == ii ? ax : nax;
    ay = j == ii ? ay : nay;
    az = j == ii ? az : naz;
  }
  if (i < n) {
    a[i] = ax;
    a[n + i] = ay;
    a[2 * n + i] = az;
  }
}

// ----------------------------------------------------------------- EIH ----
//
// Einstein-Infeld-Hoffmann 1PN accelerations (harmonic gauge) in the form of
// Newhall, Standish & Williams (1983, A&A 125, 150, eq. 1) with beta = gamma
// = 1, mu = G m, r from the softened r^2 = d.d + softening:
//   a_i = sum_{j != i} mu_j d_ij / r^3 [1 - (4 phi_i + phi_j)/c^2
//           + (v_i^2 + 2 v_j^2 - 4 v_i.v_j)/c^2 - 3/2 (d_ij.v_j / r)^2/c^2
//           + d_ij.aN_j / (2 c^2)]
//       + sum_{j != i} mu_j / (c^2 r^3) [-d_ij.(4 v_i - 3 v_j)] (v_i - v_j)
//       + 7/(2 c^2) sum_{j != i} mu_j aN_j / r
// with d_ij = x_j - x_i, phi_i = sum_{k != i} mu_k / r_ik and aN the
// Newtonian accelerations. For two bodies it is the pairwise 1PN law.
// Pass 1 forms phi and aN, pass 2 the accelerations; one thread per body,
// sources in index order (the oracle's order, so the two agree bit for bit).
__global__ __launch_bounds__(256) void k_eih_pass1(int n, const double* __restrict__ x, const double* __restrict__ m,
                                                   double G, double soft, double* __restrict__ phi,
                                                   double* __restrict__ an) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double p1x = x[i], p1y = x[n + i], p1z = x[2 * n + i];
  double ph = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    const double dx = x[j] - p1x, dy = x[n + j] - p1y, dz = x[2 * n + j] - p1z;
    const double r2 = dx * dx + dy * dy + dz * dz + soft;
    const double inv_r = 1.0 / sqrt(r2);
    const double inv_r3 = inv_r * inv_r * inv_r;
    const double mu = G * m[j];
    ph += mu * inv_r;
    ax += mu * dx * inv_r3;
    ay += mu * dy * inv_r3;
    az += mu * dz * inv_r3;
  }
  phi[i] = ph;
  an[i] = ax;
  an[n + i] = ay;
  an[2 * n + i] = az;
}

__global__ __launch_bounds__(256) void k_eih_pass2(int n, const double* __restrict__ x, const double* __restrict__ v,
                                                   const double* __restrict__ m, double G, double ic, double soft,
                                                   const double* __restrict__ phi, const double* __restrict__ an,
                                                   double* __restrict__ a) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double ic2 = ic * ic;
  const double p1x = x[i], p1y = x[n + i], p1z = x[2 * n + i];
  const double v1x = v[i], v1y = v[n + i], v1z = v[2 * n + i];
  const double vi2 = v1x * v1x + v1y * v1y + v1z * v1z;
  const double phi_i = phi[i];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int j = 0; j < n; ++j) {
    if (j == i) continue;
    const double dx = x[j] - p1x, dy = x[n + j] - p1y, dz = x[2 * n + j] - p1z;
    const double r2 = dx * dx + dy * dy + dz * dz + soft;
    const double inv_r = 1.0 / sqrt(r2);
    const double inv_r3 = inv_r * inv_r * inv_r;
    const double mu = G * m[j];
    const double v2x = v[j], v2y = v[n + j], v2z = v[2 * n + j];
    const double anx = an[j], any = an[n + j], anz = an[2 * n + j];
    const double vj2 = v2x * v2x + v2y * v2y + v2z * v2z;
    const double vij = v1x * v2x + v1y * v2y + v1z * v2z;
    const double rv = (dx * v2x + dy * v2y + dz * v2z) * inv_r;
    const double dan = dx * anx + dy * any + dz * anz;
    const double B = 1.0 - (4.0 * phi_i + phi[j]) * ic2 + (vi2 + 2.0 * vj2 - 4.0 * vij) * ic2 -
                     1.5 * (rv * rv) * ic2 + 0.5 * dan * ic2;
    const double s1 = mu * inv_r3 * B;
    const double w = -(dx * (4.0 * v1x - 3.0 * v2x) + dy * (4.0 * v1y - 3.0 * v2y) + dz * (4.0 * v1z - 3.0 * v2z));
    const double s2 = mu * inv_r3 * ic2 * w;
    const double s3 = 3.5 * mu * inv_r * ic2;
    ax += dx * s1;
    ax += (v1x - v2x) * s2;
    ax += anx * s3;
    ay += dy * s1;
    ay += (v1y - v2y) * s2;
    ay += any * s3;
    az += dz * s1;
    az += (v1z - v2z) * s2;
    az += anz * s3;
  }
  a[i] = ax;
  a[n + i] = ay;
  a[2 * n + i] = az;
}

// --------------------------------------------------------------- launch ----

// Bit-exact FP64 sums: the tiled kernel below kF64TiledMax targets (few
// wavefronts of one target per lane would leave most SIMDs idle), one target
// per lane above.
#ifndef ASTER_F64_TILED_MAX
#define ASTER_F64_TILED_MAX 32768
#endif
constexpr int kF64TiledMax = ASTER_F64_TILED_MAX;

static int launch_allpairs_f64(aster_ctx* c, bool pn, int mode, int n, int lower, int upper, double ic, int pnf,
                               double* out) {
  const bool tiled = n <= kF64TiledMax;
  const dim3 grid(tiled ? (n + kF64Targets - 1) / kF64Targets : (n + 255) / 256), block(256);
  const double G = c->P.G, soft = c->P.softening;
  if (tiled) {
    if (mode == 1)
      hipLaunchKernelGGL((k_allpairs_f64_tiled<true, 1>), grid, block, 0, c->stream, n, lower, upper, c->x.p,
                         c->v.p, c->m.p, G, ic, pnf, soft, out);
    else if (pn)
      hipLaunchKernelGGL((k_allpairs_f64_tiled<true, 0>), grid, block, 0, c->stream, n, lower, upper, c->x.p,
                         c->v.p, c->m.p, G, ic, pnf, soft, out);
    else
      hipLaunchKernelGGL((k_allpairs_f64_tiled<false, 0>), grid, block, 0, c->stream, n, lower, upper, c->x.p,
                         c->v.p, c->m.p, G, ic, pnf, soft, out);
  } else {
    if (mode == 1)
      hipLaunchKernelGGL((k_allpairs_f64<true, 1>), grid, block, 0, c->stream, n, lower, upper, c->x.p, c->v.p,
                         c->m.p, G, ic, pnf, soft, out);
    else if (pn)
      hipLaunchKernelGGL((k_allpairs_f64<true, 0>), grid, block, 0, c->stream, n, lower, upper, c->x.p, c->v.p,
                         c->m.p, G, ic, pnf, soft, out);
    else
      hipLaunchKernelGGL((k_allpairs_f64<false, 0>), grid, block, 0, c->stream, n, lower, upper, c->x.p, c->v.p,
                         c->m.p, G, ic, pnf, soft, out);
  }
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

// Source slices for `target` waves of target-slice work over the 1024 SIMDs,
// each slice at least 1024 sources, at most 64 slices.
static int split_for_waves(int64_t npad, int64_t target) {
  const int64_t waves = npad / 64;
  int64_t S = (target + waves - 1) / waves;
  const int64_t max_by_len = npad / 1024;
  if (S > max_by_len) S = max_by_len;
  if (S < 1) S = 1;
  if (S > 64) S = 64;
  return (int)S;
}

// About 64 waves per SIMD. Measured against 8 per SIMD (the FP32 kernels,
// FP64 slice partials and their combine included): n = 65536 plain 1.41 ->
// 1.16 ms, mixed 1.99 -> 1.49 ms, adjoint backward kernel 3.73 -> 3.19 ms;
// n = 262144 plain 20.1 -> 17.5 ms, mixed 28.3 -> 22.4 ms. Shorter slices keep
// a wavefront's sources, read through the scalar cache, shared by more of
// the waves on a CU.
int default_split(int64_t npad) { return split_for_waves(npad, 64 * 1024); }

// The allocation-free entry point keeps about 8 waves per SIMD, which bounds
// its slice partials (S n) for a static scratch array of modest size.
static int noalloc_split(int64_t npad) { return split_for_waves(npad, 8 * 1024); }

// Slices of a span of sources (a multiple of kApTile) for the FP32 kernels:
// the requested count, at least 1024 sources each, rounded to whole tiles.
static void slice_span(int span, int requested, int* S, int* slice) {
  int s = requested;
  if ((int64_t)s * 1024 > span) s = std::max(1, span / 1024);
  const int64_t sl = round_up((span + s - 1) / s, kApTile);
  *slice = (int)sl;
  *S = (int)((span + sl - 1) / sl);
}

int choose_split(aster_ctx* c) {
  if (c->P.split > 0) return c->P.split;
  return default_split(c->npad);
}

int fp32_scales(aster_ctx* c, Fp32Scales* sc) {
  double M = 0.0, R = 0.0;
  int rc = total_mass(c, &M, &R);
  if (rc) return rc;
  const float e2 = (float)c->P.softening;
  sc->mask = !(e2 >= FLT_MIN);
  sc->eps2 = sc->mask ? 0.f : e2;
  const int k = fp32_mass_shift(M, R, sc->eps2, sc->mask);
  if (k > kMaxMassShift)
    return fail(c, ASTER_ERR_INVALID,
                "masses and softening are beyond the FP32 force range (use an FP64 method or a larger softening)");
  sc->msc = std::ldexp(1.0f, -k);
  sc->gsc = std::ldexp(c->P.G, k);
  return ASTER_OK;
}

int launch_allpairs(aster_ctx* c, int method) {
  const int n = (int)c->n;
  const int npad = (int)c->npad;
  const int lower = c->P.lower < 0 ? 0 : c->P.lower;
  const int upper = (c->P.upper < 0 || c->P.upper > n) ? n : c->P.upper;
  if (lower > upper) return fail(c, ASTER_ERR_INVALID, "source bounds: lower > upper");
  const int pnf = c->P.pn_flags & ASTER_PN_TERMS;
  const bool pn = pnf != 0;
  const double ic = 1.0 / c->P.c;
  if (pnf & ASTER_EIH) {
    if (method != ASTER_ALLPAIRS_FP64) return fail(c, ASTER_ERR_INVALID, "ASTER_EIH needs ASTER_ALLPAIRS_FP64");
    if (lower != 0 || upper != n) return fail(c, ASTER_ERR_INVALID, "ASTER_EIH covers the full source range only");
    ASTER_HIP(c, c->apn.reserve((size_t)4 * n));
    double* an = c->apn.p;
    double* phi = c->apn.p + 3 * (size_t)n;
    const dim3 grid((n + 255) / 256);
    phase_begin(c, ASTER_PHASE_ALLPAIRS);
    hipLaunchKernelGGL(k_eih_pass1, grid, dim3(256), 0, c->stream, n, c->x.p, c->m.p, c->P.G, c->P.softening, phi, an);
    hipLaunchKernelGGL(k_eih_pass2, grid, dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->m.p, c->P.G, ic,
                       c->P.softening, (const double*)phi, (const double*)an, c->a.p);
    ASTER_HIP(c, hipGetLastError());
    phase_end(c, ASTER_PHASE_ALLPAIRS);
    return ASTER_OK;
  }
  if (method == ASTER_ALLPAIRS_FP64) {
    phase_begin(c, ASTER_PHASE_ALLPAIRS);
    if (const int e = launch_allpairs_f64(c, pn, 0, n, lower, upper, ic, pnf, c->a.p)) return e;
    ASTER_HIP(c, hipGetLastError());
    phase_end(c, ASTER_PHASE_ALLPAIRS);
    return ASTER_OK;
  }
  if (method != ASTER_ALLPAIRS_FP32 && method != ASTER_ALLPAIRS_FP32_KAHAN)
    return fail(c, ASTER_ERR_INVALID, "unknown all-pairs method");
  Fp32Scales sc;
  int rc = fp32_scales(c, &sc);
  if (rc) return rc;
  // Padded source array: FP32 for the plain method, FP64 for the mixed one.
  const bool src64 = method == ASTER_ALLPAIRS_FP32_KAHAN;
  phase_begin(c, ASTER_PHASE_TILES);
  if (src64) {
    ASTER_HIP(c, c->src64.reserve((size_t)npad));
    hipLaunchKernelGGL(k_src64, dim3(npad / 256), dim3(256), 0, c->stream, n, npad, lower, upper, sc.msc, c->x.p,
                       c->m.p, c->src64.p);
  } else {
    hipLaunchKernelGGL(k_src32, dim3(npad / 256), dim3(256), 0, c->stream, n, npad, lower, upper, sc.msc, c->x.p,
                       c->m.p, c->sh.p);
  }
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_TILES);
  // Slices over the tiles that hold the source range (sources outside it and
  // the padding bodies have zero mass).
  const int jb = lower / kApTile * kApTile;
  const int je = std::max<int>(jb + kApTile, (int)round_up(upper, kApTile));
  const int span = je - jb;
  int S = 0, slice = 0;
  slice_span(span, choose_split(c), &S, &slice);
  ASTER_HIP(c, c->apart.reserve((size_t)S * 3 * n));
  const float eps2 = sc.eps2;
  const bool mask = sc.mask;
  dim3 grid(npad / kApBlock, S);
  double* out = c->apart.p;
  phase_begin(c, ASTER_PHASE_ALLPAIRS);
  if (src64) {
    const double4* sp = c->src64.p + jb;
    if (mask)
      hipLaunchKernelGGL((k_allpairs_f32<ACC_TILE_KAHAN, true, true>), grid, dim3(kApBlock), 0, c->stream, sp,
                         c->x.p, n, span, (int)slice, eps2, out);
    else
      hipLaunchKernelGGL((k_allpairs_f32<ACC_TILE_KAHAN, false, true>), grid, dim3(kApBlock), 0, c->stream, sp,
                         c->x.p, n, span, (int)slice, eps2, out);
  } else {
    const float4* sp = c->sh.p + jb;
    if (mask)
      hipLaunchKernelGGL((k_allpairs_f32<ACC_PLAIN, true, false>), grid, dim3(kApBlock), 0, c->stream, sp, c->x.p, n,
                         span, (int)slice, eps2, out);
    else
      hipLaunchKernelGGL((k_allpairs_f32<ACC_PLAIN, false, false>), grid, dim3(kApBlock), 0, c->stream, sp, c->x.p,
                         n, span, (int)slice, eps2, out);
  }
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_ALLPAIRS);
  const double* pnp = nullptr;
  if (pn) {
    ASTER_HIP(c, c->apn.reserve((size_t)3 * n));
    if ((rc = launch_allpairs_f64(c, true, 1, n, lower, upper, ic, pnf, c->apn.p))) return rc;
    ASTER_HIP(c, hipGetLastError());
    pnp = c->apn.p;
  }
  phase_begin(c, ASTER_PHASE_COMBINE);
  hipLaunchKernelGGL(k_combine, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, S, c->apart.p, pnp, sc.gsc,
                     c->a.p);
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_COMBINE);
  return ASTER_OK;
}

}  // namespace aster

// ------------------------------------------------------ allocation-free ----
//
// aster_forces_noalloc: the ASTER_ALLPAIRS_FP32_KAHAN evaluation (same
// kernels and mass scale as aster_forces; the source split of about 8 waves
// per SIMD, which aster_forces uses with aster_params.split set to it) on
// caller-owned device arrays, with no device allocation. Its working memory is the file-scope
// array `scratch` below: each GPU the library runs on holds one copy until
// the process ends, and every call on that GPU uses it, so calls that
// overlap on one GPU race. Nothing in it needs clearing: every call writes
// each word it reads first.
namespace aster {

constexpr int64_t kNaMax = ASTER_NOALLOC_MAX_BODIES;
static_assert(kNaMax % kApBlock == 0, "the body limit must be a whole number of all-pairs blocks");
// Bound on S n of noalloc_split over n <= kNaMax (the largest is 587520, at
// n = 65280 with S = 9).
constexpr int64_t kNaPartMax = 589824;
// Layout in doubles: SoA positions | double4 sources | slice partials |
// mass-sum partials | the two scales.
constexpr size_t kNaX = 0;
constexpr size_t kNaSrc = kNaX + 3 * kNaMax;
constexpr size_t kNaPart = kNaSrc + 4 * kNaMax;
constexpr size_t kNaRed = kNaPart + 3 * kNaPartMax;
constexpr size_t kNaScale = kNaRed + (size_t)kComQ * (kComBlocks + 1);
constexpr size_t kNaTotal = kNaScale + 2;

static __device__ __align__(256) double scratch[kNaTotal];

// AoS-d3 positions -> SoA planes.
__global__ __launch_bounds__(256) void k_na_planes(int n, const double* __restrict__ pos, double* __restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) x[(size_t)c * n + i] = pos[3 * (size_t)i + c];
}

// The mass scale of fp32_scales, formed on the device from the mass sums:
// scale[0] = 2^-k, scale[1] = G 2^k. k beyond the FP32 range is clamped (the
// result is then not finite; aster_forces reports an error instead).
__global__ void k_na_scales(const double* __restrict__ sums, float eps2, int mask, double G,
                            double* __restrict__ scale) {
  if (threadIdx.x != 0) return;
  int k = fp32_mass_shift(sums[0], sums[4], eps2, mask != 0);
  if (k > kMaxMassShift) k = kMaxMassShift;
  scale[0] = (double)ldexpf(1.0f, -k);
  scale[1] = ldexp(G, k);
}

// k_src64 with the mass scale read from the device.
__global__ __launch_bounds__(256) void k_na_src(int n, int npad, const double* __restrict__ scale,
                                                const double* __restrict__ x, const double* __restrict__ m,
                                                double4* __restrict__ src) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  const float msc = (float)scale[0];
  src[i] = i < n ? make_double4(x[i], x[n + i], x[2 * n + i], pack_mass((float)m[i] * msc))
                 : make_double4(0.0, 0.0, 0.0, pack_mass(0.f));
}

// k_combine (no post-Newtonian term) writing AoS-d3.
__global__ void k_na_combine(int n, int S, const double* __restrict__ part, const double* __restrict__ scale,
                             double* __restrict__ acc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double G = scale[1];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double s = 0.0;
    for (int k = 0; k < S; ++k) s += part[((size_t)k * 3 + c) * n + i];
    acc[3 * (size_t)i + c] = G * s;
  }
}

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_lock;
static double* g_na_ptr[kNaMaxDevices] = {};

// The device address of `scratch` on device `dev` (current), looked up once
// per device. In host code the array's name is only a stand-in: kernels get
// this address.
static hipError_t na_scratch(int dev, double** out) {
  std::lock_guard<std::mutex> guard(g_na_lock);
  if (!g_na_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_na_ptr[dev] = static_cast<double*>(p);
  }
  *out = g_na_ptr[dev];
  return hipSuccess;
}

static int na_fail(int code, const std::string& msg) {
  set_global_error(msg);
  return code;
}

}  // namespace aster

extern "C" int aster_forces_noalloc(int64_t n, const double* d_pos, const double* d_mass, double G, double softening,
                                    double* d_acc, void* stream) {
  using namespace aster;
  if (n < 1 || !d_pos || !d_mass || !d_acc)
    return na_fail(ASTER_ERR_INVALID, "aster_forces_noalloc: need n >= 1 and device positions, masses, output");
  if (n > kNaMax)
    return na_fail(ASTER_ERR_CAPACITY, "aster_forces_noalloc: at most ASTER_NOALLOC_MAX_BODIES bodies");
  if (!(softening >= 0.0) || !std::isfinite(G)) return na_fail(ASTER_ERR_INVALID, "invalid G or softening");
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return na_fail(ASTER_ERR_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kNaMaxDevices) return na_fail(ASTER_ERR_CAPACITY, "device index beyond 63");
  double* base = nullptr;
  if ((e = na_scratch(dev, &base)) != hipSuccess)
    return na_fail(ASTER_ERR_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
  const int ni = (int)n;
  const int npad = (int)round_up(n, kApBlock);
  // the slicing of launch_allpairs over the full source range
  const int span = (int)round_up(n, kApTile);
  int S = 0, slice = 0;
  slice_span(span, noalloc_split(npad), &S, &slice);
  if ((int64_t)S * n > kNaPartMax) return na_fail(ASTER_ERR_CAPACITY, "slice partials beyond the scratch layout");
  double* x = base + kNaX;
  double4* src = reinterpret_cast<double4*>(base + kNaSrc);
  double* part = base + kNaPart;
  double* red = base + kNaRed;
  double* sums = red + (size_t)kComQ * kComBlocks;
  double* scale = base + kNaScale;
  const float e2 = (float)softening;
  const bool mask = !(e2 >= FLT_MIN);
  const float eps2 = mask ? 0.f : e2;
  hipStream_t st = static_cast<hipStream_t>(stream);
  const dim3 b256(256), gn((ni + 255) / 256);
  const int rblocks = (int)std::min<int64_t>((n + 255) / 256, kComBlocks);
  hipLaunchKernelGGL(k_na_planes, gn, b256, 0, st, ni, d_pos, x);
  hipLaunchKernelGGL(k_com_partial, dim3(rblocks), b256, 0, st, n, (const double*)x, d_mass, red);
  hipLaunchKernelGGL(k_com_final, dim3(1), dim3(64), 0, st, rblocks, (const double*)red, sums);
  hipLaunchKernelGGL(k_na_scales, dim3(1), dim3(64), 0, st, (const double*)sums, eps2, mask ? 1 : 0, G, scale);
  hipLaunchKernelGGL(k_na_src, dim3(npad / 256), b256, 0, st, ni, npad, (const double*)scale, (const double*)x,
                     d_mass, src);
  const dim3 grid(npad / kApBlock, S);
  if (mask)
    hipLaunchKernelGGL((k_allpairs_f32<ACC_TILE_KAHAN, true, true>), grid, dim3(kApBlock), 0, st,
                       (const double4*)src, (const double*)x, ni, span, slice, eps2, part);
  else
    hipLaunchKernelGGL((k_allpairs_f32<ACC_TILE_KAHAN, false, true>), grid, dim3(kApBlock), 0, st,
                       (const double4*)src, (const double*)x, ni, span, slice, eps2, part);
  hipLaunchKernelGGL(k_na_combine, gn, b256, 0, st, ni, S, (const double*)part, (const double*)scale, d_acc);
  if ((e = hipGetLastError()) != hipSuccess)
    return na_fail(ASTER_ERR_HIP, std::string("aster_forces_noalloc launch: ") + hipGetErrorString(e));
  return ASTER_OK;
}
