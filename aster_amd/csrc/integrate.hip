// Kick/drift updates on the FP64 state, the FP32 shadow, energy and angular
// momentum reductions, the automatic Morton box and device error flags.
//
// The kick/drift arithmetic keeps the operation order of update_leapfrog
// (src/physics/tool_chain_impl.cpp:169-185):
//   x += v*dt + ((0.5*a)*dt)*dt ;  v += (0.5*a)*dt
// with contraction disabled, so given identical accelerations the device
// trajectory is bit-identical to the CPU path. The second half-kick of a step
// and the first half-kick of the next use the same accelerations, so they are
// applied in one pass (k_kick_drift with dt_prev != 0).
#include <algorithm>

#include "common.hpp"
#include "device.hpp"

namespace aster {

// BOX: the block also writes the largest |coordinate| of its drifted bodies
// per axis to part[3 block + axis], the first stage of the automatic Morton
// box of the next force evaluation (k_box's partials for these positions).
template <bool BOX>
__global__ __launch_bounds__(256) void k_kick_drift(int n, double* __restrict__ x, double* __restrict__ v,
                                                    const double* __restrict__ a, double dt_prev, double dt,
                                                    uint32_t* __restrict__ flags, double* __restrict__ part) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (!BOX && i >= n) return;
  double mx[3] = {0.0, 0.0, 0.0};
  if (i < n) {
    bool finite = true;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const size_t k = (size_t)c * n + i;
      double xc = x[k], vc = v[k];
      kick_drift_1(xc, vc, a[k], dt_prev, dt);
      x[k] = xc;
      v[k] = vc;
      mx[c] = fabs(xc);
      finite = finite && isfinite(xc) && isfinite(vc);
    }
    if (!finite) atomicOr(flags, (uint32_t)FLAG_NONFINITE);
  }
  if (BOX) {
    __shared__ double wm[3][4];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      double r = mx[c];
      for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
      if ((threadIdx.x & 63) == 0) wm[c][w] = r;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
      const int c = threadIdx.x;
      part[3 * blockIdx.x + c] = fmax(fmax(wm[c][0], wm[c][1]), fmax(wm[c][2], wm[c][3]));
    }
  }
}

// One SABA stage after its force evaluation, as the per-body lambdas of
// update_SABAn (src/integrators/SABA.cpp:19-553): v += (a d) dt, then
// x += (v c) dt with the kicked v. kick = 0 is the sequence's opening drift.
// update_euler (tool_chain_impl.cpp:158-167) is d = c = 1, since (a 1) dt and
// (v 1) dt round exactly like a dt and v dt.
__global__ __launch_bounds__(256) void k_saba_stage(int n, double* __restrict__ x, double* __restrict__ v,
                                                    const double* __restrict__ a, int kick, double d, double cdrift,
                                                    double dt, uint32_t* __restrict__ flags) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool finite = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * n + i;
    double vc = v[k];
    if (kick) {
      vc = vc + a[k] * d * dt;
      v[k] = vc;
    }
    const double xc = x[k] + vc * cdrift * dt;
    x[k] = xc;
    finite = finite && isfinite(xc) && isfinite(vc);
  }
  if (!finite) atomicOr(flags, (uint32_t)FLAG_NONFINITE);
}

__global__ __launch_bounds__(256) void k_kick(int n, double* __restrict__ v, const double* __restrict__ a,
                                              double dt) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const size_t k = (size_t)c * n + i;
    v[k] = v[k] + 0.5 * a[k] * dt;
  }
}

// Per-body energy terms: out[0][i] = kinetic, out[1][i] = sum_j G m_i m_j /
// (|r| + soft), out[2][i] = sum_j G m_i m_j / sqrt(r^2 + soft), out[3..5][i] =
// m x cross v. Pair sums run over all j != i (each pair counted twice).
__global__ __launch_bounds__(256) void k_energy_terms(int n, const double* __restrict__ x,
                                                      const double* __restrict__ v, const double* __restrict__ m,
                                                      double G, double soft, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const double px = x[ii], py = x[n + ii], pz = x[2 * n + ii];
  const double mi = m[ii];
  double pref = 0.0, ppl = 0.0;
  for (int j = 0; j < n; ++j) {
    const double dx = x[j] - px, dy = x[n + j] - py, dz = x[2 * n + j] - pz;
    const double mj = m[j];
    if (j == ii) continue;
    const double r2 = dx * dx + dy * dy + dz * dz;
    pref += mj / (sqrt(r2) + soft);
    ppl += mj / sqrt(r2 + soft);
  }
  if (i < n) {
    const double vx = v[i], vy = v[n + i], vz = v[2 * n + i];
    out[i] = 0.5 * mi * (vx * vx + vy * vy + vz * vz);
    out[(size_t)n + i] = G * mi * pref;
    out[(size_t)2 * n + i] = G * mi * ppl;
    out[(size_t)3 * n + i] = mi * (py * vz - pz * vy);
    out[(size_t)4 * n + i] = mi * (pz * vx - px * vz);
    out[(size_t)5 * n + i] = mi * (px * vy - py * vx);
  }
}

// Deterministic reduction of `rows` rows of length n: one block, fixed
// strided order, fixed tree.
__global__ __launch_bounds__(1024) void k_reduce_rows(int n, int rows, const double* __restrict__ in,
                                                      double* __restrict__ out) {
  __shared__ double sm[1024];
  for (int r = 0; r < rows; ++r) {
    double s = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) s += in[(size_t)r * n + i];
    sm[threadIdx.x] = s;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) sm[threadIdx.x] += sm[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[r] = sm[0];
    __syncthreads();
  }
}

// Symmetric box: max |coordinate| per axis, in two stages (per-block maxima,
// then one block over those), so no two blocks contend for an address. fmax
// is exact and order-independent.
constexpr int kBoxBlocks = 1024;

__global__ __launch_bounds__(256) void k_box(int n, const double* __restrict__ x, double* __restrict__ part) {
  __shared__ double wm[3][4];
  double mx[3] = {0.0, 0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
#pragma unroll
    for (int c = 0; c < 3; ++c) mx[c] = fmax(mx[c], fabs(x[(size_t)c * n + i]));
  const int w = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double r = mx[c];
    for (int off = 32; off > 0; off >>= 1) r = fmax(r, __shfl_xor(r, off, 64));
    if ((threadIdx.x & 63) == 0) wm[c][w] = r;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    const int c = threadIdx.x;
    part[3 * blockIdx.x + c] = fmax(fmax(wm[c][0], wm[c][1]), fmax(wm[c][2], wm[c][3]));
  }
}

// One thread per partial (one round of loads), then a block maximum; fmax is
// exact and order-independent.
constexpr int kBoxFinishThreads = 1024;
__global__ __launch_bounds__(kBoxFinishThreads) void k_box_finish(const double* __restrict__ part, int nparts,
                                                                  const double* __restrict__ user, int use_user,
                                                                  double* __restrict__ box) {
#pragma clang fp contract(off)
  __shared__ double wm[3][kBoxFinishThreads / 64];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double r[3] = {0.0, 0.0, 0.0};
  // 8 partials per thread and round, all loaded before the first is used
  // (clamped indices), so a round costs one memory round trip
  if (!use_user)
    for (int i0 = t; i0 < nparts; i0 += 8 * kBoxFinishThreads) {
      double q[8][3];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int i = min(i0 + u * kBoxFinishThreads, nparts - 1);
#pragma unroll
        for (int c = 0; c < 3; ++c) q[u][c] = part[3 * i + c];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
#pragma unroll
        for (int c = 0; c < 3; ++c) r[c] = fmax(r[c], q[u][c]);
    }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    for (int off = 32; off > 0; off >>= 1) r[c] = fmax(r[c], __shfl_xor(r[c], off, 64));
    if (lane == 0) wm[c][w] = r[c];
  }
  __syncthreads();
  if (t != 0) return;
  double b[3];
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double m = 0.0;
    for (int q = 0; q < kBoxFinishThreads / 64; ++q) m = fmax(m, wm[c][q]);
    b[c] = use_user ? user[c] : m;
  }
  box[0] = b[0];
  box[1] = b[1];
  box[2] = b[2];
  // bd_size = |bounding_box| (barnes_hut.cpp:405; vec3::magnitude, vectors.h:93-99)
  box[3] = sqrt(b[0] * b[0] + b[1] * b[1] + b[2] * b[2]);
}

__global__ void k_zero_u32(uint32_t* p, int k) {
  if ((int)threadIdx.x < k) p[threadIdx.x] = 0u;
}

// --------------------------------------------------------------- launch ----

int launch_reset_flags(aster_ctx* c) {
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(64), 0, c->stream, c->flags.p, 4);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

int launch_kick_drift(aster_ctx* c, double dt_prev, double dt, bool box) {
  const int n = (int)c->n;
  const int blocks = (n + 255) / 256;
  c->box_parts = 0;
  // (at least what launch_box reserves, so that it never reallocates the partials)
  if (box) ASTER_HIP(c, c->boxpart.reserve((size_t)3 * std::max(blocks, kBoxBlocks)));
  phase_begin(c, ASTER_PHASE_KICKDRIFT);
  if (box)
    hipLaunchKernelGGL(k_kick_drift<true>, dim3(blocks), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->a.p, dt_prev,
                       dt, c->flags.p, c->boxpart.p);
  else
    hipLaunchKernelGGL(k_kick_drift<false>, dim3(blocks), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->a.p, dt_prev,
                       dt, c->flags.p, nullptr);
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_KICKDRIFT);
  if (box) {
    c->box_parts = blocks;
    c->box_parts_x = c->x.p;
    c->box_parts_n = c->n;
  }
  return ASTER_OK;
}

int launch_saba_stage(aster_ctx* c, bool kick, double d, double cdrift, double dt) {
  const int n = (int)c->n;
  phase_begin(c, ASTER_PHASE_KICKDRIFT);
  hipLaunchKernelGGL(k_saba_stage, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->a.p,
                     kick ? 1 : 0, d, cdrift, dt, c->flags.p);
  ASTER_HIP(c, hipGetLastError());
  phase_end(c, ASTER_PHASE_KICKDRIFT);
  return ASTER_OK;
}

int launch_kick(aster_ctx* c, double dt) {
  const int n = (int)c->n;
  hipLaunchKernelGGL(k_kick, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->v.p, c->a.p, dt);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

int launch_energy(aster_ctx* c, double* out6) {
  const int n = (int)c->n;
  ASTER_HIP(c, c->ework.reserve((size_t)6 * n));
  ASTER_HIP(c, c->eout.reserve(8));
  hipLaunchKernelGGL(k_energy_terms, dim3((n + 255) / 256), dim3(256), 0, c->stream, n, c->x.p, c->v.p, c->m.p,
                     c->P.G, c->P.softening, c->ework.p);
  ASTER_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_reduce_rows, dim3(1), dim3(1024), 0, c->stream, n, 6, c->ework.p, c->eout.p);
  ASTER_HIP(c, hipGetLastError());
  ASTER_HIP(c, hipMemcpyAsync(out6, c->eout.p, 6 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return ASTER_OK;
}

int launch_box(aster_ctx* c, bool user_box) {
  const int n = (int)c->n;
  ASTER_HIP(c, c->boxpart.reserve(3 * kBoxBlocks));
  ASTER_HIP(c, c->boxd.reserve(8));
  // A user box was uploaded to boxd[4..6] by aster_set_params.
  const bool user = user_box && (c->P.box[0] != 0.0 || c->P.box[1] != 0.0 || c->P.box[2] != 0.0);
  int blocks = 0;
  // Partials of these very positions left by the drift just before (same
  // state buffer and body count): only their final maximum remains.
  const bool fused = c->box_parts > 0 && c->box_parts_x == c->x.p && c->box_parts_n == c->n;
  if (!user && fused) {
    blocks = c->box_parts;
  } else if (!user) {
    blocks = (n + 255) / 256;
    if (blocks > kBoxBlocks) blocks = kBoxBlocks;
    hipLaunchKernelGGL(k_box, dim3(blocks), dim3(256), 0, c->stream, n, c->x.p, c->boxpart.p);
    ASTER_HIP(c, hipGetLastError());
  }
  c->box_parts = 0;
  hipLaunchKernelGGL(k_box_finish, dim3(1), dim3(kBoxFinishThreads), 0, c->stream, c->boxpart.p, blocks, c->boxd.p + 4,
                     user ? 1 : 0, c->boxd.p);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

}  // namespace aster
