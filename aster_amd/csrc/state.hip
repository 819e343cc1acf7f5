// State transfer between the caller's AoS-d3 layout (BodyArray: packed
// double[3] per body, include/Aster/physics/body.h:9-24) and the device SoA
// planes, done on the device, plus the Graphs collectors
// (src/utils/collectors.cpp:19-79) evaluated there.
#include <cmath>

#include "common.hpp"

namespace aster {

// AoS-d3 position/velocity and masses -> SoA planes, with the input checks
// aster_set_state makes on the host: finite values, finite positive masses.
__global__ __launch_bounds__(256) void k_import(int64_t n, const double* __restrict__ pos,
                                                const double* __restrict__ vel, const double* __restrict__ mass,
                                                double* __restrict__ x, double* __restrict__ v,
                                                double* __restrict__ m, uint32_t* __restrict__ flags) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  bool finite = true;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const double p = pos[3 * i + c];
    const double u = vel ? vel[3 * i + c] : 0.0;
    x[c * n + i] = p;
    v[c * n + i] = u;
    finite = finite && isfinite(p) && isfinite(u);
  }
  const double mi = mass[i];
  m[i] = mi;
  uint32_t f = 0;
  if (!finite) f |= FLAG_NONFINITE;
  if (!(mi > 0.0) || !isfinite(mi)) f |= FLAG_BADMASS;
  if (f) atomicOr(flags, f);
}

// SoA planes -> AoS-d3.
__global__ __launch_bounds__(256) void k_export(int64_t n, const double* __restrict__ soa, double* __restrict__ aos) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
#pragma unroll
  for (int c = 0; c < 3; ++c) aos[3 * i + c] = soa[c * n + i];
}

// Per-block sums of m and m x (x, y, z), in a fixed order (each thread sums a
// strided set of bodies, then the block's waves are summed in lane order),
// and the largest |coordinate|.
__global__ __launch_bounds__(256) void k_com_partial(int64_t n, const double* __restrict__ x,
                                                     const double* __restrict__ m, double* __restrict__ part) {
  __shared__ double red[kComQ][256];
  double s[kComQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double mi = m[i];
    s[0] += mi;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double xc = x[c * n + i];
      s[1 + c] += mi * xc;
      s[4] = fmax(s[4], fabs(xc));
    }
  }
#pragma unroll
  for (int q = 0; q < kComQ; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
#pragma unroll
      for (int q = 0; q < 4; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
      red[4][threadIdx.x] = fmax(red[4][threadIdx.x], red[4][threadIdx.x + w]);
    }
    __syncthreads();
  }
  if (threadIdx.x < kComQ) part[kComQ * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// One block of 64 threads: the partials are staged in LDS with one load
// round per thread, then thread 0 sums them in block order (the order, and so
// the bits, of a single thread reading them one by one).
__global__ __launch_bounds__(64) void k_com_final(int nparts, const double* __restrict__ part,
                                                  double* __restrict__ out) {
  __shared__ double sp[kComQ * kComBlocks];
  const int t = threadIdx.x;
  const int nb = min(nparts, kComBlocks);  // callers pass at most kComBlocks partials
  const int total = kComQ * nb;
  constexpr int kPer = (kComQ * kComBlocks + 63) / 64;
  double v[kPer];
#pragma unroll
  for (int r = 0; r < kPer; ++r) v[r] = part[min(t + 64 * r, max(total - 1, 0))];
#pragma unroll
  for (int r = 0; r < kPer; ++r)
    if (t + 64 * r < total) sp[t + 64 * r] = v[r];
  __syncthreads();
  if (t != 0) return;
  double s[kComQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = 0; b < nb; ++b) {
    for (int q = 0; q < 4; ++q) s[q] += sp[kComQ * b + q];
    s[4] = fmax(s[4], sp[kComQ * b + 4]);
  }
  for (int q = 0; q < kComQ; ++q) out[q] = s[q];
}

static int reserve_state(aster_ctx* c, int64_t n) {
  const int64_t npad = round_up(n, kApBlock);
  ASTER_HIP(c, c->x.reserve(3 * n));
  ASTER_HIP(c, c->v.reserve(3 * n));
  ASTER_HIP(c, c->a.reserve(3 * n));
  ASTER_HIP(c, c->m.reserve(n));
  ASTER_HIP(c, c->sh.reserve(npad));
  return ASTER_OK;
}

// Loads n bodies from AoS-d3 device arrays (vel may be null), checks them
// and resets the context's derived state.
int import_state(aster_ctx* c, int64_t n, const double* d_pos, const double* d_vel, const double* d_mass) {
  int rc = reserve_state(c, n);
  if (rc) return rc;
  if ((rc = launch_reset_flags(c))) return rc;
  hipLaunchKernelGGL(k_import, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, d_pos, d_vel, d_mass,
                     c->x.p, c->v.p, c->m.p, c->flags.p);
  ASTER_HIP(c, hipGetLastError());
  ASTER_HIP(c, hipMemsetAsync(c->a.p, 0, 3 * n * sizeof(double), c->stream));
  uint32_t f = 0;
  ASTER_HIP(c, hipMemcpyAsync(&f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  if (f) {
    if ((rc = launch_reset_flags(c))) return rc;
    c->have_state = false;
    if (f & FLAG_BADMASS) return fail(c, ASTER_ERR_INVALID, "masses must be finite and > 0 (accelerations divide by m_i)");
    return fail(c, ASTER_ERR_NONFINITE, "non-finite position or velocity in input");
  }
  c->n = n;
  c->npad = round_up(n, kApBlock);
  c->have_state = true;
  c->acc_valid = false;
  c->acc_stale = false;
  c->tree_valid = false;
  c->adj.ready = false;
  c->e0_valid = false;
  c->mass_total_valid = false;
  c->box_parts = 0;
  return ASTER_OK;
}

// Grows the state by `add` bodies after the current ones (whose planes are
// copied into the wider layout) and returns the first new index in *base.
// The new bodies are left for the caller to write; derived state
// (accelerations, tree, energy reference, adjoint) is invalidated.
int grow_state(aster_ctx* c, int64_t add, int64_t* base) {
  const int64_t n0 = c->have_state ? c->n : 0;
  const int64_t n1 = n0 + add;
  DBuf<double> nx, nv, nm;
  hipError_t e = nx.reserve(3 * n1);
  if (e == hipSuccess) e = nv.reserve(3 * n1);
  if (e == hipSuccess) e = nm.reserve(n1);
  if (e != hipSuccess) {
    nx.release();
    nv.release();
    nm.release();
    return hip_fail(c, e, "grow_state allocation");
  }
  if (n0 > 0) {
    for (int q = 0; q < 3; ++q) {
      ASTER_HIP(c, hipMemcpyAsync(nx.p + q * n1, c->x.p + q * n0, n0 * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
      ASTER_HIP(c, hipMemcpyAsync(nv.p + q * n1, c->v.p + q * n0, n0 * sizeof(double), hipMemcpyDeviceToDevice,
                                  c->stream));
    }
    ASTER_HIP(c, hipMemcpyAsync(nm.p, c->m.p, n0 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  }
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  c->x.release();
  c->v.release();
  c->m.release();
  c->x = nx;
  c->v = nv;
  c->m = nm;
  ASTER_HIP(c, c->a.reserve(3 * n1));
  ASTER_HIP(c, c->sh.reserve(round_up(n1, kApBlock)));
  ASTER_HIP(c, hipMemsetAsync(c->a.p, 0, 3 * n1 * sizeof(double), c->stream));
  c->n = n1;
  c->npad = round_up(n1, kApBlock);
  c->have_state = true;
  c->acc_valid = false;
  c->acc_stale = false;
  c->tree_valid = false;
  c->adj.ready = false;
  c->e0_valid = false;
  c->mass_total_valid = false;
  c->box_parts = 0;
  *base = n0;
  return ASTER_OK;
}

// SoA plane -> AoS-d3 device array.
int export_plane(aster_ctx* c, const double* soa, double* d_aos) {
  const int64_t n = c->n;
  hipLaunchKernelGGL(k_export, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, soa, d_aos);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

// Total mass and mass-weighted position sums, summed in a fixed order, and
// the largest |coordinate|.
static int mass_sums(aster_ctx* c, double* s) {
  const int64_t n = c->n;
  int blocks = (int)std::min<int64_t>((n + 255) / 256, kComBlocks);
  ASTER_HIP(c, c->ework.reserve((size_t)kComQ * kComBlocks + kComQ));
  double* part = c->ework.p;
  double* out = c->ework.p + kComQ * kComBlocks;
  hipLaunchKernelGGL(k_com_partial, dim3(blocks), dim3(256), 0, c->stream, n, c->x.p, c->m.p, part);
  hipLaunchKernelGGL(k_com_final, dim3(1), dim3(64), 0, c->stream, blocks, part, out);
  ASTER_HIP(c, hipGetLastError());
  ASTER_HIP(c, hipMemcpyAsync(s, out, kComQ * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  c->mass_total = s[0];
  c->extent = s[4];
  c->mass_total_valid = true;
  return ASTER_OK;
}

int center_of_mass(aster_ctx* c, double* com) {
  double s[kComQ];
  int rc = mass_sums(c, s);
  if (rc) return rc;
  for (int q = 0; q < 3; ++q) com[q] = s[1 + q] / s[0];
  return ASTER_OK;
}

int total_mass(aster_ctx* c, double* mtot, double* extent) {
  if (!c->mass_total_valid) {
    double s[kComQ];
    int rc = mass_sums(c, s);
    if (rc) return rc;
  }
  *mtot = c->mass_total;
  if (extent) *extent = c->extent;
  return ASTER_OK;
}

}  // namespace aster
