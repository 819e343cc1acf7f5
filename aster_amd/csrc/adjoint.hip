// Hand-written adjoint of the leapfrog (kick-drift-kick) step for trajectory
// gradients with respect to initial positions, velocities and masses.
//
// Forward step s (h = dt/2, a_s = a(x_s), Plummer-softened Newton):
//   vh = v_s + h a_s ;  x_{s+1} = x_s + dt vh ;  v_{s+1} = vh + h a_{s+1}
// (update_leapfrog, src/physics/tool_chain_impl.cpp:169-185, in the same
// expression order as aster_step). With lx, lv the adjoints of x, v:
//   lvh = lv_{s+1} + dt lx_{s+1}
//   lx_s = lx_{s+1} + J(x_s)^T (c_s lvh) [+ dL/dx_s at snapshots],  lv_s = lvh
//   lm  += M(x_s)^T (c_s lvh)
// where J = da/dx, M = da/dm and c_s = dt for s > 0 (the first half-kick of
// step s and the second half-kick of step s-1 act at the same x_s with the
// same weight, so their vector-Jacobian products merge) and c_0 = h. The last
// step's second half-kick has weight h lv_N = 0 because the loss depends on
// positions only.
//
// Time reversal: the state is recovered backwards from (x_{s+1}, v_{s+1},
// a_{s+1}) as vh = v_{s+1} - h a_{s+1}, x_s = x_{s+1} - dt vh, and
// v_s = vh - h a(x_s); every ckpt_every steps the state is re-anchored to the
// checkpoint written by the forward pass, so reconstruction error cannot grow
// over more than one segment. a(x_s) comes out of the same fused kernel as the
// vector-Jacobian products:
//   a_k   = G sum_j m_j d / s^3
//   gx_k  = G sum_j [u / s^3 - 3 d (d . u) / s^5],  u = m_k w_j - m_j w_k
//   gm_k  = -G sum_j (w_j . d) / s^3
// with d = x_j - x_k, s^2 = |d|^2 + eps^2, FP32 pair arithmetic on separations
// formed in FP64 (as ASTER_ALLPAIRS_FP32_KAHAN), Kahan-compensated FP32 sums
// over 64-source tiles and an FP64 combine over source slices.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace aster {

__device__ __forceinline__ float unpack_mass_adj(double w) {
  return __uint_as_float((unsigned)(unsigned long long)__double_as_longlong(w));
}

// Sources for the fused kernel: positions (FP64, FP32 mass in the low word of
// w) and {w, m} in FP32, padded with zero-mass bodies. Masses are scaled by
// msc (Fp32Scales), so a and J^T w come out scaled by 2^-k.
__global__ __launch_bounds__(256) void k_adj_src(int n, int npad, float msc, const double* __restrict__ x,
                                                 const double* __restrict__ m, const double* __restrict__ w,
                                                 double4* __restrict__ src, float4* __restrict__ wsrc) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npad) return;
  if (i < n) {
    const float mf = (float)m[i] * msc;
    src[i] = make_double4(x[i], x[n + i], x[2 * n + i],
                          __longlong_as_double((long long)(unsigned long long)__float_as_uint(mf)));
    wsrc[i] = make_float4((float)w[i], (float)w[n + i], (float)w[2 * n + i], mf);
  } else {
    src[i] = make_double4(0.0, 0.0, 0.0, __longlong_as_double(0));
    wsrc[i] = make_float4(0.f, 0.f, 0.f, 0.f);
  }
}

// One 64-source tile of the fused kernel. SELF: the tile holds the targets of
// this wavefront (source k is lane k's own body), whose pair is skipped: its
// d = 0, but u = m_k w_k - m_k w_k comes out of the FMA as the rounding error
// of m_k w_k, which eps^-3 would amplify.
template <bool MASK, bool SELF>
__device__ __forceinline__ void adj_tile(const double4* __restrict__ src, const float4* __restrict__ wsrc, int t,
                                         double px, double py, double pz, const float4& wk, float mk, float eps2,
                                         int lane, float* __restrict__ S, float* __restrict__ C) {
  float ax = 0.f, ay = 0.f, az = 0.f, gx = 0.f, gy = 0.f, gz = 0.f, gm = 0.f;
#pragma unroll 8
  for (int k = 0; k < kApTile; ++k) {
    const double4 q = src[t + k];
    const float4 p = wsrc[t + k];
    const float dx = (float)(q.x - px), dy = (float)(q.y - py), dz = (float)(q.z - pz);
    float r2 = fmaf(dx, dx, eps2);
    r2 = fmaf(dy, dy, r2);
    r2 = fmaf(dz, dz, r2);
    float inv = __builtin_amdgcn_rsqf(r2);
    if (MASK) inv = (r2 > 0.f) ? inv : 0.f;
    if (SELF) inv = (k == lane) ? 0.f : inv;
    const float inv2 = inv * inv;
    const float inv3 = inv2 * inv;
    const float mj = unpack_mass_adj(q.w);
    const float f = mj * inv3;
    ax = fmaf(f, dx, ax);
    ay = fmaf(f, dy, ay);
    az = fmaf(f, dz, az);
    const float ux = fmaf(mk, p.x, -(mj * wk.x));
    const float uy = fmaf(mk, p.y, -(mj * wk.y));
    const float uz = fmaf(mk, p.z, -(mj * wk.z));
    const float du = fmaf(dx, ux, fmaf(dy, uy, dz * uz));
    // J^T w: inv3 u - 3 (d.u) inv5 d = inv3 (u - s d), s = 3 (d.u) inv^2
    const float s3 = (du * inv2) * 3.f;
    gx = fmaf(fmaf(-s3, dx, ux), inv3, gx);
    gy = fmaf(fmaf(-s3, dy, uy), inv3, gy);
    gz = fmaf(fmaf(-s3, dz, uz), inv3, gz);
    const float dw = fmaf(dx, p.x, fmaf(dy, p.y, dz * p.z));
    gm = fmaf(-dw, inv3, gm);
  }
  kahan_add32(S[0], C[0], ax);
  kahan_add32(S[1], C[1], ay);
  kahan_add32(S[2], C[2], az);
  kahan_add32(S[3], C[3], gx);
  kahan_add32(S[4], C[4], gy);
  kahan_add32(S[5], C[5], gz);
  kahan_add32(S[6], C[6], gm);
}

template <bool MASK>
__global__ __launch_bounds__(256) void k_adj_fused(const double4* __restrict__ src, const float4* __restrict__ wsrc,
                                                   const double* __restrict__ x, int n, int npad, int slice_len,
                                                   float eps2, double* __restrict__ out) {
  const int i = blockIdx.x * kApBlock + threadIdx.x;
  const int ii = i < n ? i : n - 1;
  const int sl = blockIdx.y;
  const int jb = sl * slice_len;
  const int je = min(jb + slice_len, npad);
  const double px = x[ii], py = x[n + ii], pz = x[2 * n + ii];
  const float4 wk = wsrc[ii];
  const float mk = wk.w;
  const int lane = threadIdx.x & 63;
  // the wavefront's targets i0 .. i0 + 63 are sources of tile i0 (tiles are
  // 64-aligned, as are the slices)
  const int self_tile = __builtin_amdgcn_readfirstlane(i - lane);
  float S[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  float C[7] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  for (int t = jb; t < je; t += kApTile) {
    if (t == self_tile)
      adj_tile<MASK, true>(src, wsrc, t, px, py, pz, wk, mk, eps2, lane, S, C);
    else
      adj_tile<MASK, false>(src, wsrc, t, px, py, pz, wk, mk, eps2, lane, S, C);
  }
  if (i < n) {
    double* o = out + (size_t)sl * 7 * n;
#pragma unroll
    for (int q = 0; q < 7; ++q) o[(size_t)q * n + i] = (double)S[q] - (double)C[q];
  }
}

// Sum slice partials in slice order and scale: a (3 planes) and gx (3), which
// the scaled masses enter once, by G 2^k (gsc); gm, free of masses, by G.
__global__ void k_adj_combine(int n, int S, const double* __restrict__ part, double gsc, double G,
                              double* __restrict__ a, double* __restrict__ gx, double* __restrict__ gm) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  // 4 slices per round, their 28 loads issued before the first add (clamped
  // slice index); the adds stay in slice order
  double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  for (int k0 = 0; k0 < S; k0 += 4) {
    double v[4][7];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int k = min(k0 + u, S - 1);
#pragma unroll
      for (int q = 0; q < 7; ++q) v[u][q] = part[((size_t)k * 7 + q) * n + i];
    }
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (k0 + u < S)
#pragma unroll
        for (int q = 0; q < 7; ++q) s[q] += v[u][q];
  }
#pragma unroll
  for (int q = 0; q < 7; ++q) {
    const double r = s[q] * (q < 6 ? gsc : G);
    if (q < 3) a[(size_t)q * n + i] = r;
    else if (q < 6) gx[(size_t)(q - 3) * n + i] = r;
    else gm[i] = r;
  }
}

// Reverse half of a step: lvh = lv + dt lx; vh = v - h a; x -= dt vh; v = vh;
// w = c lvh; lv = lvh. Optionally re-anchor x to the forward checkpoint.
__global__ __launch_bounds__(256) void k_adj_pre(int n, double dt, double c, double* __restrict__ x,
                                                 double* __restrict__ v, const double* __restrict__ a,
                                                 const double* __restrict__ lx, double* __restrict__ lv,
                                                 double* __restrict__ w, const double* __restrict__ ckx) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = 0.5 * dt;
  for (int q = 0; q < 3; ++q) {
    const size_t k = (size_t)q * n + i;
    const double lvh = lv[k] + dt * lx[k];
    const double vh = v[k] - h * a[k];
    x[k] = ckx ? ckx[k] : x[k] - dt * vh;
    v[k] = vh;
    lv[k] = lvh;
    w[k] = c * lvh;
  }
}

// v_s = vh - h a(x_s) (or the checkpoint velocity); lx += gx (+ snapshot
// gradient); lm += gm.
__global__ __launch_bounds__(256) void k_adj_post(int n, double dt, double* __restrict__ v,
                                                  const double* __restrict__ a, const double* __restrict__ ckv,
                                                  double* __restrict__ lx, const double* __restrict__ gx,
                                                  const double* __restrict__ gsnap, double* __restrict__ lm,
                                                  const double* __restrict__ gm) {
#pragma clang fp contract(off)
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = 0.5 * dt;
  for (int q = 0; q < 3; ++q) {
    const size_t k = (size_t)q * n + i;
    v[k] = ckv ? ckv[k] : v[k] - h * a[k];
    double g = lx[k] + gx[k];
    if (gsnap) g += gsnap[k];
    lx[k] = g;
  }
  lm[i] += gm[i];
}

__global__ void k_copy(size_t cnt, const double* __restrict__ s, double* __restrict__ d) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < cnt) d[i] = s[i];
}

// ----------------------------------------------------------------- host ----

static int dcopy(aster_ctx* c, const double* s, double* d, size_t cnt) {
  hipLaunchKernelGGL(k_copy, dim3((unsigned)((cnt + 255) / 256)), dim3(256), 0, c->stream, cnt, s, d);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

int traj_forward(aster_ctx* c, double dt, int64_t nsteps, int64_t ckpt, int32_t nsnap, const int64_t* snaps,
                 double* snap_pos) {
  AdjointState& S = c->adj;
  S.ready = false;
  if (!std::isfinite(dt) || nsteps < 1 || ckpt < 1 || nsnap < 0 || (nsnap > 0 && !snaps))
    return fail(c, ASTER_ERR_INVALID, "need dt finite, nsteps >= 1, ckpt_every >= 1, snapshot steps");
  for (int q = 0; q < nsnap; ++q)
    if (snaps[q] < 1 || snaps[q] > nsteps || (q > 0 && snaps[q] <= snaps[q - 1]))
      return fail(c, ASTER_ERR_INVALID, "snapshot steps must be ascending within [1, nsteps]");
  if (c->P.pn_flags & ASTER_PN_TERMS) return fail(c, ASTER_ERR_INVALID, "the adjoint covers Newtonian forces");
  if (c->P.lower > 0 || (c->P.upper >= 0 && c->P.upper < c->n))
    return fail(c, ASTER_ERR_INVALID, "the adjoint evaluates the full body range only");
  const int64_t n = c->n;
  const size_t n3 = (size_t)3 * n;
  const int64_t nck = nsteps / ckpt + 1;
  if ((double)nck * 2.0 * n3 * 8.0 > 64e9) return fail(c, ASTER_ERR_CAPACITY, "checkpoints would exceed 64 GB");
  ASTER_HIP(c, S.ckpt_x.reserve(nck * n3));
  ASTER_HIP(c, S.ckpt_v.reserve(nck * n3));
  ASTER_HIP(c, S.snapg.reserve(std::max<size_t>(1, (size_t)nsnap * n3)));
  S.nsteps = nsteps;
  S.ckpt_every = ckpt;
  S.nsnap = nsnap;
  S.dt = dt;
  S.snap_steps.assign(snaps, snaps + nsnap);
  int rc;
  const int method = ASTER_ALLPAIRS_FP32_KAHAN;
  if ((rc = compute_forces(c, method))) return rc;
  double pending = 0.0;
  int q = 0;
  for (int64_t s = 0; s < nsteps; ++s) {
    if (s % ckpt == 0) {
      if (pending != 0.0 && (rc = launch_kick(c, pending))) return rc;
      pending = 0.0;
      const size_t o = (size_t)(s / ckpt) * n3;
      if ((rc = dcopy(c, c->x.p, S.ckpt_x.p + o, n3)) || (rc = dcopy(c, c->v.p, S.ckpt_v.p + o, n3))) return rc;
    }
    if ((rc = launch_kick_drift(c, pending, dt))) return rc;
    if ((rc = compute_forces(c, method))) return rc;
    pending = dt;
    if (q < nsnap && snaps[q] == s + 1) {
      if ((rc = dcopy(c, c->x.p, S.snapg.p + (size_t)q * n3, n3))) return rc;
      ++q;
    }
  }
  if ((rc = launch_kick(c, pending))) return rc;
  if (snap_pos && nsnap > 0) {
    std::vector<double> h((size_t)nsnap * n3);
    ASTER_HIP(c, hipMemcpyAsync(h.data(), S.snapg.p, h.size() * 8, hipMemcpyDeviceToHost, c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
    for (int k = 0; k < nsnap; ++k)
      for (int64_t i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d)
          snap_pos[(size_t)k * n3 + 3 * i + d] = h[(size_t)k * n3 + (size_t)d * n + i];
  }
  S.ready = true;
  return ASTER_OK;
}

int traj_backward(aster_ctx* c, const double* gsnap, double* gx, double* gv, double* gm) {
  AdjointState& S = c->adj;
  if (!S.ready) return fail(c, ASTER_ERR_STATE, "call aster_traj_forward first (and do not step in between)");
  if (S.nsnap > 0 && !gsnap) return fail(c, ASTER_ERR_INVALID, "snapshot gradients required");
  const int n = (int)c->n;
  const int npad = (int)c->npad;
  const size_t n3 = (size_t)3 * n;
  const double dt = S.dt, h = 0.5 * dt;
  ASTER_HIP(c, S.lx.reserve(n3));
  ASTER_HIP(c, S.lv.reserve(n3));
  ASTER_HIP(c, S.lm.reserve(n));
  ASTER_HIP(c, S.w.reserve(n3));
  ASTER_HIP(c, S.g.reserve(n3 + n));
  ASTER_HIP(c, c->src64.reserve(npad));
  ASTER_HIP(c, S.wsrc.reserve(npad));
  // Snapshot gradients to device planes (reusing the snapshot buffer).
  if (S.nsnap > 0) {
    std::vector<double> hbuf((size_t)S.nsnap * n3);
    for (int k = 0; k < S.nsnap; ++k)
      for (int i = 0; i < n; ++i)
        for (int d = 0; d < 3; ++d) hbuf[(size_t)k * n3 + (size_t)d * n + i] = gsnap[(size_t)k * n3 + 3 * i + d];
    ASTER_HIP(c, hipMemcpyAsync(S.snapg.p, hbuf.data(), hbuf.size() * 8, hipMemcpyHostToDevice, c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
  }
  ASTER_HIP(c, hipMemsetAsync(S.lx.p, 0, n3 * 8, c->stream));
  ASTER_HIP(c, hipMemsetAsync(S.lv.p, 0, n3 * 8, c->stream));
  ASTER_HIP(c, hipMemsetAsync(S.lm.p, 0, n * 8, c->stream));
  int q = S.nsnap - 1;
  if (q >= 0 && S.snap_steps[q] == S.nsteps) {
    int rc = dcopy(c, S.snapg.p + (size_t)q * n3, S.lx.p, n3);
    if (rc) return rc;
    --q;
  }
  int Sl = choose_split(c);
  int64_t slice = round_up((npad + Sl - 1) / Sl, kApTile);
  Sl = (int)((npad + slice - 1) / slice);
  ASTER_HIP(c, S.part.reserve((size_t)Sl * 7 * n));
  Fp32Scales sc;
  int rc0 = fp32_scales(c, &sc);
  if (rc0) return rc0;
  const float eps2 = sc.eps2;
  const bool mask = sc.mask;
  float4* wsrc = S.wsrc.p;
  double* gxp = S.g.p;
  double* gmp = S.g.p + n3;
  const dim3 eg((n + 255) / 256), eb(256);
  for (int64_t s = S.nsteps - 1; s >= 0; --s) {
    const bool anchor = (s % S.ckpt_every) == 0;
    const size_t co = (size_t)(s / S.ckpt_every) * n3;
    const double cw = (s > 0) ? dt : h;
    phase_begin(c, ASTER_PHASE_KICKDRIFT);
    hipLaunchKernelGGL(k_adj_pre, eg, eb, 0, c->stream, n, dt, cw, c->x.p, c->v.p, c->a.p, S.lx.p, S.lv.p, S.w.p,
                       anchor ? S.ckpt_x.p + co : (const double*)nullptr);
    phase_end(c, ASTER_PHASE_KICKDRIFT);
    phase_begin(c, ASTER_PHASE_TILES);
    hipLaunchKernelGGL(k_adj_src, dim3(npad / 256), eb, 0, c->stream, n, npad, sc.msc, c->x.p, c->m.p, S.w.p,
                       c->src64.p, wsrc);
    phase_end(c, ASTER_PHASE_TILES);
    const dim3 grid(npad / kApBlock, Sl);
    phase_begin(c, ASTER_PHASE_ADJOINT);
    if (mask)
      hipLaunchKernelGGL((k_adj_fused<true>), grid, dim3(kApBlock), 0, c->stream, c->src64.p, wsrc, c->x.p, n, npad,
                         (int)slice, eps2, S.part.p);
    else
      hipLaunchKernelGGL((k_adj_fused<false>), grid, dim3(kApBlock), 0, c->stream, c->src64.p, wsrc, c->x.p, n, npad,
                         (int)slice, eps2, S.part.p);
    phase_end(c, ASTER_PHASE_ADJOINT);
    phase_begin(c, ASTER_PHASE_COMBINE);
    hipLaunchKernelGGL(k_adj_combine, eg, eb, 0, c->stream, n, Sl, S.part.p, sc.gsc, c->P.G, c->a.p, gxp, gmp);
    phase_end(c, ASTER_PHASE_COMBINE);
    const double* gs = nullptr;
    if (q >= 0 && S.snap_steps[q] == s) {
      gs = S.snapg.p + (size_t)q * n3;
      --q;
    }
    hipLaunchKernelGGL(k_adj_post, eg, eb, 0, c->stream, n, dt, c->v.p, c->a.p,
                       anchor ? S.ckpt_v.p + co : (const double*)nullptr, S.lx.p, gxp, gs, S.lm.p, gmp);
    ASTER_HIP(c, hipGetLastError());
  }
  // The state is back at step 0 (re-anchored to the first checkpoint) with
  // a = a(x_0) from the mixed-precision kernel.
  c->acc_valid = true;
  c->acc_stale = false;
  c->acc_method = ASTER_ALLPAIRS_FP32_KAHAN;
  S.ready = false;
  auto down = [&](const double* dev, double* host) -> int {
    std::vector<double> t(n3);
    ASTER_HIP(c, hipMemcpyAsync(t.data(), dev, n3 * 8, hipMemcpyDeviceToHost, c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
    for (int i = 0; i < n; ++i)
      for (int d = 0; d < 3; ++d) host[3 * (size_t)i + d] = t[(size_t)d * n + i];
    return ASTER_OK;
  };
  int rc;
  if (gx && (rc = down(S.lx.p, gx))) return rc;
  if (gv && (rc = down(S.lv.p, gv))) return rc;
  if (gm) {
    ASTER_HIP(c, hipMemcpyAsync(gm, S.lm.p, n * 8, hipMemcpyDeviceToHost, c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
  }
  return ASTER_OK;
}

}  // namespace aster
