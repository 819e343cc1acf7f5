// Batched ensembles: thousands of small independent systems (1..64 bodies)
// integrated by one persistent kernel per launch.
//
// Layout: every system lives in the lanes of one wavefront; with packing,
// several systems share a wavefront (systems are ordered by size, largest
// first, and filled next-fit into 64 lanes), so a 3-body system does not idle
// 61 lanes. Each lane keeps its body's FP64 position, velocity and
// acceleration in registers for the whole launch; the wavefront exchanges
// positions (and velocities for 1PN) through LDS once per force evaluation.
// Force per pair: all-pairs within the system, Newton (+ pairwise 1PN),
// optionally Plummer-softened.
//
// Two arithmetic modes:
//   fast  (precision != ASTER_FP64): FP64 with an rsqrt estimate refined by
//         two Newton steps; contraction allowed.
//   exact (precision == ASTER_FP64): compute_gravitational in the reference's
//         operation order and j order (src/physics/tool_chain_impl.cpp:120-137,
//         Parallelized::compute_forces, src/utils/builder_endpoint.cpp:729-757),
//         so each system's trajectory is bit-identical to the CPU path.
// Integrators as in aster_step: leapfrog (update_leapfrog,
// tool_chain_impl.cpp:169-185) or Yoshida-4, accelerations cached between
// steps and across launches.
//
// A step sequence is split into launches of steps_per_launch steps; the
// launches are captured once into a HIP graph and replayed.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "common.hpp"
#include "device.hpp"

namespace aster {

constexpr int kBatchMaxWpb = 4;

struct BatchArgs {
  int nwaves;
  const int32_t* lane_body;
  const int32_t* lane_seg;  // (segment base lane << 8) | segment size
  const int32_t* wave_max;  // largest segment in the wavefront
  double* x;
  double* v;
  double* a;
  const double* m;
  int64_t nsteps;
  int nsub;
  double h[3];
  int init;
  double G, ic2, ic, soft;
  int pnf;  // post-Newtonian terms (exact mode; the fast mode has PN1 only)
};

template <bool PN, bool EXACT>
__device__ __forceinline__ void batch_accel(int lane, int base, int cnt, int maxcnt, bool live, const double* lx,
                                            const double* ly, const double* lz, const double* lm, const double* lvx,
                                            const double* lvy, const double* lvz, const double* lv22,
                                            const double* X, const double* V, double mi, double G, double ic2,
                                            double ic, int pnf, double soft, double* A) {
  double ax = 0.0, ay = 0.0, az = 0.0;
  // fast mode, 1PN: this body's |v1|^2 and 5 G m1 (get_pn1's per-body parts)
  const double v11 = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
  const double g5m1 = 5.0 * G * mi;
  // Four sources per iteration: independent pair forces, sums still taken
  // one by one in j order (the exact mode's results are unchanged)
#pragma unroll 4
  for (int j = 0; j < maxcnt; ++j) {
    const int src = min(base + j, 63);
    const bool valid = live && j < cnt && src != lane;
    const double x2 = lx[src], y2 = ly[src], z2 = lz[src], m2 = lm[src];
    double v2x = 0.0, v2y = 0.0, v2z = 0.0;
    if (PN) {
      v2x = lvx[src];
      v2y = lvy[src];
      v2z = lvz[src];
    }
    if (EXACT) {
      double fx, fy, fz;
      grav64(mi, m2, V[0], V[1], V[2], v2x, v2y, v2z, X[0], X[1], X[2], x2, y2, z2, G, ic, soft, PN ? pnf : 0, fx, fy,
             fz);
      {
#pragma clang fp contract(off)
        const double nax = ax + fx / mi, nay = ay + fy / mi, naz = az + fz / mi;
        ax = valid ? nax : ax;
        ay = valid ? nay : ay;
        az = valid ? naz : az;
      }
    } else {
      const double dx = x2 - X[0], dy = y2 - X[1], dz = z2 - X[2];
      const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, soft)));
      // one Newton step: 1/r to 4.1e-15 (two steps, 1.4e-16, cost 4 more
      // instructions per pair; the exact mode divides as the reference does)
      const double inv = rsqrt64_n<1>(valid ? r2 : 1.0);
      const double inv2 = inv * inv;
      // a pair that does not count gets G m2 = 0 (its inv is finite), so
      // both force terms vanish without selects of their own
      const double gm2 = valid ? G * m2 : 0.0;
      const double f = gm2 * (inv2 * inv);
      ax = fma(f, dx, ax);
      ay = fma(f, dy, ay);
      az = fma(f, dz, az);
      if (PN) {
        // get_pn1 (tool_chain_impl.cpp:110-118) with n = (x1 - x2)/r, the same
        // terms grouped for fused multiply-adds:
        //   G m2/r^2 [n (-v1^2 - 2 v2^2 + 4 v1.v2 + 3/2 (n.v2)^2 + (5 G m1 + 4 G m2)/r)
        //             + (v1 - v2)(4 n.v1 - 3 n.v2)]
        const double nx = -dx * inv, ny = -dy * inv, nz = -dz * inv;
        const double nv1 = fma(nx, V[0], fma(ny, V[1], nz * V[2]));
        const double nv2 = fma(nx, v2x, fma(ny, v2y, nz * v2z));
        const double v12 = fma(V[0], v2x, fma(V[1], v2y, V[2] * v2z));
        const double v22 = lv22[src];  // |v2|^2, formed once per body when published
        double sc = fma(4.0, v12, -v11);
        sc = fma(-2.0, v22, sc);
        sc = fma(1.5 * nv2, nv2, sc);
        sc = fma(fma(4.0 * G, m2, g5m1), inv, sc);
        const double t = fma(4.0, nv1, -3.0 * nv2);
        const double k = gm2 * inv2 * ic2;
        ax = fma(fma(nx, sc, (V[0] - v2x) * t), k, ax);
        ay = fma(fma(ny, sc, (V[1] - v2y) * t), k, ay);
        az = fma(fma(nz, sc, (V[2] - v2z) * t), k, az);
      }
    }
  }
  A[0] = ax;
  A[1] = ay;
  A[2] = az;
}

// Full N-body 1PN (EIH) accelerations of a lane's body within its system
// (exact mode): pass 1 forms each body's potential and Newtonian
// acceleration and publishes them in LDS; pass 2 is the EIH sum, in the
// order and arithmetic of the oracle's EIH (sources in system order), so a
// system's trajectory equals the oracle's bit for bit.
__device__ __forceinline__ void batch_eih(int lane, int base, int cnt, int maxcnt, bool live, const double* lx,
                                          const double* ly, const double* lz, const double* lm, const double* lvx,
                                          const double* lvy, const double* lvz, const double* X, const double* V,
                                          double G, double ic, double soft, double* lphi, double* lanx, double* lany,
                                          double* lanz, double* A) {
#pragma clang fp contract(off)
  double ph = 0.0, nx = 0.0, ny = 0.0, nz = 0.0;
  for (int j = 0; j < maxcnt; ++j) {
    const int src = min(base + j, 63);
    const bool valid = live && j < cnt && src != lane;
    const double dx = lx[src] - X[0], dy = ly[src] - X[1], dz = lz[src] - X[2];
    const double r2 = dx * dx + dy * dy + dz * dz + soft;
    const double inv_r = 1.0 / sqrt(r2);
    const double inv_r3 = inv_r * inv_r * inv_r;
    const double mu = G * lm[src];
    if (valid) {
      ph += mu * inv_r;
      nx += mu * dx * inv_r3;
      ny += mu * dy * inv_r3;
      nz += mu * dz * inv_r3;
    }
  }
  lphi[lane] = ph;
  lanx[lane] = nx;
  lany[lane] = ny;
  lanz[lane] = nz;
  wave_sync();
  const double ic2 = ic * ic;
  const double vi2 = V[0] * V[0] + V[1] * V[1] + V[2] * V[2];
  double ax = 0.0, ay = 0.0, az = 0.0;
  for (int j = 0; j < maxcnt; ++j) {
    const int src = min(base + j, 63);
    const bool valid = live && j < cnt && src != lane;
    const double dx = lx[src] - X[0], dy = ly[src] - X[1], dz = lz[src] - X[2];
    const double r2 = dx * dx + dy * dy + dz * dz + soft;
    const double inv_r = 1.0 / sqrt(r2);
    const double inv_r3 = inv_r * inv_r * inv_r;
    const double mu = G * lm[src];
    const double v2x = lvx[src], v2y = lvy[src], v2z = lvz[src];
    const double anx = lanx[src], any = lany[src], anz = lanz[src];
    const double vj2 = v2x * v2x + v2y * v2y + v2z * v2z;
    const double vij = V[0] * v2x + V[1] * v2y + V[2] * v2z;
    const double rv = (dx * v2x + dy * v2y + dz * v2z) * inv_r;
    const double dan = dx * anx + dy * any + dz * anz;
    const double B = 1.0 - (4.0 * ph + lphi[src]) * ic2 + (vi2 + 2.0 * vj2 - 4.0 * vij) * ic2 - 1.5 * (rv * rv) * ic2 +
                     0.5 * dan * ic2;
    const double s1 = mu * inv_r3 * B;
    const double w = -(dx * (4.0 * V[0] - 3.0 * v2x) + dy * (4.0 * V[1] - 3.0 * v2y) + dz * (4.0 * V[2] - 3.0 * v2z));
    const double s2 = mu * inv_r3 * ic2 * w;
    const double s3 = 3.5 * mu * inv_r * ic2;
    if (valid) {
      ax += dx * s1;
      ax += (V[0] - v2x) * s2;
      ax += anx * s3;
      ay += dy * s1;
      ay += (V[1] - v2y) * s2;
      ay += any * s3;
      az += dz * s1;
      az += (V[2] - v2z) * s2;
      az += anz * s3;
    }
  }
  wave_sync();  // the next evaluation rewrites the LDS potentials
  A[0] = ax;
  A[1] = ay;
  A[2] = az;
}

__device__ __forceinline__ void kick_drift_reg(double* X, double* V, const double* A, double dt_prev, double dt) {
#pragma clang fp contract(off)
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double vc = V[c];
    if (dt_prev != 0.0) vc = vc + 0.5 * A[c] * dt_prev;
    X[c] = X[c] + (vc * dt + 0.5 * A[c] * dt * dt);
    V[c] = vc + 0.5 * A[c] * dt;
  }
}

__device__ __forceinline__ void kick_reg(double* V, const double* A, double dt) {
#pragma clang fp contract(off)
#pragma unroll
  for (int c = 0; c < 3; ++c) V[c] = V[c] + 0.5 * A[c] * dt;
}

template <bool PN, bool EXACT>
__global__ __launch_bounds__(256) void k_batch(BatchArgs P) {
  // positions, masses (and velocities) of each wavefront's bodies, one plane
  // per quantity in a single array, so a source's values sit at one address
  // plus constant plane offsets
  // (the fast 1PN mode also keeps each body's |v|^2, a per-source term of get_pn1)
  constexpr int kPlanes = PN ? (EXACT ? 7 : 8) : 4;
  __shared__ double planes[kPlanes][kBatchMaxWpb][64];
  auto sx = planes[0];
  auto sy = planes[1];
  auto sz = planes[2];
  auto sm = planes[3];
  auto svx = planes[PN ? 4 : 0];
  auto svy = planes[PN ? 5 : 0];
  auto svz = planes[PN ? 6 : 0];
  auto sv22 = planes[(PN && !EXACT) ? 7 : 0];
  constexpr int kE = (PN && EXACT) ? kBatchMaxWpb : 1;  // EIH potentials and Newtonian accelerations
  __shared__ double sphi[kE][64], sanx[kE][64], sany[kE][64], sanz[kE][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wl;
  if (wave >= P.nwaves) return;
  const int slot = wave * 64 + lane;
  const int b = P.lane_body[slot];
  const int seg = P.lane_seg[slot];
  const int base = seg >> 8, cnt = seg & 255;
  const int maxcnt = P.wave_max[wave];
  const bool live = b >= 0;
  double X[3] = {0.0, 0.0, 0.0}, V[3] = {0.0, 0.0, 0.0}, A[3] = {0.0, 0.0, 0.0};
  double mi = 0.0;
  if (live) {
    for (int c = 0; c < 3; ++c) {
      X[c] = P.x[3 * (size_t)b + c];
      V[c] = P.v[3 * (size_t)b + c];
      A[c] = P.a[3 * (size_t)b + c];
    }
    mi = P.m[b];
  }
  double* lx = sx[wl];
  double* ly = sy[wl];
  double* lz = sz[wl];
  double* lvx = svx[PN ? wl : 0];
  double* lvy = svy[PN ? wl : 0];
  double* lvz = svz[PN ? wl : 0];
  double* lv22 = sv22[wl];
  sm[wl][lane] = mi;
  auto publish = [&]() {
    lx[lane] = X[0];
    ly[lane] = X[1];
    lz[lane] = X[2];
    if (PN) {
      lvx[lane] = V[0];
      lvy[lane] = V[1];
      lvz[lane] = V[2];
      if (!EXACT) lv22[lane] = fma(V[0], V[0], fma(V[1], V[1], V[2] * V[2]));
    }
    wave_sync();
  };
  const bool eih = PN && EXACT && (P.pnf & ASTER_EIH);
  auto accel = [&]() {
    if (eih) {
      const int e = (PN && EXACT) ? wl : 0;
      batch_eih(lane, base, cnt, maxcnt, live, lx, ly, lz, sm[wl], lvx, lvy, lvz, X, V, P.G, P.ic, P.soft, sphi[e],
                sanx[e], sany[e], sanz[e], A);
    } else {
      batch_accel<PN, EXACT>(lane, base, cnt, maxcnt, live, lx, ly, lz, sm[wl], lvx, lvy, lvz, lv22, X, V, mi, P.G,
                             P.ic2, P.ic, P.pnf, P.soft, A);
    }
  };
  if (P.init) {
    publish();
    accel();
    wave_sync();
  }
  double pending = 0.0;
  for (int64_t s = 0; s < P.nsteps; ++s) {
    for (int k = 0; k < P.nsub; ++k) {
      const double h = P.h[k];
      if (PN) {
        // Velocity-dependent forces: re-evaluate after the final kick, as
        // update_leapfrog's first compute_forces does.
        if (pending != 0.0) kick_reg(V, A, pending);
        pending = 0.0;
        publish();
        accel();
        wave_sync();
      }
      kick_drift_reg(X, V, A, pending, h);
      publish();
      accel();
      wave_sync();
      pending = h;
    }
  }
  if (P.nsteps > 0) kick_reg(V, A, pending);
  if (live) {
    for (int c = 0; c < 3; ++c) {
      P.x[3 * (size_t)b + c] = X[c];
      P.v[3 * (size_t)b + c] = V[c];
      P.a[3 * (size_t)b + c] = A[c];
    }
  }
}

// Per-body energy terms: kinetic - 1/2 G m_i sum_j m_j / sqrt(r^2 + soft),
// and m x cross v; out [4][nbodies].
__global__ __launch_bounds__(256) void k_batch_energy(BatchArgs P, int nbodies, double* __restrict__ out) {
  __shared__ double sx[kBatchMaxWpb][64], sy[kBatchMaxWpb][64], sz[kBatchMaxWpb][64], sm[kBatchMaxWpb][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wl;
  if (wave >= P.nwaves) return;
  const int slot = wave * 64 + lane;
  const int b = P.lane_body[slot];
  const int seg = P.lane_seg[slot];
  const int base = seg >> 8, cnt = seg & 255;
  const int maxcnt = P.wave_max[wave];
  const bool live = b >= 0;
  double X[3] = {0.0, 0.0, 0.0}, V[3] = {0.0, 0.0, 0.0}, mi = 0.0;
  if (live) {
    for (int c = 0; c < 3; ++c) {
      X[c] = P.x[3 * (size_t)b + c];
      V[c] = P.v[3 * (size_t)b + c];
    }
    mi = P.m[b];
  }
  sx[wl][lane] = X[0];
  sy[wl][lane] = X[1];
  sz[wl][lane] = X[2];
  sm[wl][lane] = mi;
  wave_sync();
  double pot = 0.0;
  for (int j = 0; j < maxcnt; ++j) {
    const int src = min(base + j, 63);
    const bool valid = live && j < cnt && src != lane;
    const double dx = sx[wl][src] - X[0], dy = sy[wl][src] - X[1], dz = sz[wl][src] - X[2];
    const double r2 = dx * dx + dy * dy + dz * dz + P.soft;
    if (valid) pot += sm[wl][src] / sqrt(r2);
  }
  if (live) {
    out[b] = 0.5 * mi * (V[0] * V[0] + V[1] * V[1] + V[2] * V[2]) - 0.5 * P.G * mi * pot;
    out[(size_t)nbodies + b] = mi * (X[1] * V[2] - X[2] * V[1]);
    out[(size_t)2 * nbodies + b] = mi * (X[2] * V[0] - X[0] * V[2]);
    out[(size_t)3 * nbodies + b] = mi * (X[0] * V[1] - X[1] * V[0]);
  }
}

__global__ void k_batch_sum(int64_t nsys, const int64_t* __restrict__ off, int64_t nbodies,
                            const double* __restrict__ eb, double* __restrict__ es) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  for (int q = 0; q < 4; ++q) {
    double acc = 0.0;
    for (int64_t k = off[s]; k < off[s + 1]; ++k) acc += eb[(size_t)q * nbodies + k];
    es[(size_t)q * nsys + s] = acc;
  }
}

// ----------------------------------------------------------------- host ----

// ---- Wisdom-Holman for central-mass systems ------------------------------
//
// Democratic-heliocentric Wisdom-Holman (Wisdom & Holman 1991; Duncan, Levison
// & Lee 1998). The reference's update_WH_planetary (tool_chain_impl.cpp:
// 235-274) drifts each body i >= 1 on a Kepler orbit about body 0 and then
// kicks with the full accelerations, which include body 0's pull again; this
// is the splitting that scheme is meant to be: H = H_Kepler(mu = G m_0) +
// H_interaction(planet pairs) + |sum m_i u_i|^2 / (2 m_0) in heliocentric
// positions Q_i = x_i - x_0 and barycentric velocities u_i. One step: half
// interaction kick, half jump Q_i += (sum m_j u_j / m_0) dt/2, Kepler drift,
// half jump, half kick. Body 0 of each system is the central mass. During a
// call the batch buffers hold the heliocentric state (planets: Q, u, and
// their interaction accelerations; body 0: the barycentre and its velocity),
// so launches can split the steps anywhere without changing a bit. Segment
// sums are taken by the head lane in body order, and pair forces are
// compute_gravitational (tool_chain_impl.cpp:120-137) in j order, Newtonian,
// so every system is bit-identical to the CPU restatement.
enum { WH_ENTER = 0, WH_STEPS = 1, WH_EXIT = 2 };

struct WhArgs {
  int nwaves;
  const int32_t* lane_body;
  const int32_t* lane_seg;
  double* x;
  double* v;
  double* a;
  const double* m;
  int64_t nsteps;
  double dt, G, soft;
  int phase;
};

__global__ __launch_bounds__(256) void k_batch_wh(WhArgs P) {
#pragma clang fp contract(off)
  __shared__ double sx[kBatchMaxWpb][64], sy[kBatchMaxWpb][64], sz[kBatchMaxWpb][64], sm[kBatchMaxWpb][64];
  __shared__ double tx[kBatchMaxWpb][64], ty[kBatchMaxWpb][64], tz[kBatchMaxWpb][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wl;
  if (wave >= P.nwaves) return;
  const int slot = wave * 64 + lane;
  const int b = P.lane_body[slot];
  const int seg = P.lane_seg[slot];
  const int base = seg >> 8, cnt = seg & 255;
  const bool live = b >= 0;
  const bool head = live && lane == base;   // the system's central body
  const bool planet = live && lane != base;
  double* lx = sx[wl];
  double* ly = sy[wl];
  double* lz = sz[wl];
  double* lm = sm[wl];
  double* ux = tx[wl];
  double* uy = ty[wl];
  double* uz = tz[wl];
  double X[3] = {0.0, 0.0, 0.0}, V[3] = {0.0, 0.0, 0.0}, A[3] = {0.0, 0.0, 0.0};
  double mi = 0.0;
  if (live) {
    mi = P.m[b];
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      X[c] = P.x[3 * (size_t)b + c];
      V[c] = P.v[3 * (size_t)b + c];
      A[c] = P.a[3 * (size_t)b + c];
    }
  }
  lm[lane] = mi;
  wave_sync();
  const double m0 = live ? lm[base] : 1.0;
  const double mu = P.G * m0;
  const double h = 0.5 * P.dt;

  // interaction accelerations of the planets at the positions in X
  auto interaction = [&]() {
    lx[lane] = X[0];
    ly[lane] = X[1];
    lz[lane] = X[2];
    wave_sync();
    double ax = 0.0, ay = 0.0, az = 0.0;
    if (planet)
      for (int j = 1; j < cnt; ++j) {
        const int src = base + j;
        if (src == lane) continue;
        double fx, fy, fz;
        grav64(mi, lm[src], 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, X[0], X[1], X[2], lx[src], ly[src], lz[src], P.G, 0.0,
               P.soft, false, fx, fy, fz);
        ax += fx / mi;
        ay += fy / mi;
        az += fz / mi;
      }
    A[0] = ax;
    A[1] = ay;
    A[2] = az;
    wave_sync();
  };
  // sum over the system's planets of m_j w_j (j order), by the head lane
  auto planet_sum = [&](const double* W, double* out) {
    ux[lane] = planet ? mi * W[0] : 0.0;
    uy[lane] = planet ? mi * W[1] : 0.0;
    uz[lane] = planet ? mi * W[2] : 0.0;
    wave_sync();
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    if (head)
      for (int j = 1; j < cnt; ++j) {
        s0 += ux[base + j];
        s1 += uy[base + j];
        s2 += uz[base + j];
      }
    wave_sync();
    if (head) {
      ux[lane] = s0;
      uy[lane] = s1;
      uz[lane] = s2;
    }
    wave_sync();
    if (live) {
      out[0] = ux[base];
      out[1] = uy[base];
      out[2] = uz[base];
    }
    wave_sync();
  };
  auto jump = [&]() {
    double S[3] = {0.0, 0.0, 0.0};
    planet_sum(V, S);
    if (planet)
#pragma unroll
      for (int c = 0; c < 3; ++c) X[c] = X[c] + S[c] / m0 * h;
  };

  if (P.phase == WH_ENTER) {
    // M, barycentre and its velocity over all bodies (j order), then
    // Q = x - x_0, u = v - V for the planets; body 0 keeps (X, V).
    lx[lane] = mi * X[0];
    ly[lane] = mi * X[1];
    lz[lane] = mi * X[2];
    ux[lane] = mi * V[0];
    uy[lane] = mi * V[1];
    uz[lane] = mi * V[2];
    wave_sync();
    double M = 0.0, Sx[3] = {0.0, 0.0, 0.0}, Sv[3] = {0.0, 0.0, 0.0};
    if (head)
      for (int j = 0; j < cnt; ++j) {
        M += lm[base + j];
        Sx[0] += lx[base + j];
        Sx[1] += ly[base + j];
        Sx[2] += lz[base + j];
        Sv[0] += ux[base + j];
        Sv[1] += uy[base + j];
        Sv[2] += uz[base + j];
      }
    wave_sync();
    if (head) {
      // publish x_0 (this lane's own position) and the barycentre velocity
      lx[lane] = X[0];
      ly[lane] = X[1];
      lz[lane] = X[2];
      ux[lane] = Sv[0] / M;
      uy[lane] = Sv[1] / M;
      uz[lane] = Sv[2] / M;
    }
    wave_sync();
    if (planet) {
      const double x0[3] = {lx[base], ly[base], lz[base]};
      const double vc[3] = {ux[base], uy[base], uz[base]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        X[c] = X[c] - x0[c];
        V[c] = V[c] - vc[c];
      }
    }
    if (head)
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        X[c] = Sx[c] / M;
        V[c] = Sv[c] / M;
      }
    wave_sync();
    interaction();
  } else if (P.phase == WH_STEPS) {
    for (int64_t s = 0; s < P.nsteps; ++s) {
      if (planet)
#pragma unroll
        for (int c = 0; c < 3; ++c) V[c] = V[c] + A[c] * h;
      jump();
      if (planet) kepler_drift64(X, V, mu, P.dt);
      jump();
      interaction();
      if (planet)
#pragma unroll
        for (int c = 0; c < 3; ++c) V[c] = V[c] + A[c] * h;
      if (head)
#pragma unroll
        for (int c = 0; c < 3; ++c) X[c] = X[c] + V[c] * P.dt;
    }
  } else {
    // x_0 = X - sum m_j Q_j / M, x_i = Q_i + x_0, v_0 = V - sum m_j u_j / m_0,
    // v_i = u_i + V
    double SQ[3] = {0.0, 0.0, 0.0}, SU[3] = {0.0, 0.0, 0.0};
    planet_sum(X, SQ);
    planet_sum(V, SU);
    if (head) {
      double M = 0.0;
      for (int j = 0; j < cnt; ++j) M += lm[base + j];
#pragma unroll
      for (int c = 0; c < 3; ++c) X[c] = X[c] - SQ[c] / M;
      lx[lane] = X[0];
      ly[lane] = X[1];
      lz[lane] = X[2];
      ux[lane] = V[0];
      uy[lane] = V[1];
      uz[lane] = V[2];
    }
    wave_sync();
    if (planet) {
      const double x0[3] = {lx[base], ly[base], lz[base]};
      const double vb[3] = {ux[base], uy[base], uz[base]};
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        X[c] = X[c] + x0[c];
        V[c] = V[c] + vb[c];
      }
    }
    if (head)
#pragma unroll
      for (int c = 0; c < 3; ++c) V[c] = V[c] - SU[c] / m0;
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] = 0.0;
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      P.x[3 * (size_t)b + c] = X[c];
      P.v[3 * (size_t)b + c] = V[c];
      P.a[3 * (size_t)b + c] = A[c];
    }
  }
}

// Wavefront plan. pack = 0: one system per wavefront. pack = 1: first fit
// decreasing: each wavefront opens with the largest system left (its size is
// the wavefront's loop length) and fills its free lanes with the largest
// systems left that fit. Against next fit over the sorted sizes this fills the
// lanes beside a 33..63-body system instead of leaving them idle (lane use of
// 8192 systems of 3..64 bodies: 0.87 instead of 0.78).
static int upload_plan(aster_ctx* c) {
  BatchState& B = c->batch;
  const int64_t nsys = B.nsys;
  std::vector<std::vector<int64_t>> by_size(65);  // systems of each size, in index order
  for (int64_t s = 0; s < nsys; ++s) by_size[(size_t)(B.offsets[s + 1] - B.offsets[s])].push_back(s);
  std::vector<size_t> next(65, 0);  // first unplaced system of each size
  std::vector<int32_t> lb, ls, wm;
  auto place = [&](int64_t s, int used) {
    const int sz = (int)(B.offsets[s + 1] - B.offsets[s]);
    const size_t w = wm.size() - 1;
    for (int k = 0; k < sz; ++k) {
      lb[w * 64 + used + k] = (int32_t)(B.offsets[s] + k);
      ls[w * 64 + used + k] = (used << 8) | sz;
    }
    wm[w] = std::max(wm[w], sz);
    return used + sz;
  };
  for (int top = 64; top >= 1;) {
    if (next[top] == by_size[top].size()) {
      --top;
      continue;
    }
    lb.resize(lb.size() + 64, -1);
    ls.resize(ls.size() + 64, 0);
    wm.push_back(0);
    int used = place(by_size[top][next[top]++], 0);
    if (B.pack != 0)
      for (int sz = std::min(64 - used, top); sz >= 1; --sz)
        while (used + sz <= 64 && next[sz] < by_size[sz].size()) used = place(by_size[sz][next[sz]++], used);
  }
  B.nwaves = (int)wm.size();
  B.lane_slots = 0;
  for (int32_t mx : wm) B.lane_slots += 64 * (int64_t)mx;
  ASTER_HIP(c, B.lane_body.reserve(lb.size()));
  ASTER_HIP(c, B.lane_seg.reserve(ls.size()));
  ASTER_HIP(c, B.wave_max.reserve(wm.size()));
  ASTER_HIP(c, hipMemcpyAsync(B.lane_body.p, lb.data(), lb.size() * 4, hipMemcpyHostToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.lane_seg.p, ls.data(), ls.size() * 4, hipMemcpyHostToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.wave_max.p, wm.data(), wm.size() * 4, hipMemcpyHostToDevice, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return ASTER_OK;
}

static void drop_graph(BatchState& B) {
  if (B.exec) (void)hipGraphExecDestroy(B.exec);
  if (B.graph) (void)hipGraphDestroy(B.graph);
  B.exec = nullptr;
  B.graph = nullptr;
  B.g_steps = -1;
}

// New parameters (G, c, softening, force law, precision): the cached
// accelerations and the captured launches hold the old ones.
void batch_params_changed(aster_ctx* c) {
  if (c->batch.exec) (void)hipStreamSynchronize(c->stream);  // a replay may still run
  drop_graph(c->batch);
  c->batch.a_valid = false;
}

static BatchArgs make_args(aster_ctx* c, int64_t nsteps, int nsub, const double* h, int init) {
  BatchState& B = c->batch;
  BatchArgs P;
  P.nwaves = B.nwaves;
  P.lane_body = B.lane_body.p;
  P.lane_seg = B.lane_seg.p;
  P.wave_max = B.wave_max.p;
  P.x = B.x.p;
  P.v = B.v.p;
  P.a = B.a.p;
  P.m = B.m.p;
  P.nsteps = nsteps;
  P.nsub = nsub;
  for (int k = 0; k < 3; ++k) P.h[k] = h[k];
  P.init = init;
  P.G = c->P.G;
  P.ic2 = (1.0 / c->P.c) * (1.0 / c->P.c);
  P.ic = 1.0 / c->P.c;
  P.pnf = c->P.pn_flags & ASTER_PN_TERMS;
  P.soft = c->P.softening;
  return P;
}

static int launch_batch(aster_ctx* c, const BatchArgs& P) {
  if ((c->P.pn_flags & ASTER_EIH) && c->P.precision != ASTER_FP64)
    return fail(c, ASTER_ERR_INVALID, "ASTER_EIH in ensembles needs precision ASTER_FP64 (exact mode)");
  const int wpb = c->batch.waves_per_block;
  const dim3 grid((P.nwaves + wpb - 1) / wpb), block(64 * wpb);
  const bool pn = (c->P.pn_flags & ASTER_PN_TERMS) != 0;
  const bool exact = c->P.precision == ASTER_FP64;
  if (pn && exact) hipLaunchKernelGGL((k_batch<true, true>), grid, block, 0, c->stream, P);
  else if (pn) hipLaunchKernelGGL((k_batch<true, false>), grid, block, 0, c->stream, P);
  else if (exact) hipLaunchKernelGGL((k_batch<false, true>), grid, block, 0, c->stream, P);
  else hipLaunchKernelGGL((k_batch<false, false>), grid, block, 0, c->stream, P);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

static int launch_wh(aster_ctx* c, int phase, double dt, int64_t nsteps) {
  BatchState& B = c->batch;
  WhArgs P;
  P.nwaves = B.nwaves;
  P.lane_body = B.lane_body.p;
  P.lane_seg = B.lane_seg.p;
  P.x = B.x.p;
  P.v = B.v.p;
  P.a = B.a.p;
  P.m = B.m.p;
  P.nsteps = nsteps;
  P.dt = dt;
  P.G = c->P.G;
  P.soft = c->P.softening;
  P.phase = phase;
  const int wpb = B.waves_per_block;
  hipLaunchKernelGGL(k_batch_wh, dim3((P.nwaves + wpb - 1) / wpb), dim3(64 * wpb), 0, c->stream, P);
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

static int substeps(int integrator, double dt, double* h) {
  if (integrator == ASTER_LEAPFROG) {
    h[0] = dt;
    h[1] = h[2] = 0.0;
    return 1;
  }
  double w[3];
  aster_yoshida_weights(w);
  for (int k = 0; k < 3; ++k) h[k] = w[k] * dt;
  return 3;
}

int batch_step(aster_ctx* c, int integrator, double dt, int64_t nsteps, int spl, int use_graph) {
  BatchState& B = c->batch;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  const bool wh = integrator == ASTER_WH_PLANETARY;
  if (integrator != ASTER_LEAPFROG && integrator != ASTER_YOSHIDA4 && !wh)
    return fail(c, ASTER_ERR_INVALID, "integrator must be ASTER_LEAPFROG, ASTER_YOSHIDA4 or ASTER_WH_PLANETARY");
  if (!std::isfinite(dt) || nsteps < 0 || spl < 0) return fail(c, ASTER_ERR_INVALID, "invalid dt, nsteps or spl");
  const bool pn = (c->P.pn_flags & ASTER_PN_TERMS) != 0;
  if ((c->P.pn_flags & ASTER_EIH) && c->P.precision != ASTER_FP64)
    return fail(c, ASTER_ERR_INVALID, "ASTER_EIH in ensembles needs precision ASTER_FP64 (exact mode)");
  if (wh && pn) return fail(c, ASTER_ERR_INVALID, "ASTER_WH_PLANETARY is Newtonian: post-Newtonian terms are not supported");
  if ((c->P.pn_flags & (ASTER_PN2 | ASTER_PN25)) && c->P.precision != ASTER_FP64)
    return fail(c, ASTER_ERR_INVALID, "ASTER_PN2 / ASTER_PN25 in ensembles need precision ASTER_FP64 (exact mode)");
  double h[3] = {0.0, 0.0, 0.0};
  const int nsub = wh ? 1 : substeps(integrator, dt, h);
  const double h0[3] = {0.0, 0.0, 0.0};
  int rc;
  if (!wh && !B.a_valid) {
    if ((rc = launch_batch(c, make_args(c, 0, 1, h0, 1)))) return rc;
    B.a_valid = true;
  }
  if (nsteps == 0) return ASTER_OK;
  if (spl == 0) spl = B.steps_per_launch > 0 ? B.steps_per_launch : 256;
  const int64_t full = nsteps / spl, rem = nsteps % spl;
  const int exact = c->P.precision == ASTER_FP64;
  // every launch of this call; Wisdom-Holman converts to heliocentric
  // coordinates before and back after (the accelerations are then stale)
  auto enqueue = [&]() -> int {
    int e = ASTER_OK;
    if (wh && (e = launch_wh(c, WH_ENTER, dt, 0))) return e;
    for (int64_t l = 0; l < full && !e; ++l)
      e = wh ? launch_wh(c, WH_STEPS, dt, spl) : launch_batch(c, make_args(c, spl, nsub, h, 0));
    if (!e && rem) e = wh ? launch_wh(c, WH_STEPS, dt, rem) : launch_batch(c, make_args(c, rem, nsub, h, 0));
    if (!e && wh) e = launch_wh(c, WH_EXIT, dt, 0);
    return e;
  };
  if (wh) B.a_valid = false;
  if (!use_graph) return enqueue();
  const bool same = B.exec && B.g_steps == nsteps && B.g_integrator == integrator && B.g_spl == spl &&
                    B.g_dt == dt && B.g_pn == c->P.pn_flags && B.g_exact == exact && B.g_wpb == B.waves_per_block &&
                    B.g_pack == B.pack && B.g_G == c->P.G && B.g_c == c->P.c && B.g_soft == c->P.softening;
  if (!same) {
    if (B.exec) ASTER_HIP(c, hipStreamSynchronize(c->stream));  // a replay of the old graph may still run
    drop_graph(B);
    ASTER_HIP(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int lrc = enqueue();
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (lrc) {
      if (g) (void)hipGraphDestroy(g);
      return lrc;
    }
    if (e != hipSuccess) return hip_fail(c, e, "hipStreamEndCapture");
    B.graph = g;
    ASTER_HIP(c, hipGraphInstantiate(&B.exec, B.graph, nullptr, nullptr, 0));
    B.g_steps = nsteps;
    B.g_integrator = integrator;
    B.g_spl = spl;
    B.g_dt = dt;
    B.g_pn = c->P.pn_flags;  // the captured arguments carry every force-law flag
    B.g_exact = exact;
    B.g_wpb = B.waves_per_block;
    B.g_pack = B.pack;
    B.g_G = c->P.G;
    B.g_c = c->P.c;
    B.g_soft = c->P.softening;
  }
  ASTER_HIP(c, hipGraphLaunch(B.exec, c->stream));
  return ASTER_OK;
}

int batch_energy(aster_ctx* c, double* e, double* L) {
  BatchState& B = c->batch;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  ASTER_HIP(c, B.ebody.reserve((size_t)4 * B.nbodies));
  ASTER_HIP(c, B.esys.reserve((size_t)4 * B.nsys));
  const double h0[3] = {0.0, 0.0, 0.0};
  const BatchArgs P = make_args(c, 0, 1, h0, 0);
  const int wpb = B.waves_per_block;
  hipLaunchKernelGGL(k_batch_energy, dim3((P.nwaves + wpb - 1) / wpb), dim3(64 * wpb), 0, c->stream, P,
                     (int)B.nbodies, B.ebody.p);
  ASTER_HIP(c, hipGetLastError());
  hipLaunchKernelGGL(k_batch_sum, dim3((unsigned)((B.nsys + 255) / 256)), dim3(256), 0, c->stream, B.nsys, B.doff.p,
                     B.nbodies, B.ebody.p, B.esys.p);
  ASTER_HIP(c, hipGetLastError());
  std::vector<double> h((size_t)4 * B.nsys);
  ASTER_HIP(c, hipMemcpyAsync(h.data(), B.esys.p, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t s = 0; s < B.nsys; ++s) {
    if (e) e[s] = h[s];
    if (L)
      for (int q = 0; q < 3; ++q) L[3 * s + q] = h[(size_t)(1 + q) * B.nsys + s];
  }
  return ASTER_OK;
}

// Times 16 steps of every (packing, wavefronts-per-block) candidate on the
// loaded batch, restoring the state after each trial, and keeps the fastest.
// Steps per launch are then chosen so that one launch lasts about 20 ms.
int batch_autotune(aster_ctx* c, int integrator, double dt, int32_t* spb, int32_t* spl) {
  BatchState& B = c->batch;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  const size_t nb3 = (size_t)3 * B.nbodies;
  DBuf<double> save;
  ASTER_HIP(c, save.reserve(3 * nb3));
  int rc;
  const double h0[3] = {0.0, 0.0, 0.0};
  if (!B.a_valid) {
    if ((rc = launch_batch(c, make_args(c, 0, 1, h0, 1)))) return rc;
    B.a_valid = true;
  }
  ASTER_HIP(c, hipMemcpyAsync(save.p, B.x.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(save.p + nb3, B.v.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(save.p + 2 * nb3, B.a.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  const bool wh = integrator == ASTER_WH_PLANETARY;
  double h[3] = {0.0, 0.0, 0.0};
  const int nsub = wh ? 1 : substeps(integrator, dt, h);
  const int trial = 16;
  float best = 1e30f;
  int best_pack = B.pack, best_wpb = B.waves_per_block;
  for (int pack = 0; pack <= 1; ++pack) {
    B.pack = pack;
    if ((rc = upload_plan(c))) return rc;
    for (int wpb = 1; wpb <= kBatchMaxWpb; wpb *= 2) {
      B.waves_per_block = wpb;
      if (wh && (rc = launch_wh(c, WH_ENTER, dt, 0))) return rc;
      rc = wh ? launch_wh(c, WH_STEPS, dt, 2) : launch_batch(c, make_args(c, 2, nsub, h, 0));  // warm
      if (rc) return rc;
      ASTER_HIP(c, hipEventRecord(c->ev[14], c->stream));
      rc = wh ? launch_wh(c, WH_STEPS, dt, trial) : launch_batch(c, make_args(c, trial, nsub, h, 0));
      if (rc) return rc;
      ASTER_HIP(c, hipEventRecord(c->ev[15], c->stream));
      ASTER_HIP(c, hipEventSynchronize(c->ev[15]));
      float ms = 0.f;
      ASTER_HIP(c, hipEventElapsedTime(&ms, c->ev[14], c->ev[15]));
      if (ms < best) {
        best = ms;
        best_pack = pack;
        best_wpb = wpb;
      }
      ASTER_HIP(c, hipMemcpyAsync(B.x.p, save.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
      ASTER_HIP(c, hipMemcpyAsync(B.v.p, save.p + nb3, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
      ASTER_HIP(c, hipMemcpyAsync(B.a.p, save.p + 2 * nb3, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
    }
  }
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  save.release();
  B.pack = best_pack;
  B.waves_per_block = best_wpb;
  if ((rc = upload_plan(c))) return rc;
  const double ms_per_step = best / trial;
  int s = (int)std::max(1.0, std::floor(20.0 / std::max(ms_per_step, 1e-6)));
  B.steps_per_launch = std::min(s, 100000);
  drop_graph(B);
  if (spb) *spb = B.waves_per_block;
  if (spl) *spl = B.steps_per_launch;
  return ASTER_OK;
}

// ------------------------------------------------- ensemble trajectory adjoint
// Leapfrog with Newtonian FP64 forces (fast mode arithmetic); the positions of
// every step are stored (small systems: 4096 three-body systems x 513 steps is
// 150 MB), so the backward pass evaluates each vector-Jacobian product at the
// exact forward positions. Recursion as in adjoint.hip:
//   lvh = lv + dt lx ; w = c_s lvh (c_s = dt, c_0 = dt/2)
//   lx += J(x_s)^T w [+ dL/dx_s] ; lm += M(x_s)^T w ; lv = lvh
// with, per body k and partner j in its system (d = x_j - x_k, s^2 = |d|^2 + eps^2):
//   J^T w : G sum_j [u/s^3 - 3 d (d.u)/s^5],  u = m_k w_j - m_j w_k
//   M^T w : -G sum_j (w_j . d)/s^3

__global__ __launch_bounds__(256) void k_batch_traj_fwd(BatchArgs P, double* __restrict__ traj, int64_t nbodies) {
  __shared__ double sx[kBatchMaxWpb][64], sy[kBatchMaxWpb][64], sz[kBatchMaxWpb][64], sm[kBatchMaxWpb][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wl;
  if (wave >= P.nwaves) return;
  const int slot = wave * 64 + lane;
  const int b = P.lane_body[slot];
  const int seg = P.lane_seg[slot];
  const int base = seg >> 8, cnt = seg & 255;
  const int maxcnt = P.wave_max[wave];
  const bool live = b >= 0;
  double X[3] = {0.0, 0.0, 0.0}, V[3] = {0.0, 0.0, 0.0}, A[3] = {0.0, 0.0, 0.0};
  double mi = 0.0;
  if (live) {
    for (int c = 0; c < 3; ++c) {
      X[c] = P.x[3 * (size_t)b + c];
      V[c] = P.v[3 * (size_t)b + c];
    }
    mi = P.m[b];
  }
  sm[wl][lane] = mi;
  auto publish = [&]() {
    sx[wl][lane] = X[0];
    sy[wl][lane] = X[1];
    sz[wl][lane] = X[2];
    wave_sync();
  };
  auto store = [&](int64_t s) {
    if (live)
      for (int c = 0; c < 3; ++c) traj[((size_t)s * nbodies + b) * 3 + c] = X[c];
  };
  publish();
  batch_accel<false, false>(lane, base, cnt, maxcnt, live, sx[wl], sy[wl], sz[wl], sm[wl], nullptr, nullptr,
                            nullptr, nullptr, X, V, mi, P.G, P.ic2, P.ic, P.pnf, P.soft, A);
  wave_sync();
  store(0);
  const double dt = P.h[0];
  double pending = 0.0;
  for (int64_t s = 0; s < P.nsteps; ++s) {
    kick_drift_reg(X, V, A, pending, dt);
    publish();
    batch_accel<false, false>(lane, base, cnt, maxcnt, live, sx[wl], sy[wl], sz[wl], sm[wl], nullptr, nullptr,
                              nullptr, nullptr, X, V, mi, P.G, P.ic2, P.ic, P.pnf, P.soft, A);
    wave_sync();
    pending = dt;
    store(s + 1);
  }
  if (P.nsteps > 0) kick_reg(V, A, pending);
  if (live)
    for (int c = 0; c < 3; ++c) {
      P.x[3 * (size_t)b + c] = X[c];
      P.v[3 * (size_t)b + c] = V[c];
      P.a[3 * (size_t)b + c] = A[c];
    }
}

__global__ __launch_bounds__(256) void k_batch_traj_bwd(BatchArgs P, const double* __restrict__ traj,
                                                        int64_t nbodies, int64_t nsteps, double dt,
                                                        const double* __restrict__ gsnap,
                                                        const int64_t* __restrict__ snaps, int nsnap,
                                                        double* __restrict__ out) {
  __shared__ double sx[kBatchMaxWpb][64], sy[kBatchMaxWpb][64], sz[kBatchMaxWpb][64], sm[kBatchMaxWpb][64];
  __shared__ double swx[kBatchMaxWpb][64], swy[kBatchMaxWpb][64], swz[kBatchMaxWpb][64];
  const int wl = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int wave = blockIdx.x * (blockDim.x >> 6) + wl;
  if (wave >= P.nwaves) return;
  const int slot = wave * 64 + lane;
  const int b = P.lane_body[slot];
  const int seg = P.lane_seg[slot];
  const int base = seg >> 8, cnt = seg & 255;
  const int maxcnt = P.wave_max[wave];
  const bool live = b >= 0;
  const double mk = live ? P.m[b] : 0.0;
  sm[wl][lane] = mk;
  double lx[3] = {0.0, 0.0, 0.0}, lv[3] = {0.0, 0.0, 0.0}, lm = 0.0;
  int q = nsnap - 1;
  if (q >= 0 && snaps[q] == nsteps) {
    if (live)
      for (int c = 0; c < 3; ++c) lx[c] = gsnap[((size_t)q * nbodies + b) * 3 + c];
    --q;
  }
  for (int64_t s = nsteps - 1; s >= 0; --s) {
    const double cw = (s > 0) ? dt : 0.5 * dt;
    double w[3], X[3];
    for (int c = 0; c < 3; ++c) {
      lv[c] = lv[c] + dt * lx[c];
      w[c] = cw * lv[c];
      X[c] = live ? traj[((size_t)s * nbodies + b) * 3 + c] : 0.0;
    }
    sx[wl][lane] = X[0];
    sy[wl][lane] = X[1];
    sz[wl][lane] = X[2];
    swx[wl][lane] = w[0];
    swy[wl][lane] = w[1];
    swz[wl][lane] = w[2];
    wave_sync();
    double gx = 0.0, gy = 0.0, gz = 0.0, gm = 0.0;
    for (int j = 0; j < maxcnt; ++j) {
      const int src = min(base + j, 63);
      const bool valid = live && j < cnt && src != lane;
      const double dx = sx[wl][src] - X[0], dy = sy[wl][src] - X[1], dz = sz[wl][src] - X[2];
      const double r2 = fma(dx, dx, fma(dy, dy, fma(dz, dz, P.soft)));
      const double inv = rsqrt64(valid ? r2 : 1.0);
      const double inv3 = valid ? inv * inv * inv : 0.0;
      const double inv5 = inv3 * inv * inv;
      const double mj = sm[wl][src];
      const double wjx = swx[wl][src], wjy = swy[wl][src], wjz = swz[wl][src];
      const double ux = mk * wjx - mj * w[0], uy = mk * wjy - mj * w[1], uz = mk * wjz - mj * w[2];
      const double du = dx * ux + dy * uy + dz * uz;
      gx += ux * inv3 - 3.0 * dx * du * inv5;
      gy += uy * inv3 - 3.0 * dy * du * inv5;
      gz += uz * inv3 - 3.0 * dz * du * inv5;
      gm -= (wjx * dx + wjy * dy + wjz * dz) * inv3;
    }
    wave_sync();
    lx[0] += P.G * gx;
    lx[1] += P.G * gy;
    lx[2] += P.G * gz;
    lm += P.G * gm;
    if (q >= 0 && snaps[q] == s) {
      if (live)
        for (int c = 0; c < 3; ++c) lx[c] += gsnap[((size_t)q * nbodies + b) * 3 + c];
      --q;
    }
  }
  if (live) {
    for (int c = 0; c < 3; ++c) {
      out[3 * (size_t)b + c] = lx[c];
      out[3 * (size_t)nbodies + 3 * (size_t)b + c] = lv[c];
    }
    out[6 * (size_t)nbodies + b] = lm;
  }
}

int batch_traj_forward(aster_ctx* c, double dt, int64_t nsteps, int32_t nsnap, const int64_t* snaps,
                       double* snap_pos) {
  BatchState& B = c->batch;
  B.traj_ready = false;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  if (!std::isfinite(dt) || nsteps < 1 || nsnap < 0 || (nsnap > 0 && !snaps))
    return fail(c, ASTER_ERR_INVALID, "need dt finite, nsteps >= 1 and snapshot steps");
  for (int q = 0; q < nsnap; ++q)
    if (snaps[q] < 1 || snaps[q] > nsteps || (q > 0 && snaps[q] <= snaps[q - 1]))
      return fail(c, ASTER_ERR_INVALID, "snapshot steps must be ascending within [1, nsteps]");
  if (c->P.pn_flags & ASTER_PN_TERMS) return fail(c, ASTER_ERR_INVALID, "the adjoint covers Newtonian forces");
  const size_t nb3 = (size_t)3 * B.nbodies;
  if ((double)(nsteps + 1) * nb3 * 8.0 > 64e9) return fail(c, ASTER_ERR_CAPACITY, "trajectory would exceed 64 GB");
  ASTER_HIP(c, B.traj.reserve((size_t)(nsteps + 1) * nb3));
  ASTER_HIP(c, B.x0.reserve(nb3));
  ASTER_HIP(c, B.v0.reserve(nb3));
  ASTER_HIP(c, hipMemcpyAsync(B.x0.p, B.x.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.v0.p, B.v.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  const double h[3] = {dt, 0.0, 0.0};
  const BatchArgs P = make_args(c, nsteps, 1, h, 0);
  const int wpb = B.waves_per_block;
  hipLaunchKernelGGL(k_batch_traj_fwd, dim3((P.nwaves + wpb - 1) / wpb), dim3(64 * wpb), 0, c->stream, P, B.traj.p,
                     B.nbodies);
  ASTER_HIP(c, hipGetLastError());
  B.a_valid = true;
  if (snap_pos)
    for (int q = 0; q < nsnap; ++q)
      ASTER_HIP(c, hipMemcpyAsync(snap_pos + (size_t)q * nb3, B.traj.p + (size_t)snaps[q] * nb3, nb3 * 8,
                                  hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  B.traj_snaps.assign(snaps, snaps + nsnap);
  B.traj_steps = nsteps;
  B.traj_dt = dt;
  B.traj_ready = true;
  return ASTER_OK;
}

int batch_traj_backward(aster_ctx* c, const double* gsnap, double* gx, double* gv, double* gm) {
  BatchState& B = c->batch;
  if (!B.traj_ready) return fail(c, ASTER_ERR_STATE, "call aster_batch_traj_forward first");
  const int nsnap = (int)B.traj_snaps.size();
  if (nsnap > 0 && !gsnap) return fail(c, ASTER_ERR_INVALID, "snapshot gradients required");
  const size_t nb3 = (size_t)3 * B.nbodies;
  ASTER_HIP(c, B.tgrad.reserve(std::max<size_t>(1, nsnap * nb3)));
  ASTER_HIP(c, B.tsnap.reserve(std::max(1, nsnap)));
  ASTER_HIP(c, B.tout.reserve(7 * (size_t)B.nbodies));
  if (nsnap > 0) {
    ASTER_HIP(c, hipMemcpyAsync(B.tgrad.p, gsnap, nsnap * nb3 * 8, hipMemcpyHostToDevice, c->stream));
    ASTER_HIP(c, hipMemcpyAsync(B.tsnap.p, B.traj_snaps.data(), nsnap * 8, hipMemcpyHostToDevice, c->stream));
  }
  const double h0[3] = {0.0, 0.0, 0.0};
  const BatchArgs P = make_args(c, 0, 1, h0, 0);
  const int wpb = B.waves_per_block;
  hipLaunchKernelGGL(k_batch_traj_bwd, dim3((P.nwaves + wpb - 1) / wpb), dim3(64 * wpb), 0, c->stream, P, B.traj.p,
                     B.nbodies, B.traj_steps, B.traj_dt, B.tgrad.p, B.tsnap.p, nsnap, B.tout.p);
  ASTER_HIP(c, hipGetLastError());
  if (gx) ASTER_HIP(c, hipMemcpyAsync(gx, B.tout.p, nb3 * 8, hipMemcpyDeviceToHost, c->stream));
  if (gv) ASTER_HIP(c, hipMemcpyAsync(gv, B.tout.p + nb3, nb3 * 8, hipMemcpyDeviceToHost, c->stream));
  if (gm) ASTER_HIP(c, hipMemcpyAsync(gm, B.tout.p + 2 * nb3, B.nbodies * 8, hipMemcpyDeviceToHost, c->stream));
  // restore the initial state
  ASTER_HIP(c, hipMemcpyAsync(B.x.p, B.x0.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.v.p, B.v0.p, nb3 * 8, hipMemcpyDeviceToDevice, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  B.a_valid = false;
  B.traj_ready = false;
  return ASTER_OK;
}

}  // namespace aster

using namespace aster;

extern "C" {

int aster_batch_set(aster_ctx* c, int64_t nsys, const int64_t* offsets, const double* pos, const double* vel,
                    const double* mass) {
  if (!c) return ASTER_ERR_INVALID;
  if (nsys < 1 || !offsets || !pos || !mass) return fail(c, ASTER_ERR_INVALID, "need nsys >= 1, offsets, pos, mass");
  if (offsets[0] != 0) return fail(c, ASTER_ERR_INVALID, "offsets[0] must be 0");
  for (int64_t s = 0; s < nsys; ++s) {
    const int64_t k = offsets[s + 1] - offsets[s];
    if (k < 1 || k > 64) return fail(c, ASTER_ERR_CAPACITY, "each system needs 1..64 bodies");
  }
  const int64_t nb = offsets[nsys];
  if (nb > (int64_t(1) << 28)) return fail(c, ASTER_ERR_CAPACITY, "at most 2^28 bodies per batch");
  for (int64_t i = 0; i < nb; ++i) {
    if (!(mass[i] > 0.0) || !std::isfinite(mass[i])) return fail(c, ASTER_ERR_INVALID, "masses must be > 0");
    for (int q = 0; q < 3; ++q)
      if (!std::isfinite(pos[3 * i + q]) || (vel && !std::isfinite(vel[3 * i + q])))
        return fail(c, ASTER_ERR_NONFINITE, "non-finite position or velocity in input");
  }
  ASTER_HIP(c, hipSetDevice(c->device));
  BatchState& B = c->batch;
  drop_graph(B);
  B.nsys = nsys;
  B.nbodies = nb;
  B.offsets.assign(offsets, offsets + nsys + 1);
  ASTER_HIP(c, B.x.reserve(3 * nb));
  ASTER_HIP(c, B.v.reserve(3 * nb));
  ASTER_HIP(c, B.a.reserve(3 * nb));
  ASTER_HIP(c, B.m.reserve(nb));
  ASTER_HIP(c, B.doff.reserve(nsys + 1));
  ASTER_HIP(c, hipMemcpyAsync(B.x.p, pos, 3 * nb * 8, hipMemcpyHostToDevice, c->stream));
  if (vel)
    ASTER_HIP(c, hipMemcpyAsync(B.v.p, vel, 3 * nb * 8, hipMemcpyHostToDevice, c->stream));
  else
    ASTER_HIP(c, hipMemsetAsync(B.v.p, 0, 3 * nb * 8, c->stream));
  ASTER_HIP(c, hipMemsetAsync(B.a.p, 0, 3 * nb * 8, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.m.p, mass, nb * 8, hipMemcpyHostToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(B.doff.p, offsets, (nsys + 1) * 8, hipMemcpyHostToDevice, c->stream));
  B.a_valid = false;
  return upload_plan(c);
}

int aster_batch_traj_forward(aster_ctx* c, double dt, int64_t nsteps, int32_t nsnap, const int64_t* snap_steps,
                             double* snap_pos) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return batch_traj_forward(c, dt, nsteps, nsnap, snap_steps, snap_pos);
}

int aster_batch_traj_backward(aster_ctx* c, const double* grad_snap_pos, double* g_pos, double* g_vel,
                              double* g_mass) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return batch_traj_backward(c, grad_snap_pos, g_pos, g_vel, g_mass);
}

int aster_batch_get(aster_ctx* c, double* pos, double* vel) {
  if (!c) return ASTER_ERR_INVALID;
  BatchState& B = c->batch;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  if (pos) ASTER_HIP(c, hipMemcpyAsync(pos, B.x.p, 3 * B.nbodies * 8, hipMemcpyDeviceToHost, c->stream));
  if (vel) ASTER_HIP(c, hipMemcpyAsync(vel, B.v.p, 3 * B.nbodies * 8, hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return ASTER_OK;
}

}  // extern "C"
