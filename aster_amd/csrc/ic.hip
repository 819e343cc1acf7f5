// The context's cluster queue, generated on the device: ClusterQueue
// (src/clusters/clusters_endpoint.cpp:150-207) with the add_disk preset
// (src/clusters/presets_impl.cpp:207-255) and the Plummer/cube generators of
// ic.cpp. One thread draws one body from its own counter-based stream
// (icgen.hpp), so a queue of 10^8 bodies loads at HBM write speed; the
// cluster sums (mass normalisation, barycentre) use a fixed reduction order
// that ic.cpp reproduces on the host.
#include <cmath>

#include "common.hpp"
#include "icgen.hpp"

namespace aster {
namespace {

constexpr int kGenThreads = 256;
constexpr int64_t kMaxBodiesPerContext = int64_t(1) << 28;

// The state planes ([3][n] positions and velocities, [n] masses) and the
// first index of the cluster being generated.
struct Planes {
  double* x;
  double* v;
  double* m;
  int64_t n;
  int64_t base;
};

__device__ __forceinline__ void put(const Planes& P, int64_t i, const double* p, const double* u, double m) {
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    P.x[c * P.n + i] = p[c];
    P.v[c * P.n + i] = u[c];
  }
  P.m[i] = m;
}

__global__ __launch_bounds__(kGenThreads) void k_gen_plummer(int64_t count, uint64_t seed, double b, double rmax,
                                                             double vscale, int mass_mode, double mconst, Planes P) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  double p[3], u[3], m;
  icg::plummer_body(seed, (uint64_t)k, b, rmax, vscale, mass_mode, p, u, &m);
  put(P, P.base + k, p, u, mass_mode == 1 ? m : mconst);
}

__global__ __launch_bounds__(kGenThreads) void k_gen_cube(int64_t count, uint64_t seed, double h, double mconst,
                                                          Planes P) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  double p[3], u[3];
  icg::cube_body(seed, (uint64_t)k, h, p, u);
  put(P, P.base + k, p, u, mconst);
}

__global__ __launch_bounds__(kGenThreads) void k_gen_disk(int64_t count, uint64_t seed, icg::Disk d, Planes P) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  double p[3], u[3], m;
  icg::disk_body(seed, (uint64_t)k, d, p, u, &m);
  put(P, P.base + k, p, u, m);
}

// Per-block sums over the cluster of m (NQ = 1) or of m, m x and m v
// (NQ = 7), in the order icg::ordered_sums evaluates on the host.
template <int NQ>
__global__ __launch_bounds__(icg::kSumThreads) void k_cluster_sums(int64_t count, Planes P, double* part) {
  __shared__ double red[NQ][icg::kSumThreads];
  double s[NQ];
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = 0.0;
  for (int64_t k = (int64_t)blockIdx.x * icg::kSumThreads + threadIdx.x; k < count;
       k += (int64_t)gridDim.x * icg::kSumThreads) {
    const int64_t i = P.base + k;
    const double mi = P.m[i];
    s[0] += mi;
    if (NQ == 7) {
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        s[1 + c] += mi * P.x[c * P.n + i];
        s[4 + c] += mi * P.v[c * P.n + i];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < NQ; ++q) red[q][threadIdx.x] = s[q];
  __syncthreads();
  for (int w = icg::kSumThreads / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
    __syncthreads();
  }
  if ((int)threadIdx.x < NQ) part[NQ * blockIdx.x + threadIdx.x] = red[threadIdx.x][0];
}

// Block totals, padded with zeros to kSumBlocksMax, added pairwise with
// halving strides (one block of kSumBlocksMax threads).
template <int NQ>
__global__ __launch_bounds__(icg::kSumBlocksMax) void k_cluster_sums_final(int nparts, const double* __restrict__ part,
                                                                         double* __restrict__ out) {
  __shared__ double red[NQ][icg::kSumBlocksMax];
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < NQ; ++q) red[q][t] = t < nparts ? part[NQ * t + q] : 0.0;
  __syncthreads();
  for (int w = icg::kSumBlocksMax / 2; w > 0; w >>= 1) {
    if (t < w)
#pragma unroll
      for (int q = 0; q < NQ; ++q) red[q][t] += red[q][t + w];
    __syncthreads();
  }
  if (t < NQ) out[t] = red[t][0];
}

__global__ __launch_bounds__(kGenThreads) void k_scale_mass(int64_t count, Planes P, double mtot,
                                                            const double* __restrict__ S) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  const int64_t i = P.base + k;
  P.m[i] = P.m[i] * (mtot / S[0]);
}

// Barycentre and mean velocity to zero (S = sums of m, m x, m v), then the
// cluster's translation and bulk velocity when it has them.
__global__ __launch_bounds__(kGenThreads) void k_recentre(int64_t count, Planes P, const double* __restrict__ S,
                                                          int translate, double ox, double oy, double oz, double ux,
                                                          double uy, double uz) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  const int64_t i = P.base + k;
  const double o[3] = {ox, oy, oz}, u[3] = {ux, uy, uz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    double x = P.x[c * P.n + i] - S[1 + c] / S[0];
    double v = P.v[c * P.n + i] - S[4 + c] / S[0];
    if (translate) {
      x = x + o[c];
      v = v + u[c];
    }
    P.x[c * P.n + i] = x;
    P.v[c * P.n + i] = v;
  }
}

__global__ __launch_bounds__(kGenThreads) void k_translate(int64_t count, Planes P, double ox, double oy, double oz,
                                                           double ux, double uy, double uz) {
  const int64_t k = (int64_t)blockIdx.x * kGenThreads + threadIdx.x;
  if (k >= count) return;
  const int64_t i = P.base + k;
  const double o[3] = {ox, oy, oz}, u[3] = {ux, uy, uz};
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    P.x[c * P.n + i] = P.x[c * P.n + i] + o[c];
    P.v[c * P.n + i] = P.v[c * P.n + i] + u[c];
  }
}

// One body written by one lane (add_disk's central body).
__global__ void k_put_body(Planes P, int64_t i, double px, double py, double pz, double vx, double vy, double vz,
                           double m) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const double p[3] = {px, py, pz}, u[3] = {vx, vy, vz};
  put(P, i, p, u, m);
}

bool finite3(const double* a) { return std::isfinite(a[0]) && std::isfinite(a[1]) && std::isfinite(a[2]); }

bool has_offset(const aster_cluster& k) {
  for (int c = 0; c < 3; ++c)
    if (k.position[c] != 0.0 || k.velocity[c] != 0.0) return true;
  return false;
}

int validate(aster_ctx* c, const aster_cluster& k) {
  if (k.number < 1) return fail(c, ASTER_ERR_INVALID, "a cluster needs at least one body");
  if (!(k.mass > 0.0) || !std::isfinite(k.mass)) return fail(c, ASTER_ERR_INVALID, "cluster mass must be finite and > 0");
  if (!finite3(k.position) || !finite3(k.velocity) || !finite3(k.rotation))
    return fail(c, ASTER_ERR_INVALID, "cluster position, velocity and rotation must be finite");
  const bool rotated = k.rotation[0] != 0.0 || k.rotation[1] != 0.0 || k.rotation[2] != 0.0;
  switch (k.kind) {
    case ASTER_CLUSTER_PLUMMER:
      if (!(k.size[0] > 0.0) || !(k.size[1] > 0.0) || !std::isfinite(k.size[0]) || !std::isfinite(k.size[1]))
        return fail(c, ASTER_ERR_INVALID, "Plummer cluster needs b > 0 and rmax > 0 (size[0], size[1])");
      if (k.mass_mode != 0 && k.mass_mode != 1) return fail(c, ASTER_ERR_INVALID, "mass_mode must be 0 or 1");
      if (rotated) return fail(c, ASTER_ERR_INVALID, "rotation applies to disk clusters only");
      return ASTER_OK;
    case ASTER_CLUSTER_CUBE:
      if (!(k.size[0] > 0.0) || !std::isfinite(k.size[0]))
        return fail(c, ASTER_ERR_INVALID, "cube cluster needs a half-extent > 0 (size[0])");
      if (rotated) return fail(c, ASTER_ERR_INVALID, "rotation applies to disk clusters only");
      return ASTER_OK;
    case ASTER_CLUSTER_DISK:
      // add_disk's own checks (presets_impl.cpp:211-221), and
      // rng_point_in_cylinder's radius check (:131), which the reference
      // only warns about before dividing by a zero radius
      if (!(k.size[0] > 0.0) || !std::isfinite(k.size[0]))
        return fail(c, ASTER_ERR_INVALID, "disk cannot have a non-positive radius");
      if (!(k.size[1] > 0.0) || !std::isfinite(k.size[1]))
        return fail(c, ASTER_ERR_INVALID, "disk cannot have a non-positive thickness");
      if (!(k.size[2] > 0.0) || k.size[2] > k.size[0])
        return fail(c, ASTER_ERR_INVALID, "disk inner radius must be in (0, radius]");
      if (!(k.center_mass > 0.0) || !std::isfinite(k.center_mass))
        return fail(c, ASTER_ERR_INVALID, "disk central mass must be finite and > 0");
      if (!(k.scale > 0.0) || !std::isfinite(k.scale)) return fail(c, ASTER_ERR_INVALID, "disk scale must be > 0");
      return ASTER_OK;
    default:
      return fail(c, ASTER_ERR_INVALID, "unknown cluster kind");
  }
}

// add_disk's geometry with the angles of rotate_point (phi = rotation.x,
// theta = rotation.z, both in degrees, times PI/180 as presets_impl.cpp:51-52).
icg::Disk disk_params(const aster_cluster& k, double G) {
  icg::Disk d;
  d.radius = k.size[0];
  d.thickness = k.size[1];
  d.inner = k.size[2];
  d.scale = k.scale;
  d.gm = G * k.center_mass;
  d.avr_mass = k.mass;
  for (int c = 0; c < 3; ++c) {
    d.center[c] = k.position[c];
    d.vel[c] = k.velocity[c];
  }
  const double deg = icm::kPi / 180.0;
  icm::sincos(k.rotation[0] * deg, &d.sphi, &d.cphi);
  icm::sincos(k.rotation[2] * deg, &d.sth, &d.cth);
  return d;
}

unsigned blocks_for(int64_t n) { return (unsigned)((n + kGenThreads - 1) / kGenThreads); }

// Sums of the cluster into out (device, nq doubles), through ctx scratch.
int cluster_sums(aster_ctx* c, const Planes& P, int64_t count, int nq, double** out) {
  ASTER_HIP(c, c->ework.reserve((size_t)8 * icg::kSumBlocksMax + 8));
  double* part = c->ework.p;
  double* res = c->ework.p + 8 * icg::kSumBlocksMax;
  const int B = icg::sum_blocks(count);
  if (nq == 1) {
    hipLaunchKernelGGL(k_cluster_sums<1>, dim3(B), dim3(icg::kSumThreads), 0, c->stream, count, P, part);
    hipLaunchKernelGGL(k_cluster_sums_final<1>, dim3(1), dim3(icg::kSumBlocksMax), 0, c->stream, B,
                       (const double*)part, res);
  } else {
    hipLaunchKernelGGL(k_cluster_sums<7>, dim3(B), dim3(icg::kSumThreads), 0, c->stream, count, P, part);
    hipLaunchKernelGGL(k_cluster_sums_final<7>, dim3(1), dim3(icg::kSumBlocksMax), 0, c->stream, B,
                       (const double*)part, res);
  }
  ASTER_HIP(c, hipGetLastError());
  *out = res;
  return ASTER_OK;
}

int generate(aster_ctx* c, const aster_cluster& k, int64_t base) {
  const Planes P{c->x.p, c->v.p, c->m.p, c->n, base};
  const int64_t cnt = k.number;
  const unsigned g = blocks_for(cnt);
  const double* o = k.position;
  const double* u = k.velocity;
  switch (k.kind) {
    case ASTER_CLUSTER_PLUMMER: {
      const double b = k.size[0], mtot = k.mass;
      const double vscale = std::sqrt(c->P.G * mtot / b);
      hipLaunchKernelGGL(k_gen_plummer, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, k.seed, b, k.size[1], vscale,
                         k.mass_mode, mtot / (double)cnt, P);
      ASTER_HIP(c, hipGetLastError());
      double* S = nullptr;
      int rc;
      if (k.mass_mode == 1) {
        if ((rc = cluster_sums(c, P, cnt, 1, &S))) return rc;
        hipLaunchKernelGGL(k_scale_mass, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, P, mtot, (const double*)S);
      }
      if ((rc = cluster_sums(c, P, cnt, 7, &S))) return rc;
      hipLaunchKernelGGL(k_recentre, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, P, (const double*)S,
                         has_offset(k) ? 1 : 0, o[0], o[1], o[2], u[0], u[1], u[2]);
      break;
    }
    case ASTER_CLUSTER_CUBE:
      hipLaunchKernelGGL(k_gen_cube, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, k.seed, k.size[0],
                         k.mass / (double)cnt, P);
      if (has_offset(k))
        hipLaunchKernelGGL(k_translate, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, P, o[0], o[1], o[2], u[0], u[1],
                           u[2]);
      break;
    case ASTER_CLUSTER_DISK:
      hipLaunchKernelGGL(k_gen_disk, dim3(g), dim3(kGenThreads), 0, c->stream, cnt, k.seed, disk_params(k, c->P.G),
                         P);
      break;
    default:
      return fail(c, ASTER_ERR_INVALID, "unknown cluster kind");
  }
  ASTER_HIP(c, hipGetLastError());
  return ASTER_OK;
}

int64_t queued_bodies(const aster_ctx* c) {
  int64_t s = 0;
  for (const aster_cluster& k : c->clusters) s += k.number;
  return s;
}

}  // namespace
}  // namespace aster

using namespace aster;

extern "C" {

int aster_cluster_add(aster_ctx* c, const aster_cluster* k) {
  if (!c) return ASTER_ERR_INVALID;
  if (!k) return fail(c, ASTER_ERR_INVALID, "cluster is NULL");
  int rc = validate(c, *k);
  if (rc) return rc;
  const int64_t have = c->have_state ? c->n : 0;
  const int64_t extra = k->kind == ASTER_CLUSTER_DISK ? 1 : 0;
  if (k->number > kMaxBodiesPerContext || have + queued_bodies(c) + k->number + extra > kMaxBodiesPerContext)
    return fail(c, ASTER_ERR_CAPACITY, "at most 2^28 bodies per context");
  ASTER_HIP(c, hipSetDevice(c->device));
  if (k->kind == ASTER_CLUSTER_DISK) {
    // add_disk places the central body at once (presets_impl.cpp:227-231)
    const icg::Disk d = disk_params(*k, c->P.G);
    double p[3], v[3];
    icg::rotate_point(k->position, d, p);
    icg::rotate_point(k->velocity, d, v);
    int64_t base = 0;
    if ((rc = grow_state(c, 1, &base))) return rc;
    const Planes P{c->x.p, c->v.p, c->m.p, c->n, base};
    hipLaunchKernelGGL(k_put_body, dim3(1), dim3(64), 0, c->stream, P, base, p[0], p[1], p[2], v[0], v[1], v[2],
                       k->center_mass);
    ASTER_HIP(c, hipGetLastError());
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->clusters.push_back(*k);
  return ASTER_OK;
}

int aster_cluster_pop(aster_ctx* c) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->clusters.empty()) c->clusters.pop_back();
  return ASTER_OK;
}

int aster_cluster_queued(const aster_ctx* c, int64_t* nclusters, int64_t* nbodies) {
  if (!c) return ASTER_ERR_INVALID;
  if (nclusters) *nclusters = (int64_t)c->clusters.size();
  if (nbodies) *nbodies = queued_bodies(c);
  return ASTER_OK;
}

int aster_cluster_load(aster_ctx* c) {
  if (!c) return ASTER_ERR_INVALID;
  if (c->clusters.empty()) return ASTER_OK;  // the reference warns and loads nothing
  ASTER_HIP(c, hipSetDevice(c->device));
  int64_t base = 0;
  int rc = grow_state(c, queued_bodies(c), &base);
  if (rc) return rc;
  // last added first, as ClusterQueue::load takes data.back() (:189-205)
  for (size_t q = c->clusters.size(); q-- > 0;) {
    if ((rc = generate(c, c->clusters[q], base))) return rc;
    base += c->clusters[q].number;
  }
  c->clusters.clear();
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return ASTER_OK;
}

}  // extern "C"
