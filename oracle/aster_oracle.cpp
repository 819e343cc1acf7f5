// aster_oracle.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++ FP64 restatement of Aster's CPU force/integration path. It is the
// checker for the HIP library, nothing more: only tests/, the smoke() entry of
// __graft_entry__.py and the cpu_baseline leg of bench.py may load it. The
// product library never links or calls it.
//
// Every function names the reference file:line it follows (paths relative to
// the Aster repository root). Expression order is kept exactly as in the
// reference because the FP64 GPU paths are checked against it bit for bit;
// build with -ffp-contract=off and without -ffast-math (the reference library
// flags, CMakeLists.txt:58-65).
//
// Deviations from the reference, each deliberate and documented:
//  * TBB parallel_for is replaced by std::thread over static chunks of bodies;
//    each body's sum is sequential in j, as in the reference, so the result does
//    not depend on the thread count.
//  * tbb::parallel_sort (barnes_hut.cpp:324) compares codes only, leaving ties
//    in an unspecified order; the oracle sorts by (code, body index), the one
//    stable order, which equals the reference wherever codes are unique.
//  * The traversal stack (barnes_hut.cpp:204, 64 entries, unchecked) is 512
//    entries here and overflow is reported instead of corrupting memory.
//  * get_total_energy (collectors.cpp:33-50) reduces with tbb::parallel_reduce
//    in an unspecified order; the oracle sums serially in i order.
//  * Yoshida-4 has no counterpart in the reference; it is the composition of
//    three reference leapfrog steps (tool_chain_impl.cpp:169-185) with the
//    Yoshida (1990) weights.
//  * update_SABA6 (SABA.cpp:176-233) kicks with d2 in its fifth stage
//    (SABA.cpp:216) where the symmetric sequence needs d3, so its kick weights
//    sum to 0.946 instead of 1; the oracle uses d3 there. The other nine
//    sequences are followed as written.
//
// Parity status of each function is stated in its comment: "pinned by" names
// the independent checks in tests/ that would catch a plausible mistake.

#include "../aster_amd/csrc/saba_coeffs.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <thread>
#include <vector>

namespace {

constexpr int kNewton = 0x1;   // basic.h:39
constexpr int kPN1 = 0x10;     // basic.h:40
constexpr int kPN2 = 0x100;    // basic.h:41
constexpr int kPN25 = 0x1000;  // basic.h:42
constexpr int kEIH = 0x20;     // full N-body 1PN; no counterpart in the reference

struct V3 {
  double x, y, z;
};

inline V3 load3(const double* a, int64_t i) { return {a[3 * i], a[3 * i + 1], a[3 * i + 2]}; }
inline void store3(double* a, int64_t i, V3 v) {
  a[3 * i] = v.x;
  a[3 * i + 1] = v.y;
  a[3 * i + 2] = v.z;
}

template <class F>
void parallel_for(int64_t n, int nthreads, F&& f) {
  if (nthreads <= 0) nthreads = (int)std::max(1u, std::thread::hardware_concurrency());
  if (nthreads == 1 || n < 64) {
    for (int64_t i = 0; i < n; ++i) f(i);
    return;
  }
  std::vector<std::thread> th;
  int64_t chunk = (n + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    int64_t a = t * chunk, b = std::min(n, a + chunk);
    if (a >= b) break;
    th.emplace_back([a, b, &f] {
      for (int64_t i = a; i < b; ++i) f(i);
    });
  }
  for (auto& t : th) t.join();
}

struct Law {
  double G, c, soft;
  int pn;
};

// get_pn1, tool_chain_impl.cpp:110-118. `norm` is the caller's -norm (:133).
// vec3*vec3 is a dot product (physics_endpoint.cpp:187-189); int*vec3 goes
// through operator*(REAL, vec3) (physics_endpoint.cpp:216-218).
inline V3 pn1_term(double m1, double m2, V3 v1, V3 v2, double inv_r, V3 n, double G) {
  const double v2_sq = v2.x * v2.x + v2.y * v2.y + v2.z * v2.z;
  const double v1_sq = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
  const double dot4 = (4.0 * v1.x) * v2.x + (4.0 * v1.y) * v2.y + (4.0 * v1.z) * v2.z;
  const double nv2 = n.x * v2.x + n.y * v2.y + n.z * v2.z;
  double s = -v1_sq - 2.0 * v2_sq;
  s = s + dot4;
  s = s + 1.5 * (nv2 * nv2);
  s = s + 5.0 * G * m1 * inv_r;
  s = s + 4.0 * G * m2 * inv_r;
  V3 a{n.x * s, n.y * s, n.z * s};
  const double n4v1 = (n.x * 4.0) * v1.x + (n.y * 4.0) * v1.y + (n.z * 4.0) * v1.z;
  const double n3v2 = (n.x * 3.0) * v2.x + (n.y * 3.0) * v2.y + (n.z * 3.0) * v2.z;
  const double t = n4v1 - n3v2;
  a.x = a.x + (v1.x - v2.x) * t;
  a.y = a.y + (v1.y - v2.y) * t;
  a.z = a.z + (v1.z - v2.z) * t;
  a.x = a.x * G * m2 * inv_r * inv_r;
  a.y = a.y * G * m2 * inv_r * inv_r;
  a.z = a.z * G * m2 * inv_r * inv_r;
  return a;
}

// get_pn2, tool_chain_impl.cpp:87-107: the two-body 2PN acceleration of body
// 1 (Blanchet, Living Rev. Rel. 17 (2014) 2, eq. 203, 2PN part). `n` is the
// caller's -norm. Every product and sum left to right, fraction constants
// folded as written.
inline V3 pn2_term(double m1, double m2, V3 v1, V3 v2, double inv_r, V3 n, double G) {
  const double v2_sq = v2.x * v2.x + v2.y * v2.y + v2.z * v2.z;
  const double v1_sq = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
  const double v_prod = v1.x * v2.x + v1.y * v2.y + v1.z * v2.z;
  const double nv1 = n.x * v1.x + n.y * v1.y + n.z * v1.z;
  const double nv2 = n.x * v2.x + n.y * v2.y + n.z * v2.z;
  const double nv22 = nv2 * nv2;
  const double pn_1 = G * m1 * inv_r;
  const double pn_2 = G * m2 * inv_r;
  const double nv12 = nv1 * nv1;
  const double A = -2 * v2_sq * v2_sq + 4 * v2_sq * v_prod - 2 * v_prod * v_prod + 3.0 / 2.0 * v1_sq * nv22 +
                   9.0 / 2.0 * v2_sq * nv22 - 6 * v_prod * nv22 - 15.0 / 8.0 * nv22 * nv22 +
                   pn_1 * (-15.0 / 4.0 * v1_sq + 5.0 / 4.0 * v2_sq - 5.0 / 2.0 * v_prod + 39.0 / 2.0 * nv12 -
                           39.0 * nv1 * nv2 + 17.0 / 2.0 * nv22) +
                   pn_2 * (4 * v2_sq - 8 * v_prod + 2 * nv12 - 4 * nv1 * nv2 - 6 * nv22);
  const double B = v1_sq * nv2 + 4 * v2_sq * nv1 - 5 * v2_sq * nv2 - 4 * v_prod * nv1 + 4 * v_prod * nv2 -
                   6 * nv1 * nv22 + 9.0 / 2.0 * nv22 * nv2 + pn_1 * (-63.0 / 4.0 * nv1 + 55.0 / 4.0 * nv2) +
                   pn_2 * (-2 * nv1 - 2 * nv2);
  const double C = -57.0 / 4.0 * m1 * m1 - 9 * m2 * m2 - 69.0 / 2.0 * m1 * m2;
  const double s1 = G * m2 * inv_r * inv_r;
  const double s2 = G * G * G * m2 * inv_r * inv_r * inv_r * inv_r;
  // s1 * (n A + (v1 - v2) B) + (s2 n) C, scalar*vec3 being vec3*scalar
  V3 u{n.x * A + (v1.x - v2.x) * B, n.y * A + (v1.y - v2.y) * B, n.z * A + (v1.z - v2.z) * B};
  V3 w{n.x * s2, n.y * s2, n.z * s2};
  return V3{u.x * s1 + w.x * C, u.y * s1 + w.y * C, u.z * s1 + w.z * C};
}

// get_pn25, tool_chain_impl.cpp:73-84: the two-body 2.5PN radiation-reaction
// acceleration of body 1 (Blanchet 2014, eq. 203, 2.5PN part). `n` is the
// caller's -norm.
inline V3 pn25_term(double m1, double m2, V3 v1, V3 v2, double inv_r, V3 n, double G) {
  const double pn_1 = G * m1 * inv_r;
  const double pn_2 = G * m2 * inv_r;
  const V3 dv{v1.x - v2.x, v1.y - v2.y, v1.z - v2.z};
  const double sq = dv.x * dv.x + dv.y * dv.y + dv.z * dv.z;
  const double k1 = -sq + 2 * pn_1 - 8 * pn_2;
  const double ndv = (n.x * v1.x + n.y * v1.y + n.z * v1.z) - (n.x * v2.x + n.y * v2.y + n.z * v2.z);
  const double k2 = 3 * sq - 6 * pn_1 + 52.0 / 3.0 * pn_2;
  V3 in{dv.x * k1 + (n.x * ndv) * k2, dv.y * k1 + (n.y * ndv) * k2, dv.z * k1 + (n.z * ndv) * k2};
  auto scale = [&](double c) { return c * 4.0 / 5.0 * G * G * m1 * m2 * inv_r * inv_r * inv_r; };
  return V3{scale(in.x), scale(in.y), scale(in.z)};
}

// compute_gravitational, tool_chain_impl.cpp:120-137: force on body 1 from
// body 2 (G m1 m2 / r^2 along d = p2 - p1, Plummer-softened with
// softening = eps^2), plus the pairwise 1PN, 2PN and 2.5PN terms the flags
// select, each times 1/c^k and m1.
// The post-Newtonian terms of compute_gravitational (tool_chain_impl.cpp:126-136)
// added to f in the reference's order; mn is the caller's -norm.
inline void add_pn(V3& f, double m1, double m2, V3 v1, V3 v2, double inv_r, V3 mn, const Law& L) {
  const double inv_c = 1.0 / L.c;
  if (L.pn & kPN1) {
    double ic2 = inv_c * inv_c;
    V3 p = pn1_term(m1, m2, v1, v2, inv_r, mn, L.G);
    V3 a1{p.x * ic2, p.y * ic2, p.z * ic2};
    f.x = f.x + a1.x * m1;
    f.y = f.y + a1.y * m1;
    f.z = f.z + a1.z * m1;
  }
  if (L.pn & kPN2) {
    double ic4 = inv_c * inv_c * inv_c * inv_c;
    V3 p = pn2_term(m1, m2, v1, v2, inv_r, mn, L.G);
    V3 a2{p.x * ic4, p.y * ic4, p.z * ic4};
    f.x = f.x + a2.x * m1;
    f.y = f.y + a2.y * m1;
    f.z = f.z + a2.z * m1;
  }
  if (L.pn & kPN25) {
    double ic5 = inv_c * inv_c * inv_c * inv_c * inv_c;
    V3 p = pn25_term(m1, m2, v1, v2, inv_r, mn, L.G);
    V3 a25{p.x * ic5, p.y * ic5, p.z * ic5};
    f.x = f.x + a25.x * m1;
    f.y = f.y + a25.y * m1;
    f.z = f.z + a25.z * m1;
  }
}

inline V3 gravitational(double m1, double m2, V3 v1, V3 v2, V3 p1, V3 p2, const Law& L) {
  V3 d{p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
  double r2 = d.x * d.x + d.y * d.y + d.z * d.z + L.soft;
  double inv_r = 1.0 / std::sqrt(r2);
  double inv_r2 = inv_r * inv_r;
  V3 norm{d.x * inv_r, d.y * inv_r, d.z * inv_r};
  double scal = L.G * m2 * m1 * inv_r2;
  V3 f{norm.x * scal, norm.y * scal, norm.z * scal};
  add_pn(f, m1, m2, v1, v2, inv_r, V3{-norm.x, -norm.y, -norm.z}, L);
  return f;
}

// compute_gravitational's post-Newtonian part alone (its terms added to zero).
inline V3 pn_part(double m1, double m2, V3 v1, V3 v2, V3 p1, V3 p2, const Law& L) {
  V3 d{p2.x - p1.x, p2.y - p1.y, p2.z - p1.z};
  double r2 = d.x * d.x + d.y * d.y + d.z * d.z + L.soft;
  double inv_r = 1.0 / std::sqrt(r2);
  V3 norm{d.x * inv_r, d.y * inv_r, d.z * inv_r};
  V3 f{0.0, 0.0, 0.0};
  add_pn(f, m1, m2, v1, v2, inv_r, V3{-norm.x, -norm.y, -norm.z}, L);
  return f;
}

// Parallelized::compute_forces, builder_endpoint.cpp:729-757: every target i in
// [0,n) sums F(i,j)/m_i over sources j in [lower, upper), j != i, in j order,
// then assigns accs[i].
void eih(int64_t n, const double* pos, const double* vel, const double* mass, const Law& L, double* acc,
         int nthreads);

void allpairs(int64_t n, const double* pos, const double* vel, const double* mass, const Law& L,
              int64_t lower, int64_t upper, double* acc, int nthreads) {
  if (L.pn & kEIH) {  // the full source range (orc_allpairs rejects bounds)
    eih(n, pos, vel, mass, L, acc, nthreads);
    return;
  }
  if (upper < 0) upper = n;
  parallel_for(n, nthreads, [&](int64_t i) {
    V3 a{0.0, 0.0, 0.0};
    const double m1 = mass[i];
    const V3 v1 = load3(vel, i), p1 = load3(pos, i);
    for (int64_t j = lower; j < upper; ++j) {
      if (i == j) continue;
      V3 f = gravitational(m1, mass[j], v1, load3(vel, j), p1, load3(pos, j), L);
      a.x += f.x / m1;
      a.y += f.y / m1;
      a.z += f.z / m1;
    }
    store3(acc, i, a);
  });
}

// Einstein-Infeld-Hoffmann 1PN accelerations (no counterpart in the
// reference, whose get_pn1 is the two-body law summed over pairs), in the
// form of Newhall, Standish & Williams (1983, A&A 125, 150, eq. 1) with
// beta = gamma = 1, mu = G m, r from the softened r^2 = d.d + softening:
//   a_i = sum_{j != i} mu_j d_ij / r^3 [1 - (4 phi_i + phi_j)/c^2
//           + (v_i^2 + 2 v_j^2 - 4 v_i.v_j)/c^2 - 3/2 (d_ij.v_j / r)^2/c^2
//           + d_ij.aN_j / (2 c^2)]
//       + sum_{j != i} mu_j / (c^2 r^3) [-d_ij.(4 v_i - 3 v_j)] (v_i - v_j)
//       + 7/(2 c^2) sum_{j != i} mu_j aN_j / r
// with d_ij = x_j - x_i, phi_i = sum_{k != i} mu_k / r_ik and aN the Newtonian
// accelerations; sources in index order. Pinned by: for two bodies it equals
// the pairwise 1PN law (Blanchet); an independent numpy evaluation of the
// same published equation; conservation of the EIH (Landau-Lifshitz sec. 106)
// 1PN energy along trajectories to O(1/c^4).
void eih(int64_t n, const double* pos, const double* vel, const double* mass, const Law& L, double* acc,
         int nthreads) {
  std::vector<double> phi(n), an(3 * n);
  parallel_for(n, nthreads, [&](int64_t i) {
    const V3 p1 = load3(pos, i);
    double ph = 0.0, ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const V3 p2 = load3(pos, j);
      const double dx = p2.x - p1.x, dy = p2.y - p1.y, dz = p2.z - p1.z;
      const double r2 = dx * dx + dy * dy + dz * dz + L.soft;
      const double inv_r = 1.0 / std::sqrt(r2);
      const double inv_r3 = inv_r * inv_r * inv_r;
      const double mu = L.G * mass[j];
      ph += mu * inv_r;
      ax += mu * dx * inv_r3;
      ay += mu * dy * inv_r3;
      az += mu * dz * inv_r3;
    }
    phi[i] = ph;
    an[3 * i] = ax;
    an[3 * i + 1] = ay;
    an[3 * i + 2] = az;
  });
  const double ic = 1.0 / L.c;
  const double ic2 = ic * ic;
  parallel_for(n, nthreads, [&](int64_t i) {
    const V3 p1 = load3(pos, i), v1 = load3(vel, i);
    const double vi2 = v1.x * v1.x + v1.y * v1.y + v1.z * v1.z;
    double ax = 0.0, ay = 0.0, az = 0.0;
    for (int64_t j = 0; j < n; ++j) {
      if (j == i) continue;
      const V3 p2 = load3(pos, j), v2 = load3(vel, j);
      const double dx = p2.x - p1.x, dy = p2.y - p1.y, dz = p2.z - p1.z;
      const double r2 = dx * dx + dy * dy + dz * dz + L.soft;
      const double inv_r = 1.0 / std::sqrt(r2);
      const double inv_r3 = inv_r * inv_r * inv_r;
      const double mu = L.G * mass[j];
      const double anx = an[3 * j], any = an[3 * j + 1], anz = an[3 * j + 2];
      const double vj2 = v2.x * v2.x + v2.y * v2.y + v2.z * v2.z;
      const double vij = v1.x * v2.x + v1.y * v2.y + v1.z * v2.z;
      const double rv = (dx * v2.x + dy * v2.y + dz * v2.z) * inv_r;
      const double dan = dx * anx + dy * any + dz * anz;
      const double B = 1.0 - (4.0 * phi[i] + phi[j]) * ic2 + (vi2 + 2.0 * vj2 - 4.0 * vij) * ic2 -
                       1.5 * (rv * rv) * ic2 + 0.5 * dan * ic2;
      const double s1 = mu * inv_r3 * B;
      const double w = -(dx * (4.0 * v1.x - 3.0 * v2.x) + dy * (4.0 * v1.y - 3.0 * v2.y) + dz * (4.0 * v1.z - 3.0 * v2.z));
      const double s2 = mu * inv_r3 * ic2 * w;
      const double s3 = 3.5 * mu * inv_r * ic2;
      ax += dx * s1;
      ax += (v1.x - v2.x) * s2;
      ax += anx * s3;
      ay += dy * s1;
      ay += (v1.y - v2.y) * s2;
      ay += any * s3;
      az += dz * s1;
      az += (v1.z - v2.z) * s2;
      az += anz * s3;
    }
    store3(acc, i, V3{ax, ay, az});
  });
}

// part1by2 + get_morton, bh_specialization.cpp:30-46 (PRECISION_BITS = 10,
// barnes-hut.h:16). `box` is Barnes_Hut::bounding_box, the half-extent.
inline uint32_t part1by2(uint32_t n) {
  n = (n | (n << 16)) & 0x030000FFu;
  n = (n | (n << 8)) & 0x0300F00Fu;
  n = (n | (n << 4)) & 0x030C30C3u;
  n = (n | (n << 2)) & 0x09249249u;
  return n;
}

inline uint32_t morton(V3 p, const double* box) {
  double px = (p.x + box[0]) / (2 * box[0]);
  double py = (p.y + box[1]) / (2 * box[1]);
  double pz = (p.z + box[2]) / (2 * box[2]);
  uint32_t ix = std::min((uint32_t)(px * (1 << 10)), (1u << 10) - 1);
  uint32_t iy = std::min((uint32_t)(py * (1 << 10)), (1u << 10) - 1);
  uint32_t iz = std::min((uint32_t)(pz * (1 << 10)), (1u << 10) - 1);
  return (part1by2(iz) << 2) | (part1by2(iy) << 1) | part1by2(ix);
}

// Karras (2012) split as written in Barnes_Hut::make_tree,
// barnes_hut.cpp:338-396: internal node of split i is stored at n + i; leaf k
// is sorted position k; equal codes tie-break on 32 + clz(i ^ j) of sorted
// positions (:340-342).
void karras(int64_t n, const uint32_t* codes, int32_t* left, int32_t* right, int nthreads) {
  const size_t num_leaves = (size_t)n;
  for (int64_t k = 0; k < 2 * n - 1; ++k) left[k] = right[k] = -1;
  auto delta = [&](size_t i, size_t j) -> int {
    if (j >= num_leaves) return -1;
    if (codes[i] == codes[j]) return 32 + __builtin_clz((unsigned)(i ^ j));
    return __builtin_clz(codes[i] ^ codes[j]);
  };
  parallel_for(n - 1, nthreads, [&](int64_t ii) {
    size_t i = (size_t)ii;
    int d_left = (i == 0) ? -1 : delta(i, i - 1);
    int d_right = delta(i, i + 1);
    int d = (d_right > d_left) ? 1 : -1;
    int d_min = (d == 1) ? d_left : d_right;
    size_t l_max = 2;
    while ((int)(i + l_max * d) >= 0 && i + l_max * d < num_leaves && delta(i, i + l_max * d) > d_min)
      l_max *= 2;
    size_t l = 0;
    for (size_t t = l_max / 2; t >= 1; t /= 2) {
      size_t test_pos = i + (l + t) * d;
      if ((int)test_pos >= 0 && test_pos < num_leaves && delta(i, test_pos) > d_min) l += t;
    }
    size_t j = i + l * d;
    int d_node = delta(i, j);
    size_t s = 0;
    size_t range_size = (j > i) ? j - i : i - j;
    for (size_t t = (range_size + 1) / 2; t >= 1; t = (t + 1) / 2) {
      size_t test_pos = i + (s + t) * d;
      if ((int)test_pos >= 0 && test_pos < num_leaves && delta(i, test_pos) > d_node) s += t;
      if (t == 1) break;
    }
    size_t gamma = i + s * d + std::min(d, 0);
    size_t node = num_leaves + i;
    size_t lo = std::min(i, j), hi = std::max(i, j);
    left[node] = (int32_t)(lo == gamma ? gamma : num_leaves + gamma);
    right[node] = (int32_t)(hi == gamma + 1 ? gamma + 1 : num_leaves + gamma + 1);
  });
}

struct Tree {
  std::vector<int32_t> left, right;
  std::vector<double> mass, com, vel;  // com/vel: 3 per node
};

// NodesArray::unite, barnes_hut.cpp:137-152, recursive from `start`; written as
// an explicit post-order walk so deep trees cannot overflow the C stack. The
// zero-mass guard (:146) leaves the node at its first-build value (zero).
void unite(Tree& T, int64_t start) {
  const int64_t size = (int64_t)T.mass.size();
  std::vector<std::pair<int64_t, int>> st;
  st.push_back({start, 0});
  while (!st.empty()) {
    auto& top = st.back();
    int64_t k = top.first;
    int32_t L = T.left[k], R = T.right[k];
    if (L == -1 || R == -1 || L < 0 || L >= size || R < 0 || R >= size) {
      st.pop_back();
      continue;
    }
    if (top.second == 0) {
      top.second = 1;
      st.push_back({L, 0});
      continue;
    }
    if (top.second == 1) {
      top.second = 2;
      st.push_back({R, 0});
      continue;
    }
    st.pop_back();
    double total = T.mass[R] + T.mass[L];
    if (total <= 0) continue;
    T.mass[k] = total;
    for (int c = 0; c < 3; ++c) {
      T.com[3 * k + c] = (T.com[3 * R + c] * T.mass[R] + T.com[3 * L + c] * T.mass[L]) / T.mass[k];
      T.vel[3 * k + c] = (T.vel[3 * R + c] * T.mass[R] + T.vel[3 * L + c] * T.mass[L]) / T.mass[k];
    }
  }
}

struct BuiltTree {
  std::vector<uint32_t> codes;   // per body
  std::vector<uint32_t> perm;    // sorted position -> body
  std::vector<uint32_t> sorted;  // codes in sorted order
  Tree T;
};

// Barnes_Hut::make_tree, barnes_hut.cpp:309-400 (leaf init :326-336).
void make_tree(int64_t n, const double* pos, const double* vel, const double* mass, const double* box,
               BuiltTree& B, int nthreads) {
  B.codes.resize(n);
  parallel_for(n, nthreads, [&](int64_t i) { B.codes[i] = morton(load3(pos, i), box); });
  B.perm.resize(n);
  std::iota(B.perm.begin(), B.perm.end(), 0u);
  std::stable_sort(B.perm.begin(), B.perm.end(),
                   [&](uint32_t a, uint32_t b) { return B.codes[a] < B.codes[b]; });
  B.sorted.resize(n);
  for (int64_t k = 0; k < n; ++k) B.sorted[k] = B.codes[B.perm[k]];
  const int64_t nn = 2 * n - 1;
  B.T.left.assign(nn, -1);
  B.T.right.assign(nn, -1);
  B.T.mass.assign(nn, 0.0);
  B.T.com.assign(3 * nn, 0.0);
  B.T.vel.assign(3 * nn, 0.0);
  for (int64_t k = 0; k < n; ++k) {
    uint32_t b = B.perm[k];
    B.T.mass[k] = mass[b];
    for (int c = 0; c < 3; ++c) {
      B.T.com[3 * k + c] = pos[3 * b + c];
      B.T.vel[3 * k + c] = vel[3 * b + c];
    }
  }
  if (n > 1) {
    karras(n, B.sorted.data(), B.T.left.data(), B.T.right.data(), nthreads);
    unite(B.T, n);
  }
}

// Barnes_Hut::get_node_body, barnes_hut.cpp:202-261: right-first DFS from the
// root with size bd_size halved per level; skip d2 == 0; accept on
// s^2 < theta^2 d^2 or leaf; accumulate F/m_i in visit order.
int traverse(const Tree& T, int64_t root, int64_t i, double bd_size, const double* pos, const double* vel,
             const double* mass, double theta, const Law& L, double* acc, int64_t* count) {
  struct Task {
    int64_t node;
    double size;
  };
  Task stack[512];
  int sp = 0;
  stack[sp++] = {root, bd_size};
  const V3 pi = load3(pos, i), vi = load3(vel, i);
  const double mi = mass[i];
  const double inv_m = 1.0 / mi;
  const double theta2 = theta * theta;
  V3 a{0.0, 0.0, 0.0};
  int64_t cnt = 0;
  const size_t nn = T.mass.size();
  int overflow = 0;
  while (sp > 0) {
    Task t = stack[--sp];
    if ((size_t)t.node >= nn) continue;
    V3 com = load3(T.com.data(), t.node);
    V3 vn = load3(T.vel.data(), t.node);
    double mn = T.mass[t.node];
    V3 d{pi.x - com.x, pi.y - com.y, pi.z - com.z};
    double d2 = d.x * d.x + d.y * d.y + d.z * d.z;
    if (d2 == 0.0) continue;
    bool leaf = T.right[t.node] == -1 && T.left[t.node] == -1;
    if ((t.size * t.size < d2 * theta2) || leaf) {
      V3 f = gravitational(mi, mn, vi, vn, pi, com, L);
      a.x += f.x * inv_m;
      a.y += f.y * inv_m;
      a.z += f.z * inv_m;
      ++cnt;
    } else {
      int64_t l = T.left[t.node], r = T.right[t.node];
      if (sp + 2 > 512) {
        overflow = 1;
        break;
      }
      if ((size_t)l < nn) stack[sp++] = {l, t.size * 0.5};
      if ((size_t)r < nn) stack[sp++] = {r, t.size * 0.5};
    }
  }
  store3(acc, i, a);
  if (count) count[i] = cnt;
  return overflow;
}

}  // namespace

extern "C" {

int orc_version() { return 1; }

// Parallelized::compute_forces (builder_endpoint.cpp:729-757) with
// compute_gravitational (tool_chain_impl.cpp:120-137). Pinned by: two-body
// closed form, Newton's third law for equal masses, brute-force restatement in
// tests, Plummer-field statistic.
int orc_allpairs(int64_t n, const double* pos, const double* vel, const double* mass, double G, double c,
                 double soft, int pn, int64_t lower, int64_t upper, double* acc, int nthreads) {
  if (n < 0 || !pos || !vel || !mass || !acc) return -1;
  if ((pn & kEIH) && (lower != 0 || (upper >= 0 && upper != n))) return -1;
  allpairs(n, pos, vel, mass, Law{G, c, soft, pn}, lower, upper, acc, nthreads);
  return 0;
}

// Single pair, for closed-form checks of compute_gravitational.
int orc_pair_force(double m1, double m2, const double* v1, const double* v2, const double* p1, const double* p2,
                   double G, double c, double soft, int pn, double* f_out) {
  V3 f = gravitational(m1, m2, V3{v1[0], v1[1], v1[2]}, V3{v2[0], v2[1], v2[2]}, V3{p1[0], p1[1], p1[2]},
                       V3{p2[0], p2[1], p2[2]}, Law{G, c, soft, pn});
  f_out[0] = f.x;
  f_out[1] = f.y;
  f_out[2] = f.z;
  return 0;
}

// get_morton, bh_specialization.cpp:38-46. Pinned by: per-bit interleave brute
// force over all 2^10 cell indices per axis, corner table, clamp at +box.
int orc_morton(int64_t n, const double* pos, const double* box, uint32_t* codes) {
  for (int64_t i = 0; i < n; ++i) codes[i] = morton(load3(pos, i), box);
  return 0;
}

// compare_bounding_vectors (bh_specialization.cpp:23-27) folded over all
// bodies: symmetric box half-extent = max |coordinate| per axis.
int orc_box_auto(int64_t n, const double* pos, double* box) {
  box[0] = box[1] = box[2] = 0.0;
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) box[c] = std::max(box[c], std::fabs(pos[3 * i + c]));
  return 0;
}

// Sort step of make_tree (barnes_hut.cpp:324) in the canonical (code, index)
// order. Pinned by: permutation + sortedness + tie order invariants.
int orc_sort(int64_t n, const uint32_t* codes, uint32_t* perm) {
  std::iota(perm, perm + n, 0u);
  std::stable_sort(perm, perm + n, [&](uint32_t a, uint32_t b) { return codes[a] < codes[b]; });
  return 0;
}

// Karras split, barnes_hut.cpp:338-396. left/right hold 2n-1 entries (leaves
// -1). Pinned by: the 8-key example of Karras (2012) Fig. 3, structural
// invariants and a top-down brute-force split on small inputs.
int orc_karras(int64_t n, const uint32_t* sorted_codes, int32_t* left, int32_t* right) {
  if (n < 1) return -1;
  karras(n, sorted_codes, left, right, 1);
  return 0;
}

// make_tree + unite: full tree for parity (barnes_hut.cpp:137-152, 309-400).
// Pinned by: root mass = sum m, root COM = sum m x / sum m, per-node brute
// force over each node's leaf range.
int orc_build_tree(int64_t n, const double* pos, const double* vel, const double* mass, const double* box,
                   uint32_t* codes_sorted, uint32_t* perm, int32_t* left, int32_t* right, double* nmass,
                   double* ncom, double* nvel) {
  if (n < 1) return -1;
  BuiltTree B;
  make_tree(n, pos, vel, mass, box, B, 0);
  const int64_t nn = 2 * n - 1;
  if (codes_sorted) std::memcpy(codes_sorted, B.sorted.data(), n * sizeof(uint32_t));
  if (perm) std::memcpy(perm, B.perm.data(), n * sizeof(uint32_t));
  if (left) std::memcpy(left, B.T.left.data(), nn * sizeof(int32_t));
  if (right) std::memcpy(right, B.T.right.data(), nn * sizeof(int32_t));
  if (nmass) std::memcpy(nmass, B.T.mass.data(), nn * sizeof(double));
  if (ncom) std::memcpy(ncom, B.T.com.data(), 3 * nn * sizeof(double));
  if (nvel) std::memcpy(nvel, B.T.vel.data(), 3 * nn * sizeof(double));
  return 0;
}

// Barnes_Hut::compute_forces, barnes_hut.cpp:402-418: make_tree, then every
// body traverses from root n with bd_size = |box|. `count` (optional) receives
// the number of accepted nodes per body. Returns 1 if a traversal stack
// overflowed. Pinned by: theta -> 0 equals all-pairs up to summation order;
// brute-force restatement on tiny inputs.
int orc_bh(int64_t n, const double* pos, const double* vel, const double* mass, const double* box,
           double theta, double G, double c, double soft, int pn, double* acc, int64_t* count, int nthreads) {
  if (n < 2) {
    for (int64_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
    if (count && n == 1) count[0] = 0;
    return 0;
  }
  BuiltTree B;
  make_tree(n, pos, vel, mass, box, B, nthreads);
  const double bd_size = std::sqrt(box[0] * box[0] + box[1] * box[1] + box[2] * box[2]);
  Law L{G, c, soft, pn};
  std::vector<int> ovf(n, 0);
  parallel_for(n, nthreads, [&](int64_t i) { ovf[i] = traverse(B.T, n, i, bd_size, pos, vel, mass, theta, L, acc, count); });
  for (int64_t i = 0; i < n; ++i)
    if (ovf[i]) return 1;
  return 0;
}

// Group traversal (no counterpart in the reference; the library's
// ASTER_BH_GROUP): the tree of make_tree, groups = subtrees of at most 64
// leaves whose parent has more. For each group, starting at the root, a node
// is accepted when it is a leaf or when the squared distance from its FP32
// COM to the group's FP32 box exceeds crit2 = (l / theta + delta)^2 (l the
// longest side of the node's FP32 box, delta = |COM - box centre|), and opened
// otherwise. The decisions use the same FP32 operations as the library (boxes
// from FP32-rounded positions, fmaf in the distance), so the accepted sets are
// the same; every body of the group then sums G m (r - x) / (d^2 + eps)^(3/2)
// over them in FP64 from the FP32 records and the FP32 body position. Counts:
// interactions = sum over groups of accepted x bodies - bodies (self pairs),
// visits = sum over groups of 1 + 2 x opened. Pinned by: theta -> 0 equals
// all-pairs on the FP32-rounded input with n (n - 1) interactions; the accepted
// sets agree with an ancestor-chain brute force (a node is accepted iff it
// passes and every ancestor was opened).
namespace {
struct GroupTree {
  std::vector<int32_t> parent;
  std::vector<int64_t> size;                 // leaves below each node
  std::vector<float> lo, hi, rec, crit2;     // 3, 3, 4 per node and 1 per node
};

float geo_crit2_f(const float* lo, const float* hi, const float* r, double theta) {
  const float l = std::max(hi[0] - lo[0], std::max(hi[1] - lo[1], hi[2] - lo[2]));
  const float cx = 0.5f * (hi[0] + lo[0]);
  const float cy = 0.5f * (hi[1] + lo[1]);
  const float cz = 0.5f * (hi[2] + lo[2]);
  const float ex = r[0] - cx, ey = r[1] - cy, ez = r[2] - cz;
  const float delta = std::sqrt(ex * ex + ey * ey + ez * ez);
  const float q = l / (float)theta + delta;
  return q * q;
}

void group_tree(int64_t n, const Tree& T, double theta, GroupTree& G) {
  const int64_t nn = 2 * n - 1;
  G.parent.assign(nn, -1);
  G.size.assign(nn, 1);
  G.lo.assign(3 * nn, 0.f);
  G.hi.assign(3 * nn, 0.f);
  G.rec.assign(4 * nn, 0.f);
  G.crit2.assign(nn, 0.f);
  for (int64_t k = n; k < nn; ++k) {
    G.parent[T.left[k]] = (int32_t)k;
    G.parent[T.right[k]] = (int32_t)k;
  }
  for (int64_t k = 0; k < nn; ++k) {
    for (int c = 0; c < 3; ++c) G.rec[4 * k + c] = (float)T.com[3 * k + c];
    G.rec[4 * k + 3] = (float)T.mass[k];
  }
  for (int64_t k = 0; k < n; ++k)
    for (int c = 0; c < 3; ++c) G.lo[3 * k + c] = G.hi[3 * k + c] = (float)T.com[3 * k + c];
  // internal nodes bottom-up: post-order from the root
  std::vector<std::pair<int64_t, int>> st{{n, 0}};
  while (!st.empty()) {
    auto& top = st.back();
    const int64_t k = top.first;
    if (k < n) {
      st.pop_back();
      continue;
    }
    if (top.second < 2) {
      const int side = top.second++;
      st.push_back({side == 0 ? T.left[k] : T.right[k], 0});
      continue;
    }
    st.pop_back();
    const int64_t L = T.left[k], R = T.right[k];
    G.size[k] = G.size[L] + G.size[R];
    for (int c = 0; c < 3; ++c) {
      G.lo[3 * k + c] = std::min(G.lo[3 * L + c], G.lo[3 * R + c]);
      G.hi[3 * k + c] = std::max(G.hi[3 * L + c], G.hi[3 * R + c]);
    }
  }
  for (int64_t k = 0; k < nn; ++k) G.crit2[k] = geo_crit2_f(&G.lo[3 * k], &G.hi[3 * k], &G.rec[4 * k], theta);
}

// squared distance from the FP32 point r to the box [blo, bhi], as the library forms it
float box_d2_f(const float* blo, const float* bhi, const float* r) {
  float e[3];
  for (int c = 0; c < 3; ++c) e[c] = std::max(std::max(blo[c] - r[c], r[c] - bhi[c]), 0.f);
  return std::fmaf(e[0], e[0], std::fmaf(e[1], e[1], e[2] * e[2]));
}
}  // namespace

extern "C" {
int orc_bh_group(int64_t n, const double* pos, const double* mass, const double* box, double theta, double G,
                 double soft, double* acc, int64_t* interactions, int64_t* visits, int nthreads) {
  if (n < 2) {
    for (int64_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
    if (interactions) *interactions = 0;
    if (visits) *visits = 0;
    return 0;
  }
  std::vector<double> zero(3 * n, 0.0);
  BuiltTree B;
  make_tree(n, pos, zero.data(), mass, box, B, nthreads);
  GroupTree GT;
  group_tree(n, B.T, theta, GT);
  std::vector<int64_t> roots;
  for (int64_t k = 0; k < 2 * n - 1; ++k)
    if (GT.size[k] <= 64 && (GT.parent[k] < 0 || GT.size[GT.parent[k]] > 64)) roots.push_back(k);
  // the leaf range of a subtree: leftmost and rightmost descendants
  auto first_leaf = [&](int64_t k) { while (k >= n) k = B.T.left[k]; return k; };
  const double eps = (double)(float)soft;
  std::vector<int64_t> g_inter(roots.size()), g_vis(roots.size());
  parallel_for((int64_t)roots.size(), nthreads, [&](int64_t g) {
    const int64_t root = roots[g], lo = first_leaf(root), cnt = GT.size[root];
    const float* blo = &GT.lo[3 * root];
    const float* bhi = &GT.hi[3 * root];
    std::vector<int64_t> list, stack{n};
    int64_t opened = 0;
    while (!stack.empty()) {
      const int64_t k = stack.back();
      stack.pop_back();
      if (k < n || box_d2_f(blo, bhi, &GT.rec[4 * k]) > GT.crit2[k]) {
        list.push_back(k);
      } else {
        ++opened;
        stack.push_back(B.T.left[k]);
        stack.push_back(B.T.right[k]);
      }
    }
    std::sort(list.begin(), list.end());  // a fixed summation order
    for (int64_t s = lo; s < lo + cnt; ++s) {
      const double xf = (double)(float)B.T.com[3 * s], yf = (double)(float)B.T.com[3 * s + 1],
                   zf = (double)(float)B.T.com[3 * s + 2];
      double ax = 0.0, ay = 0.0, az = 0.0;
      for (int64_t k : list) {
        const float* q = &GT.rec[4 * k];
        const double dx = (double)q[0] - xf, dy = (double)q[1] - yf, dz = (double)q[2] - zf;
        const double r2 = dx * dx + dy * dy + dz * dz + eps;
        if (r2 == 0.0) continue;
        const double f = (double)q[3] / (r2 * std::sqrt(r2));
        ax += f * dx;
        ay += f * dy;
        az += f * dz;
      }
      const uint32_t b = B.perm[s];
      acc[3 * b] = G * ax;
      acc[3 * b + 1] = G * ay;
      acc[3 * b + 2] = G * az;
    }
    g_inter[g] = (int64_t)list.size() * cnt - cnt;
    g_vis[g] = 1 + 2 * opened;
  });
  if (interactions) *interactions = std::accumulate(g_inter.begin(), g_inter.end(), (int64_t)0);
  if (visits) *visits = std::accumulate(g_vis.begin(), g_vis.end(), (int64_t)0);
  return 0;
}

// The post-Newtonian part of the group traversal's accelerations: the
// pairwise 1PN / 2PN / 2.5PN terms `pn` selects (compute_gravitational's PN
// part, tool_chain_impl.cpp:126-136) between each body and every node of its
// group's list (the lists of orc_bh_group), with the node's FP64 centre of
// mass, mass and mass-weighted velocity (get_node_body,
// barnes_hut.cpp:202-261), skipping nodes at distance 0 (:235), times 1/m_i.
// FP64 sums in node order. The library adds this to the FP32 Newtonian sums.
int orc_bh_group_pn(int64_t n, const double* pos, const double* vel, const double* mass, const double* box,
                    double theta, double G, double c, double soft, int pn, double* acc, int nthreads) {
  for (int64_t i = 0; i < 3 * n; ++i) acc[i] = 0.0;
  if (n < 2) return 0;
  BuiltTree B;
  make_tree(n, pos, vel, mass, box, B, nthreads);
  GroupTree GT;
  group_tree(n, B.T, theta, GT);
  std::vector<int64_t> roots;
  for (int64_t k = 0; k < 2 * n - 1; ++k)
    if (GT.size[k] <= 64 && (GT.parent[k] < 0 || GT.size[GT.parent[k]] > 64)) roots.push_back(k);
  auto first_leaf = [&](int64_t k) { while (k >= n) k = B.T.left[k]; return k; };
  const Law L{G, c, soft, pn & (kPN1 | kPN2 | kPN25)};
  parallel_for((int64_t)roots.size(), nthreads, [&](int64_t g) {
    const int64_t root = roots[g], lo = first_leaf(root), cnt = GT.size[root];
    const float* blo = &GT.lo[3 * root];
    const float* bhi = &GT.hi[3 * root];
    std::vector<int64_t> list, stack{n};
    while (!stack.empty()) {
      const int64_t k = stack.back();
      stack.pop_back();
      if (k < n || box_d2_f(blo, bhi, &GT.rec[4 * k]) > GT.crit2[k]) {
        list.push_back(k);
      } else {
        stack.push_back(B.T.left[k]);
        stack.push_back(B.T.right[k]);
      }
    }
    std::sort(list.begin(), list.end());
    for (int64_t s = lo; s < lo + cnt; ++s) {
      const uint32_t b = B.perm[s];
      const V3 pi = load3(pos, b), vi = load3(vel, b);
      const double mi = mass[b];
      const double inv_m = 1.0 / mi;
      V3 a{0.0, 0.0, 0.0};
      for (int64_t k : list) {
        const V3 com = load3(B.T.com.data(), k), vn = load3(B.T.vel.data(), k);
        const double mn = B.T.mass[k];
        const V3 d{pi.x - com.x, pi.y - com.y, pi.z - com.z};
        if (d.x * d.x + d.y * d.y + d.z * d.z == 0.0) continue;
        const V3 f = pn_part(mi, mn, vi, vn, pi, com, L);
        a.x += f.x * inv_m;
        a.y += f.y * inv_m;
        a.z += f.z * inv_m;
      }
      store3(acc, b, a);
    }
  });
  return 0;
}

// The accepted sets of orc_bh_group counted without a walk: node k is accepted
// for group g iff it passes (leaf or outside crit2) and every ancestor of k
// fails. Returns the interaction count; test infrastructure for the walk.
int64_t orc_bh_group_bruteforce_count(int64_t n, const double* pos, const double* mass, const double* box,
                                      double theta) {
  if (n < 2) return 0;
  std::vector<double> zero(3 * n, 0.0);
  BuiltTree B;
  make_tree(n, pos, zero.data(), mass, box, B, 0);
  GroupTree GT;
  group_tree(n, B.T, theta, GT);
  int64_t total = 0;
  for (int64_t g = 0; g < 2 * n - 1; ++g) {
    if (!(GT.size[g] <= 64 && (GT.parent[g] < 0 || GT.size[GT.parent[g]] > 64))) continue;
    auto pass = [&](int64_t k) {
      return k < n || box_d2_f(&GT.lo[3 * g], &GT.hi[3 * g], &GT.rec[4 * k]) > GT.crit2[k];
    };
    int64_t acc = 0;
    for (int64_t k = 0; k < 2 * n - 1; ++k) {
      if (!pass(k)) continue;
      bool ok = true;
      for (int64_t a = GT.parent[k]; a >= 0 && ok; a = GT.parent[a]) ok = a >= n && !pass(a);
      acc += ok;
    }
    total += acc * GT.size[g] - GT.size[g];
  }
  return total;
}
}  // extern "C"

// Graphs::get_total_energy, collectors.cpp:19-53 (serial order), plus the
// Plummer-consistent potential -G m_i m_j / sqrt(r^2 + eps^2) and the total
// angular momentum sum m x cross v (no reference counterpart).
int orc_energy(int64_t n, const double* pos, const double* vel, const double* mass, double G, double soft,
               double* e_ref, double* e_plummer, double* L3) {
  double e = 0.0, ep = 0.0, kin = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    V3 vi = load3(vel, i);
    double k = 0.5 * mass[i] * (vi.x * vi.x + vi.y * vi.y + vi.z * vi.z);
    e += k;
    kin += k;
    V3 pi = load3(pos, i);
    double pp = 0.0;
    for (int64_t j = i + 1; j < n; ++j) {
      V3 pj = load3(pos, j);
      V3 dr{pj.x - pi.x, pj.y - pi.y, pj.z - pi.z};
      double r2 = dr.x * dr.x + dr.y * dr.y + dr.z * dr.z;
      double inv_r = 1.0 / (std::sqrt(r2) + soft);
      e -= G * mass[i] * mass[j] * inv_r;
      pp += mass[j] / std::sqrt(r2 + soft);
    }
    ep -= G * mass[i] * pp;
  }
  if (e_ref) *e_ref = e;
  if (e_plummer) *e_plummer = kin + ep;
  if (L3) {
    double lx = 0, ly = 0, lz = 0;
    for (int64_t i = 0; i < n; ++i) {
      V3 p = load3(pos, i), v = load3(vel, i);
      lx += mass[i] * (p.y * v.z - p.z * v.y);
      ly += mass[i] * (p.z * v.x - p.x * v.z);
      lz += mass[i] * (p.x * v.y - p.y * v.x);
    }
    L3[0] = lx;
    L3[1] = ly;
    L3[2] = lz;
  }
  return 0;
}

// Yoshida (1990) fourth-order weights: w1 = 1/(2 - 2^(1/3)), w0 = 1 - 2 w1,
// evaluated in FP64 from the correctly rounded cube root of 2 (the test pins
// the closed form and the order conditions).
int orc_yoshida_weights(double* w) {
  const double cbrt2 = 0x1.428a2f98d728bp+0;
  const double w1 = 1.0 / (2.0 - cbrt2);
  w[0] = w1;
  w[1] = 1.0 - 2.0 * w1;
  w[2] = w1;
  return 0;
}

// update_leapfrog, tool_chain_impl.cpp:169-185: F(x); x += v dt + (0.5 a) dt dt;
// v += (0.5 a) dt; F(x); v += (0.5 a) dt. integrator 2 = leapfrog, 5 =
// Yoshida-4 (three leapfrog steps of w_k dt). method 0 = all-pairs, 1 =
// Barnes-Hut with `box`/`theta`, 2 = the group traversal (orc_bh_group,
// Newtonian) with `theta`; for methods 1 and 2 an all-zero box is the
// symmetric box of the positions at each evaluation, as the library's
// automatic box. `acc` (3n) receives the last accelerations.
// Pinned by: order-2/order-4 convergence on a Kepler orbit, time reversibility,
// energy drift bounds.
int orc_integrate(int64_t n, double* pos, double* vel, const double* mass, double G, double c, double soft,
                  int pn, int method, const double* box, double theta, int integrator, double dt, int64_t nsteps,
                  double* acc_out, int nthreads) {
  if (integrator != 2 && integrator != 5) return -1;
  if (method < 0 || method > 2) return -1;
  std::vector<double> acc(3 * n, 0.0);
  Law L{G, c, soft, pn};
  const bool auto_box = method != 0 && box[0] == 0.0 && box[1] == 0.0 && box[2] == 0.0;
  double bx[3] = {0.0, 0.0, 0.0};
  auto forces = [&]() {
    if (method == 0) {
      allpairs(n, pos, vel, mass, L, 0, n, acc.data(), nthreads);
      return;
    }
    if (auto_box)
      orc_box_auto(n, pos, bx);
    else
      for (int k = 0; k < 3; ++k) bx[k] = box[k];
    if (method == 1)
      orc_bh(n, pos, vel, mass, bx, theta, G, c, soft, pn, acc.data(), nullptr, nthreads);
    else
      orc_bh_group(n, pos, mass, bx, theta, G, soft, acc.data(), nullptr, nullptr, nthreads);
  };
  double w[3] = {1.0, 0.0, 0.0};
  int nsub = 1;
  if (integrator == 5) {
    orc_yoshida_weights(w);
    nsub = 3;
  }
  for (int64_t s = 0; s < nsteps; ++s) {
    for (int k = 0; k < nsub; ++k) {
      const double h = (integrator == 5) ? w[k] * dt : dt;
      forces();
      parallel_for(n, nthreads, [&](int64_t i) {
        for (int c3 = 0; c3 < 3; ++c3) {
          double a = acc[3 * i + c3];
          pos[3 * i + c3] += vel[3 * i + c3] * h + 0.5 * a * h * h;
          vel[3 * i + c3] += 0.5 * a * h;
        }
      });
      forces();
      parallel_for(n, nthreads, [&](int64_t i) {
        for (int c3 = 0; c3 < 3; ++c3) vel[3 * i + c3] += 0.5 * acc[3 * i + c3] * h;
      });
    }
  }
  if (acc_out) std::memcpy(acc_out, acc.data(), 3 * n * sizeof(double));
  return 0;
}

// update_euler (src/physics/tool_chain_impl.cpp:158-167) for order 1, which
// is what update_SABA1 calls (SABA.cpp:14-16), and update_SABA2..10
// (SABA.cpp:19-553) for orders 2..10: drift c_1, then per stage one force
// evaluation, a kick v += (a d_k) dt and a drift x += (v c_(k+1)) dt. The
// drifts/kicks are the library's SABAn table (Gauss-Legendre nodes and
// weights, with the reference's doubles where its literals differ).
// Pinned by: the table against numpy's Gauss-Legendre rule and the quadrature
// exactness conditions; order 1 / order 2 convergence on a Kepler orbit; the
// order-2n error of SABAn on a weakly interacting fly-by; time reversibility.
int orc_integrate_saba(int64_t n, double* pos, double* vel, const double* mass, double G, double c, double soft,
                       int pn, int method, const double* box, double theta, int order, double dt, int64_t nsteps,
                       double* acc_out, int nthreads) {
  if (order < 1 || order > aster::kSabaMaxOrder) return -1;
  std::vector<double> acc(3 * n, 0.0);
  Law L{G, c, soft, pn};
  auto forces = [&]() {
    if (method == 0)
      allpairs(n, pos, vel, mass, L, 0, n, acc.data(), nthreads);
    else
      orc_bh(n, pos, vel, mass, box, theta, G, c, soft, pn, acc.data(), nullptr, nthreads);
  };
  // one stage: v += a * d * dt; x += v * cdrift * dt (the reference's lambdas)
  auto kick_drift = [&](double d, double cdrift) {
    parallel_for(n, nthreads, [&](int64_t i) {
      for (int k = 0; k < 3; ++k) {
        vel[3 * i + k] += acc[3 * i + k] * d * dt;
        pos[3 * i + k] += vel[3 * i + k] * cdrift * dt;
      }
    });
  };
  for (int64_t s = 0; s < nsteps; ++s) {
    if (order == 1) {
      forces();
      parallel_for(n, nthreads, [&](int64_t i) {
        for (int k = 0; k < 3; ++k) {
          vel[3 * i + k] += acc[3 * i + k] * dt;
          pos[3 * i + k] += vel[3 * i + k] * dt;
        }
      });
      continue;
    }
    const double* cs = aster::kSabaC[order];
    const double* ds = aster::kSabaD[order];
    parallel_for(n, nthreads, [&](int64_t i) {
      for (int k = 0; k < 3; ++k) pos[3 * i + k] += vel[3 * i + k] * cs[0] * dt;
    });
    for (int st = 0; st < order; ++st) {
      forces();
      kick_drift(ds[st], cs[st + 1]);
    }
  }
  if (acc_out) std::memcpy(acc_out, acc.data(), 3 * n * sizeof(double));
  return 0;
}

// ---- Wisdom-Holman (democratic heliocentric) for a central-mass system ----
//
// The reference's update_WH_planetary (tool_chain_impl.cpp:235-274) drifts
// every body i >= 1 on a Kepler orbit about body 0 (update_kepler, :187-233)
// and then kicks with the full solver accelerations, which contain body 0's
// pull again, so the central force is applied twice per step. The oracle
// restates the scheme that splitting is meant to be: Wisdom & Holman (1991) in
// democratic heliocentric coordinates (Duncan, Levison & Lee 1998), with
// heliocentric positions Q_i = x_i - x_0 and barycentric velocities u_i
// (i >= 1) and H = H_Kepler(mu = G m_0) + H_interaction(planet pairs) +
// |sum m_i u_i|^2 / (2 m_0). One step: half interaction kick, half "jump"
// Q_i += (sum m_j u_j / m_0) dt/2, Kepler drift dt, half jump, half kick.
// Newtonian pair forces (compute_gravitational, softened) between planets.

// Stumpff functions c0..c3 of z: quarter z until |z| <= 0.1, series (8
// nested terms), then c3 = (c2 + c0 c3)/4, c2 = c1^2/2, c1 = c0 c1,
// c0 = 2 c0^2 - 1 once per quartering. Only +, -, *, / (bit-reproducible).
static void stumpff(double z, double& c0, double& c1, double& c2, double& c3) {
  int nq = 0;
  while (std::fabs(z) > 0.1) {
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

// Kepler drift of (q, u) about mu for dt in universal variables (G-functions
// G_k = s^k c_k(beta s^2)). The universal Kepler equation
// F(s) = r0 G1 + eta0 G2 + mu G3 - dt is increasing (F' = r > 0), so the root
// is bracketed: on a bound orbit dt is reduced modulo the period (fmod is
// exact) and s lies in [0, 2 pi / sqrt(beta)] (one revolution), otherwise the
// bracket is grown by doubling. Halley steps from s = dt / r0, replaced by
// bisection whenever they leave the bracket, at most 100, until
// |ds| <= 1e-15 |s|; then the f and g functions.
struct KepState {
  double r0, eta0, zeta0, beta, mu;
};

static void kepler_eval(const KepState& K, double s, double dt, double& F, double& Fp, double& Fpp) {
  double c0, c1, c2, c3;
  stumpff(K.beta * s * s, c0, c1, c2, c3);
  const double g1 = s * c1, g2 = s * s * c2, g3 = s * s * s * c3;
  F = K.r0 * g1 + K.eta0 * g2 + K.mu * g3 - dt;
  Fp = K.r0 * c0 + K.eta0 * g1 + K.mu * g2;
  Fpp = K.eta0 * c0 + K.zeta0 * g1;
}

static void kepler_drift(double* q, double* u, double mu, double dt) {
  KepState K;
  K.mu = mu;
  K.r0 = std::sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2]);
  const double v2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
  K.eta0 = q[0] * u[0] + q[1] * u[1] + q[2] * u[2];
  K.beta = 2.0 * mu / K.r0 - v2;
  K.zeta0 = mu - K.beta * K.r0;
  const double sgn = dt < 0.0 ? -1.0 : 1.0;
  double F, Fp, Fpp;
  double lo = 0.0, hi;  // the root lies between lo and hi (in the direction of dt)
  if (K.beta > 0.0) {
    const double sq = std::sqrt(K.beta);
    dt = std::fmod(dt, 6.283185307179586 * mu / (K.beta * sq));
    hi = sgn * (6.283185307179586 / sq);
  } else {
    hi = dt / K.r0;
    for (int k = 0; k < 200; ++k) {
      kepler_eval(K, hi, dt, F, Fp, Fpp);
      if (sgn * F >= 0.0) break;
      lo = hi;
      hi = 2.0 * hi;
    }
  }
  double s = dt / K.r0;
  if (!(sgn * (s - lo) > 0.0 && sgn * (hi - s) > 0.0)) s = 0.5 * (lo + hi);
  for (int it = 0; it < 100; ++it) {
    kepler_eval(K, s, dt, F, Fp, Fpp);
    if (sgn * F < 0.0)
      lo = s;
    else
      hi = s;
    double sn = s - F / (Fp - 0.5 * F * Fpp / Fp);
    if (!(sgn * (sn - lo) > 0.0 && sgn * (hi - sn) > 0.0)) sn = 0.5 * (lo + hi);
    const double ds = sn - s;
    s = sn;
    if (std::fabs(ds) <= 1e-15 * std::fabs(s)) break;
  }
  double c0, c1, c2, c3;
  stumpff(K.beta * s * s, c0, c1, c2, c3);
  const double g1 = s * c1, g2 = s * s * c2;
  const double r = K.r0 * c0 + K.eta0 * g1 + mu * g2;
  const double f = 1.0 - mu * g2 / K.r0;
  const double g = K.r0 * g1 + K.eta0 * g2;
  const double fd = -mu * g1 / (r * K.r0);
  const double gd = 1.0 - mu * g2 / r;
  double qn[3], un[3];
  for (int k = 0; k < 3; ++k) {
    qn[k] = f * q[k] + g * u[k];
    un[k] = fd * q[k] + gd * u[k];
  }
  for (int k = 0; k < 3; ++k) {
    q[k] = qn[k];
    u[k] = un[k];
  }
}

int orc_kepler_drift(double* q, double* u, double mu, double dt) {
  kepler_drift(q, u, mu, dt);
  return 0;
}

// One system, body 0 central, n >= 1 (pos/vel AoS-d3, updated in place).
// Conversions: M = sum m_j (j order); X = sum m_j x_j / M, V likewise;
// Q_i = x_i - x_0, u_i = v_i - V. Back: x_0 = X - sum_{j>=1} m_j Q_j / M,
// x_i = Q_i + x_0, v_0 = V - sum_{j>=1} m_j u_j / m_0, v_i = u_i + V; X
// advances by V dt per step.
int orc_integrate_wh(int64_t n, double* pos, double* vel, const double* mass, double G, double soft, double dt,
                     int64_t nsteps) {
  if (n < 1) return -1;
  Law L{G, 0.0, soft, kNewton};
  const double m0 = mass[0];
  const double mu = G * m0;
  double M = 0.0, X[3] = {0.0, 0.0, 0.0}, Vc[3] = {0.0, 0.0, 0.0};
  for (int64_t j = 0; j < n; ++j) {
    M += mass[j];
    for (int k = 0; k < 3; ++k) {
      X[k] += mass[j] * pos[3 * j + k];
      Vc[k] += mass[j] * vel[3 * j + k];
    }
  }
  for (int k = 0; k < 3; ++k) {
    X[k] = X[k] / M;
    Vc[k] = Vc[k] / M;
  }
  std::vector<double> Q(3 * n, 0.0), U(3 * n, 0.0), A(3 * n, 0.0);
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      Q[3 * i + k] = pos[3 * i + k] - pos[k];
      U[3 * i + k] = vel[3 * i + k] - Vc[k];
    }
  auto interaction = [&]() {
    for (int64_t i = 1; i < n; ++i) {
      V3 a{0.0, 0.0, 0.0};
      const V3 zero{0.0, 0.0, 0.0};
      for (int64_t j = 1; j < n; ++j) {
        if (j == i) continue;
        V3 f = gravitational(mass[i], mass[j], zero, zero, load3(Q.data(), i), load3(Q.data(), j), L);
        a.x += f.x / mass[i];
        a.y += f.y / mass[i];
        a.z += f.z / mass[i];
      }
      store3(A.data(), i, a);
    }
  };
  const double h = 0.5 * dt;
  auto kick = [&]() {
    for (int64_t i = 1; i < n; ++i)
      for (int k = 0; k < 3; ++k) U[3 * i + k] = U[3 * i + k] + A[3 * i + k] * h;
  };
  auto jump = [&]() {
    double P[3] = {0.0, 0.0, 0.0};
    for (int64_t j = 1; j < n; ++j)
      for (int k = 0; k < 3; ++k) P[k] += mass[j] * U[3 * j + k];
    for (int64_t i = 1; i < n; ++i)
      for (int k = 0; k < 3; ++k) Q[3 * i + k] = Q[3 * i + k] + P[k] / m0 * h;
  };
  interaction();
  for (int64_t s = 0; s < nsteps; ++s) {
    kick();
    jump();
    for (int64_t i = 1; i < n; ++i) kepler_drift(&Q[3 * i], &U[3 * i], mu, dt);
    jump();
    interaction();
    kick();
    for (int k = 0; k < 3; ++k) X[k] = X[k] + Vc[k] * dt;
  }
  double SQ[3] = {0.0, 0.0, 0.0}, SU[3] = {0.0, 0.0, 0.0};
  for (int64_t j = 1; j < n; ++j)
    for (int k = 0; k < 3; ++k) {
      SQ[k] += mass[j] * Q[3 * j + k];
      SU[k] += mass[j] * U[3 * j + k];
    }
  for (int k = 0; k < 3; ++k) {
    pos[k] = X[k] - SQ[k] / M;
    vel[k] = Vc[k] - SU[k] / m0;
  }
  for (int64_t i = 1; i < n; ++i)
    for (int k = 0; k < 3; ++k) {
      pos[3 * i + k] = Q[3 * i + k] + pos[k];
      vel[3 * i + k] = U[3 * i + k] + Vc[k];
    }
  return 0;
}

// The SABAn drift (order + 1) and kick (order) coefficients the oracle uses.
int orc_saba_coeffs(int order, double* c_out, double* d_out) {
  if (order < 2 || order > aster::kSabaMaxOrder) return -1;
  for (int i = 0; i <= order; ++i) c_out[i] = aster::kSabaC[order][i];
  for (int i = 0; i < order; ++i) d_out[i] = aster::kSabaD[order][i];
  return 0;
}

// Discrete adjoint of `nsteps` leapfrog steps with Newtonian all-pairs forces
// (the FP64 CPU reference gradient; no reference counterpart). Stores the whole
// trajectory, so it is meant for small n. loss = sum over snapshots s and
// bodies i of |x_i(t_s) - target_s,i|^2 / (nsnap * n). Pinned by: central
// finite differences of the same forward map.
int orc_traj_grad(int64_t n, const double* pos0, const double* vel0, const double* mass, double G, double soft,
                  double dt, int64_t nsteps, int nsnap, const int64_t* snap_steps, const double* target,
                  double* loss, double* gx, double* gv, double* gm) {
  const int64_t n3 = 3 * n;
  std::vector<double> X((nsteps + 1) * n3), V((nsteps + 1) * n3), A((nsteps + 1) * n3);
  std::memcpy(X.data(), pos0, n3 * sizeof(double));
  std::memcpy(V.data(), vel0, n3 * sizeof(double));
  std::vector<double> zero(n3, 0.0);
  Law L{G, 0.0, soft, kNewton};
  allpairs(n, X.data(), zero.data(), mass, L, 0, n, A.data(), 1);
  const double h = 0.5 * dt;
  for (int64_t s = 0; s < nsteps; ++s) {
    const double *x = &X[s * n3], *v = &V[s * n3], *a = &A[s * n3];
    double *xn = &X[(s + 1) * n3], *vn = &V[(s + 1) * n3], *an = &A[(s + 1) * n3];
    std::vector<double> vh(n3);
    for (int64_t k = 0; k < n3; ++k) {
      xn[k] = x[k] + v[k] * dt + 0.5 * a[k] * dt * dt;
      vh[k] = v[k] + 0.5 * a[k] * dt;
    }
    allpairs(n, xn, zero.data(), mass, L, 0, n, an, 1);
    for (int64_t k = 0; k < n3; ++k) vn[k] = vh[k] + 0.5 * an[k] * dt;
  }
  double lsum = 0.0;
  const double norm = 1.0 / ((double)nsnap * (double)n);
  std::vector<double> lx(n3, 0.0), lv(n3, 0.0), lm(n, 0.0);
  // dL/dx at snapshot steps
  auto add_snap = [&](int64_t step) {
    for (int q = 0; q < nsnap; ++q)
      if (snap_steps[q] == step) {
        const double* x = &X[step * n3];
        const double* tg = &target[q * n3];
        for (int64_t k = 0; k < n3; ++k) {
          double dd = x[k] - tg[k];
          lsum += dd * dd * norm;
          lx[k] += 2.0 * dd * norm;
        }
      }
  };
  // Vector-Jacobian product of a(x) (Plummer-softened Newton) with w:
  // gx_k += G sum_j K_kj (m_k w_j - m_j w_k), gm_k += -G sum_j (w_j . d_kj)/s^3,
  // where d_kj = x_j - x_k, s^2 = |d|^2 + soft, K = I/s^3 - 3 d d^T / s^5.
  auto vjp = [&](const double* x, const std::vector<double>& w, std::vector<double>& ox, std::vector<double>& om) {
    for (int64_t k = 0; k < n; ++k) {
      double gxx = 0, gxy = 0, gxz = 0, gmk = 0;
      for (int64_t j = 0; j < n; ++j) {
        if (j == k) continue;
        double dx = x[3 * j] - x[3 * k], dy = x[3 * j + 1] - x[3 * k + 1], dz = x[3 * j + 2] - x[3 * k + 2];
        double s2 = dx * dx + dy * dy + dz * dz + soft;
        double is = 1.0 / std::sqrt(s2);
        double is3 = is * is * is, is5 = is3 * is * is;
        double ux = mass[k] * w[3 * j] - mass[j] * w[3 * k];
        double uy = mass[k] * w[3 * j + 1] - mass[j] * w[3 * k + 1];
        double uz = mass[k] * w[3 * j + 2] - mass[j] * w[3 * k + 2];
        double du = dx * ux + dy * uy + dz * uz;
        gxx += ux * is3 - 3.0 * dx * du * is5;
        gxy += uy * is3 - 3.0 * dy * du * is5;
        gxz += uz * is3 - 3.0 * dz * du * is5;
        gmk -= (w[3 * j] * dx + w[3 * j + 1] * dy + w[3 * j + 2] * dz) * is3;
      }
      ox[3 * k] += G * gxx;
      ox[3 * k + 1] += G * gxy;
      ox[3 * k + 2] += G * gxz;
      om[k] += G * gmk;
    }
  };
  add_snap(nsteps);
  std::vector<double> w(n3);
  for (int64_t s = nsteps - 1; s >= 0; --s) {
    // v_{s+1} = vh + h a(x_{s+1})
    for (int64_t k = 0; k < n3; ++k) w[k] = h * lv[k];
    vjp(&X[(s + 1) * n3], w, lx, lm);
    // vh = lv (unchanged); x_{s+1} = x_s + dt vh  -> lvh += dt lx
    std::vector<double> lvh(n3);
    for (int64_t k = 0; k < n3; ++k) lvh[k] = lv[k] + dt * lx[k];
    // vh = v_s + h a(x_s)
    for (int64_t k = 0; k < n3; ++k) w[k] = h * lvh[k];
    vjp(&X[s * n3], w, lx, lm);
    lv = lvh;
    add_snap(s);
  }
  if (loss) *loss = lsum;
  if (gx) std::memcpy(gx, lx.data(), n3 * sizeof(double));
  if (gv) std::memcpy(gv, lv.data(), n3 * sizeof(double));
  if (gm) std::memcpy(gm, lm.data(), n * sizeof(double));
  return 0;
}

// CPU baseline, bounded sample: Parallelized::compute_forces for targets
// [t0, t1) over all n sources; wall seconds in *seconds.
int orc_allpairs_sample(int64_t n, const double* pos, const double* vel, const double* mass, double G, double c,
                        double soft, int pn, int64_t t0, int64_t t1, int nthreads, double* seconds, double* acc) {
  if (t0 < 0 || t1 > n || t0 > t1) return -1;
  Law L{G, c, soft, pn};
  auto a = std::chrono::steady_clock::now();
  parallel_for(t1 - t0, nthreads, [&](int64_t k) {
    const int64_t i = t0 + k;
    V3 s{0.0, 0.0, 0.0};
    const double m1 = mass[i];
    const V3 v1 = load3(vel, i), p1 = load3(pos, i);
    for (int64_t j = 0; j < n; ++j) {
      if (i == j) continue;
      V3 f = gravitational(m1, mass[j], v1, load3(vel, j), p1, load3(pos, j), L);
      s.x += f.x / m1;
      s.y += f.y / m1;
      s.z += f.z / m1;
    }
    store3(acc, k, s);
  });
  *seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - a).count();
  return 0;
}

// CPU baseline, bounded sample: Barnes_Hut::make_tree over all n bodies, then
// get_node_body for bodies with i % stride == 0. Reports build and traversal
// wall seconds and the accepted interactions of the sampled bodies.
int orc_bh_sample(int64_t n, const double* pos, const double* vel, const double* mass, const double* box,
                  double theta, double G, double c, double soft, int pn, int64_t stride, int nthreads,
                  double* t_build, double* t_trav, int64_t* interactions) {
  if (n < 2 || stride < 1) return -1;
  auto a = std::chrono::steady_clock::now();
  BuiltTree B;
  make_tree(n, pos, vel, mass, box, B, nthreads);
  auto b = std::chrono::steady_clock::now();
  const double bd_size = std::sqrt(box[0] * box[0] + box[1] * box[1] + box[2] * box[2]);
  Law L{G, c, soft, pn};
  const int64_t ns = (n + stride - 1) / stride;
  std::vector<double> acc(3 * n);
  std::vector<int64_t> cnt(n, 0);
  parallel_for(ns, nthreads, [&](int64_t k) {
    traverse(B.T, n, k * stride, bd_size, pos, vel, mass, theta, L, acc.data(), cnt.data());
  });
  auto e = std::chrono::steady_clock::now();
  *t_build = std::chrono::duration<double>(b - a).count();
  *t_trav = std::chrono::duration<double>(e - b).count();
  int64_t tot = 0;
  for (int64_t i = 0; i < n; ++i) tot += cnt[i];
  *interactions = tot;
  return 0;
}

// Loss only (same forward map as orc_traj_grad), for finite differences.
int orc_traj_loss(int64_t n, const double* pos0, const double* vel0, const double* mass, double G, double soft,
                  double dt, int64_t nsteps, int nsnap, const int64_t* snap_steps, const double* target,
                  double* loss) {
  return orc_traj_grad(n, pos0, vel0, mass, G, soft, dt, nsteps, nsnap, snap_steps, target, loss, nullptr,
                       nullptr, nullptr);
}

// ---- seeded initial conditions ------------------------------------------
// The reference draws its presets from std::rand (presets_impl.cpp:70-108),
// which no test can reproduce; what is restated here is the seeded stream
// the library defines (SplitMix64, Steele, Lea & Flood 2014) and the body
// laws: the Plummer sphere of Aarseth, Henon & Wielen (1974), the uniform
// cube, and add_disk's per-body transform (presets_impl.cpp:47-64, 130-147,
// 207-255). Elementary functions come from the C library, and sums are
// sequential, so the library's generators (own exp/log/sincos, fixed
// parallel reduction order) agree with these to rounding, not bit for bit.
// Pinned by: the published SplitMix64 output sequence for state 1234567;
// the Plummer cumulative mass profile M(<r) = r^3 (1 + r^2)^(-3/2) (KS test),
// the virial ratio, zero barycentre and momentum; the disk's radius range,
// circular speed and rotation norms.

static uint64_t orc_mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

namespace {
struct OrcStream {
  uint64_t state;
  OrcStream(uint64_t seed, uint64_t i) : state(orc_mix64(seed ^ orc_mix64(i + 0x632BE59BD9B4E019ull))) {}
  uint64_t next() {
    state += 0x9E3779B97F4A7C15ull;
    return orc_mix64(state);
  }
  double u() { return (double)(next() >> 11) / 9007199254740992.0; }
  double u_open0() { return (double)((next() >> 11) + 1) / 9007199254740992.0; }
};

void orc_iso(OrcStream& s, double r, double* out) {
  const double ct = 1.0 - 2.0 * s.u();
  const double st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
  const double ph = 2.0 * M_PI * s.u();
  out[0] = r * st * std::cos(ph);
  out[1] = r * st * std::sin(ph);
  out[2] = r * ct;
}

void orc_recentre(int64_t n, double* pos, double* vel, const double* mass) {
  double M = 0.0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    M += mass[i];
    for (int c = 0; c < 3; ++c) {
      cx[c] += mass[i] * pos[3 * i + c];
      cv[c] += mass[i] * vel[3 * i + c];
    }
  }
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] -= cx[c] / M;
      vel[3 * i + c] -= cv[c] / M;
    }
}

// rotate_point (presets_impl.cpp:47-64): single-precision intermediates.
void orc_rotate_point(const double* v, double phi, double theta, double* out) {
  float x1, z1, y1, z2;
  theta *= M_PI / 180;
  phi *= M_PI / 180;
  x1 = v[0] * std::cos(phi) + v[2] * std::sin(phi);
  z1 = -v[0] * std::sin(phi) + v[2] * std::cos(phi);
  y1 = v[1] * std::cos(theta) + z1 * std::sin(theta);
  z2 = -v[1] * std::sin(theta) + z1 * std::cos(theta);
  out[0] = x1;
  out[1] = y1;
  out[2] = z2;
}
}  // namespace

// Raw SplitMix64 outputs 1..k from `state`.
int orc_splitmix64(uint64_t state, int k, uint64_t* out) {
  for (int j = 0; j < k; ++j) {
    state += 0x9E3779B97F4A7C15ull;
    out[j] = orc_mix64(state);
  }
  return 0;
}

// First k uniforms [0, 1) of the stream of item i under `seed`.
int orc_stream_uniforms(uint64_t seed, uint64_t i, int k, double* out) {
  OrcStream s(seed, i);
  for (int j = 0; j < k; ++j) out[j] = s.u();
  return 0;
}

int orc_ic_plummer(int64_t n, uint64_t seed, double G, double mtot, double b, double rmax, int mass_mode,
                   double* pos, double* vel, double* mass) {
  const double vscale = std::sqrt(G * mtot / b);
  for (int64_t i = 0; i < n; ++i) {
    OrcStream s(seed, (uint64_t)i);
    double r = rmax;
    for (int tries = 0; tries < 1000; ++tries) {
      const double x1 = s.u();
      if (x1 <= 0.0) continue;
      const double rt = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
      if (rt <= rmax) {
        r = rt;
        break;
      }
    }
    orc_iso(s, r * b, pos + 3 * i);
    double q = 0.0;
    for (int tries = 0; tries < 1000; ++tries) {
      const double x4 = s.u(), x5 = s.u();
      q = x4;
      if (0.1 * x5 < x4 * x4 * std::pow(1.0 - x4 * x4, 3.5)) break;
    }
    orc_iso(s, q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25) * vscale, vel + 3 * i);
    mass[i] = mass_mode == 1 ? std::pow(10.0, 2.0 * s.u()) : 1.0;
  }
  double msum = 0.0;
  for (int64_t i = 0; i < n; ++i) msum += mass[i];
  for (int64_t i = 0; i < n; ++i) mass[i] = mass_mode == 1 ? mass[i] * mtot / msum : mtot / (double)n;
  orc_recentre(n, pos, vel, mass);
  return 0;
}

int orc_ic_cube(int64_t n, uint64_t seed, double h, double mtot, double* pos, double* vel, double* mass) {
  for (int64_t i = 0; i < n; ++i) {
    OrcStream s(seed, (uint64_t)i);
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] = -h + 2.0 * h * s.u();
      vel[3 * i + c] = 0.0;
    }
    mass[i] = mtot / (double)n;
  }
  return 0;
}

// add_disk (presets_impl.cpp:207-255) with seeded draws: body 0 is the
// central body, bodies 1..n the disk (stream k - 1 for body k).
int orc_cluster_disk(int64_t n, uint64_t seed, double G, double radius, double thickness, double inner, double scale,
                     double avr_mass, double center_mass, const double* center, const double* v,
                     const double* rotation, double* pos, double* vel, double* mass) {
  mass[0] = center_mass;
  orc_rotate_point(center, rotation[0], rotation[2], pos);
  orc_rotate_point(v, rotation[0], rotation[2], vel);
  for (int64_t k = 0; k < n; ++k) {
    OrcStream s(seed, (uint64_t)k);
    // rng_point_in_cylinder (:130-147); its thickness sign flip (:134-135) is kept
    const double th = 2.0 * M_PI * s.u();
    const double rr = inner + (radius - inner) * s.u();
    const double t = -thickness;
    const double cyl[3] = {rr * std::cos(th), rr * std::sin(th), (s.u() - 0.5) * t};
    const double p[3] = {cyl[0] * scale, cyl[1] * scale, cyl[2] * scale};
    const double r = std::sqrt(p[0] * p[0] + p[1] * p[1] + p[2] * p[2]);
    const double mv = std::sqrt(G * center_mass / r);
    const double u[3] = {p[1] / r * mv, -p[0] / r * mv, 0.0};
    mass[1 + k] = s.u_open0() * avr_mass;
    const double pc[3] = {p[0] + center[0], p[1] + center[1], p[2] + center[2]};
    const double uc[3] = {u[0] + v[0], u[1] + v[1], u[2] + v[2]};
    orc_rotate_point(pc, rotation[0], rotation[2], pos + 3 * (1 + k));
    orc_rotate_point(uc, rotation[0], rotation[2], vel + 3 * (1 + k));
  }
  return 0;
}

}  // extern "C"
