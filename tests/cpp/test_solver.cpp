// Drives the C++ Solver/updater shims of include/aster_hip_solver.hpp the way
// Aster's Simulation drives its solvers, and checks them bit for bit against
// the oracle (the CPU restatement of Aster's path). Exit code 0 = pass.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "aster_hip_solver.hpp"

extern "C" {
int orc_allpairs(int64_t, const double*, const double*, const double*, double, double, double, int, int64_t, int64_t,
                 double*, int);
int orc_box_auto(int64_t, const double*, double*);
int orc_bh(int64_t, const double*, const double*, const double*, const double*, double, double, double, double, int,
           double*, int64_t*, int);
int orc_integrate(int64_t, double*, double*, const double*, double, double, double, int, int, const double*, double,
                  int, double, int64_t, double*, int);
int orc_integrate_saba(int64_t, double*, double*, const double*, double, double, double, int, int, const double*,
                       double, int, double, int64_t, double*, int);
}

using namespace aster_hip;

static bool same(const std::vector<vec3>& a, const std::vector<double>& b) {
  return std::memcmp(a.data(), b.data(), b.size() * sizeof(double)) == 0;
}

// update_leapfrog (src/physics/tool_chain_impl.cpp:169-185) as Aster's host
// updater runs it: forces, then the drift and half kick on the host, forces
// at the drifted positions, the second half kick. The vec3 expressions
// evaluate left to right per component as the reference's operators do.
static void host_leapfrog(Solver& s, BodyArray& b, double dt, int steps) {
  const size_t n = b.positions.size();
  for (int st = 0; st < steps; ++st) {
    s.compute_forces();
    for (size_t i = 0; i < n; ++i) {
      double* x = &b.positions[i].x;
      double* v = &b.velocities[i].x;
      const double* a = &b.accs[i].x;
      for (int c = 0; c < 3; ++c) {
        x[c] = x[c] + (v[c] * dt + .5 * a[c] * dt * dt);
        v[c] = v[c] + .5 * a[c] * dt;
      }
    }
    s.compute_forces();
    for (size_t i = 0; i < n; ++i)
      for (int c = 0; c < 3; ++c) (&b.velocities[i].x)[c] = (&b.velocities[i].x)[c] + .5 * (&b.accs[i].x)[c] * dt;
  }
}

int main() {
  const int64_t n = 1024;
  BodyArray b;
  b.positions.resize(n);
  b.velocities.resize(n);
  b.masses.resize(n);
  if (aster_ic_plummer(n, 7, 1.0, 1.0, 10.0, 0, &b.positions[0].x, &b.velocities[0].x, b.masses.data()) != 0) return 2;
  Meta meta;
  meta.G = 1.0;
  meta.softening = 1e-4;
  meta.theta = 0.5;
  meta.dt = (double)(float)1e-3;  // Simulation::set_dt stores dt through float
  std::vector<double> ref(3 * n);
  const double* p0 = &b.positions[0].x;
  const double* v0 = &b.velocities[0].x;
  int fails = 0;
  try {
    HipAllPairs ap(&b, &meta, ASTER_ALLPAIRS_FP64);
    ap.load();
    ap.compute_forces();
    orc_allpairs(n, p0, v0, b.masses.data(), 1.0, meta.c, 1e-4, ASTER_NEWTON, 0, n, ref.data(), 0);
    if (!same(b.accs, ref)) { std::printf("all-pairs FP64 differs\n"); ++fails; }
    ap.set_bounds(100, 700);
    ap.compute_forces();
    orc_allpairs(n, p0, v0, b.masses.data(), 1.0, meta.c, 1e-4, ASTER_NEWTON, 100, 700, ref.data(), 0);
    if (!same(b.accs, ref)) { std::printf("all-pairs FP64 with bounds differs\n"); ++fails; }

    HipBarnesHut bh(&b, &meta, ASTER_BH_PARITY, ASTER_FP64);
    bh.load();
    bh.compute_forces();
    double box[3];
    orc_box_auto(n, p0, box);
    orc_bh(n, p0, v0, b.masses.data(), box, 0.5, 1.0, meta.c, 1e-4, ASTER_NEWTON, ref.data(), nullptr, 0);
    if (!same(b.accs, ref)) { std::printf("Barnes-Hut FP64 differs\n"); ++fails; }

    std::vector<double> px(p0, p0 + 3 * n), vx(v0, v0 + 3 * n);
    HipAllPairs ap2(&b, &meta, ASTER_ALLPAIRS_FP64);
    ap2.load();
    update_leapfrog_hip(ap2, meta, 5);
    orc_integrate(n, px.data(), vx.data(), b.masses.data(), 1.0, meta.c, 1e-4, ASTER_NEWTON, 0, box, 0.5, 2, meta.dt,
                  5, nullptr, 0);
    if (!same(b.positions, px) || !same(b.velocities, vx)) { std::printf("leapfrog differs\n"); ++fails; }

    update_saba_hip(ap2, meta, 4, 2);
    orc_integrate_saba(n, px.data(), vx.data(), b.masses.data(), 1.0, meta.c, 1e-4, ASTER_NEWTON, 0, box, 0.5, 4,
                       meta.dt, 2, nullptr, 0);
    if (!same(b.positions, px) || !same(b.velocities, vx)) { std::printf("SABA4 differs\n"); ++fails; }

    // Aster's host updater over the HIP solvers: the host drifts and kicks,
    // compute_forces uploads the moved bodies each time. Bit-identical to
    // update_leapfrog over Parallelized and Barnes_Hut (automatic box).
    const double zero_box[3] = {0.0, 0.0, 0.0};
    for (int kind = 0; kind < 2; ++kind) {
      BodyArray hb;
      hb.positions.assign(b.positions.begin(), b.positions.end());
      hb.velocities.assign(b.velocities.begin(), b.velocities.end());
      hb.masses = b.masses;
      std::vector<double> rx(&hb.positions[0].x, &hb.positions[0].x + 3 * n);
      std::vector<double> rv(&hb.velocities[0].x, &hb.velocities[0].x + 3 * n);
      if (kind == 0) {
        HipAllPairs s(&hb, &meta, ASTER_ALLPAIRS_FP64);
        s.load();
        host_leapfrog(s, hb, meta.dt, 5);
      } else {
        HipBarnesHut s(&hb, &meta, ASTER_BH_PARITY, ASTER_FP64);
        s.load();
        host_leapfrog(s, hb, meta.dt, 5);
      }
      orc_integrate(n, rx.data(), rv.data(), hb.masses.data(), 1.0, meta.c, 1e-4, ASTER_NEWTON, kind, zero_box, 0.5,
                    2, meta.dt, 5, nullptr, 0);
      if (!same(hb.positions, rx) || !same(hb.velocities, rv)) {
        std::printf("host leapfrog over %s differs\n", kind == 0 ? "HipAllPairs(FP64)" : "HipBarnesHut(PARITY, FP64)");
        ++fails;
      }
    }

    // Barnes-Hut with bounds: the tree holds bodies [300, 800) and only they
    // get new accelerations (Barnes_Hut::compute_forces).
    {
      HipBarnesHut s(&b, &meta, ASTER_BH_PARITY, ASTER_FP64);
      s.load();
      b.accs.assign(n, vec3{-7.0, -7.0, -7.0});
      s.set_bounds(300, 800);
      s.compute_forces();
      const int64_t m = 500;
      std::vector<double> subacc(3 * m), sbox(3);
      orc_box_auto(m, &b.positions[300].x, sbox.data());
      orc_bh(m, &b.positions[300].x, &b.velocities[300].x, b.masses.data() + 300, sbox.data(), 0.5, 1.0, meta.c, 1e-4,
             ASTER_NEWTON, subacc.data(), nullptr, 0);
      bool ok = std::memcmp(&b.accs[300].x, subacc.data(), 3 * m * sizeof(double)) == 0;
      for (int64_t i = 0; i < n && ok; ++i)
        if ((i < 300 || i >= 800) && (b.accs[i].x != -7.0 || b.accs[i].y != -7.0 || b.accs[i].z != -7.0)) ok = false;
      if (!ok) { std::printf("Barnes-Hut with bounds differs\n"); ++fails; }
    }
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 3;
  }
  std::printf(fails ? "cpp shim: FAIL\n" : "cpp shim: ok\n");
  return fails ? 1 : 0;
}
