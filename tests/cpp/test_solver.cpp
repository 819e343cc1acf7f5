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
  } catch (const std::exception& e) {
    std::printf("exception: %s\n", e.what());
    return 3;
  }
  std::printf(fails ? "cpp shim: FAIL\n" : "cpp shim: ok\n");
  return fails ? 1 : 0;
}
