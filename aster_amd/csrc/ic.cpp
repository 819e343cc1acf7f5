// Seeded initial conditions. The reference builds its presets with std::rand
// (src/clusters/presets_impl.cpp:70-108), which is neither seedable per call
// nor reproducible across machines; these generators replace them for the
// workloads this library is measured on.
//
// Random numbers: counter-based SplitMix64 streams, one per body or system
// (icmath.hpp). The Plummer sphere and the cube draw each body with the
// functions the device cluster loader uses (icgen.hpp) and reduce in the
// device's order, so aster_ic_plummer / aster_ic_cube and a cluster of the
// same kind and seed loaded on the device hold the same bits.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/aster_hip.h"
#include "icgen.hpp"

namespace {

using aster::icm::Stream;
constexpr double kPi = aster::icm::kPi;

// Sequential barycentric recentring of one small system.
void recentre(int64_t n, double* pos, double* vel, const double* mass) {
  double M = 0.0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    M += mass[i];
    for (int c = 0; c < 3; ++c) {
      cx[c] += mass[i] * pos[3 * i + c];
      cv[c] += mass[i] * vel[3 * i + c];
    }
  }
  for (int c = 0; c < 3; ++c) {
    cx[c] /= M;
    cv[c] /= M;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] -= cx[c];
      vel[3 * i + c] -= cv[c];
    }
}

}  // namespace

extern "C" {

int aster_ic_plummer(int64_t n, uint64_t seed, double mtot, double b, double rmax, int mass_mode, double* pos,
                     double* vel, double* mass) {
  if (n < 1 || !pos || !mass || !(mtot > 0) || !(b > 0) || !(rmax > 0) || (mass_mode != 0 && mass_mode != 1))
    return ASTER_ERR_INVALID;
  std::vector<double> vtmp;
  if (!vel) {
    vtmp.resize((size_t)3 * n);
    vel = vtmp.data();
  }
  const double vscale = std::sqrt(1.0 * mtot / b);
  for (int64_t i = 0; i < n; ++i)
    aster::icg::plummer_body(seed, (uint64_t)i, b, rmax, vscale, mass_mode, pos + 3 * i, vel + 3 * i, mass + i);
  if (mass_mode == 1) {
    double msum;
    aster::icg::ordered_sums(n, 1, [&](int64_t i, double* v) { v[0] = mass[i]; }, &msum);
    for (int64_t i = 0; i < n; ++i) mass[i] = mass[i] * (mtot / msum);
  } else {
    for (int64_t i = 0; i < n; ++i) mass[i] = mtot / (double)n;
  }
  // zero centre of mass and momentum
  double S[7];
  aster::icg::ordered_sums(
      n, 7,
      [&](int64_t i, double* v) {
        v[0] = mass[i];
        for (int c = 0; c < 3; ++c) {
          v[1 + c] = mass[i] * pos[3 * i + c];
          v[4 + c] = mass[i] * vel[3 * i + c];
        }
      },
      S);
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] = pos[3 * i + c] - S[1 + c] / S[0];
      vel[3 * i + c] = vel[3 * i + c] - S[4 + c] / S[0];
    }
  return ASTER_OK;
}

int aster_ic_cube(int64_t n, uint64_t seed, double h, double mtot, double* pos, double* vel, double* mass) {
  if (n < 1 || !pos || !mass || !(h > 0) || !(mtot > 0)) return ASTER_ERR_INVALID;
  double vz[3];
  for (int64_t i = 0; i < n; ++i) {
    aster::icg::cube_body(seed, (uint64_t)i, h, pos + 3 * i, vel ? vel + 3 * i : vz);
    mass[i] = mtot / (double)n;
  }
  return ASTER_OK;
}

int aster_ic_ensemble(int64_t nsys, uint64_t seed, int32_t nmin, int32_t nmax, int64_t* offsets, int64_t* total,
                      double* pos, double* vel, double* mass) {
  if (nsys < 1 || nmin < 1 || nmax < nmin || nmax > 64 || !offsets) return ASTER_ERR_INVALID;
  offsets[0] = 0;
  for (int64_t s = 0; s < nsys; ++s) {
    Stream st(seed, (uint64_t)s);
    const int32_t span = nmax - nmin + 1;
    int32_t ns = nmin + (int32_t)(st.u() * span);
    if (ns > nmax) ns = nmax;
    offsets[s + 1] = offsets[s] + ns;
  }
  if (total) *total = offsets[nsys];
  if (!pos) return ASTER_OK;
  if (!vel || !mass) return ASTER_ERR_INVALID;
  const double deg = kPi / 180.0;
  for (int64_t s = 0; s < nsys; ++s) {
    Stream st(seed, (uint64_t)s);
    (void)st.u();  // size draw
    const int64_t o = offsets[s], ns = offsets[s + 1] - offsets[s];
    mass[o] = 1.0;
    for (int c = 0; c < 3; ++c) pos[3 * o + c] = vel[3 * o + c] = 0.0;
    for (int64_t k = 1; k < ns; ++k) {
      const int64_t i = o + k;
      const double m = aster::icm::exp((-7.0 + 4.0 * st.u()) * aster::icm::kLn10);
      const double a = 0.4 * aster::icm::pow(30.0 / 0.4, st.u());
      const double e = 0.3 * st.u();
      const double inc = 5.0 * deg * st.u();
      const double Om = 2.0 * kPi * st.u(), om = 2.0 * kPi * st.u(), M = 2.0 * kPi * st.u();
      // Kepler's equation E - e sin E = M by Newton iteration.
      double E = M;
      for (int it = 0; it < 50; ++it) {
        double sE, cE;
        aster::icm::sincos(E, &sE, &cE);
        const double f = E - e * sE - M;
        const double dE = f / (1.0 - e * cE);
        E -= dE;
        if (std::fabs(dE) < 1e-15) break;
      }
      const double mu = 1.0 + m;
      const double nmo = std::sqrt(mu / (a * a * a));
      double sE, cE;
      aster::icm::sincos(E, &sE, &cE);
      const double q = std::sqrt(1.0 - e * e);
      const double xo = a * (cE - e), yo = a * q * sE;
      const double den = 1.0 - e * cE;
      const double vxo = -a * nmo * sE / den, vyo = a * nmo * q * cE / den;
      // Rotation R_z(Om) R_x(inc) R_z(om).
      double sO, cO, si, ci, so, co;
      aster::icm::sincos(Om, &sO, &cO);
      aster::icm::sincos(inc, &si, &ci);
      aster::icm::sincos(om, &so, &co);
      const double r11 = cO * co - sO * so * ci, r12 = -cO * so - sO * co * ci;
      const double r21 = sO * co + cO * so * ci, r22 = -sO * so + cO * co * ci;
      const double r31 = so * si, r32 = co * si;
      pos[3 * i] = r11 * xo + r12 * yo;
      pos[3 * i + 1] = r21 * xo + r22 * yo;
      pos[3 * i + 2] = r31 * xo + r32 * yo;
      vel[3 * i] = r11 * vxo + r12 * vyo;
      vel[3 * i + 1] = r21 * vxo + r22 * vyo;
      vel[3 * i + 2] = r31 * vxo + r32 * vyo;
      mass[i] = m;
    }
    recentre(ns, pos + 3 * o, vel + 3 * o, mass + o);
  }
  return ASTER_OK;
}

int aster_ic_threebody(int64_t nsys, uint64_t seed, double* pos, double* vel, double* mass) {
  if (nsys < 1 || !pos || !vel || !mass) return ASTER_ERR_INVALID;
  for (int64_t s = 0; s < nsys; ++s) {
    Stream st(seed, (uint64_t)s);
    double M = 0.0, p[3] = {0, 0, 0};
    for (int k = 0; k < 3; ++k) {
      const int64_t i = 3 * s + k;
      mass[i] = 0.5 + st.u();
      for (int c = 0; c < 3; ++c) pos[3 * i + c] = st.u();
      for (int c = 0; c < 3; ++c) vel[3 * i + c] = st.u() - 0.5;
      M += mass[i];
      for (int c = 0; c < 3; ++c) p[c] += mass[i] * vel[3 * i + c];
    }
    for (int k = 0; k < 3; ++k)
      for (int c = 0; c < 3; ++c) vel[3 * (3 * s + k) + c] -= p[c] / M;
  }
  return ASTER_OK;
}

}  // extern "C"
