// Seeded initial conditions. The reference builds its presets with std::rand
// (src/clusters/presets_impl.cpp:70-108), which is neither seedable per call
// nor reproducible across machines; these generators replace them for the
// workloads this library is measured on.
//
// Random numbers: counter-based SplitMix64. Stream key for body/system i is
// mix64(seed ^ mix64(i + C1)); draw k of that stream is
// mix64(key + (k + 1) * 0x9E3779B97F4A7C15) >> 11 scaled by 2^-53, a double in
// [0, 1). Every body is generated independently of the others, so any host
// (or a future device generator) reproduces identical inputs from the seed.
#include <cmath>
#include <cstdint>
#include <vector>

#include "../../include/aster_hip.h"

namespace {

constexpr double kPi = 3.14159265358979323846;

inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

struct Stream {
  uint64_t key, k = 0;
  Stream(uint64_t seed, uint64_t i) : key(mix64(seed ^ mix64(i + 0x632BE59BD9B4E019ull))) {}
  double u() {
    ++k;
    return (double)(mix64(key + k * 0x9E3779B97F4A7C15ull) >> 11) * 0x1.0p-53;
  }
};

void recentre(int64_t n, double* pos, double* vel, const double* mass) {
  double M = 0.0, cx[3] = {0, 0, 0}, cv[3] = {0, 0, 0};
  for (int64_t i = 0; i < n; ++i) {
    M += mass[i];
    for (int c = 0; c < 3; ++c) {
      cx[c] += mass[i] * pos[3 * i + c];
      if (vel) cv[c] += mass[i] * vel[3 * i + c];
    }
  }
  for (int c = 0; c < 3; ++c) {
    cx[c] /= M;
    cv[c] /= M;
  }
  for (int64_t i = 0; i < n; ++i)
    for (int c = 0; c < 3; ++c) {
      pos[3 * i + c] -= cx[c];
      if (vel) vel[3 * i + c] -= cv[c];
    }
}

void iso(Stream& s, double r, double* out) {
  const double ct = 1.0 - 2.0 * s.u();
  const double st = std::sqrt(std::fmax(0.0, 1.0 - ct * ct));
  const double ph = 2.0 * kPi * s.u();
  out[0] = r * st * std::cos(ph);
  out[1] = r * st * std::sin(ph);
  out[2] = r * ct;
}

}  // namespace

extern "C" {

int aster_ic_plummer(int64_t n, uint64_t seed, double mtot, double b, double rmax, int mass_mode, double* pos,
                     double* vel, double* mass) {
  if (n < 1 || !pos || !mass || !(mtot > 0) || !(b > 0) || !(rmax > 0) || (mass_mode != 0 && mass_mode != 1))
    return ASTER_ERR_INVALID;
  const double vscale = std::sqrt(mtot / b);
  for (int64_t i = 0; i < n; ++i) {
    Stream s(seed, (uint64_t)i);
    double r = 0.0;
    for (int tries = 0; tries < 1000; ++tries) {
      const double x1 = s.u();
      if (x1 <= 0.0) continue;
      r = 1.0 / std::sqrt(std::pow(x1, -2.0 / 3.0) - 1.0);
      if (r <= rmax) break;
      r = rmax;  // only reached if all tries fail (probability ~ 0)
    }
    iso(s, r * b, pos + 3 * i);
    double q = 0.0;
    for (int tries = 0; tries < 1000; ++tries) {
      const double x4 = s.u(), x5 = s.u();
      const double g = x4 * x4 * std::pow(1.0 - x4 * x4, 3.5);
      q = x4;
      if (0.1 * x5 < g) break;
    }
    const double ve = std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25) * vscale;
    if (vel) iso(s, q * ve, vel + 3 * i);
    mass[i] = (mass_mode == 0) ? 1.0 : std::pow(10.0, 2.0 * s.u());
  }
  double msum = 0.0;
  for (int64_t i = 0; i < n; ++i) msum += mass[i];
  for (int64_t i = 0; i < n; ++i) mass[i] = (mass_mode == 0) ? mtot / (double)n : mass[i] * (mtot / msum);
  recentre(n, pos, vel, mass);
  return ASTER_OK;
}

int aster_ic_cube(int64_t n, uint64_t seed, double h, double mtot, double* pos, double* vel, double* mass) {
  if (n < 1 || !pos || !mass || !(h > 0) || !(mtot > 0)) return ASTER_ERR_INVALID;
  for (int64_t i = 0; i < n; ++i) {
    Stream s(seed, (uint64_t)i);
    for (int c = 0; c < 3; ++c) pos[3 * i + c] = -h + 2.0 * h * s.u();
    if (vel)
      for (int c = 0; c < 3; ++c) vel[3 * i + c] = 0.0;
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
      const double m = std::pow(10.0, -7.0 + 4.0 * st.u());
      const double a = 0.4 * std::pow(30.0 / 0.4, st.u());
      const double e = 0.3 * st.u();
      const double inc = 5.0 * deg * st.u();
      const double Om = 2.0 * kPi * st.u(), om = 2.0 * kPi * st.u(), M = 2.0 * kPi * st.u();
      // Kepler's equation E - e sin E = M by Newton iteration.
      double E = M;
      for (int it = 0; it < 50; ++it) {
        const double f = E - e * std::sin(E) - M;
        const double dE = f / (1.0 - e * std::cos(E));
        E -= dE;
        if (std::fabs(dE) < 1e-15) break;
      }
      const double mu = 1.0 + m;
      const double nmo = std::sqrt(mu / (a * a * a));
      const double cE = std::cos(E), sE = std::sin(E), q = std::sqrt(1.0 - e * e);
      const double xo = a * (cE - e), yo = a * q * sE;
      const double den = 1.0 - e * cE;
      const double vxo = -a * nmo * sE / den, vyo = a * nmo * q * cE / den;
      // Rotation R_z(Om) R_x(inc) R_z(om).
      const double cO = std::cos(Om), sO = std::sin(Om), ci = std::cos(inc), si = std::sin(inc);
      const double co = std::cos(om), so = std::sin(om);
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
