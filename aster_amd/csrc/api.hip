// C ABI of libaster_hip: context lifetime, state transfer, dispatch of force
// methods and integrators, diagnostics and timers.
//
// Replaces the reference's OpenCL glue (src/utils/tc_impl_cl.cpp,
// src/utils/kernel_uploaders.cpp) and the global device state of
// include/Aster/building-api/GPU_endpoint.h:25-31 with one context object per
// stream; errors are returned, never exit()ed.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "common.hpp"
#include "saba_coeffs.hpp"

namespace aster {

int fail(aster_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(aster_ctx* c, hipError_t e, const char* what) {
  if (c) c->err = std::string("HIP error: ") + hipGetErrorString(e) + " (" + what + ")";
  return ASTER_ERR_HIP;
}

static thread_local std::string g_create_err = "no error";

void set_global_error(const std::string& msg) { g_create_err = msg; }

static int check_flags(aster_ctx* c) {
  uint32_t f[4] = {0, 0, 0, 0};
  ASTER_HIP(c, hipMemcpyAsync(f, c->flags.p, sizeof(f), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  if (f[0] == 0) return ASTER_OK;
  int rc = launch_reset_flags(c);
  if (rc) return rc;
  if (f[0] & FLAG_TREE) return fail(c, ASTER_ERR_STATE, "tree traversal exceeded its bound (corrupt tree)");
  if (f[0] & FLAG_OUT_OF_BOX)
    return fail(c, ASTER_ERR_OUT_OF_BOX, "a body lies outside the Morton box (set aster_params.box or use the auto box)");
  return fail(c, ASTER_ERR_NONFINITE, "non-finite position or velocity produced");
}

static bool is_allpairs(int m) {
  return m == ASTER_ALLPAIRS_FP32 || m == ASTER_ALLPAIRS_FP32_KAHAN || m == ASTER_ALLPAIRS_FP64;
}
static bool is_bh(int m) { return m == ASTER_BH_PARITY || m == ASTER_BH_GEOMETRIC || m == ASTER_BH_GROUP; }

int compute_forces(aster_ctx* c, int method) {
  int rc;
  if (is_allpairs(method))
    rc = launch_allpairs(c, method);
  else if (is_bh(method))
    rc = bh_forces(c, method);
  else
    return fail(c, ASTER_ERR_INVALID, "unknown force method");
  if (rc) return rc;
  c->acc_valid = true;
  c->acc_stale = false;
  c->acc_method = method;
  return ASTER_OK;
}

// SoA plane on the device -> AoS-d3 host array (transposed on the device).
static int download3(aster_ctx* c, const double* dev, double* host_aos) {
  const int64_t n = c->n;
  ASTER_HIP(c, c->stage.reserve((size_t)7 * n));
  int rc = export_plane(c, dev, c->stage.p);
  if (rc) return rc;
  ASTER_HIP(c, hipMemcpyAsync(host_aos, c->stage.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return ASTER_OK;
}

}  // namespace aster

using namespace aster;

extern "C" {

int aster_version(void) { return 1 * 10000 + 0 * 100 + 0; }

int aster_create(aster_ctx** out, int device) {
  if (!out) {
    g_create_err = "out pointer is NULL";
    return ASTER_ERR_INVALID;
  }
  *out = nullptr;
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count < 1) {
    g_create_err = std::string("no HIP device: ") + hipGetErrorString(e);
    return ASTER_ERR_HIP;
  }
  if (device < 0 || device >= count) {
    g_create_err = "device index out of range";
    return ASTER_ERR_INVALID;
  }
  e = hipSetDevice(device);
  if (e != hipSuccess) {
    g_create_err = std::string("hipSetDevice: ") + hipGetErrorString(e);
    return ASTER_ERR_HIP;
  }
  aster_ctx* c = new aster_ctx();
  c->device = device;
  c->P.G = 6e-11;
  c->P.c = 299792458.0;
  c->P.softening = 1e-14;
  c->P.theta = 0.2;
  c->P.pn_flags = ASTER_NEWTON;
  c->P.precision = ASTER_FP32;
  c->P.lower = 0;
  c->P.upper = -1;
  c->P.split = 0;
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  for (int k = 0; k < 16 && e == hipSuccess; ++k) e = hipEventCreate(&c->ev[k]);
  if (e == hipSuccess) e = c->flags.reserve(4);
  if (e == hipSuccess) e = c->boxd.reserve(8);
  if (e == hipSuccess) e = hipMemsetAsync(c->flags.p, 0, 4 * sizeof(uint32_t), c->stream);
  if (e == hipSuccess) e = hipMemsetAsync(c->boxd.p, 0, 8 * sizeof(double), c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) {
    g_create_err = std::string("context setup: ") + hipGetErrorString(e);
    aster_destroy(c);
    return ASTER_ERR_HIP;
  }
  c->err = "no error";
  *out = c;
  return ASTER_OK;
}

int aster_destroy(aster_ctx* c) {
  if (!c) return ASTER_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->batch.exec) (void)hipGraphExecDestroy(c->batch.exec);
  if (c->batch.graph) (void)hipGraphDestroy(c->batch.graph);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->prof_pool)
    if (e) (void)hipEventDestroy(e);
  if (c->tree_ev) (void)hipEventDestroy(c->tree_ev);
  if (c->host_words) (void)hipHostFree(c->host_words);
  c->x.release(); c->v.release(); c->m.release(); c->a.release(); c->sh.release(); c->src64.release();
  c->apart.release(); c->apn.release(); c->ework.release(); c->eout.release(); c->flags.release();
  c->boxpart.release(); c->boxd.release(); c->stage.release();
  c->codes.release(); c->xm4.release(); c->idx.release(); c->codes2.release(); c->idx2.release(); c->hist.release();
  c->left.release(); c->right.release(); c->parent.release(); c->escape.release(); c->scanpart.release();
  c->depth.release(); c->nmc.release(); c->nvel.release(); c->nbox.release();
  c->levels.release(); c->bucket.release(); c->grange.release(); c->gscan.release();
  c->gstart.release(); c->gcount.release(); c->splitnode.release(); c->groot.release(); c->climb.release();
  c->dmax.release(); c->censusctr.release();
  c->nrec.release(); c->nlink.release(); c->kids.release(); c->topslot.release(); c->bh_counters.release(); c->gsh.release();
  c->sub_x.release(); c->sub_v.release(); c->sub_a.release(); c->sub_m.release();
  auto& Bt = c->batch;
  Bt.x.release(); Bt.v.release(); Bt.a.release(); Bt.m.release(); Bt.ebody.release(); Bt.esys.release();
  Bt.doff.release(); Bt.lane_body.release(); Bt.lane_seg.release(); Bt.wave_max.release();
  Bt.traj.release(); Bt.x0.release(); Bt.v0.release(); Bt.tgrad.release(); Bt.tout.release(); Bt.tsnap.release();
  auto& A = c->adj;
  A.ckpt_x.release(); A.ckpt_v.release(); A.lx.release(); A.lv.release(); A.lm.release(); A.w.release();
  A.snapg.release(); A.g.release(); A.wsrc.release(); A.part.release();
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return ASTER_OK;
}

const char* aster_last_error(const aster_ctx* c) { return c ? c->err.c_str() : g_create_err.c_str(); }

int aster_set_params(aster_ctx* c, const aster_params* p) {
  if (!c || !p) return ASTER_ERR_INVALID;
  if (!(p->softening >= 0.0) || !(p->theta >= 0.0) || !(p->c > 0.0) || !std::isfinite(p->G))
    return fail(c, ASTER_ERR_INVALID, "invalid G, c, softening or theta");
  if (p->precision != ASTER_FP32 && p->precision != ASTER_FP64)
    return fail(c, ASTER_ERR_INVALID, "precision must be ASTER_FP32 or ASTER_FP64");
  if ((p->pn_flags & ~(ASTER_NEWTON | ASTER_PN_TERMS)) != 0)
    return fail(c, ASTER_ERR_INVALID, "force terms are ASTER_NEWTON, ASTER_PN1, ASTER_PN2, ASTER_PN25 and ASTER_EIH");
  if ((p->pn_flags & ASTER_EIH) && (p->pn_flags & (ASTER_PN1 | ASTER_PN2 | ASTER_PN25 | ASTER_NEWTON)) != ASTER_NEWTON)
    return fail(c, ASTER_ERR_INVALID, "ASTER_EIH goes with ASTER_NEWTON alone (it replaces the pairwise terms)");
  const bool any = p->box[0] != 0.0 || p->box[1] != 0.0 || p->box[2] != 0.0;
  if (any && !(p->box[0] > 0.0 && p->box[1] > 0.0 && p->box[2] > 0.0))
    return fail(c, ASTER_ERR_INVALID, "box half-extents must all be > 0 (or all 0 for the automatic box)");
  if (p->split < 0 || p->split > 1024) return fail(c, ASTER_ERR_INVALID, "split out of range");
  ASTER_HIP(c, hipSetDevice(c->device));
  c->P = *p;
  if (any) {
    double ub[3] = {p->box[0], p->box[1], p->box[2]};
    ASTER_HIP(c, hipMemcpyAsync(c->boxd.p + 4, ub, sizeof(ub), hipMemcpyHostToDevice, c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
  }
  c->acc_valid = false;
  c->tree_valid = false;
  batch_params_changed(c);
  return ASTER_OK;
}

int aster_get_params(const aster_ctx* c, aster_params* p) {
  if (!c || !p) return ASTER_ERR_INVALID;
  *p = c->P;
  return ASTER_OK;
}

int aster_set_state(aster_ctx* c, int64_t n, const double* pos, const double* vel, const double* mass) {
  if (!c) return ASTER_ERR_INVALID;
  if (n < 1 || !pos || !mass) return fail(c, ASTER_ERR_INVALID, "need n >= 1, positions and masses");
  if (n > (int64_t(1) << 28)) return fail(c, ASTER_ERR_CAPACITY, "at most 2^28 bodies per context");
  ASTER_HIP(c, hipSetDevice(c->device));
  // the host arrays go over as they are (AoS-d3) and are transposed and
  // checked on the device
  ASTER_HIP(c, c->stage.reserve((size_t)7 * n));
  double* sp = c->stage.p;
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  ASTER_HIP(c, hipMemcpyAsync(sp, pos, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  if (vel) ASTER_HIP(c, hipMemcpyAsync(sp + 3 * n, vel, 3 * n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  ASTER_HIP(c, hipMemcpyAsync(sp + 6 * n, mass, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return import_state(c, n, sp, vel ? sp + 3 * n : nullptr, sp + 6 * n);
}

int aster_set_state_device(aster_ctx* c, int64_t n, const double* d_pos, const double* d_vel, const double* d_mass) {
  if (!c) return ASTER_ERR_INVALID;
  if (n < 1 || !d_pos || !d_mass) return fail(c, ASTER_ERR_INVALID, "need n >= 1, positions and masses");
  if (n > (int64_t(1) << 28)) return fail(c, ASTER_ERR_CAPACITY, "at most 2^28 bodies per context");
  ASTER_HIP(c, hipSetDevice(c->device));
  return import_state(c, n, d_pos, d_vel, d_mass);
}

int aster_get_state(aster_ctx* c, double* pos, double* vel) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  const int64_t n = c->n;
  ASTER_HIP(c, c->stage.reserve((size_t)7 * n));
  int rc;
  if (pos) {
    if ((rc = export_plane(c, c->x.p, c->stage.p))) return rc;
    ASTER_HIP(c, hipMemcpyAsync(pos, c->stage.p, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  if (vel) {
    if ((rc = export_plane(c, c->v.p, c->stage.p + 3 * n))) return rc;
    ASTER_HIP(c, hipMemcpyAsync(vel, c->stage.p + 3 * n, 3 * n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  }
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int aster_get_state_device(aster_ctx* c, double* d_pos, double* d_vel) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc;
  if (d_pos && (rc = export_plane(c, c->x.p, d_pos))) return rc;
  if (d_vel && (rc = export_plane(c, c->v.p, d_vel))) return rc;
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int aster_collect(aster_ctx* c, int kind, int64_t body, double* out) {
  if (!c || !out) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc;
  if (kind == ASTER_COLLECT_HAMILTONIAN || kind == ASTER_COLLECT_ERROR || kind == ASTER_COLLECT_ANGULAR_MOMENTUM) {
    double e_ref, L[3];
    if ((rc = aster_energy(c, &e_ref, nullptr, L))) return rc;
    if (kind == ASTER_COLLECT_HAMILTONIAN) {
      *out = e_ref;
    } else if (kind == ASTER_COLLECT_ANGULAR_MOMENTUM) {
      *out = std::sqrt(L[0] * L[0] + L[1] * L[1] + L[2] * L[2]);
    } else if (!c->e0_valid) {  // error_collector: the first call records E0 and reports 0
      c->e0 = e_ref;
      c->e0_valid = true;
      *out = 0.0;
    } else {
      *out = std::fabs(c->e0 - e_ref) / std::fabs(c->e0);
    }
    return ASTER_OK;
  }
  if (kind == ASTER_COLLECT_DISTANCE) {
    if (body < 0 || body >= c->n) return fail(c, ASTER_ERR_INVALID, "body index out of range");
    double com[3], p[3];
    if ((rc = center_of_mass(c, com))) return rc;
    for (int k = 0; k < 3; ++k)
      ASTER_HIP(c, hipMemcpyAsync(&p[k], c->x.p + (size_t)k * c->n + body, sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
    ASTER_HIP(c, hipStreamSynchronize(c->stream));
    const double dx = com[0] - p[0] + 0.1, dy = com[1] - p[1] + 0.1, dz = com[2] - p[2] + 0.1;
    *out = std::sqrt(dx * dx + dy * dy + dz * dz);
    return ASTER_OK;
  }
  return fail(c, ASTER_ERR_INVALID, "unknown collector");
}

int aster_center_of_mass(aster_ctx* c, double* com) {
  if (!c || !com) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  return center_of_mass(c, com);
}

int64_t aster_num_bodies(const aster_ctx* c) { return c ? c->n : 0; }

int aster_forces(aster_ctx* c, int method, double* acc) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc = compute_forces(c, method);
  if (rc) return rc;
  if (acc) {
    if ((rc = download3(c, c->a.p, acc))) return rc;
    return check_flags(c);
  }
  return ASTER_OK;
}

int aster_get_acc(aster_ctx* c, double* acc) {
  if (!c || !acc) return ASTER_ERR_INVALID;
  if (!c->acc_valid) return fail(c, ASTER_ERR_STATE, "no accelerations computed yet");
  ASTER_HIP(c, hipSetDevice(c->device));
  return download3(c, c->a.p, acc);
}

int aster_yoshida_weights(double* w) {
  if (!w) return ASTER_ERR_INVALID;
  // Yoshida (1990): w1 = 1 / (2 - 2^(1/3)), w0 = 1 - 2 w1, evaluated in FP64
  // from the correctly rounded cube root 0x1.428a2f98d728bp+0. Literals, so
  // that no libm cbrt (whose last bit differs between toolchains) enters.
  w[0] = 0x1.59e8b6eb96339p+0;
  w[1] = -0x1.b3d16dd72c672p+0;
  w[2] = 0x1.59e8b6eb96339p+0;
  return ASTER_OK;
}

int aster_step(aster_ctx* c, int method, int integrator, double dt, int64_t nsteps) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  if (integrator == ASTER_EULER) return aster_integrate(c, method, integrator, 1, dt, nsteps);
  if (integrator == ASTER_SABA)
    return fail(c, ASTER_ERR_INVALID, "ASTER_SABA needs an order: call aster_integrate");
  if (integrator != ASTER_LEAPFROG && integrator != ASTER_YOSHIDA4)
    return fail(c, ASTER_ERR_INVALID, "integrator must be ASTER_EULER, ASTER_SABA, ASTER_LEAPFROG or ASTER_YOSHIDA4");
  if (!std::isfinite(dt) || nsteps < 0) return fail(c, ASTER_ERR_INVALID, "invalid dt or nsteps");
  if (nsteps == 0) return ASTER_OK;
  ASTER_HIP(c, hipSetDevice(c->device));
  double w[3] = {1.0, 0.0, 0.0};
  int nsub = 1;
  if (integrator == ASTER_YOSHIDA4) {
    aster_yoshida_weights(w);
    nsub = 3;
  }
  int rc;
  // With velocity-dependent (1PN) forces the reference's first force
  // evaluation of a step sees the fully kicked velocities, so it is not the
  // previous step's second evaluation: re-evaluate, as update_leapfrog does.
  // Newtonian forces depend on positions only and are reused.
  const bool vdep = (c->P.pn_flags & ASTER_PN_TERMS) != 0;
  if (!vdep && (!c->acc_valid || c->acc_stale || c->acc_method != method))
    if ((rc = compute_forces(c, method))) return rc;
  double pending = 0.0;
  for (int64_t s = 0; s < nsteps; ++s)
    for (int k = 0; k < nsub; ++k) {
      const double h = (integrator == ASTER_YOSHIDA4) ? w[k] * dt : dt;
      if (vdep) {
        if (pending != 0.0 && (rc = launch_kick(c, pending))) return rc;
        if ((rc = compute_forces(c, method))) return rc;
        pending = 0.0;
      }
      // Barnes-Hut: the drift also forms the first stage of the next tree's
      // automatic box, consumed by that force evaluation alone
      if ((rc = launch_kick_drift(c, pending, h, is_bh(method)))) return rc;
      rc = compute_forces(c, method);
      c->box_parts = 0;
      if (rc) return rc;
      pending = h;
    }
  if ((rc = launch_kick(c, pending))) return rc;
  c->adj.ready = false;
  return ASTER_OK;
}

int aster_integrate(aster_ctx* c, int method, int integrator, int order, double dt, int64_t nsteps) {
  if (!c) return ASTER_ERR_INVALID;
  if (integrator == ASTER_LEAPFROG || integrator == ASTER_YOSHIDA4)
    return aster_step(c, method, integrator, dt, nsteps);
  if (integrator != ASTER_EULER && integrator != ASTER_SABA)
    return fail(c, ASTER_ERR_INVALID, "integrator must be ASTER_EULER, ASTER_SABA, ASTER_LEAPFROG or ASTER_YOSHIDA4");
  if (integrator == ASTER_EULER) order = 1;
  if (order < 1 || order > kSabaMaxOrder) return fail(c, ASTER_ERR_INVALID, "SABA order must be 1..10");
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  if (!std::isfinite(dt) || nsteps < 0) return fail(c, ASTER_ERR_INVALID, "invalid dt or nsteps");
  if (nsteps == 0) return ASTER_OK;
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc;
  for (int64_t s = 0; s < nsteps; ++s) {
    if (order == 1) {
      if ((rc = compute_forces(c, method))) return rc;
      if ((rc = launch_saba_stage(c, true, 1.0, 1.0, dt))) return rc;
      continue;
    }
    const double* cs = kSabaC[order];
    const double* ds = kSabaD[order];
    if ((rc = launch_saba_stage(c, false, 0.0, cs[0], dt))) return rc;
    for (int st = 0; st < order; ++st) {
      if ((rc = compute_forces(c, method))) return rc;
      if ((rc = launch_saba_stage(c, true, ds[st], cs[st + 1], dt))) return rc;
    }
  }
  // every stage ends with a drift: the accelerations are those of the last
  // evaluation, not of the current positions
  c->acc_stale = true;
  c->adj.ready = false;
  return ASTER_OK;
}

int aster_saba_coeffs(int order, double* cs, double* ds) {
  if (order < 2 || order > kSabaMaxOrder || !cs || !ds) return ASTER_ERR_INVALID;
  for (int i = 0; i <= order; ++i) cs[i] = kSabaC[order][i];
  for (int i = 0; i < order; ++i) ds[i] = kSabaD[order][i];
  return ASTER_OK;
}

static int energy_ref(aster_ctx* c, double* e) {
  double s[6];
  int rc = launch_energy(c, s);
  if (rc) return rc;
  *e = s[0] - 0.5 * s[1];
  return check_flags(c);
}

// Shared by the two IntegrationReport entry points: halt == nullptr runs
// nsteps in one aster_integrate call; otherwise one step per check.
static int integrate_report(aster_ctx* c, int method, int integrator, int order, double dt, int64_t nsteps,
                            aster_halting_condition halt, void* user, int precision_test,
                            aster_integration_report* rep) {
  if (!c) return ASTER_ERR_INVALID;
  if (!rep) return fail(c, ASTER_ERR_INVALID, "report is NULL");
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  *rep = aster_integration_report{};
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc;
  double e0 = 0.0, e1 = 0.0;
  if (precision_test && (rc = energy_ref(c, &e0))) return rc;
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  const auto t0 = std::chrono::steady_clock::now();
  int64_t done = 0;
  if (!halt) {
    if (nsteps > 0) {
      if ((rc = aster_integrate(c, method, integrator, order, dt, nsteps))) return rc;
      done = nsteps;
    }
  } else {
    while ((nsteps < 0 || done < nsteps) && !halt(c, user)) {
      if ((rc = aster_integrate(c, method, integrator, order, dt, 1))) return rc;
      ASTER_HIP(c, hipStreamSynchronize(c->stream));
      ++done;
    }
  }
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  if (precision_test && (rc = energy_ref(c, &e1))) return rc;
  rep->total_time_ms = ms;
  rep->time_per_step_ms = done > 0 ? ms / (double)done : 0.0;
  rep->initial_energy = e0;
  rep->final_energy = e1;
  rep->precision = (precision_test && e0 != 0.0) ? std::fabs(e0 - e1) / std::fabs(e0) : 0.0;
  rep->tested_precision = precision_test ? 1 : 0;
  rep->steps = (int32_t)std::min<int64_t>(done, INT32_MAX);
  return check_flags(c);
}

int aster_integrate_report(aster_ctx* c, int method, int integrator, int order, double dt, int64_t nsteps,
                           int precision_test, aster_integration_report* report) {
  if (c && nsteps < 0) return fail(c, ASTER_ERR_INVALID, "nsteps must be >= 0");
  return integrate_report(c, method, integrator, order, dt, nsteps, nullptr, nullptr, precision_test, report);
}

int aster_integrate_until(aster_ctx* c, int method, int integrator, int order, double dt,
                          aster_halting_condition halt, void* user, int64_t max_steps, int precision_test,
                          aster_integration_report* report) {
  if (c && !halt) return fail(c, ASTER_ERR_INVALID, "halting condition is NULL");
  return integrate_report(c, method, integrator, order, dt, max_steps, halt, user, precision_test, report);
}

int aster_energy(aster_ctx* c, double* e_ref, double* e_plummer, double* L) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  double s[6];
  int rc = launch_energy(c, s);
  if (rc) return rc;
  if (e_ref) *e_ref = s[0] - 0.5 * s[1];
  if (e_plummer) *e_plummer = s[0] - 0.5 * s[2];
  if (L) {
    L[0] = s[3];
    L[1] = s[4];
    L[2] = s[5];
  }
  return check_flags(c);
}

int aster_tree_dump(aster_ctx* c, uint32_t* codes, uint32_t* perm, int32_t* left, int32_t* right, double* node_mass,
                    double* node_com) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->tree_valid) return fail(c, ASTER_ERR_STATE, "no Barnes-Hut tree built yet");
  ASTER_HIP(c, hipSetDevice(c->device));
  const int64_t n = c->n, nn = 2 * n - 1;
  auto cp = [&](void* dst, const void* src, size_t bytes) -> int {
    ASTER_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return ASTER_OK;
  };
  int rc = 0;
  if (codes && (rc = cp(codes, c->codes.p, n * sizeof(uint32_t)))) return rc;
  if (perm && (rc = cp(perm, c->idx.p, n * sizeof(uint32_t)))) return rc;
  if (left && (rc = cp(left, c->left.p, nn * sizeof(int32_t)))) return rc;
  if (right && (rc = cp(right, c->right.p, nn * sizeof(int32_t)))) return rc;
  std::vector<double4> mc;
  if (node_mass || node_com) {
    mc.resize((size_t)nn);
    if ((rc = cp(mc.data(), c->nmc.p, nn * sizeof(double4)))) return rc;
  }
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  for (int64_t k = 0; k < (int64_t)mc.size(); ++k) {
    if (node_mass) node_mass[k] = mc[k].w;
    if (node_com) {
      node_com[3 * k] = mc[k].x;
      node_com[3 * k + 1] = mc[k].y;
      node_com[3 * k + 2] = mc[k].z;
    }
  }
  return ASTER_OK;
}

int aster_morton(aster_ctx* c, uint32_t* codes, double* box_out) {
  if (!c || !codes) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  int rc = bh_morton_only(c);
  if (rc) return rc;
  ASTER_HIP(c, hipMemcpyAsync(codes, c->codes.p, c->n * sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
  double b[4];
  ASTER_HIP(c, hipMemcpyAsync(b, c->boxd.p, sizeof(b), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  if (box_out) {
    box_out[0] = b[0];
    box_out[1] = b[1];
    box_out[2] = b[2];
  }
  c->tree_valid = false;  // the sort buffers now hold unsorted codes
  return check_flags(c);
}

int aster_bh_stats(aster_ctx* c, double* interactions, double* node_visits) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->bh_counters.p) return fail(c, ASTER_ERR_STATE, "no Barnes-Hut evaluation yet");
  ASTER_HIP(c, hipSetDevice(c->device));
  unsigned long long v[BHC_COUNT];
  ASTER_HIP(c, hipMemcpyAsync(v, c->bh_counters.p, sizeof(v), hipMemcpyDeviceToHost, c->stream));
  ASTER_HIP(c, hipMemsetAsync(c->bh_counters.p, 0, sizeof(v), c->stream));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  if (interactions) *interactions = (double)v[BHC_INTERACTIONS];
  if (node_visits) *node_visits = (double)v[BHC_VISITS];
  return ASTER_OK;
}

int aster_sync(aster_ctx* c) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int aster_timer_mark(aster_ctx* c, int slot) {
  if (!c || slot < 0 || slot >= 16) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipEventRecord(c->ev[slot], c->stream));
  return ASTER_OK;
}

int aster_timer_elapsed(aster_ctx* c, int a, int b, float* ms) {
  if (!c || !ms || a < 0 || a >= 16 || b < 0 || b >= 16) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipEventSynchronize(c->ev[b]));
  ASTER_HIP(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return ASTER_OK;
}

int aster_batch_set(aster_ctx* c, int64_t nsys, const int64_t* offsets, const double* pos, const double* vel,
                    const double* mass);
int aster_batch_get(aster_ctx* c, double* pos, double* vel);

int aster_batch_step(aster_ctx* c, int integrator, double dt, int64_t nsteps, int32_t spl, int32_t use_graph) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return batch_step(c, integrator, dt, nsteps, spl, use_graph);
}

int aster_batch_energy(aster_ctx* c, double* e, double* L) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return batch_energy(c, e, L);
}

int aster_batch_info(aster_ctx* c, int64_t* pair_interactions, int64_t* lane_slots, int32_t* nwaves) {
  if (!c) return ASTER_ERR_INVALID;
  const BatchState& B = c->batch;
  if (B.nsys == 0) return fail(c, ASTER_ERR_STATE, "no batch loaded");
  int64_t pairs = 0;
  for (int64_t s = 0; s < B.nsys; ++s) {
    const int64_t k = B.offsets[s + 1] - B.offsets[s];
    pairs += k * (k - 1);
  }
  if (pair_interactions) *pair_interactions = pairs;
  if (lane_slots) *lane_slots = B.lane_slots;
  if (nwaves) *nwaves = B.nwaves;
  return ASTER_OK;
}

int aster_batch_autotune(aster_ctx* c, int integrator, double dt, int32_t* spb, int32_t* spl) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return batch_autotune(c, integrator, dt, spb, spl);
}

int aster_traj_forward(aster_ctx* c, double dt, int64_t nsteps, int64_t ckpt_every, int32_t nsnap,
                       const int64_t* snap_steps, double* snap_pos) {
  if (!c) return ASTER_ERR_INVALID;
  if (!c->have_state) return fail(c, ASTER_ERR_STATE, "no state loaded");
  ASTER_HIP(c, hipSetDevice(c->device));
  return traj_forward(c, dt, nsteps, ckpt_every, nsnap, snap_steps, snap_pos);
}

int aster_traj_backward(aster_ctx* c, const double* grad_snap_pos, double* g_pos, double* g_vel, double* g_mass) {
  if (!c) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  return traj_backward(c, grad_snap_pos, g_pos, g_vel, g_mass);
}

}  // extern "C"

namespace aster {

constexpr size_t kProfPairs = 8192;

void phase_begin(aster_ctx* c, int phase) {
  if (!c->profiling || c->prof_used >= kProfPairs) return;
  if (c->prof_pool.size() < 2 * (c->prof_used + 1)) {
    hipEvent_t a = nullptr, b = nullptr;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return;
    c->prof_pool.push_back(a);
    c->prof_pool.push_back(b);
  }
  if (hipEventRecord(c->prof_pool[2 * c->prof_used], c->stream) != hipSuccess) return;
  c->prof_open = phase;
}

void phase_end(aster_ctx* c, int phase) {
  if (!c->profiling || c->prof_open != phase) return;
  if (hipEventRecord(c->prof_pool[2 * c->prof_used + 1], c->stream) != hipSuccess) return;
  if (c->prof_phase.size() <= c->prof_used) c->prof_phase.resize(c->prof_used + 1);
  c->prof_phase[c->prof_used] = phase;
  c->prof_used++;
  c->prof_open = -1;
}

}  // namespace aster

extern "C" {

int aster_profile(aster_ctx* c, int enable) {
  if (!c) return ASTER_ERR_INVALID;
  c->profiling = enable != 0;
  c->prof_used = 0;
  c->prof_open = -1;
  return ASTER_OK;
}

int aster_phase_times(aster_ctx* c, double* ms, int64_t* counts) {
  if (!c || !ms) return ASTER_ERR_INVALID;
  ASTER_HIP(c, hipSetDevice(c->device));
  ASTER_HIP(c, hipStreamSynchronize(c->stream));
  for (int k = 0; k < ASTER_NUM_PHASES; ++k) {
    ms[k] = 0.0;
    if (counts) counts[k] = 0;
  }
  for (size_t i = 0; i < c->prof_used; ++i) {
    float t = 0.f;
    ASTER_HIP(c, hipEventElapsedTime(&t, c->prof_pool[2 * i], c->prof_pool[2 * i + 1]));
    const int ph = c->prof_phase[i];
    ms[ph] += t;
    if (counts) counts[ph] += 1;
  }
  c->prof_used = 0;
  return ASTER_OK;
}

}  // extern "C"
