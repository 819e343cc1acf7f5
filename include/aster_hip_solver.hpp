// aster_hip_solver.hpp -- header-only C++ shims that put the MI355X library
// behind Aster's solver/integrator interface.
//
// Aster's contract (include/Aster/simulations/sim_obj.h:34-77):
//   class Solver { virtual void load(); virtual void compute_forces();
//                  void set_bounds(int lower, int upper = -1);
//                  PN_expansion_type force_law; ... };
//   postcondition of compute_forces(): bodies.accs[i] holds the acceleration
//   of body i; Updater::update_bodies() advances one step.
// These classes keep that contract over a BodyArray-layout struct (masses,
// positions, velocities, accs as in include/Aster/physics/body.h:9-24, vec3 =
// three packed doubles as in include/Aster/physics/vectors.h:44-48) and call
// the C ABI of include/aster_hip.h. Errors throw std::runtime_error with the
// library's message instead of exit(-1) (Check(), basic.h:25-29).
//
// Inside Aster, HipAllPairs/HipBarnesHut derive from Aster::Solver instead of
// aster_hip::Solver and read Simulation::bodies in place of the BodyArray
// struct below; the bodies of load() and compute_forces() stay as written.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "aster_hip.h"

namespace aster_hip {

struct vec3 {
  double x = 0, y = 0, z = 0;
};
static_assert(sizeof(vec3) == 24, "vec3 must be three packed doubles like Aster::vec3");

// Same members and layout as Aster::BodyArray.
struct BodyArray {
  std::vector<double> masses, temps;
  std::vector<vec3> positions, velocities, accs;
};

// Physical parameters a Simulation carries (sim_meta.h:23-24, sim_obj.h:180).
struct Meta {
  double G = 6e-11;
  double c = 299792458.0;
  double softening = 1e-14;  // eps^2
  double theta = 0.2;
  double dt = 1.0;
  int force_law = ASTER_NEWTON;  // NEWTON | PN1 (basic.h:39-40)
};

inline void check(aster_ctx* ctx, int rc) {
  if (rc != ASTER_OK) throw std::runtime_error(std::string("aster_hip: ") + aster_last_error(ctx));
}

// Owns one device context. Which copy of the bodies is current:
//  - compute_forces() uploads the host bodies before it evaluates, because
//    Aster's host updaters (update_leapfrog, tool_chain_impl.cpp:169-185)
//    move positions and velocities on the host between evaluations; the one
//    exception is a device updater that ran without reading the state back,
//    after which the device copy is the current one;
//  - the device updaters (update_*_hip below) advance the device copy and
//    read it back; they upload first only after load() or mark_dirty() (the
//    host bodies were edited).
class Solver {
 public:
  Solver(BodyArray* bodies, Meta* meta, int method, int device = 0) : b_(bodies), meta_(meta), method_(method) {
    aster_ctx* c = nullptr;
    if (aster_create(&c, device) != ASTER_OK) throw std::runtime_error(aster_last_error(nullptr));
    ctx_ = c;
  }
  virtual ~Solver() { aster_destroy(ctx_); }
  Solver(const Solver&) = delete;
  Solver& operator=(const Solver&) = delete;

  // Solver::set_bounds (builder_endpoint.cpp:64-67): the source range of the
  // all-pairs methods (Parallelized::compute_forces); the body range of
  // Barnes-Hut, whose tree holds those bodies and whose evaluation writes
  // their accelerations only (Barnes_Hut::compute_forces,
  // barnes_hut.cpp:309-322,414-417).
  void set_bounds(int lower, int upper = -1) {
    lower_ = lower;
    upper_ = upper;
    params();
  }
  int force_law = ASTER_NEWTON;

  virtual void load() {
    params();
    upload();
  }
  // Solver::compute_forces postcondition: accs[i] = acceleration of body i
  // (for Barnes-Hut with bounds: of the bodies in range; the others keep
  // their accs).
  virtual void compute_forces() {
    if (!device_ahead_) upload();
    const size_t n = b_->positions.size();
    b_->accs.resize(n);
    const size_t lo = lower_ > 0 ? (size_t)lower_ : 0;
    const size_t hi = (upper_ < 0 || (size_t)upper_ > n) ? n : (size_t)upper_;
    if (bounds_select_targets() && (lo > 0 || hi < n)) {
      scratch_.resize(n);
      check(ctx_, aster_forces(ctx_, method_, &scratch_[0].x));
      for (size_t i = lo; i < hi; ++i) b_->accs[i] = scratch_[i];
    } else {
      check(ctx_, aster_forces(ctx_, method_, &b_->accs[0].x));
    }
  }
  // Host positions/velocities changed outside the device (e.g. user edits):
  // the next device updater uploads them first.
  void mark_dirty() { dirty_ = true; }
  void download() {
    check(ctx_, aster_get_state(ctx_, &b_->positions[0].x, &b_->velocities[0].x));
    device_ahead_ = false;
  }
  // Device updaters: the host copy, if edited, goes up first; afterwards
  // either the state is read back or the device copy is ahead of the host.
  void before_device_step() {
    if (dirty_) upload();
  }
  void after_device_step(bool read_back) {
    if (read_back)
      download();
    else
      device_ahead_ = true;
  }
  aster_ctx* context() { return ctx_; }
  int method() const { return method_; }

 protected:
  void params() {
    aster_params p;
    check(ctx_, aster_get_params(ctx_, &p));
    p.G = meta_->G;
    p.c = meta_->c;
    p.softening = meta_->softening;
    p.theta = meta_->theta;
    p.pn_flags = force_law | meta_->force_law;
    p.lower = lower_;
    p.upper = upper_;
    configure(p);
    check(ctx_, aster_set_params(ctx_, &p));
  }
  virtual void configure(aster_params&) {}
  // true when bounds select the bodies that get accelerations (Barnes-Hut)
  virtual bool bounds_select_targets() const { return false; }
  void upload() {
    const int64_t n = (int64_t)b_->positions.size();
    if (b_->velocities.size() != (size_t)n || b_->masses.size() != (size_t)n)
      throw std::runtime_error("aster_hip: positions, velocities and masses must have the same length");
    check(ctx_, aster_set_state(ctx_, n, &b_->positions[0].x, &b_->velocities[0].x, b_->masses.data()));
    dirty_ = false;
    device_ahead_ = false;
  }

  BodyArray* b_;
  Meta* meta_;
  int method_;
  aster_ctx* ctx_ = nullptr;
  int lower_ = 0, upper_ = -1;
  bool dirty_ = true;          // host bodies edited since the last upload
  bool device_ahead_ = false;  // a device updater advanced the state without reading it back
  std::vector<vec3> scratch_;
};

// Replaces SimpleGPU (src/solvers/SimpleGPU.cpp) and Parallelized
// (src/utils/builder_endpoint.cpp:729-757). method: ASTER_ALLPAIRS_FP64
// (bit-identical to Parallelized), ASTER_ALLPAIRS_FP32_KAHAN or
// ASTER_ALLPAIRS_FP32.
class HipAllPairs : public Solver {
 public:
  HipAllPairs(BodyArray* b, Meta* m, int method = ASTER_ALLPAIRS_FP32_KAHAN, int device = 0)
      : Solver(b, m, method, device) {}
};

// Replaces BH_hyper (src/solvers/full_GPU.cpp) and Barnes_Hut
// (src/solvers/barnes_hut.cpp). method: ASTER_BH_PARITY (reference opening
// test; FP64 precision is bit-identical to Barnes_Hut), ASTER_BH_GEOMETRIC or
// ASTER_BH_GROUP. box all zero = symmetric box of the data.
class HipBarnesHut : public Solver {
 public:
  HipBarnesHut(BodyArray* b, Meta* m, int method = ASTER_BH_GROUP, int precision = ASTER_FP32, int device = 0)
      : Solver(b, m, method, device), precision_(precision) {}
  void set_box(double bx, double by, double bz) {
    box_[0] = bx;
    box_[1] = by;
    box_[2] = bz;
    params();
  }

 protected:
  void configure(aster_params& p) override {
    p.precision = precision_;
    for (int k = 0; k < 3; ++k) p.box[k] = box_[k];
  }
  bool bounds_select_targets() const override { return true; }
  int precision_;
  double box_[3] = {0, 0, 0};
};

// Updater functions (func_ptr update, sim_obj.h:64-77): advance the bodies on
// the device and copy positions/velocities back, as read_bodies_gpu does
// when always_read_pos is set (builder_endpoint.cpp:280-281).
inline void update_leapfrog_hip(Solver& s, const Meta& meta, int64_t steps = 1, bool read_back = true) {
  s.before_device_step();
  check(s.context(), aster_step(s.context(), s.method(), ASTER_LEAPFROG, meta.dt, steps));
  s.after_device_step(read_back);
}

inline void update_yoshida4_hip(Solver& s, const Meta& meta, int64_t steps = 1, bool read_back = true) {
  s.before_device_step();
  check(s.context(), aster_step(s.context(), s.method(), ASTER_YOSHIDA4, meta.dt, steps));
  s.after_device_step(read_back);
}

// update_SABA1..10 (src/integrators/SABA.cpp:19-553), picked by order as
// bake_update_function(SABA, ord) does (src/utils/updater.cpp:144-152);
// order 1 is update_euler (tool_chain_impl.cpp:158-167).
inline void update_saba_hip(Solver& s, const Meta& meta, int order, int64_t steps = 1, bool read_back = true) {
  s.before_device_step();
  check(s.context(), aster_integrate(s.context(), s.method(), ASTER_SABA, order, meta.dt, steps));
  s.after_device_step(read_back);
}

inline void update_euler_hip(Solver& s, const Meta& meta, int64_t steps = 1, bool read_back = true) {
  s.before_device_step();
  check(s.context(), aster_integrate(s.context(), s.method(), ASTER_EULER, 1, meta.dt, steps));
  s.after_device_step(read_back);
}

}  // namespace aster_hip
