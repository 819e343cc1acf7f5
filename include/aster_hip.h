/* aster_hip.h -- C ABI of the MI355X force and integration library.
 *
 * One opaque context owns all device memory and one HIP stream. Every entry
 * point returns 0 on success or a negative aster_status; aster_last_error()
 * gives the message. Nothing here exits the process (the reference's Check()
 * macro, include/Aster/simulations/basic.h:25-29, calls exit(-1)).
 *
 * Host arrays use the reference layout: positions, velocities and
 * accelerations are packed double[3] per body ("AoS-d3", the std::vector<vec3>
 * of BodyArray, include/Aster/physics/body.h:9-24), masses double[n].
 *
 * Calls that take or fill host arrays are synchronous for those arrays; calls
 * that only enqueue device work (aster_forces with a NULL output, aster_step,
 * aster_batch_step) return once the work is queued. aster_sync() waits and
 * reports any deferred device-side error (out-of-box Morton input, non-finite
 * state).
 *
 * A context is single-thread-affine; distinct contexts are independent.
 */
#ifndef ASTER_HIP_H
#define ASTER_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct aster_ctx aster_ctx;

enum aster_status {
  ASTER_OK = 0,
  ASTER_ERR_INVALID = -1,    /* bad argument or unsupported combination */
  ASTER_ERR_HIP = -2,        /* HIP runtime error */
  ASTER_ERR_OUT_OF_BOX = -3, /* Morton input outside the box (reference UB) */
  ASTER_ERR_NONFINITE = -4,  /* NaN/Inf in state */
  ASTER_ERR_CAPACITY = -5,   /* size beyond a documented limit */
  ASTER_ERR_STATE = -6,      /* call order: e.g. no state loaded yet */
  ASTER_ERR_NOMEM = -7       /* device allocation failed */
};

/* Force-law flags: same values as the reference's NEWTON / PN1
 * (include/Aster/simulations/basic.h:39-40). */
#define ASTER_NEWTON 0x1
#define ASTER_PN1 0x10
#define ASTER_PN2 0x100
#define ASTER_PN25 0x1000
/* Full N-body 1PN (Einstein-Infeld-Hoffmann) accelerations, including the
 * cross terms the pairwise ASTER_PN1 leaves out; no counterpart in the
 * reference (its get_pn1 is the two-body closed form summed over pairs,
 * tool_chain_impl.cpp:110-118). Replaces ASTER_PN1/PN2/PN25 (combining them
 * is an error); ASTER_ALLPAIRS_FP64 only. */
#define ASTER_EIH 0x20
/* every post-Newtonian term (velocity-dependent forces) */
#define ASTER_PN_TERMS (ASTER_PN1 | ASTER_PN2 | ASTER_PN25 | ASTER_EIH)

/* Force methods. */
enum aster_method {
  /* FP32 pair forces, plain FP32 accumulation. */
  ASTER_ALLPAIRS_FP32 = 0,
  /* FP32 pair forces; FP32 sums over 64-source tiles, Kahan-compensated FP32
   * accumulation across tiles. Replaces SimpleGPU::compute_forces
   * (src/solvers/SimpleGPU.cpp:17-71). */
  ASTER_ALLPAIRS_FP32_KAHAN = 1,
  /* FP64 in the reference's operation order: bit-identical to
   * Parallelized::compute_forces (src/utils/builder_endpoint.cpp:729-757). */
  ASTER_ALLPAIRS_FP64 = 2,
  /* LBVH Barnes-Hut with the reference opening test s^2 < theta^2 d^2 and
   * sizes halved per level (src/solvers/barnes_hut.cpp:202-261); per-body
   * stackless right-first traversal. Precision from aster_params.precision:
   * FP64 is bit-identical to Barnes_Hut::compute_forces. Replaces
   * BH_hyper::compute_forces (src/solvers/full_GPU.cpp:519-527). */
  ASTER_BH_PARITY = 3,
  /* LBVH Barnes-Hut, FP32, geometric opening test on node bounding boxes:
   * accept when d > l/theta + delta (l = longest box side, delta = distance
   * of the centre of mass from the box centre). Per-body stackless traversal.
   * On every FP32 traversal the post-Newtonian terms of pn_flags are added in
   * FP64 for each node taken, with its FP64 centre of mass, mass and
   * mass-weighted velocity as in Barnes_Hut::get_node_body. */
  ASTER_BH_GEOMETRIC = 4,
  /* As ASTER_BH_GEOMETRIC, but the bodies of each subtree of at most 64
   * leaves (a group, Morton-consecutive) share one traversal against the
   * group's bounding box, and the wavefront evaluates the shared interaction
   * list cooperatively (FP32 sums, Kahan-compensated across list flushes).
   * The post-Newtonian terms of pn_flags are added in FP64 over the same
   * lists, with each node's FP64 centre of mass, mass and mass-weighted
   * velocity as in Barnes_Hut::get_node_body (barnes_hut.cpp:202-261). */
  ASTER_BH_GROUP = 5
};

enum aster_precision { ASTER_FP32 = 0, ASTER_FP64 = 1 };

/* Integrators: values follow the reference's update_type where one exists
 * (EULER = 0, SABA = 1, LEAPFROG = 2, WH_PLANETARY = 3,
 * include/Aster/simulations/basic.h:56). */
enum aster_integrator {
  ASTER_EULER = 0,
  ASTER_SABA = 1,
  ASTER_LEAPFROG = 2,
  ASTER_WH_PLANETARY = 3, /* batched ensembles only (aster_batch_step) */
  ASTER_YOSHIDA4 = 5
};

typedef struct aster_params {
  double G;         /* gravitational constant (reference default 6e-11) */
  double c;         /* speed of light for 1PN (reference default 299792458) */
  double softening; /* eps^2, added to r^2 (tool_chain_impl.cpp:122) */
  double theta;     /* opening parameter (reference default 0.2) */
  double box[3];    /* Morton box half-extent (Barnes_Hut::bounding_box);
                       all zero = symmetric max |coordinate| of the state */
  int32_t pn_flags; /* ASTER_NEWTON | ASTER_PN1 | ASTER_PN2 | ASTER_PN25, or ASTER_NEWTON | ASTER_EIH */
  int32_t precision;/* ASTER_FP32 / ASTER_FP64 for ASTER_BH_PARITY */
  int32_t lower;    /* Solver::set_bounds (sim_obj.h:45), upper < 0 = n:  */
  int32_t upper;    /* all-pairs: source range [lower, upper) (Parallelized::compute_forces,
                       builder_endpoint.cpp:729-757); Barnes-Hut: body range, the tree holds
                       bodies [lower, upper) and only they get new accelerations
                       (Barnes_Hut::compute_forces, barnes_hut.cpp:309-322,414-417) */
  int32_t split;    /* all-pairs source slices (0 = autotuned choice) */
  int32_t reserved;
} aster_params;

/* Library version (major*10000 + minor*100 + patch). */
int aster_version(void);

/* Create a context on HIP device `device`; fills *out. Parameters start at
 * the reference defaults (G = 6e-11, c = 299792458, softening = 1e-14,
 * theta = 0.2, sim_meta.h:23-24, sim_obj.h:180). */
int aster_create(aster_ctx** out, int device);
int aster_destroy(aster_ctx* ctx);
/* Message for the last failing call on ctx (never NULL). */
const char* aster_last_error(const aster_ctx* ctx);

int aster_set_params(aster_ctx* ctx, const aster_params* p);
int aster_get_params(const aster_ctx* ctx, aster_params* p);

/* Upload n bodies (Simulation::load_gpu_buffers, builder_endpoint.cpp:102-119).
 * vel may be NULL (zero velocities). Invalidates cached accelerations. */
int aster_set_state(aster_ctx* ctx, int64_t n, const double* pos, const double* vel, const double* mass);
/* Download state (Simulation::read_bodies_gpu, builder_endpoint.cpp:424-434);
 * either pointer may be NULL. */
int aster_get_state(aster_ctx* ctx, double* pos, double* vel);
/* The same for bodies already in device memory on this context's device (a
 * framework tensor, another HIP buffer): AoS-d3 device pointers, transposed
 * and checked on the device, with no host copy (the GPU state transfer of
 * load_gpu_buffers, builder_endpoint.cpp:102-119, and read_bodies_gpu,
 * :424-434). vel may be NULL (zero velocities). The call orders itself on the
 * context's stream and returns when the copy is done. */
int aster_set_state_device(aster_ctx* ctx, int64_t n, const double* d_pos, const double* d_vel,
                           const double* d_mass);
int aster_get_state_device(aster_ctx* ctx, double* d_pos, double* d_vel);
int64_t aster_num_bodies(const aster_ctx* ctx);

/* Evaluate accelerations with `method` (Solver::compute_forces postcondition:
 * accs[i] holds the acceleration of body i). acc may be NULL: the result then
 * stays on the device (enqueue only). */
int aster_forces(aster_ctx* ctx, int method, double* acc);
/* Copy the last accelerations to the host. */
int aster_get_acc(aster_ctx* ctx, double* acc);

/* Allocation-free all-pairs. No counterpart in the reference: for callers
 * that cannot allocate device memory (code under HIP graph capture, serving
 * loops, harnesses that forbid hipMalloc). The ASTER_ALLPAIRS_FP32_KAHAN
 * evaluation of aster_forces (the same kernels and FP32 mass scale) with the
 * source split S = min(64, npad / 1024, ceil(8192 / (npad / 64))), at least 1,
 * npad = n rounded up to 256: the same bits as aster_forces with those G and
 * softening, aster_params.split = S, full range, Newtonian. (aster_forces's
 * own default split is finer; S bounds the static scratch below.) It works
 * on caller-owned device arrays of the current device:
 * d_pos and d_acc AoS-d3 (3n doubles), d_mass n doubles. It allocates
 * nothing: its working memory is a static device array of about 17 MiB per
 * GPU, held until the process ends, shared by every call on that GPU, so
 * calls that overlap on one GPU race. Work is enqueued on `stream` (a
 * hipStream_t; NULL = the default stream) and the call returns without
 * waiting; inputs are not checked on the device (aster_set_state_device
 * does that). n above ASTER_NOALLOC_MAX_BODIES is ASTER_ERR_CAPACITY.
 * Errors read through aster_last_error(NULL). */
#define ASTER_NOALLOC_MAX_BODIES 65536
int aster_forces_noalloc(int64_t n, const double* d_pos, const double* d_mass, double G, double softening,
                         double* d_acc, void* stream);

/* Advance nsteps with `integrator` and force `method` entirely on the device
 * (update_leapfrog, tool_chain_impl.cpp:169-185; Yoshida-4 is the composition
 * of three such steps with w1 dt, w0 dt, w1 dt). Newtonian accelerations
 * depend on positions only and are cached between steps, so a leapfrog step
 * costs one force evaluation; with ASTER_PN1 (velocity-dependent) each step
 * keeps the reference's two evaluations. Positions and velocities are
 * synchronised at the end of the call. */
int aster_step(aster_ctx* ctx, int method, int integrator, double dt, int64_t nsteps);

/* aster_step with an order, for the integrators that take one (the
 * reference's bake_update_function(update_type, int ord),
 * src/utils/updater.cpp:144-152):
 *  - ASTER_SABA, order 2..10: update_SABAn (src/integrators/SABA.cpp:19-553),
 *    drift c1, then per stage one force evaluation, a kick (a d_k) dt and a
 *    drift (v c_(k+1)) dt; n force evaluations per step. Order 6 uses the
 *    symmetric kick sequence (the reference's fifth stage kicks with d2,
 *    SABA.cpp:216, so its kicks sum to 0.946);
 *  - ASTER_SABA order 1 and ASTER_EULER: update_euler (tool_chain_impl.cpp:
 *    158-167), which update_SABA1 calls: v += a dt, then x += v dt;
 *  - ASTER_LEAPFROG and ASTER_YOSHIDA4 ignore `order` (aster_step).
 * The accelerations left on the device are those of the last evaluation. */
int aster_integrate(aster_ctx* ctx, int method, int integrator, int order, double dt, int64_t nsteps);

/* Yoshida weights used by ASTER_YOSHIDA4: w[0] = w1, w[1] = w0, w[2] = w1. */
int aster_yoshida_weights(double* w);

/* Graphs collectors (src/utils/collectors.cpp:19-79), evaluated on the device.
 * ASTER_COLLECT_HAMILTONIAN: get_total_energy (the reference's definition,
 *   e_ref of aster_energy);
 * ASTER_COLLECT_ERROR: error_collector, |E0 - E| / |E0| with E0 the first
 *   value collected after the state was loaded (that call returns 0; the
 *   reference keeps E0 in a function static);
 * ASTER_COLLECT_DISTANCE: distance_collector, |COM - x_body + (0.1, 0.1, 0.1)|;
 *   COM is the true barycentre (the reference's get_center_of_mass,
 *   builder_endpoint.cpp:262-269, returns its x component in all three);
 * ASTER_COLLECT_ANGULAR_MOMENTUM: |L|. `body` is used by DISTANCE only. */
enum aster_collector {
  ASTER_COLLECT_HAMILTONIAN = 0,
  ASTER_COLLECT_ERROR = 1,
  ASTER_COLLECT_DISTANCE = 2,
  ASTER_COLLECT_ANGULAR_MOMENTUM = 3
};
int aster_collect(aster_ctx* ctx, int kind, int64_t body, double* out);
/* Barycentre sum m x / sum m, summed on the device in a fixed order
 * (Simulation::get_center_of_mass, builder_endpoint.cpp:262-269). */
int aster_center_of_mass(aster_ctx* ctx, double* com);

/* IntegrationReport (include/Aster/simulations/sim_obj.h:80-87): wall time
 * of the steps, and with tested_precision the energy (the reference's
 * definition, Graphs::get_total_energy) before and after and
 * precision = |E_after - E_before| / |E_before| (0 when E_before is 0). */
typedef struct aster_integration_report {
  double total_time_ms;
  double time_per_step_ms;
  double initial_energy;
  double final_energy;
  double precision;
  int32_t tested_precision;
  int32_t steps; /* steps taken (aster_integrate_until may stop early) */
} aster_integration_report;
/* Simulation::integrate(times, precision_test) (src/utils/builder_endpoint.cpp:
 * 287-352): aster_integrate for nsteps (order as there), timed on the host
 * around the device work (synchronised before and after), energies only
 * when precision_test. */
int aster_integrate_report(aster_ctx* ctx, int method, int integrator, int order, double dt, int64_t nsteps,
                           int precision_test, aster_integration_report* report);
/* Simulation::integrate(halting_condition, precision_test, max_steps)
 * (builder_endpoint.cpp:354-410): one step at a time until halt(ctx, user)
 * returns nonzero (checked before each step, as the reference's while loop)
 * or max_steps steps (max_steps < 0: no limit) have run. The state is
 * synchronised before each check, so halt may read it (aster_get_state). */
typedef int (*aster_halting_condition)(aster_ctx* ctx, void* user);
int aster_integrate_until(aster_ctx* ctx, int method, int integrator, int order, double dt,
                          aster_halting_condition halt, void* user, int64_t max_steps, int precision_test,
                          aster_integration_report* report);

/* SABAn coefficients used by ASTER_SABA (order 2..10): drifts c[0..order]
 * and kicks d[0..order-1] (GPU::saba_coeffs, include/Aster/physics/
 * coeffs.h:8-161, stores them interleaved). They are the Gauss-Legendre
 * nodes (as drift increments) and weights on [0, 1]. */
int aster_saba_coeffs(int order, double* c, double* d);

/* Energies and angular momentum of the current state, FP64:
 * e_ref     = sum 1/2 m v^2 - sum_{i<j} G m_i m_j / (|r_ij| + softening)
 *             (Graphs::get_total_energy, src/utils/collectors.cpp:19-53);
 * e_plummer = kinetic - sum_{i<j} G m_i m_j / sqrt(r_ij^2 + softening);
 * L         = sum m x cross v (3 doubles). Any pointer may be NULL. */
int aster_energy(aster_ctx* ctx, double* e_ref, double* e_plummer, double* L);

/* Tree of the last Barnes-Hut evaluation, in the reference's NodesArray
 * indexing (src/solvers/barnes_hut.cpp:309-400): leaf k = sorted position k,
 * internal node of split i at n + i, root n, children -1 for leaves.
 * codes: sorted Morton codes (n); perm: body at sorted position k (n);
 * left/right: 2n-1; node_mass: 2n-1; node_com: 3(2n-1). NULL skips a field. */
int aster_tree_dump(aster_ctx* ctx, uint32_t* codes, uint32_t* perm, int32_t* left, int32_t* right,
                    double* node_mass, double* node_com);
/* Morton codes of the current state, unsorted, with the box in use
 * (get_morton, src/solvers/bh_specialization.cpp:38-46). box_out may be NULL. */
int aster_morton(aster_ctx* ctx, uint32_t* codes, double* box_out);
/* Barnes-Hut counters accumulated over all evaluations since the previous
 * call (which resets them): accepted body-node interactions (self and
 * coincident nodes excluded) and node visits, summed over bodies. */
int aster_bh_stats(aster_ctx* ctx, double* interactions, double* node_visits);

/* Wait for all queued work; reports deferred device errors. */
int aster_sync(aster_ctx* ctx);
/* HIP-event timers on the context's stream: mark records event `slot`
 * (0..15); elapsed waits for `b` and returns ms between marks a and b. */
int aster_timer_mark(aster_ctx* ctx, int slot);
int aster_timer_elapsed(aster_ctx* ctx, int a, int b, float* ms);

/* Per-phase device timing. When enabled, every launch phase below is
 * bracketed by HIP events on the context's stream; aster_phase_times waits,
 * writes the summed milliseconds and launch counts per phase (arrays of
 * ASTER_NUM_PHASES) and clears the record. */
enum aster_phase {
  ASTER_PHASE_MORTON = 0,   /* box reduction + Morton codes */
  ASTER_PHASE_SORT = 1,     /* 4-pass LDS radix sort */
  ASTER_PHASE_BUILD = 2,    /* Karras split + leaf init */
  ASTER_PHASE_MOMENTS = 3,  /* bottom-up moments */
  ASTER_PHASE_LINKS = 4,    /* escape links + opening thresholds */
  ASTER_PHASE_TRAVERSE = 5, /* Barnes-Hut traversal */
  ASTER_PHASE_ALLPAIRS = 6, /* all-pairs pair kernel */
  ASTER_PHASE_COMBINE = 7,  /* slice combine / 1PN add */
  ASTER_PHASE_KICKDRIFT = 8,/* kick/drift update */
  ASTER_PHASE_ADJOINT = 9,  /* fused force + adjoint kernel (its sources: TILES, its slice sums: COMBINE) */
  ASTER_PHASE_TILES = 10,   /* anchored FP32 source tiles for all-pairs */
  ASTER_NUM_PHASES = 12
};
int aster_profile(aster_ctx* ctx, int enable);
int aster_phase_times(aster_ctx* ctx, double* ms, int64_t* counts);

/* ---- batched ensembles: many small independent systems ------------------ */
/* Upload nsys systems; system s owns bodies [offsets[s], offsets[s+1]).
 * Each system holds 1..64 bodies. Arrays are AoS-d3 over all bodies. */
int aster_batch_set(aster_ctx* ctx, int64_t nsys, const int64_t* offsets, const double* pos, const double* vel,
                    const double* mass);
int aster_batch_get(aster_ctx* ctx, double* pos, double* vel);
/* Integrate every system nsteps with all-pairs FP64 forces (Newton, plus
 * pairwise 1PN if pn_flags has ASTER_PN1, Plummer-softened by softening) in a
 * persistent kernel: each system lives in the lanes of one wavefront (several
 * small systems may share one). precision ASTER_FP64 evaluates
 * compute_gravitational in the reference's operation order (bit-identical to
 * the CPU path per system); ASTER_FP32 selects the fast FP64 path (rsqrt
 * estimate + Newton steps). Launches of steps_per_launch steps (0 = tuned) are
 * captured once in a HIP graph and replayed when use_graph != 0.
 * integrator: ASTER_LEAPFROG, ASTER_YOSHIDA4, or ASTER_WH_PLANETARY, which
 * replaces update_WH_planetary (src/physics/tool_chain_impl.cpp:235-274):
 * body 0 of each system is the central mass; democratic-heliocentric
 * Wisdom-Holman (half kick with the planet-planet forces, half jump, exact
 * Kepler drift about G m_0, half jump, half kick), Newtonian, in the
 * reference's pair-force order whatever the precision. The reference's
 * version kicks with the full accelerations after the Kepler drift, counting
 * the central pull twice; this one does not. */
int aster_batch_step(aster_ctx* ctx, int integrator, double dt, int64_t nsteps, int32_t steps_per_launch,
                     int32_t use_graph);
/* Per-system energy (kinetic + Plummer-softened potential) and angular
 * momentum: e[nsys], L[3 nsys]; either may be NULL. */
int aster_batch_energy(aster_ctx* ctx, double* e, double* L);
/* Time candidate launch shapes (packed or one system per wavefront; 1, 2 or
 * 4 wavefronts per block) on the loaded batch, restoring its state, and keep
 * the fastest. Writes the chosen wavefronts per block and steps per launch. */
int aster_batch_autotune(aster_ctx* ctx, int integrator, double dt, int32_t* systems_per_block,
                         int32_t* steps_per_launch);

/* Work of one force evaluation over the loaded batch, for throughput and
 * lane-use figures (no reference counterpart): ordered pair interactions
 * sum n (n - 1) over the systems, and lane-pair slots, the iterations the
 * lanes run (each wavefront loops over its largest system, so a packed
 * wavefront's smaller systems idle part of it), and the wavefront count of
 * the current packing. Any pointer may be NULL. */
int aster_batch_info(aster_ctx* ctx, int64_t* pair_interactions, int64_t* lane_slots, int32_t* nwaves);

/* Trajectory adjoint of the loaded ensemble (no reference counterpart):
 * leapfrog with Newtonian FP64 forces (fast mode) for nsteps, keeping the
 * positions of every step on the device; snapshot positions (nsnap * 3 *
 * nbodies, AoS-d3) go to snap_pos if not NULL. The backward pass takes
 * dLoss/d(snapshot positions) and writes dLoss/d(initial positions,
 * velocities) (3 * nbodies each) and d(masses) (nbodies) for every system at
 * once, then restores the initial state. */
int aster_batch_traj_forward(aster_ctx* ctx, double dt, int64_t nsteps, int32_t nsnap, const int64_t* snap_steps,
                             double* snap_pos);
int aster_batch_traj_backward(aster_ctx* ctx, const double* grad_snap_pos, double* g_pos, double* g_vel,
                              double* g_mass);

/* ---- trajectory gradients (adjoint of leapfrog) ------------------------ */
/* Forward: leapfrog with ASTER_ALLPAIRS_FP32_KAHAN forces (FP64 state) for
 * nsteps from the current state, keeping a checkpoint every ckpt_every steps.
 * snap_steps (ascending, in [1, nsteps]) select nsnap snapshots whose
 * positions are written to snap_pos (nsnap * 3n, AoS-d3) if not NULL. The
 * state is left at step nsteps. */
int aster_traj_forward(aster_ctx* ctx, double dt, int64_t nsteps, int64_t ckpt_every, int32_t nsnap,
                       const int64_t* snap_steps, double* snap_pos);
/* Backward: given dLoss/d(snapshot positions) (nsnap * 3n), write
 * dLoss/d(initial positions) (3n), d(initial velocities) (3n) and d(masses)
 * (n) by the hand-written adjoint (transposed pair-force Jacobian-vector
 * product fused with force recomputation, time-reversed recomputation,
 * re-anchored at checkpoints). Restores the initial state on return. */
int aster_traj_backward(aster_ctx* ctx, const double* grad_snap_pos, double* g_pos, double* g_vel,
                        double* g_mass);

/* ---- initial conditions (seeded, counter-based, reproducible) ---------- */
/* Plummer sphere (Aarseth, Henon & Wielen 1974), G = 1, total mass `mtot`,
 * scale radius b, truncated at r <= rmax b, recentred to zero centre of mass
 * and momentum. mass_mode 0: equal masses; 1: log-uniform over two decades
 * normalised to mtot. */
int aster_ic_plummer(int64_t n, uint64_t seed, double mtot, double b, double rmax, int mass_mode, double* pos,
                     double* vel, double* mass);
/* Uniform cube [-h, h]^3 at rest, equal masses summing to mtot. */
int aster_ic_cube(int64_t n, uint64_t seed, double h, double mtot, double* pos, double* vel, double* mass);
/* Planetary ensembles: per system a central mass 1 and n-1 bodies on Kepler
 * orbits (G = 1), barycentric. sizes[nsys] drawn uniform in [nmin, nmax]
 * unless sizes_in is given; offsets[nsys+1]. Call once with pos == NULL to
 * get sizes/offsets (and the total body count in *total), then again. */
int aster_ic_ensemble(int64_t nsys, uint64_t seed, int32_t nmin, int32_t nmax, int64_t* offsets, int64_t* total,
                      double* pos, double* vel, double* mass);
/* Three-body systems: masses U[0.5,1.5], positions U[0,1]^3, velocities
 * U[-0.5,0.5]^3 shifted to zero total momentum. Arrays of 3*nsys bodies. */
int aster_ic_threebody(int64_t nsys, uint64_t seed, double* pos, double* vel, double* mass);

/* ---- cluster queue, generated on the device ---------------------------- */
/* Cluster (include/Aster/building-api/clusters.h:17-116) and the presets
 * that fill it (add_disk, src/clusters/presets_impl.cpp:207-255). Bodies are
 * drawn from counter-based seeded streams on the device, so a queue of
 * clusters loads 10^8 bodies without a host round trip; a Plummer or cube
 * cluster holds the same bits as aster_ic_plummer / aster_ic_cube with the
 * same seed (plus its translation). */
enum aster_cluster_kind {
  ASTER_CLUSTER_PLUMMER = 1, /* aster_ic_plummer body law; size = {b, rmax / b, -}; mass = total mass */
  ASTER_CLUSTER_CUBE = 2,    /* uniform cube at rest; size = {half-extent, -, -}; mass = total mass */
  ASTER_CLUSTER_DISK = 3     /* add_disk; size = {radius, thickness, inner radius}; mass = avr_mass */
};
typedef struct aster_cluster {
  int32_t kind;        /* aster_cluster_kind */
  int32_t mass_mode;   /* Plummer: 0 equal masses, 1 log-uniform over two decades (normalised) */
  int64_t number;      /* bodies drawn (Cluster::number); a disk also has its central body */
  uint64_t seed;       /* stream seed; body k of the cluster uses stream (seed, k) */
  double position[3];  /* translation (Cluster::position; add_disk's center) */
  double velocity[3];  /* bulk velocity (add_disk's v) */
  double rotation[3];  /* degrees, disks only: rotate_point(p, rotation[0], rotation[2]) as add_disk */
  double size[3];      /* per kind, see aster_cluster_kind */
  double mass;         /* per kind, see aster_cluster_kind */
  double center_mass;  /* disk: central body mass (add_disk's center_mass) */
  double scale;        /* disk: Simulation::get_scale(), multiplies the cylinder point */
} aster_cluster;
/* ClusterQueue::add_cluster (clusters_endpoint.cpp:150-156) via a preset.
 * Rejects what the presets reject with a warning (no bodies, non-positive
 * mass, radius or thickness; presets_impl.cpp:208-221) and an inner radius
 * above the radius (rng_point_in_cylinder, :131). A disk adds its central
 * body to the state at once, as add_disk does (:227-231). */
int aster_cluster_add(aster_ctx* ctx, const aster_cluster* cluster);
/* ClusterQueue::pop_cluster (clusters_endpoint.cpp:163-174): drops the last
 * queued cluster; on an empty queue it does nothing. */
int aster_cluster_pop(aster_ctx* ctx);
/* Queued clusters and the bodies they will add. */
int aster_cluster_queued(const aster_ctx* ctx, int64_t* nclusters, int64_t* nbodies);
/* ClusterQueue::load (clusters_endpoint.cpp:182-207): generates every queued
 * cluster on the device, last added first, appends the bodies after the
 * current state and empties the queue. Disk velocities use the G set by
 * aster_set_params at this call, as the reference's builder reads get_G(). */
int aster_cluster_load(aster_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* ASTER_HIP_H */
