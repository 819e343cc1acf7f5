// Internal declarations shared by the HIP translation units of libaster_hip.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/aster_hip.h"

namespace aster {

// Growable device allocation. Contents are not preserved on growth.
template <class T>
struct DBuf {
  T* p = nullptr;
  size_t cap = 0;  // elements
  hipError_t reserve(size_t n) {
    if (n <= cap) return hipSuccess;
    if (p) {
      hipError_t e = hipFree(p);
      if (e != hipSuccess) return e;
      p = nullptr;
      cap = 0;
    }
    if (n == 0) return hipSuccess;
    hipError_t e = hipMalloc(&p, n * sizeof(T));
    if (e == hipSuccess) cap = n;
    return e;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
  }
};

// Device-side error word bits (ctx->flags[0]).
enum : uint32_t { FLAG_OUT_OF_BOX = 1u, FLAG_NONFINITE = 2u, FLAG_TREE = 4u, FLAG_BADMASS = 8u };

// Sizes of the Barnes-Hut counters block (ctx->bh_counters, u64).
enum { BHC_INTERACTIONS = 0, BHC_VISITS = 1, BHC_COUNT = 4 };

struct BatchState {
  int64_t nsys = 0, nbodies = 0;
  std::vector<int64_t> offsets;      // host copy, nsys+1
  DBuf<double> x, v, a, m;           // AoS-d3 over bodies (x, v, a), m per body
  DBuf<double> ebody;                // per-body energy terms [4][nbodies]
  DBuf<double> esys;                 // per-system energy + L [4][nsys]
  DBuf<int64_t> doff;                // device offsets
  // Launch plan: wavefront w owns lanes [64 w, 64 w + 64).
  int pack = 1;                      // 0: one system per wavefront; 1: systems packed into wavefronts
  int waves_per_block = 4;
  int steps_per_launch = 0;          // 0: not tuned yet
  int nwaves = 0;
  int64_t lane_slots = 0;            // sum over wavefronts of 64 x its largest system: lane-pair slots per evaluation
  DBuf<int32_t> lane_body, lane_seg, wave_max;
  bool a_valid = false;
  // Captured launch sequence and the arguments it was captured with.
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int64_t g_steps = -1;
  int g_integrator = -1, g_spl = 0, g_pn = -1, g_exact = -1, g_wpb = 0, g_pack = -1;
  double g_dt = 0.0, g_G = 0.0, g_c = 0.0, g_soft = -1.0;
  // Trajectory adjoint of the ensemble: positions at every step are kept.
  DBuf<double> traj;                 // (nsteps + 1) * 3 * nbodies, AoS-d3 per step
  DBuf<double> x0, v0;               // initial state, restored after the backward pass
  DBuf<double> tgrad;                // snapshot gradients, nsnap * 3 * nbodies
  DBuf<double> tout;                 // gx, gv (AoS-d3) and gm per body
  DBuf<int64_t> tsnap;               // snapshot steps on the device
  std::vector<int64_t> traj_snaps;
  int64_t traj_steps = 0;
  double traj_dt = 0.0;
  bool traj_ready = false;
};

struct AdjointState {
  int64_t nsteps = 0, ckpt_every = 0;
  int32_t nsnap = 0;
  double dt = 0.0;
  std::vector<int64_t> snap_steps;
  DBuf<double> ckpt_x, ckpt_v;  // (nsteps/ckpt_every + 1) * 3n each
  DBuf<double> lx, lv, lm;      // adjoint accumulators (3n, 3n, n)
  DBuf<double> w;               // weights for the fused VJP (3n)
  DBuf<double> snapg;           // snapshot positions, then their gradients (nsnap * 3n)
  DBuf<double> g;               // J^T w (3n) and M^T w (n) of the current step
  DBuf<float4> wsrc;            // FP32 {w, m} per source, npad
  DBuf<double> part;            // slice partials: slices * 7 * n
  bool ready = false;
};

}  // namespace aster

struct aster_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  aster_params P{};
  int64_t n = 0;
  int64_t npad = 0;  // n rounded up to the all-pairs block
  bool have_state = false;
  bool acc_valid = false;
  bool acc_stale = false;  // positions moved after the last evaluation
  int acc_method = -1;

  // State, device resident. x, v, a: SoA planes [3][n]; m: [n].
  aster::DBuf<double> x, v, m, a;
  aster::DBuf<float4> sh;       // FP32 sources {x, y, z, m}, npad entries
  aster::DBuf<double4> src64;   // FP64 sources {x, y, z, m}, npad entries
  aster::DBuf<double> apart;    // all-pairs slice partials [S][3][n]
  aster::DBuf<double> apn;      // FP64 1PN correction [3][n]
  aster::DBuf<double> ework;    // energy work [6][n] + reduction scratch
  aster::DBuf<double> eout;     // 8 doubles
  aster::DBuf<uint32_t> flags;  // device error word(s), 4 words
  aster::DBuf<double> boxpart;  // per-block |x| maxima for the auto box [blocks][3]
  // boxpart holds box_parts partials of the positions a fused kick/drift just
  // wrote (state buffer box_parts_x, box_parts_n bodies); 0 = none pending.
  int box_parts = 0;
  const double* box_parts_x = nullptr;
  int64_t box_parts_n = 0;
  aster::DBuf<double> stage;    // AoS-d3 staging for host transfers (7 n)
  double e0 = 0.0;              // first Hamiltonian after a state load (error collector)
  bool e0_valid = false;
  double mass_total = 0.0;      // sum of the masses and largest |coordinate| when they were taken (FP32
  double extent = 0.0;          // force scaling), valid while mass_total_valid (the state was not reloaded)
  bool mass_total_valid = false;
  aster::DBuf<double> boxd;     // box half-extent used (3) + |box| (1)

  // Barnes-Hut / LBVH.
  aster::DBuf<uint32_t> codes, idx, codes2, idx2;  // sort ping-pong, n
  aster::DBuf<double4> xm4;                        // {x, y, z, m} per body (body order), n
  aster::DBuf<uint32_t> hist;                      // per-block digit counts
  aster::DBuf<int32_t> left, right, parent, escape;// 2n-1
  aster::DBuf<uint32_t> scanpart;                  // per-chunk totals of the device scan
  aster::DBuf<int32_t> depth;                      // 2n-1
  aster::DBuf<int32_t> splitnode;                  // internal node whose split lies after sorted position k (n-1)
  aster::DBuf<int32_t> groot;                      // at a group's first position: the group's root node (n)
  aster::DBuf<int32_t> climb;                      // leaf k: levels from the leaf up to its group's root (n)
  aster::DBuf<int32_t> dmax;                       // per-block largest internal-node depth
  aster::DBuf<uint32_t> censusctr;                 // census: per-depth counts and cursors (256)
  aster::DBuf<int32_t> levels;                     // per-depth bucket offsets, level count, max depth
  aster::DBuf<int32_t> bucket;                     // internal nodes grouped by depth, in-group ones last (n-1)
  aster::DBuf<int2> grange;                        // sorted-position range of internal node n+i
  aster::DBuf<uint32_t> gscan;                     // group-start flags, then their exclusive scan (n)
  aster::DBuf<int32_t> gstart;                     // first sorted position of each group (+ n)
  aster::DBuf<uint32_t> gcount;                    // number of groups (1)
  aster::DBuf<double4> nmc;                        // FP64 {com, mass} per node (2n-1)
  aster::DBuf<double4> nvel;                       // FP64 {mass-weighted velocity, 0} per node (1PN only)
  aster::DBuf<float4> nbox;                        // FP32 AABB: {lo, 0}, {hi, 0} per node (2 (2n-1))
  aster::DBuf<float4> nrec;                        // FP32 node record: com, mass (2n-1)
  aster::DBuf<int4> nlink;                         // right, escape, crit2 bits, leaf body (2n-1)
  aster::DBuf<int32_t> topslot;                    // top-levels moments: node - n -> list position (n)
  aster::DBuf<float4> kids;                        // group traversal: per child slot {record}, {node, crit2} (2(2n-1))
  aster::DBuf<unsigned long long> bh_counters;     // BHC_COUNT
  aster::DBuf<float4> gsh;                         // bodies in sorted order (FP32), n
  aster::DBuf<double> sub_x, sub_v, sub_a, sub_m;  // body range [lower, upper) of a bounded Barnes-Hut evaluation
  bool tree_valid = false;
  int tree_mac = -1;
  int tree_levels = 0;                             // depths of internal nodes: 0 .. tree_levels-1
  int32_t* host_words = nullptr;                   // pinned: census level bounds + largest depth
  hipEvent_t tree_ev = nullptr;                    // recorded after their readback

  // Timers.
  hipEvent_t ev[16] = {};
  // Phase profiling: event pairs recorded around launch phases.
  bool profiling = false;
  std::vector<hipEvent_t> prof_pool;       // created lazily, reused
  std::vector<int> prof_phase;             // phase of each recorded pair
  size_t prof_used = 0;                    // pairs in use
  int prof_open = -1;                      // phase currently open

  aster::BatchState batch;
  aster::AdjointState adj;
  std::vector<aster_cluster> clusters;  // queued, not yet generated (aster_cluster_load)
};

namespace aster {

int fail(aster_ctx* c, int code, const std::string& msg);
int hip_fail(aster_ctx* c, hipError_t e, const char* what);

#define ASTER_HIP(ctx, expr)                              \
  do {                                                    \
    hipError_t _e = (expr);                               \
    if (_e != hipSuccess) return ::aster::hip_fail((ctx), _e, #expr); \
  } while (0)

constexpr int kApBlock = 256;  // all-pairs targets per block
constexpr int kApTile = 64;    // sources per accumulation tile

inline int64_t round_up(int64_t a, int64_t b) { return (a + b - 1) / b * b; }

// Scaling of the FP32 pair-force kernels. They sum m_j d / (d.d + eps^2)^1.5
// with G factored out; in SI units (m ~ 1e30 kg, eps^2 ~ 1e-14 m^2) a body's
// own pair (d = 0, always in the group lists) or two coincident bodies give
// m eps^-3 ~ 1e51, beyond FP32, and inf * 0 = NaN. The kernels therefore see
// masses times msc = 2^-k and the sums are multiplied by gsc = G 2^k, with
// k >= 0 the smallest power that keeps (total mass) msc max(r^-3, r^-2) below
// 2^120 for every r >= eps. Powers of two scale exactly, so k = 0 (every
// problem in code units) leaves the arithmetic unchanged, and a larger k only
// moves far, light contributions toward FP32 underflow. With eps^2 not a
// positive normal FP32 number the kernels run their masked variant (zero
// force at r = 0) with eps^2 = 0, and the bound takes r >= 2^-40 times the
// largest |coordinate| instead of eps.
struct Fp32Scales {
  float eps2;   // FP32 softening eps^2 (0 when masked)
  bool mask;    // zero the pair term at r^2 = 0
  float msc;    // mass multiplier 2^-k
  double gsc;   // result multiplier G 2^k
};

// k of Fp32Scales from the total mass, the largest |coordinate| and the FP32
// softening, in integer exponent arithmetic (ilogb is exact), so the host and
// the device derive the same k: mtot < 2^em, every separation r >= 2^er (eps
// >= 2^floor(e2 / 2) with eps^2 >= 2^e2; masked: 2^-40 of the extent), and
// max(r^-3, r^-2) <= 2^max(-3 er, -2 er).
__host__ __device__ inline int fp32_mass_shift(double mtot, double extent, float eps2, bool mask) {
  const int em = ilogb(mtot) + 1;
  int er;
  if (mask) {
    er = (extent > 0.0 ? ilogb(extent) : 0) - 40;
  } else {
    const int e2 = ilogb((double)eps2);
    er = e2 >= 0 ? e2 / 2 : -((1 - e2) / 2);
  }
  const int k = em + (er < 0 ? -3 * er : -2 * er) - 120;
  return k < 0 ? 0 : k;
}
constexpr int kMaxMassShift = 126;  // 2^-k must stay a normal FP32 number
// Launchers (each returns an aster_status).
int launch_allpairs(aster_ctx* c, int method);           // writes c->a
int launch_kick_drift(aster_ctx* c, double dt_prev, double dt, bool box = false);
int launch_kick(aster_ctx* c, double dt);
int launch_saba_stage(aster_ctx* c, bool kick, double d, double cdrift, double dt);
int launch_energy(aster_ctx* c, double* out6_host);
int import_state(aster_ctx* c, int64_t n, const double* d_pos, const double* d_vel, const double* d_mass);
int export_plane(aster_ctx* c, const double* soa, double* d_aos);
int grow_state(aster_ctx* c, int64_t add, int64_t* base);  // appends `add` bodies; *base = first new index
int center_of_mass(aster_ctx* c, double* com);
int total_mass(aster_ctx* c, double* mtot, double* extent);  // sum of the masses, largest |coordinate| (cached)
int fp32_scales(aster_ctx* c, Fp32Scales* s);
int launch_reset_flags(aster_ctx* c);
int launch_box(aster_ctx* c, bool user_box);             // fills c->boxd
int bh_forces(aster_ctx* c, int method);                 // writes c->a
int bh_morton_only(aster_ctx* c);
int morton_sort(aster_ctx* c, bool user_box);            // c->codes sorted, c->idx = perm
int batch_step(aster_ctx* c, int integrator, double dt, int64_t nsteps, int spl, int use_graph);
int batch_energy(aster_ctx* c, double* e, double* L);
void batch_params_changed(aster_ctx* c);
int batch_autotune(aster_ctx* c, int integrator, double dt, int32_t* spb, int32_t* spl);
int traj_forward(aster_ctx* c, double dt, int64_t nsteps, int64_t ckpt, int32_t nsnap, const int64_t* snaps,
                 double* snap_pos);
int traj_backward(aster_ctx* c, const double* gsnap, double* gx, double* gv, double* gm);
int batch_traj_forward(aster_ctx* c, double dt, int64_t nsteps, int32_t nsnap, const int64_t* snaps,
                       double* snap_pos);
int batch_traj_backward(aster_ctx* c, const double* gsnap, double* gx, double* gv, double* gm);
// Mass sums (fixed order) and the largest |coordinate| over SoA planes:
// per-block partials of kComQ values (m, m x, m y, m z, max |x|) over at most
// kComBlocks blocks of 256 threads, then one thread sums them in block order.
constexpr int kComBlocks = 512;
constexpr int kComQ = 5;
__global__ void k_com_partial(int64_t n, const double* __restrict__ x, const double* __restrict__ m,
                              double* __restrict__ part);
__global__ void k_com_final(int nparts, const double* __restrict__ part, double* __restrict__ out);

int choose_split(aster_ctx* c);
int default_split(int64_t npad);                           // source slices of the FP32 all-pairs kernels
void set_global_error(const std::string& msg);            // aster_last_error(NULL) for calls without a context
int compute_forces(aster_ctx* c, int method);             // all-pairs or Barnes-Hut into c->a
void phase_begin(aster_ctx* c, int phase);
void phase_end(aster_ctx* c, int phase);

}  // namespace aster
