"""Python host binding of libaster_hip (the MI355X force/integration library).

A thin ctypes layer over the C ABI declared in include/aster_hip.h. It mirrors
the reference's Solver/Updater vocabulary: `Context.set_state` plays
Simulation::load_gpu_buffers (src/utils/builder_endpoint.cpp:102-119),
`Context.forces` plays Solver::compute_forces (include/Aster/simulations/
sim_obj.h:34-61) and `Context.step` plays Updater::update_bodies with the
leapfrog of src/physics/tool_chain_impl.cpp:169-185.

Host arrays are numpy float64 arrays in the reference layout: positions,
velocities and accelerations shaped (n, 3) ("AoS-d3"), masses shaped (n,).

The library must be built (python -m aster_amd.build, or __graft_entry__.build());
loading fails loudly otherwise. There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libaster_hip.so")

# Force methods (enum aster_method).
ALLPAIRS_FP32 = 0
ALLPAIRS_FP32_KAHAN = 1
ALLPAIRS_FP64 = 2
BH_PARITY = 3
BH_GEOMETRIC = 4
BH_GROUP = 5
# Precision (enum aster_precision).
FP32 = 0
FP64 = 1
# Integrators (enum aster_integrator; EULER = 0, SABA = 1, LEAPFROG = 2 as in
# basic.h:56).
EULER = 0
SABA = 1
LEAPFROG = 2
WH_PLANETARY = 3  # batched ensembles only

# Graphs collectors (enum aster_collector).
COLLECT_HAMILTONIAN = 0
COLLECT_ERROR = 1
COLLECT_DISTANCE = 2
COLLECT_ANGULAR_MOMENTUM = 3

CLUSTER_PLUMMER = 1
CLUSTER_CUBE = 2
CLUSTER_DISK = 3
YOSHIDA4 = 5
# Force-law flags (basic.h:39-40).
NEWTON = 0x1
PN1 = 0x10
PN2 = 0x100
PN25 = 0x1000
EIH = 0x20
# Profiling phases (enum aster_phase).
PHASES = ["morton", "sort", "build", "moments", "links", "traverse", "allpairs", "combine", "kickdrift",
          "adjoint", "tiles", "phase11"]
NUM_PHASES = 12

STATUS = {
    0: "ASTER_OK", -1: "ASTER_ERR_INVALID", -2: "ASTER_ERR_HIP", -3: "ASTER_ERR_OUT_OF_BOX",
    -4: "ASTER_ERR_NONFINITE", -5: "ASTER_ERR_CAPACITY", -6: "ASTER_ERR_STATE", -7: "ASTER_ERR_NOMEM",
}


class AsterError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS.get(code, code)}: {msg}")
        self.code = code


class Params(ctypes.Structure):
    _fields_ = [
        ("G", ctypes.c_double), ("c", ctypes.c_double), ("softening", ctypes.c_double),
        ("theta", ctypes.c_double), ("box", ctypes.c_double * 3), ("pn_flags", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("lower", ctypes.c_int32), ("upper", ctypes.c_int32),
        ("split", ctypes.c_int32), ("reserved", ctypes.c_int32),
    ]


class Cluster(ctypes.Structure):
    """aster_cluster (include/aster_hip.h)."""
    _fields_ = [
        ("kind", ctypes.c_int32), ("mass_mode", ctypes.c_int32), ("number", ctypes.c_int64),
        ("seed", ctypes.c_uint64), ("position", ctypes.c_double * 3), ("velocity", ctypes.c_double * 3),
        ("rotation", ctypes.c_double * 3), ("size", ctypes.c_double * 3), ("mass", ctypes.c_double),
        ("center_mass", ctypes.c_double), ("scale", ctypes.c_double),
    ]


class IntegrationReport(ctypes.Structure):
    """aster_integration_report (include/aster_hip.h)."""
    _fields_ = [
        ("total_time_ms", ctypes.c_double), ("time_per_step_ms", ctypes.c_double),
        ("initial_energy", ctypes.c_double), ("final_energy", ctypes.c_double), ("precision", ctypes.c_double),
        ("tested_precision", ctypes.c_int32), ("steps", ctypes.c_int32),
    ]


HALTING_CONDITION = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p)

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_U64 = ctypes.c_uint64
_PI64 = ctypes.POINTER(ctypes.c_int64)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PF = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes): every symbol declared in include/aster_hip.h.
SIGNATURES = {
    "aster_version": (ctypes.c_int, []),
    "aster_create": (ctypes.c_int, [ctypes.POINTER(_P), ctypes.c_int]),
    "aster_destroy": (ctypes.c_int, [_P]),
    "aster_last_error": (ctypes.c_char_p, [_P]),
    "aster_set_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "aster_get_params": (ctypes.c_int, [_P, ctypes.POINTER(Params)]),
    "aster_set_state": (ctypes.c_int, [_P, _I64, _D, _D, _D]),
    "aster_get_state": (ctypes.c_int, [_P, _D, _D]),
    "aster_set_state_device": (ctypes.c_int, [_P, _I64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    "aster_batch_info": (ctypes.c_int, [_P, ctypes.POINTER(_I64), ctypes.POINTER(_I64), ctypes.POINTER(ctypes.c_int32)]),
    "aster_forces_noalloc": (ctypes.c_int, [_I64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_void_p, ctypes.c_void_p]),
    "aster_get_state_device": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_void_p]),
    "aster_collect": (ctypes.c_int, [_P, ctypes.c_int, _I64, _D]),
    "aster_center_of_mass": (ctypes.c_int, [_P, _D]),
    "aster_num_bodies": (ctypes.c_int64, [_P]),
    "aster_forces": (ctypes.c_int, [_P, ctypes.c_int, _D]),
    "aster_get_acc": (ctypes.c_int, [_P, _D]),
    "aster_step": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I64]),
    "aster_integrate": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I64]),
    "aster_yoshida_weights": (ctypes.c_int, [_D]),
    "aster_saba_coeffs": (ctypes.c_int, [ctypes.c_int, _D, _D]),
    "aster_energy": (ctypes.c_int, [_P, _D, _D, _D]),
    "aster_tree_dump": (ctypes.c_int, [_P, _PU32, _PU32, _PI32, _PI32, _D, _D]),
    "aster_morton": (ctypes.c_int, [_P, _PU32, _D]),
    "aster_bh_stats": (ctypes.c_int, [_P, _D, _D]),
    "aster_sync": (ctypes.c_int, [_P]),
    "aster_timer_mark": (ctypes.c_int, [_P, ctypes.c_int]),
    "aster_timer_elapsed": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, _PF]),
    "aster_profile": (ctypes.c_int, [_P, ctypes.c_int]),
    "aster_phase_times": (ctypes.c_int, [_P, _D, _PI64]),
    "aster_batch_set": (ctypes.c_int, [_P, _I64, _PI64, _D, _D, _D]),
    "aster_batch_get": (ctypes.c_int, [_P, _D, _D]),
    "aster_batch_step": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, _I64, _I32, _I32]),
    "aster_batch_energy": (ctypes.c_int, [_P, _D, _D]),
    "aster_batch_autotune": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_double, _PI32, _PI32]),
    "aster_batch_traj_forward": (ctypes.c_int, [_P, ctypes.c_double, _I64, _I32, _PI64, _D]),
    "aster_batch_traj_backward": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "aster_traj_forward": (ctypes.c_int, [_P, ctypes.c_double, _I64, _I64, _I32, _PI64, _D]),
    "aster_traj_backward": (ctypes.c_int, [_P, _D, _D, _D, _D]),
    "aster_ic_plummer": (ctypes.c_int, [_I64, _U64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int, _D, _D, _D]),
    "aster_ic_cube": (ctypes.c_int, [_I64, _U64, ctypes.c_double, ctypes.c_double, _D, _D, _D]),
    "aster_ic_ensemble": (ctypes.c_int, [_I64, _U64, _I32, _I32, _PI64, _PI64, _D, _D, _D]),
    "aster_ic_threebody": (ctypes.c_int, [_I64, _U64, _D, _D, _D]),
    "aster_integrate_report": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _I64,
                                              ctypes.c_int, ctypes.POINTER(IntegrationReport)]),
    "aster_integrate_until": (ctypes.c_int, [_P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                             HALTING_CONDITION, ctypes.c_void_p, _I64, ctypes.c_int,
                                             ctypes.POINTER(IntegrationReport)]),
    "aster_cluster_add": (ctypes.c_int, [_P, ctypes.POINTER(Cluster)]),
    "aster_cluster_pop": (ctypes.c_int, [_P]),
    "aster_cluster_queued": (ctypes.c_int, [_P, _PI64, _PI64]),
    "aster_cluster_load": (ctypes.c_int, [_P]),
}

_lib = None


def lib():
    """Load libaster_hip.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m aster_amd.build`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None and a.shape != shape:
        a = a.reshape(shape)
    return a


def _check(code, ctx=None):
    if code != 0:
        msg = lib().aster_last_error(ctx).decode()
        raise AsterError(code, msg)


class Context:
    """One device context (owns device memory and one HIP stream)."""

    def __init__(self, device=0):
        h = _P()
        _check(lib().aster_create(ctypes.byref(h), device))
        self.h = h
        self.n = 0

    def close(self):
        if getattr(self, "h", None):
            lib().aster_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ck(self, code):
        _check(code, self.h)

    # ---- parameters and state
    def params(self):
        p = Params()
        self._ck(lib().aster_get_params(self.h, ctypes.byref(p)))
        return p

    def set_params(self, **kw):
        p = self.params()
        for k, v in kw.items():
            if k == "box":
                for i in range(3):
                    p.box[i] = float(v[i])
            else:
                setattr(p, k, v)
        self._ck(lib().aster_set_params(self.h, ctypes.byref(p)))
        return p

    def set_state(self, pos, vel, mass):
        pos = _f64(pos)
        n = pos.shape[0]
        pos = pos.reshape(n, 3)
        vel = None if vel is None else _f64(vel).reshape(n, 3)
        mass = _f64(mass).reshape(n)
        self._ck(lib().aster_set_state(self.h, n, _dp(pos), _dp(vel), _dp(mass)))
        self.n = n

    def get_state(self):
        pos = np.empty((self.n, 3))
        vel = np.empty((self.n, 3))
        self._ck(lib().aster_get_state(self.h, _dp(pos), _dp(vel)))
        return pos, vel

    def set_state_device(self, n, pos_ptr, vel_ptr, mass_ptr):
        """Load n bodies from AoS-d3 device memory (raw device addresses, vel may be None)."""
        self._ck(lib().aster_set_state_device(self.h, int(n), pos_ptr, vel_ptr, mass_ptr))
        self.n = int(n)

    def get_state_device(self, pos_ptr, vel_ptr):
        """Write positions/velocities as AoS-d3 into device memory (raw addresses, either may be None)."""
        self._ck(lib().aster_get_state_device(self.h, pos_ptr, vel_ptr))

    def collect(self, kind, body=0):
        """Graphs collector on the device: COLLECT_HAMILTONIAN, COLLECT_ERROR, COLLECT_DISTANCE, COLLECT_ANGULAR_MOMENTUM."""
        out = ctypes.c_double()
        self._ck(lib().aster_collect(self.h, int(kind), int(body), ctypes.byref(out)))
        return out.value

    def center_of_mass(self):
        com = np.empty(3)
        self._ck(lib().aster_center_of_mass(self.h, _dp(com)))
        return com

    # ---- cluster queue (generated on the device)
    def cluster_add(self, kind, number, seed, position=(0, 0, 0), velocity=(0, 0, 0), rotation=(0, 0, 0),
                    size=(1.0, 10.0, 0.0), mass=1.0, center_mass=0.0, scale=1.0, mass_mode=0):
        k = Cluster(kind=kind, mass_mode=mass_mode, number=int(number), seed=int(seed), mass=float(mass),
                    center_mass=float(center_mass), scale=float(scale))
        for i in range(3):
            k.position[i] = float(position[i])
            k.velocity[i] = float(velocity[i])
            k.rotation[i] = float(rotation[i])
            k.size[i] = float(size[i])
        self._ck(lib().aster_cluster_add(self.h, ctypes.byref(k)))
        self.n = int(lib().aster_num_bodies(self.h))

    def add_disk(self, nums, center, radius, thickness, rotation, avr_mass=10e6, center_mass=10e20, v=(0, 0, 0),
                 seed=0, scale=1.0, inner=10.0):
        """The reference's add_disk (presets_impl.cpp:207-255) on this context's queue."""
        self.cluster_add(CLUSTER_DISK, nums, seed, position=center, velocity=v, rotation=rotation,
                         size=(radius, thickness, inner), mass=avr_mass, center_mass=center_mass, scale=scale)

    def cluster_pop(self):
        self._ck(lib().aster_cluster_pop(self.h))

    def cluster_queued(self):
        nc = ctypes.c_int64()
        nb = ctypes.c_int64()
        self._ck(lib().aster_cluster_queued(self.h, ctypes.byref(nc), ctypes.byref(nb)))
        return nc.value, nb.value

    def cluster_load(self):
        self._ck(lib().aster_cluster_load(self.h))
        self.n = int(lib().aster_num_bodies(self.h))

    # ---- forces and integration
    def forces(self, method, out=True):
        acc = np.empty((self.n, 3)) if out else None
        self._ck(lib().aster_forces(self.h, method, _dp(acc)))
        return acc

    def get_acc(self):
        acc = np.empty((self.n, 3))
        self._ck(lib().aster_get_acc(self.h, _dp(acc)))
        return acc

    def step(self, method, integrator, dt, nsteps):
        self._ck(lib().aster_step(self.h, method, integrator, float(dt), int(nsteps)))

    def integrate(self, method, integrator, order, dt, nsteps):
        """aster_integrate: SABA of `order` (1..10), Euler, or the step integrators."""
        self._ck(lib().aster_integrate(self.h, method, integrator, int(order), float(dt), int(nsteps)))

    def integrate_report(self, method, integrator, dt, nsteps, order=0, precision_test=True):
        """Simulation::integrate(times, precision_test): IntegrationReport fields as a dict."""
        r = IntegrationReport()
        self._ck(lib().aster_integrate_report(self.h, method, integrator, int(order), float(dt), int(nsteps),
                                              1 if precision_test else 0, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in IntegrationReport._fields_}

    def integrate_until(self, method, integrator, dt, halt, max_steps=-1, order=0, precision_test=True):
        """Simulation::integrate(halting_condition, ...): halt(ctx) -> bool is checked before each step."""
        cb = HALTING_CONDITION(lambda _c, _u: 1 if halt(self) else 0)
        r = IntegrationReport()
        self._ck(lib().aster_integrate_until(self.h, method, integrator, int(order), float(dt), cb, None,
                                             int(max_steps), 1 if precision_test else 0, ctypes.byref(r)))
        return {k: getattr(r, k) for k, _ in IntegrationReport._fields_}

    def energy(self):
        e = ctypes.c_double()
        ep = ctypes.c_double()
        L = np.empty(3)
        self._ck(lib().aster_energy(self.h, ctypes.byref(e), ctypes.byref(ep), _dp(L)))
        return e.value, ep.value, L

    def tree(self):
        n = self.n
        nn = 2 * n - 1
        codes = np.empty(n, np.uint32)
        perm = np.empty(n, np.uint32)
        left = np.empty(nn, np.int32)
        right = np.empty(nn, np.int32)
        nmass = np.empty(nn)
        ncom = np.empty((nn, 3))
        self._ck(lib().aster_tree_dump(self.h, codes.ctypes.data_as(_PU32), perm.ctypes.data_as(_PU32),
                                       left.ctypes.data_as(_PI32), right.ctypes.data_as(_PI32),
                                       _dp(nmass), _dp(ncom)))
        return dict(codes=codes, perm=perm, left=left, right=right, mass=nmass, com=ncom)

    def morton(self):
        codes = np.empty(self.n, np.uint32)
        box = np.empty(3)
        self._ck(lib().aster_morton(self.h, codes.ctypes.data_as(_PU32), _dp(box)))
        return codes, box

    def bh_stats(self):
        a = ctypes.c_double()
        b = ctypes.c_double()
        self._ck(lib().aster_bh_stats(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def sync(self):
        self._ck(lib().aster_sync(self.h))

    def mark(self, slot):
        self._ck(lib().aster_timer_mark(self.h, slot))

    def elapsed(self, a, b):
        ms = ctypes.c_float()
        self._ck(lib().aster_timer_elapsed(self.h, a, b, ctypes.byref(ms)))
        return ms.value

    def profile(self, enable=True):
        self._ck(lib().aster_profile(self.h, 1 if enable else 0))

    def phase_times(self):
        """{phase name: (summed ms, launches)} since the last call."""
        ms = np.zeros(NUM_PHASES)
        cnt = np.zeros(NUM_PHASES, np.int64)
        self._ck(lib().aster_phase_times(self.h, _dp(ms), cnt.ctypes.data_as(_PI64)))
        return {name: (float(ms[k]), int(cnt[k])) for k, name in enumerate(PHASES) if cnt[k]}

    # ---- batched ensembles
    def batch_set(self, offsets, pos, vel, mass):
        offsets = np.ascontiguousarray(offsets, np.int64)
        nsys = offsets.shape[0] - 1
        pos = _f64(pos).reshape(-1, 3)
        vel = _f64(vel).reshape(-1, 3)
        mass = _f64(mass).reshape(-1)
        self._ck(lib().aster_batch_set(self.h, nsys, offsets.ctypes.data_as(_PI64), _dp(pos), _dp(vel),
                                       _dp(mass)))
        self.batch_n = pos.shape[0]
        self.batch_nsys = nsys

    def batch_get(self):
        pos = np.empty((self.batch_n, 3))
        vel = np.empty((self.batch_n, 3))
        self._ck(lib().aster_batch_get(self.h, _dp(pos), _dp(vel)))
        return pos, vel

    def batch_step(self, integrator, dt, nsteps, steps_per_launch=0, use_graph=True):
        self._ck(lib().aster_batch_step(self.h, integrator, float(dt), int(nsteps), int(steps_per_launch),
                                        1 if use_graph else 0))

    def batch_info(self):
        """(pair interactions, lane-pair slots, wavefronts) of one force evaluation over the batch."""
        p, l, w = _I64(), _I64(), ctypes.c_int32()
        self._ck(lib().aster_batch_info(self.h, ctypes.byref(p), ctypes.byref(l), ctypes.byref(w)))
        return p.value, l.value, w.value

    def batch_energy(self):
        e = np.empty(self.batch_nsys)
        L = np.empty((self.batch_nsys, 3))
        self._ck(lib().aster_batch_energy(self.h, _dp(e), _dp(L)))
        return e, L

    def batch_autotune(self, integrator, dt):
        spb = ctypes.c_int32()
        spl = ctypes.c_int32()
        self._ck(lib().aster_batch_autotune(self.h, integrator, float(dt), ctypes.byref(spb), ctypes.byref(spl)))
        return spb.value, spl.value

    def batch_traj_forward(self, dt, nsteps, snap_steps):
        snap = np.ascontiguousarray(snap_steps, np.int64)
        out = np.empty((len(snap), self.batch_n, 3))
        self._ck(lib().aster_batch_traj_forward(self.h, float(dt), int(nsteps), len(snap),
                                                snap.ctypes.data_as(_PI64), _dp(out)))
        return out

    def batch_traj_backward(self, grad_snap):
        g = _f64(grad_snap)
        gx = np.empty((self.batch_n, 3))
        gv = np.empty((self.batch_n, 3))
        gm = np.empty(self.batch_n)
        self._ck(lib().aster_batch_traj_backward(self.h, _dp(g), _dp(gx), _dp(gv), _dp(gm)))
        return gx, gv, gm

    # ---- trajectory adjoint
    def traj_forward(self, dt, nsteps, ckpt_every, snap_steps, want_pos=True):
        snap = np.ascontiguousarray(snap_steps, np.int64)
        out = np.empty((len(snap), self.n, 3)) if want_pos else None
        self._ck(lib().aster_traj_forward(self.h, float(dt), int(nsteps), int(ckpt_every), len(snap),
                                          snap.ctypes.data_as(_PI64), _dp(out)))
        return out

    def traj_backward(self, grad_snap):
        g = _f64(grad_snap)
        gx = np.empty((self.n, 3))
        gv = np.empty((self.n, 3))
        gm = np.empty(self.n)
        self._ck(lib().aster_traj_backward(self.h, _dp(g), _dp(gx), _dp(gv), _dp(gm)))
        return gx, gv, gm


NOALLOC_MAX_BODIES = 65536  # ASTER_NOALLOC_MAX_BODIES


def forces_noalloc(n, pos_ptr, mass_ptr, G, softening, acc_ptr, stream=None):
    """aster_forces_noalloc: ASTER_ALLPAIRS_FP32_KAHAN accelerations of n
    bodies on device arrays of the current device (AoS-d3 positions and
    output, n masses), enqueued on `stream` (None = the default stream) with
    no device allocation."""
    _check(lib().aster_forces_noalloc(int(n), pos_ptr, mass_ptr, float(G), float(softening), acc_ptr, stream))


def yoshida_weights():
    w = np.empty(3)
    _check(lib().aster_yoshida_weights(_dp(w)))
    return w


def saba_coeffs(order):
    """(drifts c, order + 1 entries; kicks d, order entries) of SABAn, order 2..10."""
    cs = np.empty(order + 1)
    ds = np.empty(order)
    _check(lib().aster_saba_coeffs(int(order), _dp(cs), _dp(ds)))
    return cs, ds


# ---- seeded initial conditions -------------------------------------------

def ic_plummer(n, seed, mtot=1.0, b=1.0, rmax=10.0, mass_mode=0):
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    _check(lib().aster_ic_plummer(n, seed, mtot, b, rmax, mass_mode, _dp(pos), _dp(vel), _dp(mass)))
    return pos, vel, mass


def ic_cube(n, seed, h=4.0, mtot=1.0):
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    _check(lib().aster_ic_cube(n, seed, h, mtot, _dp(pos), _dp(vel), _dp(mass)))
    return pos, vel, mass


def ic_ensemble(nsys, seed, nmin=3, nmax=64):
    offsets = np.empty(nsys + 1, np.int64)
    total = ctypes.c_int64()
    _check(lib().aster_ic_ensemble(nsys, seed, nmin, nmax, offsets.ctypes.data_as(_PI64), ctypes.byref(total),
                                   None, None, None))
    n = total.value
    pos = np.empty((n, 3))
    vel = np.empty((n, 3))
    mass = np.empty(n)
    _check(lib().aster_ic_ensemble(nsys, seed, nmin, nmax, offsets.ctypes.data_as(_PI64), ctypes.byref(total),
                                   _dp(pos), _dp(vel), _dp(mass)))
    return offsets, pos, vel, mass


def ic_threebody(nsys, seed):
    pos = np.empty((3 * nsys, 3))
    vel = np.empty((3 * nsys, 3))
    mass = np.empty(3 * nsys)
    _check(lib().aster_ic_threebody(nsys, seed, _dp(pos), _dp(vel), _dp(mass)))
    return pos, vel, mass


def ic_mixed(n, seed):
    """Half Plummer sphere (mass 1/2, b = 1, r <= 10) and half uniform cube
    [-4, 4]^3 at rest; equal masses 1/n; positions and masses rounded to FP32."""
    h = n // 2
    p1, v1, m1 = ic_plummer(h, seed, mtot=0.5, b=1.0, rmax=10.0)
    p2, v2, m2 = ic_cube(n - h, seed + 1, h=4.0, mtot=(n - h) / n)
    pos = np.concatenate([p1, p2]).astype(np.float32).astype(np.float64)
    vel = np.concatenate([v1, v2])
    mass = np.full(n, 1.0 / n).astype(np.float32).astype(np.float64)
    return pos, vel, mass
