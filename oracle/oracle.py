"""ctypes wrapper of liboracle.so -- TEST INFRASTRUCTURE ONLY.

The oracle is a C++ FP64 restatement of Aster's CPU path (see
oracle/aster_oracle.cpp for the reference file:line each function follows).
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
it, and only as the checker. Arrays follow the reference layout: (n, 3)
float64 positions/velocities/accelerations, (n,) masses.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_PU32 = ctypes.POINTER(ctypes.c_uint32)
_PI32 = ctypes.POINTER(ctypes.c_int32)
_PI64 = ctypes.POINTER(ctypes.c_int64)
_I64 = ctypes.c_int64
_dbl = ctypes.c_double

_SIG = {
    "orc_allpairs": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, ctypes.c_int, _I64, _I64, _D, ctypes.c_int],
    "orc_pair_force": [_dbl, _dbl, _D, _D, _D, _D, _dbl, _dbl, _dbl, ctypes.c_int, _D],
    "orc_morton": [_I64, _D, _D, _PU32],
    "orc_box_auto": [_I64, _D, _D],
    "orc_sort": [_I64, _PU32, _PU32],
    "orc_karras": [_I64, _PU32, _PI32, _PI32],
    "orc_build_tree": [_I64, _D, _D, _D, _D, _PU32, _PU32, _PI32, _PI32, _D, _D, _D],
    "orc_bh": [_I64, _D, _D, _D, _D, _dbl, _dbl, _dbl, _dbl, ctypes.c_int, _D, _PI64, ctypes.c_int],
    "orc_bh_group": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, _D, _PI64, _PI64, ctypes.c_int],
    "orc_bh_group_bruteforce_count": [_I64, _D, _D, _D, _dbl],
    "orc_bh_group_pn": [_I64, _D, _D, _D, _D, _dbl, _dbl, _dbl, _dbl, ctypes.c_int, _D, ctypes.c_int],
    "orc_energy": [_I64, _D, _D, _D, _dbl, _dbl, _D, _D, _D],
    "orc_yoshida_weights": [_D],
    "orc_integrate": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, ctypes.c_int, ctypes.c_int, _D, _dbl, ctypes.c_int,
                      _dbl, _I64, _D, ctypes.c_int],
    "orc_integrate_saba": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, ctypes.c_int, ctypes.c_int, _D, _dbl, ctypes.c_int,
                           _dbl, _I64, _D, ctypes.c_int],
    "orc_saba_coeffs": [ctypes.c_int, _D, _D],
    "orc_kepler_drift": [_D, _D, _dbl, _dbl],
    "orc_integrate_wh": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, _I64],
    "orc_traj_grad": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, _I64, ctypes.c_int, _PI64, _D, _D, _D, _D, _D],
    "orc_traj_loss": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, _I64, ctypes.c_int, _PI64, _D, _D],
    "orc_allpairs_sample": [_I64, _D, _D, _D, _dbl, _dbl, _dbl, ctypes.c_int, _I64, _I64, ctypes.c_int, _D, _D],
    "orc_bh_sample": [_I64, _D, _D, _D, _D, _dbl, _dbl, _dbl, _dbl, ctypes.c_int, _I64, ctypes.c_int, _D, _D,
                      _PI64],
    "orc_splitmix64": [ctypes.c_uint64, ctypes.c_int, ctypes.POINTER(ctypes.c_uint64)],
    "orc_stream_uniforms": [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, _D],
    "orc_ic_plummer": [_I64, ctypes.c_uint64, _dbl, _dbl, _dbl, _dbl, ctypes.c_int, _D, _D, _D],
    "orc_ic_cube": [_I64, ctypes.c_uint64, _dbl, _dbl, _D, _D, _D],
    "orc_cluster_disk": [_I64, ctypes.c_uint64, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _dbl, _D, _D, _D, _D, _D, _D],
}


def allpairs_sample(pos, vel, mass, t0, t1, G=1.0, c=1e4, soft=0.0, pn=1, threads=0):
    """(accelerations of targets [t0, t1), wall seconds)."""
    pos, vel, mass = _c(pos), _c(vel), _c(mass)
    acc = np.empty((t1 - t0, 3))
    sec = ctypes.c_double()
    assert lib().orc_allpairs_sample(len(mass), _dp(pos), _dp(vel), _dp(mass), G, c, soft, pn, t0, t1, threads,
                                     ctypes.byref(sec), _dp(acc)) == 0
    return acc, sec.value


def bh_sample(pos, vel, mass, box, theta, stride, G=1.0, c=1e4, soft=0.0, pn=1, threads=0):
    """(build seconds, traversal seconds of the sampled bodies, their interactions)."""
    pos, vel, mass, box = _c(pos), _c(vel), _c(mass), _c(box)
    tb = ctypes.c_double()
    tt = ctypes.c_double()
    cnt = ctypes.c_int64()
    assert lib().orc_bh_sample(len(mass), _dp(pos), _dp(vel), _dp(mass), _dp(box), theta, G, c, soft, pn, stride,
                               threads, ctypes.byref(tb), ctypes.byref(tt), ctypes.byref(cnt)) == 0
    return tb.value, tt.value, cnt.value

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        for k, args in _SIG.items():
            f = getattr(L, k)
            f.argtypes = args
            f.restype = ctypes.c_int64 if k == "orc_bh_group_bruteforce_count" else ctypes.c_int
        _lib = L
    return _lib


def _dp(a):
    return None if a is None else a.ctypes.data_as(_D)


def _c(a, dt=np.float64):
    return np.ascontiguousarray(a, dtype=dt)


def allpairs(pos, vel, mass, G=1.0, c=1e4, soft=0.0, pn=1, lower=0, upper=-1, threads=0):
    pos, vel, mass = _c(pos), _c(vel), _c(mass)
    acc = np.empty_like(pos)
    r = lib().orc_allpairs(len(mass), _dp(pos), _dp(vel), _dp(mass), G, c, soft, pn, lower, upper, _dp(acc), threads)
    assert r == 0
    return acc


def pair_force(m1, m2, v1, v2, p1, p2, G=1.0, c=1e4, soft=0.0, pn=1):
    out = np.empty(3)
    args = [_c(v1), _c(v2), _c(p1), _c(p2)]
    lib().orc_pair_force(m1, m2, *[_dp(a) for a in args], G, c, soft, pn, _dp(out))
    return out


def morton(pos, box):
    pos, box = _c(pos), _c(box)
    codes = np.empty(len(pos), np.uint32)
    lib().orc_morton(len(pos), _dp(pos), _dp(box), codes.ctypes.data_as(_PU32))
    return codes


def box_auto(pos):
    pos = _c(pos)
    box = np.empty(3)
    lib().orc_box_auto(len(pos), _dp(pos), _dp(box))
    return box


def sort(codes):
    codes = _c(codes, np.uint32)
    perm = np.empty(len(codes), np.uint32)
    lib().orc_sort(len(codes), codes.ctypes.data_as(_PU32), perm.ctypes.data_as(_PU32))
    return perm


def karras(sorted_codes):
    sc = _c(sorted_codes, np.uint32)
    n = len(sc)
    left = np.empty(2 * n - 1, np.int32)
    right = np.empty(2 * n - 1, np.int32)
    assert lib().orc_karras(n, sc.ctypes.data_as(_PU32), left.ctypes.data_as(_PI32), right.ctypes.data_as(_PI32)) == 0
    return left, right


def build_tree(pos, vel, mass, box):
    pos, vel, mass, box = _c(pos), _c(vel), _c(mass), _c(box)
    n = len(mass)
    nn = 2 * n - 1
    codes = np.empty(n, np.uint32)
    perm = np.empty(n, np.uint32)
    left = np.empty(nn, np.int32)
    right = np.empty(nn, np.int32)
    nmass = np.empty(nn)
    ncom = np.empty((nn, 3))
    nvel = np.empty((nn, 3))
    assert lib().orc_build_tree(n, _dp(pos), _dp(vel), _dp(mass), _dp(box), codes.ctypes.data_as(_PU32),
                                perm.ctypes.data_as(_PU32), left.ctypes.data_as(_PI32), right.ctypes.data_as(_PI32),
                                _dp(nmass), _dp(ncom), _dp(nvel)) == 0
    return dict(codes=codes, perm=perm, left=left, right=right, mass=nmass, com=ncom, vel=nvel)


def bh(pos, vel, mass, box, theta, G=1.0, c=1e4, soft=0.0, pn=1, threads=0, counts=False):
    pos, vel, mass, box = _c(pos), _c(vel), _c(mass), _c(box)
    n = len(mass)
    acc = np.empty((n, 3))
    cnt = np.empty(n, np.int64) if counts else None
    r = lib().orc_bh(n, _dp(pos), _dp(vel), _dp(mass), _dp(box), theta, G, c, soft, pn, _dp(acc),
                     None if cnt is None else cnt.ctypes.data_as(_PI64), threads)
    assert r == 0, "traversal stack overflow"
    return (acc, cnt) if counts else acc


def bh_group(pos, mass, box, theta, G=1.0, soft=0.0, threads=0):
    """Group traversal restated (the library's ASTER_BH_GROUP): accelerations
    (n, 3) in FP64 from the FP32 records, plus the interaction and node-visit
    counts, which equal the library's counters exactly."""
    pos, mass, box = _c(pos), _c(mass), _c(box)
    n = len(mass)
    acc = np.empty((n, 3))
    inter = ctypes.c_int64()
    vis = ctypes.c_int64()
    lib().orc_bh_group(n, _dp(pos), _dp(mass), _dp(box), theta, G, soft, _dp(acc), ctypes.byref(inter),
                       ctypes.byref(vis), threads)
    return acc, inter.value, vis.value


def bh_group_pn(pos, vel, mass, box, theta, G=1.0, c=1e4, soft=0.0, pn=0x11, threads=0):
    """The post-Newtonian part of the group traversal's accelerations (n, 3):
    the terms `pn` selects between each body and its group's list, with the
    nodes' FP64 centres of mass, masses and mass-weighted velocities."""
    pos, vel, mass, box = _c(pos), _c(vel), _c(mass), _c(box)
    n = len(mass)
    acc = np.empty((n, 3))
    assert lib().orc_bh_group_pn(n, _dp(pos), _dp(vel), _dp(mass), _dp(box), theta, G, c, soft, pn, _dp(acc),
                                 threads) == 0
    return acc


def bh_group_bruteforce_count(pos, mass, box, theta):
    """Interaction count of the group traversal from an ancestor-chain brute force (no walk)."""
    pos, mass, box = _c(pos), _c(mass), _c(box)
    return int(lib().orc_bh_group_bruteforce_count(len(mass), _dp(pos), _dp(mass), _dp(box), theta))


def energy(pos, vel, mass, G=1.0, soft=0.0):
    pos, vel, mass = _c(pos), _c(vel), _c(mass)
    e = ctypes.c_double()
    ep = ctypes.c_double()
    L = np.empty(3)
    lib().orc_energy(len(mass), _dp(pos), _dp(vel), _dp(mass), G, soft, ctypes.byref(e), ctypes.byref(ep), _dp(L))
    return e.value, ep.value, L


def yoshida_weights():
    w = np.empty(3)
    lib().orc_yoshida_weights(_dp(w))
    return w


def integrate(pos, vel, mass, dt, nsteps, integrator=2, method=0, G=1.0, c=1e4, soft=0.0, pn=1, box=None,
              theta=0.5, threads=0, order=0):
    """integrator: 0 Euler, 1 SABA (order 1..10), 2 leapfrog, 5 Yoshida-4."""
    pos = _c(pos).copy()
    vel = _c(vel).copy()
    mass = _c(mass)
    acc = np.empty_like(pos)
    boxa = _c(box if box is not None else np.zeros(3))
    if integrator in (0, 1):
        order = 1 if integrator == 0 else order
        r = lib().orc_integrate_saba(len(mass), _dp(pos), _dp(vel), _dp(mass), G, c, soft, pn, method, _dp(boxa),
                                     theta, order, dt, nsteps, _dp(acc), threads)
    else:
        r = lib().orc_integrate(len(mass), _dp(pos), _dp(vel), _dp(mass), G, c, soft, pn, method, _dp(boxa), theta,
                                integrator, dt, nsteps, _dp(acc), threads)
    assert r == 0
    return pos, vel, acc


def kepler_drift(q, u, mu, dt):
    """Kepler drift of relative position q and velocity u about mu for dt."""
    q = _c(q).copy()
    u = _c(u).copy()
    assert lib().orc_kepler_drift(_dp(q), _dp(u), mu, dt) == 0
    return q, u


def integrate_wh(pos, vel, mass, dt, nsteps, G=1.0, soft=0.0):
    """Democratic-heliocentric Wisdom-Holman steps of one system (body 0 central)."""
    pos = _c(pos).copy()
    vel = _c(vel).copy()
    mass = _c(mass)
    assert lib().orc_integrate_wh(len(mass), _dp(pos), _dp(vel), _dp(mass), G, soft, dt, nsteps) == 0
    return pos, vel


def saba_coeffs(order):
    """(drifts c, order + 1 entries; kicks d, order entries) of SABAn, n >= 2."""
    cs = np.empty(order + 1)
    ds = np.empty(order)
    assert lib().orc_saba_coeffs(order, _dp(cs), _dp(ds)) == 0
    return cs, ds


def traj_grad(pos, vel, mass, dt, nsteps, snap_steps, target, G=1.0, soft=0.0):
    pos, vel, mass, target = _c(pos), _c(vel), _c(mass), _c(target)
    snaps = _c(snap_steps, np.int64)
    n = len(mass)
    loss = ctypes.c_double()
    gx = np.empty((n, 3))
    gv = np.empty((n, 3))
    gm = np.empty(n)
    lib().orc_traj_grad(n, _dp(pos), _dp(vel), _dp(mass), G, soft, dt, nsteps, len(snaps),
                        snaps.ctypes.data_as(_PI64), _dp(target), ctypes.byref(loss), _dp(gx), _dp(gv), _dp(gm))
    return loss.value, gx, gv, gm


def traj_loss(pos, vel, mass, dt, nsteps, snap_steps, target, G=1.0, soft=0.0):
    pos, vel, mass, target = _c(pos), _c(vel), _c(mass), _c(target)
    snaps = _c(snap_steps, np.int64)
    loss = ctypes.c_double()
    lib().orc_traj_loss(len(mass), _dp(pos), _dp(vel), _dp(mass), G, soft, dt, nsteps, len(snaps),
                        snaps.ctypes.data_as(_PI64), _dp(target), ctypes.byref(loss))
    return loss.value


def splitmix64(state, k):
    """Raw SplitMix64 outputs 1..k from `state`."""
    out = np.empty(k, np.uint64)
    lib().orc_splitmix64(state, k, out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)))
    return out


def stream_uniforms(seed, i, k):
    out = np.empty(k)
    lib().orc_stream_uniforms(seed, i, k, _dp(out))
    return out


def ic_plummer(n, seed, mtot=1.0, b=1.0, rmax=10.0, mass_mode=0, G=1.0):
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    lib().orc_ic_plummer(n, seed, G, mtot, b, rmax, mass_mode, _dp(pos), _dp(vel), _dp(mass))
    return pos, vel, mass


def ic_cube(n, seed, h=4.0, mtot=1.0):
    pos, vel, mass = np.empty((n, 3)), np.empty((n, 3)), np.empty(n)
    lib().orc_ic_cube(n, seed, h, mtot, _dp(pos), _dp(vel), _dp(mass))
    return pos, vel, mass


def cluster_disk(n, seed, center, radius, thickness, rotation, avr_mass, center_mass, v=(0, 0, 0), scale=1.0,
                 inner=10.0, G=1.0):
    """add_disk with seeded draws: body 0 central, bodies 1..n the disk."""
    pos, vel, mass = np.empty((n + 1, 3)), np.empty((n + 1, 3)), np.empty(n + 1)
    lib().orc_cluster_disk(n, seed, G, radius, thickness, inner, scale, avr_mass, center_mass,
                           _dp(_c(center)), _dp(_c(v)), _dp(_c(rotation)), _dp(pos), _dp(vel), _dp(mass))
    return pos, vel, mass
