"""GPU state transfer (host and device AoS-d3 import/export) and the Graphs
collectors, through the C ABI. Transfers are exact copies, so they are checked
bit for bit; the collectors against numpy closed forms with the tolerance
written in each test."""
import ctypes
import os

import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyKind


class DeviceArray:
    """Device memory from the HIP runtime the library itself uses (a second
    runtime in the process, such as a framework's bundled one, would not see
    the library's device state)."""
    _hip = None

    def __init__(self, host):
        if DeviceArray._hip is None:
            DeviceArray._hip = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                                        "libamdhip64.so"))
        self.host = np.ascontiguousarray(host, np.float64)
        self.ptr = ctypes.c_void_p()
        assert self._hip.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(self.host.nbytes)) == 0
        assert self._hip.hipMemcpy(self.ptr, self.host.ctypes.data_as(ctypes.c_void_p),
                                   ctypes.c_size_t(self.host.nbytes), H2D) == 0

    def read(self):
        out = np.empty_like(self.host)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes),
                                   D2H) == 0
        return out

    def free(self):
        self._hip.hipFree(self.ptr)


def plummer(n, seed=1):
    return A.ic_plummer(n, seed)


def test_host_round_trip_bit_exact(ctx):
    pos, vel, mass = plummer(1000, 2)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    p, v = ctx.get_state()
    assert np.array_equal(p, pos) and np.array_equal(v, vel)
    ctx.set_state(pos, None, mass)  # zero velocities
    _, v = ctx.get_state()
    assert not v.any()


def test_device_import_export_bit_exact(ctx, orc):
    pos, vel, mass = plummer(777, 3)
    dp, dv, dm = DeviceArray(pos), DeviceArray(vel), DeviceArray(mass)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state_device(777, dp.ptr, dv.ptr, dm.ptr)
    p, v = ctx.get_state()
    assert np.array_equal(p, pos) and np.array_equal(v, vel)
    # forces from a device-loaded state equal the CPU path
    assert np.array_equal(ctx.forces(A.ALLPAIRS_FP64), orc.allpairs(pos, vel, mass, soft=1e-4))
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, 1e-3, 3)
    op, ov = DeviceArray(np.zeros_like(pos)), DeviceArray(np.zeros_like(vel))
    ctx.get_state_device(op.ptr, ov.ptr)
    p, v = ctx.get_state()
    assert np.array_equal(op.read(), p) and np.array_equal(ov.read(), v)
    for d in (dp, dv, dm, op, ov):
        d.free()


def test_device_import_rejects_bad_input(ctx):
    pos, vel, mass = plummer(64, 4)
    bad_mass = mass.copy()
    bad_mass[17] = 0.0
    dp, dm = DeviceArray(pos), DeviceArray(bad_mass)
    with pytest.raises(A.AsterError):
        ctx.set_state_device(64, dp.ptr, None, dm.ptr)
    bad_pos = pos.copy()
    bad_pos[5, 1] = np.nan
    dq, dn = DeviceArray(bad_pos), DeviceArray(mass)
    with pytest.raises(A.AsterError):
        ctx.set_state_device(64, dq.ptr, None, dn.ptr)
    ctx.set_state_device(64, dp.ptr, None, dn.ptr)  # good input loads after a rejection
    assert np.array_equal(ctx.get_state()[0], pos)
    for d in (dp, dm, dq, dn):
        d.free()


def test_collectors(ctx, orc):
    # Tolerances: Hamiltonian and |L| equal aster_energy's values exactly (same
    # reduction); the barycentre within 1e-14 of numpy's sum (summation order).
    pos, vel, mass = plummer(2048, 5)
    pos = pos + np.array([0.3, -0.2, 0.1])
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    e_ref, _, L = ctx.energy()
    assert ctx.collect(A.COLLECT_HAMILTONIAN) == e_ref
    assert ctx.collect(A.COLLECT_ANGULAR_MOMENTUM) == pytest.approx(np.sqrt((L ** 2).sum()), rel=1e-15)
    assert ctx.collect(A.COLLECT_ERROR) == 0.0  # the first call records E0
    com = (mass[:, None] * pos).sum(0) / mass.sum()
    np.testing.assert_allclose(ctx.center_of_mass(), com, rtol=0, atol=1e-14)
    d = ctx.collect(A.COLLECT_DISTANCE, 123)
    assert d == pytest.approx(np.linalg.norm(com - pos[123] + 0.1), rel=1e-13)
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, 1e-3, 20)
    e1 = ctx.energy()[0]
    assert ctx.collect(A.COLLECT_ERROR) == pytest.approx(abs(e_ref - e1) / abs(e_ref), rel=1e-12)
    with pytest.raises(A.AsterError):
        ctx.collect(A.COLLECT_DISTANCE, 2048)
    ctx.set_state(pos, vel, mass)  # a new state resets E0
    assert ctx.collect(A.COLLECT_ERROR) == 0.0


def test_integration_report(ctx, orc):
    # Simulation::integrate(times, precision_test): energies before/after in
    # the reference's definition, |dE|/|E0|, step count; the state equals a
    # plain aster_integrate of the same steps.
    pos, vel, mass = A.ic_plummer(512, 3)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    e0, _, _ = ctx.energy()
    rep = ctx.integrate_report(A.ALLPAIRS_FP64, A.LEAPFROG, 1e-3, 20)
    e1, _, _ = ctx.energy()
    p, v = ctx.get_state()
    assert rep["steps"] == 20 and rep["tested_precision"] == 1
    assert rep["initial_energy"] == e0 and rep["final_energy"] == e1
    assert rep["precision"] == abs(e0 - e1) / abs(e0)
    assert rep["total_time_ms"] > 0 and abs(rep["time_per_step_ms"] * 20 - rep["total_time_ms"]) < 1e-9
    rp, rv, _ = orc.integrate(pos, vel, mass, 1e-3, 20, integrator=2, soft=1e-4)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    rep = ctx.integrate_report(A.ALLPAIRS_FP64, A.SABA, 1e-3, 3, order=4, precision_test=False)
    assert rep["tested_precision"] == 0 and rep["precision"] == 0.0 and rep["steps"] == 3


def test_integrate_until_halting_condition(ctx, orc):
    # Simulation::integrate(halting_condition, ...): the condition is checked
    # before every step and may read the state; max_steps caps the run.
    pos, vel, mass = A.ic_plummer(256, 4)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    seen = []

    def halt(c):
        p, _ = c.get_state()
        seen.append(p[0, 0])
        return len(seen) > 7

    rep = ctx.integrate_until(A.ALLPAIRS_FP64, A.YOSHIDA4, 1e-3, halt)
    assert rep["steps"] == 7 and len(seen) == 8
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 1e-3, 7, integrator=5, soft=1e-4)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    assert seen[0] == pos[0, 0] and seen[-1] == rp[0, 0]
    rep = ctx.integrate_until(A.ALLPAIRS_FP64, A.LEAPFROG, 1e-3, lambda c: False, max_steps=3, precision_test=False)
    assert rep["steps"] == 3
