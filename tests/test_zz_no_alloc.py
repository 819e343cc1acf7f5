"""aster_forces_noalloc (the allocation-free all-pairs entry point) against
aster_forces with ASTER_ALLPAIRS_FP32_KAHAN and the entry point's source split,
through the C ABI: the same bits on two calls in a row, the second with a
smaller input (and an SI-scale system, whose mass scale the entry point forms
on the device), and the FP32 bar against the oracle."""
import ctypes
import os

import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyKind


class DeviceArray:
    """Device memory from the HIP runtime the library itself uses."""
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
        assert self._hip.hipDeviceSynchronize() == 0
        out = np.empty_like(self.host)
        assert self._hip.hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), self.ptr, ctypes.c_size_t(out.nbytes),
                                   D2H) == 0
        return out

    def free(self):
        self._hip.hipFree(self.ptr)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def si_system(n, seed):
    pos, vel, _ = A.ic_plummer(n, seed)
    mass = 10.0 ** np.random.default_rng(seed).uniform(28.0, 31.0, n)
    return pos * 1e11, vel, mass


CASES = [
    # n, G, softening, system
    (A.NOALLOC_MAX_BODIES, 1.0, 2.5e-5, lambda n: A.ic_plummer(n, 3, mass_mode=1)),
    (3000, 6.674e-11, 1e-14, lambda n: si_system(n, 4)),
]


def noalloc_split(n):
    """The entry point's source split as its declaration states it:
    S = min(64, npad / 1024, ceil(8192 / (npad / 64))), at least 1, npad = n
    rounded up to 256."""
    npad = -(-n // 256) * 256
    waves = npad // 64
    return max(1, min(64, npad // 1024, -(-8192 // waves)))


def test_noalloc_matches_aster_forces(ctx, orc):
    for n, G, soft, make in CASES:
        pos, vel, mass = make(n)
        ctx.set_params(G=G, softening=soft, lower=0, upper=-1, pn_flags=A.NEWTON, split=noalloc_split(n))
        ctx.set_state(pos, vel, mass)
        ref = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
        dp, dm, da = DeviceArray(pos), DeviceArray(mass), DeviceArray(np.full((n, 3), np.nan))
        try:
            print(f"aster_forces_noalloc n={n} G={G} softening={soft}", flush=True)
            A.forces_noalloc(n, dp.ptr, dm.ptr, G, soft, da.ptr)
            out = da.read()
        finally:
            for d in (dp, dm, da):
                d.free()
        assert np.array_equal(out, ref), n
        if n <= 4096:
            assert rel_err(out, orc.allpairs(pos, vel, mass, G=G, soft=soft)).max() <= 1e-5


def test_noalloc_rejects_beyond_its_limit():
    with pytest.raises(A.AsterError) as ei:
        A.forces_noalloc(A.NOALLOC_MAX_BODIES + 1, ctypes.c_void_p(16), ctypes.c_void_p(16), 1.0, 0.0,
                         ctypes.c_void_p(16))
    assert ei.value.code == -5
    with pytest.raises(A.AsterError) as ei:
        A.forces_noalloc(10, None, None, 1.0, 0.0, None)
    assert ei.value.code == -1
