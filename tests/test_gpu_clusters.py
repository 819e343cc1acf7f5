"""Cluster queue generated on the device, through the C ABI.

Plummer and cube clusters must hold the same bits as the host generators of
the same seed (same per-body draws and the same reduction order); disks are
checked against the oracle's restatement of add_disk (single-precision
rotation outputs: within 3 FP32 ulps) and against the preset's geometry."""
import numpy as np
import pytest

import aster_amd as A
from tests.test_ic import check_disk, plummer_cdf

pytestmark = pytest.mark.gpu

DISK = dict(center=(3.0, -2.0, 1.0), radius=50.0, thickness=4.0, rotation=(30.0, 99.0, -45.0), avr_mass=2.0,
            center_mass=1e4, v=(0.5, 0.0, -0.25), scale=1.5, inner=10.0)


def fp32_close(a, b, ulps=3):
    tol = ulps * np.spacing(np.abs(b).astype(np.float32)).astype(np.float64)
    return np.all(np.abs(a - b) <= tol)


@pytest.mark.parametrize("mode", [0, 1])
def test_plummer_cluster_equals_host_generator(ctx, mode):
    n = 100_003
    ctx.set_params(G=1.0)
    ctx.cluster_add(A.CLUSTER_PLUMMER, n, 21, size=(1.0, 10.0, 0.0), mass=1.0, mass_mode=mode)
    assert ctx.cluster_queued() == (1, n)
    ctx.cluster_load()
    assert ctx.cluster_queued() == (0, 0) and ctx.n == n
    p, v = ctx.get_state()
    hp, hv, hm = A.ic_plummer(n, 21, mass_mode=mode)
    assert np.array_equal(p, hp) and np.array_equal(v, hv)
    ctx.set_params(G=1.0, softening=0.0)
    e, _, L = ctx.energy()
    # masses went over too: the energy of the host state equals the device's
    ctx2 = A.Context(0)
    ctx2.set_params(G=1.0, softening=0.0)
    ctx2.set_state(hp, hv, hm)
    e2, _, L2 = ctx2.energy()
    ctx2.close()
    assert e == e2 and np.array_equal(L, L2)


def test_cube_cluster_translated(ctx):
    n = 50_000
    ctx.cluster_add(A.CLUSTER_CUBE, n, 5, position=(10.0, -3.0, 0.5), velocity=(0.1, 0.0, -0.2), size=(2.0, 0, 0),
                    mass=3.0)
    ctx.cluster_load()
    p, v = ctx.get_state()
    hp, hv, hm = A.ic_cube(n, 5, h=2.0, mtot=3.0)
    assert np.array_equal(p, hp + np.array([10.0, -3.0, 0.5]))
    assert np.array_equal(v, hv + np.array([0.1, 0.0, -0.2]))


def test_disk_cluster_against_oracle(ctx, orc):
    n = 20_000
    ctx.set_params(G=1.0)
    ctx.add_disk(n, seed=4, **DISK)
    assert ctx.n == 1 and ctx.cluster_queued() == (1, n)  # the central body is placed at once
    ctx.cluster_load()
    p, v = ctx.get_state()
    rp, rv, rm = orc.cluster_disk(n, 4, G=1.0, **DISK)
    assert fp32_close(p, rp) and fp32_close(v, rv)
    check_disk(p, v, rm, n, G=1.0, **DISK)
    # masses: the energy of (device positions, oracle masses) equals the
    # device energy only if the device masses are the oracle's
    ctx.set_params(G=1.0, softening=1e-2)
    e, _, _ = ctx.energy()
    c2 = A.Context(0)
    c2.set_params(G=1.0, softening=1e-2)
    c2.set_state(p, v, rm)
    e2, _, _ = c2.energy()
    c2.close()
    assert e == e2


def test_queue_order_append_and_pop(ctx, orc):
    # existing bodies stay first; the disk's central body is appended when the
    # disk is added; load generates the last added cluster first.
    base_p, base_v, base_m = A.ic_cube(7, 99, h=1.0)
    ctx.set_params(G=1.0)
    ctx.set_state(base_p, base_v, base_m)
    ctx.cluster_add(A.CLUSTER_PLUMMER, 3000, 8, size=(1.0, 10.0, 0.0), mass=0.5, position=(100.0, 0.0, 0.0))
    ctx.add_disk(2000, seed=6, **DISK)
    ctx.cluster_add(A.CLUSTER_CUBE, 500, 2, size=(1.0, 0, 0), mass=1.0)
    assert ctx.cluster_queued() == (3, 5500) and ctx.n == 8
    ctx.cluster_pop()  # drops the cube
    assert ctx.cluster_queued() == (2, 5000)
    ctx.cluster_load()
    assert ctx.n == 8 + 5000
    p, v = ctx.get_state()
    assert np.array_equal(p[:7], base_p) and np.array_equal(v[:7], base_v)
    rp, rv, _ = orc.cluster_disk(2000, 6, G=1.0, **DISK)
    assert fp32_close(p[7], rp[0]) and fp32_close(p[8:2008], rp[1:]) and fp32_close(v[8:2008], rv[1:])
    hp, hv, _ = A.ic_plummer(3000, 8, mtot=0.5)
    assert np.array_equal(p[2008:], hp + np.array([100.0, 0.0, 0.0])) and np.array_equal(v[2008:], hv)
    # popping an empty queue and loading nothing are no-ops
    ctx.cluster_pop()
    ctx.cluster_load()
    assert ctx.n == 5008
    # forces run on the appended state
    acc = ctx.forces(A.ALLPAIRS_FP64)
    assert np.all(np.isfinite(acc))


def test_cluster_errors(ctx):
    with pytest.raises(A.AsterError):
        ctx.cluster_add(A.CLUSTER_PLUMMER, 0, 1)  # no bodies
    with pytest.raises(A.AsterError):
        ctx.cluster_add(A.CLUSTER_PLUMMER, 10, 1, mass=-1.0)
    with pytest.raises(A.AsterError):
        ctx.cluster_add(A.CLUSTER_CUBE, 10, 1, size=(1, 0, 0), rotation=(10, 0, 0))  # disks only
    with pytest.raises(A.AsterError):
        ctx.add_disk(10, (0, 0, 0), radius=5.0, thickness=1.0, rotation=(0, 0, 0))  # inner radius 10 > radius
    with pytest.raises(A.AsterError):
        ctx.add_disk(10, (0, 0, 0), radius=50.0, thickness=0.0, rotation=(0, 0, 0))
    with pytest.raises(A.AsterError):
        ctx.cluster_add(7, 10, 1)
    with pytest.raises(A.AsterError) as ei:
        ctx.cluster_add(A.CLUSTER_CUBE, 1 << 29, 1, size=(1, 0, 0))
    assert ei.value.code == -5
    assert ctx.cluster_queued() == (0, 0) and ctx.n == 0


def test_large_plummer_cluster_on_device(ctx):
    # 2^24 bodies generated on the device; radial profile (KS on a sample),
    # barycentre at the origin.
    from scipy import stats
    n = 1 << 24
    ctx.set_params(G=1.0)
    ctx.cluster_add(A.CLUSTER_PLUMMER, n, 77, size=(1.0, 10.0, 0.0), mass=1.0)
    ctx.cluster_load()
    com = ctx.center_of_mass()
    assert np.abs(com).max() < 1e-14
    p, _ = ctx.get_state()
    r = np.linalg.norm(p[:: 1 << 10], axis=1)
    assert stats.kstest(r, lambda x: plummer_cdf(x, 10.0)).pvalue > 1e-3
    # truncated at r <= 10 before recentring; the shift is the sample's
    # barycentre, of order rms(r) / sqrt(n)
    assert np.linalg.norm(p, axis=1).max() <= 10.0 + 1e-2
