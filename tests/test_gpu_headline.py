"""The timed Barnes-Hut path of bench.py (ASTER_BH_GROUP forces under
aster_step's leapfrog) against the oracle, at the workload's full size.

Workload: half Plummer sphere + half uniform cube [-4, 4]^3 (seeded), equal
masses summing to 1, eps = 0.005 (softening 2.5e-5), automatic Morton box.

- Steps: aster_step(BH_GROUP, LEAPFROG) against the oracle's update_leapfrog
  over the same group traversal summed in FP64 (orc.integrate method 2).
- Full size (2^20 bodies), theta 0.3 / 0.5 / 0.7: interaction and node-visit
  counts equal to the oracle's group traversal; forces equal to the oracle's
  FP64 sums of the same lists within FP32 rounding; force error against FP64
  all-pairs within the stated median and p99 bars.
"""
import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu

SOFT = 0.005 ** 2
N_FULL = 1 << 20
SEED_FULL = 20240  # bench.py's headline state

# Force error against FP64 all-pairs at 2^20 (relative, per body): bars on the
# median and the 99th percentile, per opening angle.
ERROR_BARS = {0.3: (3e-4, 1e-3), 0.5: (1e-3, 4e-3), 0.7: (5e-3, 0.15)}


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def reset_counts(ctx):
    try:
        ctx.bh_stats()
    except A.AsterError:
        pass  # no evaluation on this context yet


@pytest.mark.parametrize("n,theta", [(1 << 18, 0.5), (50000, 0.7)])
def test_bh_group_leapfrog_steps(ctx, orc, n, theta):
    """5 leapfrog steps, dt = 5e-4. The library sums each group list in FP32
    (forces within 1e-5 of the FP64 sums of the same list for all but a
    handful of bodies) and caches the accelerations between steps; the
    oracle re-evaluates twice per step in FP64. Over s steps a force error
    da moves a body by at most ~ s^2 dt^2 |da|, so positions must agree to
    1e-5 s^2 dt^2 max|a| for all but 0.1% of the bodies (whose MAC decision
    may flip with a last-bit position difference) and to 1e-3 s^2 dt^2 max|a|
    for every body; velocities to the same bars over s dt."""
    dt, steps = 5e-4, 5
    pos, vel, mass = A.ic_mixed(n, 31)
    ctx.set_params(G=1.0, softening=SOFT, theta=theta, precision=A.FP32, box=(0.0, 0.0, 0.0), lower=0, upper=-1)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.BH_GROUP, A.LEAPFROG, dt, steps)
    p, v = ctx.get_state()
    a_end = ctx.get_acc()
    rp, rv, ra = orc.integrate(pos, vel, mass, dt, steps, integrator=2, method=2, soft=SOFT, theta=theta)
    amax = np.abs(ra).max()
    ex = np.abs(p - rp).max(1) / (steps ** 2 * dt ** 2 * amax)
    ev = np.abs(v - rv).max(1) / (steps * dt * amax)
    print(f"\nn={n} theta={theta}: position err / (s^2 dt^2 max|a|) median {np.median(ex):.2e} "
          f"p99.9 {np.quantile(ex, 0.999):.2e} max {ex.max():.2e}; velocity median {np.median(ev):.2e} "
          f"max {ev.max():.2e}")
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert np.quantile(ex, 0.999) <= 1e-5 and ex.max() <= 1e-3
    assert np.quantile(ev, 0.999) <= 1e-5 and ev.max() <= 1e-3
    # the accelerations left on the device are those of the final positions
    scale = np.linalg.norm(ra, axis=1)
    e = np.linalg.norm(a_end - ra, axis=1) / np.maximum(scale, np.median(scale))
    assert np.quantile(e, 0.999) < 1e-4


@pytest.mark.parametrize("integrator", [A.LEAPFROG, A.YOSHIDA4])
@pytest.mark.parametrize("soft", [SOFT, 0.0])
def test_bh_group_steps_split_same_bits(ctx, integrator, soft):
    """Several steps in one aster_step call (each kick/drift pass finishes the
    previous step's half-kick, and each drift leaves the next automatic box's
    partials) against one step per call (a separate closing half-kick, the box
    formed from scratch): positions, velocities and accelerations must be the
    same bits (softening 0 takes the masked traversal)."""
    dt, steps = 5e-4, 4
    pos, vel, mass = A.ic_mixed(40000, 77)
    vel = vel + 0.05 * np.random.default_rng(3).standard_normal(vel.shape)  # bodies cross Morton cells
    out = []
    for split in (False, True):
        ctx.set_params(G=1.0, softening=soft, theta=0.5, precision=A.FP32, box=(0.0, 0.0, 0.0), lower=0, upper=-1)
        ctx.set_state(pos, vel, mass)
        if split:
            for _ in range(steps):
                ctx.step(A.BH_GROUP, integrator, dt, 1)
        else:
            ctx.step(A.BH_GROUP, integrator, dt, steps)
        p, v = ctx.get_state()
        out.append((p, v, ctx.get_acc()))
    for q in range(3):
        assert np.array_equal(out[0][q], out[1][q]), q
    assert np.all(np.isfinite(out[0][0]))


@pytest.fixture(scope="module")
def full_state():
    """The headline state and its FP64 all-pairs accelerations (the device
    path that is bit-identical to the reference's Parallelized solver)."""
    pos, vel, mass = A.ic_mixed(N_FULL, SEED_FULL)
    with A.Context(0) as c:
        c.set_params(G=1.0, softening=SOFT, theta=0.5, precision=A.FP32)
        c.set_state(pos, vel, mass)
        ref = c.forces(A.ALLPAIRS_FP64)
    return pos, vel, mass, ref


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_bh_group_full_size(ctx, orc, full_state, theta):
    pos, vel, mass, ref64 = full_state
    ctx.set_params(G=1.0, softening=SOFT, theta=theta, precision=A.FP32, box=(0.0, 0.0, 0.0), lower=0, upper=-1)
    ctx.set_state(pos, vel, mass)
    reset_counts(ctx)
    acc = ctx.forces(A.BH_GROUP)
    inter, visits = ctx.bh_stats()
    ref, ri, rv = orc.bh_group(pos, mass, orc.box_auto(pos), theta, soft=SOFT)
    assert (inter, visits) == (ri, rv)
    scale = np.linalg.norm(ref, axis=1)
    e_list = np.linalg.norm(acc - ref, axis=1) / np.maximum(scale, np.median(scale))
    assert np.median(e_list) < 1e-6 and e_list.max() < 1e-4, (np.median(e_list), e_list.max())
    e = rel_err(acc, ref64)
    med, p99 = np.median(e), np.quantile(e, 0.99)
    print(f"\ntheta={theta}: {inter / N_FULL:.1f} interactions per body; error vs FP64 all-pairs median "
          f"{med:.3e} p99 {p99:.3e} max {e.max():.3e}")
    bar_med, bar_p99 = ERROR_BARS[theta]
    assert med <= bar_med and p99 <= bar_p99
