"""GPU parity for all-pairs forces, integrators and energy, through the C ABI.

FP64 paths must be bit-identical to the oracle (the reference's CPU path).
FP32 paths are checked against FP64 with the tolerances written in each test."""
import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu

DT_PARITY = float(np.float32(1e-3))  # Simulation::set_dt stores dt through float


def plummer(n, seed=1, mass_mode=0):
    return A.ic_plummer(n, seed, mass_mode=mass_mode)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


@pytest.mark.parametrize("n", [1, 2, 3, 63, 257, 4096])
def test_allpairs_fp64_bit_identical(ctx, orc, n):
    pos, vel, mass = plummer(n, seed=n)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(A.ALLPAIRS_FP64)
    ref = orc.allpairs(pos, vel, mass, G=1.0, soft=1e-4)
    assert np.array_equal(acc, ref)


def test_allpairs_fp64_pn1_and_bounds_bit_identical(ctx, orc):
    pos, vel, mass = plummer(700, seed=3, mass_mode=1)
    vel = vel * 50.0
    ctx.set_params(G=1.0, c=1e2, softening=2.5e-5, pn_flags=A.NEWTON | A.PN1, lower=100, upper=600)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(A.ALLPAIRS_FP64)
    ref = orc.allpairs(pos, vel, mass, G=1.0, c=1e2, soft=2.5e-5, pn=0x11, lower=100, upper=600)
    assert np.array_equal(acc, ref)


ALL_PN = A.NEWTON | A.PN1 | A.PN2 | A.PN25


@pytest.mark.parametrize("flags", [A.NEWTON | A.PN2, A.NEWTON | A.PN25, ALL_PN])
def test_allpairs_fp64_pn2_pn25_bit_identical(ctx, orc, flags):
    # get_pn2 / get_pn25 in the reference's operation order, FP64.
    pos, vel, mass = plummer(600, seed=13, mass_mode=1)
    vel = vel * 30.0
    ctx.set_params(G=1.0, c=20.0, softening=2.5e-5, pn_flags=flags)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(A.ALLPAIRS_FP64)
    ref = orc.allpairs(pos, vel, mass, G=1.0, c=20.0, soft=2.5e-5, pn=flags)
    assert np.array_equal(acc, ref)


def test_all_pn_trajectory_bit_identical(ctx, orc):
    pos, vel, mass = plummer(256, seed=14)
    vel = vel * 20.0
    ctx.set_params(G=1.0, c=20.0, softening=1e-4, pn_flags=ALL_PN)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, A.YOSHIDA4, 1e-3, 3)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 1e-3, 3, integrator=5, c=20.0, soft=1e-4, pn=ALL_PN)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)


def test_mixed_precision_with_pn_terms(ctx, orc):
    # FP32 Newtonian part + FP64 post-Newtonian correction; tolerance 1e-5
    # max relative error against the FP64 path.
    pos, vel, mass = plummer(4096, seed=15, mass_mode=1)
    vel = vel * 30.0
    ctx.set_params(G=1.0, c=20.0, softening=2.5e-5, pn_flags=ALL_PN)
    ctx.set_state(pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    assert np.array_equal(ref, orc.allpairs(pos, vel, mass, G=1.0, c=20.0, soft=2.5e-5, pn=ALL_PN))
    assert rel_err(ctx.forces(A.ALLPAIRS_FP32_KAHAN), ref).max() < 1e-5


def test_allpairs_no_softening_bit_identical(ctx, orc):
    pos, vel, mass = plummer(300, seed=4)
    ctx.set_params(G=1.0, softening=0.0)
    ctx.set_state(pos, vel, mass)
    assert np.array_equal(ctx.forces(A.ALLPAIRS_FP64), orc.allpairs(pos, vel, mass, soft=0.0))
    a32 = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
    assert np.all(np.isfinite(a32))
    assert np.median(rel_err(a32, orc.allpairs(pos, vel, mass, soft=0.0))) < 1e-6


def test_allpairs_fp32_vs_oracle(ctx, orc):
    # FP32 pair forces, FP32 Kahan accumulation vs the FP64 CPU path.
    # Tolerances: max relative error per body <= 1e-5, median <= 1e-6.
    pos, vel, mass = plummer(8192, seed=5, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5)
    ctx.set_state(pos, vel, mass)
    ref = orc.allpairs(pos, vel, mass, soft=2.5e-5)
    e = rel_err(ctx.forces(A.ALLPAIRS_FP32_KAHAN), ref)
    print(f"\nFP32+Kahan vs FP64 oracle n=8192: median {np.median(e):.2e} p99 {np.quantile(e, .99):.2e} max {e.max():.2e}")
    assert np.median(e) <= 1e-6 and e.max() <= 1e-5
    e2 = rel_err(ctx.forces(A.ALLPAIRS_FP32), ref)
    print(f"FP32 plain vs FP64 oracle n=8192: median {np.median(e2):.2e} max {e2.max():.2e}")
    assert np.median(e2) <= 1e-6 and e2.max() <= 1e-4


def test_allpairs_fp32_c3_size(ctx):
    # N = 65536 log-uniform masses (the mixed-precision configuration); FP64
    # device path is the reference here (it is bit-identical to the oracle).
    pos, vel, mass = plummer(65536, seed=6, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5)
    ctx.set_state(pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    e = rel_err(ctx.forces(A.ALLPAIRS_FP32_KAHAN), ref)
    e2 = rel_err(ctx.forces(A.ALLPAIRS_FP32), ref)
    print(f"\nn=65536 Kahan: median {np.median(e):.2e} p99 {np.quantile(e, .99):.2e} max {e.max():.2e}; "
          f"plain: median {np.median(e2):.2e} max {e2.max():.2e}")
    assert np.median(e) <= 1e-6 and e.max() <= 1e-5
    # Newton's third law on the whole system (momentum change rate)
    p = (mass[:, None] * ctx.get_acc()).sum(0)
    assert np.abs(p).max() < 1e-6 * np.abs(mass[:, None] * ref).sum()


def test_split_does_not_change_fp64_parity_of_combine(ctx, orc):
    pos, vel, mass = plummer(2000, seed=7)
    for split in (1, 3, 8):
        ctx.set_params(G=1.0, softening=1e-4, split=split)
        ctx.set_state(pos, vel, mass)
        a = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
        assert np.median(rel_err(a, orc.allpairs(pos, vel, mass, soft=1e-4))) < 1e-6


@pytest.mark.parametrize("n,lower,upper", [(1000, 0, -1), (5000, 0, -1), (100003, 0, -1), (100003, 1234, 77777)])
def test_fp32_source_slices_at_the_split_rules_edges(ctx, n, lower, upper):
    """The default source split (one slice below 2048 bodies, a short last
    slice, 42 slices at 100003 bodies, a source range) against one slice and
    against the FP64 path: every split must give the FP32 bars (Kahan: 1e-5,
    plain 1e-4, per body relative to max(|a_i|, median |a|), since a source
    range leaves some bodies with nearly cancelled forces), and the splits
    differ only by FP32 rounding."""
    pos, vel, mass = plummer(n, seed=n % 97, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5, lower=lower, upper=upper, split=0)
    ctx.set_state(pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    scale = np.linalg.norm(ref, axis=1)
    scale = np.maximum(scale, np.median(scale))
    for method, bar in ((A.ALLPAIRS_FP32_KAHAN, 1e-5), (A.ALLPAIRS_FP32, 1e-4)):
        out = []
        for split in (0, 1):
            ctx.set_params(G=1.0, softening=2.5e-5, lower=lower, upper=upper, split=split)
            a = ctx.forces(method)
            e = np.linalg.norm(a - ref, axis=1) / scale
            assert np.all(np.isfinite(a)) and e.max() <= bar, (n, method, split, e.max())
            out.append(a)
        assert (np.linalg.norm(out[0] - out[1], axis=1) / scale).max() <= bar, (n, method)


def test_leapfrog_trajectory_bit_identical(ctx, orc):
    # Config 1 shape (Plummer, eps = 0.01 -> softening 1e-4), FP64 all-pairs.
    pos, vel, mass = plummer(1024, seed=8)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, DT_PARITY, 10)
    p, v = ctx.get_state()
    rp, rv, ra = orc.integrate(pos, vel, mass, DT_PARITY, 10, integrator=2, soft=1e-4)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    assert np.array_equal(ctx.get_acc(), ra)
    # continuing in two calls gives the same trajectory as one call
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, DT_PARITY, 5)
    rp2, rv2, _ = orc.integrate(rp, rv, mass, DT_PARITY, 5, integrator=2, soft=1e-4)
    p, v = ctx.get_state()
    assert np.array_equal(p, rp2) and np.array_equal(v, rv2)


def test_yoshida_trajectory_bit_identical(ctx, orc):
    pos, vel, mass = plummer(512, seed=9, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, A.YOSHIDA4, 2.5e-4, 6)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 2.5e-4, 6, integrator=5, soft=2.5e-5)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    assert np.array_equal(A.yoshida_weights(), orc.yoshida_weights())


@pytest.mark.parametrize("integrator,oi", [(A.LEAPFROG, 2), (A.YOSHIDA4, 5)])
def test_pn1_trajectory_bit_identical(ctx, orc, integrator, oi):
    # Velocity-dependent forces: two force evaluations per (sub)step, as the
    # reference's update_leapfrog.
    pos, vel, mass = plummer(256, seed=12)
    vel = vel * 20.0
    ctx.set_params(G=1.0, c=40.0, softening=1e-4, pn_flags=A.NEWTON | A.PN1)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, integrator, 1e-3, 4)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 1e-3, 4, integrator=oi, c=40.0, soft=1e-4, pn=0x11)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)


@pytest.mark.parametrize("order", range(1, 11))
def test_saba_trajectory_bit_identical(ctx, orc, order):
    # update_SABAn (order 1: update_euler) with FP64 all-pairs, Newton + 1PN:
    # n force evaluations per step, each after a drift, so velocity-dependent
    # forces need no special handling.
    pos, vel, mass = plummer(384, seed=20 + order)
    vel = vel * 20.0
    ctx.set_params(G=1.0, c=40.0, softening=1e-4, pn_flags=A.NEWTON | A.PN1)
    ctx.set_state(pos, vel, mass)
    ctx.integrate(A.ALLPAIRS_FP64, A.SABA, order, 1e-3, 3)
    p, v = ctx.get_state()
    rp, rv, ra = orc.integrate(pos, vel, mass, 1e-3, 3, integrator=1, order=order, c=40.0, soft=1e-4, pn=0x11)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    assert np.array_equal(ctx.get_acc(), ra)  # accelerations of the last evaluation


def test_euler_and_saba_then_leapfrog(ctx, orc):
    # ASTER_EULER through aster_step equals SABA order 1; a leapfrog step after
    # a SABA call re-evaluates forces (the cached ones are from before the
    # last drift).
    pos, vel, mass = plummer(256, seed=31)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, A.EULER, DT_PARITY, 4)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, DT_PARITY, 4, integrator=0, soft=1e-4)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    ctx.integrate(A.ALLPAIRS_FP64, A.SABA, 4, DT_PARITY, 2)
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, DT_PARITY, 2)
    rp, rv, _ = orc.integrate(rp, rv, mass, DT_PARITY, 2, integrator=1, order=4, soft=1e-4)
    rp, rv, _ = orc.integrate(rp, rv, mass, DT_PARITY, 2, integrator=2, soft=1e-4)
    p, v = ctx.get_state()
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    with pytest.raises(A.AsterError):
        ctx.step(A.ALLPAIRS_FP64, A.SABA, DT_PARITY, 1)  # needs an order
    with pytest.raises(A.AsterError):
        ctx.integrate(A.ALLPAIRS_FP64, A.SABA, 11, DT_PARITY, 1)


def test_saba_barnes_hut_bit_identical(ctx, orc):
    # fixed Morton box, so the box does not move with the drifts
    pos, vel, mass = A.ic_mixed(2048, 33)
    box = np.array([20.0, 20.0, 20.0])
    ctx.set_params(G=1.0, softening=2.5e-5, theta=0.5, precision=A.FP64, box=tuple(box))
    ctx.set_state(pos, vel, mass)
    ctx.integrate(A.BH_PARITY, A.SABA, 3, 5e-4, 2)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 5e-4, 2, integrator=1, order=3, method=1, soft=2.5e-5, box=box,
                              theta=0.5)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)


def test_mixed_precision_trajectory(ctx, orc):
    # FP32 forces + FP64 state vs the all-FP64 path over 50 Yoshida steps.
    pos, vel, mass = plummer(4096, seed=10, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP64, A.YOSHIDA4, 2.5e-4, 50)
    p64, _ = ctx.get_state()
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP32_KAHAN, A.YOSHIDA4, 2.5e-4, 50)
    p32, _ = ctx.get_state()
    disp = np.linalg.norm(p64 - pos, axis=1)
    err = np.linalg.norm(p32 - p64, axis=1)
    assert np.median(err / disp) < 1e-5


def test_energy_matches_oracle(ctx, orc):
    pos, vel, mass = plummer(3000, seed=11, mass_mode=1)
    ctx.set_params(G=1.3, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    e, ep, L = ctx.energy()
    re, rep, rL = orc.energy(pos, vel, mass, G=1.3, soft=1e-4)
    assert e == pytest.approx(re, rel=1e-12)
    assert ep == pytest.approx(rep, rel=1e-12)
    np.testing.assert_allclose(L, rL, rtol=1e-10, atol=1e-14)


def test_energy_drift_leapfrog_config1(ctx):
    pos, vel, mass = plummer(4096, seed=1)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    e0 = ctx.energy()[1]
    L0 = ctx.energy()[2]
    ctx.step(A.ALLPAIRS_FP64, A.LEAPFROG, DT_PARITY, 100)
    e1, L1 = ctx.energy()[1], ctx.energy()[2]
    print(f"\nconfig1 leapfrog 100 steps: |dE/E| = {abs(e1 - e0) / abs(e0):.2e}, |dL| = {np.abs(L1 - L0).max():.2e}")
    assert abs(e1 - e0) / abs(e0) < 1e-4
    assert np.abs(L1 - L0).max() < 1e-10


def test_errors_are_reported(ctx):
    import aster_amd
    with pytest.raises(aster_amd.AsterError):
        ctx.forces(A.ALLPAIRS_FP64)  # no state yet
    pos, vel, mass = plummer(64)
    bad = mass.copy()
    bad[3] = 0.0
    with pytest.raises(aster_amd.AsterError):
        ctx.set_state(pos, vel, bad)
    nanpos = pos.copy()
    nanpos[5, 1] = np.nan
    with pytest.raises(aster_amd.AsterError) as ei:
        ctx.set_state(nanpos, vel, mass)
    assert ei.value.code == -4
    ctx.set_state(pos, vel, mass)
    with pytest.raises(aster_amd.AsterError):
        ctx.forces(99)
    with pytest.raises(aster_amd.AsterError):
        ctx.step(A.ALLPAIRS_FP64, 3, 0.1, 1)
    with pytest.raises(aster_amd.AsterError):
        ctx.set_params(softening=-1.0)
    ctx.set_params(lower=10, upper=5)
    with pytest.raises(aster_amd.AsterError):
        ctx.forces(A.ALLPAIRS_FP32_KAHAN)  # an inverted range
    with pytest.raises(aster_amd.AsterError):
        ctx.forces(A.BH_GROUP)
