"""FP32 force paths at SI scale, source and body bounds on every force method,
and batched-ensemble parameters changed between graph replays, through the C
ABI against the oracle."""
import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu

G_SI = 6.674e-11


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def si_system(n, seed, coincident=True):
    """A Plummer sphere scaled to 1e11 m with masses log-uniform over
    1e28..1e31 kg (a star cluster in SI units); bodies 0 and 1 share one
    position when `coincident` (each then gets zero force from the other, as
    in the reference, where d = 0 makes the pair term zero)."""
    pos, vel, _ = A.ic_plummer(n, seed)
    rng = np.random.default_rng(seed)
    mass = 10.0 ** rng.uniform(28.0, 31.0, n)
    pos = pos * 1e11
    if coincident:
        pos[1] = pos[0]
    return pos, vel * 1e4, mass


SOFT_SI = 1e-14  # the reference's default softening (eps^2, m^2)


@pytest.mark.parametrize("method", [A.ALLPAIRS_FP32, A.ALLPAIRS_FP32_KAHAN])
def test_fp32_allpairs_si_scale(ctx, orc, method):
    """m eps^-3 ~ 1e51 overflows FP32; the library scales masses by a power of
    two so the self pair and the coincident pair stay finite (zero force).
    Bars as in code units: mixed max <= 1e-5; plain median <= 1e-6, max <= 1e-4."""
    pos, vel, mass = si_system(4096, 1)
    ctx.set_params(G=G_SI, softening=SOFT_SI, theta=0.5, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(method)
    ref = orc.allpairs(pos, vel, mass, G=G_SI, soft=SOFT_SI)
    assert np.all(np.isfinite(acc))
    e = rel_err(acc, ref)
    print(f"\nSI scale method {method}: median {np.median(e):.2e} max {e.max():.2e}")
    if method == A.ALLPAIRS_FP32_KAHAN:
        assert e.max() <= 1e-5
    else:
        assert np.median(e) <= 1e-6 and e.max() <= 1e-4


def test_fp32_allpairs_si_scale_unsoftened(ctx, orc):
    """Softening 0 in SI units: the masked kernels (zero force at r = 0) with
    distinct positions; mixed max <= 1e-5."""
    pos, vel, mass = si_system(4096, 2, coincident=False)
    ctx.set_params(G=G_SI, softening=0.0, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
    e = rel_err(acc, orc.allpairs(pos, vel, mass, G=G_SI, soft=0.0))
    assert np.all(np.isfinite(acc)) and e.max() <= 1e-5


def test_bh_group_si_scale(ctx, orc):
    """Every body's own leaf is in its group's list (d = 0): at SI scale the
    group traversal stays finite and sums the oracle's lists (equal counts;
    median <= 1e-6, max <= 1e-4 of the larger of |a| and the median |a|)."""
    pos, vel, mass = si_system(20000, 3)
    box = orc.box_auto(pos)
    ctx.set_params(G=G_SI, softening=SOFT_SI, theta=0.5, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    try:
        ctx.bh_stats()
    except A.AsterError:
        pass
    acc = ctx.forces(A.BH_GROUP)
    inter, visits = ctx.bh_stats()
    ref, ri, rv = orc.bh_group(pos, mass, box, 0.5, G=G_SI, soft=SOFT_SI)
    assert np.all(np.isfinite(acc))
    assert (inter, visits) == (ri, rv)
    scale = np.linalg.norm(ref, axis=1)
    e = np.linalg.norm(acc - ref, axis=1) / np.maximum(scale, np.median(scale))
    assert np.median(e) < 1e-6 and e.max() < 1e-4, (np.median(e), e.max())


@pytest.mark.parametrize("method", [A.BH_GEOMETRIC, A.BH_PARITY])
def test_bh_per_body_si_scale(ctx, orc, method):
    """The per-body FP32 traversals at SI scale: finite; the reference MAC in
    FP32 against the oracle's FP64 Barnes_Hut: median <= 1e-5 (as in code
    units); the geometric MAC against all-pairs: no less accurate than on the
    same system in code units (lengths / 1e11, masses / 1e30, G = 1), whose
    error comes from the opening test alone."""
    pos, vel, mass = si_system(8192, 4)
    ctx.set_params(G=G_SI, softening=SOFT_SI, theta=0.5, precision=A.FP32)
    ctx.set_state(pos, vel, mass)
    acc = ctx.forces(method)
    assert np.all(np.isfinite(acc))
    if method == A.BH_PARITY:
        e = rel_err(acc, orc.bh(pos, vel, mass, orc.box_auto(pos), 0.5, G=G_SI, soft=SOFT_SI))
        assert np.median(e) < 1e-5
    else:
        e = rel_err(acc, orc.allpairs(pos, vel, mass, G=G_SI, soft=SOFT_SI))
        pc, mc = pos / 1e11, mass / 1e30
        ctx.set_params(G=1.0, softening=SOFT_SI / 1e22)
        ctx.set_state(pc, vel, mc)
        ec = rel_err(ctx.forces(method), orc.allpairs(pc, vel, mc, soft=SOFT_SI / 1e22))
        print(f"\ngeometric MAC median error: SI {np.median(e):.3e}, code units {np.median(ec):.3e}")
        assert np.median(e) <= 1.01 * np.median(ec) + 1e-6


@pytest.mark.parametrize("lower,upper", [(100, 3000), (64, 4096), (0, 1), (1, -1), (4000, 4095)])
def test_fp32_allpairs_source_bounds(ctx, orc, lower, upper):
    """Solver::set_bounds on the FP32 all-pairs: sources [lower, upper), every
    target (Parallelized::compute_forces); bars as without bounds."""
    pos, vel, mass = A.ic_plummer(4096, 9, mass_mode=1)
    ctx.set_params(G=1.0, softening=2.5e-5, lower=lower, upper=upper, pn_flags=A.NEWTON)
    ctx.set_state(pos, vel, mass)
    ref = orc.allpairs(pos, vel, mass, soft=2.5e-5, lower=lower, upper=upper)
    mixed = ctx.forces(A.ALLPAIRS_FP32_KAHAN)
    plain = ctx.forces(A.ALLPAIRS_FP32)
    # a target whose only source is itself has zero acceleration
    live = np.linalg.norm(ref, axis=1) > 0
    assert np.all(mixed[~live] == 0.0) and np.all(plain[~live] == 0.0)
    e = rel_err(mixed[live], ref[live])
    e2 = rel_err(plain[live], ref[live])
    assert e.max() <= 1e-5 and np.median(e2) <= 1e-6 and e2.max() <= 1e-4, (e.max(), e2.max())


def test_fp32_allpairs_source_bounds_with_pn1(ctx, orc):
    """The FP64 post-Newtonian correction of the FP32 methods takes the same
    source range."""
    pos, vel, mass = A.ic_plummer(2048, 10, mass_mode=1)
    vel = vel * 30.0
    ctx.set_params(G=1.0, c=20.0, softening=2.5e-5, lower=200, upper=1500, pn_flags=A.NEWTON | A.PN1)
    ctx.set_state(pos, vel, mass)
    ref = orc.allpairs(pos, vel, mass, c=20.0, soft=2.5e-5, pn=0x11, lower=200, upper=1500)
    assert rel_err(ctx.forces(A.ALLPAIRS_FP32_KAHAN), ref).max() <= 1e-5


def test_bh_body_bounds(ctx, orc):
    """Solver::set_bounds on Barnes-Hut: the tree holds bodies [lower, upper)
    and only they get new accelerations (Barnes_Hut::compute_forces); the
    FP64 reference MAC is bit-identical to the oracle's Barnes_Hut over those
    bodies alone, the group traversal sums the oracle's group lists over them,
    and every other body keeps its previous acceleration."""
    n, lo, hi = 6000, 1000, 4500
    pos, vel, mass = A.ic_mixed(n, 12)
    sp, sv, sm = pos[lo:hi], vel[lo:hi], mass[lo:hi]
    sbox = orc.box_auto(sp)
    ctx.set_params(G=1.0, softening=2.5e-5, theta=0.5, precision=A.FP64, lower=0, upper=-1)
    ctx.set_state(pos, vel, mass)
    full = ctx.forces(A.BH_PARITY)
    ctx.set_params(lower=lo, upper=hi)
    acc = ctx.forces(A.BH_PARITY)
    assert np.array_equal(acc[lo:hi], orc.bh(sp, sv, sm, sbox, 0.5, soft=2.5e-5))
    assert np.array_equal(acc[:lo], full[:lo]) and np.array_equal(acc[hi:], full[hi:])
    ctx.set_params(precision=A.FP32)
    try:
        ctx.bh_stats()
    except A.AsterError:
        pass
    acc = ctx.forces(A.BH_GROUP)
    inter, visits = ctx.bh_stats()
    ref, ri, rv = orc.bh_group(sp, sm, sbox, 0.5, soft=2.5e-5)
    assert (inter, visits) == (ri, rv)
    scale = np.linalg.norm(ref, axis=1)
    e = np.linalg.norm(acc[lo:hi] - ref, axis=1) / np.maximum(scale, np.median(scale))
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    assert np.array_equal(acc[:lo], full[:lo]) and np.array_equal(acc[hi:], full[hi:])
    # an empty range evaluates nothing
    ctx.set_params(lower=100, upper=100)
    assert np.array_equal(ctx.forces(A.BH_GROUP), acc)


def test_bh_body_bounds_in_steps(ctx, orc):
    """Leapfrog steps with a bounded FP64 parity traversal: the bodies in range
    move as update_leapfrog over Barnes_Hut on those bodies alone (bit for
    bit); the others drift with the accelerations they had."""
    n, lo, hi = 3000, 500, 2500
    pos, vel, mass = A.ic_mixed(n, 13)
    ctx.set_params(G=1.0, softening=2.5e-5, theta=0.5, precision=A.FP64, lower=lo, upper=hi)
    ctx.set_state(pos, vel, mass)
    ctx.step(A.BH_PARITY, A.LEAPFROG, 5e-4, 3)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos[lo:hi], vel[lo:hi], mass[lo:hi], 5e-4, 3, integrator=2, method=1, soft=2.5e-5,
                              theta=0.5)
    assert np.array_equal(p[lo:hi], rp) and np.array_equal(v[lo:hi], rv)


def test_batch_graph_after_parameter_changes(ctx):
    """Graph-mode batch steps with the force law changed in between equal the
    same sequence run without graphs, bit for bit (the captured launches
    carry every force-law flag); after G and the force law change again, the
    continuing run equals a fresh start from the same state (new parameters
    drop the cached Newtonian accelerations of the first half kick)."""
    off, pos, vel, mass = A.ic_ensemble(96, 21, 3, 40)
    runs, mids = [], []
    for graph in (True, False):
        ctx.set_params(G=1.0, softening=0.0, c=30.0, pn_flags=A.NEWTON | A.PN1, precision=A.FP64)
        ctx.batch_set(off, pos, vel, mass)
        ctx.batch_step(A.LEAPFROG, 1e-3, 8, steps_per_launch=4, use_graph=graph)
        ctx.set_params(pn_flags=A.NEWTON | A.PN1 | A.PN25)
        ctx.batch_step(A.LEAPFROG, 1e-3, 8, steps_per_launch=4, use_graph=graph)
        mids.append(ctx.batch_get())
        ctx.set_params(G=1.5, pn_flags=A.NEWTON)
        ctx.batch_step(A.LEAPFROG, 1e-3, 8, steps_per_launch=4, use_graph=graph)
        runs.append(ctx.batch_get())
    (pg, vg), (pe, ve) = runs
    assert np.array_equal(mids[0][0], mids[1][0]) and np.array_equal(mids[0][1], mids[1][1])
    assert np.array_equal(pg, pe) and np.array_equal(vg, ve)
    ctx.set_params(G=1.5, softening=0.0, c=30.0, pn_flags=A.NEWTON, precision=A.FP64)
    ctx.batch_set(off, mids[1][0], mids[1][1], mass)
    ctx.batch_step(A.LEAPFROG, 1e-3, 8, steps_per_launch=4, use_graph=False)
    pf, vf = ctx.batch_get()
    assert np.array_equal(pf, pe) and np.array_equal(vf, ve)
