"""GPU parity for the LBVH Barnes-Hut pipeline, through the C ABI.

Morton codes, sort permutation, tree topology, node masses/centres of mass and
the FP64 reference-MAC forces must be bit-identical to the oracle (Aster's CPU
path); FP32 traversals are checked against FP64 all-pairs with the stated
tolerances."""
import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu


def mixed(n, seed=1):
    return A.ic_mixed(n, seed)


def rel_err(a, b):
    return np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)


def load(ctx, pos, vel, mass, **kw):
    p = dict(G=1.0, softening=2.5e-5, theta=0.5)
    p.update(kw)
    ctx.set_params(**p)
    ctx.set_state(pos, vel, mass)


@pytest.mark.parametrize("n", [2, 3, 100, 4097, 1 << 16])
def test_morton_bit_identical(ctx, orc, n):
    pos, vel, mass = mixed(n, seed=n)
    load(ctx, pos, vel, mass)
    codes, box = ctx.morton()
    assert np.array_equal(box, orc.box_auto(pos))
    assert np.array_equal(codes, orc.morton(pos, box))


def test_morton_user_box_and_out_of_box(ctx, orc):
    pos, vel, mass = mixed(5000, seed=2)
    load(ctx, pos, vel, mass, box=(12.0, 11.0, 10.5))
    codes, box = ctx.morton()
    assert np.array_equal(box, [12.0, 11.0, 10.5])
    assert np.array_equal(codes, orc.morton(pos, box))
    # the reference default box (1366, sim_meta.h:14-18) puts everything in a few cells
    load(ctx, pos, vel, mass, box=(1366.0, 1366.0, 1366.0))
    codes, box = ctx.morton()
    assert np.array_equal(codes, orc.morton(pos, box))
    assert len(np.unique(codes)) < len(codes) // 10
    load(ctx, pos, vel, mass, box=(1.0, 1.0, 1.0))
    with pytest.raises(A.AsterError) as ei:
        ctx.morton()
    assert ei.value.code == -3


@pytest.mark.parametrize("n,dup", [(2, False), (3, False), (1000, False), (4096, True), (1 << 16, False),
                                   (1 << 20, False)])
def test_tree_bit_identical(ctx, orc, n, dup):
    pos, vel, mass = mixed(n, seed=n + 1)
    if dup:  # many equal Morton codes: cluster a third of the bodies in one cell
        pos[: n // 3] = pos[0] + 1e-9 * np.arange(n // 3)[:, None]
    load(ctx, pos, vel, mass)
    ctx.forces(A.BH_PARITY, out=False)
    T = ctx.tree()
    box = orc.box_auto(pos)
    R = orc.build_tree(pos, vel, mass, box)
    assert np.array_equal(T["codes"], R["codes"])
    assert np.array_equal(T["perm"], R["perm"])
    assert np.array_equal(T["left"], R["left"])
    assert np.array_equal(T["right"], R["right"])
    assert np.array_equal(T["mass"], R["mass"])
    assert np.array_equal(T["com"], R["com"])


@pytest.mark.parametrize("n,dup", [(5000, True), (1 << 18, False), (300001, True), (1 << 20, True),
                                   ((1 << 20) + 1, False), (3 << 20, True)])
def test_sort_order(ctx, n, dup):
    """The (code, body index) order of the radix sort against numpy, on each
    tile layout: 4096- and 8192-key tiles with the offsets formed in the
    scatter (up to 2^20 keys), 2048-key tiles with the scanned offsets
    (above)."""
    pos, vel, mass = mixed(n, seed=n + 7)
    if dup:
        pos[: n // 3] = pos[0] + 1e-9 * np.arange(n // 3)[:, None]
    load(ctx, pos, vel, mass, precision=A.FP32, box=(0.0, 0.0, 0.0))
    codes, _ = ctx.morton()
    ctx.forces(A.BH_GROUP, out=False)
    T = ctx.tree()
    ref = np.lexsort((np.arange(n), codes))
    assert np.array_equal(T["perm"], ref)
    assert np.array_equal(T["codes"], codes[ref])


@pytest.mark.parametrize("theta", [0.3, 0.5, 0.7])
def test_bh_fp64_parity_bit_identical(ctx, orc, theta):
    pos, vel, mass = mixed(4096, seed=3)
    load(ctx, pos, vel, mass, theta=theta, precision=A.FP64)
    acc = ctx.forces(A.BH_PARITY)
    ref, cnt = orc.bh(pos, vel, mass, orc.box_auto(pos), theta, soft=2.5e-5, counts=True)
    assert np.array_equal(acc, ref)
    inter, visits = ctx.bh_stats()
    assert inter == cnt.sum()


def test_bh_fp64_parity_pn1_bit_identical(ctx, orc):
    pos, vel, mass = mixed(2048, seed=4)
    load(ctx, pos, vel, mass, theta=0.4, precision=A.FP64, c=30.0, pn_flags=A.NEWTON | A.PN1)
    acc = ctx.forces(A.BH_PARITY)
    ref = orc.bh(pos, vel, mass, orc.box_auto(pos), 0.4, c=30.0, soft=2.5e-5, pn=0x11)
    assert np.array_equal(acc, ref)


def test_bh_fp64_parity_pn2_pn25_bit_identical(ctx, orc):
    # node velocities (mass-weighted) feed the 2PN and 2.5PN terms as in
    # get_node_body
    flags = A.NEWTON | A.PN1 | A.PN2 | A.PN25
    pos, vel, mass = mixed(2048, seed=5)
    load(ctx, pos, vel, mass, theta=0.4, precision=A.FP64, c=30.0, pn_flags=flags)
    acc = ctx.forces(A.BH_PARITY)
    ref = orc.bh(pos, vel, mass, orc.box_auto(pos), 0.4, c=30.0, soft=2.5e-5, pn=flags)
    assert np.array_equal(acc, ref)
    load(ctx, pos, vel, mass, theta=0.4, precision=A.FP32, c=30.0, pn_flags=A.NEWTON | A.EIH)
    with pytest.raises(A.AsterError):
        ctx.forces(A.BH_GEOMETRIC)  # the N-body EIH law runs on FP64 all-pairs only


@pytest.mark.parametrize("method", [A.BH_GEOMETRIC, A.BH_PARITY])
@pytest.mark.parametrize("flags", [A.NEWTON | A.PN1 | A.PN2 | A.PN25, A.NEWTON | A.PN1])
def test_bh_per_body_fp32_pn(ctx, orc, method, flags):
    """The per-body FP32 walks add the post-Newtonian terms in FP64 for every
    node they take. With every node opened (theta = 0) the post-Newtonian part
    is the all-pairs one (to FP64 summation-order rounding); at theta = 0.5 the
    FP32 reference-MAC walk takes the nodes of the FP64 oracle walk, so most
    bodies get its post-Newtonian part to the same rounding."""
    pos, vel, mass = mixed(3000, seed=44)

    def pn_part(theta):
        load(ctx, pos, vel, mass, theta=theta, precision=A.FP32, c=30.0, pn_flags=A.NEWTON)
        newton = ctx.forces(method)
        load(ctx, pos, vel, mass, theta=theta, precision=A.FP32, c=30.0, pn_flags=flags)
        return ctx.forces(method) - newton

    got = pn_part(0.0)
    ref = orc.bh_group_pn(pos, vel, mass, orc.box_auto(pos), 0.0, c=30.0, soft=2.5e-5, pn=flags)
    scale = np.linalg.norm(ref, axis=1)
    assert np.all(np.linalg.norm(got - ref, axis=1) <= 1e-9 * np.maximum(scale, np.median(scale)))
    if method == A.BH_PARITY:
        got = pn_part(0.5)
        box = orc.box_auto(pos)
        ref = (orc.bh(pos, vel, mass, box, 0.5, c=30.0, soft=2.5e-5, pn=flags)
               - orc.bh(pos, vel, mass, box, 0.5, c=30.0, soft=2.5e-5, pn=A.NEWTON))
        scale = np.linalg.norm(ref, axis=1)
        e = np.linalg.norm(got - ref, axis=1) / np.maximum(scale, np.median(scale))
        print(f"\nparity MAC FP32 walk: PN part vs oracle median {np.median(e):.2e} p90 {np.quantile(e, .9):.2e}")
        assert np.median(e) < 1e-9 and np.quantile(e, 0.9) < 1e-9


@pytest.mark.parametrize("n,theta,flags", [(20000, 0.5, A.NEWTON | A.PN1 | A.PN2 | A.PN25),
                                           (3000, 0.0, A.NEWTON | A.PN1 | A.PN2 | A.PN25),
                                           (130, 0.7, A.NEWTON | A.PN1 | A.PN2 | A.PN25),
                                           (20000, 0.5, A.NEWTON | A.PN1), (3000, 0.0, A.NEWTON | A.PN1)])
def test_bh_group_pn_matches_oracle(ctx, orc, n, theta, flags):
    """Group traversal with the post-Newtonian terms: the FP32 Newtonian sums
    as without them, plus the FP64 post-Newtonian part over the same lists
    (node centres of mass, masses and mass-weighted velocities), which must
    equal the oracle's to FP64 rounding (the 1PN term alone is evaluated
    grouped for fused multiply-adds, the full set in the reference's order)."""
    pos, vel, mass = mixed(n, seed=n + 3)
    box = orc.box_auto(pos)
    load(ctx, pos, vel, mass, theta=theta, precision=A.FP32, c=30.0, pn_flags=A.NEWTON)
    newton = ctx.forces(A.BH_GROUP)
    load(ctx, pos, vel, mass, theta=theta, precision=A.FP32, c=30.0, pn_flags=flags)
    acc = ctx.forces(A.BH_GROUP)
    ref_pn = orc.bh_group_pn(pos, vel, mass, box, theta, c=30.0, soft=2.5e-5, pn=flags)
    scale = np.linalg.norm(ref_pn, axis=1)
    err = np.linalg.norm((acc - newton) - ref_pn, axis=1)
    assert np.all(err <= 1e-9 * np.maximum(scale, np.median(scale)) + 1e-15 * np.abs(acc).max())
    # and the whole force against the oracle's FP64 sums of the same lists
    ref_n, _, _ = orc.bh_group(pos, mass, box, theta, soft=2.5e-5)
    ref = ref_n + ref_pn
    s = np.linalg.norm(ref, axis=1)
    e = np.linalg.norm(acc - ref, axis=1) / np.maximum(s, np.median(s))
    assert np.median(e) < 1e-6 and e.max() < 1e-4


def test_bh_group_pn_leapfrog_steps(ctx, orc):
    """Leapfrog with the velocity-dependent 1PN law over the group traversal
    with every node opened (theta = 0) against update_leapfrog over FP64
    all-pairs: the positions agree to the FP32 rounding of the Newtonian sums
    (bars as for the Newtonian steps)."""
    n, dt, steps = 3000, 5e-4, 4
    pos, vel, mass = mixed(n, seed=33)
    load(ctx, pos, vel, mass, theta=0.0, precision=A.FP32, c=30.0, pn_flags=A.NEWTON | A.PN1, box=(0.0, 0.0, 0.0))
    ctx.step(A.BH_GROUP, A.LEAPFROG, dt, steps)
    p, v = ctx.get_state()
    rp, rv, ra = orc.integrate(pos, vel, mass, dt, steps, integrator=2, method=0, c=30.0, soft=2.5e-5, pn=0x11)
    amax = np.abs(ra).max()
    ex = np.abs(p - rp).max(1) / (steps ** 2 * dt ** 2 * amax)
    ev = np.abs(v - rv).max(1) / (steps * dt * amax)
    assert np.all(np.isfinite(p)) and np.all(np.isfinite(v))
    assert np.quantile(ex, 0.999) <= 1e-5 and ex.max() <= 1e-3
    assert np.quantile(ev, 0.999) <= 1e-5 and ev.max() <= 1e-3


def test_bh_fp64_parity_node_velocities_in_moment_bands(ctx, orc):
    """At 2^18 bodies most nodes above the traversal groups are formed by the
    multi-level moment bands (not the top-levels workgroup), velocities
    included; their mass-weighted node velocities feed the 2PN and 2.5PN
    terms, so the forces are the same bits only if those moments are."""
    flags = A.NEWTON | A.PN1 | A.PN2 | A.PN25
    pos, vel, mass = mixed(1 << 18, seed=6)
    load(ctx, pos, vel, mass, theta=0.4, precision=A.FP64, c=30.0, pn_flags=flags)
    acc = ctx.forces(A.BH_PARITY)
    ref = orc.bh(pos, vel, mass, orc.box_auto(pos), 0.4, c=30.0, soft=2.5e-5, pn=flags)
    assert np.array_equal(acc, ref)


def test_bh_fp64_user_box_bit_identical(ctx, orc):
    pos, vel, mass = mixed(3000, seed=5)
    box = np.array([10.5, 10.25, 11.0])
    load(ctx, pos, vel, mass, theta=0.5, precision=A.FP64, box=tuple(box))
    acc = ctx.forces(A.BH_PARITY)
    assert np.array_equal(acc, orc.bh(pos, vel, mass, box, 0.5, soft=2.5e-5))


def test_bh_fp32_parity_close_to_oracle(ctx, orc):
    pos, vel, mass = mixed(1 << 14, seed=6)
    load(ctx, pos, vel, mass, theta=0.5, precision=A.FP32)
    acc = ctx.forces(A.BH_PARITY)
    ref = orc.bh(pos, vel, mass, orc.box_auto(pos), 0.5, soft=2.5e-5)
    e = rel_err(acc, ref)
    print(f"\nBH parity FP32 vs FP64 oracle: median {np.median(e):.2e} p99 {np.quantile(e, .99):.2e}")
    assert np.median(e) < 1e-5 and np.quantile(e, 0.99) < 1e-3


def test_bh_group_theta_zero_is_allpairs(ctx, orc):
    pos, vel, mass = mixed(3000, seed=9)
    load(ctx, pos, vel, mass, theta=0.0)
    acc = ctx.forces(A.BH_GROUP)
    inter, _ = ctx.bh_stats()
    ref = orc.allpairs(pos, vel, mass, soft=2.5e-5)
    e = rel_err(acc, ref)
    assert np.median(e) < 1e-6 and e.max() < 1e-4
    assert inter == 3000 * 2999


def test_bh_group_handles_partial_wavefronts_and_tiny_n(ctx, orc):
    for n in (2, 3, 65, 130):
        pos, vel, mass = mixed(n, seed=n)
        load(ctx, pos, vel, mass, theta=0.5)
        e = rel_err(ctx.forces(A.BH_GROUP), orc.allpairs(pos, vel, mass, soft=2.5e-5))
        assert np.median(e) < 1e-2, n


@pytest.mark.parametrize("n,seed", [(2, 1), (3, 2), (64, 3), (65, 4), (100, 5), (4097, 6), (20000, 7),
                                    (1 << 18, 8)])
def test_bh_group_matches_oracle_lists(ctx, orc, n, seed):
    """Same accepted node sets as the oracle's restatement of the group
    traversal (interaction and node-visit counts equal exactly, for every
    group size the input produces), and the same forces within FP32 rounding
    (the library sums FP32 terms, Kahan-compensated across list flushes; the
    oracle sums the same terms in FP64)."""
    pos, vel, mass = mixed(n, seed=seed)
    box = orc.box_auto(pos)
    for theta in (0.3, 0.5, 0.7):
        load(ctx, pos, vel, mass, theta=theta)
        try:
            ctx.bh_stats()  # clears counts left by earlier evaluations
        except A.AsterError:
            pass  # none yet on this context
        acc = ctx.forces(A.BH_GROUP)
        inter, visits = ctx.bh_stats()
        ref, ref_inter, ref_visits = orc.bh_group(pos, mass, box, theta, soft=2.5e-5)
        assert (inter, visits) == (ref_inter, ref_visits), (n, theta)
        scale = np.linalg.norm(ref, axis=1)
        e = np.linalg.norm(acc - ref, axis=1) / np.maximum(scale, np.median(scale))
        assert np.median(e) < 1e-6 and e.max() < 1e-4, (n, theta, np.median(e), e.max())


def test_bh_group_matches_oracle_user_box_and_wide_angles(ctx, orc):
    """The same exact-list check with a user Morton box (a different tree)
    and opening angles of 1 and more."""
    n = 30000
    pos, vel, mass = mixed(n, seed=31)
    box = np.array([11.0, 10.5, 12.0])
    for theta in (1.0, 1.5):
        load(ctx, pos, vel, mass, theta=theta, box=tuple(box))
        try:
            ctx.bh_stats()
        except A.AsterError:
            pass
        acc = ctx.forces(A.BH_GROUP)
        inter, visits = ctx.bh_stats()
        ref, ref_inter, ref_visits = orc.bh_group(pos, mass, box, theta, soft=2.5e-5)
        assert (inter, visits) == (ref_inter, ref_visits), theta
        scale = np.linalg.norm(ref, axis=1)
        e = np.linalg.norm(acc - ref, axis=1) / np.maximum(scale, np.median(scale))
        assert np.median(e) < 1e-6 and e.max() < 1e-4, (theta, np.median(e), e.max())


def force_errors(ctx, pos, vel, mass, method, thetas):
    load(ctx, pos, vel, mass)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    out = {}
    for theta in thetas:
        load(ctx, pos, vel, mass, theta=theta)
        e = rel_err(ctx.forces(method), ref)
        inter, _ = ctx.bh_stats()
        out[theta] = dict(median=np.median(e), p99=np.quantile(e, 0.99), rms=np.sqrt(np.mean(e ** 2)),
                          inter=inter / len(mass))
        print(f"\nmethod {method} theta {theta}: " + " ".join(f"{k} {v:.3e}" for k, v in out[theta].items()))
    return out


def test_bh_accuracy_vs_allpairs(ctx):
    # Force error of the FP32 traversals against FP64 all-pairs at each theta
    # (mixed Plummer + cube, 2^16 bodies). Tolerances: errors grow with theta;
    # the group traversal (box-to-node test) is at least as accurate as the
    # per-body geometric test; group median error at theta = 0.5 <= 1e-3.
    pos, vel, mass = mixed(1 << 16, seed=7)
    geo = force_errors(ctx, pos, vel, mass, A.BH_GEOMETRIC, (0.3, 0.5, 0.7))
    grp = force_errors(ctx, pos, vel, mass, A.BH_GROUP, (0.3, 0.5, 0.7))
    for out in (geo, grp):
        assert out[0.3]["median"] < out[0.5]["median"] < out[0.7]["median"]
    for t in (0.3, 0.5, 0.7):
        assert grp[t]["median"] <= geo[t]["median"]
    assert grp[0.5]["median"] <= 1e-3


def test_bh_leapfrog_steps(ctx, orc):
    # One BH evaluation per leapfrog step (accelerations cached across steps);
    # with the reference MAC in FP64 and a fixed box the trajectory is
    # bit-identical to update_leapfrog over the CPU Barnes_Hut solver.
    pos, vel, mass = mixed(2048, seed=8)
    box = (20.0, 20.0, 20.0)
    load(ctx, pos, vel, mass, theta=0.5, precision=A.FP64, box=box)
    ctx.step(A.BH_PARITY, A.LEAPFROG, 5e-4, 3)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 5e-4, 3, integrator=2, method=1, soft=2.5e-5, box=np.array(box),
                              theta=0.5)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)


@pytest.mark.parametrize("integrator", [A.LEAPFROG, A.YOSHIDA4])
def test_bh_steps_automatic_box(ctx, orc, integrator):
    # With the automatic box each force evaluation takes the box of the
    # positions the drift just wrote (the drift forms its first stage):
    # bit-identical to the CPU path, which recomputes the box per evaluation.
    pos, vel, mass = mixed(3000, seed=9)
    load(ctx, pos, vel, mass, theta=0.5, precision=A.FP64, box=(0.0, 0.0, 0.0))
    ctx.step(A.BH_PARITY, integrator, 5e-4, 4)
    p, v = ctx.get_state()
    rp, rv, _ = orc.integrate(pos, vel, mass, 5e-4, 4, integrator=2 if integrator == A.LEAPFROG else 5, method=1,
                              soft=2.5e-5, theta=0.5)
    assert np.array_equal(p, rp) and np.array_equal(v, rv)
    # a state loaded afterwards gets its own box, not the drift's
    pos2 = pos * 3.0
    load(ctx, pos2, vel, mass, theta=0.5, precision=A.FP64, box=(0.0, 0.0, 0.0))
    _, box = ctx.morton()
    assert np.array_equal(box, orc.box_auto(pos2))
    assert np.array_equal(ctx.forces(A.BH_PARITY), orc.bh(pos2, vel, mass, orc.box_auto(pos2), 0.5, soft=2.5e-5))


def test_bh_group_duplicates_and_unsoftened(ctx):
    # A third of the bodies share one Morton cell (deep duplicate-code
    # subtrees, many tiny groups that split their lists over 2-64 lanes) and a
    # few share exact positions with softening on; then distinct positions
    # with softening 0 (the masked self pair). Group traversal against FP64
    # all-pairs: finite everywhere; median relative error <= 1e-3 at
    # theta 0.5 with softening, <= 2e-3 without (no smoothing of the near
    # field: 1.07e-3 measured), and smaller at theta 0.3.
    pos, vel, mass = mixed(1 << 14, seed=21)
    r = np.random.default_rng(5)
    k = len(mass) // 3
    pos[:k] = 0.25 + r.uniform(0, 1e-4, size=(k, 3))
    pos[k:k + 8] = pos[k + 8:k + 16]
    load(ctx, pos, vel, mass, theta=0.5)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    acc = ctx.forces(A.BH_GROUP)
    assert np.all(np.isfinite(acc))
    assert np.median(rel_err(acc, ref)) <= 1e-3
    pos, vel, mass = mixed(1 << 14, seed=22)
    load(ctx, pos, vel, mass, theta=0.5, softening=0.0)
    ref = ctx.forces(A.ALLPAIRS_FP64)
    acc = ctx.forces(A.BH_GROUP)
    assert np.all(np.isfinite(acc))
    e05 = np.median(rel_err(acc, ref))
    load(ctx, pos, vel, mass, theta=0.3, softening=0.0)
    e03 = np.median(rel_err(ctx.forces(A.BH_GROUP), ref))
    assert e05 <= 2e-3 and e03 < e05
