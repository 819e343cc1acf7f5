"""GPU checks of the trajectory adjoint (FP32 pair forces, FP64 state).

References: the oracle's FP64 discrete-adjoint gradient (itself pinned by
central finite differences in test_oracle.py) and central finite differences
of the device forward map. Tolerances are written in each test."""
import numpy as np
import pytest

import aster_amd as A

pytestmark = pytest.mark.gpu


def small_system(n, seed):
    r = np.random.default_rng(seed)
    pos, vel, mass = A.ic_plummer(n, seed)
    target_shift = r.normal(scale=1e-3, size=(n, 3))
    return pos, vel, mass, target_shift


def loss_and_grad_snap(snaps, target):
    d = snaps - target
    nsnap, n = snaps.shape[:2]
    return (d ** 2).sum() / (nsnap * n), 2.0 * d / (nsnap * n)


def relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_forward_matches_step(ctx):
    pos, vel, mass, _ = small_system(512, 1)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    snaps = ctx.traj_forward(1e-3, 30, 8, [10, 30])
    p1, v1 = ctx.get_state()
    ctx.set_state(pos, vel, mass)
    ctx.step(A.ALLPAIRS_FP32_KAHAN, A.LEAPFROG, 1e-3, 10)
    p10, _ = ctx.get_state()
    ctx.step(A.ALLPAIRS_FP32_KAHAN, A.LEAPFROG, 1e-3, 20)
    p2, v2 = ctx.get_state()
    assert np.array_equal(snaps[0], p10) and np.array_equal(snaps[1], p2)
    assert np.array_equal(p1, p2) and np.array_equal(v1, v2)


def test_gradient_matches_fp64_oracle(ctx, orc):
    # Tolerance: relative 2-norm difference <= 1e-4 per gradient block.
    n, steps, ckpt = 200, 48, 8
    pos, vel, mass, shift = small_system(n, 2)
    snaps_idx = np.array([12, 24, 48])
    ref_traj = [orc.integrate(pos, vel, mass, 1e-3, s, soft=1e-4)[0] for s in snaps_idx]
    target = np.stack(ref_traj) + shift[None]
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    snaps = ctx.traj_forward(1e-3, steps, ckpt, snaps_idx)
    loss, gsnap = loss_and_grad_snap(snaps, target)
    gx, gv, gm = ctx.traj_backward(gsnap)
    rloss, rgx, rgv, rgm = orc.traj_grad(pos, vel, mass, 1e-3, steps, snaps_idx, target, soft=1e-4)
    print(f"\nloss {loss:.6e} vs {rloss:.6e}; dx {relnorm(gx, rgx):.2e} dv {relnorm(gv, rgv):.2e} "
          f"dm {relnorm(gm, rgm):.2e}")
    assert loss == pytest.approx(rloss, rel=1e-6)
    assert relnorm(gx, rgx) < 1e-4 and relnorm(gv, rgv) < 1e-4 and relnorm(gm, rgm) < 1e-4
    # the state is restored to the initial conditions
    p, v = ctx.get_state()
    assert np.array_equal(p, pos) and np.array_equal(v, vel)


def test_gradient_finite_differences_device(ctx):
    # Directional central differences of the device forward map; tolerance
    # 1e-4 relative on each directional derivative.
    n, steps, ckpt = 1024, 64, 16
    pos, vel, mass, shift = small_system(n, 3)
    snaps_idx = np.array([16, 32, 64])
    ctx.set_params(G=1.0, softening=1e-4)

    def run(p, v, m):
        ctx.set_state(p, v, m)
        return ctx.traj_forward(1e-3, steps, ckpt, snaps_idx)

    target = run(pos, vel, mass) + shift[None]
    snaps = run(pos, vel, mass)
    _, gsnap = loss_and_grad_snap(snaps, target)
    gx, gv, gm = ctx.traj_backward(gsnap)
    r = np.random.default_rng(4)
    # step sizes keep the loss change well above the FP32 force rounding of
    # the forward map (the velocity gradient is ~T/dt times smaller, the mass
    # gradient smaller still: eps 1e-4 moves each mass by ~0.3%, central
    # truncation ~1e-5). Masses enter the pair forces as FP32: the mass step
    # is snapped to whole FP32 ulps of each mass, so m +- step converts to
    # FP32 exactly and the difference sees the step that was asked for.
    for which, g, eps in (("x", gx, 1e-5), ("v", gv, 1e-4), ("m", gm, 1e-4)):
        d = r.normal(size=g.shape)
        d /= np.linalg.norm(d)
        if which == "m":
            ulp = np.spacing(mass.astype(np.float32)).astype(np.float64)
            d = np.round(eps * d / ulp) * ulp / eps
        args_p = [pos, vel, mass]
        args_m = [pos, vel, mass]
        k = "xvm".index(which)
        args_p[k] = args_p[k] + eps * d
        args_m[k] = args_m[k] - eps * d
        lp, _ = loss_and_grad_snap(run(*args_p), target)
        lm, _ = loss_and_grad_snap(run(*args_m), target)
        fd = (lp - lm) / (2 * eps)
        ad = float((g * d).sum())
        print(f"\n{which}: adjoint {ad:.8e} fd {fd:.8e} rel {abs(ad - fd) / abs(fd):.2e}")
        assert ad == pytest.approx(fd, rel=1e-4)


def test_torch_autograd_function(ctx, orc):
    torch = pytest.importorskip("torch")
    from tests.adjoint_fn import TrajectorySnapshots
    n, steps = 64, 20
    pos, vel, mass, shift = small_system(n, 5)
    snaps_idx = [10, 20]
    target = np.stack([orc.integrate(pos, vel, mass, 1e-3, s, soft=1e-4)[0] for s in snaps_idx]) + shift[None]
    ctx.set_params(G=1.0, softening=1e-4)
    x0 = torch.tensor(pos, requires_grad=True)
    v0 = torch.tensor(vel, requires_grad=True)
    m = torch.tensor(mass, requires_grad=True)
    snaps = TrajectorySnapshots.apply(ctx, x0, v0, m, 1e-3, steps, 5, snaps_idx)
    loss = ((snaps - torch.tensor(target)) ** 2).mean() * 3.0  # mean over (nsnap, n, 3) -> per-body sum
    loss.backward()
    rloss, rgx, rgv, rgm = orc.traj_grad(pos, vel, mass, 1e-3, steps, np.array(snaps_idx), target, soft=1e-4)
    assert float(loss) == pytest.approx(rloss, rel=1e-6)
    assert relnorm(x0.grad.numpy(), rgx) < 1e-4
    assert relnorm(v0.grad.numpy(), rgv) < 1e-4
    assert relnorm(m.grad.numpy(), rgm) < 1e-4


def test_adjoint_errors(ctx):
    pos, vel, mass, _ = small_system(32, 6)
    ctx.set_params(G=1.0, softening=1e-4)
    ctx.set_state(pos, vel, mass)
    with pytest.raises(A.AsterError):
        ctx.traj_backward(np.zeros((1, 32, 3)))  # no forward yet
    with pytest.raises(A.AsterError):
        ctx.traj_forward(1e-3, 10, 4, [11])  # snapshot beyond the horizon
    with pytest.raises(A.AsterError):
        ctx.traj_forward(1e-3, 10, 4, [5, 5])  # not ascending
    ctx.set_params(G=1.0, softening=1e-4, pn_flags=A.NEWTON | A.PN1)
    with pytest.raises(A.AsterError):
        ctx.traj_forward(1e-3, 10, 4, [5])


def test_batched_threebody_fits(ctx, orc):
    # 4096 three-body systems (masses U[0.5,1.5], unit box, zero momentum); the
    # targets come from initial conditions perturbed by N(0, 1e-3^2). One
    # backward pass gives every system's gradient; checked against the FP64
    # oracle adjoint (relative 2-norm <= 1e-8) and against central differences
    # of the device forward map (relative 1e-6).
    nsys, steps, dt, soft = 4096, 200, 1e-3, 1e-2
    pos, vel, mass = A.ic_threebody(nsys, 1)
    off = np.arange(nsys + 1) * 3
    snaps = np.array([50, 100, 200])
    rng = np.random.default_rng(7)
    ctx.set_params(G=1.0, softening=soft, precision=A.FP32)
    ctx.batch_set(off, pos + rng.normal(scale=1e-3, size=pos.shape), vel + rng.normal(scale=1e-3, size=vel.shape),
                  mass)
    target = ctx.batch_traj_forward(dt, steps, snaps)

    def losses(p, v, m):
        ctx.batch_set(off, p, v, m)
        s = ctx.batch_traj_forward(dt, steps, snaps)
        d = (s - target) ** 2
        return d.reshape(len(snaps), nsys, 3, 3).sum(axis=(0, 2, 3)) / (len(snaps) * 3)

    ctx.batch_set(off, pos, vel, mass)
    s = ctx.batch_traj_forward(dt, steps, snaps)
    gsnap = 2.0 * (s - target) / (len(snaps) * 3)
    gx, gv, gm = ctx.batch_traj_backward(gsnap)
    for sidx in range(0, nsys, 512):
        b = slice(3 * sidx, 3 * sidx + 3)
        _, rgx, rgv, rgm = orc.traj_grad(pos[b], vel[b], mass[b], dt, steps, snaps, target[:, b], soft=soft)
        assert relnorm(gx[b], rgx) < 1e-8 and relnorm(gv[b], rgv) < 1e-8 and relnorm(gm[b], rgm) < 1e-8, sidx
    eps = 1e-6
    for sidx, comp in ((3, (0, 1)), (1000, (2, 2)), (4095, (1, 0))):
        i = 3 * sidx + comp[0]
        for which, g in (("x", gx), ("v", gv)):
            p, v = pos.copy(), vel.copy()
            arr = p if which == "x" else v
            arr[i, comp[1]] += eps
            lp = losses(p, v, mass)[sidx]
            arr[i, comp[1]] -= 2 * eps
            lm = losses(p, v, mass)[sidx]
            fd = (lp - lm) / (2 * eps)
            assert g[i, comp[1]] == pytest.approx(fd, rel=1e-6, abs=1e-12), (sidx, which)
        m = mass.copy()
        m[i] += eps
        lp = losses(pos, vel, m)[sidx]
        m[i] -= 2 * eps
        lm = losses(pos, vel, m)[sidx]
        assert gm[i] == pytest.approx((lp - lm) / (2 * eps), rel=1e-6, abs=1e-12), (sidx, "m")
