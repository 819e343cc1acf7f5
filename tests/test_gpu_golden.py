"""Device paths against the committed golden fixtures (tests/golden/*.npz),
through the C ABI. FP64 paths must match bit for bit; the batched adjoint
(FP32 pair forces inside an FP64 state, as in the library) within a relative
2-norm of 1e-8."""
import numpy as np
import pytest

import aster_amd as A
from tests.golden.make_fixtures import INTEGRATE_RUNS, PN_SETS
from tests.test_golden import fixture

pytestmark = pytest.mark.gpu

INTEGRATORS = {0: A.EULER, 1: A.SABA, 2: A.LEAPFROG, 5: A.YOSHIDA4}


def relnorm(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.mark.parametrize("n", [8, 64, 1024])
def test_allpairs_fp64_golden(ctx, n):
    f = fixture("forces.npz")
    for flags in PN_SETS:
        ctx.set_params(G=1.0, c=20.0, softening=1e-4, pn_flags=flags)
        ctx.set_state(f[f"pos{n}"], f[f"vel{n}"], f[f"mass{n}"])
        acc = ctx.forces(A.ALLPAIRS_FP64)
        assert np.array_equal(acc, f[f"acc{n}_{flags:x}"]), hex(flags)


def test_tree_and_bh_golden(ctx):
    f = fixture("tree_bh_n1024.npz")
    ctx.set_params(G=1.0, c=20.0, softening=2.5e-5, theta=0.5, pn_flags=A.NEWTON | A.PN1, precision=A.FP64)
    ctx.set_state(f["pos"], f["vel"], f["mass"])
    acc = ctx.forces(A.BH_PARITY)
    t = ctx.tree()
    for key, ref in (("codes", "codes"), ("perm", "perm"), ("left", "left"), ("right", "right"),
                     ("mass", "node_mass"), ("com", "node_com")):
        assert np.array_equal(t[key], f[ref]), key
    assert np.array_equal(acc, f["acc_bh"])
    codes, box = ctx.morton()
    assert np.array_equal(box, f["box"]) and np.array_equal(codes[t["perm"]], f["codes"])


def test_morton_corners_golden(ctx):
    f = fixture("morton_corners.npz")
    n = len(f["pos"])
    ctx.set_params(box=tuple(f["box"]), precision=A.FP64)
    ctx.set_state(f["pos"], np.zeros((n, 3)), np.ones(n))
    codes, box = ctx.morton()
    assert np.array_equal(codes, f["codes"])


@pytest.mark.parametrize("run", sorted(INTEGRATE_RUNS))
def test_integrate_golden(ctx, run):
    f = fixture("integrate_n64.npz")
    kw = INTEGRATE_RUNS[run]
    ctx.set_params(G=1.0, c=20.0, softening=1e-4, pn_flags=kw["pn"])
    ctx.set_state(f["pos"], f["vel"], f["mass"])
    if kw["integrator"] == 1:
        ctx.integrate(A.ALLPAIRS_FP64, A.SABA, kw["order"], 1e-3, kw["nsteps"])
    else:
        ctx.step(A.ALLPAIRS_FP64, INTEGRATORS[kw["integrator"]], 1e-3, kw["nsteps"])
    p, v = ctx.get_state()
    assert np.array_equal(p, f[run + "_pos"]) and np.array_equal(v, f[run + "_vel"])


@pytest.mark.parametrize("precision", [A.FP64, A.FP32])
def test_wisdom_holman_golden(ctx, precision):
    f = fixture("wh_planetary.npz")
    n = len(f["mass"])
    ctx.set_params(G=1.0, softening=0.0, pn_flags=A.NEWTON, precision=precision)
    ctx.batch_set(np.array([0, n]), f["pos"], f["vel"], f["mass"])
    ctx.batch_step(A.WH_PLANETARY, float(f["dt"]), int(f["nsteps"]), use_graph=False)
    p, v = ctx.batch_get()
    assert np.array_equal(p, f["pos_out"]) and np.array_equal(v, f["vel_out"])


def test_threebody_adjoint_golden(ctx):
    f = fixture("threebody_adjoint.npz")
    snaps = f["snaps"]
    ctx.set_params(G=1.0, softening=float(f["soft"]), pn_flags=A.NEWTON, precision=A.FP32)
    ctx.batch_set(np.array([0, 3]), f["pos"], f["vel"], f["mass"])
    s = ctx.batch_traj_forward(float(f["dt"]), int(f["nsteps"]), snaps)
    gsnap = 2.0 * (s - f["target"]) / (len(snaps) * 3)
    gx, gv, gm = ctx.batch_traj_backward(gsnap)
    loss = ((s - f["target"]) ** 2).sum() / (len(snaps) * 3)
    assert loss == pytest.approx(float(f["loss"]), rel=1e-10)
    assert relnorm(gx, f["gx"]) < 1e-8 and relnorm(gv, f["gv"]) < 1e-8 and relnorm(gm, f["gm"]) < 1e-8
