"""The oracle against the committed golden fixtures (tests/golden/*.npz).

Each fixture holds seeded inputs and the FP64 outputs the oracle gave for them
when the file was written (tests/golden/make_fixtures.py). Recomputing them
bit for bit guards the oracle itself against regressions; the Karras and
Morton files are also checked against published or closed-form values, and
the coefficient file against the library's own tables."""
import os

import numpy as np
import pytest

from tests.golden.make_fixtures import INTEGRATE_RUNS, PN_SETS

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def fixture(name):
    with np.load(os.path.join(GOLDEN, name), allow_pickle=False) as f:
        return {k: f[k] for k in f.files}


def test_karras_fig3_fixture(orc):
    f = fixture("karras_fig3.npz")
    n = 8
    I = lambda i: n + i  # internal node i
    # Karras (2012) Fig. 3 topology
    expect = {0: (I(3), I(4)), 1: (0, 1), 2: (2, 3), 3: (I(1), I(2)), 4: (4, I(5)), 5: (I(6), 7), 6: (5, 6)}
    for i, (l, r) in expect.items():
        assert (f["left"][I(i)], f["right"][I(i)]) == (l, r), i
    left, right = orc.karras(f["keys"])
    assert np.array_equal(left, f["left"]) and np.array_equal(right, f["right"])


def test_morton_corner_fixture(orc):
    f = fixture("morton_corners.npz")
    assert np.array_equal(orc.morton(f["pos"], f["box"]), f["codes"])
    # closed forms: -B -> cell 0 on every axis; +B and just below +B -> 1023
    # on every axis (the clamp); the centre -> cell 512 on every axis.
    assert f["codes"][0] == 0
    assert f["codes"][1] == f["codes"][5] == (1 << 30) - 1
    c512 = sum(1 << (3 * 9 + a) for a in range(3))
    assert f["codes"][2] == c512


@pytest.mark.parametrize("n", [8, 64, 1024])
def test_forces_fixture(orc, n):
    f = fixture("forces.npz")
    p, v, m = f[f"pos{n}"], f[f"vel{n}"], f[f"mass{n}"]
    for flags in PN_SETS:
        acc = orc.allpairs(p, v, m, G=1.0, c=20.0, soft=1e-4, pn=flags)
        assert np.array_equal(acc, f[f"acc{n}_{flags:x}"]), hex(flags)


def test_tree_and_bh_fixture(orc):
    f = fixture("tree_bh_n1024.npz")
    box = orc.box_auto(f["pos"])
    assert np.array_equal(box, f["box"])
    t = orc.build_tree(f["pos"], f["vel"], f["mass"], box)
    for key, ref in (("codes", "codes"), ("perm", "perm"), ("left", "left"), ("right", "right"),
                     ("mass", "node_mass"), ("com", "node_com")):
        assert np.array_equal(t[key], f[ref]), key
    acc = orc.bh(f["pos"], f["vel"], f["mass"], box, 0.5, G=1.0, c=20.0, soft=2.5e-5, pn=0x11)
    assert np.array_equal(acc, f["acc_bh"])


@pytest.mark.parametrize("run", sorted(INTEGRATE_RUNS))
def test_integrate_fixture(orc, run):
    f = fixture("integrate_n64.npz")
    kw = INTEGRATE_RUNS[run]
    p, v, _ = orc.integrate(f["pos"], f["vel"], f["mass"], 1e-3, kw["nsteps"], integrator=kw["integrator"],
                            order=kw.get("order", 0), c=20.0, soft=1e-4, pn=kw["pn"])
    assert np.array_equal(p, f[run + "_pos"]) and np.array_equal(v, f[run + "_vel"])


def test_coefficient_fixture(orc):
    import aster_amd
    f = fixture("coeffs.npz")
    assert np.array_equal(orc.yoshida_weights(), f["yoshida"])
    assert np.array_equal(aster_amd.yoshida_weights(), f["yoshida"])
    for k in range(2, 11):
        c, d = orc.saba_coeffs(k)
        assert np.array_equal(c, f[f"saba{k}_c"]) and np.array_equal(d, f[f"saba{k}_d"]), k
        c, d = aster_amd.saba_coeffs(k)
        assert np.array_equal(c, f[f"saba{k}_c"]) and np.array_equal(d, f[f"saba{k}_d"]), k


def test_wisdom_holman_fixture(orc):
    f = fixture("wh_planetary.npz")
    p, v = orc.integrate_wh(f["pos"], f["vel"], f["mass"], float(f["dt"]), int(f["nsteps"]))
    assert np.array_equal(p, f["pos_out"]) and np.array_equal(v, f["vel_out"])


def test_threebody_adjoint_fixture(orc):
    f = fixture("threebody_adjoint.npz")
    loss, gx, gv, gm = orc.traj_grad(f["pos"], f["vel"], f["mass"], float(f["dt"]), int(f["nsteps"]), f["snaps"],
                                     f["target"], soft=float(f["soft"]))
    assert loss == float(f["loss"])
    assert np.array_equal(gx, f["gx"]) and np.array_equal(gv, f["gv"]) and np.array_equal(gm, f["gm"])
