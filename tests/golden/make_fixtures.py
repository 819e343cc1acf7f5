"""Writes the golden fixtures in this folder: seeded inputs and the oracle's
FP64 outputs for them (Aster's CPU path restated in oracle/aster_oracle.cpp).

The inputs come from numpy's seeded generator, not from the library's own
initial-condition generators, so a fixture stays valid when those change.
Each file holds inputs and expected outputs only. Run from the repository
root after building the oracle:

    python tests/golden/make_fixtures.py

tests/test_golden.py checks that the oracle still reproduces every file bit
for bit; tests/test_gpu_golden.py checks the device paths against them."""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

# Flag sets: Newton; Newton + 1PN; Newton + 1PN + 2PN + 2.5PN.
PN_SETS = (0x1, 0x11, 0x1111)


# Trajectory fixtures of integrate_n64.npz: integrator id (0 Euler, 1 SABA of
# `order`, 2 leapfrog, 5 Yoshida-4), PN flags and step count; dt = 1e-3,
# G = 1, c = 20, softening 1e-4.
INTEGRATE_RUNS = {"leapfrog_newton": dict(integrator=2, pn=0x1, nsteps=10),
                  "leapfrog_pn1": dict(integrator=2, pn=0x11, nsteps=10),
                  "yoshida_pn1": dict(integrator=5, pn=0x11, nsteps=10),
                  "saba4_newton": dict(integrator=1, order=4, pn=0x1, nsteps=5),
                  "euler_newton": dict(integrator=0, pn=0x1, nsteps=5)}


def cluster(n, seed, vscale=0.3):
    """Gaussian blob with random velocities and masses in [0.5, 1.5], normalised to total mass 1."""
    r = np.random.default_rng(seed)
    pos = r.normal(size=(n, 3))
    vel = r.normal(scale=vscale, size=(n, 3))
    mass = r.uniform(0.5, 1.5, size=n)
    return pos, vel, mass / mass.sum()


def mixed(n, seed):
    """Half Gaussian blob, half uniform cube [-4, 4]^3 at rest; equal masses; FP32-representable positions."""
    r = np.random.default_rng(seed)
    h = n // 2
    pos = np.concatenate([r.normal(size=(h, 3)), r.uniform(-4.0, 4.0, size=(n - h, 3))])
    pos = pos.astype(np.float32).astype(np.float64)
    vel = np.concatenate([r.normal(scale=0.3, size=(h, 3)), np.zeros((n - h, 3))])
    return pos, vel, np.full(n, 1.0 / n)


def planetary(seed, nplanets=5):
    """Central mass 1 and low-mass planets on near-circular, near-coplanar orbits (G = 1)."""
    r = np.random.default_rng(seed)
    pos = np.zeros((nplanets + 1, 3))
    vel = np.zeros((nplanets + 1, 3))
    mass = np.empty(nplanets + 1)
    mass[0] = 1.0
    for k in range(1, nplanets + 1):
        a = 0.5 * 1.6 ** k
        ph = r.uniform(0, 2 * np.pi)
        inc = r.uniform(-0.03, 0.03)
        v = np.sqrt(1.0 / a) * (1.0 + r.uniform(-0.05, 0.05))
        pos[k] = a * np.array([np.cos(ph), np.sin(ph), inc])
        vel[k] = v * np.array([-np.sin(ph), np.cos(ph), 0.0])
        mass[k] = 10.0 ** r.uniform(-6, -3)
    return pos, vel, mass


def write(name, **arrays):
    path = os.path.join(HERE, name)
    np.savez_compressed(path, **arrays)
    print(f"{name}: {os.path.getsize(path)} bytes")


def main():
    orc.lib()

    # Karras (2012), "Maximizing parallelism in the construction of BVHs,
    # octrees, and k-d trees", Fig. 3: eight sorted 5-bit keys.
    keys = np.array([0b00001, 0b00010, 0b00100, 0b00101, 0b10011, 0b11000, 0b11001, 0b11110], np.uint32)
    left, right = orc.karras(keys)
    write("karras_fig3.npz", keys=keys, left=left, right=right)

    # Morton codes at the box corners, faces and centre, and just inside the top clamp.
    box = np.array([4.0, 2.0, 1.0])
    edge = np.nextafter(1.0, 0.0)
    unit = np.array([[-1, -1, -1], [1, 1, 1], [0, 0, 0], [-1, 1, -1], [1, -1, 1], [edge, edge, edge],
                     [-edge, -edge, -edge], [0.5, -0.25, 0.125], [1.0 / 1024, -1.0 / 1024, 0.0]])
    pos = unit * box
    write("morton_corners.npz", pos=pos, box=box, codes=orc.morton(pos, box))

    # FP64 all-pairs accelerations, G = 1, c = 20, softening 1e-4, each PN flag set.
    out = {}
    for n in (8, 64, 1024):
        p, v, m = cluster(n, 100 + n)
        out[f"pos{n}"], out[f"vel{n}"], out[f"mass{n}"] = p, v, m
        for f in PN_SETS:
            out[f"acc{n}_{f:x}"] = orc.allpairs(p, v, m, G=1.0, c=20.0, soft=1e-4, pn=f)
    write("forces.npz", **out)

    # LBVH of 1024 mixed bodies in the automatic box: Morton codes, sort
    # permutation, Karras topology, node masses and centres of mass; then the
    # reference-MAC Barnes-Hut accelerations at theta = 0.5 with 1PN.
    p, v, m = mixed(1024, 7)
    bx = orc.box_auto(p)
    t = orc.build_tree(p, v, m, bx)
    bh = orc.bh(p, v, m, bx, 0.5, G=1.0, c=20.0, soft=2.5e-5, pn=0x11)
    write("tree_bh_n1024.npz", pos=p, vel=v, mass=m, box=bx, codes=t["codes"], perm=t["perm"], left=t["left"],
          right=t["right"], node_mass=t["mass"], node_com=t["com"], acc_bh=bh)

    # Trajectories of 64 bodies on FP64 all-pairs forces (dt = 1e-3).
    p, v, m = cluster(64, 64)
    out = dict(pos=p, vel=v, mass=m)
    for name, kw in INTEGRATE_RUNS.items():
        xp, xv, _ = orc.integrate(p, v, m, 1e-3, kw["nsteps"], integrator=kw["integrator"], order=kw.get("order", 0),
                                  c=20.0, soft=1e-4, pn=kw["pn"])
        out[name + "_pos"], out[name + "_vel"] = xp, xv
    write("integrate_n64.npz", **out)

    # Integrator coefficients: Yoshida-4 weights and SABA2..10 drifts/kicks.
    out = dict(yoshida=orc.yoshida_weights())
    for k in range(2, 11):
        out[f"saba{k}_c"], out[f"saba{k}_d"] = orc.saba_coeffs(k)
    write("coeffs.npz", **out)

    # Wisdom-Holman (democratic heliocentric) on a six-body planetary system.
    p, v, m = planetary(5)
    wp, wv = orc.integrate_wh(p, v, m, 0.01, 200)
    write("wh_planetary.npz", pos=p, vel=v, mass=m, dt=0.01, nsteps=200, pos_out=wp, vel_out=wv)

    # Discrete adjoint of a 3-body leapfrog trajectory (softening 1e-2, dt 1e-3,
    # 200 steps, snapshots at 50/100/200; loss = mean over snapshots and bodies
    # of |x - target|^2).
    r = np.random.default_rng(3)
    p = r.uniform(0.0, 1.0, size=(3, 3))
    m = r.uniform(0.5, 1.5, size=3)
    v = r.normal(scale=0.3, size=(3, 3))
    v -= (m[:, None] * v).sum(0) / m.sum()
    snaps = np.array([50, 100, 200])
    target = np.stack([orc.integrate(p, v, m, 1e-3, s, soft=1e-2)[0] for s in snaps])
    target = target + r.normal(scale=1e-3, size=target.shape)
    loss, gx, gv, gm = orc.traj_grad(p, v, m, 1e-3, 200, snaps, target, soft=1e-2)
    write("threebody_adjoint.npz", pos=p, vel=v, mass=m, dt=1e-3, nsteps=200, soft=1e-2, snaps=snaps,
          target=target, loss=loss, gx=gx, gv=gv, gm=gm)


if __name__ == "__main__":
    main()
