"""Full N-body 1PN (Einstein-Infeld-Hoffmann) force law of the oracle, pinned
by the two-body limit (the pairwise 1PN law), an independent numpy
evaluation of the published equation (Newhall, Standish & Williams 1983,
eq. 1, beta = gamma = 1) and conservation of the EIH energy (Landau &
Lifshitz, Classical Theory of Fields, sec. 106) along trajectories."""
import numpy as np
import pytest

EIH = 0x21  # ASTER_NEWTON | ASTER_EIH


def eih_numpy(pos, vel, mass, G, c):
    """Vectorised Newhall et al. (1983) eq. 1 with beta = gamma = 1."""
    mu = G * mass
    d = pos[None, :, :] - pos[:, None, :]          # d[i, j] = x_j - x_i
    r = np.linalg.norm(d, axis=2)
    np.fill_diagonal(r, np.inf)
    inv = 1.0 / r
    phi = (mu[None, :] * inv).sum(1)               # phi_i
    an = (mu[None, :, None] * d * inv[:, :, None] ** 3).sum(1)
    v2 = (vel ** 2).sum(1)
    vij = vel @ vel.T
    rv = np.einsum("ijk,jk->ij", d, vel) * inv     # d_ij . v_j / r
    dan = np.einsum("ijk,jk->ij", d, an)
    ic2 = 1.0 / c ** 2
    B = 1 - (4 * phi[:, None] + phi[None, :]) * ic2 + (v2[:, None] + 2 * v2[None, :] - 4 * vij) * ic2 \
        - 1.5 * rv ** 2 * ic2 + 0.5 * dan * ic2
    s1 = mu[None, :] * inv ** 3 * B
    w = -(np.einsum("ijk,ik->ij", d, 4 * vel) - np.einsum("ijk,jk->ij", d, 3 * vel))
    s2 = mu[None, :] * inv ** 3 * ic2 * w
    s3 = 3.5 * mu[None, :] * inv * ic2
    a = (s1[:, :, None] * d).sum(1) + (s2[:, :, None] * (vel[:, None, :] - vel[None, :, :])).sum(1) \
        + (s3[:, :, None] * an[None, :, :]).sum(1)
    return a


def energy_1pn(pos, vel, mass, G, c):
    """E from the EIH Lagrangian: kinetic + potential + 1/c^2 [3/8 m v^4
    + sum_{a<b} G m_a m_b/(2 r)(3(v_a^2 + v_b^2) - 7 v_a.v_b - (n.v_a)(n.v_b))
    + sum_a sum_{b != a} sum_{c != a} G^2 m_a m_b m_c / (2 r_ab r_ac)]."""
    n = len(mass)
    v2 = (vel ** 2).sum(1)
    E0 = 0.5 * (mass * v2).sum()
    E1 = (3.0 / 8.0) * (mass * v2 ** 2).sum()
    for a in range(n):
        for b in range(a + 1, n):
            d = pos[b] - pos[a]
            r = np.linalg.norm(d)
            nn = d / r
            E0 -= G * mass[a] * mass[b] / r
            E1 += G * mass[a] * mass[b] / (2 * r) * (3 * (v2[a] + v2[b]) - 7 * vel[a] @ vel[b]
                                                    - (nn @ vel[a]) * (nn @ vel[b]))
    for a in range(n):
        s = sum(G * mass[b] / np.linalg.norm(pos[b] - pos[a]) for b in range(n) if b != a)
        E1 += 0.5 * mass[a] * s * s
    return E0 + E1 / c ** 2


def test_eih_two_bodies_is_pairwise_1pn(orc):
    r = np.random.default_rng(1)
    for _ in range(20):
        pos, vel = r.normal(size=(2, 3)), r.normal(scale=0.3, size=(2, 3))
        mass = r.uniform(0.1, 1.0, 2)
        a = orc.allpairs(pos, vel, mass, G=1.0, c=7.0, soft=0.0, pn=EIH)
        b = orc.allpairs(pos, vel, mass, G=1.0, c=7.0, soft=0.0, pn=0x11)
        assert np.allclose(a, b, rtol=1e-13, atol=0)


def test_eih_matches_numpy(orc):
    r = np.random.default_rng(2)
    pos, vel = r.normal(size=(7, 3)), r.normal(scale=0.4, size=(7, 3))
    mass = r.uniform(0.2, 2.0, 7)
    a = orc.allpairs(pos, vel, mass, G=1.3, c=5.0, soft=0.0, pn=EIH)
    ref = eih_numpy(pos, vel, mass, 1.3, 5.0)
    assert np.allclose(a, ref, rtol=1e-12, atol=1e-14)
    # the 1PN part is a real change of the Newtonian acceleration here
    newton = orc.allpairs(pos, vel, mass, G=1.3, c=5.0, soft=0.0, pn=0x1)
    assert np.abs(a - newton).max() > 1e-3 * np.abs(newton).max()


def figure_eight():
    """Chenciner & Montgomery (2000) figure-eight: G = m = 1, all three bodies
    close together, so the cross terms of the 1PN law matter."""
    m = np.ones(3)
    pos = np.array([[0.97000436, -0.24308753, 0.0], [-0.97000436, 0.24308753, 0.0], [0.0, 0.0, 0.0]])
    v3 = np.array([-0.93240737, -0.86473146, 0.0])
    return pos, np.array([-v3 / 2, -v3 / 2, v3]), m


def rk4_drift(orc, pos, vel, m, flags, c, dt=2e-3, T=3.0):
    """Largest |E_1PN / E_1PN(0) - 1| along an RK4 trajectory (fourth order
    for velocity-dependent forces, unlike the leapfrog)."""
    f = lambda p, v: orc.allpairs(p, v, m, G=1.0, c=c, soft=0.0, pn=flags)
    p, v = pos.copy(), vel.copy()
    E0 = energy_1pn(p, v, m, 1.0, c)
    worst = 0.0
    for s in range(int(round(T / dt))):
        k1x, k1v = v, f(p, v)
        k2x, k2v = v + 0.5 * dt * k1v, f(p + 0.5 * dt * k1x, v + 0.5 * dt * k1v)
        k3x, k3v = v + 0.5 * dt * k2v, f(p + 0.5 * dt * k2x, v + 0.5 * dt * k2v)
        k4x, k4v = v + dt * k3v, f(p + dt * k3x, v + dt * k3v)
        p = p + dt / 6 * (k1x + 2 * k2x + 2 * k3x + k4x)
        v = v + dt / 6 * (k1v + 2 * k2v + 2 * k3v + k4v)
        if s % 50 == 49:
            worst = max(worst, abs(energy_1pn(p, v, m, 1.0, c) / E0 - 1.0))
    return worst


def test_eih_conserves_its_energy(orc):
    # The EIH energy is conserved to O(1/c^4): doubling c divides its drift
    # by ~16. The pairwise law lacks the cross terms and breaks it at
    # O(1/c^2) (ratio -> 4). Measured on the figure-eight: EIH 1.3e-5 / 8.4e-7
    # at c = 40 / 80 (ratio 15.9), pairwise 1.8e-4 / 3.7e-5 (ratio 4.9).
    pos, vel, m = figure_eight()
    e40, e80 = (rk4_drift(orc, pos, vel, m, EIH, c) for c in (40.0, 80.0))
    p40, p80 = (rk4_drift(orc, pos, vel, m, 0x11, c) for c in (40.0, 80.0))
    print(f"EIH {e40:.3e} {e80:.3e}; pairwise {p40:.3e} {p80:.3e}")
    assert e40 / e80 > 12.0 and e80 < 2e-6
    assert p40 / p80 < 8.0 and p80 > 10 * e80
