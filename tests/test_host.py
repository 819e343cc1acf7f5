"""Host-side checks that need no GPU: the C-ABI library loads and exports every
declared symbol, the seeded initial-condition generators, and argument
validation that happens before any device call."""
import ctypes
import os
import re

import numpy as np
import pytest

import aster_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "aster_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(aster_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    L = aster_amd.lib()
    syms = declared_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(aster_amd.SIGNATURES), set(syms) ^ set(aster_amd.SIGNATURES)


def test_version_and_weights(orc):
    assert aster_amd.lib().aster_version() == 10000
    w = aster_amd.yoshida_weights()
    assert w[0] == w[2] and 2 * w[0] + w[1] == pytest.approx(1.0, abs=1e-15)
    assert np.array_equal(w, orc.yoshida_weights())  # same bits as the oracle
    for n in range(2, 11):
        cs, ds = aster_amd.saba_coeffs(n)
        ocs, ods = orc.saba_coeffs(n)
        assert np.array_equal(cs, ocs) and np.array_equal(ds, ods), n


def test_params_struct_layout():
    assert ctypes.sizeof(aster_amd.Params) == 8 * 7 + 4 * 6


def test_plummer_ic():
    pos, vel, mass = aster_amd.ic_plummer(4096, 1)
    assert mass.sum() == pytest.approx(1.0, rel=1e-14)
    assert np.all(mass == mass[0])
    assert np.abs((mass[:, None] * pos).sum(0)).max() < 1e-14
    assert np.abs((mass[:, None] * vel).sum(0)).max() < 1e-14
    r = np.linalg.norm(pos, axis=1)
    assert r.max() <= 10.0 + 1e-9
    # half-mass radius of a Plummer sphere: b / sqrt(2^{2/3} - 1) = 1.305 b
    assert np.median(r) == pytest.approx(1.305, rel=0.05)
    # virial equilibrium: 2K / |W| ~ 1 (W for G = 1, unsoftened)
    K = 0.5 * (mass * (vel ** 2).sum(1)).sum()
    d = pos[:, None, :] - pos[None, :, :]
    rr = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(rr, np.inf)
    W = -0.5 * (mass[:, None] * mass[None, :] / rr).sum()
    assert 2 * K / abs(W) == pytest.approx(1.0, abs=0.06)
    # deterministic per seed, different across seeds
    p2, _, _ = aster_amd.ic_plummer(4096, 1)
    p3, _, _ = aster_amd.ic_plummer(4096, 2)
    assert np.array_equal(pos, p2) and not np.array_equal(pos, p3)


def test_plummer_lognormal_masses():
    _, _, mass = aster_amd.ic_plummer(20000, 5, mass_mode=1)
    assert mass.sum() == pytest.approx(1.0, rel=1e-13)
    assert mass.max() / mass.min() == pytest.approx(100.0, rel=0.02)
    lg = np.log10(mass / mass.min())
    assert np.histogram(lg, bins=4, range=(0, 2))[0].min() > 0.2 * len(mass)  # ~uniform in log


def test_cube_and_mixed():
    pos, vel, mass = aster_amd.ic_cube(1000, 3, h=4.0)
    assert np.all(np.abs(pos) <= 4.0) and np.all(vel == 0) and mass.sum() == pytest.approx(1.0)
    pos, vel, mass = aster_amd.ic_mixed(1 << 12, 4)
    assert np.array_equal(pos, pos.astype(np.float32).astype(np.float64))
    assert np.array_equal(mass, mass.astype(np.float32).astype(np.float64))
    assert np.all(np.abs(pos[2048:]) <= 4.0)


def test_ensemble_ic():
    off, pos, vel, mass = aster_amd.ic_ensemble(500, 9, 3, 64)
    sizes = np.diff(off)
    assert sizes.min() >= 3 and sizes.max() <= 64 and len(sizes) == 500
    assert sizes.min() <= 5 and sizes.max() >= 60
    for s in range(0, 500, 50):
        a, b = off[s], off[s + 1]
        m, x, v = mass[a:b], pos[a:b], vel[a:b]
        assert m[0] == 1.0 and np.all((m[1:] >= 1e-7) & (m[1:] <= 1e-3))
        assert np.abs((m[:, None] * x).sum(0)).max() < 1e-14
        assert np.abs((m[:, None] * v).sum(0)).max() < 1e-14
        # osculating elements of each planet about the central body
        rel = x[1:] - x[0]
        rv = v[1:] - v[0]
        mu = 1.0 + m[1:]
        rn = np.linalg.norm(rel, axis=1)
        en = 0.5 * (rv ** 2).sum(1) - mu / rn
        a_ = -mu / (2 * en)
        h = np.cross(rel, rv)
        e_ = np.sqrt(np.maximum(0.0, 1 + 2 * en * (h ** 2).sum(1) / mu ** 2))
        inc = np.degrees(np.arccos(h[:, 2] / np.linalg.norm(h, axis=1)))
        assert np.all((a_ > 0.4 - 1e-9) & (a_ < 30 + 1e-9))
        assert np.all(e_ < 0.3 + 1e-9) and np.all(inc < 5 + 1e-9)


def test_threebody_ic():
    pos, vel, mass = aster_amd.ic_threebody(100, 2)
    assert np.all((mass >= 0.5) & (mass <= 1.5))
    assert np.all((pos >= 0) & (pos <= 1))
    p = (mass[:, None] * vel).reshape(100, 3, 3).sum(1)
    assert np.abs(p).max() < 1e-15


def test_invalid_arguments_rejected_without_device():
    L = aster_amd.lib()
    assert L.aster_ic_plummer(0, 1, 1.0, 1.0, 10.0, 0, None, None, None) == -1
    assert L.aster_ic_ensemble(10, 1, 3, 65, None, None, None, None, None) == -1
    assert L.aster_create(None, 0) == -1
    assert L.aster_yoshida_weights(None) == -1
    c11, d11 = np.empty(12), np.empty(11)
    for order in (0, 1, 11):  # SABA tables exist for orders 2..10
        assert L.aster_saba_coeffs(order, c11.ctypes.data_as(aster_amd._D), d11.ctypes.data_as(aster_amd._D)) == -1
    assert L.aster_integrate(None, 0, aster_amd.SABA, 4, 1e-3, 1) == -1
    assert L.aster_set_state_device(None, 4, None, None, None) == -1
    assert L.aster_get_state_device(None, None, None) == -1
    assert L.aster_collect(None, aster_amd.COLLECT_HAMILTONIAN, 0, None) == -1
    assert L.aster_center_of_mass(None, None) == -1
    assert L.aster_integrate_report(None, 0, aster_amd.LEAPFROG, 0, 1e-3, 1, 1, None) == -1
    assert L.aster_integrate_until(None, 0, aster_amd.LEAPFROG, 0, 1e-3, aster_amd.HALTING_CONDITION(), None, 1, 0,
                                   None) == -1
    assert L.aster_cluster_add(None, None) == -1
    assert L.aster_cluster_pop(None) == -1
    assert L.aster_cluster_queued(None, None, None) == -1
    assert L.aster_cluster_load(None) == -1
    assert L.aster_destroy(None) == 0
    assert L.aster_last_error(None)  # message for failed creation


def test_create_reports_missing_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(aster_amd.AsterError) as ei:
        aster_amd.Context(0)
    assert ei.value.code in (-1, -2)
