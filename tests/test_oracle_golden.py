"""The oracle against the golden fixtures in tests/golden/ (outputs of the
reference's own code, written by tests/golden/make_golden.py).  These run
everywhere, also where the reference checkout is absent.

Phases replayed from the fixture's own inputs are bit-exact; the deposit is
compared within FP64 rounding (the reference sums run-length groups in array
order, the oracle per cell).
"""
import os

import numpy as np
import pytest

from oracle_ctypes import (AX1, AY1, AZ1, BTX, BTY, BTZ, BX, BY, BZ, EX, EY, EZ, JX, JY, JZ, PHI1, PHI2,
                           RHO, Case, Oracle, cell_ids_from_x, oracle_rng_gaussian, oracle_rng_uniform)

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys_path_golden = HERE
import sys  # noqa: E402

sys.path.insert(0, HERE)
import make_golden as mg  # noqa: E402

FIELDS = mg.FIELDS


def load(name):
    return np.load(os.path.join(HERE, name))


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def test_g1_rng():
    g = load("g1_rng.npz")
    for seed in (-1, 1233):
        assert np.array_equal(oracle_rng_uniform(seed, 1000), g[f"uniform_{seed}"])
        assert np.array_equal(oracle_rng_gaussian(seed, 1000), g[f"gauss_{seed}"])


def _oracle_from(g, case):
    o = Oracle(case)
    o.set_species(list(g["q"]), list(g["m"]))
    o.set_particles(g["load_x"], g["load_v"], g["load_gamma"], g["load_species"], g["load_id"])
    return o


def _by_id(o):
    return o.particles().by_id()


@pytest.mark.parametrize("relativity,fname", [(0, "g2_boris.npz"), (1, "g2_vay.npz")])
def test_g2_load_and_g3_step(relativity, fname):
    g = load(fname)
    case = Case(**mg.CASE_G2, relativity=relativity, E0=(0.01, -0.02, 0.03), B0=(0.05, 0.02, -0.04))
    o = Oracle(case)
    assert o.load() == 0
    p = o.particles()
    for k in ("x", "v", "gamma", "species", "id"):
        assert np.array_equal(getattr(p, k), g[f"load_{k}"]), k
    assert np.array_equal(o.q, g["q"]) and np.array_equal(o.m, g["m"])
    dt = float(g["dt"])
    o.initialize()
    assert np.array_equal(_by_id(o).field, g["init_field"])
    o.push(dt)
    p = _by_id(o)
    for k in ("x", "v", "gamma"):
        assert np.array_equal(getattr(p, k), g[f"push_{k}"]), k
    o.particle_bc()
    p = _by_id(o)
    assert np.array_equal(p.x, g["bc_x"]) and np.array_equal(p.v, g["bc_v"])
    o.sort()
    o.deposit(True)
    for k in ("Jx", "Jy", "Jz", "rho"):
        assert rel(o.field(FIELDS[k]), g[f"dep_{k}"]) < 1e-14
        o.set_field(FIELDS[k], g[f"dep_{k}"])
    o.fold()
    for k in ("Jx", "Jy", "Jz", "rho"):
        assert np.array_equal(o.field(FIELDS[k]), g[f"fold_{k}"]), k
    o.evolve(dt)
    for k, fid in FIELDS.items():
        assert np.array_equal(o.field(fid), g[f"evolve_{k}"]), k
    o.refresh(0.0, dt)
    for k, fid in FIELDS.items():
        assert np.array_equal(o.field(fid), g[f"refresh_{k}"]), k
    o.gather()
    assert np.array_equal(_by_id(o).field, g["gather_field"])


@pytest.mark.parametrize("relativity,fname", [(0, "g2_boris.npz"), (1, "g2_vay.npz")])
def test_g5_history_and_20_steps(relativity, fname):
    g = load(fname)
    case = Case(**mg.CASE_G2, relativity=relativity, E0=(0.01, -0.02, 0.03), B0=(0.05, 0.02, -0.04))
    o = Oracle(case)
    o.load()
    o.initialize()
    dt = float(g["dt"])
    hist = []
    t = 0.0
    for ti in range(20):
        if ti % 10 == 0:
            hist.append(o.diag()[:7])
        o.step(t, dt)
        t += dt
    hist.append(o.diag()[:7])
    # history.dat carries 16 significant digits
    np.testing.assert_allclose(np.array(hist), g["history"][:, :7], rtol=1e-12)
    p = _by_id(o)
    assert np.array_equal(cell_ids_from_x(case, p.x), cell_ids_from_x(case, g["steps20_x"]))
    assert rel(p.x, g["steps20_x"]) < 1e-13 and rel(p.v, g["steps20_v"]) < 1e-12
    for k in ("Ex", "By", "Jz", "rho"):
        assert rel(o.field(FIELDS[k]), g[f"steps20_{k}"]) < 1e-12, k


@pytest.mark.parametrize("tag", ["one", "conv"])
def test_g4_poisson(tag):
    g = load("g4_poisson.npz")
    bphi = (0.0, 1e12, 0.0, 0.0, 0.0, 0.0) if tag == "one" else (0.0, 3e-2, 0.0, 0.0, 0.0, 0.0)
    case = Case(**{**mg.CASE_G2, "n": (4, 5, 3), "L": (4.0, 5.0, 3.0)}, poisson=True, bound_phi=bphi,
                bound_Ax=(0.0, 1e12, 0.0, 0.0, 0.0, 0.0) if tag == "one" else (0.0,) * 6)
    o = Oracle(case)
    for k, fid in (("rho", RHO), ("Jx", JX), ("Jy", JY), ("Jz", JZ)):
        o.set_field(fid, g[f"{tag}_src_{k}"])
    sw = o.poisson_init()
    if tag == "one":
        assert list(sw) == [1, 1, 1, 1]
    for k, fid in (("phi1", PHI1), ("phi2", PHI2), ("Ax1", AX1), ("Ay1", AY1), ("Az1", AZ1), ("Ex", EX),
                   ("Ey", EY), ("Ez", EZ), ("Bx", BX), ("By", BY), ("Bz", BZ)):
        assert np.array_equal(o.field(fid), g[f"{tag}_{k}"]), k


def test_g6_lightbc_reflecting():
    g = load("g6_lightbc.npz")
    case = Case(**mg.CASE_G6)
    o = _oracle_from(g, case)
    o.initialize()
    dt = float(g["dt"])
    t = 0.0
    for _ in range(3):
        o.step(t, dt)
        t += dt
    for k, fid in FIELDS.items():
        assert rel(o.field(fid), g[f"step3_{k}"]) < 1e-12, k
    # replay step 4 from the fixture's own state, bit-exact
    o2 = Oracle(case)
    o2.set_species(list(g["q"]), list(g["m"]))
    o2.set_particles(g["step3_x"], g["step3_v"], g["step3_gamma"], g["load_species"][np.argsort(g["load_id"])],
                     np.sort(g["load_id"]))
    o2.set_particle_fields(g["step3_field"])
    o2.set_b_alias(True)  # three evolveFields swaps so far
    for k, fid in FIELDS.items():
        o2.set_field(fid, g[f"step3_{k}"])
    o2.push(dt)
    o2.particle_bc()
    p = _by_id(o2)
    assert np.array_equal(p.x, g["bc4_x"]) and np.array_equal(p.v, g["bc4_v"])
    for k in ("Jx", "Jy", "Jz", "rho"):
        o2.set_field(FIELDS[k], g[f"fold4_{k}"])
    o2.evolve(dt)
    o2.refresh(t, dt)
    for k, fid in FIELDS.items():
        assert np.array_equal(o2.field(fid), g[f"refresh4_{k}"]), k
