"""The oracle against the reference's own compiled code (oracle/_ref).

oracle/_ref/libemo_ref.so is built by oracle/Makefile from the reference
sources where they lie; these tests skip when it has not been built (e.g.
on a machine without the reference checkout).  Per-phase results on the
same inputs are bit-identical; the deposit differs only in summation order
(the reference flushes run-length groups in array order), so J/rho and
everything downstream of them over many steps agree to FP64 rounding.
"""
import numpy as np
import pytest

from oracle_ctypes import (AX1, AZ2, BTX, BTY, BTZ, BX, BY, BZ, EX, EXTERNAL, EY, EZ, JX, JY, JZ,
                           PERIODIC, PHI1, PHI2, REFLECTING, RHO, Case, Oracle, Reference, Wave,
                           cell_ids_from_x, oracle_lib, oracle_rng_gaussian, oracle_rng_uniform,
                           ref_available, ref_rng_gaussian, ref_rng_uniform)

pytestmark = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (no reference checkout)")

EB = [EX, EY, EZ, BX, BY, BZ]
ALL = EB + [JX, JY, JZ, RHO, BTX, BTY, BTZ]


def rel(a, b):
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


THERMAL = dict(n=(4, 5, 6), L=(4.0, 5.0, 3.0), nparticles=480, dens_phys=2.824e9,
               E0=(0.01, -0.02, 0.03), B0=(0.05, 0.02, -0.04), temp=(5000.0, 5000.0))
LIGHT = dict(n=(8, 4, 5), L=(8.0, 4.0, 5.0), nparticles=640, dens_phys=4.355e9, temp=(5000.0, 5000.0),
             waves=[Wave(-1, 2, 0.3, 0.39, 30.0, 0.05, 5.0), Wave(1, 3, 0.1, 0.7, -10.0, 0.1, 2.0),
                    Wave(-1, 3, 0.05, 0.5, 90.0, 0.2, 1.0)],
             fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
             pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC), B0=(0.0, 0.0, 0.01))


@pytest.mark.parametrize("seed", [-1, 1233, -5, 0])
def test_rng_streams_bit_exact(seed):
    assert np.array_equal(ref_rng_uniform(seed, 2000), oracle_rng_uniform(seed, 2000))
    assert np.array_equal(ref_rng_gaussian(seed, 2001, 0.3, 0.7), oracle_rng_gaussian(seed, 2001, 0.3, 0.7))


@pytest.mark.parametrize("poisson,relativity", [(True, 0), (False, 1), (False, 0)])
def test_loader_bit_exact(poisson, relativity):
    if poisson:
        # zeroAvgChargeAndVelocity_ asserts neutrality (particle_handler.cpp:435-438)
        sp = dict(mass_ratio=(1.0, 1836.0), charge_ratio=(-1.0, 1.0), dens_frac=(0.5, 0.5),
                  temp=(500.0, 20.0), is_test=(0, 0))
    else:
        sp = dict(mass_ratio=(1.0, 1836.0, 4.0), charge_ratio=(-1.0, 1.0, 0.0), dens_frac=(0.5, 0.3, 0.2),
                  temp=(500.0, 20.0, 3.0), is_test=(0, 0, 1))
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=300, poisson=poisson, relativity=relativity,
                dens_phys=2.824e9, **sp)
    r = Reference(case)
    o = Oracle(case)
    assert o.load() == 0
    pr, po = r.particles(), o.particles()
    for k in ("x", "v", "gamma", "species", "id"):
        assert np.array_equal(getattr(pr, k), getattr(po, k)), k
    q, m = r.species()
    assert np.array_equal(q, o.q) and np.array_equal(m, o.m)


@pytest.mark.parametrize("params,relativity", [(THERMAL, 0), (THERMAL, 1), (LIGHT, 1)])
def test_phases_bit_exact(params, relativity):
    case = Case(**params, relativity=relativity)
    r = Reference(case, nstep_sort=1)
    o = Oracle(case)
    o.load()
    r.initialize()
    o.initialize()
    assert np.array_equal(r.particles().field, o.particles().field)
    dt = case.dt_cfl()
    for step in range(3):
        t = step * dt
        r.push(dt)
        o.push(dt)
        pr, po = r.particles().by_id(), o.particles().by_id()
        assert np.array_equal(pr.x, po.x) and np.array_equal(pr.v, po.v) and np.array_equal(pr.gamma, po.gamma)
        r.particle_bc()
        o.particle_bc()
        pr, po = r.particles().by_id(), o.particles().by_id()
        assert np.array_equal(pr.x, po.x) and np.array_equal(pr.v, po.v)
        assert len(pr.x) == case.nparticles
        o.sort()
        r.deposit()
        o.deposit()
        for f in (JX, JY, JZ, RHO):
            assert rel(o.field(f), r.field(f)) < 1e-14
            r.set_field(f, o.field(f))  # continue from identical sources
        r.fold()
        o.fold()
        for f in (JX, JY, JZ, RHO):
            assert np.array_equal(r.field(f), o.field(f))
        r.evolve(dt)
        o.evolve(dt)
        for f in ALL:
            assert np.array_equal(r.field(f), o.field(f)), f
        r.refresh(t, dt)
        o.refresh(t, dt)
        for f in ALL:
            assert np.array_equal(r.field(f), o.field(f)), f
        r.gather()
        o.gather()
        assert np.array_equal(r.particles().by_id().field, o.particles().by_id().field)


def test_stale_fieldptr_after_swap():
    # Q23: after an odd number of evolveFields calls Grid::fieldPtr_[B] is the
    # member B_tm1; the oracle keeps the same aliasing
    case = Case(**THERMAL)
    r = Reference(case)
    r.initialize()
    r.deposit()
    r.evolve(0.1)
    assert np.array_equal(r.field_ptr(BX), r.field(BTX))
    assert not np.array_equal(r.field_ptr(BX), r.field(BX))
    r.evolve(0.1)
    assert np.array_equal(r.field_ptr(BX), r.field(BX))


@pytest.mark.parametrize("params,relativity,steps", [(THERMAL, 0, 40), (THERMAL, 1, 40), (LIGHT, 1, 30)])
def test_trajectory_matches(params, relativity, steps):
    case = Case(**params, relativity=relativity)
    r = Reference(case, nstep_sort=100)
    o = Oracle(case)
    o.load()
    r.initialize()
    o.initialize()
    dt = case.dt_cfl()
    t = 0.0
    for ti in range(steps):
        r.step(ti, t, dt)
        o.step(t, dt)
        t += dt
    pr, po = r.particles().by_id(), o.particles().by_id()
    assert np.array_equal(cell_ids_from_x(case, pr.x), cell_ids_from_x(case, po.x))
    assert rel(po.x, pr.x) < 1e-13 and rel(po.v, pr.v) < 1e-12
    for f in ALL:
        assert rel(o.field(f), r.field(f)) < 1e-12, f
    d_r, d_o = r.diag(t, 0), o.diag()
    np.testing.assert_allclose(d_o[:7], d_r[:7], rtol=1e-12)


def test_poisson_bit_exact_on_same_sources():
    case = Case(n=(4, 5, 6), L=(4.0, 5.0, 3.0), nparticles=240, poisson=True, dens_phys=2.824e9,
                bound_phi=(0.0, 0.02, 0.0, 0.0, 0.0, 0.0))
    r = Reference(case)
    o = Oracle(case)
    o.load()
    o.sort()
    o.deposit(True)
    o.fold()
    for f in (RHO, JX, JY, JZ):
        r.set_field(f, o.field(f))
    r.poisson_only()
    o.poisson_init()
    for f in EB + [PHI1, PHI2, AX1, AZ2]:
        assert np.array_equal(r.field(f), o.field(f)), f


def test_poisson_external_bit_exact():
    case = Case(n=(6, 4, 4), L=(6.0, 4.0, 4.0), nparticles=192, poisson=True, dens_phys=2.824e9,
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                bound_phi=(0.1, -0.2, 0.0, 0.0, 0.0, 0.0), bound_Ay=(0.05, 0.0, 0.0, 0.0, 0.0, 0.0),
                waves=[Wave(-1, 2, 0.3, 0.39, 30.0, 0.05, 5.0)])
    r = Reference(case)
    o = Oracle(case)
    o.load()
    o.sort()
    o.deposit(True)
    o.fold()
    for f in (RHO, JX, JY, JZ):
        r.set_field(f, o.field(f))
    r.poisson_only()
    o.poisson_init()
    for f in EB + [PHI1, PHI2, AX1, AZ2]:
        assert np.array_equal(r.field(f), o.field(f)), f


def test_initialize_poisson_path_matches():
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=128, poisson=True, dens_phys=2.824e9,
                bound_phi=(0.0, 3e-2, 0.0, 0.0, 0.0, 0.0))
    r = Reference(case)
    o = Oracle(case)
    o.load()
    r.initialize()
    o.initialize()
    for f in EB + [RHO, JX]:
        assert rel(o.field(f), r.field(f)) < 1e-12, f
    assert rel(o.particles().by_id().field, r.particles().by_id().field) < 1e-12


def test_cell_id_closed_forms_match_reference():
    # oGrid_unittests.cc:43-70 geometry (5,10,20 cells on a unit box)
    case = Case(n=(5, 10, 20), L=(1.0, 1.0, 1.0), nparticles=2)
    r = Reference(case)
    o = Oracle(case)
    pts = [(0.501, 0.501, 0.501), (0, 0, 0), (0, 0, 1), (0, 0, 0.5), (0, 1, 0.5), (0, 0.5, 0.5),
           (1, 0.5, 0.5), (0.2, 0.3, 0.35), (0.9999999, 0.1, 0.05)]
    for p in pts:
        assert r.cell_id(*p) == o.cell_id(*p)


def test_gaussian_pulses_bit_exact():
    import ctypes as C

    amp = np.array([0.1, 0.3, -0.2])
    omg = np.array([1.2, 5.6, 3.7])
    pha = np.array([90.0, 0.0, 45.0])
    dwt = np.array([0.2, 0.0, 0.8])
    dl = np.array([-0.2, 0.0, 1.2])
    from oracle_ctypes import _p, ref_lib

    for t in np.linspace(-1, 20, 57):
        for dt in (0.0, 0.37):
            a = ref_lib().rh_gaussian_pulses(3, _p(amp), _p(omg), _p(pha), _p(dl), _p(dwt), C.c_double(t), C.c_double(dt))
            b = oracle_lib().eo_gaussian_pulses(3, _p(amp), _p(omg), _p(pha), _p(dl), _p(dwt), t, dt)
            assert a == b
