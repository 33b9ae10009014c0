"""Known-answer tests that pin the oracle by mathematics, independently of the
reference build: closed forms the reference's own unit tests use
(test/test_src/*_unittests.cc) plus ones chosen so that a plausible slip
(a dropped term, a wrong sign or index, the wrong cell corner) fails.
"""
import math

import numpy as np
import pytest

from oracle_ctypes import (BTX, BTY, BTZ, BX, BY, BZ, EX, EY, EZ, JX, JY, JZ, PERIODIC, PHI1, RHO,
                           UNIT_ACC, UNIT_RHOJ, Case, Oracle, cell_ids_from_x)

RNG = np.random.default_rng(12345)


def make(case, x, v, gamma=None, species=None, q=(-1.0, 1.0), m=(1.0, 1836.0)):
    o = Oracle(case, max_particles=len(x))
    o.set_species(list(q), list(m))
    n = len(x)
    o.set_particles(x, v, np.ones(n) if gamma is None else gamma,
                    np.zeros(n, np.int32) if species is None else species, np.arange(n))
    return o


def test_cell_id_closed_forms():
    # oGrid_unittests.cc:43-70: 5x10x20 cells on a unit box, ghost offset +1
    case = Case(n=(5, 10, 20), L=(1.0, 1.0, 1.0), nparticles=1)
    o = Oracle(case)
    nx, ny, nz = 7, 12, 22
    x = y = z = 0.501
    ix, iy, iz = int(x * 5 + 1), int(y * 10 + 1), int(z * 20 + 1)
    assert o.cell_id(x, y, z) == ix * ny * nz + nz * iy + iz
    assert o.cell_id(0, 0, 0) == 1 * ny * nz + 1 * nz + 1
    assert o.cell_id(0, 0, 1) == 1 * ny * nz + 1 * nz + 21
    assert o.cell_id(0, 0, 0.5) == 1 * ny * nz + 1 * nz + 11
    assert o.cell_id(0, 1, 0.5) == 1 * ny * nz + 11 * nz + 11
    assert o.cell_id(0, 0.5, 0.5) == 1 * ny * nz + 6 * nz + 11
    assert o.cell_id(1, 0.5, 0.5) == 6 * ny * nz + 6 * nz + 11


def _gather_const(case, E, B):
    x = RNG.uniform(0, 1, (200, 3)) * np.array(case.L)
    o = make(case, x, np.zeros_like(x))
    sh = o.shape
    for c in range(3):
        o.set_field(EX + c, np.full(sh, E[c]))
        o.set_field(BX + c, np.full(sh, B[c]))
        o.set_field(BTX + c, np.full(sh, B[c]))
    o.gather()
    return o.particles().field


def test_constant_fields_gather_exactly():
    # depositJ_unittests.cc:133-164 (weights sum to one)
    case = Case(n=(5, 4, 6), L=(5.0, 1.0, 3.0), nparticles=200)
    f = _gather_const(case, (1.0, 2.0, 3.0), (-0.5, 0.25, 4.0))
    np.testing.assert_allclose(f, np.tile([1.0, 2.0, 3.0, -0.5, 0.25, 4.0], (len(f), 1)), rtol=1e-14)


def test_linear_field_is_sampled_one_cell_low():
    # Q22: with the "least vertex" at the cell's upper corner, a field linear
    # in y, z (Ex(j,k) = a + b*Y_j + c*Z_k on the x-edges) is returned as
    # a + b*(y-dy) + c*(z-dz); the lower-corner weights would give a + b*y + c*z.
    case = Case(n=(4, 5, 6), L=(4.0, 2.5, 4.5), nparticles=300)
    d = np.array(case.L) / np.array(case.n)
    x = RNG.uniform(0.05, 0.95, (300, 3)) * np.array(case.L)
    o = make(case, x, np.zeros_like(x))
    sh = o.shape
    a, b, c = 0.3, -1.7, 0.45
    J = np.arange(sh[1])[None, :, None]
    K = np.arange(sh[2])[None, None, :]
    Y = (J - 1) * d[1]
    Z = (K - 1) * d[2]
    o.set_field(EX, np.broadcast_to(a + b * Y + c * Z, sh))
    o.gather()
    ex = o.particles().field[:, 0]
    want = a + b * (x[:, 1] - d[1]) + c * (x[:, 2] - d[2])
    np.testing.assert_allclose(ex, want, rtol=0, atol=1e-12)
    assert np.abs(ex - (a + b * x[:, 1] + c * x[:, 2])).min() > 1e-3


def test_deposit_partition_of_unity_and_edge_map():
    # depositJ_unittests.cc:66-130: before the fold, sum J = UNIT_RHOJ sum(q v)/V
    # and sum rho = UNIT_RHOJ sum(q)/V over the whole padded grid
    case = Case(n=(5, 4, 3), L=(5.0, 2.0, 1.5), nparticles=500)
    V = np.prod(np.array(case.L) / np.array(case.n))
    x = RNG.uniform(0, 1, (500, 3)) * np.array(case.L)
    v = RNG.normal(0, 0.1, (500, 3))
    sp = RNG.integers(0, 2, 500).astype(np.int32)
    o = make(case, x, v, species=sp, q=(-2.0, 3.0), m=(1.0, 5.0))
    o.sort()
    o.deposit(True)
    qv = np.array([-2.0, 3.0])[o.particles().species][:, None] * o.particles().v
    for c, f in enumerate((JX, JY, JZ)):
        np.testing.assert_allclose(o.field(f).sum(), UNIT_RHOJ * qv[:, c].sum() / V, rtol=1e-12)
    q = np.array([-2.0, 3.0])[sp]
    np.testing.assert_allclose(o.field(RHO).sum(), UNIT_RHOJ * q.sum() / V, rtol=1e-12)


def test_single_particle_deposit_targets():
    # one particle: the 12 edge targets of Grid::addJ (oGrid.cpp:98-111),
    # including the duplicate Jx[ix][iy+1][iz] (Q21) and the never-written
    # Jx[ix][iy][iz+1]; weights in terms of the true in-cell fraction f are
    # (2-f_j)(2-f_k), (f_j-1)(2-f_k), (f_j-1)(f_k-1), (2-f_j)(f_k-1) (Q22)
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=1)
    x = np.array([[1.3, 2.6, 0.2]])
    v = np.array([[0.1, -0.2, 0.3]])
    o = make(case, x, v, q=(1.0, 1.0), m=(1.0, 1.0))
    o.sort()
    o.deposit(False)
    ix, iy, iz = (x[0] // 1).astype(int) + 1
    fy, fz, fx = 0.6, 0.2, 0.3
    w = lambda fj, fk: [(2 - fj) * (2 - fk), (fj - 1) * (2 - fk), (fj - 1) * (fk - 1), (2 - fj) * (fk - 1)]
    s = UNIT_RHOJ * v[0]
    Jx = o.field(JX)
    wx = w(fy, fz)
    np.testing.assert_allclose(Jx[ix, iy, iz], s[0] * wx[0], rtol=1e-12)
    np.testing.assert_allclose(Jx[ix, iy + 1, iz], s[0] * (wx[1] + wx[3]), rtol=1e-12)
    np.testing.assert_allclose(Jx[ix, iy + 1, iz + 1], s[0] * wx[2], rtol=1e-12)
    assert Jx[ix, iy, iz + 1] == 0.0
    assert np.count_nonzero(Jx) == 3
    Jy = o.field(JY)
    wy = w(fz, fx)
    np.testing.assert_allclose([Jy[ix, iy, iz], Jy[ix, iy, iz + 1], Jy[ix + 1, iy, iz + 1], Jy[ix + 1, iy, iz]],
                               s[1] * np.array(wy), rtol=1e-12)
    Jz = o.field(JZ)
    wz = w(fx, fy)
    np.testing.assert_allclose([Jz[ix, iy, iz], Jz[ix + 1, iy, iz], Jz[ix + 1, iy + 1, iz], Jz[ix, iy + 1, iz]],
                               s[2] * np.array(wz), rtol=1e-12)


def _push_const(case, fld, v0, steps, gamma=None, q=-1.0, m=1.0, dt=0.1):
    x = np.tile([[2.0, 2.0, 2.0]], (len(v0), 1))
    o = make(case, x, v0, gamma=gamma, q=(q, q), m=(m, m))
    o.set_particle_fields(np.tile(fld, (len(v0), 1)))
    out = []
    for _ in range(steps):
        o.push(dt)
        out.append(o.particles())
    return o, out


def test_boris_pure_b_rotation():
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=1)
    dt, q, m, Bz = 0.1, -1.0, 1.0, 2.0
    v0 = np.array([[0.05, 0.0, 0.01]])
    _, out = _push_const(case, [0, 0, 0, 0, 0, Bz], v0, 50, q=q, m=m, dt=dt)
    qp = 0.5 * dt * q / m * UNIT_ACC
    theta = 2 * math.atan(abs(qp * Bz))
    for n, p in enumerate(out, 1):
        vp = p.v[0]
        np.testing.assert_allclose(np.hypot(vp[0], vp[1]), 0.05, rtol=1e-14)
        assert vp[2] == 0.01
        ang = math.atan2(vp[1], vp[0]) % (2 * math.pi)
        want = (theta * n * (1 if qp * Bz < 0 else -1)) % (2 * math.pi)
        assert abs(((ang - want + math.pi) % (2 * math.pi)) - math.pi) < 1e-12


def test_boris_exb_drift():
    # ConstEBTest: E x B / B^2 drift
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=1)
    E, B = np.array([0.0, 0.001, 0.0]), np.array([0.0, 0.0, 1.0])
    _, out = _push_const(case, np.r_[E, B], np.array([[0.0, 0.0, 0.0]]), 4000, dt=0.05)
    vs = np.array([p.v[0] for p in out])
    drift = np.cross(E, B) / B.dot(B)
    np.testing.assert_allclose(vs.mean(0), drift, atol=2e-5)


def test_vay_pure_e_momentum_is_linear_in_time():
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=1, relativity=1)
    dt, q, m = 0.1, -1.0, 1.0
    E = np.array([0.3, -0.2, 0.1])
    v0 = np.array([[0.2, 0.1, -0.3]])
    g0 = 1.0 / math.sqrt(1 - (v0 ** 2).sum())
    _, out = _push_const(case, np.r_[E, 0, 0, 0], v0, 200, gamma=np.array([g0]), q=q, m=m, dt=dt)
    qp = 0.5 * dt * q / m * UNIT_ACC
    u0 = g0 * v0[0]
    for n in (1, 10, 200):
        p = out[n - 1]
        u = p.gamma[0] * p.v[0]
        np.testing.assert_allclose(u, u0 + 2 * n * qp * E, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(p.gamma[0], math.sqrt(1 + (u ** 2).sum()), rtol=1e-13)
        assert np.linalg.norm(p.v[0]) < 1.0


def test_vay_pure_b_gamma_constant_and_larmor():
    # particles_unittests.cc:34-56 (Larmor orbit radius ~1)
    case = Case(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=1, relativity=1)
    dt, q, m = 0.01, 1.0, 1.0
    v0 = np.array([[0.5, 0.0, 0.0]])
    g0 = 1.0 / math.sqrt(1 - 0.25)
    Bz = 1.0
    o, out = _push_const(case, [0, 0, 0, 0, 0, Bz], v0, 3000, gamma=np.array([g0]), q=q, m=m, dt=dt)
    gam = np.array([p.gamma[0] for p in out])
    np.testing.assert_allclose(gam, g0, rtol=1e-12)
    xy = np.array([p.x[0][:2] for p in out])
    r_larmor = g0 * 0.5 / (UNIT_ACC * abs(q / m) * Bz)
    ctr = 0.5 * (xy.max(0) + xy.min(0))
    np.testing.assert_allclose(np.linalg.norm(xy - ctr, axis=1), r_larmor, rtol=2e-3)


def test_jacobi_geometric_series_and_odd_iterate():
    # From a zero guess with a periodic eigenmode source R, sweep n gives
    # u_n = 4 pi af R sum_{m<n} lam^m, lam = 2(ax cos kx dx + ay cos ky dy + az cos kz dz);
    # phi1 holds the last odd-numbered iterate (Q9).
    case = Case(n=(8, 6, 4), L=(8.0, 3.0, 4.0), nparticles=1, poisson=True)
    o = Oracle(case)
    sh = o.shape
    d = np.array(case.L) / np.array(case.n)
    I = np.arange(sh[0])[:, None, None]
    Jj = np.arange(sh[1])[None, :, None]
    K = np.arange(sh[2])[None, None, :]
    kx, ky, kz = 2 * np.pi / case.L[0], 2 * np.pi / case.L[1], 0.0
    R = np.cos(kx * (I - 1) * d[0]) * np.cos(ky * (Jj - 1) * d[1]) * np.ones_like(K, dtype=float)
    o.set_field(RHO, R)
    for n_sweeps in (1, 2, 7, 8):
        sw = o.poisson_init(iter_max=n_sweeps - 1, tol=0.0)
        assert sw[0] == n_sweeps
        dx2, dy2, dz2 = d ** 2
        cd2 = dx2 * dy2 + dy2 * dz2 + dx2 * dz2
        ax, ay, az = dy2 * dz2 / (2 * cd2), dx2 * dz2 / (2 * cd2), dx2 * dy2 / (2 * cd2)
        af = dx2 * dy2 * dz2 / (2 * cd2)
        lam = 2 * (ax * math.cos(kx * d[0]) + ay * math.cos(ky * d[1]) + az * math.cos(kz * d[2]))
        n_eff = n_sweeps if n_sweeps % 2 == 0 else n_sweeps  # phi1 = u_n for even n ...
        if n_sweeps % 2 == 1:
            n_eff = n_sweeps - 1  # ... and u_{n-1} for odd n (last odd iternum)
        series = sum(lam ** m for m in range(n_eff))
        want = 4 * math.pi * af * R * series
        phys = (slice(1, -2), slice(1, -2), slice(1, -2))
        np.testing.assert_allclose(o.field(PHI1)[phys], want[phys], rtol=1e-12, atol=1e-14)


def test_periodic_ghost_rows_equal_opposite_physical_rows():
    # spookyGrid_unittests.cc:95-284: after the refresh, ghost planes hold the
    # opposite physical planes over the physical range; physical rows untouched
    case = Case(n=(4, 5, 3), L=(4.0, 5.0, 3.0), nparticles=1)
    o = Oracle(case)
    sh = o.shape
    before = {}
    for f in (EX, EY, EZ, BX, BY, BZ):
        a = RNG.normal(size=sh)
        o.set_field(f, a)
        before[f] = a
    o.refresh(0.0, 0.1)
    n = case.n
    for f in (EX, EY, EZ, BX, BY, BZ):
        a = o.field(f)
        np.testing.assert_array_equal(a[1:-2, 1:-2, 1:-2], before[f][1:-2, 1:-2, 1:-2])
        np.testing.assert_array_equal(a[0, 1:n[1] + 1, 1:n[2] + 1], a[n[0], 1:n[1] + 1, 1:n[2] + 1])
        np.testing.assert_array_equal(a[n[0] + 1, 1:n[1] + 1, 1:n[2] + 1], a[1, 1:n[1] + 1, 1:n[2] + 1])
        np.testing.assert_array_equal(a[1:n[0] + 1, 0, 1:n[2] + 1], a[1:n[0] + 1, n[1], 1:n[2] + 1])
        np.testing.assert_array_equal(a[1:n[0] + 1, 1:n[1] + 1, n[2] + 1], a[1:n[0] + 1, 1:n[1] + 1, 1])
        # edge/corner ghosts are never refreshed (Q6)
        assert a[0, 0, 1] == before[f][0, 0, 1]


def test_fold_conserves_the_folded_sum():
    # the fold adds the upper ghost plane of (J_perp, rho) into plane 1 over
    # the physical range; the physical-range sum gains exactly that plane
    case = Case(n=(4, 5, 3), L=(4.0, 5.0, 3.0), nparticles=1)
    o = Oracle(case)
    sh = o.shape
    a = RNG.normal(size=sh)
    o.set_field(RHO, a)
    o.fold()
    n = case.n
    b = o.field(RHO)
    # x fold: plane n+1 (y,z physical) added to plane 1
    want = a.copy()
    want[1, 1:n[1] + 1, 1:n[2] + 1] += a[n[0] + 1, 1:n[1] + 1, 1:n[2] + 1]
    want[1:n[0] + 1, 1, 1:n[2] + 1] += want[1:n[0] + 1, n[1] + 1, 1:n[2] + 1]
    want[1:n[0] + 1, 1:n[1] + 1, 1] += want[1:n[0] + 1, 1:n[1] + 1, n[2] + 1]
    np.testing.assert_allclose(b, want, rtol=1e-15, atol=0)


def test_poisson_residual_below_tolerance_at_convergence():
    # depositJ_unittests.cc:167-219: after convergence the discrete Poisson
    # residual is below 1e-9 at every physical point (rho = +-1 at two points)
    case = Case(n=(6, 6, 6), L=(6.0, 6.0, 6.0), nparticles=1, poisson=True,
                bound_phi=(0.0, 1e-1, 0.0, 0.0, 0.0, 0.0))
    o = Oracle(case)
    sh = o.shape
    R = np.zeros(sh)
    R[2, 2, 2], R[4, 4, 4] = 1.0, -1.0
    o.set_field(RHO, R)
    sw = o.poisson_init(iter_max=10000, tol=1e-13)
    assert sw[0] < 10001
    u = o.field(PHI1)
    c = u[1:-2, 1:-2, 1:-2]
    lap = (u[:-3, 1:-2, 1:-2] + u[2:-1, 1:-2, 1:-2] + u[1:-2, :-3, 1:-2] + u[1:-2, 2:-1, 1:-2]
           + u[1:-2, 1:-2, :-3] + u[1:-2, 1:-2, 2:-1] - 6 * c)
    res = lap + 4 * math.pi * R[1:-2, 1:-2, 1:-2]
    assert np.abs(res).max() < 1e-9
