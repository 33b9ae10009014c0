"""Pins of the oracle's three extension functions -- the displacement loader
(a density perturbation), the drift velocity and the transverse Gaussian
waist of an injected pulse -- none of which the reference has, so no
reference run can pin them.  Each check follows from the mathematics of the
stated convention and fails on a plausible slip:

* displacement x = x_u - (eps/k) sin(k (x_u - x0)):
  - the inverse map x + (eps/k) sin(k (x_u - x0)) = x_u (mod L) holds to
    roundoff per particle (a flipped sign or a missing x0 fails);
  - the loaded density is n0 (1 + eps cos(k x) + O(eps^2)): the particle mean
    of cos(k x) - cos(k x_u) is the Bessel value J1(eps) (closed form of the
    uniform average of cos(t - eps sin t)), so the density peaks at x = x0;
  - the sensitivity emo_loader_dx_deps is the derivative of that map,
    -sin(k (x_u - x0))/k, also at eps = 0, and agrees with central
    differences of the oracle's loader;
* drift: v = v_thermal + drift exactly (one rounding), gamma = sqrt(1 - v^2)
  of the drifted v, species without drift unchanged bit for bit, and the
  sample mean within shot noise of the drift;
* waist: the injected wave part is multiplied by exp(-r^2 / w^2), r the
  transverse distance of each staggered point from the box centre: exactly
  1 at a point on the axis, e^-1 at distance w, e^-2 on the diagonal, for
  both polarisations' staggering (E_i offset d_i/2 along i; B_i offset d/2
  along the two other axes).
"""
import math

import numpy as np
import pytest
from scipy.special import j1

from oracle_ctypes import (BX, BY, BZ, EX, EY, EZ, EXTERNAL, PERIODIC, REFLECTING, Case, Oracle,
                           Wave)

K_BOX = 16.0


def _load(case):
    o = Oracle(case)
    assert o.load() == 0
    return o.particles()


def _pert_case(eps, k=2 * math.pi / K_BOX, x0=-3.0, n=1 << 17, drift=()):
    return Case(n=(16, 2, 2), x0=(x0, 0.5, 0.25), L=(K_BOX, 2.0, 2.0), nparticles=n,
                temp=(500.0, 500.0), zero_avg=0, seed=-7,
                pert_eps=(eps, 0.0), pert_k=(k, 0.0), drift=drift)


def test_displacement_inverse_map_and_species_split():
    k, eps, x0 = 2 * math.pi / K_BOX, 0.2, -3.0
    pu = _load(_pert_case(0.0))
    pd = _load(_pert_case(eps))
    e = pu.species == 0
    assert e.sum() == len(e) // 2
    # the displacement draws no random numbers: everything else is unchanged
    np.testing.assert_array_equal(pd.v, pu.v)
    np.testing.assert_array_equal(pd.x[:, 1:], pu.x[:, 1:])
    np.testing.assert_array_equal(pd.x[~e], pu.x[~e])
    xu = pu.x[e, 0] - x0
    x = pd.x[e, 0] - x0
    back = np.mod(x + (eps / k) * np.sin(k * xu), K_BOX)
    d = np.abs(back - xu)
    d = np.minimum(d, K_BOX - d)
    assert d.max() < 1e-13
    assert np.all((x >= 0) & (x < K_BOX))
    # the displaced particles are not where a plus sign would put them
    assert np.abs(np.mod(xu + (eps / k) * np.sin(k * xu), K_BOX) - x).max() > 0.1 / k


def test_displacement_density_mode_peaks_at_x0():
    k, eps, x0 = 2 * math.pi / K_BOX, 0.2, -3.0
    pu = _load(_pert_case(0.0))
    pd = _load(_pert_case(eps))
    e = pu.species == 0
    xu = pu.x[e, 0] - x0
    x = pd.x[e, 0] - x0
    n = e.sum()
    # <cos(k x) - cos(k x_u)> -> J1(eps); per-particle spread ~ eps/sqrt(8)
    got = np.mean(np.cos(k * x) - np.cos(k * xu))
    sigma = eps / math.sqrt(8 * n)
    assert abs(got - j1(eps)) < 6 * sigma, (got, j1(eps), sigma)
    assert got > 0  # a density peak at x0, not a trough
    # the sine mode stays zero (the peak is at x0, not shifted)
    gs = np.mean(np.sin(k * x) - np.sin(k * xu))
    assert abs(gs) < 6 * sigma


def test_displacement_sensitivity_at_and_away_from_zero():
    from apc_524_project_amd._lib import loader_dx_deps
    from emo_case import emo_config, emo_loader

    k, x0 = 2 * math.pi / K_BOX, -3.0
    n = 1 << 12
    pu = _load(_pert_case(0.0, n=n)).by_id()
    e = pu.species == 0
    want = np.where(e, -np.sin(k * (pu.x[:, 0] - x0)) / k, 0.0)
    c0 = _pert_case(0.0, n=n)
    d0 = loader_dx_deps(emo_loader(c0), emo_config(c0))
    np.testing.assert_allclose(d0, want, rtol=0, atol=1e-14)
    assert np.abs(d0[e]).max() > 0.9 / k
    c1 = _pert_case(1e-12, n=n)
    np.testing.assert_array_equal(loader_dx_deps(emo_loader(c1), emo_config(c1)), d0)
    # central differences of the oracle's loader about eps = 0.05
    h, eps = 1e-6, 0.05
    xp = _load(_pert_case(eps + h, n=n)).by_id().x[:, 0]
    xm = _load(_pert_case(eps - h, n=n)).by_id().x[:, 0]
    fd = (xp - xm) / (2 * h)
    ce = _pert_case(eps, n=n)
    de = loader_dx_deps(emo_loader(ce), emo_config(ce))
    inner = np.abs(fd) < 10.0  # particles wrapped across x0 or x0 + L under the perturbation
    assert inner.sum() > n - 8
    np.testing.assert_allclose(fd[inner], de[inner], rtol=0, atol=1e-7)


def test_displacement_without_wavenumber_is_rejected():
    from apc_524_project_amd._lib import EmoError, loader_dx_deps
    from emo_case import emo_config, emo_loader

    c = _pert_case(0.01, k=0.0, n=64)
    with pytest.raises(EmoError):
        loader_dx_deps(emo_loader(c), emo_config(c))


def test_drift_is_added_once_per_component():
    drift = ((0.1, -0.05, 0.02), (0.0, 0.0, 0.0))
    p0 = _load(_pert_case(0.0))
    pd = _load(_pert_case(0.0, drift=drift))
    e = p0.species == 0
    np.testing.assert_array_equal(pd.x, p0.x)
    np.testing.assert_array_equal(pd.v[~e], p0.v[~e])
    np.testing.assert_array_equal(pd.gamma[~e], p0.gamma[~e])
    dv = pd.v[e] - p0.v[e]
    ulp = np.spacing(np.abs(pd.v[e]))
    assert np.all(np.abs(dv - np.array(drift[0])) <= ulp)
    # Q3's stored gamma is that of the drifted velocity
    np.testing.assert_allclose(pd.gamma[e], np.sqrt(1 - (pd.v[e] ** 2).sum(axis=1)), rtol=1e-15)
    # sample mean within shot noise of the drift
    sig = p0.v[e].std(axis=0)
    mean = pd.v[e].mean(axis=0)
    assert np.all(np.abs(mean - np.array(drift[0])) < 6 * sig / math.sqrt(e.sum()))


# staggering offsets in units of the cell, per field (E_i: 1/2 along i;
# B_i: 1/2 along the other two axes)
def _offset(fid, axis):
    comp = fid % 3
    return (0.5 if axis == comp else 0.0) if fid < BX else (0.0 if axis == comp else 0.5)


@pytest.mark.parametrize("pol,efid,bfid,n", [(2, EY, BZ, (4, 5, 4)), (3, EZ, BY, (4, 4, 5))])
def test_waist_profile_on_staggered_planes(pol, efid, bfid, n):
    w = 1.0
    E0, B0 = (0.1, 0.2, 0.3), (0.05, -0.1, 0.15)
    wave = Wave(side=-1, pol=pol, amp=0.7, omega=0.39, phase=30.0, inv_width=0.08, delay=3.0)

    def planes(waist):
        case = Case(n=n, L=tuple(float(v) for v in n), fbc=(EXTERNAL, EXTERNAL) + (PERIODIC,) * 4,
                    pbc=(REFLECTING, REFLECTING) + (PERIODIC,) * 4, nparticles=1, E0=E0, B0=B0,
                    waves=[wave], waist=waist)
        o = Oracle(case)
        o.const_init()
        o.refresh(3.1, 0.25)
        return o.field(efid)[1], o.field(bfid)[1]  # the x- injection plane (ix = 1)

    flat = planes(0.0)
    prof = planes(w)
    yc, zc = n[1] / 2.0, n[2] / 2.0
    for fid, f0, fw, base in ((efid, flat[0], prof[0], E0[efid]), (bfid, flat[1], prof[1], B0[bfid - BX])):
        wave0 = f0 - base
        assert np.abs(wave0[1:-2, 1:-2]).min() > 1e-3  # the pulse is on
        ratio = {}
        for j in range(1, n[1] + 1):
            for kk in range(1, n[2] + 1):
                y = (j - 1) + _offset(fid, 1)
                z = (kk - 1) + _offset(fid, 2)
                r2 = (y - yc) ** 2 + (z - zc) ** 2
                ratio[(y - yc, z - zc)] = (fw[j, kk] - base) / wave0[j, kk]
                np.testing.assert_allclose(fw[j, kk] - base, wave0[j, kk] * math.exp(-r2 / (w * w)),
                                           rtol=1e-12, atol=1e-15)
        # the closed-form points: on the axis, at distance w, on the diagonal
        assert ratio[(0.0, 0.0)] == 1.0
        for dyz in ((w, 0.0), (-w, 0.0), (0.0, w), (0.0, -w)):
            assert abs(ratio[dyz] - math.exp(-1.0)) < 1e-14
        assert abs(ratio[(w, w)] - math.exp(-2.0)) < 1e-14
