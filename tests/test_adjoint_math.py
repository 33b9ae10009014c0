"""Reverse-mode derivatives of the per-particle arithmetic (interpolation,
Boris and relativistic pushes, current deposit) against central finite
differences, on the host build of the same headers the GPU kernels use.

Each check draws random inputs and a random output adjoint w and compares
<w, f(x + h e_i) - f(x - h e_i)> / 2h with the i-th entry of the vector-
Jacobian product, for every input direction e_i.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "native", "adjoint_math_check.cpp")


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("am") / "libam.so")
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-shared", "-fPIC", "-std=c++17", SRC, "-o", so],
                   check=True)
    L = C.CDLL(so)
    return L


def p(a):
    return a.ctypes.data_as(C.c_void_p)


def fd_check(f, vjp, x, wout, h=1e-6, rtol=1e-6):
    """f(x)->y; vjp(x, wout)->g; compares g with central differences."""
    g = vjp(x, wout)
    fd = np.zeros_like(x)
    for i in range(x.size):
        xp = x.copy()
        xm = x.copy()
        xp.flat[i] += h
        xm.flat[i] -= h
        fd.flat[i] = (np.dot(wout, f(xp)) - np.dot(wout, f(xm))) / (2 * h)
    scale = max(np.abs(fd).max(), 1e-30)
    assert np.abs(g - fd).max() / scale < rtol, (g, fd)


@pytest.mark.parametrize("seed", range(5))
def test_interp_vjp(lib, seed):
    rng = np.random.default_rng(seed)
    d = rng.uniform(0.5, 2.0, 3)
    cv = rng.normal(size=21)
    cv[:3] = rng.uniform(1, 5, 3)
    pos = cv[:3] + rng.uniform(-1, 0, 3) * d  # relpos in [-d, 0): the Q22 vertex convention
    w = rng.normal(size=6)

    def f(x):
        out = np.zeros(6)
        c = cv.copy()
        c[3:] = x[3:]
        lib.am_interp(p(d), p(c), p(np.ascontiguousarray(x[:3])), p(out))
        return out

    def vjp(x, wout):
        posb = np.zeros(3)
        cvb = np.zeros(21)
        c = cv.copy()
        c[3:] = x[3:]
        lib.am_interp_vjp(p(d), p(c), p(np.ascontiguousarray(x[:3])), p(wout), p(posb), p(cvb))
        return np.concatenate([posb, cvb[3:]])

    x = np.concatenate([pos, cv[3:]])
    fd_check(f, vjp, x, w)


@pytest.mark.parametrize("rel", [0, 1])
@pytest.mark.parametrize("seed", range(5))
def test_push_vjp(lib, rel, seed):
    rng = np.random.default_rng(100 + seed)
    v = rng.normal(scale=0.2, size=3)
    gam = 1.0 / np.sqrt(1 - v @ v) if rel else 1.0
    state = np.concatenate([rng.uniform(1, 5, 3), v, [gam]])
    fld = rng.normal(scale=0.3, size=6)
    q, m, dt = (-1.0, 1.0, 0.3) if seed % 2 == 0 else (2.5e3, 1836 * 2.5e3, 0.29)
    w = rng.normal(size=7)
    if not rel:
        w[6] = 0.0  # Boris leaves gamma untouched

    def f(x):
        st = np.ascontiguousarray(x[:7]).copy()
        lib.am_push(rel, p(st), p(np.ascontiguousarray(x[7:])), C.c_double(q), C.c_double(m), C.c_double(dt))
        return st

    def vjp(x, wout):
        sb = np.zeros(7)
        fb = np.zeros(6)
        lib.am_push_vjp(rel, p(np.ascontiguousarray(x[:7])), p(np.ascontiguousarray(x[7:])), C.c_double(q),
                        C.c_double(m), C.c_double(dt), p(np.ascontiguousarray(wout)), p(sb), p(fb))
        if not rel:
            sb[6] = 0.0
        return np.concatenate([sb, fb])

    x = np.concatenate([state, fld])
    if not rel:
        # gamma is neither read nor written by Boris: compare the other inputs
        gfun = vjp(x, w)
        assert gfun[6] == 0.0
    fd_check(f, vjp, x, w, rtol=2e-6)


@pytest.mark.parametrize("seed", range(5))
def test_deposit_vjp(lib, seed):
    rng = np.random.default_rng(200 + seed)
    d = rng.uniform(0.5, 2.0, 3)
    cv3 = rng.uniform(1, 5, 3)
    pos = cv3 + rng.uniform(-1, 0, 3) * d
    vel = rng.normal(scale=0.2, size=3)
    q = rng.uniform(-3e3, 3e3)
    w = rng.normal(size=12)

    def f(x):
        r = np.zeros(12)
        lib.am_deposit(p(d), p(cv3), p(np.ascontiguousarray(x[:3])), p(np.ascontiguousarray(x[3:])),
                       C.c_double(q), p(r))
        return r

    def vjp(x, wout):
        pb = np.zeros(3)
        vb = np.zeros(3)
        lib.am_deposit_vjp(p(d), p(cv3), p(np.ascontiguousarray(x[:3])), p(np.ascontiguousarray(x[3:])),
                           C.c_double(q), p(wout), p(pb), p(vb))
        return np.concatenate([pb, vb])

    fd_check(f, vjp, np.concatenate([pos, vel]), w, h=1e-7)
