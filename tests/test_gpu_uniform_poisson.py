"""The device loader (emo_load_uniform), the neutralising background of the
Poisson initialisation, and the FP32 chosen form's Poisson initialisation
(FP32 Jacobi sweeps with FP64 residual norms) against the FP64 plain form,
which is bit-exact with the oracle (test_gpu_fp64_parity.py).  All calls go
through the C ABI."""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import (BC_PERIODIC, EX, EY, EZ, BX, BY, BZ, FP32_CHOSEN, FP64_PLAIN, JX, RHO,
                                      Context, EmoConfig, EmoLoader, lib)

pytestmark = pytest.mark.gpu

UNIT_VTH = 0.002232
WPE2_PER_DENS = 4 * math.pi * 4.80320427e-13 * 0.58668774


def make(n, precision, species, ppc, relativity=1, neutralize=0, seed=-7):
    """species: (mass_ratio, charge_ratio, sigma in c, drift_x) per species,
    ppc particles per cell and species"""
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = n[d]
        c.Lxyz[d] = float(n[d])
    c.relativity = relativity
    c.precision = precision
    for s in range(6):
        c.fields_bound[s] = BC_PERIODIC
        c.parts_bound[s] = BC_PERIODIC
    c.neutralize = neutralize
    ld = EmoLoader()
    ld.nspecies = len(species)
    ld.nparticles_tot = ppc * len(species) * n[0] * n[1] * n[2]
    for i, (m, q, sigma, drift) in enumerate(species):
        ld.mass_ratio[i] = m
        ld.charge_ratio[i] = q
        ld.dens_frac[i] = 1.0
        ld.temp[i] = (sigma / UNIT_VTH) ** 2 * m
        ld.drift[i][0] = drift
    ld.dens_phys = 0.01 / WPE2_PER_DENS
    ld.seed = seed
    ld.zero_avg = 0
    ld.poisson_init = 1
    c.max_particles = ld.nparticles_tot
    ctx = Context(c)
    ctx.load_uniform(ld)
    return ctx, ld


def by_id(ctx):
    p = ctx.particles(fields=False)
    o = np.argsort(p["id"])
    return {k: v[o] for k, v in p.items() if v is not None}


BEAMS = [(1.0, -1.0, 0.01, 0.2), (1.0, -1.0, 0.01, -0.2)]
PLASMA = [(1.0, -1.0, 0.05, 0.0), (1836.0, 1.0, 0.05 / math.sqrt(1836.0), 0.0)]


def test_uniform_loader_counts_moments_and_both_forms():
    n = (8, 6, 4)
    ppc = 64
    c64, _ = make(n, FP64_PLAIN, BEAMS, ppc)
    c32, _ = make(n, FP32_CHOSEN, BEAMS, ppc)
    p64, p32 = by_id(c64), by_id(c32)
    assert np.array_equal(p64["id"], np.arange(2 * ppc * 8 * 6 * 4))
    assert np.array_equal(p64["species"], p32["species"])
    # every physical cell holds ppc particles of each species
    cells = c64.cell_ids()
    nc = (n[0] + 2, n[1] + 2, n[2] + 2)
    ix, rem = np.divmod(cells, nc[1] * nc[2])
    iy, iz = np.divmod(rem, nc[2])
    assert ix.min() >= 1 and ix.max() <= n[0] and iy.min() >= 1 and iy.max() <= n[1] and iz.max() <= n[2]
    counts = np.bincount(cells, minlength=nc[0] * nc[1] * nc[2])
    assert set(np.unique(counts[counts > 0])) == {2 * ppc}
    # the FP32 store holds the same particles up to its offset / momentum rounding
    assert np.max(np.abs(p32["x"] - p64["x"])) < 1e-6
    assert np.max(np.abs(p32["v"] - p64["v"])) < 1e-7
    # moments: drift +-0.2 along x, sigma 0.01 per component (4096 particles per beam)
    for s, (m, q, sigma, drift) in enumerate(BEAMS):
        v = p64["v"][p64["species"] == s]
        assert abs(v[:, 0].mean() - drift) < 5 * sigma / math.sqrt(len(v))
        assert abs(v[:, 1].mean()) < 5 * sigma / math.sqrt(len(v))
        assert abs(v.std(axis=0) / sigma - 1).max() < 0.05
    # the stored gamma is the reference loader's sqrt(1 - v^2)
    assert np.allclose(p64["gamma"], np.sqrt(1 - (p64["v"] ** 2).sum(axis=1)), rtol=0, atol=1e-15)
    # deterministic per seed; another seed gives other particles
    c64b, _ = make(n, FP64_PLAIN, BEAMS, ppc)
    assert np.array_equal(by_id(c64b)["x"], p64["x"])
    c64c, _ = make(n, FP64_PLAIN, BEAMS, ppc, seed=-8)
    assert not np.array_equal(by_id(c64c)["x"], p64["x"])
    for c in (c64, c32, c64b, c64c):
        c.close()


def test_uniform_loader_rejects_fractional_cells():
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = 4
        c.Lxyz[d] = 4.0
    c.precision = FP32_CHOSEN
    c.max_particles = 1000
    ld = EmoLoader()
    ld.nspecies = 1
    ld.nparticles_tot = 100  # not a multiple of 64 cells
    ld.mass_ratio[0] = 1.0
    ld.charge_ratio[0] = -1.0
    ld.dens_frac[0] = 1.0
    ctx = Context(c)
    with pytest.raises(Exception, match="whole number"):
        ctx.load_uniform(ld)
    ctx.close()


def inner(n):
    return (slice(1, n[0] + 1), slice(1, n[1] + 1), slice(1, n[2] + 1))


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("species,neutralize", [(PLASMA, 0), (BEAMS, 1)])
def test_fp32_poisson_init_tracks_fp64(species, neutralize):
    """Same particles, 50 + 1 sweeps per solve (the reference's rule with
    iter_max 50): E from phi and B from AToB of the FP32 form agree with the
    FP64 form to 1e-4 relative L2; the sweep counts agree; the residual norms
    are FP64 values of the FP32 iterates."""
    n = (8, 8, 8)
    c64, _ = make(n, FP64_PLAIN, species, 16, neutralize=neutralize)
    c32, _ = make(n, FP32_CHOSEN, species, 16, neutralize=neutralize)
    sw64, res64 = c64.init_poisson(iter_max=50, tol=-1.0)
    sw32, res32 = c32.init_poisson(iter_max=50, tol=-1.0)
    assert list(sw64) == list(sw32) == [51, 51, 51, 51]
    I = inner(n)
    e64 = np.concatenate([c64.field(f)[I].ravel() for f in (EX, EY, EZ)])
    e32 = np.concatenate([c32.field(f)[I].ravel() for f in (EX, EY, EZ)])
    assert np.linalg.norm(e64) > 0
    assert rel(e32, e64) < 1e-4
    b64 = np.concatenate([c64.field(f)[I].ravel() for f in (BX, BY, BZ)])
    b32 = np.concatenate([c32.field(f)[I].ravel() for f in (BX, BY, BZ)])
    assert rel(b32, b64) < 1e-4
    assert rel(c32.field(JX)[I], c64.field(JX)[I]) < 1e-5
    assert np.all(res32[::2] > 0) and abs(res32[0] / res64[0] - 1) < 1e-2
    for c in (c64, c32):
        c.close()


def test_neutralize_removes_the_mean_charge():
    """A one-sign plasma between periodic boundaries: with the neutralising
    background the Poisson source has zero mean, in both forms."""
    n = (8, 8, 8)
    for prec in (FP64_PLAIN, FP32_CHOSEN):
        ctx, _ = make(n, prec, BEAMS, 16, neutralize=1)
        ctx.init_poisson(iter_max=20, tol=-1.0)
        rho = ctx.field(RHO)[inner(n)]
        scale = np.abs(rho).max()
        assert scale > 0
        assert abs(rho.mean()) < (1e-12 if prec == FP64_PLAIN else 1e-6) * scale
        ctx.close()


def test_fp32_poisson_then_steps_is_deterministic():
    n = (8, 8, 8)
    out = []
    for _ in range(2):
        ctx, _ = make(n, FP32_CHOSEN, BEAMS, 16, neutralize=1)
        ctx.init_poisson(iter_max=30, tol=-1.0)
        ctx.step(10, 0.0, 0.5 / math.sqrt(3.0))
        out.append((ctx.field(EX), by_id(ctx)["x"]))
        ctx.close()
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
