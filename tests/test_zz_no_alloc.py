"""emo_init_poisson_noalloc (the static-workspace variant of
emo_init_poisson) against emo_init_poisson, through the C ABI, in both
precision forms: identical particles, identical fields (bit for bit: the
same kernels in the same order on a workspace of the same layout),
identical sweep counts and residual norms.  Two calls in a row on the GPU's
one static workspace, the second on a smaller grid, so that the second
starts from the first call's leftover contents; and a grid whose workspace
exceeds the static array is refused before any work."""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import (BC_PERIODIC, ERR_CONFIG, FP32_CHOSEN, FP64_PLAIN, Context, EmoConfig,
                                      EmoError, EmoLoader)

pytestmark = pytest.mark.gpu

UNIT_VTH = 0.002232


def make(n, precision, ppc=8, seed=-1):
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = n[d]
        c.Lxyz[d] = float(n[d])
    c.relativity = 0
    c.precision = precision
    for s in range(6):
        c.fields_bound[s] = BC_PERIODIC
        c.parts_bound[s] = BC_PERIODIC
    ld = EmoLoader()
    ld.nspecies = 2
    ld.nparticles_tot = 2 * ppc * n[0] * n[1] * n[2]
    for i, (m, q) in enumerate(((1.0, -1.0), (1836.0, 1.0))):
        ld.mass_ratio[i] = m
        ld.charge_ratio[i] = q
        ld.dens_frac[i] = 0.5
        ld.temp[i] = 501.8
    ld.dens_phys = 2.824e9
    ld.seed = seed
    ld.zero_avg = -1
    ld.poisson_init = 1
    c.max_particles = ld.nparticles_tot
    ctx = Context(c)
    ctx.load_maxwellian(ld)
    return ctx


def compare(n, precision, iter_max):
    a = make(n, precision)
    b = make(n, precision)
    sa, ra = a.init_poisson(iter_max=iter_max, tol=-1.0)
    sb, rb = b.init_poisson_noalloc(iter_max=iter_max, tol=-1.0)
    print(f"no-alloc Poisson: n={n} precision={precision} sweeps={list(sb)}", flush=True)
    assert np.array_equal(sa, sb)
    assert np.array_equal(ra, rb)
    for f in range(13):
        assert np.array_equal(a.field(f), b.field(f)), f
    pa, pb = a.particles(fields=False), b.particles(fields=False)
    for k in ("x", "v", "id"):
        assert np.array_equal(pa[k], pb[k]), k
    a.close()
    b.close()


@pytest.mark.parametrize("precision", [FP64_PLAIN, FP32_CHOSEN])
def test_noalloc_poisson_matches_allocating_entry_point(precision):
    compare((16, 16, 16), precision, 50)
    compare((8, 6, 4), precision, 20)


def test_noalloc_poisson_refuses_a_grid_beyond_the_workspace():
    n = (100, 100, 100)  # FP64: 8 x 103^3 doubles = 70 MB > 64 MiB
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = n[d]
        c.Lxyz[d] = float(n[d])
    c.precision = FP64_PLAIN
    c.max_particles = 1
    ctx = Context(c)
    with pytest.raises(EmoError) as e:
        ctx.init_poisson_noalloc(iter_max=5)
    assert e.value.code == ERR_CONFIG
    assert "workspace" in ctx.L.emo_last_error(ctx.h).decode()
    ctx.close()
