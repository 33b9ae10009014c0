"""The wave-injection configurations on their own workloads, through the C
ABI:

- c3 at its full size (128x64x64 cells; x: external fields with a y-polarised
  Gaussian pulse from x- of a0 = 0.5, wavelength 16 cells and waist 10
  cells, reflecting particle walls; y, z periodic; 16 particles per cell and
  species, n = 0.1 n_crit, relativistic; 16,777,216 particles): the FP64
  plain form bit-exact against the oracle over 3 steps, and the FP32 chosen
  form within the tolerances written below of the FP64 form over 10 steps;
- c4 at its full size (64x32x32 cells, 8 particles per cell and species,
  1,048,576 particles, the electron density perturbation 0.01 cos(2 pi x/64)
  from the displacement loader): the FP64 plain form bit-exact against the
  oracle over 20 steps; the gradient of the final field energy with respect
  to the pulse amplitude, the pulse phase and the perturbation amplitude
  against central differences of the same forward run.  The configuration's
  step h = 1e-4 moves particles across cell faces, where the derivative of
  the reference's weights jumps (the adjoint is the derivative between such
  crossings), so at h = 1e-4 the difference is measured and printed, not
  bounded.  With a step small enough that no particle changes cell (1e-9
  relative for the amplitude, 1e-9 rad for the phase, 1e-10 for eps) the
  agreement is within 1e-4 plus the resolution of the difference quotient:
  four units in the last place of F divided by 2h.
"""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import FP32_CHOSEN, Context, loader_dx_deps
from apc_524_project_amd.workloads import UNIT_ACC, WPE2_PER_DENS, make_workload, temp_for_sigma
from emo_case import emo_config, emo_loader
from oracle_ctypes import (BX, BY, BZ, EX, EY, EZ, EXTERNAL, JX, JY, JZ, PERIODIC, REFLECTING, Case, Oracle,
                           Wave)
from test_gpu_fp64_parity import FIELDS, assert_fields_equal, assert_particles_equal, by_id, pair

pytestmark = pytest.mark.gpu


def wave_case(name, amp_scale=1.0, phase=0.0, eps=None):
    """c3 or c4 in the oracle's vocabulary (the workload definitions of
    apc_524_project_amd.workloads)."""
    n = (128, 64, 64) if name == "c3" else (64, 32, 32)
    ppc = 16 if name == "c3" else 8
    omega = 2 * math.pi / 16.0
    w = Wave(side=-1, pol=2, amp=amp_scale * 0.5 * omega / UNIT_ACC, omega=omega, phase=phase,
             inv_width=1.0 / 40.0, delay=60.0)
    t = temp_for_sigma(0.01, 1.0)
    kw = {}
    if name == "c4":
        kw = dict(pert_eps=(0.01 if eps is None else eps, 0.0), pert_k=(2 * math.pi / n[0], 0.0))
    return Case(n=n, L=tuple(float(v) for v in n), nparticles=2 * ppc * n[0] * n[1] * n[2], temp=(t, t),
                dens_phys=0.1 * omega * omega / WPE2_PER_DENS, relativity=1, waves=[w], waist=10.0,
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                seed=-3 if name == "c3" else -4, zero_avg=0, **kw)


def test_wave_case_matches_the_workload_definition():
    """The oracle-side description above and the benchmark's workload are the
    same run (geometry, pulse, species, loader)."""
    for name in ("c3", "c4"):
        w = make_workload(name)
        cfg = emo_config(wave_case(name))
        for f in ("nCell", "Lxyz", "fields_bound", "parts_bound"):
            assert list(getattr(cfg, f)) == list(getattr(w.cfg, f)), (name, f)
        for f in ("peakamps", "omegas", "phases", "invWidths", "delays", "inSide", "inPolE"):
            assert getattr(cfg, f)[0] == getattr(w.cfg, f)[0], (name, f)
        assert cfg.waist == w.cfg.waist and cfg.relativity == w.cfg.relativity
        ld = emo_loader(wave_case(name))
        for f in ("nparticles_tot", "seed", "dens_phys"):
            assert getattr(ld, f) == getattr(w.loader, f), (name, f)
        assert abs(ld.temp[0] - w.loader.temp[0]) <= 1e-12 * w.loader.temp[0]
        assert ld.pert_eps[0] == w.loader.pert_eps[0] and ld.pert_k[0] == w.loader.pert_k[0]


def rel(a, b):
    n = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (n if n > 0 else 1.0))


def test_c3_full_size_fp64_bit_exact_and_fp32_tolerances():
    case = wave_case("c3")
    ctx, o = pair(case)
    print(f"c3 full size: {case.nparticles} particles", flush=True)
    ctx.init_constant()
    o.initialize()
    dt = case.dt_cfl()
    t = 0.0
    for _ in range(3):
        o.step(t, dt)
        t += dt
    ctx.step(3, 0.0, dt)
    assert_particles_equal(ctx, o, fields=False)
    assert_fields_equal(ctx, o)
    # the FP32 chosen form from the same loaded particles, 10 steps, against
    # the FP64 form (bit-exact with the oracle)
    p0 = Oracle(case)
    assert p0.load() == 0
    p = p0.particles()
    qabs = float(np.abs(p0.q[p.species]).sum())  # the scale of the (cancelling) total charge
    c64 = Context(emo_config(case, q=p0.q, m=p0.m))
    c32 = Context(emo_config(case, precision=FP32_CHOSEN, q=p0.q, m=p0.m))
    for c in (c64, c32):
        c.load_particles(p.x, p.v, p.gamma, p.species, p.id)
        c.init_constant()
    del p, p0
    c64.interpolate()
    c64.sort()
    c64.step(10, 0.0, dt)
    c32.step(10, 0.0, dt)
    n = case.n
    inner = (slice(1, n[0] + 1), slice(1, n[1] + 1), slice(1, n[2] + 1))
    for fids, tol in (((EX, EY, EZ), 5e-4), ((BX, BY, BZ), 5e-4), ((JX, JY, JZ), 1e-2)):
        a = np.concatenate([c32.field(f)[inner].ravel() for f in fids])
        b = np.concatenate([c64.field(f)[inner].ravel() for f in fids])
        assert rel(a, b) < tol, (fids, rel(a, b))
    g32 = by_id(c32.particles(fields=False))
    g64 = by_id(c64.particles(fields=False))
    assert np.array_equal(g32["id"], g64["id"])
    assert rel(g32["x"], g64["x"]) < 1e-6
    assert rel(g32["v"], g64["v"]) < 1e-3
    d32, d64 = c32.diag(), c64.diag()
    assert abs(d32[4] + d32[5] - d64[4] - d64[5]) < 1e-5 * abs(d64[4] + d64[5])
    assert abs(d32[7] - d64[7]) <= 1e-12 * qabs


def test_c4_full_size_fp64_bit_exact_20_steps():
    case = wave_case("c4")
    ctx, o = pair(case)
    ctx.init_constant()
    o.initialize()
    dt = case.dt_cfl()
    t = 0.0
    for _ in range(20):
        o.step(t, dt)
        t += dt
    ctx.step(20, 0.0, dt)
    assert_particles_equal(ctx, o, fields=False)
    assert_fields_equal(ctx, o)


C4_STEPS = 400


def c4_energy(amp_scale=1.0, phase=0.0, eps=0.01):
    case = wave_case("c4", amp_scale, phase, eps)
    ctx = Context(emo_config(case))
    ctx.load_maxwellian(emo_loader(case))
    ctx.init_constant()
    ctx.interpolate()
    ctx.sort()
    ctx.step(C4_STEPS, 0.0, case.dt_cfl())
    d = ctx.diag()
    ctx.close()
    return d[4] + d[5]


@pytest.fixture(scope="module")
def c4_gradient():
    case = wave_case("c4")
    ctx = Context(emo_config(case))
    ctx.load_maxwellian(emo_loader(case))
    ctx.init_constant()
    ctx.interpolate()
    out = ctx.adjoint_field_energy(C4_STEPS, 0.0, case.dt_cfl(), 20)
    p = ctx.particles(fields=False)
    ctx.close()
    species = np.empty(len(p["id"]), np.int32)
    species[p["id"]] = p["species"]
    dxde = loader_dx_deps(emo_loader(case), emo_config(case))
    d_eps = float((out["xbar0"][:, 0] * dxde)[species == 0].sum())
    amp = case.waves[0].amp
    return {"F": out["energy"], "amp": amp, "d_amp": float(out["d_amp"][0]),
            "d_phase_rad": float(out["d_phase"][0]), "d_eps": d_eps}


def test_c4_gradient_against_central_differences(c4_gradient):
    """dF/d(amplitude), dF/d(phase) per radian and dF/d(eps) of the 400-step
    c4 run (checkpoint every 20 steps) against central differences.  F is a
    double near 463, so a difference quotient with step h resolves no finer
    than ulp(F)/(2h): at eps's h = 1e-10 that is 1.2e-3 of |dF/deps| = 0.24
    per ulp.  Steps of 1e-8 relative (amplitude), 1e-8 rad (phase) or 1e-9
    (eps) already move particles across faces in this run."""
    g = c4_gradient
    amp = g["amp"]
    F = c4_energy()
    assert F == pytest.approx(g["F"], rel=1e-14)
    ulp = math.ulp(F)
    rows = {}
    small = {"amp": 1e-9 * amp, "phase": 1e-9, "eps": 1e-10}
    for name in ("amp", "phase", "eps"):
        for h in (small[name], 1e-4):
            if name == "amp":
                fd = (c4_energy(amp_scale=1 + h / amp) - c4_energy(amp_scale=1 - h / amp)) / (2 * h)
                ad = g["d_amp"]
            elif name == "phase":
                hd = h * 180.0 / math.pi  # the input phase is in degrees
                fd = (c4_energy(phase=hd) - c4_energy(phase=-hd)) / (2 * h)
                ad = g["d_phase_rad"]
            else:
                fd = (c4_energy(eps=0.01 + h) - c4_energy(eps=0.01 - h)) / (2 * h)
                ad = g["d_eps"]
            rows[(name, h)] = (ad, fd, h)
            print(f"c4 dF/d{name}: adjoint {ad:.10e}, central difference (h={h:.1e}) {fd:.10e}, "
                  f"rel diff {abs(ad - fd) / abs(fd):.2e}", flush=True)
    for name, h in small.items():
        ad, fd, _ = rows[(name, h)]
        assert abs(ad - fd) <= 1e-4 * abs(fd) + 4 * ulp / (2 * h), (name, rows[(name, h)])
    for name in small:
        assert math.isfinite(rows[(name, 1e-4)][1]), (name, rows[(name, 1e-4)])
