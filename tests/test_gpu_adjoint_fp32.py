"""The discrete adjoint of the FP32 chosen form (emo_adjoint_field_energy on
an FP32 context): FP32 particle adjoints (the push, BCs, sort and deposit
transposed, with the forward recomputed from the taped pre-push records) and
the FP64 field adjoints.  Checked against the FP64 plain form's adjoint on the
same input (which test_gpu_adjoint.py pins to central differences), against
central differences of the FP32 forward itself, for determinism across
checkpoint intervals, and for the state the call leaves behind.  All calls go
through the C ABI."""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import FP32_CHOSEN, FP64_PLAIN, Context, loader_dx_deps
from emo_case import emo_config, emo_loader
from oracle_ctypes import EXTERNAL, PERIODIC, REFLECTING, Case, Oracle, Wave

pytestmark = pytest.mark.gpu

NSTEPS = 16


def case_with(amp=0.3, phase=30.0, eps=0.05, relativity=1):
    w = Wave(side=-1, pol=2, amp=amp, omega=0.39, phase=phase, inv_width=0.08, delay=3.0)
    return Case(n=(12, 6, 6), L=(12.0, 6.0, 6.0), nparticles=2 * 2 * 432, temp=(501.8, 501.8),
                dens_phys=4.355e9, relativity=relativity, waves=[w],
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pert_eps=(eps, 0.0), pert_k=(2 * math.pi / 12.0, 0.0))


def setup(case, precision):
    ctx = Context(emo_config(case, precision=precision))
    ctx.load_maxwellian(emo_loader(case))
    ctx.init_constant()
    if precision == FP64_PLAIN:
        ctx.interpolate()
    return ctx


def energy32(case, dt, nsteps=NSTEPS):
    ctx = setup(case, FP32_CHOSEN)
    ctx.step(nsteps, 0.0, dt)
    d = ctx.diag()
    ctx.close()
    return d[4] + d[5]


def adjoint(case, dt, k, precision, nsteps=NSTEPS):
    ctx = setup(case, precision)
    out = ctx.adjoint_field_energy(nsteps, 0.0, dt, k)
    ctx.close()
    return out


def rel_err(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def base():
    case = case_with()
    dt = case.dt_cfl()
    return case, dt, adjoint(case, dt, 4, FP32_CHOSEN), adjoint(case, dt, 4, FP64_PLAIN)


def test_fp32_adjoint_objective_is_the_forward_energy(base):
    case, dt, a32, a64 = base
    assert a32["energy"] == energy32(case, dt)
    assert rel_err(a32["energy"], a64["energy"]) < 1e-5


def test_fp32_checkpoint_interval_does_not_change_the_gradient(base):
    """The backward is deterministic (fixed reduction orders) and the
    recomputation is bitwise the forward, so any interval gives the same bits."""
    case, dt, a32, _ = base
    for k in (1, 3, NSTEPS, 0):
        other = adjoint(case, dt, k, FP32_CHOSEN)
        assert other["energy"] == a32["energy"], k
        for key in ("d_amp", "d_phase", "xbar0", "vbar0", "gbar0"):
            assert np.array_equal(other[key], a32[key]), (k, key)


def test_fp32_wave_gradients_match_the_fp64_adjoint(base):
    """FP32 rounding of the trajectory (relative ~1e-6 over these steps) and of
    the particle adjoint: the wave gradients agree with the FP64 adjoint to
    1e-4 relative."""
    _, _, a32, a64 = base
    assert a64["d_amp"][0] != 0 and a64["d_phase"][0] != 0
    assert rel_err(a32["d_amp"][0], a64["d_amp"][0]) < 1e-4, (a32["d_amp"], a64["d_amp"])
    assert rel_err(a32["d_phase"][0], a64["d_phase"][0]) < 1e-4, (a32["d_phase"], a64["d_phase"])


def test_fp32_particle_gradients_match_the_fp64_adjoint(base):
    """Per-particle adjoints of the initial state, by id: relative L2 over all
    particles within 1e-3 (a particle within FP32 resolution of a cell face
    may take the other cell's weights in one form), and the density
    perturbation's gradient (a sum over particles) within 1e-4."""
    case, _, a32, a64 = base
    for key in ("xbar0", "vbar0", "gbar0"):
        assert np.linalg.norm(a64[key]) > 0, key
        assert rel_l2(a32[key], a64[key]) < 1e-3, (key, rel_l2(a32[key], a64[key]))
    dxde = loader_dx_deps(emo_loader(case), emo_config(case))
    g32 = float(np.dot(a32["xbar0"][:, 0], dxde))
    g64 = float(np.dot(a64["xbar0"][:, 0], dxde))
    assert rel_err(g32, g64) < 1e-4, (g32, g64)


def test_fp32_amplitude_gradient_matches_central_differences(base):
    """Central differences of the FP32 forward: F is nearly quadratic in the
    amplitude, so a step of 1e-3 keeps the truncation error far below the
    FP32 noise of F (relative ~1e-7)."""
    case, dt, a32, _ = base
    h = 1e-3
    fd = (energy32(case_with(amp=0.3 + h), dt) - energy32(case_with(amp=0.3 - h), dt)) / (2 * h)
    assert rel_err(a32["d_amp"][0], fd) < 1e-3, (a32["d_amp"][0], fd)


@pytest.mark.parametrize("relativity", [0, 1])
def test_fp32_initial_velocity_gradient_matches_the_fp64_adjoint(relativity):
    """dF/dv0 and dF/dgamma0 through emo_load_particles (u = gamma0 v in the
    FP32 store), along a random direction, against the FP64 adjoint."""
    case = Case(n=(8, 6, 6), L=(8.0, 6.0, 6.0), nparticles=2 * 2 * 288, temp=(501.8, 501.8),
                dens_phys=4.355e9, relativity=relativity, E0=(0.01, 0.0, 0.0), B0=(0.0, 0.0, 0.05))
    o = Oracle(case)
    assert o.load() == 0
    p = o.particles()
    dt = case.dt_cfl()
    out = {}
    for prec in (FP64_PLAIN, FP32_CHOSEN):
        ctx = Context(emo_config(case, precision=prec, q=o.q, m=o.m))
        ctx.load_particles(p.x, p.v, p.gamma, p.species, p.id)
        ctx.init_constant()
        if prec == FP64_PLAIN:
            ctx.interpolate()
        out[prec] = ctx.adjoint_field_energy(10, 0.0, dt, 3)
        ctx.close()
    rng = np.random.default_rng(7)
    dv = rng.normal(size=p.v.shape)
    dg = rng.normal(size=p.gamma.shape)
    g64 = float(np.sum(out[FP64_PLAIN]["vbar0"] * dv))
    g32 = float(np.sum(out[FP32_CHOSEN]["vbar0"] * dv))
    assert g64 != 0
    assert rel_err(g32, g64) < 1e-3, (g32, g64)
    if relativity:
        h64 = float(np.sum(out[FP64_PLAIN]["gbar0"] * dg))
        h32 = float(np.sum(out[FP32_CHOSEN]["gbar0"] * dg))
        assert h64 != 0
        assert rel_err(h32, h64) < 1e-3, (h32, h64)
    else:
        assert not np.any(out[FP32_CHOSEN]["gbar0"])


def test_fp32_adjoint_leaves_the_final_state(base):
    """After the gradient call the context holds the state emo_step reaches
    (records, fields, currents), bit for bit, and steps on from it alike."""
    case, dt, _, _ = base
    a = setup(case, FP32_CHOSEN)
    a.adjoint_field_energy(NSTEPS, 0.0, dt, 3)
    b = setup(case, FP32_CHOSEN)
    b.step(NSTEPS, 0.0, dt)
    for _ in range(2):
        pa, pb = a.particles(fields=False), b.particles(fields=False)
        for k in ("x", "v", "species", "id"):
            assert np.array_equal(pa[k], pb[k]), k
        for f in range(13):
            assert np.array_equal(a.field(f), b.field(f)), f
        a.step(3, NSTEPS * dt, dt)
        b.step(3, NSTEPS * dt, dt)


def test_fp32_autograd_function(base):
    import torch

    from apc_524_project_amd.adjoint import RunSpec, field_energy

    case, dt, a32, _ = base
    spec = RunSpec(cfg=emo_config(case, precision=FP32_CHOSEN), loader=emo_loader(case), nsteps=NSTEPS, dt=dt,
                   checkpoint_every=4)
    amps = torch.tensor([0.3], dtype=torch.float64, requires_grad=True)
    phases = torch.tensor([30.0], dtype=torch.float64, requires_grad=True)
    eps = torch.tensor([0.05, 0.0], dtype=torch.float64, requires_grad=True)
    F = field_energy(amps, phases, eps, spec)
    assert F.item() == a32["energy"]
    F.backward()
    assert amps.grad[0].item() == a32["d_amp"][0]


def test_fp32_adjoint_of_zero_steps():
    """nsteps = 0: the objective is the current field energy and every
    gradient is zero."""
    case = case_with()
    ctx = setup(case, FP32_CHOSEN)
    d = ctx.diag()
    out = ctx.adjoint_field_energy(0, 0.0, case.dt_cfl(), 4)
    assert out["energy"] == d[4] + d[5]
    assert out["d_amp"][0] == 0 and out["d_phase"][0] == 0
    for key in ("xbar0", "vbar0", "gbar0"):
        assert not np.any(out[key]), key
    ctx.close()


def test_fp32_adjoint_calls_in_a_row_continue_the_run():
    """Two gradient calls of 8 steps on one context (the second reuses the
    first's device pool and starts from its final state) reach the state
    and energy of 16 plain steps."""
    case = case_with()
    dt = case.dt_cfl()
    a = setup(case, FP32_CHOSEN)
    a.adjoint_field_energy(8, 0.0, dt, 3)
    second = a.adjoint_field_energy(8, 8 * dt, dt, 3)
    b = setup(case, FP32_CHOSEN)
    b.step(16, 0.0, dt)
    d = b.diag()
    assert second["energy"] == d[4] + d[5]
    assert np.array_equal(a.particles(fields=False)["x"], b.particles(fields=False)["x"])
    assert second["d_amp"][0] != 0
    a.close()
    b.close()


def test_fp32_boris_lightbc_gradients_match_the_fp64_adjoint():
    """The non-relativistic (Boris) particle adjoint with reflecting walls."""
    case = case_with(relativity=0)
    dt = case.dt_cfl()
    a32 = adjoint(case, dt, 4, FP32_CHOSEN)
    a64 = adjoint(case, dt, 4, FP64_PLAIN)
    assert a64["d_amp"][0] != 0
    assert rel_err(a32["d_amp"][0], a64["d_amp"][0]) < 1e-4
    assert rel_err(a32["d_phase"][0], a64["d_phase"][0]) < 1e-4
    for key in ("xbar0", "vbar0"):
        assert rel_l2(a32[key], a64[key]) < 1e-3, (key, rel_l2(a32[key], a64[key]))


def test_fp32_adjoint_without_particles_matches_fp64():
    """A LightBC pulse into vacuum (no particles): the gradient of the field
    energy with respect to the pulse is the field adjoint's alone, in both
    forms."""
    w = Wave(side=-1, pol=2, amp=0.3, omega=0.39, phase=30.0, inv_width=0.08, delay=3.0)
    case = Case(n=(12, 4, 4), L=(12.0, 4.0, 4.0), nparticles=0, temp=(0.0, 0.0), dens_phys=0.0,
                relativity=1, waves=[w],
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC))
    out = {}
    for prec in (FP64_PLAIN, FP32_CHOSEN):
        ctx = Context(emo_config(case, precision=prec, max_particles=1))
        ctx.init_constant()
        if prec == FP64_PLAIN:
            ctx.interpolate()
        out[prec] = ctx.adjoint_field_energy(NSTEPS, 0.0, case.dt_cfl(), 4)
        ctx.close()
    assert out[FP64_PLAIN]["d_amp"][0] != 0
    assert rel_err(out[FP32_CHOSEN]["d_amp"][0], out[FP64_PLAIN]["d_amp"][0]) < 1e-5
    assert rel_err(out[FP32_CHOSEN]["d_phase"][0], out[FP64_PLAIN]["d_phase"][0]) < 1e-5
