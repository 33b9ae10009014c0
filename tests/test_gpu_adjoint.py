"""The discrete adjoint of the FP64 step (emo_adjoint_field_energy) against
central finite differences of the same GPU forward run, on a small LightBC
case with reflecting particle walls and a displacement-loaded density
perturbation (the geometry of the differentiable-run workload, reduced).

The objective is the final field energy E_en + B_en of the reference's
history diagnostics.  Finite differences use the forward path the adjoint
differentiates: sort, then emo_step.  The steps are few and the step sizes
small so that no particle changes cell under the perturbation (the
reference's weights jump at cell faces; the adjoint is the derivative away
from those faces).
"""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import Context, loader_dx_deps
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


def setup(case):
    ctx = Context(emo_config(case))
    ctx.load_maxwellian(emo_loader(case))
    ctx.init_constant()
    ctx.interpolate()
    return ctx


def energy(case, dt, nsteps=NSTEPS):
    ctx = setup(case)
    ctx.sort()
    ctx.step(nsteps, 0.0, dt)
    d = ctx.diag()
    return d[4] + d[5]


def adjoint(case, dt, k, nsteps=NSTEPS):
    ctx = setup(case)
    return ctx.adjoint_field_energy(nsteps, 0.0, dt, k)


@pytest.fixture(scope="module")
def base():
    case = case_with()
    dt = case.dt_cfl()
    return case, dt, adjoint(case, dt, 4)


def test_adjoint_objective_is_the_forward_energy(base):
    case, dt, adj = base
    assert adj["energy"] == energy(case, dt)
    assert adj["energy"] > 0


def test_checkpoint_interval_does_not_change_the_gradient(base):
    case, dt, adj = base
    for k in (1, 3, NSTEPS):
        other = adjoint(case, dt, k)
        assert np.array_equal(other["d_amp"], adj["d_amp"]), k
        assert np.array_equal(other["d_phase"], adj["d_phase"]), k
        assert np.array_equal(other["xbar0"], adj["xbar0"]), k


def rel_err(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_amplitude_gradient_matches_central_differences(base):
    case, dt, adj = base
    h = 1e-8
    fd = (energy(case_with(amp=0.3 + h), dt) - energy(case_with(amp=0.3 - h), dt)) / (2 * h)
    assert adj["d_amp"][0] != 0
    assert rel_err(adj["d_amp"][0], fd) < 1e-5, (adj["d_amp"][0], fd)


def test_phase_gradient_matches_central_differences(base):
    case, dt, adj = base
    h = 1e-6  # degrees
    fd_deg = (energy(case_with(phase=30.0 + h), dt) - energy(case_with(phase=30.0 - h), dt)) / (2 * h)
    fd_rad = fd_deg * 180.0 / math.pi
    assert rel_err(adj["d_phase"][0], fd_rad) < 1e-5, (adj["d_phase"][0], fd_rad)


def test_density_perturbation_gradient_matches_central_differences(base):
    case, dt, adj = base
    cfg = emo_config(case)
    dxde = loader_dx_deps(emo_loader(case), cfg)
    g = float(np.dot(adj["xbar0"][:, 0], dxde))
    h = 1e-7
    fd = (energy(case_with(eps=0.05 + h), dt) - energy(case_with(eps=0.05 - h), dt)) / (2 * h)
    assert rel_err(g, fd) < 1e-4, (g, fd)


@pytest.mark.parametrize("relativity", [0, 1])
def test_initial_velocity_gradient_matches_central_differences(relativity):
    """dF/dv0 through emo_load_particles, along a random direction in the
    velocities of all particles (the stored gamma held fixed)."""
    case = Case(n=(8, 6, 6), L=(8.0, 6.0, 6.0), nparticles=2 * 2 * 288, temp=(501.8, 501.8),
                dens_phys=4.355e9, relativity=relativity, E0=(0.01, 0.0, 0.0), B0=(0.0, 0.0, 0.05))
    o = Oracle(case)
    assert o.load() == 0
    p = o.particles()
    cfg = emo_config(case, q=o.q, m=o.m)
    dt = case.dt_cfl()

    def run(v, adj=False):
        ctx = Context(cfg)
        ctx.load_particles(p.x, v, p.gamma, p.species, p.id)
        ctx.init_constant()
        ctx.interpolate()
        if adj:
            return ctx.adjoint_field_energy(10, 0.0, dt, 3)
        ctx.sort()
        ctx.step(10, 0.0, dt)
        d = ctx.diag()
        return d[4] + d[5]

    adj = run(p.v, adj=True)
    assert adj["energy"] == run(p.v)
    direction = np.random.default_rng(7).normal(size=p.v.shape)
    g = float(np.sum(adj["vbar0"][p.id] * direction))
    h = 1e-8
    fd = (run(p.v + h * direction) - run(p.v - h * direction)) / (2 * h)
    assert g != 0
    assert rel_err(g, fd) < 1e-4, (g, fd)


def test_autograd_function_returns_the_adjoint_gradient(base):
    import torch

    from apc_524_project_amd.adjoint import RunSpec, field_energy

    case, dt, adj = base
    spec = RunSpec(cfg=emo_config(case), loader=emo_loader(case), nsteps=NSTEPS, dt=dt, checkpoint_every=4)
    amps = torch.tensor([0.3], dtype=torch.float64, requires_grad=True)
    phases = torch.tensor([30.0], dtype=torch.float64, requires_grad=True)
    eps = torch.tensor([0.05, 0.0], dtype=torch.float64, requires_grad=True)
    F = field_energy(amps, phases, eps, spec)
    assert F.item() == adj["energy"]
    (2.0 * F).backward()
    assert amps.grad[0].item() == 2.0 * adj["d_amp"][0]
    assert phases.grad[0].item() == pytest.approx(2.0 * adj["d_phase"][0] * math.pi / 180.0, rel=1e-14)
    dxde = loader_dx_deps(emo_loader(case), emo_config(case))
    g_eps = 2.0 * float(np.dot(adj["xbar0"][:, 0], dxde))
    assert eps.grad[0].item() == pytest.approx(g_eps, rel=1e-12)
    assert eps.grad[1].item() == 0.0


def test_adjoint_leaves_the_final_state(base):
    """After the gradient call the context holds the state emo_step reaches
    (particles with their own ids, fields), bit for bit."""
    case, dt, _ = base
    a = setup(case)
    a.adjoint_field_energy(NSTEPS, 0.0, dt, 3)
    b = setup(case)
    b.sort()
    b.step(NSTEPS, 0.0, dt)
    pa, pb = a.particles(), b.particles()
    oa, ob = np.argsort(pa["id"]), np.argsort(pb["id"])
    for k in ("x", "v", "gamma", "species", "field"):
        assert np.array_equal(pa[k][oa], pb[k][ob]), k
    for f in range(13):
        assert np.array_equal(a.field(f), b.field(f)), f


def test_autograd_on_device_at_zero_perturbation():
    """Inputs on cuda:0: the output and the gradients come back on that
    device.  At eps = 0 with a wavenumber set, dF/deps is the linear
    response about the uniform load (nonzero), the dot product of the
    position adjoint with the loader's sensitivity, and central differences
    in eps agree with it.  Without a gradient the forward alone gives F."""
    import torch

    from apc_524_project_amd.adjoint import RunSpec, field_energy

    case = case_with(eps=0.0)
    dt = case.dt_cfl()
    spec = RunSpec(cfg=emo_config(case), loader=emo_loader(case), nsteps=NSTEPS, dt=dt, checkpoint_every=4)
    dev = torch.device("cuda:0")
    amps = torch.tensor([0.3], dtype=torch.float64, device=dev, requires_grad=True)
    phases = torch.tensor([30.0], dtype=torch.float64, device=dev, requires_grad=True)
    eps = torch.tensor([0.0, 0.0], dtype=torch.float64, device=dev, requires_grad=True)
    F = field_energy(amps, phases, eps, spec)
    assert F.device == dev
    F.backward()
    assert amps.grad.device == dev and eps.grad.device == dev
    adj = adjoint(case, dt, 4)
    assert F.item() == adj["energy"]
    dxde = loader_dx_deps(emo_loader(case), emo_config(case))
    assert np.abs(dxde).max() > 1.0
    g_eps = float(np.dot(adj["xbar0"][:, 0], dxde))
    assert g_eps != 0.0
    assert eps.grad[0].item() == pytest.approx(g_eps, rel=1e-12)
    h = 1e-6
    fd = (energy(case_with(eps=h), dt) - energy(case_with(eps=-h), dt)) / (2 * h)
    assert rel_err(g_eps, fd) < 1e-4, (g_eps, fd)
    with torch.no_grad():
        F0 = field_energy(amps.detach(), phases.detach(), eps.detach(), spec)
    assert F0.item() == pytest.approx(adj["energy"], rel=1e-14)
