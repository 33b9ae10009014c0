"""torch.autograd.Function over libemopic's discrete adjoint.

    energy = FieldEnergy.apply(amps, phases, eps, spec)
    energy.backward()   # amps.grad, phases.grad, eps.grad

`energy` is the final field energy E_en + B_en of the reference's history
diagnostics (src/IO/output.cpp:90-121) after `spec.nsteps` steps of the form
`spec.cfg.precision` names (FP64 plain, or the FP32 chosen form with its
FP32 particle adjoint and FP64 field adjoints), as a function of the injected pulses' peak amplitudes and phases
(LightBC, src/grid/lightBC.cpp:165-198; phases in degrees as in the input
file) and of the per-species density-perturbation amplitudes of the
displacement loader.  When an input requires a gradient, the forward pass
runs emo_adjoint_field_energy, which does the forward steps and the adjoint
sweep (checkpoint and recompute every `spec.checkpoint_every` steps) in one
call, and backward scales the stored gradient; otherwise it runs the forward
steps alone.  The computation runs on the GPU through the C ABI; there is no
CPU path.
"""
from __future__ import annotations

import copy
import math
from dataclasses import dataclass

import numpy as np
import torch

from ._lib import FP64_PLAIN, Context, EmoConfig, EmoLoader, loader_dx_deps


@dataclass
class RunSpec:
    cfg: EmoConfig          # geometry, BCs, waves (amplitudes/phases are overridden)
    loader: EmoLoader       # species and loader (pert_eps is overridden)
    nsteps: int
    dt: float
    t0: float = 0.0
    checkpoint_every: int = 20


def _setup(spec: RunSpec, amps, phases, eps):
    cfg = copy.deepcopy(spec.cfg)
    ld = copy.deepcopy(spec.loader)
    for w in range(int(cfg.nwaves)):
        cfg.peakamps[w] = float(amps[w])
        cfg.phases[w] = float(phases[w])
    for s in range(int(ld.nspecies)):
        ld.pert_eps[s] = float(eps[s])
    ctx = Context(cfg)
    ctx.load_maxwellian(ld)
    ctx.init_constant()
    if cfg.precision == FP64_PLAIN:
        ctx.interpolate()  # the FP32 form gathers from the grid inside its step
    return ctx, cfg, ld


def field_energy_and_grad(spec: RunSpec, amps, phases, eps):
    """(F, dF/damps, dF/dphases [per degree], dF/deps) as numpy arrays."""
    ctx, cfg, ld = _setup(spec, amps, phases, eps)
    ids0 = np.arange(int(ld.nparticles_tot))
    out = ctx.adjoint_field_energy(spec.nsteps, spec.t0, spec.dt, spec.checkpoint_every)
    nsp = int(ld.nspecies)
    d_eps = np.zeros(nsp)
    if any(float(eps[s]) != 0.0 or ld.pert_k[s] != 0.0 for s in range(nsp)):
        dxde = loader_dx_deps(ld, cfg)
        p = ctx.particles(fields=False)
        species_by_id = np.empty(len(ids0), np.int32)
        species_by_id[p["id"]] = p["species"]
        contrib = out["xbar0"][:, 0] * dxde
        for s in range(nsp):
            d_eps[s] = float(contrib[species_by_id == s].sum())
    ctx.close()
    return (out["energy"], np.asarray(out["d_amp"]), np.asarray(out["d_phase"]) * math.pi / 180.0, d_eps)


def field_energy_only(spec: RunSpec, amps, phases, eps) -> float:
    """F alone: the forward steps and the diagnostics, no adjoint sweep."""
    ctx, cfg, ld = _setup(spec, amps, phases, eps)
    if cfg.precision == FP64_PLAIN:
        ctx.sort()  # as the gradient call does before its first push (src/main.cpp:225-226)
    ctx.step(spec.nsteps, spec.t0, spec.dt)
    d = ctx.diag()
    ctx.close()
    return float(d[5] + d[4])


class FieldEnergy(torch.autograd.Function):
    """F(amps, phases, eps); the output and the saved gradients live on the
    inputs' device and dtype.  The gradient call (forward steps plus the
    adjoint sweep) runs only when an input requires a gradient; otherwise
    the forward steps alone give F."""

    @staticmethod
    def forward(ctx, amps, phases, eps, spec):
        a = amps.detach().cpu().numpy()
        p = phases.detach().cpu().numpy()
        e = eps.detach().cpu().numpy()
        if any(ctx.needs_input_grad[:3]):
            F, da, dp, de = field_energy_and_grad(spec, a, p, e)
        else:
            F = field_energy_only(spec, a, p, e)
            da, dp, de = np.zeros(len(a)), np.zeros(len(p)), np.zeros(len(e))
        ctx.save_for_backward(torch.as_tensor(da, dtype=amps.dtype, device=amps.device),
                              torch.as_tensor(dp, dtype=phases.dtype, device=phases.device),
                              torch.as_tensor(de, dtype=eps.dtype, device=eps.device))
        return torch.tensor(F, dtype=amps.dtype, device=amps.device)

    @staticmethod
    def backward(ctx, g):
        da, dp, de = ctx.saved_tensors
        return g * da, g * dp, g * de, None


def field_energy(amps: torch.Tensor, phases: torch.Tensor, eps: torch.Tensor, spec: RunSpec) -> torch.Tensor:
    return FieldEnergy.apply(amps, phases, eps, spec)
