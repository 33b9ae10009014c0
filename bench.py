#!/usr/bin/env python3
"""Benchmark: particle-steps/s of EMOOPIC's per-timestep PIC loop on one MI355X.

    python bench.py [--steps K] [--warmup W] [--mode fp64|fp32|both]

Workload c2 (64^3 periodic, 32 particles/cell/species, 16.8M particles,
relativistic Boris, electron sigma 0.1c with a 0.01 cos(2 pi x/64)
density perturbation), synthetic seeded input, inputs resident in HBM
before the timed region.  A step is one full iteration of the loop (push,
particle BCs, sort, deposit, fold, Yee, ghost refresh, gather); value is
particles x steps / wall time of K steps.

Prints one JSON line.  The roofline entry is for the step's dominant
kernel, timed with HIP events on the library stream during the timed
steps; cpu_baseline times the reference's own serial step loop
(oracle/_ref, built from the reference sources) on one host core over a
bounded sample.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
METRIC = "particle-steps/s (push+gather+deposit) FP32 vs FP64; adjoint/forward time"


def algorithmic_bytes(kernel: str, np_: int, ncell: int, relativity: int) -> float | None:
    """Bytes a kernel must move per launch at minimum (each operand read or
    written once): the per-particle and per-cell figures of the table below
    times the particle and cell counts."""
    ng = ncell  # physical cells ~ padded points for the grid kernels
    table = {
        # FP64 plain form
        "gather_f64": 72 * np_ + 72 * ncell,
        "push_f64": (164 if relativity else 148) * np_,
        "particle_bc_f64": 24 * np_,
        "sort_keys": 32 * np_,
        "sort_permute": 236 * np_,
        "pgd_f64": (117 if relativity else 101) * np_ + 72 * ncell,
        "rank_f64": 3 * np_ + 8 * ncell,
        "base_f64": 4 * 27 * ncell,
        "scatter_f64": 145 * np_ + 8 * ncell,
        "deposit_cells_f64": 52 * np_ + 168 * ncell,
        "deposit_grid_f64": 160 * ncell + 32 * ng,
        "yee_E_f64": 96 * ncell,
        "yee_B_f64": 72 * ncell,
        # FP32 chosen form: 32-byte particles (cell word, offset[3], u[3], int32 id)
        "push_deposit_f32": 64 * np_ + 140 * ncell,
        "jsum_f32": 48 * ncell + 12 * ng,
        "yee_E_f32": 48 * ncell,
        "yee_B_f32": 36 * ncell,
    }
    return table.get(kernel)


KERNELS = {
    "fp64": ["pgd_f64", "far_f64", "rank_f64", "base_f64", "scan_f64", "scatter_f64", "gather_f64", "push_f64",
             "particle_bc_f64", "sort_keys", "sort_radix", "sort_permute", "deposit_cells_f64", "deposit_grid_f64",
             "yee_E_f64", "yee_B_f64"],
    "fp32": ["cellvec_f32", "push_deposit_f32", "check_f32", "relayout_f32", "arrive_f32", "jsum_f32",
             "yee_E_f32", "yee_B_f32"],
}


def run_gpu(mode: str, steps: int, warmup: int, scale: float):
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN, FP64_PLAIN

    prec = FP64_PLAIN if mode == "fp64" else FP32_CHOSEN
    w = emo.make_workload("c2", precision=prec, scale=scale)
    t0 = time.perf_counter()
    ctx = emo.Context(w.cfg)
    ctx.load_maxwellian(w.loader)
    ctx.init_constant()
    ctx.synchronize()
    t_load = time.perf_counter() - t0
    print(f"[bench] {mode}: loaded {w.nparticles} particles in {t_load:.1f}s", file=sys.stderr, flush=True)
    t = 0.0
    ctx.step(warmup, t, w.dt)
    t += warmup * w.dt
    ctx.synchronize()
    # per-kernel breakdown: HIP events around every launch of a few steps
    nb = max(1, min(steps, 10))
    ctx.timing(True)
    ctx.timing_reset()
    a = time.perf_counter()
    ctx.step(nb, t, w.dt)
    ctx.synchronize()
    el_ev = time.perf_counter() - a
    t += nb * w.dt
    ktimes = {}
    for k in KERNELS[mode]:
        ms, n = ctx.kernel_time(k)
        if n:
            ktimes[k] = (ms, n)
    own = {k: v for k, v in ktimes.items() if k not in LIBRARY_CALLS}
    dom = max(own, key=lambda k: own[k][0]) if own else None
    # the timed steps: events around the dominant kernel only (two per step)
    ctx.timing_select(dom)
    ctx.timing_reset()
    ctx.synchronize()
    a = time.perf_counter()
    ctx.step(steps, t, w.dt)
    ctx.synchronize()
    el = time.perf_counter() - a
    ctx.timing(False)
    if dom:
        ktimes[dom] = ctx.kernel_time(dom)
    ctx.timing_select(None)
    rate = w.nparticles * steps / el
    print(f"[bench] {mode}: {steps} steps in {el:.3f}s -> {rate:.4g} particle-steps/s "
          f"({1e3 * el_ev / nb:.4f} ms/step with events around every launch)", file=sys.stderr, flush=True)
    ctx.close()
    floor = copy_floor(w.nparticles, 32) if mode == "fp32" else None
    return dict(rate=rate, elapsed=el, ms_per_step=1e3 * el / steps, ktimes=ktimes, workload=w, mode=mode,
                scale=scale, dominant=dom, ms_per_step_all_events=1e3 * el_ev / nb, copy_floor=floor)


def copy_floor(np_: int, record_bytes: int, reps: int = 20):
    """A plain device copy of the particle records' bytes (np_ records of
    record_bytes, as two arrays of 16-byte vectors, read once and written
    once) timed with events on torch's stream, right after the timed steps:
    the time a pure stream of the particle pass's own particle bytes takes on
    this GPU in its current clock state."""
    import torch

    n = np_ * record_bytes // 16
    half = n // 2
    src = [torch.empty((half, 4), dtype=torch.int32, device="cuda") for _ in range(2)]
    dst = [torch.empty((half, 4), dtype=torch.int32, device="cuda") for _ in range(2)]
    for t in src:
        t.fill_(1)
    for d, s_ in zip(dst, src):
        d.copy_(s_)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        for d, s_ in zip(dst, src):
            d.copy_(s_)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    nbytes = 2 * 2 * half * 16
    del src, dst
    torch.cuda.empty_cache()
    return {"what": f"device copy of {np_} x {record_bytes} B records (read + written)", "bytes": int(nbytes),
            "ms": round(ms, 5), "GBs": round(nbytes / ms / 1e6, 1)}


def run_fp64_per_phase(steps: int, warmup: int, scale: float):
    """The FP64 plain form driven phase by phase through the C ABI in the
    reference loop's order (src/main.cpp:225-275: SortParticles, Push,
    particle BCs + clearGhosts, depositRhoJ, executeBC(-2), evolveFields,
    executeBC(-1), InterpolateEB): the reference's pass structure, one
    library call per phase."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP64_PLAIN

    w = emo.make_workload("c2", precision=FP64_PLAIN, scale=scale)
    ctx = emo.Context(w.cfg)
    ctx.load_maxwellian(w.loader)
    ctx.init_constant()
    ctx.interpolate()
    t = 0.0

    def one(t):
        ctx.sort()
        ctx.push(w.dt)
        ctx.particle_bc()
        ctx.deposit(True)
        ctx.fold()
        ctx.evolve(w.dt)
        ctx.refresh(t, w.dt)
        ctx.interpolate()

    for _ in range(warmup):
        one(t)
        t += w.dt
    ctx.synchronize()
    a = time.perf_counter()
    for _ in range(steps):
        one(t)
        t += w.dt
    ctx.synchronize()
    el = time.perf_counter() - a
    ctx.close()
    rate = w.nparticles * steps / el
    print(f"[bench] fp64 per-phase: {steps} steps in {el:.3f}s -> {rate:.4g} particle-steps/s", file=sys.stderr,
          flush=True)
    return rate, 1e3 * el / steps


# library calls (several rocprim kernels each): reported, never the roofline kernel
LIBRARY_CALLS = {"sort_radix", "scan_f32", "scan_f64"}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "c2_fp32_step_hbm_traffic.json")


def measured_traffic(kernel: str, mode: str, scale: float):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE
    and WRITE_SIZE runs of the full-size c2 FP32 step (each counter scaled by
    what it reads for a copy kernel of known size); None where not measured."""
    if mode != "fp32" or scale != 1.0 or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        d = json.load(f)
    e = d.get("per_launch", {}).get(kernel)
    return int(e["hbm_bytes"]) if e else None


def path_bytes(mode: str, np_: int, ncell: int, relativity: int) -> int:
    """Algorithmic bytes of one pass of the path (push + gather + deposit)
    over np_ particles and ncell cells: each particle's state read once and
    written once, each cell's field block read once and its current written
    once.
      FP32 chosen form: 56 B per particle (a 28-byte state: int32 cell,
      fp32 offset[3], fp32 u[3]) and 60 B per cell (E, B, B_tm1: 9 fp32
      values read; J: 3 two-float values written).
      FP64 plain form: 112 B per particle relativistic (x, v, gamma: 56 B),
      96 B for Boris (x, v), and 96 B per cell (9 fp64 values read, 3
      written)."""
    if mode == "fp32":
        return 56 * np_ + 60 * ncell
    return (112 if relativity else 96) * np_ + 96 * ncell


def roofline_for(res):
    """The dominant hand-written kernel of the step (longest average
    duration, HIP events around its launches on the library stream during
    the timed steps) against HBM: achieved = the path's algorithmic bytes
    per launch / its average duration.  Also the whole step against the same
    bytes, and the kernel's own operand bytes as a secondary figure."""
    w = res["workload"]
    kt = res["ktimes"]
    own = {k: v for k, v in kt.items() if k not in LIBRARY_CALLS}
    if not own:
        return None
    dom = res.get("dominant") or max(own, key=lambda k: own[k][0])
    ms, n = kt[dom]
    avg_s = ms / n / 1e3
    rel = int(w.cfg.relativity)
    b = path_bytes(res["mode"], w.nparticles, w.ncells, rel)
    ach = b / avg_s / 1e9
    traffic = measured_traffic(dom, res["mode"], res["scale"])
    step_s = res["ms_per_step"] / 1e3
    kown = algorithmic_bytes(dom, w.nparticles, w.ncells, rel)
    return {"kernel": dom, "bound": "hbm", "achieved": round(ach, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": traffic, "avg_ms": round(ms / n, 5),
            "algorithmic_bytes": int(b),
            "traffic_over_algorithmic": round(traffic / b, 3) if traffic else None,
            "whole_step": {"ms": round(res["ms_per_step"], 5), "achieved": round(b / step_s / 1e9, 1),
                           "frac": round(b / step_s / 1e9 / HBM_PEAK_GBS, 4)},
            "kernel_own_bytes": int(kown) if kown else None,
            "kernel_own_frac": round(kown / avg_s / 1e9 / HBM_PEAK_GBS, 4) if kown else None,
            "copy_floor": res.get("copy_floor")}


def run_yee(mode: str, name: str, calls: int, scale: float = 1.0):
    """The Yee update alone (emo_evolve_fields: the E and B stencil kernels)
    on workload `name`'s grid: wall time of `calls` calls between stream
    synchronisations.  Cell-updates/s = physical cells x calls / time; the
    achieved bandwidth counts 84 B (FP32) or 168 B (FP64) per cell-update
    (E: read E, B, J, write E; B: read B, E, write B)."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN, FP64_PLAIN

    w = emo.make_workload(name, precision=FP32_CHOSEN if mode == "fp32" else FP64_PLAIN, scale=scale)
    cfg = w.cfg
    cfg.max_particles = 1
    ctx = emo.Context(cfg)
    ctx.init_constant()
    for _ in range(3):
        ctx.evolve(w.dt)
    ctx.synchronize()
    a = time.perf_counter()
    for _ in range(calls):
        ctx.evolve(w.dt)
    ctx.synchronize()
    el = time.perf_counter() - a
    ctx.close()
    rate = w.ncells * calls / el
    per = 84 if mode == "fp32" else 168
    return {"grid": name, "cells": w.ncells, "calls": calls, "ms_per_update": round(1e3 * el / calls, 5),
            "cell_updates_per_s": round(rate, 1), "achieved_GBs": round(rate * per / 1e9, 1),
            "frac": round(rate * per / 1e9 / HBM_PEAK_GBS, 4)}


def run_adjoint_c4(steps: int | None, k: int, fp32: bool = False, fp64_leg: dict | None = None):
    """Workload c4 (64x32x32, x: LightBC pulse + reflecting walls, 8 ppc/species):
    wall time of the forward steps alone vs. the gradient call (the same
    forward with checkpoints, then the adjoint: the backward sweep with the
    segments the forward did not tape re-run from their checkpoints), each
    the median of three alternating runs, and a central-difference check of
    dF/d(pulse amplitude).  adjoint_s = gradient_s - forward_s is the time the
    derivative adds to a forward run.  FP64 plain form, or the FP32 chosen
    form (FP32 particle adjoint, FP64 field adjoints).  The FP32 gradient is
    checked against the FP64 leg's central difference (fp64_leg): central
    differences of the FP32 forward itself are not a derivative over hundreds
    of steps, since the rounding of the two perturbed FP32 runs differs and
    the plasma amplifies that difference."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN, FP64_PLAIN

    w = emo.make_workload("c4", precision=FP32_CHOSEN if fp32 else FP64_PLAIN)
    n = steps or w.steps

    def reset(c):
        c.load_maxwellian(w.loader)
        c.init_constant()
        if not fp32:
            c.interpolate()
        c.synchronize()
        return c

    def ready(amp=None, phase=None, eps=None):
        import copy

        cfg = copy.deepcopy(w.cfg)
        if amp is not None:
            cfg.peakamps[0] = amp
        if phase is not None:
            cfg.phases[0] = phase
        c = emo.Context(cfg)
        if eps is None:
            return reset(c)
        ld = copy.deepcopy(w.loader)
        ld.pert_eps[0] = eps
        c.load_maxwellian(ld)
        c.init_constant()
        c.interpolate()
        c.synchronize()
        return c

    def forward(amp=None, phase=None, eps=None):
        c = ready(amp, phase, eps)
        if not fp32:
            c.sort()
        c.synchronize()
        a = time.perf_counter()
        c.step(n, 0.0, w.dt)
        c.synchronize()
        el = time.perf_counter() - a
        d = c.diag()
        c.close()
        return d[4] + d[5], el

    forward()  # warm-up (first launches, allocations)
    # the first gradient call on a context allocates its checkpoint and tape
    # pool, which later calls reuse; then three forward runs and three
    # gradient calls from the same initial state, alternating: the medians
    c = ready()
    c.adjoint_field_energy(n, 0.0, w.dt, k, particles=False)
    t_f, t_g = [], []
    for _ in range(3):
        f0, el = forward()
        t_f.append(el)
        reset(c)
        a = time.perf_counter()
        out = c.adjoint_field_energy(n, 0.0, w.dt, k, particles=False)
        c.synchronize()
        t_g.append(time.perf_counter() - a)
    c.close()
    t_fwd, t_adj = sorted(t_f)[1], sorted(t_g)[1]
    amp = float(w.cfg.peakamps[0])
    if fp32:
        fd = fp64_leg["dF_damp_central_diff"] if fp64_leg else None
    else:
        h = 1e-9 * amp
        fd = (forward(amp + h)[0] - forward(amp - h)[0]) / (2 * h)
    res = {
        "workload": f"c4: 64x32x32, x external+LightBC pulse (waist 10), reflecting walls, 8 ppc/species "
                    f"({w.nparticles} particles), relativistic, {'FP32 chosen' if fp32 else 'FP64 plain'} form, "
                    f"{n} steps, checkpoint every {k}",
        "forward_s": round(t_fwd, 4),
        "gradient_s": round(t_adj, 4),
        "forward_runs_s": [round(t, 4) for t in t_f],
        "gradient_runs_s": [round(t, 4) for t in t_g],
        "adjoint_s": round(t_adj - t_fwd, 4),
        "adjoint_over_forward": round((t_adj - t_fwd) / t_fwd, 3),
        "gradient_over_forward": round(t_adj / t_fwd, 3),
        "energy_matches_forward": bool(out["energy"] == f0),
        "dF_damp": float(out["d_amp"][0]),
        ("dF_damp_fp64_central_diff" if fp32 else "dF_damp_central_diff"): float(fd) if fd else None,
        "fd_rel_err": float(abs(out["d_amp"][0] - fd) / abs(fd)) if fd else None,
    }
    if not fp32 and n == w.steps:
        res["fd_check"] = c4_fd_check(w, ready, forward)
    if fp32 and fp64_leg:
        res["dF_damp_rel_diff_vs_fp64_adjoint"] = float(abs(out["d_amp"][0] - fp64_leg["dF_damp"]) /
                                                        abs(fp64_leg["dF_damp"]))
    print(f"[bench] adjoint c4 {'fp32' if fp32 else 'fp64'}: forward {t_fwd:.3f}s, gradient {t_adj:.3f}s, adjoint/forward "
          f"{(t_adj - t_fwd) / t_fwd:.2f}, gradient/forward {t_adj / t_fwd:.2f}, fd rel err {res['fd_rel_err']}",
          file=sys.stderr, flush=True)
    return res


def c4_fd_check(w, ready, forward):
    """The c4 gradient (amplitude, phase per radian, eps) against central
    differences of the FP64 forward at the configuration's step h = 1e-4 and
    at a step small enough that no particle changes cell (1e-9 relative for
    the amplitude, 1e-9 rad, 1e-10 for eps).  At h = 1e-4 particles cross
    cell faces, where the derivative of the linear weights jumps, so the
    quotient there is not the derivative the adjoint computes; both are
    reported.  `resolution` is ulp(F)/(2h) relative to the derivative: the
    finest the small-step quotient can resolve.  Untimed."""
    import numpy as np

    from apc_524_project_amd._lib import loader_dx_deps

    c = ready()
    out = c.adjoint_field_energy(w.steps, 0.0, w.dt, 20)
    p = c.particles(fields=False)
    c.close()
    species = np.empty(len(p["id"]), np.int32)
    species[p["id"]] = p["species"]
    d_eps = float((out["xbar0"][:, 0] * loader_dx_deps(w.loader, w.cfg))[species == 0].sum())
    F = out["energy"]
    amp = float(w.cfg.peakamps[0])
    ph0 = float(w.cfg.phases[0])
    eps0 = float(w.loader.pert_eps[0])
    adj = {"amp": float(out["d_amp"][0]), "phase_per_rad": float(out["d_phase"][0]), "eps": d_eps}
    rows = {}
    for name, small in (("amp", 1e-9 * amp), ("phase_per_rad", 1e-9), ("eps", 1e-10)):
        r = {"adjoint": adj[name]}
        for tag, h in (("h_small", small), ("h_1e-4", 1e-4)):
            if name == "amp":
                fd = (forward(amp=amp + h)[0] - forward(amp=amp - h)[0]) / (2 * h)
            elif name == "phase_per_rad":
                hd = h * 180.0 / math.pi  # phases are given in degrees
                fd = (forward(phase=ph0 + hd)[0] - forward(phase=ph0 - hd)[0]) / (2 * h)
            else:
                fd = (forward(eps=eps0 + h)[0] - forward(eps=eps0 - h)[0]) / (2 * h)
            r[tag] = {"h": h, "central_diff": float(fd), "rel_diff": float(abs(fd - adj[name]) / abs(fd))}
        r["h_small"]["resolution"] = float(math.ulp(F) / (2 * small) / abs(adj[name]))
        rows[name] = r
        print(f"[bench] c4 fd {name}: adjoint {adj[name]:.8e}, small h {r['h_small']['central_diff']:.8e} "
              f"(rel {r['h_small']['rel_diff']:.2e}), h=1e-4 {r['h_1e-4']['central_diff']:.8e} "
              f"(rel {r['h_1e-4']['rel_diff']:.2e})", file=sys.stderr, flush=True)
    return rows


def run_adjoint_c5s(steps: int, k: int):
    """c5's physics on the 128x32x32 box (16,777,216 particles: two electron
    beams at +-0.2c, sigma 0.01c, 64 ppc each, neutralising background,
    Poisson-initialised), `steps` steps with a checkpoint every k: the FP32
    gradient call against its own forward run and against the FP64 adjoint
    from the same loaded particles (the FP64 adjoint is checked against
    central differences at c4 and in the tests).  Both hold the initial
    fields fixed.  Reported gradient: dF/d(a uniform shift of beam 1's v_x),
    the sum of its particles' x-velocity adjoints (stored gamma held fixed)."""
    import numpy as np

    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN, FP64_PLAIN

    out = {"workload": f"c5s: 128x32x32 periodic, two electron beams +-0.2c (sigma 0.01c, 64 ppc each), "
                       f"16,777,216 particles, neutralising background, Poisson init; {steps} steps, "
                       f"checkpoint every {k}"}
    grads = {}
    for prec, tag in ((FP64_PLAIN, "fp64"), (FP32_CHOSEN, "fp32")):
        w = emo.make_workload("c5s", precision=prec)

        def reset(c):
            load(c, w)
            init_fields(c, w)
            if prec == FP64_PLAIN:
                c.interpolate()
            c.synchronize()
            return c

        def ready():
            return reset(emo.Context(w.cfg))

        c = ready()
        if prec == FP64_PLAIN:
            c.sort()
        c.step(5, 0.0, w.dt)  # warm-up (first launches, allocations)
        c.close()
        # the gradient calls are timed without the per-particle outputs
        # (by-id copies to the host of ~1 GB at this size) on one context
        # (its first call allocates the checkpoint and tape pool, which later
        # calls reuse): three forward runs and three gradient calls from the
        # same initial state, alternating, and their medians; the gradient
        # itself comes from a last call
        cg = ready()
        cg.adjoint_field_energy(steps, 0.0, w.dt, k, particles=False)
        t_f, t_g = [], []
        for _ in range(3):
            c = ready()
            if prec == FP64_PLAIN:
                c.sort()
            c.synchronize()
            a = time.perf_counter()
            c.step(steps, 0.0, w.dt)
            c.synchronize()
            t_f.append(time.perf_counter() - a)
            d = c.diag()
            f_fwd = d[4] + d[5]
            c.close()
            reset(cg)
            a = time.perf_counter()
            cg.adjoint_field_energy(steps, 0.0, w.dt, k, particles=False)
            cg.synchronize()
            t_g.append(time.perf_counter() - a)
        t_fwd, t_grad = sorted(t_f)[1], sorted(t_g)[1]
        c = cg
        reset(c)
        g = c.adjoint_field_energy(steps, 0.0, w.dt, k)
        p = c.particles(fields=False)
        c.close()
        sp = np.empty(len(p["id"]), np.int32)
        sp[p["id"]] = p["species"]
        grads[tag] = g
        out[tag] = {"forward_s": round(t_fwd, 4), "gradient_s": round(t_grad, 4),
                    "forward_runs_s": [round(t, 4) for t in t_f], "gradient_runs_s": [round(t, 4) for t in t_g],
                    "adjoint_over_forward": round((t_grad - t_fwd) / t_fwd, 3),
                    "gradient_over_forward": round(t_grad / t_fwd, 3),
                    "energy_matches_forward": bool(g["energy"] == f_fwd), "energy": g["energy"],
                    "dF_dvx_beam1": float(g["vbar0"][sp == 0, 0].sum())}
        out[tag]["dF_dvx_beam1_over_F"] = out[tag]["dF_dvx_beam1"] / g["energy"]
        out[tag]["energy_per_cell"] = g["energy"] / w.ncells
        out[tag]["dF_dvx_beam1_per_cell"] = out[tag]["dF_dvx_beam1"] / w.ncells
        print(f"[bench] adjoint c5s {tag}: forward {t_fwd:.3f}s, gradient {t_grad:.3f}s, adjoint/forward "
              f"{(t_grad - t_fwd) / t_fwd:.2f}, dF/dvx(beam 1) {out[tag]['dF_dvx_beam1']:.6e}", file=sys.stderr,
              flush=True)
    a32, a64 = out["fp32"]["dF_dvx_beam1"], out["fp64"]["dF_dvx_beam1"]
    out["fp32_vs_fp64"] = {
        "energy_rel": abs(out["fp32"]["energy"] - out["fp64"]["energy"]) / abs(out["fp64"]["energy"]),
        "dF_dvx_beam1_rel": abs(a32 - a64) / abs(a64) if a64 else None,
        "xbar0_rel_l2": float(np.linalg.norm(grads["fp32"]["xbar0"] - grads["fp64"]["xbar0"]) /
                              np.linalg.norm(grads["fp64"]["xbar0"])),
        "vbar0_rel_l2": float(np.linalg.norm(grads["fp32"]["vbar0"] - grads["fp64"]["vbar0"]) /
                              np.linalg.norm(grads["fp64"]["vbar0"])),
    }
    print(f"[bench] adjoint c5s fp32 vs fp64: {out['fp32_vs_fp64']}", file=sys.stderr, flush=True)
    return out


def run_adjoint_c5(steps: int, k: int):
    """Workload c5 at full size (1.07e9 particles) in the FP32 chosen form:
    `steps` forward steps timed alone, then the gradient call from the same
    initial state (reloaded) with checkpoint_every = k.  The K-step segments
    and their tape do not fit beside the state in device memory, so the call
    takes binomial checkpointing with the snapshots that fit; the schedule it
    ran (snapshots, forward steps, restores) is reported with the times.  The
    time of the call includes allocating its snapshot pool (the first call on
    the context).  Reported gradient: dF/d(a uniform shift of beam 1's v_x),
    the sum of its particles' x-velocity adjoints (stored gamma held fixed),
    reduced on the device (emo_adjoint_stats); (1/F) dF/dv is intensive and
    is set beside c5s's (the same physics on 1/64 of the volume)."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN

    w = emo.make_workload("c5", precision=FP32_CHOSEN)
    ctx = emo.Context(w.cfg)

    def reset():
        ctx.load_uniform(w.loader)
        init_fields(ctx, w)
        ctx.synchronize()

    reset()
    ctx.step(2, 0.0, w.dt)  # first launches
    reset()
    a = time.perf_counter()
    ctx.step(steps, 0.0, w.dt)
    ctx.synchronize()
    t_fwd = time.perf_counter() - a
    d = ctx.diag()
    f_fwd = d[4] + d[5]
    print(f"[bench] adjoint c5: forward {steps} steps {t_fwd:.2f}s; gradient call running", file=sys.stderr,
          flush=True)
    reset()
    a = time.perf_counter()
    g = ctx.adjoint_field_energy(steps, 0.0, w.dt, k, particles=False)
    ctx.synchronize()
    t_grad = time.perf_counter() - a
    st = ctx.adjoint_stats()
    ctx.close()
    dvx = float(st["sums"][0, 3])
    out = {"workload": f"c5: 512x128x128 periodic, two electron beams +-0.2c (sigma 0.01c), 64 ppc per beam "
                       f"({w.nparticles} particles), FP32 chosen form, {steps} steps",
           "forward_s": round(t_fwd, 3), "gradient_s": round(t_grad, 3),
           "gradient_over_forward": round(t_grad / t_fwd, 3),
           "adjoint_over_forward": round((t_grad - t_fwd) / t_fwd, 3),
           "checkpointing": st["checkpointing"], "snapshots": st["snapshots"],
           "forward_steps_run": st["forward_steps"], "restores": st["restores"],
           "energy": g["energy"], "energy_matches_forward": bool(g["energy"] == f_fwd),
           "dF_dvx_beam1": dvx, "dF_dvx_beam1_over_F": dvx / g["energy"],
           "energy_per_cell": g["energy"] / w.ncells, "dF_dvx_beam1_per_cell": dvx / w.ncells}
    print(f"[bench] adjoint c5: gradient {t_grad:.2f}s ({out['gradient_over_forward']}x forward), "
          f"{st['checkpointing']} with {st['snapshots']} snapshots, {st['forward_steps']} forward steps, "
          f"{st['restores']} restores; dF/dvx(beam 1) {dvx:.6e}", file=sys.stderr, flush=True)
    return out


ACCURACY_CHECKPOINTS = {
    "c1": (10, 100),
    "c2": (10, 100, 1000),
    "c3": (10, 100, 1000, 2000),
    "c4": (10, 100, 400),
    "c5s": (10, 100, 1000),
}
C5_POISSON_ITER_MAX = 499  # at most 500 Jacobi sweeps per solve (reference convergence rule)


def init_fields(ctx, w):
    """The workload's field initialisation: Poisson (c1: 50 sweeps; c5: the
    reference's convergence rule, at most 500 sweeps) or constant."""
    if not w.poisson:
        ctx.init_constant()
        return None
    if w.poisson_sweeps:
        sw, res = ctx.init_poisson(iter_max=w.poisson_sweeps - 1, tol=0.0)
    else:
        sw, res = ctx.init_poisson(iter_max=C5_POISSON_ITER_MAX, tol=-1.0)
    return {"sweeps": [int(x) for x in sw], "resid_max": [float(x) for x in res[::2]],
            "resid_l2": [float(x) for x in res[1::2]]}


def load(ctx, w):
    if w.device_loader:
        ctx.load_uniform(w.loader)
    else:
        ctx.load_maxwellian(w.loader)


def run_accuracy(name, checkpoints, control=False, scale: float = 1.0):
    """Workload `name` in both forms from the same loaded particles: relative
    L2 error of E, B, J (physical points) of the FP32 chosen form against the
    FP64 plain form (bit-exact with the oracle) after each checkpoint step
    count, of the field energies, of the k = 2 pi / L_x component of the y-z
    mean of E_x, and of the particles' positions and velocities at the last
    checkpoint.

    With control, the same errors are given for an FP64 plain run started
    from the FP32 form's initial state (its particles and fields as the FP32
    store holds them): it isolates how fast the loop itself amplifies a
    perturbation of FP32 representation size, independent of FP32
    arithmetic during the steps."""
    import numpy as np

    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN, FP64_PLAIN
    from apc_524_project_amd import _lib as L_

    ctxs = {}
    poisson = {}
    for form, prec in (("fp64", FP64_PLAIN), ("fp32", FP32_CHOSEN)):
        w = emo.make_workload(name, precision=prec, scale=scale)
        c = emo.Context(w.cfg)
        load(c, w)
        poisson[form] = init_fields(c, w)
        ctxs[form] = c
    if control:
        # the FP32 form's initial state (particles and fields, as the FP32
        # store holds them) stepped in FP64
        p32 = ctxs["fp32"].particles(fields=False)
        w64 = emo.make_workload(name, precision=FP64_PLAIN, scale=scale)
        c = emo.Context(w64.cfg)
        c.load_particles(p32["x"], p32["v"], p32["gamma"], p32["species"], p32["id"])
        c.init_constant()
        for fid in (L_.EX, L_.EY, L_.EZ, L_.BX, L_.BY, L_.BZ, L_.BX_TM1, L_.BY_TM1, L_.BZ_TM1):
            c.set_field(fid, ctxs["fp32"].field(fid))
        c.interpolate()
        ctxs["fp64_from_fp32_state"] = c
        del p32
    n = w.cfg.nCell
    inner = (slice(1, n[0] + 1), slice(1, n[1] + 1), slice(1, n[2] + 1))

    def group(c, fids):
        return np.concatenate([c.field(f)[inner].ravel() for f in fids])

    def rel(a, b):
        nb = float(np.linalg.norm(b))
        return float(np.linalg.norm(a - b) / nb) if nb > 0 else None

    def mode(c):
        ex = c.field(L_.EX)[inner].mean(axis=(1, 2))
        return np.fft.fft(ex)[1]

    groups = {"E": (L_.EX, L_.EY, L_.EZ), "B": (L_.BX, L_.BY, L_.BZ), "J": (L_.JX, L_.JY, L_.JZ)}
    rows = {k: [] for k in ctxs if k != "fp64"}
    done = 0
    for cp in checkpoints:
        for c in ctxs.values():
            c.step(cp - done, done * w.dt, w.dt)
            c.synchronize()
        done = cp
        ref = {g: group(ctxs["fp64"], fids) for g, fids in groups.items()}
        d64 = ctxs["fp64"].diag()
        m64 = mode(ctxs["fp64"])
        for form in rows:
            row = {"steps": cp}
            for g, fids in groups.items():
                row[g] = rel(group(ctxs[form], fids), ref[g])
            d = ctxs[form].diag()
            row["field_energy"] = rel(np.array(d[4:6]), np.array(d64[4:6]))
            row["ex_mode_k1"] = float(abs(mode(ctxs[form]) - m64) / abs(m64)) if abs(m64) > 0 else None
            rows[form].append(row)
            print(f"[bench] accuracy {name} {form} vs fp64 after {cp} steps: " +
                  ", ".join(f"{k} {v:.3g}" for k, v in row.items() if k != "steps" and v is not None),
                  file=sys.stderr, flush=True)
    parts = {}
    for form, c in ctxs.items():
        p = c.particles(fields=False)
        o = np.argsort(p["id"])
        parts[form] = (p["x"][o], p["v"][o])
        c.close()
    L = np.array(w.cfg.Lxyz)
    res = {"workload": f"{name} ({w.nparticles} particles, {n[0]}x{n[1]}x{n[2]} cells): FP32 chosen form vs FP64 "
                       f"plain form from the same loaded particles",
           "dt": w.dt}
    if poisson["fp64"]:
        res["poisson_init"] = poisson
    for form in rows:
        dx = parts[form][0] - parts["fp64"][0]
        dx -= L * np.round(dx / L)  # periodic images
        res[form] = {
            "rel_l2": rows[form],
            "positions_rel_l2": float(np.linalg.norm(dx) / np.linalg.norm(parts["fp64"][0])),
            "velocities_rel_l2": rel(parts[form][1], parts["fp64"][1]),
        }
    if "fp64_from_fp32_state" in rows:
        # FP32 error / control error per checkpoint: about 1 where the loop's
        # amplification of the initial rounding explains the FP32 error,
        # above 1 where FP32 arithmetic adds its own
        ratios = []
        for r32, rc in zip(rows["fp32"], rows["fp64_from_fp32_state"]):
            row = {"steps": r32["steps"]}
            for k in ("E", "B", "J", "field_energy", "ex_mode_k1"):
                row[k] = (r32[k] / rc[k]) if (r32[k] is not None and rc[k]) else None
            ratios.append(row)
        res["fp32_over_control"] = ratios
        print(f"[bench] accuracy {name} fp32/control: " +
              "; ".join(f"{r['steps']}: E {r['E']:.3g} J {r['J']:.3g}" for r in ratios
                        if r["E"] is not None and r["J"] is not None), file=sys.stderr, flush=True)
    return res


def run_c5(steps: int, warmup: int):
    """Workload c5 at full size in the FP32 chosen form: 1.07e9 particles from
    the device loader, FP32 Poisson initialisation with FP64 residual norms,
    then `steps` timed steps; FP64 diagnostics before and after."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN

    w = emo.make_workload("c5", precision=FP32_CHOSEN)
    t0 = time.perf_counter()
    ctx = emo.Context(w.cfg)
    ctx.load_uniform(w.loader)
    ctx.synchronize()
    t_load = time.perf_counter() - t0
    ctx.timing(True)
    ctx.timing_select("jacobi_f32")
    ctx.timing_reset()
    a = time.perf_counter()
    pois = init_fields(ctx, w)
    ctx.synchronize()
    t_pois = time.perf_counter() - a
    jac_ms, jac_n = ctx.kernel_time("jacobi_f32")
    ctx.timing(False)
    ctx.timing_select(None)
    d0 = ctx.diag()
    t = 0.0
    ctx.step(warmup, t, w.dt)
    t += warmup * w.dt
    ctx.synchronize()
    a = time.perf_counter()
    ctx.step(steps, t, w.dt)
    ctx.synchronize()
    el = time.perf_counter() - a
    d1 = ctx.diag()
    ctx.close()
    rate = w.nparticles * steps / el
    print(f"[bench] c5 fp32: {w.nparticles} particles, {steps} steps in {el:.3f}s -> {rate:.4g} particle-steps/s; "
          f"load {t_load:.1f}s, poisson {t_pois:.2f}s ({pois['sweeps']} sweeps)", file=sys.stderr, flush=True)
    keys = ("pMx", "pMy", "pMz", "kinetic", "B_energy", "E_energy", "total_energy", "charge")
    return {"workload": "c5: 512x128x128 periodic, two electron beams +-0.2c (sigma 0.01c), 64 ppc per beam "
                        f"({w.nparticles} particles), fixed neutralising ion background, relativistic Boris, "
                        "device loader (seed -5)",
            "particle_steps_per_s": round(rate, 1), "ms_per_step": round(1e3 * el / steps, 4), "steps": steps,
            "warmup": warmup, "load_s": round(t_load, 2), "poisson_s": round(t_pois, 3), "poisson_init": pois,
            "jacobi_sweep": jacobi_roof(jac_ms, jac_n, w.ncells),
            "diag_initial": dict(zip(keys, map(float, d0))), "diag_final": dict(zip(keys, map(float, d1)))}


def run_fields_c5(yee_calls: int = 50):
    """The field kernels at c5 size (512x128x128): the FP32 Poisson
    initialisation of the c5 load (Jacobi sweeps timed with events around
    their launches) and the Yee update alone in both precisions."""
    import apc_524_project_amd as emo
    from apc_524_project_amd import FP32_CHOSEN

    w = emo.make_workload("c5", precision=FP32_CHOSEN)
    ctx = emo.Context(w.cfg)
    ctx.load_uniform(w.loader)
    ctx.synchronize()
    ctx.timing(True)
    ctx.timing_select("jacobi_f32")
    ctx.timing_reset()
    a = time.perf_counter()
    pois = init_fields(ctx, w)
    ctx.synchronize()
    t_pois = time.perf_counter() - a
    jac_ms, jac_n = ctx.kernel_time("jacobi_f32")
    ctx.timing(False)
    ctx.timing_select(None)
    ctx.close()
    out = {"poisson_s": round(t_pois, 3), "poisson_init": pois, "jacobi_sweep": jacobi_roof(jac_ms, jac_n, w.ncells),
           "yee_fp32": run_yee("fp32", "c5", yee_calls), "yee_fp64": run_yee("fp64", "c5", yee_calls)}
    print(f"[bench] fields c5: jacobi {out['jacobi_sweep']}, yee fp32 {out['yee_fp32']}, yee fp64 {out['yee_fp64']}",
          file=sys.stderr, flush=True)
    return out


def host_cpu():
    """The host CPU's model name, family/model/stepping and clock from
    /proc/cpuinfo, and the logical CPUs this process may use."""
    info = {}
    try:
        for line in open("/proc/cpuinfo"):
            if ":" in line:
                k, v = (x.strip() for x in line.split(":", 1))
                if k in ("model name", "cpu family", "model", "stepping", "cpu MHz") and k not in info:
                    info[k] = v
            if not line.strip() and info:
                break
    except OSError:
        pass
    try:
        ncpu = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        ncpu = os.cpu_count()
    return (f"{info.get('model name', '?')} (family {info.get('cpu family', '?')}, model {info.get('model', '?')}, "
            f"stepping {info.get('stepping', '?')}, {info.get('cpu MHz', '?')} MHz), {ncpu} logical CPUs available")


def jacobi_roof(ms, n, ncells):
    """One FP32 Jacobi sweep (HIP events around its launches): 12 B per
    physical point (the iterate and the source read, the new iterate
    written)."""
    if not n:
        return None
    avg = ms / n / 1e3
    gbs = 12 * ncells / avg / 1e9
    return {"sweeps": n, "avg_ms": round(ms / n, 5), "achieved_GBs": round(gbs, 1),
            "frac": round(gbs / HBM_PEAK_GBS, 4)}


def cpu_baseline(sample_cells: int = 64, sample_steps: int = 4, runs: int = 3):
    """The reference's serial loop (oracle/_ref) on one core, c2 physics on a
    sample_cells^3 box at the same density (32 ppc/species).  The loop's
    first step (which sorts the freshly loaded particles: the reference sorts
    when ti % 100 == 0) is timed apart; then `runs` runs of `sample_steps`
    steps (no sort among them), whose median gives the rate.  The rate with
    the sort amortised over its 100-step period is given beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_ctypes as oc

    n = sample_cells
    wpe2 = 4 * math.pi * oc.UNIT_RHOJ * oc.UNIT_ACC
    case = oc.Case(n=(n, n, n), L=(float(n),) * 3, nparticles=2 * 32 * n ** 3, relativity=1,
                   temp=((0.1 / oc.UNIT_VTH) ** 2, (0.1 / oc.UNIT_VTH) ** 2), dens_phys=0.01 / wpe2,
                   zero_avg=0)
    cpu = host_cpu()
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    dt = 0.5 / math.sqrt(3.0)
    runs_s = []
    if oc.ref_available():
        r = oc.Reference(case, nstep_sort=100)
        r.initialize()
        first = r.time_steps_from(0, 1, dt)
        for k in range(runs):
            runs_s.append(r.time_steps_from(1 + k * sample_steps, sample_steps, dt))
        kind = "reference"
    else:
        o = oc.Oracle(case)
        o.load()
        o.initialize()
        t = 0.0
        a = time.perf_counter()
        o.step(t, dt)
        first = time.perf_counter() - a
        t += dt
        for _ in range(runs):
            a = time.perf_counter()
            for _ in range(sample_steps):
                o.step(t, dt)
                t += dt
            runs_s.append(time.perf_counter() - a)
        kind = "port"
    med = sorted(runs_s)[len(runs_s) // 2] / sample_steps
    sort_s = max(first - med, 0.0)
    return {"value": case.nparticles / med, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "value_with_sort_amortised": case.nparticles / (med + sort_s / 100.0),
            "sample": f"c2 physics on a {n}^3 box{' (the full c2 box)' if n == 64 else ''} "
                      f"(32 ppc/species, {case.nparticles} particles), reference serial loop (nstep_sort=100): "
                      f"first step (with the sort of the loaded particles) {first:.2f}s, then {runs} runs of "
                      f"{sample_steps} steps: {', '.join(f'{x:.2f}s' for x in runs_s)} (median used); "
                      f"1 thread pinned to one core; host CPU: {cpu}"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000, help="timed steps (1000: the c2 configuration's step count)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--mode", default="auto", choices=["auto", "fp64", "fp32", "both"])
    ap.add_argument("--scale", type=float, default=1.0, help="particle-count scale (1 = full c2)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true", help="only the c4 adjoint leg (for profiling)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run an elementwise copy of 256 MiB (a counter reference of known bytes)")
    ap.add_argument("--adjoint-steps", type=int, default=0, help="c4 steps for the adjoint leg (0: 400)")
    ap.add_argument("--checkpoint-every", type=int, default=20)
    ap.add_argument("--accuracy", default="all",
                    help="configs of the FP32-vs-FP64 error leg: all, none, or a comma list of c1,c2,c3,c4,c5s")
    ap.add_argument("--c5-steps", type=int, default=20, help="timed c5 steps (0: skip the c5 leg)")
    ap.add_argument("--c5s-adjoint-steps", type=int, default=200,
                    help="steps of the c5s FP32-vs-FP64 adjoint leg (0: skip it)")
    ap.add_argument("--c5s-adjoint-only", action="store_true", help="only the c5s adjoint leg")
    ap.add_argument("--c5-adjoint-steps", type=int, default=200,
                    help="steps of the full-size c5 FP32 adjoint leg (0: skip it)")
    ap.add_argument("--c5-adjoint-only", action="store_true", help="only the c5 adjoint leg")
    ap.add_argument("--accuracy-only", action="store_true", help="only the FP32-vs-FP64 error leg")
    ap.add_argument("--fields-only", action="store_true", help="only the c5 field kernels (Jacobi, Yee)")
    ap.add_argument("--cpu-cells", type=int, default=64)
    ap.add_argument("--cpu-steps", type=int, default=4, help="steps per timed run of the CPU baseline (3 runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the multi-GPU decomposition is a design, "
              "not part of this build (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    import apc_524_project_amd as emo

    if args.calibrate_copy:
        import torch

        a = torch.ones(64 << 20, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        del a, b
    if args.adjoint_only:
        a64 = run_adjoint_c4(args.adjoint_steps or None, args.checkpoint_every)
        a32 = run_adjoint_c4(args.adjoint_steps or None, args.checkpoint_every, True, a64)
        print(json.dumps({"adjoint": a64, "adjoint_fp32": a32}), flush=True)
        return
    if args.c5s_adjoint_only:
        print(json.dumps({"adjoint_c5s": run_adjoint_c5s(args.c5s_adjoint_steps or 200, args.checkpoint_every)}),
              flush=True)
        return
    if args.fields_only:
        print(json.dumps({"fields_c5": run_fields_c5()}), flush=True)
        return
    if args.accuracy_only:
        names = list(ACCURACY_CHECKPOINTS) if args.accuracy in ("all", "none") else args.accuracy.split(",")
        print(json.dumps({"fp32_accuracy": {nm: run_accuracy(nm, ACCURACY_CHECKPOINTS[nm], control=True,
                                                             scale=args.scale) for nm in names}}), flush=True)
        return
    if args.c5_adjoint_only:
        print(json.dumps({"adjoint_c5": run_adjoint_c5(args.c5_adjoint_steps or 200, args.checkpoint_every)}),
              flush=True)
        return

    have_fp32 = True
    try:
        w = emo.make_workload("c2", precision=emo.FP32_CHOSEN, scale=0.001)
        c = emo.Context(w.cfg)
        c.close()
    except emo.EmoError:
        have_fp32 = False
    mode = args.mode
    if mode == "auto":
        mode = "both" if have_fp32 else "fp64"
    results = {}
    for m in (["fp64", "fp32"] if mode == "both" else [mode]):
        results[m] = run_gpu(m, args.steps, args.warmup, args.scale)
    head = "fp32" if "fp32" in results else "fp64"
    r = results[head]
    w = r["workload"]
    out = {
        "metric": METRIC,
        "value": round(r["rate"], 1),
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if head == "fp32" else "f64",
        "data": "synthetic",
        "config": {
            "workload": "c2: 64^3 periodic Yee grid, dx=1, dt=0.5 CFL, 32 ppc/species "
                        "(16,777,216 particles), relativistic Boris, sigma_e=0.1c, "
                        "0.01 cos(2 pi x/64) electron density perturbation, seed -2",
            "form": "FP32 chosen form" if head == "fp32" else "FP64 plain form",
            "particles": w.nparticles,
            "cells": w.ncells,
            "scale": args.scale,
        },
        "roofline": roofline_for(r),
    }
    if head == "fp32":
        out["config"]["rho"] = ("the FP32 step deposits J only; rho, which no phase of the step reads, is "
                                "deposited from the current particles when it is read (emo_get_field)")
    for m, rr in results.items():
        out[f"{m}_particle_steps_per_s"] = round(rr["rate"], 1)
        out[f"{m}_kernel_ms_per_step"] = {k: round(v[0] / v[1], 5) for k, v in sorted(rr["ktimes"].items())}
        out[f"{m}_ms_per_step_events_every_launch"] = round(rr["ms_per_step_all_events"], 4)
        out[f"{m}_yee_c2"] = run_yee(m, "c2", 200, args.scale)
    if "fp64" in results:
        out["fp64_roofline"] = roofline_for(results["fp64"])
    if "fp32" in results and "fp64" in results:
        out["fp32_over_fp64"] = round(results["fp32"]["rate"] / results["fp64"]["rate"], 3)
        rate_pp, ms_pp = run_fp64_per_phase(max(args.steps // 5, 5), 2, args.scale)
        out["fp64_per_phase_particle_steps_per_s"] = round(rate_pp, 1)
        out["fp64_per_phase_ms_per_step"] = round(ms_pp, 4)
        out["fp32_over_fp64_per_phase"] = round(results["fp32"]["rate"] / rate_pp, 3)
    if args.accuracy != "none" and "fp32" in results:
        names = list(ACCURACY_CHECKPOINTS) if args.accuracy == "all" else args.accuracy.split(",")
        out["fp32_accuracy"] = {nm: run_accuracy(nm, ACCURACY_CHECKPOINTS[nm], control=True, scale=args.scale)
                                for nm in names}
    if args.c5_steps > 0 and "fp32" in results:
        out["c5"] = run_c5(args.c5_steps, 2)
        out["c5"]["yee_fp32"] = run_yee("fp32", "c5", 50)
        out["c5"]["yee_fp64"] = run_yee("fp64", "c5", 50)
    if not args.no_adjoint:
        out["adjoint"] = run_adjoint_c4(args.adjoint_steps or None, args.checkpoint_every)
        if "fp32" in results:
            out["adjoint_fp32"] = run_adjoint_c4(args.adjoint_steps or None, args.checkpoint_every, True,
                                                 out["adjoint"])
            if args.c5s_adjoint_steps > 0:
                out["adjoint_c5s"] = run_adjoint_c5s(args.c5s_adjoint_steps, args.checkpoint_every)
            if args.c5_adjoint_steps > 0:
                a5 = run_adjoint_c5(args.c5_adjoint_steps, args.checkpoint_every)
                if "adjoint_c5s" in out and args.c5s_adjoint_steps == args.c5_adjoint_steps:
                    # per cell (and per unit of F) the figures are intensive: c5 beside
                    # c5s's FP64 and FP32 adjoints (c5's longer box holds longer modes,
                    # so they agree to the growth those modes add, not exactly)
                    for tag in ("fp64", "fp32"):
                        for key in ("dF_dvx_beam1_per_cell", "energy_per_cell", "dF_dvx_beam1_over_F"):
                            a5[f"c5s_{tag}_{key}"] = out["adjoint_c5s"][tag][key]
                out["adjoint_c5"] = a5
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(args.cpu_cells, args.cpu_steps)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
