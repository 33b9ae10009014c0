"""Binomial checkpointing of the FP32 chosen form's gradient call
(emo_adjoint_field_energy with checkpoint_every < 0: dense snapshots, particle
adjoints by id, no tape), through the C ABI.

The backward step is the same arithmetic on the same state whichever way the
state was reached, and every reduction has a fixed order, so the binomial
form gives the same bits as the K-step segments for any number of snapshots:
the energy, the wave gradients and every particle's initial-state gradient
are compared for equality.  Also: the final state the call leaves, the
per-species sums of emo_adjoint_stats against the sums of the by-id outputs
(both forms and both precisions), the steps and restores the schedule took,
and the edge cases of 0 and 1 steps.
"""
import math

import numpy as np
import pytest

from apc_524_project_amd._lib import FP32_CHOSEN, FP64_PLAIN, Context
from emo_case import emo_config, emo_loader
from oracle_ctypes import EXTERNAL, PERIODIC, REFLECTING, Case, Wave

pytestmark = pytest.mark.gpu

NSTEPS = 13


def lightbc_case(relativity=1):
    w = Wave(side=-1, pol=2, amp=0.3, omega=0.39, phase=30.0, inv_width=0.08, delay=3.0)
    return Case(n=(12, 6, 6), L=(12.0, 6.0, 6.0), nparticles=2 * 2 * 432, temp=(501.8, 501.8),
                dens_phys=4.355e9, relativity=relativity, waves=[w],
                fbc=(EXTERNAL, EXTERNAL, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pbc=(REFLECTING, REFLECTING, PERIODIC, PERIODIC, PERIODIC, PERIODIC),
                pert_eps=(0.05, 0.0), pert_k=(2 * math.pi / 12.0, 0.0))


def periodic_case():
    return Case(n=(8, 6, 5), L=(8.0, 6.0, 5.0), nparticles=2 * 8 * 240, temp=(5000.0, 5000.0),
                dens_phys=3.0e9, relativity=1, B0=(0.0, 0.0, 0.02), E0=(0.003, 0.0, 0.0),
                drift=((0.2, 0.0, 0.0), (-0.2, 0.0, 0.0)))


CASES = {"lightbc_rel": lightbc_case(1), "lightbc_boris": lightbc_case(0), "periodic_drift": periodic_case()}


def setup(case, precision=FP32_CHOSEN):
    ctx = Context(emo_config(case, precision=precision))
    ctx.load_maxwellian(emo_loader(case))
    ctx.init_constant()
    if precision == FP64_PLAIN:
        ctx.interpolate()
    return ctx


def run(case, k, nsteps=NSTEPS, precision=FP32_CHOSEN):
    ctx = setup(case, precision)
    print(f"[binomial] k={k} nsteps={nsteps}", flush=True)
    out = ctx.adjoint_field_energy(nsteps, 0.0, case.dt_cfl(), k)
    out["stats"] = ctx.adjoint_stats()
    p = ctx.particles(fields=False)
    order = np.argsort(p["id"])
    out["final"] = {key: p[key][order] for key in ("x", "v", "gamma", "id")}
    ctx.close()
    return out


def assert_same_gradient(a, b):
    assert a["energy"] == b["energy"]
    assert np.array_equal(a["d_amp"], b["d_amp"]) and np.array_equal(a["d_phase"], b["d_phase"])
    for key in ("xbar0", "vbar0", "gbar0"):
        assert np.array_equal(a[key], b[key]), key


@pytest.fixture(scope="module", params=sorted(CASES))
def segments(request):
    case = CASES[request.param]
    return case, run(case, 4)


@pytest.mark.parametrize("snapshots", [1, 2, 3])
def test_binomial_equals_segments(segments, snapshots):
    """-k asks for at most k snapshots besides the initial one.  With one,
    the part after the first split point has no free snapshot left, so each
    of its backward steps re-runs from that point (the quadratic form)."""
    case, ref = segments
    out = run(case, -snapshots)
    assert out["stats"]["checkpointing"] == "binomial"
    assert out["stats"]["snapshots"] == snapshots + 1
    assert_same_gradient(out, ref)


def test_binomial_leaves_the_final_state(segments):
    case, ref = segments
    out = run(case, -2)
    fwd = setup(case)
    fwd.step(NSTEPS, 0.0, case.dt_cfl())
    p = fwd.particles(fields=False)
    fwd.close()
    order = np.argsort(p["id"])
    for key in ("x", "v", "gamma", "id"):
        assert np.array_equal(out["final"][key], p[key][order]), key
        assert np.array_equal(ref["final"][key], p[key][order]), key


def test_binomial_schedule_counts(segments):
    """The first pass (NSTEPS steps), the re-runs of the schedule and the
    final re-run (NSTEPS steps): with more snapshots, fewer steps; every
    backward step but those on the first pass's live state needs a restore."""
    case, _ = segments
    steps = [run(case, -s)["stats"] for s in (1, 2, 3)]
    for st in steps:
        assert st["forward_steps"] >= 2 * NSTEPS
        assert st["restores"] >= 1
    assert steps[0]["forward_steps"] > steps[1]["forward_steps"] >= steps[2]["forward_steps"]
    # the quadratic bound of a single snapshot: every step re-run from 0
    assert steps[0]["forward_steps"] < 2 * NSTEPS + NSTEPS * (NSTEPS - 1) // 2


def expected_sums(out, species, ns):
    s = np.zeros((ns, 7))
    for q in range(ns):
        m = species == q
        s[q, :3] = out["xbar0"][m].sum(axis=0)
        s[q, 3:6] = out["vbar0"][m].sum(axis=0)
        s[q, 6] = out["gbar0"][m].sum()
    return s


@pytest.mark.parametrize("k,precision", [(4, FP32_CHOSEN), (-2, FP32_CHOSEN), (4, FP64_PLAIN)],
                         ids=["fp32_segments", "fp32_binomial", "fp64_segments"])
def test_species_sums_are_the_sums_of_the_outputs(k, precision):
    case = CASES["periodic_drift"]
    ctx = setup(case, precision)
    p0 = ctx.particles(fields=False)
    species = np.empty(len(p0["id"]), np.int32)
    species[p0["id"]] = p0["species"]
    out = ctx.adjoint_field_energy(NSTEPS, 0.0, case.dt_cfl(), k)
    st = ctx.adjoint_stats()
    ctx.close()
    exp = expected_sums(out, species, case.nspecies)
    scale = np.array([[np.abs(out["xbar0"]).sum()] * 3 + [np.abs(out["vbar0"]).sum()] * 3 +
                      [np.abs(out["gbar0"]).sum()]])
    assert np.all(np.abs(st["sums"] - exp) <= 1e-12 * scale + 1e-300), (st["sums"], exp)
    assert st["checkpointing"] == ("binomial" if k < 0 else "segments")


def test_binomial_zero_and_one_step():
    case = CASES["lightbc_rel"]
    for n in (0, 1):
        a, b = run(case, -3, nsteps=n), run(case, 4, nsteps=n)
        assert_same_gradient(a, b)
        for key in ("x", "v", "gamma", "id"):
            assert np.array_equal(a["final"][key], b["final"][key]), (n, key)
    # after zero steps the state still carries the stored gamma of the first
    # push: a step after the call equals a step without it
    c1 = setup(case)
    c1.adjoint_field_energy(0, 0.0, case.dt_cfl(), -2)
    c1.step(3, 0.0, case.dt_cfl())
    c2 = setup(case)
    c2.step(3, 0.0, case.dt_cfl())
    p1, p2 = c1.particles(fields=False), c2.particles(fields=False)
    c1.close()
    c2.close()
    o1, o2 = np.argsort(p1["id"]), np.argsort(p2["id"])
    assert np.array_equal(p1["v"][o1], p2["v"][o2])


def test_stats_before_any_gradient_call():
    ctx = setup(CASES["periodic_drift"])
    with pytest.raises(Exception):
        ctx.adjoint_stats()
    ctx.close()
