"""Writes the golden fixtures under tests/golden/ by running the reference's
own code (oracle/_ref/libemo_ref.so, compiled by oracle/Makefile from the
reference sources).  Each fixture is data only: inputs and the outputs the
reference produced for them.

    make -C oracle && python tests/golden/make_golden.py

G1  ran2 / Box-Muller streams (seeds -1, 1233)
G2  Load + ProcessInfo + zeroAvg state (4^3 cells, 64 particles), Boris and Vay
G3  one full step, per phase, from the G2 state with constant E0, B0
G4  Poisson initialisation: potentials and fields after 1 sweep and at convergence
G5  20-step history columns (OutputBoxQuantities) from the G2 state
G6  one LightBC / reflecting step (x external, y/z periodic)
"""
from __future__ import annotations

import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import oracle_ctypes as oc  # noqa: E402

CASE_G2 = dict(n=(4, 4, 4), L=(4.0, 4.0, 4.0), nparticles=64, temp=(500.0, 500.0), dens_phys=2.824e9)
CASE_G6 = dict(n=(8, 4, 4), L=(8.0, 4.0, 4.0), nparticles=128, temp=(2000.0, 2000.0), dens_phys=4.355e9,
               relativity=1, B0=(0.0, 0.0, 0.01),
               fbc=(oc.EXTERNAL, oc.EXTERNAL, oc.PERIODIC, oc.PERIODIC, oc.PERIODIC, oc.PERIODIC),
               pbc=(oc.REFLECTING, oc.REFLECTING, oc.PERIODIC, oc.PERIODIC, oc.PERIODIC, oc.PERIODIC),
               waves=[oc.Wave(-1, 2, 0.3, 0.39, 30.0, 0.05, 5.0), oc.Wave(1, 3, 0.1, 0.7, -10.0, 0.1, 2.0)])
FIELDS = {"Ex": oc.EX, "Ey": oc.EY, "Ez": oc.EZ, "Bx": oc.BX, "By": oc.BY, "Bz": oc.BZ, "Jx": oc.JX,
          "Jy": oc.JY, "Jz": oc.JZ, "rho": oc.RHO, "Btx": oc.BTX, "Bty": oc.BTY, "Btz": oc.BTZ}


def parts(r, prefix):
    p = r.particles().by_id()
    return {f"{prefix}_x": p.x, f"{prefix}_v": p.v, f"{prefix}_gamma": p.gamma, f"{prefix}_field": p.field}


def fields(r, prefix, names=FIELDS):
    return {f"{prefix}_{k}": r.field(fid) for k, fid in names.items()}


def g1():
    out = {}
    for seed in (-1, 1233):
        out[f"uniform_{seed}"] = oc.ref_rng_uniform(seed, 1000)
        out[f"gauss_{seed}"] = oc.ref_rng_gaussian(seed, 1000, 0.0, 1.0)
    return out


def g2_g3_g5(relativity):
    case = oc.Case(**CASE_G2, relativity=relativity, E0=(0.01, -0.02, 0.03), B0=(0.05, 0.02, -0.04))
    r = oc.Reference(case, nstep_sort=1)
    q, m = r.species()
    out = {"q": q, "m": m, "dt": np.array(case.dt_cfl())}
    p = r.particles()
    out.update(load_x=p.x, load_v=p.v, load_gamma=p.gamma, load_species=p.species, load_id=p.id)
    r.initialize()
    out.update(parts(r, "init"))
    dt = case.dt_cfl()
    r.push(dt)
    out.update(parts(r, "push"))
    r.particle_bc()
    out.update(parts(r, "bc"))
    r.deposit()
    out.update(fields(r, "dep", {"Jx": oc.JX, "Jy": oc.JY, "Jz": oc.JZ, "rho": oc.RHO}))
    r.fold()
    out.update(fields(r, "fold", {"Jx": oc.JX, "Jy": oc.JY, "Jz": oc.JZ, "rho": oc.RHO}))
    r.evolve(dt)
    out.update(fields(r, "evolve"))
    r.refresh(0.0, dt)
    out.update(fields(r, "refresh"))
    r.gather()
    out.update(parts(r, "gather"))
    # G5: history over 20 steps of the reference loop from a fresh load
    r2 = oc.Reference(case, nstep_sort=100)
    r2.initialize()
    hist = []
    t = 0.0
    for ti in range(20):
        if ti % 10 == 0:
            hist.append(r2.diag(t, ti))
        r2.step(ti, t, dt)
        t += dt
    hist.append(r2.diag(t, 20))
    out["history"] = np.array(hist)
    r3 = oc.Reference(case, nstep_sort=100)
    r3.initialize()
    t = 0.0
    for ti in range(20):
        r3.step(ti, t, dt)
        t += dt
    out.update(parts(r3, "steps20"))
    out.update(fields(r3, "steps20", {"Ex": oc.EX, "By": oc.BY, "Jz": oc.JZ, "rho": oc.RHO}))
    return out


def g4():
    out = {}
    for tag, bphi in (("one", (0.0, 1e12, 0.0, 0.0, 0.0, 0.0)), ("conv", (0.0, 3e-2, 0.0, 0.0, 0.0, 0.0))):
        case = oc.Case(**{**CASE_G2, "n": (4, 5, 3), "L": (4.0, 5.0, 3.0)}, poisson=True, bound_phi=bphi,
                       bound_Ax=(0.0, 1e12, 0.0, 0.0, 0.0, 0.0) if tag == "one" else (0.0,) * 6)
        r = oc.Reference(case)
        p = r.particles()
        out[f"{tag}_load_x"] = p.x
        out[f"{tag}_load_v"] = p.v
        # sources: deposit + fold as the reference does them, kept as inputs
        r.deposit()
        r.fold()
        for k, fid in (("rho", oc.RHO), ("Jx", oc.JX), ("Jy", oc.JY), ("Jz", oc.JZ)):
            out[f"{tag}_src_{k}"] = r.field(fid)
        r.poisson_only()
        for k, fid in (("phi1", oc.PHI1), ("phi2", oc.PHI2), ("Ax1", oc.AX1), ("Ay1", oc.AY1), ("Az1", oc.AZ1),
                       ("Ex", oc.EX), ("Ey", oc.EY), ("Ez", oc.EZ), ("Bx", oc.BX), ("By", oc.BY), ("Bz", oc.BZ)):
            out[f"{tag}_{k}"] = r.field(fid)
    return out


def g6():
    case = oc.Case(**CASE_G6)
    r = oc.Reference(case, nstep_sort=1)
    q, m = r.species()
    out = {"q": q, "m": m, "dt": np.array(case.dt_cfl())}
    p = r.particles()
    out.update(load_x=p.x, load_v=p.v, load_gamma=p.gamma, load_species=p.species, load_id=p.id)
    r.initialize()
    dt = case.dt_cfl()
    t = 0.0
    for ti in range(3):
        r.step(ti, t, dt)
        t += dt
    out.update(parts(r, "step3"))
    out.update(fields(r, "step3"))
    # one more step phase by phase, for bit-exact replays
    r.push(dt)
    r.particle_bc()
    out.update(parts(r, "bc4"))
    r.deposit()
    r.fold()
    out.update(fields(r, "fold4", {"Jx": oc.JX, "Jy": oc.JY, "Jz": oc.JZ, "rho": oc.RHO}))
    r.evolve(dt)
    r.refresh(t, dt)
    out.update(fields(r, "refresh4"))
    return out


def main():
    if not oc.ref_available():
        sys.exit("oracle/_ref/libemo_ref.so is not built (make -C oracle)")
    np.savez_compressed(os.path.join(HERE, "g1_rng.npz"), **g1())
    np.savez_compressed(os.path.join(HERE, "g2_boris.npz"), **g2_g3_g5(0))
    np.savez_compressed(os.path.join(HERE, "g2_vay.npz"), **g2_g3_g5(1))
    np.savez_compressed(os.path.join(HERE, "g4_poisson.npz"), **g4())
    np.savez_compressed(os.path.join(HERE, "g6_lightbc.npz"), **g6())
    for f in sorted(os.listdir(HERE)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(HERE, f)))


if __name__ == "__main__":
    main()
