"""The benchmark workloads as concrete inputs in the reference's units.

Units follow the reference (src/globals.hpp): length in cm (dx = 1),
c = 1, acceleration factor UNIT_ACC, current unit UNIT_RHOJ.  "0.5 CFL" is
dt = 0.5 dx / (c sqrt 3).  A species' thermal spread sigma (in c) maps to the
reference's temperature input T[eV] = (sigma / UNIT_VTH)^2 m.
Plasma frequency: w_pe^2 = 4 pi UNIT_RHOJ UNIT_ACC dens_phys.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

from ._lib import BC_EXTERNAL, BC_PERIODIC, BC_REFLECTING, FP64_PLAIN, EmoConfig, EmoLoader

UNIT_ACC = 0.58668774
UNIT_RHOJ = 4.80320427e-13
UNIT_VTH = 0.002232
WPE2_PER_DENS = 4 * math.pi * UNIT_RHOJ * UNIT_ACC


def temp_for_sigma(sigma: float, mass_ratio: float = 1.0) -> float:
    return (sigma / UNIT_VTH) ** 2 * mass_ratio


@dataclass
class Workload:
    name: str
    cfg: EmoConfig
    loader: EmoLoader
    dt: float
    steps: int
    poisson: bool
    poisson_sweeps: int  # fixed sweep count (0: the reference's convergence rule)
    device_loader: bool = False  # emo_load_uniform instead of the reference's ran2 stream

    @property
    def nparticles(self) -> int:
        return int(self.loader.nparticles_tot)

    @property
    def ncells(self) -> int:
        return int(self.cfg.nCell[0]) * int(self.cfg.nCell[1]) * int(self.cfg.nCell[2])


def _base(n, precision, relativity):
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = n[d]
        c.Lxyz[d] = float(n[d])
        c.xyz0[d] = 0.0
    c.relativity = relativity
    c.precision = precision
    for s in range(6):
        c.fields_bound[s] = BC_PERIODIC
        c.parts_bound[s] = BC_PERIODIC
    return c


def _species(ld: EmoLoader, specs):
    ld.nspecies = len(specs)
    for i, (m, q, frac, sigma) in enumerate(specs):
        ld.mass_ratio[i] = m
        ld.charge_ratio[i] = q
        ld.dens_frac[i] = frac
        ld.temp[i] = temp_for_sigma(sigma, m) if sigma > 0 else 0.0


def _process_info(ld: EmoLoader, c: EmoConfig) -> None:
    """Input::ProcessInfo's super-particle q, m (computed by libemopic)."""
    import ctypes

    from ._lib import lib

    rc = lib().emo_process_info(ctypes.byref(ld), ctypes.byref(c))
    if rc != 0:
        raise ValueError(f"emo_process_info failed ({rc})")


def make_workload(name: str, precision: int = FP64_PLAIN, scale: float = 1.0) -> Workload:
    """c1 .. c5 of the benchmark set, and c5s, c5's physics on a 128 x 32 x 32
    box (1/64 of c5's volume, the FP64 comparison of c5).  scale < 1 shrinks
    the particle count (same grid) for quick parity runs (c1 .. c4);
    benchmarks use scale = 1."""
    dt = 0.5 / math.sqrt(3.0)
    if name == "c1":
        # 16^3 periodic thermal plasma, 8 ppc/species, Boris, Poisson init (50 sweeps)
        n = (16, 16, 16)
        c = _base(n, precision, 0)
        ld = EmoLoader()
        ppc = 8
        npart = int(2 * ppc * 16 ** 3 * scale)
        ld.nparticles_tot = npart
        # electrons sigma 0.05c; ions the same temperature
        _species(ld, [(1.0, -1.0, 0.5, 0.05), (1836.0, 1.0, 0.5, 0.0)])
        ld.temp[1] = ld.temp[0]
        ld.dens_phys = 0.01 / WPE2_PER_DENS  # w_pe = 0.1
        ld.seed = -1
        ld.zero_avg = -1
        ld.poisson_init = 1
        steps, poisson, sweeps = 100, True, 50
    elif name == "c2":
        # 64^3 periodic, 32 ppc/species, relativistic Boris, sigma_e = 0.1c,
        # electron density perturbation 0.01 cos(2 pi x / 64)
        n = (64, 64, 64)
        c = _base(n, precision, 1)
        ld = EmoLoader()
        ppc = 32
        npart = int(2 * ppc * 64 ** 3 * scale)
        ld.nparticles_tot = npart
        _species(ld, [(1.0, -1.0, 0.5, 0.1), (1836.0, 1.0, 0.5, 0.0)])
        ld.temp[1] = ld.temp[0]
        ld.dens_phys = 0.01 / WPE2_PER_DENS
        ld.seed = -2
        ld.zero_avg = 0
        ld.poisson_init = 0
        ld.pert_eps[0] = 0.01
        ld.pert_k[0] = 2 * math.pi / 64
        steps, poisson, sweeps = 1000, False, 0
    elif name in ("c3", "c4"):
        # x: external fields / reflecting particles, y,z periodic; y-polarised
        # pulse from x-, lambda 16 cells, a0 0.5, waist 10 cells; n = 0.1 n_crit
        n = (128, 64, 64) if name == "c3" else (64, 32, 32)
        ppc = 16 if name == "c3" else 8
        c = _base(n, precision, 1)
        for s in (0, 1):
            c.fields_bound[s] = BC_EXTERNAL
            c.parts_bound[s] = BC_REFLECTING
        omega = 2 * math.pi / 16.0
        c.nwaves = 1
        c.inSide[0] = -1
        c.inPolE[0] = 2
        c.peakamps[0] = 0.5 * omega / UNIT_ACC
        c.omegas[0] = omega
        c.phases[0] = 0.0
        c.invWidths[0] = 1.0 / 40.0
        c.delays[0] = 60.0
        c.waist = 10.0
        ld = EmoLoader()
        npart = int(2 * ppc * n[0] * n[1] * n[2] * scale)
        ld.nparticles_tot = npart
        _species(ld, [(1.0, -1.0, 0.5, 0.01), (1836.0, 1.0, 0.5, 0.0)])
        ld.temp[1] = ld.temp[0]
        ld.dens_phys = 0.1 * omega * omega / WPE2_PER_DENS
        ld.seed = -3 if name == "c3" else -4
        if name == "c4":
            # the density-perturbation amplitude the adjoint differentiates
            ld.pert_eps[0] = 0.01
            ld.pert_k[0] = 2 * math.pi / n[0]
        ld.zero_avg = 0
        ld.poisson_init = 0
        steps, poisson, sweeps = (2000 if name == "c3" else 400), False, 0
    elif name in ("c5", "c5s"):
        # two-stream instability: two electron beams at +-0.2c along x
        # (sigma 0.01c), 64 ppc per beam, on a fixed uniform ion background
        # (the Poisson source minus its mean); periodic; relativistic Boris;
        # w_pe = 0.1.  Particles from the device loader (uniform in each cell).
        n = (512, 128, 128) if name == "c5" else (128, 32, 32)
        c = _base(n, precision, 1)
        c.neutralize = 1
        ld = EmoLoader()
        ppc = 64
        ld.nparticles_tot = 2 * ppc * n[0] * n[1] * n[2]
        _species(ld, [(1.0, -1.0, 0.5, 0.01), (1.0, -1.0, 0.5, 0.01)])
        ld.drift[0][0] = 0.2
        ld.drift[1][0] = -0.2
        ld.dens_phys = 0.01 / WPE2_PER_DENS
        ld.seed = -5
        ld.zero_avg = 0
        ld.poisson_init = 1
        steps, poisson, sweeps = 1000, True, 0
        c.max_particles = ld.nparticles_tot
        _process_info(ld, c)
        return Workload(name, c, ld, dt, steps, poisson, sweeps, device_loader=True)
    else:
        raise ValueError(f"unknown workload {name}")
    c.max_particles = ld.nparticles_tot
    _process_info(ld, c)
    return Workload(name, c, ld, dt, steps, poisson, sweeps)
