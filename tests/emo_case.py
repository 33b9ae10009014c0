"""Builds libemopic's emo_config / emo_loader from a tests.oracle_ctypes.Case,
so that the HIP path and the oracle run the same input."""
from __future__ import annotations

from apc_524_project_amd._lib import FP64_PLAIN, EmoConfig, EmoLoader


def emo_config(case, precision=FP64_PLAIN, q=None, m=None, max_particles=None) -> EmoConfig:
    c = EmoConfig()
    for d in range(3):
        c.nCell[d] = case.n[d]
        c.xyz0[d] = case.x0[d]
        c.Lxyz[d] = case.L[d]
        c.E0[d] = case.E0[d]
        c.B0[d] = case.B0[d]
    c.relativity = case.relativity
    c.precision = precision
    for s in range(6):
        c.fields_bound[s] = case.fbc[s]
        c.parts_bound[s] = case.pbc[s]
        c.bound_phi[s] = case.bound_phi[s]
        c.bound_Ax[s] = case.bound_Ax[s]
        c.bound_Ay[s] = case.bound_Ay[s]
        c.bound_Az[s] = case.bound_Az[s]
    c.nspecies = case.nspecies
    for i in range(case.nspecies):
        if q is not None:
            c.charge[i] = q[i]
            c.mass[i] = m[i]
        c.isTestParticle[i] = case.is_test[i]
    c.nwaves = len(case.waves)
    for i, w in enumerate(case.waves):
        c.inSide[i] = w.side
        c.inPolE[i] = w.pol
        c.peakamps[i] = w.amp
        c.omegas[i] = w.omega
        c.phases[i] = w.phase
        c.invWidths[i] = w.inv_width
        c.delays[i] = w.delay
    c.waist = case.waist
    c.max_particles = case.nparticles if max_particles is None else max_particles
    return c


def emo_loader(case) -> EmoLoader:
    ld = EmoLoader()
    ld.nparticles_tot = case.nparticles
    ld.nspecies = case.nspecies
    for i in range(case.nspecies):
        ld.mass_ratio[i] = case.mass_ratio[i]
        ld.charge_ratio[i] = case.charge_ratio[i]
        ld.dens_frac[i] = case.dens_frac[i]
        ld.temp[i] = case.temp[i]
        ld.isTestParticle[i] = case.is_test[i]
        if case.pert_eps:
            ld.pert_eps[i] = case.pert_eps[i]
            ld.pert_k[i] = case.pert_k[i]
        if case.drift:
            for d in range(3):
                ld.drift[i][d] = case.drift[i][d]
    ld.dens_phys = case.dens_phys
    ld.seed = case.seed
    ld.zero_avg = case.zero_avg
    ld.poisson_init = 1 if case.poisson else 0
    return ld
