"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and for the
reference harness (oracle/_ref/libemo_ref.so).

Test infrastructure only: used by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, always as the checker.
"""
from __future__ import annotations

import ctypes as C
import math
import os
from dataclasses import dataclass, field

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libemo_ref.so")

NSPEC = 32
NWAVE = 32

# reference constants, src/globals.hpp:12,16,20
UNIT_ACC = 0.58668774
UNIT_RHOJ = 4.80320427e-13
UNIT_VTH = 0.002232

# field ids (src/grid/grid.cpp:45-60, src/poisson/poisson.cpp:18-26)
EX, EY, EZ, BX, BY, BZ, JX, JY, JZ, RHO, BTX, BTY, BTZ = range(13)
PHI1, AX1, AY1, AZ1, PHI2, AX2, AY2, AZ2 = range(13, 21)

PERIODIC, EXTERNAL, REFLECTING = 0, 1, 1


class EoConfig(C.Structure):
    _fields_ = [
        ("n", C.c_int32 * 3),
        ("x0", C.c_double * 3),
        ("L", C.c_double * 3),
        ("relativity", C.c_int32),
        ("fbc", C.c_int32 * 6),
        ("pbc", C.c_int32 * 6),
        ("nspecies", C.c_int32),
        ("q", C.c_double * NSPEC),
        ("m", C.c_double * NSPEC),
        ("is_test", C.c_int32 * NSPEC),
        ("E0", C.c_double * 3),
        ("B0", C.c_double * 3),
        ("nwaves", C.c_int32),
        ("in_side", C.c_int32 * NWAVE),
        ("in_pol", C.c_int32 * NWAVE),
        ("peakamps", C.c_double * NWAVE),
        ("omegas", C.c_double * NWAVE),
        ("phases", C.c_double * NWAVE),
        ("inv_widths", C.c_double * NWAVE),
        ("delays", C.c_double * NWAVE),
        ("waist", C.c_double),
        ("bound_phi", C.c_double * 6),
        ("bound_Ax", C.c_double * 6),
        ("bound_Ay", C.c_double * 6),
        ("bound_Az", C.c_double * 6),
    ]


class EoLoadConfig(C.Structure):
    _fields_ = [
        ("nparticles_tot", C.c_int64),
        ("nspecies", C.c_int32),
        ("mass_ratio", C.c_double * NSPEC),
        ("charge_ratio", C.c_double * NSPEC),
        ("dens_frac", C.c_double * NSPEC),
        ("temp", C.c_double * NSPEC),
        ("is_test", C.c_int32 * NSPEC),
        ("dens_phys", C.c_double),
        ("seed", C.c_int64),
        ("zero_avg", C.c_int32),
        ("poisson_init", C.c_int32),
        ("pert_eps", C.c_double * NSPEC),
        ("pert_k", C.c_double * NSPEC),
        ("drift", (C.c_double * 3) * NSPEC),
    ]


@dataclass
class Wave:
    side: int
    pol: int
    amp: float
    omega: float
    phase: float
    inv_width: float
    delay: float


@dataclass
class Case:
    """A run configuration in the reference's input vocabulary."""

    n: tuple = (4, 4, 4)
    x0: tuple = (0.0, 0.0, 0.0)
    L: tuple = (4.0, 4.0, 4.0)
    relativity: int = 0
    fbc: tuple = (PERIODIC,) * 6
    pbc: tuple = (PERIODIC,) * 6
    nparticles: int = 64
    mass_ratio: tuple = (1.0, 1836.0)
    charge_ratio: tuple = (-1.0, 1.0)
    dens_frac: tuple = (0.5, 0.5)
    temp: tuple = (500.0, 500.0)
    is_test: tuple = (0, 0)
    dens_phys: float = 0.0
    seed: int = -1
    poisson: bool = False
    E0: tuple = (0.0, 0.0, 0.0)
    B0: tuple = (0.0, 0.0, 0.0)
    waves: list = field(default_factory=list)
    waist: float = 0.0
    bound_phi: tuple = (0.0,) * 6
    bound_Ax: tuple = (0.0,) * 6
    bound_Ay: tuple = (0.0,) * 6
    bound_Az: tuple = (0.0,) * 6
    zero_avg: int = -1
    pert_eps: tuple = ()
    pert_k: tuple = ()
    drift: tuple = ()

    @property
    def nspecies(self):
        return len(self.mass_ratio)

    def dt_cfl(self, frac=0.5):
        d = [self.L[i] / self.n[i] for i in range(3)]
        return frac / math.sqrt(sum(1.0 / (x * x) for x in d))

    def eo_config(self) -> EoConfig:
        c = EoConfig()
        for i in range(3):
            c.n[i] = self.n[i]
            c.x0[i] = self.x0[i]
            c.L[i] = self.L[i]
            c.E0[i] = self.E0[i]
            c.B0[i] = self.B0[i]
        c.relativity = self.relativity
        for i in range(6):
            c.fbc[i] = self.fbc[i]
            c.pbc[i] = self.pbc[i]
            c.bound_phi[i] = self.bound_phi[i]
            c.bound_Ax[i] = self.bound_Ax[i]
            c.bound_Ay[i] = self.bound_Ay[i]
            c.bound_Az[i] = self.bound_Az[i]
        c.nwaves = len(self.waves)
        for i, w in enumerate(self.waves):
            c.in_side[i] = w.side
            c.in_pol[i] = w.pol
            c.peakamps[i] = w.amp
            c.omegas[i] = w.omega
            c.phases[i] = w.phase
            c.inv_widths[i] = w.inv_width
            c.delays[i] = w.delay
        c.waist = self.waist
        return c

    def load_config(self) -> EoLoadConfig:
        lc = EoLoadConfig()
        lc.nparticles_tot = self.nparticles
        lc.nspecies = self.nspecies
        for i in range(self.nspecies):
            lc.mass_ratio[i] = self.mass_ratio[i]
            lc.charge_ratio[i] = self.charge_ratio[i]
            lc.dens_frac[i] = self.dens_frac[i]
            lc.temp[i] = self.temp[i]
            lc.is_test[i] = self.is_test[i]
            if self.pert_eps:
                lc.pert_eps[i] = self.pert_eps[i]
                lc.pert_k[i] = self.pert_k[i]
            if self.drift:
                for d in range(3):
                    lc.drift[i][d] = self.drift[i][d]
        lc.dens_phys = self.dens_phys
        lc.seed = self.seed
        lc.zero_avg = self.zero_avg
        lc.poisson_init = 1 if self.poisson else 0
        return lc


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


_oracle = None
_ref = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError(f"oracle not built: {ORACLE_SO} (run make -C oracle)")
        lib = C.CDLL(ORACLE_SO)
        vp = C.c_void_p
        lib.eo_create.restype = vp
        lib.eo_create.argtypes = [C.POINTER(EoConfig), C.c_int64]
        lib.eo_destroy.argtypes = [vp]
        lib.eo_set_particles.argtypes = [vp, C.c_int64, vp, vp, vp, vp, vp]
        lib.eo_num_particles.restype = C.c_int64
        lib.eo_num_particles.argtypes = [vp]
        lib.eo_get_particles.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        lib.eo_set_particle_fields.argtypes = [vp, vp]
        lib.eo_field_points.restype = C.c_int64
        lib.eo_field_points.argtypes = [vp]
        lib.eo_get_field.argtypes = [vp, C.c_int, vp]
        lib.eo_set_field.argtypes = [vp, C.c_int, vp]
        lib.eo_process_and_load.argtypes = [vp, C.POINTER(EoConfig), C.POINTER(EoLoadConfig)]
        lib.eo_rng_uniform.restype = C.c_double
        lib.eo_rng_gaussian.restype = C.c_double
        lib.eo_rng_init.argtypes = [vp, C.c_int64]
        lib.eo_rng_uniform.argtypes = [vp]
        lib.eo_rng_gaussian.argtypes = [vp, C.c_double, C.c_double]
        for fn in ("eo_push", "eo_evolve"):
            getattr(lib, fn).argtypes = [vp, C.c_double]
        for fn in ("eo_particle_bc", "eo_sort", "eo_fold", "eo_gather", "eo_const_init"):
            getattr(lib, fn).argtypes = [vp]
        lib.eo_deposit.argtypes = [vp, C.c_int]
        lib.eo_refresh.argtypes = [vp, C.c_double, C.c_double]
        lib.eo_step.argtypes = [vp, C.c_double, C.c_double]
        lib.eo_cell_id.restype = C.c_int32
        lib.eo_cell_id.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        lib.eo_poisson_init.argtypes = [vp, C.c_int32, C.c_double, vp, C.c_double, C.c_double]
        lib.eo_diag.argtypes = [vp, vp]
        lib.eo_set_b_alias.argtypes = [vp, C.c_int]
        lib.eo_gaussian_pulses.restype = C.c_double
        lib.eo_gaussian_pulses.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_double]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref_lib():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError(f"reference harness not built: {REF_SO}")
        lib = C.CDLL(REF_SO)
        vp = C.c_void_p
        lib.rh_create.restype = vp
        lib.rh_create.argtypes = [C.POINTER(EoConfig), C.POINTER(EoLoadConfig), C.c_int, C.c_int, C.c_int]
        lib.rh_destroy.argtypes = [vp]
        lib.rh_species.argtypes = [vp, vp, vp]
        lib.rh_num_particles.restype = C.c_int64
        lib.rh_num_particles.argtypes = [vp]
        lib.rh_get_particles.argtypes = [vp, vp, vp, vp, vp, vp, vp]
        lib.rh_set_particle_state.argtypes = [vp, C.c_int64, vp, vp, vp, vp]
        lib.rh_field_points.restype = C.c_int64
        lib.rh_field_points.argtypes = [vp]
        lib.rh_get_field.argtypes = [vp, C.c_int, vp]
        lib.rh_set_field.argtypes = [vp, C.c_int, vp]
        lib.rh_get_field_ptr.argtypes = [vp, C.c_int, vp]
        for fn in ("rh_initialize", "rh_particle_bc", "rh_sort", "rh_deposit", "rh_fold", "rh_gather"):
            getattr(lib, fn).argtypes = [vp]
        lib.rh_poisson_only.argtypes = [vp]
        lib.rh_push.argtypes = [vp, C.c_double]
        lib.rh_evolve.argtypes = [vp, C.c_double]
        lib.rh_refresh.argtypes = [vp, C.c_double, C.c_double]
        lib.rh_step.argtypes = [vp, C.c_int, C.c_double, C.c_double]
        lib.rh_diag.argtypes = [vp, C.c_double, C.c_long, vp]
        lib.rh_cell_id.restype = C.c_int32
        lib.rh_cell_id.argtypes = [vp, C.c_double, C.c_double, C.c_double]
        lib.rh_rng_uniform.argtypes = [C.c_int64, C.c_int64, vp]
        lib.rh_rng_gaussian.argtypes = [C.c_int64, C.c_int64, C.c_double, C.c_double, vp]
        lib.rh_gaussian_pulses.restype = C.c_double
        lib.rh_gaussian_pulses.argtypes = [C.c_int, vp, vp, vp, vp, vp, C.c_double, C.c_double]
        lib.rh_time_steps.restype = C.c_double
        lib.rh_time_steps.argtypes = [vp, C.c_int, C.c_double]
        lib.rh_time_steps_from.restype = C.c_double
        lib.rh_time_steps_from.argtypes = [vp, C.c_int, C.c_int, C.c_double]
        _ref = lib
    return _ref


@dataclass
class Particles:
    x: np.ndarray
    v: np.ndarray
    gamma: np.ndarray
    species: np.ndarray
    id: np.ndarray
    field: np.ndarray

    def by_id(self) -> "Particles":
        o = np.argsort(self.id, kind="stable")
        return Particles(self.x[o], self.v[o], self.gamma[o], self.species[o], self.id[o], self.field[o])


class Oracle:
    """One oracle state (grid + particles)."""

    def __init__(self, case: Case, max_particles: int | None = None):
        self.case = case
        self.lib = oracle_lib()
        self.cfg = case.eo_config()
        cap = max_particles if max_particles is not None else max(case.nparticles, 1)
        self.h = self.lib.eo_create(C.byref(self.cfg), cap)
        self.shape = tuple(n + 3 for n in case.n)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.eo_destroy(self.h)
            self.h = None

    def load(self) -> int:
        lc = self.case.load_config()
        return self.lib.eo_process_and_load(self.h, C.byref(self.cfg), C.byref(lc))

    @property
    def q(self):
        return np.array(self.cfg.q[: self.case.nspecies])

    @property
    def m(self):
        return np.array(self.cfg.m[: self.case.nspecies])

    def set_particles(self, x, v, gamma, species, ids):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        gamma = np.ascontiguousarray(gamma, dtype=np.float64)
        species = np.ascontiguousarray(species, dtype=np.int32)
        ids = np.ascontiguousarray(ids, dtype=np.int64)
        rc = self.lib.eo_set_particles(self.h, len(gamma), _p(x), _p(v), _p(gamma), _p(species), _p(ids))
        assert rc == 0

    def set_species(self, q, m, is_test=None):
        """Set super-particle q/m (needs a fresh state: recreates it)."""
        for i in range(len(q)):
            self.cfg.q[i] = q[i]
            self.cfg.m[i] = m[i]
            self.cfg.is_test[i] = 0 if is_test is None else is_test[i]
        self.cfg.nspecies = len(q)
        self.lib.eo_destroy(self.h)
        cap = max(self.case.nparticles, 1)
        self.h = self.lib.eo_create(C.byref(self.cfg), cap)

    def particles(self) -> Particles:
        n = self.lib.eo_num_particles(self.h)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        g = np.zeros(n)
        s = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int64)
        f = np.zeros((n, 6))
        self.lib.eo_get_particles(self.h, _p(x), _p(v), _p(g), _p(s), _p(i), _p(f))
        return Particles(x, v, g, s, i, f)

    def set_particle_fields(self, f6):
        f6 = np.ascontiguousarray(f6, dtype=np.float64)
        self.lib.eo_set_particle_fields(self.h, _p(f6))

    def field(self, fid) -> np.ndarray:
        out = np.zeros(self.shape)
        assert self.lib.eo_get_field(self.h, fid, _p(out)) == 0
        return out

    def set_field(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        assert self.lib.eo_set_field(self.h, fid, _p(a)) == 0

    def push(self, dt):
        self.lib.eo_push(self.h, dt)

    def set_b_alias(self, odd: bool):
        """Q23: fieldPtr_[B] aliases the member B_tm1 after an odd number of swaps."""
        self.lib.eo_set_b_alias(self.h, 1 if odd else 0)

    def particle_bc(self):
        self.lib.eo_particle_bc(self.h)

    def sort(self):
        if self.lib.eo_sort(self.h) != 0:
            raise RuntimeError("oracle: a particle left the padded grid (or is NaN)")

    def deposit(self, rho=True):
        self.lib.eo_deposit(self.h, 1 if rho else 0)

    def fold(self):
        self.lib.eo_fold(self.h)

    def evolve(self, dt):
        self.lib.eo_evolve(self.h, dt)

    def refresh(self, t, dt):
        self.lib.eo_refresh(self.h, t, dt)

    def gather(self):
        self.lib.eo_gather(self.h)

    def step(self, t, dt):
        if self.lib.eo_step(self.h, t, dt) != 0:
            raise RuntimeError("oracle: a particle left the padded grid (or is NaN)")

    def const_init(self):
        self.lib.eo_const_init(self.h)

    def poisson_init(self, iter_max=10000, tol=-1.0, t=0.0, dt=0.0):
        sw = np.zeros(4, dtype=np.int32)
        self.lib.eo_poisson_init(self.h, iter_max, tol, _p(sw), t, dt)
        return sw

    def initialize(self, iter_max=10000, tol=-1.0):
        """src/main.cpp:159-179 in the oracle's terms."""
        if self.case.poisson:
            self.sort()
            self.deposit(True)
            self.fold()
            sw = self.poisson_init(iter_max, tol)
        else:
            self.const_init()
            sw = None
        self.gather()
        return sw

    def diag(self):
        out = np.zeros(8)
        self.lib.eo_diag(self.h, _p(out))
        return out

    def cell_id(self, x, y, z):
        return self.lib.eo_cell_id(self.h, x, y, z)


class Reference:
    """The reference's own classes driven through oracle/_ref/libemo_ref.so."""

    def __init__(self, case: Case, nstep_sort: int = 100, debug: int = 0):
        self.case = case
        self.lib = ref_lib()
        self.cfg = case.eo_config()
        self.lc = case.load_config()
        self.h = self.lib.rh_create(C.byref(self.cfg), C.byref(self.lc), 1 if case.poisson else 0, nstep_sort, debug)
        if not self.h:
            raise RuntimeError("reference rejected the input")
        self.shape = tuple(n + 3 for n in case.n)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.rh_destroy(self.h)
            self.h = None

    def species(self):
        q = np.zeros(NSPEC)
        m = np.zeros(NSPEC)
        self.lib.rh_species(self.h, _p(q), _p(m))
        return q[: self.case.nspecies], m[: self.case.nspecies]

    def particles(self) -> Particles:
        n = self.lib.rh_num_particles(self.h)
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        g = np.zeros(n)
        s = np.zeros(n, dtype=np.int32)
        i = np.zeros(n, dtype=np.int64)
        f = np.zeros((n, 6))
        self.lib.rh_get_particles(self.h, _p(x), _p(v), _p(g), _p(s), _p(i), _p(f))
        return Particles(x, v, g, s, i, f)

    def set_particle_state(self, x, v, gamma, field6=None):
        x = np.ascontiguousarray(x, dtype=np.float64)
        v = np.ascontiguousarray(v, dtype=np.float64)
        gamma = np.ascontiguousarray(gamma, dtype=np.float64)
        f = None if field6 is None else np.ascontiguousarray(field6, dtype=np.float64)
        rc = self.lib.rh_set_particle_state(self.h, len(gamma), _p(x), _p(v), _p(gamma), None if f is None else _p(f))
        assert rc == 0

    def field(self, fid):
        out = np.zeros(self.shape)
        assert self.lib.rh_get_field(self.h, fid, _p(out)) == 0
        return out

    def set_field(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        assert self.lib.rh_set_field(self.h, fid, _p(a)) == 0

    def field_ptr(self, fid):
        """Grid::getFieldPtr() view (stale for B after an odd number of swaps)."""
        out = np.zeros(self.shape)
        assert self.lib.rh_get_field_ptr(self.h, fid, _p(out)) == 0
        return out

    def initialize(self):
        self.lib.rh_initialize(self.h)

    def poisson_only(self):
        assert self.lib.rh_poisson_only(self.h) == 0

    def push(self, dt):
        self.lib.rh_push(self.h, dt)

    def particle_bc(self):
        self.lib.rh_particle_bc(self.h)

    def sort(self):
        self.lib.rh_sort(self.h)

    def deposit(self):
        self.lib.rh_deposit(self.h)

    def fold(self):
        self.lib.rh_fold(self.h)

    def evolve(self, dt):
        self.lib.rh_evolve(self.h, dt)

    def refresh(self, t, dt):
        self.lib.rh_refresh(self.h, t, dt)

    def gather(self):
        self.lib.rh_gather(self.h)

    def step(self, ti, t, dt):
        self.lib.rh_step(self.h, ti, t, dt)

    def diag(self, t=0.0, i=0):
        out = np.zeros(8)
        got = self.lib.rh_diag(self.h, t, i, _p(out))
        return out if got else None

    def cell_id(self, x, y, z):
        return self.lib.rh_cell_id(self.h, x, y, z)

    def time_steps(self, nsteps, dt):
        return self.lib.rh_time_steps(self.h, nsteps, dt)

    def time_steps_from(self, ti0, nsteps, dt):
        """Wall time of the loop's steps ti0 .. ti0 + nsteps - 1."""
        return self.lib.rh_time_steps_from(self.h, ti0, nsteps, dt)


def ref_rng_uniform(seed, n):
    out = np.zeros(n)
    ref_lib().rh_rng_uniform(seed, n, _p(out))
    return out


def ref_rng_gaussian(seed, n, mu=0.0, sigma=1.0):
    out = np.zeros(n)
    ref_lib().rh_rng_gaussian(seed, n, mu, sigma, _p(out))
    return out


class _Rng(C.Structure):
    _fields_ = [
        ("idum", C.c_int64),
        ("idum2", C.c_int64),
        ("iy", C.c_int64),
        ("iv", C.c_int64 * 32),
        ("z0", C.c_double),
        ("z1", C.c_double),
        ("generate", C.c_int32),
    ]


def oracle_rng_uniform(seed, n):
    lib = oracle_lib()
    r = _Rng()
    lib.eo_rng_init(C.byref(r), seed)
    return np.array([lib.eo_rng_uniform(C.byref(r)) for _ in range(n)])


def oracle_rng_gaussian(seed, n, mu=0.0, sigma=1.0):
    lib = oracle_lib()
    r = _Rng()
    lib.eo_rng_init(C.byref(r), seed)
    return np.array([lib.eo_rng_gaussian(C.byref(r), mu, sigma) for _ in range(n)])


def cell_ids_from_x(case: Case, x: np.ndarray) -> np.ndarray:
    """Grid::getCellID (src/grid/oGrid.cpp:208-222) vectorised."""
    n = np.array(case.n)
    d = np.array(case.L, dtype=np.float64) / n
    idx = 1.0 / d
    ijk = ((x - np.array(case.x0)) * idx).astype(np.int64) + 1
    nc = n + 2
    return (nc[1] * nc[2]) * ijk[:, 0] + nc[2] * ijk[:, 1] + ijk[:, 2]
