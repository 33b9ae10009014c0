"""ctypes binding of libemopic's C ABI (include/emo_api.h).

The library is the product: this module only marshals arrays.  Loading fails
loudly when the in-tree libemopic.so is missing; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libemopic.so")

NSPEC = 32
NWAVE = 32

OK = 0
ERR_CONFIG, ERR_ALLOC, ERR_HIP, ERR_NUMERIC, ERR_INVARIANT, ERR_UNSUPPORTED = -1, -2, -3, -4, -5, -6
BC_PERIODIC, BC_EXTERNAL, BC_REFLECTING = 0, 1, 1
FP64_PLAIN, FP32_CHOSEN = 0, 1
EX, EY, EZ, BX, BY, BZ, JX, JY, JZ, RHO, BX_TM1, BY_TM1, BZ_TM1 = range(13)

# every function declared in include/emo_api.h
EXPORTS = (
    "emo_create", "emo_destroy", "emo_last_error", "emo_version",
    "emo_load_particles", "emo_process_info", "emo_load_maxwellian", "emo_load_uniform", "emo_num_particles",
    "emo_get_particles", "emo_get_cell_ids", "emo_field_points", "emo_get_field", "emo_set_field",
    "emo_init_constant", "emo_init_poisson", "emo_sort", "emo_push", "emo_particle_bc",
    "emo_deposit", "emo_fold", "emo_evolve_fields", "emo_refresh", "emo_interpolate",
    "emo_step", "emo_diag", "emo_synchronize", "emo_timing_enable", "emo_timing_select", "emo_kernel_time",
    "emo_timing_reset", "emo_adjoint_field_energy", "emo_loader_dx_deps", "emo_init_poisson_noalloc",
    "emo_layout_info", "emo_get_tracked", "emo_adjoint_stats",
)


class EmoConfig(C.Structure):
    _fields_ = [
        ("nCell", C.c_int32 * 3),
        ("xyz0", C.c_double * 3),
        ("Lxyz", C.c_double * 3),
        ("relativity", C.c_int32),
        ("precision", C.c_int32),
        ("fields_bound", C.c_int32 * 6),
        ("parts_bound", C.c_int32 * 6),
        ("nspecies", C.c_int32),
        ("charge", C.c_double * NSPEC),
        ("mass", C.c_double * NSPEC),
        ("isTestParticle", C.c_int32 * NSPEC),
        ("E0", C.c_double * 3),
        ("B0", C.c_double * 3),
        ("nwaves", C.c_int32),
        ("inSide", C.c_int32 * NWAVE),
        ("inPolE", C.c_int32 * NWAVE),
        ("peakamps", C.c_double * NWAVE),
        ("omegas", C.c_double * NWAVE),
        ("phases", C.c_double * NWAVE),
        ("invWidths", C.c_double * NWAVE),
        ("delays", C.c_double * NWAVE),
        ("waist", C.c_double),
        ("bound_phi", C.c_double * 6),
        ("bound_Ax", C.c_double * 6),
        ("bound_Ay", C.c_double * 6),
        ("bound_Az", C.c_double * 6),
        ("max_particles", C.c_int64),
        ("neutralize", C.c_int32),
    ]


class EmoLoader(C.Structure):
    _fields_ = [
        ("nparticles_tot", C.c_int64),
        ("nspecies", C.c_int32),
        ("mass_ratio", C.c_double * NSPEC),
        ("charge_ratio", C.c_double * NSPEC),
        ("dens_frac", C.c_double * NSPEC),
        ("temp", C.c_double * NSPEC),
        ("isTestParticle", C.c_int32 * NSPEC),
        ("dens_phys", C.c_double),
        ("seed", C.c_int64),
        ("zero_avg", C.c_int32),
        ("poisson_init", C.c_int32),
        ("pert_eps", C.c_double * NSPEC),
        ("pert_k", C.c_double * NSPEC),
        ("drift", (C.c_double * 3) * NSPEC),
    ]


class EmoError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libemopic error {code}: {msg}")
        self.code = code


_lib = None


def _torch_runtime_first():
    """torch-ROCm ships its own HIP runtime beside the system one that
    libemopic links.  The two coexist in one process only when torch's
    initialises first (the other order leaves torch with no devices), so
    when torch is installed its runtime is initialised before libemopic's
    first HIP call.  Without torch this does nothing."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def lib():
    """Load the in-tree libemopic.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -m apc_524_project_amd.build` "
            "(the HIP extension is required; there is no CPU fallback)"
        )
    _torch_runtime_first()
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int32, C.c_int64, C.c_double
    L.emo_create.argtypes = [C.POINTER(EmoConfig), C.POINTER(vp)]
    L.emo_destroy.argtypes = [vp]
    L.emo_destroy.restype = None
    L.emo_last_error.argtypes = [vp]
    L.emo_last_error.restype = C.c_char_p
    L.emo_version.restype = C.c_char_p
    L.emo_load_particles.argtypes = [vp, i64, vp, vp, vp, vp, vp]
    L.emo_process_info.argtypes = [C.POINTER(EmoLoader), C.POINTER(EmoConfig)]
    L.emo_load_maxwellian.argtypes = [vp, C.POINTER(EmoLoader)]
    L.emo_load_uniform.argtypes = [vp, C.POINTER(EmoLoader)]
    L.emo_num_particles.argtypes = [vp]
    L.emo_num_particles.restype = i64
    L.emo_get_particles.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.emo_get_cell_ids.argtypes = [vp, vp]
    L.emo_get_tracked.argtypes = [vp, i64, vp, vp, vp]
    L.emo_adjoint_stats.argtypes = [vp, vp, vp]
    L.emo_field_points.argtypes = [vp]
    L.emo_field_points.restype = i64
    L.emo_get_field.argtypes = [vp, C.c_int, vp]
    L.emo_set_field.argtypes = [vp, C.c_int, vp]
    L.emo_init_constant.argtypes = [vp]
    L.emo_init_poisson.argtypes = [vp, i32, dbl, vp, vp]
    L.emo_init_poisson_noalloc.argtypes = [vp, i32, dbl, vp, vp]
    for fn in ("emo_sort", "emo_particle_bc", "emo_fold", "emo_interpolate", "emo_synchronize",
               "emo_timing_reset"):
        getattr(L, fn).argtypes = [vp]
    L.emo_push.argtypes = [vp, dbl]
    L.emo_deposit.argtypes = [vp, i32]
    L.emo_evolve_fields.argtypes = [vp, dbl]
    L.emo_refresh.argtypes = [vp, dbl, dbl]
    L.emo_step.argtypes = [vp, i32, dbl, dbl]
    L.emo_diag.argtypes = [vp, vp]
    L.emo_timing_enable.argtypes = [vp, i32]
    L.emo_timing_select.argtypes = [vp, C.c_char_p]
    L.emo_timing_select.restype = i32
    L.emo_kernel_time.argtypes = [vp, C.c_char_p, C.POINTER(dbl), C.POINTER(i64)]
    L.emo_adjoint_field_energy.argtypes = [vp, i32, dbl, dbl, i32, vp, vp, vp, vp, vp, vp]
    L.emo_loader_dx_deps.argtypes = [C.POINTER(EmoLoader), C.POINTER(EmoConfig), vp]
    L.emo_layout_info.argtypes = [vp, vp]
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


class Context:
    """One libemopic context: a padded grid, a particle store and a stream."""

    def __init__(self, cfg: EmoConfig):
        self.L = lib()
        self.cfg = cfg
        h = C.c_void_p()
        rc = self.L.emo_create(C.byref(cfg), C.byref(h))
        if rc != OK:
            raise EmoError(rc, "emo_create failed (see stderr)")
        self.h = h
        self.shape = tuple(int(cfg.nCell[d]) + 3 for d in range(3))

    def close(self):
        if getattr(self, "h", None):
            self.L.emo_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc):
        if rc != OK:
            raise EmoError(rc, self.L.emo_last_error(self.h).decode())
        return rc

    # particles
    def load_particles(self, x, v, gamma=None, species=None, ids=None):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 3)
        n = x.shape[0]
        v = np.ascontiguousarray(v, dtype=np.float64).reshape(n, 3)
        g = np.ones(n) if gamma is None else np.ascontiguousarray(gamma, dtype=np.float64)
        s = np.zeros(n, np.int32) if species is None else np.ascontiguousarray(species, dtype=np.int32)
        i = np.arange(n, dtype=np.int64) if ids is None else np.ascontiguousarray(ids, dtype=np.int64)
        self._chk(self.L.emo_load_particles(self.h, n, _p(x), _p(v), _p(g), _p(s), _p(i)))

    def load_maxwellian(self, ld: EmoLoader):
        self._chk(self.L.emo_load_maxwellian(self.h, C.byref(ld)))

    def load_uniform(self, ld: EmoLoader):
        """Device-side loader (emo_load_uniform): per-cell uniform positions,
        counter-based random velocities."""
        self._chk(self.L.emo_load_uniform(self.h, C.byref(ld)))

    @property
    def num_particles(self):
        return int(self.L.emo_num_particles(self.h))

    def particles(self, fields=True):
        n = self.num_particles
        x = np.zeros((n, 3))
        v = np.zeros((n, 3))
        g = np.zeros(n)
        s = np.zeros(n, np.int32)
        i = np.zeros(n, np.int64)
        f = np.zeros((n, 6)) if fields else None
        self._chk(self.L.emo_get_particles(self.h, _p(x), _p(v), _p(g), _p(s), _p(i),
                                           _p(f) if fields else None))
        return dict(x=x, v=v, gamma=g, species=s, id=i, field=f)

    def layout_info(self):
        """FP32 chosen form: slots per record buffer, slots in the cells'
        ranges, layout rebuilds so far."""
        out = np.zeros(3, np.int64)
        self._chk(self.L.emo_layout_info(self.h, _p(out)))
        return {"slots": int(out[0]), "used": int(out[1]), "rebuilds": int(out[2])}

    def tracked(self, nid):
        """x, v of the particles with id < nid, by id, and whether each id exists
        (the rows of the reference's particle track files)."""
        x = np.zeros((nid, 3))
        v = np.zeros((nid, 3))
        found = np.zeros(nid, np.int32)
        self._chk(self.L.emo_get_tracked(self.h, nid, _p(x), _p(v), _p(found)))
        return x, v, found.astype(bool)

    def cell_ids(self):
        out = np.zeros(self.num_particles, np.int32)
        self._chk(self.L.emo_get_cell_ids(self.h, _p(out)))
        return out

    # fields
    def field(self, fid):
        out = np.zeros(self.shape)
        self._chk(self.L.emo_get_field(self.h, fid, _p(out)))
        return out

    def set_field(self, fid, a):
        a = np.ascontiguousarray(a, dtype=np.float64).reshape(self.shape)
        self._chk(self.L.emo_set_field(self.h, fid, _p(a)))

    # initialisation
    def init_constant(self):
        self._chk(self.L.emo_init_constant(self.h))

    def init_poisson(self, iter_max=10000, tol=-1.0):
        sw = np.zeros(4, np.int32)
        res = np.zeros(8)
        self._chk(self.L.emo_init_poisson(self.h, iter_max, tol, _p(sw), _p(res)))
        return sw, res

    def init_poisson_noalloc(self, iter_max=10000, tol=-1.0):
        """emo_init_poisson on the library's static device workspace."""
        sw = np.zeros(4, np.int32)
        res = np.zeros(8)
        self._chk(self.L.emo_init_poisson_noalloc(self.h, iter_max, tol, _p(sw), _p(res)))
        return sw, res

    # phases (FP64 plain form; evolve and refresh in both forms)
    def sort(self):
        self._chk(self.L.emo_sort(self.h))

    def push(self, dt):
        self._chk(self.L.emo_push(self.h, dt))

    def particle_bc(self):
        self._chk(self.L.emo_particle_bc(self.h))

    def deposit(self, rho=True):
        self._chk(self.L.emo_deposit(self.h, 1 if rho else 0))

    def fold(self):
        self._chk(self.L.emo_fold(self.h))

    def evolve(self, dt):
        self._chk(self.L.emo_evolve_fields(self.h, dt))

    def refresh(self, t, dt):
        self._chk(self.L.emo_refresh(self.h, t, dt))

    def interpolate(self):
        self._chk(self.L.emo_interpolate(self.h))

    def step(self, nsteps, t0, dt):
        self._chk(self.L.emo_step(self.h, nsteps, t0, dt))

    def diag(self):
        out = np.zeros(8)
        self._chk(self.L.emo_diag(self.h, _p(out)))
        return out

    def synchronize(self):
        self._chk(self.L.emo_synchronize(self.h))

    # timing
    def timing(self, on=True):
        self._chk(self.L.emo_timing_enable(self.h, 1 if on else 0))

    def timing_select(self, name=None):
        """Time only the launches of kernel `name` (None: all of them)."""
        self._chk(self.L.emo_timing_select(self.h, name.encode() if name else None))

    def timing_reset(self):
        self._chk(self.L.emo_timing_reset(self.h))

    def kernel_time(self, name):
        ms = C.c_double()
        n = C.c_int64()
        self._chk(self.L.emo_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    # adjoint (both forms)
    def adjoint_field_energy(self, nsteps, t0, dt, checkpoint_every=0, particles=True):
        """Runs nsteps forward and the adjoint back; returns the final field
        energy and its gradient (see emo_adjoint_field_energy)."""
        nw = int(self.cfg.nwaves)
        n = self.num_particles
        e = np.zeros(1)
        da = np.zeros(max(nw, 1))
        dp = np.zeros(max(nw, 1))
        xb = np.zeros((n, 3)) if particles else None
        vb = np.zeros((n, 3)) if particles else None
        gb = np.zeros(n) if particles else None
        self._chk(self.L.emo_adjoint_field_energy(
            self.h, nsteps, t0, dt, checkpoint_every, _p(e), _p(da), _p(dp),
            _p(xb) if particles else None, _p(vb) if particles else None, _p(gb) if particles else None))
        return dict(energy=float(e[0]), d_amp=da[:nw], d_phase=da[:0] if nw == 0 else dp[:nw],
                    xbar0=xb, vbar0=vb, gbar0=gb)

    def adjoint_stats(self):
        """The last gradient call: checkpointing ("segments" or "binomial"),
        snapshots held, forward steps run, restores, and per species the sums
        of xbar0[3], vbar0[3], gbar0 (an array nspecies x 7)."""
        info = np.zeros(4, np.int64)
        ns = max(int(self.cfg.nspecies), 1)
        sums = np.zeros(7 * ns)
        self._chk(self.L.emo_adjoint_stats(self.h, _p(info), _p(sums)))
        return dict(checkpointing="binomial" if info[0] == 1 else "segments", snapshots=int(info[1]),
                    forward_steps=int(info[2]), restores=int(info[3]), sums=sums.reshape(ns, 7))


def loader_dx_deps(ld: EmoLoader, cfg: EmoConfig) -> np.ndarray:
    """d x0 / d pert_eps of every loaded particle (id order)."""
    out = np.zeros(int(ld.nparticles_tot))
    rc = lib().emo_loader_dx_deps(C.byref(ld), C.byref(cfg), _p(out))
    if rc != OK:
        raise EmoError(rc, "emo_loader_dx_deps failed")
    return out
