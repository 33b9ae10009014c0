/*
 * emo_api.h -- C ABI of libemopic, the MI355X (gfx950) implementation of
 * EMOOPIC's per-timestep particle-in-cell loop.
 *
 * Plain pointers and sizes only.  A context owns its device memory and one
 * HIP stream; every call is stream-ordered and returns 0 on success or a
 * negative EMO_ERR_* code, with a message from emo_last_error().  Host arrays
 * passed in are copied during the call and never retained.
 *
 * The entry points stand in for the methods that the reference's driver
 * loop calls (DenSto/APC_524_Project src/main.cpp:221-287); each declaration
 * names the reference method it replaces.
 *
 * Array conventions (the reference's, src/grid/grid.cpp:161-181):
 *   fields are padded (nx+3)*(ny+3)*(nz+3) doubles, z fastest; physical
 *   cells are indices 1..n; field ids follow src/grid/grid.cpp:45-60.
 *   particle vectors x3, v3 are [n][3] (Particle::x/v, src/particles/particle.hpp:10-11).
 */
#ifndef EMO_API_H
#define EMO_API_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EMO_NSPEC 32 /* Input_Info_t NSPEC, src/IO/input.hpp:10 */
#define EMO_NWAVE 32 /* Input_Info_t NWAVE, src/IO/input.hpp:12 */

/* return codes */
#define EMO_OK 0
#define EMO_ERR_CONFIG (-1)      /* invalid configuration or argument */
#define EMO_ERR_ALLOC (-2)       /* device or host allocation failed */
#define EMO_ERR_HIP (-3)         /* a HIP runtime call failed */
#define EMO_ERR_NUMERIC (-4)     /* NaN/Inf, Poisson non-convergence when required */
#define EMO_ERR_INVARIANT (-5)   /* particle outside the padded grid, net charge, ... */
#define EMO_ERR_UNSUPPORTED (-6) /* entry point not available in this precision mode */

/* boundary kinds (Input_Info_t parts_bound / fields_bound strings) */
#define EMO_BC_PERIODIC 0
#define EMO_BC_EXTERNAL 1   /* fields: "external" (LightBC / PoissonBC) */
#define EMO_BC_REFLECTING 1 /* particles: "reflecting" */

/* precision modes */
#define EMO_FP64_PLAIN 0  /* FP64 throughout, reference pass structure, bit-exact to the oracle */
#define EMO_FP32_CHOSEN 1 /* FP32 cell+offset SoA, fused gather/push/deposit */

/* field ids (src/grid/grid.cpp:45-60; src/poisson/poisson.cpp:18-26) */
enum {
  EMO_EX = 0, EMO_EY, EMO_EZ, EMO_BX, EMO_BY, EMO_BZ, EMO_JX, EMO_JY, EMO_JZ, EMO_RHO,
  EMO_BX_TM1, EMO_BY_TM1, EMO_BZ_TM1,
  EMO_PHI = 13, EMO_AX, EMO_AY, EMO_AZ
};

/* Configuration: the subset of Input_Info_t (src/IO/input.hpp:15-126) the
 * step needs, after Input::ProcessInfo (q, m are super-particle scaled). */
typedef struct emo_config {
  int32_t nCell[3];
  double xyz0[3];
  double Lxyz[3];
  int32_t relativity;          /* 0 Boris, 1 relativistic (Vay) Boris */
  int32_t precision;           /* EMO_FP64_PLAIN / EMO_FP32_CHOSEN */
  int32_t fields_bound[6];     /* x-,x+,y-,y+,z-,z+ */
  int32_t parts_bound[6];
  int32_t nspecies;
  double charge[EMO_NSPEC];    /* super-particle charge per species */
  double mass[EMO_NSPEC];      /* super-particle mass per species */
  int32_t isTestParticle[EMO_NSPEC];
  double E0[3], B0[3];
  int32_t nwaves;
  int32_t inSide[EMO_NWAVE];
  int32_t inPolE[EMO_NWAVE];
  double peakamps[EMO_NWAVE], omegas[EMO_NWAVE], phases[EMO_NWAVE];
  double invWidths[EMO_NWAVE], delays[EMO_NWAVE];
  double waist;                /* 0: plane-uniform pulse (reference); >0: transverse Gaussian waist */
  double bound_phi[6], bound_Ax[6], bound_Ay[6], bound_Az[6];
  int64_t max_particles;       /* capacity of the particle store */
  int32_t neutralize;          /* extension (0 = reference): the Poisson initialisation subtracts the
                                  mean charge density, i.e. adds a fixed uniform neutralising
                                  background, so that a net-charge plasma between periodic
                                  boundaries (where the reference asserts, particle_handler.cpp:
                                  435-438) has a solution */
} emo_config;

typedef struct emo_ctx emo_ctx;

/* ---- lifecycle ---------------------------------------------------------- */
/* Builds the Grid/Domain/Particle_Handler state (src/main.cpp:115-150). */
int emo_create(const emo_config *cfg, emo_ctx **out);
void emo_destroy(emo_ctx *ctx);
const char *emo_last_error(const emo_ctx *ctx);
/* library version string */
const char *emo_version(void);

/* ---- particles ---------------------------------------------------------- */
/* Replaces Particle_Handler::Load's result (src/particles/particle_handler.cpp:34-122):
 * copies n particles in; gamma may be NULL for Boris. */
int emo_load_particles(emo_ctx *ctx, int64_t n, const double *x3, const double *v3,
                       const double *gamma, const int32_t *species, const int64_t *id);
/* Particle_Handler::Load with the reference's ran2 + Box-Muller stream
 * (src/utils/RNG.cpp:23-141) after Input::ProcessInfo (src/IO/check.cpp:53-111).
 * Writes the super-particle charge/mass into cfg_inout. */
typedef struct emo_loader {
  int64_t nparticles_tot;
  int32_t nspecies;
  double mass_ratio[EMO_NSPEC], charge_ratio[EMO_NSPEC], dens_frac[EMO_NSPEC], temp[EMO_NSPEC];
  int32_t isTestParticle[EMO_NSPEC];
  double dens_phys;
  int64_t seed;                /* the reference uses -1 */
  int32_t zero_avg;            /* -1: reference rule; 0/1 force */
  int32_t poisson_init;        /* fields_init == "poisson" */
  double pert_eps[EMO_NSPEC];  /* extension: x -= (eps/k) sin(k x) displacement loader */
  double pert_k[EMO_NSPEC];
  double drift[EMO_NSPEC][3];  /* extension: drift velocity added after sampling */
} emo_loader;
int emo_process_info(const emo_loader *ld, emo_config *cfg_inout);
int emo_load_maxwellian(emo_ctx *ctx, const emo_loader *ld);
/* No counterpart in the reference (its loader is the serial ran2 stream of
 * emo_load_maxwellian); for loads too large for a host stream (10^9
 * particles).  Generates the particles on the device, per physical cell and
 * species nparticles_tot * dens_frac / ncells of them (which must be whole),
 * uniformly distributed in the cell, velocities Maxwellian (UNIT_VTH
 * sqrt(T/m) per component, T and m of Input::ProcessInfo) plus drift, stored
 * gamma = sqrt(1 - v^2) as the reference loader stores it.  The random
 * numbers come from a counter-based generator (SplitMix64 of seed, particle
 * and draw index; Box-Muller pairs), so both precisions load the same
 * particles.  Ids are the generation order (cell by cell).  Writes q, m
 * into the context as emo_load_maxwellian does. */
int emo_load_uniform(emo_ctx *ctx, const emo_loader *ld);
int64_t emo_num_particles(const emo_ctx *ctx);
/* copy out in the current array order; any pointer may be NULL */
int emo_get_particles(emo_ctx *ctx, double *x3, double *v3, double *gamma, int32_t *species,
                      int64_t *id, double *field6);
/* per-particle cell ids (Grid::getCellID, src/grid/oGrid.cpp:208-222) in array order */
int emo_get_cell_ids(emo_ctx *ctx, int32_t *cells);
/* The rows of the particle track files (Particle_Handler::outputParticles,
 * src/particles/particle_handler.cpp:308-365: x and v of each particle with
 * my_id < output_pCount): x3[3 id + d], v3[3 id + d] for id < nid, found[id]
 * = 1 where a particle with that id exists.  Selected on the device, so only
 * 7 doubles per tracked id cross to the host. */
int emo_get_tracked(emo_ctx *ctx, int64_t nid, double *x3, double *v3, int32_t *found);

/* ---- fields ------------------------------------------------------------- */
int64_t emo_field_points(const emo_ctx *ctx);
/* B/B_tm1 ids address the current member arrays (src/grid/oGrid.cpp:18-20) */
int emo_get_field(emo_ctx *ctx, int fid, double *padded);
int emo_set_field(emo_ctx *ctx, int fid, const double *padded);

/* ---- initialisation ----------------------------------------------------- */
/* Grid::InitializeFields constant branch (src/grid/grid.cpp:554-574) */
int emo_init_constant(emo_ctx *ctx);
/* src/main.cpp:160-174 on the Poisson path: deposit, executeBC(-2),
 * Poisson_Solver::InitializeFields (src/poisson/poisson.cpp:143-170).
 * iter_max mirrors ITER_MAX (src/poisson/poisson.cpp:14; the loop runs up to
 * iter_max+1 sweeps); tol < 0 selects the reference tolerance rule.
 * sweeps_out[4] (phi, Ax, Ay, Az) and resid_out[8] (FP64 max|du| and L2 of
 * the last sweep per solve) may be NULL. */
int emo_init_poisson(emo_ctx *ctx, int32_t iter_max, double tol, int32_t *sweeps_out,
                     double *resid_out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): emo_init_poisson with the same arguments and the same
 * results, whose workspace -- the Jacobi iterates of the four solves, the
 * residual partial sums and, in the FP32 form, the cell charge vectors --
 * lies in a static device array of 64 MiB per GPU instead of an allocation.
 * It allocates nothing.  A grid whose workspace exceeds 64 MiB (FP64: about
 * 1.0e6 padded points; FP32: about 2.0e6 minus the cell vectors) returns
 * EMO_ERR_CONFIG before any work.  The array is shared by every context on
 * the GPU and keeps the last call's contents: calls that overlap on one GPU
 * (two host threads, two contexts' streams) race. */
int emo_init_poisson_noalloc(emo_ctx *ctx, int32_t iter_max, double tol, int32_t *sweeps_out,
                             double *resid_out);

/* ---- the step's phases (FP64 plain form; in FP32 the particle phases and
 * the fold return EMO_ERR_UNSUPPORTED, the field-only phases emo_evolve_fields
 * and emo_refresh run on the FP32 fields) ------------------------------- */
int emo_sort(emo_ctx *ctx);                  /* Particle_Handler::SortParticles, particle_handler.cpp:180-183 */
int emo_push(emo_ctx *ctx, double dt);       /* Particle_Handler::Push, particle_handler.cpp:124-128 */
int emo_particle_bc(emo_ctx *ctx);           /* executeParticleBoundaryConditions + clearGhosts, :271-295 */
int emo_deposit(emo_ctx *ctx, int32_t rho);  /* Particle_Handler::depositRhoJ, :185-239 */
int emo_fold(emo_ctx *ctx);                  /* Grid::executeBC(-2), src/grid/grid.cpp:577-585 */
int emo_evolve_fields(emo_ctx *ctx, double dt); /* Grid::evolveFields, src/grid/oGrid.cpp:15-60 */
int emo_refresh(emo_ctx *ctx, double t, double dt); /* Grid::executeBC(-1) incl. LightBC::applyBCs */
int emo_interpolate(emo_ctx *ctx);           /* Particle_Handler::InterpolateEB, :138-173 */

/* ---- whole steps -------------------------------------------------------- */
/* nsteps iterations of src/main.cpp:231-275 starting at time t0 */
int emo_step(emo_ctx *ctx, int32_t nsteps, double t0, double dt);

/* ---- diagnostics -------------------------------------------------------- */
/* OutputBoxQuantities::output sums (src/IO/output.cpp:76-118) in FP64:
 * out8 = {pMx, pMy, pMz, pEn, B_en, E_en, E_tot, total charge} */
int emo_diag(emo_ctx *ctx, double *out8);

/* ---- adjoint (both forms) ------------------------------------------------
 * No counterpart in the reference (it has no derivatives); for callers that
 * differentiate the step, e.g. the torch.autograd.Function in
 * apc_524_project_amd/adjoint.py.
 *
 * Runs nsteps steps of emo_step from time t0, then the exact discrete
 * adjoint of those steps backward in time, re-running each K-step segment
 * from a checkpoint.  The objective is the final field energy of
 * OutputBoxQuantities, F = E_en + B_en (emo_diag out8[5] + out8[4]).
 * FP64 plain form: the particles are sorted by cell first (as SortParticles
 * before the reference loop's first push, src/main.cpp:225-226), and the
 * stored interpolated fields must be current (emo_interpolate after the
 * field initialisation).  FP32 chosen form: the adjoint of the FP32 step
 * itself (FP32 particle adjoints, FP64 field adjoints).  Outputs (any may be
 * NULL):
 *   energy        F
 *   d_amp[nw]     dF/d peakamps[w]
 *   d_phase[nw]   dF/d phases[w], per radian
 *   xbar0[n][3], vbar0[n][3], gbar0[n]   dF/d the initial positions,
 *                 velocities and stored gamma, in particle-id order
 * The context is left at the final state.
 * Checkpointing: checkpoint_every = K > 0 keeps a checkpoint every K steps
 * and tapes one K-step segment at a time (each segment re-run once); K = 0
 * is one segment.  FP32 chosen form: K < 0 asks for binomial checkpointing
 * with at most -K snapshots besides the initial state (particle adjoints by
 * id, no tape: the backward step of step s runs on the state at s, reached
 * by a restore and a re-run), which the call also takes by itself, with as
 * many snapshots as fit, when the segments do not fit in free device memory.
 * Both give the same bits; emo_adjoint_stats reports which ran. */
int emo_adjoint_field_energy(emo_ctx *ctx, int32_t nsteps, double t0, double dt, int32_t checkpoint_every,
                             double *energy, double *d_amp, double *d_phase, double *xbar0, double *vbar0,
                             double *gbar0);
/* The last gradient call on ctx (no counterpart in the reference):
 *   info4 = {checkpointing (0: K-step segments, 1: binomial), snapshots
 *            held (with the initial state), forward steps run (the first
 *            pass and every re-run), snapshot restores}
 *   sums[7 * nspecies] = per species the sums over its particles of
 *            xbar0[3], vbar0[3] and gbar0 (the by-id outputs), reduced on
 *            the device, so that runs too large for per-particle outputs
 *            still report ensemble gradients (a beam's dF/d drift).
 * Either pointer may be NULL.  EMO_ERR_CONFIG before any gradient call. */
int emo_adjoint_stats(emo_ctx *ctx, int64_t *info4, double *sums);
/* d x0 / d pert_eps[species] of emo_load_maxwellian's displacement loader for
 * every particle (load = id order; 0 for unperturbed species), n doubles.
 * No counterpart in the reference (the loader extension is this library's). */
int emo_loader_dx_deps(const emo_loader *ld, const emo_config *cfg, double *dx_deps);

/* No counterpart in the reference (whose particle vector has no layout).
 * FP32 chosen form: out3 = {slots in each particle record buffer, slots
 * covered by the cells' ranges, layout rebuilds so far}.  A cell owns a slot
 * range of its count plus headroom; a step whose arrivals exceed a cell's
 * headroom rebuilds the ranges on the device.  EMO_ERR_UNSUPPORTED in the
 * FP64 plain form. */
int emo_layout_info(emo_ctx *ctx, int64_t *out3);

/* ---- stream / timing ---------------------------------------------------- */
int emo_synchronize(emo_ctx *ctx);
/* Record HIP events around every launch of the step's dominant kernels on
 * the context stream; emo_kernel_time returns the summed milliseconds and
 * launch count per kernel name since the last reset. */
int emo_timing_enable(emo_ctx *ctx, int32_t on);
/* Restrict the event timing to the launches of one kernel name (NULL or ""
 * times them all), so that a timed run of many steps carries two events per
 * step instead of two per launch.  No counterpart in the reference. */
int emo_timing_select(emo_ctx *ctx, const char *name);
int emo_kernel_time(emo_ctx *ctx, const char *name, double *ms_total, int64_t *launches);
int emo_timing_reset(emo_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
