/*
 * emo_oracle.h -- CPU oracle for the EMOOPIC per-timestep PIC loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, FP64, single-threaded
 * restatement of the reference's serial algorithm (DenSto/APC_524_Project,
 * src/main.cpp:221-287 and the files it calls), defects included.  It is the
 * checker that the HIP product path is compared against.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it; the
 * product library never links or calls it.
 *
 * Deliberate, documented differences from the reference's control flow that
 * do not change per-particle arithmetic:
 *   - particles are re-sorted every step by a stable counting sort on the
 *     ascending cell id (the reference sorts every nstep_sort steps, std::sort,
 *     descending, unstable: src/particles/particle_handler.cpp:180-183);
 *   - current/charge are accumulated per cell in array order and added to the
 *     grid in ascending cell order, every cell visited (the reference flushes
 *     run-length groups in array order: particle_handler.cpp:203-236);
 *   - periodic particle BCs are applied in place, side by side, instead of
 *     through ghost copies + swap-pop (bc_p_periodic.cpp:44-87,
 *     particle_handler.cpp:271-284); the per-particle result is identical.
 *
 * Parity status of each function against the reference is recorded in the
 * tests that pin it (tests/test_oracle_*.py).
 */
#ifndef EMO_ORACLE_H
#define EMO_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define EO_NSPEC 32
#define EO_NWAVE 32

/* boundary kinds */
#define EO_BC_PERIODIC 0
#define EO_BC_EXTERNAL 1   /* fields */
#define EO_BC_REFLECTING 1 /* particles */

typedef struct {
  int32_t n[3];
  double x0[3];
  double L[3];
  int32_t relativity;      /* 0 Boris, 1 Vay relativistic Boris */
  int32_t fbc[6];          /* field BC per side x-,x+,y-,y+,z-,z+ */
  int32_t pbc[6];          /* particle BC per side */
  int32_t nspecies;
  double q[EO_NSPEC];      /* super-particle charge (after ProcessInfo) */
  double m[EO_NSPEC];      /* super-particle mass */
  int32_t is_test[EO_NSPEC];
  double E0[3], B0[3];
  int32_t nwaves;
  int32_t in_side[EO_NWAVE];
  int32_t in_pol[EO_NWAVE];
  double peakamps[EO_NWAVE], omegas[EO_NWAVE], phases[EO_NWAVE];
  double inv_widths[EO_NWAVE], delays[EO_NWAVE];
  double waist;            /* 0: plane-uniform pulse (reference); >0 transverse Gaussian (extension) */
  double bound_phi[6], bound_Ax[6], bound_Ay[6], bound_Az[6];
} eo_config;

typedef struct eo_state eo_state;

/* ---- lifecycle ---------------------------------------------------------- */
eo_state *eo_create(const eo_config *cfg, int64_t max_particles);
void eo_destroy(eo_state *s);

/* ---- particles ------------------------------------------------------------ */
/* x, v: [np][3]; gamma: [np]; species: [np]; id: [np] */
int eo_set_particles(eo_state *s, int64_t np, const double *x, const double *v,
                     const double *gamma, const int32_t *species, const int64_t *id);
int64_t eo_num_particles(const eo_state *s);
void eo_get_particles(const eo_state *s, double *x, double *v, double *gamma,
                      int32_t *species, int64_t *id, double *field6);
void eo_set_particle_fields(eo_state *s, const double *field6);

/* ---- fields: ids 0..2 E, 3..5 B, 6..8 J, 9 rho, 10..12 B_tm1,
 *      13 phi1, 14..16 A1, 17 phi2, 18..20 A2 (reference ids, src/grid/grid.cpp:45-60,
 *      src/poisson/poisson.cpp:18-26); arrays are padded (n+3)^3, z fastest. */
int64_t eo_field_points(const eo_state *s);
int eo_get_field(const eo_state *s, int fid, double *out);
int eo_set_field(eo_state *s, int fid, const double *in);

/* Q23 aliasing: odd != 0 makes Grid::fieldPtr_[B] alias the member B_tm1
 * (the state after an odd number of evolveFields calls), 0 the member B. */
void eo_set_b_alias(eo_state *s, int odd);

/* ---- loader: Particle_Handler::Load + Input::ProcessInfo + zeroAvg ------ */
typedef struct {
  int64_t nparticles_tot;
  int32_t nspecies;
  double mass_ratio[EO_NSPEC], charge_ratio[EO_NSPEC], dens_frac[EO_NSPEC], temp[EO_NSPEC];
  int32_t is_test[EO_NSPEC];
  double dens_phys;
  int64_t seed;                /* reference hard-codes -1 */
  int32_t zero_avg;            /* -1: reference rule (all field BCs periodic && poisson init) */
  int32_t poisson_init;        /* fields_init == "poisson" */
  /* extensions (not in the reference; default 0 = off) */
  double pert_eps[EO_NSPEC];   /* displacement loader x -= (eps/k) sin(k x) along x */
  double pert_k[EO_NSPEC];
  double drift[EO_NSPEC][3];   /* added to v after sampling */
} eo_load_config;

/* fills cfg->q/m/is_test (super-particle scaled) and loads particles */
int eo_process_and_load(eo_state *s, eo_config *cfg_inout, const eo_load_config *lc);

/* ---- RNG (NR ran2 + Box-Muller, src/utils/RNG.cpp:23-141) ------------- */
typedef struct {
  int64_t idum, idum2, iy, iv[32];
  double z0, z1;
  int32_t generate;
} eo_rng;
void eo_rng_init(eo_rng *r, int64_t seed);
double eo_rng_uniform(eo_rng *r);
double eo_rng_gaussian(eo_rng *r, double mu, double sigma);

/* ---- the per-step phases ------------------------------------------------ */
void eo_push(eo_state *s, double dt);
void eo_particle_bc(eo_state *s);
/* -4 (nothing sorted) when a particle lies outside the padded grid or is NaN */
int eo_sort(eo_state *s);
void eo_deposit(eo_state *s, int deposit_rho);
void eo_fold(eo_state *s);
void eo_evolve(eo_state *s, double dt);
void eo_refresh(eo_state *s, double t, double dt);
void eo_gather(eo_state *s);
/* single-cell helpers used by the known-answer tests */
int32_t eo_cell_id(const eo_state *s, double x, double y, double z);

/* Poisson initialisation: zeroes E,B,B_tm1,phi,A (Q8), solves phi and A with
 * the reference's Jacobi loop, phiToE, AToB, executeBC(-1).  sweeps_out[4]
 * gets the number of sweeps run per solve.  iter_max mirrors ITER_MAX
 * (src/poisson/poisson.cpp:14): the loop runs while not converged and
 * iternum < iter_max, i.e. at most iter_max+1 sweeps (reference: 10000).
 * tol < 0 selects the reference tolerance rule (poisson.cpp:151,157). */
void eo_poisson_init(eo_state *s, int32_t iter_max, double tol, int32_t *sweeps_out,
                     double t, double dt);
/* constant initialisation: Grid::InitializeFields (src/grid/grid.cpp:554-574) */
void eo_const_init(eo_state *s);

/* history.dat columns (src/IO/output.cpp:76-118):
 * out = {pMx, pMy, pMz, pEn, B_en, E_en, E_tot, total_charge(extension)} */
void eo_diag(const eo_state *s, double *out8);

/* one full step in the reference order (src/main.cpp:231-275) */
int eo_step(eo_state *s, double t, double dt); /* -4: see eo_sort */

/* the pulse scalar used by LightBC (src/utils/evaluateGaussianPulse.cpp:7-30) */
double eo_gaussian_pulses(int nw, const double *amp, const double *omg, const double *pha,
                          const double *del, const double *dwt, double t, double dt);

#ifdef __cplusplus
}
#endif
#endif
