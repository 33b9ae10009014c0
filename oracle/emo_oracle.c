/*
 * emo_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY, see emo_oracle.h).
 *
 * FP64 restatement of EMOOPIC's serial per-timestep loop.  Built with
 * -O2 -ffp-contract=off so that every expression rounds exactly as the
 * reference's -O3 build without FMA does (src/Makefile:5 has no -march).
 * Each function cites the reference file:line it restates.  Reference paths
 * are relative to the DenSto/APC_524_Project checkout.
 */
#define _GNU_SOURCE
#include "emo_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* src/globals.hpp:12,16 */
#define UNIT_ACC 0.58668774
#define UNIT_RHOJ 4.80320427e-13
#define UNIT_VTH 0.002232
#define EO_PI 3.14159265358979323846

/* field ids, src/grid/grid.cpp:45-60 and src/poisson/poisson.cpp:18-26 */
enum { FEX, FEY, FEZ, FBX, FBY, FBZ, FJX, FJY, FJZ, FRHO, FBTX, FBTY, FBTZ,
       FPHI1, FAX1, FAY1, FAZ1, FPHI2, FAX2, FAY2, FAZ2, NFIELD };

struct eo_state {
  eo_config c;
  int n[3];
  int nt[3];     /* n+3: padded points per dim (src/grid/grid.cpp:25-27) */
  int nc[3];     /* n+2: cell-id dims nx_,ny_,nz_ (src/grid/grid.cpp:22-24) */
  int64_t ntot;  /* padded points */
  int64_t ncell; /* nc0*nc1*nc2 */
  double d[3], id[3];
  double *f[NFIELD];
  /* Grid::fieldPtr_[BxID_..BzID_]: fixed allocations that evolveFields'
   * member swap never updates (src/grid/oGrid.cpp:18-20 vs
   * src/grid/grid.cpp:321-342); the ghost refresh and the diagnostics go
   * through them (spookyGrid.cpp:87,302; output.cpp:99) */
  double *fBfix[3];
  /* particles, [np][3] like Particle::x (src/particles/particle.hpp:10-13) */
  int64_t cap, np;
  double *x, *v, *gam, *fld;
  int32_t *sp;
  int64_t *pid;
  /* scratch */
  double *tx, *tv, *tg, *tf;
  int32_t *tsp;
  int64_t *tid;
  int32_t *key;
  int64_t *cstart; /* ncell+1 */
  double *cj;      /* 12 per cell */
  double *cr;      /* 8 per cell */
};

static inline int64_t IX(const eo_state *s, int i, int j, int k) {
  return ((int64_t)i * s->nt[1] + j) * s->nt[2] + k;
}

eo_state *eo_create(const eo_config *cfg, int64_t max_particles) {
  eo_state *s = (eo_state *)calloc(1, sizeof(eo_state));
  s->c = *cfg;
  for (int d = 0; d < 3; d++) {
    s->n[d] = cfg->n[d];
    s->nt[d] = cfg->n[d] + 3;
    s->nc[d] = cfg->n[d] + 2;
    /* dx_(Lxyz[0]/nxReal_), idx_(1.0/dx_): src/grid/grid.cpp:37-42 */
    s->d[d] = cfg->L[d] / cfg->n[d];
    s->id[d] = 1.0 / s->d[d];
  }
  s->ntot = (int64_t)s->nt[0] * s->nt[1] * s->nt[2];
  s->ncell = (int64_t)s->nc[0] * s->nc[1] * s->nc[2];
  for (int f = 0; f < NFIELD; f++) s->f[f] = (double *)calloc(s->ntot, sizeof(double));
  for (int c = 0; c < 3; c++) s->fBfix[c] = s->f[FBX + c];
  s->cap = max_particles > 0 ? max_particles : 1;
  s->x = (double *)calloc(3 * s->cap, sizeof(double));
  s->v = (double *)calloc(3 * s->cap, sizeof(double));
  s->gam = (double *)calloc(s->cap, sizeof(double));
  s->fld = (double *)calloc(6 * s->cap, sizeof(double));
  s->sp = (int32_t *)calloc(s->cap, sizeof(int32_t));
  s->pid = (int64_t *)calloc(s->cap, sizeof(int64_t));
  s->tx = (double *)calloc(3 * s->cap, sizeof(double));
  s->tv = (double *)calloc(3 * s->cap, sizeof(double));
  s->tg = (double *)calloc(s->cap, sizeof(double));
  s->tf = (double *)calloc(6 * s->cap, sizeof(double));
  s->tsp = (int32_t *)calloc(s->cap, sizeof(int32_t));
  s->tid = (int64_t *)calloc(s->cap, sizeof(int64_t));
  s->key = (int32_t *)calloc(s->cap, sizeof(int32_t));
  s->cstart = (int64_t *)calloc(s->ncell + 1, sizeof(int64_t));
  s->cj = (double *)calloc(12 * s->ncell, sizeof(double));
  s->cr = (double *)calloc(8 * s->ncell, sizeof(double));
  return s;
}

void eo_destroy(eo_state *s) {
  if (!s) return;
  for (int f = 0; f < NFIELD; f++) free(s->f[f]);
  free(s->x); free(s->v); free(s->gam); free(s->fld); free(s->sp); free(s->pid);
  free(s->tx); free(s->tv); free(s->tg); free(s->tf); free(s->tsp); free(s->tid);
  free(s->key); free(s->cstart); free(s->cj); free(s->cr);
  free(s);
}

int eo_set_particles(eo_state *s, int64_t np, const double *x, const double *v,
                     const double *gamma, const int32_t *species, const int64_t *id) {
  if (np > s->cap) return -1;
  s->np = np;
  memcpy(s->x, x, 3 * np * sizeof(double));
  memcpy(s->v, v, 3 * np * sizeof(double));
  memcpy(s->gam, gamma, np * sizeof(double));
  memcpy(s->sp, species, np * sizeof(int32_t));
  memcpy(s->pid, id, np * sizeof(int64_t));
  memset(s->fld, 0, 6 * np * sizeof(double));
  return 0;
}

int64_t eo_num_particles(const eo_state *s) { return s->np; }

void eo_get_particles(const eo_state *s, double *x, double *v, double *gamma,
                      int32_t *species, int64_t *id, double *field6) {
  int64_t np = s->np;
  if (x) memcpy(x, s->x, 3 * np * sizeof(double));
  if (v) memcpy(v, s->v, 3 * np * sizeof(double));
  if (gamma) memcpy(gamma, s->gam, np * sizeof(double));
  if (species) memcpy(species, s->sp, np * sizeof(int32_t));
  if (id) memcpy(id, s->pid, np * sizeof(int64_t));
  if (field6) memcpy(field6, s->fld, 6 * np * sizeof(double));
}

void eo_set_particle_fields(eo_state *s, const double *field6) {
  memcpy(s->fld, field6, 6 * s->np * sizeof(double));
}

int64_t eo_field_points(const eo_state *s) { return s->ntot; }

int eo_get_field(const eo_state *s, int fid, double *out) {
  if (fid < 0 || fid >= NFIELD) return -1;
  memcpy(out, s->f[fid], s->ntot * sizeof(double));
  return 0;
}

void eo_set_b_alias(eo_state *s, int odd) {
  for (int c = 0; c < 3; c++) s->fBfix[c] = odd ? s->f[FBTX + c] : s->f[FBX + c];
}

int eo_set_field(eo_state *s, int fid, const double *in) {
  if (fid < 0 || fid >= NFIELD) return -1;
  memcpy(s->f[fid], in, s->ntot * sizeof(double));
  return 0;
}

/* ------------------------------------------------------------------------ */
/* RNG: Numerical Recipes ran2 (Press et al., NR in C 2nd ed. s7.1) as used  */
/* by src/utils/RNG.cpp:23-47 (seeding) and :68-88 (draw); Box-Muller with   */
/* the cached second value, src/utils/RNG.cpp:118-141.                       */
/* ------------------------------------------------------------------------ */
#define R_IM1 2147483563L
#define R_IM2 2147483399L
#define R_IMM1 (R_IM1 - 1)
#define R_IA1 40014L
#define R_IA2 40692L
#define R_IQ1 53668L
#define R_IQ2 52774L
#define R_IR1 12211L
#define R_IR2 3791L
#define R_NTAB 32
#define R_NDIV (1 + R_IMM1 / R_NTAB)

void eo_rng_init(eo_rng *r, int64_t seed) {
  int64_t idum = seed;
  if (idum == 0) idum = 1;
  else if (idum < 0) idum = -idum;
  r->idum2 = idum;
  for (int j = R_NTAB + 7; j >= 0; j--) {
    int64_t k = idum / R_IQ1;
    idum = R_IA1 * (idum - k * R_IQ1) - k * R_IR1;
    if (idum < 0) idum += R_IM1;
    if (j < R_NTAB) r->iv[j] = idum;
  }
  r->idum = idum;
  r->iy = r->iv[0];
  r->generate = 1;
  r->z0 = r->z1 = 0.0;
}

double eo_rng_uniform(eo_rng *r) {
  int64_t k = r->idum / R_IQ1;
  r->idum = R_IA1 * (r->idum - k * R_IQ1) - k * R_IR1;
  if (r->idum < 0) r->idum += R_IM1;
  k = r->idum2 / R_IQ2;
  r->idum2 = R_IA2 * (r->idum2 - k * R_IQ2) - k * R_IR2;
  if (r->idum2 < 0) r->idum2 += R_IM2;
  int j = (int)(r->iy / R_NDIV);
  r->iy = r->iv[j] - r->idum2;
  r->iv[j] = r->idum;
  if (r->iy < 1) r->iy += R_IMM1;
  double temp = (1.0 / 2147483563) * r->iy;
  const double rnmx = 1.0 - DBL_EPSILON;
  return temp > rnmx ? rnmx : temp;
}

double eo_rng_gaussian(eo_rng *r, double mu, double sigma) {
  const double epsilon = DBL_MIN;
  const double two_pi = 2.0 * 3.14159265358979323846;
  if (!r->generate) {
    r->generate = !r->generate;
    return r->z1 * sigma + mu;
  }
  double u1, u2;
  do {
    u1 = eo_rng_uniform(r);
    u2 = eo_rng_uniform(r);
  } while (u1 <= epsilon);
  /* g++ -O3 (the reference's build, src/Makefile:5) merges the cos and sin
   * of one argument into a single glibc sincos call; do so explicitly */
  double sn, cs;
  sincos(two_pi * u2, &sn, &cs);
  r->z0 = sqrt(-2.0 * log(u1)) * cs;
  r->z1 = sqrt(-2.0 * log(u1)) * sn;
  r->generate = !r->generate;
  return r->z0 * sigma + mu;
}

/* ------------------------------------------------------------------------ */
/* Loader: Input::ProcessInfo scaling (src/IO/check.cpp:53-111), Domain's    */
/* serial particle count (src/domain/domain.cpp:96-100), Particle_Handler::  */
/* Load (src/particles/particle_handler.cpp:34-85) and                      */
/* zeroAvgChargeAndVelocity_ (:379-444).                                     */
/* ------------------------------------------------------------------------ */
static int zero_avg(eo_state *s) {
  int64_t np = s->np;
  double qTot = 0.0, vx, vy, vz;
  int64_t qRefIdx, qCount;
  double min_q = (double)FLT_MAX;
  unsigned char *counted = (unsigned char *)calloc(np > 0 ? np : 1, 1);
  do {
    qRefIdx = 0;
    qCount = 0;
    vx = vy = vz = 0.0;
    for (int64_t i = 0; i < np; i++) {
      if (counted[i] == 0) {
        if (qRefIdx == 0) qRefIdx = i;
        if (s->c.q[s->sp[i]] == s->c.q[s->sp[qRefIdx]]) {
          counted[i] = 1;
          qCount += 1;
          vx += s->v[3 * i + 0];
          vy += s->v[3 * i + 1];
          vz += s->v[3 * i + 2];
        }
      }
    }
    if (qCount > 0) {
      vx /= qCount;
      vy /= qCount;
      vz /= qCount;
      for (int64_t i = 0; i < np; i++) {
        if (counted[i] == 1) {
          counted[i] = 2;
          s->v[3 * i + 0] -= vx;
          s->v[3 * i + 1] -= vy;
          s->v[3 * i + 2] -= vz;
          qTot += s->c.q[s->sp[i]];
          double aq = fabs(s->c.q[s->sp[i]]);
          min_q = aq < min_q ? aq : min_q;
        }
      }
    }
  } while (qCount > 0);
  free(counted);
  /* the reference asserts here (particle_handler.cpp:435-438) */
  return (fabs(qTot) >= min_q / np) ? -2 : 0;
}

int eo_process_and_load(eo_state *s, eo_config *cfg, const eo_load_config *lc) {
  int nspec = lc->nspecies;
  double total_volume = 1.0;
  for (int i = 0; i < 3; i++) total_volume *= cfg->L[i];
  int64_t nparticles_tot = lc->nparticles_tot;
  double n0 = nparticles_tot / total_volume;
  double dens_phys = lc->dens_phys;
  if (dens_phys < 0) dens_phys = -dens_phys;
  else if (dens_phys == 0) dens_phys = n0;
  double super_ratio = n0 > 0 ? dens_phys / n0 : -1.0;
  double dens[EO_NSPEC], mass[EO_NSPEC], charge[EO_NSPEC], temp[EO_NSPEC];
  double cden = 0.0;
  for (int i = 0; i < nspec; i++) {
    dens[i] = lc->dens_frac[i];
    mass[i] = lc->mass_ratio[i];
    charge[i] = lc->charge_ratio[i];
    temp[i] = lc->temp[i];
    cden += dens[i];
  }
  for (int i = 0; i < nspec; i++) {
    if (temp[i] < 0) temp[i] = -temp[i];
    if (mass[i] < 0) mass[i] = -mass[i];
    if (mass[i] == 0) return -1;
  }
  if (super_ratio > 0) {
    for (int i = 0; i < nspec; i++) {
      dens[i] /= cden;
      mass[i] *= super_ratio;
      charge[i] *= super_ratio;
      temp[i] *= super_ratio;
    }
  }
  cfg->nspecies = nspec;
  for (int i = 0; i < nspec; i++) {
    cfg->q[i] = charge[i];
    cfg->m[i] = mass[i];
    cfg->is_test[i] = lc->is_test[i];
  }
  s->c.nspecies = nspec;
  for (int i = 0; i < nspec; i++) {
    s->c.q[i] = charge[i];
    s->c.m[i] = mass[i];
    s->c.is_test[i] = lc->is_test[i];
  }
  /* serial domain: nparticles_domain = nparticles_tot (domain.cpp:96-100) */
  int64_t npart = nparticles_tot;
  if (npart > s->cap) return -3;
  eo_rng rng;
  eo_rng_init(&rng, lc->seed);
  int ispec = 0;
  double cd = dens[0];
  for (int64_t ip = 0; ip < npart; ip++) {
    if (ip >= cd * npart) {
      ispec += 1;
      cd += dens[ispec];
    }
    double m = mass[ispec];
    double vth = UNIT_VTH * sqrt(temp[ispec] / m);
    double *x = s->x + 3 * ip, *v = s->v + 3 * ip;
    x[0] = eo_rng_uniform(&rng) * cfg->L[0] + cfg->x0[0];
    x[1] = eo_rng_uniform(&rng) * cfg->L[1] + cfg->x0[1];
    x[2] = eo_rng_uniform(&rng) * cfg->L[2] + cfg->x0[2];
    v[0] = eo_rng_gaussian(&rng, 0.0, vth);
    v[1] = eo_rng_gaussian(&rng, 0.0, vth);
    v[2] = eo_rng_gaussian(&rng, 0.0, vth);
    /* extensions (off by default): displacement loader and drift */
    if (lc->pert_eps[ispec] != 0.0) {
      double k = lc->pert_k[ispec];
      double xr = x[0] - cfg->x0[0];
      xr = xr - (lc->pert_eps[ispec] / k) * sin(k * xr);
      if (xr < 0) xr += cfg->L[0];
      if (xr >= cfg->L[0]) xr -= cfg->L[0];
      x[0] = xr + cfg->x0[0];
    }
    v[0] += lc->drift[ispec][0];
    v[1] += lc->drift[ispec][1];
    v[2] += lc->drift[ispec][2];
    /* Q3: gamma = sqrt(1 - v^2), particle_handler.cpp:77 */
    s->gam[ip] = sqrt(1 - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
    s->sp[ip] = ispec;
    s->pid[ip] = ip;
  }
  s->np = npart;
  memset(s->fld, 0, 6 * npart * sizeof(double));
  int all_periodic = 1;
  for (int i = 0; i < 6; i++)
    if (cfg->fbc[i] != EO_BC_PERIODIC) all_periodic = 0;
  int za = lc->zero_avg < 0 ? (all_periodic && lc->poisson_init) : lc->zero_avg;
  if (za) return zero_avg(s);
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Cell id: Grid::getCellID, src/grid/oGrid.cpp:208-222                      */
/* ------------------------------------------------------------------------ */
int32_t eo_cell_id(const eo_state *s, double x, double y, double z) {
  int ix = (int)((x - s->c.x0[0]) * s->id[0]);
  ix += 1;
  int iy = (int)((y - s->c.x0[1]) * s->id[1]);
  iy += 1;
  int iz = (int)((z - s->c.x0[2]) * s->id[2]);
  iz += 1;
  return (s->nc[1] * s->nc[2]) * ix + s->nc[2] * iy + iz;
}

/* -1 for a position outside the padded grid (or NaN): the reference would
   index out of bounds there; the oracle reports it instead. */
static int32_t cell_checked(const eo_state *s, const double *p) {
  for (int d = 0; d < 3; d++) {
    double u = (p[d] - s->c.x0[d]) * s->id[d];
    if (!(u >= 0.0 && u < (double)(s->c.n[d] + 1))) return -1;
  }
  return eo_cell_id(s, p[0], p[1], p[2]);
}

static inline void cell_decode(const eo_state *s, int32_t cid, int *ix, int *iy, int *iz) {
  /* src/grid/oGrid.cpp:155-157 */
  int z = cid % s->nc[2];
  int y = ((cid - z) / s->nc[2]) % s->nc[1];
  int x = (((cid - z) / s->nc[2]) - y) / s->nc[1];
  *ix = x;
  *iy = y;
  *iz = z;
}

/* ------------------------------------------------------------------------ */
/* Push: Boris::Step src/pusher/boris.cpp:12-75; Relativistic_Boris::Step   */
/* src/pusher/relativisticBoris.cpp:28-140.  Particle_Handler::Push loops   */
/* over all particles, src/particles/particle_handler.cpp:124-128.          */
/* ------------------------------------------------------------------------ */
static void boris_step(double *xp, double *vp, const double *fld, double q, double m, double dt) {
  double x = xp[0], y = xp[1], z = xp[2];
  double vx = vp[0], vy = vp[1], vz = vp[2];
  double q_p = (0.5 * dt * q) / m;
  q_p *= UNIT_ACC;
  double ex = fld[0], ey = fld[1], ez = fld[2];
  double bx = fld[3], by = fld[4], bz = fld[5];
  double tx = q_p * bx, ty = q_p * by, tz = q_p * bz;
  double bsqr = bx * bx + by * by + bz * bz;
  double tsqr = 2.0 / (1.0 + q_p * q_p * bsqr);
  double sx = tsqr * tx, sy = tsqr * ty, sz = tsqr * tz;
  double v_mx = vx + q_p * ex;
  double v_my = vy + q_p * ey;
  double v_mz = vz + q_p * ez;
  double v_ix = v_mx + v_my * tz - v_mz * ty;
  double v_iy = v_my + v_mz * tx - v_mx * tz;
  double v_iz = v_mz + v_mx * ty - v_my * tx;
  double v_px = v_mx + v_iy * sz - v_iz * sy;
  double v_py = v_my + v_iz * sx - v_ix * sz;
  double v_pz = v_mz + v_ix * sy - v_iy * sx;
  double vx_new = v_px + q_p * ex;
  double vy_new = v_py + q_p * ey;
  double vz_new = v_pz + q_p * ez;
  xp[0] = x + dt * vx_new;
  xp[1] = y + dt * vy_new;
  xp[2] = z + dt * vz_new;
  vp[0] = vx_new;
  vp[1] = vy_new;
  vp[2] = vz_new;
}

static void vay_step(double *xp, double *vp, double *gp, const double *fld, double q, double m,
                     double dt) {
  double x = xp[0], y = xp[1], z = xp[2];
  double vx = vp[0], vy = vp[1], vz = vp[2];
  double gamma = *gp;
  double q_p = (0.5 * dt * q) / m;
  q_p *= UNIT_ACC;
  double ex = fld[0], ey = fld[1], ez = fld[2];
  double bx = fld[3], by = fld[4], bz = fld[5];
  double vXbx = vy * bz - vz * by;
  double vXby = vz * bx - vx * bz;
  double vXbz = vx * by - vy * bx;
  double uhx = gamma * vx + q_p * (ex + vXbx);
  double uhy = gamma * vy + q_p * (ey + vXby);
  double uhz = gamma * vz + q_p * (ez + vXbz);
  double upx = uhx + q_p * ex;
  double upy = uhy + q_p * ey;
  double upz = uhz + q_p * ez;
  double gammap_sq = 1.0 + upx * upx + upy * upy + upz * upz;
  double taux = q_p * bx, tauy = q_p * by, tauz = q_p * bz;
  double tau_sq = taux * taux + tauy * tauy + tauz * tauz;
  double sigma = gammap_sq - tau_sq;
  double ustar = upx * taux + upy * tauy + upz * tauz;
  double gamma_int = sqrt(sigma * sigma + 4 * (tau_sq + ustar * ustar));
  gamma_int = sqrt(0.5 * (sigma + gamma_int));
  double tx = taux / gamma_int, ty = tauy / gamma_int, tz = tauz / gamma_int;
  double s = 1.0 / (1.0 + tx * tx + ty * ty + tz * tz);
  double udott = ustar / gamma_int;
  double upXt_x = upy * tz - upz * ty;
  double upXt_y = upz * tx - upx * tz;
  double upXt_z = upx * ty - upy * tx;
  double uint_x = s * (upx + udott * tx + upXt_x);
  double uint_y = s * (upy + udott * ty + upXt_y);
  double uint_z = s * (upz + udott * tz + upXt_z);
  double vx_new = uint_x / gamma_int;
  double vy_new = uint_y / gamma_int;
  double vz_new = uint_z / gamma_int;
  xp[0] = x + dt * vx_new;
  xp[1] = y + dt * vy_new;
  xp[2] = z + dt * vz_new;
  vp[0] = vx_new;
  vp[1] = vy_new;
  vp[2] = vz_new;
  *gp = gamma_int;
}

void eo_push(eo_state *s, double dt) {
  for (int64_t i = 0; i < s->np; i++) {
    int sp = s->sp[i];
    if (s->c.relativity)
      vay_step(s->x + 3 * i, s->v + 3 * i, s->gam + i, s->fld + 6 * i, s->c.q[sp], s->c.m[sp], dt);
    else
      boris_step(s->x + 3 * i, s->v + 3 * i, s->fld + 6 * i, s->c.q[sp], s->c.m[sp], dt);
  }
}

/* ------------------------------------------------------------------------ */
/* Particle BCs: six sides in the order x-,x+,y-,y+,z-,z+                    */
/* (src/particles/particle_handler.cpp:287-295, particle_bc_factory.cpp:    */
/* 98-104).  Periodic: strict compares, shift by (xMax-xMin)                 */
/* (bc_p_periodic.cpp:44-87); reflecting: mirror x and flip v                */
/* (bc_p_reflecting.cpp:43-53).  Applied in place.                           */
/* ------------------------------------------------------------------------ */
void eo_particle_bc(eo_state *s) {
  for (int side = 0; side < 6; side++) {
    int d = side / 2, right = side & 1;
    double xMin = s->c.x0[d];
    double xMax = xMin + s->c.L[d];
    for (int64_t i = 0; i < s->np; i++) {
      double *xp = s->x + 3 * i, *vp = s->v + 3 * i;
      if (s->c.pbc[side] == EO_BC_PERIODIC) {
        if (!right && xp[d] < xMin) xp[d] += (xMax - xMin);
        if (right && xp[d] > xMax) xp[d] -= (xMax - xMin);
      } else {
        if (right && xp[d] > xMax) {
          xp[d] = 2.0 * xMax - xp[d];
          vp[d] = -vp[d];
        }
        if (!right && xp[d] < xMin) {
          xp[d] = 2.0 * xMin - xp[d];
          vp[d] = -vp[d];
        }
      }
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Stable counting sort by ascending cell id (replaces Particle_Compare +   */
/* std::sort, src/particles/particle_utils.hpp:31-50; only cell membership  */
/* is part of the contract).                                                 */
/* ------------------------------------------------------------------------ */
int eo_sort(eo_state *s) {
  int64_t np = s->np;
  for (int64_t i = 0; i < np; i++)
    if (cell_checked(s, s->x + 3 * i) < 0) return -4;
  memset(s->cstart, 0, (s->ncell + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < np; i++) {
    int32_t c = eo_cell_id(s, s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2]);
    s->key[i] = c;
    s->cstart[c + 1]++;
  }
  for (int64_t c = 0; c < s->ncell; c++) s->cstart[c + 1] += s->cstart[c];
  int64_t *pos = (int64_t *)malloc((s->ncell + 1) * sizeof(int64_t));
  memcpy(pos, s->cstart, (s->ncell + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < np; i++) {
    int64_t dst = pos[s->key[i]]++;
    memcpy(s->tx + 3 * dst, s->x + 3 * i, 3 * sizeof(double));
    memcpy(s->tv + 3 * dst, s->v + 3 * i, 3 * sizeof(double));
    memcpy(s->tf + 6 * dst, s->fld + 6 * i, 6 * sizeof(double));
    s->tg[dst] = s->gam[i];
    s->tsp[dst] = s->sp[i];
    s->tid[dst] = s->pid[i];
  }
  free(pos);
  double *t;
  t = s->x; s->x = s->tx; s->tx = t;
  t = s->v; s->v = s->tv; s->tv = t;
  t = s->fld; s->fld = s->tf; s->tf = t;
  t = s->gam; s->gam = s->tg; s->tg = t;
  int32_t *ti = s->sp; s->sp = s->tsp; s->tsp = ti;
  int64_t *tl = s->pid; s->pid = s->tid; s->tid = tl;
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Deposit: Particle_Handler::depositRhoJ (particle_handler.cpp:185-236),   */
/* Depositor::deposit_particle_J/_Rho (src/particles/deposit.cpp:12-62),    */
/* Grid::getCellVertex (src/grid/oGrid.cpp:233-244), Grid::addJ/addRho      */
/* (oGrid.cpp:91-139) including the duplicate Jx[ix][iy+1][iz] target.      */
/* Requires the particles sorted by cell (eo_sort).                          */
/* ------------------------------------------------------------------------ */
static void deposit_J(const double *x, const double *v, double q, const double *lcell,
                      const double *cellverts, double *JObj) {
  double cellVolume = lcell[0] * lcell[1] * lcell[2];
  double relpos[3];
  for (int i = 0; i < 3; i++) relpos[i] = x[i] - cellverts[i];
  double pcharge = q;
  for (int i = 0; i < 3; i++) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    double cellArea = lcell[j] * lcell[k];
    double curIn1D = UNIT_RHOJ * pcharge * v[i];
    double cur2AreaRatio = (curIn1D / cellVolume) / cellArea;
    JObj[4 * i] += cur2AreaRatio * (lcell[j] - relpos[j]) * (lcell[k] - relpos[k]);
    JObj[4 * i + 1] += cur2AreaRatio * relpos[j] * (lcell[k] - relpos[k]);
    JObj[4 * i + 2] += cur2AreaRatio * relpos[j] * relpos[k];
    JObj[4 * i + 3] += cur2AreaRatio * (lcell[j] - relpos[j]) * relpos[k];
  }
}

static void deposit_Rho(const double *x, double q, const double *lcell, const double *cellverts,
                        double *R) {
  double cellVolume = lcell[0] * lcell[1] * lcell[2];
  double relpos[3];
  for (int i = 0; i < 3; i++) relpos[i] = x[i] - cellverts[i];
  double pcharge = q;
  double c = UNIT_RHOJ * (pcharge / cellVolume) / cellVolume;
  R[0] += c * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
  R[1] += c * relpos[0] * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
  R[2] += c * relpos[0] * relpos[1] * (lcell[2] - relpos[2]);
  R[3] += c * (lcell[0] - relpos[0]) * relpos[1] * (lcell[2] - relpos[2]);
  R[4] += c * (lcell[0] - relpos[0]) * relpos[1] * relpos[2];
  R[5] += c * relpos[0] * relpos[1] * relpos[2];
  R[6] += c * relpos[0] * (lcell[1] - relpos[1]) * relpos[2];
  R[7] += c * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * relpos[2];
}

void eo_deposit(eo_state *s, int deposit_rho) {
  double *Jx = s->f[FJX], *Jy = s->f[FJY], *Jz = s->f[FJZ], *rho = s->f[FRHO];
  /* constJ(0,0,0); constRho(0): whole padded arrays, src/grid/grid.cpp:509-549 */
  memset(Jx, 0, s->ntot * sizeof(double));
  memset(Jy, 0, s->ntot * sizeof(double));
  memset(Jz, 0, s->ntot * sizeof(double));
  if (deposit_rho) memset(rho, 0, s->ntot * sizeof(double));
  /* per-cell accumulation in array order */
  memset(s->cstart, 0, (s->ncell + 1) * sizeof(int64_t));
  for (int64_t i = 0; i < s->np; i++) {
    int32_t c = eo_cell_id(s, s->x[3 * i], s->x[3 * i + 1], s->x[3 * i + 2]);
    s->key[i] = c;
  }
  memset(s->cj, 0, 12 * s->ncell * sizeof(double));
  memset(s->cr, 0, 8 * s->ncell * sizeof(double));
  for (int64_t i = 0; i < s->np; i++) {
    int sp = s->sp[i];
    if (s->c.is_test[sp]) continue; /* particle_handler.cpp:205 */
    int32_t c = s->key[i];
    int ix, iy, iz;
    cell_decode(s, c, &ix, &iy, &iz);
    double cellverts[3];
    cellverts[0] = s->c.x0[0] + s->d[0] * ix;
    cellverts[1] = s->c.x0[1] + s->d[1] * iy;
    cellverts[2] = s->c.x0[2] + s->d[2] * iz;
    deposit_J(s->x + 3 * i, s->v + 3 * i, s->c.q[sp], s->d, cellverts, s->cj + 12 * (int64_t)c);
    if (deposit_rho) deposit_Rho(s->x + 3 * i, s->c.q[sp], s->d, cellverts, s->cr + 8 * (int64_t)c);
  }
  /* addJ / addRho for every cell in ascending id */
  for (int64_t c = 0; c < s->ncell; c++) {
    int ix, iy, iz;
    cell_decode(s, (int32_t)c, &ix, &iy, &iz);
    if (ix + 1 >= s->nt[0] || iy + 1 >= s->nt[1] || iz + 1 >= s->nt[2]) continue;
    const double *Jv = s->cj + 12 * c;
    Jx[IX(s, ix, iy, iz)] += Jv[0];
    Jx[IX(s, ix, iy + 1, iz)] += Jv[1];
    Jx[IX(s, ix, iy + 1, iz + 1)] += Jv[2];
    Jx[IX(s, ix, iy + 1, iz)] += Jv[3];
    Jy[IX(s, ix, iy, iz)] += Jv[4];
    Jy[IX(s, ix, iy, iz + 1)] += Jv[5];
    Jy[IX(s, ix + 1, iy, iz + 1)] += Jv[6];
    Jy[IX(s, ix + 1, iy, iz)] += Jv[7];
    Jz[IX(s, ix, iy, iz)] += Jv[8];
    Jz[IX(s, ix + 1, iy, iz)] += Jv[9];
    Jz[IX(s, ix + 1, iy + 1, iz)] += Jv[10];
    Jz[IX(s, ix, iy + 1, iz)] += Jv[11];
    if (deposit_rho) {
      const double *R = s->cr + 8 * c;
      rho[IX(s, ix, iy, iz)] += R[0];
      rho[IX(s, ix + 1, iy, iz)] += R[1];
      rho[IX(s, ix + 1, iy + 1, iz)] += R[2];
      rho[IX(s, ix, iy + 1, iz)] += R[3];
      rho[IX(s, ix, iy + 1, iz + 1)] += R[4];
      rho[IX(s, ix + 1, iy + 1, iz + 1)] += R[5];
      rho[IX(s, ix + 1, iy, iz + 1)] += R[6];
      rho[IX(s, ix, iy, iz + 1)] += R[7];
    }
  }
}

/* ------------------------------------------------------------------------ */
/* Ghost planes: Grid::sliceMatToVec_/unsliceMatToVec_ index ranges          */
/* (src/grid/spookyGrid.cpp:286-425): the plane at index `dex` along d,      */
/* over the physical range [1..n] of the two other dims.                     */
/* ------------------------------------------------------------------------ */
static void plane_copy(eo_state *s, double *f, int d, int src, int dst, int add) {
  int a = (d + 1) % 3, b = (d + 2) % 3;
  if (a > b) { int t = a; a = b; b = t; }
  for (int p = 1; p <= s->n[a]; p++) {
    for (int q = 1; q <= s->n[b]; q++) {
      int is[3], id[3];
      is[d] = src; id[d] = dst;
      is[a] = p; id[a] = p;
      is[b] = q; id[b] = q;
      double val = f[IX(s, is[0], is[1], is[2])];
      if (add) f[IX(s, id[0], id[1], id[2])] += val;
      else f[IX(s, id[0], id[1], id[2])] = val;
    }
  }
}

/* J/rho fold, Grid::executeBC(-2) (src/grid/grid.cpp:577-585) ->            */
/* BC_F_Periodic::completeBC(-2) on the left side only                        */
/* (src/boundaries/b_fields/bc_f_periodic.cpp:40-46); the two J components   */
/* orthogonal to d plus rho (spookyGrid.cpp:54-72).  External sides do      */
/* nothing (bc_f_external.cpp:45-58).                                        */
void eo_fold(eo_state *s) {
  for (int d = 0; d < 3; d++) {
    if (s->c.fbc[2 * d] != EO_BC_PERIODIC) continue;
    int f0 = (d == 0) ? FJY : FJX;
    int f1 = (d == 2) ? FJY : FJZ;
    plane_copy(s, s->f[f0], d, s->n[d] + 1, 1, 1);
    plane_copy(s, s->f[f1], d, s->n[d] + 1, 1, 1);
    plane_copy(s, s->f[FRHO], d, s->n[d] + 1, 1, 1);
  }
}

/* ------------------------------------------------------------------------ */
/* Yee update, Grid::evolveFields (src/grid/oGrid.cpp:15-60).                */
/* ------------------------------------------------------------------------ */
void eo_evolve(eo_state *s, double dt) {
  double *t;
  for (int c = 0; c < 3; c++) {
    t = s->f[FBTX + c]; s->f[FBTX + c] = s->f[FBX + c]; s->f[FBX + c] = t;
  }
  double *Ex = s->f[FEX], *Ey = s->f[FEY], *Ez = s->f[FEZ];
  double *Bx = s->f[FBX], *By = s->f[FBY], *Bz = s->f[FBZ];
  double *Btx = s->f[FBTX], *Bty = s->f[FBTY], *Btz = s->f[FBTZ];
  double *Jx = s->f[FJX], *Jy = s->f[FJY], *Jz = s->f[FJZ];
  double dx_ = s->d[0], dy_ = s->d[1], dz_ = s->d[2];
  int64_t sy = s->nt[2], sx = (int64_t)s->nt[1] * s->nt[2];
  for (int ix = 1; ix <= s->n[0]; ix++)
    for (int iy = 1; iy <= s->n[1]; iy++)
      for (int iz = 1; iz <= s->n[2]; iz++) {
        int64_t c = IX(s, ix, iy, iz);
        Ex[c] = Ex[c] - dt * ((Bty[c] - Bty[c - 1]) / dz_ - (Btz[c] - Btz[c - sy]) / dy_ +
                              4 * EO_PI * Jx[c]);
        Ey[c] = Ey[c] - dt * ((Btz[c] - Btz[c - sx]) / dx_ - (Btx[c] - Btx[c - 1]) / dz_ +
                              4 * EO_PI * Jy[c]);
        Ez[c] = Ez[c] - dt * ((Btx[c] - Btx[c - sy]) / dy_ - (Bty[c] - Bty[c - sx]) / dx_ +
                              4 * EO_PI * Jz[c]);
      }
  for (int ix = 1; ix <= s->n[0]; ix++)
    for (int iy = 1; iy <= s->n[1]; iy++)
      for (int iz = 1; iz <= s->n[2]; iz++) {
        int64_t c = IX(s, ix, iy, iz);
        Bx[c] = Btx[c] + dt * ((Ey[c + 1] - Ey[c]) / dz_ - (Ez[c + sy] - Ez[c]) / dy_);
        By[c] = Bty[c] + dt * ((Ez[c + sx] - Ez[c]) / dx_ - (Ex[c + 1] - Ex[c]) / dz_);
        Bz[c] = Btz[c] + dt * ((Ex[c + sy] - Ex[c]) / dy_ - (Ey[c + sx] - Ey[c]) / dx_);
      }
}

/* ------------------------------------------------------------------------ */
/* LightBC (src/grid/lightBC.cpp:10-142 constructor, :165-198 applyBCs),    */
/* GaussianPulses (src/utils/evaluateGaussianPulse.cpp:7-30), and           */
/* Grid::setFieldAlongEdge/InPlane (src/grid/oGrid.cpp:291-476).             */
/* ------------------------------------------------------------------------ */
double eo_gaussian_pulses(int nw, const double *amp, const double *omg, const double *pha,
                          const double *del, const double *dwt, double t, double dt) {
  double val = 0.0;
  for (int i = 0; i < nw; i++) {
    double tmp = amp[i] * cos(omg[i] * (t + dt) + pha[i] * EO_PI / 180.0);
    val += tmp * exp(-pow((t - del[i]) * dwt[i], 2));
  }
  return val;
}

/* plane index for setFieldAlongEdge: fid 0..5 = Ex..Bz, dim, upper side */
static int edge_plane_index(const eo_state *s, int fid, int dim, int upper) {
  if (!upper) return 1;
  int n_ = s->nc[dim]; /* nx_ = n + 2 */
  /* along-dim staggered (E_dim, B faces not normal to dim) use n_-1-1 = n */
  int comp = fid % 3;
  int is_e = fid < 3;
  if (is_e) return (comp == dim) ? n_ - 1 - 1 : n_ - 1;
  return (comp == dim) ? n_ - 1 : n_ - 1 - 1;
}

/* transverse profile of the injected wave part; waist<=0 -> 1 (reference) */
static double waist_profile(const eo_state *s, int fid, int dim, int i1, int i2) {
  if (s->c.waist <= 0) return 1.0;
  int a = (dim + 1) % 3, b = (dim + 2) % 3;
  if (a > b) { int t = a; a = b; b = t; }
  /* staggering offsets of Grid::getGridPhys (src/grid/grid.cpp:415-440) */
  int comp = fid % 3, is_e = fid < 3;
  double off[3] = {0, 0, 0};
  for (int d = 0; d < 3; d++) {
    if (is_e) off[d] = (d == comp) ? 0.5 * s->d[d] : 0.0;
    else off[d] = (d == comp) ? 0.0 : 0.5 * s->d[d];
  }
  double ya = s->c.x0[a] + s->d[a] * (i1 - 1) + off[a] - (s->c.x0[a] + 0.5 * s->c.L[a]);
  double yb = s->c.x0[b] + s->d[b] * (i2 - 1) + off[b] - (s->c.x0[b] + 0.5 * s->c.L[b]);
  double w = s->c.waist;
  return exp(-(ya * ya + yb * yb) / (w * w));
}

static void light_bc(eo_state *s, int side, double t, double dt) {
  int sign = (side & 1) ? 1 : -1;
  int dim = side / 2;
  int sidev = sign * (dim + 1);
  /* registry of waves injected on this side per polarisation */
  int nw[3] = {0, 0, 0};
  double amp[3][EO_NWAVE], omg[3][EO_NWAVE], pha[3][EO_NWAVE], del[3][EO_NWAVE], dwt[3][EO_NWAVE];
  for (int i = 0; i < s->c.nwaves; i++) {
    if (s->c.in_side[i] == sidev) {
      int pid = s->c.in_pol[i] - 1;
      int k = nw[pid]++;
      amp[pid][k] = s->c.peakamps[i];
      omg[pid][k] = s->c.omegas[i];
      pha[pid][k] = s->c.phases[i];
      del[pid][k] = s->c.delays[i];
      dwt[pid][k] = s->c.inv_widths[i];
    }
  }
  int Epol[3], Bpol[3];
  if (dim == 0) { Epol[0] = 1; Bpol[0] = 2; Epol[1] = 2; Bpol[1] = 1; Epol[2] = 0; Bpol[2] = 0; }
  else if (dim == 1) { Epol[0] = 2; Bpol[0] = 0; Epol[1] = 0; Bpol[1] = 2; Epol[2] = 1; Bpol[2] = 1; }
  else { Epol[0] = 0; Bpol[0] = 1; Epol[1] = 1; Bpol[1] = 0; Epol[2] = 2; Bpol[2] = 2; }
  double E1[3], B1[3], Ew[3] = {0, 0, 0}, Bw[3] = {0, 0, 0};
  for (int i = 0; i < 3; i++) { E1[i] = s->c.E0[i]; B1[i] = s->c.B0[i]; }
  double dx = s->d[dim];
  double bPhase = (dx + sign * dt) / 2;
  for (int i = 0; i < 3; i++) {
    int eid = Epol[i], bid = Bpol[i];
    if (nw[eid] > 0 && eid != dim) {
      double eFieldVal = eo_gaussian_pulses(nw[eid], amp[eid], omg[eid], pha[eid], del[eid], dwt[eid], t, 0.0);
      E1[eid] += eFieldVal;
      Ew[eid] = eFieldVal;
      int bsign = sign * (int)pow(-1, i + 1);
      double bFieldVal = bsign * eo_gaussian_pulses(nw[eid], amp[eid], omg[eid], pha[eid], del[eid], dwt[eid], t, bPhase);
      B1[bid] += bFieldVal;
      Bw[bid] = bFieldVal;
    }
    /* setFieldAlongEdge(eString_[i], ...), then bString_[i]: full planes */
    for (int pass = 0; pass < 2; pass++) {
      int fid = pass == 0 ? eid : 3 + bid;
      double val = pass == 0 ? E1[eid] : B1[bid];
      double base = pass == 0 ? s->c.E0[eid] : s->c.B0[bid];
      double wave = pass == 0 ? Ew[eid] : Bw[bid];
      int idx = edge_plane_index(s, fid, dim, sign > 0);
      int a = (dim + 1) % 3, b = (dim + 2) % 3;
      if (a > b) { int tt = a; a = b; b = tt; }
      double *f = s->f[fid];
      for (int p = 0; p < s->nt[a]; p++)
        for (int q = 0; q < s->nt[b]; q++) {
          int ii[3];
          ii[dim] = idx; ii[a] = p; ii[b] = q;
          double v = val;
          if (s->c.waist > 0) v = base + wave * waist_profile(s, fid, dim, p, q);
          f[IX(s, ii[0], ii[1], ii[2])] = v;
        }
    }
  }
}

/* E/B ghost refresh, Grid::executeBC(-1): per side in order x-,x+,...       */
/* periodic -> BC_F_Periodic::completeBC (bc_f_periodic.cpp:47-52):          */
/* get the physical plane of the opposite side, replace this side's ghost.  */
/* external -> LightBC::applyBCs(time_phys, dt_phys).                        */
static void refresh_fields(eo_state *s, const int *fids, int nf, double t, double dt, int light) {
  for (int side = 0; side < 6; side++) {
    int d = side / 2, right = side & 1;
    if (s->c.fbc[side] == EO_BC_PERIODIC) {
      int src = right ? 1 : s->n[d];
      int dst = right ? s->n[d] + 1 : 0;
      for (int f = 0; f < nf; f++) {
        /* Q23: B goes through the stale fieldPtr_ allocation */
        int fid = fids[f];
        double *arr = (fid >= FBX && fid <= FBZ) ? s->fBfix[fid - FBX] : s->f[fid];
        plane_copy(s, arr, d, src, dst, 0);
      }
    } else if (light) {
      light_bc(s, side, t, dt);
    }
  }
}

void eo_refresh(eo_state *s, double t, double dt) {
  static const int eb[6] = {FEX, FEY, FEZ, FBX, FBY, FBZ};
  refresh_fields(s, eb, 6, t, dt, 1);
}

/* ------------------------------------------------------------------------ */
/* Gather: Particle_Handler::InterpolateEB (particle_handler.cpp:138-171),   */
/* Grid::getFieldInterpolatorVec (oGrid.cpp:153-197),                        */
/* Interpolator::interpolate_fields (src/particles/interpolate.cpp:11-47).  */
/* ------------------------------------------------------------------------ */
static void interp_vec(const eo_state *s, int32_t cid, double *V) {
  int ix, iy, iz;
  cell_decode(s, cid, &ix, &iy, &iz);
  const double *Ex = s->f[FEX], *Ey = s->f[FEY], *Ez = s->f[FEZ];
  const double *Bx = s->f[FBX], *By = s->f[FBY], *Bz = s->f[FBZ];
  const double *Btx = s->f[FBTX], *Bty = s->f[FBTY], *Btz = s->f[FBTZ];
  V[0] = s->c.x0[0] + s->d[0] * ix;
  V[1] = s->c.x0[1] + s->d[1] * iy;
  V[2] = s->c.x0[2] + s->d[2] * iz;
  V[3] = Ex[IX(s, ix, iy, iz)];
  V[4] = Ex[IX(s, ix, iy + 1, iz)];
  V[5] = Ex[IX(s, ix, iy + 1, iz + 1)];
  V[6] = Ex[IX(s, ix, iy, iz + 1)];
  V[7] = Ey[IX(s, ix, iy, iz)];
  V[8] = Ey[IX(s, ix, iy, iz + 1)];
  V[9] = Ey[IX(s, ix + 1, iy, iz + 1)];
  V[10] = Ey[IX(s, ix + 1, iy, iz)];
  V[11] = Ez[IX(s, ix, iy, iz)];
  V[12] = Ez[IX(s, ix + 1, iy, iz)];
  V[13] = Ez[IX(s, ix + 1, iy + 1, iz)];
  V[14] = Ez[IX(s, ix, iy + 1, iz)];
  V[15] = (Btx[IX(s, ix, iy, iz)] + Bx[IX(s, ix, iy, iz)]) / 2;
  V[16] = (Btx[IX(s, ix + 1, iy, iz)] + Bx[IX(s, ix + 1, iy, iz)]) / 2;
  V[17] = (Bty[IX(s, ix, iy, iz)] + By[IX(s, ix, iy, iz)]) / 2;
  V[18] = (Bty[IX(s, ix, iy + 1, iz)] + By[IX(s, ix, iy + 1, iz)]) / 2;
  V[19] = (Btz[IX(s, ix, iy, iz)] + Bz[IX(s, ix, iy, iz)]) / 2;
  V[20] = (Btz[IX(s, ix, iy, iz + 1)] + Bz[IX(s, ix, iy, iz + 1)]) / 2;
}

static void interpolate_fields(const double *pos, const double *lcell, const double *cv, double *field) {
  double ds[3];
  for (int i = 0; i < 3; i++) ds[i] = pos[i] - cv[i];
  for (int i = 0; i < 3; i++) {
    int j = (i + 1) % 3, k = (i + 2) % 3;
    double eField = 0;
    eField += (lcell[j] - ds[j]) * (lcell[k] - ds[k]) * cv[3 + 4 * i];
    eField += ds[j] * (lcell[k] - ds[k]) * cv[3 + 4 * i + 1];
    eField += ds[j] * ds[k] * cv[3 + 4 * i + 2];
    eField += (lcell[j] - ds[j]) * ds[k] * cv[3 + 4 * i + 3];
    eField /= lcell[j] * lcell[k];
    double bField = 0;
    bField += (lcell[i] - ds[i]) * cv[3 + 12 + 2 * i];
    bField += ds[i] * cv[3 + 12 + 2 * i + 1];
    bField /= lcell[i];
    field[i] = eField;
    field[3 + i] = bField;
  }
}

void eo_gather(eo_state *s) {
  double cv[21];
  int32_t icell = -1;
  for (int64_t i = 0; i < s->np; i++) {
    const double *pos = s->x + 3 * i;
    int32_t c = eo_cell_id(s, pos[0], pos[1], pos[2]);
    if (c != icell) {
      icell = c;
      interp_vec(s, c, cv);
    }
    interpolate_fields(pos, s->d, cv, s->fld + 6 * i);
  }
}

/* ------------------------------------------------------------------------ */
/* Poisson initialisation: Poisson_Solver::InitializeFields                  */
/* (src/poisson/poisson.cpp:143-170), run_poisson_solver_ (:172-287),         */
/* phiToE (src/poisson/convertFields.cpp:10-68), AToB (:73-154), PoissonBC   */
/* (src/poisson/poissonBC.cpp:10-80).                                        */
/* ------------------------------------------------------------------------ */
static void set_plane_full(eo_state *s, double *f, int dim, int idx, double val) {
  int a = (dim + 1) % 3, b = (dim + 2) % 3;
  if (a > b) { int t = a; a = b; b = t; }
  for (int p = 0; p < s->nt[a]; p++)
    for (int q = 0; q < s->nt[b]; q++) {
      int ii[3];
      ii[dim] = idx; ii[a] = p; ii[b] = q;
      f[IX(s, ii[0], ii[1], ii[2])] = val;
    }
}

/* executeBC(fieldID) for one potential array: periodic copy, or the        */
/* PoissonBC constant plane (left: first physical plane; right: ghost n+1). */
static void potential_bc(eo_state *s, double *u, int which /*0 phi,1..3 A*/) {
  for (int side = 0; side < 6; side++) {
    int d = side / 2, right = side & 1;
    if (s->c.fbc[side] == EO_BC_PERIODIC) {
      int src = right ? 1 : s->n[d];
      int dst = right ? s->n[d] + 1 : 0;
      plane_copy(s, u, d, src, dst, 0);
    } else {
      /* convert side to bound_* index: (-1,+1,-2,+2,-3,+3) -> (0,1,2,3,4,5) */
      double val;
      if (which == 0) val = s->c.bound_phi[side];
      else if (which == 1) val = s->c.bound_Ax[side];
      else if (which == 2) val = s->c.bound_Ay[side];
      else val = s->c.bound_Az[side];
      int idx = right ? 1 + s->n[d] : 1;
      set_plane_full(s, u, d, idx, val);
    }
  }
}

static int jacobi(eo_state *s, double *u1, double *u2, const double *R, double tol,
                  double sourceMult, int which, int iter_max) {
  double dx_ = s->d[0], dy_ = s->d[1], dz_ = s->d[2];
  double celldist2 = pow(dx_, 2.0) * pow(dy_, 2.0) + pow(dy_, 2.0) * pow(dz_, 2.0) +
                     pow(dx_, 2.0) * pow(dz_, 2.0);
  double ax = pow(dy_, 2.0) * pow(dz_, 2.0) / (2.0 * celldist2);
  double ay = pow(dx_, 2.0) * pow(dz_, 2.0) / (2.0 * celldist2);
  double az = pow(dx_, 2.0) * pow(dy_, 2.0) / (2.0 * celldist2);
  double af = pow(dx_, 2.0) * pow(dy_, 2.0) * pow(dz_, 2.0) / (2.0 * celldist2);
  int64_t sy = s->nt[2], sx = (int64_t)s->nt[1] * s->nt[2];
  potential_bc(s, u1, which);
  int iternum = -1;
  int converged = 0;
  double maxDiff;
  do {
    iternum++;
    maxDiff = 0.0;
    double *src = (iternum % 2 == 0) ? u1 : u2;
    double *dst = (iternum % 2 == 0) ? u2 : u1;
    for (int i = 1; i <= s->n[0]; i++)
      for (int j = 1; j <= s->n[1]; j++)
        for (int k = 1; k <= s->n[2]; k++) {
          int64_t c = IX(s, i, j, k);
          dst[c] = ax * (src[c - sx] + src[c + sx]) + ay * (src[c - sy] + src[c + sy]) +
                   az * (src[c - 1] + src[c + 1]) - af * R[c] * sourceMult;
        }
    potential_bc(s, dst, which);
    for (int i = 1; i <= s->n[0]; i++)
      for (int j = 1; j <= s->n[1]; j++)
        for (int k = 1; k <= s->n[2]; k++) {
          int64_t c = IX(s, i, j, k);
          double absDiff = fabs(u2[c] - u1[c]);
          if (absDiff > maxDiff) maxDiff = absDiff;
        }
    if (maxDiff < tol) converged = 1;
  } while (!converged && iternum < iter_max);
  return iternum + 1;
}

void eo_poisson_init(eo_state *s, int32_t iter_max, double tol, int32_t *sweeps_out, double t,
                     double dt) {
  /* Q8: the reference never initialises these; the harness zeroes them */
  for (int f = 0; f < NFIELD; f++)
    if (f != FJX && f != FJY && f != FJZ && f != FRHO) memset(s->f[f], 0, s->ntot * sizeof(double));
  double sourceMult = -4 * EO_PI;
  double vmax, vmin;
  vmax = vmin = s->c.bound_phi[0];
  for (int i = 1; i < 6; i++) {
    if (s->c.bound_phi[i] > vmax) vmax = s->c.bound_phi[i];
    if (s->c.bound_phi[i] < vmin) vmin = s->c.bound_phi[i];
  }
  double conv_phi = vmax - vmin;
  vmax = vmin = s->c.bound_Ax[0];
  for (int i = 0; i < 6; i++) {
    const double a3[3] = {s->c.bound_Ax[i], s->c.bound_Ay[i], s->c.bound_Az[i]};
    for (int c = 0; c < 3; c++) {
      if (a3[c] > vmax) vmax = a3[c];
      if (a3[c] < vmin) vmin = a3[c];
    }
  }
  double conv_A = vmax - vmin;
  double tphi = tol >= 0 ? tol : fmax(conv_phi * 1e-9, 1e-12);
  double tA = tol >= 0 ? tol : fmax(conv_A * 1e-9, 1e-12);
  int32_t sw[4];
  sw[0] = jacobi(s, s->f[FPHI1], s->f[FPHI2], s->f[FRHO], tphi, sourceMult, 0, iter_max);
  /* phiToE on the physical points of each E component */
  double *phi = s->f[FPHI1];
  int64_t sy = s->nt[2], sx = (int64_t)s->nt[1] * s->nt[2];
  for (int c = 0; c < 3; c++) {
    double h = s->d[c];
    int64_t st = c == 0 ? sx : (c == 1 ? sy : 1);
    double *E = s->f[FEX + c];
    for (int i = 1; i <= s->n[0]; i++)
      for (int j = 1; j <= s->n[1]; j++)
        for (int k = 1; k <= s->n[2]; k++) {
          int64_t p = IX(s, i, j, k);
          E[p] = -(phi[p + st] - phi[p]) / h;
        }
  }
  sw[1] = jacobi(s, s->f[FAX1], s->f[FAX2], s->f[FJX], tA, sourceMult, 1, iter_max);
  sw[2] = jacobi(s, s->f[FAY1], s->f[FAY2], s->f[FJY], tA, sourceMult, 2, iter_max);
  sw[3] = jacobi(s, s->f[FAZ1], s->f[FAZ2], s->f[FJZ], tA, sourceMult, 3, iter_max);
  /* AToB, Q10: (nextValk - Ak)/hj + (nextValj - Aj)/hk */
  for (int c = 0; c < 3; c++) {
    double *Ak, *Aj, hk, hj;
    int64_t sk, sj;
    if (c == 0) { Ak = s->f[FAZ1]; hk = s->d[2]; Aj = s->f[FAY1]; hj = s->d[1]; sk = sy; sj = 1; }
    else if (c == 1) { Ak = s->f[FAX1]; hk = s->d[0]; Aj = s->f[FAZ1]; hj = s->d[2]; sk = 1; sj = sx; }
    else { Ak = s->f[FAY1]; hk = s->d[1]; Aj = s->f[FAX1]; hj = s->d[0]; sk = sx; sj = sy; }
    double *B = s->f[FBX + c];
    for (int i = 1; i <= s->n[0]; i++)
      for (int j = 1; j <= s->n[1]; j++)
        for (int k = 1; k <= s->n[2]; k++) {
          int64_t p = IX(s, i, j, k);
          double nextValk = Ak[p + sk];
          double nextValj = Ak[p + sj];
          B[p] = (nextValk - Ak[p]) / hj + (nextValj - Aj[p]) / hk;
        }
  }
  eo_refresh(s, t, dt);
  if (sweeps_out) memcpy(sweeps_out, sw, sizeof(sw));
}

void eo_const_init(eo_state *s) {
  for (int c = 0; c < 3; c++) {
    for (int64_t p = 0; p < s->ntot; p++) {
      s->f[FEX + c][p] = s->c.E0[c];
      s->f[FBX + c][p] = s->c.B0[c];
      s->f[FBTX + c][p] = s->c.B0[c];
      s->f[FJX + c][p] = 0.0;
    }
  }
  memset(s->f[FRHO], 0, s->ntot * sizeof(double));
}

/* ------------------------------------------------------------------------ */
/* Diagnostics: OutputBoxQuantities::output sums (src/IO/output.cpp:76-118), */
/* with SQR(x) expanding to (x*x) textually (output.cpp:9).                   */
/* ------------------------------------------------------------------------ */
void eo_diag(const eo_state *s, double *out) {
  double pEn = 0.0, pM_x = 0.0, pM_y = 0.0, pM_z = 0.0, qtot = 0.0;
  for (int64_t i = 0; i < s->np; i++) {
    double m = s->c.m[s->sp[i]];
    const double *v = s->v + 3 * i;
    pEn += 0.5 * m * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    pM_x += m * v[0];
    pM_y += m * v[1];
    pM_z += m * v[2];
    qtot += s->c.q[s->sp[i]];
  }
  double EEx = 0, EEy = 0, EEz = 0, EBx = 0, EBy = 0, EBz = 0;
  double vol = s->d[0] * s->d[1] * s->d[2];
  int64_t sy = s->nt[2], sx = (int64_t)s->nt[1] * s->nt[2];
  const double *Ex = s->f[FEX], *Ey = s->f[FEY], *Ez = s->f[FEZ];
  /* Q23: fieldPtr_[B_X..B_Z] (output.cpp:99) */
  const double *Bx = s->fBfix[0], *By = s->fBfix[1], *Bz = s->fBfix[2];
#define SQR(x) (x * x)
  for (int i = 1; i < 1 + s->n[0]; i++)
    for (int j = 1; j < 1 + s->n[1]; j++)
      for (int k = 1; k < 1 + s->n[2]; k++) {
        int64_t p = IX(s, i, j, k);
        EEx += 0.5 * vol * SQR(0.5 * (Ex[p] + Ex[p + sx]));
        EEy += 0.5 * vol * SQR(0.5 * (Ey[p] + Ey[p + sy]));
        EEz += 0.5 * vol * SQR(0.5 * (Ez[p] + Ez[p + 1]));
        EBx += 0.5 * vol * SQR(0.5 * (Bx[p] + Bx[p + sx]));
        EBy += 0.5 * vol * SQR(0.5 * (By[p] + By[p + sy]));
        EBz += 0.5 * vol * SQR(0.5 * (Bz[p] + Bz[p + 1]));
      }
#undef SQR
  double B_en = EBx + EBy + EBz;
  double E_en = EEx + EEy + EEz;
  out[0] = pM_x;
  out[1] = pM_y;
  out[2] = pM_z;
  out[3] = pEn;
  out[4] = B_en;
  out[5] = E_en;
  out[6] = B_en + E_en + pEn;
  out[7] = qtot;
}

/* one step, src/main.cpp:231-272 (sort placed after the particle BCs) */
int eo_step(eo_state *s, double t, double dt) {
  eo_push(s, dt);
  eo_particle_bc(s);
  if (eo_sort(s) != 0) return -4;
  eo_deposit(s, 1);
  eo_fold(s);
  eo_evolve(s, dt);
  eo_refresh(s, t, dt);
  eo_gather(s);
  return 0;
}
