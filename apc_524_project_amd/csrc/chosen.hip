// chosen.hip -- the FP32 chosen form of the PIC step for gfx950.
//
// Particles are a cell-sorted structure of arrays: one int32 word with the
// species in the top 5 bits and the reference's ghost-inclusive cell id
// ((ny+2)(nz+2) ix + (nz+2) iy + iz, src/grid/oGrid.cpp:208-222) in the low
// 26 bits, the in-cell offset f in [0,1)^3 (FP32) and u (FP32: v for Boris,
// gamma*v for the relativistic pusher).  28 bytes per particle; optional
// int64 ids ride along for comparisons against the FP64 oracle.
//
// The algorithm is the reference's (same step order, same defects); only the
// representation and precision change.  With x = x0 + dx (ix - 1 + f) the
// reference's "least vertex" (src/grid/oGrid.cpp:160-162, 239-241) gives
// relpos / l = f - 1, so its linear weights (l - r)/l, r/l become 2 - f and
// f - 1 (src/particles/interpolate.cpp:25-33, src/particles/deposit.cpp:94-97).
//
// Per step:
//   pgd_f32        gather (E, (B + B_tm1)/2) + push + particle BCs, one thread
//                  per particle, new state written beside the old one
//   sort           stable radix sort of the new cell ids, payload permutation
//   deposit_f32    one wavefront per cell: each lane sums its particles' 12
//                  edge currents with compensated (two-float) accumulation,
//                  then a fixed butterfly combines the lanes' (hi, lo) pairs:
//                  deterministic, no atomics
//   jsum_f32       per grid point, the fixed-order sum of the cell vectors
//   fold, Yee E/B, ghost refresh   shared FP32 instances of the grid kernels
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "deposit_grid.hpp"
#include "emo_internal.hpp"

namespace emo {

constexpr int kSpeciesShift = 26;
constexpr int32_t kCellMask = (1 << kSpeciesShift) - 1;

struct ChosenState {
  int64_t cap = 0;
  int32_t *cs = nullptr, *cs_alt = nullptr;
  float *f[3] = {}, *f_alt[3] = {};
  float *u[3] = {}, *u_alt[3] = {};
  int64_t *id = nullptr, *id_alt = nullptr;
  float *g0 = nullptr;  // stored gamma for the first relativistic push (Q3)
  bool g0_active = false;
  uint32_t *key = nullptr, *key_alt = nullptr, *perm = nullptr, *perm_alt = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int sort_bits = 0;
  int32_t *cbeg = nullptr, *cend = nullptr;
  float *cj = nullptr;  // 12 per cell, component-major
  float *cr = nullptr;  // 8 per cell
  FieldSet<float> fs;
  float *Btfix[3] = {};
};

namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

struct SpeciesF32 {
  float qp[EMO_NSPEC];  // ((0.5 dt q)/m) UNIT_ACC
  float jc[EMO_NSPEC];  // UNIT_RHOJ q / V: current per unit velocity per particle
  int test[EMO_NSPEC];
};

SpeciesF32 species_f32(const emo_ctx *c, double dt) {
  SpeciesF32 s;
  double V = c->geo.d[0] * c->geo.d[1] * c->geo.d[2];
  for (int k = 0; k < EMO_NSPEC; k++) {
    double q = c->species.q[k], m = c->species.m[k];
    double qp = (0.5 * dt * q) / m;
    qp *= kUnitAcc;
    s.qp[k] = (float)qp;
    s.jc[k] = (float)(kUnitRhoJ * q / V);
    s.test[k] = c->species.test[k];
  }
  return s;
}

__device__ inline void decode_cell(const Geo &g, int cell, int *ix, int *iy, int *iz) {
  *iz = cell % g.nc[2];
  int t = cell / g.nc[2];
  *iy = t % g.nc[1];
  *ix = t / g.nc[1];
}

// The particle BCs of the six sides in the reference's order
// (src/particles/particle_handler.cpp:287-295) on one dimension in cell
// units: p = ix-1 is the physical cell, P = p + f the position / dx.
// Periodic (bc_p_periodic.cpp:44-87): P < 0 -> P + n; P > n -> P - n.
// Reflecting (bc_p_reflecting.cpp:43-53): mirror at 0 or n and flip u.
__device__ inline void bc_dim(int &ix, float &f, float &u, int n, int left, int right) {
  int p = ix - 1;
  if (p < 0) {
    if (left == EMO_BC_PERIODIC) {
      p += n;
    } else {
      if (f == 0.f) {
        p = -p;
      } else {
        p = -p - 1;
        f = 1.f - f;
      }
      u = -u;
    }
  }
  if (p > n || (p == n && f > 0.f)) {
    if (right == EMO_BC_PERIODIC) {
      p -= n;
    } else {
      if (f == 0.f) {
        p = 2 * n - p;
      } else {
        p = 2 * n - p - 1;
        f = 1.f - f;
      }
      u = -u;
    }
  }
  ix = p + 1;
}

// ---------------------------------------------------------------------------
// Fused gather + push + particle BCs
// ---------------------------------------------------------------------------
struct PgdArgs {
  Geo g;
  const int32_t *cs;
  const float *f[3], *u[3];
  const float *g0;
  int32_t *cs_out;
  float *f_out[3], *u_out[3];
  uint32_t *key, *perm;
  const float *E[3], *B[3], *Bt[3];
  SpeciesF32 s;
  float dtid[3];  // dt / dx per dim
  int pbc[6];
  unsigned long long *flag;
  long long np;
};

template <bool REL>
__global__ void __launch_bounds__(kBlock) k_pgd_f32(PgdArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const Geo &g = a.g;
  int w = a.cs[i];
  int sp = (int)((uint32_t)w >> kSpeciesShift);
  int ix, iy, iz;
  decode_cell(g, w & kCellMask, &ix, &iy, &iz);
  float fx = a.f[0][i], fy = a.f[1][i], fz = a.f[2][i];
  float ux = a.u[0][i], uy = a.u[1][i], uz = a.u[2][i];
  // gather (Grid::getFieldInterpolatorVec + Interpolator::interpolate_fields)
  const float ax = 2.f - fx, bx = fx - 1.f;
  const float ay = 2.f - fy, by = fy - 1.f;
  const float az = 2.f - fz, bz = fz - 1.f;
  const long long c000 = gidx(g, ix, iy, iz);
  const long long sx = g.sx, sy = g.sy;
  const float *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
  float ex = ay * az * Ex[c000] + by * az * Ex[c000 + sy] + by * bz * Ex[c000 + sy + 1] +
             ay * bz * Ex[c000 + 1];
  float ey = az * ax * Ey[c000] + bz * ax * Ey[c000 + 1] + bz * bx * Ey[c000 + sx + 1] +
             az * bx * Ey[c000 + sx];
  float ez = ax * ay * Ez[c000] + bx * ay * Ez[c000 + sx] + bx * by * Ez[c000 + sx + sy] +
             ax * by * Ez[c000 + sy];
  float bxf = ax * 0.5f * (a.Bt[0][c000] + a.B[0][c000]) + bx * 0.5f * (a.Bt[0][c000 + sx] + a.B[0][c000 + sx]);
  float byf = ay * 0.5f * (a.Bt[1][c000] + a.B[1][c000]) + by * 0.5f * (a.Bt[1][c000 + sy] + a.B[1][c000 + sy]);
  float bzf = az * 0.5f * (a.Bt[2][c000] + a.B[2][c000]) + bz * 0.5f * (a.Bt[2][c000 + 1] + a.B[2][c000 + 1]);
  const float q_p = a.s.qp[sp];
  float vx, vy, vz;
  if (REL) {
    // Relativistic_Boris::Step (src/pusher/relativisticBoris.cpp:28-140) on u = gamma v
    float gam = a.g0 ? a.g0[i] : sqrtf(1.f + ux * ux + uy * uy + uz * uz);
    float ig = 1.f / gam;
    float vx0 = ux * ig, vy0 = uy * ig, vz0 = uz * ig;
    float upx = ux + q_p * (ex + (vy0 * bzf - vz0 * byf)) + q_p * ex;
    float upy = uy + q_p * (ey + (vz0 * bxf - vx0 * bzf)) + q_p * ey;
    float upz = uz + q_p * (ez + (vx0 * byf - vy0 * bxf)) + q_p * ez;
    float taux = q_p * bxf, tauy = q_p * byf, tauz = q_p * bzf;
    float tau_sq = taux * taux + tauy * tauy + tauz * tauz;
    float sigma = 1.f + upx * upx + upy * upy + upz * upz - tau_sq;
    float ustar = upx * taux + upy * tauy + upz * tauz;
    float gi = sqrtf(0.5f * (sigma + sqrtf(sigma * sigma + 4.f * (tau_sq + ustar * ustar))));
    float igi = 1.f / gi;
    float tx = taux * igi, ty = tauy * igi, tz = tauz * igi;
    float s = 1.f / (1.f + tx * tx + ty * ty + tz * tz);
    float udott = ustar * igi;
    ux = s * (upx + udott * tx + (upy * tz - upz * ty));
    uy = s * (upy + udott * ty + (upz * tx - upx * tz));
    uz = s * (upz + udott * tz + (upx * ty - upy * tx));
    vx = ux * igi;
    vy = uy * igi;
    vz = uz * igi;
  } else {
    // Boris::Step (src/pusher/boris.cpp:12-75)
    float tx = q_p * bxf, ty = q_p * byf, tz = q_p * bzf;
    float bsqr = bxf * bxf + byf * byf + bzf * bzf;
    float tsqr = 2.f / (1.f + q_p * q_p * bsqr);
    float sx_ = tsqr * tx, sy_ = tsqr * ty, sz_ = tsqr * tz;
    float v_mx = ux + q_p * ex, v_my = uy + q_p * ey, v_mz = uz + q_p * ez;
    float v_ix = v_mx + v_my * tz - v_mz * ty;
    float v_iy = v_my + v_mz * tx - v_mx * tz;
    float v_iz = v_mz + v_mx * ty - v_my * tx;
    ux = v_mx + v_iy * sz_ - v_iz * sy_ + q_p * ex;
    uy = v_my + v_iz * sx_ - v_ix * sz_ + q_p * ey;
    uz = v_mz + v_ix * sy_ - v_iy * sx_ + q_p * ez;
    vx = ux;
    vy = uy;
    vz = uz;
  }
  // x += dt v, in cell units, split into cell index and offset
  float pf[3] = {fx + a.dtid[0] * vx, fy + a.dtid[1] * vy, fz + a.dtid[2] * vz};
  int ic[3] = {ix, iy, iz};
  float uu[3] = {ux, uy, uz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    float fl = floorf(pf[d]);
    ic[d] += (int)fl;
    float fr = pf[d] - fl;
    if (fr >= 1.f) {  // rounding of a tiny negative offset
      fr = 0.f;
      ic[d] += 1;
    }
    pf[d] = fr;
    bc_dim(ic[d], pf[d], uu[d], g.n[d], a.pbc[2 * d], a.pbc[2 * d + 1]);
  }
  if (ic[0] < 0 || ic[0] >= g.nc[0] || ic[1] < 0 || ic[1] >= g.nc[1] || ic[2] < 0 || ic[2] >= g.nc[2]) {
    atomicOr(a.flag, kFlagOutOfGrid);
    ic[0] = min(max(ic[0], 1), g.n[0]);
    ic[1] = min(max(ic[1], 1), g.n[1]);
    ic[2] = min(max(ic[2], 1), g.n[2]);
  }
  int cell = (ic[0] * g.nc[1] + ic[1]) * g.nc[2] + ic[2];
  a.cs_out[i] = (sp << kSpeciesShift) | cell;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.f_out[d][i] = pf[d];
    a.u_out[d][i] = uu[d];
  }
  a.key[i] = (uint32_t)cell;
  a.perm[i] = (uint32_t)i;
}

struct PermF32Args {
  const int32_t *cs;
  const float *f[3], *u[3];
  const int64_t *id;
  int32_t *cs_out;
  float *f_out[3], *u_out[3];
  int64_t *id_out;
  const uint32_t *perm;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_permute_f32(PermF32Args a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  uint32_t s = a.perm[i];
  a.cs_out[i] = a.cs[s];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.f_out[d][i] = a.f[d][s];
    a.u_out[d][i] = a.u[d][s];
  }
  if (a.id) a.id_out[i] = a.id[s];
}

__global__ void __launch_bounds__(kBlock) k_cell_bounds_f32(const uint32_t *key, int32_t *beg, int32_t *end,
                                                             long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint32_t k = key[i];
  if (i == 0 || key[i - 1] != k) beg[k] = (int32_t)i;
  if (i == np - 1 || key[i + 1] != k) end[k] = (int32_t)(i + 1);
}

// ---------------------------------------------------------------------------
// Deposit: one wavefront per cell.  Knuth's TwoSum keeps the rounding error
// of every addition; the (hi, lo) pairs are combined across lanes by a fixed
// xor butterfly, so the result does not depend on scheduling.
// ---------------------------------------------------------------------------
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

struct DepF32Args {
  Geo g;
  const int32_t *cs;
  const float *f[3], *u[3];
  const float *g0;
  const int32_t *beg, *end;
  SpeciesF32 s;
  float *cj, *cr;
  int rel;
  int rho;
};

template <bool RHO>
__global__ void __launch_bounds__(kBlock) k_deposit_f32(DepF32Args a) {
  const int lane = threadIdx.x & 63;
  const long long c = (long long)blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
  const Geo &g = a.g;
  if (c >= g.ncell) return;
  constexpr int NV = RHO ? 20 : 12;
  float hi[NV], lo[NV];
#pragma unroll
  for (int q = 0; q < NV; q++) hi[q] = lo[q] = 0.f;
  const int b = a.beg[c], e = a.end[c];
  for (int p = b + lane; p < e; p += 64) {
    int sp = (int)((uint32_t)a.cs[p] >> kSpeciesShift);
    if (a.s.test[sp]) continue;
    float f[3] = {a.f[0][p], a.f[1][p], a.f[2][p]};
    float ux = a.u[0][p], uy = a.u[1][p], uz = a.u[2][p];
    float ig = a.rel ? 1.f / sqrtf(1.f + ux * ux + uy * uy + uz * uz) : 1.f;
    float v[3] = {ux * ig, uy * ig, uz * ig};
    float A[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
    float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
    float jc = a.s.jc[sp];
#pragma unroll
    for (int q = 0; q < 3; q++) {
      int j = (q + 1) % 3, k = (q + 2) % 3;
      float cq = jc * v[q];
      float w4[4] = {A[j] * A[k], Bw[j] * A[k], Bw[j] * Bw[k], A[j] * Bw[k]};
#pragma unroll
      for (int r = 0; r < 4; r++) {
        float s, err;
        two_sum(hi[4 * q + r], cq * w4[r], s, err);
        hi[4 * q + r] = s;
        lo[4 * q + r] += err;
      }
    }
    if (RHO) {
      float w8[8] = {A[0] * A[1] * A[2], Bw[0] * A[1] * A[2], Bw[0] * Bw[1] * A[2], A[0] * Bw[1] * A[2],
                     A[0] * Bw[1] * Bw[2], Bw[0] * Bw[1] * Bw[2], Bw[0] * A[1] * Bw[2], A[0] * A[1] * Bw[2]};
#pragma unroll
      for (int r = 0; r < 8; r++) {
        float s, err;
        two_sum(hi[12 + r], jc * w8[r], s, err);
        hi[12 + r] = s;
        lo[12 + r] += err;
      }
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < NV; q++) {
      float oh = __shfl_xor(hi[q], off, 64);
      float ol = __shfl_xor(lo[q], off, 64);
      float s, err;
      two_sum(hi[q], oh, s, err);
      hi[q] = s;
      lo[q] = lo[q] + ol + err;
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < 12; q++) a.cj[q * g.ncell + c] = hi[q] + lo[q];
    if (RHO) {
#pragma unroll
      for (int q = 0; q < 8; q++) a.cr[q * g.ncell + c] = hi[12 + q] + lo[12 + q];
    }
  }
}

// FP64 reconstruction of positions and velocities (global-position output)
struct OutArgs {
  Geo g;
  const int32_t *cs;
  const float *f[3], *u[3];
  const float *g0;
  int rel;
  double *x[3], *v[3], *gam;
  int32_t *sp;
  int32_t *cell;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_output_f32(OutArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const Geo &g = a.g;
  int w = a.cs[i];
  int cell = w & kCellMask;
  int ic[3];
  decode_cell(g, cell, &ic[0], &ic[1], &ic[2]);
  double ux = a.u[0][i], uy = a.u[1][i], uz = a.u[2][i];
  double gam = 1.0;
  if (a.rel) gam = a.g0 ? (double)a.g0[i] : sqrt(1.0 + ux * ux + uy * uy + uz * uz);
  double uu[3] = {ux, uy, uz};
  for (int d = 0; d < 3; d++) {
    if (a.x[d]) a.x[d][i] = g.x0[d] + g.d[d] * ((double)(ic[d] - 1) + (double)a.f[d][i]);
    if (a.v[d]) a.v[d][i] = a.rel ? uu[d] / gam : uu[d];
  }
  if (a.gam) a.gam[i] = gam;
  if (a.sp) a.sp[i] = (int)((uint32_t)w >> kSpeciesShift);
  if (a.cell) a.cell[i] = cell;
}

struct DiagF32Args {
  const int32_t *cs;
  const float *u[3];
  const float *g0;
  int rel;
  SpeciesParams s;
  long long np;
  double *partials;
};

__global__ void __launch_bounds__(kBlock) k_diag_particles_f32(DiagF32Args a) {
  __shared__ double sh[5][kBlock];
  double acc[5] = {0, 0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.np;
       i += (long long)gridDim.x * blockDim.x) {
    int sp = (int)((uint32_t)a.cs[i] >> kSpeciesShift);
    double m = a.s.m[sp];
    double ux = a.u[0][i], uy = a.u[1][i], uz = a.u[2][i];
    double ig = 1.0;
    if (a.rel) ig = 1.0 / (a.g0 ? (double)a.g0[i] : sqrt(1.0 + ux * ux + uy * uy + uz * uz));
    double vx = ux * ig, vy = uy * ig, vz = uz * ig;
    acc[0] += m * vx;
    acc[1] += m * vy;
    acc[2] += m * vz;
    acc[3] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    acc[4] += a.s.q[sp];
  }
  for (int q = 0; q < 5; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int q = 0; q < 5; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < 5; q++) a.partials[5 * blockIdx.x + q] = sh[q][0];
}

__global__ void k_gather_f32(const float *in, float *out, const uint32_t *perm, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[perm[i]];
}

int set_err(emo_ctx *c, int code, const std::string &m) {
  c->err = m;
  return code;
}

template <typename T>
int dalloc(emo_ctx *c, T **p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    return set_err(c, EMO_ERR_ALLOC, std::string("hipMalloc: ") + hipGetErrorString(e));
  }
  return EMO_OK;
}

void dfree(void *p) {
  if (p) hipFree(p);
}

int check_flag(emo_ctx *c) {
  hipError_t e = hipMemcpyAsync(c->hflag, c->dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, hipGetErrorString(e));
  if (c->hflag[0] & kFlagOutOfGrid)
    return set_err(c, EMO_ERR_INVARIANT, "particle outside the padded grid");
  return EMO_OK;
}

int sort_f32(emo_ctx *c) {
  ChosenState *s = c->ch;
  if (c->np == 0) return EMO_OK;
  size_t bytes = s->sort_tmp_bytes;
  hipError_t e;
  {
    TimedLaunch t(c, "sort_f32_radix");
    e = hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->key, s->key_alt, s->perm, s->perm_alt,
                                           (int)c->np, 0, s->sort_bits, c->stream);
  }
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, std::string("radix sort: ") + hipGetErrorString(e));
  PermF32Args pa;
  pa.cs = s->cs_alt;
  pa.cs_out = s->cs;
  for (int d = 0; d < 3; d++) {
    pa.f[d] = s->f_alt[d];
    pa.u[d] = s->u_alt[d];
    pa.f_out[d] = s->f[d];
    pa.u_out[d] = s->u[d];
  }
  pa.id = s->id;
  pa.id_out = s->id_alt;
  pa.perm = s->perm_alt;
  pa.np = c->np;
  {
    TimedLaunch t(c, "sort_f32_scatter");
    hipLaunchKernelGGL(k_permute_f32, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, pa);
  }
  if (s->id) std::swap(s->id, s->id_alt);
  EMO_HIP_CHECK_RET(c, hipMemsetAsync(s->cbeg, 0, c->geo.ncell * sizeof(int32_t), c->stream));
  EMO_HIP_CHECK_RET(c, hipMemsetAsync(s->cend, 0, c->geo.ncell * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_cell_bounds_f32, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, s->key_alt,
                     s->cbeg, s->cend, (long long)c->np);
  return EMO_OK;
}

// initial stable sort of a freshly loaded store (cs/f/u in the main arrays)
__global__ void k_keys_from_cs(const int32_t *cs, uint32_t *key, uint32_t *perm, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  key[i] = (uint32_t)(cs[i] & kCellMask);
  perm[i] = (uint32_t)i;
}

int resort_loaded(emo_ctx *c) {
  ChosenState *s = c->ch;
  if (c->np == 0) return EMO_OK;
  hipLaunchKernelGGL(k_keys_from_cs, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, s->cs, s->key,
                     s->perm, (long long)c->np);
  // move the loaded state to the alt arrays so sort_f32 permutes alt -> main
  std::swap(s->cs, s->cs_alt);
  for (int d = 0; d < 3; d++) {
    std::swap(s->f[d], s->f_alt[d]);
    std::swap(s->u[d], s->u_alt[d]);
  }
  int rc = sort_f32(c);
  if (rc) return rc;
  if (s->g0_active) {
    // the stored gamma follows its particles through the same permutation
    float *tmp = nullptr;
    rc = dalloc(c, &tmp, c->np);
    if (rc) return rc;
    hipLaunchKernelGGL(k_gather_f32, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream,
                       (const float *)s->g0, tmp, (const uint32_t *)s->perm_alt, (long long)c->np);
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    dfree(s->g0);
    s->g0 = tmp;
  }
  return EMO_OK;
}

void launch_pgd(emo_ctx *c, double dt) {
  ChosenState *s = c->ch;
  PgdArgs a;
  a.g = c->geo;
  a.cs = s->cs;
  a.cs_out = s->cs_alt;
  for (int d = 0; d < 3; d++) {
    a.f[d] = s->f[d];
    a.u[d] = s->u[d];
    a.f_out[d] = s->f_alt[d];
    a.u_out[d] = s->u_alt[d];
    a.E[d] = s->fs.E[d];
    a.B[d] = s->fs.B[d];
    a.Bt[d] = s->fs.Bt[d];
    a.dtid[d] = (float)(dt * c->geo.id[d]);
  }
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.key = s->key;
  a.perm = s->perm;
  a.s = species_f32(c, dt);
  for (int k = 0; k < 6; k++) a.pbc[k] = c->cfg.parts_bound[k];
  a.flag = c->dflag;
  a.np = c->np;
  TimedLaunch t(c, "pgd_f32");
  if (c->cfg.relativity)
    hipLaunchKernelGGL(k_pgd_f32<true>, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_pgd_f32<false>, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
}

void launch_deposit_f32(emo_ctx *c, int rho) {
  ChosenState *s = c->ch;
  DepF32Args a;
  a.g = c->geo;
  a.cs = s->cs;
  for (int d = 0; d < 3; d++) {
    a.f[d] = s->f[d];
    a.u[d] = s->u[d];
  }
  a.g0 = nullptr;
  a.beg = s->cbeg;
  a.end = s->cend;
  a.s = species_f32(c, 0.0);
  a.cj = s->cj;
  a.cr = s->cr;
  a.rel = c->cfg.relativity;
  a.rho = rho;
  unsigned blocks = (unsigned)((c->geo.ncell + (kBlock / 64) - 1) / (kBlock / 64));
  {
    TimedLaunch t(c, "deposit_f32");
    if (rho)
      hipLaunchKernelGGL(k_deposit_f32<true>, dim3(blocks), dim3(kBlock), 0, c->stream, a);
    else
      hipLaunchKernelGGL(k_deposit_f32<false>, dim3(blocks), dim3(kBlock), 0, c->stream, a);
  }
  DepGridArgs<float> b;
  b.g = c->geo;
  b.cj = s->cj;
  b.cr = s->cr;
  for (int d = 0; d < 3; d++) b.J[d] = s->fs.J[d];
  b.rho = s->fs.rho;
  b.do_rho = rho;
  TimedLaunch t(c, "jsum_f32");
  hipLaunchKernelGGL(k_deposit_grid<float>, dim3(grid_for(c->geo.ntot)), dim3(kBlock), 0, c->stream, b);
}

}  // namespace

int fp32_create(emo_ctx *c) {
  ChosenState *s = new (std::nothrow) ChosenState();
  if (!s) return EMO_ERR_ALLOC;
  c->ch = s;
  const Geo &g = c->geo;
  if (g.ncell > kCellMask) return set_err(c, EMO_ERR_CONFIG, "grid too large for the FP32 cell word");
  int64_t cap = std::max<int64_t>(c->cfg.max_particles, 1);
  s->cap = cap;
  c->cap = cap;
  int rc = 0;
  if (!rc) rc = dalloc(c, &s->cs, cap);
  if (!rc) rc = dalloc(c, &s->cs_alt, cap);
  for (int d = 0; d < 3 && !rc; d++) {
    rc = dalloc(c, &s->f[d], cap);
    if (!rc) rc = dalloc(c, &s->f_alt[d], cap);
    if (!rc) rc = dalloc(c, &s->u[d], cap);
    if (!rc) rc = dalloc(c, &s->u_alt[d], cap);
  }
  if (!rc) rc = dalloc(c, &s->key, cap);
  if (!rc) rc = dalloc(c, &s->key_alt, cap);
  if (!rc) rc = dalloc(c, &s->perm, cap);
  if (!rc) rc = dalloc(c, &s->perm_alt, cap);
  if (!rc) rc = dalloc(c, &s->cbeg, g.ncell);
  if (!rc) rc = dalloc(c, &s->cend, g.ncell);
  if (!rc) rc = dalloc(c, &s->cj, 12 * g.ncell);
  if (!rc) rc = dalloc(c, &s->cr, 8 * g.ncell);
  size_t nt = (size_t)g.ntot;
  for (int q = 0; q < 3 && !rc; q++) {
    rc = dalloc(c, &s->fs.E[q], nt);
    if (!rc) rc = dalloc(c, &s->fs.Bfix[q], nt);
    if (!rc) rc = dalloc(c, &s->Btfix[q], nt);
    if (!rc) rc = dalloc(c, &s->fs.J[q], nt);
  }
  if (!rc) rc = dalloc(c, &s->fs.rho, nt);
  if (rc) return rc;
  for (int q = 0; q < 3; q++) {
    s->fs.B[q] = s->fs.Bfix[q];
    s->fs.Bt[q] = s->Btfix[q];
  }
  float *all[13] = {s->fs.E[0], s->fs.E[1], s->fs.E[2], s->fs.Bfix[0], s->fs.Bfix[1], s->fs.Bfix[2],
                    s->Btfix[0], s->Btfix[1], s->Btfix[2], s->fs.J[0], s->fs.J[1], s->fs.J[2], s->fs.rho};
  for (int q = 0; q < 13; q++) EMO_HIP_CHECK_RET(c, hipMemset(all[q], 0, nt * sizeof(float)));
  int bits = 1;
  while ((1ll << bits) < g.ncell) bits++;
  s->sort_bits = bits;
  s->sort_tmp_bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, s->sort_tmp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                     (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)cap, 0, bits);
  if (hipMalloc(&s->sort_tmp, std::max<size_t>(s->sort_tmp_bytes, 16)) != hipSuccess)
    return set_err(c, EMO_ERR_ALLOC, "sort scratch");
  return EMO_OK;
}

void fp32_destroy(emo_ctx *c) {
  ChosenState *s = c->ch;
  if (!s) return;
  dfree(s->cs);
  dfree(s->cs_alt);
  for (int d = 0; d < 3; d++) {
    dfree(s->f[d]);
    dfree(s->f_alt[d]);
    dfree(s->u[d]);
    dfree(s->u_alt[d]);
    dfree(s->fs.E[d]);
    dfree(s->fs.Bfix[d]);
    dfree(s->Btfix[d]);
    dfree(s->fs.J[d]);
  }
  dfree(s->fs.rho);
  dfree(s->id);
  dfree(s->id_alt);
  dfree(s->g0);
  dfree(s->key);
  dfree(s->key_alt);
  dfree(s->perm);
  dfree(s->perm_alt);
  dfree(s->sort_tmp);
  dfree(s->cbeg);
  dfree(s->cend);
  dfree(s->cj);
  dfree(s->cr);
  delete s;
  c->ch = nullptr;
}

// Converts FP64 particles to the chosen representation on the host (the same
// cell-id arithmetic as Grid::getCellID) and sorts them by cell.
int fp32_load(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
              const int32_t *species, const int64_t *id) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  std::vector<int32_t> cs(n);
  std::vector<float> f[3], u[3], g0;
  for (int d = 0; d < 3; d++) {
    f[d].resize(n);
    u[d].resize(n);
  }
  const bool rel = c->cfg.relativity != 0;
  if (rel) g0.resize(n);
  for (int64_t i = 0; i < n; i++) {
    int ic[3];
    for (int d = 0; d < 3; d++) {
      double r = (x3[3 * i + d] - g.x0[d]) * g.id[d];
      int ix = (int)r;
      ic[d] = ix + 1;
      if (ic[d] < 0 || ic[d] >= g.nc[d]) return set_err(c, EMO_ERR_INVARIANT, "particle outside the padded grid");
      // true in-cell fraction of the reference's cell (floor via the int cast)
      double fr = (x3[3 * i + d] - g.x0[d]) * g.id[d] - (double)ix;
      float ff = (float)fr;
      if (ff >= 1.f) ff = std::nextafter(1.f, 0.f);
      if (ff < 0.f) ff = 0.f;
      f[d][i] = ff;
    }
    cs[i] = (species[i] << kSpeciesShift) | ((ic[0] * g.nc[1] + ic[1]) * g.nc[2] + ic[2]);
    double gm = (rel && gamma) ? gamma[i] : 1.0;
    for (int d = 0; d < 3; d++) u[d][i] = (float)(rel ? gm * v3[3 * i + d] : v3[3 * i + d]);
    if (rel) g0[i] = (float)gm;
  }
  EMO_HIP_CHECK_RET(c, hipMemcpy(s->cs, cs.data(), n * 4, hipMemcpyHostToDevice));
  for (int d = 0; d < 3; d++) {
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->f[d], f[d].data(), n * 4, hipMemcpyHostToDevice));
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->u[d], u[d].data(), n * 4, hipMemcpyHostToDevice));
  }
  dfree(s->id);
  dfree(s->id_alt);
  s->id = s->id_alt = nullptr;
  if (id) {
    int rc = dalloc(c, &s->id, n);
    if (!rc) rc = dalloc(c, &s->id_alt, n);
    if (rc) return rc;
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->id, id, n * 8, hipMemcpyHostToDevice));
  }
  dfree(s->g0);
  s->g0 = nullptr;
  s->g0_active = false;
  if (rel) {
    int rc = dalloc(c, &s->g0, n);
    if (rc) return rc;
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->g0, g0.data(), n * 4, hipMemcpyHostToDevice));
    s->g0_active = true;
  }
  c->np = n;
  int rc = resort_loaded(c);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return EMO_OK;
}

int fp32_get_particles(emo_ctx *c, double *x3, double *v3, double *gamma, int32_t *species, int64_t *id) {
  ChosenState *s = c->ch;
  int64_t n = c->np;
  if (n == 0) return EMO_OK;
  double *dx[3] = {}, *dv[3] = {}, *dg = nullptr;
  int32_t *dsp = nullptr;
  int rc = 0;
  for (int d = 0; d < 3 && !rc; d++) {
    rc = dalloc(c, &dx[d], n);
    if (!rc) rc = dalloc(c, &dv[d], n);
  }
  if (!rc) rc = dalloc(c, &dg, n);
  if (!rc) rc = dalloc(c, &dsp, n);
  if (!rc) {
    OutArgs a;
    a.g = c->geo;
    a.cs = s->cs;
    for (int d = 0; d < 3; d++) {
      a.f[d] = s->f[d];
      a.u[d] = s->u[d];
      a.x[d] = dx[d];
      a.v[d] = dv[d];
    }
    a.g0 = s->g0_active ? s->g0 : nullptr;
    a.rel = c->cfg.relativity;
    a.gam = dg;
    a.sp = dsp;
    a.cell = nullptr;
    a.np = n;
    hipLaunchKernelGGL(k_output_f32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
    std::vector<double> buf(n);
    auto down = [&](double *dst, const double *src, int stride, int comp) -> int {
      EMO_HIP_CHECK_RET(c, hipMemcpyAsync(buf.data(), src, n * 8, hipMemcpyDeviceToHost, c->stream));
      EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
      for (int64_t i = 0; i < n; i++) dst[stride * i + comp] = buf[i];
      return EMO_OK;
    };
    for (int d = 0; d < 3 && !rc; d++) {
      if (x3) rc = down(x3, dx[d], 3, d);
      if (v3 && !rc) rc = down(v3, dv[d], 3, d);
    }
    if (gamma && !rc) rc = down(gamma, dg, 1, 0);
    if (species && !rc && hipMemcpy(species, dsp, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMO_ERR_HIP;
    if (id && !rc) {
      if (s->id) {
        if (hipMemcpy(id, s->id, n * 8, hipMemcpyDeviceToHost) != hipSuccess) rc = EMO_ERR_HIP;
      } else {
        for (int64_t i = 0; i < n; i++) id[i] = -1;
      }
    }
  }
  for (int d = 0; d < 3; d++) {
    dfree(dx[d]);
    dfree(dv[d]);
  }
  dfree(dg);
  dfree(dsp);
  return rc;
}

int fp32_get_cells(emo_ctx *c, int32_t *cells) {
  ChosenState *s = c->ch;
  int64_t n = c->np;
  if (n == 0) return EMO_OK;
  std::vector<int32_t> cs(n);
  EMO_HIP_CHECK_RET(c, hipMemcpy(cs.data(), s->cs, n * 4, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < n; i++) cells[i] = cs[i] & kCellMask;
  return EMO_OK;
}

static float *f32_field(ChosenState *s, int fid) {
  switch (fid) {
    case EMO_EX: case EMO_EY: case EMO_EZ: return s->fs.E[fid];
    case EMO_BX: case EMO_BY: case EMO_BZ: return s->fs.B[fid - EMO_BX];
    case EMO_JX: case EMO_JY: case EMO_JZ: return s->fs.J[fid - EMO_JX];
    case EMO_RHO: return s->fs.rho;
    case EMO_BX_TM1: case EMO_BY_TM1: case EMO_BZ_TM1: return s->fs.Bt[fid - EMO_BX_TM1];
    default: return nullptr;
  }
}

int fp32_get_field(emo_ctx *c, int fid, double *out) {
  float *p = f32_field(c->ch, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  std::vector<float> h(c->geo.ntot);
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(h.data(), p, c->geo.ntot * 4, hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  for (long long i = 0; i < c->geo.ntot; i++) out[i] = h[i];
  return EMO_OK;
}

int fp32_set_field(emo_ctx *c, int fid, const double *in) {
  float *p = f32_field(c->ch, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  std::vector<float> h(c->geo.ntot);
  for (long long i = 0; i < c->geo.ntot; i++) h[i] = (float)in[i];
  EMO_HIP_CHECK_RET(c, hipMemcpy(p, h.data(), c->geo.ntot * 4, hipMemcpyHostToDevice));
  return EMO_OK;
}

int fp32_init_constant(emo_ctx *c) {
  ChosenState *s = c->ch;
  long long nt = c->geo.ntot;
  for (int q = 0; q < 3; q++) {
    fk_const_fill<float>(c, s->fs.E[q], c->cfg.E0[q], nt);
    fk_const_fill<float>(c, s->fs.B[q], c->cfg.B0[q], nt);
    fk_const_fill<float>(c, s->fs.Bt[q], c->cfg.B0[q], nt);
    fk_const_fill<float>(c, s->fs.J[q], 0.0, nt);
  }
  fk_const_fill<float>(c, s->fs.rho, 0.0, nt);
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return check_flag(c);
}

int fp32_init_poisson(emo_ctx *c, int32_t, double, int32_t *, double *) {
  return set_err(c, EMO_ERR_UNSUPPORTED, "Poisson initialisation is implemented in the FP64 plain form");
}

int fp32_step(emo_ctx *c, int32_t nsteps, double t0, double dt) {
  ChosenState *s = c->ch;
  double t = t0;
  for (int k = 0; k < nsteps; k++) {
    launch_pgd(c, dt);
    int rc = sort_f32(c);
    if (rc) return rc;
    if (s->g0_active) {
      // the stored gamma was used by this first push only (Q3); sort_f32
      // permuted the state, so from now on gamma = sqrt(1 + u^2)
      s->g0_active = false;
      dfree(s->g0);
      s->g0 = nullptr;
    }
    launch_deposit_f32(c, 0);
    fk_fold<float>(c, s->fs);
    fk_evolve<float>(c, s->fs, dt);
    refresh_fields<float>(c, s->fs, t, dt);
    t += dt;
  }
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return EMO_OK;
}

int fp32_diag(emo_ctx *c, double *out8) {
  ChosenState *s = c->ch;
  int nb = 256;
  double *pp = c->red, *pf = c->red + 5 * nb, *res = c->red + 11 * nb;
  DiagF32Args a;
  a.cs = s->cs;
  for (int d = 0; d < 3; d++) a.u[d] = s->u[d];
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.rel = c->cfg.relativity;
  a.s = c->species;
  a.np = c->np;
  a.partials = pp;
  hipLaunchKernelGGL(k_diag_particles_f32, dim3(nb), dim3(kBlock), 0, c->stream, a);
  launch_reduce_partials(c, pp, nb, 5, res);
  fk_diag_fields<float>(c, s->fs, pf, nb);
  launch_reduce_partials(c, pf, nb, 6, res + 5);
  double h[11];
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(h, res, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  double E_en = h[5] + h[6] + h[7];
  double B_en = h[8] + h[9] + h[10];
  out8[0] = h[0];
  out8[1] = h[1];
  out8[2] = h[2];
  out8[3] = h[3];
  out8[4] = B_en;
  out8[5] = E_en;
  out8[6] = B_en + E_en + h[3];
  out8[7] = h[4];
  return EMO_OK;
}

}  // namespace emo
