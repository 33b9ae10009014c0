// chosen.hip -- the FP32 chosen form of the PIC step for gfx950.
//
// Particles are a cell-sorted structure of arrays: one int32 word with the
// species in the top 5 bits and the reference's ghost-inclusive cell id
// ((ny+2)(nz+2) ix + (nz+2) iy + iz, src/grid/oGrid.cpp:208-222) in the low
// 26 bits, the in-cell offset f in [0,1)^3 (FP32) and u (FP32: v for Boris,
// gamma*v for the relativistic pusher).  28 bytes per particle, plus an
// optional int32 id that rides along for comparisons against the FP64 oracle.
// Cell c holds the particles [off[c], off[c+1]) of the arrays.
//
// The algorithm is the reference's (same step order, same defects); only the
// representation and precision change.  With x = x0 + dx (ix - 1 + f) the
// reference's "least vertex" (src/grid/oGrid.cpp:160-162, 239-241) gives
// relpos / l = f - 1, so its linear weights (l - r)/l, r/l become 2 - f and
// f - 1 (src/particles/interpolate.cpp:25-33, src/particles/deposit.cpp:94-97).
//
// Per step:
//   pgd_f32      gather (E, (B + B_tm1)/2) + push + particle BCs, one thread
//                per particle; writes the new state beside the old one and
//                the particle's move code: which of the <= 6 x 6 x 6 one-cell
//                moves (periodic wraps included) took it to its new cell
//   far_f32      (only when some particle moved further than one cell) an
//                ordered list of those particles
//   rank_f32     per source cell and move code, the particles' ranks in array
//                order and the group sizes (8 source cells per wavefront)
//   base_f32     per destination cell, the offsets of its (source, code)
//                groups in ascending source id; its particle count
//   scan_f32     exclusive scan of the counts: the new cell ranges
//   scatter_f32  every particle to off[dst] + base + rank -- the stable
//                counting sort by cell of the previous order, deterministic
//                and without radix passes
//   deposit_f32  16 lanes per cell: each lane sums its particles' 12 edge
//                currents with compensated (two-float) accumulation, a fixed
//                butterfly combines the lanes: deterministic, no atomics
//   jsum_f32     per grid point, the fixed-order sum of the cell vectors
//   fold, Yee E/B, ghost refresh   shared FP32 instances of the grid kernels
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "cellsort.hpp"
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
  int32_t *id = nullptr, *id_alt = nullptr;
  float *g0 = nullptr;  // stored gamma for the first relativistic push (Q3)
  bool g0_active = false;
  CellSortBufs srt;  // the counting sort's buffers (cellsort.hpp)
  // initial sort of loaded particles (arbitrary order)
  uint32_t *key = nullptr, *key_alt = nullptr, *perm = nullptr, *perm_alt = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int sort_bits = 0;
  float *cj = nullptr;  // 12 per cell, component-major
  FieldSet<float> fs;
  float *Btfix[3] = {};
};

namespace {


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
// Fused gather + push + particle BCs + move code + arrival counts
// ---------------------------------------------------------------------------
struct PgdArgs {
  Geo g;
  const int32_t *cs;
  const float *f[3], *u[3];
  const float *g0;
  int32_t *cs_out;
  float *f_out[3], *u_out[3];
  uint8_t *code;
  int32_t *nfar;   // particles that moved beyond one cell
  const float *E[3], *B[3], *Bt[3];
  SpeciesF32 s;
  float dtid[3];  // dt / dx per dim
  int pbc[6];
  int per[3];     // periodic particle BCs per dim
  unsigned long long *flag;
  long long np;
};

template <bool REL>
__global__ void __launch_bounds__(kBlock) k_pgd_f32(PgdArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < a.np;
  const Geo &g = a.g;
  int cell = -1;
  int mcode = 0;
  if (live) {
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
  int src[3] = {ix, iy, iz};
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
  cell = (ic[0] * g.nc[1] + ic[1]) * g.nc[2] + ic[2];
  int k0 = move_code(ic[0] - src[0], g.n[0], a.per[0]);
  int k1 = move_code(ic[1] - src[1], g.n[1], a.per[1]);
  int k2 = move_code(ic[2] - src[2], g.n[2], a.per[2]);
  mcode = (k0 < 0 || k1 < 0 || k2 < 0) ? (int)kFarCode : (k0 * 6 + k1) * 6 + k2;
  a.cs_out[i] = (sp << kSpeciesShift) | cell;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.f_out[d][i] = pf[d];
    a.u_out[d][i] = uu[d];
  }
  a.code[i] = (uint8_t)mcode;
  }
  const int nf = __popcll(__ballot(live && mcode == kFarCode));
  if (lane == 0 && nf) atomicAdd(a.nfar, nf);
}

// ---------------------------------------------------------------------------
// Far movers (rare: |v| dt > dx in some dimension).  An ordered list, in
// particle order (= ascending source cell), built in two passes that return
// at once when there are none.
// ---------------------------------------------------------------------------
// ---------------------------------------------------------------------------
// Migration + deposit: one wavefront per destination cell.
// ---------------------------------------------------------------------------
__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// the 12 edge currents of one particle (Depositor::deposit_particle_J,
// src/particles/deposit.cpp:12-40, weights 2 - f and f - 1), accumulated
// with TwoSum into (hi, lo)
__device__ inline void deposit_j(float hi[12], float lo[12], const float f[3], const float v[3], float jc) {
  float A[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
  float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
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
}

// ---------------------------------------------------------------------------
// Stable counting sort by destination cell, in four streaming passes.
// Within a destination cell the order is: source cells in ascending id, and
// within one source cell the particles' previous order -- the stable sort of
// the previous order by the new cell.  A particle's position is
//   off_new[dst] + base[code][src] + rank
// with rank its index among the particles of its (source cell, move code)
// group and base the number of particles that reach dst from (source, code)
// groups of lower source id.
// ---------------------------------------------------------------------------
struct ScatterArgs {
  Geo g;
  CellSortBufs b;
  const int32_t *cs_in;
  const float *f_in[3], *u_in[3];
  const int32_t *id_in;
  int32_t *cs_out;
  float *f_out[3], *u_out[3];
  int32_t *id_out;
  unsigned long long *flag;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_scatter_f32(ScatterArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int w = a.cs_in[i];
  const long long pos = cellsort_pos(a.g, a.b, i, a.b.code[i], w & kCellMask, a.flag);
  if (pos < 0) return;
  a.cs_out[pos] = w;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.f_out[d][pos] = a.f_in[d][i];
    a.u_out[d][pos] = a.u_in[d][i];
  }
  if (a.id_in) a.id_out[pos] = a.id_in[i];
}

// source and destination cells of an FP32 particle (far-mover list)
struct CellsF32 {
  const int32_t *src, *dst;
  __device__ void operator()(long long i, int &s, int &d) const {
    s = src[i] & kCellMask;
    d = dst[i] & kCellMask;
  }
};

// ---------------------------------------------------------------------------
// Deposit: 16 lanes per cell, 4 cells per wavefront.  Each lane sums its
// particles' 12 edge currents with TwoSum; a fixed xor butterfly within the
// 16-lane group combines the (hi, lo) pairs: deterministic, no atomics.
// ---------------------------------------------------------------------------
struct DepF32Args {
  Geo g;
  const int32_t *cs;
  const float *f[3], *u[3];
  const int32_t *off;
  SpeciesF32 s;
  float *cj;
};

template <bool REL>
__global__ void __launch_bounds__(kBlock) k_deposit_f32(DepF32Args a) {
  const int sub = threadIdx.x & 15;
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  const Geo &g = a.g;
  float hi[12], lo[12];
#pragma unroll
  for (int q = 0; q < 12; q++) hi[q] = lo[q] = 0.f;
  if (c < g.ncell) {
    const int b = a.off[c], e = a.off[c + 1];
    for (int p = b + sub; p < e; p += 16) {
      const int w = a.cs[p];
      const int sp = (int)((uint32_t)w >> kSpeciesShift);
      if (a.s.test[sp]) continue;
      float f[3] = {a.f[0][p], a.f[1][p], a.f[2][p]};
      float uq[3] = {a.u[0][p], a.u[1][p], a.u[2][p]};
      float ig = REL ? 1.f / sqrtf(1.f + uq[0] * uq[0] + uq[1] * uq[1] + uq[2] * uq[2]) : 1.f;
      float v[3] = {uq[0] * ig, uq[1] * ig, uq[2] * ig};
      deposit_j(hi, lo, f, v, a.s.jc[sp]);
    }
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 12; q++) {
      float oh = __shfl_xor(hi[q], off, 64);
      float ol = __shfl_xor(lo[q], off, 64);
      float s, err;
      two_sum(hi[q], oh, s, err);
      hi[q] = s;
      lo[q] = lo[q] + ol + err;
    }
  }
  if (sub == 0 && c < g.ncell) {
#pragma unroll
    for (int q = 0; q < 12; q++) a.cj[q * g.ncell + c] = hi[q] + lo[q];
  }
}

// ---------------------------------------------------------------------------
// Initial sort of freshly loaded particles (arbitrary order): radix sort of
// the cell ids, payload permutation, arrival counts.
// ---------------------------------------------------------------------------
__global__ void k_keys_from_cs(const int32_t *cs, uint32_t *key, uint32_t *perm, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  key[i] = (uint32_t)(cs[i] & kCellMask);
  perm[i] = (uint32_t)i;
}

struct PermF32Args {
  const int32_t *cs;
  const float *f[3], *u[3];
  const int32_t *id;
  const float *g0;
  int32_t *cs_out;
  float *f_out[3], *u_out[3];
  int32_t *id_out;
  float *g0_out;
  const uint32_t *perm;
  int32_t *count;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_permute_f32(PermF32Args a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < a.np;
  int cell = -1;
  if (live) {
    uint32_t s = a.perm[i];
    int w = a.cs[s];
    a.cs_out[i] = w;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      a.f_out[d][i] = a.f[d][s];
      a.u_out[d][i] = a.u[d][s];
    }
    if (a.id) a.id_out[i] = a.id[s];
    if (a.g0) a.g0_out[i] = a.g0[s];
    cell = w & kCellMask;
  }
  // particles per cell: one atomic per run of equal cells (the output is sorted)
  const int prev = __shfl_up(cell, 1, 64);
  const bool head = live && (lane == 0 || prev != cell);
  const unsigned long long heads = __ballot(head);
  const int nlive = __popcll(__ballot(live));
  if (head) {
    unsigned long long above = heads & ~(lanes_below(lane) | (1ull << lane));
    int next = above ? __ffsll((long long)above) - 1 : 64;
    atomicAdd(&a.count[cell], min(next, nlive) - lane);
  }
}

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
  if (c->hflag[0] & kFlagFarMovers)
    return set_err(c, EMO_ERR_NUMERIC,
                   "too many particles moved more than one cell in a step (|v| dt > dx): reduce dt");
  if (c->hflag[0] & kFlagMigrate) return set_err(c, EMO_ERR_INVARIANT, "cell migration lost particles");
  return EMO_OK;
}

// Sorts freshly loaded particles (main arrays, any order) by cell, stably.
int sort_loaded(emo_ctx *c) {
  ChosenState *s = c->ch;
  const long long np = c->np;
  cs_fill_i32(c, s->srt.count, 0, c->geo.ncell + 1);
  if (np > 0) {
    hipLaunchKernelGGL(k_keys_from_cs, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, s->cs, s->key, s->perm,
                       np);
    size_t bytes = s->sort_tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->key, s->key_alt, s->perm,
                                                      s->perm_alt, (int)np, 0, s->sort_bits, c->stream);
    if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, std::string("radix sort: ") + hipGetErrorString(e));
    float *g0_new = nullptr;
    if (s->g0_active) {
      int rc = dalloc(c, &g0_new, np);
      if (rc) return rc;
    }
    PermF32Args pa;
    pa.cs = s->cs;
    pa.cs_out = s->cs_alt;
    for (int d = 0; d < 3; d++) {
      pa.f[d] = s->f[d];
      pa.u[d] = s->u[d];
      pa.f_out[d] = s->f_alt[d];
      pa.u_out[d] = s->u_alt[d];
    }
    pa.id = s->id;
    pa.id_out = s->id_alt;
    pa.g0 = s->g0_active ? s->g0 : nullptr;
    pa.g0_out = g0_new;
    pa.perm = s->perm_alt;
    pa.count = s->srt.count;
    pa.np = np;
    hipLaunchKernelGGL(k_permute_f32, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, pa);
    std::swap(s->cs, s->cs_alt);
    for (int d = 0; d < 3; d++) {
      std::swap(s->f[d], s->f_alt[d]);
      std::swap(s->u[d], s->u_alt[d]);
    }
    if (s->id) std::swap(s->id, s->id_alt);
    if (g0_new) {
      EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
      dfree(s->g0);
      s->g0 = g0_new;
    }
  }
  int rc = cellsort_scan(c, s->srt, s->srt.off, "scan_f32");
  if (rc) return rc;
  cs_fill_i32(c, s->srt.count, 0, c->geo.ncell + 1);
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
    a.per[d] = c->cfg.parts_bound[2 * d] == EMO_BC_PERIODIC || c->cfg.parts_bound[2 * d + 1] == EMO_BC_PERIODIC;
  }
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.code = s->srt.code;
  a.nfar = s->srt.misc;
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

// far-mover list, ranks, group offsets, new cell ranges, scatter, deposit,
// grid sum
int launch_sort_deposit(emo_ctx *c) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  const long long np = c->np;
  static const char *const names[4] = {"far_f32", "rank_f32", "base_f32", "scan_f32"};
  CellsF32 cells{s->cs, s->cs_alt};
  int rc = cellsort_offsets(c, s->srt, np, cells, names);
  if (rc) return rc;
  {
    ScatterArgs a;
    a.g = g;
    a.b = s->srt;
    a.cs_in = s->cs_alt;
    a.cs_out = s->cs;
    for (int d = 0; d < 3; d++) {
      a.f_in[d] = s->f_alt[d];
      a.u_in[d] = s->u_alt[d];
      a.f_out[d] = s->f[d];
      a.u_out[d] = s->u[d];
    }
    a.id_in = s->id;
    a.id_out = s->id_alt;
    a.flag = c->dflag;
    a.np = np;
    TimedLaunch t(c, "scatter_f32");
    hipLaunchKernelGGL(k_scatter_f32, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, a);
  }
  std::swap(s->srt.off, s->srt.off_new);
  if (s->id) std::swap(s->id, s->id_alt);
  // the far-mover count starts from zero again
  cs_fill_i32(c, s->srt.misc, 0, 1);
  {
    DepF32Args d;
    d.g = g;
    d.cs = s->cs;
    for (int q = 0; q < 3; q++) {
      d.f[q] = s->f[q];
      d.u[q] = s->u[q];
    }
    d.off = s->srt.off;
    d.s = species_f32(c, 0.0);
    d.cj = s->cj;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    TimedLaunch t(c, "deposit_f32");
    if (c->cfg.relativity)
      hipLaunchKernelGGL(k_deposit_f32<true>, dim3(blocks), dim3(kBlock), 0, c->stream, d);
    else
      hipLaunchKernelGGL(k_deposit_f32<false>, dim3(blocks), dim3(kBlock), 0, c->stream, d);
  }
  DepGridArgs<float> b;
  b.g = g;
  b.cj = s->cj;
  b.cr = nullptr;
  for (int d = 0; d < 3; d++) b.J[d] = s->fs.J[d];
  b.rho = s->fs.rho;
  b.do_rho = 0;
  TimedLaunch t(c, "jsum_f32");
  hipLaunchKernelGGL(k_deposit_grid<float>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, c->stream, b);
  return EMO_OK;
}

}  // namespace

int fp32_create(emo_ctx *c) {
  ChosenState *s = new (std::nothrow) ChosenState();
  if (!s) return EMO_ERR_ALLOC;
  c->ch = s;
  const Geo &g = c->geo;
  if (g.ncell > kCellMask) return set_err(c, EMO_ERR_CONFIG, "grid too large for the FP32 cell word");
  int64_t cap = std::max<int64_t>(c->cfg.max_particles, 1);
  if (cap > INT_MAX) return set_err(c, EMO_ERR_CONFIG, "the FP32 chosen form holds at most 2^31-1 particles");
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
  if (!rc) rc = cellsort_alloc(c, s->srt, cap, g.ncell);
  if (!rc) rc = dalloc(c, &s->key, cap);
  if (!rc) rc = dalloc(c, &s->key_alt, cap);
  if (!rc) rc = dalloc(c, &s->perm, cap);
  if (!rc) rc = dalloc(c, &s->perm_alt, cap);
  if (!rc) rc = dalloc(c, &s->cj, 12 * g.ncell);
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
  cellsort_free(s->srt);
  dfree(s->key);
  dfree(s->key_alt);
  dfree(s->perm);
  dfree(s->perm_alt);
  dfree(s->sort_tmp);
  dfree(s->cj);
  delete s;
  c->ch = nullptr;
}

// Converts FP64 particles to the chosen representation on the host (the same
// cell-id arithmetic as Grid::getCellID) and sorts them by cell.
int fp32_load(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
              const int32_t *species, const int64_t *id) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  std::vector<int32_t> cs(n), ids;
  std::vector<float> f[3], u[3], g0;
  for (int d = 0; d < 3; d++) {
    f[d].resize(n);
    u[d].resize(n);
  }
  const bool rel = c->cfg.relativity != 0;
  if (rel) g0.resize(n);
  if (id) ids.resize(n);
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
    if (id) {
      if (id[i] < INT_MIN || id[i] > INT_MAX)
        return set_err(c, EMO_ERR_UNSUPPORTED, "particle ids must fit in int32 in the FP32 chosen form");
      ids[i] = (int32_t)id[i];
    }
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
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->id, ids.data(), n * 4, hipMemcpyHostToDevice));
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
  int rc = sort_loaded(c);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return check_flag(c);
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
        std::vector<int32_t> h(n);
        if (hipMemcpy(h.data(), s->id, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMO_ERR_HIP;
        for (int64_t i = 0; i < n && !rc; i++) id[i] = h[i];
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
    if (s->g0_active) {
      // the stored gamma was used by this first push only (Q3): from now on
      // gamma = sqrt(1 + u^2)
      EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
      s->g0_active = false;
      dfree(s->g0);
      s->g0 = nullptr;
    }
    int rc = launch_sort_deposit(c);
    if (rc) return rc;
    fk_fold<float>(c, s->fs);
    fk_evolve<float>(c, s->fs, dt);
    refresh_fields<float>(c, s->fs, t, dt);
    t += dt;
  }
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return check_flag(c);
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
