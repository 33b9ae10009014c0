// plain_kernels.hip -- FP64 "plain" particle kernels for gfx950.
//
// The plain form keeps the reference's pass structure (gather, push,
// particle BCs, sort, deposit as separate passes over an SoA particle store)
// and its exact floating-point operation order, so that with
// -ffp-contract=off every result is bit-identical to the CPU oracle.  One
// thread per particle for gather/push/BC, one thread per cell for the
// deposit's per-cell accumulation, one thread per grid point for the
// fixed-order edge/vertex sums.  No atomics on floating-point data.
#include <hipcub/hipcub.hpp>

#include "deposit_grid.hpp"
#include "emo_internal.hpp"

namespace emo {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// Grid::getCellID (src/grid/oGrid.cpp:208-222); returns -1 outside the
// padded grid (where the reference asserts).
__device__ inline int cell_of(const Geo &g, double x, double y, double z, int *ix, int *iy,
                              int *iz) {
  int a = (int)((x - g.x0[0]) * g.id[0]);
  a += 1;
  int b = (int)((y - g.x0[1]) * g.id[1]);
  b += 1;
  int c = (int)((z - g.x0[2]) * g.id[2]);
  c += 1;
  *ix = a;
  *iy = b;
  *iz = c;
  if (a < 0 || a >= g.nc[0] || b < 0 || b >= g.nc[1] || c < 0 || c >= g.nc[2]) return -1;
  return (g.nc[1] * g.nc[2]) * a + g.nc[2] * b + c;
}

// ---------------------------------------------------------------------------
// Gather: Grid::getFieldInterpolatorVec (src/grid/oGrid.cpp:153-197) and
// Interpolator::interpolate_fields (src/particles/interpolate.cpp:11-47).
// ---------------------------------------------------------------------------
struct GatherArgs {
  Geo g;
  const double *x[3];
  double *fld[6];
  const double *E[3], *B[3], *Bt[3];
  unsigned long long *flag;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_gather(GatherArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const Geo &g = a.g;
  double pos[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
  int ix, iy, iz;
  if (cell_of(g, pos[0], pos[1], pos[2], &ix, &iy, &iz) < 0) {
    atomicOr(a.flag, kFlagOutOfGrid);
    return;
  }
  double cv[21];
  cv[0] = g.x0[0] + g.d[0] * ix;
  cv[1] = g.x0[1] + g.d[1] * iy;
  cv[2] = g.x0[2] + g.d[2] * iz;
  const double *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
  cv[3] = Ex[gidx(g, ix, iy, iz)];
  cv[4] = Ex[gidx(g, ix, iy + 1, iz)];
  cv[5] = Ex[gidx(g, ix, iy + 1, iz + 1)];
  cv[6] = Ex[gidx(g, ix, iy, iz + 1)];
  cv[7] = Ey[gidx(g, ix, iy, iz)];
  cv[8] = Ey[gidx(g, ix, iy, iz + 1)];
  cv[9] = Ey[gidx(g, ix + 1, iy, iz + 1)];
  cv[10] = Ey[gidx(g, ix + 1, iy, iz)];
  cv[11] = Ez[gidx(g, ix, iy, iz)];
  cv[12] = Ez[gidx(g, ix + 1, iy, iz)];
  cv[13] = Ez[gidx(g, ix + 1, iy + 1, iz)];
  cv[14] = Ez[gidx(g, ix, iy + 1, iz)];
  cv[15] = (a.Bt[0][gidx(g, ix, iy, iz)] + a.B[0][gidx(g, ix, iy, iz)]) / 2;
  cv[16] = (a.Bt[0][gidx(g, ix + 1, iy, iz)] + a.B[0][gidx(g, ix + 1, iy, iz)]) / 2;
  cv[17] = (a.Bt[1][gidx(g, ix, iy, iz)] + a.B[1][gidx(g, ix, iy, iz)]) / 2;
  cv[18] = (a.Bt[1][gidx(g, ix, iy + 1, iz)] + a.B[1][gidx(g, ix, iy + 1, iz)]) / 2;
  cv[19] = (a.Bt[2][gidx(g, ix, iy, iz)] + a.B[2][gidx(g, ix, iy, iz)]) / 2;
  cv[20] = (a.Bt[2][gidx(g, ix, iy, iz + 1)] + a.B[2][gidx(g, ix, iy, iz + 1)]) / 2;
  const double *lcell = g.d;
  double ds[3];
#pragma unroll
  for (int q = 0; q < 3; q++) ds[q] = pos[q] - cv[q];
#pragma unroll
  for (int q = 0; q < 3; q++) {
    int j = (q + 1) % 3, k = (q + 2) % 3;
    double eField = 0;
    eField += (lcell[j] - ds[j]) * (lcell[k] - ds[k]) * cv[3 + 4 * q];
    eField += ds[j] * (lcell[k] - ds[k]) * cv[3 + 4 * q + 1];
    eField += ds[j] * ds[k] * cv[3 + 4 * q + 2];
    eField += (lcell[j] - ds[j]) * ds[k] * cv[3 + 4 * q + 3];
    eField /= lcell[j] * lcell[k];
    double bField = 0;
    bField += (lcell[q] - ds[q]) * cv[3 + 12 + 2 * q];
    bField += ds[q] * cv[3 + 12 + 2 * q + 1];
    bField /= lcell[q];
    a.fld[q][i] = eField;
    a.fld[3 + q][i] = bField;
  }
}

// ---------------------------------------------------------------------------
// Push: Boris::Step (src/pusher/boris.cpp:12-75) and Relativistic_Boris::Step
// (src/pusher/relativisticBoris.cpp:28-140), operation for operation.
// ---------------------------------------------------------------------------
struct PushArgs {
  double *x[3], *v[3], *gam;
  const double *fld[6];
  const int32_t *sp;
  SpeciesParams s;
  double dt;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_push_boris(PushArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  int sp = a.sp[i];
  double dt = a.dt;
  double x = a.x[0][i], y = a.x[1][i], z = a.x[2][i];
  double vx = a.v[0][i], vy = a.v[1][i], vz = a.v[2][i];
  double q = a.s.q[sp], m = a.s.m[sp];
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  double ex = a.fld[0][i], ey = a.fld[1][i], ez = a.fld[2][i];
  double bx = a.fld[3][i], by = a.fld[4][i], bz = a.fld[5][i];
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
  a.x[0][i] = x + dt * vx_new;
  a.x[1][i] = y + dt * vy_new;
  a.x[2][i] = z + dt * vz_new;
  a.v[0][i] = vx_new;
  a.v[1][i] = vy_new;
  a.v[2][i] = vz_new;
}

__global__ void __launch_bounds__(kBlock) k_push_vay(PushArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  int sp = a.sp[i];
  double dt = a.dt;
  double x = a.x[0][i], y = a.x[1][i], z = a.x[2][i];
  double vx = a.v[0][i], vy = a.v[1][i], vz = a.v[2][i];
  double gamma = a.gam[i];
  double q = a.s.q[sp], m = a.s.m[sp];
  double q_p = (0.5 * dt * q) / m;
  q_p *= kUnitAcc;
  double ex = a.fld[0][i], ey = a.fld[1][i], ez = a.fld[2][i];
  double bx = a.fld[3][i], by = a.fld[4][i], bz = a.fld[5][i];
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
  a.x[0][i] = x + dt * vx_new;
  a.x[1][i] = y + dt * vy_new;
  a.x[2][i] = z + dt * vz_new;
  a.v[0][i] = vx_new;
  a.v[1][i] = vy_new;
  a.v[2][i] = vz_new;
  a.gam[i] = gamma_int;
}

// ---------------------------------------------------------------------------
// Particle BCs: the six sides in the order x-,x+,y-,y+,z-,z+
// (src/particles/particle_handler.cpp:287-295).  Periodic: strict compare and
// shift by (xMax-xMin) (src/boundaries/b_particles/bc_p_periodic.cpp:44-87);
// reflecting: mirror and flip v (bc_p_reflecting.cpp:43-53).  In place; the
// reference's ghost copies + swap-pop leave the same per-particle state.
// ---------------------------------------------------------------------------
struct BcArgs {
  double *x[3], *v[3];
  double xmin[3], xmax[3];
  int pbc[6];
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_particle_bc(BcArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
#pragma unroll
  for (int side = 0; side < 6; side++) {
    int d = side >> 1, right = side & 1;
    double xMin = a.xmin[d], xMax = a.xmax[d];
    double xp = a.x[d][i];
    if (a.pbc[side] == EMO_BC_PERIODIC) {
      if (!right && xp < xMin) a.x[d][i] = xp + (xMax - xMin);
      if (right && xp > xMax) a.x[d][i] = xp - (xMax - xMin);
    } else {
      if (right && xp > xMax) {
        a.x[d][i] = 2.0 * xMax - xp;
        a.v[d][i] = -a.v[d][i];
      }
      if (!right && xp < xMin) {
        a.x[d][i] = 2.0 * xMin - xp;
        a.v[d][i] = -a.v[d][i];
      }
    }
  }
}

// ---------------------------------------------------------------------------
// Sort support: cell keys, payload permutation, per-cell ranges.
// ---------------------------------------------------------------------------
struct KeyArgs {
  Geo g;
  const double *x[3];
  uint32_t *key;
  uint32_t *perm;
  unsigned long long *flag;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_cell_keys(KeyArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  int ix, iy, iz;
  int c = cell_of(a.g, a.x[0][i], a.x[1][i], a.x[2][i], &ix, &iy, &iz);
  if (c < 0) {
    atomicOr(a.flag, kFlagOutOfGrid);
    c = 0;
  }
  a.key[i] = (uint32_t)c;
  if (a.perm) a.perm[i] = (uint32_t)i;
}

struct PermuteArgs {
  ParticlesF64 src, dst;
  const uint32_t *perm;
  long long np;
  int fields;  // 0: the gathered fields are stale (regathered later), not moved
};

__global__ void __launch_bounds__(kBlock) k_permute(PermuteArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  uint32_t s = a.perm[i];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.dst.x[d][i] = a.src.x[d][s];
    a.dst.v[d][i] = a.src.v[d][s];
  }
  if (a.fields) {
#pragma unroll
    for (int d = 0; d < 6; d++) a.dst.fld[d][i] = a.src.fld[d][s];
  }
  a.dst.gam[i] = a.src.gam[s];
  a.dst.sp[i] = a.src.sp[s];
  a.dst.id[i] = a.src.id[s];
}

__global__ void __launch_bounds__(kBlock) k_cell_bounds(const uint32_t *key, int32_t *beg, int32_t *end,
                                                         long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  uint32_t k = key[i];
  if (i == 0 || key[i - 1] != k) beg[k] = (int32_t)i;
  if (i == np - 1 || key[i + 1] != k) end[k] = (int32_t)(i + 1);
}

// ---------------------------------------------------------------------------
// Deposit, stage 1: per cell, the sums of Depositor::deposit_particle_J and
// _Rho (src/particles/deposit.cpp:12-62) over the cell's particles in array
// order, with the cell's vertex from Grid::getCellVertex
// (src/grid/oGrid.cpp:233-244); test particles are skipped
// (src/particles/particle_handler.cpp:205).
//
// One wavefront per cell: the 64 lanes load 64 consecutive particles
// (coalesced) and compute their 20 contributions in parallel; the rows go
// to LDS and lanes 0..19 then add column q row by row, so every component is
// summed in exactly the particle order of the sequential loop.
// ---------------------------------------------------------------------------
struct DepCellArgs {
  Geo g;
  const double *x[3], *v[3];
  const int32_t *sp;
  const int32_t *beg, *end;
  SpeciesParams s;
  double *cj, *cr;
  int rho;
};

constexpr int kDepWaves = kBlock / 64;
constexpr int kDepNV = 20;       // 12 J + 8 rho
constexpr int kDepStride = 21;   // padded row length in LDS

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__global__ void __launch_bounds__(kBlock) k_deposit_cells(DepCellArgs a) {
  __shared__ double rows[kDepWaves][64 * kDepStride];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  const Geo &g = a.g;
  const int nv = a.rho ? 20 : 12;
  double *row = rows[w];
  for (long long c = (long long)blockIdx.x * kDepWaves + w; c < g.ncell;
       c += (long long)gridDim.x * kDepWaves) {
    const int b = a.beg[c], e = a.end[c];
    double acc = 0.0;
    if (e > b) {
      int iz = (int)(c % g.nc[2]);
      int iy = (int)(((c - iz) / g.nc[2]) % g.nc[1]);
      int ix = (int)((((c - iz) / g.nc[2]) - iy) / g.nc[1]);
      double cellverts[3];
      cellverts[0] = g.x0[0] + g.d[0] * ix;
      cellverts[1] = g.x0[1] + g.d[1] * iy;
      cellverts[2] = g.x0[2] + g.d[2] * iz;
      const double *lcell = g.d;
      const double cellVolume = lcell[0] * lcell[1] * lcell[2];
      for (int base = b; base < e; base += 64) {
        const int p = base + lane;
        bool valid = false;
        if (p < e) {
          int sp = a.sp[p];
          if (!a.s.test[sp]) {
            valid = true;
            double pcharge = a.s.q[sp];
            double relpos[3] = {a.x[0][p] - cellverts[0], a.x[1][p] - cellverts[1], a.x[2][p] - cellverts[2]};
            double vel[3] = {a.v[0][p], a.v[1][p], a.v[2][p]};
            double *r = row + lane * kDepStride;
#pragma unroll
            for (int q = 0; q < 3; q++) {
              int j = (q + 1) % 3, k = (q + 2) % 3;
              double cellArea = lcell[j] * lcell[k];
              double curIn1D = kUnitRhoJ * pcharge * vel[q];
              double cur2AreaRatio = (curIn1D / cellVolume) / cellArea;
              r[4 * q] = cur2AreaRatio * (lcell[j] - relpos[j]) * (lcell[k] - relpos[k]);
              r[4 * q + 1] = cur2AreaRatio * relpos[j] * (lcell[k] - relpos[k]);
              r[4 * q + 2] = cur2AreaRatio * relpos[j] * relpos[k];
              r[4 * q + 3] = cur2AreaRatio * (lcell[j] - relpos[j]) * relpos[k];
            }
            if (a.rho) {
              double cc = kUnitRhoJ * (pcharge / cellVolume) / cellVolume;
              r[12] = cc * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
              r[13] = cc * relpos[0] * (lcell[1] - relpos[1]) * (lcell[2] - relpos[2]);
              r[14] = cc * relpos[0] * relpos[1] * (lcell[2] - relpos[2]);
              r[15] = cc * (lcell[0] - relpos[0]) * relpos[1] * (lcell[2] - relpos[2]);
              r[16] = cc * (lcell[0] - relpos[0]) * relpos[1] * relpos[2];
              r[17] = cc * relpos[0] * relpos[1] * relpos[2];
              r[18] = cc * relpos[0] * (lcell[1] - relpos[1]) * relpos[2];
              r[19] = cc * (lcell[0] - relpos[0]) * (lcell[1] - relpos[1]) * relpos[2];
            }
          }
        }
        const unsigned long long m = __ballot(valid);
        wave_sync();
        const int cnt = min(64, e - base);
        if (lane < nv) {
          for (int rr = 0; rr < cnt; rr++)
            if ((m >> rr) & 1ull) acc += row[rr * kDepStride + lane];
        }
        wave_sync();
      }
    }
    // per-cell vectors, component-major so that stage 2 reads coalesce
    if (lane < 12) a.cj[lane * g.ncell + c] = acc;
    else if (lane < nv) a.cr[(lane - 12) * g.ncell + c] = acc;
  }
}

__global__ void __launch_bounds__(kBlock) k_cell_ids(Geo g, const double *x0, const double *x1,
                                                      const double *x2, int32_t *out, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int ix, iy, iz;
  out[i] = cell_of(g, x0[i], x1[i], x2[i], &ix, &iy, &iz);
}

}  // namespace

void launch_gather_f64(emo_ctx *c) {
  if (c->np == 0) return;
  GatherArgs a;
  a.g = c->geo;
  for (int d = 0; d < 3; d++) {
    a.x[d] = c->p.x[d];
    a.E[d] = c->E[d];
    a.B[d] = c->B[d];
    a.Bt[d] = c->Bt[d];
  }
  for (int d = 0; d < 6; d++) a.fld[d] = c->p.fld[d];
  a.flag = c->dflag;
  a.np = c->np;
  TimedLaunch t(c, "gather_f64");
  hipLaunchKernelGGL(k_gather, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
}

void launch_push_f64(emo_ctx *c, double dt) {
  if (c->np == 0) return;
  PushArgs a;
  for (int d = 0; d < 3; d++) {
    a.x[d] = c->p.x[d];
    a.v[d] = c->p.v[d];
  }
  for (int d = 0; d < 6; d++) a.fld[d] = c->p.fld[d];
  a.gam = c->p.gam;
  a.sp = c->p.sp;
  a.s = c->species;
  a.dt = dt;
  a.np = c->np;
  TimedLaunch t(c, "push_f64");
  if (c->cfg.relativity)
    hipLaunchKernelGGL(k_push_vay, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
  else
    hipLaunchKernelGGL(k_push_boris, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
}

void launch_particle_bc_f64(emo_ctx *c) {
  if (c->np == 0) return;
  BcArgs a;
  for (int d = 0; d < 3; d++) {
    a.x[d] = c->p.x[d];
    a.v[d] = c->p.v[d];
    // BC_P_*: xMin_ = xyz0, xMax_ = xMin_ + Lxyz (bc_p_periodic.cpp:41-42)
    a.xmin[d] = c->cfg.xyz0[d];
    a.xmax[d] = a.xmin[d] + c->cfg.Lxyz[d];
  }
  for (int s = 0; s < 6; s++) a.pbc[s] = c->cfg.parts_bound[s];
  a.np = c->np;
  TimedLaunch t(c, "particle_bc_f64");
  hipLaunchKernelGGL(k_particle_bc, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, a);
}

size_t sort_tmp_bytes_for(int64_t n, int bits) {
  size_t bytes = 0;
  hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                     (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)n, 0, bits);
  return bytes;
}

// Stable sort of the particle store by ascending cell id: keys, a stable
// LSD radix sort of (key, index) pairs over the cell-id bits only, a payload
// permutation into the alternate buffers, and per-cell [beg,end) ranges.
int sort_f64(emo_ctx *c, int with_fields) {
  if (c->np == 0) {
    c->sorted = true;
    return EMO_OK;
  }
  KeyArgs k;
  k.g = c->geo;
  for (int d = 0; d < 3; d++) k.x[d] = c->p.x[d];
  k.key = c->key;
  k.perm = c->perm;
  k.flag = c->dflag;
  k.np = c->np;
  {
    TimedLaunch t(c, "sort_keys");
    hipLaunchKernelGGL(k_cell_keys, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, k);
  }
  size_t bytes = c->sort_tmp_bytes;
  hipError_t e;
  {
    TimedLaunch t(c, "sort_radix");
    e = hipcub::DeviceRadixSort::SortPairs(c->sort_tmp, bytes, c->key, c->key_alt, c->perm,
                                           c->perm_alt, (int)c->np, 0, c->sort_bits, c->stream);
  }
  if (e != hipSuccess) {
    c->err = std::string("hipcub radix sort: ") + hipGetErrorString(e);
    return EMO_ERR_HIP;
  }
  PermuteArgs pa;
  pa.src = c->p;
  pa.dst = c->palt;
  pa.perm = c->perm_alt;
  pa.np = c->np;
  pa.fields = with_fields;
  {
    TimedLaunch t(c, "sort_permute");
    hipLaunchKernelGGL(k_permute, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, pa);
  }
  std::swap(c->p, c->palt);
  if (!with_fields)
    for (int d = 0; d < 6; d++) std::swap(c->p.fld[d], c->palt.fld[d]);
  std::swap(c->key, c->key_alt);
  EMO_HIP_CHECK_RET(c, hipMemsetAsync(c->cbeg, 0, c->geo.ncell * sizeof(int32_t), c->stream));
  EMO_HIP_CHECK_RET(c, hipMemsetAsync(c->cend, 0, c->geo.ncell * sizeof(int32_t), c->stream));
  hipLaunchKernelGGL(k_cell_bounds, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, c->key,
                     c->cbeg, c->cend, (long long)c->np);
  c->sorted = true;
  return EMO_OK;
}

void launch_deposit_f64(emo_ctx *c, int rho) {
  DepCellArgs a;
  a.g = c->geo;
  for (int d = 0; d < 3; d++) {
    a.x[d] = c->p.x[d];
    a.v[d] = c->p.v[d];
  }
  a.sp = c->p.sp;
  a.beg = c->cbeg;
  a.end = c->cend;
  a.s = c->species;
  a.cj = c->cj;
  a.cr = c->cr;
  a.rho = rho;
  {
    TimedLaunch t(c, "deposit_cells_f64");
    long long blocks = (c->geo.ncell + kDepWaves - 1) / kDepWaves;
    hipLaunchKernelGGL(k_deposit_cells, dim3((unsigned)std::min<long long>(blocks, 1 << 20)), dim3(kBlock), 0,
                       c->stream, a);
  }
  DepGridArgs<double> b;
  b.g = c->geo;
  b.cj = c->cj;
  b.cr = c->cr;
  for (int d = 0; d < 3; d++) b.J[d] = c->J[d];
  b.rho = c->rho;
  b.do_rho = rho;
  TimedLaunch t(c, "deposit_grid_f64");
  hipLaunchKernelGGL(k_deposit_grid<double>, dim3(grid_for(c->geo.ntot)), dim3(kBlock), 0, c->stream, b);
}

void launch_cell_ids(emo_ctx *c, int32_t *out) {
  if (c->np == 0) return;
  hipLaunchKernelGGL(k_cell_ids, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, c->geo,
                     (const double *)c->p.x[0], (const double *)c->p.x[1],
                     (const double *)c->p.x[2], out, (long long)c->np);
}

}  // namespace emo
