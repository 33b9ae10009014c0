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
#include "pic_math.hpp"

namespace emo {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
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
  double cv[21], fld[6];
  cell_vector(g, a.E, a.B, a.Bt, ix, iy, iz, cv);
  interp_fields(g, cv, pos, fld);
#pragma unroll
  for (int q = 0; q < 6; q++) a.fld[q][i] = fld[q];
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
  double xp[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
  double vp[3] = {a.v[0][i], a.v[1][i], a.v[2][i]};
  double f[6];
#pragma unroll
  for (int q = 0; q < 6; q++) f[q] = a.fld[q][i];
  boris_step(xp, vp, f, a.s.q[sp], a.s.m[sp], a.dt);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.x[d][i] = xp[d];
    a.v[d][i] = vp[d];
  }
}

__global__ void __launch_bounds__(kBlock) k_push_vay(PushArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  int sp = a.sp[i];
  double xp[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
  double vp[3] = {a.v[0][i], a.v[1][i], a.v[2][i]};
  double gam = a.gam[i];
  double f[6];
#pragma unroll
  for (int q = 0; q < 6; q++) f[q] = a.fld[q][i];
  vay_step(xp, vp, gam, f, a.s.q[sp], a.s.m[sp], a.dt);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.x[d][i] = xp[d];
    a.v[d][i] = vp[d];
  }
  a.gam[i] = gam;
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
  double xp[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
  double vp[3] = {a.v[0][i], a.v[1][i], a.v[2][i]};
  int flips[3];
  particle_bcs(xp, vp, a.xmin, a.xmax, a.pbc, flips);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    a.x[d][i] = xp[d];
    a.v[d][i] = vp[d];
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
      for (int base = b; base < e; base += 64) {
        const int p = base + lane;
        bool valid = false;
        if (p < e) {
          int sp = a.sp[p];
          if (!a.s.test[sp]) {
            valid = true;
            double pos[3] = {a.x[0][p], a.x[1][p], a.x[2][p]};
            double vel[3] = {a.v[0][p], a.v[1][p], a.v[2][p]};
            double r20[20];
            deposit_row(g, cellverts, pos, vel, a.s.q[sp], a.rho != 0, r20);
            double *r = row + lane * kDepStride;
#pragma unroll
            for (int q = 0; q < 20; q++)
              if (q < nv) r[q] = r20[q];
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
