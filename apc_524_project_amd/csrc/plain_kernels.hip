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

#include "cellsort.hpp"
#include "deposit_grid.hpp"
#include "emo_internal.hpp"
#include "pic_math.hpp"
#include "uniform_load.hpp"

namespace emo {
namespace {


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

// particles per cell of a sorted key array: one atomic per run of equal keys
__global__ void __launch_bounds__(kBlock) k_count_sorted(const uint32_t *key, int32_t *count, long long np) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < np;
  const int cell = live ? (int)key[i] : -1;
  const int prev = __shfl_up(cell, 1, 64);
  const bool head = live && (lane == 0 || prev != cell);
  const unsigned long long heads = __ballot(head);
  const int nlive = __popcll(__ballot(live));
  if (head) {
    unsigned long long above = heads & ~(lanes_below(lane) | (1ull << lane));
    int next = above ? __ffsll((long long)above) - 1 : 64;
    atomicAdd(&count[cell], min(next, nlive) - lane);
  }
}

// ---------------------------------------------------------------------------
// Deposit, stage 1: per cell, the sums of Depositor::deposit_particle_J and
// _Rho (src/particles/deposit.cpp:12-62) over the cell's particles in array
// order, with the cell's vertex from Grid::getCellVertex
// (src/grid/oGrid.cpp:233-244); test particles are skipped
// (src/particles/particle_handler.cpp:205).
//
// Sixteen lanes per cell, four cells per wavefront: in batches of 16
// consecutive particles (coalesced) each lane computes one particle's 20
// contributions; the rows go to LDS, and lane l then adds column l (and
// column 16 + l) of the batch row by row, so every component is summed in
// exactly the particle order of the sequential loop.  Rows past the cell's
// end and rows of test particles are zero: adding +0.0 to a sum that starts
// at +0.0 leaves every bit unchanged (such a sum is never -0.0).
// ---------------------------------------------------------------------------
struct DepCellArgs {
  Geo g;
  const double *x[3], *v[3];
  const int32_t *sp;
  const int32_t *off;  // cell c holds [off[c], off[c+1])
  SpeciesParams s;
  double *cj, *cr;
  int rho;
};

constexpr int kDepLanes = 16;                       // lanes per cell
constexpr int kDepCells = kBlock / kDepLanes;       // cells per block
constexpr int kDepStride = 21;                      // padded row length in LDS

__global__ void __launch_bounds__(kBlock) k_deposit_cells(DepCellArgs a) {
  __shared__ double rows[kDepCells][kDepLanes * kDepStride];
  const int l = threadIdx.x & (kDepLanes - 1);
  const int slot = threadIdx.x / kDepLanes;
  const Geo &g = a.g;
  const int nv = a.rho ? 20 : 12;
  double *row = rows[slot];
  const long long c = (long long)blockIdx.x * kDepCells + slot;
  const bool cell_ok = c < g.ncell;
  int b = 0, e = 0;
  double cellverts[3] = {0.0, 0.0, 0.0};
  if (cell_ok) {
    b = a.off[c];
    e = a.off[c + 1];
    int iz = (int)(c % g.nc[2]);
    int iy = (int)(((c - iz) / g.nc[2]) % g.nc[1]);
    int ix = (int)((((c - iz) / g.nc[2]) - iy) / g.nc[1]);
    cellverts[0] = g.x0[0] + g.d[0] * ix;
    cellverts[1] = g.x0[1] + g.d[1] * iy;
    cellverts[2] = g.x0[2] + g.d[2] * iz;
  }
  // the wavefront runs as many batches as its fullest cell needs
  int nb = (e - b + kDepLanes - 1) / kDepLanes;
  for (int o = kDepLanes; o < 64; o <<= 1) nb = max(nb, __shfl_xor(nb, o));
  double acc0 = 0.0, acc1 = 0.0;
  for (int t = 0; t < nb; t++) {
    const int p = b + t * kDepLanes + l;
    double r20[20];
#pragma unroll
    for (int q = 0; q < 20; q++) r20[q] = 0.0;
    if (p < e) {
      const int sp = a.sp[p];
      if (!a.s.test[sp]) {
        double pos[3] = {a.x[0][p], a.x[1][p], a.x[2][p]};
        double vel[3] = {a.v[0][p], a.v[1][p], a.v[2][p]};
        deposit_row(g, cellverts, pos, vel, a.s.q[sp], a.rho != 0, r20);
      }
    }
    double *r = row + l * kDepStride;
#pragma unroll
    for (int q = 0; q < 20; q++) r[q] = r20[q];
    wave_sync();
#pragma unroll
    for (int rr = 0; rr < kDepLanes; rr++) acc0 += row[rr * kDepStride + l];
    if (l < 4) {
#pragma unroll
      for (int rr = 0; rr < kDepLanes; rr++) acc1 += row[rr * kDepStride + 16 + l];
    }
    wave_sync();
  }
  // per-cell vectors, component-major so that stage 2 reads coalesce
  if (cell_ok) {
    if (l < 12) a.cj[l * g.ncell + c] = acc0;
    else if (l < nv) a.cr[(l - 12) * g.ncell + c] = acc0;
    if (l < 4 && nv == 20) a.cr[(l + 4) * g.ncell + c] = acc1;
  }
}

__global__ void __launch_bounds__(kBlock) k_cell_ids(Geo g, const double *x0, const double *x1,
                                                      const double *x2, int32_t *out, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  int ix, iy, iz;
  out[i] = cell_of(g, x0[i], x1[i], x2[i], &ix, &iy, &iz);
}

// ---------------------------------------------------------------------------
// Fused step of the plain form (emo_step on a cell-sorted store):
// gather + push + particle BCs + move code in one pass, then the stable
// counting sort of cellsort.hpp and a scatter.  The per-particle arithmetic
// is pic_math.hpp's, so the result equals the separate phases bit for bit;
// the gathered fields are not stored (the next push gathers them again from
// the same grid).
// ---------------------------------------------------------------------------
struct Pgd64Args {
  Geo g;
  const double *x[3], *v[3], *gam;
  const int32_t *sp;
  double *fld[6];  // written only with write_fld
  uint8_t *flips;  // when not null (the adjoint's tape): bit d = odd number of reflections in dim d
  double *xo[3], *vo[3], *go;
  const double *E[3], *B[3], *Bt[3];
  SpeciesParams s;
  double dt;
  double xmin[3], xmax[3];
  int pbc[6];
  int per[3];
  uint8_t *code;
  int32_t *nfar;
  int write_fld;
  unsigned long long *flag;
  long long np;
};

template <bool REL>
__global__ void __launch_bounds__(kBlock) k_pgd64(Pgd64Args a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = threadIdx.x & 63;
  const bool live = i < a.np;
  const Geo &g = a.g;
  int mcode = 0;
  if (live) {
    double pos[3] = {a.x[0][i], a.x[1][i], a.x[2][i]};
    double vel[3] = {a.v[0][i], a.v[1][i], a.v[2][i]};
    double gam = a.gam[i];
    int sx, sy, sz;
    double f[6] = {0, 0, 0, 0, 0, 0};
    if (cell_of(g, pos[0], pos[1], pos[2], &sx, &sy, &sz) < 0) {
      atomicOr(a.flag, kFlagOutOfGrid);
      sx = sy = sz = 1;
    } else {
      double cv[21];
      cell_vector(g, a.E, a.B, a.Bt, sx, sy, sz, cv);
      interp_fields(g, cv, pos, f);
    }
    if (a.write_fld)
      for (int q = 0; q < 6; q++) a.fld[q][i] = f[q];
    const int sp = a.sp[i];
    if (REL) vay_step(pos, vel, gam, f, a.s.q[sp], a.s.m[sp], a.dt);
    else boris_step(pos, vel, f, a.s.q[sp], a.s.m[sp], a.dt);
    int flips[3];
    particle_bcs(pos, vel, a.xmin, a.xmax, a.pbc, flips);
    if (a.flips) a.flips[i] = (uint8_t)((flips[0] & 1) | (flips[1] & 1) << 1 | (flips[2] & 1) << 2);
    int dx, dy, dz;
    if (cell_of(g, pos[0], pos[1], pos[2], &dx, &dy, &dz) < 0) {
      atomicOr(a.flag, kFlagOutOfGrid);
      mcode = kFarCode;
    } else {
      int k0 = move_code(dx - sx, g.n[0], a.per[0]);
      int k1 = move_code(dy - sy, g.n[1], a.per[1]);
      int k2 = move_code(dz - sz, g.n[2], a.per[2]);
      mcode = (k0 < 0 || k1 < 0 || k2 < 0) ? (int)kFarCode : (k0 * 6 + k1) * 6 + k2;
    }
    for (int d = 0; d < 3; d++) {
      a.xo[d][i] = pos[d];
      a.vo[d][i] = vel[d];
    }
    a.go[i] = gam;
    a.code[i] = (uint8_t)mcode;
  }
  const int nf = __popcll(__ballot(live && mcode == kFarCode));
  if (lane == 0 && nf) atomicAdd(a.nfar, nf);
}

// a particle's cell from its position, clamped into the grid (an
// out-of-grid particle has already raised the flag)
__device__ inline int cell_clamped(const Geo &g, double x, double y, double z) {
  int ix, iy, iz;
  int c = cell_of(g, x, y, z, &ix, &iy, &iz);
  if (c >= 0) return c;
  ix = min(max(ix, 1), g.n[0]);
  iy = min(max(iy, 1), g.n[1]);
  iz = min(max(iz, 1), g.n[2]);
  return (g.nc[1] * g.nc[2]) * ix + g.nc[2] * iy + iz;
}

// source (old position) and destination (new position) cells
struct CellsF64 {
  Geo g;
  const double *xs[3], *xd[3];
  __device__ void operator()(long long i, int &s, int &d) const {
    s = cell_clamped(g, xs[0][i], xs[1][i], xs[2][i]);
    d = cell_clamped(g, xd[0][i], xd[1][i], xd[2][i]);
  }
};

struct Scatter64Args {
  Geo g;
  CellSortBufs b;
  const double *xi[3], *vi[3], *gi;  // new state, previous order
  const int32_t *spi;
  const int64_t *idi;
  double *xo[3], *vo[3], *go;        // new state, sorted
  int32_t *spo;
  int64_t *ido;
  uint32_t *perm;                    // sorted position -> previous index
  uint32_t *inv;                     // previous index -> sorted position (optional)
  unsigned long long *flag;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_invert_perm(const uint32_t *perm, uint32_t *inv, long long np) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < np) inv[perm[j]] = (uint32_t)j;
}

__global__ void __launch_bounds__(kBlock) k_scatter64(Scatter64Args a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const double x0 = a.xi[0][i], x1 = a.xi[1][i], x2 = a.xi[2][i];
  const int dst = cell_clamped(a.g, x0, x1, x2);
  const long long pos = cellsort_pos(a.g, a.b, i, a.b.code[i], dst, a.flag);
  if (pos < 0) return;
  a.xo[0][pos] = x0;
  a.xo[1][pos] = x1;
  a.xo[2][pos] = x2;
  for (int d = 0; d < 3; d++) a.vo[d][pos] = a.vi[d][i];
  a.go[pos] = a.gi[i];
  a.spo[pos] = a.spi[i];
  a.ido[pos] = a.idi[i];
  a.perm[pos] = (uint32_t)i;
  if (a.inv) a.inv[i] = (uint32_t)pos;
}

}  // namespace

void launch_gather_f64(emo_ctx *c) {
  c->fld_state = kFldFresh;
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
    cs_fill_i32(c, c->srt.off, 0, c->geo.ncell + 1);
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
  cs_fill_i32(c, c->srt.count, 0, c->geo.ncell + 1);
  hipLaunchKernelGGL(k_count_sorted, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, (const uint32_t *)c->key,
                     c->srt.count, (long long)c->np);
  int rc = cellsort_scan(c, c->srt, c->srt.off, "sort_ranges");
  if (rc) return rc;
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
  a.off = c->srt.off;
  a.s = c->species;
  a.cj = c->cj;
  a.cr = c->cr;
  a.rho = rho;
  {
    TimedLaunch t(c, "deposit_cells_f64");
    const long long blocks = (c->geo.ncell + kDepCells - 1) / kDepCells;
    hipLaunchKernelGGL(k_deposit_cells, dim3((unsigned)blocks), dim3(kBlock), 0, c->stream, a);
  }
  DepGridArgs<double> b;
  b.g = c->geo;
  b.cj = c->cj;
  b.cr = c->cr;
  b.sj = b.sr = c->geo.ncell;
  for (int d = 0; d < 3; d++) b.J[d] = c->J[d];
  b.rho = c->rho;
  b.do_rho = rho;
  TimedLaunch t(c, "deposit_grid_f64");
  hipLaunchKernelGGL(k_deposit_grid<double>, dim3(grid_for(c->geo.ntot)), dim3(kBlock), 0, c->stream, b);
}

// emo_load_uniform's particles in the FP64 store (generation order)
struct GenF64Args {
  UniformGen u;
  ParticlesF64 p;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_uniform_f64(GenF64Args a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  double x[3], v[3], gam;
  int sp;
  ugen_particle(a.u, i, x, v, &gam, &sp);
  for (int d = 0; d < 3; d++) {
    a.p.x[d][i] = x[d];
    a.p.v[d][i] = v[d];
  }
  a.p.gam[i] = gam;
  a.p.sp[i] = sp;
  a.p.id[i] = i;
  for (int q = 0; q < 6; q++) a.p.fld[q][i] = 0.0;
}

void launch_uniform_f64(emo_ctx *c, const UniformGen &u, int64_t n) {
  if (n <= 0) return;
  GenF64Args a;
  a.u = u;
  a.p = c->p;
  a.np = n;
  hipLaunchKernelGGL(k_uniform_f64, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
}

// The particles with id < nid, by id (Particle_Handler::outputParticles,
// src/particles/particle_handler.cpp:308-365, writes these rows): out[7 id +
// 0..2] = x, [3..5] = v, [6] = 1 (found); the caller zeroes out.
__global__ void __launch_bounds__(kBlock) k_tracked_f64(ParticlesF64 p, long long np, long long nid, double *out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const long long k = p.id[i];
  if (k < 0 || k >= nid) return;
  double *o = out + 7 * k;
  for (int d = 0; d < 3; d++) {
    o[d] = p.x[d][i];
    o[3 + d] = p.v[d][i];
  }
  o[6] = 1.0;
}

void launch_tracked_f64(emo_ctx *c, long long nid, double *out) {
  launch_const_fill_f64(c, out, 0.0, 7 * nid);
  if (c->np == 0) return;
  hipLaunchKernelGGL(k_tracked_f64, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, c->p, (long long)c->np, nid,
                     out);
}

void launch_cell_ids(emo_ctx *c, int32_t *out) {
  if (c->np == 0) return;
  hipLaunchKernelGGL(k_cell_ids, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, c->geo,
                     (const double *)c->p.x[0], (const double *)c->p.x[1],
                     (const double *)c->p.x[2], out, (long long)c->np);
}


void fld_materialize(emo_ctx *c) {
  if (c->fld_state != kFldDeferred) return;
  launch_gather_f64(c);
  c->fld_state = kFldFresh;
}

void fld_invalidate(emo_ctx *c) {
  fld_materialize(c);
  c->fld_state = kFldStale;
}

int plain_sort_create(emo_ctx *c) { return cellsort_alloc(c, c->srt, c->cap, c->geo.ncell); }

void plain_sort_destroy(emo_ctx *c) { cellsort_free(c->srt); }

// The fused particle phase of the plain form on a cell-sorted store, part 1:
// push (gathering the fields from the grid) and BCs into palt, with the move
// code of every particle.  With fld_out the gathered fields are stored there
// (six arrays, the pre-push order); with flips_out the reflections of the
// particle BCs.
void fused_push_f64(emo_ctx *c, double dt, double *const *fld_out, uint8_t *flips_out) {
  const Geo &g = c->geo;
  const long long np = c->np;
  Pgd64Args a;
  a.g = g;
  for (int d = 0; d < 3; d++) {
    a.x[d] = c->p.x[d];
    a.v[d] = c->p.v[d];
    a.xo[d] = c->palt.x[d];
    a.vo[d] = c->palt.v[d];
    a.E[d] = c->E[d];
    a.B[d] = c->B[d];
    a.Bt[d] = c->Bt[d];
    a.xmin[d] = c->cfg.xyz0[d];
    a.xmax[d] = a.xmin[d] + c->cfg.Lxyz[d];
    a.per[d] = c->cfg.parts_bound[2 * d] == EMO_BC_PERIODIC || c->cfg.parts_bound[2 * d + 1] == EMO_BC_PERIODIC;
  }
  for (int q = 0; q < 6; q++) a.fld[q] = fld_out ? fld_out[q] : nullptr;
  a.gam = c->p.gam;
  a.go = c->palt.gam;
  a.sp = c->p.sp;
  a.s = c->species;
  a.dt = dt;
  for (int s = 0; s < 6; s++) a.pbc[s] = c->cfg.parts_bound[s];
  a.code = c->srt.code;
  a.nfar = c->srt.misc;
  a.write_fld = fld_out != nullptr;
  a.flips = flips_out;
  a.flag = c->dflag;
  a.np = np;
  {
    TimedLaunch t(c, "pgd_f64");
    if (c->cfg.relativity)
      hipLaunchKernelGGL(k_pgd64<true>, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, a);
    else
      hipLaunchKernelGGL(k_pgd64<false>, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, a);
  }
}

// part 2: the stable counting sort by the new cell and the scatter into the
// main store; the permutation (new position -> previous index) goes to
// perm_out (c->srt.perm if null), its inverse to inv_out (if given), and the
// new cell ranges to c->srt.off.
int fused_sort_f64(emo_ctx *c, uint32_t *perm_out, uint32_t *inv_out) {
  if (!perm_out) perm_out = c->srt.perm;
  const Geo &g = c->geo;
  const long long np = c->np;
  static const char *const names[4] = {"far_f64", "rank_f64", "base_f64", "scan_f64"};
  CellsF64 cells;
  cells.g = g;
  for (int d = 0; d < 3; d++) {
    cells.xs[d] = c->p.x[d];
    cells.xd[d] = c->palt.x[d];
  }
  int rc = cellsort_offsets(c, c->srt, np, cells, names);
  if (rc) return rc;
  // The counting sort holds at most kFarCap particles that move further than
  // one cell and 65535 per (cell, move) group.  Its flags are read back
  // before the scatter: on overflow this step sorts with the radix sort
  // instead (the same stable order), from the pushed state in palt.
  if (hipMemcpyAsync(c->hflag, c->dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess) {
    c->err = "flag readback failed";
    return EMO_ERR_HIP;
  }
  const unsigned long long fl = c->hflag[0];
  if (fl & kFlagOutOfGrid) {
    c->err = "particle outside the padded grid (Grid::getCellID asserts, src/grid/oGrid.cpp:212-220)";
    return EMO_ERR_INVARIANT;
  }
  if (fl & (kFlagFarMovers | kFlagMigrate)) {
    c->hflag[0] = fl & ~(kFlagFarMovers | kFlagMigrate);
    if (hipMemcpyAsync(c->dflag, c->hflag, sizeof(unsigned long long), hipMemcpyHostToDevice, c->stream) !=
            hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess) {
      c->err = "flag reset failed";
      return EMO_ERR_HIP;
    }
    cs_fill_i32(c, c->srt.misc, 0, 1);
    for (int d = 0; d < 3; d++) {
      std::swap(c->p.x[d], c->palt.x[d]);
      std::swap(c->p.v[d], c->palt.v[d]);
    }
    std::swap(c->p.gam, c->palt.gam);
    rc = sort_f64(c, 0);
    if (rc) return rc;
    if (hipMemcpyAsync(perm_out, c->perm_alt, (size_t)np * 4, hipMemcpyDeviceToDevice, c->stream) !=
        hipSuccess) {
      c->err = "permutation copy failed";
      return EMO_ERR_HIP;
    }
    if (inv_out)
      hipLaunchKernelGGL(k_invert_perm, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, (const uint32_t *)perm_out,
                         inv_out, (long long)np);
    c->fallback_sorts++;
    return EMO_OK;
  }
  Scatter64Args sa;
  sa.g = g;
  sa.b = c->srt;
  for (int d = 0; d < 3; d++) {
    sa.xi[d] = c->palt.x[d];
    sa.vi[d] = c->palt.v[d];
    sa.xo[d] = c->p.x[d];
    sa.vo[d] = c->p.v[d];
  }
  sa.gi = c->palt.gam;
  sa.go = c->p.gam;
  sa.spi = c->p.sp;
  sa.idi = c->p.id;
  sa.spo = c->palt.sp;
  sa.ido = c->palt.id;
  sa.perm = perm_out;
  sa.inv = inv_out;
  sa.flag = c->dflag;
  sa.np = np;
  {
    TimedLaunch t(c, "scatter_f64");
    hipLaunchKernelGGL(k_scatter64, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, sa);
  }
  std::swap(c->p.sp, c->palt.sp);
  std::swap(c->p.id, c->palt.id);
  std::swap(c->srt.off, c->srt.off_new);
  cs_fill_i32(c, c->srt.misc, 0, 1);
  c->sorted = true;
  return EMO_OK;
}

}  // namespace emo
