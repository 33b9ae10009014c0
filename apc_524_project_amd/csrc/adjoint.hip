// adjoint.hip -- discrete adjoint of the FP64 plain-form step for gfx950.
//
// The objective is the final field energy of OutputBoxQuantities
// (src/IO/output.cpp:90-121): F = E_en + B_en from E and the fixed B
// allocations (Grid::fieldPtr_).  Its gradient with respect to the injected
// pulses' amplitudes and phases (LightBC, src/grid/lightBC.cpp:165-198) and
// to the initial particle state is computed by running the exact transpose
// of every operation of the step backward in time:
//
//   forward step s:  push -> particle BCs -> stable sort -> deposit J -> fold
//                    -> Yee (swap, E, B) -> ghost refresh / LightBC -> gather
//   adjoint step s:  gather^T -> refresh^T -> Yee^T -> fold^T -> deposit^T
//                    -> sort^T -> BCs^T -> push^T
//
// Forward states are checkpointed every K steps; each segment is re-run from
// its checkpoint with a tape (pre-push particle state, the sort permutation
// and cell ranges, the fields at the start of the step), then walked
// backward.  Every accumulation is in a fixed order (per-cell sums in array
// order, per-point sums over neighbouring cells in a fixed order, fixed
// reduction trees): the gradient is deterministic.
//
// Cell changes of particles are piecewise constant in the inputs and the
// reference's linear weights jump at cell faces (its least vertex is the
// cell's upper corner), so the adjoint is the derivative of the step away
// from those faces -- the derivative that central differences reproduce when
// no particle changes cell under the perturbation.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint_math.hpp"
#include "chosen_math.hpp"
#include "emo_internal.hpp"

namespace emo {
namespace {

constexpr int kBlock = 256;
constexpr int kWaves = kBlock / 64;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline bool is_phys(const Geo &g, int i, int j, int k) {
  return i >= 1 && i <= g.n[0] && j >= 1 && j <= g.n[1] && k >= 1 && k <= g.n[2];
}

__device__ inline void point_ijk(const Geo &g, long long p, int *i, int *j, int *k) {
  *k = (int)(p % g.nt[2]);
  long long r = p / g.nt[2];
  *j = (int)(r % g.nt[1]);
  *i = (int)(r / g.nt[1]);
}

// ---------------------------------------------------------------------------
// Objective seed: dF/dE and dF/dB for F = sum over physical cells of
// 0.5 V (0.5 (f[p] + f[p+s]))^2 per component (s along its own axis), the
// expression of OutputBoxQuantities with the SQR macro expanded.
// ---------------------------------------------------------------------------
template <typename T>
struct SeedArgs {
  Geo g;
  const T *E[3], *B[3];
  double *Eb[3], *Bb[3];
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_seed_energy(SeedArgs<T> a) {
  const Geo &g = a.g;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ntot) return;
  int i, j, k;
  point_ijk(g, p, &i, &j, &k);
  const double vol = g.d[0] * g.d[1] * g.d[2];
  const long long st[3] = {g.sx, g.sy, 1};
  for (int q = 0; q < 3; q++) {
    int di = q == 0, dj = q == 1, dk = q == 2;
    double ge = 0.0, gb = 0.0;
    if (is_phys(g, i, j, k)) {
      ge += 0.25 * vol * ((double)a.E[q][p] + (double)a.E[q][p + st[q]]);
      gb += 0.25 * vol * ((double)a.B[q][p] + (double)a.B[q][p + st[q]]);
    }
    if (is_phys(g, i - di, j - dj, k - dk)) {
      ge += 0.25 * vol * ((double)a.E[q][p - st[q]] + (double)a.E[q][p]);
      gb += 0.25 * vol * ((double)a.B[q][p - st[q]] + (double)a.B[q][p]);
    }
    a.Eb[q][p] += ge;
    a.Bb[q][p] += gb;
  }
}

// ---------------------------------------------------------------------------
// Ghost refresh^T.  Periodic copy dst <- src over the physical range of the
// other two dims: src_b += dst_b, dst_b = 0.  LightBC plane fill over the
// full plane: the plane's adjoint, weighted by the transverse profile, is the
// adjoint of the pulse value; the plane's adjoint is then zero.
// ---------------------------------------------------------------------------
struct FillAdjArgs {
  Geo g;
  double *f;
  int dim, idx;
  const double *prof;  // the plane's transverse profile (nt[u] x nt[w]), or null: uniform
  double *out;         // the weighted plane sum
};

// ---------------------------------------------------------------------------
// Yee^T (Grid::evolveFields, src/grid/oGrid.cpp:15-60).  With Bn the array
// the B update writes and Bo the previous B (B_tm1 after the swap):
//   E' = E - dt (curl_back(Bo) + 4 pi J)     (physical points)
//   Bn = Bo + dt curl_fwd(E')                (physical points)
// Reverse: (1) E'_b += dt curl_fwd^T(Bn_b); (2) Bo_b += Bn_b, Bn_b = 0 on the
// physical points; (3) Bo_b -= dt curl_back^T(E'_b), J_b = -dt 4 pi E'_b.
// Gather form: every point reads its neighbours' adjoints, no atomics.
// ---------------------------------------------------------------------------
struct YeeAdjArgs {
  Geo g;
  double *Eb[3];
  double *Bn[3];
  double *Bo[3];
  double *Jb[3];
  double dt;
};

__device__ inline double phys_val(const Geo &g, const double *f, int i, int j, int k) {
  return is_phys(g, i, j, k) ? f[gidx(g, i, j, k)] : 0.0;
}

__global__ void __launch_bounds__(kBlock) k_yee_b_adj(YeeAdjArgs a) {
  const Geo &g = a.g;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ntot) return;
  int i, j, k;
  point_ijk(g, p, &i, &j, &k);
  const double dt = a.dt, dx = g.d[0], dy = g.d[1], dz = g.d[2];
  const double *bx = a.Bn[0], *by = a.Bn[1], *bz = a.Bn[2];
  // Bx = .. + dt((Ey[c+1]-Ey)/dz - (Ez[c+sy]-Ez)/dy)
  // By = .. + dt((Ez[c+sx]-Ez)/dx - (Ex[c+1]-Ex)/dz)
  // Bz = .. + dt((Ex[c+sy]-Ex)/dy - (Ey[c+sx]-Ey)/dx)
  double ex = (phys_val(g, by, i, j, k) - phys_val(g, by, i, j, k - 1)) / dz +
              (phys_val(g, bz, i, j - 1, k) - phys_val(g, bz, i, j, k)) / dy;
  double ey = (phys_val(g, bx, i, j, k - 1) - phys_val(g, bx, i, j, k)) / dz +
              (phys_val(g, bz, i, j, k) - phys_val(g, bz, i - 1, j, k)) / dx;
  double ez = (phys_val(g, bx, i, j, k) - phys_val(g, bx, i, j - 1, k)) / dy +
              (phys_val(g, by, i - 1, j, k) - phys_val(g, by, i, j, k)) / dx;
  a.Eb[0][p] += dt * ex;
  a.Eb[1][p] += dt * ey;
  a.Eb[2][p] += dt * ez;
}

__global__ void __launch_bounds__(kBlock) k_yee_move_adj(YeeAdjArgs a) {
  const Geo &g = a.g;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ntot) return;
  int i, j, k;
  point_ijk(g, p, &i, &j, &k);
  if (!is_phys(g, i, j, k)) return;
  for (int q = 0; q < 3; q++) {
    a.Bo[q][p] += a.Bn[q][p];
    a.Bn[q][p] = 0.0;
  }
}

__global__ void __launch_bounds__(kBlock) k_yee_e_adj(YeeAdjArgs a) {
  const Geo &g = a.g;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ntot) return;
  int i, j, k;
  point_ijk(g, p, &i, &j, &k);
  const double dt = a.dt, dx = g.d[0], dy = g.d[1], dz = g.d[2];
  const double *ex = a.Eb[0], *ey = a.Eb[1], *ez = a.Eb[2];
  // Ex -= dt((By-By[c-1])/dz - (Bz-Bz[c-sy])/dy + 4pi Jx)
  // Ey -= dt((Bz-Bz[c-sx])/dx - (Bx-Bx[c-1])/dz + 4pi Jy)
  // Ez -= dt((Bx-Bx[c-sy])/dy - (By-By[c-sx])/dx + 4pi Jz)
  double bx = (phys_val(g, ey, i, j, k) - phys_val(g, ey, i, j, k + 1)) / dz -
              (phys_val(g, ez, i, j, k) - phys_val(g, ez, i, j + 1, k)) / dy;
  double by = -(phys_val(g, ex, i, j, k) - phys_val(g, ex, i, j, k + 1)) / dz +
              (phys_val(g, ez, i, j, k) - phys_val(g, ez, i + 1, j, k)) / dx;
  double bz = (phys_val(g, ex, i, j, k) - phys_val(g, ex, i, j + 1, k)) / dy -
              (phys_val(g, ey, i, j, k) - phys_val(g, ey, i + 1, j, k)) / dx;
  a.Bo[0][p] += dt * bx;
  a.Bo[1][p] += dt * by;
  a.Bo[2][p] += dt * bz;
  const double fpi = 4 * kPi;
  const bool ph = is_phys(g, i, j, k);
  for (int q = 0; q < 3; q++) a.Jb[q][p] = ph ? -dt * fpi * a.Eb[q][p] : 0.0;
}

// ---------------------------------------------------------------------------
// Per-particle adjoints.
// ---------------------------------------------------------------------------
struct PartView {
  const double *x[3], *v[3], *gam, *fld[6];
  const int32_t *sp;
};

struct PartAdj {
  double *x[3], *v[3], *gam, *fld[6];
};

// gather^T, stage 2: every point adds the entries of the cells whose
// cell_vector reads it, in a fixed order
struct GatherGridArgs {
  Geo g;
  const double *cg;
  double *Eb[3], *Bb[3], *Btb[3];
};

__device__ inline double cg_at(const Geo &g, const double *cg, int q, int i, int j, int k) {
  if (i < 0 || j < 0 || k < 0 || i >= g.nc[0] || j >= g.nc[1] || k >= g.nc[2]) return 0.0;
  return cg[(long long)q * g.ncell + ((long long)i * g.nc[1] + j) * g.nc[2] + k];
}

__global__ void __launch_bounds__(kBlock) k_gather_adj_grid(GatherGridArgs a) {
  const Geo &g = a.g;
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= g.ntot) return;
  int i, j, k;
  point_ijk(g, p, &i, &j, &k);
  const double *cg = a.cg;
  // Ex: cells (i,j,k)#0, (i,j-1,k)#1, (i,j-1,k-1)#2, (i,j,k-1)#3
  a.Eb[0][p] += cg_at(g, cg, 0, i, j, k) + cg_at(g, cg, 1, i, j - 1, k) + cg_at(g, cg, 2, i, j - 1, k - 1) +
                cg_at(g, cg, 3, i, j, k - 1);
  // Ey: (i,j,k)#4, (i,j,k-1)#5, (i-1,j,k-1)#6, (i-1,j,k)#7
  a.Eb[1][p] += cg_at(g, cg, 4, i, j, k) + cg_at(g, cg, 5, i, j, k - 1) + cg_at(g, cg, 6, i - 1, j, k - 1) +
                cg_at(g, cg, 7, i - 1, j, k);
  // Ez: (i,j,k)#8, (i-1,j,k)#9, (i-1,j-1,k)#10, (i,j-1,k)#11
  a.Eb[2][p] += cg_at(g, cg, 8, i, j, k) + cg_at(g, cg, 9, i - 1, j, k) + cg_at(g, cg, 10, i - 1, j - 1, k) +
                cg_at(g, cg, 11, i, j - 1, k);
  // B faces are (B_tm1 + B)/2: half of the entry to each array
  double bx = cg_at(g, cg, 12, i, j, k) + cg_at(g, cg, 13, i - 1, j, k);
  double by = cg_at(g, cg, 14, i, j, k) + cg_at(g, cg, 15, i, j - 1, k);
  double bz = cg_at(g, cg, 16, i, j, k) + cg_at(g, cg, 17, i, j, k - 1);
  a.Bb[0][p] += bx / 2;
  a.Btb[0][p] += bx / 2;
  a.Bb[1][p] += by / 2;
  a.Btb[1][p] += by / 2;
  a.Bb[2][p] += bz / 2;
  a.Btb[2][p] += bz / 2;
}

__global__ void k_zero(double *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Batched device copies (the tape): up to kMaxCopies arrays in one launch,
// 16 bytes per thread.  Every array starts 16-byte aligned and holds a whole
// number of 4-byte words.
// ---------------------------------------------------------------------------
constexpr int kMaxCopies = 32;
struct CopyList {
  const uint32_t *src[kMaxCopies];
  uint32_t *dst[kMaxCopies];
  long long words[kMaxCopies];
  long long start[kMaxCopies + 1];  // 16-byte chunks before array i
  int n;
};

__global__ void __launch_bounds__(kBlock) k_copy_many(CopyList a) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.start[a.n]) return;
  int s = 0;
  while (t >= a.start[s + 1]) s++;
  long long w = (t - a.start[s]) * 4;
  const long long nw = a.words[s];
  if (w + 4 <= nw) {
    *(uint4 *)(a.dst[s] + w) = *(const uint4 *)(a.src[s] + w);
  } else {
    for (; w < nw; w++) a.dst[s][w] = a.src[s][w];
  }
}

// ---------------------------------------------------------------------------
// refresh^T of one run of periodic sides (one forward launch), gather form:
// a thread per physical point adds the ghost adjoints of the sides whose
// source plane holds it, in reverse side order, and zeroes those ghosts
// (each ghost point has exactly one source point).
// ---------------------------------------------------------------------------
struct SidesAdjArgs {
  Geo g;
  double *f[6];
  int nsides;                      // in reverse forward order
  int d[6], src[6], dst[6];
};

__global__ void __launch_bounds__(kBlock) k_copy_sides_adj(SidesAdjArgs a) {
  const Geo &g = a.g;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  const int at[3] = {1 + (int)(t / ((long long)g.n[1] * g.n[2])), 1 + (int)((t / g.n[2]) % g.n[1]),
                     1 + (int)(t % g.n[2])};
  bool any = false;
  for (int s = 0; s < a.nsides; s++) any |= at[a.d[s]] == a.src[s];
  if (!any) return;
  const long long p = gidx(g, at[0], at[1], at[2]);
  for (int f = 0; f < 6; f++) {
    double val = a.f[f][p];
    for (int s = 0; s < a.nsides; s++) {
      if (at[a.d[s]] != a.src[s]) continue;
      int gh[3] = {at[0], at[1], at[2]};
      gh[a.d[s]] = a.dst[s];
      const long long q = gidx(g, gh[0], gh[1], gh[2]);
      val += a.f[f][q];
      a.f[f][q] = 0.0;
    }
    a.f[f][p] = val;
  }
}

// fill^T of the (up to six) planes of one LightBC side: one block per plane
struct FillAdjMulti {
  FillAdjArgs p[6];
};

__global__ void __launch_bounds__(kBlock) k_plane_fill_adj_multi(FillAdjMulti m) {
  __shared__ double sh[kBlock];
  const FillAdjArgs &a = m.p[blockIdx.x];
  const Geo &g = a.g;
  int d = a.dim;
  int u = d == 0 ? 1 : 0;
  int w = d == 2 ? 1 : 2;
  long long n = (long long)g.nt[u] * g.nt[w];
  double s = 0.0;
  for (long long t = threadIdx.x; t < n; t += kBlock) {
    int pu = (int)(t / g.nt[w]);
    int pw = (int)(t % g.nt[w]);
    int ii[3];
    ii[d] = a.idx;
    ii[u] = pu;
    ii[w] = pw;
    long long p = gidx(g, ii[0], ii[1], ii[2]);
    const double prof = a.prof ? a.prof[t] : 1.0;
    s += a.f[p] * prof;
    a.f[p] = 0.0;
  }
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int st = kBlock / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) *a.out = sh[0];
}

// ---------------------------------------------------------------------------
// The particle part of one adjoint step, per cell of the pre-push order
// (sixteen lanes per cell, four cells per wavefront, batches of 16
// consecutive particles):
//   deposit^T  the current adjoint at the particle's deposit targets, read at
//              its post-sort slot j = inv[i];
//   BCs^T + push^T  from the taped pre-push state i;
//   gather^T   of the gather that produced the push's fields (the previous
//              step's last phase, on this same state): the position adjoint
//              gains the interpolation term, and lane l adds column l (and
//              16 + l) of the batch's 18 field-entry adjoints row by row, so
//              each cell's sums follow the particle order.
// ---------------------------------------------------------------------------
constexpr int kGLanes = 16;                 // lanes per cell
constexpr int kGCells = kBlock / kGLanes;   // cells per block
constexpr int kGStride = 19;

struct CellStepAdjArgs {
  Geo g;
  PartView post;        // post-BC state, sorted order (x, v, sp)
  PartView pre;         // pre-push state, cell-sorted (x, v, gam, fld, sp)
  const uint8_t *flips; // the push's reflections, pre-push order
  const int32_t *off;   // cell c of the pre-push order holds off[c]..off[c+1]
  const uint32_t *inv;  // pre-push index -> post-sort slot
  const double *E[3], *B[3], *Bt[3];  // the grid the pre-push fields came from
  const double *bx[3], *bv[3], *bg;   // adjoint, sorted order (in)
  double *ox[3], *ov[3], *og;         // adjoint, pre-push order (out)
  const double *Jb[3];
  double *cg;  // 18 per cell, component-major
  SpeciesParams sp;
  double dt;
  int rel;
};

__global__ void __launch_bounds__(kBlock, 4) k_cell_step_adj(CellStepAdjArgs a) {
  __shared__ double rows[kGCells][kGLanes * kGStride];
  const int l = threadIdx.x & (kGLanes - 1);
  const int slot = threadIdx.x / kGLanes;
  const Geo &g = a.g;
  const long long c = (long long)blockIdx.x * kGCells + slot;
  const bool cell_ok = c < g.ncell;
  double *row = rows[slot];
  int b = 0, e = 0;
  if (cell_ok) {
    b = a.off[c];
    e = a.off[c + 1];
  }
  int nb = (e - b + kGLanes - 1) / kGLanes;
  for (int o = kGLanes; o < 64; o <<= 1) nb = max(nb, __shfl_xor(nb, o));
  double acc0 = 0.0, acc1 = 0.0;
  for (int t = 0; t < nb; t++) {
    const int i = b + t * kGLanes + l;
    double cvb[21];
#pragma unroll
    for (int q = 0; q < 21; q++) cvb[q] = 0.0;
    if (i < e) {
      const long long j = a.inv[i];
      double xb1[3], vb1[3];
      for (int d = 0; d < 3; d++) {
        xb1[d] = a.bx[d][j];
        vb1[d] = a.bv[d][j];
      }
      const double gb1 = a.bg[j];
      const int sp = a.pre.sp[i];
      const double q = a.sp.q[sp], m = a.sp.m[sp];
      if (!a.sp.test[sp]) {
        double pos[3] = {a.post.x[0][j], a.post.x[1][j], a.post.x[2][j]};
        double vel[3] = {a.post.v[0][j], a.post.v[1][j], a.post.v[2][j]};
        int ix, iy, iz;
        if (cell_of(g, pos[0], pos[1], pos[2], &ix, &iy, &iz) >= 0) {
          double cellverts[3] = {g.x0[0] + g.d[0] * ix, g.x0[1] + g.d[1] * iy, g.x0[2] + g.d[2] * iz};
          long long idx[12];
          deposit_target_index(g, ix, iy, iz, idx);
          double rb[12];
          for (int k = 0; k < 12; k++) rb[k] = a.Jb[k / 4][idx[k]];
          double posb[3] = {0, 0, 0}, velb[3] = {0, 0, 0};
          deposit_row_vjp(g, cellverts, pos, vel, q, rb, posb, velb);
          for (int d = 0; d < 3; d++) {
            xb1[d] += posb[d];
            vb1[d] += velb[d];
          }
        }
      }
      double x[3] = {a.pre.x[0][i], a.pre.x[1][i], a.pre.x[2][i]};
      double v[3] = {a.pre.v[0][i], a.pre.v[1][i], a.pre.v[2][i]};
      const double gam = a.pre.gam[i];
      double f[6];
      for (int k = 0; k < 6; k++) f[k] = a.pre.fld[k][i];
      const int fl = a.flips[i];
      const int flips[3] = {fl & 1, (fl >> 1) & 1, (fl >> 2) & 1};
      particle_bcs_vjp(flips, xb1, vb1);
      double xb[3] = {0, 0, 0}, vb[3] = {0, 0, 0}, gb = 0.0, fb[6] = {0, 0, 0, 0, 0, 0};
      if (a.rel) {
        vay_step_vjp(v, gam, f, q, m, a.dt, xb1, vb1, gb1, xb, vb, gb, fb);
      } else {
        boris_step_vjp(v, f, q, m, a.dt, xb1, vb1, xb, vb, fb);
        gb = gb1;  // Boris leaves the stored gamma untouched
      }
      // gather^T at the pre-push position
      int ix, iy, iz;
      cell_of(g, x[0], x[1], x[2], &ix, &iy, &iz);
      double cv[21], posb[3] = {0, 0, 0};
      cell_vector(g, a.E, a.B, a.Bt, ix, iy, iz, cv);
      interp_fields_vjp(g, cv, x, fb, posb, cvb);
      for (int d = 0; d < 3; d++) {
        a.ox[d][i] = xb[d] + posb[d];
        a.ov[d][i] = vb[d];
      }
      a.og[i] = gb;
    }
    double *r = row + l * kGStride;
#pragma unroll
    for (int q = 0; q < 18; q++) r[q] = cvb[3 + q];
    wave_sync();
#pragma unroll
    for (int rr = 0; rr < kGLanes; rr++) acc0 += row[rr * kGStride + l];
    if (l < 2) {
#pragma unroll
      for (int rr = 0; rr < kGLanes; rr++) acc1 += row[rr * kGStride + 16 + l];
    }
    wave_sync();
  }
  if (cell_ok) {
    a.cg[l * g.ncell + c] = acc0;
    if (l < 2) a.cg[(16 + l) * g.ncell + c] = acc1;
  }
}

// fold^T of every periodic dim in one launch: the ghost plane n_d + 1 gains
// the adjoint of plane 1 (physical range); each ghost point has one source
template <typename T>
struct FoldAdjArgs {
  Geo g;
  T *J[3];
  int per[3];
};

__global__ void __launch_bounds__(kBlock) k_fold_adj_all(FoldAdjArgs<double> a) {
  const Geo &g = a.g;
  const long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t < n1 * n2) {
    i = 1;
    j = 1 + (int)(t / n2);
    k = 1 + (int)(t % n2);
  } else if ((t -= n1 * n2) < (n0 - 1) * n2) {
    i = 2 + (int)(t / n2);
    j = 1;
    k = 1 + (int)(t % n2);
  } else if ((t -= (n0 - 1) * n2) < (n0 - 1) * (n1 - 1)) {
    i = 2 + (int)(t / (n1 - 1));
    j = 2 + (int)(t % (n1 - 1));
    k = 1;
  } else {
    return;
  }
  const long long p = gidx(g, i, j, k);
  const int at[3] = {i, j, k};
  for (int d = 0; d < 3; d++) {
    if (!a.per[d] || at[d] != 1) continue;
    int gh[3] = {i, j, k};
    gh[d] = g.n[d] + 1;
    const long long q = gidx(g, gh[0], gh[1], gh[2]);
    for (int c = 0; c < 3; c++)
      if (c != d) a.J[c][q] += a.J[c][p];
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
// Every buffer of a gradient call comes from one device block that the
// context keeps between calls; a pass without a base sizes it.
struct Pool {
  char *base = nullptr;
  size_t off = 0;
  template <typename T>
  void take(T **p, size_t n) {
    const size_t bytes = (std::max<size_t>(n, 1) * sizeof(T) + 255) & ~(size_t)255;
    *p = base ? (T *)(base + off) : nullptr;
    off += bytes;
  }
};

// One taped state: the particles as the push of a step reads them (the
// state right after the previous gather), the fields at that time, the cell
// ranges of the particle order, and the sort that follows.
struct Snap {
  double *x[3] = {}, *v[3] = {}, *gam = nullptr, *fld[6] = {};
  int32_t *sp = nullptr;
  int64_t *id = nullptr;     // checkpoints and the final state only
  double *J[3] = {}, *rho = nullptr;  // the final state only
  double *E[3] = {}, *A[2][3] = {};  // A[0] = Bfix, A[1] = Btfix
  int32_t *off = nullptr;    // cell ranges of this order (ncell + 1)
  uint32_t *inv = nullptr;   // the following sort: old index i -> new position inv[i]
  uint8_t *flips = nullptr;  // the push's reflections (bit d: dim d)
  bool b_is_a0 = true;       // member B is the Bfix allocation
  double t = 0.0;
};

void snap_alloc(Pool &P, const emo_ctx *c, Snap &s, bool ids, bool perm, bool currents = false) {
  const size_t n = (size_t)std::max<int64_t>(c->np, 1);
  const size_t nt = (size_t)c->geo.ntot;
  for (int d = 0; d < 3; d++) {
    P.take(&s.x[d], n);
    P.take(&s.v[d], n);
    P.take(&s.E[d], nt);
    P.take(&s.A[0][d], nt);
    P.take(&s.A[1][d], nt);
  }
  P.take(&s.gam, n);
  for (int k = 0; k < 6 && perm; k++) P.take(&s.fld[k], n);
  P.take(&s.sp, n);
  if (ids) P.take(&s.id, n);
  P.take(&s.off, (size_t)c->geo.ncell + 1);
  if (perm) P.take(&s.inv, n);
  if (perm) P.take(&s.flips, n);
  for (int d = 0; d < 3 && currents; d++) P.take(&s.J[d], nt);
  if (currents) P.take(&s.rho, nt);
}

// Collects device copies and issues them as k_copy_many launches.
struct CopyBatch {
  emo_ctx *c;
  CopyList l;
  explicit CopyBatch(emo_ctx *ctx) : c(ctx) {
    l.n = 0;
    l.start[0] = 0;
  }
  void add(void *dst, const void *src, size_t bytes) {
    if (!bytes) return;
    if (l.n == kMaxCopies) flush();
    l.src[l.n] = (const uint32_t *)src;
    l.dst[l.n] = (uint32_t *)dst;
    l.words[l.n] = (long long)(bytes / 4);
    l.start[l.n + 1] = l.start[l.n] + (l.words[l.n] + 3) / 4;
    l.n++;
  }
  void flush() {
    if (!l.n) return;
    hipLaunchKernelGGL(k_copy_many, dim3(grid_for(l.start[l.n])), dim3(kBlock), 0, c->stream, l);
    l.n = 0;
  }
};

// ctx -> snapshot (particles, fields, ranges)
int snap_take(emo_ctx *c, Snap &s, double t) {
  const size_t n8 = (size_t)c->np * 8, n4 = (size_t)c->np * 4, nt8 = (size_t)c->geo.ntot * 8;
  CopyBatch cb(c);
  for (int d = 0; d < 3; d++) {
    cb.add(s.x[d], c->p.x[d], n8);
    cb.add(s.v[d], c->p.v[d], n8);
    cb.add(s.E[d], c->E[d], nt8);
    cb.add(s.A[0][d], c->Bfix[d], nt8);
    cb.add(s.A[1][d], c->Btfix[d], nt8);
  }
  cb.add(s.gam, c->p.gam, n8);
  cb.add(s.sp, c->p.sp, n4);
  if (s.id) cb.add(s.id, c->p.id, (size_t)c->np * 8);
  for (int d = 0; d < 3 && s.rho; d++) cb.add(s.J[d], c->J[d], nt8);
  if (s.rho) cb.add(s.rho, c->rho, nt8);
  cb.add(s.off, c->srt.off, (size_t)(c->geo.ncell + 1) * 4);
  cb.flush();
  s.b_is_a0 = c->B[0] == c->Bfix[0];
  s.t = t;
  return EMO_OK;
}

// checkpoint -> ctx (the particles' gathered fields are not restored: the
// fused step gathers them from the restored grid)
int snap_restore(emo_ctx *c, const Snap &s) {
  const size_t n8 = (size_t)c->np * 8, n4 = (size_t)c->np * 4, nt8 = (size_t)c->geo.ntot * 8;
  CopyBatch cb(c);
  for (int d = 0; d < 3; d++) {
    cb.add(c->p.x[d], s.x[d], n8);
    cb.add(c->p.v[d], s.v[d], n8);
    cb.add(c->E[d], s.E[d], nt8);
    cb.add(c->Bfix[d], s.A[0][d], nt8);
    cb.add(c->Btfix[d], s.A[1][d], nt8);
  }
  cb.add(c->p.gam, s.gam, n8);
  cb.add(c->p.sp, s.sp, n4);
  if (s.id) cb.add(c->p.id, s.id, (size_t)c->np * 8);
  for (int d = 0; d < 3 && s.rho; d++) cb.add(c->J[d], s.J[d], nt8);
  if (s.rho) cb.add(c->rho, s.rho, nt8);
  cb.add(c->srt.off, s.off, (size_t)(c->geo.ncell + 1) * 4);
  cb.flush();
  for (int q = 0; q < 3; q++) {
    c->B[q] = s.b_is_a0 ? c->Bfix[q] : c->Btfix[q];
    c->Bt[q] = s.b_is_a0 ? c->Btfix[q] : c->Bfix[q];
  }
  c->sorted = true;
  c->fld_state = kFldDeferred;  // the particles' fields: the gather of this grid
  return EMO_OK;
}

// One forward step of emo_step.  With a tape entry, the state the push reads
// (with the fields it gathers) goes to the entry before the sort, and the
// sort's inverse permutation after it.
int forward_step(emo_ctx *c, double t, double dt, Snap *tape) {
  int rc;
  if (c->np > 0) {
    fused_push_f64(c, dt, tape ? tape->fld : nullptr, tape ? tape->flips : nullptr);
    if (tape) {
      rc = snap_take(c, *tape, t);
      if (rc) return rc;
    }
    rc = fused_sort_f64(c, nullptr, tape ? tape->inv : nullptr);
    if (rc) return rc;
  } else {
    if (tape) {
      rc = snap_take(c, *tape, t);
      if (rc) return rc;
    }
    rc = sort_f64(c, 0);
    if (rc) return rc;
  }
  launch_deposit_f64(c, 1);
  launch_fold_f64(c);
  launch_evolve_f64(c, dt);
  FieldSet<double> v = f64_view(c);
  refresh_fields<double>(c, v, t, dt);
  c->fld_state = kFldDeferred;
  return EMO_OK;
}

// the LightBC planes of one side, in the forward fill order
struct FillPlane {
  int side, i, pass;  // pass 0: E, 1: B
  int comp;           // field component
  int dim, idx;
  double off[3];
};

int edge_plane_index(const Geo &g, int fid, int dim, bool upper) {
  if (!upper) return 1;
  int n_ = g.nc[dim];
  int comp = fid % 3;
  bool is_e = fid < 3;
  if (is_e) return comp == dim ? n_ - 1 - 1 : n_ - 1;
  return comp == dim ? n_ - 1 : n_ - 1 - 1;
}

std::vector<FillPlane> fill_planes(const emo_ctx *c) {
  static const int EpolT[3][3] = {{1, 2, 0}, {2, 0, 1}, {0, 1, 2}};
  static const int BpolT[3][3] = {{2, 1, 0}, {0, 2, 1}, {1, 0, 2}};
  std::vector<FillPlane> out;
  for (int side = 0; side < 6; side++) {
    if (c->cfg.fields_bound[side] == EMO_BC_PERIODIC) continue;
    int dim = side / 2;
    bool upper = side & 1;
    for (int i = 0; i < 3; i++) {
      for (int pass = 0; pass < 2; pass++) {
        FillPlane f;
        f.side = side;
        f.i = i;
        f.pass = pass;
        int fid = pass == 0 ? EpolT[dim][i] : 3 + BpolT[dim][i];
        f.comp = fid % 3;
        f.dim = dim;
        f.idx = edge_plane_index(c->geo, fid, dim, upper);
        for (int q = 0; q < 3; q++) {
          if (fid < 3) f.off[q] = q == f.comp ? 0.5 * c->geo.d[q] : 0.0;
          else f.off[q] = q == f.comp ? 0.0 : 0.5 * c->geo.d[q];
        }
        out.push_back(f);
      }
    }
  }
  return out;
}

// d(pulse value)/d(amplitude), d/d(phase in radians) of one LightBC plane at
// time t, for every wave (GaussianPulses, src/utils/evaluateGaussianPulse.cpp:7-30)
void pulse_derivs(const emo_ctx *c, const FillPlane &f, double t, double dt, std::vector<double> &damp,
                  std::vector<double> &dpha) {
  static const int EpolT[3][3] = {{1, 2, 0}, {2, 0, 1}, {0, 1, 2}};
  const emo_config &cf = c->cfg;
  const int dim = f.side / 2;
  const int sign = (f.side & 1) ? 1 : -1;
  const int sidev = sign * (dim + 1);
  const int eid = EpolT[dim][f.i];
  damp.assign(cf.nwaves, 0.0);
  dpha.assign(cf.nwaves, 0.0);
  if (eid == dim) return;  // the normal component carries no wave
  const double bPhase = (c->geo.d[dim] + sign * dt) / 2;
  const double shift = f.pass == 0 ? 0.0 : bPhase;
  const double scale = f.pass == 0 ? 1.0 : (double)(sign * (int)std::pow(-1, f.i + 1));
  for (int w = 0; w < cf.nwaves; w++) {
    if (cf.inSide[w] != sidev || cf.inPolE[w] - 1 != eid) continue;
    const double th = cf.omegas[w] * (t + shift) + cf.phases[w] * kPi / 180.0;
    const double G = std::exp(-std::pow((t - cf.delays[w]) * cf.invWidths[w], 2));
    damp[w] = scale * std::cos(th) * G;
    dpha[w] = -scale * cf.peakamps[w] * std::sin(th) * G;
  }
}

struct AdjFields {
  double *Eb[3] = {}, *Ab[2][3] = {}, *Jb[3] = {};
  double *cg = nullptr;
};

void launch_zero(hipStream_t st, double *p, long long n) {
  hipLaunchKernelGGL(k_zero, dim3(grid_for(n)), dim3(kBlock), 0, st, p, n);
}

// gather^T, stage 2, for a gather that ran on state s: the per-cell sums in
// af.cg go to the points (B roles of that state)
void gather_adj_grid_roles(emo_ctx *c, hipStream_t st, bool b_is_a0, AdjFields &af) {
  GatherGridArgs b;
  b.g = c->geo;
  b.cg = af.cg;
  const int mb = b_is_a0 ? 0 : 1;
  for (int d = 0; d < 3; d++) {
    b.Eb[d] = af.Eb[d];
    b.Bb[d] = af.Ab[mb][d];
    b.Btb[d] = af.Ab[1 - mb][d];
  }
  TimedLaunch t(c, "adj_gather_grid", st);
  hipLaunchKernelGGL(k_gather_adj_grid, dim3(grid_for(c->geo.ntot)), dim3(kBlock), 0, st, b);
}

void gather_adj_grid(emo_ctx *c, hipStream_t st, const Snap &s, AdjFields &af) {
  gather_adj_grid_roles(c, st, s.b_is_a0, af);
}

// refresh^T, Yee^T and fold^T of one step on stream `as` (the field part of
// the adjoint step, common to both precisions' adjoints: these operations
// are linear and read no forward state); dsums_row receives the LightBC
// plane sums of the step, pre_b_is_a0 the member-B role at the step's start.
void fields_step_adj(emo_ctx *c, hipStream_t as, AdjFields &af, const std::vector<FillPlane> &planes,
                     double *dsums_row, double dt, bool pre_b_is_a0) {
  const Geo &g = c->geo;
  const int nplanes = (int)planes.size();
  // refresh^T, the forward's launches in reverse: a LightBC side's six
  // fills as one launch, a run of periodic sides as one gather pass
  const int mnew = pre_b_is_a0 ? 1 : 0;  // the array the B update writes
  int pi = nplanes - 1;
  for (int side = 5; side >= 0;) {
    if (c->cfg.fields_bound[side] == EMO_BC_PERIODIC) {
      SidesAdjArgs sa;
      sa.g = g;
      for (int q = 0; q < 3; q++) {
        sa.f[q] = af.Eb[q];
        sa.f[3 + q] = af.Ab[0][q];  // Q23: the refresh writes B through Bfix
      }
      sa.nsides = 0;
      for (; side >= 0 && c->cfg.fields_bound[side] == EMO_BC_PERIODIC; side--) {
        const int d = side / 2, right = side & 1;
        sa.d[sa.nsides] = d;
        sa.src[sa.nsides] = right ? 1 : g.n[d];
        sa.dst[sa.nsides] = right ? g.n[d] + 1 : 0;
        sa.nsides++;
      }
      hipLaunchKernelGGL(k_copy_sides_adj, dim3(grid_for((long long)g.n[0] * g.n[1] * g.n[2])), dim3(kBlock), 0,
                         as, sa);
    } else {
      FillAdjMulti fm;
      int nb = 0;
      for (; pi >= 0 && planes[pi].side == side; pi--) {
        const FillPlane &f = planes[pi];
        FillAdjArgs &fa = fm.p[nb++];
        fa.g = g;
        fa.f = f.pass == 0 ? af.Eb[f.comp] : af.Ab[mnew][f.comp];
        fa.dim = f.dim;
        fa.idx = f.idx;
        fa.prof = c->wprof[f.dim][f.pass == 0 ? f.comp : 3 + f.comp];
        fa.out = dsums_row + pi;
      }
      if (nb) hipLaunchKernelGGL(k_plane_fill_adj_multi, dim3(nb), dim3(kBlock), 0, as, fm);
      side--;
    }
  }
  // Yee^T
  YeeAdjArgs ya;
  ya.g = g;
  ya.dt = dt;
  for (int q = 0; q < 3; q++) {
    ya.Eb[q] = af.Eb[q];
    ya.Bn[q] = af.Ab[mnew][q];
    ya.Bo[q] = af.Ab[1 - mnew][q];
    ya.Jb[q] = af.Jb[q];
  }
  {
    TimedLaunch tl(c, "adj_yee", as);
    hipLaunchKernelGGL(k_yee_b_adj, dim3(grid_for(g.ntot)), dim3(kBlock), 0, as, ya);
    hipLaunchKernelGGL(k_yee_move_adj, dim3(grid_for(g.ntot)), dim3(kBlock), 0, as, ya);
    hipLaunchKernelGGL(k_yee_e_adj, dim3(grid_for(g.ntot)), dim3(kBlock), 0, as, ya);
  }
  // fold^T: each ghost plane receives its first plane's adjoint
  {
    FoldAdjArgs<double> fa;
    fa.g = g;
    int any = 0;
    for (int d = 0; d < 3; d++) {
      fa.J[d] = af.Jb[d];
      fa.per[d] = c->cfg.fields_bound[2 * d] == EMO_BC_PERIODIC;
      any |= fa.per[d];
    }
    if (any) {
      const long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
      const long long pts = n1 * n2 + (n0 - 1) * n2 + (n0 - 1) * (n1 - 1);
      hipLaunchKernelGGL(k_fold_adj_all, dim3(grid_for(pts)), dim3(kBlock), 0, as, fa);
    }
  }
}

// Everything one gradient call needs besides the context.

// Tape sets of the K-step adjoint: two let the backward of one segment run
// beside the recomputation of the next, and each set holds one segment the
// forward pass tapes on its way (so the backward does not recompute it).
// Beyond two, as many as fit in free device memory with 1 GiB to spare, at
// most kMaxTapeSets.
constexpr int kMaxTapeSets = 6;
template <class Layout>
int more_tape_sets(emo_ctx *c, int nseg, int nsets, Layout &&layout_of) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return nsets;
  for (int ns = std::min(nseg, kMaxTapeSets); ns > nsets; ns--) {
    Pool d;
    layout_of(d, ns);
    if (d.off + ((size_t)1 << 30) <= fr + c->adj_pool_bytes) return ns;
  }
  return nsets;
}

struct AdjRun {
  emo_ctx *c;
  hipStream_t as;  // the backward sweep's stream; the context's stream recomputes
  double dt;
  int nsteps, K, nseg;
  std::vector<Snap> ck, tape[kMaxTapeSets];
  Snap fin;
  AdjFields af;
  double *dsums = nullptr;  // nsteps x nplanes
  PartAdj pb, pb2;
  std::vector<FillPlane> planes;
};

// Segment seg re-run from its checkpoint on the context's stream, taping
// into tape set `set`.
int recompute_segment(AdjRun &R, int seg, int set) {
  emo_ctx *c = R.c;
  const int s0 = seg * R.K, s1 = std::min(R.nsteps, s0 + R.K);
  int rc = snap_restore(c, R.ck[seg]);
  double tt = R.ck[seg].t;
  for (int s = s0; s < s1 && !rc; s++) {
    rc = forward_step(c, tt, R.dt, &R.tape[set][s - s0]);
    tt += R.dt;
  }
  if (!rc) rc = snap_take(c, R.tape[set][s1 - s0], tt);
  return rc;
}

// The adjoint of segment seg's steps, newest first, on the adjoint stream.
void backward_segment(AdjRun &R, int seg, int set) {
  emo_ctx *c = R.c;
  const Geo &g = c->geo;
  hipStream_t as = R.as;
  AdjFields &af = R.af;
  const int s0 = seg * R.K, s1 = std::min(R.nsteps, s0 + R.K);
  const int nplanes = (int)R.planes.size();
  for (int s = s1 - 1; s >= s0; s--) {
    const Snap &pre = R.tape[set][s - s0];      // the state the push of step s reads
    const Snap &post = R.tape[set][s + 1 - s0];  // the state after step s's sort
    fields_step_adj(c, as, af, R.planes, R.dsums + (size_t)s * nplanes, R.dt, pre.b_is_a0);
    // deposit^T + sort^T + BCs^T + push^T, then gather^T of the gather that
    // fed this push
    CellStepAdjArgs ca;
    ca.g = g;
    for (int d = 0; d < 3; d++) {
      ca.post.x[d] = post.x[d];
      ca.post.v[d] = post.v[d];
      ca.pre.x[d] = pre.x[d];
      ca.pre.v[d] = pre.v[d];
      ca.E[d] = pre.E[d];
      ca.B[d] = pre.b_is_a0 ? pre.A[0][d] : pre.A[1][d];
      ca.Bt[d] = pre.b_is_a0 ? pre.A[1][d] : pre.A[0][d];
      ca.bx[d] = R.pb.x[d];
      ca.bv[d] = R.pb.v[d];
      ca.ox[d] = R.pb2.x[d];
      ca.ov[d] = R.pb2.v[d];
      ca.Jb[d] = af.Jb[d];
    }
    ca.post.sp = post.sp;
    ca.pre.sp = pre.sp;
    ca.pre.gam = pre.gam;
    for (int k = 0; k < 6; k++) ca.pre.fld[k] = pre.fld[k];
    ca.off = pre.off;
    ca.inv = pre.inv;
    ca.flips = pre.flips;
    ca.bg = R.pb.gam;
    ca.og = R.pb2.gam;
    ca.cg = af.cg;
    ca.sp = c->species;
    ca.dt = R.dt;
    ca.rel = c->cfg.relativity;
    {
      TimedLaunch tl(c, "adj_cells", as);
      hipLaunchKernelGGL(k_cell_step_adj, dim3((unsigned)((g.ncell + kGCells - 1) / kGCells)), dim3(kBlock), 0,
                         as, ca);
    }
    gather_adj_grid(c, as, pre, af);
    for (int d = 0; d < 3; d++) {
      std::swap(R.pb.x[d], R.pb2.x[d]);
      std::swap(R.pb.v[d], R.pb2.v[d]);
    }
    std::swap(R.pb.gam, R.pb2.gam);
  }
}


// ===========================================================================
// The adjoint of the FP32 chosen form (chosen.hip).  The field part of an
// adjoint step is the FP64 one above (fields_step_adj): those operations
// are linear and read no forward state, so the field adjoints are carried
// in FP64.  The particle part is the transpose of the chosen form's push,
// BCs, sort and deposit in FP32 arithmetic (push_deposit_vjp,
// chosen_math.hpp), per cell of the pre-push order, with the particle's
// forward recomputed from the taped pre-push state.
// ===========================================================================

// Particle adjoints of the FP32 form: the offsets' and u's adjoints, SoA.
struct PartAdj32 {
  float *f[3], *u[3];
};

struct CellStepAdj32Args {
  PushK k;                      // geometry, species q_p, dt / dx, BCs
  const int4 *pa, *pb;          // pre-push state of step s, cell-sorted
  const float *g0;              // the stored gamma of the first push, or null
  const int32_t *base, *cnt;    // cell c holds cnt[c] particles from slot base[c]
  const uint32_t *fwd;          // pre-push slot -> post-step slot
  const float *E[3], *B[3], *Bt[3];  // the fields the push gathered from
  const float *Jb[3];           // current adjoint (after fold^T), rounded to FP32
  PartAdj32 in, out;            // adjoints: post-sort order in, pre-push order out
  float *g0b;                   // adjoint of the stored gamma (first step), or null
  double *cg;                   // 18 per cell, component-major
  long long ncell;
  uint32_t nslots;
  // BYID: the adjoints by particle id, 6 per id {f[3], u[3]}, updated in
  // place (no fwd map); g0b by id
  float *adj;
  long long np;
};

template <bool REL, bool PER, bool BYID>
__global__ void __launch_bounds__(kBlock) k_cell_step_adj32(CellStepAdj32Args a) {
  __shared__ float qp_s[EMO_NSPEC], jc_s[EMO_NSPEC];
  if (threadIdx.x < EMO_NSPEC) {
    qp_s[threadIdx.x] = a.k.s.qp[threadIdx.x];
    jc_s[threadIdx.x] = a.k.s.test[threadIdx.x] ? 0.f : a.k.s.jc[threadIdx.x];
  }
  __syncthreads();
  const int l = threadIdx.x & 15;
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  double acc[18];
#pragma unroll
  for (int q = 0; q < 18; q++) acc[q] = 0.0;
  if (c < a.ncell) {
    const int b = a.base[c], e = b + a.cnt[c];
    if (b < e) {
      // the cell's gather vector (the cell is the same for all its particles)
      const Geo &g = a.k.g;
      int ix, iy, iz;
      decode_fast(a.k, (int)c, ix, iy, iz);
      const long long o = gidx(g, ix, iy, iz), X = g.sx, Y = g.sy;
      CellVec V;
      V.Ex = make_float4(a.E[0][o], a.E[0][o + Y], a.E[0][o + Y + 1], a.E[0][o + 1]);
      V.Ey = make_float4(a.E[1][o], a.E[1][o + 1], a.E[1][o + X + 1], a.E[1][o + X]);
      V.Ez = make_float4(a.E[2][o], a.E[2][o + X], a.E[2][o + X + Y], a.E[2][o + Y]);
      V.B = make_float4(0.5f * (a.Bt[0][o] + a.B[0][o]), 0.5f * (a.Bt[0][o + X] + a.B[0][o + X]),
                        0.5f * (a.Bt[1][o] + a.B[1][o]), 0.5f * (a.Bt[1][o + Y] + a.B[1][o + Y]));
      V.Bz = make_float4(0.5f * (a.Bt[2][o] + a.B[2][o]), 0.5f * (a.Bt[2][o + 1] + a.B[2][o + 1]), 0.f, 0.f);
      // software pipeline: the next particle's records and slot are in
      // flight during the current one (loads past the cell's end read its
      // last particle, so that they are unconditional)
      const int last = e - 1;
      int i = b + l;
      int4 A = a.pa[min(i, last)], Bv = a.pb[min(i, last)];
      // (a slot the step did not map -- only after a flagged error -- reads
      // slot 0; by id, an id outside [0, np) reads and writes nothing)
      uint32_t j = BYID ? (uint32_t)Bv.w : a.fwd[min(i, last)];
      if (!BYID) j = j < a.nslots ? j : 0u;
      for (; i < e; i += 16) {
        const int inx = min(i + 16, last);
        const int4 An = a.pa[inx], Bn = a.pb[inx];
        uint32_t jn = BYID ? (uint32_t)Bn.w : a.fwd[inx];
        if (!BYID) jn = jn < a.nslots ? jn : 0u;
        const bool ok = !BYID || (long long)j < a.np;
        const int sp = (int)((uint32_t)A.x >> kSpeciesShift);
        const float fin[3] = {as_f(A.y), as_f(A.z), as_f(A.w)};
        const float uin[3] = {as_f(Bv.x), as_f(Bv.y), as_f(Bv.z)};
        float fbo[3] = {0.f, 0.f, 0.f}, ubo[3] = {0.f, 0.f, 0.f};
        if (BYID) {
          if (ok) {
            const float2 *q = reinterpret_cast<const float2 *>(a.adj + 6 * (long long)j);
            const float2 q0 = q[0], q1 = q[1], q2 = q[2];
            fbo[0] = q0.x;
            fbo[1] = q0.y;
            fbo[2] = q1.x;
            ubo[0] = q1.y;
            ubo[1] = q2.x;
            ubo[2] = q2.y;
          }
        } else {
          for (int d = 0; d < 3; d++) {
            fbo[d] = a.in.f[d][j];
            ubo[d] = a.in.u[d][j];
          }
        }
        float fb[3], ub[3], g0b, cvb[18];
        push_deposit_vjp<REL, PER>(a.k, V, A.x, fin, uin, qp_s[sp], a.g0 != nullptr, a.g0 ? a.g0[i] : 0.f,
                                   jc_s[sp], a.Jb, fbo, ubo, fb, ub, g0b, cvb);
        if (BYID) {
          if (ok) {
            float2 *q = reinterpret_cast<float2 *>(a.adj + 6 * (long long)j);
            q[0] = make_float2(fb[0], fb[1]);
            q[1] = make_float2(fb[2], ub[0]);
            q[2] = make_float2(ub[1], ub[2]);
            if (a.g0b) a.g0b[j] = g0b;
          }
        } else {
          for (int d = 0; d < 3; d++) {
            a.out.f[d][i] = fb[d];
            a.out.u[d][i] = ub[d];
          }
          if (a.g0b) a.g0b[i] = g0b;
        }
#pragma unroll
        for (int q = 0; q < 18; q++) acc[q] += (double)cvb[q];
        A = An;
        Bv = Bn;
        j = jn;
      }
    }
  }
  // fixed butterfly over the cell's 16 lanes: deterministic
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 18; q++) acc[q] += __shfl_xor(acc[q], off, 64);
  }
  if (l == 0 && c < a.ncell) {
#pragma unroll
    for (int q = 0; q < 18; q++) a.cg[q * a.ncell + c] = acc[q];
  }
}

// One FP32 snapshot: the particle records (every slot), the stored gamma
// (first push), the fields, the slot layout; on the tape also the step's
// map source slot -> new slot.
struct Snap32 {
  int4 *pa = nullptr, *pb = nullptr;
  float *g0 = nullptr;
  bool has_g0 = false;
  float *E[3] = {}, *A[2][3] = {};  // A[0] = Bfix, A[1] = Btfix
  float *J[3] = {};                 // the final state only
  int32_t *base = nullptr, *cnt = nullptr;
  uint32_t *fwd = nullptr;          // tape: source slot -> new slot
  bool b_is_a0 = true;
  double t = 0.0;
};

void snap32_alloc(Pool &P, emo_ctx *c, Snap32 &s, bool particles, bool g0, bool tape, bool currents) {
  const size_t n = (size_t)fp32_view(c).nslots;
  const size_t nt = (size_t)c->geo.ntot;
  if (particles) {
    P.take(&s.pa, n);
    P.take(&s.pb, n);
  }
  if (g0) P.take(&s.g0, n);
  for (int d = 0; d < 3; d++) {
    P.take(&s.E[d], nt);
    P.take(&s.A[0][d], nt);
    P.take(&s.A[1][d], nt);
    if (currents) P.take(&s.J[d], nt);
  }
  P.take(&s.base, (size_t)c->geo.ncell + 1);
  P.take(&s.cnt, (size_t)c->geo.ncell + 1);
  if (tape) P.take(&s.fwd, n);
}

// the fields and layout (and, with particles, the records) of the context
void snap32_take(emo_ctx *c, Snap32 &s, double t, bool particles) {
  Fp32View v = fp32_view(c);
  const size_t np = (size_t)v.nslots, nt4 = (size_t)c->geo.ntot * 4;
  CopyBatch cb(c);
  if (particles && s.pa) {
    cb.add(s.pa, *v.pa, np * 16);
    cb.add(s.pb, *v.pb, np * 16);
  }
  s.has_g0 = s.g0 && *v.g0_active;
  if (s.has_g0) cb.add(s.g0, *v.g0, np * 4);
  for (int d = 0; d < 3; d++) {
    cb.add(s.E[d], v.fs->E[d], nt4);
    cb.add(s.A[0][d], v.fs->Bfix[d], nt4);
    cb.add(s.A[1][d], v.Btfix[d], nt4);
    if (s.J[d]) cb.add(s.J[d], v.fs->J[d], nt4);
  }
  cb.add(s.base, *v.base, (size_t)(c->geo.ncell + 1) * 4);
  cb.add(s.cnt, *v.cnt, (size_t)(c->geo.ncell + 1) * 4);
  cb.flush();
  s.b_is_a0 = v.fs->B[0] == v.fs->Bfix[0];
  s.t = t;
}

// snapshot -> context: the records into pa/pb (which the caller may point
// at tape buffers), fields, ranges, B roles; the stored gamma by reference
void snap32_restore(emo_ctx *c, const Snap32 &s, int4 *pa, int4 *pb) {
  Fp32View v = fp32_view(c);
  const size_t np = (size_t)v.nslots, nt4 = (size_t)c->geo.ntot * 4;
  CopyBatch cb(c);
  cb.add(pa, s.pa, np * 16);
  cb.add(pb, s.pb, np * 16);
  for (int d = 0; d < 3; d++) {
    cb.add(v.fs->E[d], s.E[d], nt4);
    cb.add(v.fs->Bfix[d], s.A[0][d], nt4);
    cb.add(v.Btfix[d], s.A[1][d], nt4);
    if (s.J[d]) cb.add(v.fs->J[d], s.J[d], nt4);
  }
  cb.add(*v.base, s.base, (size_t)(c->geo.ncell + 1) * 4);
  cb.add(*v.cnt, s.cnt, (size_t)(c->geo.ncell + 1) * 4);
  cb.flush();
  *v.pa = pa;
  *v.pb = pb;
  for (int q = 0; q < 3; q++) {
    v.fs->B[q] = s.b_is_a0 ? v.fs->Bfix[q] : v.Btfix[q];
    v.fs->Bt[q] = s.b_is_a0 ? v.Btfix[q] : v.fs->Bfix[q];
  }
  *v.g0_active = s.has_g0;
  *v.g0 = s.has_g0 ? s.g0 : nullptr;
  *v.g0_owned = false;
}

// FP64 -> FP32 copy of the three current adjoints: the particle VJP reads
// them in FP32 (it rounded each value it read to FP32 before)
__global__ void k_round_f32(const double *a, const double *b, const double *c, float *x, float *y, float *z,
                            long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    x[i] = (float)a[i];
    y[i] = (float)b[i];
    z[i] = (float)c[i];
  }
}

struct AdjRun32 {
  emo_ctx *c;
  hipStream_t as;
  double dt;
  int nsteps, K, nseg;
  std::vector<Snap32> ck, tape[kMaxTapeSets];
  Snap32 fin;
  AdjFields af;
  double *dsums = nullptr;
  PartAdj32 pa, pa2;
  float *g0b = nullptr;
  float *Jb32[3] = {};
  std::vector<FillPlane> planes;
  int4 *own[4] = {};  // the context's own record buffers (pa, pb, pa_alt, pb_alt)
  int4 *stage_a = nullptr, *stage_b = nullptr;  // the recomputed movers' new records
  int32_t *mvsrc = nullptr;                       // and their source slots
};

void restore_own_buffers(AdjRun32 &R) {
  if (!R.own[0]) return;  // never replaced
  Fp32View v = fp32_view(R.c);
  *v.pa = R.own[0];
  *v.pb = R.own[1];
  *v.pa_alt = R.own[2];
  *v.pb_alt = R.own[3];
}

// Segment seg re-run from its checkpoint on the context's stream.  Each
// step writes its new records straight into the next tape entry's buffers
// (the state's alternate buffers point there), so the tape holds every
// step's pre-push records without copies.
int recompute32(AdjRun32 &R, int seg, int set) {
  emo_ctx *c = R.c;
  Fp32View v = fp32_view(c);
  const int s0 = seg * R.K, s1 = std::min(R.nsteps, s0 + R.K);
  std::vector<Snap32> &tp = R.tape[set];
  snap32_restore(c, R.ck[seg], tp[0].pa, tp[0].pb);
  double t = R.ck[seg].t;
  int rc = 0;
  for (int st = s0; st < s1 && !rc; st++) {
    snap32_take(c, tp[st - s0], t, false);
    *v.pa_alt = tp[st - s0 + 1].pa;
    *v.pb_alt = tp[st - s0 + 1].pb;
    const Fp32Tape tape{tp[st - s0].fwd, R.stage_a, R.stage_b, R.mvsrc};
    rc = fp32_step_one(c, t, R.dt, &tape);
    t += R.dt;
  }
  restore_own_buffers(R);
  *v.g0 = nullptr;
  *v.g0_active = false;
  *v.g0_owned = true;
  return rc;
}

void launch_cell_step_adj32(emo_ctx *c, hipStream_t as, const CellStepAdj32Args &ca, bool all_per, unsigned blocks,
                            bool byid) {
  const bool rel = c->cfg.relativity != 0;
#define EMO_ADJ32(R, P, I) hipLaunchKernelGGL((k_cell_step_adj32<R, P, I>), dim3(blocks), dim3(kBlock), 0, as, ca)
  if (byid) {
    if (rel && all_per) EMO_ADJ32(true, true, true);
    else if (rel) EMO_ADJ32(true, false, true);
    else if (all_per) EMO_ADJ32(false, true, true);
    else EMO_ADJ32(false, false, true);
  } else {
    if (rel && all_per) EMO_ADJ32(true, true, false);
    else if (rel) EMO_ADJ32(true, false, false);
    else if (all_per) EMO_ADJ32(false, true, false);
    else EMO_ADJ32(false, false, false);
  }
#undef EMO_ADJ32
}

void backward32(AdjRun32 &R, int seg, int set) {
  emo_ctx *c = R.c;
  const Geo &g = c->geo;
  hipStream_t as = R.as;
  const int s0 = seg * R.K, s1 = std::min(R.nsteps, s0 + R.K);
  const int nplanes = (int)R.planes.size();
  bool all_per = true;
  for (int q = 0; q < 6; q++) all_per = all_per && c->cfg.parts_bound[q] == EMO_BC_PERIODIC;
  const PushK k = fp32_push_consts(c, R.dt);
  for (int st = s1 - 1; st >= s0; st--) {
    const Snap32 &pre = R.tape[set][st - s0];
    fields_step_adj(c, as, R.af, R.planes, R.dsums + (size_t)st * nplanes, R.dt, pre.b_is_a0);
    hipLaunchKernelGGL(k_round_f32, dim3(grid_for(g.ntot)), dim3(kBlock), 0, as, (const double *)R.af.Jb[0],
                       (const double *)R.af.Jb[1], (const double *)R.af.Jb[2], R.Jb32[0], R.Jb32[1], R.Jb32[2],
                       (long long)g.ntot);
    CellStepAdj32Args ca;
    ca.k = k;
    ca.pa = pre.pa;
    ca.pb = pre.pb;
    ca.g0 = (st == 0 && R.ck[0].has_g0) ? R.ck[0].g0 : nullptr;
    ca.base = pre.base;
    ca.cnt = pre.cnt;
    ca.fwd = pre.fwd;
    for (int d = 0; d < 3; d++) {
      ca.E[d] = pre.E[d];
      ca.B[d] = pre.b_is_a0 ? pre.A[0][d] : pre.A[1][d];
      ca.Bt[d] = pre.b_is_a0 ? pre.A[1][d] : pre.A[0][d];
      ca.Jb[d] = R.Jb32[d];
    }
    ca.in = R.pa;
    ca.out = R.pa2;
    ca.g0b = ca.g0 ? R.g0b : nullptr;
    ca.cg = R.af.cg;
    ca.ncell = g.ncell;
    ca.nslots = (uint32_t)fp32_view(c).nslots;
    ca.adj = nullptr;
    ca.np = 0;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    {
      TimedLaunch tl(c, "adj_cells_f32", as);
      launch_cell_step_adj32(c, as, ca, all_per, blocks, false);
    }
    gather_adj_grid_roles(c, as, pre.b_is_a0, R.af);
    std::swap(R.pa, R.pa2);
  }
}

__global__ void k_zero_f(float *p, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}


// ---------------------------------------------------------------------------
// Per-species sums of the initial-state gradients (emo_adjoint_stats): the
// sums of what the by-id outputs hold, reduced on the device.
// ---------------------------------------------------------------------------
struct SumsArgs32 {
  const int4 *ra, *rb;        // the initial state's records
  const int32_t *rbase;       // cell c's records from ra[rbase[c]] (slots, or dense offsets)
  const int32_t *sbase;       // cell c's slots (the stored gamma's layout)
  const int32_t *cnt;
  const float *g0;            // the stored gamma by slot, or null
  const float *adj;           // by id: 6 per id {f[3], u[3]}; else null
  PartAdj32 pa;               // by slot (adj null)
  const float *g0b;           // by id (adj) or slot, or null
  double idx[3];              // 1 / dx
  int rel, ns;
  long long ncell, np;
  double *part;               // [gridDim.x][ns * 7]
};

__global__ void __launch_bounds__(kBlock) k_species_sums32(SumsArgs32 a) {
  __shared__ double sh[7][kBlock];
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  const int l = threadIdx.x & 15;
  for (int q = 0; q < a.ns; q++) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    if (c < a.ncell) {
      const int n = a.cnt[c];
      for (int k = l; k < n; k += 16) {
        const long long r = (long long)a.rbase[c] + k, s = (long long)a.sbase[c] + k;
        const int4 A = a.ra[r], B = a.rb[r];
        if ((int)((uint32_t)A.x >> kSpeciesShift) != q) continue;
        const long long id = B.w;
        float fb[3], ub[3], gb0 = 0.f;
        if (a.adj) {
          if (id < 0 || id >= a.np) continue;
          const float *p = a.adj + 6 * id;
          for (int d = 0; d < 3; d++) {
            fb[d] = p[d];
            ub[d] = p[3 + d];
          }
          if (a.g0b) gb0 = a.g0b[id];
        } else {
          for (int d = 0; d < 3; d++) {
            fb[d] = a.pa.f[d][s];
            ub[d] = a.pa.u[d][s];
          }
          if (a.g0b) gb0 = a.g0b[s];
        }
        const double gm = a.g0 ? (double)a.g0[s] : 1.0;
        const float u[3] = {as_f(B.x), as_f(B.y), as_f(B.z)};
        double gb = a.g0 ? (double)gb0 : 0.0;
        for (int d = 0; d < 3; d++) {
          acc[d] += (double)fb[d] * a.idx[d];
          acc[3 + d] += (double)ub[d] * (a.rel ? gm : 1.0);
          gb += (double)ub[d] * ((double)u[d] / gm);
        }
        acc[6] += a.rel ? gb : 0.0;
      }
    }
    for (int j = 0; j < 7; j++) sh[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int j = 0; j < 7; j++) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x < 7) a.part[((long long)blockIdx.x * a.ns + q) * 7 + threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

// device buffers outside the pool (hipMalloc, hipFree)
template <typename T>
int adj_alloc(emo_ctx *c, T **p, size_t n) {
  if (hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)) != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    c->err = "hipMalloc failed in the adjoint";
    return EMO_ERR_ALLOC;
  }
  return EMO_OK;
}
void adj_free(void *p) {
  if (p) (void)hipFree(p);
}

// FP64 plain form: the particle adjoints by position i of the initial order
// (ids id[i], species sp[i])
struct SumsArgs64 {
  PartAdj pb;
  const int32_t *sp;
  const int64_t *id;
  int ns;
  long long np;
  double *part;  // [gridDim.x][ns * 7]
};

__global__ void __launch_bounds__(kBlock) k_species_sums64(SumsArgs64 a) {
  __shared__ double sh[7][kBlock];
  for (int q = 0; q < a.ns; q++) {
    double acc[7] = {0, 0, 0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * kBlock + threadIdx.x; i < a.np; i += (long long)gridDim.x * kBlock) {
      if (a.sp[i] != q || a.id[i] < 0 || a.id[i] >= a.np) continue;
      for (int d = 0; d < 3; d++) {
        acc[d] += a.pb.x[d][i];
        acc[3 + d] += a.pb.v[d][i];
      }
      acc[6] += a.pb.gam[i];
    }
    for (int j = 0; j < 7; j++) sh[j][threadIdx.x] = acc[j];
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w)
        for (int j = 0; j < 7; j++) sh[j][threadIdx.x] += sh[j][threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x < 7) a.part[((long long)blockIdx.x * a.ns + q) * 7 + threadIdx.x] = sh[threadIdx.x][0];
    __syncthreads();
  }
}

// block partials [nb][ns * 7] -> c->adj_sums, summed in block order
int finish_species_sums(emo_ctx *c, double *part, long long nb, int ns) {
  std::vector<double> h((size_t)(nb * ns * 7));
  hipError_t e = hipMemcpyAsync(h.data(), part, h.size() * 8, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  adj_free(part);
  if (e != hipSuccess) {
    c->err = hipGetErrorString(e);
    return EMO_ERR_HIP;
  }
  std::fill(c->adj_sums, c->adj_sums + 7 * EMO_NSPEC, 0.0);
  for (long long b = 0; b < nb; b++)
    for (int q = 0; q < ns * 7; q++) c->adj_sums[q] += h[(size_t)(b * ns * 7 + q)];
  return EMO_OK;
}

int species_sums64(emo_ctx *c, SumsArgs64 a) {
  const long long nb = 1024;
  const int ns = std::max(1, (int)c->cfg.nspecies);
  double *part = nullptr;
  int rc = adj_alloc(c, &part, (size_t)(nb * ns * 7));
  if (rc) return rc;
  a.ns = ns;
  a.np = c->np;
  a.part = part;
  hipLaunchKernelGGL(k_species_sums64, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, a);
  return finish_species_sums(c, part, nb, ns);
}

// the sums into c->adj_sums (partials on the device, summed in block order)
int species_sums32(emo_ctx *c, SumsArgs32 a) {
  const long long nb = (c->geo.ncell + (kBlock / 16) - 1) / (kBlock / 16);
  const int ns = std::max(1, (int)c->cfg.nspecies);
  double *part = nullptr;
  int rc = adj_alloc(c, &part, (size_t)(nb * ns * 7));
  if (rc) return rc;
  for (int d = 0; d < 3; d++) a.idx[d] = c->geo.id[d];
  a.rel = c->cfg.relativity;
  a.ns = ns;
  a.ncell = c->geo.ncell;
  a.np = c->np;
  a.part = part;
  hipLaunchKernelGGL(k_species_sums32, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, a);
  return finish_species_sums(c, part, nb, ns);
}

// ---------------------------------------------------------------------------
// Binomial checkpointing: the low-memory gradient of the FP32 form
// ---------------------------------------------------------------------------
// Snapshots hold the state densely (the records in cell order, 32 bytes a
// particle, with the fields and the slot layout); the particle adjoints are
// indexed by particle id and updated in place; no step keeps a tape.  The
// backward step of step s runs on the live state at s, which a restore and a
// re-run reach.  The split points minimise the re-run steps for the number
// of snapshots that fit in device memory (binomial checkpointing, Griewank
// and Walther, "Algorithm 799: revolve", ACM TOMS 26, 2000), found by
// dynamic programming over (steps, free snapshots).

struct Dense32 {
  int4 *ra = nullptr, *rb = nullptr;
  float *E[3] = {}, *A[2][3] = {};  // A[0] = Bfix, A[1] = Btfix
  int32_t *base = nullptr, *cnt = nullptr, *doff = nullptr;
  bool b_is_a0 = true;
  int step = -1;                    // -1: free
};

void dense32_alloc(Pool &P, const emo_ctx *c, Dense32 &s) {
  const size_t n = (size_t)std::max<int64_t>(c->np, 1), nt = (size_t)c->geo.ntot;
  const size_t nc1 = (size_t)c->geo.ncell + 1;
  P.take(&s.ra, n);
  P.take(&s.rb, n);
  for (int d = 0; d < 3; d++) {
    P.take(&s.E[d], nt);
    P.take(&s.A[0][d], nt);
    P.take(&s.A[1][d], nt);
  }
  P.take(&s.base, nc1);
  P.take(&s.cnt, nc1);
  P.take(&s.doff, nc1);
}

struct PackArgs {
  int4 *pa, *pb;             // the slot layout
  int4 *ra, *rb;             // dense, cell order
  const int32_t *base, *cnt, *doff;
  long long ncell, np;
};

// cell c's cnt[c] records between the slots from base[c] and the dense
// positions from doff[c]; 16 lanes per cell
template <bool UNPACK>
__global__ void __launch_bounds__(kBlock) k_pack_dense32(PackArgs a) {
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  const int l = threadIdx.x & 15;
  if (c >= a.ncell) return;
  const int n = a.cnt[c];
  const long long b = a.base[c], d = a.doff[c];
  for (int k = l; k < n; k += 16) {
    if (d + k >= a.np) break;  // (only where particles were lost, which is flagged)
    if (UNPACK) {
      a.pa[b + k] = a.ra[d + k];
      a.pb[b + k] = a.rb[d + k];
    } else {
      a.ra[d + k] = a.pa[b + k];
      a.rb[d + k] = a.pb[b + k];
    }
  }
}

// Split points of binomial checkpointing: arg(n, s) for reversing n steps
// with s free snapshots, minimising re-run steps plus kRestoreCost per
// restore (a restore moves about as many bytes as half a step).
struct BinomialPlan {
  int N = 0, S = 0;
  std::vector<double> cost;
  std::vector<int> arg;
  static constexpr double kRestoreCost = 0.5;
  double &C(int n, int s) { return cost[(size_t)s * (N + 1) + n]; }
  int &A(int n, int s) { return arg[(size_t)s * (N + 1) + n]; }
  void build(int n_, int s_) {
    N = n_;
    S = s_;
    cost.assign((size_t)(S + 1) * (N + 1), 0.0);
    arg.assign((size_t)(S + 1) * (N + 1), 1);
    for (int n = 2; n <= N; n++) C(n, 0) = C(n - 1, 0) + (n - 1) + kRestoreCost;
    const bool exact = (double)N * N * S <= 2e8;
    for (int s = 1; s <= S; s++) {
      C(1, s) = 0.0;
      for (int n = 2; n <= N; n++) {
        if (!exact) {  // long runs: equal parts
          const int k = std::max(1, n / (s + 1));
          A(n, s) = k;
          C(n, s) = k + C(n - k, s - 1) + kRestoreCost + C(k, s);
          continue;
        }
        double best = -1.0;
        int bk = 1;
        for (int k = 1; k < n; k++) {
          const double v = k + C(n - k, s - 1) + kRestoreCost + C(k, s);
          if (best < 0 || v < best) {
            best = v;
            bk = k;
          }
        }
        C(n, s) = best;
        A(n, s) = bk;
      }
    }
  }
};

struct Lean32 {
  emo_ctx *c = nullptr;
  double dt = 0.0;
  int N = 0, S = 0;
  bool has_g0 = false, all_per = true;
  std::vector<double> tstep;
  std::vector<Dense32> slot;     // slot[0]: the initial state
  std::vector<int> free_slots;
  std::vector<int> stored;       // step -> slot holding it, or -1
  int live = 0;                  // the step of the live state
  float *g0slots = nullptr;      // the stored gamma of the initial layout (owned here)
  float *adj = nullptr;          // 6 per particle id
  float *g0b = nullptr;          // by id (first step)
  float *Jb32[3] = {};
  AdjFields af;
  double *dsums = nullptr;
  std::vector<FillPlane> planes;
  BinomialPlan plan;
  int64_t nfwd = 0, nrestore = 0;
  int rc = 0;
};

void dense32_take(Lean32 &L, int sl, int step) {
  emo_ctx *c = L.c;
  Dense32 &s = L.slot[sl];
  Fp32View v = fp32_view(c);
  const size_t nc1 = (size_t)c->geo.ncell + 1, nt4 = (size_t)c->geo.ntot * 4;
  CopyBatch cb(c);
  cb.add(s.base, *v.base, nc1 * 4);
  cb.add(s.cnt, *v.cnt, nc1 * 4);
  for (int d = 0; d < 3; d++) {
    cb.add(s.E[d], v.fs->E[d], nt4);
    cb.add(s.A[0][d], v.fs->Bfix[d], nt4);
    cb.add(s.A[1][d], v.Btfix[d], nt4);
  }
  cb.flush();
  if (!L.rc) L.rc = fp32_scan_i32(c, s.cnt, s.doff, (long long)nc1);
  PackArgs pk{*v.pa, *v.pb, s.ra, s.rb, s.base, s.cnt, s.doff, c->geo.ncell, c->np};
  const unsigned blocks = (unsigned)((c->geo.ncell + (kBlock / 16) - 1) / (kBlock / 16));
  hipLaunchKernelGGL(k_pack_dense32<false>, dim3(blocks), dim3(kBlock), 0, c->stream, pk);
  s.b_is_a0 = v.fs->B[0] == v.fs->Bfix[0];
  s.step = step;
  L.stored[step] = sl;
}

void dense32_restore(Lean32 &L, int sl) {
  emo_ctx *c = L.c;
  const Dense32 &s = L.slot[sl];
  Fp32View v = fp32_view(c);
  const size_t nc1 = (size_t)c->geo.ncell + 1, nt4 = (size_t)c->geo.ntot * 4;
  CopyBatch cb(c);
  cb.add(*v.base, s.base, nc1 * 4);
  cb.add(*v.cnt, s.cnt, nc1 * 4);
  for (int d = 0; d < 3; d++) {
    cb.add(v.fs->E[d], s.E[d], nt4);
    cb.add(v.fs->Bfix[d], s.A[0][d], nt4);
    cb.add(v.Btfix[d], s.A[1][d], nt4);
  }
  cb.flush();
  PackArgs pk{*v.pa, *v.pb, s.ra, s.rb, s.base, s.cnt, s.doff, c->geo.ncell, c->np};
  const unsigned blocks = (unsigned)((c->geo.ncell + (kBlock / 16) - 1) / (kBlock / 16));
  hipLaunchKernelGGL(k_pack_dense32<true>, dim3(blocks), dim3(kBlock), 0, c->stream, pk);
  for (int q = 0; q < 3; q++) {
    v.fs->B[q] = s.b_is_a0 ? v.fs->Bfix[q] : v.Btfix[q];
    v.fs->Bt[q] = s.b_is_a0 ? v.Btfix[q] : v.fs->Bfix[q];
  }
  const bool g0 = s.step == 0 && L.has_g0;
  *v.g0_active = g0;
  *v.g0 = g0 ? L.g0slots : nullptr;
  *v.g0_owned = false;
  L.live = s.step;
  L.nrestore++;
}

void lean_advance(Lean32 &L, int to) {
  while (!L.rc && L.live < to) {
    L.rc = fp32_step_one(L.c, L.tstep[L.live], L.dt, nullptr);
    L.live++;
    L.nfwd++;
  }
}

// the live state at step `to` (a <= to), from the snapshot sa of step a
void lean_reach(Lean32 &L, int sa, int to) {
  if (L.rc) return;
  const int a = L.slot[sa].step;
  if (L.live > to || L.live < a) dense32_restore(L, sa);
  lean_advance(L, to);
}

int lean_store(Lean32 &L, int step) {
  const int sl = L.free_slots.back();
  L.free_slots.pop_back();
  dense32_take(L, sl, step);
  return sl;
}

// the backward step of step st on the live state (at st)
void lean_backward(Lean32 &L, int st) {
  if (L.rc) return;
  emo_ctx *c = L.c;
  const Geo &g = c->geo;
  Fp32View v = fp32_view(c);
  const bool b_is_a0 = v.fs->B[0] == v.fs->Bfix[0];
  const int nplanes = (int)L.planes.size();
  fields_step_adj(c, c->stream, L.af, L.planes, L.dsums + (size_t)st * nplanes, L.dt, b_is_a0);
  hipLaunchKernelGGL(k_round_f32, dim3(grid_for(g.ntot)), dim3(kBlock), 0, c->stream, (const double *)L.af.Jb[0],
                     (const double *)L.af.Jb[1], (const double *)L.af.Jb[2], L.Jb32[0], L.Jb32[1], L.Jb32[2],
                     (long long)g.ntot);
  CellStepAdj32Args ca{};
  ca.k = fp32_push_consts(c, L.dt);
  ca.pa = *v.pa;
  ca.pb = *v.pb;
  ca.g0 = (st == 0 && L.has_g0) ? L.g0slots : nullptr;
  ca.base = *v.base;
  ca.cnt = *v.cnt;
  ca.fwd = nullptr;
  for (int d = 0; d < 3; d++) {
    ca.E[d] = v.fs->E[d];
    ca.B[d] = b_is_a0 ? v.fs->Bfix[d] : v.Btfix[d];
    ca.Bt[d] = b_is_a0 ? v.Btfix[d] : v.fs->Bfix[d];
    ca.Jb[d] = L.Jb32[d];
  }
  ca.g0b = ca.g0 ? L.g0b : nullptr;
  if (ca.g0b) {
    // (g0b may share a snapshot's memory, free from this step on: cleared here)
    hipLaunchKernelGGL(k_zero_f, dim3(grid_for(std::max<int64_t>(c->np, 1))), dim3(kBlock), 0, c->stream, L.g0b,
                       (long long)c->np);
  }
  ca.cg = L.af.cg;
  ca.ncell = g.ncell;
  ca.nslots = (uint32_t)v.nslots;
  ca.adj = L.adj;
  ca.np = c->np;
  const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
  {
    TimedLaunch tl(c, "adj_cells_f32", c->stream);
    launch_cell_step_adj32(c, c->stream, ca, L.all_per, blocks, true);
  }
  gather_adj_grid_roles(c, c->stream, b_is_a0, L.af);
}

// reverses steps [a, b) with s free snapshots; snapshot sa holds step a
void lean_reverse(Lean32 &L, int a, int b, int s, int sa) {
  if (L.rc) return;
  if (b - a == 1) {
    lean_reach(L, sa, a);
    lean_backward(L, a);
    return;
  }
  const int m = s > 0 ? a + L.plan.A(b - a, s) : a;
  // no free snapshot (or, defensively, none left to store the split point in):
  // every backward step re-runs from a
  if (s == 0 || (L.stored[m] < 0 && L.free_slots.empty())) {
    for (int st = b - 1; st >= a && !L.rc; st--) {
      lean_reach(L, sa, st);
      lean_backward(L, st);
    }
    return;
  }
  int sm = L.stored[m];
  if (sm < 0) {
    lean_reach(L, sa, m);
    if (L.rc) return;
    sm = lean_store(L, m);
  }
  lean_reverse(L, m, b, s - 1, sm);
  L.stored[m] = -1;
  L.slot[sm].step = -1;
  L.free_slots.push_back(sm);
  lean_reverse(L, a, m, s, sa);
}

}  // namespace

// the most snapshots binomial checkpointing takes when it picks the number
constexpr int kMaxBinomialSnapshots = 32;

// The gradient call of the FP32 chosen form with binomial checkpointing:
// snapshots in device memory, as many as fit (at most smax besides the
// initial state), particle adjoints by id, no tape.  For runs whose K-step
// segments do not fit (checkpoint_every < 0 asks for it).
int fp32_adjoint_binomial(emo_ctx *c, int32_t nsteps, double t0, double dt, int smax, double *energy, double *d_amp,
                          double *d_phase, double *xbar0, double *vbar0, double *gbar0) {
  const Geo &g = c->geo;
  const int64_t np = c->np;
  const int nw = c->cfg.nwaves;
  const bool rel = c->cfg.relativity != 0;
  Fp32View v = fp32_view(c);
  Lean32 L;
  L.c = c;
  L.dt = dt;
  L.N = nsteps;
  L.planes = fill_planes(c);
  L.has_g0 = *v.g0_active && rel;
  for (int q = 0; q < 6; q++) L.all_per = L.all_per && c->cfg.parts_bound[q] == EMO_BC_PERIODIC;
  // the times of the steps, accumulated as emo_step advances them
  L.tstep.resize(nsteps + 1);
  L.tstep[0] = t0;
  for (int s = 0; s < nsteps; s++) L.tstep[s + 1] = L.tstep[s] + dt;
  const size_t nt = (size_t)g.ntot;
  const int nplanes = (int)L.planes.size();
  auto layout = [&](Pool &P, int S) {
    L.slot.assign(S + 1, Dense32());
    P.take(&L.adj, 6 * (size_t)std::max<int64_t>(np, 1));
    for (int d = 0; d < 3; d++) {
      P.take(&L.af.Eb[d], nt);
      P.take(&L.af.Ab[0][d], nt);
      P.take(&L.af.Ab[1][d], nt);
      P.take(&L.af.Jb[d], nt);
      P.take(&L.Jb32[d], nt);
    }
    P.take(&L.af.cg, 18 * (size_t)g.ncell);
    P.take(&L.dsums, (size_t)std::max(1, nsteps * nplanes));
    // with a free snapshot the stored gamma's adjoint (written by the first
    // step's backward step, when every other snapshot is free) uses one
    if (L.has_g0 && S == 0) P.take(&L.g0b, (size_t)std::max<int64_t>(np, 1));
    for (int k = 0; k <= S; k++) dense32_alloc(P, c, L.slot[k]);
  };
  // the snapshots that fit in what the device has free (the held pool counts)
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    c->err = "hipMemGetInfo failed";
    return EMO_ERR_HIP;
  }
  const size_t avail = fr + c->adj_pool_bytes, margin = (size_t)1 << 30;
  const int scap = std::max(0, std::min(smax, nsteps - 1));
  int S = 0;
  {
    Pool d0;
    layout(d0, 0);
    if (d0.off + margin > avail) {
      c->err = "not enough device memory for the adjoint (particle adjoints and one snapshot)";
      return EMO_ERR_ALLOC;
    }
    while (S < scap) {
      Pool dp;
      layout(dp, S + 1);
      if (dp.off + margin > avail) break;
      S++;
    }
  }
  for (;;) {
    Pool dry;
    layout(dry, S);
    if (dry.off <= c->adj_pool_bytes) break;
    if (c->adj_pool) hipFree(c->adj_pool);
    c->adj_pool = nullptr;
    c->adj_pool_bytes = 0;
    if (hipMalloc(&c->adj_pool, dry.off) == hipSuccess) {
      c->adj_pool_bytes = dry.off;
      break;
    }
    c->adj_pool = nullptr;
    (void)hipGetLastError();
    if (S == 0) {
      c->err = "hipMalloc failed for the adjoint's snapshots";
      return EMO_ERR_ALLOC;
    }
    S--;
  }
  {
    Pool real;
    real.base = (char *)c->adj_pool;
    layout(real, S);
  }
  L.S = S;
  for (int k = S; k >= 1; k--) L.free_slots.push_back(k);
  L.stored.assign((size_t)nsteps + 1, -1);
  if (nsteps > 0) L.plan.build(nsteps, S);
  if (L.has_g0 && S >= 1) L.g0b = reinterpret_cast<float *>(L.slot[1].ra);
  // the stored gamma of the initial layout, kept here for the re-runs of the
  // first step (the context's copy, or a copy of a borrowed one)
  if (L.has_g0) {
    if (*v.g0_owned) {
      L.g0slots = *v.g0;
    } else {
      if (adj_alloc(c, &L.g0slots, (size_t)v.nslots)) return EMO_ERR_ALLOC;
      CopyBatch cb(c);
      cb.add(L.g0slots, *v.g0, (size_t)v.nslots * 4);
      cb.flush();
    }
    *v.g0 = L.g0slots;
    *v.g0_owned = false;
  }
  // ---- first pass: the initial snapshot, the rightmost split points, F ----
  L.live = 0;
  dense32_take(L, 0, 0);
  {
    int a = 0, b = nsteps, s = S;
    while (!L.rc && s > 0 && b - a > 1) {
      const int m = a + L.plan.A(b - a, s);
      lean_advance(L, m);
      if (L.rc) break;
      lean_store(L, m);
      a = m;
      s--;
    }
  }
  lean_advance(L, nsteps);
  if (!L.rc) L.rc = fp32_check(c);
  double F = 0.0;
  if (!L.rc) {
    double out8[8];
    L.rc = emo_diag(c, out8);
    F = out8[4] + out8[5];
  }
  // ---- seed and reverse ----------------------------------------------------
  if (!L.rc) {
    for (int d = 0; d < 3; d++) {
      launch_zero(c->stream, L.af.Eb[d], nt);
      launch_zero(c->stream, L.af.Ab[0][d], nt);
      launch_zero(c->stream, L.af.Ab[1][d], nt);
      launch_zero(c->stream, L.af.Jb[d], nt);
    }
    hipLaunchKernelGGL(k_zero_f, dim3(grid_for(6 * std::max<int64_t>(np, 1))), dim3(kBlock), 0, c->stream, L.adj,
                       (long long)(6 * np));
    SeedArgs<float> sa;
    sa.g = g;
    for (int d = 0; d < 3; d++) {
      sa.E[d] = v.fs->E[d];
      sa.B[d] = v.fs->Bfix[d];
      sa.Eb[d] = L.af.Eb[d];
      sa.Bb[d] = L.af.Ab[0][d];
    }
    hipLaunchKernelGGL(k_seed_energy<float>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, c->stream, sa);
  }
  if (!L.rc && nsteps > 0) lean_reverse(L, 0, nsteps, S, 0);
  if (!L.rc) L.rc = fp32_check(c);
  int rc = L.rc;
  const Dense32 &s0 = L.slot[0];
  // ---- outputs ---------------------------------------------------------------
  if (!rc) {
    if (energy) *energy = F;
    std::vector<double> sums((size_t)std::max(1, nsteps * nplanes));
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(sums.data(), L.dsums, sums.size() * 8, hipMemcpyDeviceToHost, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    std::vector<double> da(nw, 0.0), dp(nw, 0.0), ta, tp;
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < nplanes; k++) {
        pulse_derivs(c, L.planes[k], L.tstep[s], dt, ta, tp);
        for (int w = 0; w < nw; w++) {
          da[w] += sums[(size_t)s * nplanes + k] * ta[w];
          dp[w] += sums[(size_t)s * nplanes + k] * tp[w];
        }
      }
    for (int w = 0; w < nw; w++) {
      if (d_amp) d_amp[w] = da[w];
      if (d_phase) d_phase[w] = dp[w];
    }
    SumsArgs32 sa{};
    sa.ra = s0.ra;
    sa.rb = s0.rb;
    sa.rbase = s0.doff;
    sa.sbase = s0.base;
    sa.cnt = s0.cnt;
    sa.g0 = L.has_g0 ? L.g0slots : nullptr;
    sa.adj = L.adj;
    sa.g0b = (L.has_g0 && nsteps > 0) ? L.g0b : nullptr;
    rc = species_sums32(c, sa);
    // particle adjoints of the initial state, by id (the offsets f = (x -
    // x0) / dx - ix and, relativistic, u = gamma0 v with the stored gamma0)
    if (!rc && (xbar0 || vbar0 || gbar0) && np > 0) {
      const size_t nc1 = (size_t)g.ncell + 1;
      std::vector<int32_t> hbase(nc1), hcnt(nc1), hdoff(nc1);
      std::vector<float> adj((size_t)(6 * np)), g0h, g0bh;
      std::vector<int4> rb((size_t)np);
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, hbase.data(), s0.base, nc1 * 4));
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, hcnt.data(), s0.cnt, nc1 * 4));
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, hdoff.data(), s0.doff, nc1 * 4));
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, rb.data(), s0.rb, (size_t)np * 16));
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, adj.data(), L.adj, (size_t)np * 24));
      if (L.has_g0) {
        g0h.resize((size_t)v.nslots);
        g0bh.assign((size_t)np, 0.f);
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, g0h.data(), L.g0slots, (size_t)v.nslots * 4));
        if (nsteps > 0) EMO_HIP_CHECK_RET(c, ctx_d2h(c, g0bh.data(), L.g0b, (size_t)np * 4));
      }
      for (long long cell = 0; cell < g.ncell; cell++)
        for (int j = 0; j < hcnt[cell]; j++) {
          const int64_t r = (int64_t)hdoff[cell] + j, sl = (int64_t)hbase[cell] + j;
          if (r >= np) break;
          const int64_t k = rb[r].w;
          if (k < 0 || k >= np) continue;
          float u[3];
          std::memcpy(u, &rb[r], 12);
          const double gm = L.has_g0 ? (double)g0h[sl] : 1.0;
          double gb = L.has_g0 ? (double)g0bh[k] : 0.0;
          for (int d = 0; d < 3; d++) {
            const double fb = adj[6 * k + d], ub = adj[6 * k + 3 + d];
            if (xbar0) xbar0[3 * k + d] = fb * g.id[d];
            if (vbar0) vbar0[3 * k + d] = ub * (rel ? gm : 1.0);
            gb += ub * ((double)u[d] / gm);
          }
          if (gbar0) gbar0[k] = rel ? gb : 0.0;
        }
    }
  }
  // ---- the final state: re-run from the initial snapshot --------------------
  if (nsteps > 0 && !rc) {
    lean_reach(L, 0, nsteps);
    if (!L.rc) L.rc = fp32_check(c);
    rc = L.rc;
  }
  if (L.has_g0) {
    if (nsteps == 0) {
      // the state is still the initial one: its stored gamma stays with it
      *v.g0 = L.g0slots;
      *v.g0_active = true;
      *v.g0_owned = true;
    } else {
      // past the first step the stored gamma is no longer part of the state
      (void)hipStreamSynchronize(c->stream);
      if (*v.g0 == L.g0slots) {
        *v.g0 = nullptr;
        *v.g0_active = false;
        *v.g0_owned = true;
      }
      adj_free(L.g0slots);
    }
  }
  c->adj_info[0] = 1;
  c->adj_info[1] = S + 1;
  c->adj_info[2] = L.nfwd;
  c->adj_info[3] = L.nrestore;
  return rc;
}

// The gradient call of the FP32 chosen form: forward with checkpoints every
// K steps (the last segment taped on the way), then segment by segment
// backwards, each segment's recomputation (context stream) overlapping the
// adjoint of the segment after it (adjoint stream).
int fp32_adjoint_field_energy(emo_ctx *c, int32_t nsteps, double t0, double dt, int32_t K, double *energy,
                              double *d_amp, double *d_phase, double *xbar0, double *vbar0, double *gbar0) {
  if (K < 0)
    return fp32_adjoint_binomial(c, nsteps, t0, dt, -K, energy, d_amp, d_phase, xbar0, vbar0, gbar0);
  if (K == 0) K = std::max(nsteps, 1);
  const Geo &g = c->geo;
  const int64_t np = c->np;
  const int nw = c->cfg.nwaves;
  const bool rel = c->cfg.relativity != 0;
  int rc = 0;
  if (!c->adj_stream && hipStreamCreate(&c->adj_stream) != hipSuccess) {
    c->adj_stream = nullptr;
    c->err = "hipStreamCreate failed for the adjoint";
    return EMO_ERR_HIP;
  }
  Fp32View v = fp32_view(c);
  AdjRun32 R;
  R.c = c;
  R.as = c->adj_stream;
  R.dt = dt;
  R.nsteps = nsteps;
  R.K = K;
  R.nseg = (nsteps + K - 1) / K;
  R.planes = fill_planes(c);
  R.ck.resize(std::max(R.nseg, 1));
  int nsets = R.nseg > 1 ? 2 : 1;
  for (int st = 0; st < kMaxTapeSets; st++) R.tape[st].resize(K + 1);
  const size_t nt = (size_t)g.ntot;
  const size_t npa = (size_t)v.nslots;  // particle arrays are by slot
  const int nplanes = (int)R.planes.size();
  const bool has_g0 = *v.g0_active && rel;
  auto layout = [&](Pool &P) {
    for (size_t k = 0; k < R.ck.size(); k++) snap32_alloc(P, c, R.ck[k], true, k == 0 && has_g0, false, false);
    snap32_alloc(P, c, R.fin, false, false, false, true);
    for (int st = 0; st < nsets; st++)
      for (auto &sn : R.tape[st]) snap32_alloc(P, c, sn, true, false, true, false);
    for (int d = 0; d < 3; d++) {
      P.take(&R.af.Eb[d], nt);
      P.take(&R.af.Ab[0][d], nt);
      P.take(&R.af.Ab[1][d], nt);
      P.take(&R.af.Jb[d], nt);
      P.take(&R.Jb32[d], nt);
      P.take(&R.pa.f[d], npa);
      P.take(&R.pa.u[d], npa);
      P.take(&R.pa2.f[d], npa);
      P.take(&R.pa2.u[d], npa);
    }
    P.take(&R.af.cg, 18 * (size_t)g.ncell);
    P.take(&R.dsums, (size_t)std::max(1, nsteps * nplanes));
    P.take(&R.g0b, npa);
    P.take(&R.stage_a, npa);
    P.take(&R.stage_b, npa);
    P.take(&R.mvsrc, npa);
  };
  nsets = more_tape_sets(c, R.nseg, nsets, [&](Pool &P, int ns) {
    const int keep = nsets;
    nsets = ns;
    layout(P);
    nsets = keep;
  });
  {
    Pool dry;
    layout(dry);
    if (dry.off > c->adj_pool_bytes) {
      // segments and tape that do not fit in device memory: binomial
      // checkpointing, with what fits
      size_t fr = 0, tot = 0;
      if (hipMemGetInfo(&fr, &tot) == hipSuccess && dry.off + ((size_t)1 << 30) > fr + c->adj_pool_bytes)
        return fp32_adjoint_binomial(c, nsteps, t0, dt, kMaxBinomialSnapshots, energy, d_amp, d_phase, xbar0, vbar0,
                                     gbar0);
      if (c->adj_pool) hipFree(c->adj_pool);
      c->adj_pool = nullptr;
      c->adj_pool_bytes = 0;
      if (hipMalloc(&c->adj_pool, dry.off) != hipSuccess) {
        c->adj_pool = nullptr;
        c->err = "hipMalloc failed for the adjoint's checkpoints and tape";
        return EMO_ERR_ALLOC;
      }
      c->adj_pool_bytes = dry.off;
    }
    Pool real;
    real.base = (char *)c->adj_pool;
    layout(real);
  }
  // ---- forward with checkpoints -------------------------------------------
  double t = t0;
  // the times of the steps, accumulated as emo_step advances them
  std::vector<double> tstep(nsteps + 1);
  tstep[0] = t0;
  for (int s = 0; s < nsteps; s++) tstep[s + 1] = tstep[s] + dt;
  // The last nsets segments are taped on the way, one into each tape set, so
  // that the backward recomputes only the others.
  const int taped_from = std::max(R.nseg - nsets, 0);
  for (int seg = 0; seg < R.nseg && !rc; seg++) {
    snap32_take(c, R.ck[seg], t, true);
    if (seg == 0 && has_g0) {
      // the first push reads the checkpoint's copy of the stored gamma
      if (*v.g0_owned) {
        EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
        hipFree(*v.g0);
      }
      *v.g0 = R.ck[0].g0;
      *v.g0_owned = false;
    }
    const int s0 = seg * K, s1 = std::min(nsteps, s0 + K);
    if (seg >= taped_from) {
      // the context's own record buffers (the segment's end state is copied
      // back into them from the tape)
      R.own[0] = *v.pa;
      R.own[1] = *v.pb;
      R.own[2] = *v.pa_alt;
      R.own[3] = *v.pb_alt;
      rc = recompute32(R, seg, seg % nsets);
      if (!rc) {
        CopyBatch cb(c);
        cb.add(R.own[0], R.tape[seg % nsets][s1 - s0].pa, npa * 16);
        cb.add(R.own[1], R.tape[seg % nsets][s1 - s0].pb, npa * 16);
        cb.flush();
      }
    } else {
      for (int st = s0; st < s1 && !rc; st++) rc = fp32_step_one(c, tstep[st], dt, nullptr);
    }
    t = tstep[s1];
  }
  if (R.nseg == 0) {
    snap32_take(c, R.ck[0], t, true);  // the ids of the outputs
    R.own[0] = *v.pa;
    R.own[1] = *v.pb;
    R.own[2] = *v.pa_alt;
    R.own[3] = *v.pb_alt;
  }
  if (!rc) rc = fp32_check(c);
  double F = 0.0;
  if (!rc) {
    double out8[8];
    rc = emo_diag(c, out8);
    F = out8[4] + out8[5];
  }
  bool fin_taken = false;
  if (!rc) {
    snap32_take(c, R.fin, t, false);
    fin_taken = true;
  }
  std::vector<hipEvent_t> ev;
  auto record = [&](hipStream_t st) -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    ev.push_back(e);
    if (hipEventRecord(e, st) != hipSuccess) return nullptr;
    return e;
  };
  auto wait = [&](hipStream_t st, hipEvent_t e) -> int {
    if (!e || hipStreamWaitEvent(st, e, 0) != hipSuccess) {
      c->err = "stream ordering of the adjoint failed";
      return EMO_ERR_HIP;
    }
    return EMO_OK;
  };
  if (!rc) rc = wait(R.as, record(c->stream));
  if (!rc) {
    for (int d = 0; d < 3; d++) {
      launch_zero(R.as, R.af.Eb[d], nt);
      launch_zero(R.as, R.af.Ab[0][d], nt);
      launch_zero(R.as, R.af.Ab[1][d], nt);
      launch_zero(R.as, R.af.Jb[d], nt);
      hipLaunchKernelGGL(k_zero_f, dim3(grid_for(npa)), dim3(kBlock), 0, R.as, R.pa.f[d], (long long)npa);
      hipLaunchKernelGGL(k_zero_f, dim3(grid_for(npa)), dim3(kBlock), 0, R.as, R.pa.u[d], (long long)npa);
    }
    hipLaunchKernelGGL(k_zero_f, dim3(grid_for(npa)), dim3(kBlock), 0, R.as, R.g0b, (long long)npa);
    SeedArgs<float> sa;
    sa.g = g;
    for (int d = 0; d < 3; d++) {
      sa.E[d] = R.fin.E[d];
      sa.B[d] = R.fin.A[0][d];
      sa.Eb[d] = R.af.Eb[d];
      sa.Bb[d] = R.af.Ab[0][d];
    }
    hipLaunchKernelGGL(k_seed_energy<float>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, R.as, sa);
  }
  if (!rc && nsteps > 0) {
    std::vector<hipEvent_t> done_r(R.nseg, nullptr), done_a(R.nseg, nullptr);
    done_r[R.nseg - 1] = record(c->stream);
    for (int seg = taped_from; seg < R.nseg - 1; seg++) done_r[seg] = done_r[R.nseg - 1];
    for (int seg = R.nseg - 1; seg >= 0 && !rc; seg--) {
      rc = wait(R.as, done_r[seg]);
      if (rc) break;
      backward32(R, seg, seg % nsets);
      done_a[seg] = record(R.as);
      if (seg > 0 && seg - 1 < taped_from) {
        // into the set the backward of segment seg - 1 + nsets read
        if (seg - 1 + nsets < R.nseg) rc = wait(c->stream, done_a[seg - 1 + nsets]);
        if (!rc) rc = recompute32(R, seg - 1, (seg - 1) % nsets);
        if (!rc) done_r[seg - 1] = record(c->stream);
      }
    }
  }
  if (!rc && hipStreamSynchronize(R.as) != hipSuccess) {
    c->err = "adjoint stream failed";
    rc = EMO_ERR_HIP;
  }
  if (!rc) rc = fp32_check(c);
  // ---- outputs -------------------------------------------------------------
  if (!rc) {
    if (energy) *energy = F;
    std::vector<double> sums((size_t)std::max(1, nsteps * nplanes));
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(sums.data(), R.dsums, sums.size() * 8, hipMemcpyDeviceToHost, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    std::vector<double> da(nw, 0.0), dp(nw, 0.0), ta, tp;
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < nplanes; k++) {
        pulse_derivs(c, R.planes[k], tstep[s], dt, ta, tp);
        for (int w = 0; w < nw; w++) {
          da[w] += sums[(size_t)s * nplanes + k] * ta[w];
          dp[w] += sums[(size_t)s * nplanes + k] * tp[w];
        }
      }
    for (int w = 0; w < nw; w++) {
      if (d_amp) d_amp[w] = da[w];
      if (d_phase) d_phase[w] = dp[w];
    }
    {
      SumsArgs32 sa{};
      sa.ra = R.ck[0].pa;
      sa.rb = R.ck[0].pb;
      sa.rbase = R.ck[0].base;
      sa.sbase = R.ck[0].base;
      sa.cnt = R.ck[0].cnt;
      sa.g0 = has_g0 ? R.ck[0].g0 : nullptr;
      sa.adj = nullptr;
      sa.pa = R.pa;
      sa.g0b = has_g0 ? R.g0b : nullptr;
      rc = species_sums32(c, sa);
    }
    // particle adjoints of the initial state, by id: the offsets f = (x -
    // x0) / dx - ix and, relativistic, u = gamma0 v with the stored gamma0
    if (!rc && (xbar0 || vbar0 || gbar0) && np > 0) {
      // the initial state's slots: cell c holds cnt[c] particles from base[c]
      const size_t nc1 = (size_t)g.ncell + 1;
      std::vector<int32_t> hbase(nc1), hcnt(nc1);
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, hbase.data(), R.ck[0].base, nc1 * 4));
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, hcnt.data(), R.ck[0].cnt, nc1 * 4));
      std::vector<int4> rb(npa);
      std::vector<float> g0h;
      std::vector<float> fb[3], ub[3], g0bh;
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, rb.data(), R.ck[0].pb, npa * 16));
      for (int d = 0; d < 3; d++) {
        fb[d].resize(npa);
        ub[d].resize(npa);
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, fb[d].data(), R.pa.f[d], npa * 4));
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, ub[d].data(), R.pa.u[d], npa * 4));
      }
      if (has_g0) {
        g0h.resize(npa);
        g0bh.resize(npa);
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, g0h.data(), R.ck[0].g0, npa * 4));
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, g0bh.data(), R.g0b, npa * 4));
      }
      for (long long cell = 0; cell < g.ncell; cell++)
        for (int j = 0; j < hcnt[cell]; j++) {
          const int64_t i = (int64_t)hbase[cell] + j;
          const int64_t k = rb[i].w;
          if (k < 0 || k >= np) continue;
          float u[3];
          std::memcpy(u, &rb[i], 12);
          const double gm = has_g0 ? (double)g0h[i] : 1.0;
          double gb = has_g0 ? (double)g0bh[i] : 0.0;
          for (int d = 0; d < 3; d++) {
            if (xbar0) xbar0[3 * k + d] = (double)fb[d][i] * g.id[d];
            if (vbar0) vbar0[3 * k + d] = (double)ub[d][i] * (rel ? gm : 1.0);
            gb += (double)ub[d][i] * ((double)u[d] / gm);
          }
          if (gbar0) gbar0[k] = rel ? gb : 0.0;
        }
    }
  }
  // checkpointing of this call: K-step segments, each re-run once
  c->adj_info[0] = 0;
  c->adj_info[1] = R.nseg;
  c->adj_info[2] = (int64_t)nsteps + (int64_t)std::max(R.nseg - 1, 0) * R.K;
  c->adj_info[3] = R.nseg;
  // ---- the final state back (the records stayed in the own buffers) --------
  hipStreamSynchronize(R.as);
  if (R.nseg > 0) restore_own_buffers(R);
  if (nsteps > 0 && fin_taken) {
    Fp32View w = fp32_view(c);
    const size_t nt4 = (size_t)g.ntot * 4;
    CopyBatch cb(c);
    for (int d = 0; d < 3; d++) {
      cb.add(w.fs->E[d], R.fin.E[d], nt4);
      cb.add(w.fs->Bfix[d], R.fin.A[0][d], nt4);
      cb.add(w.Btfix[d], R.fin.A[1][d], nt4);
      cb.add(w.fs->J[d], R.fin.J[d], nt4);
    }
    cb.add(*w.base, R.fin.base, (size_t)(g.ncell + 1) * 4);
    cb.add(*w.cnt, R.fin.cnt, (size_t)(g.ncell + 1) * 4);
    cb.flush();
    for (int q = 0; q < 3; q++) {
      w.fs->B[q] = R.fin.b_is_a0 ? w.fs->Bfix[q] : w.Btfix[q];
      w.fs->Bt[q] = R.fin.b_is_a0 ? w.Btfix[q] : w.fs->Bfix[q];
    }
    *w.g0 = nullptr;
    *w.g0_active = false;
    *w.g0_owned = true;
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      c->err = hipGetErrorString(e);
      rc = EMO_ERR_HIP;
    }
  }
  for (hipEvent_t e : ev) hipEventDestroy(e);
  return rc;
}

// dx0/d pert_eps of the displacement loader, per particle in load order
int loader_dx_deps(const emo_config *cfg, const emo_loader *ld, std::vector<double> &out);

}  // namespace emo

using namespace emo;

extern "C" int emo_adjoint_field_energy(emo_ctx *c, int32_t nsteps, double t0, double dt, int32_t K,
                                        double *energy, double *d_amp, double *d_phase, double *xbar0,
                                        double *vbar0, double *gbar0) {
  if (!c || nsteps < 0) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN)
    return fp32_adjoint_field_energy(c, nsteps, t0, dt, K, energy, d_amp, d_phase, xbar0, vbar0, gbar0);
  if (K <= 0) K = std::max(nsteps, 1);
  const Geo &g = c->geo;
  const int64_t np = c->np;
  const int nw = c->cfg.nwaves;
  int rc = 0;
  // the first push's fields enter the derivative as the gather of the grid
  if (c->fld_state == kFldStale) {
    c->err = "the adjoint starts from particle fields gathered from the grid (emo_interpolate)";
    return EMO_ERR_UNSUPPORTED;
  }
  if (!c->adj_stream && hipStreamCreate(&c->adj_stream) != hipSuccess) {
    c->adj_stream = nullptr;
    c->err = "hipStreamCreate failed for the adjoint";
    return EMO_ERR_HIP;
  }
  // the reference loop sorts before its first push (src/main.cpp:225-226)
  rc = sort_f64(c, c->fld_state == kFldFresh ? 1 : 0);
  if (rc) return rc;
  // ---- buffers: checkpoints, final state, two tape sets, adjoints -----------
  AdjRun R;
  R.c = c;
  R.as = c->adj_stream;
  R.dt = dt;
  R.nsteps = nsteps;
  R.K = K;
  R.nseg = (nsteps + K - 1) / K;
  R.planes = fill_planes(c);
  R.ck.resize(std::max(R.nseg, 1));
  for (int st = 0; st < kMaxTapeSets; st++) R.tape[st].resize(K + 1);
  const size_t nt = (size_t)g.ntot;
  const int nplanes = (int)R.planes.size();
  int nsets = R.nseg > 1 ? 2 : 1;
  auto layout = [&](Pool &P) {
    for (auto &sn : R.ck) snap_alloc(P, c, sn, true, false);
    snap_alloc(P, c, R.fin, true, false, true);
    for (int st = 0; st < nsets; st++)
      for (auto &sn : R.tape[st]) snap_alloc(P, c, sn, false, true);
    for (int d = 0; d < 3; d++) {
      P.take(&R.af.Eb[d], nt);
      P.take(&R.af.Ab[0][d], nt);
      P.take(&R.af.Ab[1][d], nt);
      P.take(&R.af.Jb[d], nt);
      P.take(&R.pb.x[d], (size_t)np);
      P.take(&R.pb.v[d], (size_t)np);
      P.take(&R.pb2.x[d], (size_t)np);
      P.take(&R.pb2.v[d], (size_t)np);
    }
    P.take(&R.af.cg, 18 * (size_t)g.ncell);
    P.take(&R.dsums, (size_t)std::max(1, nsteps * nplanes));
    P.take(&R.pb.gam, (size_t)np);
    P.take(&R.pb2.gam, (size_t)np);
  };
  nsets = more_tape_sets(c, R.nseg, nsets, [&](Pool &P, int ns) {
    const int keep = nsets;
    nsets = ns;
    layout(P);
    nsets = keep;
  });
  {
    Pool dry;
    layout(dry);
    if (dry.off > c->adj_pool_bytes) {
      if (c->adj_pool) hipFree(c->adj_pool);
      c->adj_pool = nullptr;
      c->adj_pool_bytes = 0;
      if (hipMalloc(&c->adj_pool, dry.off) != hipSuccess) {
        c->adj_pool = nullptr;
        c->err = "hipMalloc failed for the adjoint's checkpoints and tape";
        return EMO_ERR_ALLOC;
      }
      c->adj_pool_bytes = dry.off;
    }
    Pool real;
    real.base = (char *)c->adj_pool;
    layout(real);
  }
  // ---- forward with checkpoints -------------------------------------------
  double t = t0;
  std::vector<double> tstep(nsteps + 1);
  // The last nsets segments are taped on the way, one into each tape set:
  // the backward recomputes only the others.
  const int last0 = nsteps > 0 ? (R.nseg - 1) * K : 0;
  const int last_set = nsteps > 0 ? (R.nseg - 1) % nsets : 0;
  const int taped_from = std::max(R.nseg - nsets, 0);
  for (int s = 0; s < nsteps; s++) {
    const int seg = s / K;
    if (s % K == 0) {
      // the end state of the previous segment, when that one is taped
      if (seg > 0 && seg - 1 >= taped_from) rc = snap_take(c, R.tape[(seg - 1) % nsets][K], t);
      if (!rc) rc = snap_take(c, R.ck[seg], t);
      if (rc) break;
    }
    tstep[s] = t;
    rc = forward_step(c, t, dt, seg >= taped_from ? &R.tape[seg % nsets][s - seg * K] : nullptr);
    if (rc) break;
    t += dt;
  }
  tstep[nsteps] = t;
  if (!rc && nsteps > 0) rc = snap_take(c, R.tape[last_set][nsteps - last0], t);
  if (!rc && nsteps == 0) rc = snap_take(c, R.ck[0], t);
  // objective
  double F = 0.0;
  if (!rc) {
    double out8[8];
    rc = emo_diag(c, out8);
    F = out8[4] + out8[5];
  }
  // final state (for the last segment's tape and the objective seed)
  bool fin_taken = false;
  if (!rc) {
    rc = snap_take(c, R.fin, t);
    fin_taken = rc == 0;
  }
  std::vector<hipEvent_t> ev;  // every event of this call, destroyed at the end
  auto record = [&](hipStream_t st) -> hipEvent_t {
    hipEvent_t e = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
    ev.push_back(e);
    if (hipEventRecord(e, st) != hipSuccess) return nullptr;
    return e;
  };
  auto wait = [&](hipStream_t st, hipEvent_t e) -> int {
    if (!e || hipStreamWaitEvent(st, e, 0) != hipSuccess) {
      c->err = "stream ordering of the adjoint failed";
      return EMO_ERR_HIP;
    }
    return EMO_OK;
  };
  if (!rc) {
    // the seed reads the final fields: the adjoint stream waits for the forward
    rc = wait(R.as, record(c->stream));
  }
  if (!rc) {
    for (int d = 0; d < 3; d++) {
      launch_zero(R.as, R.af.Eb[d], nt);
      launch_zero(R.as, R.af.Ab[0][d], nt);
      launch_zero(R.as, R.af.Ab[1][d], nt);
      launch_zero(R.as, R.af.Jb[d], nt);
      launch_zero(R.as, R.pb.x[d], np);
      launch_zero(R.as, R.pb.v[d], np);
    }
    launch_zero(R.as, R.pb.gam, np);
    // from the final-state snapshot: the context's grid is overwritten by
    // the recomputation running beside this stream
    SeedArgs<double> sa;
    sa.g = g;
    for (int d = 0; d < 3; d++) {
      sa.E[d] = R.fin.E[d];
      sa.B[d] = R.fin.A[0][d];
      sa.Eb[d] = R.af.Eb[d];
      sa.Bb[d] = R.af.Ab[0][d];
    }
    hipLaunchKernelGGL(k_seed_energy<double>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, R.as, sa);
  }
  // ---- backward: segment seg's adjoint (adjoint stream) runs while segment
  // seg - 1 is recomputed (context stream) into the other tape set ----------
  if (!rc && nsteps > 0) {
    std::vector<hipEvent_t> done_r(R.nseg, nullptr), done_a(R.nseg, nullptr);
    done_r[R.nseg - 1] = record(c->stream);  // taped by the forward pass
    for (int seg = taped_from; seg < R.nseg - 1; seg++) done_r[seg] = done_r[R.nseg - 1];
    for (int seg = R.nseg - 1; seg >= 0 && !rc; seg--) {
      rc = wait(R.as, done_r[seg]);
      if (rc) break;
      backward_segment(R, seg, seg % nsets);
      done_a[seg] = record(R.as);
      if (seg > 0 && seg - 1 < taped_from) {
        // the set seg - 1 tapes into was last read by segment seg - 1 + nsets's adjoint
        if (seg - 1 + nsets < R.nseg) rc = wait(c->stream, done_a[seg - 1 + nsets]);
        if (!rc) rc = recompute_segment(R, seg - 1, (seg - 1) % nsets);
        if (!rc) done_r[seg - 1] = record(c->stream);
      }
    }
  }
  if (!rc && hipStreamSynchronize(R.as) != hipSuccess) {
    c->err = "adjoint stream failed";
    rc = EMO_ERR_HIP;
  }
  // ---- outputs -------------------------------------------------------------
  if (!rc) {
    if (energy) *energy = F;
    std::vector<double> sums((size_t)std::max(1, nsteps * nplanes));
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(sums.data(), R.dsums, sums.size() * 8, hipMemcpyDeviceToHost, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    std::vector<double> da(nw, 0.0), dp(nw, 0.0), ta, tp;
    for (int s = 0; s < nsteps; s++)
      for (int k = 0; k < nplanes; k++) {
        pulse_derivs(c, R.planes[k], tstep[s], dt, ta, tp);
        for (int w = 0; w < nw; w++) {
          da[w] += sums[(size_t)s * nplanes + k] * ta[w];
          dp[w] += sums[(size_t)s * nplanes + k] * tp[w];
        }
      }
    for (int w = 0; w < nw; w++) {
      if (d_amp) d_amp[w] = da[w];
      if (d_phase) d_phase[w] = dp[w];
    }
    if (np > 0) {
      SumsArgs64 sa{};
      sa.pb = R.pb;
      sa.sp = R.ck[0].sp;
      sa.id = R.ck[0].id;
      rc = species_sums64(c, sa);
    } else {
      std::fill(c->adj_sums, c->adj_sums + 7 * EMO_NSPEC, 0.0);
    }
    // particle adjoints of the initial state, by id
    if (!rc && (xbar0 || vbar0 || gbar0) && np > 0) {
      std::vector<int64_t> ids(np);
      std::vector<double> h(np);
      EMO_HIP_CHECK_RET(c, ctx_d2h(c, ids.data(), R.ck[0].id, np * 8));
      auto put = [&](double *dst, const double *src, int stride, int comp) -> int {
        EMO_HIP_CHECK_RET(c, ctx_d2h(c, h.data(), src, np * 8));
        for (int64_t i = 0; i < np; i++) {
          int64_t k = ids[i];
          if (k >= 0 && k < np) dst[stride * k + comp] = h[i];
        }
        return EMO_OK;
      };
      for (int d = 0; d < 3 && !rc; d++) {
        if (xbar0) rc = put(xbar0, R.pb.x[d], 3, d);
        if (vbar0 && !rc) rc = put(vbar0, R.pb.v[d], 3, d);
      }
      if (gbar0 && !rc) rc = put(gbar0, R.pb.gam, 1, 0);
    }
  }
  // checkpointing of this call: K-step segments, each re-run once
  c->adj_info[0] = 0;
  c->adj_info[1] = R.nseg;
  c->adj_info[2] = (int64_t)nsteps + (int64_t)std::max(R.nseg - 1, 0) * R.K;
  c->adj_info[3] = R.nseg;
  // ---- restore the final state (the buffers stay with the context) --------
  hipStreamSynchronize(R.as);
  if (nsteps > 0 && fin_taken) {
    int rc2 = snap_restore(c, R.fin);
    if (!rc) rc = rc2;
  }
  if (!rc) {
    hipError_t e = hipStreamSynchronize(c->stream);
    if (e != hipSuccess) {
      c->err = hipGetErrorString(e);
      rc = EMO_ERR_HIP;
    }
  }
  for (hipEvent_t e : ev) hipEventDestroy(e);
  return rc;
}

extern "C" int emo_adjoint_stats(emo_ctx *c, int64_t *info4, double *sums) {
  if (!c) return EMO_ERR_CONFIG;
  if (c->adj_info[0] < 0) {
    c->err = "no gradient call has run on this context";
    return EMO_ERR_CONFIG;
  }
  if (info4)
    for (int k = 0; k < 4; k++) info4[k] = c->adj_info[k];
  if (sums)
    for (int k = 0; k < 7 * std::max(1, (int)c->cfg.nspecies) && k < 7 * EMO_NSPEC; k++) sums[k] = c->adj_sums[k];
  return EMO_OK;
}

extern "C" int emo_loader_dx_deps(const emo_loader *ld, const emo_config *cfg, double *dx_deps) {
  if (!ld || !cfg || !dx_deps) return EMO_ERR_CONFIG;
  std::vector<double> out;
  int rc = emo::loader_dx_deps(cfg, ld, out);
  if (rc) return rc;
  std::memcpy(dx_deps, out.data(), out.size() * sizeof(double));
  return EMO_OK;
}
