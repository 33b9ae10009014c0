// field_kernels.hip -- FP64 grid kernels for gfx950: Yee E/B stencils, ghost
// plane fold/refresh, LightBC plane fill, Jacobi sweeps, phiToE/AToB and the
// FP64 diagnostic reductions.
//
// All stencils read the padded (n+3)^3 arrays with z fastest, so a wavefront
// covers 64 consecutive k and every load coalesces.  Operation order follows
// the reference expressions exactly (built with -ffp-contract=off).
#include "emo_internal.hpp"

namespace emo {
namespace {

constexpr int kBlock = 256;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

// decode a linear index over the physical box [1..n]^3 (k fastest)
__device__ inline void phys_ijk(const Geo &g, long long t, int *i, int *j, int *k) {
  *k = 1 + (int)(t % g.n[2]);
  long long r = t / g.n[2];
  *j = 1 + (int)(r % g.n[1]);
  *i = 1 + (int)(r / g.n[1]);
}

// ---------------------------------------------------------------------------
// Fold, Grid::executeBC(-2) -> BC_F_Periodic::completeBC(-2)
// (src/boundaries/b_fields/bc_f_periodic.cpp:40-46): add the upper ghost
// plane n_d+1 of (J_j, J_k, rho) into plane 1 over the physical range of the
// other two dims (src/grid/spookyGrid.cpp:54-72, 286-425).
// ---------------------------------------------------------------------------
template <typename T>
struct PlaneArgs {
  Geo g;
  T *f[6];
  int nf;
  int d, src, dst, add;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_plane_copy(PlaneArgs<T> a) {
  const Geo &g = a.g;
  int d = a.d;
  int u = d == 0 ? 1 : 0;
  int w = d == 2 ? 1 : 2;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long np = (long long)g.n[u] * g.n[w];
  if (t >= np) return;
  int pu = 1 + (int)(t / g.n[w]);
  int pw = 1 + (int)(t % g.n[w]);
  int is[3], id[3];
  is[d] = a.src;
  id[d] = a.dst;
  is[u] = pu;
  id[u] = pu;
  is[w] = pw;
  id[w] = pw;
  long long ps = gidx(g, is[0], is[1], is[2]);
  long long pd = gidx(g, id[0], id[1], id[2]);
  for (int f = 0; f < a.nf; f++) {
    T val = a.f[f][ps];
    if (a.add) a.f[f][pd] += val;
    else a.f[f][pd] = val;
  }
}

// All periodic sides of one refresh in one launch: each side's ghost plane
// takes the opposite physical plane over the physical range of the other
// two dims.  No side reads a point another side writes (the sources are
// physical points, the destinations ghost points), so their order is free.
template <typename T>
struct SidesArgs {
  Geo g;
  T *f[6];
  int nf, nsides;
  int d[6], src[6], dst[6];
  long long start[7];  // thread ranges of the sides
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_plane_copy_sides(SidesArgs<T> a) {
  const Geo &g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.start[a.nsides]) return;
  int s = 0;
  while (t >= a.start[s + 1]) s++;
  t -= a.start[s];
  const int d = a.d[s];
  const int u = d == 0 ? 1 : 0;
  const int w = d == 2 ? 1 : 2;
  int is[3], id[3];
  is[d] = a.src[s];
  id[d] = a.dst[s];
  is[u] = id[u] = 1 + (int)(t / g.n[w]);
  is[w] = id[w] = 1 + (int)(t % g.n[w]);
  const long long ps = gidx(g, is[0], is[1], is[2]);
  const long long pd = gidx(g, id[0], id[1], id[2]);
  for (int f = 0; f < a.nf; f++) a.f[f][pd] = a.f[f][ps];
}

// The fold of every periodic dim in one launch.  A thread owns one point of
// the union of the planes x=1, y=1, z=1 (physical range) and adds the ghost
// values of the dims that fold onto it in the order d = 0, 1, 2, as the
// per-dim passes do; the ghost sources are never destinations.
template <typename T>
struct FoldArgs {
  Geo g;
  T *J[3];
  T *rho;
  int per[3];
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_fold_all(FoldArgs<T> a) {
  const Geo &g = a.g;
  const long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  int i, j, k;
  if (t < n1 * n2) {  // x = 1
    i = 1;
    j = 1 + (int)(t / n2);
    k = 1 + (int)(t % n2);
  } else if ((t -= n1 * n2) < (n0 - 1) * n2) {  // y = 1, x >= 2
    i = 2 + (int)(t / n2);
    j = 1;
    k = 1 + (int)(t % n2);
  } else if ((t -= (n0 - 1) * n2) < (n0 - 1) * (n1 - 1)) {  // z = 1, x, y >= 2
    i = 2 + (int)(t / (n1 - 1));
    j = 2 + (int)(t % (n1 - 1));
    k = 1;
  } else {
    return;
  }
  const long long p = gidx(g, i, j, k);
  const int at[3] = {i, j, k};
  // arrays folded along d: J of the two transverse components, and rho
  T *arr[4] = {a.J[0], a.J[1], a.J[2], a.rho};
  for (int q = 0; q < 4; q++) {
    if (!arr[q]) continue;
    T val = arr[q][p];
    bool touched = false;
    for (int d = 0; d < 3; d++) {
      if (!a.per[d] || at[d] != 1 || q == d) continue;
      int src[3] = {i, j, k};
      src[d] = g.n[d] + 1;
      val += arr[q][gidx(g, src[0], src[1], src[2])];
      touched = true;
    }
    if (touched) arr[q][p] = val;
  }
}

// Several plane fills in one launch (the six of one LightBC side write six
// different arrays).
struct FillMultiArgs {
  Geo g;
  PlaneFill pf[6];
  int npf;
  long long start[7];
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_plane_fill_multi(FillMultiArgs a) {
  const Geo &g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= a.start[a.npf]) return;
  int s = 0;
  while (t >= a.start[s + 1]) s++;
  t -= a.start[s];
  const PlaneFill &pf = a.pf[s];
  const int d = pf.dim;
  const int u = d == 0 ? 1 : 0;
  const int w = d == 2 ? 1 : 2;
  const int pu = (int)(t / g.nt[w]);
  const int pw = (int)(t % g.nt[w]);
  int ii[3];
  ii[d] = pf.idx;
  ii[u] = pu;
  ii[w] = pw;
  double v;
  if (pf.prof) {
    v = pf.base + pf.wave * pf.prof[(long long)pu * g.nt[w] + pw];
  } else {
    v = pf.base + pf.wave;
  }
  ((T *)pf.f)[gidx(g, ii[0], ii[1], ii[2])] = (T)v;
}

// ---------------------------------------------------------------------------
// Yee update, Grid::evolveFields (src/grid/oGrid.cpp:23-57): E on the
// physical points from B_tm1 and J; then B from B_tm1 and the new E.  Two
// kernels: the B kernel must see every new E, including the stale ghost
// planes that are only refreshed afterwards.
// ---------------------------------------------------------------------------
// The Yee grids are multiples of the 8 XCDs, which take blocks round-robin:
// XCD x runs the x-th eighth of the points (a contiguous run of rows and
// planes), so the neighbouring rows and planes a stencil reads are hits in
// its own L2.
__device__ inline long long yee_block() {
  const unsigned per = gridDim.x / 8;
  return (long long)(blockIdx.x % 8) * per + blockIdx.x / 8;
}
inline unsigned yee_grid(long long n) { return (unsigned)((grid_for(n) + 7) / 8 * 8); }

template <typename T>
struct YeeArgs {
  Geo g;
  T *E[3];
  T *B[3];
  const T *Bt[3];
  const T *J[3];
  T dt;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_evolve_E(YeeArgs<T> a) {
  const Geo &g = a.g;
  long long t = (long long)yee_block() * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long c = gidx(g, i, j, k);
  const long long sx = g.sx, sy = g.sy;
  const T dt = a.dt, dx_ = (T)g.d[0], dy_ = (T)g.d[1], dz_ = (T)g.d[2];
  const T fpi = (T)(4 * kPi);
  const T *Btx = a.Bt[0], *Bty = a.Bt[1], *Btz = a.Bt[2];
  a.E[0][c] = a.E[0][c] - dt * ((Bty[c] - Bty[c - 1]) / dz_ - (Btz[c] - Btz[c - sy]) / dy_ +
                                fpi * a.J[0][c]);
  a.E[1][c] = a.E[1][c] - dt * ((Btz[c] - Btz[c - sx]) / dx_ - (Btx[c] - Btx[c - 1]) / dz_ +
                                fpi * a.J[1][c]);
  a.E[2][c] = a.E[2][c] - dt * ((Btx[c] - Btx[c - sy]) / dy_ - (Bty[c] - Bty[c - sx]) / dx_ +
                                fpi * a.J[2][c]);
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_evolve_B(YeeArgs<T> a) {
  const Geo &g = a.g;
  long long t = (long long)yee_block() * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long c = gidx(g, i, j, k);
  const long long sx = g.sx, sy = g.sy;
  const T dt = a.dt, dx_ = (T)g.d[0], dy_ = (T)g.d[1], dz_ = (T)g.d[2];
  const T *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
  a.B[0][c] = a.Bt[0][c] + dt * ((Ey[c + 1] - Ey[c]) / dz_ - (Ez[c + sy] - Ez[c]) / dy_);
  a.B[1][c] = a.Bt[1][c] + dt * ((Ez[c + sx] - Ez[c]) / dx_ - (Ex[c + 1] - Ex[c]) / dz_);
  a.B[2][c] = a.Bt[2][c] + dt * ((Ex[c + sy] - Ex[c]) / dy_ - (Ey[c + sx] - Ey[c]) / dx_);
}

// ---------------------------------------------------------------------------
// LightBC plane fill: Grid::setFieldInPlane over the full (n+3)^2 plane
// (src/grid/oGrid.cpp:446-476) with the value of LightBC::applyBCs
// (src/grid/lightBC.cpp:165-198).  With a profile table the wave part is
// shaped by a transverse Gaussian (an extension; the reference is
// plane-uniform).
// ---------------------------------------------------------------------------
struct FillArgs {
  Geo g;
  PlaneFill pf;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_plane_fill(FillArgs a) {
  const Geo &g = a.g;
  const PlaneFill &pf = a.pf;
  int d = pf.dim;
  int u = d == 0 ? 1 : 0;
  int w = d == 2 ? 1 : 2;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.nt[u] * g.nt[w]) return;
  int pu = (int)(t / g.nt[w]);
  int pw = (int)(t % g.nt[w]);
  int ii[3];
  ii[d] = pf.idx;
  ii[u] = pu;
  ii[w] = pw;
  double v;
  if (pf.prof) {
    v = pf.base + pf.wave * pf.prof[(long long)pu * g.nt[w] + pw];
  } else {
    v = pf.base + pf.wave;
  }
  ((T *)pf.f)[gidx(g, ii[0], ii[1], ii[2])] = (T)v;
}

template <typename T>
__global__ void __launch_bounds__(kBlock) k_const_fill(T *f, T v, long long n) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) f[t] = v;
}

// ---------------------------------------------------------------------------
// Jacobi sweep, Poisson_Solver::run_poisson_solver_ (src/poisson/poisson.cpp:
// 223-240) over the physical points.
// ---------------------------------------------------------------------------
template <typename T>
struct JacobiArgs {
  Geo g;
  const T *src;
  T *dst;
  const T *R;
  T ax, ay, az, af, sourceMult;
};

template <typename T>
__global__ void __launch_bounds__(kBlock) k_jacobi(JacobiArgs<T> a) {
  const Geo &g = a.g;
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long c = gidx(g, i, j, k);
  const T *u = a.src;
  a.dst[c] = a.ax * (u[c - g.sx] + u[c + g.sx]) + a.ay * (u[c - g.sy] + u[c + g.sy]) +
             a.az * (u[c - 1] + u[c + 1]) - a.af * a.R[c] * a.sourceMult;
}

// The sweep fused with its residual, tiled for reuse: a block holds a
// kJacTK x kJacTJ (z, y) tile and marches over `ni` x planes; each thread
// keeps its column's x neighbours in registers and the tile of the current
// plane with its y/z halo in LDS, so each value of the iterate comes from
// HBM about once.  The arithmetic is k_jacobi's, in the same order.  Per
// block: max|dst - src| and sum (dst - src)^2 over its points in FP64
// (partials[2b], [2b + 1]), the residual of the sweep (the ghost refresh
// after it touches no physical point).
constexpr int kJacTK = 32, kJacTJ = kBlock / kJacTK;

template <typename T>
__global__ void __launch_bounds__(kBlock) k_jacobi_tiled(JacobiArgs<T> a, int ni, int nb, double *partials) {
  __shared__ T L[kJacTJ + 2][kJacTK + 2];
  __shared__ double smax[kBlock], ssum[kBlock];
  const Geo &g = a.g;
  const int tk = threadIdx.x % kJacTK, tj = threadIdx.x / kJacTK;
  const int tilesK = (g.n[2] + kJacTK - 1) / kJacTK, tilesJ = (g.n[1] + kJacTJ - 1) / kJacTJ;
  // the grid is a multiple of the 8 XCDs, which take blocks round-robin:
  // XCD x runs the x-th eighth of the tiles (whole runs of neighbouring
  // tiles and x chunks), so halo and chunk-edge planes are hits in its L2
  const int per = (int)gridDim.x / 8;
  const int t = (int)(blockIdx.x % 8) * per + (int)(blockIdx.x / 8);
  if (t >= nb) return;
  int b = t;
  const int bK = b % tilesK;
  b /= tilesK;
  const int bJ = b % tilesJ;
  const int chunk = b / tilesJ;
  const int k0 = 1 + bK * kJacTK, j0 = 1 + bJ * kJacTJ;
  const int kB = min(k0 + kJacTK - 1, g.n[2]), jB = min(j0 + kJacTJ - 1, g.n[1]);
  const int k = k0 + tk, j = j0 + tj;
  const bool in = k <= kB && j <= jB;
  const int i0 = 1 + chunk * ni, i1 = min(g.n[0], i0 + ni - 1);
  const long long sx = g.sx, sy = g.sy;
  // the halo values this thread loads for the tile (at most one per side)
  const bool hN = in && j == j0, hS = in && j == jB, hW = in && k == k0, hE = in && k == kB;
  long long c = gidx(g, i0, in ? j : j0, in ? k : k0);
  const T *u = a.src;
  // registers: plane i - 1 and i of this column, plane i's halo, plane i's
  // R; the next plane's values load one iteration ahead
  T up = in ? u[c - sx] : (T)0, uc = in ? u[c] : (T)0;
  T n0 = hN ? u[c - sy] : (T)0, s0 = hS ? u[c + sy] : (T)0, w0 = hW ? u[c - 1] : (T)0, e0 = hE ? u[c + 1] : (T)0;
  T rr = in ? a.R[c] : (T)0;
  T un = in ? u[c + sx] : (T)0;
  double mx = 0.0, sm = 0.0;
  for (int i = i0; i <= i1; i++) {
    // prefetch the next plane (its ghost plane at the chunk's end is valid)
    const long long cn = c + sx;
    T un2 = (T)0, n1 = (T)0, s1 = (T)0, w1 = (T)0, e1 = (T)0, rr1 = (T)0;
    if (i < i1 && in) {
      un2 = u[cn + sx];
      rr1 = a.R[cn];
      if (hN) n1 = u[cn - sy];
      if (hS) s1 = u[cn + sy];
      if (hW) w1 = u[cn - 1];
      if (hE) e1 = u[cn + 1];
    }
    __syncthreads();  // the previous plane's LDS reads are done
    if (in) {
      L[tj + 1][tk + 1] = uc;
      if (hN) L[0][tk + 1] = n0;
      if (hS) L[tj + 2][tk + 1] = s0;
      if (hW) L[tj + 1][0] = w0;
      if (hE) L[tj + 1][tk + 2] = e0;
    }
    __syncthreads();
    if (in) {
      const T v = a.ax * (up + un) + a.ay * (L[tj][tk + 1] + L[tj + 2][tk + 1]) +
                  a.az * (L[tj + 1][tk] + L[tj + 1][tk + 2]) - a.af * rr * a.sourceMult;
      a.dst[c] = v;
      const double dd = (double)v - (double)uc;
      const double ad = fabs(dd);
      mx = ad > mx ? ad : mx;
      sm += dd * dd;
    }
    up = uc;
    uc = un;
    un = un2;
    n0 = n1;
    s0 = s1;
    w0 = w1;
    e0 = e1;
    rr = rr1;
    c = cn;
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * t] = smax[0];
    partials[2 * t + 1] = ssum[0];
  }
}

// the block partials of a residual in block order: out[0] = max, out[1] = sum
__global__ void __launch_bounds__(kBlock) k_reduce_resid(const double *partials, int nb, double *out) {
  __shared__ double smax[kBlock], ssum[kBlock];
  double mx = 0.0, sm = 0.0;
  for (int b = threadIdx.x; b < nb; b += kBlock) {
    mx = fmax(mx, partials[2 * b]);
    sm += partials[2 * b + 1];
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int w = kBlock / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + w]);
      ssum[threadIdx.x] += ssum[threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = smax[0];
    out[1] = ssum[0];
  }
}

// max|u2-u1| (:250-260) and sum (u2-u1)^2 per block in FP64 (for either
// precision of u): partials[2*b], [2*b+1].  The max is exact in any order;
// the L2 partials are reduced in a fixed order by launch_reduce_partials.
template <typename T>
__global__ void __launch_bounds__(kBlock) k_residual(Geo g, const T *u1, const T *u2, double *partials) {
  __shared__ double smax[kBlock];
  __shared__ double ssum[kBlock];
  long long nphys = (long long)g.n[0] * g.n[1] * g.n[2];
  double mx = 0.0, sm = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nphys;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    phys_ijk(g, t, &i, &j, &k);
    long long c = gidx(g, i, j, k);
    double dd = (double)u2[c] - (double)u1[c];
    double ad = fabs(dd);
    mx = ad > mx ? ad : mx;
    sm += dd * dd;
  }
  smax[threadIdx.x] = mx;
  ssum[threadIdx.x] = sm;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      smax[threadIdx.x] = fmax(smax[threadIdx.x], smax[threadIdx.x + s]);
      ssum[threadIdx.x] += ssum[threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[2 * blockIdx.x] = smax[0];
    partials[2 * blockIdx.x + 1] = ssum[0];
  }
}

// phiToE, src/poisson/convertFields.cpp:20-67
template <typename T>
__global__ void __launch_bounds__(kBlock) k_phi_to_e(Geo g, const T *phi, T *Ex, T *Ey, T *Ez) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long p = gidx(g, i, j, k);
  Ex[p] = -(phi[p + g.sx] - phi[p]) / (T)g.d[0];
  Ey[p] = -(phi[p + g.sy] - phi[p]) / (T)g.d[1];
  Ez[p] = -(phi[p + 1] - phi[p]) / (T)g.d[2];
}

// AToB, src/poisson/convertFields.cpp:83-153 (not a curl: Q10)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_a_to_b(Geo g, const T *Ax, const T *Ay, const T *Az, T *Bx, T *By,
                                                   T *Bz) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)g.n[0] * g.n[1] * g.n[2]) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long p = gidx(g, i, j, k);
  const T dx = (T)g.d[0], dy = (T)g.d[1], dz = (T)g.d[2];
  // Bx: Ak=Az (hk=dz), Aj=Ay (hj=dy); nextValk=Ak[j+1], nextValj=Ak[k+1]
  Bx[p] = (Az[p + g.sy] - Az[p]) / dy + (Az[p + 1] - Ay[p]) / dz;
  // By: Ak=Ax (hk=dx), Aj=Az (hj=dz); nextValk=Ak[k+1], nextValj=Ak[i+1]
  By[p] = (Ax[p + 1] - Ax[p]) / dz + (Ax[p + g.sx] - Az[p]) / dx;
  // Bz: Ak=Ay (hk=dy), Aj=Ax (hj=dx); nextValk=Ak[i+1], nextValj=Ak[j+1]
  Bz[p] = (Ay[p + g.sx] - Ay[p]) / dx + (Ay[p + g.sy] - Ax[p]) / dy;
}

// FP64 sum of a field over the physical points: per-block partials
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sum_phys(Geo g, const T *f, double *partials) {
  __shared__ double sh[kBlock];
  long long nphys = (long long)g.n[0] * g.n[1] * g.n[2];
  double sm = 0.0;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nphys;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    phys_ijk(g, t, &i, &j, &k);
    sm += (double)f[gidx(g, i, j, k)];
  }
  sh[threadIdx.x] = sm;
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) partials[blockIdx.x] = sh[0];
}

// f -= mean over the physical points, the mean read from the device
template <typename T>
__global__ void __launch_bounds__(kBlock) k_sub_mean(Geo g, T *f, const double *sum) {
  long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long nphys = (long long)g.n[0] * g.n[1] * g.n[2];
  if (t >= nphys) return;
  int i, j, k;
  phys_ijk(g, t, &i, &j, &k);
  long long p = gidx(g, i, j, k);
  f[p] = (T)((double)f[p] - *sum / (double)nphys);
}

// ---------------------------------------------------------------------------
// Diagnostics: OutputBoxQuantities::output (src/IO/output.cpp:76-118).
// Per-block partial sums in a fixed grid-stride order, then a fixed-order
// final sum: deterministic run to run (not the reference's sequential
// order, so equal to it within FP64 rounding).
// ---------------------------------------------------------------------------
struct DiagPArgs {
  const double *v[3];
  const int32_t *sp;
  SpeciesParams s;
  long long np;
  double *partials;  // 5 per block
};

template <int NC>
__device__ void block_reduce_store(double (&vals)[NC], double *out) {
  __shared__ double sh[NC][kBlock];
#pragma unroll
  for (int q = 0; q < NC; q++) sh[q][threadIdx.x] = vals[q];
  __syncthreads();
  for (int s = kBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
#pragma unroll
      for (int q = 0; q < NC; q++) sh[q][threadIdx.x] += sh[q][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NC; q++) out[NC * blockIdx.x + q] = sh[q][0];
  }
}

__global__ void __launch_bounds__(kBlock) k_diag_particles(DiagPArgs a) {
  double acc[5] = {0, 0, 0, 0, 0};
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.np;
       i += (long long)gridDim.x * blockDim.x) {
    int sp = a.sp[i];
    double m = a.s.m[sp];
    double vx = a.v[0][i], vy = a.v[1][i], vz = a.v[2][i];
    acc[0] += m * vx;
    acc[1] += m * vy;
    acc[2] += m * vz;
    acc[3] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
    acc[4] += a.s.q[sp];
  }
  block_reduce_store<5>(acc, a.partials);
}

template <typename T>
struct DiagFArgs {
  Geo g;
  const T *E[3];
  const T *B[3];
  double *partials;  // 6 per block
};

#define EMO_SQR(x) (x * x)
template <typename T>
__global__ void __launch_bounds__(kBlock) k_diag_fields(DiagFArgs<T> a) {
  const Geo &g = a.g;
  double acc[6] = {0, 0, 0, 0, 0, 0};
  double vol = g.d[0] * g.d[1] * g.d[2];
  long long nphys = (long long)g.n[0] * g.n[1] * g.n[2];
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < nphys;
       t += (long long)gridDim.x * blockDim.x) {
    int i, j, k;
    phys_ijk(g, t, &i, &j, &k);
    long long p = gidx(g, i, j, k);
    acc[0] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.E[0][p] + (double)a.E[0][p + g.sx]));
    acc[1] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.E[1][p] + (double)a.E[1][p + g.sy]));
    acc[2] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.E[2][p] + (double)a.E[2][p + 1]));
    acc[3] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.B[0][p] + (double)a.B[0][p + g.sx]));
    acc[4] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.B[1][p] + (double)a.B[1][p + g.sy]));
    acc[5] += 0.5 * vol * EMO_SQR(0.5 * ((double)a.B[2][p] + (double)a.B[2][p + 1]));
  }
  block_reduce_store<6>(acc, a.partials);
}
#undef EMO_SQR

// one block: out[q] = sum over blocks b (ascending, fixed tree) of partials[ncomp*b+q]
__global__ void __launch_bounds__(kBlock) k_reduce_partials(const double *partials, int nblocks,
                                                            int ncomp, double *out) {
  __shared__ double sh[kBlock];
  for (int q = 0; q < ncomp; q++) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += kBlock) s += partials[ncomp * b + q];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int st = kBlock / 2; st > 0; st >>= 1) {
      if ((int)threadIdx.x < st) sh[threadIdx.x] += sh[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[q] = sh[0];
    __syncthreads();
  }
}

}  // namespace

template <typename T>
void fk_plane_copy(emo_ctx *c, T *const *fields, int nf, int d, int src, int dst, int add) {
  PlaneArgs<T> a;
  a.g = c->geo;
  for (int q = 0; q < nf; q++) a.f[q] = fields[q];
  a.nf = nf;
  a.d = d;
  a.src = src;
  a.dst = dst;
  a.add = add;
  int u = d == 0 ? 1 : 0, w = d == 2 ? 1 : 2;
  hipLaunchKernelGGL(k_plane_copy<T>, dim3(grid_for((long long)c->geo.n[u] * c->geo.n[w])),
                     dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_plane_fill(emo_ctx *c, const PlaneFill &pf) {
  FillArgs a;
  a.g = c->geo;
  a.pf = pf;
  int d = pf.dim;
  int u = d == 0 ? 1 : 0, w = d == 2 ? 1 : 2;
  hipLaunchKernelGGL(k_plane_fill<T>, dim3(grid_for((long long)c->geo.nt[u] * c->geo.nt[w])),
                     dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_plane_fill_multi(emo_ctx *c, const PlaneFill *pf, int n) {
  if (n <= 0) return;
  FillMultiArgs a;
  a.g = c->geo;
  a.npf = n;
  a.start[0] = 0;
  for (int s = 0; s < n; s++) {
    a.pf[s] = pf[s];
    const int d = pf[s].dim;
    const int u = d == 0 ? 1 : 0, w = d == 2 ? 1 : 2;
    a.start[s + 1] = a.start[s] + (long long)c->geo.nt[u] * c->geo.nt[w];
  }
  hipLaunchKernelGGL(k_plane_fill_multi<T>, dim3(grid_for(a.start[n])), dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_plane_copy_sides(emo_ctx *c, T *const *fields, int nf, int nsides, const int *d, const int *src,
                         const int *dst) {
  if (nsides <= 0) return;
  SidesArgs<T> a;
  a.g = c->geo;
  for (int q = 0; q < nf; q++) a.f[q] = fields[q];
  a.nf = nf;
  a.nsides = nsides;
  a.start[0] = 0;
  for (int s = 0; s < nsides; s++) {
    a.d[s] = d[s];
    a.src[s] = src[s];
    a.dst[s] = dst[s];
    const int u = d[s] == 0 ? 1 : 0, w = d[s] == 2 ? 1 : 2;
    a.start[s + 1] = a.start[s] + (long long)c->geo.n[u] * c->geo.n[w];
  }
  hipLaunchKernelGGL(k_plane_copy_sides<T>, dim3(grid_for(a.start[nsides])), dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_const_fill(emo_ctx *c, T *f, double v, long long n) {
  hipLaunchKernelGGL(k_const_fill<T>, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, f, (T)v, n);
}

template <typename T>
void fk_fold(emo_ctx *c, FieldSet<T> &fs) {
  const Geo &g = c->geo;
  FoldArgs<T> a;
  a.g = g;
  int any = 0;
  for (int d = 0; d < 3; d++) {
    a.J[d] = fs.J[d];
    a.per[d] = c->cfg.fields_bound[2 * d] == EMO_BC_PERIODIC;
    any |= a.per[d];
  }
  a.rho = fs.rho;
  if (!any) return;
  const long long n0 = g.n[0], n1 = g.n[1], n2 = g.n[2];
  const long long pts = n1 * n2 + (n0 - 1) * n2 + (n0 - 1) * (n1 - 1);
  hipLaunchKernelGGL(k_fold_all<T>, dim3(grid_for(pts)), dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_evolve(emo_ctx *c, FieldSet<T> &fs, double dt) {
  for (int q = 0; q < 3; q++) std::swap(fs.B[q], fs.Bt[q]);
  YeeArgs<T> a;
  a.g = c->geo;
  for (int q = 0; q < 3; q++) {
    a.E[q] = fs.E[q];
    a.B[q] = fs.B[q];
    a.Bt[q] = fs.Bt[q];
    a.J[q] = fs.J[q];
  }
  a.dt = (T)dt;
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  const bool f64 = sizeof(T) == 8;
  {
    TimedLaunch t(c, f64 ? "yee_E_f64" : "yee_E_f32");
    hipLaunchKernelGGL(k_evolve_E<T>, dim3(yee_grid(nphys)), dim3(kBlock), 0, c->stream, a);
  }
  TimedLaunch t(c, f64 ? "yee_B_f64" : "yee_B_f32");
  hipLaunchKernelGGL(k_evolve_B<T>, dim3(yee_grid(nphys)), dim3(kBlock), 0, c->stream, a);
}

template <typename T>
void fk_jacobi(emo_ctx *c, const T *src, T *dst, const T *R, double ax, double ay, double az, double af,
               double sourceMult) {
  JacobiArgs<T> a;
  a.g = c->geo;
  a.src = src;
  a.dst = dst;
  a.R = R;
  a.ax = (T)ax;
  a.ay = (T)ay;
  a.az = (T)az;
  a.af = (T)af;
  a.sourceMult = (T)sourceMult;
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  TimedLaunch t(c, sizeof(T) == 8 ? "jacobi_f64" : "jacobi_f32");
  hipLaunchKernelGGL(k_jacobi<T>, dim3(grid_for(nphys)), dim3(kBlock), 0, c->stream, a);
}

// One sweep src -> dst and its residual (max|dst - src|, sum (dst - src)^2
// over the physical points) in partials[2 * kMaxResidualBlocks + 0, 1];
// partials holds 2 * kMaxResidualBlocks + 2 doubles.
template <typename T>
void fk_jacobi_resid(emo_ctx *c, const T *src, T *dst, const T *R, double ax, double ay, double az, double af,
                     double sourceMult, double *partials) {
  const Geo &g = c->geo;
  const long long tiles = (long long)((g.n[2] + kJacTK - 1) / kJacTK) * ((g.n[1] + kJacTJ - 1) / kJacTJ);
  double *out = partials + 2 * kMaxResidualBlocks;
  if (tiles <= kMaxResidualBlocks) {
    JacobiArgs<T> a;
    a.g = g;
    a.src = src;
    a.dst = dst;
    a.R = R;
    a.ax = (T)ax;
    a.ay = (T)ay;
    a.az = (T)az;
    a.af = (T)af;
    a.sourceMult = (T)sourceMult;
    // x planes per block: at least 8, at most kMaxResidualBlocks blocks
    const long long chunks_max = std::max<long long>(1, kMaxResidualBlocks / tiles);
    const int ni = (int)std::max<long long>(8, (g.n[0] + chunks_max - 1) / chunks_max);
    const int chunks = (g.n[0] + ni - 1) / ni;
    const int nb = (int)(tiles * chunks);
    {
      TimedLaunch t(c, sizeof(T) == 8 ? "jacobi_f64" : "jacobi_f32");
      hipLaunchKernelGGL(k_jacobi_tiled<T>, dim3((nb + 7) / 8 * 8), dim3(kBlock), 0, c->stream, a, ni, nb,
                         partials);
    }
    hipLaunchKernelGGL(k_reduce_resid, dim3(1), dim3(kBlock), 0, c->stream, (const double *)partials, nb, out);
    return;
  }
  // very wide transverse grids: the sweep, then the residual pass
  fk_jacobi<T>(c, src, dst, R, ax, ay, az, af, sourceMult);
  const int nb = residual_blocks(c);
  hipLaunchKernelGGL(k_residual<T>, dim3(nb), dim3(kBlock), 0, c->stream, g, (const T *)src, (const T *)dst,
                     partials);
  hipLaunchKernelGGL(k_reduce_resid, dim3(1), dim3(kBlock), 0, c->stream, (const double *)partials, nb, out);
}

template <typename T>
void fk_residual(emo_ctx *c, const T *u1, const T *u2, double *partials, int nblocks) {
  hipLaunchKernelGGL(k_residual<T>, dim3(nblocks), dim3(kBlock), 0, c->stream, c->geo, u1, u2, partials);
}

template <typename T>
void fk_phi_to_e(emo_ctx *c, const T *phi, T *const *E) {
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  hipLaunchKernelGGL(k_phi_to_e<T>, dim3(grid_for(nphys)), dim3(kBlock), 0, c->stream, c->geo, phi, E[0], E[1],
                     E[2]);
}

template <typename T>
void fk_a_to_b(emo_ctx *c, const T *Ax, const T *Ay, const T *Az, T *const *B) {
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  hipLaunchKernelGGL(k_a_to_b<T>, dim3(grid_for(nphys)), dim3(kBlock), 0, c->stream, c->geo, Ax, Ay, Az, B[0],
                     B[1], B[2]);
}

template <typename T>
void fk_subtract_mean(emo_ctx *c, T *f, double *partials, int nblocks, double *sum_out) {
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  hipLaunchKernelGGL(k_sum_phys<T>, dim3(nblocks), dim3(kBlock), 0, c->stream, c->geo, f, partials);
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, partials, nblocks, 1, sum_out);
  hipLaunchKernelGGL(k_sub_mean<T>, dim3(grid_for(nphys)), dim3(kBlock), 0, c->stream, c->geo, f,
                     (const double *)sum_out);
}

template <typename T>
void fk_diag_fields(emo_ctx *c, const FieldSet<T> &fs, double *partials, int nblocks) {
  DiagFArgs<T> a;
  a.g = c->geo;
  for (int q = 0; q < 3; q++) {
    a.E[q] = fs.E[q];
    a.B[q] = fs.Bfix[q];  // Q23: the diagnostics read Grid::fieldPtr_
  }
  a.partials = partials;
  hipLaunchKernelGGL(k_diag_fields<T>, dim3(nblocks), dim3(kBlock), 0, c->stream, a);
}

#define EMO_INSTANTIATE(T)                                                                       \
  template void fk_plane_copy<T>(emo_ctx *, T *const *, int, int, int, int, int);                \
  template void fk_plane_fill<T>(emo_ctx *, const PlaneFill &);                                  \
  template void fk_plane_fill_multi<T>(emo_ctx *, const PlaneFill *, int);                       \
  template void fk_plane_copy_sides<T>(emo_ctx *, T *const *, int, int, const int *, const int *, \
                                       const int *);                                             \
  template void fk_const_fill<T>(emo_ctx *, T *, double, long long);                             \
  template void fk_fold<T>(emo_ctx *, FieldSet<T> &);                                            \
  template void fk_evolve<T>(emo_ctx *, FieldSet<T> &, double);                                  \
  template void fk_diag_fields<T>(emo_ctx *, const FieldSet<T> &, double *, int);                \
  template void fk_jacobi<T>(emo_ctx *, const T *, T *, const T *, double, double, double, double, double); \
  template void fk_residual<T>(emo_ctx *, const T *, const T *, double *, int);                  \
  template void fk_jacobi_resid<T>(emo_ctx *, const T *, T *, const T *, double, double, double, double, double, \
                                   double *);                                                    \
  template void fk_phi_to_e<T>(emo_ctx *, const T *, T *const *);                                \
  template void fk_a_to_b<T>(emo_ctx *, const T *, const T *, const T *, T *const *);            \
  template void fk_subtract_mean<T>(emo_ctx *, T *, double *, int, double *);
EMO_INSTANTIATE(double)
EMO_INSTANTIATE(float)
#undef EMO_INSTANTIATE

FieldSet<double> f64_view(emo_ctx *c) {
  FieldSet<double> v;
  for (int q = 0; q < 3; q++) {
    v.E[q] = c->E[q];
    v.B[q] = c->B[q];
    v.Bt[q] = c->Bt[q];
    v.Bfix[q] = c->Bfix[q];
    v.J[q] = c->J[q];
  }
  v.rho = c->rho;
  return v;
}

void launch_fold_f64(emo_ctx *c) {
  FieldSet<double> v = f64_view(c);
  fk_fold<double>(c, v);
}

void launch_evolve_f64(emo_ctx *c, double dt) {
  FieldSet<double> v = f64_view(c);
  fk_evolve<double>(c, v, dt);
  for (int q = 0; q < 3; q++) {
    c->B[q] = v.B[q];
    c->Bt[q] = v.Bt[q];
  }
}

void launch_plane_copy_f64(emo_ctx *c, double *const *fields, int nf, int d, int src, int dst,
                           int add) {
  fk_plane_copy<double>(c, fields, nf, d, src, dst, add);
}

void launch_plane_fill_f64(emo_ctx *c, const PlaneFill &pf) { fk_plane_fill<double>(c, pf); }

void launch_const_fill_f64(emo_ctx *c, double *f, double v, long long n) {
  fk_const_fill<double>(c, f, v, n);
}

void launch_jacobi_f64(emo_ctx *c, const double *src, double *dst, const double *R, double ax,
                       double ay, double az, double af, double sourceMult) {
  fk_jacobi<double>(c, src, dst, R, ax, ay, az, af, sourceMult);
}

void launch_jacobi_resid_f64(emo_ctx *c, const double *src, double *dst, const double *R, double ax, double ay,
                             double az, double af, double sourceMult, double *partials) {
  fk_jacobi_resid<double>(c, src, dst, R, ax, ay, az, af, sourceMult, partials);
}

int residual_blocks(emo_ctx *c) {
  long long nphys = (long long)c->geo.n[0] * c->geo.n[1] * c->geo.n[2];
  long long b = (nphys + kBlock - 1) / kBlock;
  return (int)(b < kMaxResidualBlocks ? (b > 0 ? b : 1) : kMaxResidualBlocks);
}

void launch_residual_f64(emo_ctx *c, const double *u1, const double *u2, double *partials,
                         int nblocks) {
  fk_residual<double>(c, u1, u2, partials, nblocks);
}

void launch_phi_to_e_f64(emo_ctx *c, const double *phi) { fk_phi_to_e<double>(c, phi, c->E); }

void launch_a_to_b_f64(emo_ctx *c, const double *Ax, const double *Ay, const double *Az) {
  fk_a_to_b<double>(c, Ax, Ay, Az, c->B);
}

void launch_diag_particles_f64(emo_ctx *c, double *partials, int nblocks) {
  DiagPArgs a;
  for (int q = 0; q < 3; q++) a.v[q] = c->p.v[q];
  a.sp = c->p.sp;
  a.s = c->species;
  a.np = c->np;
  a.partials = partials;
  hipLaunchKernelGGL(k_diag_particles, dim3(nblocks), dim3(kBlock), 0, c->stream, a);
}

void launch_diag_fields_f64(emo_ctx *c, double *partials, int nblocks) {
  FieldSet<double> v = f64_view(c);
  fk_diag_fields<double>(c, v, partials, nblocks);
}

void launch_reduce_partials(emo_ctx *c, const double *partials, int nblocks, int ncomp,
                            double *out) {
  hipLaunchKernelGGL(k_reduce_partials, dim3(1), dim3(kBlock), 0, c->stream, partials, nblocks,
                     ncomp, out);
}

}  // namespace emo
