// deposit_grid.hpp -- second stage of the deterministic deposit, shared by
// both precisions: every padded grid point sums the per-cell 12-edge (J) and
// 8-vertex (rho) vectors that Grid::addJ / Grid::addRho
// (src/grid/oGrid.cpp:91-139) would add to it, in ascending cell-id order,
// starting from zero (Grid::constJ / constRho).  Includes addJ's duplicate
// Jx[ix][iy+1][iz] target (entries 1 and 3) and its never-written
// Jx[ix][iy][iz+1].  Cell vectors are component-major: v[q * stride + cell]
// (sj for the J vectors, sr for the rho vectors, at least ncell).
#pragma once

#include "emo_internal.hpp"

namespace emo {

template <typename T>
struct DepGridArgs {
  Geo g;
  const T *cj, *cr;
  long long sj, sr;  // component strides of cj and cr
  T *J[3], *rho;
  int do_rho;
  // optional (FP32 step): base[c] = base_new[c] for c <= ncell when lay[0]
  // is set, and *nfar = 0
  int32_t *commit_base = nullptr;
  const int32_t *commit_new = nullptr, *commit_lay = nullptr;
  int32_t *nfar = nullptr;
};

__device__ inline bool dg_cell_ok(const Geo &g, int i, int j, int k) {
  return i >= 0 && j >= 0 && k >= 0 && i < g.nc[0] && j < g.nc[1] && k < g.nc[2];
}
__device__ inline long long dg_cid(const Geo &g, int i, int j, int k) {
  return ((long long)i * g.nc[1] + j) * g.nc[2] + k;
}

template <typename T>
__global__ void __launch_bounds__(256) k_deposit_grid(DepGridArgs<T> a) {
  long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Geo &g = a.g;
  if (a.commit_base && p <= g.ncell && a.commit_lay[0] != 0) a.commit_base[p] = a.commit_new[p];
  if (a.nfar && p == 0) *a.nfar = 0;
  if (p >= g.ntot) return;
  int k = (int)(p % g.nt[2]);
  int j = (int)((p / g.nt[2]) % g.nt[1]);
  int i = (int)(p / ((long long)g.nt[1] * g.nt[2]));
  const long long N = a.sj;
  const T *cj = a.cj;
  T acc;
  // Jx(i,j,k) <- J2(i,j-1,k-1), J1(i,j-1,k), J3(i,j-1,k), J0(i,j,k)
  acc = 0;
  if (dg_cell_ok(g, i, j - 1, k - 1)) acc += cj[2 * N + dg_cid(g, i, j - 1, k - 1)];
  if (dg_cell_ok(g, i, j - 1, k)) {
    acc += cj[1 * N + dg_cid(g, i, j - 1, k)];
    acc += cj[3 * N + dg_cid(g, i, j - 1, k)];
  }
  if (dg_cell_ok(g, i, j, k)) acc += cj[0 * N + dg_cid(g, i, j, k)];
  a.J[0][p] = acc;
  // Jy(i,j,k) <- J6(i-1,j,k-1), J7(i-1,j,k), J5(i,j,k-1), J4(i,j,k)
  acc = 0;
  if (dg_cell_ok(g, i - 1, j, k - 1)) acc += cj[6 * N + dg_cid(g, i - 1, j, k - 1)];
  if (dg_cell_ok(g, i - 1, j, k)) acc += cj[7 * N + dg_cid(g, i - 1, j, k)];
  if (dg_cell_ok(g, i, j, k - 1)) acc += cj[5 * N + dg_cid(g, i, j, k - 1)];
  if (dg_cell_ok(g, i, j, k)) acc += cj[4 * N + dg_cid(g, i, j, k)];
  a.J[1][p] = acc;
  // Jz(i,j,k) <- J10(i-1,j-1,k), J9(i-1,j,k), J11(i,j-1,k), J8(i,j,k)
  acc = 0;
  if (dg_cell_ok(g, i - 1, j - 1, k)) acc += cj[10 * N + dg_cid(g, i - 1, j - 1, k)];
  if (dg_cell_ok(g, i - 1, j, k)) acc += cj[9 * N + dg_cid(g, i - 1, j, k)];
  if (dg_cell_ok(g, i, j - 1, k)) acc += cj[11 * N + dg_cid(g, i, j - 1, k)];
  if (dg_cell_ok(g, i, j, k)) acc += cj[8 * N + dg_cid(g, i, j, k)];
  a.J[2][p] = acc;
  if (a.do_rho) {
    // rho(i,j,k) <- R5(i-1,j-1,k-1), R2(i-1,j-1,k), R6(i-1,j,k-1), R1(i-1,j,k),
    //               R4(i,j-1,k-1), R3(i,j-1,k), R7(i,j,k-1), R0(i,j,k)
    const T *cr = a.cr;
    const long long N = a.sr;
    acc = 0;
    if (dg_cell_ok(g, i - 1, j - 1, k - 1)) acc += cr[5 * N + dg_cid(g, i - 1, j - 1, k - 1)];
    if (dg_cell_ok(g, i - 1, j - 1, k)) acc += cr[2 * N + dg_cid(g, i - 1, j - 1, k)];
    if (dg_cell_ok(g, i - 1, j, k - 1)) acc += cr[6 * N + dg_cid(g, i - 1, j, k - 1)];
    if (dg_cell_ok(g, i - 1, j, k)) acc += cr[1 * N + dg_cid(g, i - 1, j, k)];
    if (dg_cell_ok(g, i, j - 1, k - 1)) acc += cr[4 * N + dg_cid(g, i, j - 1, k - 1)];
    if (dg_cell_ok(g, i, j - 1, k)) acc += cr[3 * N + dg_cid(g, i, j - 1, k)];
    if (dg_cell_ok(g, i, j, k - 1)) acc += cr[7 * N + dg_cid(g, i, j, k - 1)];
    if (dg_cell_ok(g, i, j, k)) acc += cr[0 * N + dg_cid(g, i, j, k)];
    a.rho[p] = acc;
  }
}

}  // namespace emo
