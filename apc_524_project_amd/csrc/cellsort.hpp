// cellsort.hpp -- the stable counting sort by destination cell shared by
// both precisions.  Particles are kept sorted by cell; a step moves each by
// at most one cell per dimension (periodic wraps included), which the push
// records as a move code.  From the codes:
//   far lists  an ordered list of the particles that moved further (rare;
//              two passes that return at once when there are none)
//   k_rank     each particle's rank within its (source cell, code) group, in
//              array order, and the group sizes
//   k_base     per destination cell, the offsets of its groups in ascending
//              source id (far movers merged in by source id) and its count
//   scan       the new cell ranges
// The caller's scatter then writes particle i to off[dst] + base + rank:
// the stable sort of the previous order by the new cell, deterministic and
// without radix passes.
#pragma once

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <string>

#include "emo_internal.hpp"
#include "movecode.hpp"

namespace emo {
namespace {

constexpr int kBlock = 256;
constexpr int kFarCap = 1 << 16;   // ordered far-mover list capacity per step
constexpr int kFarTile = 4096;

inline unsigned grid_for(long long n, int block = kBlock) {
  long long g = (n + block - 1) / block;
  return (unsigned)(g > 0 ? g : 1);
}

__device__ inline void decode_cell(const Geo &g, int cell, int *ix, int *iy, int *iz) {
  *iz = cell % g.nc[2];
  int t = cell / g.nc[2];
  *iy = t % g.nc[1];
  *ix = t / g.nc[1];
}

__device__ inline void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ inline unsigned long long lanes_below(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

__device__ inline int block_sum_256(int v, int *sh) {
  // sh: 4 ints of LDS
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) sh[wv] = v;
  __syncthreads();
  int t = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return t;
}

__global__ void __launch_bounds__(kBlock) k_far_count(const uint8_t *code, long long np, const int32_t *nfar,
                                                      int32_t *tiles) {
  __shared__ int sh[4];
  if (*nfar == 0) return;
  const long long b = (long long)blockIdx.x * kFarTile;
  int n = 0;
  for (int r = threadIdx.x; r < kFarTile; r += kBlock) {
    long long i = b + r;
    n += (i < np && code[i] == kFarCode) ? 1 : 0;
  }
  n = block_sum_256(n, sh);
  if (threadIdx.x == 0) tiles[blockIdx.x] = n;
}

// Ordered compaction of the far movers: entry (index, source cell,
// destination cell, 0) in ascending particle index.  CellPair gives a
// particle's source and destination cells.
template <class CellPair>
__global__ void __launch_bounds__(kBlock) k_far_compact(const uint8_t *code, CellPair cells, long long np,
                                                        const int32_t *nfar, const int32_t *tiles, int4 *far,
                                                        int far_cap, unsigned long long *flag) {
  __shared__ int sh[4];
  __shared__ int wsum[4];
  if (*nfar == 0) return;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  int base = 0;
  for (int t = threadIdx.x; t < (int)blockIdx.x; t += kBlock) base += tiles[t];
  base = block_sum_256(base, sh);
  const long long b = (long long)blockIdx.x * kFarTile;
  for (int r0 = 0; r0 < kFarTile; r0 += kBlock) {
    long long i = b + r0 + threadIdx.x;
    bool m = i < np && code[i] == kFarCode;
    unsigned long long bal = __ballot(m);
    if (lane == 0) wsum[wv] = __popcll(bal);
    __syncthreads();
    int wbase = 0;
    for (int q = 0; q < wv; q++) wbase += wsum[q];
    int total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
    if (m) {
      int pos = base + wbase + __popcll(bal & lanes_below(lane));
      if (pos < far_cap) {
        int src, dst;
        cells(i, src, dst);
        far[pos] = make_int4((int)i, src, dst, 0);
      } else {
        atomicOr(flag, kFlagFarMovers);
      }
    }
    base += total;
    __syncthreads();
  }
}

// One-dimensional source candidates of destination index t: the move codes
// in the order 3, 2, 1, 0, 4, 5, whose sources t - sigma ascend; a code is a
// candidate if it can occur and its source index lies in [1, n+1].  At most
// four are ever valid, so a cell has at most 64 (source cell, code) pairs.
struct Cand1 {
  int src[6];
  bool ok[6];
  int n;
};
__device__ inline void cand1(int t, int n, int per, Cand1 &o) {
  o.n = 0;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const int k = q == 0 ? 3 : q == 1 ? 2 : q == 2 ? 1 : q == 3 ? 0 : q == 4 ? 4 : 5;
    const int s = t - move_sigma(k, n);
    o.ok[q] = move_live(k, n, per) && s >= 1 && s <= n + 1;
    o.src[q] = s;
    o.n += o.ok[q] ? 1 : 0;
  }
}
// the a-th valid candidate (source index, code)
__device__ inline void cand1_pick(const Cand1 &o, int a, int &src, int &k) {
  int cnt = 0;
  src = 0;
  k = 0;
#pragma unroll
  for (int q = 0; q < 6; q++) {
    if (o.ok[q]) {
      if (cnt == a) {
        src = o.src[q];
        k = q == 0 ? 3 : q == 1 ? 2 : q == 2 ? 1 : q == 3 ? 0 : q == 4 ? 4 : 5;
      }
      cnt++;
    }
  }
}

constexpr int kCodes = 216;
constexpr int kRankCells = 8;   // source cells per wavefront in the rank pass
constexpr int kRankUnroll = 8;  // 64-particle chunks loaded before use
constexpr int kMaxGroup = 65535;

// Rank of each participating lane's key among the lanes of the wavefront
// with the same key (in lane order), offset by the running count ct[key]
// (LDS, one table per wavefront), which advances by the number of lanes
// holding the key.  One ballot per distinct key.
__device__ inline int wave_rank(int key, bool part, int *ct, int lane) {
  unsigned long long rem = __ballot(part);
  int myrank = 0;
  while (rem) {
    const int leader = __ffsll((long long)rem) - 1;
    const int kl = __builtin_amdgcn_readlane(key, leader);
    const bool mine = part && key == kl;
    const unsigned long long bal = __ballot(mine);
    const int base = ct[kl];
    if (mine) myrank = base + __popcll(bal & lanes_below(lane));
    wave_sync();
    if (lane == 0) ct[kl] = base + __popcll(bal);
    wave_sync();
    rem &= ~bal;
  }
  return myrank;
}

struct RankArgs {
  const uint8_t *code;
  const int32_t *off;  // ranges of the (sorted) source cells
  long long ncell;
  uint16_t *rank;
  uint16_t *nsk;  // [code * ncell + source]: particles per group
  unsigned long long *flag;
};

__global__ void __launch_bounds__(kBlock) k_rank(RankArgs a) {
  __shared__ int cnt[kBlock / 64][kRankCells * kCodes];
  __shared__ int bnd[kBlock / 64][kRankCells + 1];
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long s0 = ((long long)blockIdx.x * (kBlock / 64) + wv) * kRankCells;
  if (s0 >= a.ncell) return;
  const int nc = (int)min((long long)kRankCells, a.ncell - s0);
  int *ct = cnt[wv];
  for (int e = lane; e < nc * kCodes; e += 64) ct[e] = 0;
  if (lane <= nc) bnd[wv][lane] = a.off[s0 + lane];
  wave_sync();
  const int pb = bnd[wv][0], pe = bnd[wv][nc];
  for (int p0 = pb; p0 < pe; p0 += 64 * kRankUnroll) {
    uint8_t cc[kRankUnroll];
#pragma unroll
    for (int u = 0; u < kRankUnroll; u++) {
      const int p = p0 + u * 64 + lane;
      cc[u] = p < pe ? a.code[p] : kFarCode;
    }
#pragma unroll
    for (int u = 0; u < kRankUnroll; u++) {
      const int p = p0 + u * 64 + lane;
      if (p0 + u * 64 >= pe) break;  // uniform
      const bool valid = p < pe && cc[u] != kFarCode;
      int j = 0;
      for (int t = 1; t < nc; t++) j += bnd[wv][t] <= p ? 1 : 0;
      const int myrank = wave_rank(j * kCodes + cc[u], valid, ct, lane);
      if (valid) {
        if (myrank > kMaxGroup) atomicOr(a.flag, kFlagMigrate);
        a.rank[p] = (uint16_t)myrank;
      }
    }
  }
  wave_sync();
  for (int e = lane; e < nc * kCodes; e += 64) {
    const int v = ct[e];
    if (v) a.nsk[(long long)(e % kCodes) * a.ncell + s0 + e / kCodes] = (uint16_t)min(v, kMaxGroup);
  }
}

struct BaseArgs {
  Geo g;
  uint16_t *nsk;
  uint16_t *base;  // [code * ncell + source]
  int32_t *count;  // particles per destination cell
  int per[3];
  const int32_t *nfar;
  const int4 *far;
  int far_cap;
  int32_t *far_pos;
  unsigned long long *flag;
};

// One thread per destination cell: its (source, code) groups in ascending
// source id, far movers merged in by source id.  The group sizes of the
// candidates with one source x are loaded together (independent loads),
// then walked in order; 16 live sizes per thread keep the occupancy up.
__global__ void __launch_bounds__(kBlock) k_base(BaseArgs a) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const Geo &g = a.g;
  if (c >= g.ncell) return;
  int cx, cy, cz;
  decode_cell(g, (int)c, &cx, &cy, &cz);
  Cand1 X, Y, Z;
  cand1(cx, g.n[0], a.per[0], X);
  cand1(cy, g.n[1], a.per[1], Y);
  cand1(cz, g.n[2], a.per[2], Z);
  // the valid candidates of each dimension, in ascending source index
  int xs[4], xk[4], ys[4], yk[4], zs[4], zk[4];
#pragma unroll
  for (int q = 0; q < 4; q++) {
    cand1_pick(X, q, xs[q], xk[q]);
    cand1_pick(Y, q, ys[q], yk[q]);
    cand1_pick(Z, q, zs[q], zk[q]);
  }
  const uint16_t *__restrict__ nsk_in = a.nsk;
  const int nf = min(*a.nfar, a.far_cap);
  int fp = 0;
  int run = 0;
#pragma unroll
  for (int ia = 0; ia < 4; ia++) {
    if (ia >= X.n) break;
    int nn[16];
#pragma unroll
    for (int ib = 0; ib < 4; ib++)
#pragma unroll
      for (int ic = 0; ic < 4; ic++) {
        const int slot = ib * 4 + ic;
        nn[slot] = 0;
        if (ib < Y.n && ic < Z.n) {
          const int s = (xs[ia] * g.nc[1] + ys[ib]) * g.nc[2] + zs[ic];
          nn[slot] = nsk_in[(long long)((xk[ia] * 6 + yk[ib]) * 6 + zk[ic]) * g.ncell + s];
        }
      }
#pragma unroll
    for (int ib = 0; ib < 4; ib++)
#pragma unroll
      for (int ic = 0; ic < 4; ic++) {
        const int slot = ib * 4 + ic;
        if (!(ib < Y.n && ic < Z.n)) continue;
        const int s = (xs[ia] * g.nc[1] + ys[ib]) * g.nc[2] + zs[ic];
        const long long idx = (long long)((xk[ia] * 6 + yk[ib]) * 6 + zk[ic]) * g.ncell + s;
        while (fp < nf) {
          const int4 e = a.far[fp];
          if (e.y >= s) break;
          if (e.z == (int)c) a.far_pos[fp] = run++;
          fp++;
        }
        const int n = nn[slot];
        if (n) {
          a.nsk[idx] = 0;
          a.base[idx] = (uint16_t)min(run, kMaxGroup);
          run += n;
        }
      }
  }
  for (; fp < nf; fp++)
    if (a.far[fp].z == (int)c) a.far_pos[fp] = run++;
  if (run > kMaxGroup) atomicOr(a.flag, kFlagMigrate);
  a.count[c] = run;
}

__global__ void k_fill_i32(int32_t *p, int32_t v, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// FP64 reconstruction of positions and velocities (global-position output)

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
inline void cs_fill_i32(emo_ctx *c, int32_t *p, int32_t v, long long n) {
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, p, v, n);
}

inline int cellsort_alloc(emo_ctx *c, CellSortBufs &b, int64_t cap, long long ncell) {
  auto al = [&](void **p, size_t bytes) -> int {
    *p = nullptr;
    if (hipMalloc(p, std::max<size_t>(bytes, 16)) != hipSuccess) {
      *p = nullptr;
      c->err = "hipMalloc failed (cell sort buffers)";
      return EMO_ERR_ALLOC;
    }
    return EMO_OK;
  };
  int rc = 0;
  // (8 entries of padding: the FP32 inverse pass reads codes and ranks four
  // at a time from 4-aligned indices)
  if (!rc) rc = al((void **)&b.code, (size_t)cap + 8);
  if (!rc) rc = al((void **)&b.rank, ((size_t)cap + 8) * 2);
  if (!rc) rc = al((void **)&b.nsk, (size_t)kCodes * ncell * 2);
  if (!rc) rc = al((void **)&b.gbase, (size_t)kCodes * ncell * 2);
  if (!rc) rc = al((void **)&b.count, (size_t)(ncell + 1) * 4);
  if (!rc) rc = al((void **)&b.off, (size_t)(ncell + 1) * 4);
  if (!rc) rc = al((void **)&b.off_new, (size_t)(ncell + 1) * 4);
  if (!rc) rc = al((void **)&b.misc, 16);
  if (!rc) rc = al((void **)&b.far_tiles, (size_t)((cap + kFarTile - 1) / kFarTile + 1) * 4);
  if (!rc) rc = al((void **)&b.far, (size_t)kFarCap * sizeof(int4));
  if (!rc) rc = al((void **)&b.far_pos, (size_t)kFarCap * 4);
  if (!rc) rc = al((void **)&b.perm, (size_t)cap * 4);
  if (rc) return rc;
  if (ctx_zero(c, b.nsk, (size_t)kCodes * ncell * 2) != hipSuccess ||
      ctx_zero(c, b.count, (size_t)(ncell + 1) * 4) != hipSuccess ||
      ctx_zero(c, b.off, (size_t)(ncell + 1) * 4) != hipSuccess || ctx_zero(c, b.misc, 16) != hipSuccess) {
    c->err = "hipMemset failed (cell sort buffers)";
    return EMO_ERR_HIP;
  }
  b.scan_tmp_bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, b.scan_tmp_bytes, (const int32_t *)nullptr, (int32_t *)nullptr,
                                   (int)(ncell + 1));
  return al(&b.scan_tmp, b.scan_tmp_bytes);
}

inline void cellsort_free(CellSortBufs &b) {
  void *ps[] = {b.code, b.rank, b.nsk, b.gbase, b.count, b.off, b.off_new, b.misc,
                b.far_tiles, b.far, b.far_pos, b.perm, b.scan_tmp};
  for (void *p : ps)
    if (p) hipFree(p);
  b = CellSortBufs();
}

// exclusive scan of b.count into out
inline int cellsort_scan(emo_ctx *c, CellSortBufs &b, int32_t *out, const char *name) {
  size_t bytes = b.scan_tmp_bytes;
  TimedLaunch t(c, name);
  hipError_t e =
      hipcub::DeviceScan::ExclusiveSum(b.scan_tmp, bytes, b.count, out, (int)(c->geo.ncell + 1), c->stream);
  if (e != hipSuccess) {
    c->err = std::string("scan: ") + hipGetErrorString(e);
    return EMO_ERR_HIP;
  }
  return EMO_OK;
}

// the ordered far-mover list of the codes written by the push
template <class CellPair>
void cellsort_far(emo_ctx *c, CellSortBufs &b, long long np, CellPair cells, const char *name) {
  const unsigned ntiles = (unsigned)std::max<long long>((np + kFarTile - 1) / kFarTile, 1);
  TimedLaunch t(c, name);
  hipLaunchKernelGGL(k_far_count, dim3(ntiles), dim3(kBlock), 0, c->stream, (const uint8_t *)b.code, np,
                     (const int32_t *)b.misc, b.far_tiles);
  hipLaunchKernelGGL(k_far_compact<CellPair>, dim3(ntiles), dim3(kBlock), 0, c->stream, (const uint8_t *)b.code,
                     cells, np, (const int32_t *)b.misc, (const int32_t *)b.far_tiles, b.far, kFarCap, c->dflag);
}

inline void cellsort_periodic(const emo_ctx *c, int per[3]) {
  for (int d = 0; d < 3; d++)
    per[d] = c->cfg.parts_bound[2 * d] == EMO_BC_PERIODIC || c->cfg.parts_bound[2 * d + 1] == EMO_BC_PERIODIC;
}

// group offsets and the new ranges (b.off_new) from the group sizes b.nsk
// and the far-mover list
inline int cellsort_base_scan(emo_ctx *c, CellSortBufs &b, const char *base_name, const char *scan_name) {
  const Geo &g = c->geo;
  BaseArgs a;
  a.g = g;
  a.nsk = b.nsk;
  a.base = b.gbase;
  a.count = b.count;
  cellsort_periodic(c, a.per);
  a.nfar = b.misc;
  a.far = b.far;
  a.far_cap = kFarCap;
  a.far_pos = b.far_pos;
  a.flag = c->dflag;
  {
    TimedLaunch t(c, base_name);
    hipLaunchKernelGGL(k_base, dim3(grid_for(g.ncell)), dim3(kBlock), 0, c->stream, a);
  }
  return cellsort_scan(c, b, b.off_new, scan_name);
}

// far lists, ranks, group offsets and the new ranges (b.off_new) for the
// codes written by the push; b.off holds the current (source) ranges
template <class CellPair>
int cellsort_offsets(emo_ctx *c, CellSortBufs &b, long long np, CellPair cells, const char *const names[4]) {
  const Geo &g = c->geo;
  cellsort_far(c, b, np, cells, names[0]);
  {
    RankArgs r;
    r.code = b.code;
    r.off = b.off;
    r.ncell = g.ncell;
    r.rank = b.rank;
    r.nsk = b.nsk;
    r.flag = c->dflag;
    const long long waves = (g.ncell + kRankCells - 1) / kRankCells;
    TimedLaunch t(c, names[1]);
    hipLaunchKernelGGL(k_rank, dim3((unsigned)((waves + 3) / 4)), dim3(kBlock), 0, c->stream, r);
  }
  return cellsort_base_scan(c, b, names[2], names[3]);
}

// the position of particle i with move code k and destination cell dst
__device__ inline long long cellsort_pos(const Geo &g, const CellSortBufs &b, long long i, int k, int dst,
                                         unsigned long long *flag) {
  int within;
  if (k != kFarCode) {
    int dx, dy, dz;
    decode_cell(g, dst, &dx, &dy, &dz);
    const int kx = k / 36, ky = (k / 6) % 6, kz = k % 6;
    const int s = ((dx - move_sigma(kx, g.n[0])) * g.nc[1] + (dy - move_sigma(ky, g.n[1]))) * g.nc[2] +
                  (dz - move_sigma(kz, g.n[2]));
    if (s < 0 || s >= g.ncell) {
      atomicOr(flag, kFlagMigrate);
      return -1;
    }
    within = (int)b.gbase[(long long)k * g.ncell + s] + (int)b.rank[i];
  } else {
    const int nf = min(*b.misc, kFarCap);
    int lo = 0, hi = nf;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (b.far[mid].x < i) lo = mid + 1;
      else hi = mid;
    }
    if (lo >= nf || b.far[lo].x != i) {
      atomicOr(flag, kFlagFarMovers);
      return -1;
    }
    within = b.far_pos[lo];
  }
  const long long pos = (long long)b.off_new[dst] + within;
  if (pos >= b.off_new[dst + 1]) {
    atomicOr(flag, kFlagMigrate);
    return -1;
  }
  return pos;
}

}  // namespace
}  // namespace emo
