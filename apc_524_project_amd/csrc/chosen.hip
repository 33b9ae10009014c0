// chosen.hip -- the FP32 chosen form of the PIC step for gfx950.
//
// Particles are kept sorted by cell as two arrays of 16-byte records:
//   A = {word, fx, fy, fz}: word holds the species in the top 6 bits and the
//       reference's ghost-inclusive cell id ((ny+2)(nz+2) ix + (nz+2) iy + iz,
//       src/grid/oGrid.cpp:208-222) in the low 26; f in [0,1)^3 is the
//       in-cell offset (FP32);
//   B = {ux, uy, uz, id}: u (FP32) is v for Boris and gamma*v for the
//       relativistic pusher; id is the particle's int32 id (-1 if none).
// 32 bytes per particle, each record one 16-byte vector access.  Cell c holds
// the particles [off[c], off[c+1]).
//
// The algorithm is the reference's (same step order, same defects); only the
// representation and precision change.  With x = x0 + dx (ix - 1 + f) the
// reference's "least vertex" (src/grid/oGrid.cpp:160-162, 239-241) gives
// relpos / l = f - 1, so its linear weights (l - r)/l, r/l become 2 - f and
// f - 1 (src/particles/interpolate.cpp:25-33, src/particles/deposit.cpp:94-97).
//
// Per step, with the state double-buffered (old -> new):
//   cellvec_f32       per cell, the 18 values the gather reads (E on the 4
//                     edges of each component, (B + B_tm1)/2 on the 2 faces):
//                     five 16-byte loads per particle instead of 24 scattered
//                     ones
//   push_rank_f32     source order, one wavefront per 4 source cells: gather
//                     + push + particle BCs of every particle, which yield its
//                     move code (which of the <= 6 x 6 x 6 one-cell moves,
//                     periodic wraps included, took it to its new cell) and
//                     its rank among the particles of its (source cell, code)
//                     group; the group sizes.  Nothing else is written.
//   far_f32           (only when some particle moved further than one cell)
//                     an ordered list of those particles
//   base_f32, scan    per destination cell, the offsets of its (source,
//                     code) groups in ascending source id, its count, the new
//                     cell ranges
//   inv_f32           source order: new position off[dst] + base + rank of
//                     every particle -> the inverse permutation
//   move_deposit_f32  destination order, 16 lanes per cell: each new slot
//                     takes its particle from the old state, repeats the
//                     push of push_rank_f32 (the same function on the same
//                     inputs, so the same result bit for bit; a disagreement
//                     on the cell would be reported), writes the new state
//                     and sums the 12 edge currents with compensated
//                     (two-float) accumulation; a fixed butterfly combines
//                     the lanes: deterministic, no atomics
//   jsum_f32          per grid point, the fixed-order sum of the cell vectors
//   fold, Yee E/B, ghost refresh   shared FP32 instances of the grid kernels
// The permutation is the stable counting sort of the previous order by the
// new cell.  Repeating the push costs arithmetic but saves writing and
// re-reading an intermediate state: particle traffic is 32 B read in the
// first pass and 4 + 32 + 32 B in the last.  Both particle passes run
// software-pipelined (the next particle's loads are in flight during a
// push), with the per-species constants in LDS, and are specialised for
// all-periodic particle BCs.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "cellsort.hpp"
#include "deposit_grid.hpp"
#include "emo_internal.hpp"
#include "uniform_load.hpp"
#include "chosen_math.hpp"

namespace emo {

struct ChosenState {
  int64_t cap = 0;
  int4 *pa = nullptr, *pa_alt = nullptr;  // {word, f[3]}
  int4 *pb = nullptr, *pb_alt = nullptr;  // {u[3], id}
  bool have_id = false;
  float *g0 = nullptr;  // stored gamma for the first relativistic push (Q3)
  bool g0_active = false;
  bool g0_owned = true;  // g0 is this state's allocation (not a snapshot's)
  // the step deposits J only; rho (an output, not read by the step) is
  // deposited from the current particles when it is read
  bool rho_stale = false;
  // non-relativistic forms: the loaded gamma by particle id.  No Boris step
  // reads or updates it (the reference keeps the loader's value, Q3); it is
  // only handed back.  Kept in id order, so the sort never moves it.
  float *gam_id = nullptr;
  CellSortBufs srt;  // the counting sort's buffers (cellsort.hpp); srt.perm is the inverse permutation
  int32_t *fdst = nullptr;  // destination cell of a far mover
  // initial sort of loaded particles (arbitrary order)
  uint32_t *key = nullptr, *key_alt = nullptr, *perm = nullptr, *perm_alt = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int sort_bits = 0;
  float *cj = nullptr;  // 12 per cell, component-major
  float4 *cv = nullptr;  // gather vectors, kCellVec per cell
  FieldSet<float> fs;
  float *Btfix[3] = {};
};

namespace {

// Work-groups go to the 8 XCDs round-robin (block b to XCD b % 8).  This
// renumbers them so that each XCD takes a contiguous eighth of the tiles,
// and its L2 holds only the gather vectors and cell ranges of its own part
// of the grid.  The grid is a multiple of 8 blocks; blocks past the last
// tile find no work.
constexpr unsigned kXcds = 8;
__device__ inline unsigned xcd_tile() {
  const unsigned per = gridDim.x / kXcds;
  return (blockIdx.x % kXcds) * per + blockIdx.x / kXcds;
}
inline unsigned xcd_grid(unsigned blocks) { return (blocks + kXcds - 1) / kXcds * kXcds; }

// The gather vector of a cell (Grid::getFieldInterpolatorVec,
// src/grid/oGrid.cpp:153-197, without the vertex): the 4 edge values of each
// E component and the 2 face values of each (B + B_tm1)/2 component, as five
// 16-byte vectors:
//   {Ex(i,j,k), Ex(i,j+1,k), Ex(i,j+1,k+1), Ex(i,j,k+1)}
//   {Ey(i,j,k), Ey(i,j,k+1), Ey(i+1,j,k+1), Ey(i+1,j,k)}
//   {Ez(i,j,k), Ez(i+1,j,k), Ez(i+1,j+1,k), Ez(i,j+1,k)}
//   {Bx(i), Bx(i+1), By(j), By(j+1)}   (face averages)
//   {Bz(k), Bz(k+1), 0, 0}
struct CellVecArgs {
  Geo g;
  const float *E[3], *B[3], *Bt[3];
  float4 *cv;
};

// One thread per cell; the block's vectors are staged in LDS and stored as
// one contiguous run (a thread's own five vectors lie 80 bytes apart).
__global__ void __launch_bounds__(kBlock) k_cellvec_f32(CellVecArgs a) {
  __shared__ float4 buf[kBlock * kCellVec];
  const long long c0 = (long long)blockIdx.x * blockDim.x;
  const long long c = c0 + threadIdx.x;
  const Geo &g = a.g;
  if (c < g.ncell) {
    int ix, iy, iz;
    decode_cell(g, (int)c, &ix, &iy, &iz);
    const long long o = gidx(g, ix, iy, iz), X = g.sx, Y = g.sy;
    const float *Ex = a.E[0], *Ey = a.E[1], *Ez = a.E[2];
    float4 *v = buf + threadIdx.x * kCellVec;
    v[0] = make_float4(Ex[o], Ex[o + Y], Ex[o + Y + 1], Ex[o + 1]);
    v[1] = make_float4(Ey[o], Ey[o + 1], Ey[o + X + 1], Ey[o + X]);
    v[2] = make_float4(Ez[o], Ez[o + X], Ez[o + X + Y], Ez[o + Y]);
    v[3] = make_float4(0.5f * (a.Bt[0][o] + a.B[0][o]), 0.5f * (a.Bt[0][o + X] + a.B[0][o + X]),
                       0.5f * (a.Bt[1][o] + a.B[1][o]), 0.5f * (a.Bt[1][o + Y] + a.B[1][o + Y]));
    v[4] = make_float4(0.5f * (a.Bt[2][o] + a.B[2][o]), 0.5f * (a.Bt[2][o + 1] + a.B[2][o + 1]), 0.f, 0.f);
  }
  __syncthreads();
  const int nv = (int)min((long long)kBlock, g.ncell - c0) * kCellVec;
  float4 *out = a.cv + c0 * kCellVec;
  for (int i = threadIdx.x; i < nv; i += kBlock) out[i] = buf[i];
}

// ---------------------------------------------------------------------------
// Pass 1: push, move codes, ranks and group sizes.  One wavefront owns 8
// consecutive source cells, so the ranks within each (source cell, code)
// group are complete when it finishes.
// ---------------------------------------------------------------------------
constexpr int kCellsPerWave = 8;  // source cells per wavefront of the inverse-permutation pass
constexpr int kPushCells = 4;     // source cells per wavefront of the push pass

struct PushRankArgs {
  PushK k;
  const int4 *pa, *pb;
  const float *g0;
  const int32_t *off;
  long long ncell;
  uint8_t *code;
  uint16_t *rank;
  uint16_t *nsk;   // [code * ncell + source]
  int32_t *fdst;   // destination cells of far movers
  int32_t *nfar;
  unsigned long long *flag;
};

template <bool REL, bool PER>
__global__ void __launch_bounds__(kBlock) k_push_rank_f32(PushRankArgs a) {
  __shared__ int cnt[kBlock / 64][kPushCells * kCodes];
  __shared__ int bnd[kBlock / 64][kPushCells + 1];
  __shared__ float4 cvs[kBlock / 64][kPushCells * kCellVec];  // the wavefront's gather vectors
  __shared__ float qp_s[EMO_NSPEC];
  // species constants in LDS: a per-lane load from the argument block would
  // make every iteration wait for all loads in flight
  if (threadIdx.x < EMO_NSPEC) qp_s[threadIdx.x] = a.k.s.qp[threadIdx.x];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long s0 = ((long long)xcd_tile() * (kBlock / 64) + wv) * kPushCells;
  if (s0 >= a.ncell) return;
  const int nc = (int)min((long long)kPushCells, a.ncell - s0);
  int *ct = cnt[wv];
  for (int e = lane; e < nc * kCodes; e += 64) ct[e] = 0;
  if (lane <= nc) bnd[wv][lane] = a.off[s0 + lane];
  // every particle of this wavefront lies in its own nc source cells: their
  // gather vectors are read once, into LDS
  if (lane < nc * kCellVec) cvs[wv][lane] = a.k.cv[s0 * kCellVec + lane];
  wave_sync();
  const int pb = bnd[wv][0], pe = bnd[wv][nc];
  // the inner cell boundaries, wave-uniform (scalar registers); past the
  // last cell they never compare below a particle index
  int sb[kPushCells - 1];
#pragma unroll
  for (int t = 1; t < kPushCells; t++) sb[t - 1] = __builtin_amdgcn_readfirstlane(t < nc ? bnd[wv][t] : INT_MAX);
  int nfar = 0;
  const bool use_g0 = a.g0 != nullptr;
  // software pipeline: the next chunk's states are in flight during a push.
  // Lanes past the end load (and push) the last particle again, so that the
  // loads are unconditional; only their stores are skipped.
  int4 A = make_int4(0, 0, 0, 0), Bv = A;
  float G = 0.f;
  if (pb < pe) {
    const int q = min(pb + lane, pe - 1);
    A = a.pa[q];
    Bv = a.pb[q];
    if (use_g0) G = a.g0[q];
  }
  for (int p0 = pb; p0 < pe; p0 += 64) {
    const int p = p0 + lane;
    const bool valid = p < pe;
    // the source cell of the state held (lanes past the end hold the last
    // particle)
    const int pq = min(p, pe - 1);
    int j = 0;
#pragma unroll
    for (int t = 0; t < kPushCells - 1; t++) j += sb[t] <= pq ? 1 : 0;
    const float4 *cvl = cvs[wv] + j * kCellVec;
    const CellVec V{cvl[0], cvl[1], cvl[2], cvl[3], cvl[4]};
    const int pn = min(p + 64, pe - 1);
    const int4 An = a.pa[pn], Bn = a.pb[pn];
    float Gn = 0.f;
    if (use_g0) Gn = a.g0[pn];
    const int sp = (int)((uint32_t)A.x >> kSpeciesShift);
    Pushed o;
    push_particle<REL, PER>(a.k, V, A.x, as_f(A.y), as_f(A.z), as_f(A.w), as_f(Bv.x), as_f(Bv.y), as_f(Bv.z),
                       qp_s[sp], use_g0, G, o, a.flag);
    const int mc = valid ? o.code : (int)kFarCode;
    if (valid) {
      if (mc == kFarCode) a.fdst[p] = o.cell;
    }
    A = An;
    Bv = Bn;
    G = Gn;
    const bool ranked = valid && mc != kFarCode;
    const int myrank = wave_rank(j * kCodes + mc, ranked, ct, lane);
    if (valid) {
      a.code[p] = (uint8_t)mc;
      if (ranked) {
        if (myrank > kMaxGroup) atomicOr(a.flag, kFlagMigrate);
        a.rank[p] = (uint16_t)myrank;
      }
    }
    nfar += __popcll(__ballot(valid && mc == kFarCode));
  }
  wave_sync();
  for (int e = lane; e < nc * kCodes; e += 64) {
    const int v = ct[e];
    if (v) a.nsk[(long long)(e % kCodes) * a.ncell + s0 + e / kCodes] = (uint16_t)min(v, kMaxGroup);
  }
  if (lane == 0 && nfar) atomicAdd(a.nfar, nfar);
}

// source and destination cells of a far mover (far-mover list)
struct CellsF32 {
  const int4 *pa;
  const int32_t *fdst;
  __device__ void operator()(long long i, int &s, int &d) const {
    s = pa[i].x & kCellMask;
    d = fdst[i];
  }
};

// ---------------------------------------------------------------------------
// Pass 2: the inverse permutation.  Same wavefront-to-cells mapping as pass 1;
// particle p of source cell s with move code k goes to
// off_new[s + shift(k)] + base[k][s] + rank[p].
// ---------------------------------------------------------------------------
struct InvArgs {
  Geo g;
  CellSortBufs b;
  const int32_t *fdst;
  int per[3];
  uint32_t *inv;
  unsigned long long *flag;
};

constexpr int kInvUnroll = 4;  // particles per lane in flight (2 and 8 measured the same)
constexpr int kStayCode = 43;  // (1 * 6 + 1) * 6 + 1: no move in any dimension

// New slot of particle p (move code k, rank r) of source cell src, or -1
// (flagged) if the group tables do not hold it.
__device__ inline long long inv_pos(const InvArgs &a, const int *shift, int src, int p, int k, int r) {
  const Geo &g = a.g;
  long long pos;
  if (k != kFarCode) {
    const int dst = src + shift[k];
    pos = -1;
    if (dst >= 0 && dst < g.ncell) {
      pos = (long long)a.b.off_new[dst] + a.b.gbase[(long long)k * g.ncell + src] + r;
      if (pos >= a.b.off_new[dst + 1]) pos = -1;
    }
    if (pos < 0) atomicOr(a.flag, kFlagMigrate);
  } else {
    pos = cellsort_pos(g, a.b, p, k, a.fdst[p], a.flag);
  }
  return pos;
}

// The particles that stay in their cell (most of them) take the cell's
// stayer base plus their rank, from a per-wavefront table; the others are
// collected in LDS and placed 64 at a time, one per lane, so that their
// dependent table loads are not waited for once per particle batch.
__global__ void __launch_bounds__(kBlock) k_inv_f32(InvArgs a) {
  __shared__ int bnd[kBlock / 64][kCellsPerWave + 1];
  __shared__ int stay0[kBlock / 64][kCellsPerWave];  // first new slot of each source cell's stayers
  __shared__ int stay1[kBlock / 64][kCellsPerWave];  // end of that cell's new range
  __shared__ int mov[kBlock / 64][128];              // movers waiting for their slot
  __shared__ int shift[kCodes];                      // cell-id shift of each move code
  const Geo &g = a.g;
  for (int k = threadIdx.x; k < kCodes; k += kBlock) {
    const int kx = k / 36, ky = (k / 6) % 6, kz = k % 6;
    shift[k] = (move_sigma(kx, g.n[0]) * g.nc[1] + move_sigma(ky, g.n[1])) * g.nc[2] + move_sigma(kz, g.n[2]);
  }
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long s0 = ((long long)xcd_tile() * (kBlock / 64) + wv) * kCellsPerWave;
  if (s0 >= g.ncell) return;
  const int nc = (int)min((long long)kCellsPerWave, g.ncell - s0);
  if (lane <= nc) bnd[wv][lane] = a.b.off[s0 + lane];
  if (lane < nc) {
    // (the base is stale for a cell without stayers; it is then not read)
    stay0[wv][lane] = a.b.off_new[s0 + lane] + a.b.gbase[(long long)kStayCode * g.ncell + s0 + lane];
    stay1[wv][lane] = a.b.off_new[s0 + lane + 1];
  }
  wave_sync();
  const int pb = bnd[wv][0], pe = bnd[wv][nc];
  // the inner cell boundaries in scalar registers (past the last cell they
  // never compare below a particle index)
  int sb[kCellsPerWave - 1];
#pragma unroll
  for (int t = 1; t < kCellsPerWave; t++) sb[t - 1] = __builtin_amdgcn_readfirstlane(t < nc ? bnd[wv][t] : INT_MAX);
  auto cell_of = [&](int p) {
    int j = 0;
#pragma unroll
    for (int t = 0; t < kCellsPerWave - 1; t++) j += sb[t] <= p ? 1 : 0;
    return j;
  };
  // the first n (<= 64) waiting movers, one per lane
  auto place_movers = [&](int n) {
    if (lane < n) {
      const int p = mov[wv][lane];
      const long long pos = inv_pos(a, shift, (int)s0 + cell_of(p), p, (int)a.b.code[p], (int)a.b.rank[p]);
      if (pos >= 0) a.inv[pos] = (uint32_t)p;
    }
  };
  int nmov = 0;  // movers waiting in mov[wv] (wavefront-uniform, below 64 between batches)
  // a wavefront-uniform loop: every lane takes part in the batches
  // each lane takes 4 consecutive particles: their codes are one 4-byte
  // load and their ranks one 8-byte load, from 4-aligned indices (the first
  // chunk starts at pb rounded down; indices outside [pb, pe) are skipped)
  static_assert(kInvUnroll == 4, "four particles per lane");
  for (int q0 = pb & ~3; q0 < pe; q0 += 64 * kInvUnroll) {
    const int p0 = q0 + 4 * lane;
    uint32_t kw = 0xffffffffu;
    uint2 rw = make_uint2(0u, 0u);
    if (p0 < pe) {
      kw = *(const uint32_t *)(a.b.code + p0);
      rw = *(const uint2 *)(a.b.rank + p0);
    }
    int k[kInvUnroll], r[kInvUnroll];
#pragma unroll
    for (int u = 0; u < kInvUnroll; u++) {
      const int p = p0 + u;
      k[u] = (p >= pb && p < pe) ? (int)((kw >> (8 * u)) & 0xffu) : -1;
      r[u] = (int)(((u < 2 ? rw.x : rw.y) >> (16 * (u & 1))) & 0xffffu);
    }
#pragma unroll
    for (int u = 0; u < kInvUnroll; u++) {
      const int p = p0 + u;
      if (k[u] == kStayCode) {
        const int j = cell_of(p);
        const long long pos = (long long)stay0[wv][j] + r[u];
        if (pos < stay1[wv][j])
          a.inv[pos] = (uint32_t)p;
        else
          atomicOr(a.flag, kFlagMigrate);
      }
      const bool mover = k[u] >= 0 && k[u] != kStayCode;
      const unsigned long long bal = __ballot(mover);
      if (mover) mov[wv][nmov + __popcll(bal & lanes_below(lane))] = p;
      nmov += __popcll(bal);
      if (nmov >= 64) {
        wave_sync();
        place_movers(64);
        wave_sync();
        // the remainder (fewer than 64) moves to the front
        if (lane < nmov - 64) mov[wv][lane] = mov[wv][64 + lane];
        nmov -= 64;
        wave_sync();
      }
    }
  }
  wave_sync();
  place_movers(nmov);
}

// ---------------------------------------------------------------------------
// Pass 3: move + deposit in destination order, 16 lanes per cell, 4 cells per
// wavefront.  (8 lanes per cell, with a cheaper butterfly, measured 25%
// slower on the 64^3, 64-particles-per-cell step.)
// ---------------------------------------------------------------------------
constexpr int kDepLanes = 16;

__device__ inline void two_sum(float a, float b, float &s, float &e) {
  s = a + b;
  float bb = s - a;
  e = (a - (s - bb)) + (b - bb);
}

// One level of the deposit butterfly: the partner's (hi, lo) by a DPP move
// (CTRL), added with TwoSum.  Levels, within each group of 16 lanes:
// row_mirror pairs lane i with 15 - i, row_half_mirror lane i with 7 - i
// within each 8, quad_perm [2,3,0,1] and [1,0,3,2] the quad halves and
// neighbours.  Each level pairs lanes whose partial sums cover disjoint
// lane sets, so after the last level every lane of the group holds the sum
// of all of them, combined in a fixed order.
template <int CTRL>
__device__ inline void dep_level(float hi[12], float lo[12]) {
#pragma unroll
  for (int q = 0; q < 12; q++) {
    const float oh = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(hi[q]), CTRL, 0xF, 0xF, false));
    const float ol = __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(lo[q]), CTRL, 0xF, 0xF, false));
    float s, err;
    two_sum(hi[q], oh, s, err);
    hi[q] = s;
    lo[q] = lo[q] + ol + err;
  }
}

__device__ inline void dep_reduce16(float hi[12], float lo[12]) {
  dep_level<0x140>(hi, lo);
  dep_level<0x141>(hi, lo);
  dep_level<0x4E>(hi, lo);
  dep_level<0xB1>(hi, lo);
}

// the 12 edge currents of one particle (Depositor::deposit_particle_J,
// src/particles/deposit.cpp:12-40, weights 2 - f and f - 1), accumulated
// with TwoSum into (hi, lo)
__device__ inline void deposit_j(float hi[12], float lo[12], const float f[3], const float v[3], float jc) {
  const float A[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
  const float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int j = (q + 1) % 3, k = (q + 2) % 3;
    const float cq = jc * v[q];
    const float ca = cq * A[j], cb = cq * Bw[j];
    const float w4[4] = {ca * A[k], cb * A[k], cb * Bw[k], ca * Bw[k]};
#pragma unroll
    for (int r = 0; r < 4; r++) {
      float s, err;
      two_sum(hi[4 * q + r], w4[r], s, err);
      hi[4 * q + r] = s;
      lo[4 * q + r] += err;
    }
  }
}

struct MoveDepArgs {
  PushK k;
  const int4 *pa, *pb;  // old state, source order
  const float *g0;
  int4 *pa_out, *pb_out;
  const uint32_t *inv;
  const int32_t *off;  // new cell ranges
  long long ncell;
  long long np;
  float *cj;
  int32_t *nfar;  // the far-mover count, read by the earlier passes: reset here for the next step
  unsigned long long *flag;
};

template <bool REL, bool PER>
__global__ void __launch_bounds__(kBlock) k_move_deposit_f32(MoveDepArgs a) {
  __shared__ float qp_s[EMO_NSPEC], jc_s[EMO_NSPEC];
  // species constants in LDS (a per-lane load from the argument block would
  // make every iteration wait for all loads in flight); test particles
  // deposit nothing (src/particles/particle_handler.cpp:205)
  if (threadIdx.x < EMO_NSPEC) {
    qp_s[threadIdx.x] = a.k.s.qp[threadIdx.x];
    jc_s[threadIdx.x] = a.k.s.test[threadIdx.x] ? 0.f : a.k.s.jc[threadIdx.x];
  }
  __syncthreads();
  if (blockIdx.x == 0 && threadIdx.x == 0) *a.nfar = 0;
  const int sub = threadIdx.x & (kDepLanes - 1);
  const long long c = (long long)xcd_tile() * (kBlock / kDepLanes) + threadIdx.x / kDepLanes;
  float hi[12], lo[12];
#pragma unroll
  for (int q = 0; q < 12; q++) hi[q] = lo[q] = 0.f;
  const int b = c < a.ncell ? a.off[c] : 0, e = c < a.ncell ? a.off[c + 1] : 0;
  if (b < e) {  // (an empty cell reads nothing: with no particles, slot 0 holds no record)
    // software pipeline: the source index two slots ahead and the state one
    // slot ahead are in flight while a particle is pushed.  A source index
    // outside the store (a slot that the permutation did not fill) is an
    // error; it is replaced by 0 where it is used, so that no access leaves
    // the store.  Loads past the cell's end read its last slot (b < e
    // there), so that the loads are unconditional.
    const uint32_t np = (uint32_t)a.np;
    const bool use_g0 = a.g0 != nullptr;
    const int last = e - 1;
    int j = b + sub;
    bool bad = false;
    uint32_t src = 0u, src_next = 0u;
    if (j < e) {
      src = a.inv[j];
      src_next = a.inv[min(j + kDepLanes, last)];
    }
    bad = src >= np;
    src = bad ? 0u : src;
    int4 A = a.pa[src], Bv = a.pb[src];
    float G = 0.f;
    if (use_g0) G = a.g0[src];
    CellVec V = load_cellvec(a.k, A.x);
    for (; j < e; j += kDepLanes) {
      const uint32_t src_next2 = a.inv[min(j + 2 * kDepLanes, last)];
      bad = bad || src_next >= np;
      const uint32_t sn = src_next < np ? src_next : 0u;
      const int4 An = a.pa[sn], Bn = a.pb[sn];
      float Gn = 0.f;
      if (use_g0) Gn = a.g0[sn];
      const int sp = (int)((uint32_t)A.x >> kSpeciesShift);
      Pushed o;
      push_particle<REL, PER>(a.k, V, A.x, as_f(A.y), as_f(A.z), as_f(A.w), as_f(Bv.x), as_f(Bv.y), as_f(Bv.z),
                         qp_s[sp], use_g0, G, o, nullptr);
      bad = bad || o.cell != (int)c;
      // the next particle's gather vector: in flight during the deposit
      V = load_cellvec(a.k, An.x);
      a.pa_out[j] = make_int4((int)(((uint32_t)sp << kSpeciesShift) | (uint32_t)c), as_i(o.f[0]), as_i(o.f[1]),
                              as_i(o.f[2]));
      a.pb_out[j] = make_int4(as_i(o.u[0]), as_i(o.u[1]), as_i(o.u[2]), Bv.w);
      deposit_j(hi, lo, o.f, o.v, jc_s[sp]);
      src_next = src_next2;
      A = An;
      Bv = Bn;
      G = Gn;
    }
    if (bad) atomicOr(a.flag, kFlagMigrate);
  }
  // all 64 lanes take part (the DPP moves read every lane of the group)
  static_assert(kDepLanes == 16, "the butterfly covers 16 lanes");
  dep_reduce16(hi, lo);
  if (sub == 0 && c < a.ncell) {
#pragma unroll
    for (int q = 0; q < 12; q++) a.cj[q * a.ncell + c] = hi[q] + lo[q];
  }
}

// the sort as pre-push index -> new slot, from the inverse permutation
__global__ void k_fwd_from_inv(const uint32_t *inv, uint32_t *fwd, long long np) {
  const long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (j < np) fwd[inv[j]] = (uint32_t)j;
}

// ---------------------------------------------------------------------------
// rho + J of the current (sorted) state, for the Poisson initialisation
// (src/main.cpp:160-166): the 12 edge currents and the 8 vertex charges of
// Depositor::deposit_particle_J / _Rho (src/particles/deposit.cpp:12-62)
// with the weights 2 - f and f - 1, 16 lanes per cell, compensated sums and a
// fixed butterfly as in the step's deposit.
// ---------------------------------------------------------------------------
struct RhoJArgs {
  const int4 *pa, *pb;
  const float *g0;
  const int32_t *off;
  SpeciesF32 s;
  int rel;
  long long ncell;
  float *cj, *cr;
};

__global__ void __launch_bounds__(kBlock) k_deposit_rhoj_f32(RhoJArgs a) {
  const int sub = threadIdx.x & 15;
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  float hi[20], lo[20];
#pragma unroll
  for (int q = 0; q < 20; q++) hi[q] = lo[q] = 0.f;
  if (c < a.ncell) {
    const int b = a.off[c], e = a.off[c + 1];
    for (int j = b + sub; j < e; j += 16) {
      const int4 A = a.pa[j], Bv = a.pb[j];
      const int sp = (int)((uint32_t)A.x >> kSpeciesShift);
      if (a.s.test[sp]) continue;
      const float f[3] = {as_f(A.y), as_f(A.z), as_f(A.w)};
      const float ux = as_f(Bv.x), uy = as_f(Bv.y), uz = as_f(Bv.z);
      float ig = 1.f;
      if (a.rel) ig = a.g0 ? __builtin_amdgcn_rcpf(a.g0[j]) : __builtin_amdgcn_rsqf(1.f + ux * ux + uy * uy + uz * uz);
      const float v[3] = {ux * ig, uy * ig, uz * ig};
      const float jc = a.s.jc[sp];
      deposit_j(hi, lo, f, v, jc);
      const float Aw[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
      const float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
      // vertex order of Depositor::deposit_particle_Rho
      const float w8[8] = {Aw[0] * Aw[1] * Aw[2], Bw[0] * Aw[1] * Aw[2], Bw[0] * Bw[1] * Aw[2],
                           Aw[0] * Bw[1] * Aw[2], Aw[0] * Bw[1] * Bw[2], Bw[0] * Bw[1] * Bw[2],
                           Bw[0] * Aw[1] * Bw[2], Aw[0] * Aw[1] * Bw[2]};
#pragma unroll
      for (int r = 0; r < 8; r++) {
        float s2, err;
        two_sum(hi[12 + r], jc * w8[r], s2, err);
        hi[12 + r] = s2;
        lo[12 + r] += err;
      }
    }
  }
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 20; q++) {
      float oh = __shfl_xor(hi[q], off, 64);
      float ol = __shfl_xor(lo[q], off, 64);
      float s2, err;
      two_sum(hi[q], oh, s2, err);
      hi[q] = s2;
      lo[q] = lo[q] + ol + err;
    }
  }
  if (sub == 0 && c < a.ncell) {
#pragma unroll
    for (int q = 0; q < 12; q++) a.cj[q * a.ncell + c] = hi[q] + lo[q];
#pragma unroll
    for (int q = 0; q < 8; q++) a.cr[q * a.ncell + c] = hi[12 + q] + lo[12 + q];
  }
}

// emo_load_uniform's particles in the chosen representation (generation
// order; sorted afterwards like any load): the cell and offset from the FP64
// position with Grid::getCellID's arithmetic, as fp32_load does on the host
struct GenF32Args {
  UniformGen u;
  int rel;
  int4 *pa, *pb;
  float *g0;
  float *gam_id;  // non-relativistic: the loaded gamma by id
  long long np;
  unsigned long long *flag;
};

__global__ void __launch_bounds__(kBlock) k_uniform_f32(GenF32Args a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const Geo &g = a.u.g;
  double x[3], v[3], gam;
  int sp;
  ugen_particle(a.u, i, x, v, &gam, &sp);
  int ic[3];
  int fb[3];
  for (int d = 0; d < 3; d++) {
    const double r = (x[d] - g.x0[d]) * g.id[d];
    const int ix = (int)r;
    ic[d] = ix + 1;
    if (ic[d] < 0 || ic[d] >= g.nc[d]) {
      atomicOr(a.flag, kFlagOutOfGrid);
      ic[d] = 1;
    }
    float f = (float)(r - (double)ix);
    f = f >= 1.f ? 0x1.fffffep-1f : (f < 0.f ? 0.f : f);
    fb[d] = as_i(f);
  }
  const double gm = a.rel ? gam : 1.0;
  a.pa[i] = make_int4((int)(((uint32_t)sp << kSpeciesShift) | (uint32_t)((ic[0] * g.nc[1] + ic[1]) * g.nc[2] + ic[2])),
                      fb[0], fb[1], fb[2]);
  a.pb[i] = make_int4(as_i((float)(gm * v[0])), as_i((float)(gm * v[1])), as_i((float)(gm * v[2])), (int)i);
  if (a.rel) a.g0[i] = (float)gm;
  if (!a.rel) a.gam_id[i] = (float)gam;  // the id of load index i is i
}

// ---------------------------------------------------------------------------
// Initial sort of freshly loaded particles (arbitrary order): radix sort of
// the cell ids, payload permutation, particles per cell.
// ---------------------------------------------------------------------------
__global__ void k_keys_from_pa(const int4 *pa, uint32_t *key, uint32_t *perm, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  key[i] = (uint32_t)(pa[i].x & kCellMask);
  perm[i] = (uint32_t)i;
}

struct PermF32Args {
  const int4 *pa, *pb;
  const float *g0;
  int4 *pa_out, *pb_out;
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
    const int4 A = a.pa[s];
    a.pa_out[i] = A;
    a.pb_out[i] = a.pb[s];
    if (a.g0) a.g0_out[i] = a.g0[s];
    cell = A.x & kCellMask;
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
  const int4 *pa, *pb;
  const float *g0;
  const float *gam_id;  // non-relativistic: the loaded gamma by id, or null
  const float4 *cv;     // with fld: the current grid's gather vectors
  int rel;
  double *x[3], *v[3], *gam;
  double *fld[6];       // the gathered E, B per particle (Field_part), or null
  int32_t *sp;
  int32_t *id;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_output_f32(OutArgs a) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const Geo &g = a.g;
  const int4 A = a.pa[i], Bv = a.pb[i];
  int cell = A.x & kCellMask;
  int ic[3];
  decode_cell(g, cell, &ic[0], &ic[1], &ic[2]);
  const double uu[3] = {as_f(Bv.x), as_f(Bv.y), as_f(Bv.z)};
  const double ff[3] = {as_f(A.y), as_f(A.z), as_f(A.w)};
  double gam = 1.0;
  if (a.rel) gam = a.g0 ? (double)a.g0[i] : sqrt(1.0 + uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
  for (int d = 0; d < 3; d++) {
    if (a.x[d]) a.x[d][i] = g.x0[d] + g.d[d] * ((double)(ic[d] - 1) + ff[d]);
    if (a.v[d]) a.v[d][i] = a.rel ? uu[d] / gam : uu[d];
  }
  if (!a.rel && a.gam_id && Bv.w >= 0 && Bv.w < a.np) gam = a.gam_id[Bv.w];
  if (a.gam) a.gam[i] = gam;
  if (a.fld[0]) {
    const float4 *cv = a.cv + (long long)cell * kCellVec;
    const CellVec V{cv[0], cv[1], cv[2], cv[3], cv[4]};
    float e[3], b[3];
    gather_cell(V, as_f(A.y), as_f(A.z), as_f(A.w), e, b);
    for (int d = 0; d < 3; d++) {
      a.fld[d][i] = e[d];
      a.fld[3 + d][i] = b[d];
    }
  }
  if (a.sp) a.sp[i] = (int)((uint32_t)A.x >> kSpeciesShift);
  if (a.id) a.id[i] = Bv.w;
}

__global__ void k_cells_f32(const int4 *pa, int32_t *out, long long np) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < np) out[i] = pa[i].x & kCellMask;
}

struct DiagF32Args {
  const int4 *pa, *pb;
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
    int sp = (int)((uint32_t)a.pa[i].x >> kSpeciesShift);
    const int4 Bv = a.pb[i];
    double m = a.s.m[sp];
    double ux = as_f(Bv.x), uy = as_f(Bv.y), uz = as_f(Bv.z);
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
    hipLaunchKernelGGL(k_keys_from_pa, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, s->pa, s->key, s->perm, np);
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
    pa.pa = s->pa;
    pa.pb = s->pb;
    pa.pa_out = s->pa_alt;
    pa.pb_out = s->pb_alt;
    pa.g0 = s->g0_active ? s->g0 : nullptr;
    pa.g0_out = g0_new;
    pa.perm = s->perm_alt;
    pa.count = s->srt.count;
    pa.np = np;
    hipLaunchKernelGGL(k_permute_f32, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, pa);
    std::swap(s->pa, s->pa_alt);
    std::swap(s->pb, s->pb_alt);
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

PushK push_consts(emo_ctx *c, double dt) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  PushK k;
  k.g = g;
  k.cv = s->cv;
  for (int d = 0; d < 3; d++) {
    k.dtid[d] = (float)(dt * g.id[d]);
    k.lper[d] = c->cfg.parts_bound[2 * d] == EMO_BC_PERIODIC;
    k.rper[d] = c->cfg.parts_bound[2 * d + 1] == EMO_BC_PERIODIC;
  }
  cellsort_periodic(c, k.per);
  k.s = species_f32(c, dt);
  k.div1 = make_fastdiv((uint32_t)g.nc[1]);
  k.div2 = make_fastdiv((uint32_t)g.nc[2]);
  return k;
}

unsigned cell_wave_blocks(long long ncell, int cells_per_wave) {
  const long long waves = (ncell + cells_per_wave - 1) / cells_per_wave;
  return (unsigned)std::max<long long>((waves + kBlock / 64 - 1) / (kBlock / 64), 1);
}

// One step's particle work: push + ranks, far list, group offsets, new
// ranges, inverse permutation, move + deposit, grid sum of the currents.
int particle_step(emo_ctx *c, double dt) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  const long long np = c->np;
  {
    CellVecArgs a;
    a.g = g;
    for (int d = 0; d < 3; d++) {
      a.E[d] = s->fs.E[d];
      a.B[d] = s->fs.B[d];
      a.Bt[d] = s->fs.Bt[d];
    }
    a.cv = s->cv;
    TimedLaunch t(c, "cellvec_f32");
    hipLaunchKernelGGL(k_cellvec_f32, dim3(grid_for(g.ncell)), dim3(kBlock), 0, c->stream, a);
  }
  const PushK k = push_consts(c, dt);
  const float *g0 = s->g0_active ? s->g0 : nullptr;
  bool all_per = true;
  for (int q = 0; q < 6; q++) all_per = all_per && c->cfg.parts_bound[q] == EMO_BC_PERIODIC;
  {
    PushRankArgs a;
    a.k = k;
    a.pa = s->pa;
    a.pb = s->pb;
    a.g0 = g0;
    a.off = s->srt.off;
    a.ncell = g.ncell;
    a.code = s->srt.code;
    a.rank = s->srt.rank;
    a.nsk = s->srt.nsk;
    a.fdst = s->fdst;
    a.nfar = s->srt.misc;
    a.flag = c->dflag;
    TimedLaunch t(c, "push_rank_f32");
    const dim3 grid(xcd_grid(cell_wave_blocks(g.ncell, kPushCells)));
    if (c->cfg.relativity && all_per)
      hipLaunchKernelGGL((k_push_rank_f32<true, true>), grid, dim3(kBlock), 0, c->stream, a);
    else if (c->cfg.relativity)
      hipLaunchKernelGGL((k_push_rank_f32<true, false>), grid, dim3(kBlock), 0, c->stream, a);
    else if (all_per)
      hipLaunchKernelGGL((k_push_rank_f32<false, true>), grid, dim3(kBlock), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_push_rank_f32<false, false>), grid, dim3(kBlock), 0, c->stream, a);
  }
  cellsort_far(c, s->srt, np, CellsF32{s->pa, s->fdst}, "far_f32");
  int rc = cellsort_base_scan(c, s->srt, "base_f32", "scan_f32");
  if (rc) return rc;
  {
    InvArgs a;
    a.g = g;
    a.b = s->srt;
    a.fdst = s->fdst;
    cellsort_periodic(c, a.per);
    a.inv = s->srt.perm;
    a.flag = c->dflag;
    TimedLaunch t(c, "inv_f32");
    hipLaunchKernelGGL(k_inv_f32, dim3(xcd_grid(cell_wave_blocks(g.ncell, kCellsPerWave))), dim3(kBlock), 0,
                       c->stream, a);
  }
  {
    MoveDepArgs a;
    a.k = k;
    a.pa = s->pa;
    a.pb = s->pb;
    a.g0 = g0;
    a.pa_out = s->pa_alt;
    a.pb_out = s->pb_alt;
    a.inv = s->srt.perm;
    a.off = s->srt.off_new;
    a.ncell = g.ncell;
    a.np = np;
    a.cj = s->cj;
    a.nfar = s->srt.misc;
    a.flag = c->dflag;
    const unsigned blocks = xcd_grid((unsigned)((g.ncell + (kBlock / kDepLanes) - 1) / (kBlock / kDepLanes)));
    TimedLaunch t(c, "move_deposit_f32");
    if (c->cfg.relativity && all_per)
      hipLaunchKernelGGL((k_move_deposit_f32<true, true>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
    else if (c->cfg.relativity)
      hipLaunchKernelGGL((k_move_deposit_f32<true, false>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
    else if (all_per)
      hipLaunchKernelGGL((k_move_deposit_f32<false, true>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
    else
      hipLaunchKernelGGL((k_move_deposit_f32<false, false>), dim3(blocks), dim3(kBlock), 0, c->stream, a);
  }
  std::swap(s->pa, s->pa_alt);
  std::swap(s->pb, s->pb_alt);
  std::swap(s->srt.off, s->srt.off_new);
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
  if (!rc) rc = dalloc(c, &s->pa, cap);
  if (!rc) rc = dalloc(c, &s->pa_alt, cap);
  if (!rc) rc = dalloc(c, &s->pb, cap);
  if (!rc) rc = dalloc(c, &s->pb_alt, cap);
  if (!rc) rc = dalloc(c, &s->fdst, cap);
  if (!rc) rc = cellsort_alloc(c, s->srt, cap, g.ncell);
  if (!rc) rc = dalloc(c, &s->key, cap);
  if (!rc) rc = dalloc(c, &s->key_alt, cap);
  if (!rc) rc = dalloc(c, &s->perm, cap);
  if (!rc) rc = dalloc(c, &s->perm_alt, cap);
  if (!rc) rc = dalloc(c, &s->cj, 12 * g.ncell);
  if (!rc) rc = dalloc(c, &s->cv, (size_t)kCellVec * g.ncell);
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
  dfree(s->pa);
  dfree(s->pa_alt);
  dfree(s->pb);
  dfree(s->pb_alt);
  dfree(s->fdst);
  for (int d = 0; d < 3; d++) {
    dfree(s->fs.E[d]);
    dfree(s->fs.Bfix[d]);
    dfree(s->Btfix[d]);
    dfree(s->fs.J[d]);
  }
  dfree(s->fs.rho);
  dfree(s->g0);
  dfree(s->gam_id);
  cellsort_free(s->srt);
  dfree(s->key);
  dfree(s->key_alt);
  dfree(s->perm);
  dfree(s->perm_alt);
  dfree(s->sort_tmp);
  dfree(s->cj);
  dfree(s->cv);
  delete s;
  c->ch = nullptr;
}

// Converts FP64 particles to the chosen representation on the host (the same
// cell-id arithmetic as Grid::getCellID) and sorts them by cell.
int fp32_load(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
              const int32_t *species, const int64_t *id) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  std::vector<int4> ha(n), hb(n);
  std::vector<float> g0;
  const bool rel = c->cfg.relativity != 0;
  if (rel) g0.resize(n);
  for (int64_t i = 0; i < n; i++) {
    int ic[3];
    float ff[3];
    for (int d = 0; d < 3; d++) {
      double r = (x3[3 * i + d] - g.x0[d]) * g.id[d];
      int ix = (int)r;
      ic[d] = ix + 1;
      if (ic[d] < 0 || ic[d] >= g.nc[d]) return set_err(c, EMO_ERR_INVARIANT, "particle outside the padded grid");
      // true in-cell fraction of the reference's cell (floor via the int cast)
      double fr = (x3[3 * i + d] - g.x0[d]) * g.id[d] - (double)ix;
      float f = (float)fr;
      if (f >= 1.f) f = std::nextafter(1.f, 0.f);
      if (f < 0.f) f = 0.f;
      ff[d] = f;
    }
    int32_t w = (species[i] << kSpeciesShift) | ((ic[0] * g.nc[1] + ic[1]) * g.nc[2] + ic[2]);
    double gm = (rel && gamma) ? gamma[i] : 1.0;
    float u[3];
    for (int d = 0; d < 3; d++) u[d] = (float)(rel ? gm * v3[3 * i + d] : v3[3 * i + d]);
    if (rel) g0[i] = (float)gm;
    int32_t pid = -1;
    if (id) {
      if (id[i] < INT_MIN || id[i] > INT_MAX)
        return set_err(c, EMO_ERR_UNSUPPORTED, "particle ids must fit in int32 in the FP32 chosen form");
      pid = (int32_t)id[i];
    }
    int fb[3], ub[3];
    std::memcpy(fb, ff, 12);
    std::memcpy(ub, u, 12);
    ha[i] = make_int4(w, fb[0], fb[1], fb[2]);
    hb[i] = make_int4(ub[0], ub[1], ub[2], pid);
  }
  if (n > 0) {
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->pa, ha.data(), n * sizeof(int4), hipMemcpyHostToDevice));
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->pb, hb.data(), n * sizeof(int4), hipMemcpyHostToDevice));
  }
  s->have_id = id != nullptr;
  dfree(s->g0);
  s->g0 = nullptr;
  s->g0_active = false;
  dfree(s->gam_id);
  s->gam_id = nullptr;
  if (rel && n > 0) {
    int rc = dalloc(c, &s->g0, n);
    if (rc) return rc;
    EMO_HIP_CHECK_RET(c, hipMemcpy(s->g0, g0.data(), n * 4, hipMemcpyHostToDevice));
    s->g0_active = true;
  }
  if (!rel && gamma && id && n > 0) {
    // by id, when the ids are 0 .. n-1
    std::vector<float> gi(n, 1.f);
    bool ok = true;
    for (int64_t i = 0; i < n && ok; i++) {
      ok = id[i] >= 0 && id[i] < n;
      if (ok) gi[id[i]] = (float)gamma[i];
    }
    if (ok) {
      int rc = dalloc(c, &s->gam_id, n);
      if (rc) return rc;
      EMO_HIP_CHECK_RET(c, hipMemcpy(s->gam_id, gi.data(), n * 4, hipMemcpyHostToDevice));
    }
  }
  c->np = n;
  int rc = sort_loaded(c);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return check_flag(c);
}

int fp32_get_particles(emo_ctx *c, double *x3, double *v3, double *gamma, int32_t *species, int64_t *id,
                       double *field6) {
  ChosenState *s = c->ch;
  int64_t n = c->np;
  if (n == 0) return EMO_OK;
  double *dx[3] = {}, *dv[3] = {}, *dg = nullptr, *df[6] = {};
  int32_t *dsp = nullptr, *did = nullptr;
  int rc = 0;
  for (int d = 0; d < 3 && !rc; d++) {
    rc = dalloc(c, &dx[d], n);
    if (!rc) rc = dalloc(c, &dv[d], n);
  }
  for (int d = 0; d < 6 && !rc && field6; d++) rc = dalloc(c, &df[d], n);
  if (!rc && field6) {
    // the gather of the current grid (Particle_Handler::InterpolateEB, which
    // ends the reference's step): the gather vectors of this grid first
    CellVecArgs ca;
    ca.g = c->geo;
    for (int d = 0; d < 3; d++) {
      ca.E[d] = s->fs.E[d];
      ca.B[d] = s->fs.B[d];
      ca.Bt[d] = s->fs.Bt[d];
    }
    ca.cv = s->cv;
    hipLaunchKernelGGL(k_cellvec_f32, dim3(grid_for(c->geo.ncell)), dim3(kBlock), 0, c->stream, ca);
  }
  if (!rc) rc = dalloc(c, &dg, n);
  if (!rc) rc = dalloc(c, &dsp, n);
  if (!rc) rc = dalloc(c, &did, n);
  if (!rc) {
    OutArgs a;
    a.g = c->geo;
    a.pa = s->pa;
    a.pb = s->pb;
    for (int d = 0; d < 3; d++) {
      a.x[d] = dx[d];
      a.v[d] = dv[d];
    }
    a.g0 = s->g0_active ? s->g0 : nullptr;
    a.gam_id = s->gam_id;
    a.cv = s->cv;
    for (int d = 0; d < 6; d++) a.fld[d] = df[d];
    a.rel = c->cfg.relativity;
    a.gam = dg;
    a.sp = dsp;
    a.id = did;
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
    for (int d = 0; d < 6 && field6 && !rc; d++) rc = down(field6, df[d], 6, d);
    if (species && !rc && hipMemcpy(species, dsp, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMO_ERR_HIP;
    if (id && !rc) {
      std::vector<int32_t> h(n);
      if (hipMemcpy(h.data(), did, n * 4, hipMemcpyDeviceToHost) != hipSuccess) rc = EMO_ERR_HIP;
      for (int64_t i = 0; i < n && !rc; i++) id[i] = s->have_id ? h[i] : -1;
    }
  }
  for (int d = 0; d < 3; d++) {
    dfree(dx[d]);
    dfree(dv[d]);
  }
  dfree(dg);
  for (int d = 0; d < 6; d++) dfree(df[d]);
  dfree(dsp);
  dfree(did);
  return rc;
}

int fp32_get_cells(emo_ctx *c, int32_t *cells) {
  ChosenState *s = c->ch;
  int64_t n = c->np;
  if (n == 0) return EMO_OK;
  int32_t *d = nullptr;
  int rc = dalloc(c, &d, n);
  if (rc) return rc;
  hipLaunchKernelGGL(k_cells_f32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, s->pa, d, (long long)n);
  hipError_t e = hipMemcpyAsync(cells, d, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d);
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, hipGetErrorString(e));
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

// rho of the current particles (Depositor::deposit_particle_Rho and the fold,
// as the reference's depositRhoJ + executeBC(-2) leave it after a step);
// the J the step deposited is left as it is
int fp32_deposit_rho(emo_ctx *c) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  float *cj = nullptr, *cr = nullptr, *jt = nullptr;
  int rc = dalloc(c, &cj, 12 * g.ncell);
  if (!rc) rc = dalloc(c, &cr, 8 * g.ncell);
  if (!rc) rc = dalloc(c, &jt, 3 * g.ntot);
  if (!rc) {
    RhoJArgs a;
    a.pa = s->pa;
    a.pb = s->pb;
    a.g0 = s->g0_active ? s->g0 : nullptr;
    a.off = s->srt.off;
    a.s = species_f32(c, 0.0);
    a.rel = c->cfg.relativity;
    a.ncell = g.ncell;
    a.cj = cj;
    a.cr = cr;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    hipLaunchKernelGGL(k_deposit_rhoj_f32, dim3(blocks), dim3(kBlock), 0, c->stream, a);
    DepGridArgs<float> b;
    b.g = g;
    b.cj = cj;
    b.cr = cr;
    for (int d = 0; d < 3; d++) b.J[d] = jt + d * g.ntot;  // discarded
    b.rho = s->fs.rho;
    b.do_rho = 1;
    hipLaunchKernelGGL(k_deposit_grid<float>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, c->stream, b);
    FieldSet<float> fr = s->fs;
    for (int d = 0; d < 3; d++) fr.J[d] = nullptr;
    fk_fold<float>(c, fr);
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = set_err(c, EMO_ERR_HIP, "rho deposit failed");
  }
  dfree(cj);
  dfree(cr);
  dfree(jt);
  if (!rc) s->rho_stale = false;
  return rc;
}

int fp32_get_field(emo_ctx *c, int fid, double *out) {
  float *p = f32_field(c->ch, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  if (fid == EMO_RHO && c->ch->rho_stale) {
    int rc = fp32_deposit_rho(c);
    if (rc) return rc;
  }
  std::vector<float> h(c->geo.ntot);
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(h.data(), p, c->geo.ntot * 4, hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  for (long long i = 0; i < c->geo.ntot; i++) out[i] = h[i];
  return EMO_OK;
}

int fp32_set_field(emo_ctx *c, int fid, const double *in) {
  float *p = f32_field(c->ch, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  if (fid == EMO_RHO) c->ch->rho_stale = false;
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
  s->rho_stale = false;
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return check_flag(c);
}

int fp32_load_uniform(emo_ctx *c, const UniformGen &u, int64_t n) {
  ChosenState *s = c->ch;
  const bool rel = c->cfg.relativity != 0;
  dfree(s->g0);
  s->g0 = nullptr;
  s->g0_active = false;
  dfree(s->gam_id);
  s->gam_id = nullptr;
  if (rel && n > 0) {
    int rc = dalloc(c, &s->g0, n);
    if (rc) return rc;
    s->g0_active = true;
  }
  if (!rel && n > 0) {
    int rc = dalloc(c, &s->gam_id, n);
    if (rc) return rc;
  }
  if (n > 0) {
    GenF32Args a;
    a.u = u;
    a.rel = rel;
    a.pa = s->pa;
    a.pb = s->pb;
    a.g0 = s->g0;
    a.gam_id = s->gam_id;
    a.np = n;
    a.flag = c->dflag;
    hipLaunchKernelGGL(k_uniform_f32, dim3(grid_for(n)), dim3(kBlock), 0, c->stream, a);
  }
  s->have_id = true;
  c->np = n;
  int rc = sort_loaded(c);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return check_flag(c);
}

// src/main.cpp:160-174 on the Poisson path in the chosen precision: deposit
// (rho and J) + fold, E = B = B_tm1 = 0 (Q8), the four Jacobi solves of
// Poisson_Solver::run_poisson_solver_ on FP32 arrays with the FP64 residual
// norms deciding convergence (the reference's rule), phiToE, AToB, the ghost
// refresh.  sweeps/resid as emo_init_poisson.
int fp32_init_poisson(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps_out, double *resid_out,
                      const PoissonWs &ws) {
  if (iter_max < 0) return set_err(c, EMO_ERR_CONFIG, "iter_max must be >= 0");
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  const size_t nt = (size_t)g.ntot;
  float *cr = ws.cr;
  int rc = 0;
  {
    RhoJArgs a;
    a.pa = s->pa;
    a.pb = s->pb;
    a.g0 = s->g0_active ? s->g0 : nullptr;
    a.off = s->srt.off;
    a.s = species_f32(c, 0.0);
    a.rel = c->cfg.relativity;
    a.ncell = g.ncell;
    a.cj = s->cj;
    a.cr = cr;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    hipLaunchKernelGGL(k_deposit_rhoj_f32, dim3(blocks), dim3(kBlock), 0, c->stream, a);
    DepGridArgs<float> b;
    b.g = g;
    b.cj = s->cj;
    b.cr = cr;
    for (int d = 0; d < 3; d++) b.J[d] = s->fs.J[d];
    b.rho = s->fs.rho;
    b.do_rho = 1;
    hipLaunchKernelGGL(k_deposit_grid<float>, dim3(grid_for(g.ntot)), dim3(kBlock), 0, c->stream, b);
  }
  fk_fold<float>(c, s->fs);
  s->rho_stale = false;
  for (int q = 0; q < 3; q++) {
    fk_const_fill<float>(c, s->fs.E[q], 0.0, (long long)nt);
    fk_const_fill<float>(c, s->fs.B[q], 0.0, (long long)nt);
    fk_const_fill<float>(c, s->fs.Bt[q], 0.0, (long long)nt);
  }
  const int nb = residual_blocks(c);
  double *partials = ws.partials;
  float *u[8];
  for (int q = 0; q < 8; q++) u[q] = (float *)ws.u[q];
  if (c->cfg.neutralize) fk_subtract_mean<float>(c, s->fs.rho, partials, nb, partials + 2 * nb);
  for (int q = 0; q < 8; q++) fk_const_fill<float>(c, u[q], 0.0, (long long)nt);
  // run_poisson_solver_ coefficients (src/poisson/poisson.cpp:181-185)
  const double dx_ = g.d[0], dy_ = g.d[1], dz_ = g.d[2];
  const double celldist2 = dx_ * dx_ * dy_ * dy_ + dy_ * dy_ * dz_ * dz_ + dx_ * dx_ * dz_ * dz_;
  const double ax = dy_ * dy_ * dz_ * dz_ / (2.0 * celldist2);
  const double ay = dx_ * dx_ * dz_ * dz_ / (2.0 * celldist2);
  const double az = dx_ * dx_ * dy_ * dy_ / (2.0 * celldist2);
  const double af = dx_ * dx_ * dy_ * dy_ * dz_ * dz_ / (2.0 * celldist2);
  const double sourceMult = -4 * kPi;
  const emo_config &cf = c->cfg;
  double pmax = cf.bound_phi[0], pmin = cf.bound_phi[0];
  for (int i = 1; i < 6; i++) {
    pmax = std::max(pmax, cf.bound_phi[i]);
    pmin = std::min(pmin, cf.bound_phi[i]);
  }
  double amax = cf.bound_Ax[0], amin = cf.bound_Ax[0];
  for (int i = 0; i < 6; i++) {
    for (double v : {cf.bound_Ax[i], cf.bound_Ay[i], cf.bound_Az[i]}) {
      amax = std::max(amax, v);
      amin = std::min(amin, v);
    }
  }
  const double tphi = tol >= 0 ? tol : std::max((pmax - pmin) * 1e-9, 1e-12);
  const double tA = tol >= 0 ? tol : std::max((amax - amin) * 1e-9, 1e-12);
  const double *bounds[4] = {cf.bound_phi, cf.bound_Ax, cf.bound_Ay, cf.bound_Az};
  const float *srcs[4] = {s->fs.rho, s->fs.J[0], s->fs.J[1], s->fs.J[2]};
  double resid[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t sweeps[4] = {0, 0, 0, 0};
  double hpart[2 * kMaxResidualBlocks];
  for (int sv = 0; sv < 4 && !rc; sv++) {
    float *u1 = u[2 * sv], *u2 = u[2 * sv + 1];
    const double tl = sv == 0 ? tphi : tA;
    potential_bc<float>(c, u1, bounds[sv]);
    int iternum = -1;
    bool converged = false;
    do {
      iternum++;
      const float *src = (iternum % 2 == 0) ? u1 : u2;
      float *dst = (iternum % 2 == 0) ? u2 : u1;
      fk_jacobi<float>(c, src, dst, srcs[sv], ax, ay, az, af, sourceMult);
      potential_bc<float>(c, dst, bounds[sv]);
      fk_residual<float>(c, u1, u2, partials, nb);
      hipError_t e = hipMemcpyAsync(hpart, partials, 2 * nb * sizeof(double), hipMemcpyDeviceToHost,
                                    c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        rc = set_err(c, EMO_ERR_HIP, std::string("poisson residual: ") + hipGetErrorString(e));
        break;
      }
      double maxDiff = 0.0, l2 = 0.0;
      for (int b = 0; b < nb; b++) {
        maxDiff = std::max(maxDiff, hpart[2 * b]);
        l2 += hpart[2 * b + 1];
      }
      resid[2 * sv] = maxDiff;
      resid[2 * sv + 1] = std::sqrt(l2);
      if (maxDiff < tl) converged = true;
    } while (!converged && iternum < iter_max);
    sweeps[sv] = iternum + 1;
    if (sv == 0) fk_phi_to_e<float>(c, u1, s->fs.E);  // phi1_ holds the odd-numbered iterate (Q9)
  }
  if (!rc) {
    fk_a_to_b<float>(c, u[2], u[4], u[6], s->fs.B);
    refresh_fields<float>(c, s->fs, 0.0, 0.0);
  }
  hipStreamSynchronize(c->stream);
  if (rc) return rc;
  if (sweeps_out) memcpy(sweeps_out, sweeps, sizeof(sweeps));
  if (resid_out) memcpy(resid_out, resid, sizeof(resid));
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return check_flag(c);
}

PushK fp32_push_consts(emo_ctx *c, double dt) { return push_consts(c, dt); }

int fp32_step_one(emo_ctx *c, double t, double dt, uint32_t *fwd_out) {
  ChosenState *s = c->ch;
  int rc = particle_step(c, dt);
  if (rc) return rc;
  if (fwd_out && c->np > 0)
    hipLaunchKernelGGL(k_fwd_from_inv, dim3(grid_for(c->np)), dim3(kBlock), 0, c->stream, (const uint32_t *)s->srt.perm,
                       fwd_out, (long long)c->np);
  if (s->g0_active) {
    // the stored gamma was used by this first push only (Q3): from now on
    // gamma = sqrt(1 + u^2)
    if (s->g0_owned) {
      EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
      dfree(s->g0);
    }
    s->g0_active = false;
    s->g0 = nullptr;
    s->g0_owned = true;
  }
  {
    FieldSet<float> fj = s->fs;
    fj.rho = nullptr;  // not deposited by the step (rho_stale)
    fk_fold<float>(c, fj);
  }
  s->rho_stale = true;
  fk_evolve<float>(c, s->fs, dt);
  refresh_fields<float>(c, s->fs, t, dt);
  return EMO_OK;
}

int fp32_check(emo_ctx *c) {
  EMO_HIP_CHECK_RET(c, hipGetLastError());
  return check_flag(c);
}

Fp32View fp32_view(emo_ctx *c) {
  ChosenState *s = c->ch;
  Fp32View v;
  v.pa = &s->pa;
  v.pb = &s->pb;
  v.pa_alt = &s->pa_alt;
  v.pb_alt = &s->pb_alt;
  v.g0 = &s->g0;
  v.g0_active = &s->g0_active;
  v.g0_owned = &s->g0_owned;
  v.fs = &s->fs;
  for (int q = 0; q < 3; q++) v.Btfix[q] = s->Btfix[q];
  v.off = &s->srt.off;
  return v;
}

int fp32_step(emo_ctx *c, int32_t nsteps, double t0, double dt) {
  double t = t0;
  for (int k = 0; k < nsteps; k++) {
    int rc = fp32_step_one(c, t, dt, nullptr);
    if (rc) return rc;
    t += dt;
  }
  return fp32_check(c);
}

int fp32_diag(emo_ctx *c, double *out8) {
  ChosenState *s = c->ch;
  int nb = 256;
  double *pp = c->red, *pf = c->red + 5 * nb, *res = c->red + 11 * nb;
  DiagF32Args a;
  a.pa = s->pa;
  a.pb = s->pb;
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
