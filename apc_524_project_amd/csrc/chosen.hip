// chosen.hip -- the FP32 chosen form of the PIC step for gfx950.
//
// Particles are two arrays of 16-byte records:
//   A = {word, fx, fy, fz}: word holds the species in the top 6 bits and the
//       reference's ghost-inclusive cell id ((ny+2)(nz+2) ix + (nz+2) iy + iz,
//       src/grid/oGrid.cpp:208-222) in the low 26; f in [0,1)^3 is the
//       in-cell offset (FP32);
//   B = {ux, uy, uz, id}: u (FP32) is v for Boris and gamma*v for the
//       relativistic pusher; id is the particle's int32 id.
// 32 bytes per particle, each record one 16-byte vector access.
//
// Slot layout: cell c owns the slots [base[c], base[c+1]) -- its count plus
// headroom, each range starting on a 128-byte boundary -- and holds its
// cnt[c] particles in [base[c], base[c] + cnt[c]), in ascending cell order
// overall.  The layout is rebuilt on the device, within the step, only when
// a cell's arrivals exceed its headroom.
//
// The algorithm is the reference's (same step order, same defects); only the
// representation and precision change.  With x = x0 + dx (ix - 1 + f) the
// reference's "least vertex" (src/grid/oGrid.cpp:160-162, 239-241) gives
// relpos / l = f - 1, so its linear weights (l - r)/l, r/l become 2 - f and
// f - 1 (src/particles/interpolate.cpp:25-33, src/particles/deposit.cpp:94-97).
//
// Per step, from the record buffers S (before) to D (after):
//   cellvec_f32       per cell, the 18 values the gather reads (E on the 4
//                     edges of each component, (B + B_tm1)/2 on the 2 faces)
//   push_deposit_f32  16 lanes per cell, 32 cells per block: gather + push +
//                     particle BCs of every particle, once.  A particle that
//                     stays in its cell (~93% at 64 particles per cell) is
//                     written to D at the cell's next free slot (in order)
//                     and its current is summed as 12 FP64 moments
//                     (deposit_moments), combined over the lanes in lane
//                     order.  A particle that leaves is staged in S at the
//                     front of its source range (slots this pass has already
//                     read) and adds one to its destination's arrival count
//                     and its step's bit to the destination's arrival mask.
//                     Every line written is written whole (runs padded with
//                     zeros, per-cell outputs gathered in LDS).
//   check_f32         does every cell's stayers + arrivals fit its range?
//   rebuild1_f32,     only when one does not (early exits otherwise): a new
//   relayout_f32      layout from the new counts (a scan of the check's
//                     block sums, then one per block of cells), the stayers
//                     copied to it (the staged movers stay where they are)
//   arrive_f32        one lane per cell: the cell's arrivals found in the
//                     staging ranges of the neighbours its mask names, in
//                     ascending order of their staging slot (a canonical
//                     order, independent of the atomics: the sources visited
//                     in ascending cell id, each source's movers in staging
//                     order; far movers and periodic dimensions of at most 2
//                     cells collected and sorted instead), records to D
//                     after the stayers, their currents added in that order
//   jsum_f32          per grid point, the fixed-order sum of the cell
//                     vectors; a rebuilt layout becomes current
//   fold, Yee E/B, ghost refresh   shared FP32 instances of the grid kernels
// Within a cell the new order is the stayers in their previous order, then
// the arrivals by the slot they were staged in (source cells in ascending
// order, each in its previous order): deterministic, independent of the
// headroom, with no atomics on floating-point data.  Particle traffic is
// 32 B read and 32 B written per particle, plus about 64 B per mover and the
// padding of each run to a whole line.
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
  int64_t cap = 0;        // particle capacity
  int64_t nslots = 0;     // slots of each record buffer
  int4 *pa = nullptr, *pa_alt = nullptr;  // {word, f[3]}
  int4 *pb = nullptr, *pb_alt = nullptr;  // {u[3], id}
  bool have_id = false;
  float *g0 = nullptr;  // stored gamma for the first relativistic push (Q3), by slot
  bool g0_active = false;
  bool g0_owned = true;  // g0 is this state's allocation (not a snapshot's)
  // the step deposits J only; rho (an output, not read by the step) is
  // deposited from the current particles when it is read
  bool rho_stale = false;
  // non-relativistic forms: the loaded gamma by particle id.  No Boris step
  // reads or updates it (the reference keeps the loader's value, Q3); it is
  // only handed back.  Kept in id order, so the layout never moves it.
  float *gam_id = nullptr;
  // slot layout (ncell + 1 bases, ncell counts) and its rebuild target
  int32_t *base = nullptr, *base_new = nullptr, *cnt = nullptr;
  // per step: stayers, movers, arrivals and the arrivals' direction mask
  // per cell; lay[0] = rebuild pending, lay[1] = rebuilds so far
  int32_t *ns = nullptr, *nm = nullptr, *arr = nullptr, *arrm = nullptr;
  unsigned long long *mcode = nullptr;  // the movers' codes per cell
  long long *rpart = nullptr;  // the check's per-block sums of the new ranges (2 x nblk), then the rebuild's block offsets
  // movers further than one cell: (destination, staging slot), and their count
  int2 *far = nullptr;
  int32_t *nfar = nullptr;
  // the arrival slots of crowded cells and their sorted order, at the top of
  // each range
  int32_t *akey = nullptr, *akey2 = nullptr;
  int32_t *lay = nullptr;
  // where the particle pass's lanes without a particle send their stores
  // (written, never read), so that every round issues the same stores
  int4 *sink = nullptr;
  // per-cell dense offsets (readback in array order) and scan scratch
  int32_t *dcount = nullptr, *doff = nullptr;
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
  // initial sort of loaded particles (arbitrary order)
  // the load's radix sort (keys, permutation, scratch): allocated by
  // sort_loaded for the sort and freed after it
  uint32_t *key = nullptr, *key_alt = nullptr, *perm = nullptr, *perm_alt = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int sort_bits = 0;
  float *cj = nullptr;  // 12 per cell, component-major with stride cstride (a multiple of 32)
  long long cstride = 0;
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

// A cell's slot range for n particles: n plus 6 sqrt(n) plus 16, rounded up
// to 8 slots (so every range starts on a 128-byte boundary of both record
// arrays).  A cell's count fluctuates by about sqrt(n) around its mean, so
// headroom of 6 sqrt(n) rarely fills (a 256k-cell thermal plasma at 64
// particles per cell rebuilds about once per hundred steps, against one in
// seven steps with n/4 + 8).  The tight form (n rounded up, plus 8) is used
// when the generous layout does not fit the buffers; the buffers hold
// cap + 6 sqrt(cap ncell) + 29 ncell slots, more than either needs
// (sum sqrt(n) <= sqrt(cap ncell)).
__host__ __device__ inline long long isqrt_ceil(long long n) {
  long long r = (long long)sqrt((double)n);
  while (r * r < n) r++;
  while (r > 0 && (r - 1) * (r - 1) >= n) r--;
  return r;
}
__host__ __device__ inline long long slot_cap(long long n, bool tight) {
  return tight ? ((n + 7) & ~7ll) + 8 : (n + 6 * isqrt_ceil(n) + 16 + 7) & ~7ll;
}
inline long long slots_for(long long cap_particles, long long ncell) {
  return cap_particles + 6 * isqrt_ceil(cap_particles * ncell) + 29 * ncell + 64;
}

constexpr int kCellLanes = 16;                   // lanes per cell
constexpr int kSinkVecs = 2;                     // the particle pass's sink: 2 records
constexpr long long kFarListCap = 1 << 20;       // movers further than one cell, per step
constexpr int kTileCells = kBlock / kCellLanes;  // cells per block

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
// Deposit helpers
// ---------------------------------------------------------------------------
// the 12 edge currents of one particle (Depositor::deposit_particle_J,
// src/particles/deposit.cpp:12-40, weights 2 - f and f - 1) as fp32
// products: 4 per component, the x edges first
__device__ inline void edge_terms(const float f[3], const float v[3], float jc, float w[12]) {
  const float A[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
  const float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int j = (q + 1) % 3, k = (q + 2) % 3;
    const float cq = jc * v[q];
    const float ca = cq * A[j], cb = cq * Bw[j];
    w[4 * q + 0] = ca * A[k];
    w[4 * q + 1] = cb * A[k];
    w[4 * q + 2] = cb * Bw[k];
    w[4 * q + 3] = ca * Bw[k];
  }
}

// the same, accumulated in FP64: each fp32 term is exact in a double, and a
// lane's few terms per cell sum without loss (v_cvt_f64_f32 and v_add_f64
// issue at the full vector rate: two instructions a term, where a
// compensated fp32 sum takes six)
__device__ inline void deposit_j64(double acc[12], const float f[3], const float v[3], float jc) {
  float w[12];
  edge_terms(f, v, jc, w);
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] += (double)w[q];
}

// The step's deposit as moments.  With g = f - 1 the weights are A = 2 - f
// = 1 - g and B = f - 1 = g, so the four edge currents of component q (j, k
// the other two dimensions; edge order A_j A_k, B_j A_k, B_j B_k, A_j B_k)
// are linear in four moments of c = jc v_q over the cell's particles:
// sum c, sum c g_j, sum c g_k and sum c g_j g_k.  The moments are summed in
// FP64: g = f - 1 is exact in a double, c is the fp32 product jc v_q, and
// each term is rounded once (1e-16 relative), so a cell's edge currents,
// formed from its moment sums by edges_from_moments, carry no rounding of the
// fp32 weight products.  27 vector instructions a particle (FP64 fused
// multiply-adds issue at the full vector rate).
__device__ inline void deposit_moments(double m[12], const float f[3], const float v[3], float jc) {
  const double g[3] = {(double)f[0] - 1.0, (double)f[1] - 1.0, (double)f[2] - 1.0};
  const double gp[3] = {g[1] * g[2], g[2] * g[0], g[0] * g[1]};  // g_j g_k of component q
#pragma unroll
  for (int q = 0; q < 3; q++) {
    const int j = (q + 1) % 3, k = (q + 2) % 3;
    const double cq = (double)(jc * v[q]);
    m[4 * q + 0] += cq;
    m[4 * q + 1] = fma(cq, g[j], m[4 * q + 1]);
    m[4 * q + 2] = fma(cq, g[k], m[4 * q + 2]);
    m[4 * q + 3] = fma(cq, gp[q], m[4 * q + 3]);
  }
}

// edge e (0..3) of a component from its moments (M0, Mj, Mk, Mjk)
__device__ inline double edge_from_moments(int e, double M0, double Mj, double Mk, double Mjk) {
  return e == 0 ? ((M0 - Mj) - Mk) + Mjk : e == 1 ? Mj - Mjk : e == 2 ? Mjk : Mk - Mjk;
}

// one level of a fixed-order butterfly on FP64 values within quads (DPP
// quad_perm CTRL on the two halves)
template <int CTRL>
__device__ inline void dep_level64(double acc[12]) {
#pragma unroll
  for (int q = 0; q < 12; q++) {
    const long long bits = __double_as_longlong(acc[q]);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(bits & 0xffffffffll), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(bits >> 32), CTRL, 0xF, 0xF, false);
    acc[q] += __longlong_as_double(((long long)hi << 32) | (long long)(uint32_t)lo);
  }
}

// ---------------------------------------------------------------------------
// The step's particle pass.  16 lanes per cell, 16 cells per tile, two tiles
// per block (32 cells); the lanes take a cell's particles 16 at a time, in
// order, with the records of the next two rounds in flight during a push.
// Stayers go to the output buffer at the cell's next free slots.  Movers go,
// in order, to the front of the cell's own range of the staging buffers (the
// input buffer itself, whose slots up to the current particle have been
// read, or, with a tape, separate arrays).  Each mover adds one to its
// destination's arrival count and sets the bit of its direction in the
// destination's arrival mask -- atomics whose results are never read, so no
// round waits for them; a mover that is not a neighbour of its source cell
// (far) is also listed, with its staging slot, in a global list.  The
// arrival pass finds a cell's arrivals in the staging ranges of the
// neighbours its mask names.
//
// Every 128-byte line the pass writes is written whole, because a partial
// line costs the memory a read of the line: the stayers' and the movers'
// runs are padded with zeros to the next 8-slot boundary (slots the pass has
// read, or headroom of the cell's own range), and the per-cell outputs (the
// stayer and mover counts and the stayers' current vector) are gathered in
// LDS and stored 32 cells at a time.  The cell's gather vector is read from
// LDS every round rather than held in registers, and the lanes' currents are
// summed as FP64 moments.
// ---------------------------------------------------------------------------
constexpr int kTilesPerBlock = 2;
constexpr int kBlockCells = kTilesPerBlock * kTileCells;  // 32: a 128-byte line of 32-bit per-cell values
// a cell's arrival mask: bit k = 9 (dx + 1) + 3 (dy + 1) + (dz + 1) for a
// neighbour step (dx, dy, dz) (bit 13, no step, is never set), kFarBit for a
// mover from further away
constexpr int kFarBit = 27;
// A cell's mover codes: 3 bits for each of its first kCodedMovers staged
// movers, in staging order -- 1..6 for the face steps (the mask bits of
// kFaceBit, in that order), 7 for any other move -- so that the arrival
// pass reads a face neighbour's movers for its cell without reading the
// others.
constexpr int kCodedMovers = 21;
constexpr int kFaceBit[6] = {4, 10, 12, 14, 16, 22};  // steps -x, -y, -z, +z, +y, +x

struct PushDepArgs {
  PushK k;
  const int4 *pa, *pb;        // records before the step, source layout
  int4 *pa_out, *pb_out;      // records after the step (the stayers here)
  int4 *ma, *mb;              // staging: cell c's movers' new records from slot base[c] on
  int32_t *mvsrc;             // with a tape: their source slots, or null
  const float *g0;            // the stored gamma of the first push, by source slot, or null
  const int32_t *base, *cnt;  // source layout
  int32_t *ns, *nm;           // stayers and movers per cell
  unsigned long long *mcode;  // the movers' codes per cell
  int32_t *arr, *arrm;        // arrivals per cell and their direction mask (zero on entry)
  int2 *far;                  // far movers: (destination, staging slot)
  int32_t *nfar;              // far movers listed this step (zero on entry)
  long long far_cap;
  float *cj;                  // 12 per cell, component-major with stride cstride: the stayers' currents
  long long cstride;
  uint32_t *fwd;              // with a tape: source slot -> new slot, or null
  int32_t *lay;               // lay[0]: rebuild pending (cleared here)
  int4 *sink;                 // kSinkVecs vectors, written (never read) by lanes without a particle
  long long ncell, nblocks;   // cells; blocks of kBlockCells cells
  unsigned long long *flag;
};

// A move from cell index i to j in a dimension of n cells as a neighbour
// step: -1, 0 or 1, with the periodic wrap (from index 1 to index n is a
// step of -1) where the dimension's particle BCs are periodic; 2 for a
// larger move.  Where the BCs are periodic, a move from or to a cell with
// the upper ghost index n + 1 in that dimension is also 2: a particle
// exactly at the upper boundary stays in that ghost cell (the reference's
// periodic BC moves P > n, not P = n), and the arrival pass, which wraps
// source indices into [1, n], would look for it in the cell at index 1.
__device__ inline int neighbour_step(int i, int j, int n, bool per) {
  int d = j - i;
  if (per) {
    if (i < 1 || i > n || j < 1 || j > n) return 2;
    if (n > 2) d = d == n - 1 ? -1 : (d == 1 - n ? 1 : d);
  }
  return (d >= -1 && d <= 1) ? d : 2;
}

// The cell sums of the particle pass, one component (4 moments) a round:
// each lane's FP64 moment sums in LDS, lane-major, a group of 16 lanes (one
// cell) padded so that the 16 reading lanes (4 cells x 4 moments) fall on
// different banks.
constexpr int kRedQ = 4;                             // moments per round
constexpr int kRedGroup = kCellLanes * kRedQ + 4;    // doubles per cell group
constexpr int kRedWave = 4 * kRedGroup;              // doubles per wavefront


// G0: the first relativistic push after a load, which reads the stored
// gamma; TAPE: a recomputed step writing the adjoint's map and mover
// sources.  Without either, every round issues the same stores whatever its
// particles do (stores of lanes without a particle go to the sink), so that
// the compiler's wait for the next rounds' records counts exactly.  Four
// wavefronts per SIMD (104 registers: the two rounds of prefetched records
// do not fit in the 102 of five without spilling).
template <bool REL, bool PER, bool G0, bool TAPE>
__global__ void __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(4))) k_push_deposit_f32(PushDepArgs a) {
  __shared__ float qp_s[EMO_NSPEC], jc_s[EMO_NSPEC];
  __shared__ double red[(kBlock / 64) * kRedWave];
  __shared__ float4 cvs[kTileCells][kCellVec];  // the tile's gather vectors
  __shared__ float cj_s[12][kBlockCells];       // the block's per-cell outputs
  __shared__ int ns_s[kBlockCells], nm_s[kBlockCells];
  __shared__ unsigned long long mc_s[kBlockCells];
  // species constants in LDS (a per-lane load from the argument block would
  // make every iteration wait for all loads in flight); test particles
  // deposit nothing (src/particles/particle_handler.cpp:205)
  if (threadIdx.x < EMO_NSPEC) {
    qp_s[threadIdx.x] = a.k.s.qp[threadIdx.x];
    jc_s[threadIdx.x] = a.k.s.test[threadIdx.x] ? 0.f : a.k.s.jc[threadIdx.x];
  }
  if (threadIdx.x < kBlockCells) mc_s[threadIdx.x] = 0ull;
  if (blockIdx.x == 0 && threadIdx.x == 0) a.lay[0] = 0;
  __syncthreads();
  const long long blk = xcd_tile();
  if (blk >= a.nblocks) return;
  const int sub = threadIdx.x & (kCellLanes - 1);
  const int lane = threadIdx.x & 63;
  const int gsh = lane & ~(kCellLanes - 1);  // the cell's first lane in the wavefront
  const unsigned below = (1u << sub) - 1u;
  const bool per[3] = {PER || a.k.per[0] != 0, PER || a.k.per[1] != 0, PER || a.k.per[2] != 0};
  bool outside = false;
  // (one sink address for all lanes: wavefront-uniform, so it stays in
  // scalar registers)
  int4 *const sink_a = a.sink, *const sink_b = a.sink + 1;
  const int4 zero4 = make_int4(0, 0, 0, 0);
  for (int t = 0; t < kTilesPerBlock; t++) {
    const int lc = t * kTileCells + (threadIdx.x >> 4);  // the cell among the block's 32
    const long long c = blk * kBlockCells + lc;
    const bool in = c < a.ncell;
    // every particle of the cell has its padded indices
    int cix, ciy, ciz;
    decode_fast(a.k, (int)c, cix, ciy, ciz);
    const int b = in ? a.base[c] : 0;
    const int n = in ? a.cnt[c] : 0;
    // the wavefront runs to its largest cell (ballots need every lane)
    int nmax = n;
    nmax = max(nmax, __shfl_xor(nmax, 16, 64));
    nmax = max(nmax, __shfl_xor(nmax, 32, 64));
    // the cell's gather vector in LDS, read every round rather than held in
    // 20 registers (fewer registers, more wavefronts in flight); the
    // wavefront's reads of the previous tile's vectors are done (LDS
    // operations of a wavefront complete in order)
    if (sub < kCellVec)
      cvs[threadIdx.x >> 4][sub] = n > 0 ? a.k.cv[c * kCellVec + sub] : make_float4(0.f, 0.f, 0.f, 0.f);
    wave_sync();
    const float4 *cvl = cvs[threadIdx.x >> 4];
    double acc[12];
#pragma unroll
    for (int q = 0; q < 12; q++) acc[q] = 0.0;
    int nstay = 0, nmove = 0;  // the same in all lanes of the cell
    // loads past the cell's end read its last particle (a slot of the cell's
    // own range even when it is empty), so that they are unconditional
    const int last = b + max(n - 1, 0);
    int4 A = a.pa[min(b + sub, last)], Bv = a.pb[min(b + sub, last)];
    float G = G0 ? a.g0[min(b + sub, last)] : 0.f;
    // the records two rounds ahead are in flight during a round (a staging
    // store of this round goes to a slot of this round or an earlier one)
    int4 A1 = a.pa[min(b + kCellLanes + sub, last)], B1 = a.pb[min(b + kCellLanes + sub, last)];
    float G1 = G0 ? a.g0[min(b + kCellLanes + sub, last)] : 0.f;
    // the first records in before the loop: the compiler's waits in the loop
    // then count only the loop's own loads (otherwise every round would wait
    // as the first one must)
    __builtin_amdgcn_s_waitcnt(0);
    for (int j0 = 0; j0 < nmax; j0 += kCellLanes) {
      // (keeps the gather vector's LDS reads in the loop)
      asm volatile("" ::: "memory");
      const CellVec V{cvl[0], cvl[1], cvl[2], cvl[3], cvl[4]};
      const int jn = min(b + j0 + 2 * kCellLanes + sub, last);
      const int4 An = a.pa[jn], Bn = a.pb[jn];
      const float Gn = G0 ? a.g0[jn] : 0.f;
      const int src = b + j0 + sub;
      const bool valid = j0 + sub < n;
      const int sp = (int)((uint32_t)A.x >> kSpeciesShift);
      Pushed o;
      push_particle<REL, PER>(a.k, V, cix, ciy, ciz, as_f(A.y), as_f(A.z), as_f(A.w), as_f(Bv.x), as_f(Bv.y),
                              as_f(Bv.z), qp_s[sp], G0, G, o, nullptr);
      outside = outside || (valid && o.outside);
      const bool stay = valid && o.cell == (int)c;
      const bool move = valid && !stay;
      const unsigned gm = (unsigned)(__ballot(stay) >> gsh) & 0xffffu;
      const unsigned mm = (unsigned)(__ballot(move) >> gsh) & 0xffffu;
      const int4 Ao = make_int4((int)(((uint32_t)sp << kSpeciesShift) | (uint32_t)o.cell), as_i(o.f[0]),
                                as_i(o.f[1]), as_i(o.f[2]));
      const int4 Bo = make_int4(as_i(o.u[0]), as_i(o.u[1]), as_i(o.u[2]), Bv.w);
      // a stayer to the output at the cell's next free slot, a mover to its
      // staging slot (at most this particle's own: already read); lanes
      // without a particle to the sink
      const int rs = stay ? b + nstay + __popc(gm & below) : b + nmove + __popc(mm & below);
      int4 *const da = valid ? (stay ? a.pa_out : a.ma) + rs : sink_a;
      int4 *const db = valid ? (stay ? a.pb_out : a.mb) + rs : sink_b;
      *da = Ao;
      *db = Bo;
      if (TAPE) {
        // (a mover's map entry is written by the arrival pass; this one
        // only completes the line)
        if (valid) a.fwd[src] = stay ? (uint32_t)rs : 0xffffffffu;
        if (move) a.mvsrc[rs] = src;
      }
      if (stay) deposit_moments(acc, o.f, o.v, jc_s[sp]);
      if (move) {
        // the destination's arrival count and mask (results not read)
        const int sx = neighbour_step(cix, o.ijk[0], a.k.g.n[0], per[0]);
        const int sy = neighbour_step(ciy, o.ijk[1], a.k.g.n[1], per[1]);
        const int sz = neighbour_step(ciz, o.ijk[2], a.k.g.n[2], per[2]);
        const bool far = sx == 2 || sy == 2 || sz == 2;
        atomicAdd(&a.arr[o.cell], 1);
        atomicOr(&a.arrm[o.cell], far ? 1 << kFarBit : 1 << (9 * (sx + 1) + 3 * (sy + 1) + (sz + 1)));
        // the mover's code in its cell's word (an LDS atomic, no wait)
        const int mi = rs - b;
        if (mi < kCodedMovers) {
          const int nz = (sx != 0) + (sy != 0) + (sz != 0);
          const int code = (far || nz != 1) ? 7 : (sx < 0 ? 1 : sy < 0 ? 2 : sz < 0 ? 3 : sz > 0 ? 4 : sy > 0 ? 5 : 6);
          atomicOr(&mc_s[lc], (unsigned long long)code << (3 * mi));
        }
        if (far) {
          // (only a particle moving more than a cell in one step)
          const int k = atomicAdd(a.nfar, 1);
          if (k < a.far_cap)
            a.far[k] = make_int2(o.cell, rs);
          else
            atomicOr(a.flag, kFlagFarList);
        }
      }
      nstay += __popc(gm);
      nmove += __popc(mm);
      A = A1;
      Bv = B1;
      G = G1;
      A1 = An;
      B1 = Bn;
      G1 = Gn;
    }
    // zeros to the next 8-slot boundary after the stayers (output) and the
    // staged movers: whole 128-byte lines (every range starts on an 8-slot
    // boundary and holds at least round8(n) slots)
    {
      const int es = b + nstay, pe = (es + 7) & ~7;
      if (in && es + sub < pe) {
        a.pa_out[es + sub] = zero4;
        a.pb_out[es + sub] = zero4;
      }
      const int em = b + nmove, pm = (em + 7) & ~7;
      if (in && em + sub < pm) {
        a.ma[em + sub] = zero4;
        a.mb[em + sub] = zero4;
      }
      if (TAPE) {
        // the map and the movers' sources, to whole 32-byte sectors
        const int en = b + n, pn = (en + 7) & ~7;
        if (in && en + sub < pn) a.fwd[en + sub] = 0xffffffffu;
        if (in && em + sub < pm) a.mvsrc[em + sub] = -1;
      }
    }
    // the cell's sum over its 16 lanes in lane order, in FP64: the lanes' sums
    // through LDS one component (4 moments) at a time, then one lane per
    // (cell, moment), and the quad of a cell's 4 moments forms its 4 edges
    static_assert(kRedQ == 4, "a reduction round holds one component's 4 moments");
    double *rw = red + (threadIdx.x >> 6) * kRedWave;
#pragma unroll
    for (int q0 = 0; q0 < 12; q0 += kRedQ) {
      double *mine = rw + (lane >> 4) * kRedGroup + (lane & 15) * kRedQ;
#pragma unroll
      for (int q = 0; q < kRedQ; q++) mine[q] = acc[q0 + q];
      wave_sync();
      if (lane < 4 * kRedQ) {
        const int grp = lane / kRedQ, q = lane - kRedQ * grp;
        const double *col = rw + grp * kRedGroup + q;
        double sum = col[0];
#pragma unroll
        for (int l = 1; l < kCellLanes; l++) sum += col[kRedQ * l];
        const int qb = lane & ~3;
        const double M0 = __shfl(sum, qb, 64), Mj = __shfl(sum, qb + 1, 64);
        const double Mk = __shfl(sum, qb + 2, 64), Mjk = __shfl(sum, qb + 3, 64);
        cj_s[q0 + q][t * kTileCells + (threadIdx.x >> 6) * 4 + grp] = (float)edge_from_moments(q, M0, Mj, Mk, Mjk);
      }
      wave_sync();
    }
    if (sub == 0) {
      ns_s[lc] = nstay;
      nm_s[lc] = nmove;
    }
  }
  if (__ballot(outside) != 0ull && lane == 0) atomicOr(a.flag, kFlagOutOfGrid);
  __syncthreads();
  // the block's per-cell outputs, 32 cells (128 bytes) per array
  const long long c0 = blk * kBlockCells;
  const int ncb = (int)min((long long)kBlockCells, a.ncell - c0);
  for (int i = threadIdx.x; i < 12 * kBlockCells; i += kBlock) {
    const int q = i / kBlockCells, l = i - q * kBlockCells;
    if (l < ncb) a.cj[q * a.cstride + c0 + l] = cj_s[q][l];
  }
  if (threadIdx.x < ncb) {
    a.ns[c0 + threadIdx.x] = ns_s[threadIdx.x];
    a.nm[c0 + threadIdx.x] = nm_s[threadIdx.x];
    a.mcode[c0 + threadIdx.x] = mc_s[threadIdx.x];
  }
}

// Does every cell's range hold its stayers and arrivals?  If not, a rebuild
// is pending for this step.  Also the sums, per block of cells, of the
// generous and of the tight ranges of the new counts (for a rebuild).
constexpr int kRebaseThreads = 1024;
__global__ void __launch_bounds__(kBlock) k_check_f32(const int32_t *ns, const int32_t *arr, const int32_t *base,
                                                      long long ncell, long long *part, long long nblk,
                                                      int32_t *lay) {
  __shared__ long long sh[2][kBlock / 64];
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  long long g = 0, t = 0;
  if (c < ncell) {
    const int n = ns[c] + arr[c];
    if (n > base[c + 1] - base[c]) atomicOr(&lay[0], 1);
    g = slot_cap(n, false);
    t = slot_cap(n, true);
  }
  for (int o = 32; o > 0; o >>= 1) {
    g += __shfl_xor(g, o, 64);
    t += __shfl_xor(t, o, 64);
  }
  if ((threadIdx.x & 63) == 0) {
    sh[0][threadIdx.x >> 6] = g;
    sh[1][threadIdx.x >> 6] = t;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    part[blockIdx.x] = sh[0][0] + sh[0][1] + sh[0][2] + sh[0][3];
    part[nblk + blockIdx.x] = sh[1][0] + sh[1][1] + sh[1][2] + sh[1][3];
  }
}

// Rebuild, step 1 of 2 (one launch; every block returns at once unless a
// rebuild is pending).  Block 0 scans the check's block sums -- generous
// ranges if they fit the buffers, else tight ones (which always fit) -- into
// the first new slot of each check block's 256 cells.  The other blocks, 64
// cells at a time, copy each cell's stayers from the output buffer (old
// layout) to the staging buffers' free slots behind its movers -- exactly as
// many as it has stayers, since its range there held stayers + movers.
struct RelayoutArgs {
  const int32_t *base, *cnt, *ns, *nm, *arr, *lay;
  int32_t *base_new;
  const long long *boff;  // the first new slot of each check block's cells
  int4 *pa_out, *pb_out;  // the step's output buffer
  int4 *ma, *mb;          // staging
  uint32_t *fwd;          // with a tape: source slot -> new slot, or null
  long long ncell;
};
struct RebuildArgs {
  RelayoutArgs r;
  long long *part;  // the check's per-block sums of the new ranges (generous, tight); then the blocks' offsets
  long long nblk, nslots;
  int32_t *lay;     // lay[2]: the rebuilt ranges are tight
  unsigned long long *flag;
};
constexpr int kRestageCells = kRebaseThreads / kCellLanes;
constexpr unsigned kRestageBlocks = 2048;  // at most (the copy loops over the cells)

__global__ void __launch_bounds__(kRebaseThreads) k_rebuild1_f32(RebuildArgs a) {
  if (a.lay[0] == 0) return;
  if (blockIdx.x > 0) {
    const RelayoutArgs &r = a.r;
    const int sub = threadIdx.x & (kCellLanes - 1);
    const long long step = (long long)(gridDim.x - 1) * kRestageCells;
    for (long long c = (long long)(blockIdx.x - 1) * kRestageCells + (threadIdx.x >> 4); c < r.ncell; c += step) {
      const int b = r.base[c], sm = b + r.nm[c], n = r.ns[c];
      for (int k = sub; k < n; k += kCellLanes) {
        r.ma[sm + k] = r.pa_out[b + k];
        r.mb[sm + k] = r.pb_out[b + k];
      }
    }
    return;
  }
  __shared__ long long sh[kRebaseThreads];
  __shared__ int use_tight;
  const long long nblk = a.nblk, ncell = a.r.ncell;
  const int t = threadIdx.x;
  const long long per = (nblk + kRebaseThreads - 1) / kRebaseThreads;
  const long long b0 = min((long long)t * per, nblk), b1 = min(b0 + per, nblk);
  long long total = 0, s = 0;
  int pass = 0;
  for (; pass < 2; pass++) {
    s = 0;
    for (long long b = b0; b < b1; b++) s += a.part[pass * nblk + b];
    sh[t] = s;
    __syncthreads();
    for (int o = 1; o < kRebaseThreads; o <<= 1) {
      const long long v = t >= o ? sh[t - o] : 0;
      __syncthreads();
      sh[t] += v;
      __syncthreads();
    }
    total = sh[kRebaseThreads - 1];
    if (t == 0) use_tight = pass;
    __syncthreads();
    if (total <= a.nslots) break;
    __syncthreads();
  }
  if (total > a.nslots) {
    // (cannot happen: the buffers hold the tight layout of any counts) no
    // rebuild, so that every later access stays in the current layout (the
    // other blocks' copies went to free staging slots)
    if (t == 0) {
      atomicOr(a.flag, kFlagLayout);
      a.lay[0] = 0;
    }
    return;
  }
  const int used = use_tight;
  long long run = sh[t] - s;
  for (long long b = b0; b < b1; b++) {
    a.part[2 * nblk + b] = run;
    run += a.part[used * nblk + b];
  }
  if (t == kRebaseThreads - 1) {
    a.r.base_new[ncell] = (int32_t)total;
    a.lay[1] += 1;
    a.lay[2] = used;
  }
}

// Rebuild, step 2 of 2, per check block of 256 cells: each cell's new base
// (the block's offset plus a scan of its cells' new ranges), then the
// stayers to their new slots in the output buffer (with a tape, their map
// entries moved along).  The movers stay where they are staged (the old
// layout's ranges), where the arrival pass finds them.
__global__ void __launch_bounds__(kBlock) k_relayout_f32(RelayoutArgs a) {
  if (a.lay[0] == 0) return;
  __shared__ int bn_s[kBlock];
  __shared__ long long wsum[kBlock / 64];
  const bool tight = a.lay[2] != 0;
  const long long c0 = (long long)blockIdx.x * kBlock;
  {
    const long long c = c0 + threadIdx.x;
    const int lane = threadIdx.x & 63;
    const long long cap = c < a.ncell ? slot_cap(a.ns[c] + a.arr[c], tight) : 0;
    long long x = cap;  // inclusive scan over the wavefront
    for (int o = 1; o < 64; o <<= 1) {
      const long long y = __shfl_up(x, o, 64);
      if (lane >= o) x += y;
    }
    if (lane == 63) wsum[threadIdx.x >> 6] = x;
    __syncthreads();
    long long off = a.boff[blockIdx.x];
    for (int w = 0; w < (int)(threadIdx.x >> 6); w++) off += wsum[w];
    const int bn = (int)(off + x - cap);
    bn_s[threadIdx.x] = bn;
    if (c < a.ncell) a.base_new[c] = bn;
    __syncthreads();
  }
  const int sub = threadIdx.x & (kCellLanes - 1);
  for (int lc = threadIdx.x >> 4; lc < kBlock; lc += kBlock / kCellLanes) {
    const long long c = c0 + lc;
    if (c >= a.ncell) break;
    const int b = a.base[c], m = a.nm[c], n = a.ns[c], bn = bn_s[lc];
    for (int k = sub; k < n; k += kCellLanes) {
      a.pa_out[bn + k] = a.ma[b + m + k];
      a.pb_out[bn + k] = a.mb[b + m + k];
    }
    if (a.fwd) {
      // stayers' entries point into [b, b + n); movers' entries are
      // rewritten by arrive_f32
      const int cn = a.cnt[c];
      for (int j = sub; j < cn; j += kCellLanes) {
        const uint32_t f = a.fwd[b + j];
        if (f >= (uint32_t)b && f < (uint32_t)(b + n)) a.fwd[b + j] = f - (uint32_t)b + (uint32_t)bn;
      }
    }
  }
}

// Per cell (one lane, 128 cells per block, XCD-aware order): the cell's
// arrivals are the movers staged in the ranges of the neighbours its arrival
// mask names (bit k: the neighbour one step of -(dx, dy, dz) away, the
// periodic wrap applied where the particle BCs are periodic) whose new cell
// is this one, and, when a far mover was flagged, the far list's entries for
// this cell.  They are taken in ascending order of their staging slot
// (source cell, then order within it: the previous array order), sorted by
// insertion in LDS for up to kArrFast arrivals, ranked in global memory
// beyond.  The rank-r arrival is written after the stayers and its current
// (v = u / gamma, gamma from u as the next push computes it) added to the
// cell's moments in that order; the last line is padded with zeros.  Then
// the cell's current vector (stayers + arrivals) and new count, every cell
// of the block (whole lines); the per-step counters are reset for the next
// step.
constexpr int kArrBlock = 128;
constexpr int kArrFast = 16;

struct ArriveArgs {
  Geo g;
  int per[3];
  int rel;
  int dedupe;  // a periodic dimension of at most 2 cells: two steps can reach one neighbour
  const int32_t *base, *base_new, *lay;
  const int32_t *ns, *nm;
  const unsigned long long *mcode;
  int32_t *arr, *arrm, *cnt;
  int32_t *akey, *akey2;  // more than kArrFast arrivals: their slots, and sorted, at the top of the range
  const int2 *far;
  const int32_t *nfar;
  long long far_cap;
  const int4 *ma, *mb;
  const int32_t *mvsrc;  // with a tape
  int4 *pa_out, *pb_out;
  uint32_t *fwd;
  float *cj;
  long long cstride;
  float jc[EMO_NSPEC];  // 0 for test particles
  long long ncell;
  unsigned long long *flag;
};

// the padded index of the neighbour one step of -s away in a dimension (the
// source of an arrival that took step s), wrapped where periodic (the
// caller checks it against the padded grid)
__device__ inline int source_index(int i, int s, int n, bool per) {
  int j = i - s;
  if (per) j = j < 1 ? j + n : (j > n ? j - n : j);
  return j;
}

__global__ void __launch_bounds__(kArrBlock) k_arrive_f32(ArriveArgs a) {
  __shared__ int cand[kArrBlock][kArrFast + 1];
  __shared__ float jc_s[EMO_NSPEC];
  if (threadIdx.x < EMO_NSPEC) jc_s[threadIdx.x] = a.jc[threadIdx.x];
  __syncthreads();
  const long long c = (long long)xcd_tile() * kArrBlock + threadIdx.x;
  const bool in = c < a.ncell;
  // the loads that do not depend on each other issued together: both
  // layouts' bounds (the rebuilt one is current after a rebuild), the counts,
  // and the stayers' currents this pass adds to
  const bool rebuilt = a.lay[0] != 0;
  const int bo0 = in ? a.base[c] : 0, bo1 = in ? a.base[c + 1] : 0;
  const int bn0 = in ? a.base_new[c] : 0, bn1 = in ? a.base_new[c + 1] : 0;
  int na = in ? a.arr[c] : 0;
  const int mask = in ? a.arrm[c] : 0;
  const int nsc = in ? a.ns[c] : 0;
  float cj0[12];
#pragma unroll
  for (int q = 0; q < 12; q++) cj0[q] = in ? a.cj[q * a.cstride + c] : 0.f;
  const int s0 = (rebuilt ? bn0 : bo0) + nsc;
  const int e1 = rebuilt ? bn1 : bo1;
  if (in && s0 + na > e1) {  // (the check has rebuilt the layout, so only after a flagged error)
    atomicOr(a.flag, kFlagMigrate);
    na = max(e1 - s0, 0);
  }
  double acc[12];
#pragma unroll
  for (int q = 0; q < 12; q++) acc[q] = 0.0;
  if (na > 0) {
    int ix, iy, iz;
    decode_cell(a.g, (int)c, &ix, &iy, &iz);
    // the source cell of step k, or -1 outside the padded grid
    auto source_of = [&](int k) -> int {
      const int sx = source_index(ix, k / 9 - 1, a.g.n[0], a.per[0] != 0);
      const int sy = source_index(iy, (k / 3) % 3 - 1, a.g.n[1], a.per[1] != 0);
      const int sz = source_index(iz, k % 3 - 1, a.g.n[2], a.per[2] != 0);
      if (sx < 0 || sy < 0 || sz < 0 || sx >= a.g.nc[0] || sy >= a.g.nc[1] || sz >= a.g.nc[2]) return -1;
      return (sx * a.g.nc[1] + sy) * a.g.nc[2] + sz;
    };
    // the face code (1..6) of step k, 0 for another step
    auto face_of = [&](int k) -> int {
      int face = 0;
#pragma unroll
      for (int f = 0; f < 6; f++) face = k == kFaceBit[f] ? f + 1 : face;
      return face;
    };
    // bit 3i for each of source s's first kCodedMovers movers with code face
    auto coded_hits = [&](int s, int ms, int face) -> unsigned long long {
      constexpr unsigned long long kLow = 0x1249249249249249ull;  // bit 3i, i < 21
      const unsigned long long x = a.mcode[s] ^ (kLow * (unsigned long long)face);
      const unsigned long long hit = ~(x | (x >> 1) | (x >> 2)) & kLow;
      const int coded = min(ms, kCodedMovers);
      return hit & (coded >= kCodedMovers ? ~0ull : (1ull << (3 * coded)) - 1ull);
    };
    // the arrival of rank r: its staged record written after the stayers,
    // its current (v = u / gamma, gamma from u as the next push computes
    // it) added to the cell's moments
    auto take = [&](int st, int r) {
      const int4 A = a.ma[st], Bv = a.mb[st];
      const int pos = s0 + r;
      a.pa_out[pos] = A;
      a.pb_out[pos] = Bv;
      if (a.fwd) a.fwd[a.mvsrc[st]] = (uint32_t)pos;
      const float f[3] = {as_f(A.y), as_f(A.z), as_f(A.w)};
      const float ux = as_f(Bv.x), uy = as_f(Bv.y), uz = as_f(Bv.z);
      const float ig = a.rel ? __builtin_amdgcn_rsqf(fmaf(ux, ux, fmaf(uy, uy, fmaf(uz, uz, 1.f)))) : 1.f;
      const float v[3] = {ux * ig, uy * ig, uz * ig};
      deposit_moments(acc, f, v, jc_s[(uint32_t)A.x >> kSpeciesShift]);
    };
    int nc = 0;
    if (!a.dedupe && !(mask & (1 << kFarBit))) {
      // The sources in ascending cell-id order -- per dimension the three
      // steps ordered by their (wrapped) source index, the x steps outermost
      // -- and each source's movers in staging order: the arrivals come out
      // in ascending staging slot, the canonical order, and are placed as
      // they are found (a face source's by their codes).
      int ord[3][3];
      const int idx[3] = {ix, iy, iz};
#pragma unroll
      for (int d = 0; d < 3; d++) {
        const bool per = a.per[d] != 0;
        const bool lo = per && idx[d] == 1, hi = per && idx[d] == a.g.n[d];
        ord[d][0] = lo ? 0 : (hi ? -1 : 1);   // sources i - 1, i, i + 1 in ascending order,
        ord[d][1] = lo ? -1 : (hi ? 1 : 0);   // the wrapped one (index n or 1) moved to its
        ord[d][2] = lo ? 1 : (hi ? 0 : -1);   // place (every periodic dimension has n >= 3)
      }
      for (int ua = 0; ua < 3; ua++)
        for (int ub = 0; ub < 3; ub++)
          for (int uc = 0; uc < 3; uc++) {
            const int k = 9 * (ord[0][ua] + 1) + 3 * (ord[1][ub] + 1) + (ord[2][uc] + 1);
            if (!((mask >> k) & 1)) continue;
            const int s = source_of(k);
            if (s < 0) continue;
            const int bs = a.base[s], ms = a.nm[s];
            const int face = face_of(k);
            int j0 = 0;
            if (face) {
              for (unsigned long long hit = coded_hits(s, ms, face); hit != 0ull; hit &= hit - 1ull) {
                if (nc < na) take(bs + (__ffsll((long long)hit) - 1) / 3, nc);
                nc++;
              }
              j0 = kCodedMovers;
            }
            for (int j = j0; j < ms; j++)
              if ((a.ma[bs + j].x & kCellMask) == (int)c) {
                if (nc < na) take(bs + j, nc);
                nc++;
              }
          }
      if (nc != na) {  // (the counts and the staged movers disagree: flagged)
        atomicOr(a.flag, kFlagArrivals);
        na = min(na, nc);
      }
    } else {
      // A far mover (listed by slot) or a periodic dimension of at most 2
      // cells (where two steps can reach one source): the candidates
      // collected, sorted by insertion in LDS for up to kArrFast, ranked in
      // global memory beyond, then taken in ascending order.
      const bool fast = na <= kArrFast;
      const int k0 = e1 - na;  // (slow path) the range's top na slots of akey and akey2
      int *cl = cand[threadIdx.x];
      auto add = [&](int slot) {
        if (nc < na) {
          if (fast) {
            int i = nc;
            while (i > 0 && cl[i - 1] > slot) {
              cl[i] = cl[i - 1];
              i--;
            }
            cl[i] = slot;
          } else {
            a.akey[k0 + nc] = slot;
          }
        }
        nc++;
      };
      for (int m = mask & ((1 << kFarBit) - 1); m != 0; m &= m - 1) {
        const int k = __ffs(m) - 1;
        const int s = source_of(k);
        if (s < 0) continue;
        if (a.dedupe) {
          // another step of the mask that reaches the same source was scanned
          bool seen = false;
          for (int m2 = mask & ((1 << k) - 1); m2 != 0 && !seen; m2 &= m2 - 1) seen = source_of(__ffs(m2) - 1) == s;
          if (seen) continue;
        }
        const int bs = a.base[s], ms = a.nm[s];
        const int face = a.dedupe ? 0 : face_of(k);
        int j0 = 0;
        if (face) {
          for (unsigned long long hit = coded_hits(s, ms, face); hit != 0ull; hit &= hit - 1ull)
            add(bs + (__ffsll((long long)hit) - 1) / 3);
          j0 = kCodedMovers;
        }
        for (int j = j0; j < ms; j++)
          if ((a.ma[bs + j].x & kCellMask) == (int)c) add(bs + j);
      }
      if (mask & (1 << kFarBit)) {
        const int nf = (int)min((long long)*a.nfar, a.far_cap);
        for (int j = 0; j < nf; j++) {
          const int2 e = a.far[j];
          if (e.x == (int)c) add(e.y);
        }
      }
      if (nc != na) {  // (the counts and the staged movers disagree: flagged)
        atomicOr(a.flag, kFlagArrivals);
        na = min(na, nc);
      }
      if (!fast) {
        for (int i = 0; i < na; i++) {
          const int key = a.akey[k0 + i];
          int r = 0;
          for (int jj = 0; jj < na; jj++) r += a.akey[k0 + jj] < key ? 1 : 0;
          a.akey2[k0 + r] = key;
        }
      }
      for (int r = 0; r < na; r++) take(fast ? cl[r] : a.akey2[k0 + r], r);
    }
    // zeros to the next 8-slot boundary: a whole last line
    const int4 zero4 = make_int4(0, 0, 0, 0);
    for (int p = s0 + na; p & 7; p++) {
      a.pa_out[p] = zero4;
      a.pb_out[p] = zero4;
    }
  }
  if (in) {
#pragma unroll
    for (int q = 0; q < 12; q++) {
      const int q0 = q & ~3;
      const double w = edge_from_moments(q & 3, acc[q0], acc[q0 + 1], acc[q0 + 2], acc[q0 + 3]);
      a.cj[q * a.cstride + c] = (float)((double)cj0[q] + w);
    }
    a.cnt[c] = nsc + na;
    a.arr[c] = 0;
    a.arrm[c] = 0;
  }
}

// ---------------------------------------------------------------------------
// rho + J of the current state, for the Poisson initialisation
// (src/main.cpp:160-166): the 12 edge currents and the 8 vertex charges of
// Depositor::deposit_particle_J / _Rho (src/particles/deposit.cpp:12-62)
// with the weights 2 - f and f - 1, 16 lanes per cell, FP64 sums and a
// fixed butterfly.
// ---------------------------------------------------------------------------
struct RhoJArgs {
  const int4 *pa, *pb;
  const float *g0;
  const int32_t *base, *cnt;
  SpeciesF32 s;
  int rel;
  long long ncell;
  float *cj, *cr;
  long long sj;  // component stride of cj (cr: ncell)
};

__global__ void __launch_bounds__(kBlock) k_deposit_rhoj_f32(RhoJArgs a) {
  const int sub = threadIdx.x & 15;
  const long long c = (long long)blockIdx.x * (kBlock / 16) + (threadIdx.x >> 4);
  double acc[20];
#pragma unroll
  for (int q = 0; q < 20; q++) acc[q] = 0.0;
  if (c < a.ncell) {
    const int b = a.base[c], e = b + a.cnt[c];
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
      deposit_j64(acc, f, v, jc);
      const float Aw[3] = {2.f - f[0], 2.f - f[1], 2.f - f[2]};
      const float Bw[3] = {f[0] - 1.f, f[1] - 1.f, f[2] - 1.f};
      // vertex order of Depositor::deposit_particle_Rho
      const float w8[8] = {Aw[0] * Aw[1] * Aw[2], Bw[0] * Aw[1] * Aw[2], Bw[0] * Bw[1] * Aw[2],
                           Aw[0] * Bw[1] * Aw[2], Aw[0] * Bw[1] * Bw[2], Bw[0] * Bw[1] * Bw[2],
                           Bw[0] * Aw[1] * Bw[2], Aw[0] * Aw[1] * Bw[2]};
#pragma unroll
      for (int r = 0; r < 8; r++) acc[12 + r] += (double)(jc * w8[r]);
    }
  }
  // the 16 lanes' sums in a fixed order (xor butterfly), in FP64
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) {
#pragma unroll
    for (int q = 0; q < 20; q++) acc[q] += __shfl_xor(acc[q], off, 64);
  }
  if (sub == 0 && c < a.ncell) {
#pragma unroll
    for (int q = 0; q < 12; q++) a.cj[q * a.sj + c] = (float)acc[q];
#pragma unroll
    for (int q = 0; q < 8; q++) a.cr[q * a.ncell + c] = (float)acc[12 + q];
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

// The sorted loaded particles (dense, cell c at doff[c] .. doff[c] + cnt[c])
// into the slot layout, with the stored gamma.
struct ScatterArgs {
  const int4 *pa, *pb;
  const float *g0;
  int4 *pa_out, *pb_out;
  float *g0_out;
  const int32_t *doff, *base;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_scatter_slots_f32(ScatterArgs a) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= a.np) return;
  const int4 A = a.pa[i];
  const int c = A.x & kCellMask;
  const long long s = (long long)a.base[c] + (i - a.doff[c]);
  a.pa_out[s] = A;
  a.pb_out[s] = a.pb[i];
  if (a.g0) a.g0_out[s] = a.g0[i];
}

// Per-cell iteration over the particles in array order (cell by cell, in
// ascending cell order): 16 lanes per cell, particle k of cell c at slot
// base[c] + k and array index doff[c] + k.
struct OutArgs {
  Geo g;
  const int4 *pa, *pb;
  const float *g0;
  const float *gam_id;  // non-relativistic: the loaded gamma by id, or null
  const float4 *cv;     // with fld: the current grid's gather vectors
  const int32_t *base, *cnt, *doff;
  int rel;
  double *x[3], *v[3], *gam;
  double *fld[6];       // the gathered E, B per particle (Field_part), or null
  int32_t *sp;
  int32_t *id;
  int32_t *cell;
  double *trk;          // the particles with id < ntrk by id: 7 per id (x, v, found), or null
  long long ntrk;
  long long np;
};

__global__ void __launch_bounds__(kBlock) k_output_f32(OutArgs a) {
  const long long c = (long long)blockIdx.x * kTileCells + (threadIdx.x >> 4);
  const int sub = threadIdx.x & (kCellLanes - 1);
  const Geo &g = a.g;
  if (c >= g.ncell) return;
  const int n = a.cnt[c];
  if (n == 0) return;
  int ic[3];
  decode_cell(g, (int)c, &ic[0], &ic[1], &ic[2]);
  for (int k = sub; k < n; k += kCellLanes) {
    const long long s = (long long)a.base[c] + k, i = (long long)a.doff[c] + k;
    if (i >= a.np) break;  // (only if particles were lost, which is flagged)
    const int4 A = a.pa[s], Bv = a.pb[s];
    if (a.cell) a.cell[i] = (int32_t)c;
    const double uu[3] = {as_f(Bv.x), as_f(Bv.y), as_f(Bv.z)};
    const double ff[3] = {as_f(A.y), as_f(A.z), as_f(A.w)};
    double gam = 1.0;
    if (a.rel) gam = a.g0 ? (double)a.g0[s] : sqrt(1.0 + uu[0] * uu[0] + uu[1] * uu[1] + uu[2] * uu[2]);
    double xo[3], vo[3];
    for (int d = 0; d < 3; d++) {
      xo[d] = g.x0[d] + g.d[d] * ((double)(ic[d] - 1) + ff[d]);
      vo[d] = a.rel ? uu[d] / gam : uu[d];
      if (a.x[d]) a.x[d][i] = xo[d];
      if (a.v[d]) a.v[d][i] = vo[d];
    }
    if (a.trk && Bv.w >= 0 && Bv.w < a.ntrk) {
      double *o = a.trk + 7 * (long long)Bv.w;
      for (int d = 0; d < 3; d++) {
        o[d] = xo[d];
        o[3 + d] = vo[d];
      }
      o[6] = 1.0;
    }
    if (!a.rel && a.gam_id && Bv.w >= 0 && Bv.w < a.np) gam = a.gam_id[Bv.w];
    if (a.gam) a.gam[i] = gam;
    if (a.fld[0]) {
      const float4 *cv = a.cv + c * kCellVec;
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
}

struct DiagF32Args {
  const int4 *pa, *pb;
  const float *g0;
  const int32_t *base, *cnt;
  int rel;
  SpeciesParams s;
  long long ncell;
  double *partials;
};

// FP64 sums over the particles (OutputBoxQuantities::output, src/IO/output.cpp:
// 76-88), cell by cell, 16 lanes per cell, a fixed tree per block
__global__ void __launch_bounds__(kBlock) k_diag_particles_f32(DiagF32Args a) {
  __shared__ double sh[5][kBlock];
  double acc[5] = {0, 0, 0, 0, 0};
  const int sub = threadIdx.x & (kCellLanes - 1);
  for (long long c = (long long)blockIdx.x * kTileCells + (threadIdx.x >> 4); c < a.ncell;
       c += (long long)gridDim.x * kTileCells) {
    const int b = a.base[c], n = a.cnt[c];
    for (int k = sub; k < n; k += kCellLanes) {
      const int i = b + k;
      const int sp = (int)((uint32_t)a.pa[i].x >> kSpeciesShift);
      const int4 Bv = a.pb[i];
      const double m = a.s.m[sp];
      const double ux = as_f(Bv.x), uy = as_f(Bv.y), uz = as_f(Bv.z);
      double ig = 1.0;
      if (a.rel) ig = 1.0 / (a.g0 ? (double)a.g0[i] : sqrt(1.0 + ux * ux + uy * uy + uz * uz));
      const double vx = ux * ig, vy = uy * ig, vz = uz * ig;
      acc[0] += m * vx;
      acc[1] += m * vy;
      acc[2] += m * vz;
      acc[3] += 0.5 * m * (vx * vx + vy * vy + vz * vz);
      acc[4] += a.s.q[sp];
    }
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

// counts per cell -> slot ranges (before the scan into bases)
__global__ void k_slot_caps(const int32_t *count, int32_t *caps, long long ncell, int tight) {
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncell) caps[c] = (int32_t)slot_cap(count[c], tight != 0);
  if (c == ncell) caps[c] = 0;
}

__global__ void k_copy_i32(const int32_t *src, int32_t *dst, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
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
  if (c->hflag[0] & kFlagMigrate)
    return set_err(c, EMO_ERR_INVARIANT, "a moving particle found no slot in its new cell");
  if (c->hflag[0] & kFlagFarList)
    return set_err(c, EMO_ERR_INVARIANT, "more particles moved further than one cell in a step than the far list holds");
  if (c->hflag[0] & kFlagArrivals)
    return set_err(c, EMO_ERR_INVARIANT, "a cell's arrivals found in its neighbours disagree with its arrival count");
  if (c->hflag[0] & kFlagLayout)
    return set_err(c, EMO_ERR_INVARIANT, "the rebuilt slot layout does not fit the particle buffers");
  return EMO_OK;
}

// exclusive scan of n int32 values (device) into out
int scan_i32(emo_ctx *c, const int32_t *in, int32_t *out, long long n) {
  ChosenState *s = c->ch;
  size_t bytes = s->scan_tmp_bytes;
  hipError_t e = hipcub::DeviceScan::ExclusiveSum(s->scan_tmp, bytes, in, out, (int)n, c->stream);
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, std::string("scan: ") + hipGetErrorString(e));
  return EMO_OK;
}

// the array index of each cell's first particle (readback in array order)
int dense_offsets(emo_ctx *c) {
  ChosenState *s = c->ch;
  return scan_i32(c, s->cnt, s->doff, c->geo.ncell + 1);
}

// A layout for the counts count[0..ncell) (count[ncell] = 0): the generous
// ranges if they fit the buffers, else the tight ones.  Writes s->base.
int build_layout(emo_ctx *c, const int32_t *count) {
  ChosenState *s = c->ch;
  const long long nc = c->geo.ncell;
  for (int tight = 0; tight < 2; tight++) {
    hipLaunchKernelGGL(k_slot_caps, dim3(grid_for(nc + 1)), dim3(kBlock), 0, c->stream, count, s->base_new, nc, tight);
    int rc = scan_i32(c, s->base_new, s->base, nc + 1);
    if (rc) return rc;
    int32_t total = 0;
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(&total, s->base + nc, 4, hipMemcpyDeviceToHost, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    if (total >= 0 && total <= s->nslots) return EMO_OK;
  }
  return set_err(c, EMO_ERR_INVARIANT, "the slot layout does not fit the particle buffers");
}

// Lays out freshly loaded particles: s->pa / s->pb hold np records in any
// order (and s->g0, when active, their stored gamma in the same order).
// They are sorted by cell (radix sort of the cell ids, stable), then
// scattered into the slot layout of their counts.
int sort_loaded(emo_ctx *c) {
  ChosenState *s = c->ch;
  const long long np = c->np;
  const long long nc = c->geo.ncell;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(nc + 1)), dim3(kBlock), 0, c->stream, s->dcount, 0, nc + 1);
  float *g0_dense = nullptr, *g0_slots = nullptr;
  if (s->g0_active) {
    int rc = dalloc(c, &g0_dense, std::max<long long>(np, 1));
    if (!rc) rc = dalloc(c, &g0_slots, s->nslots);
    if (rc) {
      dfree(g0_dense);
      return rc;
    }
  }
  auto free_sort = [&]() {
    hipStreamSynchronize(c->stream);
    void *ps[] = {s->key, s->key_alt, s->perm, s->perm_alt, s->sort_tmp};
    for (void *q : ps) dfree(q);
    s->key = s->key_alt = s->perm = s->perm_alt = nullptr;
    s->sort_tmp = nullptr;
    s->sort_tmp_bytes = 0;
  };
  if (np > 0) {
    int rc = dalloc(c, &s->key, np);
    if (!rc) rc = dalloc(c, &s->key_alt, np);
    if (!rc) rc = dalloc(c, &s->perm, np);
    if (!rc) rc = dalloc(c, &s->perm_alt, np);
    if (!rc) {
      s->sort_tmp_bytes = 0;
      hipcub::DeviceRadixSort::SortPairs(nullptr, s->sort_tmp_bytes, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                         (const uint32_t *)nullptr, (uint32_t *)nullptr, (int)np, 0, s->sort_bits);
      if (hipMalloc(&s->sort_tmp, std::max<size_t>(s->sort_tmp_bytes, 16)) != hipSuccess) {
        s->sort_tmp = nullptr;
        rc = set_err(c, EMO_ERR_ALLOC, "sort scratch");
      }
    }
    if (rc) {
      free_sort();
      dfree(g0_dense);
      dfree(g0_slots);
      return rc;
    }
    hipLaunchKernelGGL(k_keys_from_pa, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, s->pa, s->key, s->perm, np);
    size_t bytes = s->sort_tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(s->sort_tmp, bytes, s->key, s->key_alt, s->perm,
                                                      s->perm_alt, (int)np, 0, s->sort_bits, c->stream);
    if (e != hipSuccess) {
      free_sort();
      dfree(g0_dense);
      dfree(g0_slots);
      return set_err(c, EMO_ERR_HIP, std::string("radix sort: ") + hipGetErrorString(e));
    }
    PermF32Args pa;
    pa.pa = s->pa;
    pa.pb = s->pb;
    pa.pa_out = s->pa_alt;
    pa.pb_out = s->pb_alt;
    pa.g0 = s->g0_active ? s->g0 : nullptr;
    pa.g0_out = g0_dense;
    pa.perm = s->perm_alt;
    pa.count = s->dcount;
    pa.np = np;
    hipLaunchKernelGGL(k_permute_f32, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, pa);
  }
  int rc = scan_i32(c, s->dcount, s->doff, nc + 1);
  if (!rc) rc = build_layout(c, s->dcount);
  if (!rc && np > 0) {
    ScatterArgs a;
    a.pa = s->pa_alt;
    a.pb = s->pb_alt;
    a.g0 = g0_dense;
    a.pa_out = s->pa;
    a.pb_out = s->pb;
    a.g0_out = g0_slots;
    a.doff = s->doff;
    a.base = s->base;
    a.np = np;
    hipLaunchKernelGGL(k_scatter_slots_f32, dim3(grid_for(np)), dim3(kBlock), 0, c->stream, a);
  }
  if (!rc)
    hipLaunchKernelGGL(k_copy_i32, dim3(grid_for(nc + 1)), dim3(kBlock), 0, c->stream, (const int32_t *)s->dcount,
                       s->cnt, nc + 1);
  if (hipStreamSynchronize(c->stream) != hipSuccess && !rc) rc = set_err(c, EMO_ERR_HIP, "layout of the load failed");
  free_sort();
  dfree(g0_dense);
  if (s->g0_active) {
    if (s->g0_owned) dfree(s->g0);
    s->g0 = g0_slots;
    s->g0_owned = true;
  } else {
    dfree(g0_slots);
  }
  return rc;
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

void launch_cellvec(emo_ctx *c) {
  ChosenState *s = c->ch;
  CellVecArgs a;
  a.g = c->geo;
  for (int d = 0; d < 3; d++) {
    a.E[d] = s->fs.E[d];
    a.B[d] = s->fs.B[d];
    a.Bt[d] = s->fs.Bt[d];
  }
  a.cv = s->cv;
  TimedLaunch t(c, "cellvec_f32");
  hipLaunchKernelGGL(k_cellvec_f32, dim3(grid_for(c->geo.ncell)), dim3(kBlock), 0, c->stream, a);
}

template <bool REL, bool PER>
void launch_push_deposit_bc(emo_ctx *c, const PushDepArgs &a, dim3 grid) {
  const bool g0 = a.g0 != nullptr, tape = a.fwd != nullptr;
  if (g0 && tape)
    hipLaunchKernelGGL((k_push_deposit_f32<REL, PER, true, true>), grid, dim3(kBlock), 0, c->stream, a);
  else if (g0)
    hipLaunchKernelGGL((k_push_deposit_f32<REL, PER, true, false>), grid, dim3(kBlock), 0, c->stream, a);
  else if (tape)
    hipLaunchKernelGGL((k_push_deposit_f32<REL, PER, false, true>), grid, dim3(kBlock), 0, c->stream, a);
  else
    hipLaunchKernelGGL((k_push_deposit_f32<REL, PER, false, false>), grid, dim3(kBlock), 0, c->stream, a);
}

void launch_push_deposit(emo_ctx *c, const PushDepArgs &a, dim3 grid) {
  bool all_per = true;
  for (int q = 0; q < 6; q++) all_per = all_per && c->cfg.parts_bound[q] == EMO_BC_PERIODIC;
  if (c->cfg.relativity && all_per)
    launch_push_deposit_bc<true, true>(c, a, grid);
  else if (c->cfg.relativity)
    launch_push_deposit_bc<true, false>(c, a, grid);
  else if (all_per)
    launch_push_deposit_bc<false, true>(c, a, grid);
  else
    launch_push_deposit_bc<false, false>(c, a, grid);
}

// One step's particle work, from the current record buffers into the
// alternate ones (which then become current): gather vectors, push +
// deposit, the layout check (and the rebuild it may need), arrivals, grid
// sum of the currents.  With a tape, the step's map source slot -> new slot,
// and the movers staged in the tape's arrays instead of the input buffer,
// which then stays untouched.
int particle_step(emo_ctx *c, double dt, const Fp32Tape *tape) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  launch_cellvec(c);
  const long long nblocks = (g.ncell + kBlockCells - 1) / kBlockCells;
  uint32_t *fwd = tape ? tape->fwd : nullptr;
  PushDepArgs a;
  a.k = push_consts(c, dt);
  a.pa = s->pa;
  a.pb = s->pb;
  a.pa_out = s->pa_alt;
  a.pb_out = s->pb_alt;
  a.ma = tape ? tape->stage_a : s->pa;
  a.mb = tape ? tape->stage_b : s->pb;
  a.mvsrc = tape ? tape->mvsrc : nullptr;
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.base = s->base;
  a.cnt = s->cnt;
  a.ns = s->ns;
  a.nm = s->nm;
  a.mcode = s->mcode;
  a.arr = s->arr;
  a.arrm = s->arrm;
  a.far = s->far;
  a.nfar = s->nfar;
  a.far_cap = kFarListCap;
  a.cj = s->cj;
  a.cstride = s->cstride;
  a.fwd = fwd;
  a.lay = s->lay;
  a.sink = s->sink;
  a.ncell = g.ncell;
  a.nblocks = nblocks;
  a.flag = c->dflag;
  {
    TimedLaunch t(c, "push_deposit_f32");
    launch_push_deposit(c, a, dim3(xcd_grid((unsigned)nblocks)));
  }
  const long long nblk = (g.ncell + kBlock - 1) / kBlock;
  {
    TimedLaunch t(c, "check_f32");
    hipLaunchKernelGGL(k_check_f32, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream, (const int32_t *)s->ns,
                       (const int32_t *)s->arr, (const int32_t *)s->base, g.ncell, s->rpart, nblk, s->lay);
  }
  {
    RebuildArgs rb;
    RelayoutArgs &r = rb.r;
    r.base = s->base;
    r.cnt = s->cnt;
    r.ns = s->ns;
    r.nm = s->nm;
    r.arr = s->arr;
    r.lay = s->lay;
    r.base_new = s->base_new;
    r.boff = s->rpart + 2 * nblk;
    r.pa_out = s->pa_alt;
    r.pb_out = s->pb_alt;
    r.ma = a.ma;
    r.mb = a.mb;
    r.fwd = fwd;
    r.ncell = g.ncell;
    rb.part = s->rpart;
    rb.nblk = nblk;
    rb.nslots = s->nslots;
    rb.lay = s->lay;
    rb.flag = c->dflag;
    const long long nrestage = (g.ncell + kRestageCells - 1) / kRestageCells;
    TimedLaunch t(c, "relayout_f32");
    hipLaunchKernelGGL(k_rebuild1_f32, dim3(1u + (unsigned)std::min<long long>(nrestage, kRestageBlocks)),
                       dim3(kRebaseThreads), 0, c->stream, rb);
    hipLaunchKernelGGL(k_relayout_f32, dim3((unsigned)nblk), dim3(kBlock), 0, c->stream, r);
  }
  {
    ArriveArgs r;
    r.g = g;
    for (int d = 0; d < 3; d++) r.per[d] = a.k.per[d];
    r.rel = c->cfg.relativity ? 1 : 0;
    r.dedupe = 0;
    for (int d = 0; d < 3; d++) r.dedupe |= (a.k.per[d] && g.n[d] <= 2) ? 1 : 0;
    r.base = s->base;
    r.base_new = s->base_new;
    r.lay = s->lay;
    r.ns = s->ns;
    r.nm = s->nm;
    r.mcode = s->mcode;
    r.arr = s->arr;
    r.arrm = s->arrm;
    r.cnt = s->cnt;
    r.akey = s->akey;
    r.akey2 = s->akey2;
    r.far = s->far;
    r.nfar = s->nfar;
    r.far_cap = kFarListCap;
    r.ma = a.ma;
    r.mb = a.mb;
    r.mvsrc = a.mvsrc;
    r.pa_out = s->pa_alt;
    r.pb_out = s->pb_alt;
    r.fwd = fwd;
    r.cj = s->cj;
    r.cstride = s->cstride;
    for (int q = 0; q < EMO_NSPEC; q++) r.jc[q] = a.k.s.test[q] ? 0.f : a.k.s.jc[q];
    r.ncell = g.ncell;
    r.flag = c->dflag;
    TimedLaunch t(c, "arrive_f32");
    hipLaunchKernelGGL(k_arrive_f32, dim3(xcd_grid((unsigned)((g.ncell + kArrBlock - 1) / kArrBlock))),
                       dim3(kArrBlock), 0, c->stream, r);
  }
  std::swap(s->pa, s->pa_alt);
  std::swap(s->pb, s->pb_alt);
  DepGridArgs<float> b;
  b.g = g;
  b.cj = s->cj;
  b.cr = nullptr;
  b.sj = s->cstride;
  b.sr = g.ncell;
  for (int d = 0; d < 3; d++) b.J[d] = s->fs.J[d];
  b.rho = s->fs.rho;
  b.do_rho = 0;
  // the rebuilt layout becomes current (after a rebuild) and the far list
  // is emptied, in this launch
  b.commit_base = s->base;
  b.commit_new = s->base_new;
  b.commit_lay = s->lay;
  b.nfar = s->nfar;
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
  const long long nslots = slots_for(cap, g.ncell);
  if (nslots > INT_MAX)
    return set_err(c, EMO_ERR_CONFIG, "the FP32 chosen form's slot layout needs int32 slot indices (fewer particles)");
  s->cap = cap;
  s->nslots = nslots;
  c->cap = cap;
  const size_t nc1 = (size_t)g.ncell + 1;
  int rc = 0;
  if (!rc) rc = dalloc(c, &s->pa, nslots);
  if (!rc) rc = dalloc(c, &s->pa_alt, nslots);
  if (!rc) rc = dalloc(c, &s->pb, nslots);
  if (!rc) rc = dalloc(c, &s->pb_alt, nslots);
  if (!rc) rc = dalloc(c, &s->akey, nslots);
  if (!rc) rc = dalloc(c, &s->akey2, nslots);
  if (!rc) rc = dalloc(c, &s->base, nc1);
  if (!rc) rc = dalloc(c, &s->base_new, nc1);
  if (!rc) rc = dalloc(c, &s->cnt, nc1);
  if (!rc) rc = dalloc(c, &s->ns, nc1);
  if (!rc) rc = dalloc(c, &s->nm, nc1);
  if (!rc) rc = dalloc(c, &s->mcode, nc1);
  if (!rc) rc = dalloc(c, &s->rpart, 3 * (nc1 / kBlock + 1));
  if (!rc) rc = dalloc(c, &s->arr, nc1);
  if (!rc) rc = dalloc(c, &s->arrm, nc1);
  if (!rc) rc = dalloc(c, &s->far, kFarListCap);
  if (!rc) rc = dalloc(c, &s->nfar, 1);
  if (!rc) rc = dalloc(c, &s->dcount, nc1);
  if (!rc) rc = dalloc(c, &s->doff, nc1);
  if (!rc) rc = dalloc(c, &s->lay, 4);
  if (!rc) rc = dalloc(c, &s->sink, kSinkVecs);
  s->cstride = (g.ncell + kBlockCells - 1) / kBlockCells * kBlockCells;
  if (!rc) rc = dalloc(c, &s->cj, 12 * s->cstride);
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
  for (int q = 0; q < 13; q++) EMO_HIP_CHECK_RET(c, ctx_zero(c, all[q], nt * sizeof(float)));
  // per-step counters start (and end every step) at zero; an empty layout
  int32_t *zero[] = {s->cnt, s->ns, s->nm, s->arr, s->arrm, s->dcount};
  for (int32_t *p : zero) EMO_HIP_CHECK_RET(c, ctx_zero(c, p, nc1 * 4));
  EMO_HIP_CHECK_RET(c, ctx_zero(c, s->nfar, 4));
  // every arrival key ever read is a slot number (also where a flagged error
  // left a range short)
  EMO_HIP_CHECK_RET(c, ctx_zero(c, s->akey, (size_t)nslots * 4));
  EMO_HIP_CHECK_RET(c, ctx_zero(c, s->akey2, (size_t)nslots * 4));
  EMO_HIP_CHECK_RET(c, ctx_zero(c, s->lay, 16));
  int bits = 1;
  while ((1ll << bits) < g.ncell) bits++;
  s->sort_bits = bits;
  s->scan_tmp_bytes = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, s->scan_tmp_bytes, (const int32_t *)nullptr, (int32_t *)nullptr,
                                   (int)nc1);
  if (hipMalloc(&s->scan_tmp, std::max<size_t>(s->scan_tmp_bytes, 16)) != hipSuccess)
    return set_err(c, EMO_ERR_ALLOC, "scan scratch");
  return build_layout(c, s->cnt);
}

void fp32_destroy(emo_ctx *c) {
  ChosenState *s = c->ch;
  if (!s) return;
  void *ps[] = {s->pa, s->pa_alt, s->pb, s->pb_alt, s->akey, s->akey2, s->base, s->base_new, s->cnt,
                s->ns, s->nm, s->mcode, s->rpart, s->arr, s->arrm, s->far, s->nfar, s->dcount, s->doff, s->lay, s->sink, s->key, s->key_alt, s->perm,
                s->perm_alt, s->sort_tmp, s->scan_tmp, s->cj, s->cv, s->gam_id, s->fs.rho};
  for (void *p : ps) dfree(p);
  if (s->g0_owned) dfree(s->g0);
  for (int d = 0; d < 3; d++) {
    dfree(s->fs.E[d]);
    dfree(s->fs.Bfix[d]);
    dfree(s->Btfix[d]);
    dfree(s->fs.J[d]);
  }
  delete s;
  c->ch = nullptr;
}

// drops the stored gamma (and frees it when it is this state's own)
void drop_g0(ChosenState *s) {
  if (s->g0_owned) dfree(s->g0);
  s->g0 = nullptr;
  s->g0_active = false;
  s->g0_owned = true;
}

// Converts FP64 particles to the chosen representation on the host (the same
// cell-id arithmetic as Grid::getCellID) and lays them out by cell.
int fp32_load(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
              const int32_t *species, const int64_t *id) {
  ChosenState *s = c->ch;
  const Geo &g = c->geo;
  if (n > s->cap) return set_err(c, EMO_ERR_CONFIG, "more particles than max_particles");
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
    if (species[i] < 0 || species[i] >= std::max(c->cfg.nspecies, 1))
      return set_err(c, EMO_ERR_CONFIG, "species index out of range");
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
    EMO_HIP_CHECK_RET(c, ctx_h2d(c, s->pa, ha.data(), n * sizeof(int4)));
    EMO_HIP_CHECK_RET(c, ctx_h2d(c, s->pb, hb.data(), n * sizeof(int4)));
  }
  s->have_id = id != nullptr;
  drop_g0(s);
  dfree(s->gam_id);
  s->gam_id = nullptr;
  if (rel && n > 0) {
    int rc = dalloc(c, &s->g0, n);
    if (rc) return rc;
    EMO_HIP_CHECK_RET(c, ctx_h2d(c, s->g0, g0.data(), n * 4));
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
      EMO_HIP_CHECK_RET(c, ctx_h2d(c, s->gam_id, gi.data(), n * 4));
    }
  }
  c->np = n;
  int rc = sort_loaded(c);
  if (rc) return rc;
  return check_flag(c);
}

// output in array order (cell by cell); the caller's device buffers of np
// entries, any of them null
int launch_output(emo_ctx *c, OutArgs &a) {
  ChosenState *s = c->ch;
  int rc = dense_offsets(c);
  if (rc) return rc;
  a.g = c->geo;
  a.pa = s->pa;
  a.pb = s->pb;
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.gam_id = s->gam_id;
  a.cv = s->cv;
  a.base = s->base;
  a.cnt = s->cnt;
  a.doff = s->doff;
  a.rel = c->cfg.relativity;
  a.np = c->np;
  hipLaunchKernelGGL(k_output_f32, dim3((unsigned)((c->geo.ncell + kTileCells - 1) / kTileCells)), dim3(kBlock), 0,
                     c->stream, a);
  return EMO_OK;
}

// the particles with id < nid by id into out (7 per id: x, v, found)
int fp32_tracked(emo_ctx *c, long long nid, double *out) {
  launch_const_fill_f64(c, out, 0.0, 7 * nid);
  if (c->np == 0 || !c->ch->have_id) return EMO_OK;
  OutArgs a{};
  a.trk = out;
  a.ntrk = nid;
  return launch_output(c, a);
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
  // the gather of the current grid (Particle_Handler::InterpolateEB, which
  // ends the reference's step): the gather vectors of this grid first
  if (!rc && field6) launch_cellvec(c);
  if (!rc) rc = dalloc(c, &dg, n);
  if (!rc) rc = dalloc(c, &dsp, n);
  if (!rc) rc = dalloc(c, &did, n);
  if (!rc) {
    OutArgs a{};
    for (int d = 0; d < 3; d++) {
      a.x[d] = dx[d];
      a.v[d] = dv[d];
    }
    for (int d = 0; d < 6; d++) a.fld[d] = df[d];
    a.gam = dg;
    a.sp = dsp;
    a.id = did;
    a.cell = nullptr;
    rc = launch_output(c, a);
  }
  if (!rc) {
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
    if (species && !rc && ctx_d2h(c, species, dsp, n * 4) != hipSuccess) rc = EMO_ERR_HIP;
    if (id && !rc) {
      std::vector<int32_t> h(n);
      if (ctx_d2h(c, h.data(), did, n * 4) != hipSuccess) rc = EMO_ERR_HIP;
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
  int64_t n = c->np;
  if (n == 0) return EMO_OK;
  int32_t *d = nullptr;
  int rc = dalloc(c, &d, n);
  if (rc) return rc;
  OutArgs a{};
  a.cell = d;
  rc = launch_output(c, a);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(cells, d, n * 4, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d);
  if (rc) return rc;
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
    a.base = s->base;
    a.cnt = s->cnt;
    a.s = species_f32(c, 0.0);
    a.rel = c->cfg.relativity;
    a.ncell = g.ncell;
    a.cj = cj;
    a.cr = cr;
    a.sj = g.ncell;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    hipLaunchKernelGGL(k_deposit_rhoj_f32, dim3(blocks), dim3(kBlock), 0, c->stream, a);
    DepGridArgs<float> b;
    b.g = g;
    b.cj = cj;
    b.cr = cr;
    b.sj = b.sr = g.ncell;
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
  EMO_HIP_CHECK_RET(c, ctx_h2d(c, p, h.data(), c->geo.ntot * 4));
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
  if (n > s->cap) return set_err(c, EMO_ERR_CONFIG, "more particles than max_particles");
  drop_g0(s);
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
    a.base = s->base;
    a.cnt = s->cnt;
    a.s = species_f32(c, 0.0);
    a.rel = c->cfg.relativity;
    a.ncell = g.ncell;
    a.cj = s->cj;
    a.cr = cr;
    a.sj = s->cstride;
    const unsigned blocks = (unsigned)((g.ncell + (kBlock / 16) - 1) / (kBlock / 16));
    hipLaunchKernelGGL(k_deposit_rhoj_f32, dim3(blocks), dim3(kBlock), 0, c->stream, a);
    DepGridArgs<float> b;
    b.g = g;
    b.cj = s->cj;
    b.cr = cr;
    b.sj = s->cstride;
    b.sr = g.ncell;
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
  double hpart[2];
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
      // the sweep with its FP64 residual (max|du|, sum du^2 over the
      // physical points, which the ghost refresh does not touch)
      fk_jacobi_resid<float>(c, src, dst, srcs[sv], ax, ay, az, af, sourceMult, partials);
      potential_bc<float>(c, dst, bounds[sv]);
      hipError_t e = hipMemcpyAsync(hpart, partials + 2 * kMaxResidualBlocks, 2 * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        rc = set_err(c, EMO_ERR_HIP, std::string("poisson residual: ") + hipGetErrorString(e));
        break;
      }
      const double maxDiff = hpart[0], l2 = hpart[1];
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

int fp32_step_one(emo_ctx *c, double t, double dt, const Fp32Tape *tape) {
  ChosenState *s = c->ch;
  int rc = particle_step(c, dt, tape);
  if (rc) return rc;
  if (s->g0_active) {
    // the stored gamma was used by this first push only (Q3): from now on
    // gamma = sqrt(1 + u^2)
    if (s->g0_owned) EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    drop_g0(s);
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
  v.base = &s->base;
  v.cnt = &s->cnt;
  v.nslots = s->nslots;
  v.lay = s->lay;
  return v;
}

// The error flags are read every kStepCheck steps (one stream
// synchronisation each), so a call stops within that many steps of an error
// instead of stepping an invalid state to its end.
constexpr int kStepCheck = 64;

int fp32_step(emo_ctx *c, int32_t nsteps, double t0, double dt) {
  double t = t0;
  for (int k = 0; k < nsteps; k++) {
    int rc = fp32_step_one(c, t, dt, nullptr);
    if (!rc && (k + 1) % kStepCheck == 0 && k + 1 < nsteps) rc = fp32_check(c);
    if (rc) return rc;
    t += dt;
  }
  return fp32_check(c);
}

int fp32_scan_i32(emo_ctx *c, const int32_t *in, int32_t *out, long long n) { return scan_i32(c, in, out, n); }

int fp32_layout_info(emo_ctx *c, int64_t *out3) {
  ChosenState *s = c->ch;
  int32_t used = 0, lay[2] = {0, 0};
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(&used, s->base + c->geo.ncell, 4, hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(lay, s->lay, 8, hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  out3[0] = s->nslots;
  out3[1] = used;
  out3[2] = lay[1];
  return EMO_OK;
}

int fp32_diag(emo_ctx *c, double *out8) {
  ChosenState *s = c->ch;
  int nb = 256;
  double *pp = c->red, *pf = c->red + 5 * nb, *res = c->red + 11 * nb;
  DiagF32Args a;
  a.pa = s->pa;
  a.pb = s->pb;
  a.g0 = s->g0_active ? s->g0 : nullptr;
  a.base = s->base;
  a.cnt = s->cnt;
  a.rel = c->cfg.relativity;
  a.s = c->species;
  a.ncell = c->geo.ncell;
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
