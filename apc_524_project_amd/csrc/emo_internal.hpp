// emo_internal.hpp -- context layout and kernel entry points of libemopic.
//
// One emo_ctx owns: the padded field arrays on the device, the particle
// store, per-cell scratch for the deterministic deposit, and a HIP stream on
// which every kernel of the context is launched.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "../../include/emo_api.h"
#include "geo.hpp"

namespace emo {

struct SpeciesParams {
  double q[EMO_NSPEC];
  double m[EMO_NSPEC];
  int test[EMO_NSPEC];
};

// Pulse scalars evaluated on the host for one boundary plane fill.
struct PlaneFill {
  void *f;        // field array (member pointer), double or float
  int dim;        // plane normal
  int idx;        // plane index along dim
  double base;    // E0/B0 component
  double wave;    // pulse value added on top
  // the wave part's transverse profile over the plane (nt[u] x nt[w], u < w
  // the two other dims), device memory; null: uniform (the reference)
  const double *prof;
};

// One precision's grid arrays.  Bfix are the fixed allocations that the
// reference's Grid::fieldPtr_ keeps (src/grid/grid.cpp:321-342); B/Bt are
// the member pointers that Grid::evolveFields swaps (src/grid/oGrid.cpp:18-20).
template <typename T>
struct FieldSet {
  T *E[3] = {nullptr, nullptr, nullptr};
  T *B[3] = {nullptr, nullptr, nullptr};
  T *Bt[3] = {nullptr, nullptr, nullptr};
  T *Bfix[3] = {nullptr, nullptr, nullptr};
  T *J[3] = {nullptr, nullptr, nullptr};
  T *rho = nullptr;
};

struct ChosenState;  // FP32 chosen form (chosen.hip)

// Device buffers of the stable counting sort by destination cell
// (cellsort.hpp): per particle a move code, a rank and the permutation
// (new position -> old index); per cell the ranges off[ncell+1] and, for
// each of the 216 move codes, the (source cell, code) group sizes and
// offsets.
struct CellSortBufs {
  uint8_t *code = nullptr;
  uint16_t *rank = nullptr;
  uint16_t *nsk = nullptr;    // [code * ncell + source], zero between steps
  uint16_t *gbase = nullptr;  // [code * ncell + source]
  int32_t *count = nullptr;   // ncell + 1 particles per destination cell
  int32_t *off = nullptr, *off_new = nullptr;  // ncell + 1 cell ranges
  int32_t *misc = nullptr;       // [0]: far-mover count
  int32_t *far_tiles = nullptr;  // far movers per tile
  int4 *far = nullptr;           // (index, source cell, destination cell, 0)
  int32_t *far_pos = nullptr;    // far mover's offset in its destination
  uint32_t *perm = nullptr;      // new position -> old index
  void *scan_tmp = nullptr;
  size_t scan_tmp_bytes = 0;
};

// FP64 particle store (structure of arrays).
struct ParticlesF64 {
  double *x[3] = {nullptr, nullptr, nullptr};
  double *v[3] = {nullptr, nullptr, nullptr};
  double *gam = nullptr;
  double *fld[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  int32_t *sp = nullptr;
  int64_t *id = nullptr;
};

}  // namespace emo

struct emo_ctx {
  emo_config cfg;
  emo::Geo geo;
  emo::SpeciesParams species;
  hipStream_t stream = nullptr;
  std::string err;

  // fields, FP64. Bfix/Btfix are the fixed allocations that the reference's
  // Grid::fieldPtr_ keeps (src/grid/grid.cpp:321-342); B/Bt are the member
  // pointers that evolveFields swaps (src/grid/oGrid.cpp:18-20).
  double *E[3] = {};
  double *Bfix[3] = {};
  double *Btfix[3] = {};
  double *B[3] = {};
  double *Bt[3] = {};
  double *J[3] = {};
  double *rho = nullptr;

  // particles
  int64_t np = 0;
  int64_t cap = 0;
  emo::ParticlesF64 p, palt;
  uint32_t *key = nullptr, *key_alt = nullptr;
  uint32_t *perm = nullptr, *perm_alt = nullptr;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  int sort_bits = 0;
  emo::CellSortBufs srt;  // counting sort; srt.off[c]..srt.off[c+1] is cell c
  double *cj = nullptr;   // 12 per cell
  double *cr = nullptr;   // 8 per cell
  bool sorted = false;  // store in ascending cell order and srt.off valid
  // The per-particle fields p.fld against the grid: kFldFresh = they are the
  // gather of the grid at the particles' positions; kFldStale = the grid or
  // the positions changed since the last gather (p.fld is what the next push
  // uses); kFldDeferred = the gather of the grid is due but not stored (the
  // fused step gathers inside its push and leaves the stored copy to be made
  // when something reads it).
  int fld_state = 1;
  int64_t fallback_sorts = 0;  // fused steps that overflowed the counting sort
  // checkpoints, tape and adjoint buffers of emo_adjoint_field_energy, kept
  // for later calls
  void *adj_pool = nullptr;
  size_t adj_pool_bytes = 0;
  hipStream_t adj_stream = nullptr;  // the adjoint's backward sweep
  // the last gradient call (emo_adjoint_stats): {checkpointing (0: K-step
  // segments, 1: binomial), snapshots, forward steps run, restores}, and the
  // per-species sums of its initial-state gradients (7 per species)
  int64_t adj_info[4] = {-1, 0, 0, 0};
  double adj_sums[7 * EMO_NSPEC] = {};

  // transverse Gaussian profiles of an injected pulse (emo_config.waist > 0)
  // per plane normal and field id 0..5 (E, B), built on the host at
  // creation (boundary.cpp)
  double *wprof[3][6] = {};

  // the device this context was created on, and that device's copy of the
  // static workspace of emo_init_poisson_noalloc (emo_api.cpp)
  int device = 0;
  void *noalloc_ws = nullptr;

  // reductions / flags
  double *red = nullptr;   // scratch for FP64 reductions
  int64_t red_len = 0;
  unsigned long long *dflag = nullptr;  // [0] error flags, [1] max bits, ...
  unsigned long long *hflag = nullptr;  // pinned host mirror

  // FP32 chosen form
  emo::ChosenState *ch = nullptr;

  // timing
  bool timing = false;
  std::string timing_only;  // non-empty: time only the launches of this name
  struct Rec {
    hipEvent_t a, b;
  };
  std::map<std::string, std::vector<Rec>> pending;
  std::map<std::string, std::pair<double, int64_t>> totals;
};

namespace emo {

// error flag bits written by kernels into dflag[0]
constexpr unsigned long long kFlagOutOfGrid = 1ull;
constexpr unsigned long long kFlagNonFinite = 2ull;
// FP64 migration sort: more particles left the one-cell neighbourhood of
// their cell in one step than its side list holds
constexpr unsigned long long kFlagFarMovers = 4ull;
// FP64 migration sort: a cell's gathered particles disagree with its count;
// FP32 slot layout: a moving particle found no slot in its new cell
constexpr unsigned long long kFlagMigrate = 8ull;
// FP32 step: more particles moved further than one cell than the far list
// holds
constexpr unsigned long long kFlagFarList = 16ull;
// FP32 step: the movers found in a cell's neighbours disagree with the
// cell's arrival count
constexpr unsigned long long kFlagArrivals = 32ull;
// FP32 step: a rebuilt slot layout does not fit the particle buffers
constexpr unsigned long long kFlagLayout = 64ull;

// ---- plain FP64 particle kernels (plain_kernels.hip) ----
void launch_gather_f64(emo_ctx *c);
void launch_push_f64(emo_ctx *c, double dt);
void launch_particle_bc_f64(emo_ctx *c);
// with_fields = 0 leaves the per-particle gathered fields behind (valid
// only when a gather follows before the next push)
int sort_f64(emo_ctx *c, int with_fields);
int plain_sort_create(emo_ctx *c);
void plain_sort_destroy(emo_ctx *c);
void fused_push_f64(emo_ctx *c, double dt, double *const *fld_out, uint8_t *flips_out);
int fused_sort_f64(emo_ctx *c, uint32_t *perm_out, uint32_t *inv_out);
constexpr int kFldFresh = 0, kFldStale = 1, kFldDeferred = 2;
// stores the deferred gather (kFldDeferred -> kFldFresh)
void fld_materialize(emo_ctx *c);
// before a change of the grid or of the positions: materialize, then stale
void fld_invalidate(emo_ctx *c);
void launch_deposit_f64(emo_ctx *c, int rho);
void launch_cell_ids(emo_ctx *c, int32_t *out);
void launch_tracked_f64(emo_ctx *c, long long nid, double *out);
size_t sort_tmp_bytes_for(int64_t n, int bits);

// ---- field kernels (field_kernels.hip) ----
void launch_fold_f64(emo_ctx *c);
void launch_evolve_f64(emo_ctx *c, double dt);
void launch_plane_copy_f64(emo_ctx *c, double *const *fields, int nf, int d, int src, int dst,
                           int add);
void launch_plane_fill_f64(emo_ctx *c, const PlaneFill &pf);
void launch_const_fill_f64(emo_ctx *c, double *f, double v, long long n);
void launch_jacobi_f64(emo_ctx *c, const double *src, double *dst, const double *R, double ax,
                       double ay, double az, double af, double sourceMult);
void launch_jacobi_resid_f64(emo_ctx *c, const double *src, double *dst, const double *R, double ax, double ay,
                             double az, double af, double sourceMult, double *partials);
void launch_residual_f64(emo_ctx *c, const double *u1, const double *u2, double *partials,
                         int nblocks);
int residual_blocks(emo_ctx *c);
void launch_phi_to_e_f64(emo_ctx *c, const double *phi);
void launch_a_to_b_f64(emo_ctx *c, const double *Ax, const double *Ay, const double *Az);
void launch_diag_particles_f64(emo_ctx *c, double *partials, int nblocks);
void launch_diag_fields_f64(emo_ctx *c, double *partials, int nblocks);
void launch_reduce_partials(emo_ctx *c, const double *partials, int nblocks, int ncomp,
                            double *out);

// ---- precision-generic grid kernels (field_kernels.hip), T = double or float ----
template <typename T>
void fk_plane_copy(emo_ctx *c, T *const *fields, int nf, int d, int src, int dst, int add);
template <typename T>
void fk_plane_fill(emo_ctx *c, const PlaneFill &pf);
// up to six fills of different planes in one launch
template <typename T>
void fk_plane_fill_multi(emo_ctx *c, const PlaneFill *pf, int n);
// up to six periodic sides (plain copies over the physical range) in one launch
template <typename T>
void fk_plane_copy_sides(emo_ctx *c, T *const *fields, int nf, int nsides, const int *d, const int *src,
                         const int *dst);
template <typename T>
void fk_const_fill(emo_ctx *c, T *f, double v, long long n);
template <typename T>
void fk_fold(emo_ctx *c, FieldSet<T> &fs);
// swaps fs.B <-> fs.Bt, then the E and B stencils
template <typename T>
void fk_evolve(emo_ctx *c, FieldSet<T> &fs, double dt);
template <typename T>
void fk_diag_fields(emo_ctx *c, const FieldSet<T> &fs, double *partials, int nblocks);
// Poisson initialisation kernels (src/poisson/poisson.cpp:223-260,
// src/poisson/convertFields.cpp): one Jacobi sweep over the physical points,
// FP64 residual partials of two iterates, phiToE into E[3], AToB into B[3]
template <typename T>
void fk_jacobi(emo_ctx *c, const T *src, T *dst, const T *R, double ax, double ay, double az, double af,
               double sourceMult);
template <typename T>
void fk_residual(emo_ctx *c, const T *u1, const T *u2, double *partials, int nblocks);
// one Jacobi sweep with its residual: max|dst - src| and sum (dst - src)^2
// over the physical points land in partials[2 * kMaxResidualBlocks + 0, 1]
template <typename T>
void fk_jacobi_resid(emo_ctx *c, const T *src, T *dst, const T *R, double ax, double ay, double az, double af,
                     double sourceMult, double *partials);
template <typename T>
void fk_phi_to_e(emo_ctx *c, const T *phi, T *const *E);
template <typename T>
void fk_a_to_b(emo_ctx *c, const T *Ax, const T *Ay, const T *Az, T *const *B);
// f -= its FP64 mean over the physical points (partials: nblocks doubles,
// sum_out: one double, both device memory)
template <typename T>
void fk_subtract_mean(emo_ctx *c, T *f, double *partials, int nblocks, double *sum_out);

// ---- boundary host logic (boundary.cpp) ----
// the waist profiles of the LightBC planes (emo_config.waist > 0)
int build_waist_profiles(emo_ctx *c);
void free_waist_profiles(emo_ctx *c);
// Grid::executeBC(-1): periodic ghost copies (through Bfix for B) and LightBC
template <typename T>
void refresh_fields(emo_ctx *c, FieldSet<T> &fs, double t, double dt);
// executeBC(potential) for a Jacobi array: periodic copy / PoissonBC plane
template <typename T>
void potential_bc(emo_ctx *c, T *u, const double *bound);
FieldSet<double> f64_view(emo_ctx *c);

// ---- FP32 chosen form (chosen.hip) ----
int fp32_create(emo_ctx *c);
void fp32_destroy(emo_ctx *c);
int fp32_load(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
              const int32_t *species, const int64_t *id);
int fp32_get_particles(emo_ctx *c, double *x3, double *v3, double *gamma, int32_t *species, int64_t *id,
                       double *field6);
int fp32_get_cells(emo_ctx *c, int32_t *cells);
int fp32_tracked(emo_ctx *c, long long nid, double *out);
int fp32_get_field(emo_ctx *c, int fid, double *out);
int fp32_set_field(emo_ctx *c, int fid, const double *in);
int fp32_init_constant(emo_ctx *c);
// The Poisson initialisation's workspace: the two Jacobi iterates of each of
// the four solves (phi, Ax, Ay, Az), the FP64 residual partial sums and, in
// the FP32 form, the cell charge vectors of the initial deposit.  Both
// entry points carve it from one block with the same layout (every piece
// 256-byte aligned): emo_init_poisson from an allocation, the no-allocation
// variant from the library's static device array.
struct PoissonWs {
  void *u[8] = {};
  double *partials = nullptr;  // 2 * kMaxResidualBlocks + 2 doubles
  float *cr = nullptr;         // FP32 form: 8 per cell
};
constexpr int kMaxResidualBlocks = 4096;
size_t poisson_ws_bytes(const emo_ctx *c);
void poisson_ws_carve(const emo_ctx *c, void *base, PoissonWs &ws);
int fp32_init_poisson(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps, double *resid,
                      const PoissonWs &ws);
// The FP32 form's device state, by reference, for its adjoint (adjoint.hip):
// the particle records (current and alternate buffers, nslots slots each),
// the stored gamma of the first push (by slot), the fields and the slot
// layout (cell c holds cnt[c] particles from slot base[c]).
struct Fp32View {
  int4 **pa, **pb, **pa_alt, **pb_alt;
  float **g0;
  bool *g0_active, *g0_owned;
  FieldSet<float> *fs;
  float *Btfix[3];
  int32_t **base, **cnt;
  int64_t nslots;
  int32_t *lay;
};
Fp32View fp32_view(emo_ctx *c);
// A recomputed step's outputs for the adjoint: the map source slot -> new
// slot, and staging arrays (nslots each) for the movers' new records and
// source slots, so that the step's input records stay intact.
struct Fp32Tape {
  uint32_t *fwd;
  int4 *stage_a, *stage_b;
  int32_t *mvsrc;
};
// one step of fp32_step without the flag readback, optionally taped
int fp32_step_one(emo_ctx *c, double t, double dt, const Fp32Tape *tape);
// the device error flags of the FP32 form (stream synchronisation)
int fp32_check(emo_ctx *c);
// exclusive prefix sum of n int32 values on the context stream (n <= ncell + 1)
int fp32_scan_i32(emo_ctx *c, const int32_t *in, int32_t *out, long long n);
int fp32_adjoint_field_energy(emo_ctx *c, int32_t nsteps, double t0, double dt, int32_t K, double *energy,
                              double *d_amp, double *d_phase, double *xbar0, double *vbar0, double *gbar0);
struct UniformGen;  // uniform_load.hpp
int fp32_load_uniform(emo_ctx *c, const UniformGen &u, int64_t n);
void launch_uniform_f64(emo_ctx *c, const UniformGen &u, int64_t n);
int fp32_step(emo_ctx *c, int32_t nsteps, double t0, double dt);
int fp32_diag(emo_ctx *c, double *out8);
// slots per record buffer, slots covered by the cells' ranges, rebuilds so far
int fp32_layout_info(emo_ctx *c, int64_t *out3);

// timing helpers (emo_api.cpp)
struct TimedLaunch {
  emo_ctx *c;
  const char *name;
  hipStream_t s;
  hipEvent_t a = nullptr, b = nullptr;
  TimedLaunch(emo_ctx *ctx, const char *nm, hipStream_t stream = nullptr);
  ~TimedLaunch();
};

}  // namespace emo

namespace emo {
// Host copies and fills of context memory, ordered on the context's stream.
// That stream is non-blocking: a hipMemcpy or hipMemset on the null stream
// is not ordered with the kernels the context launches (a fill could land
// after the kernel that reads it).  The copies return once they are done.
inline hipError_t ctx_h2d(emo_ctx *c, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream);
  return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}
inline hipError_t ctx_d2h(emo_ctx *c, void *dst, const void *src, size_t bytes) {
  hipError_t e = hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream);
  return e == hipSuccess ? hipStreamSynchronize(c->stream) : e;
}
inline hipError_t ctx_zero(emo_ctx *c, void *p, size_t bytes) { return hipMemsetAsync(p, 0, bytes, c->stream); }
}  // namespace emo

#define EMO_HIP_CHECK_RET(ctx, expr)                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      (ctx)->err = std::string(#expr) + ": " + hipGetErrorString(e_);             \
      return EMO_ERR_HIP;                                                         \
    }                                                                             \
  } while (0)
