// emo_api.cpp -- the C ABI (include/emo_api.h): context lifetime, host<->device
// transfers, the step's phase ordering (src/main.cpp:231-275), the Poisson
// initialisation loop, LightBC's host-side pulse scalars and the timing hooks.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <string>
#include <vector>

#include "emo_internal.hpp"
#include "uniform_load.hpp"

namespace emo {
int process_info(const emo_loader *ld, emo_config *cfg, double *dens_out, double *temp_out);
int load_maxwellian_host(const emo_config *cfg, const emo_loader *ld, std::vector<double> &x,
                         std::vector<double> &v, std::vector<double> &gam,
                         std::vector<int32_t> &sp, std::vector<int64_t> &id,
                         std::vector<double> *dx_deps = nullptr);

TimedLaunch::TimedLaunch(emo_ctx *ctx, const char *nm, hipStream_t stream)
    : c(ctx), name(nm), s(stream ? stream : ctx->stream) {
  if (!c->timing || (!c->timing_only.empty() && c->timing_only != nm)) return;
  if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
    a = b = nullptr;
    return;
  }
  hipEventRecord(a, s);
}

TimedLaunch::~TimedLaunch() {
  if (!a) return;
  hipEventRecord(b, s);
  c->pending[name].push_back({a, b});
}

}  // namespace emo

using namespace emo;

namespace {

int set_err(emo_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
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

int check_flags(emo_ctx *c) {
  hipError_t e = hipMemcpyAsync(c->hflag, c->dflag, sizeof(unsigned long long), hipMemcpyDeviceToHost,
                                c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, std::string("flag readback: ") + hipGetErrorString(e));
  if (c->hflag[0] & kFlagOutOfGrid)
    return set_err(c, EMO_ERR_INVARIANT,
                   "particle outside the padded grid (Grid::getCellID asserts, src/grid/oGrid.cpp:212-220)");
  return EMO_OK;
}

int validate(const emo_config *cfg, std::string &msg) {
  for (int d = 0; d < 3; d++) {
    if (cfg->nCell[d] < 1) return msg = "nCell must be >= 1", EMO_ERR_CONFIG;
    if (!(cfg->Lxyz[d] > 0)) return msg = "Lxyz must be > 0", EMO_ERR_CONFIG;
  }
  if (cfg->nspecies < 0 || cfg->nspecies > EMO_NSPEC) return msg = "nspecies out of range", EMO_ERR_CONFIG;
  if (cfg->precision != EMO_FP64_PLAIN && cfg->precision != EMO_FP32_CHOSEN)
    return msg = "unknown precision mode", EMO_ERR_CONFIG;
  for (int s = 0; s < 6; s++) {
    if (cfg->fields_bound[s] != EMO_BC_PERIODIC && cfg->fields_bound[s] != EMO_BC_EXTERNAL)
      return msg = "unknown field boundary kind", EMO_ERR_CONFIG;
    if (cfg->parts_bound[s] != EMO_BC_PERIODIC && cfg->parts_bound[s] != EMO_BC_REFLECTING)
      return msg = "unknown particle boundary kind", EMO_ERR_CONFIG;
  }
  // Input::ProcessInfo (src/IO/check.cpp:117-133): per dimension either all
  // four (field, particle) x (left, right) periodic or none of them
  for (int d = 0; d < 3; d++) {
    int np = (cfg->fields_bound[2 * d] == EMO_BC_PERIODIC) + (cfg->fields_bound[2 * d + 1] == EMO_BC_PERIODIC) +
             (cfg->parts_bound[2 * d] == EMO_BC_PERIODIC) + (cfg->parts_bound[2 * d + 1] == EMO_BC_PERIODIC);
    if (np != 0 && np != 4) return msg = "inconsistent boundary conditions in one direction", EMO_ERR_CONFIG;
  }
  if (cfg->nwaves < 0 || cfg->nwaves >= EMO_NWAVE) return msg = "nwaves out of range", EMO_ERR_CONFIG;
  for (int i = 0; i < cfg->nwaves; i++)
    if (cfg->inPolE[i] < 1 || cfg->inPolE[i] > 3) return msg = "inPolE must be 1..3", EMO_ERR_CONFIG;
  long long cells = (long long)(cfg->nCell[0] + 2) * (cfg->nCell[1] + 2) * (cfg->nCell[2] + 2);
  if (cells >= (1ll << 31)) return msg = "grid too large for int32 cell ids", EMO_ERR_CONFIG;
  if (cfg->max_particles < 0 || cfg->max_particles >= (1ll << 31))
    return msg = "max_particles must be in [0, 2^31)", EMO_ERR_CONFIG;
  return EMO_OK;
}

int alloc_particles(emo_ctx *c, ParticlesF64 &p, int64_t cap) {
  int rc = 0;
  for (int d = 0; d < 3 && !rc; d++) rc = dalloc(c, &p.x[d], cap);
  for (int d = 0; d < 3 && !rc; d++) rc = dalloc(c, &p.v[d], cap);
  for (int d = 0; d < 6 && !rc; d++) rc = dalloc(c, &p.fld[d], cap);
  if (!rc) rc = dalloc(c, &p.gam, cap);
  if (!rc) rc = dalloc(c, &p.sp, cap);
  if (!rc) rc = dalloc(c, &p.id, cap);
  return rc;
}

void free_particles(ParticlesF64 &p) {
  for (int d = 0; d < 3; d++) {
    dfree(p.x[d]);
    dfree(p.v[d]);
  }
  for (int d = 0; d < 6; d++) dfree(p.fld[d]);
  dfree(p.gam);
  dfree(p.sp);
  dfree(p.id);
  p = ParticlesF64();
}

int require_plain(emo_ctx *c, const char *what) {
  if (c->cfg.precision != EMO_FP64_PLAIN)
    return set_err(c, EMO_ERR_UNSUPPORTED, std::string(what) + " is a separate phase only in the FP64 plain form");
  return EMO_OK;
}

int last_launch(emo_ctx *c) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, std::string("kernel launch: ") + hipGetErrorString(e));
  return EMO_OK;
}

// The static device workspace of emo_init_poisson_noalloc: one copy per GPU
// the library runs on, held until the process ends.  Host code never uses
// the array's name as an address: each device's copy is looked up once with
// hipGetSymbolAddress (emo_create, that device current) and kept per device.
constexpr size_t kNoAllocWsBytes = 64ull << 20;
constexpr int kMaxDevices = 64;
__device__ __align__(256) unsigned char scratch[kNoAllocWsBytes];
std::mutex g_ws_mu;
void *g_ws_ptr[kMaxDevices] = {};

int noalloc_ws_for_device(emo_ctx *c) {
  if (c->device < 0 || c->device >= kMaxDevices) return set_err(c, EMO_ERR_CONFIG, "device id out of range");
  std::lock_guard<std::mutex> lk(g_ws_mu);
  if (!g_ws_ptr[c->device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return set_err(c, EMO_ERR_HIP, std::string("hipGetSymbolAddress(workspace): ") + hipGetErrorString(e));
    g_ws_ptr[c->device] = ptr;
  }
  c->noalloc_ws = g_ws_ptr[c->device];
  return EMO_OK;
}


}  // namespace

extern "C" {

const char *emo_version(void) { return "emopic-mi355x 0.1 (gfx950)"; }

const char *emo_last_error(const emo_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int emo_create(const emo_config *cfg, emo_ctx **out) {
  if (!cfg || !out) return EMO_ERR_CONFIG;
  *out = nullptr;
  std::string msg;
  int rc = validate(cfg, msg);
  if (rc) {
    fprintf(stderr, "emo_create: %s\n", msg.c_str());
    return rc;
  }
  emo_ctx *c = new (std::nothrow) emo_ctx();
  if (!c) return EMO_ERR_ALLOC;
  c->cfg = *cfg;
  Geo &g = c->geo;
  for (int d = 0; d < 3; d++) {
    g.n[d] = cfg->nCell[d];
    g.nt[d] = cfg->nCell[d] + 3;
    g.nc[d] = cfg->nCell[d] + 2;
    g.x0[d] = cfg->xyz0[d];
    g.d[d] = cfg->Lxyz[d] / cfg->nCell[d];
    g.id[d] = 1.0 / g.d[d];
  }
  g.sy = g.nt[2];
  g.sx = (long long)g.nt[1] * g.nt[2];
  g.ntot = (long long)g.nt[0] * g.nt[1] * g.nt[2];
  g.ncell = (long long)g.nc[0] * g.nc[1] * g.nc[2];
  for (int s = 0; s < EMO_NSPEC; s++) {
    c->species.q[s] = cfg->charge[s];
    c->species.m[s] = cfg->mass[s] != 0 ? cfg->mass[s] : 1.0;
    c->species.test[s] = cfg->isTestParticle[s];
  }
  hipError_t e = hipGetDevice(&c->device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    fprintf(stderr, "emo_create: hipStreamCreate: %s\n", hipGetErrorString(e));
    delete c;
    return EMO_ERR_HIP;
  }
  // the static workspace of emo_init_poisson_noalloc on this device; a
  // failure here only disables that entry point (it reports the error)
  if (noalloc_ws_for_device(c) != EMO_OK) fprintf(stderr, "emo_create: %s\n", c->err.c_str());
  rc = 0;
  if (!rc) rc = dalloc(c, &c->dflag, 8);
  if (!rc && hipHostMalloc((void **)&c->hflag, 8 * sizeof(unsigned long long)) != hipSuccess) rc = EMO_ERR_ALLOC;
  if (!rc) rc = ctx_zero(c, c->dflag, 8 * sizeof(unsigned long long)) == hipSuccess ? 0 : EMO_ERR_HIP;
  c->red_len = 16 * 1024;
  if (!rc) rc = dalloc(c, &c->red, c->red_len);
  if (cfg->precision == EMO_FP64_PLAIN) {
    size_t nt = (size_t)g.ntot;
    for (int q = 0; q < 3 && !rc; q++) rc = dalloc(c, &c->E[q], nt);
    for (int q = 0; q < 3 && !rc; q++) rc = dalloc(c, &c->Bfix[q], nt);
    for (int q = 0; q < 3 && !rc; q++) rc = dalloc(c, &c->Btfix[q], nt);
    for (int q = 0; q < 3 && !rc; q++) rc = dalloc(c, &c->J[q], nt);
    if (!rc) rc = dalloc(c, &c->rho, nt);
    for (int q = 0; q < 3 && !rc; q++) {
      c->B[q] = c->Bfix[q];
      c->Bt[q] = c->Btfix[q];
    }
    // fields start at zero (the reference leaves them uninitialised, Q8)
    double *all[13] = {c->E[0], c->E[1], c->E[2], c->Bfix[0], c->Bfix[1], c->Bfix[2],
                       c->Btfix[0], c->Btfix[1], c->Btfix[2], c->J[0], c->J[1], c->J[2], c->rho};
    for (int q = 0; q < 13 && !rc; q++)
      if (ctx_zero(c, all[q], nt * sizeof(double)) != hipSuccess) rc = EMO_ERR_HIP;
    int64_t cap = std::max<int64_t>(cfg->max_particles, 1);
    c->cap = cap;
    if (!rc) rc = alloc_particles(c, c->p, cap);
    if (!rc) rc = alloc_particles(c, c->palt, cap);
    if (!rc) rc = dalloc(c, &c->key, cap);
    if (!rc) rc = dalloc(c, &c->key_alt, cap);
    if (!rc) rc = dalloc(c, &c->perm, cap);
    if (!rc) rc = dalloc(c, &c->perm_alt, cap);
    if (!rc) rc = plain_sort_create(c);
    if (!rc) rc = dalloc(c, &c->cj, 12 * g.ncell);
    if (!rc) rc = dalloc(c, &c->cr, 8 * g.ncell);
    if (!rc) {
      int bits = 1;
      while ((1ll << bits) < g.ncell) bits++;
      c->sort_bits = bits;
      c->sort_tmp_bytes = sort_tmp_bytes_for(cap, bits);
      if (hipMalloc(&c->sort_tmp, std::max<size_t>(c->sort_tmp_bytes, 16)) != hipSuccess) rc = EMO_ERR_ALLOC;
    }
  } else {
    rc = fp32_create(c);
  }
  if (!rc) rc = build_waist_profiles(c);
  if (rc) {
    fprintf(stderr, "emo_create: %s\n", c->err.c_str());
    emo_destroy(c);
    return rc;
  }
  if (hipDeviceSynchronize() != hipSuccess) {
    emo_destroy(c);
    return EMO_ERR_HIP;
  }
  *out = c;
  return EMO_OK;
}

void emo_destroy(emo_ctx *c) {
  if (!c) return;
  if (c->stream) hipStreamSynchronize(c->stream);
  free_waist_profiles(c);
  for (auto &kv : c->pending)
    for (auto &r : kv.second) {
      hipEventDestroy(r.a);
      hipEventDestroy(r.b);
    }
  for (int q = 0; q < 3; q++) {
    dfree(c->E[q]);
    dfree(c->Bfix[q]);
    dfree(c->Btfix[q]);
    dfree(c->J[q]);
  }
  dfree(c->rho);
  free_particles(c->p);
  free_particles(c->palt);
  dfree(c->key);
  dfree(c->key_alt);
  dfree(c->perm);
  dfree(c->perm_alt);
  dfree(c->sort_tmp);
  plain_sort_destroy(c);
  dfree(c->adj_pool);
  if (c->adj_stream) hipStreamDestroy(c->adj_stream);
  dfree(c->cj);
  dfree(c->cr);
  dfree(c->red);
  dfree(c->dflag);
  if (c->hflag) hipHostFree(c->hflag);
  fp32_destroy(c);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
}

int emo_process_info(const emo_loader *ld, emo_config *cfg) {
  if (!ld || !cfg) return EMO_ERR_CONFIG;
  return process_info(ld, cfg, nullptr, nullptr);
}

int emo_load_particles(emo_ctx *c, int64_t n, const double *x3, const double *v3, const double *gamma,
                       const int32_t *species, const int64_t *id) {
  if (!c) return EMO_ERR_CONFIG;
  if (n < 0 || n > c->cap) return set_err(c, EMO_ERR_CONFIG, "particle count exceeds max_particles");
  if (n > 0 && (!x3 || !v3 || !species)) return set_err(c, EMO_ERR_CONFIG, "null particle array");
  for (int64_t i = 0; i < n; i++)
    if (species[i] < 0 || species[i] >= std::max(c->cfg.nspecies, 1))
      return set_err(c, EMO_ERR_CONFIG, "species index out of range");
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_load(c, n, x3, v3, gamma, species, id);
  }
  std::vector<double> buf(n);
  auto up = [&](double *dst, const double *src, int stride, int comp) -> int {
    for (int64_t i = 0; i < n; i++) buf[i] = src ? src[stride * i + comp] : 1.0;
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(dst, buf.data(), n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    return EMO_OK;
  };
  int rc = 0;
  for (int d = 0; d < 3 && !rc; d++) rc = up(c->p.x[d], x3, 3, d);
  for (int d = 0; d < 3 && !rc; d++) rc = up(c->p.v[d], v3, 3, d);
  if (!rc) rc = up(c->p.gam, gamma, 1, 0);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, ctx_h2d(c, c->p.sp, species, n * sizeof(int32_t)));
  std::vector<int64_t> ids(n);
  for (int64_t i = 0; i < n; i++) ids[i] = id ? id[i] : i;
  EMO_HIP_CHECK_RET(c, ctx_h2d(c, c->p.id, ids.data(), n * sizeof(int64_t)));
  for (int d = 0; d < 6; d++) EMO_HIP_CHECK_RET(c, ctx_zero(c, c->p.fld[d], std::max<int64_t>(n, 1) * sizeof(double)));
  c->np = n;
  c->sorted = false;
  c->fld_state = kFldStale;  // zero until the first gather
  return EMO_OK;
}

int emo_load_maxwellian(emo_ctx *c, const emo_loader *ld) {
  if (!c || !ld) return EMO_ERR_CONFIG;
  std::vector<double> x, v, gam;
  std::vector<int32_t> sp;
  std::vector<int64_t> id;
  int rc = load_maxwellian_host(&c->cfg, ld, x, v, gam, sp, id);
  if (rc == EMO_ERR_INVARIANT)
    return set_err(c, rc, "Poisson initialization with periodic boundaries cannot be run with a net charge");
  if (rc) return set_err(c, rc, "invalid loader configuration");
  // the particles carry the super-particle q, m of Input::ProcessInfo
  rc = process_info(ld, &c->cfg, nullptr, nullptr);
  if (rc) return set_err(c, rc, "invalid loader configuration");
  for (int s = 0; s < EMO_NSPEC; s++) {
    c->species.q[s] = c->cfg.charge[s];
    c->species.m[s] = c->cfg.mass[s] != 0 ? c->cfg.mass[s] : 1.0;
    c->species.test[s] = c->cfg.isTestParticle[s];
  }
  return emo_load_particles(c, (int64_t)gam.size(), x.data(), v.data(), gam.data(), sp.data(), id.data());
}

int emo_load_uniform(emo_ctx *c, const emo_loader *ld) {
  if (!c || !ld) return EMO_ERR_CONFIG;
  double dens[EMO_NSPEC], temp[EMO_NSPEC];
  int rc = process_info(ld, &c->cfg, dens, temp);
  if (rc) return set_err(c, rc, "invalid loader configuration");
  for (int s = 0; s < EMO_NSPEC; s++) {
    c->species.q[s] = c->cfg.charge[s];
    c->species.m[s] = c->cfg.mass[s] != 0 ? c->cfg.mass[s] : 1.0;
    c->species.test[s] = c->cfg.isTestParticle[s];
  }
  const Geo &g = c->geo;
  const int64_t nphys = (int64_t)g.n[0] * g.n[1] * g.n[2];
  const int64_t n = ld->nparticles_tot;
  if (n < 0 || n > c->cap) return set_err(c, EMO_ERR_CONFIG, "particle count exceeds max_particles");
  UniformGen u;
  u.g = g;
  u.key = splitmix64((uint64_t)ld->seed);
  u.nspec = ld->nspecies;
  int64_t tot = 0;
  for (int s = 0; s < ld->nspecies; s++) {
    const int64_t cnt = std::llround(dens[s] * (double)n);
    if (cnt % nphys != 0)
      return set_err(c, EMO_ERR_CONFIG, "the uniform loader needs a whole number of particles per cell and species");
    tot += cnt / nphys;
    if (tot > INT_MAX) return set_err(c, EMO_ERR_CONFIG, "too many particles per cell");
    u.ppc_end[s] = (int)tot;
    u.vth[s] = kUnitVth * std::sqrt(temp[s] / c->cfg.mass[s]);
    for (int d = 0; d < 3; d++) u.drift[s][d] = ld->drift[s][d];
  }
  if (tot * nphys != n) return set_err(c, EMO_ERR_CONFIG, "species counts do not add up to nparticles_tot");
  u.ppc = (int)tot;
  if (c->cfg.precision != EMO_FP64_PLAIN) return fp32_load_uniform(c, u, n);
  launch_uniform_f64(c, u, n);
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  c->np = n;
  c->sorted = false;
  c->fld_state = kFldStale;
  return last_launch(c);
}

int64_t emo_num_particles(const emo_ctx *c) { return c ? c->np : -1; }

int emo_get_particles(emo_ctx *c, double *x3, double *v3, double *gamma, int32_t *species, int64_t *id,
                      double *field6) {
  if (!c) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_get_particles(c, x3, v3, gamma, species, id, field6);
  }
  int64_t n = c->np;
  std::vector<double> buf(std::max<int64_t>(n, 1));
  auto down = [&](double *dst, const double *src, int stride, int comp) -> int {
    EMO_HIP_CHECK_RET(c, hipMemcpyAsync(buf.data(), src, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
    for (int64_t i = 0; i < n; i++) dst[stride * i + comp] = buf[i];
    return EMO_OK;
  };
  int rc = 0;
  if (field6) fld_materialize(c);
  if (x3)
    for (int d = 0; d < 3 && !rc; d++) rc = down(x3, c->p.x[d], 3, d);
  if (v3)
    for (int d = 0; d < 3 && !rc; d++) rc = down(v3, c->p.v[d], 3, d);
  if (gamma && !rc) rc = down(gamma, c->p.gam, 1, 0);
  if (field6)
    for (int d = 0; d < 6 && !rc; d++) rc = down(field6, c->p.fld[d], 6, d);
  if (rc) return rc;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  if (species && n) EMO_HIP_CHECK_RET(c, ctx_d2h(c, species, c->p.sp, n * sizeof(int32_t)));
  if (id && n) EMO_HIP_CHECK_RET(c, ctx_d2h(c, id, c->p.id, n * sizeof(int64_t)));
  return EMO_OK;
}

int emo_get_tracked(emo_ctx *c, int64_t nid, double *x3, double *v3, int32_t *found) {
  if (!c || nid < 0) return EMO_ERR_CONFIG;
  if (nid == 0) return EMO_OK;
  if (!x3 || !v3 || !found) return set_err(c, EMO_ERR_CONFIG, "emo_get_tracked needs x3, v3 and found");
  double *d = nullptr;
  int rc = dalloc(c, &d, (size_t)(7 * nid));
  if (rc) return rc;
  if (c->cfg.precision != EMO_FP64_PLAIN) rc = fp32_tracked(c, nid, d);
  else launch_tracked_f64(c, nid, d);
  std::vector<double> h(rc ? 0 : (size_t)(7 * nid));
  hipError_t e = hipSuccess;
  if (!rc) {
    e = hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  }
  dfree(d);
  if (rc) return rc;
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, hipGetErrorString(e));
  for (int64_t k = 0; k < nid; k++) {
    for (int q = 0; q < 3; q++) {
      x3[3 * k + q] = h[7 * k + q];
      v3[3 * k + q] = h[7 * k + 3 + q];
    }
    found[k] = h[7 * k + 6] != 0.0;
  }
  return EMO_OK;
}

int emo_get_cell_ids(emo_ctx *c, int32_t *cells) {
  if (!c || !cells) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_get_cells(c, cells);
  }
  if (c->np == 0) return EMO_OK;
  int32_t *d = nullptr;
  int rc = dalloc(c, &d, c->np);
  if (rc) return rc;
  launch_cell_ids(c, d);
  hipError_t e = hipMemcpyAsync(cells, d, c->np * sizeof(int32_t), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  dfree(d);
  if (e != hipSuccess) return set_err(c, EMO_ERR_HIP, hipGetErrorString(e));
  return EMO_OK;
}

int64_t emo_field_points(const emo_ctx *c) { return c ? c->geo.ntot : -1; }

static double *field_ptr(emo_ctx *c, int fid) {
  switch (fid) {
    case EMO_EX: case EMO_EY: case EMO_EZ: return c->E[fid];
    case EMO_BX: case EMO_BY: case EMO_BZ: return c->B[fid - EMO_BX];
    case EMO_JX: case EMO_JY: case EMO_JZ: return c->J[fid - EMO_JX];
    case EMO_RHO: return c->rho;
    case EMO_BX_TM1: case EMO_BY_TM1: case EMO_BZ_TM1: return c->Bt[fid - EMO_BX_TM1];
    default: return nullptr;
  }
}

int emo_get_field(emo_ctx *c, int fid, double *out) {
  if (!c || !out) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_get_field(c, fid, out);
  }
  double *p = field_ptr(c, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(out, p, c->geo.ntot * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return EMO_OK;
}

int emo_set_field(emo_ctx *c, int fid, const double *in) {
  if (!c || !in) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_set_field(c, fid, in);
  }
  double *p = field_ptr(c, fid);
  if (!p) return set_err(c, EMO_ERR_CONFIG, "unknown field id");
  if (fid < EMO_JX || fid >= EMO_BX_TM1) fld_invalidate(c);  // E, B, B_tm1 feed the gather
  EMO_HIP_CHECK_RET(c, hipMemcpyAsync(p, in, c->geo.ntot * sizeof(double), hipMemcpyHostToDevice, c->stream));
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return EMO_OK;
}

int emo_init_constant(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_init_constant(c);
  }
  long long nt = c->geo.ntot;
  for (int q = 0; q < 3; q++) {
    launch_const_fill_f64(c, c->E[q], c->cfg.E0[q], nt);
    launch_const_fill_f64(c, c->B[q], c->cfg.B0[q], nt);
    launch_const_fill_f64(c, c->Bt[q], c->cfg.B0[q], nt);
    launch_const_fill_f64(c, c->J[q], 0.0, nt);
  }
  launch_const_fill_f64(c, c->rho, 0.0, nt);
  launch_gather_f64(c);
  int rc = last_launch(c);
  if (rc) return rc;
  return check_flags(c);
}

int emo_sort(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_sort");
  if (rc) return rc;
  fld_materialize(c);
  rc = sort_f64(c, 1);
  if (rc) return rc;
  return last_launch(c);
}

int emo_push(emo_ctx *c, double dt) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_push");
  if (rc) return rc;
  fld_materialize(c);
  launch_push_f64(c, dt);
  c->sorted = false;
  c->fld_state = kFldStale;
  return last_launch(c);
}

int emo_particle_bc(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_particle_bc");
  if (rc) return rc;
  fld_invalidate(c);
  launch_particle_bc_f64(c);
  c->sorted = false;
  return last_launch(c);
}

int emo_deposit(emo_ctx *c, int32_t rho) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_deposit");
  if (rc) return rc;
  if (!c->sorted) {
    fld_materialize(c);
    rc = sort_f64(c, 1);
    if (rc) return rc;
  }
  launch_deposit_f64(c, rho ? 1 : 0);
  return last_launch(c);
}

int emo_fold(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_fold");
  if (rc) return rc;
  launch_fold_f64(c);
  return last_launch(c);
}

int emo_evolve_fields(emo_ctx *c, double dt) {
  if (!c) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {  // field-only phase: both forms
    fk_evolve<float>(c, *fp32_view(c).fs, dt);
    return last_launch(c);
  }
  fld_invalidate(c);
  launch_evolve_f64(c, dt);
  return last_launch(c);
}

int emo_refresh(emo_ctx *c, double t, double dt) {
  if (!c) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {  // field-only phase: both forms
    refresh_fields<float>(c, *fp32_view(c).fs, t, dt);
    return last_launch(c);
  }
  fld_invalidate(c);
  FieldSet<double> v = f64_view(c);
  refresh_fields<double>(c, v, t, dt);
  return last_launch(c);
}

int emo_interpolate(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  int rc = require_plain(c, "emo_interpolate");
  if (rc) return rc;
  launch_gather_f64(c);
  return last_launch(c);
}

}  // extern "C"

namespace emo {

static size_t ws_round(size_t b) { return (b + 255) & ~(size_t)255; }

size_t poisson_ws_bytes(const emo_ctx *c) {
  const size_t el = c->cfg.precision == EMO_FP64_PLAIN ? sizeof(double) : sizeof(float);
  size_t b = ws_round((2 * kMaxResidualBlocks + 2) * sizeof(double));
  if (c->cfg.precision != EMO_FP64_PLAIN) b += ws_round((size_t)8 * c->geo.ncell * sizeof(float));
  return b + 8 * ws_round((size_t)c->geo.ntot * el);
}

void poisson_ws_carve(const emo_ctx *c, void *base, PoissonWs &ws) {
  const size_t el = c->cfg.precision == EMO_FP64_PLAIN ? sizeof(double) : sizeof(float);
  char *p = (char *)base;
  ws.partials = (double *)p;
  p += ws_round((2 * kMaxResidualBlocks + 2) * sizeof(double));
  ws.cr = nullptr;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    ws.cr = (float *)p;
    p += ws_round((size_t)8 * c->geo.ncell * sizeof(float));
  }
  for (int q = 0; q < 8; q++) {
    ws.u[q] = p;
    p += ws_round((size_t)c->geo.ntot * el);
  }
}

}  // namespace emo

namespace {

// the FP64 Poisson initialisation on a carved workspace
int init_poisson_f64(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps_out, double *resid_out,
                     const PoissonWs &ws) {
  if (iter_max < 0) return set_err(c, EMO_ERR_CONFIG, "iter_max must be >= 0");
  const Geo &g = c->geo;
  size_t nt = (size_t)g.ntot;
  // deposit + fold (src/main.cpp:160-166)
  fld_invalidate(c);
  int rc = sort_f64(c, 1);
  if (rc) return rc;
  launch_deposit_f64(c, 1);
  launch_fold_f64(c);
  // Q8: E, B, B_tm1 zeroed; potentials allocated zeroed
  for (int q = 0; q < 3; q++) {
    launch_const_fill_f64(c, c->E[q], 0.0, nt);
    launch_const_fill_f64(c, c->B[q], 0.0, nt);
    launch_const_fill_f64(c, c->Bt[q], 0.0, nt);
  }
  double *u[8];
  for (int q = 0; q < 8; q++) {
    u[q] = (double *)ws.u[q];
    launch_const_fill_f64(c, u[q], 0.0, nt);
  }
  int nb = residual_blocks(c);
  double *partials = ws.partials;
  // extension: a uniform neutralising background (emo_config.neutralize)
  if (c->cfg.neutralize) fk_subtract_mean<double>(c, c->rho, partials, nb, partials + 2 * nb);
  // run_poisson_solver_ coefficients (src/poisson/poisson.cpp:181-185)
  double dx_ = g.d[0], dy_ = g.d[1], dz_ = g.d[2];
  double celldist2 = std::pow(dx_, 2.0) * std::pow(dy_, 2.0) + std::pow(dy_, 2.0) * std::pow(dz_, 2.0) +
                     std::pow(dx_, 2.0) * std::pow(dz_, 2.0);
  double ax = std::pow(dy_, 2.0) * std::pow(dz_, 2.0) / (2.0 * celldist2);
  double ay = std::pow(dx_, 2.0) * std::pow(dz_, 2.0) / (2.0 * celldist2);
  double az = std::pow(dx_, 2.0) * std::pow(dy_, 2.0) / (2.0 * celldist2);
  double af = std::pow(dx_, 2.0) * std::pow(dy_, 2.0) * std::pow(dz_, 2.0) / (2.0 * celldist2);
  double sourceMult = -4 * kPi;
  const emo_config &cf = c->cfg;
  // tolerances (Poisson_Solver ctor :54-71, InitializeFields :151,157)
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
  double tphi = tol >= 0 ? tol : std::max((pmax - pmin) * 1e-9, 1e-12);
  double tA = tol >= 0 ? tol : std::max((amax - amin) * 1e-9, 1e-12);
  const double *bounds[4] = {cf.bound_phi, cf.bound_Ax, cf.bound_Ay, cf.bound_Az};
  const double *srcs[4] = {c->rho, c->J[0], c->J[1], c->J[2]};
  double resid[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int32_t sweeps[4] = {0, 0, 0, 0};
  double hpart[2];
  for (int s = 0; s < 4 && !rc; s++) {
    double *u1 = u[2 * s], *u2 = u[2 * s + 1];
    double tl = s == 0 ? tphi : tA;
    potential_bc<double>(c, u1, bounds[s]);
    int iternum = -1;
    bool converged = false;
    do {
      iternum++;
      const double *src = (iternum % 2 == 0) ? u1 : u2;
      double *dst = (iternum % 2 == 0) ? u2 : u1;
      // the sweep with its residual (max|du|, sum du^2 over the physical
      // points, which the ghost refresh does not touch)
      launch_jacobi_resid_f64(c, src, dst, srcs[s], ax, ay, az, af, sourceMult, partials);
      potential_bc<double>(c, dst, bounds[s]);
      hipError_t e = hipMemcpyAsync(hpart, partials + 2 * kMaxResidualBlocks, 2 * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
      if (e != hipSuccess) {
        rc = set_err(c, EMO_ERR_HIP, std::string("poisson residual: ") + hipGetErrorString(e));
        break;
      }
      const double maxDiff = hpart[0], l2 = hpart[1];
      resid[2 * s] = maxDiff;
      resid[2 * s + 1] = std::sqrt(l2);
      if (maxDiff < tl) converged = true;
    } while (!converged && iternum < iter_max);
    sweeps[s] = iternum + 1;
    if (s == 0) launch_phi_to_e_f64(c, u1);  // phi1_ holds the odd-numbered iterate (Q9)
  }
  if (!rc) {
    launch_a_to_b_f64(c, u[2], u[4], u[6]);
    FieldSet<double> v = f64_view(c);
    refresh_fields<double>(c, v, 0.0, 0.0);  // executeBC(-1) at time_phys = dt_phys = 0
    launch_gather_f64(c);         // src/main.cpp:179
  }
  hipStreamSynchronize(c->stream);
  if (rc) return rc;
  if (sweeps_out) memcpy(sweeps_out, sweeps, sizeof(sweeps));
  if (resid_out) memcpy(resid_out, resid, sizeof(resid));
  rc = last_launch(c);
  if (rc) return rc;
  return check_flags(c);
}

int init_poisson_ws(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps_out, double *resid_out,
                    const PoissonWs &ws) {
  if (c->cfg.precision != EMO_FP64_PLAIN) return fp32_init_poisson(c, iter_max, tol, sweeps_out, resid_out, ws);
  return init_poisson_f64(c, iter_max, tol, sweeps_out, resid_out, ws);
}

}  // namespace

extern "C" {

int emo_init_poisson(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps_out, double *resid_out) {
  if (!c) return EMO_ERR_CONFIG;
  void *base = nullptr;
  const size_t bytes = poisson_ws_bytes(c);
  if (hipMalloc(&base, bytes) != hipSuccess)
    return set_err(c, EMO_ERR_ALLOC, "hipMalloc failed for the Poisson initialisation's workspace");
  PoissonWs ws;
  poisson_ws_carve(c, base, ws);
  int rc = init_poisson_ws(c, iter_max, tol, sweeps_out, resid_out, ws);
  hipStreamSynchronize(c->stream);
  hipFree(base);
  return rc;
}

int emo_init_poisson_noalloc(emo_ctx *c, int32_t iter_max, double tol, int32_t *sweeps_out, double *resid_out) {
  if (!c) return EMO_ERR_CONFIG;
  const size_t bytes = poisson_ws_bytes(c);
  if (bytes > kNoAllocWsBytes)
    return set_err(c, EMO_ERR_CONFIG,
                   "emo_init_poisson_noalloc: this grid needs " + std::to_string(bytes) +
                       " bytes of workspace, more than the static workspace's " + std::to_string(kNoAllocWsBytes) +
                       " (use emo_init_poisson)");
  if (!c->noalloc_ws) return set_err(c, EMO_ERR_HIP, "emo_init_poisson_noalloc: no static workspace on this device");
  PoissonWs ws;
  poisson_ws_carve(c, c->noalloc_ws, ws);
  return init_poisson_ws(c, iter_max, tol, sweeps_out, resid_out, ws);
}

int emo_step(emo_ctx *c, int32_t nsteps, double t0, double dt) {
  if (!c || nsteps < 0) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_step(c, nsteps, t0, dt);
  }
  double t = t0;
  for (int s = 0; s < nsteps; s++) {
    int rc;
    if (c->sorted && c->fld_state != kFldStale && c->np > 0) {
      // fused: the push gathers the fields from the grid, which the stored
      // per-particle fields equal (fresh) or would equal (deferred); the
      // store is in cell order, so the counting sort keeps the radix
      // sort's stable order
      fused_push_f64(c, dt, nullptr, nullptr);
      rc = fused_sort_f64(c, nullptr, nullptr);
      if (rc) return rc;
    } else {
      fld_materialize(c);
      launch_push_f64(c, dt);
      launch_particle_bc_f64(c);
      rc = sort_f64(c, 0);
      if (rc) return rc;
    }
    launch_deposit_f64(c, 1);
    launch_fold_f64(c);
    launch_evolve_f64(c, dt);
    {
      FieldSet<double> v = f64_view(c);
      refresh_fields<double>(c, v, t, dt);
    }
    c->fld_state = kFldDeferred;  // the gather that ends the step
    t += dt;
  }
  int rc = last_launch(c);
  if (rc) return rc;
  return check_flags(c);
}

int emo_diag(emo_ctx *c, double *out8) {
  if (!c || !out8) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP64_PLAIN) {
    return fp32_diag(c, out8);
  }
  int nb = 256;
  double *pp = c->red;              // 5*nb
  double *pf = c->red + 5 * nb;     // 6*nb
  double *res = c->red + 11 * nb;   // 11
  launch_diag_particles_f64(c, pp, nb);
  launch_reduce_partials(c, pp, nb, 5, res);
  launch_diag_fields_f64(c, pf, nb);
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

int emo_layout_info(emo_ctx *c, int64_t *out3) {
  if (!c || !out3) return EMO_ERR_CONFIG;
  if (c->cfg.precision != EMO_FP32_CHOSEN)
    return set_err(c, EMO_ERR_UNSUPPORTED, "the slot layout belongs to the FP32 chosen form");
  return fp32_layout_info(c, out3);
}

int emo_synchronize(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  EMO_HIP_CHECK_RET(c, hipStreamSynchronize(c->stream));
  return check_flags(c);
}

int emo_timing_enable(emo_ctx *c, int32_t on) {
  if (!c) return EMO_ERR_CONFIG;
  c->timing = on != 0;
  return EMO_OK;
}

int emo_timing_select(emo_ctx *c, const char *name) {
  if (!c) return EMO_ERR_CONFIG;
  c->timing_only = name ? name : "";
  return EMO_OK;
}

static void drain_timing(emo_ctx *c) {
  hipStreamSynchronize(c->stream);
  for (auto &kv : c->pending) {
    auto &tot = c->totals[kv.first];
    for (auto &r : kv.second) {
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
        tot.first += ms;
        tot.second += 1;
      }
      hipEventDestroy(r.a);
      hipEventDestroy(r.b);
    }
    kv.second.clear();
  }
}

int emo_kernel_time(emo_ctx *c, const char *name, double *ms_total, int64_t *launches) {
  if (!c || !name) return EMO_ERR_CONFIG;
  drain_timing(c);
  auto it = c->totals.find(name);
  if (ms_total) *ms_total = it == c->totals.end() ? 0.0 : it->second.first;
  if (launches) *launches = it == c->totals.end() ? 0 : it->second.second;
  return EMO_OK;
}

int emo_timing_reset(emo_ctx *c) {
  if (!c) return EMO_ERR_CONFIG;
  drain_timing(c);
  c->totals.clear();
  return EMO_OK;
}

}  // extern "C"
