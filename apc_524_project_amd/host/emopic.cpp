// emopic.cpp -- the reference's class surface over libemopic (see emopic.hpp).
#include "emopic.hpp"

#include <math.h>  // sqrt for the shared per-particle arithmetic

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

#include "../csrc/pic_math.hpp"

namespace emopic {

// process-wide run state, as the reference keeps it in globals
// (src/globals.hpp:35-46)
double time_phys = 0.0;
double dt_phys = 0.0;

namespace {

constexpr double kUnitAcc = 0.58668774;  // src/globals.hpp:12

Input_Info_t *g_info = nullptr;  // registered by Domain's constructor

// raw (unscaled) species inputs kept by Input::ProcessInfo so that the
// library's loader applies the super-particle scaling exactly once
struct RawSpecies {
  double mass[NSPEC], charge[NSPEC], dens[NSPEC], temp[NSPEC];
  bool valid = false;
};
std::map<const Input_Info_t *, RawSpecies> g_raw;

int bc_kind(const char *s, bool field) {
  std::string t(s);
  if (t == "periodic") return EMO_BC_PERIODIC;
  if (field && t == "external") return EMO_BC_EXTERNAL;
  if (!field && t == "reflecting") return EMO_BC_REFLECTING;
  throw std::runtime_error(std::string("Unable to find ") + (field ? "field" : "particle") +
                           " boundary " + t);
}

// the registered boundary objects: they only carry their kind and side
class BC_F_Periodic : public BC_Field {
 public:
  explicit BC_F_Periodic(int side) { side_ = side; }
  int kind() const override { return EMO_BC_PERIODIC; }
};
class BC_F_External : public BC_Field {
 public:
  explicit BC_F_External(int side) { side_ = side; }
  int kind() const override { return EMO_BC_EXTERNAL; }
};
class BC_P_Periodic : public BC_Particle {
 public:
  BC_P_Periodic(int d, bool r) {
    dim_ = d;
    isRight_ = r;
  }
  int kind() const override { return EMO_BC_PERIODIC; }
};
class BC_P_Reflecting : public BC_Particle {
 public:
  BC_P_Reflecting(int d, bool r) {
    dim_ = d;
    isRight_ = r;
  }
  int kind() const override { return EMO_BC_REFLECTING; }
};

BC_Field *make_f_periodic(int side, Domain *, Grid *, Input_Info_t *) { return new BC_F_Periodic(side); }
BC_Field *make_f_external(int side, Domain *, Grid *, Input_Info_t *) { return new BC_F_External(side); }
BC_Particle *make_p_periodic(Domain *, int d, short r, std::string) { return new BC_P_Periodic(d, r != 0); }
BC_Particle *make_p_reflecting(Domain *, int d, short r, std::string) { return new BC_P_Reflecting(d, r != 0); }

}  // namespace

// ---- Input --------------------------------------------------------------
int Input::ProcessInfo() {
  Input_Info_t *in = &info_;
  int err = 0;
  RawSpecies raw;
  for (int i = 0; i < in->nspecies && i < NSPEC; i++) {
    raw.mass[i] = in->mass_ratio[i];
    raw.charge[i] = in->charge_ratio[i];
    raw.dens[i] = in->dens_frac[i];
    raw.temp[i] = in->temp[i];
  }
  raw.valid = true;
  // src/IO/check.cpp:12-147 through the library (same arithmetic)
  emo_loader ld;
  memset(&ld, 0, sizeof(ld));
  ld.nparticles_tot = in->nparticles_tot;
  ld.nspecies = in->nspecies;
  for (int i = 0; i < in->nspecies && i < NSPEC; i++) {
    ld.mass_ratio[i] = in->mass_ratio[i];
    ld.charge_ratio[i] = in->charge_ratio[i];
    ld.dens_frac[i] = in->dens_frac[i];
    ld.temp[i] = in->temp[i];
    ld.isTestParticle[i] = in->isTestParticle[i];
  }
  ld.dens_phys = in->dens_phys;
  emo_config cfg;
  memset(&cfg, 0, sizeof(cfg));
  for (int d = 0; d < 3; d++) cfg.Lxyz[d] = in->Lxyz[d];
  if (emo_process_info(&ld, &cfg) != EMO_OK) err += 1;
  double V = in->Lxyz[0] * in->Lxyz[1] * in->Lxyz[2];
  double n0 = in->nparticles_tot / V;
  double dens_phys = in->dens_phys < 0 ? -in->dens_phys : (in->dens_phys == 0 ? n0 : in->dens_phys);
  in->dens_phys = dens_phys;
  in->super_ratio = n0 > 0 ? dens_phys / n0 : -1.0;
  double cden = 0.0;
  for (int i = 0; i < in->nspecies; i++) cden += in->dens_frac[i];
  for (int i = 0; i < in->nspecies; i++) {
    in->mass_ratio[i] = cfg.mass[i];
    in->charge_ratio[i] = cfg.charge[i];
    in->temp[i] = std::fabs(in->temp[i]);
    if (in->super_ratio > 0) {
      in->dens_frac[i] /= cden;
      in->temp[i] *= in->super_ratio;
    }
  }
  // periodic consistency per dimension (src/IO/check.cpp:117-133)
  for (int d = 0; d < 3; d++) {
    int np = 0;
    for (int s = 0; s < 2; s++) {
      np += std::string(in->fields_bound[2 * d + s]) == "periodic";
      np += std::string(in->parts_bound[2 * d + s]) == "periodic";
    }
    if (np != 0 && np != 4) err += 1;
  }
  g_raw[in] = raw;
  return err;
}

// ---- Domain -------------------------------------------------------------
Domain::Domain(Input_Info_t *input_info) : info_(input_info) {
  for (int i = 0; i < 3; i++) {
    nxyz_[i] = input_info->nCell[i];
    Lxyz_[i] = input_info->Lxyz[i];
    xyz0_[i] = input_info->xyz0[i];
  }
  input_info->nparticles_domain = input_info->nparticles_tot;
  g_info = input_info;
}

double Domain::getmindx() {
  double m = Lxyz_[0] / nxyz_[0];
  for (int i = 1; i < 3; i++) m = std::min(m, Lxyz_[i] / nxyz_[i]);
  return m;
}

// ---- pushers on one host particle (FP64, the reference formulas) --------
int Boris::Step(Particle *part, Field_part *field, double dt) {
  double q_p = (0.5 * dt * part->q) / part->m;
  q_p *= kUnitAcc;
  double ex = field->e1, ey = field->e2, ez = field->e3;
  double bx = field->b1, by = field->b2, bz = field->b3;
  double tx = q_p * bx, ty = q_p * by, tz = q_p * bz;
  double bsqr = bx * bx + by * by + bz * bz;
  double tsqr = 2.0 / (1.0 + q_p * q_p * bsqr);
  double sx = tsqr * tx, sy = tsqr * ty, sz = tsqr * tz;
  double v_mx = part->v[0] + q_p * ex, v_my = part->v[1] + q_p * ey, v_mz = part->v[2] + q_p * ez;
  double v_ix = v_mx + v_my * tz - v_mz * ty;
  double v_iy = v_my + v_mz * tx - v_mx * tz;
  double v_iz = v_mz + v_mx * ty - v_my * tx;
  double vx = v_mx + v_iy * sz - v_iz * sy + q_p * ex;
  double vy = v_my + v_iz * sx - v_ix * sz + q_p * ey;
  double vz = v_mz + v_ix * sy - v_iy * sx + q_p * ez;
  part->x[0] += dt * vx;
  part->x[1] += dt * vy;
  part->x[2] += dt * vz;
  part->v[0] = vx;
  part->v[1] = vy;
  part->v[2] = vz;
  return 1;
}

int Relativistic_Boris::Step(Particle *part, Field_part *field, double dt) {
  double vx = part->v[0], vy = part->v[1], vz = part->v[2], gamma = part->gamma;
  double q_p = (0.5 * dt * part->q) / part->m;
  q_p *= kUnitAcc;
  double ex = field->e1, ey = field->e2, ez = field->e3;
  double bx = field->b1, by = field->b2, bz = field->b3;
  double upx = gamma * vx + q_p * (ex + (vy * bz - vz * by)) + q_p * ex;
  double upy = gamma * vy + q_p * (ey + (vz * bx - vx * bz)) + q_p * ey;
  double upz = gamma * vz + q_p * (ez + (vx * by - vy * bx)) + q_p * ez;
  double taux = q_p * bx, tauy = q_p * by, tauz = q_p * bz;
  double tau_sq = taux * taux + tauy * tauy + tauz * tauz;
  double sigma = 1.0 + upx * upx + upy * upy + upz * upz - tau_sq;
  double ustar = upx * taux + upy * tauy + upz * tauz;
  double g = std::sqrt(0.5 * (sigma + std::sqrt(sigma * sigma + 4 * (tau_sq + ustar * ustar))));
  double tx = taux / g, ty = tauy / g, tz = tauz / g;
  double s = 1.0 / (1.0 + tx * tx + ty * ty + tz * tz);
  double udott = ustar / g;
  double ux = s * (upx + udott * tx + (upy * tz - upz * ty));
  double uy = s * (upy + udott * ty + (upz * tx - upx * tz));
  double uz = s * (upz + udott * tz + (upx * ty - upy * tx));
  part->v[0] = ux / g;
  part->v[1] = uy / g;
  part->v[2] = uz / g;
  part->x[0] += dt * part->v[0];
  part->x[1] += dt * part->v[1];
  part->x[2] += dt * part->v[2];
  part->gamma = g;
  return 1;
}

// ---- single host particle: gather and deposit ---------------------------
namespace {
// the cell lengths are the only geometry the per-particle formulas read
emo::Geo cell_geo(const double *lcell) {
  emo::Geo g{};
  for (int d = 0; d < 3; d++) {
    g.d[d] = lcell[d];
    g.id[d] = 1.0 / lcell[d];
  }
  return g;
}
}  // namespace

void Interpolator::interpolate_fields(double *pos, double *lcell, double *cellvars, Field_part *field) {
  const emo::Geo g = cell_geo(lcell);
  double f[6];
  emo::interp_fields(g, cellvars, pos, f);
  field->e1 = f[0];
  field->e2 = f[1];
  field->e3 = f[2];
  field->b1 = f[3];
  field->b2 = f[4];
  field->b3 = f[5];
}

void Depositor::deposit_particle_J(Particle *part, double *lcell, double *cellverts, double *JObj) {
  const emo::Geo g = cell_geo(lcell);
  double r[20];
  emo::deposit_row(g, cellverts, part->x, part->v, part->q, false, r);
  for (int q = 0; q < 12; q++) JObj[q] += r[q];
}

void Depositor::deposit_particle_Rho(Particle *part, double *lcell, double *cellverts, double *RhoObj) {
  const emo::Geo g = cell_geo(lcell);
  double r[20];
  emo::deposit_row(g, cellverts, part->x, part->v, part->q, true, r);
  for (int q = 0; q < 8; q++) RhoObj[q] += r[12 + q];
}

// ---- factories ------------------------------------------------------------
Field_BC_Factory::Field_BC_Factory() {
  declare("periodic", make_f_periodic);
  declare("external", make_f_external);
}

Field_BC_Factory &Field_BC_Factory::getInstance() {
  static Field_BC_Factory instance;
  return instance;
}

Field_BC_Factory::Factory Field_BC_Factory::lookup(const std::string &type) {
  auto it = registry_.find(type);
  if (it == registry_.end()) throw std::runtime_error("Unable to find field boundary " + type);
  return it->second;
}

void Field_BC_Factory::Construct(Domain *domain, Grid *grids, Input_Info_t *info) {
  BC_Field **ret = new BC_Field *[6];
  for (int i = 0; i < 3; i++) {
    ret[2 * i] = lookup(info->fields_bound[2 * i])(-(i + 1), domain, grids, info);
    ret[2 * i + 1] = lookup(info->fields_bound[2 * i + 1])(i + 1, domain, grids, info);
  }
  grids->setBoundaries(ret);
}

Part_BC_Factory::Part_BC_Factory() {
  declare("periodic", make_p_periodic);
  declare("reflecting", make_p_reflecting);
}

Part_BC_Factory &Part_BC_Factory::getInstance() {
  static Part_BC_Factory instance;
  return instance;
}

Part_BC_Factory::Factory Part_BC_Factory::lookup(const std::string &type) {
  auto it = registry_.find(type);
  if (it == registry_.end()) throw std::runtime_error("Unable to find particle boundary " + type);
  return it->second;
}

BC_Particle **Part_BC_Factory::constructConditions(Domain *domain, Input_Info_t *info) {
  BC_Particle **ret = new BC_Particle *[6];
  for (int i = 0; i < 3; i++) {
    ret[2 * i] = lookup(info->parts_bound[2 * i])(domain, i, 0, info->parts_bound[2 * i]);
    ret[2 * i + 1] = lookup(info->parts_bound[2 * i + 1])(domain, i, 1, info->parts_bound[2 * i + 1]);
  }
  return ret;
}

// ---- device ---------------------------------------------------------------
emo_config make_config(const Input_Info_t &in, const int *nxyz, const double *xyz0, const double *Lxyz) {
  emo_config c;
  memset(&c, 0, sizeof(c));
  for (int d = 0; d < 3; d++) {
    c.nCell[d] = nxyz[d];
    c.xyz0[d] = xyz0[d];
    c.Lxyz[d] = Lxyz[d];
    c.E0[d] = in.E0[d];
    c.B0[d] = in.B0[d];
  }
  c.relativity = in.relativity;
  c.precision = in.precision;
  for (int s = 0; s < 6; s++) {
    c.fields_bound[s] = bc_kind(in.fields_bound[s], true);
    c.parts_bound[s] = bc_kind(in.parts_bound[s], false);
    c.bound_phi[s] = in.bound_phi[s];
    c.bound_Ax[s] = in.bound_Ax[s];
    c.bound_Ay[s] = in.bound_Ay[s];
    c.bound_Az[s] = in.bound_Az[s];
  }
  c.nspecies = in.nspecies;
  for (int i = 0; i < in.nspecies && i < NSPEC; i++) {
    c.charge[i] = in.charge_ratio[i];
    c.mass[i] = in.mass_ratio[i];
    c.isTestParticle[i] = in.isTestParticle[i];
  }
  c.nwaves = in.nwaves;
  for (int i = 0; i < in.nwaves && i < NWAVE; i++) {
    c.inSide[i] = in.inSide[i];
    c.inPolE[i] = in.inPolE[i];
    c.peakamps[i] = in.peakamps[i];
    c.omegas[i] = in.omegas[i];
    c.phases[i] = in.phases[i];
    c.invWidths[i] = in.invWidths[i];
    c.delays[i] = in.delays[i];
  }
  c.waist = in.waist;
  c.max_particles = std::max<long>(in.nparticles_tot, 1);
  return c;
}

Device::Device(const emo_config &c) : cfg(c) {
  int rc = emo_create(&cfg, &ctx_);
  if (rc != EMO_OK) throw Error(rc, "emo_create failed");
}

Device::~Device() { emo_destroy(ctx_); }

void Device::check(int rc) {
  if (rc != EMO_OK) throw Error(rc, emo_last_error(ctx_));
}

// ---- Grid -----------------------------------------------------------------
Grid::Grid(int *nxyz, int nGhosts, double *xyz0, double *Lxyz) {
  if (nGhosts != 1) throw std::runtime_error("nGhosts must be 1 (src/grid/grid.cpp:357)");
  for (int d = 0; d < 3; d++) {
    n_[d] = nxyz[d];
    x0_[d] = xyz0[d];
    L_[d] = Lxyz[d];
    d_[d] = L_[d] / n_[d];
    id_[d] = 1.0 / d_[d];
  }
}

Grid::~Grid() = default;

Device &Grid::device() {
  if (!dev_) {
    if (!g_info) throw std::runtime_error("construct a Domain from the Input_Info_t first");
    dev_ = std::make_shared<Device>(make_config(*g_info, n_, x0_, L_));
  }
  return *dev_;
}

void Grid::freeBoundaries() {
  if (!boundaries_) return;
  for (int i = 0; i < 6; i++) delete boundaries_[i];
  delete[] boundaries_;
  boundaries_ = nullptr;
}

int Grid::evolveFields(double dt) {
  device().check(emo_evolve_fields(device().ctx(), dt));
  return 0;
}

void Grid::InitializeFields(Input_Info_t *) {
  Device &d = device();
  d.check(emo_init_constant(d.ctx()));
  d.fields_ready = true;
}

void Grid::executeBC(int sendID) {
  Device &d = device();
  if (sendID == -1) d.check(emo_refresh(d.ctx(), time_phys, dt_phys));
  else if (sendID == -2) d.check(emo_fold(d.ctx()));
  else throw Error(EMO_ERR_UNSUPPORTED, "executeBC of a single field is internal to the Poisson solve");
}

int Grid::getCellID(double x, double y, double z) {
  int ix = (int)((x - x0_[0]) * id_[0]) + 1;
  int iy = (int)((y - x0_[1]) * id_[1]) + 1;
  int iz = (int)((z - x0_[2]) * id_[2]) + 1;
  return ((n_[1] + 2) * (n_[2] + 2)) * ix + (n_[2] + 2) * iy + iz;
}

double Grid::getStepSize(int dimension) {
  if (dimension < 0 || dimension > 2) return -1;
  return d_[dimension];
}

double Grid::getCellVolume() { return d_[0] * d_[1] * d_[2]; }

int Grid::getnxyzTot(int *nxyzTot) {
  for (int d = 0; d < 3; d++) nxyzTot[d] = n_[d] + 3;
  return 0;
}

int Grid::getNumberOfCells() { return (n_[0] + 2) * (n_[1] + 2) * (n_[2] + 2); }

void Grid::getField(int fieldID, double *out) {
  Device &d = device();
  d.check(emo_get_field(d.ctx(), fieldID, out));
}

void Grid::setField(int fieldID, const double *in) {
  Device &d = device();
  d.check(emo_set_field(d.ctx(), fieldID, in));
}

// ---- Poisson_Solver --------------------------------------------------------
Poisson_Solver::Poisson_Solver(Domain *domain, Input_Info_t *)
    : Grid(domain->getnxyz(), 1, domain->getxyz0(), domain->getLxyz()) {}

void Poisson_Solver::InitializeFields(Input_Info_t *info) {
  Device &d = device();
  d.check(emo_init_poisson(d.ctx(), info->poisson_iter_max, info->poisson_tol, sweeps_, nullptr));
  d.fields_ready = true;
}

// ---- Particle_Handler ----------------------------------------------------
Particle_Handler::~Particle_Handler() = default;

void Particle_Handler::Load(Input_Info_t *input_info, Domain *domain) {
  info_ = input_info;
  domain_ = domain;
  loaded_ = false;
}

void Particle_Handler::bind(Grid *grid) {
  if (dev_ && loaded_) return;
  dev_ = grid->dev_ ? grid->dev_ : (grid->device(), grid->dev_);
  if (!info_) throw std::runtime_error("Particle_Handler::Load must run before the particles are used");
  if (pusher_ && pusher_->relativity() != dev_->cfg.relativity)
    throw std::runtime_error("the pusher does not match input_info->relativity");
  emo_loader ld;
  memset(&ld, 0, sizeof(ld));
  const Input_Info_t &in = *info_;
  auto it = g_raw.find(info_);
  ld.nparticles_tot = in.nparticles_tot;
  ld.nspecies = in.nspecies;
  for (int i = 0; i < in.nspecies && i < NSPEC; i++) {
    bool raw = it != g_raw.end();
    ld.mass_ratio[i] = raw ? it->second.mass[i] : in.mass_ratio[i];
    ld.charge_ratio[i] = raw ? it->second.charge[i] : in.charge_ratio[i];
    ld.dens_frac[i] = raw ? it->second.dens[i] : in.dens_frac[i];
    ld.temp[i] = raw ? it->second.temp[i] : in.temp[i];
    ld.isTestParticle[i] = in.isTestParticle[i];
  }
  // dens_phys after ProcessInfo is the physical density; the library
  // recomputes the same super ratio from it
  ld.dens_phys = in.dens_phys;
  ld.seed = in.seed;
  ld.zero_avg = -1;
  ld.poisson_init = std::string(in.fields_init) == "poisson";
  dev_->check(emo_load_maxwellian(dev_->ctx(), &ld));
  loaded_ = true;
}

void Particle_Handler::Push(double dt) {
  if (!dev_) throw std::runtime_error("Push before the particles met a grid");
  dev_->check(emo_push(dev_->ctx(), dt));
}

long Particle_Handler::nParticles() {
  return dev_ ? (long)emo_num_particles(dev_->ctx()) : (info_ ? info_->nparticles_tot : 0);
}

void Particle_Handler::SortParticles(Particle_Compare) {
  if (dev_) dev_->check(emo_sort(dev_->ctx()));
}

void Particle_Handler::InterpolateEB(Grid *grid) {
  bind(grid);
  // the FP32 chosen form gathers inside its fused push: nothing to do here
  if (dev_->cfg.precision != EMO_FP64_PLAIN) return;
  dev_->check(emo_interpolate(dev_->ctx()));
}

void Particle_Handler::depositRhoJ(Grid *grid, bool depositRho, Domain *, Input_Info_t *) {
  bind(grid);
  dev_->check(emo_deposit(dev_->ctx(), depositRho ? 1 : 0));
}

void Particle_Handler::executeParticleBoundaryConditions() {
  if (!dev_) throw std::runtime_error("particle BCs before the particles met a grid");
  dev_->check(emo_particle_bc(dev_->ctx()));
}

void Particle_Handler::getParticles(std::vector<Particle> &out) {
  if (!dev_) throw std::runtime_error("no particles on the device yet");
  int64_t n = emo_num_particles(dev_->ctx());
  std::vector<double> x(3 * n), v(3 * n), g(n), f(6 * n);
  std::vector<int32_t> sp(n);
  std::vector<int64_t> id(n);
  bool fp64 = dev_->cfg.precision == EMO_FP64_PLAIN;
  dev_->check(emo_get_particles(dev_->ctx(), x.data(), v.data(), g.data(), sp.data(), id.data(),
                                fp64 ? f.data() : nullptr));
  out.resize(n);
  for (int64_t i = 0; i < n; i++) {
    Particle &p = out[i];
    for (int d = 0; d < 3; d++) {
      p.x[d] = x[3 * i + d];
      p.v[d] = v[3 * i + d];
    }
    p.gamma = g[i];
    p.type = sp[i];
    p.my_id = (long)id[i];
    p.q = dev_->cfg.charge[sp[i]];
    p.m = dev_->cfg.mass[sp[i]];
    p.isTestParticle = (short)dev_->cfg.isTestParticle[sp[i]];
    if (fp64) {
      p.field.e1 = f[6 * i];
      p.field.e2 = f[6 * i + 1];
      p.field.e3 = f[6 * i + 2];
      p.field.b1 = f[6 * i + 3];
      p.field.b2 = f[6 * i + 4];
      p.field.b3 = f[6 * i + 5];
    }
  }
}

void Particle_Handler::outputParticles(const char *basename, long step, Input_Info_t *input_info) {
  const double t = time_phys;
  dstep_ = input_info->nstep_parts;
  outputCount_ = input_info->output_pCount;
  bool needs = false;
  if (dT_ > 0 && t >= nextT_) {
    needs = true;
    nextT_ += dT_;
  }
  if (dstep_ > 0 && step >= nextStep_) {
    needs = true;
    nextStep_ += dstep_;
  }
  if (!tracks_init_ && !needs) return;
  if (!dev_) throw std::runtime_error("particle tracks before the particles met a grid");
  // loaded ids are 0..N-1, so no tracked id is at or above the count
  const int64_t nid = std::max<int64_t>(0, std::min<int64_t>(outputCount_, emo_num_particles(dev_->ctx())));
  std::vector<double> x(3 * nid), v(3 * nid);
  std::vector<int32_t> found(nid);
  if (nid > 0) dev_->check(emo_get_tracked(dev_->ctx(), nid, x.data(), v.data(), found.data()));
  auto name = [&](int64_t id) {
    char fname[4096];
    snprintf(fname, sizeof(fname), "%strack_%d_%ld.dat", basename, 0, (long)id);
    return std::string(fname);
  };
  if (tracks_init_) {
    tracks_init_ = false;
    for (int64_t k = 0; k < nid; k++) {
      if (!found[k]) continue;
      FILE *fp = fopen(name(k).c_str(), "w");
      if (!fp) throw std::runtime_error("cannot open " + name(k));
      fprintf(fp, "[1] time [2] x [3] y [4] z  [5] vx   [6] vy   [7] vz\n");
      fclose(fp);
    }
  }
  if (!needs) return;
  for (int64_t k = 0; k < nid; k++) {
    if (!found[k]) continue;
    FILE *fp = fopen(name(k).c_str(), "a");
    if (!fp) throw std::runtime_error("cannot open " + name(k));
    fprintf(fp, "%e %.15e %.15e %.15e %.15e %.15e %.15e\n", t, x[3 * k], x[3 * k + 1], x[3 * k + 2], v[3 * k],
            v[3 * k + 1], v[3 * k + 2]);
    fclose(fp);
  }
}

// ---- fused loop body -------------------------------------------------------
void advance(Grid *grids, Particle_Handler *handler, int nsteps) {
  if (nsteps <= 0) return;
  handler->bind(grids);
  Device &d = grids->device();
  d.check(emo_step(d.ctx(), nsteps, time_phys, dt_phys));
  for (int k = 0; k < nsteps; k++) time_phys += dt_phys;
}

// ---- OutputBoxQuantities --------------------------------------------------
OutputBoxQuantities::OutputBoxQuantities(const char *fname, Grid *grid, Particle_Handler *, Input_Info_t *)
    : filename_(fname), grid_(grid) {}

void OutputBoxQuantities::output(double t, long i) {
  if (init_) {
    init_ = false;
    FILE *fp = fopen(filename_.c_str(), "w");
    if (!fp) throw std::runtime_error("cannot open " + filename_);
    fprintf(fp, "[1] time [2] part mom x [3] part mom y [4] part mom z [5] part Energy [6] B Energy [7] E Energy [8] Total E\n");
    fclose(fp);
  }
  if (!(dStep_ > 0 && i >= nextStep_)) return;
  nextStep_ += dStep_;
  double out[8];
  Device &d = grid_->device();
  d.check(emo_diag(d.ctx(), out));
  FILE *fp = fopen(filename_.c_str(), "a");
  if (!fp) throw std::runtime_error("cannot open " + filename_);
  fprintf(fp, "%e %.15e %.15e %.15e %.15e %.15e %.15e %.15e\n", t, out[0], out[1], out[2], out[3], out[4],
          out[5], out[6]);
  fclose(fp);
}

}  // namespace emopic
