// emopic.hpp -- C++ class surface of the reference (DenSto/APC_524_Project)
// over libemopic's C ABI.
//
// The classes keep the reference's names, constructor arguments and the
// method signatures its driver loop calls (src/main.cpp:114-287), so a driver
// written against the serial build ports by swapping includes.  Their state
// lives on the GPU: one libemopic context, shared by the Grid and the
// Particle_Handler of a run, created at the first call that needs the device.
//
// Differences a caller sees:
//   - Grid::getFieldPtr() (src/grid/grid.hpp:68) cannot hand out device
//     memory; getField(fieldID, out) copies a padded array out instead.
//   - Particle_Handler::getParticleVector() (particle_handler.hpp:44) becomes
//     getParticles(std::vector<Particle>&), a copy in the current array order.
//   - SortParticles is a stable counting sort by cell every step; the
//     comparator argument is accepted for source compatibility.
//   - Pusher::Step, Interpolator and Depositor on one host particle are kept
//     for single-particle use and tests (the FP64 formulas of
//     src/pusher/boris.cpp, relativisticBoris.cpp,
//     src/particles/interpolate.cpp, deposit.cpp).
//   - Errors: the reference asserts or throws; the facade throws
//     emopic::Error (derived from std::runtime_error) with libemopic's message.
#pragma once

#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/emo_api.h"

namespace emopic {

constexpr int NDIM = 3;
constexpr int NCHAR = 64;
constexpr int NSPEC = 32;
constexpr int NWAVE = 32;

// Input_Info_t (src/IO/input.hpp:15-126): the same fields, same meaning.
struct Input_Info_t {
  int nCell[NDIM] = {0, 0, 0};
  int nProc[NDIM] = {1, 1, 1};
  int nt = 0;
  int restart = 0;
  int debug = 0;
  int relativity = 0;
  int electrostatic = 0;
  int nspecies = 0;
  int nstep_fields = 10;
  int nstep_parts = 10;
  int nstep_restart = 100;
  int nstep_sort = 100;
  int which_fields = -1;
  int output_pCount = 0;
  int nwaves = 0;
  int inSide[NWAVE] = {};
  int inPolE[NWAVE] = {};
  int isTestParticle[NSPEC] = {};
  long nparticles_tot = 0;
  long nparticles_domain = 0;
  double t0 = 0.0;
  double dens_phys = 0.0;
  double super_ratio = 0.0;
  double mass_ratio[NSPEC] = {};
  double charge_ratio[NSPEC] = {};
  double dens_frac[NSPEC] = {};
  double temp[NSPEC] = {};
  double peakamps[NWAVE] = {}, omegas[NWAVE] = {}, phases[NWAVE] = {};
  double invWidths[NWAVE] = {}, delays[NWAVE] = {};
  double E0[NDIM] = {}, B0[NDIM] = {};
  double bound_phi[2 * NDIM] = {}, bound_Ax[2 * NDIM] = {}, bound_Ay[2 * NDIM] = {},
         bound_Az[2 * NDIM] = {};
  double xyz0[NDIM] = {}, Lxyz[NDIM] = {};
  char distname[NCHAR] = "";
  char parts_init[NCHAR] = "random";
  char fields_init[NCHAR] = "constant";
  char parts_bound[2 * NDIM][NCHAR] = {"periodic", "periodic", "periodic", "periodic", "periodic", "periodic"};
  char fields_bound[2 * NDIM][NCHAR] = {"periodic", "periodic", "periodic", "periodic", "periodic", "periodic"};
  // additions of this implementation
  int precision = EMO_FP64_PLAIN;  // EMO_FP32_CHOSEN for the FP32 form
  long seed = -1;                  // Particle_Handler::Load hard-codes -1
  double waist = 0.0;              // transverse Gaussian waist of injected pulses (0: uniform)
  int poisson_iter_max = 10000;    // ITER_MAX (src/poisson/poisson.cpp:14)
  double poisson_tol = -1.0;       // < 0: the reference's tolerance rule
  double dt = -1.0;                // < 0: min(dx)/100, as src/main.cpp:213-214 sets it
};

class Error : public std::runtime_error {
 public:
  Error(int code, const std::string &msg) : std::runtime_error(msg), code_(code) {}
  int code() const { return code_; }

 private:
  int code_;
};

// Input (src/IO/input.hpp:130-144): holds the Input_Info_t; ProcessInfo
// validates it and applies the super-particle scaling (src/IO/check.cpp:12-147).
class Input {
 public:
  Input() = default;
  // readinfo (src/IO/readinfo.cpp:18-733): the reference's libconfig input
  // file; 0 on success, EXIT_FAILURE after a message on stderr
  int readinfo(const char *fname);
  int ProcessInfo();
  Input_Info_t *getinfo() { return &info_; }

 private:
  Input_Info_t info_;
};

// Domain (src/domain/domain.hpp:12-58), serial: the whole box.
class Domain {
 public:
  explicit Domain(Input_Info_t *input_info);
  int *getnxyz() { return nxyz_; }
  double *getxyz0() { return xyz0_; }
  double *getLxyz() { return Lxyz_; }
  double getmindx();
  Input_Info_t *info() { return info_; }

 private:
  int nxyz_[3];
  double xyz0_[3], Lxyz_[3];
  Input_Info_t *info_;
};

// Particle record (src/particles/particle.hpp:4-31), the fields the step uses.
struct Field_part {
  double e1 = 0, e2 = 0, e3 = 0;
  double b1 = 0, b2 = 0, b3 = 0;
};
struct Particle {
  double x[3] = {0, 0, 0};
  double v[3] = {0, 0, 0};
  double gamma = 0.0;
  int type = 0;
  long my_id = 0;
  double q = 0.0;
  double m = 1.0;
  short isGhost = 0;
  short isTestParticle = 0;
  Field_part field;
};

// Interpolator (src/particles/interpolate.hpp:6-11): the E and B of one host
// particle from its cell's 21 values (the least vertex, 4 E edges per
// component, 2 (B + B_tm1)/2 faces per component, as
// Grid::getFieldInterpolatorVec orders them); FP64, the operation order of
// src/particles/interpolate.cpp:11-47, which the device gather also follows.
class Interpolator {
 public:
  Interpolator() = default;
  ~Interpolator() = default;
  void interpolate_fields(double *pos, double *lcell, double *cellvars, Field_part *field);
};

// Depositor (src/particles/deposit.hpp:6-14): one host particle's current
// on its cell's 12 edges and charge on its 8 vertices, added into JObj and
// RhoObj; FP64, the operation order of src/particles/deposit.cpp:12-62,
// which the device deposit also follows.
class Depositor {
 public:
  Depositor() = default;
  ~Depositor() = default;
  void deposit_particle_J(Particle *part, double *lcell, double *cellverts, double *JObj);
  void deposit_particle_Rho(Particle *part, double *lcell, double *cellverts, double *RhoObj);
};

// Pushers (src/pusher/pusher.hpp:6-10): select the device kernel; Step on a
// single host particle restates the reference formulas in FP64.
class Pusher {
 public:
  virtual ~Pusher() = default;
  virtual int Step(Particle *part, Field_part *field, double dt) = 0;
  virtual int relativity() const = 0;
};
class Boris : public Pusher {
 public:
  int Step(Particle *part, Field_part *field, double dt) override;
  int relativity() const override { return 0; }
};
class Relativistic_Boris : public Pusher {
 public:
  int Step(Particle *part, Field_part *field, double dt) override;
  int relativity() const override { return 1; }
};

class Device;  // the shared libemopic context of a run

// Field boundary conditions (src/boundaries/fields_boundary.hpp:7-15) and
// their string-keyed factory (src/boundaries/field_bc_factory.hpp:20-87).
class BC_Field {
 public:
  virtual ~BC_Field() = default;
  virtual int completeBC(int sendID) { return 0; }  // applied by Grid::executeBC on the device
  virtual int kind() const = 0;                     // EMO_BC_PERIODIC / EMO_BC_EXTERNAL
  int side() const { return side_; }

 protected:
  int side_ = 0;
};

class Grid;
class Field_BC_Factory {
 public:
  typedef BC_Field *(*Factory)(int side, Domain *domain, Grid *grids, Input_Info_t *info);
  static Field_BC_Factory &getInstance();
  void Construct(Domain *domain, Grid *grids, Input_Info_t *input_info);
  void declare(const std::string &type, Factory factory) { registry_[type] = factory; }
  Factory lookup(const std::string &type);

 private:
  Field_BC_Factory();
  std::map<std::string, Factory> registry_;
};

// Particle boundary conditions (src/boundaries/particles_boundary.hpp:33-43)
// and their factory (src/boundaries/particle_bc_factory.hpp:21-88).
class BC_Particle {
 public:
  virtual ~BC_Particle() = default;
  virtual int kind() const = 0;  // EMO_BC_PERIODIC / EMO_BC_REFLECTING
  int dim() const { return dim_; }
  bool isRight() const { return isRight_; }

 protected:
  int dim_ = 0;
  bool isRight_ = false;
};

class Part_BC_Factory {
 public:
  typedef BC_Particle *(*Factory)(Domain *domain, int dim_Index, short isRight, std::string type);
  static Part_BC_Factory &getInstance();
  BC_Particle **constructConditions(Domain *domain, Input_Info_t *info);
  void declare(const std::string &type, Factory factory) { registry_[type] = factory; }
  Factory lookup(const std::string &type);

 private:
  Part_BC_Factory();
  std::map<std::string, Factory> registry_;
};

// Grid (src/grid/grid.hpp:22-232).
class Grid {
 public:
  Grid(int *nxyz, int nGhosts, double *xyz0, double *Lxyz);
  virtual ~Grid();

  int evolveFields(double dt);
  virtual void InitializeFields(Input_Info_t *input_info);
  void executeBC(int sendID);
  void setBoundaries(BC_Field **bc) { boundaries_ = bc; }
  void freeBoundaries();

  int getCellID(double x, double y, double z);
  double getStepSize(int dimension);
  double getCellVolume();
  int getnxyzTot(int *nxyzTot);
  int getNumberOfCells();

  // replaces getFieldPtr(): copies the padded array of fieldID (ids of
  // src/grid/grid.cpp:45-60) into out[(nx+3)(ny+3)(nz+3)]
  void getField(int fieldID, double *out);
  void setField(int fieldID, const double *in);

  Device &device();
  BC_Field **boundaries() { return boundaries_; }

 protected:
  int n_[3];
  double x0_[3], L_[3], d_[3], id_[3];
  BC_Field **boundaries_ = nullptr;
  std::shared_ptr<Device> dev_;
  friend class Particle_Handler;
};

// Poisson_Solver (src/poisson/poisson.hpp:8-80): InitializeFields runs the
// Jacobi solves, phiToE, AToB and executeBC(-1) on the device.
class Poisson_Solver : public Grid {
 public:
  Poisson_Solver(Domain *domain, Input_Info_t *input_info);
  void InitializeFields(Input_Info_t *input_info) override;
  const int *sweeps() const { return sweeps_; }

 private:
  int sweeps_[4] = {0, 0, 0, 0};
};

class Particle_Compare {
 public:
  explicit Particle_Compare(Grid *) {}
};

// Particle_Handler (src/particles/particle_handler.hpp:21-63).
class Particle_Handler {
 public:
  Particle_Handler() = default;
  ~Particle_Handler();
  void Load(Input_Info_t *input_info, Domain *domain);
  void Push(double dt);
  long nParticles();
  void SortParticles(Particle_Compare comp);
  void setPusher(Pusher *pusher) { pusher_ = pusher; }
  void clearGhosts() {}  // ghosts never exist: the BCs act in place
  void InterpolateEB(Grid *grid);
  void depositRhoJ(Grid *grid, bool depositRho, Domain *domain, Input_Info_t *input_info);
  void setParticleBoundaries(BC_Particle **bc) { boundaries_ = bc; }
  void executeParticleBoundaryConditions();
  void getParticles(std::vector<Particle> &out);
  // outputParticles (src/particles/particle_handler.cpp:308-365): at the
  // first call, a file <basename>track_0_<id>.dat with the reference's header
  // for each particle with my_id < output_pCount; then, every nstep_parts
  // steps, one row "%e %.15e x6" of time_phys, x, v appended to each.  The
  // tracked particles are selected on the device (emo_get_tracked).
  void outputParticles(const char *basename, long step, Input_Info_t *input_info);
  // attaches to the grid's device context and loads the particles (first use)
  void bind(Grid *grid);

 private:
  Pusher *pusher_ = nullptr;
  BC_Particle **boundaries_ = nullptr;
  std::shared_ptr<Device> dev_;
  Input_Info_t *info_ = nullptr;
  Domain *domain_ = nullptr;
  bool loaded_ = false;
  // track-file cadence (src/particles/particle_handler.cpp:24-27, 311-341)
  bool tracks_init_ = true;
  double dT_ = -1.0, nextT_ = 0.0;
  long dstep_ = -1, nextStep_ = 0;
  long outputCount_ = 0;
};

// OutputBoxQuantities (src/IO/output.hpp): history.dat every dStep_ = 10
// steps with the reference's header and column format (src/IO/output.cpp:52-123).
class OutputBoxQuantities {
 public:
  OutputBoxQuantities(const char *fname, Grid *grid, Particle_Handler *handler, Input_Info_t *info);
  void output(double t, long i);

 private:
  std::string filename_;
  Grid *grid_;
  bool init_ = true;
  long dStep_ = 10, nextStep_ = 0;
};

// The libemopic context shared by a run's Grid and Particle_Handler.
class Device {
 public:
  Device(const emo_config &cfg);
  ~Device();
  emo_ctx *ctx() { return ctx_; }
  void check(int rc);
  emo_config cfg;
  bool fields_ready = false;
  bool gathered = false;

 private:
  emo_ctx *ctx_ = nullptr;
};

// The loop body of src/main.cpp:225-275 (sort, push, particle BCs, deposit,
// fold, Yee, ghost refresh / LightBC, gather, time_phys += dt_phys) as one
// fused device call per nsteps; no counterpart in the reference, which calls
// the phases one by one.  The only loop form of the FP32 chosen precision.
void advance(Grid *grids, Particle_Handler *handler, int nsteps);

// process-wide simulation clock (src/globals.hpp:35-46)
extern double time_phys;
extern double dt_phys;

// Builds the emo_config of a run from the Input_Info_t and the geometry.
emo_config make_config(const Input_Info_t &info, const int *nxyz, const double *xyz0,
                       const double *Lxyz);

}  // namespace emopic
