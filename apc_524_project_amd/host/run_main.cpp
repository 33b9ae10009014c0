// run_main.cpp -- the reference's serial driver loop (src/main.cpp:114-287)
// written against the emopic facade, as a caller of the drop-in would.
//
//   emo_run_main <settings.txt> <outdir> [--fused]
//
// settings.txt holds one Input_Info_t field per line, `name = v1 v2 ...`
// (# starts a comment; strings unquoted), using the field names of
// src/IO/input.hpp:15-126 plus `dt` (default min(dx)/100, src/main.cpp:213-214)
// and the facade additions precision / seed / waist / poisson_iter_max /
// poisson_tol.  The libconfig front end of the reference (src/IO/readinfo.cpp)
// is not part of this library.
//
// Writes <outdir>/history.dat (OutputBoxQuantities) and <outdir>/particles.bin
// (int64 n; then x[3n], v[3n], gamma[n] as doubles and id[n] as int64, in
// ascending id order) after the last step.  --fused runs the loop body as one
// device call per step (required for the FP32 chosen form); the default calls
// the phases one by one in the reference's order.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "emopic.hpp"

using namespace emopic;

namespace {

template <class T>
void put(std::istringstream &is, T *dst, int cap, const std::string &key) {
  int k = 0;
  T v;
  while (is >> v) {
    if (k >= cap) throw std::runtime_error("too many values for " + key);
    dst[k++] = v;
  }
  if (k == 0) throw std::runtime_error("no value for " + key);
  if (cap == 3 && k == 1) dst[1] = dst[2] = dst[0];  // readinfo's single-value rule
}

void put_str(std::istringstream &is, char (*dst)[NCHAR], int cap, const std::string &key) {
  int k = 0;
  std::string v;
  while (is >> v) {
    if (k >= cap || v.size() >= NCHAR) throw std::runtime_error("bad value for " + key);
    strcpy(dst[k++], v.c_str());
  }
  if (k != cap) throw std::runtime_error(key + " needs " + std::to_string(cap) + " values");
}

void read_settings(const char *fname, Input_Info_t *in, double *dt) {
  std::ifstream f(fname);
  if (!f) throw std::runtime_error(std::string("cannot open ") + fname);
  std::string line;
  while (std::getline(f, line)) {
    auto h = line.find('#');
    if (h != std::string::npos) line.resize(h);
    auto eq = line.find('=');
    if (eq == std::string::npos) {
      if (line.find_first_not_of(" \t\r") != std::string::npos)
        throw std::runtime_error("expected name = value: " + line);
      continue;
    }
    std::string key = line.substr(0, eq);
    key.erase(std::remove_if(key.begin(), key.end(), ::isspace), key.end());
    std::istringstream is(line.substr(eq + 1));
    if (key == "nCell") put(is, in->nCell, 3, key);
    else if (key == "xyz0") put(is, in->xyz0, 3, key);
    else if (key == "Lxyz") put(is, in->Lxyz, 3, key);
    else if (key == "nt" || key == "nTimesteps") put(is, &in->nt, 1, key);
    else if (key == "t0" || key == "startTime") put(is, &in->t0, 1, key);
    else if (key == "debug") put(is, &in->debug, 1, key);
    else if (key == "relativity") put(is, &in->relativity, 1, key);
    else if (key == "nspecies") put(is, &in->nspecies, 1, key);
    else if (key == "nstep_sort") put(is, &in->nstep_sort, 1, key);
    else if (key == "nparticles_tot" || key == "nParticles_tot") put(is, &in->nparticles_tot, 1, key);
    else if (key == "dens_phys") put(is, &in->dens_phys, 1, key);
    else if (key == "mass_ratio") put(is, in->mass_ratio, NSPEC, key);
    else if (key == "charge_ratio") put(is, in->charge_ratio, NSPEC, key);
    else if (key == "dens_frac") put(is, in->dens_frac, NSPEC, key);
    else if (key == "temp" || key == "temperature") put(is, in->temp, NSPEC, key);
    else if (key == "isTestParticle") put(is, in->isTestParticle, NSPEC, key);
    else if (key == "E0") put(is, in->E0, 3, key);
    else if (key == "B0") put(is, in->B0, 3, key);
    else if (key == "fields_init" || key == "init") {
      std::string v;
      is >> v;
      if (v.empty() || v.size() >= NCHAR) throw std::runtime_error("bad fields_init");
      strcpy(in->fields_init, v.c_str());
    } else if (key == "parts_bound") put_str(is, in->parts_bound, 6, key);
    else if (key == "fields_bound") put_str(is, in->fields_bound, 6, key);
    else if (key == "nwaves") put(is, &in->nwaves, 1, key);
    else if (key == "inSide") put(is, in->inSide, NWAVE, key);
    else if (key == "inPolE") put(is, in->inPolE, NWAVE, key);
    else if (key == "peakamps") put(is, in->peakamps, NWAVE, key);
    else if (key == "omegas") put(is, in->omegas, NWAVE, key);
    else if (key == "phases") put(is, in->phases, NWAVE, key);
    else if (key == "invWidths") put(is, in->invWidths, NWAVE, key);
    else if (key == "delays") put(is, in->delays, NWAVE, key);
    else if (key == "bound_phi") put(is, in->bound_phi, 6, key);
    else if (key == "bound_Ax") put(is, in->bound_Ax, 6, key);
    else if (key == "bound_Ay") put(is, in->bound_Ay, 6, key);
    else if (key == "bound_Az") put(is, in->bound_Az, 6, key);
    else if (key == "precision") put(is, &in->precision, 1, key);
    else if (key == "seed") put(is, &in->seed, 1, key);
    else if (key == "waist") put(is, &in->waist, 1, key);
    else if (key == "poisson_iter_max") put(is, &in->poisson_iter_max, 1, key);
    else if (key == "poisson_tol") put(is, &in->poisson_tol, 1, key);
    else if (key == "dt") put(is, dt, 1, key);
    else throw std::runtime_error("unknown setting " + key);
  }
}

void write_particles(const std::string &path, Particle_Handler *ph) {
  std::vector<Particle> parts;
  ph->getParticles(parts);
  std::sort(parts.begin(), parts.end(), [](const Particle &a, const Particle &b) { return a.my_id < b.my_id; });
  int64_t n = (int64_t)parts.size();
  std::vector<double> x(3 * n), v(3 * n), g(n);
  std::vector<int64_t> id(n);
  for (int64_t i = 0; i < n; i++) {
    for (int d = 0; d < 3; d++) {
      x[3 * i + d] = parts[i].x[d];
      v[3 * i + d] = parts[i].v[d];
    }
    g[i] = parts[i].gamma;
    id[i] = parts[i].my_id;
  }
  FILE *fp = fopen(path.c_str(), "wb");
  if (!fp) throw std::runtime_error("cannot write " + path);
  bool ok = fwrite(&n, sizeof(n), 1, fp) == 1 && fwrite(x.data(), sizeof(double), 3 * n, fp) == (size_t)(3 * n) &&
            fwrite(v.data(), sizeof(double), 3 * n, fp) == (size_t)(3 * n) &&
            fwrite(g.data(), sizeof(double), n, fp) == (size_t)n &&
            fwrite(id.data(), sizeof(int64_t), n, fp) == (size_t)n;
  fclose(fp);
  if (!ok) throw std::runtime_error("short write to " + path);
}

int run(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: %s <settings.txt> <outdir> [--fused]\n", argv[0]);
    return 2;
  }
  bool fused = argc > 3 && std::string(argv[3]) == "--fused";
  std::string dir = std::string(argv[2]) + "/";

  Input *input = new Input();
  Input_Info_t *input_info = input->getinfo();
  double dt = -1.0;
  read_settings(argv[1], input_info, &dt);
  if (input->ProcessInfo() != 0) {
    fprintf(stderr, "input check failed\n");
    return 1;
  }
  if (input_info->precision != EMO_FP64_PLAIN) fused = true;
  bool poisson = strcmp(input_info->fields_init, "poisson") == 0;

  // set-up, src/main.cpp:114-189
  Domain *domain = new Domain(input_info);
  Particle_Handler *part_handler = new Particle_Handler();
  Pusher *pusher = input_info->relativity == 0 ? (Pusher *)new Boris() : (Pusher *)new Relativistic_Boris();
  part_handler->setPusher(pusher);
  BC_Particle **bc = Part_BC_Factory::getInstance().constructConditions(domain, input_info);
  part_handler->setParticleBoundaries(bc);
  Grid *grids = poisson ? new Poisson_Solver(domain, input_info)
                        : new Grid(domain->getnxyz(), 1, domain->getxyz0(), domain->getLxyz());
  Field_BC_Factory::getInstance().Construct(domain, grids, input_info);
  part_handler->Load(input_info, domain);
  Particle_Compare compare(grids);
  if (poisson && !fused) {
    part_handler->depositRhoJ(grids, true, domain, input_info);
    grids->executeBC(-2);
  }
  part_handler->bind(grids);
  grids->InitializeFields(input_info);
  part_handler->InterpolateEB(grids);
  OutputBoxQuantities *boxOutput =
      new OutputBoxQuantities((dir + "history.dat").c_str(), grids, part_handler, input_info);

  int nt = input_info->nt;
  time_phys = input_info->t0;
  dt_phys = dt > 0 ? dt : domain->getmindx() / 1 / 100.0;

  // time loop, src/main.cpp:221-287
  for (int ti = 0; ti < nt; ti++) {
    if (fused) {
      boxOutput->output(time_phys, ti);
      advance(grids, part_handler, 1);
      continue;
    }
    if (input_info->nstep_sort > 0 && ti % input_info->nstep_sort == 0) part_handler->SortParticles(compare);
    boxOutput->output(time_phys, ti);
    part_handler->Push(dt_phys);
    part_handler->executeParticleBoundaryConditions();
    part_handler->clearGhosts();
    part_handler->depositRhoJ(grids, true, domain, input_info);
    grids->executeBC(-2);
    grids->evolveFields(dt_phys);
    grids->executeBC(-1);
    part_handler->InterpolateEB(grids);
    time_phys += dt_phys;
  }
  write_particles(dir + "particles.bin", part_handler);

  delete boxOutput;
  delete part_handler;
  grids->freeBoundaries();
  delete grids;
  for (int i = 0; i < 6; i++) delete bc[i];
  delete[] bc;
  delete pusher;
  delete domain;
  delete input;
  return 0;
}

}  // namespace

int main(int argc, char **argv) {
  try {
    return run(argc, argv);
  } catch (const std::exception &e) {
    fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
}
