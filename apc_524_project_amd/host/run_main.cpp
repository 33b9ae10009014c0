// run_main.cpp -- the reference's serial driver loop (src/main.cpp:114-287)
// written against the emopic facade, as a caller of the drop-in would.
//
//   emo_run_main <input.txt> [outdir] [--fused | --print-input]
//
// input.txt is the reference's libconfig input file (Input::readinfo,
// src/IO/readinfo.cpp), read unchanged; an optional group `emopic` holds the
// facade's additions (precision, seed, waist, dt, poisson_iter_max,
// poisson_tol).  dt defaults to min(dx)/100 (src/main.cpp:213-214).
//
// Outputs go to outdir, or, as in the reference (src/main.cpp:182-185), to
// the input file's directory: history.dat (OutputBoxQuantities), the particle
// tracks track_0_<id>.dat (Particle_Handler::outputParticles, written at step
// 0 and after every step at the nstep_parts cadence), and particles.bin
// (int64 n; then x[3n], v[3n], gamma[n] as doubles and id[n] as int64, in
// ascending id order) after the last step.  The HDF5 field output
// (src/IO/hdf5io.cpp) is not written: no HDF5 library is available.
//
// --fused runs the loop body as one device call per step (required for the
// FP32 chosen form); the default calls the phases one by one in the
// reference's order.  --print-input prints what readinfo read and exits.
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

// --print-input: the Input_Info_t that readinfo produced, one `name = values`
// line per field, before ProcessInfo scales it; no device call
void print_input(const Input_Info_t &in) {
  auto ints = [](const char *name, const int *v, int n) {
    printf("%s =", name);
    for (int i = 0; i < n; i++) printf(" %d", v[i]);
    printf("\n");
  };
  auto dbls = [](const char *name, const double *v, int n) {
    printf("%s =", name);
    for (int i = 0; i < n; i++) printf(" %.17g", v[i]);
    printf("\n");
  };
  auto strs = [](const char *name, const char (*v)[NCHAR], int n) {
    printf("%s =", name);
    for (int i = 0; i < n; i++) printf(" %s", v[i]);
    printf("\n");
  };
  const int ns = std::max(0, std::min(in.nspecies, NSPEC)), nw = std::max(0, std::min(in.nwaves, NWAVE));
  ints("nCell", in.nCell, 3);
  ints("nProc", in.nProc, 3);
  dbls("xyz0", in.xyz0, 3);
  dbls("Lxyz", in.Lxyz, 3);
  ints("nt", &in.nt, 1);
  dbls("t0", &in.t0, 1);
  ints("debug", &in.debug, 1);
  ints("nstep_sort", &in.nstep_sort, 1);
  ints("restart", &in.restart, 1);
  dbls("dens_phys", &in.dens_phys, 1);
  ints("relativity", &in.relativity, 1);
  ints("nspecies", &in.nspecies, 1);
  dbls("mass_ratio", in.mass_ratio, ns);
  dbls("charge_ratio", in.charge_ratio, ns);
  dbls("temp", in.temp, ns);
  dbls("dens_frac", in.dens_frac, ns);
  ints("isTestParticle", in.isTestParticle, ns);
  printf("nparticles_tot = %ld\n", in.nparticles_tot);
  strs("parts_bound", in.parts_bound, 6);
  ints("electrostatic", &in.electrostatic, 1);
  printf("fields_init = %s\n", in.fields_init);
  dbls("B0", in.B0, 3);
  dbls("E0", in.E0, 3);
  strs("fields_bound", in.fields_bound, 6);
  ints("nwaves", &in.nwaves, 1);
  ints("inSide", in.inSide, nw);
  ints("inPolE", in.inPolE, nw);
  dbls("peakamps", in.peakamps, nw);
  dbls("omegas", in.omegas, nw);
  dbls("phases", in.phases, nw);
  dbls("invWidths", in.invWidths, nw);
  dbls("delays", in.delays, nw);
  dbls("bound_phi", in.bound_phi, 6);
  dbls("bound_Ax", in.bound_Ax, 6);
  dbls("bound_Ay", in.bound_Ay, 6);
  dbls("bound_Az", in.bound_Az, 6);
  ints("nstep_fields", &in.nstep_fields, 1);
  ints("nstep_parts", &in.nstep_parts, 1);
  ints("nstep_restart", &in.nstep_restart, 1);
  ints("output_pCount", &in.output_pCount, 1);
  ints("which_fields", &in.which_fields, 1);
  ints("precision", &in.precision, 1);
  printf("seed = %ld\n", in.seed);
  dbls("waist", &in.waist, 1);
  dbls("dt", &in.dt, 1);
  ints("poisson_iter_max", &in.poisson_iter_max, 1);
  dbls("poisson_tol", &in.poisson_tol, 1);
}

int run(int argc, char **argv) {
  std::vector<std::string> pos;
  bool fused = false, print_only = false;
  for (int k = 1; k < argc; k++) {
    if (std::string(argv[k]) == "--fused") fused = true;
    else if (std::string(argv[k]) == "--print-input") print_only = true;
    else pos.push_back(argv[k]);
  }
  if (pos.empty() || pos.size() > 2) {
    fprintf(stderr, "usage: %s <input.txt> [outdir] [--fused | --print-input]\n", argv[0]);
    return 2;
  }
  const std::string inputname = pos[0];
  std::string dir = pos.size() > 1 ? pos[1] + "/" : inputname.substr(0, inputname.find_last_of('/') + 1);

  Input *input = new Input();
  Input_Info_t *input_info = input->getinfo();
  if (input->readinfo(inputname.c_str()) != 0) {
    fprintf(stderr, "input file %s could not be read\n", inputname.c_str());
    return 1;
  }
  if (print_only) {
    print_input(*input_info);
    return 0;
  }
  if (input_info->electrostatic != 0) {
    fprintf(stderr, "electrostatic field evolution (Grid::evolveFieldsES) is not part of this library\n");
    return 1;
  }
  if (input->ProcessInfo() != 0) {
    fprintf(stderr, "input check failed\n");
    return 1;
  }
  const double dt = input_info->dt;
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
  part_handler->outputParticles(dir.c_str(), 0, input_info);

  int nt = input_info->nt;
  time_phys = input_info->t0;
  dt_phys = dt > 0 ? dt : domain->getmindx() / 1 / 100.0;

  // time loop, src/main.cpp:221-287
  for (int ti = 0; ti < nt; ti++) {
    if (fused) {
      boxOutput->output(time_phys, ti);
      advance(grids, part_handler, 1);
      part_handler->outputParticles(dir.c_str(), ti + 1, input_info);
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
    part_handler->outputParticles(dir.c_str(), ti + 1, input_info);
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
