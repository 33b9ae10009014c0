// uniform_load.hpp -- the device-side particle generator of emo_load_uniform
// (an extension without a reference counterpart): particle i of the load is
// the r-th particle of physical cell i / ppc, whose species follows from r
// and the per-cell species counts.  Random numbers are counter-based, so any
// particle can be generated independently and both precisions load the same
// particles: draw k of particle i is SplitMix64 (Steele, Lea and Flood,
// "Fast splittable pseudorandom number generators", OOPSLA 2014) of
// key + 8 i + k, its top 53 bits as a double in [0, 1).  Gaussians are
// Box-Muller pairs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "geo.hpp"
#include "../../include/emo_api.h"

namespace emo {

struct UniformGen {
  Geo g;
  uint64_t key;
  int nspec;
  int ppc;                  // particles per cell, all species
  int ppc_end[EMO_NSPEC];   // cumulative per-cell counts by species
  double vth[EMO_NSPEC];    // UNIT_VTH sqrt(T/m)
  double drift[EMO_NSPEC][3];
};

EMO_HD inline uint64_t splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double ugen_uniform(const UniformGen &u, long long i, int k) {
  return (double)(splitmix64(u.key + (uint64_t)i * 8u + (uint64_t)k) >> 11) * 0x1.0p-53;
}

// position (x0 + d (p + U)), velocity, stored gamma and species of particle i
__device__ inline void ugen_particle(const UniformGen &u, long long i, double x[3], double v[3], double *gam,
                                     int *sp) {
  const Geo &g = u.g;
  const long long cell = i / u.ppc;
  const int r = (int)(i - cell * u.ppc);
  int s = 0;
  while (s < u.nspec - 1 && r >= u.ppc_end[s]) s++;
  const int pz = (int)(cell % g.n[2]);
  const long long t = cell / g.n[2];
  const int py = (int)(t % g.n[1]);
  const int px = (int)(t / g.n[1]);
  const int pc[3] = {px, py, pz};
  for (int d = 0; d < 3; d++) x[d] = g.x0[d] + g.d[d] * ((double)pc[d] + ugen_uniform(u, i, d));
  const double two_pi = 2.0 * 3.14159265358979323846;
  const double r1 = sqrt(-2.0 * log(1.0 - ugen_uniform(u, i, 3)));
  const double a1 = two_pi * ugen_uniform(u, i, 4);
  const double r2 = sqrt(-2.0 * log(1.0 - ugen_uniform(u, i, 5)));
  const double a2 = two_pi * ugen_uniform(u, i, 6);
  const double gs[3] = {r1 * cos(a1), r1 * sin(a1), r2 * cos(a2)};
  for (int d = 0; d < 3; d++) v[d] = u.vth[s] * gs[d] + u.drift[s][d];
  // gamma = sqrt(1 - v^2) as the reference's loader stores it
  // (src/particles/particle_handler.cpp:77)
  *gam = sqrt(1.0 - v[0] * v[0] - v[1] * v[1] - v[2] * v[2]);
  *sp = s;
}

}  // namespace emo
