// geo.hpp -- the padded-grid geometry and the reference's unit constants,
// without HIP headers, so that the per-particle arithmetic that includes it
// also compiles as plain host C++.
#pragma once

#include <cstdint>

#if defined(__HIPCC__)
#define EMO_HD __host__ __device__
#else
#define EMO_HD
#endif

namespace emo {

// reference constants (src/globals.hpp:12,16,20)
constexpr double kUnitAcc = 0.58668774;
constexpr double kUnitRhoJ = 4.80320427e-13;
constexpr double kUnitVth = 0.002232;
constexpr double kPi = 3.14159265358979323846;

// Geometry of the padded grid, passed by value to every field kernel.
struct Geo {
  int n[3];       // physical cells
  int nt[3];      // n+3 padded points (src/grid/grid.cpp:25-27)
  int nc[3];      // n+2 cell-id extents nx_,ny_,nz_ (src/grid/grid.cpp:22-24)
  double x0[3];
  double d[3];    // Lxyz/nCell
  double id[3];   // 1.0/d
  long long sx, sy;  // strides of i and j in the padded arrays (k stride 1)
  long long ntot;
  long long ncell;
};

// Division by a fixed divisor d >= 1 through a multiply-high (Granlund and
// Montgomery, "Division by invariant integers using multiplication", 1994):
// with l = ceil(log2 d) and m = floor(2^32 (2^l - d) / d) + 1,
// floor(n / d) = (mulhi(m, n) + n) >> l for every n < 2^31.
struct FastDiv {
  uint32_t m;
  int l;
};

inline FastDiv make_fastdiv(uint32_t d) {
  int l = 0;
  while ((1ull << l) < d) l++;
  FastDiv f;
  f.m = (uint32_t)((((1ull << 32) * ((1ull << l) - d)) / d) + 1);
  f.l = l;
  return f;
}

EMO_HD inline uint32_t fastdiv(uint32_t n, FastDiv f) {
  const uint32_t t = (uint32_t)(((uint64_t)f.m * n) >> 32);
  return (t + n) >> f.l;
}

EMO_HD inline long long gidx(const Geo &g, int i, int j, int k) {
  return ((long long)i * g.nt[1] + j) * g.nt[2] + k;
}

}  // namespace emo
