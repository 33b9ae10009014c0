// movecode.hpp -- the one-cell move codes of the stable counting sort by
// destination cell (cellsort.hpp), shared by the push kernels of both
// precisions and the FP32 adjoint.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace emo {

constexpr uint8_t kFarCode = 255;  // moved beyond the one-cell neighbourhood

// One-dimensional move between the source index s and destination index t
// (padded indices): 0, 1, 2 for t - s = -1, 0, +1; with periodic particle
// BCs 3, 4, 5 for the wraps t - s = n - 1, 1 - n, -n (a particle leaving
// through the lower / upper face, or from the upper ghost plane where a
// particle sitting exactly on x = L lives, src/boundaries/b_particles/
// bc_p_periodic.cpp:44-87).  -1: not a one-cell move.
__device__ inline int move_code(int sig, int n, int per) {
  if (sig >= -1 && sig <= 1) return sig + 1;
  if (per) {
    // a wrap whose shift equals a one-cell move (n <= 2) takes that code
    if (n >= 3 && sig == n - 1) return 3;
    if (n >= 3 && sig == 1 - n) return 4;
    if (n >= 2 && sig == -n) return 5;
  }
  return -1;
}
// whether move code k can occur in a dimension of n cells
__device__ inline bool move_live(int k, int n, int per) {
  if (k < 3) return true;
  return per && (k == 5 ? n >= 2 : n >= 3);
}
__device__ inline int move_sigma(int k, int n) {
  switch (k) {
    case 0: return -1;
    case 1: return 0;
    case 2: return 1;
    case 3: return n - 1;
    case 4: return 1 - n;
    default: return -n;
  }
}

}  // namespace emo
