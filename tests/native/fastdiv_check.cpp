// Host build of the multiply-high division in geo.hpp: counts the n in
// [0, nmax) for which fastdiv(n, make_fastdiv(d)) differs from n / d.
#include <cstdint>

#include "../../apc_524_project_amd/csrc/geo.hpp"

extern "C" long long fastdiv_mismatches(uint32_t d, uint32_t nmax, uint32_t stride) {
  const emo::FastDiv f = emo::make_fastdiv(d);
  long long bad = 0;
  for (uint64_t n = 0; n < nmax; n += stride)
    if (emo::fastdiv((uint32_t)n, f) != (uint32_t)n / d) bad++;
  // the top of the range, one by one
  for (uint64_t n = nmax > 4096 ? nmax - 4096 : 0; n < nmax; n++)
    if (emo::fastdiv((uint32_t)n, f) != (uint32_t)n / d) bad++;
  return bad;
}
