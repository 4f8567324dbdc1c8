// model_util.h — small helpers shared by the host-side solver drivers.
#pragma once

#include <algorithm>

#include "ctc_internal.h"
#include "kernels_common.h"

namespace ctc {

inline CissArgs ciss_args(const DevCiss &d) {
  CissArgs a{};
  for (int s = 0; s < 3; ++s) a.oidx[s] = d.oidx[s];
  a.vals = d.vals;
  a.tile_slot = d.tile_slot;
  a.tile_mask = d.tile_mask;
  a.directory = d.directory;
  a.rowptr = d.rowptr;
  a.nnz = d.nnz;
  a.ntiles = d.ntiles;
  a.dlength = d.dlength;
  return a;
}

// factor pointers relative to a CISS mode: f[0] = own mode, f[1+s] = mode+1+s
inline FactorArgs rel_factors(float *const *F, uint32_t N, uint32_t mode) {
  FactorArgs f{};
  for (uint32_t s = 0; s < N; ++s) f.f[s] = F[(mode + s) % N];
  return f;
}

// tiles per span from the options' tile_len (nonzeros per wave work unit)
inline uint32_t span_tiles_of(const ctc_options *o) {
  uint32_t tl = o && o->tile_len ? o->tile_len : 1024;  // 0: the measured default
  tl = std::max<uint32_t>(64, (tl + 63) / 64 * 64);
  return tl / 64;
}

inline unsigned block_of(const ctc_options *o) {
  unsigned b = o && o->block ? o->block : 256;
  if (b < 64) b = 64;
  if (b > 256) b = 256;
  return b / 64 * 64;
}

}  // namespace ctc
