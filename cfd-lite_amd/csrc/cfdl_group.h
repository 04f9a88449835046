// cfdl_group.h -- grouped launches (HIP translation units only): one launch
// runs the same step for several independent systems ("slots").  Block b of
// the launch belongs to the slot whose range [b0[i], b0[i+1]) holds it, and
// runs as that slot's block b - b0[i].
#pragma once

#include <hip/hip_runtime.h>

#include "cfdl_internal.h"

namespace cfdl {

__device__ inline int group_slot(const GroupMap &g, int &lb) {
  int i = 0;
  while (i + 1 < g.n && static_cast<int>(blockIdx.x) >= g.b0[i + 1]) i++;
  lb = static_cast<int>(blockIdx.x) - g.b0[i];
  return g.slot[i];
}

// host: the map of slots taking part (slot ids and their block counts);
// returns the total block count
inline int group_map(GroupMap &g, const int *slot, const int *blocks, int n) {
  g.n = 0;
  g.b0[0] = 0;
  for (int i = 0; i < n; i++) {
    if (blocks[i] <= 0) continue;
    g.slot[g.n] = slot[i];
    g.b0[g.n + 1] = g.b0[g.n] + blocks[i];
    g.n++;
  }
  return g.b0[g.n];
}

}  // namespace cfdl
