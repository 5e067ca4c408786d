// gj_carry.h -- where a rebuilt neighbour list finds an edge's gap-junction
// value from the previous build (gj_refresh, cell_connection.cpp:133-152).
//
// The reference keys gj by the neighbour's identity in a per-cell hash map.
// Here a cell's previous row (olen entries) holds the neighbours' uids; the
// new row's entry e, whose neighbour has uid u, takes the value at the slot
// of the previous row that holds u, or gj_init if none does.  Kept edges
// usually sit at or near their old position, so the scan starts at e modulo
// olen and wraps once around the previous row; it never reads past it.
//
// The search is a host-and-device function so that the same code runs in the
// row-assembly kernel and in a CPU test (uid_at checks its bounds there).
#pragma once
#include <hip/hip_runtime.h>

namespace epi {

// slot of the previous row holding uid u, or -1; uid_at(q) is the uid at slot
// q, called only for 0 <= q < olen
template <class UidAt>
__host__ __device__ inline int gj_carry_slot(int e, int olen, long long u, const UidAt& uid_at) {
  int q = olen > 0 ? e % olen : 0;
  for (int f = 0; f < olen; f++) {
    if (uid_at(q) == u) return q;
    if (++q == olen) q = 0;
  }
  return -1;
}

}  // namespace epi
