// epi_group.h -- lane groups inside a 64-wide wavefront.
//
// Several kernels give one cell to a group of G lanes of the same wave: the
// lanes do the per-neighbour work in parallel and exchange results through
// LDS or cross-lane shuffles, while the order in which results are combined
// stays the serial order of the reference.  Groups never span waves, so a
// wave-scope fence plus a compiler barrier orders LDS traffic inside a group.
#pragma once
#include <hip/hip_runtime.h>

namespace epi {

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// exclusive prefix sum of v over the G lanes of a group (lane = index in group)
template <int G>
__device__ __forceinline__ int group_exclusive_scan(int v, int lane, int* total) {
  int x = v;
#pragma unroll
  for (int o = 1; o < G; o <<= 1) {
    const int y = __shfl_up(x, o, G);
    if (lane >= o) x += y;
  }
  *total = __shfl(x, G - 1, G);
  return x - v;
}

}  // namespace epi
