// tile_sort.h — bitonic sort of 256 int keys held one per thread of a
// 256-thread work-group: register shuffles for distances inside a wave64,
// LDS (sh, 256 ints) for the two cross-wave distances.  Used by the tiled
// push kernels to order deposit (and adjoint scatter) events by cell.
#pragma once

#include <hip/hip_runtime.h>

namespace qss_tile {

__device__ inline int sort256(int key, int *sh) {
  const int t = threadIdx.x;
  for (int k = 2; k <= 256; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      int other;
      if (j >= 64) {
        sh[t] = key;
        __syncthreads();
        other = sh[t ^ j];
        __syncthreads();
      } else {
        other = __shfl_xor(key, j, 64);
      }
      const bool up = (t & k) == 0;
      const bool lower = (t & j) == 0;
      // lower element keeps min if ascending block, max otherwise
      const int mn = min(key, other), mx = max(key, other);
      key = (lower == up) ? mn : mx;
    }
  }
  return key;
}

}  // namespace qss_tile
