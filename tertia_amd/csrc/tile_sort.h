// tile_sort.h — bitonic sort of 256 int keys held one per thread of a
// 256-thread work-group.  Every compare-exchange distance is a compile-time
// constant: distances 1, 2, 4 and 8 exchange through DPP lane moves (no LDS
// round trip), 16 through ds_swizzle, 32 through ds_bpermute, and the two
// cross-wave distances (64, 128) through LDS (sh, 256 ints).  Used by the
// tiled push kernels to order deposit (and adjoint scatter) events by cell.
#pragma once

#include <hip/hip_runtime.h>

namespace qss_tile {

// value held by lane (lane ^ J) of this wave64
template <int J>
__device__ inline int xor_lane(int v) {
  if constexpr (J == 1) {
    return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xF, 0xF, false);   // quad_perm [1,0,3,2]
  } else if constexpr (J == 2) {
    return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xF, 0xF, false);   // quad_perm [2,3,0,1]
  } else if constexpr (J == 4) {
    // row_ror:n makes lane i read lane (i - n) mod 16 of its row
    const int from_minus4 = __builtin_amdgcn_mov_dpp(v, 0x124, 0xF, 0xF, false);
    const int from_plus4 = __builtin_amdgcn_mov_dpp(v, 0x12C, 0xF, 0xF, false);
    return (threadIdx.x & 4) ? from_minus4 : from_plus4;
  } else if constexpr (J == 8) {
    return __builtin_amdgcn_mov_dpp(v, 0x128, 0xF, 0xF, false);  // row_ror:8
  } else if constexpr (J == 16) {
    // swizzle bit mode: and 0x1f, or 0, xor 0x10 (within 32 lanes)
    return __builtin_amdgcn_ds_swizzle(v, 0x401F);
  } else {
    return __shfl_xor(v, J, 64);
  }
}

template <int K, int J>
__device__ inline int bitonic_stage(int key, int *sh) {
  const int t = threadIdx.x;
  int other;
  if constexpr (J >= 64) {
    sh[t] = key;
    __syncthreads();
    other = sh[t ^ J];
    __syncthreads();
  } else {
    other = xor_lane<J>(key);
  }
  const bool up = (t & K) == 0;
  const bool lower = (t & J) == 0;
  // lower element keeps min if ascending block, max otherwise
  const int mn = min(key, other), mx = max(key, other);
  return (lower == up) ? mn : mx;
}

template <int K, int J>
__device__ inline int bitonic_merge(int key, int *sh) {
  key = bitonic_stage<K, J>(key, sh);
  if constexpr (J > 1) key = bitonic_merge<K, J / 2>(key, sh);
  return key;
}

template <int K>
__device__ inline int bitonic_sort(int key, int *sh) {
  if constexpr (K > 2) key = bitonic_sort<K / 2>(key, sh);
  return bitonic_merge<K, K / 2>(key, sh);
}

__device__ inline int sort256(int key, int *sh) { return bitonic_sort<256>(key, sh); }

}  // namespace qss_tile
