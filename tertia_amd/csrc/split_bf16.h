// split_bf16.h — the operand form of the split-bf16 transform GEMM: a double
// x ~ p[0] + ... + p[NPC-1], each piece the round-to-nearest bf16 of the
// remainder of fp32(x) (which fp32 holds exactly).  A plane of n values is
// stored as NPC consecutive planes of n bf16 bit patterns.  Used by the GEMM's
// own split kernel and by the kernels that produce transform inputs, so a
// sweep need not re-read those inputs just to split them.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

namespace qss_split {

template <int NPC>
__device__ inline void split(double x, uint16_t *p) {
  float r = (float)x;
  // keep the fp32 rounding of x a separate step: otherwise the compiler may
  // fold (bf16)(float)x into one direct rounding of x, which differs from the
  // two-step form when fp32(x) falls exactly halfway between two bf16 values,
  // and this split would then depend on the kernel it is inlined into
  asm volatile("" : "+v"(r));
#pragma unroll
  for (int q = 0; q < NPC; ++q) {
    const __bf16 h = (__bf16)r;
    p[q] = __builtin_bit_cast(uint16_t, h);
    r -= (float)h;
  }
}

// element i of a plane of n values, npc = 2 or 3 pieces
__device__ inline void store(uint16_t *out, long n, long i, double x, int npc) {
  uint16_t p[3];
  if (npc == 3) {
    split<3>(x, p);
    out[i] = p[0];
    out[n + i] = p[1];
    out[2 * n + i] = p[2];
  } else {
    split<2>(x, p);
    out[i] = p[0];
    out[n + i] = p[1];
  }
}

}  // namespace qss_split
