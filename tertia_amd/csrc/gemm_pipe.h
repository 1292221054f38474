// gemm_pipe.h — pieces shared by the transform GEMM kernels (gemm.hip) and
// the fused factorized transform (fact_fused.hip): the software-pipelined K
// loop and the FP64 LDS swizzle.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace {

// unpadded k-major FP64 LDS tiles with an XOR swizzle of the m (n) index by
// the k row: the 16-lane groups of an MFMA operand read (row k, k+1 in the
// other 32 banks) and the 8-lane groups that store one row of a global slab
// hit distinct banks.  Needs tile widths >= 32.
struct SwzF64 {
  __device__ static int xa(int k) { return ((k & 1) << 4) ^ (((k >> 1) & 7) << 2); }
  __device__ static int xb(int k) { return (k & 1) << 4; }
};

// The K loop shared by the kernels: slab s is loaded into register set s%2
// two slabs ahead, stored into LDS buffer s%2 one slab ahead and multiplied
// from there.  GUARD = false needs K % (2*BK) == 0: no conditional loads (a
// look-ahead past the end re-reads the last slab, which is never multiplied),
// so the waits before each LDS store only cover the older register slab.
template <bool GUARD, int BK, class L0, class L1, class S0, class S1, class C>
__device__ inline void k_loop(int K, L0 load0, L1 load1, S0 store0, S1 store1, C compute) {
  const int KT = (K + BK - 1) / BK;
  if constexpr (!GUARD) {
    const int last = (KT - 1) * BK;
    load0(0);
    load1(BK);
    store0(0);
    load0(min(2 * BK, last));
    __syncthreads();
    for (int kt = 0; kt < KT; kt += 2) {
      compute(0);
      store1(1);
      load1(min((kt + 3) * BK, last));
      __syncthreads();
      compute(1);
      store0(0);
      load0(min((kt + 4) * BK, last));
      __syncthreads();
    }
  } else {
    load0(0);
    if (KT > 1) load1(BK);
    store0(0);
    if (KT > 2) load0(2 * BK);
    __syncthreads();
    for (int kt = 0; kt < KT; kt += 2) {
      compute(0);
      if (kt + 1 < KT) {
        store1(1);
        if (kt + 3 < KT) load1((kt + 3) * BK);
      }
      __syncthreads();
      if (kt + 1 < KT) {
        compute(1);
        if (kt + 2 < KT) {
          store0(0);
          if (kt + 4 < KT) load0((kt + 4) * BK);
        }
        __syncthreads();
      }
    }
  }
}

template <int V>
using SetI = std::integral_constant<int, V>;

// The K loop with PF register sets (PF = 2 or 4): slab s is loaded into
// register set s % PF PF slabs ahead, stored into LDS buffer s % 2 one slab
// ahead and multiplied from there.  load(SetI<r>, k0), store(SetI<r>, buf)
// and compute(buf).  Needs K % (PF * BK) == 0; look-aheads past the end
// re-read the last slab, which is never multiplied.  One barrier per slab.
template <int PF, int BK, class L, class S, class C>
__device__ inline void k_loop_pf(int K, L load, S store, C compute) {
  static_assert(PF == 2 || PF == 4, "prefetch depth");
  const int KT = K / BK, last = (KT - 1) * BK;
  load(SetI<0>{}, 0);
  load(SetI<1>{}, min(BK, last));
  if constexpr (PF == 4) {
    load(SetI<2>{}, min(2 * BK, last));
    load(SetI<3>{}, min(3 * BK, last));
  }
  store(SetI<0>{}, 0);
  load(SetI<0>{}, min(PF * BK, last));
  __syncthreads();
  for (int kt = 0; kt < KT; kt += PF) {
    compute(0);
    store(SetI<1>{}, 1);
    load(SetI<1>{}, min((kt + PF + 1) * BK, last));
    __syncthreads();
    compute(1);
    if constexpr (PF == 2) {
      store(SetI<0>{}, 0);
      load(SetI<0>{}, min((kt + 4) * BK, last));
      __syncthreads();
    } else {
      store(SetI<2>{}, 0);
      load(SetI<2>{}, min((kt + 6) * BK, last));
      __syncthreads();
      compute(0);
      store(SetI<3>{}, 1);
      load(SetI<3>{}, min((kt + 7) * BK, last));
      __syncthreads();
      compute(1);
      store(SetI<0>{}, 0);
      load(SetI<0>{}, min((kt + 8) * BK, last));
      __syncthreads();
    }
  }
}

}  // namespace
