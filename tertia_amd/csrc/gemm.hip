// gemm.hip — batched FP64 MFMA GEMM for the half-integer transforms.
//
// A 2-D DCT-IV/DST-IV of an Nz x Ny plane is X_hat = T_kz . X . T_ky with the
// symmetric matrices T^c_{jk} = 2cos(pi(j+1/2)(k+1/2)/N), T^s = 2sin(...)
// (FFTW REDFT11/RODFT11 as used at movefieldshydrolinlayersplit.cpp:389-427;
// the dense-table form is the reference's `HalfInteger` +
// `fft1D_X/fft1D_Y`, CUDA_WRAP/fft_matrix.cu:10-24, surfaceFFT.cu:108-165).
// One launch runs one pass (X.T_y or T_z.Y) for up to 8 fields.
//
// Tiling for CDNA4: the block tile BM x BN is split over WGM x WGN wave64s,
// each owning (BM/WGM) x (BN/WGN) as a grid of v_mfma_f64_16x16x4_f64 tiles.
// K is staged through LDS in slabs of BK (A stored k-major so that the 16
// lanes of an MFMA operand read consecutive doubles): two slabs sit in
// registers in flight while a third is multiplied out of a double-buffered
// LDS, so the global latency is covered by two slabs of MFMA work; one
// barrier per slab.  Work-groups are numbered
// so that each XCD receives a contiguous run of tiles (the field rows it
// reads stay in that XCD's L2).  Several tile shapes are compiled; the
// autotuner of qss_api.cpp picks one per launch shape.
//
// Epilogues fused into the store: alpha scaling (the inverse's 1/(4NyNz),
// normalization.cu:241), the corrector residual max|E_new-E_old|, max|E_new|
// (movefieldshydrolinlayersplit.cpp:1584-1621) and the interleaved push field
// (P+C)/2 that the particle kernel gathers.

#include "qss_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kXcd = 8;

__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
  // non-negative IEEE doubles order like their bit patterns
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

template <int BM, int BN, int BK, int WGM, int WGN>
struct Cfg {
  static constexpr int NW = WGM * WGN;
  static constexpr int NT = NW * 64;
  static constexpr int WTM = BM / WGM, WTN = BN / WGN;   // wave tile
  static constexpr int TI = WTM / 16, TJ = WTN / 16;      // MFMA tiles per wave
  static constexpr int PAD = 2;
  static constexpr int A_PER_T = BM * BK / NT;            // doubles per thread
  static constexpr int B_PER_T = BK * BN / NT;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile");
  static_assert(BK % 4 == 0 && A_PER_T % 2 == 0 && B_PER_T % 2 == 0, "slab");
};

// GUARD = false: M, N, K are multiples of BM, BN, BK (every plane of the
// sweep is), loads are unconditional and branch-free so that the compiler
// can wait for the older register slab only (s_waitcnt vmcnt(n) instead of
// vmcnt(0)).  GUARD = true handles ragged sizes element by element.
template <int BM, int BN, int BK, int WGM, int WGN, bool GUARD>
__global__ __launch_bounds__(WGM * WGN * 64) void k_gemm_f64(GemmArgs a) {
  using C = Cfg<BM, BN, BK, WGM, WGN>;
  const int M = a.M, N = a.N, K = a.K;
  const int tiles_m = (M + BM - 1) / BM, tiles_n = (N + BN - 1) / BN;
  const int per_field = tiles_m * tiles_n;
  // XCD-aware numbering: hardware sends block b to XCD b % 8; give each XCD
  // a contiguous range of (field, tile_m, tile_n) work items
  const int nblk = gridDim.x;
  int bid = blockIdx.x;
  if (nblk % kXcd == 0) bid = (bid % kXcd) * (nblk / kXcd) + bid / kXcd;
  const int f = bid / per_field;
  const int tt = bid - f * per_field;
  const int tm = tt / tiles_n, tn = tt - tm * tiles_n;
  const GemmBatch &B = a.b[f];
  const int m0 = tm * BM, n0 = tn * BN;

  __shared__ double As[2][BK][BM + C::PAD];
  __shared__ double Bs[2][BK][BN + C::PAD];

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w / WGN, wn = w - wm * WGN;
  const int li = lane & 15, lk = lane >> 4;
  const double *Ap = B.A;
  const double *Bp = B.B;

  // two register slabs in flight + two LDS slabs: a global load has the
  // compute of about two slabs to land before it is stored to LDS
  double ra[2][C::A_PER_T], rb[2][C::B_PER_T];
  auto load = [&](int k0, double *ra_, double *rb_) {
#pragma unroll
    for (int q = 0; q < C::A_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BK, cc = e - r * BK;
      const long off = (long)(m0 + r) * K + k0 + cc;
      if constexpr (!GUARD) {
        const double2 v = *reinterpret_cast<const double2 *>(Ap + off);
        ra_[2 * q] = v.x;
        ra_[2 * q + 1] = v.y;
      } else {
        const bool rok = m0 + r < M;
        ra_[2 * q] = (rok && k0 + cc < K) ? Ap[off] : 0.0;
        ra_[2 * q + 1] = (rok && k0 + cc + 1 < K) ? Ap[off + 1] : 0.0;
      }
    }
#pragma unroll
    for (int q = 0; q < C::B_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BN, cc = e - r * BN;
      const long off = (long)(k0 + r) * N + n0 + cc;
      if constexpr (!GUARD) {
        const double2 v = *reinterpret_cast<const double2 *>(Bp + off);
        rb_[2 * q] = v.x;
        rb_[2 * q + 1] = v.y;
      } else {
        const bool rok = k0 + r < K;
        rb_[2 * q] = (rok && n0 + cc < N) ? Bp[off] : 0.0;
        rb_[2 * q + 1] = (rok && n0 + cc + 1 < N) ? Bp[off + 1] : 0.0;
      }
    }
  };

  d4 acc[C::TI][C::TJ];
#pragma unroll
  for (int i = 0; i < C::TI; ++i)
#pragma unroll
    for (int j = 0; j < C::TJ; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  auto store = [&](int buf, const double *ra_, const double *rb_) {
#pragma unroll
    for (int q = 0; q < C::A_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BK, cc = e - r * BK;
      As[buf][cc][r] = ra_[2 * q];
      As[buf][cc + 1][r] = ra_[2 * q + 1];
    }
#pragma unroll
    for (int q = 0; q < C::B_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BN, cc = e - r * BN;
      Bs[buf][r][cc] = rb_[2 * q];
      Bs[buf][r][cc + 1] = rb_[2 * q + 1];
    }
  };
  auto compute = [&](int buf) {
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + lk;
      double av[C::TI], bv[C::TJ];
#pragma unroll
      for (int i = 0; i < C::TI; ++i) av[i] = As[buf][kr][wm * C::WTM + i * 16 + li];
#pragma unroll
      for (int j = 0; j < C::TJ; ++j) bv[j] = Bs[buf][kr][wn * C::WTN + j * 16 + li];
#pragma unroll
      for (int i = 0; i < C::TI; ++i)
#pragma unroll
        for (int j = 0; j < C::TJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  };

  // pipeline: slab s is loaded into register set s%2 two slabs ahead, stored
  // into LDS buffer s%2 one slab ahead, and multiplied from there
  const int KT = (K + BK - 1) / BK;
  if constexpr (!GUARD) {
    // K % (2*BK) == 0: no conditional loads (a look-ahead past the end
    // re-reads the last slab, which is never multiplied), so the waits
    // before each LDS store only cover the older register slab
    const int last = (KT - 1) * BK;
    load(0, ra[0], rb[0]);
    load(BK, ra[1], rb[1]);
    store(0, ra[0], rb[0]);
    load(min(2 * BK, last), ra[0], rb[0]);
    __syncthreads();
    for (int kt = 0; kt < KT; kt += 2) {
      compute(0);
      store(1, ra[1], rb[1]);
      load(min((kt + 3) * BK, last), ra[1], rb[1]);
      __syncthreads();
      compute(1);
      store(0, ra[0], rb[0]);
      load(min((kt + 4) * BK, last), ra[0], rb[0]);
      __syncthreads();
    }
  } else {
  load(0, ra[0], rb[0]);
  if (KT > 1) load(BK, ra[1], rb[1]);
  store(0, ra[0], rb[0]);
  if (KT > 2) load(2 * BK, ra[0], rb[0]);
  __syncthreads();
  for (int kt = 0; kt < KT; kt += 2) {
    compute(0);
    if (kt + 1 < KT) {
      store(1, ra[1], rb[1]);
      if (kt + 3 < KT) load((kt + 3) * BK, ra[1], rb[1]);
    }
    __syncthreads();
    if (kt + 1 < KT) {
      compute(1);
      if (kt + 2 < KT) {
        store(0, ra[0], rb[0]);
        if (kt + 4 < KT) load((kt + 4) * BK, ra[0], rb[0]);
      }
      __syncthreads();
    }
  }
  }

  // epilogue: v_mfma_f64_16x16x4f64 leaves D[lane/16 + 4r][lane%16] in
  // register r of each lane (measured on gfx950: rows interleave by 4)
  double dmax = 0.0, vmax = 0.0;
  const double alpha = a.alpha;
#pragma unroll
  for (int i = 0; i < C::TI; ++i)
#pragma unroll
    for (int j = 0; j < C::TJ; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * C::WTM + i * 16 + lk + 4 * r;
        const int col = n0 + wn * C::WTN + j * 16 + li;
        if (!GUARD || (row < M && col < N)) {
          const long n = (long)row * N + col;
          const double v = alpha * acc[i][j][r];
          if (B.resid >= 0) {
            // NaN -> +inf so the non-finite guard sees it (fmax drops NaN)
            const double old = B.C[n];
            const double d = fabs(old - v), m = fabs(v);
            dmax = fmax(dmax, d == d ? d : INFINITY);
            vmax = fmax(vmax, m == m ? m : INFINITY);
          }
          B.C[n] = v;
          if (B.push >= 0) a.F[n * kPushStride + B.push] = 0.5 * (B.P[n] + v);
        }
      }
  if (B.resid >= 0) {
    dmax = wave_max(dmax);
    vmax = wave_max(vmax);
    if (lane == 0) {
      atomic_max_nonneg(&a.resid[B.resid], dmax);
      atomic_max_nonneg(&a.resid[3 + B.resid], vmax);
    }
  }
}

struct Variant {
  int bm, bn, bk, nt;
  void (*fast)(GemmArgs);
  void (*guarded)(GemmArgs);
  const char *name;
};

#define QSS_GEMM_VARIANT(BM, BN, BK, WGM, WGN)                                         \
  Variant{BM, BN, BK, Cfg<BM, BN, BK, WGM, WGN>::NT,                                   \
          k_gemm_f64<BM, BN, BK, WGM, WGN, false>, k_gemm_f64<BM, BN, BK, WGM, WGN, true>, \
          #BM "x" #BN "x" #BK "_w" #WGM "x" #WGN}

const Variant kVariants[] = {
    QSS_GEMM_VARIANT(64, 64, 16, 2, 2),    // 0: 4 waves of 32x32
    QSS_GEMM_VARIANT(64, 64, 32, 2, 2),    // 1: deeper slab
    QSS_GEMM_VARIANT(32, 32, 16, 1, 1),    // 2: 1 wave of 32x32, shallow slab
    QSS_GEMM_VARIANT(32, 64, 32, 1, 2),    // 3: 2 waves of 32x32
    QSS_GEMM_VARIANT(64, 32, 32, 2, 1),    // 4
    QSS_GEMM_VARIANT(128, 64, 16, 2, 2),   // 5: 4 waves of 64x32
    QSS_GEMM_VARIANT(64, 128, 16, 2, 2),   // 6: 4 waves of 32x64
    QSS_GEMM_VARIANT(128, 128, 16, 2, 2),  // 7: 4 waves of 64x64
    QSS_GEMM_VARIANT(32, 32, 32, 1, 1),    // 8: 1 wave of 32x32
    QSS_GEMM_VARIANT(64, 64, 16, 1, 1),    // 9: 1 wave of 64x64
};

}  // namespace

int gemm_num_variants() { return (int)(sizeof(kVariants) / sizeof(kVariants[0])); }
const char *gemm_variant_name(int v) {
  return (v >= 0 && v < gemm_num_variants()) ? kVariants[v].name : "";
}

hipError_t launch_gemm_f64(const GemmArgs &a, hipStream_t s, int variant) {
  if (a.nbatch < 1 || a.nbatch > kGemmMaxBatch || a.M < 1 || a.N < 1 || a.K < 1)
    return hipErrorInvalidValue;
  if (variant < 0 || variant >= gemm_num_variants()) variant = 0;
  const Variant &v = kVariants[variant];
  const int tiles = ((a.M + v.bm - 1) / v.bm) * ((a.N + v.bn - 1) / v.bn);
  const bool exact = a.M % v.bm == 0 && a.N % v.bn == 0 && a.K % (2 * v.bk) == 0;
  hipLaunchKernelGGL(exact ? v.fast : v.guarded, dim3(tiles * a.nbatch), dim3(v.nt), 0, s, a);
  return hipGetLastError();
}
