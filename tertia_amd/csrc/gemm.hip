// gemm.hip — batched FP64 MFMA GEMM for the half-integer transforms.
//
// A 2-D DCT-IV/DST-IV of an Nz x Ny plane is X_hat = T_kz . X . T_ky with the
// symmetric matrices T^c_{jk} = 2cos(pi(j+1/2)(k+1/2)/N), T^s = 2sin(...)
// (FFTW REDFT11/RODFT11 as used at movefieldshydrolinlayersplit.cpp:389-427;
// the dense-table form is the reference's `HalfInteger` +
// `fft1D_X/fft1D_Y`, CUDA_WRAP/fft_matrix.cu:10-24, surfaceFFT.cu:108-165).
// One call of this kernel runs one pass (X.T_y or T_z.Y) for up to 8 fields.
//
// Tiling for CDNA4: 256 threads = 4 wave64s in a 2x2 grid, block tile
// 64x64, each wave owns 32x32 = 2x2 tiles of v_mfma_f64_16x16x4_f64.  K is
// staged through LDS in slabs of 16 with one register-prefetched slab in
// flight (double-buffered LDS, one barrier per slab).
//
// Epilogues fused into the store: alpha scaling (the inverse's 1/(4NyNz),
// normalization.cu:241), the corrector residual max|E_new-E_old|, max|E_new|
// (movefieldshydrolinlayersplit.cpp:1584-1621) and the interleaved push field
// (P+C)/2 that the particle kernel gathers.

#include "qss_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int BM = 64, BN = 64, BK = 16;
constexpr int NT = 256;
constexpr int PAD = 2;

__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
  // non-negative IEEE doubles order like their bit patterns
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__global__ __launch_bounds__(NT) void k_gemm_f64(GemmArgs a) {
  const GemmBatch &B = a.b[blockIdx.z];
  const int M = a.M, N = a.N, K = a.K;
  const int tiles_n = (N + BN - 1) / BN;
  const int tm = blockIdx.x / tiles_n, tn = blockIdx.x - tm * tiles_n;
  const int m0 = tm * BM, n0 = tn * BN;

  __shared__ double As[2][BK][BM + PAD];
  __shared__ double Bs[2][BK][BN + PAD];

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int li = lane & 15, lk = lane >> 4;

  // global -> register mapping
  const int ar = t >> 2, ac = (t & 3) * 4;    // A slab: 64 rows x 16 cols
  const int br = t >> 4, bc = (t & 15) * 4;   // B slab: 16 rows x 64 cols
  const double *Ap = B.A;
  const double *Bp = B.B;

  double ra[4], rb[4];
  auto load = [&](int k0) {
    const int gr = m0 + ar;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gc = k0 + ac + q;
      ra[q] = (gr < M && gc < K) ? Ap[(long)gr * K + gc] : 0.0;
    }
    const int gk = k0 + br;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int gn = n0 + bc + q;
      rb[q] = (gk < K && gn < N) ? Bp[(long)gk * N + gn] : 0.0;
    }
  };

  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int KT = (K + BK - 1) / BK;
  load(0);
  for (int kt = 0; kt < KT; ++kt) {
    const int buf = kt & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) As[buf][ac + q][ar] = ra[q];
#pragma unroll
    for (int q = 0; q < 4; ++q) Bs[buf][br][bc + q] = rb[q];
    __syncthreads();
    if (kt + 1 < KT) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK / 4; ++kk) {
      const int kr = kk * 4 + lk;
      double av[2], bv[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) av[i] = As[buf][kr][wm * 32 + i * 16 + li];
#pragma unroll
      for (int j = 0; j < 2; ++j) bv[j] = Bs[buf][kr][wn * 32 + j * 16 + li];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(av[i], bv[j], acc[i][j], 0, 0, 0);
    }
  }

  // epilogue: v_mfma_f64_16x16x4f64 leaves D[lane/16 + 4r][lane%16] in
  // register r of each lane (measured on gfx950: rows interleave by 4)
  double dmax = 0.0, vmax = 0.0;
  const double alpha = a.alpha;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = m0 + wm * 32 + i * 16 + lk + 4 * r;
        const int col = n0 + wn * 32 + j * 16 + li;
        if (row < M && col < N) {
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

}  // namespace

hipError_t launch_gemm_f64(const GemmArgs &a, hipStream_t s) {
  if (a.nbatch < 1 || a.nbatch > kGemmMaxBatch || a.M < 1 || a.N < 1 || a.K < 1)
    return hipErrorInvalidValue;
  const int tiles = ((a.M + BM - 1) / BM) * ((a.N + BN - 1) / BN);
  dim3 grid(tiles, 1, a.nbatch);
  hipLaunchKernelGGL(k_gemm_f64, grid, dim3(NT), 0, s, a);
  return hipGetLastError();
}
