// gemm.hip — batched MFMA GEMMs for the half-integer transforms.
//
// A 2-D DCT-IV/DST-IV of an Nz x Ny plane is X_hat = T_kz . X . T_ky with the
// symmetric matrices T^c_{jk} = 2cos(pi(j+1/2)(k+1/2)/N), T^s = 2sin(...)
// (FFTW REDFT11/RODFT11 as used at movefieldshydrolinlayersplit.cpp:389-427;
// the dense-table form is the reference's `HalfInteger` +
// `fft1D_X/fft1D_Y`, CUDA_WRAP/fft_matrix.cu:10-24, surfaceFFT.cu:108-165).
// One launch runs one pass (X.T_y or T_z.Y) for up to 8 fields.  Operands and
// results are fp64 in HBM in every precision mode; what changes is the MFMA:
//   FP64     v_mfma_f64_16x16x4f64
//   MIXED32  v_mfma_f32_16x16x4f32 on operands rounded to fp32
//   BF16X3 / BF16X6  v_mfma_f32_16x16x32_bf16 on operands split into two or
//            three bf16 pieces (x = h + m + l); the products h.h go to one
//            fp32 accumulator, the cross terms (h.m, m.h; and h.l, l.h, m.m
//            for three pieces) to a second one, added in fp64 at the store.
//
// Tiling for CDNA4: the block tile BM x BN is split over WGM x WGN wave64s,
// each owning (BM/WGM) x (BN/WGN) as a grid of 16x16 MFMA tiles.  K is staged
// through LDS in slabs of BK: two slabs sit in registers in flight while a
// third is multiplied out of a double-buffered LDS, so the global latency is
// covered by two slabs of MFMA work; one barrier per slab.  Work-groups are
// numbered so that each XCD receives a contiguous run of tiles (the field
// rows it reads stay in that XCD's L2).  Several tile shapes are compiled per
// precision; the autotuner of qss_api.cpp picks one per launch shape.
//
// Epilogues fused into the store: alpha scaling (the inverse's 1/(4NyNz),
// normalization.cu:241), the corrector residual max|E_new-E_old|, max|E_new|
// (movefieldshydrolinlayersplit.cpp:1584-1621) and the push field planes
// (P+C)/2 that the particle kernel gathers (fp64, or fp32 in mixed modes).

#include <hip/hip_ext.h>

#include "gemm_pipe.h"
#include "qss_internal.h"
#include "split_bf16.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int kXcd = 8;

__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
  // non-negative IEEE doubles order like their bit patterns
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// work item of this block: wake w, field f, tile origin (m0, n0).  Hardware
// sends block b to XCD b % 8; each XCD gets a contiguous range of items.
struct Item {
  int w, f, m0, n0;
};
template <int BM, int BN>
__device__ inline Item block_item(const GemmArgs &a) {
  const int tiles_m = (a.M + BM - 1) / BM, tiles_n = (a.N + BN - 1) / BN;
  const int per_field = tiles_m * tiles_n;
  const int nblk = gridDim.x;
  int bid = blockIdx.x;
  if (nblk % kXcd == 0) bid = (bid % kXcd) * (nblk / kXcd) + bid / kXcd;
  Item it;
  const int fw = bid / per_field;
  it.w = fw / a.nbatch;
  it.f = fw - it.w * a.nbatch;
  const int tt = bid - fw * per_field;
  const int tm = tt / tiles_n;
  it.m0 = tm * BM;
  it.n0 = (tt - tm * tiles_n) * BN;
  return it;
}

// Store of one wave's TI x TJ MFMA tiles.  val(i, j, r) is the fp64 result of
// tile (i, j), register r; ROW(lk, r) the tile row that register r of a lane
// in row group lk = lane/16 holds (the column is lane % 16).  The epilogue's
// loads (the old C for the residual or the beta term, the P plane for the
// push field) of one MFMA tile are all issued before any of them is used,
// so their latencies overlap instead of adding up.
// EPI = false: a launch with no epilogue work (no beta, residual or push
// field) compiles to plain scaled stores, keeping its register count low.
// PRE: the caller loaded those operands already (pre_old(i, j, r),
// pre_p(i, j, r)), e.g. before its K loop so the loads overlap the MFMAs.
template <bool GUARD, int TI, int TJ, class ROW, class V, bool EPI = true, bool PRE = false,
          class PO = int, class PP = int>
__device__ inline void store_tiles(const GemmArgs &a, const GemmBatch &B, int w, int r0, int c0,
                                   int lane, V val, PO pre_old = 0, PP pre_p = 0) {
  const int M = a.M, N = a.N;
  const int li = lane & 15, lk = lane >> 4;
  if constexpr (!EPI) {
    double *const Cw = B.C + w * a.wC;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + i * 16 + ROW::row(lk, r);
          const int col = c0 + j * 16 + li;
          if (!GUARD || (row < M && col < N)) Cw[(long)row * N + col] = a.alpha * val(i, j, r);
        }
    return;
  }
  double dmax = 0.0, vmax = 0.0;
  const double alpha = a.alpha, beta = a.beta;
  double *const Cw = B.C + w * a.wC;
  const double *const Pw = B.P ? B.P + w * a.wP : nullptr;
  double *const Fw = a.F ? a.F + w * a.wF : nullptr;
  float *const Ffw = a.Ff ? a.Ff + w * a.wF : nullptr;
  unsigned long long *const Rw = a.resid + w * a.wR + (blockIdx.x % kResidSlots) * 8;
  const bool need_old = B.resid >= 0 || beta != 0.0;
  const bool need_p = B.push >= 0;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      double oldv[4], pv[4];
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        if constexpr (PRE) {
          oldv[r] = pre_old(i, j, r);
          pv[r] = pre_p(i, j, r);
        } else {
          const int row = r0 + i * 16 + ROW::row(lk, r);
          const int col = c0 + j * 16 + li;
          const bool ok = !GUARD || (row < M && col < N);
          const long n = (long)row * N + col;
          oldv[r] = (need_old && ok) ? Cw[n] : 0.0;
          pv[r] = (need_p && ok) ? Pw[n] : 0.0;
        }
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = r0 + i * 16 + ROW::row(lk, r);
        const int col = c0 + j * 16 + li;
        if (!GUARD || (row < M && col < N)) {
          const long n = (long)row * N + col;
          double v = alpha * val(i, j, r);
          if (beta != 0.0) v += beta * oldv[r];   // accumulate (adjoint sweep)
          if (B.resid >= 0) {
            // NaN -> +inf so the non-finite guard sees it (fmax drops NaN)
            const double d = fabs(oldv[r] - v), m = fabs(v);
            dmax = fmax(dmax, d == d ? d : INFINITY);
            vmax = fmax(vmax, m == m ? m : INFINITY);
          }
          Cw[n] = v;
          if (need_p) {
            const double q = 0.5 * (pv[r] + v);
            const long pn = (long)B.push * M * N + n;
            if (Ffw) Ffw[pn] = (float)q;
            else Fw[pn] = q;
          }
        }
      }
  }
  if (B.resid >= 0) {
    dmax = wave_max(dmax);
    vmax = wave_max(vmax);
    if (lane == 0) {
      atomic_max_nonneg(&Rw[B.resid], dmax);
      atomic_max_nonneg(&Rw[3 + B.resid], vmax);
    }
  }
}

// ---------------------------------------------------------------------------
// dense-operand kernel (one k per lane per MFMA): fp64 and fp32 MFMA
// ---------------------------------------------------------------------------
struct MathF64 {
  typedef double lds_t;
  typedef d4 acc_t;
  // unpadded k-major LDS tiles with an XOR swizzle of the m (n) index by the
  // k row: the 16-lane groups of an MFMA operand read (row k, k+1 in the
  // other 32 banks) and the 8-lane groups that store one row of A's global
  // slab (8 values of k) hit distinct banks.  Needs BM, BN >= 32.
  static constexpr int PAD = 0;
  __device__ static int xa(int k) { return SwzF64::xa(k); }
  __device__ static int xb(int k) { return SwzF64::xb(k); }
  __device__ static acc_t mma(double x, double y, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
  }
  // measured on gfx950: register r of lane l holds D[l/16 + 4r][l%16]
  __device__ static int row(int lk, int r) { return lk + 4 * r; }
};
struct MathF32 {
  typedef float lds_t;
  typedef f4 acc_t;
  static constexpr int PAD = 16;   // row pitch = 16 mod 64 banks
  __device__ static int xa(int) { return 0; }
  __device__ static int xb(int) { return 0; }
  __device__ static acc_t mma(float x, float y, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(x, y, c, 0, 0, 0);
  }
  // register r of lane l holds D[4(l/16) + r][l%16]
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};

// KS > 1: KS waves share each wave tile; wave ks multiplies the ks-th part
// of every K slab and the partial sums are added through LDS (in ks order)
// before the epilogue.  More waves per output tile keep the MFMA pipes busy
// when a launch has few tiles (one to four fields at 512^2).
template <int BM, int BN, int BK, int WGM, int WGN, int KS>
struct Cfg {
  static constexpr int NWT = WGM * WGN;   // wave tiles
  static constexpr int NW = NWT * KS;
  static constexpr int NT = NW * 64;
  static constexpr int WTM = BM / WGM, WTN = BN / WGN;   // wave tile
  static constexpr int TI = WTM / 16, TJ = WTN / 16;      // MFMA tiles per wave
  static constexpr int A_PER_T = BM * BK / NT;            // doubles per thread
  static constexpr int B_PER_T = BK * BN / NT;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0, "wave tile");
  static_assert(BK % 4 == 0 && A_PER_T % 2 == 0 && B_PER_T % 2 == 0, "slab");
  static_assert((BK / 4) % KS == 0, "K split");
  static constexpr int KK = BK / 4 / KS;   // MFMA k-steps per wave per slab
};

template <class MT, int BM, int BN, int BK, int WGM, int WGN, int KS, bool GUARD, bool EPI>
__global__ __launch_bounds__(WGM * WGN * KS * 64) void k_gemm(GemmArgs a) {
  using C = Cfg<BM, BN, BK, WGM, WGN, KS>;
  typedef typename MT::lds_t T;
  typedef typename MT::acc_t ACC;
  static_assert(BM >= 32 && BN >= 32, "the LDS swizzle permutes m, n within 32");
  const int M = a.M, N = a.N, K = a.K;
  const Item it = block_item<BM, BN>(a);
  const GemmBatch &B = a.b[it.f];
  const int m0 = it.m0, n0 = it.n0;

  // A stored k-major so that the 16 lanes of an MFMA operand read
  // consecutive elements.  The split-K partial sums reuse the slab storage
  // after the K loop (less LDS per work-group: more work-groups per CU).
  constexpr size_t kSlabA = sizeof(T) * 2 * BK * (BM + MT::PAD);
  constexpr size_t kSlabB = sizeof(T) * 2 * BK * (BN + MT::PAD);
  constexpr size_t kRed = KS > 1 ? sizeof(ACC) * (KS - 1) * C::NWT * C::TI * C::TJ * 64 : 0;
  constexpr size_t kSmem = kSlabA + kSlabB > kRed ? kSlabA + kSlabB : kRed;
  __shared__ __attribute__((aligned(32))) char smem[kSmem];
  T(*As)[BK][BM + MT::PAD] = reinterpret_cast<T(*)[BK][BM + MT::PAD]>(smem);
  T(*Bs)[BK][BN + MT::PAD] = reinterpret_cast<T(*)[BK][BN + MT::PAD]>(smem + kSlabA);

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int ks = w / C::NWT, wl = w - ks * C::NWT;
  const int wm = wl / WGN, wn = wl - wm * WGN;
  const int li = lane & 15, lk = lane >> 4;
  const double *Ap = B.A + it.w * a.wA;
  const double *Bp = B.B + it.w * a.wB;

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
  auto store = [&](int buf, const double *ra_, const double *rb_) {
#pragma unroll
    for (int q = 0; q < C::A_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BK, cc = e - r * BK;
      As[buf][cc][r ^ MT::xa(cc)] = (T)ra_[2 * q];
      As[buf][cc + 1][r ^ MT::xa(cc + 1)] = (T)ra_[2 * q + 1];
    }
#pragma unroll
    for (int q = 0; q < C::B_PER_T / 2; ++q) {
      const int e = (q * C::NT + t) * 2;
      const int r = e / BN, cc = e - r * BN;
      Bs[buf][r][cc ^ MT::xb(r)] = (T)rb_[2 * q];
      Bs[buf][r][(cc + 1) ^ MT::xb(r)] = (T)rb_[2 * q + 1];
    }
  };

  ACC acc[C::TI][C::TJ];
#pragma unroll
  for (int i = 0; i < C::TI; ++i)
#pragma unroll
    for (int j = 0; j < C::TJ; ++j) acc[i][j] = ACC{0, 0, 0, 0};

  auto compute = [&](int buf) {
#pragma unroll
    for (int kq = 0; kq < C::KK; ++kq) {
      const int kr = (ks * C::KK + kq) * 4 + lk;
      const int sa = MT::xa(kr), sb = MT::xb(kr);
      T av[C::TI], bv[C::TJ];
#pragma unroll
      for (int i = 0; i < C::TI; ++i) av[i] = As[buf][kr][(wm * C::WTM + i * 16 + li) ^ sa];
#pragma unroll
      for (int j = 0; j < C::TJ; ++j) bv[j] = Bs[buf][kr][(wn * C::WTN + j * 16 + li) ^ sb];
#pragma unroll
      for (int i = 0; i < C::TI; ++i)
#pragma unroll
        for (int j = 0; j < C::TJ; ++j) acc[i][j] = MT::mma(av[i], bv[j], acc[i][j]);
    }
  };

  // epilogue operands of this wave tile (old C, P plane), loaded now so that
  // their latency hides behind the K loop; only the waves that store need them
  double po[EPI ? C::TI : 1][EPI ? C::TJ : 1][4], pp[EPI ? C::TI : 1][EPI ? C::TJ : 1][4];
  if constexpr (EPI) {
    const int r0 = m0 + wm * C::WTM, c0 = n0 + wn * C::WTN;
    const bool need_old = B.resid >= 0 || a.beta != 0.0, need_p = B.push >= 0;
    const double *Cw = B.C + it.w * a.wC;
    const double *Pw = B.P ? B.P + it.w * a.wP : nullptr;
#pragma unroll
    for (int i = 0; i < C::TI; ++i)
#pragma unroll
      for (int j = 0; j < C::TJ; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = r0 + i * 16 + MT::row(lk, r), col = c0 + j * 16 + li;
          const bool ok = ks == 0 && (!GUARD || (row < M && col < N));
          const long n = (long)row * N + col;
          po[i][j][r] = (need_old && ok) ? Cw[n] : 0.0;
          pp[i][j][r] = (need_p && ok) ? Pw[n] : 0.0;
        }
  }

  k_loop<GUARD, BK>(
      K, [&](int k0) { load(k0, ra[0], rb[0]); }, [&](int k0) { load(k0, ra[1], rb[1]); },
      [&](int buf) { store(buf, ra[0], rb[0]); }, [&](int buf) { store(buf, ra[1], rb[1]); },
      compute);

  if constexpr (KS > 1) {
    // partial sums of waves ks = 1..KS-1 through LDS (the slab storage, once
    // every wave has finished reading its last slab), added in ks order
    ACC(*red)[C::NWT][C::TI * C::TJ][64] =
        reinterpret_cast<ACC(*)[C::NWT][C::TI * C::TJ][64]>(smem);
    __syncthreads();
    if (ks > 0) {
#pragma unroll
      for (int i = 0; i < C::TI; ++i)
#pragma unroll
        for (int j = 0; j < C::TJ; ++j) red[ks - 1][wl][i * C::TJ + j][lane] = acc[i][j];
    }
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int i = 0; i < C::TI; ++i)
#pragma unroll
        for (int j = 0; j < C::TJ; ++j) acc[i][j] += red[q][wl][i * C::TJ + j][lane];
  }

  auto val = [&](int i, int j, int r) { return (double)acc[i][j][r]; };
  auto pre_old = [&](int i, int j, int r) { return po[EPI ? i : 0][EPI ? j : 0][r]; };
  auto pre_p = [&](int i, int j, int r) { return pp[EPI ? i : 0][EPI ? j : 0][r]; };
  store_tiles<GUARD, C::TI, C::TJ, MT, decltype(val), EPI, true, decltype(pre_old),
              decltype(pre_p)>(a, B, it.w, m0 + wm * C::WTM, n0 + wn * C::WTN, lane, val,
                               pre_old, pre_p);
}


// ---------------------------------------------------------------------------
// split-bf16 kernel: v_mfma_f32_16x16x32_bf16.  Operands arrive pre-split
// (npc bf16 pieces per element, k-contiguous: A[M][K], Bt[N][K]); each lane
// supplies 8 consecutive k of one row of A and of one row of Bt, so global
// loads, LDS stores and LDS reads are all 16-byte vectors with no conversion
// in the K loop.
// ---------------------------------------------------------------------------
struct RowBF16 {
  // measured on gfx950: register r of lane l holds D[4(l/16) + r][l%16]
  __device__ static int row(int lk, int r) { return 4 * lk + r; }
};

template <int NPC>
__device__ inline void split_bf16(double x, uint16_t *p) { qss_split::split<NPC>(x, p); }

// KS > 1: KS waves share each wave tile, wave ks multiplying the ks-th part
// of every K slab; their fp64 partial sums are added through LDS in ks order
// before the store (launches of one to three fields otherwise leave most
// SIMDs without a wave)
template <int BM, int BN, int BK, int WGM, int WGN, int NPC, bool GUARD, bool TOUT,
          bool EPI = true, int KS = 1>
__global__ __launch_bounds__(WGM * WGN * KS * 64) void k_gemm_bf16(GemmArgs a) {
  constexpr int NWT = WGM * WGN;
  constexpr int NT = NWT * KS * 64;
  constexpr int KSTEP = BK / 32 / KS;          // MFMA k-steps per wave per slab
  static_assert(KSTEP * KS * 32 == BK, "K split");
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TI = WTM / 16, TJ = WTN / 16;
  constexpr int KP = BK + 8;                   // LDS row pitch (bf16): 16-byte pad
  constexpr int KG = BK / 8;                   // 8-element groups per row
  constexpr int GA = BM * KG / NT;             // groups per thread
  constexpr int GB = BN * KG / NT;
  static_assert(WTM % 16 == 0 && WTN % 16 == 0 && BK % 32 == 0, "tile");
  static_assert(GA >= 1 && GB >= 1 && GA * NT == BM * KG && GB * NT == BN * KG, "groups");
  const int M = a.M, N = a.N, K = a.K;
  const Item it = block_item<BM, BN>(a);
  const GemmBatch &B = a.b[it.f];
  const int m0 = it.m0, n0 = it.n0;
  const long pa = (long)M * K, pb = (long)N * K;   // piece strides

  constexpr size_t kSlabA = sizeof(uint16_t) * NPC * 2 * BM * KP;
  constexpr size_t kSlabB = sizeof(uint16_t) * NPC * 2 * BN * KP;
  constexpr size_t kRed = KS > 1 ? sizeof(d4) * (KS - 1) * NWT * TI * TJ * 64 : 0;
  constexpr size_t kSmem = kSlabA + kSlabB > kRed ? kSlabA + kSlabB : kRed;
  __shared__ __attribute__((aligned(16))) char smem[kSmem];
  uint16_t(*As)[2][BM][KP] = reinterpret_cast<uint16_t(*)[2][BM][KP]>(smem);
  uint16_t(*Bs)[2][BN][KP] = reinterpret_cast<uint16_t(*)[2][BN][KP]>(smem + kSlabA);

  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int ks = w / NWT, wl = w - ks * NWT;
  const int wm = wl / WGN, wn = wl - wm * WGN;
  const int li = lane & 15, lk = lane >> 4;

  // group g of an operand tile: row g / KG, k offset 8 (g % KG); consecutive
  // lanes read consecutive 16-byte chunks of a row
  auto load_op = [&](const uint16_t *base, long ps, int R0, int RMAX, int k0, int g,
                     u4 *dst) {
    const int r = g / KG, kc = (g - r * KG) * 8;
    const long off = (long)(R0 + r) * K + k0 + kc;
#pragma unroll
    for (int p = 0; p < NPC; ++p) {
      if constexpr (!GUARD) {
        dst[p] = *reinterpret_cast<const u4 *>(base + p * ps + off);
      } else {
        uint16_t e[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
          e[q] = (R0 + r < RMAX && k0 + kc + q < K) ? base[p * ps + off + q] : (uint16_t)0;
        dst[p] = u4{e[0] | (uint32_t)e[1] << 16, e[2] | (uint32_t)e[3] << 16,
                    e[4] | (uint32_t)e[5] << 16, e[6] | (uint32_t)e[7] << 16};
      }
    }
  };
  u4 ra0[GA][NPC], ra1[GA][NPC], rb0[GB][NPC], rb1[GB][NPC];
  auto load = [&](int k0, u4 (&ra_)[GA][NPC], u4 (&rb_)[GB][NPC]) {
#pragma unroll
    for (int q = 0; q < GA; ++q) load_op(B.As, pa, m0, M, k0, q * NT + t, ra_[q]);
#pragma unroll
    for (int q = 0; q < GB; ++q) load_op(B.Bts, pb, n0, N, k0, q * NT + t, rb_[q]);
  };
  auto store = [&](int buf, const u4 (&ra_)[GA][NPC], const u4 (&rb_)[GB][NPC]) {
#pragma unroll
    for (int q = 0; q < GA; ++q) {
      const int g = q * NT + t, r = g / KG, kc = (g - r * KG) * 8;
#pragma unroll
      for (int p = 0; p < NPC; ++p) *reinterpret_cast<u4 *>(&As[p][buf][r][kc]) = ra_[q][p];
    }
#pragma unroll
    for (int q = 0; q < GB; ++q) {
      const int g = q * NT + t, r = g / KG, kc = (g - r * KG) * 8;
#pragma unroll
      for (int p = 0; p < NPC; ++p) *reinterpret_cast<u4 *>(&Bs[p][buf][r][kc]) = rb_[q][p];
    }
  };

  // hi: the products of the leading pieces, summed in fp32 over one K slab
  // only and then added into fp64 (hid), so the fp32 rounding of the largest
  // terms does not grow with K; lo: the cross terms (2^-8 smaller and
  // below), summed in fp32 over all of K
  f4 hi[TI][TJ], lo[TI][TJ];
  d4 hid[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) {
      hi[i][j] = f4{0, 0, 0, 0};
      lo[i][j] = f4{0, 0, 0, 0};
      hid[i][j] = d4{0, 0, 0, 0};
    }
  auto flush_hi = [&]() {
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) hid[i][j][r] += (double)hi[i][j][r];
        hi[i][j] = f4{0, 0, 0, 0};
      }
  };

  auto compute = [&](int buf) {
    // the previous slab's hi sums (their MFMAs finished behind the barrier)
    flush_hi();
#pragma unroll
    for (int kq = 0; kq < KSTEP; ++kq) {
      const int kc = (ks * KSTEP + kq) * 32 + 8 * lk;
      bf8 av[NPC][TI], bv[NPC][TJ];
#pragma unroll
      for (int p = 0; p < NPC; ++p) {
#pragma unroll
        for (int i = 0; i < TI; ++i)
          av[p][i] = *reinterpret_cast<const bf8 *>(&As[p][buf][wm * WTM + i * 16 + li][kc]);
#pragma unroll
        for (int j = 0; j < TJ; ++j)
          bv[p][j] = *reinterpret_cast<const bf8 *>(&Bs[p][buf][wn * WTN + j * 16 + li][kc]);
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) {
          // smaller terms first into the cross-term accumulator
          if constexpr (NPC == 3) {
            lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[1][j], lo[i][j], 0, 0, 0);
            lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[2][i], bv[0][j], lo[i][j], 0, 0, 0);
            lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[2][j], lo[i][j], 0, 0, 0);
          }
          lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[1][i], bv[0][j], lo[i][j], 0, 0, 0);
          lo[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[1][j], lo[i][j], 0, 0, 0);
          hi[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[0][i], bv[0][j], hi[i][j], 0, 0, 0);
        }
    }
  };

  k_loop<GUARD, BK>(
      K, [&](int k0) { load(k0, ra0, rb0); }, [&](int k0) { load(k0, ra1, rb1); },
      [&](int buf) { store(buf, ra0, rb0); }, [&](int buf) { store(buf, ra1, rb1); },
      compute);
  flush_hi();

  if constexpr (KS > 1) {
    // each wave's fp64 sum hid + lo; waves ks > 0 hand theirs to wave 0 of
    // their wave tile through LDS (the slab storage, once every wave has
    // finished its last slab), added in ks order
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
#pragma unroll
        for (int r = 0; r < 4; ++r) hid[i][j][r] += (double)lo[i][j][r];
        lo[i][j] = f4{0, 0, 0, 0};
      }
    d4(*red)[NWT][TI * TJ][64] = reinterpret_cast<d4(*)[NWT][TI * TJ][64]>(smem);
    __syncthreads();
    if (ks > 0) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) red[ks - 1][wl][i * TJ + j][lane] = hid[i][j];
    }
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) hid[i][j] += red[q][wl][i * TJ + j][lane];
  }

  if constexpr (!TOUT) {
    auto val = [&](int i, int j, int r) { return hid[i][j][r] + (double)lo[i][j][r]; };
    store_tiles<GUARD, TI, TJ, RowBF16, decltype(val), EPI>(a, B, 0, m0 + wm * WTM,
                                                            n0 + wn * WTN, lane, val);
  } else {
    // transposed split store: Ct[col][row], the 4 rows of a lane's registers
    // are consecutive -> one 8-byte store per piece
    const long pc = (long)N * M;
    const double alpha = a.alpha;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) {
        const int row0 = m0 + wm * WTM + i * 16 + 4 * lk;
        const int col = n0 + wn * WTN + j * 16 + li;
        uint16_t pcs[4][NPC];
#pragma unroll
        for (int r = 0; r < 4; ++r)
          split_bf16<NPC>(alpha * (hid[i][j][r] + (double)lo[i][j][r]), pcs[r]);
#pragma unroll
        for (int p = 0; p < NPC; ++p) {
          uint16_t *dst = B.Cts + p * pc + (long)col * M + row0;
          if constexpr (!GUARD) {
            *reinterpret_cast<uint2 *>(dst) =
                make_uint2(pcs[0][p] | (uint32_t)pcs[1][p] << 16,
                           pcs[2][p] | (uint32_t)pcs[3][p] << 16);
          } else if (col < N) {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (row0 + r < M) dst[r] = pcs[r][p];
          }
        }
      }
  }
}

// pieces of up to 8 planes: out[b][p*n + i] = piece p of in[b][i]
struct SplitArgs {
  const double *in[kGemmMaxBatch];
  uint16_t *out[kGemmMaxBatch];
  long n;
};
template <int NPC>
__global__ __launch_bounds__(256) void k_split(SplitArgs a) {
  const int b = blockIdx.y;
  const long i = ((long)blockIdx.x * 256 + threadIdx.x) * 4;
  if (i >= a.n) return;
  const double *src = a.in[b] + i;
  uint16_t *dst = a.out[b] + i;
  uint16_t pcs[4][NPC];
  if (i + 4 <= a.n && (a.n & 3) == 0) {
    const double4 v = *reinterpret_cast<const double4 *>(src);
    split_bf16<NPC>(v.x, pcs[0]);
    split_bf16<NPC>(v.y, pcs[1]);
    split_bf16<NPC>(v.z, pcs[2]);
    split_bf16<NPC>(v.w, pcs[3]);
#pragma unroll
    for (int p = 0; p < NPC; ++p)
      *reinterpret_cast<uint2 *>(dst + p * a.n) =
          make_uint2(pcs[0][p] | (uint32_t)pcs[1][p] << 16, pcs[2][p] | (uint32_t)pcs[3][p] << 16);
  } else {
    for (int q = 0; q < 4 && i + q < a.n; ++q) {
      split_bf16<NPC>(src[q], pcs[q]);
#pragma unroll
      for (int p = 0; p < NPC; ++p) dst[p * a.n + q] = pcs[q][p];
    }
  }
}

// ---------------------------------------------------------------------------
// variant tables
// ---------------------------------------------------------------------------
typedef void (*GemmFn)(GemmArgs);
struct Variant {
  int bm, bn, bk, nt;
  // [guarded][store]: store 0 = plain, 1 = with epilogue work (beta, residual,
  // push field), 2 = transposed split store (split-bf16 y pass)
  GemmFn fn[2][3];
  const char *name;
};

#define QSS_GEMM_DENSE(MT, BM, BN, BK, WGM, WGN, KS)                                   \
  Variant{BM, BN, BK, Cfg<BM, BN, BK, WGM, WGN, KS>::NT,                               \
          {{k_gemm<MT, BM, BN, BK, WGM, WGN, KS, false, false>,                        \
            k_gemm<MT, BM, BN, BK, WGM, WGN, KS, false, true>, nullptr},               \
           {k_gemm<MT, BM, BN, BK, WGM, WGN, KS, true, false>,                         \
            k_gemm<MT, BM, BN, BK, WGM, WGN, KS, true, true>, nullptr}},               \
          #BM "x" #BN "x" #BK "_w" #WGM "x" #WGN "_k" #KS}

#define QSS_GEMM_BF16(NPC, BM, BN, BK, WGM, WGN)                                       \
  Variant{BM, BN, BK, WGM * WGN * 64,                                                  \
          {{k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, false, false>,               \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, false, true>,                \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, true>},                      \
           {k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, false, false>,                \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, false, true>,                 \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, true>}},                      \
          #BM "x" #BN "x" #BK "_w" #WGM "x" #WGN}

const Variant kF64[] = {
    QSS_GEMM_DENSE(MathF64, 64, 64, 16, 2, 2, 1),    // 0: 4 waves of 32x32
    QSS_GEMM_DENSE(MathF64, 64, 64, 32, 2, 2, 1),    // 1: deeper slab
    QSS_GEMM_DENSE(MathF64, 32, 32, 16, 1, 1, 1),    // 2: 1 wave of 32x32, shallow slab
    QSS_GEMM_DENSE(MathF64, 32, 64, 32, 1, 2, 1),    // 3: 2 waves of 32x32
    QSS_GEMM_DENSE(MathF64, 64, 32, 32, 2, 1, 1),    // 4
    QSS_GEMM_DENSE(MathF64, 128, 64, 16, 2, 2, 1),   // 5: 4 waves of 64x32
    QSS_GEMM_DENSE(MathF64, 64, 128, 16, 2, 2, 1),   // 6: 4 waves of 32x64
    QSS_GEMM_DENSE(MathF64, 128, 128, 16, 2, 2, 1),  // 7: 4 waves of 64x64
    QSS_GEMM_DENSE(MathF64, 32, 32, 32, 1, 1, 1),    // 8: 1 wave of 32x32
    QSS_GEMM_DENSE(MathF64, 64, 64, 16, 1, 1, 1),    // 9: 1 wave of 64x64
    // K split inside the work-group: more waves per output tile
    QSS_GEMM_DENSE(MathF64, 32, 32, 16, 1, 1, 2),    // 10
    QSS_GEMM_DENSE(MathF64, 32, 32, 16, 1, 1, 4),    // 11
    QSS_GEMM_DENSE(MathF64, 32, 32, 32, 1, 1, 4),    // 12
    QSS_GEMM_DENSE(MathF64, 64, 64, 16, 2, 2, 2),    // 13
    QSS_GEMM_DENSE(MathF64, 32, 64, 32, 1, 2, 2),    // 14
    QSS_GEMM_DENSE(MathF64, 64, 32, 16, 2, 1, 2),    // 15
    QSS_GEMM_DENSE(MathF64, 64, 64, 16, 1, 1, 2),    // 16: 2 waves share one 64x64 tile
    QSS_GEMM_DENSE(MathF64, 32, 32, 32, 1, 1, 2),    // 17
    QSS_GEMM_DENSE(MathF64, 64, 32, 32, 2, 1, 4),    // 18: 8 waves
    QSS_GEMM_DENSE(MathF64, 32, 64, 32, 1, 2, 4),    // 19: 8 waves
    QSS_GEMM_DENSE(MathF64, 64, 64, 32, 2, 2, 2),    // 20: 8 waves, deep slab
    QSS_GEMM_DENSE(MathF64, 32, 32, 64, 1, 1, 4),    // 21: deep slab, 4 waves
};

const Variant kF32[] = {
    QSS_GEMM_DENSE(MathF32, 64, 64, 16, 2, 2, 1),
    QSS_GEMM_DENSE(MathF32, 64, 64, 32, 2, 2, 1),
    QSS_GEMM_DENSE(MathF32, 32, 32, 16, 1, 1, 1),
    QSS_GEMM_DENSE(MathF32, 128, 64, 16, 2, 2, 1),
    QSS_GEMM_DENSE(MathF32, 64, 128, 16, 2, 2, 1),
    QSS_GEMM_DENSE(MathF32, 128, 128, 16, 2, 2, 1),
    QSS_GEMM_DENSE(MathF32, 64, 64, 16, 1, 1, 1),
    QSS_GEMM_DENSE(MathF32, 32, 32, 16, 1, 1, 2),
    QSS_GEMM_DENSE(MathF32, 32, 32, 16, 1, 1, 4),
    QSS_GEMM_DENSE(MathF32, 64, 64, 16, 2, 2, 2),
    QSS_GEMM_DENSE(MathF32, 64, 64, 32, 1, 1, 2),
    QSS_GEMM_DENSE(MathF32, 32, 64, 32, 1, 2, 2),
};

#define QSS_GEMM_BF16K(NPC, BM, BN, BK, WGM, WGN, KS)                                  \
  Variant{BM, BN, BK, WGM * WGN * KS * 64,                                             \
          {{k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, false, false, KS>,           \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, false, true, KS>,            \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, false, true, true, KS>},            \
           {k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, false, false, KS>,            \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, false, true, KS>,             \
            k_gemm_bf16<BM, BN, BK, WGM, WGN, NPC, true, true, true, KS>}},            \
          #BM "x" #BN "x" #BK "_w" #WGM "x" #WGN "_k" #KS}

#define QSS_BF16_TABLE(NPC)                     \
  {                                             \
    QSS_GEMM_BF16(NPC, 64, 64, 32, 2, 2),       \
    QSS_GEMM_BF16(NPC, 64, 64, 64, 2, 2),       \
    QSS_GEMM_BF16(NPC, 32, 32, 32, 1, 1),       \
    QSS_GEMM_BF16(NPC, 32, 32, 64, 1, 1),       \
    QSS_GEMM_BF16(NPC, 128, 64, 32, 2, 2),      \
    QSS_GEMM_BF16(NPC, 64, 128, 32, 2, 2),      \
    QSS_GEMM_BF16(NPC, 128, 128, 32, 2, 2),     \
    QSS_GEMM_BF16(NPC, 32, 64, 32, 1, 2),       \
    QSS_GEMM_BF16(NPC, 64, 32, 32, 2, 1),       \
    QSS_GEMM_BF16K(NPC, 32, 32, 64, 1, 1, 2),   \
    QSS_GEMM_BF16K(NPC, 32, 32, 128, 1, 1, 4),  \
    QSS_GEMM_BF16K(NPC, 32, 64, 64, 1, 2, 2),   \
    QSS_GEMM_BF16K(NPC, 64, 32, 64, 2, 1, 2),   \
    QSS_GEMM_BF16K(NPC, 64, 64, 64, 2, 2, 2),   \
    QSS_GEMM_BF16K(NPC, 32, 32, 128, 1, 1, 2),  \
  }

const Variant kBF16x3[] = QSS_BF16_TABLE(2);
const Variant kBF16x6[] = QSS_BF16_TABLE(3);

struct Table {
  const Variant *v;
  int n;
};
Table table(int prec) {
  switch (prec) {
    case QSS_PREC_MIXED32: return {kF32, (int)(sizeof(kF32) / sizeof(kF32[0]))};
    case QSS_PREC_MIXED_BF16X3: return {kBF16x3, (int)(sizeof(kBF16x3) / sizeof(kBF16x3[0]))};
    case QSS_PREC_MIXED_BF16X6: return {kBF16x6, (int)(sizeof(kBF16x6) / sizeof(kBF16x6[0]))};
    default: return {kF64, (int)(sizeof(kF64) / sizeof(kF64[0]))};
  }
}

}  // namespace

int gemm_num_variants(int prec) { return table(prec).n; }
const char *gemm_variant_name(int prec, int v) {
  const Table tb = table(prec);
  return (v >= 0 && v < tb.n) ? tb.v[v].name : "";
}
int gemm_pieces(int prec) {
  return prec == QSS_PREC_MIXED_BF16X3 ? 2 : prec == QSS_PREC_MIXED_BF16X6 ? 3 : 0;
}

hipError_t launch_gemm(const GemmArgs &a, hipStream_t s, int prec, int variant, hipEvent_t ev0,
                       hipEvent_t ev1) {
  if (a.nbatch < 1 || a.nbatch > kGemmMaxBatch || a.M < 1 || a.N < 1 || a.K < 1)
    return hipErrorInvalidValue;
  const Table tb = table(prec);
  if (variant < 0 || variant >= tb.n) variant = 0;
  const Variant &v = tb.v[variant];
  const int tout = a.tout ? 1 : 0;
  if (tout && !gemm_pieces(prec)) return hipErrorInvalidValue;
  const int nw = a.nw > 1 ? a.nw : 1;
  if (nw > 1 && gemm_pieces(prec)) return hipErrorInvalidValue;   // split modes: one wake
  const int tiles = ((a.M + v.bm - 1) / v.bm) * ((a.N + v.bn - 1) / v.bn);
  const bool exact = a.M % v.bm == 0 && a.N % v.bn == 0 && a.K % (2 * v.bk) == 0 &&
                     (!tout || a.M % 4 == 0);
  const dim3 grid(tiles * a.nbatch * nw), block(v.nt);
  bool epi = a.beta != 0.0;
  for (int b = 0; b < a.nbatch; ++b) epi = epi || a.b[b].resid >= 0 || a.b[b].push >= 0;
  if (tout && epi) return hipErrorInvalidValue;   // the split store has no epilogue
  const GemmFn fn = v.fn[exact ? 0 : 1][tout ? 2 : (epi ? 1 : 0)];
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(fn, grid, block, 0, s, ev0, ev1, 0, a);
  else
    hipLaunchKernelGGL(fn, grid, block, 0, s, a);
  return hipGetLastError();
}

hipError_t launch_split_bf16(const double *const *in, uint16_t *const *out, int nb, long n,
                             int npc, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
  if (nb < 1 || nb > kGemmMaxBatch || n < 1) return hipErrorInvalidValue;
  SplitArgs a{};
  for (int b = 0; b < nb; ++b) {
    a.in[b] = in[b];
    a.out[b] = out[b];
  }
  a.n = n;
  const dim3 grid((unsigned)((n + 1023) / 1024), (unsigned)nb);
  if (npc != 2 && npc != 3) return hipErrorInvalidValue;
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(npc == 2 ? k_split<2> : k_split<3>, grid, dim3(256), 0, s, ev0, ev1,
                          0, a);
  else
    hipLaunchKernelGGL(npc == 2 ? k_split<2> : k_split<3>, grid, dim3(256), 0, s, a);
  return hipGetLastError();
}
