// fact_fused.hip — the factorized 2-D half-integer transform (fact.hip) with
// its twiddles folded into the two GEMMs: two launches per batch of planes
// and no twiddle passes over HBM.
//
// Per axis of length N (h = N/2, m = N/4) a DCT-IV is (fact.hip)
//   v_n = (x~_2n + i x~_{N-1-2n}) pre_n,   E_k / O_k = DFT_m of v_even / v_odd,
//   V_k = E_k + mid_k O_k,  V_{k+m} = E_k - mid_k O_k,  w_k = V_k post_k,
//   y_2k = 2 Re w_k,  y_{N-1-2k} = -2 Im w_k
// (x~ the reversed input and the odd outputs negated for a DST-IV).  The two
// m-point DFTs are one real GEMM with K = h against G = [[Fr, Fi], [-Fi, Fr]].
// The pre-twiddle is split as pre_{2q+p} = pre_{2q} e^{-i pi p/N}: pre_{2q}
// depends on the GEMM's K index q only and is folded into the table (entries
// H_qc = pre_{2q} F^{qc}), and e^{-i pi/N} of the odd parity into the
// butterfly twiddle (mid'_k = mid_k e^{-i pi/N}), so the operand loaders only
// gather the raw pairs (x~_2n, x~_{N-1-2n}).
//
// The orders of the GEMM's K, M and N indices are free (the tables are built
// for them on the host), and are chosen so that
//   * K: k' = 4u + j pairs q = u (j = 0, 1: Re, Im) with q'' = m-1-u (j = 2, 3).
//     The pairs of both come from the same two 32-byte chunks of the input,
//     x~[4u .. 4u+3] and x~[N-4u-4 .. N-4u-1]: the operand loader reads those
//     and places them in LDS, using every byte.
//   * rows (the parity p of v_{2q+p} in the y pass, Re/Im of the output
//     frequency in the z pass): row t of a 16-row MFMA tile holds index
//     (t&3) + 4(t>>3) and parity (t>>2)&1.  The FP64 MFMA leaves D[lk+4r][li]
//     in register r of lane 16 lk + li, so registers 2h and 2h+1 of a lane
//     hold the two parities of the same index: E and O (y), Re and Im (z).
//   * columns: y pass, 64-column groups of four 16-wide slots: (Re, Im) of
//     frequency k = 16g + li, then (Re, Im) of m-1-k; z pass, 32-column groups
//     of two slots, the parities p = 0, 1 of column y = 16g + li.
// So every butterfly's four inputs sit in one lane's accumulators, and the
// epilogue writes the transformed values directly: the y pass pairs k with
// m-1-k so that its outputs form four aligned 16-byte pairs per lane, the z
// pass stores rows of consecutive y (coalesced) through the dense path's
// epilogue (alpha, beta, corrector residual, push field).
//
// Two kernel families, autotuned per pass and batch size: staged (operand
// slabs through double-buffered LDS; the B tiles 2j and 2j+1 interleaved
// there, so one 16-byte read feeds both) and direct (each wave loads its own
// fragments from L2, with its own K order, further below).
//
// Shapes: both sizes multiples of 128 (at most 1024), FP64.  numpy
// restatement of this index algebra (tables, loaders, butterflies) agrees with
// scipy's type-4 transforms to 1e-15; the GPU tests compare with the oracle.

#include <hip/hip_ext.h>

#include <cmath>
#include <utility>

#include "gemm_pipe.h"
#include "qss_internal.h"

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int kXcdF = 8;
constexpr int kMaxN = 1024;   // largest axis the tile shapes are checked for

__device__ inline double2 cmulf(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline void atomic_max_nn(unsigned long long *addr, double v) {
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_maxf(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

__device__ inline d4 mma64(double x, double y, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(x, y, c, 0, 0, 0);
}

// block -> (wake, field, tile origin)
struct FItem {
  int w, f, m0, n0;
};
// block -> (wake, field, tile origin) with XCD-contiguous numbering: the
// work-groups of one XCD take consecutive tiles, which share an operand in
// that XCD's L2.  M_FAST: consecutive tiles run down a column of tiles (the z
// pass: its row tiles all read the same columns of the plane); otherwise along
// a row of tiles (the y pass: its column tiles all read the same rows).
template <int BM, int BN, bool M_FAST = false>
__device__ inline FItem fitem(int GM, int GN, int nbatch) {
  const int tiles_m = GM / BM, tiles_n = GN / BN;
  const int per_field = tiles_m * tiles_n;
  const int nblk = gridDim.x;
  int bid = blockIdx.x;
  if (nblk % kXcdF == 0) bid = (bid % kXcdF) * (nblk / kXcdF) + bid / kXcdF;
  FItem it;
  const int fw = bid / per_field;
  it.w = fw / nbatch;
  it.f = fw - it.w * nbatch;
  const int tt = bid - fw * per_field;
  if (M_FAST) {
    const int tn = tt / tiles_m;
    it.n0 = tn * BN;
    it.m0 = (tt - tn * tiles_m) * BM;
  } else {
    const int tm = tt / tiles_n;
    it.m0 = tm * BM;
    it.n0 = (tt - tm * tiles_n) * BN;
  }
  return it;
}

// butterflies along y and the stores of one wave's accumulator tiles: tile
// (i, j) holds virtual rows mr + 16 i .. and columns nc + 16 j ..; frequency
// k = 16 g + li and k2 = m-1-k of X rows (mr + 16 i)/2 + lk + 4 hh, one
// 64-column group g per four accumulator tiles
template <int TI, int TJ>
__device__ inline void y_epilogue(const FusedArgs &a, const d4 (&acc)[TI][TJ], int mr, int nc,
                                  int fb, int w, bool rev, int li, int lk) {
  const int N = a.ny, h = N / 2, m = N / 4;
  const double2 *mid = a.tw, *post = a.tw + m;   // mid' (m), post (h)
  const double sg = rev ? 2.0 : -2.0;
  double *Y = a.out[fb] + w * a.w_out;
#pragma unroll
  for (int gg = 0; gg < TJ / 4; ++gg) {
    const int g = nc / 64 + gg;
    const int k = 16 * g + li, k2 = m - 1 - k;
    const double2 md = mid[k], md2 = mid[k2];
    const double2 pk = post[k], pkm = post[k + m], pk2 = post[k2], pk2m = post[k2 + m];
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int hh = 0; hh < 2; ++hh) {
        const int xrow = (mr + 16 * i) / 2 + lk + 4 * hh;
        const d4 *A4 = &acc[i][4 * gg];
        const double2 E = make_double2(A4[0][2 * hh], A4[1][2 * hh]);
        const double2 O = make_double2(A4[0][2 * hh + 1], A4[1][2 * hh + 1]);
        const double2 E2 = make_double2(A4[2][2 * hh], A4[3][2 * hh]);
        const double2 O2 = make_double2(A4[2][2 * hh + 1], A4[3][2 * hh + 1]);
        const double2 wo = cmulf(md, O), wo2 = cmulf(md2, O2);
        const double2 wk = cmulf(make_double2(E.x + wo.x, E.y + wo.y), pk);
        const double2 wkm = cmulf(make_double2(E.x - wo.x, E.y - wo.y), pkm);
        const double2 wk2 = cmulf(make_double2(E2.x + wo2.x, E2.y + wo2.y), pk2);
        const double2 wk2m = cmulf(make_double2(E2.x - wo2.x, E2.y - wo2.y), pk2m);
        double *yr = Y + (long)xrow * N;
        *reinterpret_cast<double2 *>(yr + 2 * k) = make_double2(2.0 * wk.x, sg * wk2m.y);
        *reinterpret_cast<double2 *>(yr + h - 2 - 2 * k) = make_double2(2.0 * wk2.x, sg * wkm.y);
        *reinterpret_cast<double2 *>(yr + h + 2 * k) = make_double2(2.0 * wkm.x, sg * wk2.y);
        *reinterpret_cast<double2 *>(yr + N - 2 - 2 * k) = make_double2(2.0 * wk2m.x, sg * wk.y);
      }
  }
}

// butterflies along z and the epilogue (alpha, beta, corrector residual,
// push field) of one wave's accumulator tiles: tile (i, j) holds virtual rows
// mr + 16 i .. and columns ncol + 16 j ..; frequency kf of rows
// (mr + 16 i)/2 + lk + 4 hh, column y = ncol/2 + 16 g + li
template <int TI, int TJ>
__device__ inline void z_epilogue(const FusedArgs &a, const d4 (&acc)[TI][TJ], int mr, int ncol,
                                  int fb, int w, bool rev, int li, int lk, int lane) {
  const int Ny = a.ny, Nz = a.nz, h = Nz / 2, m = Nz / 4;
  const double2 *mid = a.tw, *post = a.tw + m;   // mid' (m), post (h)
  const double sg = rev ? 2.0 : -2.0;
  const long nc = (long)Ny * Nz;
  double *C = a.out[fb] + w * a.w_out;
  const double *P = a.P[fb] ? a.P[fb] + w * a.w_p : nullptr;
  const int resid = a.resid_c[fb], push = a.push[fb];
  const bool need_old = resid >= 0 || a.beta != 0.0;
  double dmax = 0.0, vmax = 0.0;
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int hh = 0; hh < 2; ++hh) {
      const int kf = (mr + 16 * i) / 2 + lk + 4 * hh;
      const double2 md = mid[kf], pk = post[kf], pkm = post[kf + m];
      const int zr[4] = {2 * kf, Nz - 1 - 2 * kf, h + 2 * kf, h - 1 - 2 * kf};
#pragma unroll
      for (int g2 = 0; g2 < TJ / 2; ++g2) {
        const int y = ncol / 2 + 16 * g2 + li;
        const double2 E = make_double2(acc[i][2 * g2][2 * hh], acc[i][2 * g2][2 * hh + 1]);
        const double2 O =
            make_double2(acc[i][2 * g2 + 1][2 * hh], acc[i][2 * g2 + 1][2 * hh + 1]);
        const double2 wo = cmulf(md, O);
        const double2 wk = cmulf(make_double2(E.x + wo.x, E.y + wo.y), pk);
        const double2 wkm = cmulf(make_double2(E.x - wo.x, E.y - wo.y), pkm);
        const double val[4] = {2.0 * wk.x, sg * wk.y, 2.0 * wkm.x, sg * wkm.y};
        double oldv[4], pv[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long n = (long)zr[q] * Ny + y;
          oldv[q] = need_old ? C[n] : 0.0;
          pv[q] = push >= 0 ? P[n] : 0.0;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const long n = (long)zr[q] * Ny + y;
          double v = a.alpha * val[q];
          if (a.beta != 0.0) v += a.beta * oldv[q];   // accumulate (adjoint sweep)
          if (resid >= 0) {
            const double d = fabs(oldv[q] - v), mm = fabs(v);
            dmax = fmax(dmax, d == d ? d : INFINITY);
            vmax = fmax(vmax, mm == mm ? mm : INFINITY);
          }
          C[n] = v;
          if (push >= 0) {
            const double pq = 0.5 * (pv[q] + v);
            const long pn = (long)push * nc + n;
            if (a.Ff) a.Ff[w * a.w_f + pn] = (float)pq;
            else a.F[w * a.w_f + pn] = pq;
          }
        }
      }
    }
  if (resid >= 0) {   // uniform over the work-group (one field per block)
    dmax = wave_maxf(dmax);
    vmax = wave_maxf(vmax);
    if (lane == 0) {
      unsigned long long *R = a.rslots + w * a.w_r + (blockIdx.x % kResidSlots) * 8;
      atomic_max_nn(&R[resid], dmax);
      atomic_max_nn(&R[3 + resid], vmax);
    }
  }
}


// ---------------------------------------------------------------------------
// y pass: virtual A (2 Nz x h) = the pairs of the rows of X, B = G'_y (h x h),
// output Y (Nz x Ny, natural layout) after the butterflies along y
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WGM, int WGN, int KS, int PF>
__global__ __launch_bounds__(WGM * WGN * KS * 64) void k_fact_y(FusedArgs a) {
  constexpr int NWT = WGM * WGN, NT = NWT * KS * 64;
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TI = WTM / 16, TJ = WTN / 16;   // TJ / 4 groups of 64 columns per wave
  constexpr int KK = BK / 4 / KS;   // MFMA k-steps per wave per slab
  static_assert(WTN % 64 == 0 && WTM % 16 == 0 && BM >= 32 && KK * KS * 4 == BK, "tile");
  // A units: one complex value v each, (row, u, p, w) with w (v_{2u+p} or
  // v_{2(m-1-u)+p}) fastest: lanes 4i .. 4i+3 read the same two 32-byte chunks
  // of a row between them; whole waves take part or skip
  constexpr int UNITS = BM * BK / 2;
  static_assert(UNITS % 64 == 0, "A units");
  constexpr int UPT = (UNITS + NT - 1) / NT;
  constexpr int BCH = BK * BN / 2;   // 16-byte chunks of the B slab
  static_assert(BCH % 64 == 0, "B chunks");
  constexpr int BPT = (BCH + NT - 1) / NT;
  const int N = a.ny, h = N / 2, m = N / 4;
  const int GM = 2 * a.nz, GN = h, K = h;
  const FItem it = fitem<BM, BN>(GM, GN, a.nbatch);
  const int m0 = it.m0, n0 = it.n0, fb = it.f;
  const bool rev = a.ky[fb] == QSS_KIND_S;

  constexpr size_t kSlab = sizeof(double) * 2 * BK * (BM + BN);
  constexpr size_t kRed = KS > 1 ? sizeof(d4) * (KS - 1) * NWT * TI * TJ * 64 : 0;
  __shared__ __attribute__((aligned(32))) char smem[kSlab > kRed ? kSlab : kRed];
  double(*As)[BK][BM] = reinterpret_cast<double(*)[BK][BM]>(smem);
  double(*Bs)[BK][BN] = reinterpret_cast<double(*)[BK][BN]>(smem + sizeof(double) * 2 * BK * BM);

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int ks = wv / NWT, wl = wv - ks * NWT;
  const int wm = wl / WGN, wn = wl - wm * WGN;
  const int li = lane & 15, lk = lane >> 4;

  const double *X = a.in[fb] + it.w * a.w_in;
  const double *G = a.G;
  const int xr0 = m0 / 2;   // first row of X of this block

  // unit U: w = U & 1, p = (U >> 1) & 1, u = k0/4 + (U >> 2) % (BK/4), row
  // xl = (U >> 2) / (BK/4).  With the chunks f = x~[4u .. 4u+3] and
  // b = x~[N-4u-4 .. N-4u-1] (x~ = x, or x reversed for the DST: x~[4u+i] =
  // x[N-1-4u-i], x~[N-4u-4+i] = x[4u+3-i]), the pair of
  //   w = 0: v_{2u+p}        is (f[2p], b[3-2p])
  //   w = 1: v_{2(m-1-u)+p}  is (b[2p], f[3-2p])
  // PF register sets of the slab loads
  double ra[PF][UPT][2], rb[PF][BPT][2];
  auto load = [&](auto S, int k0) {
    constexpr int rs = decltype(S)::value;
    double(*ra_)[2] = ra[rs];
    double(*rb_)[2] = rb[rs];
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int U = q * NT + t;
      if (UNITS % NT != 0 && U >= UNITS) continue;
      const int w = U & 1, pp = (U >> 1) & 1, ul = (U >> 2) % (BK / 4), xl = (U >> 2) / (BK / 4);
      const int u = k0 / 4 + ul;
      const double *row = X + (long)(xr0 + xl) * N;
      // x~ indices of the real and imaginary input, then to x
      const int c0 = w == 0 ? 4 * u + 2 * pp : N - 4 * u - 4 + 2 * pp;
      const int c1 = w == 0 ? N - 4 * u - 1 - 2 * pp : 4 * u + 3 - 2 * pp;
      ra_[q][0] = row[rev ? N - 1 - c0 : c0];
      ra_[q][1] = row[rev ? N - 1 - c1 : c1];
    }
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      if (BCH % NT != 0 && q * NT + t >= BCH) continue;
      const int e = (q * NT + t) * 2;
      const int r = e / BN, cc = e - r * BN;
      const double2 v = *reinterpret_cast<const double2 *>(G + (long)(k0 + r) * GN + n0 + cc);
      rb_[q][0] = v.x;
      rb_[q][1] = v.y;
    }
  };
  auto store = [&](auto S, int buf) {
    constexpr int rs = decltype(S)::value;
    const double(*ra_)[2] = ra[rs];
    const double(*rb_)[2] = rb[rs];
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int U = q * NT + t;
      if (UNITS % NT != 0 && U >= UNITS) continue;
      const int w = U & 1, pp = (U >> 1) & 1, ul = (U >> 2) % (BK / 4), xl = (U >> 2) / (BK / 4);
      const int x8 = xl & 7;
      const int rr = (xl >> 3) * 16 + (x8 & 3) + 8 * (x8 >> 2) + 4 * pp;
      const int kr = 4 * ul + 2 * w;
      As[buf][kr][rr ^ SwzF64::xa(kr)] = ra_[q][0];
      As[buf][kr + 1][rr ^ SwzF64::xa(kr + 1)] = ra_[q][1];
    }
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
      if (BCH % NT != 0 && q * NT + t >= BCH) continue;
      const int e = (q * NT + t) * 2;
      const int r = e / BN, cc = e - r * BN;
      Bs[buf][r][cc ^ SwzF64::xb(r)] = rb_[q][0];
      Bs[buf][r][(cc + 1) ^ SwzF64::xb(r)] = rb_[q][1];
    }
  };

  d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = d4{0, 0, 0, 0};
  auto compute = [&](int buf) {
#pragma unroll
    for (int kq = 0; kq < KK; ++kq) {
      const int kr = (ks * KK + kq) * 4 + lk;
      const int sa = SwzF64::xa(kr), sb = SwzF64::xb(kr);
      double av[TI], bv[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) av[i] = As[buf][kr][(wm * WTM + i * 16 + li) ^ sa];
#pragma unroll
      for (int jp = 0; jp < TJ / 2; ++jp) {   // tiles 2jp, 2jp+1 interleaved: one 16-byte read
        const double2 v =
            *reinterpret_cast<const double2 *>(&Bs[buf][kr][(wn * WTN + 32 * jp + 2 * li) ^ sb]);
        bv[2 * jp] = v.x;
        bv[2 * jp + 1] = v.y;
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = mma64(av[i], bv[j], acc[i][j]);
    }
  };
  k_loop_pf<PF, BK>(K, load, store, compute);
  if constexpr (KS > 1) {
    // partial sums of waves ks = 1..KS-1 through LDS, added in ks order
    d4(*red)[NWT][TI * TJ][64] = reinterpret_cast<d4(*)[NWT][TI * TJ][64]>(smem);
    __syncthreads();
    if (ks > 0) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) red[ks - 1][wl][i * TJ + j][lane] = acc[i][j];
    }
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] += red[q][wl][i * TJ + j][lane];
  }

  y_epilogue<TI, TJ>(a, acc, m0 + wm * WTM, n0 + wn * WTN, fb, it.w, rev, li, lk);
}

// ---------------------------------------------------------------------------
// z pass: A = G'_z (h_z x h_z), virtual B (h_z x 2 Ny) = the pairs of the columns
// of Y, output C (Nz x Ny) after the butterflies along z and the epilogue
// ---------------------------------------------------------------------------
template <int BM, int BN, int BK, int WGM, int WGN, int KS, int PF>
__global__ __launch_bounds__(WGM * WGN * KS * 64) void k_fact_z(FusedArgs a) {
  constexpr int NWT = WGM * WGN, NT = NWT * KS * 64;
  constexpr int WTM = BM / WGM, WTN = BN / WGN;
  constexpr int TI = WTM / 16, TJ = WTN / 16;
  constexpr int KK = BK / 4 / KS;
  static_assert(WTM % 16 == 0 && WTN % 32 == 0 && BM >= 32 && KK * KS * 4 == BK, "tile");
  // B units: one complex value each, (y, p, w, u) with y fastest, so a wave
  // reads rows of consecutive y
  constexpr int UNITS = BK * BN / 2;
  static_assert(UNITS % 64 == 0, "B units");
  constexpr int UPT = (UNITS + NT - 1) / NT;
  constexpr int ACH = BM * BK / 2;   // 16-byte chunks of the A slab
  static_assert(ACH % 64 == 0, "A chunks");
  constexpr int APT = (ACH + NT - 1) / NT;
  const int Ny = a.ny, Nz = a.nz, h = Nz / 2, m = Nz / 4;
  const int GM = h, GN = 2 * Ny, K = h;
  const FItem it = fitem<BM, BN, true>(GM, GN, a.nbatch);
  const int m0 = it.m0, n0 = it.n0, fb = it.f;
  const bool rev = a.kz[fb] == QSS_KIND_S;

  constexpr size_t kSlab = sizeof(double) * 2 * BK * (BM + BN);
  constexpr size_t kRed = KS > 1 ? sizeof(d4) * (KS - 1) * NWT * TI * TJ * 64 : 0;
  __shared__ __attribute__((aligned(32))) char smem[kSlab > kRed ? kSlab : kRed];
  double(*As)[BK][BM] = reinterpret_cast<double(*)[BK][BM]>(smem);
  double(*Bs)[BK][BN] = reinterpret_cast<double(*)[BK][BN]>(smem + sizeof(double) * 2 * BK * BM);

  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  const int ks = wv / NWT, wl = wv - ks * NWT;
  const int wm = wl / WGN, wn = wl - wm * WGN;
  const int li = lane & 15, lk = lane >> 4;

  const double *Yp = a.in[fb] + it.w * a.w_in;
  const double *G = a.G;
  const int y0 = n0 / 2;   // first column of Y of this block

  // unit U: yl = U % (BN/2), then p, w, u from U / (BN/2) as in the y pass;
  // front rows 4u .. 4u+3 and back rows Nz-4u-4 .. Nz-4u-1 of column y0 + yl
  // (the row order reversed for the DST)
  double ra[PF][APT][2], rb[PF][UPT][2];
  auto load = [&](auto S, int k0) {
    constexpr int rs = decltype(S)::value;
    double(*ra_)[2] = ra[rs];
    double(*rb_)[2] = rb[rs];
#pragma unroll
    for (int q = 0; q < APT; ++q) {
      if (ACH % NT != 0 && q * NT + t >= ACH) continue;
      const int e = (q * NT + t) * 2;
      const int r = e / BK, cc = e - r * BK;
      const double2 v = *reinterpret_cast<const double2 *>(G + (long)(m0 + r) * K + k0 + cc);
      ra_[q][0] = v.x;
      ra_[q][1] = v.y;
    }
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int U = q * NT + t;
      if (UNITS % NT != 0 && U >= UNITS) continue;
      const int yl = U % (BN / 2), r = U / (BN / 2);
      const int w = r & 1, pp = (r >> 1) & 1, ul = r >> 2;
      const int u = k0 / 4 + ul;
      const double *col = Yp + y0 + yl;
      const int c0 = w == 0 ? 4 * u + 2 * pp : Nz - 4 * u - 4 + 2 * pp;
      const int c1 = w == 0 ? Nz - 4 * u - 1 - 2 * pp : 4 * u + 3 - 2 * pp;
      rb_[q][0] = col[(long)(rev ? Nz - 1 - c0 : c0) * Ny];
      rb_[q][1] = col[(long)(rev ? Nz - 1 - c1 : c1) * Ny];
    }
  };
  auto store = [&](auto S, int buf) {
    constexpr int rs = decltype(S)::value;
    const double(*ra_)[2] = ra[rs];
    const double(*rb_)[2] = rb[rs];
#pragma unroll
    for (int q = 0; q < APT; ++q) {
      if (ACH % NT != 0 && q * NT + t >= ACH) continue;
      const int e = (q * NT + t) * 2;
      const int r = e / BK, cc = e - r * BK;
      As[buf][cc][r ^ SwzF64::xa(cc)] = ra_[q][0];
      As[buf][cc + 1][r ^ SwzF64::xa(cc + 1)] = ra_[q][1];
    }
#pragma unroll
    for (int q = 0; q < UPT; ++q) {
      const int U = q * NT + t;
      if (UNITS % NT != 0 && U >= UNITS) continue;
      const int yl = U % (BN / 2), r = U / (BN / 2);
      const int w = r & 1, pp = (r >> 1) & 1, ul = r >> 2;
      const int cc = 32 * (yl >> 4) + 2 * (yl & 15) + pp;   // the two parities interleaved
      const int kr = 4 * ul + 2 * w;
      Bs[buf][kr][cc ^ SwzF64::xb(kr)] = rb_[q][0];
      Bs[buf][kr + 1][cc ^ SwzF64::xb(kr + 1)] = rb_[q][1];
    }
  };

  d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = d4{0, 0, 0, 0};
  auto compute = [&](int buf) {
#pragma unroll
    for (int kq = 0; kq < KK; ++kq) {
      const int kr = (ks * KK + kq) * 4 + lk;
      const int sa = SwzF64::xa(kr), sb = SwzF64::xb(kr);
      double av[TI], bv[TJ];
#pragma unroll
      for (int i = 0; i < TI; ++i) av[i] = As[buf][kr][(wm * WTM + i * 16 + li) ^ sa];
#pragma unroll
      for (int jp = 0; jp < TJ / 2; ++jp) {   // tiles 2jp, 2jp+1 interleaved: one 16-byte read
        const double2 v =
            *reinterpret_cast<const double2 *>(&Bs[buf][kr][(wn * WTN + 32 * jp + 2 * li) ^ sb]);
        bv[2 * jp] = v.x;
        bv[2 * jp + 1] = v.y;
      }
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] = mma64(av[i], bv[j], acc[i][j]);
    }
  };
  k_loop_pf<PF, BK>(K, load, store, compute);
  if constexpr (KS > 1) {
    d4(*red)[NWT][TI * TJ][64] = reinterpret_cast<d4(*)[NWT][TI * TJ][64]>(smem);
    __syncthreads();
    if (ks > 0) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) red[ks - 1][wl][i * TJ + j][lane] = acc[i][j];
    }
    __syncthreads();
    if (ks > 0) return;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] += red[q][wl][i * TJ + j][lane];
  }

  z_epilogue<TI, TJ>(a, acc, m0 + wm * WTM, n0 + wn * WTN, fb, it.w, rev, li, lk, lane);
}

// ---------------------------------------------------------------------------
// direct variants: every wave loads its own MFMA operand fragments from global
// memory (L2) into a ring of D register sets, with no LDS staging and no
// barrier in the K loop; the KS waves of a work-group take consecutive
// quarters (KS parts) of K and add their partial sums through LDS in wave
// order.  Work-group tile 16 TI x 16 TJ.
// ---------------------------------------------------------------------------
template <int I, int D, class F>
__device__ inline void for_sets(F f) {
  if constexpr (I < D) {
    f(SetI<I>{});
    for_sets<I + 1, D>(f);
  }
}

// K steps s0 .. s0+S-1 (S % D == 0): set r holds step s0 + r + D q; loads
// past the end re-read the last step, which is never multiplied
template <int D, class L, class M>
__device__ inline void k_ring(int s0, int S, L load, M mma) {
  const int slast = s0 + S - 1;
  for_sets<0, D>([&](auto R) { load(R, min(s0 + (int)decltype(R)::value, slast)); });
  for (int s = s0; s < s0 + S; s += D)
    for_sets<0, D>([&](auto R) {
      mma(R);
      load(R, min(s + D + (int)decltype(R)::value, slast));
    });
}

// partial sums of waves 1 .. KS-1 added to wave 0's through LDS; returns
// false for the waves that are done
template <int TI, int TJ, int KS>
__device__ inline bool reduce_k(d4 (&acc)[TI][TJ], int ks, int lane) {
  if constexpr (KS > 1) {
    __shared__ d4 red[KS - 1][TI * TJ][64];
    if (ks > 0) {
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) red[ks - 1][i * TJ + j][lane] = acc[i][j];
    }
    __syncthreads();
    if (ks > 0) return false;
#pragma unroll
    for (int q = 0; q < KS - 1; ++q)
#pragma unroll
      for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int j = 0; j < TJ; ++j) acc[i][j] += red[q][i * TJ + j][lane];
  }
  return true;
}

// The direct kernels use their own order of the K index, k' = 4s + j with
//   j = 0: (q = s, Re), 1: (q = m/2-1-s, Im), 2: (q = m/2+s, Re), 3: (q = m-1-s, Im)
// (tables fact_fused_matrix passes 3 and 4), so that the pair element a lane
// multiplies at K step s is x~[c_j(p) + 4s] for every j:
//   c_0 = 2p, c_1 = N/2 + 3 - 2p, c_2 = N/2 + 2p, c_3 = 3 - 2p
// and every operand address is a uniform base (stepping by a uniform amount
// per K step) plus a loop-invariant per-lane offset: no vector arithmetic in
// the K loop.
__device__ inline int pair_first(int N, int j, int p) {
  return j == 0 ? 2 * p : j == 1 ? N / 2 + 3 - 2 * p : j == 2 ? N / 2 + 2 * p : 3 - 2 * p;
}

// y pass: A fragment of tile i from X row (m0 + 16 i)/2 + (li & 3) +
// 4 (li >> 3) (virtual row m0 + 16 i + li, parity (li >> 2) & 1), B fragment
// of tile j from the table row 4s + lk, column n0 + 16 j + li
template <int TI, int TJ, int KS, int D>
__global__ __launch_bounds__(KS * 64) void k_fact_yd(FusedArgs a) {
  constexpr int BM = 16 * TI, BN = 16 * TJ;
  static_assert(TJ % 4 == 0, "64-column groups");
  const int N = a.ny, GM = 2 * a.nz, GN = N / 2, K = N / 2;
  const FItem it = fitem<BM, BN>(GM, GN, a.nbatch);
  const int m0 = it.m0, n0 = it.n0, fb = it.f;
  const bool rev = a.ky[fb] == QSS_KIND_S;
  const int t = threadIdx.x, lane = t & 63;
  const int ks = __builtin_amdgcn_readfirstlane(t >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const int c = pair_first(N, lk, (li >> 2) & 1);
  const char *Xb = reinterpret_cast<const char *>(a.in[fb] + it.w * a.w_in);
  const long xstep = rev ? -32 : 32;   // bytes per K step
  uint32_t xo[TI];
#pragma unroll
  for (int i = 0; i < TI; ++i)
    xo[i] = (uint32_t)(((m0 + 16 * i) / 2 + (li & 3) + 4 * (li >> 3)) * N + (rev ? N - 1 - c : c)) * 8u;
  const char *Gb = reinterpret_cast<const char *>(a.Gd + n0);
  const long gstep = 32L * GN;
  const uint32_t go = (uint32_t)(lk * GN + li) * 8u;
  const int S = K / 4 / KS;
  double ra[D][TI], rb[D][TJ];
  d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = d4{0, 0, 0, 0};
  auto load = [&](auto R, int s) {
    constexpr int r = decltype(R)::value;
    const char *xs = Xb + xstep * s, *gs = Gb + gstep * s;
#pragma unroll
    for (int i = 0; i < TI; ++i) ra[r][i] = *reinterpret_cast<const double *>(xs + xo[i]);
#pragma unroll
    for (int j = 0; j < TJ; ++j)
      rb[r][j] = *reinterpret_cast<const double *>(gs + go + 128 * j);
  };
  auto mma = [&](auto R) {
    constexpr int r = decltype(R)::value;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) acc[i][j] = mma64(ra[r][i], rb[r][j], acc[i][j]);
  };
  k_ring<D>(ks * S, S, load, mma);
  if (!reduce_k<TI, TJ, KS>(acc, ks, lane)) return;
  y_epilogue<TI, TJ>(a, acc, m0, n0, fb, it.w, rev, li, lk);
}

// z pass: A fragment of tile i from the k-major table row 4s + lk, column
// m0 + 16 i + li; B fragment of tile j (parity p = j & 1, columns y = n0/2 +
// 16 (j >> 1) + li) from Y row c_lk(p) + 4s (reversed for the DST)
template <int TI, int TJ, int KS, int D>
__global__ __launch_bounds__(KS * 64) void k_fact_zd(FusedArgs a) {
  constexpr int BM = 16 * TI, BN = 16 * TJ;
  static_assert(TJ % 2 == 0, "32-column groups");
  const int Ny = a.ny, Nz = a.nz, GM = Nz / 2, GN = 2 * Ny, K = Nz / 2;
  const FItem it = fitem<BM, BN, true>(GM, GN, a.nbatch);
  const int m0 = it.m0, n0 = it.n0, fb = it.f;
  const bool rev = a.kz[fb] == QSS_KIND_S;
  const int t = threadIdx.x, lane = t & 63;
  const int ks = __builtin_amdgcn_readfirstlane(t >> 6);
  const int li = lane & 15, lk = lane >> 4;
  const char *Yb = reinterpret_cast<const char *>(a.in[fb] + it.w * a.w_in + n0 / 2);
  const long ystep = (rev ? -32L : 32L) * Ny;
  uint32_t yo[2];
#pragma unroll
  for (int p = 0; p < 2; ++p) {
    const int c = pair_first(Nz, lk, p);
    yo[p] = (uint32_t)((rev ? Nz - 1 - c : c) * Ny + li) * 8u;
  }
  const char *Gb = reinterpret_cast<const char *>(a.Gd + m0);
  const long gstep = 32L * GM;
  const uint32_t go = (uint32_t)(lk * GM + li) * 8u;
  const int S = K / 4 / KS;
  double ra[D][TI], rb[D][TJ];
  d4 acc[TI][TJ];
#pragma unroll
  for (int i = 0; i < TI; ++i)
#pragma unroll
    for (int j = 0; j < TJ; ++j) acc[i][j] = d4{0, 0, 0, 0};
  auto load = [&](auto R, int s) {
    constexpr int r = decltype(R)::value;
    const char *gs = Gb + gstep * s, *ys = Yb + ystep * s;
#pragma unroll
    for (int i = 0; i < TI; ++i)
      ra[r][i] = *reinterpret_cast<const double *>(gs + go + 128 * i);
#pragma unroll
    for (int j = 0; j < TJ; ++j)
      rb[r][j] = *reinterpret_cast<const double *>(ys + yo[j & 1] + 128 * (j >> 1));
  };
  auto mma = [&](auto R) {
    constexpr int r = decltype(R)::value;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
      for (int j = 0; j < TJ; ++j) acc[i][j] = mma64(ra[r][i], rb[r][j], acc[i][j]);
  };
  k_ring<D>(ks * S, S, load, mma);
  if (!reduce_k<TI, TJ, KS>(acc, ks, lane)) return;
  z_epilogue<TI, TJ>(a, acc, m0, n0, fb, it.w, rev, li, lk, lane);
}

// tile shapes: {BM, BN, BK, WGM, WGN, KS, PF (register slabs in flight),
// direct}; direct shapes: {16 TI, 16 TJ, 0, 1, 1, KS, D}
struct FShape {
  int bm, bn, bk, wgm, wgn, ks, pf, direct;
};
constexpr FShape kShapesY[] = {
    {32, 64, 16, 2, 1, 1, 2},  {64, 64, 16, 4, 1, 1, 2},  {32, 64, 16, 2, 1, 2, 2},
    {32, 64, 16, 2, 1, 4, 2},  {32, 64, 32, 2, 1, 4, 2},  {64, 64, 16, 4, 1, 2, 2},
    {32, 64, 32, 2, 1, 2, 2},  {32, 128, 16, 2, 1, 1, 2}, {32, 128, 16, 2, 1, 2, 2},
    {64, 128, 16, 4, 1, 1, 2}, {32, 256, 16, 2, 1, 1, 2}, {32, 64, 16, 2, 1, 2, 4},
    {32, 64, 16, 2, 1, 4, 4},  {64, 64, 16, 4, 1, 2, 4},  {32, 64, 16, 2, 1, 1, 4},
    {64, 64, 16, 4, 1, 1, 4},  {32, 64, 0, 1, 1, 4, 4, 1}, {16, 64, 0, 1, 1, 4, 4, 1},
    {32, 64, 0, 1, 1, 2, 4, 1}, {16, 64, 0, 1, 1, 8, 4, 1}, {16, 64, 0, 1, 1, 2, 8, 1},
    {32, 64, 0, 1, 1, 4, 8, 1}};
constexpr FShape kShapesZ[] = {
    {32, 64, 16, 2, 1, 1, 2},  {64, 64, 16, 4, 1, 1, 2},  {32, 32, 16, 2, 1, 2, 2},
    {32, 32, 16, 2, 1, 4, 2},  {32, 64, 16, 2, 1, 2, 2},  {32, 64, 16, 2, 1, 4, 2},
    {32, 32, 32, 2, 1, 4, 2},  {64, 64, 16, 4, 1, 2, 2},  {128, 32, 16, 4, 1, 2, 2},
    {128, 32, 16, 8, 1, 1, 2}, {128, 64, 16, 4, 1, 1, 2}, {128, 32, 16, 4, 1, 1, 2},
    {32, 32, 16, 2, 1, 2, 4},  {32, 32, 16, 2, 1, 4, 4},  {32, 64, 16, 2, 1, 2, 4},
    {64, 64, 16, 4, 1, 2, 4},  {32, 64, 16, 2, 1, 1, 4},  {32, 32, 0, 1, 1, 4, 4, 1},
    {32, 64, 0, 1, 1, 4, 4, 1}, {16, 64, 0, 1, 1, 4, 4, 1}, {16, 32, 0, 1, 1, 4, 4, 1},
    {16, 32, 0, 1, 1, 8, 4, 1}, {32, 32, 0, 1, 1, 2, 8, 1}};
constexpr int kNY = sizeof(kShapesY) / sizeof(kShapesY[0]);
constexpr int kNZ = sizeof(kShapesZ) / sizeof(kShapesZ[0]);

template <class K>
hipError_t launch_fn(K fn, const FusedArgs &a, dim3 grid, dim3 blk, hipStream_t s, hipEvent_t e0,
                     hipEvent_t e1) {
  if (e0 || e1) hipExtLaunchKernelGGL(fn, grid, blk, 0, s, e0, e1, 0, a);
  else hipLaunchKernelGGL(fn, grid, blk, 0, s, a);
  return hipGetLastError();
}

template <int V>
hipError_t launch_y(const FusedArgs &a, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  constexpr FShape S = kShapesY[V];
  const dim3 blk(S.wgm * S.wgn * S.ks * 64);
  if constexpr (S.direct)
    return launch_fn(k_fact_yd<S.bm / 16, S.bn / 16, S.ks, S.pf>, a, grid, blk, s, e0, e1);
  else
    return launch_fn(k_fact_y<S.bm, S.bn, S.bk, S.wgm, S.wgn, S.ks, S.pf>, a, grid, blk, s, e0,
                     e1);
}
template <int V>
hipError_t launch_z(const FusedArgs &a, dim3 grid, hipStream_t s, hipEvent_t e0, hipEvent_t e1) {
  constexpr FShape S = kShapesZ[V];
  const dim3 blk(S.wgm * S.wgn * S.ks * 64);
  if constexpr (S.direct)
    return launch_fn(k_fact_zd<S.bm / 16, S.bn / 16, S.ks, S.pf>, a, grid, blk, s, e0, e1);
  else
    return launch_fn(k_fact_z<S.bm, S.bn, S.bk, S.wgm, S.wgn, S.ks, S.pf>, a, grid, blk, s, e0,
                     e1);
}

// the launchers of variants 0 .. n-1 of a pass
template <int... V>
hipError_t launch_y_v(int v, const FusedArgs &a, dim3 g, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1, std::integer_sequence<int, V...>) {
  hipError_t r = hipErrorInvalidValue;
  ((v == V ? (void)(r = launch_y<V>(a, g, s, e0, e1)) : (void)0), ...);
  return r;
}
template <int... V>
hipError_t launch_z_v(int v, const FusedArgs &a, dim3 g, hipStream_t s, hipEvent_t e0,
                      hipEvent_t e1, std::integer_sequence<int, V...>) {
  hipError_t r = hipErrorInvalidValue;
  ((v == V ? (void)(r = launch_z<V>(a, g, s, e0, e1)) : (void)0), ...);
  return r;
}

const FShape &shape_of(int pass, int v) { return pass == 0 ? kShapesY[v] : kShapesZ[v]; }

}  // namespace

int fact_fused_variants(int pass) { return pass == 0 ? kNY : kNZ; }

bool fact_fused_shape_ok(int pass, int v, int ny, int nz) {
  if (v < 0 || v >= fact_fused_variants(pass)) return false;
  if (ny % 128 || nz % 128 || ny > kMaxN || nz > kMaxN) return false;
  const FShape &s = shape_of(pass, v);
  const int GM = pass == 0 ? 2 * nz : nz / 2, GN = pass == 0 ? ny / 2 : 2 * ny;
  const int K = pass == 0 ? ny / 2 : nz / 2;
  if (s.direct) return GM % s.bm == 0 && GN % s.bn == 0 && (K / 4) % (s.ks * s.pf) == 0;
  return GM % s.bm == 0 && GN % s.bn == 0 && K % (s.pf * s.bk) == 0;
}

hipError_t launch_fact_fused(const FusedArgs &a, int pass, int v, hipStream_t s, hipEvent_t ev0,
                             hipEvent_t ev1) {
  if (!fact_fused_shape_ok(pass, v, a.ny, a.nz)) return hipErrorInvalidValue;
  const FShape &sh = shape_of(pass, v);
  const int GM = pass == 0 ? 2 * a.nz : a.nz / 2, GN = pass == 0 ? a.ny / 2 : 2 * a.ny;
  const dim3 grid((unsigned)((GM / sh.bm) * (GN / sh.bn) * a.nbatch * a.nw));
  if (pass == 0)
    return launch_y_v(v, a, grid, s, ev0, ev1, std::make_integer_sequence<int, kNY>{});
  return launch_z_v(v, a, grid, s, ev0, ev1, std::make_integer_sequence<int, kNZ>{});
}

// G'_y[k'][c'] (pass 0), G'_z[r'][k'] (pass 1), and with the direct kernels'
// K order G'_y (pass 3) and G'_z k-major, [k'][r'] (pass 4), for an axis of
// length n:
// blk[s_in][s_out](q, c) of [[Hr, Hi], [-Hi, Hr]], H_qc = pre_{2q} F^{qc} =
// e^{-i pi (32 qc + 8q + 1) / 4n} (F_qc = e^{-2 pi i qc/m}, the phase reduced
// exactly in integers), with k' -> (q, s_in), c' -> (c, s_out) (y columns),
// r' -> (c, s_out) (z rows) as described at the top of this file
void fact_fused_matrix(int n, int pass, std::vector<double> &g) {
  const int h = n / 2, m = n / 4;
  g.assign((size_t)h * h, 0.0);
  // K index -> (q, s_in): the staged kernels' order (passes 0, 1) or the
  // direct kernels' (passes 3, 4)
  const bool direct = pass >= 3;
  auto kmap = [&](int kp, int &q, int &si) {
    const int u = kp / 4, j = kp % 4;
    if (!direct) {
      q = j < 2 ? u : m - 1 - u;
    } else {
      const int qs[4] = {u, m / 2 - 1 - u, m / 2 + u, m - 1 - u};
      q = qs[j];
    }
    si = j & 1;
  };
  auto val = [&](int q, int c, int si, int so) {
    const long r = (32L * q * c + 8L * q + 1) % (8L * n);
    const double ang = M_PI * (double)r / (4.0 * n);
    const double fr = std::cos(ang), fi = -std::sin(ang);
    const double b[2][2] = {{fr, fi}, {-fi, fr}};
    return b[si][so];
  };
  for (int kp = 0; kp < h; ++kp) {
    int q, si;
    kmap(kp, q, si);
    for (int x = 0; x < h; ++x) {
      int c, so;
      if (pass == 0 || pass == 3) {
        // pass 0 stores the columns of tiles 2j and 2j+1 interleaved (the
        // staged kernels read both with one 16-byte LDS read); xl is the column
        // in GEMM order
        const int xl = pass == 0 ? 32 * (x / 32) + 16 * (x & 1) + (x % 32) / 2 : x;
        const int gg = xl / 64, slot = (xl % 64) / 16, l = xl % 16, k = 16 * gg + l;
        c = slot < 2 ? k : m - 1 - k;
        so = slot & 1;
        g[(size_t)kp * h + x] = val(q, c, si, so);
      } else {
        const int T = x / 16, tt = x % 16;
        c = 8 * T + (tt & 3) + 4 * (tt >> 3);
        so = (tt >> 2) & 1;
        // pass 1: row-major [r'][k'], pass 4: k-major [k'][r']
        g[pass == 1 ? (size_t)x * h + kp : (size_t)kp * h + x] = val(q, c, si, so);
      }
    }
  }
}

// the butterfly twiddles of the fused kernels for an axis of length n:
// mid'_k = e^{-i pi (16k + 4) / 4n} (k < n/4; the odd parity's e^{-i pi/n}
// included), then post_k = e^{-i pi 4k / 4n} (k < n/2), as (re, im) pairs
void fact_fused_twiddles(int n, std::vector<double> &tw) {
  tw.clear();
  auto put = [&](long a) {
    a %= 8L * n;
    const double ang = M_PI * (double)a / (4.0 * n);
    tw.push_back(std::cos(ang));
    tw.push_back(-std::sin(ang));
  };
  for (int k = 0; k < n / 4; ++k) put(16L * k + 4);
  for (int k = 0; k < n / 2; ++k) put(4L * k);
}
