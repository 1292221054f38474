// fact.hip — the 2-D half-integer transforms with half the MFMA work.
//
// A length-N DCT-IV (REDFT11: y_k = 2 sum_n x_n cos(pi (n+1/2)(k+1/2)/N),
// movefieldshydrolinlayersplit.cpp:389-427) is an N/2-point complex DFT between
// twiddles:
//   v_n = (x_{2n} + i x_{N-1-2n}) e^{-i pi (4n+1)/(4N)},  n < N/2
//   V   = DFT_{N/2}(v),  w_k = V_k e^{-i pi k/N}
//   y_{2k} = 2 Re w_k,  y_{N-1-2k} = -2 Im w_k
// and one radix-2 step of that DFT, V_k = E_k + W^k O_k, V_{k+N/4} = E_k - W^k O_k
// (E, O the N/4-point DFTs of the even and odd v, W = e^{-2 pi i/(N/2)}), leaves
// two N/4-point complex DFTs: as real products, an (N/2 x N/2) matrix applied
// to the even and the odd half of every row, N^2/2 multiply-adds per row
// instead of the dense table's N^2.  A DST-IV is the DCT-IV of the reversed
// input with the output's odd entries negated (y_k (-1)^k).
//
// One 2-D transform T_kz . X . T_ky of a batch of planes:
//   k_fact_pre   rows of X -> Vin_y (2 Nz x Ny/2): row 2r+p holds the parity-p
//                half of row r's v as [Re | Im]
//   GEMM         Out_y = Vin_y . G_y   (G_y = [[Fr, Fi], [-Fi, Fr]], F = DFT_{Ny/4})
//   k_fact_mid   butterflies and post-twiddles along y, then the same
//                pre-twiddle along z: Vin_z (Nz/2 x 2 Ny), row = z index of
//                [Re | Im], column = (parity, y)
//   GEMM         Out_z = G_z . Vin_z   (G_z = G_y^T for Nz)
//   k_fact_post  butterflies and post-twiddles along z, then the epilogue of the
//                dense path (alpha, beta accumulation, corrector residual, push
//                field (P+C)/2), stored in the natural layout
// The GEMMs are the dense path's kernel (gemm.hip) on these shapes.

#include <hip/hip_ext.h>

#include <cmath>

#include "qss_internal.h"

namespace {

constexpr int NT = 256;

__device__ inline double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

__device__ inline void atomic_max_nonneg(unsigned long long *addr, double v) {
  atomicMax(addr, (unsigned long long)__double_as_longlong(v));
}

__device__ inline double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}

// twiddles of one axis (length N): pre[n] (n < N/2), mid[k] (k < N/4), post[k] (k < N/2)
struct Tw {
  const double2 *pre, *mid, *post;
};
__device__ inline Tw tw_of(const double2 *t, int n) {
  return Tw{t, t + n / 2, t + n / 2 + n / 4};
}

// grid: (ceil(Nz * Ny/4 / NT), nbatch * nw).  Thread (r, n), n < Ny/4, forms
// v_n and v_n' for n' = Ny/2-1-n: their four inputs are the pairs
// (x_2n, x_2n+1) and (x_Ny-2-2n, x_Ny-1-2n), two 16-byte loads.
__global__ __launch_bounds__(NT) void k_fact_pre(FactArgs a) {
  const int ny = a.ny, nz = a.nz, h = ny / 2, q4 = ny / 4;
  const long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long)nz * q4) return;
  const int b = blockIdx.y % a.nbatch, w = blockIdx.y / a.nbatch;
  const long nc = (long)ny * nz;
  const int r = (int)(i / q4), n = (int)(i - (long)r * q4);
  const double *x = a.in[b] + w * a.w_in + (long)r * ny;
  const double2 lo = *reinterpret_cast<const double2 *>(x + 2 * n);
  const double2 hi = *reinterpret_cast<const double2 *>(x + ny - 2 - 2 * n);
  // x_2n = lo.x, x_Ny-1-2n = hi.y; x_2n' = hi.x, x_Ny-1-2n' = lo.y
  double a0 = lo.x, b0 = hi.y, a1 = hi.x, b1 = lo.y;
  if (a.ky[b] == QSS_KIND_S) {   // reversed input: x~_j = x_{Ny-1-j}
    a0 = hi.y; b0 = lo.x; a1 = lo.y; b1 = hi.x;
  }
  const Tw t = tw_of(a.twy, ny);
  const int n1 = h - 1 - n;
  const double2 v0 = cmul(make_double2(a0, b0), t.pre[n]);
  const double2 v1 = cmul(make_double2(a1, b1), t.pre[n1]);
  double *base = a.t0[b] + w * nc + (long)(2 * r) * h;
  double *o0 = base + (n & 1) * h, *o1 = base + (n1 & 1) * h;
  o0[n >> 1] = v0.x;
  o0[q4 + (n >> 1)] = v0.y;
  o1[n1 >> 1] = v1.x;
  o1[q4 + (n1 >> 1)] = v1.y;
}

// y-transformed values Y[row][2kk] and Y[row][Ny-1-2kk] from Out_y
__device__ inline void y_pair(const FactArgs &a, const double *Oy, const Tw &ty, int row,
                              int kk, bool sy, double &y0, double &y1) {
  const int ny = a.ny, hy = ny / 2, my = ny / 4;
  const int k = kk < my ? kk : kk - my;
  const double sg = kk < my ? 1.0 : -1.0;
  const double *e = Oy + (long)(2 * row) * hy, *o = e + hy;
  const double2 E = make_double2(e[k], e[my + k]);
  const double2 O = make_double2(o[k], o[my + k]);
  const double2 wo = cmul(ty.mid[k], O);
  const double2 V = make_double2(E.x + sg * wo.x, E.y + sg * wo.y);
  const double2 ww = cmul(V, ty.post[kk]);
  y0 = 2.0 * ww.x;
  y1 = sy ? 2.0 * ww.y : -2.0 * ww.y;   // DST: odd outputs negated
}

// grid: (ceil(Nz/2 * Ny/4 / NT), nbatch * nw).  Thread (n, kk), kk < Ny/4:
// rows 2n and Nz-1-2n; kk and kk' = Ny/2-1-kk give the columns 2kk, 2kk+1
// (= Ny-1-2kk') and Ny-2-2kk (= 2kk'), Ny-1-2kk, stored as 16-byte pairs.
__global__ __launch_bounds__(NT) void k_fact_mid(FactArgs a) {
  const int ny = a.ny, nz = a.nz, hy = ny / 2, qy = ny / 4, hz = nz / 2, mz = nz / 4;
  const long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i >= (long)hz * qy) return;
  const int b = blockIdx.y % a.nbatch, w = blockIdx.y / a.nbatch;
  const long nc = (long)ny * nz;
  const int n = (int)(i / qy), kk = (int)(i - (long)n * qy);
  const int kk1 = hy - 1 - kk;
  const double *Oy = a.t1[b] + w * nc;
  const Tw ty = tw_of(a.twy, ny), tz = tw_of(a.twz, nz);
  const bool sy = a.ky[b] == QSS_KIND_S;
  // Y[q][c]: row q (0: 2n, 1: Nz-1-2n), columns c = 2kk, 2kk+1, Ny-2-2kk, Ny-1-2kk
  double Y[2][4];
  const int rows[2] = {2 * n, nz - 1 - 2 * n};
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    y_pair(a, Oy, ty, rows[q], kk, sy, Y[q][0], Y[q][3]);
    y_pair(a, Oy, ty, rows[q], kk1, sy, Y[q][2], Y[q][1]);
  }
  // pre-twiddle along z: v_n = (x~_{2n} + i x~_{Nz-1-2n}), x~ reversed for DST
  const bool sz = a.kz[b] == QSS_KIND_S;
  const double2 twp = tz.pre[n];
  const int p = n & 1, m = n >> 1;
  double *vz = a.t0[b] + w * nc;
  double re[4], im[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const double xa = sz ? Y[1][c] : Y[0][c], xb = sz ? Y[0][c] : Y[1][c];
    const double2 v = cmul(make_double2(xa, xb), twp);
    re[c] = v.x;
    im[c] = v.y;
  }
  double *rr = vz + (long)m * (2 * ny) + (long)p * ny, *ri = vz + (long)(mz + m) * (2 * ny) + (long)p * ny;
  *reinterpret_cast<double2 *>(rr + 2 * kk) = make_double2(re[0], re[1]);
  *reinterpret_cast<double2 *>(rr + ny - 2 - 2 * kk) = make_double2(re[2], re[3]);
  *reinterpret_cast<double2 *>(ri + 2 * kk) = make_double2(im[0], im[1]);
  *reinterpret_cast<double2 *>(ri + ny - 2 - 2 * kk) = make_double2(im[2], im[3]);
}

// grid: (ceil(Nz/2 * Ny / NT), nbatch * nw); thread (kk, j): rows 2kk and Nz-1-2kk
__global__ __launch_bounds__(NT) void k_fact_post(FactArgs a) {
  const int ny = a.ny, nz = a.nz, hz = nz / 2, mz = nz / 4;
  const long i = (long)blockIdx.x * NT + threadIdx.x;
  const int b = blockIdx.y % a.nbatch, w = blockIdx.y / a.nbatch;
  const bool live = i < (long)hz * ny;
  const int kk = live ? (int)(i / ny) : 0, j = live ? (int)(i - (long)kk * ny) : 0;
  const long nc = (long)ny * nz;
  double dmax = 0.0, vmax = 0.0;
  const int resid = a.resid_c[b], push = a.push[b];
  if (live) {
    const double *Oz = a.t1[b] + w * nc;
    const Tw tz = tw_of(a.twz, nz);
    const int k = kk < mz ? kk : kk - mz;
    const double sg = kk < mz ? 1.0 : -1.0;
    const long row2 = 2L * ny;
    const double2 E = make_double2(Oz[k * row2 + j], Oz[(mz + k) * row2 + j]);
    const double2 O = make_double2(Oz[k * row2 + ny + j], Oz[(mz + k) * row2 + ny + j]);
    const double2 wo = cmul(tz.mid[k], O);
    const double2 V = make_double2(E.x + sg * wo.x, E.y + sg * wo.y);
    const double2 ww = cmul(V, tz.post[kk]);
    const bool sz = a.kz[b] == QSS_KIND_S;
    const double val[2] = {2.0 * ww.x, sz ? 2.0 * ww.y : -2.0 * ww.y};
    const long n0[2] = {(long)(2 * kk) * ny + j, (long)(nz - 1 - 2 * kk) * ny + j};
    double *C = a.out[b] + w * a.w_out;
    const double *P = a.P[b] ? a.P[b] + w * a.w_p : nullptr;
    const bool need_old = resid >= 0 || a.beta != 0.0;
    double oldv[2] = {0.0, 0.0}, pv[2] = {0.0, 0.0};
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      if (need_old) oldv[q] = C[n0[q]];
      if (push >= 0) pv[q] = P[n0[q]];
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      double v = a.alpha * val[q];
      if (a.beta != 0.0) v += a.beta * oldv[q];   // accumulate (adjoint sweep)
      if (resid >= 0) {
        const double d = fabs(oldv[q] - v), mm = fabs(v);
        dmax = fmax(dmax, d == d ? d : INFINITY);
        vmax = fmax(vmax, mm == mm ? mm : INFINITY);
      }
      C[n0[q]] = v;
      if (push >= 0) {
        const double pq = 0.5 * (pv[q] + v);
        const long pn = (long)push * nc + n0[q];
        if (a.Ff) a.Ff[w * a.w_f + pn] = (float)pq;
        else a.F[w * a.w_f + pn] = pq;
      }
    }
  }
  if (resid >= 0) {   // uniform over the work-group (one field per blockIdx.y)
    dmax = wave_max(dmax);
    vmax = wave_max(vmax);
    if ((threadIdx.x & 63) == 0) {
      unsigned long long *R = a.rslots + w * a.w_r + (blockIdx.x % kResidSlots) * 8;
      atomic_max_nonneg(&R[resid], dmax);
      atomic_max_nonneg(&R[3 + resid], vmax);
    }
  }
}

}  // namespace

hipError_t launch_fact(const FactArgs &a, int stage, hipStream_t s, hipEvent_t ev0,
                       hipEvent_t ev1) {
  const long nyz = (long)a.ny * a.nz;
  const long items = stage == 0 ? nyz / 4 : stage == 1 ? nyz / 8 : nyz / 2;
  const dim3 grid((unsigned)((items + NT - 1) / NT), (unsigned)(a.nbatch * a.nw));
  auto fn = stage == 0 ? k_fact_pre : stage == 1 ? k_fact_mid : k_fact_post;
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(fn, grid, dim3(NT), 0, s, ev0, ev1, 0, a);
  else
    hipLaunchKernelGGL(fn, grid, dim3(NT), 0, s, a);
  return hipGetLastError();
}

// twiddle table of one axis: pre, mid, post as (cos, -sin) of pi a / (4N) for
// integer a (4n+1, 16k, 4k), the angle reduced exactly through a mod 8N
void fact_twiddles(int n, std::vector<double> &tw) {
  tw.clear();
  auto put = [&](long a) {
    a %= 8L * n;
    const double ang = M_PI * (double)a / (4.0 * n);
    tw.push_back(std::cos(ang));
    tw.push_back(-std::sin(ang));
  };
  for (int q = 0; q < n / 2; ++q) put(4L * q + 1);
  for (int q = 0; q < n / 4; ++q) put(16L * q);
  for (int q = 0; q < n / 2; ++q) put(4L * q);
}

// G (N/2 x N/2, row-major) for the y pass, [[Fr, Fi], [-Fi, Fr]] with
// F_rc = e^{-2 pi i rc/(N/4)}; transpose = true gives G_z = G^T
void fact_matrix(int n, bool transpose, std::vector<double> &g) {
  const int m = n / 4, h = n / 2;
  g.assign((size_t)h * h, 0.0);
  for (int r = 0; r < m; ++r)
    for (int c = 0; c < m; ++c) {
      const long q = ((long)r * c) % m;
      const double ang = 2.0 * M_PI * (double)q / m;
      const double fr = std::cos(ang), fi = -std::sin(ang);
      const double blk[2][2] = {{fr, fi}, {-fi, fr}};
      for (int rb = 0; rb < 2; ++rb)
        for (int cb = 0; cb < 2; ++cb) {
          const size_t row = (size_t)rb * m + r, col = (size_t)cb * m + c;
          if (transpose) g[col * h + row] = blk[rb][cb];
          else g[row * h + col] = blk[rb][cb];
        }
    }
}
