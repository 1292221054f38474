// spectral.hip — per-mode predictor (Guess) and corrector (Iterate) updates,
// utility fills, and the residual finaliser.
//
// One thread per spectral mode (j,k), n1 = j + ny*k.  The arithmetic follows
// the reference CPU loops term by term:
//   Guess   movefieldshydrolinlayersplit.cpp:807-901 (GPU twin
//           CUDA_WRAP/linearized.cu:11-92)
//   Iterate movefieldshydrolinlayersplit.cpp:1373-1478 (GPU twin
//           CUDA_WRAP/linearized.cu:105-208)
// The P-layer spectra are the spectra left by the previous layer's final
// Iterate: the reference recomputes them by forward-transforming the P layer
// it has just inverse-transformed (T.T/(4NyNz) = I), so keeping them saves
// 7 transforms per layer with identical results up to rounding.

#include "qss_internal.h"
#include "spectral_mode.h"
#include "split_bf16.h"

namespace {

constexpr int NT = 256;

__global__ void k_fill_i32(int32_t *p, long n, int32_t v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double *p, long n, double v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, long n, unsigned long long v) {
  long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// zero fills of up to kFillBufs double buffers in one launch (blockIdx.y
// picks the buffer)
__global__ void k_zero_bufs(ZeroBufs a) {
  const int b = blockIdx.y;
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < a.n[b]) a.p[b][i] = 0.0;
}
// up to kCopyBufs device-to-device copies in one launch: blockIdx.y picks the
// buffer, a thread moves 16 bytes (4-byte steps for the last partial 16 bytes).
// Buffers are 16-byte aligned and their sizes multiples of 4 bytes.
__global__ void k_copy_bufs(CopyBufs a) {
  const int b = blockIdx.y;
  const long nb = a.nbytes[b];
  const long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 16;
  if (i >= nb) return;
  const char *src = static_cast<const char *>(a.src[b]);
  char *dst = static_cast<char *>(a.dst[b]);
  if (i + 16 <= nb) {
    *reinterpret_cast<uint4 *>(dst + i) = *reinterpret_cast<const uint4 *>(src + i);
  } else {
    for (long k = i; k < nb; k += 4)
      *reinterpret_cast<uint32_t *>(dst + k) = *reinterpret_cast<const uint32_t *>(src + k);
  }
}

struct SpecArgs {
  double *S[SP_N];
  int ny, nz, nw;
  double ly, lz, hx, dens;
  int flag;  // iterate: final pass
  // split-bf16 modes: the first nxs outputs (E, B, then rho, psi) also
  // written as the split operands of the inverse transform that follows
  uint16_t *xs[kGemmMaxBatch];
  int npc, nxs;
  // iterate: residual slots (kResidPerWake per wake) cleared for the inverse-GEMM
  // epilogues that follow on the stream (null: untouched)
  unsigned long long *resid;
};

__device__ inline void wavenumbers(const SpecArgs &a, int j, int k, double &aky,
                                   double &akz, double &ak2) {
  akz = M_PI / a.lz * (k + 0.5);
  aky = M_PI / a.ly * (j + 0.5);
  ak2 = aky * aky + akz * akz;
}

__device__ inline qss_spec::Mode mode_of(const SpecArgs &a, int j, int k) {
  qss_spec::Mode m;
  wavenumbers(a, j, k, m.aky, m.akz, m.ak2);
  m.dens = a.dens;
  m.hx = a.hx;
  return m;
}

__global__ __launch_bounds__(NT) void k_guess(SpecArgs a) {
  using namespace qss_spec;
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;   // wake-major mode index
  const long nc = (long)a.ny * a.nz;
  if (n1 >= nc * a.nw) return;
  const long mi = n1 % nc;
  const Mode m = mode_of(a, (int)(mi % a.ny), (int)(mi / a.ny));
  double **S = a.S;
  // P spectra = the spectra left by the previous Iterate (or zero after a
  // reset); the beam P spectrum is the previous beam spectrum (:637-646)
  double in[GI_N], out[GO_N];
  in[GI_RP] = S[SP_RHO][n1];
  in[GI_JXP] = S[SP_JX][n1];
  in[GI_JYP] = S[SP_JY][n1];
  in[GI_JZP] = S[SP_JZ][n1];
  in[GI_EXP] = S[SP_EX][n1];
  in[GI_EYP] = S[SP_EY][n1];
  in[GI_EZP] = S[SP_EZ][n1];
  in[GI_RB] = S[SP_RHOB][n1];
  in[GI_RBP] = S[SP_RHOBP][n1];
  in[GI_JXB] = S[SP_JXB][n1];
  S[SP_RHOP][n1] = in[GI_RP];
  S[SP_JXP][n1] = in[GI_JXP];
  S[SP_JYP][n1] = in[GI_JYP];
  S[SP_JZP][n1] = in[GI_JZP];
  S[SP_EXP][n1] = in[GI_EXP];
  S[SP_EYP][n1] = in[GI_EYP];
  S[SP_EZP][n1] = in[GI_EZP];
  guess(m, in, out);
  S[SP_RHO][n1] = out[GO_RHO];
  S[SP_EX][n1] = out[GO_EX];
  S[SP_EY][n1] = out[GO_EY];
  S[SP_EZ][n1] = out[GO_EZ];
  S[SP_JX][n1] = out[GO_JX];
  S[SP_JY][n1] = out[GO_JY];
  S[SP_JZ][n1] = out[GO_JZ];
  S[SP_BX][n1] = out[GO_BX];
  S[SP_BY][n1] = out[GO_BY];
  S[SP_BZ][n1] = out[GO_BZ];
  if (a.nxs) {   // single wake (split modes): n1 < nc
    const double v[6] = {out[GO_EX], out[GO_EY], out[GO_EZ],
                         out[GO_BX], out[GO_BY], out[GO_BZ]};
#pragma unroll
    for (int b = 0; b < 6; ++b) qss_split::store(a.xs[b], nc, n1, v[b], a.npc);
  }
}

__global__ __launch_bounds__(NT) void k_iterate(SpecArgs a) {
  using namespace qss_spec;
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;   // wake-major mode index
  const long nc = (long)a.ny * a.nz;
  if (a.resid && n1 < (long)kResidPerWake * a.nw) a.resid[n1] = 0ull;
  if (n1 >= nc * a.nw) return;
  const long mi = n1 % nc;
  const Mode m = mode_of(a, (int)(mi % a.ny), (int)(mi / a.ny));
  double **S = a.S;
  double in[II_N], out[IO_N];
  in[II_RP] = S[SP_RHOP][n1];
  in[II_JX] = S[SP_JX][n1];
  in[II_JY] = S[SP_JY][n1];
  in[II_JZ] = S[SP_JZ][n1];
  in[II_JXP] = S[SP_JXP][n1];
  in[II_JYP] = S[SP_JYP][n1];
  in[II_JZP] = S[SP_JZP][n1];
  in[II_RB] = S[SP_RHOB][n1];
  in[II_RBP] = S[SP_RHOBP][n1];
  in[II_EYP] = S[SP_EYP][n1];
  in[II_EZP] = S[SP_EZP][n1];
  in[II_RHO_OLD] = S[SP_RHO][n1];
  in[II_JXB] = S[SP_JXB][n1];
  iterate(m, in, out);
  S[SP_EY][n1] = out[IO_EY];
  S[SP_EZ][n1] = out[IO_EZ];
  S[SP_RHO][n1] = out[IO_RHO];
  S[SP_EX][n1] = out[IO_EX];
  S[SP_BX][n1] = out[IO_BX];
  S[SP_BY][n1] = out[IO_BY];
  S[SP_BZ][n1] = out[IO_BZ];
  if (a.flag) S[SP_PSI][n1] = out[IO_PSI];
  if (a.nxs) {   // single wake (split modes): n1 < nc
    const double v[8] = {out[IO_EX], out[IO_EY], out[IO_EZ], out[IO_BX],
                         out[IO_BY], out[IO_BZ], out[IO_RHO], a.flag ? out[IO_PSI] : 0.0};
    for (int b = 0; b < a.nxs; ++b) qss_split::store(a.xs[b], nc, n1, v[b], a.npc);
  }
}

// ---- adjoints (A = adjoint of the spectral state, same ids) ---------------
// Reverse of k_guess: the written entries' adjoints are consumed (zeroed)
// and M^T of them is added to the read entries; an entry both read and
// written (RHO, JX, ..., EZ) ends with the adjoint of its value before the
// Guess, the P copies' adjoints included.
__global__ __launch_bounds__(NT) void k_guess_adj(SpecArgs a) {
  using namespace qss_spec;
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;   // wake-major mode index
  const long nc = (long)a.ny * a.nz;
  if (n1 >= nc * a.nw) return;
  const long mi = n1 % nc;
  const Mode m = mode_of(a, (int)(mi % a.ny), (int)(mi / a.ny));
  double **A = a.S;
  double ob[GO_N], ib[GI_N];
  ob[GO_RHO] = A[SP_RHO][n1];
  ob[GO_EX] = A[SP_EX][n1];
  ob[GO_EY] = A[SP_EY][n1];
  ob[GO_EZ] = A[SP_EZ][n1];
  ob[GO_JX] = A[SP_JX][n1];
  ob[GO_JY] = A[SP_JY][n1];
  ob[GO_JZ] = A[SP_JZ][n1];
  ob[GO_BX] = A[SP_BX][n1];
  ob[GO_BY] = A[SP_BY][n1];
  ob[GO_BZ] = A[SP_BZ][n1];
  guess_adj(m, ob, ib);
  A[SP_RHO][n1] = ib[GI_RP] + A[SP_RHOP][n1];
  A[SP_JX][n1] = ib[GI_JXP] + A[SP_JXP][n1];
  A[SP_JY][n1] = ib[GI_JYP] + A[SP_JYP][n1];
  A[SP_JZ][n1] = ib[GI_JZP] + A[SP_JZP][n1];
  A[SP_EX][n1] = ib[GI_EXP] + A[SP_EXP][n1];
  A[SP_EY][n1] = ib[GI_EYP] + A[SP_EYP][n1];
  A[SP_EZ][n1] = ib[GI_EZP] + A[SP_EZP][n1];
  A[SP_RHOB][n1] += ib[GI_RB];
  A[SP_RHOBP][n1] = a.flag ? 0.0 : A[SP_RHOBP][n1] + ib[GI_RBP];   // flag: first layer
  A[SP_JXB][n1] += ib[GI_JXB];
  A[SP_RHOP][n1] = 0.0;
  A[SP_JXP][n1] = 0.0;
  A[SP_JYP][n1] = 0.0;
  A[SP_JZP][n1] = 0.0;
  A[SP_EXP][n1] = 0.0;
  A[SP_EYP][n1] = 0.0;
  A[SP_EZP][n1] = 0.0;
  A[SP_BX][n1] = 0.0;
  A[SP_BY][n1] = 0.0;
  A[SP_BZ][n1] = 0.0;
}

// Reverse of k_iterate (flag: final pass, psi written)
__global__ __launch_bounds__(NT) void k_iterate_adj(SpecArgs a) {
  using namespace qss_spec;
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;   // wake-major mode index
  const long nc = (long)a.ny * a.nz;
  if (n1 >= nc * a.nw) return;
  const long mi = n1 % nc;
  const Mode m = mode_of(a, (int)(mi % a.ny), (int)(mi / a.ny));
  double **A = a.S;
  double ob[IO_N], ib[II_N];
  ob[IO_RHO] = A[SP_RHO][n1];
  ob[IO_EX] = A[SP_EX][n1];
  ob[IO_EY] = A[SP_EY][n1];
  ob[IO_EZ] = A[SP_EZ][n1];
  ob[IO_BX] = A[SP_BX][n1];
  ob[IO_BY] = A[SP_BY][n1];
  ob[IO_BZ] = A[SP_BZ][n1];
  ob[IO_PSI] = a.flag ? A[SP_PSI][n1] : 0.0;
  iterate_adj(m, ob, ib);
  A[SP_EX][n1] = 0.0;
  A[SP_EY][n1] = 0.0;
  A[SP_EZ][n1] = 0.0;
  A[SP_BX][n1] = 0.0;
  A[SP_BY][n1] = 0.0;
  A[SP_BZ][n1] = 0.0;
  if (a.flag) A[SP_PSI][n1] = 0.0;
  A[SP_RHO][n1] = ib[II_RHO_OLD];
  A[SP_RHOP][n1] += ib[II_RP];
  A[SP_JX][n1] += ib[II_JX];
  A[SP_JY][n1] += ib[II_JY];
  A[SP_JZ][n1] += ib[II_JZ];
  A[SP_JXP][n1] += ib[II_JXP];
  A[SP_JYP][n1] += ib[II_JYP];
  A[SP_JZP][n1] += ib[II_JZP];
  A[SP_RHOB][n1] += ib[II_RB];
  A[SP_RHOBP][n1] += ib[II_RBP];
  A[SP_EYP][n1] += ib[II_EYP];
  A[SP_EZP][n1] += ib[II_EZP];
  A[SP_JXB][n1] += ib[II_JXB];
}

// Eerr = max_c dif_c / max(1e-5, max_c)  (:1541, :1584-1621, :1641-1643)
// wake w's residual slots -> dst[w * stride]; during a graph replay the
// layer slot comes from device memory: dst += *slice_dev * per_slice
__global__ void k_resid_final(const unsigned long long *r, double *dst, int nw, long stride,
                              const int *slice_dev, int per_slice) {
  // one wave per wake (blockIdx.x): max over the slots, lanes then shuffles;
  // the bit patterns of non-negative doubles (and of +inf) order like the values
  const int w = blockIdx.x, lane = threadIdx.x;
  if (w >= nw) return;
  r += (long)kResidPerWake * w;
  dst += w * stride;
  if (slice_dev) dst += (long)(*slice_dev) * per_slice;
  unsigned long long m[6] = {0, 0, 0, 0, 0, 0};
  for (int s = lane; s < kResidSlots; s += 64)
    for (int q = 0; q < 6; ++q) m[q] = m[q] > r[s * 8 + q] ? m[q] : r[s * 8 + q];
  for (int o = 32; o > 0; o >>= 1)
    for (int q = 0; q < 6; ++q) {
      const unsigned long long v = __shfl_xor(m[q], o, 64);
      m[q] = m[q] > v ? m[q] : v;
    }
  if (lane != 0) return;
  double err = 0.0;
  for (int c = 0; c < 3; ++c) {
    double dif = __longlong_as_double((long long)m[c]);
    double mx = __longlong_as_double((long long)m[3 + c]);
    if (!(mx > 1e-5)) mx = (mx != mx) ? mx : 1e-5;
    double e = dif / mx;
    err = (e > err || e != e) ? e : err;
  }
  *dst = err;
}

// F = (P + C)/2 as 6 planes, for layers edited from the host; fp32 copy in
// the mixed modes
template <typename FT>
__global__ __launch_bounds__(NT) void k_push_field(const double *p0, const double *p1,
                                                  const double *p2, const double *p3,
                                                  const double *p4, const double *p5,
                                                  const double *c0, const double *c1,
                                                  const double *c2, const double *c3,
                                                  const double *c4, const double *c5,
                                                  FT *F, long n, int nw) {
  const long i = (long)blockIdx.x * NT + threadIdx.x;   // wake-major node index
  if (i >= n * nw) return;
  FT *f = F + (i / n) * 6 * n + i % n;   // the wake's six planes
  f[0] = (FT)(0.5 * (p0[i] + c0[i]));
  f[n] = (FT)(0.5 * (p1[i] + c1[i]));
  f[2 * n] = (FT)(0.5 * (p2[i] + c2[i]));
  f[3 * n] = (FT)(0.5 * (p3[i] + c3[i]));
  f[4 * n] = (FT)(0.5 * (p4[i] + c4[i]));
  f[5 * n] = (FT)(0.5 * (p5[i] + c5[i]));
}

SpecArgs make_args(qss_ctx *c, int flag) {
  SpecArgs a;
  for (int s = 0; s < SP_N; ++s) a.S[s] = c->S[s];
  a.ny = c->ny; a.nz = c->nz; a.nw = c->nw;
  a.ly = c->p.ly; a.lz = c->p.lz;
  a.hx = c->hxs; a.dens = c->p.dens;
  a.flag = flag;
  a.resid = nullptr;
  for (int b = 0; b < kGemmMaxBatch; ++b) a.xs[b] = c->xs[b];
  a.npc = c->npc;
  a.nxs = 0;
  return a;
}

SpecArgs make_adj_args(qss_ctx *c, double *const *A, int flag) {
  SpecArgs a = make_args(c, flag);
  for (int s = 0; s < SP_N; ++s) a.S[s] = A[s];
  return a;
}

// history store inside a replayed graph: wake w's plane of `src` into
// hist + (w * nslices + *slice_dev) * n
__global__ void k_store_hist(const double *src, double *hist, long n, int nw, int nslices,
                             const int *slice_dev) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * nw) return;
  const long w = i / n;
  hist[(w * nslices + *slice_dev) * n + i % n] = src[i];
}

__global__ void k_add_i32(int *p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *p += v;
}

// slice_dev (graph replay): x is the plane of slice *slice_dev in an array of
// planes of n values
__global__ void k_axpy(double *y, const double *x, double alpha, long n, const int *slice_dev) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (slice_dev) x += (long)(*slice_dev) * n;
  if (i < n) y[i] += alpha * x[i];
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }

}  // namespace

hipError_t launch_axpy(double *y, const double *x, double a, long n, hipStream_t s,
                       const int *slice_dev) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_axpy, grid1(n), dim3(NT), 0, s, y, x, a, n, slice_dev);
  return hipGetLastError();
}

hipError_t launch_guess_adj(qss_ctx *c, double *const *A, int first_layer) {
  hipLaunchKernelGGL(k_guess_adj, grid1(c->ncell), dim3(NT), 0, c->stream,
                     make_adj_args(c, A, first_layer));
  return hipGetLastError();
}

hipError_t launch_iterate_adj(qss_ctx *c, double *const *A, int final_pass) {
  hipLaunchKernelGGL(k_iterate_adj, grid1(c->ncell), dim3(NT), 0, c->stream,
                     make_adj_args(c, A, final_pass));
  return hipGetLastError();
}

hipError_t launch_fill_i32(int32_t *p, long n, int32_t v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, grid1(n), dim3(NT), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_fill_f64(double *p, long n, double v, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, grid1(n), dim3(NT), 0, s, p, n, v);
  return hipGetLastError();
}
hipError_t launch_zero_bufs(const ZeroBufs &a, hipStream_t s) {
  long mx = 0;
  for (int b = 0; b < kFillBufs; ++b) mx = a.n[b] > mx ? a.n[b] : mx;
  if (mx == 0) return hipSuccess;
  hipLaunchKernelGGL(k_zero_bufs, dim3(grid1(mx).x, kFillBufs), dim3(NT), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_copy_bufs(const CopyBufs &a, hipStream_t s) {
  long mx = 0;
  for (int b = 0; b < kCopyBufs; ++b) {
    if (a.nbytes[b] % 4 || ((uintptr_t)a.src[b] | (uintptr_t)a.dst[b]) % 16)
      return hipErrorInvalidValue;
    mx = a.nbytes[b] > mx ? a.nbytes[b] : mx;
  }
  if (mx == 0) return hipSuccess;
  hipLaunchKernelGGL(k_copy_bufs, dim3(grid1((mx + 15) / 16).x, kCopyBufs), dim3(NT), 0, s, a);
  return hipGetLastError();
}
hipError_t launch_fill_u64(unsigned long long *p, long n, unsigned long long v,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, grid1(n), dim3(NT), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_guess_update(qss_ctx *c, int first_layer) {
  SpecArgs a = make_args(c, first_layer);
  if (c->npc) {
    a.nxs = 6;
    const int ids[6] = {SP_EX, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ};
    for (int b = 0; b < 6; ++b) c->xs_src[b] = c->S[ids[b]];
  }
  hipLaunchKernelGGL(k_guess, grid1(c->ncell * c->nw), dim3(NT), 0, c->stream, a);
  return hipGetLastError();
}

hipError_t launch_iterate_update(qss_ctx *c, int final_pass) {
  SpecArgs a = make_args(c, final_pass);
  if (c->npc) {
    a.nxs = final_pass && !c->skip_rho_psi ? 8 : 6;
    const int ids[8] = {SP_EX, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ, SP_RHO, SP_PSI};
    for (int b = 0; b < a.nxs; ++b) c->xs_src[b] = c->S[ids[b]];
  }
  a.resid = c->d_resid;
  // threads for every mode and every residual slot
  const long n = c->ncell > kResidPerWake ? c->ncell : kResidPerWake;
  hipLaunchKernelGGL(k_iterate, grid1(n * c->nw), dim3(NT), 0, c->stream, a);
  return hipGetLastError();
}

hipError_t launch_residual_final(qss_ctx *c, double *dst) {
  // capturing: dst is the slot of slice 0 and the slice is read at replay
  hipLaunchKernelGGL(k_resid_final, dim3(c->nw), dim3(64), 0, c->stream,
                     c->d_resid, dst, c->nw, (long)c->p.nslices * c->p.nx_split,
                     c->capturing ? c->d_slice : nullptr, c->p.nx_split);
  return hipGetLastError();
}

hipError_t launch_store_hist(qss_ctx *c, const double *src, double *hist) {
  hipLaunchKernelGGL(k_store_hist, grid1(c->ncell * c->nw), dim3(NT), 0, c->stream, src, hist,
                     c->ncell, c->nw, c->p.nslices, c->d_slice);
  return hipGetLastError();
}

hipError_t launch_add_i32(int *p, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_add_i32, dim3(1), dim3(64), 0, s, p, v);
  return hipGetLastError();
}

hipError_t launch_push_field(qss_ctx *c) {
  if (c->Ff)
    hipLaunchKernelGGL(k_push_field<float>, grid1(c->ncell * c->nw), dim3(NT), 0, c->stream, c->P[0],
                       c->P[1], c->P[2], c->P[3], c->P[4], c->P[5], c->C[0], c->C[1],
                       c->C[2], c->C[3], c->C[4], c->C[5], c->Ff, c->ncell, c->nw);
  else
    hipLaunchKernelGGL(k_push_field<double>, grid1(c->ncell * c->nw), dim3(NT), 0, c->stream, c->P[0],
                       c->P[1], c->P[2], c->P[3], c->P[4], c->P[5], c->C[0], c->C[1],
                       c->C[2], c->C[3], c->C[4], c->C[5], c->F, c->ncell, c->nw);
  return hipGetLastError();
}
