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

struct SpecArgs {
  double *S[SP_N];
  int ny, nz;
  double ly, lz, hx, dens;
  int flag;  // iterate: final pass
};

__device__ inline void wavenumbers(const SpecArgs &a, int j, int k, double &aky,
                                   double &akz, double &ak2) {
  akz = M_PI / a.lz * (k + 0.5);
  aky = M_PI / a.ly * (j + 0.5);
  ak2 = aky * aky + akz * akz;
}

__global__ __launch_bounds__(NT) void k_guess(SpecArgs a) {
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;
  if (n1 >= (long)a.ny * a.nz) return;
  const int j = (int)(n1 % a.ny), k = (int)(n1 / a.ny);
  double aky, akz, ak2;
  wavenumbers(a, j, k, aky, akz, ak2);
  double **S = a.S;
  const double dens = a.dens, hx = a.hx;
  // P spectra = the spectra left by the previous Iterate (or zero after a
  // reset); the beam P spectrum is the previous beam spectrum (:637-646)
  const double rp = S[SP_RHO][n1];
  const double jxp = S[SP_JX][n1];
  const double jyp = S[SP_JY][n1];
  const double jzp = S[SP_JZ][n1];
  const double exp_ = S[SP_EX][n1];
  const double eyp = S[SP_EY][n1];
  const double ezp = S[SP_EZ][n1];
  S[SP_RHOP][n1] = rp;
  S[SP_JXP][n1] = jxp;
  S[SP_JYP][n1] = jyp;
  S[SP_JZP][n1] = jzp;
  S[SP_EXP][n1] = exp_;
  S[SP_EYP][n1] = eyp;
  S[SP_EZP][n1] = ezp;
  const double rb = S[SP_RHOB][n1];
  const double rbp = S[SP_RHOBP][n1];
  const double jxb = S[SP_JXB][n1];
  const double propagator = (4. - dens * hx * hx) / (4. + dens * hx * hx);
  const double denominator = 4. + dens * hx * hx;
  const double rho = propagator * rp - (rb + rbp) * dens * hx * hx / denominator +
                     4. * exp_ * hx * (ak2 + dens) / denominator;
  const double ex = propagator * exp_ -
                    2. * (2. * rp + rb + rbp) * dens * hx / ((ak2 + dens) * denominator);
  const double ey = -eyp + aky * (rb + rbp + rp + rho) / (dens + ak2);
  const double ez = -ezp + akz * (rb + rbp + rp + rho) / (dens + ak2);
  const double jy = jyp + hx * dens * (ey + eyp) / 2.;
  const double jz = jzp + hx * dens * (ez + ezp) / 2.;
  const double jx = jxp + hx * dens * (ex + exp_) / 2.;
  S[SP_RHO][n1] = rho;
  S[SP_EX][n1] = ex;
  S[SP_EY][n1] = ey;
  S[SP_EZ][n1] = ez;
  S[SP_JX][n1] = jx;
  S[SP_JY][n1] = jy;
  S[SP_JZ][n1] = jz;
  S[SP_BX][n1] = -aky / ak2 * jz + akz / ak2 * jy;
  S[SP_BY][n1] = (-akz * (jx + jxb) + dens * ez) / ak2;
  S[SP_BZ][n1] = (aky * (jx + jxb) - dens * ey) / ak2;
}

__global__ __launch_bounds__(NT) void k_iterate(SpecArgs a) {
  const long n1 = (long)blockIdx.x * NT + threadIdx.x;
  if (n1 >= (long)a.ny * a.nz) return;
  const int j = (int)(n1 % a.ny), k = (int)(n1 / a.ny);
  double aky, akz, ak2;
  wavenumbers(a, j, k, aky, akz, ak2);
  double **S = a.S;
  const double dens = a.dens, hx = a.hx;
  const double rp = S[SP_RHOP][n1];
  const double jx = S[SP_JX][n1];
  const double jy = S[SP_JY][n1];
  const double jz = S[SP_JZ][n1];
  const double rho = rp + jx - S[SP_JXP][n1] -
                     hx * (aky * (jy + S[SP_JYP][n1]) + akz * (jz + S[SP_JZP][n1])) / 2.;
  const double rb = S[SP_RHOB][n1];
  const double rbp = S[SP_RHOBP][n1];
  const double eyp = S[SP_EYP][n1];
  const double ezp = S[SP_EZP][n1];
  const double rho_old = S[SP_RHO][n1];
  const double newEx = -(aky * jy + akz * jz) / ak2;
  const double newEy = -eyp + aky * (rb + rbp + rp + rho_old) / (dens + ak2);
  const double newEz = -ezp + akz * (rb + rbp + rp + rho_old) / (dens + ak2);
  const double jxb = S[SP_JXB][n1];
  S[SP_EY][n1] = newEy;
  S[SP_EZ][n1] = newEz;
  S[SP_RHO][n1] = rho;
  S[SP_EX][n1] = newEx;
  S[SP_BX][n1] = -aky / ak2 * jz + akz / ak2 * jy;
  S[SP_BY][n1] = (-akz * (jx + jxb) + dens * newEz) / ak2;
  S[SP_BZ][n1] = (aky * (jx + jxb) - dens * newEy) / ak2;
  if (a.flag) S[SP_PSI][n1] = (rho - jx) / ak2;
}

// Eerr = max_c dif_c / max(1e-5, max_c)  (:1541, :1584-1621, :1641-1643)
__global__ void k_resid_final(const unsigned long long *r, double *dst) {
  if (threadIdx.x != 0) return;
  double err = 0.0;
  for (int c = 0; c < 3; ++c) {
    double dif = __longlong_as_double((long long)r[c]);
    double mx = __longlong_as_double((long long)r[3 + c]);
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
                                                  FT *F, long n) {
  const long i = (long)blockIdx.x * NT + threadIdx.x;
  if (i >= n) return;
  F[i] = (FT)(0.5 * (p0[i] + c0[i]));
  F[n + i] = (FT)(0.5 * (p1[i] + c1[i]));
  F[2 * n + i] = (FT)(0.5 * (p2[i] + c2[i]));
  F[3 * n + i] = (FT)(0.5 * (p3[i] + c3[i]));
  F[4 * n + i] = (FT)(0.5 * (p4[i] + c4[i]));
  F[5 * n + i] = (FT)(0.5 * (p5[i] + c5[i]));
}

SpecArgs make_args(qss_ctx *c, int flag) {
  SpecArgs a;
  for (int s = 0; s < SP_N; ++s) a.S[s] = c->S[s];
  a.ny = c->ny; a.nz = c->nz;
  a.ly = c->p.ly; a.lz = c->p.lz;
  a.hx = c->hxs; a.dens = c->p.dens;
  a.flag = flag;
  return a;
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }

}  // namespace

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
hipError_t launch_fill_u64(unsigned long long *p, long n, unsigned long long v,
                           hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_u64, grid1(n), dim3(NT), 0, s, p, n, v);
  return hipGetLastError();
}

hipError_t launch_guess_update(qss_ctx *c, int first_layer) {
  hipLaunchKernelGGL(k_guess, grid1(c->ncell), dim3(NT), 0, c->stream,
                     make_args(c, first_layer));
  return hipGetLastError();
}

hipError_t launch_iterate_update(qss_ctx *c, int final_pass) {
  hipLaunchKernelGGL(k_iterate, grid1(c->ncell), dim3(NT), 0, c->stream,
                     make_args(c, final_pass));
  return hipGetLastError();
}

hipError_t launch_residual_final(qss_ctx *c, double *dst) {
  hipLaunchKernelGGL(k_resid_final, dim3(1), dim3(64), 0, c->stream, c->d_resid, dst);
  return hipGetLastError();
}

hipError_t launch_push_field(qss_ctx *c) {
  if (c->Ff)
    hipLaunchKernelGGL(k_push_field<float>, grid1(c->ncell), dim3(NT), 0, c->stream, c->P[0],
                       c->P[1], c->P[2], c->P[3], c->P[4], c->P[5], c->C[0], c->C[1],
                       c->C[2], c->C[3], c->C[4], c->C[5], c->Ff, c->ncell);
  else
    hipLaunchKernelGGL(k_push_field<double>, grid1(c->ncell), dim3(NT), 0, c->stream, c->P[0],
                       c->P[1], c->P[2], c->P[3], c->P[4], c->P[5], c->C[0], c->C[1],
                       c->C[2], c->C[3], c->C[4], c->C[5], c->F, c->ncell);
  return hipGetLastError();
}
