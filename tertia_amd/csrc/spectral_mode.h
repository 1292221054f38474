// spectral_mode.h — the per-mode predictor (Guess) and corrector (Iterate)
// maps of the spectral field solve and their transposes.
//
// Forward arithmetic term by term as the reference CPU loops:
//   Guess   movefieldshydrolinlayersplit.cpp:807-901 (GPU twin
//           CUDA_WRAP/linearized.cu:11-92)
//   Iterate movefieldshydrolinlayersplit.cpp:1373-1478 (GPU twin
//           CUDA_WRAP/linearized.cu:105-208)
// Both maps are linear in their inputs with coefficients fixed by the mode
// (aky, akz), dens and hx, so the adjoint of each is its transpose; *_adj
// returns in_bar = M^T out_bar.  Header-only and __host__ __device__: the
// GPU kernels and a CPU dot-product check use the same code.
#pragma once

#ifdef __HIPCC__
#define QSS_HD __host__ __device__
#else
#define QSS_HD
#endif

namespace qss_spec {

struct Mode {
  double aky, akz, ak2, dens, hx;
};

// Guess inputs: the previous layer's spectra and the beam spectra
enum { GI_RP, GI_JXP, GI_JYP, GI_JZP, GI_EXP, GI_EYP, GI_EZP, GI_RB, GI_RBP, GI_JXB, GI_N };
// Guess outputs
enum { GO_RHO, GO_EX, GO_EY, GO_EZ, GO_JX, GO_JY, GO_JZ, GO_BX, GO_BY, GO_BZ, GO_N };

QSS_HD inline void guess(const Mode &m, const double *in, double *out) {
  const double dens = m.dens, hx = m.hx, aky = m.aky, akz = m.akz, ak2 = m.ak2;
  const double rp = in[GI_RP], jxp = in[GI_JXP], jyp = in[GI_JYP], jzp = in[GI_JZP];
  const double exp_ = in[GI_EXP], eyp = in[GI_EYP], ezp = in[GI_EZP];
  const double rb = in[GI_RB], rbp = in[GI_RBP], jxb = in[GI_JXB];
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
  out[GO_RHO] = rho;
  out[GO_EX] = ex;
  out[GO_EY] = ey;
  out[GO_EZ] = ez;
  out[GO_JX] = jx;
  out[GO_JY] = jy;
  out[GO_JZ] = jz;
  out[GO_BX] = -aky / ak2 * jz + akz / ak2 * jy;
  out[GO_BY] = (-akz * (jx + jxb) + dens * ez) / ak2;
  out[GO_BZ] = (aky * (jx + jxb) - dens * ey) / ak2;
}

QSS_HD inline void guess_adj(const Mode &m, const double *ob, double *ib) {
  const double dens = m.dens, hx = m.hx, aky = m.aky, akz = m.akz, ak2 = m.ak2;
  const double D = 4. + dens * hx * hx;
  const double prop = (4. - dens * hx * hx) / D;
  for (int i = 0; i < GI_N; ++i) ib[i] = 0.;
  // B from (jx, jy, jz, ez, ey, jxb)
  double a_jx = ob[GO_JX], a_jy = ob[GO_JY], a_jz = ob[GO_JZ];
  double a_ey = ob[GO_EY], a_ez = ob[GO_EZ], a_ex = ob[GO_EX], a_rho = ob[GO_RHO];
  a_jz += -aky / ak2 * ob[GO_BX];
  a_jy += akz / ak2 * ob[GO_BX];
  a_jx += -akz / ak2 * ob[GO_BY];
  ib[GI_JXB] += -akz / ak2 * ob[GO_BY];
  a_ez += dens / ak2 * ob[GO_BY];
  a_jx += aky / ak2 * ob[GO_BZ];
  ib[GI_JXB] += aky / ak2 * ob[GO_BZ];
  a_ey += -dens / ak2 * ob[GO_BZ];
  // J = J_P + hx dens (E + E_P)/2
  const double hd = hx * dens / 2.;
  ib[GI_JXP] += a_jx;
  a_ex += hd * a_jx;
  ib[GI_EXP] += hd * a_jx;
  ib[GI_JYP] += a_jy;
  a_ey += hd * a_jy;
  ib[GI_EYP] += hd * a_jy;
  ib[GI_JZP] += a_jz;
  a_ez += hd * a_jz;
  ib[GI_EZP] += hd * a_jz;
  // Ey, Ez = -E_P + ak (rb + rbp + rp + rho)/(dens + k^2)
  ib[GI_EYP] += -a_ey;
  ib[GI_EZP] += -a_ez;
  const double a_sum = (aky * a_ey + akz * a_ez) / (dens + ak2);
  ib[GI_RB] += a_sum;
  ib[GI_RBP] += a_sum;
  ib[GI_RP] += a_sum;
  a_rho += a_sum;
  // Ex
  const double cx = -2. * dens * hx / ((ak2 + dens) * D);
  ib[GI_EXP] += prop * a_ex;
  ib[GI_RP] += 2. * cx * a_ex;
  ib[GI_RB] += cx * a_ex;
  ib[GI_RBP] += cx * a_ex;
  // rho
  ib[GI_RP] += prop * a_rho;
  ib[GI_RB] += -dens * hx * hx / D * a_rho;
  ib[GI_RBP] += -dens * hx * hx / D * a_rho;
  ib[GI_EXP] += 4. * hx * (ak2 + dens) / D * a_rho;
}

// Iterate inputs
enum {
  II_RP, II_JX, II_JY, II_JZ, II_JXP, II_JYP, II_JZP, II_RB, II_RBP, II_EYP, II_EZP,
  II_RHO_OLD, II_JXB, II_N
};
// Iterate outputs
enum { IO_RHO, IO_EX, IO_EY, IO_EZ, IO_BX, IO_BY, IO_BZ, IO_PSI, IO_N };

QSS_HD inline void iterate(const Mode &m, const double *in, double *out) {
  const double dens = m.dens, hx = m.hx, aky = m.aky, akz = m.akz, ak2 = m.ak2;
  const double rp = in[II_RP], jx = in[II_JX], jy = in[II_JY], jz = in[II_JZ];
  const double rho = rp + jx - in[II_JXP] -
                     hx * (aky * (jy + in[II_JYP]) + akz * (jz + in[II_JZP])) / 2.;
  const double rb = in[II_RB], rbp = in[II_RBP];
  const double eyp = in[II_EYP], ezp = in[II_EZP];
  const double rho_old = in[II_RHO_OLD];
  const double newEx = -(aky * jy + akz * jz) / ak2;
  const double newEy = -eyp + aky * (rb + rbp + rp + rho_old) / (dens + ak2);
  const double newEz = -ezp + akz * (rb + rbp + rp + rho_old) / (dens + ak2);
  const double jxb = in[II_JXB];
  out[IO_EY] = newEy;
  out[IO_EZ] = newEz;
  out[IO_RHO] = rho;
  out[IO_EX] = newEx;
  out[IO_BX] = -aky / ak2 * jz + akz / ak2 * jy;
  out[IO_BY] = (-akz * (jx + jxb) + dens * newEz) / ak2;
  out[IO_BZ] = (aky * (jx + jxb) - dens * newEy) / ak2;
  out[IO_PSI] = (rho - jx) / ak2;
}

QSS_HD inline void iterate_adj(const Mode &m, const double *ob, double *ib) {
  const double dens = m.dens, hx = m.hx, aky = m.aky, akz = m.akz, ak2 = m.ak2;
  for (int i = 0; i < II_N; ++i) ib[i] = 0.;
  double a_rho = ob[IO_RHO], a_ey = ob[IO_EY], a_ez = ob[IO_EZ];
  // psi
  a_rho += ob[IO_PSI] / ak2;
  ib[II_JX] += -ob[IO_PSI] / ak2;
  // B
  ib[II_JZ] += -aky / ak2 * ob[IO_BX];
  ib[II_JY] += akz / ak2 * ob[IO_BX];
  ib[II_JX] += -akz / ak2 * ob[IO_BY];
  ib[II_JXB] += -akz / ak2 * ob[IO_BY];
  a_ez += dens / ak2 * ob[IO_BY];
  ib[II_JX] += aky / ak2 * ob[IO_BZ];
  ib[II_JXB] += aky / ak2 * ob[IO_BZ];
  a_ey += -dens / ak2 * ob[IO_BZ];
  // Ey, Ez
  ib[II_EYP] += -a_ey;
  ib[II_EZP] += -a_ez;
  const double a_sum = (aky * a_ey + akz * a_ez) / (dens + ak2);
  ib[II_RB] += a_sum;
  ib[II_RBP] += a_sum;
  ib[II_RP] += a_sum;
  ib[II_RHO_OLD] += a_sum;
  // Ex
  ib[II_JY] += -aky / ak2 * ob[IO_EX];
  ib[II_JZ] += -akz / ak2 * ob[IO_EX];
  // continuity rho
  ib[II_RP] += a_rho;
  ib[II_JX] += a_rho;
  ib[II_JXP] += -a_rho;
  ib[II_JY] += -hx * aky / 2. * a_rho;
  ib[II_JYP] += -hx * aky / 2. * a_rho;
  ib[II_JZ] += -hx * akz / 2. * a_rho;
  ib[II_JZP] += -hx * akz / 2. * a_rho;
}

}  // namespace qss_spec
