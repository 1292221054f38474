// push_adj.h — one plasma particle's predictor-corrector push and its
// hand-written reverse-mode derivative (the adjoint push).
//
// Forward (the tiled push kernel's arithmetic, CIC shape, with the
// reference's sub-cycle restart, movepart_splitlayer.cpp:932-1512 and
// movepart_layer_CIC.cpp:166-203,449-496):
//   f   = sum_c w_c(y0,z0) F_c                 bilinear gather of (P+C)/2
//   e   = f_E q2m hx/2,  u = 1 - Vp_x,  K = e/u,  H = P0 + K
//   b   = (f_B + B_ext) beta,  beta = q2m hx / (2 g1 u),  g1 = sqrt(1+|H|^2)
//   co  = 2/(1+|b|^2),  p3 = co (H + H x b),  P1 = H + K + p3 x b
//   Vm  = P1/g2,  Vb = (Vm + Vp)/2,  v = 1 - Vb_x
//   dj  = w dj0 Vm / v,  drho = w drho0 / v,  d = (Vb_y hx/hy, Vb_z hx/hz)/v
//   sub-steps: s = max(|d_y|,|d_z|) if > 1 else 1; parts p_i (1/s each, the
//   last (s - n + 1)/s); position after sub-step i: start cell + fraction of
//   the previous sub-step + d p_i; deposit (dj, drho) p_i at each sub-step end.
// Reverse: given the adjoint of the deposited planes (a callback returning
// Jbar_q at a node) and, for the full step, of the new position and
// momentum, it returns the adjoint of the start position and momentum and of
// the six gathered field values (the caller scatters those with the CIC
// weights w_c).  The fraction after sub-step i is y0 + d P_i modulo cell
// shifts (P_i = p_1 + ... + p_i), so the sub-step chain is summed in one
// replay of the sub-steps without storing them.
//
// Header-only and __host__ __device__: the adjoint kernel and a CPU
// finite-difference check compile the same code.
#pragma once

#include <math.h>

#ifdef __HIPCC__
#define QSS_HD __host__ __device__
#else
#define QSS_HD
#endif

namespace qss_adj {

constexpr int kMaxSub = 64;
constexpr double kQ2m = -1.0, kDj0 = -1.0, kDrho0 = -1.0;

struct Consts {
  double hx, hy, hz, weight;
  double hxy, hxz;  // hx / hy, hx / hz
  double bext[3];   // uniform external B, added to the gathered B
  int ny, nz;
};

// CIC weights of corners (0,0), (1,0), (0,1), (1,1) and their derivatives
QSS_HD inline void cic_weights(double y, double z, double w[4], double wy[4], double wz[4]) {
  w[0] = (1. - y) * (1. - z);
  w[1] = y * (1. - z);
  w[2] = (1. - y) * z;
  w[3] = y * z;
  wy[0] = -(1. - z); wy[1] = 1. - z; wy[2] = -z; wy[3] = z;
  wz[0] = -(1. - y); wz[1] = -y; wz[2] = 1. - y; wz[3] = y;
}

// forward intermediates of the momentum update of one particle
struct Kin {
  double P0[3], g0, Vp[3];
  double fE[3], fB[3];
  double e[3], u, K[3], H[3], g1, beta, b[3], co, Q[3], p3[3], R[3], P1[3], g2, Vm[3];
  double Vb[3], v, dj[4], d[2];
  double s;    // sub-step divisor (1: no sub-cycling)
  int nsub;    // number of sub-steps
  bool frozen; // displacement zeroed by the sub-step guard
};

// Momentum update and sub-step control; the expressions and their order
// follow the push kernel (one reciprocal per denominator, then products), so
// a replay reproduces its sub-steps bit for bit.
QSS_HD inline void kinematics(const Consts &c, const double P0[3], const double f[6], Kin &k) {
  const double q2m = kQ2m, hx = c.hx;
  for (int q = 0; q < 3; ++q) k.P0[q] = P0[q];
  const double px = P0[0], py = P0[1], pz = P0[2];
  k.g0 = sqrt(1. + px * px + py * py + pz * pz);
  const double ig0 = 1. / k.g0;
  k.Vp[0] = px * ig0; k.Vp[1] = py * ig0; k.Vp[2] = pz * ig0;
  for (int q = 0; q < 3; ++q) { k.fE[q] = f[q]; k.fB[q] = f[3 + q] + c.bext[q]; }
  for (int q = 0; q < 3; ++q) k.e[q] = f[q] * (q2m * hx / 2.);
  k.u = 1. - k.Vp[0];
  const double iu = 1. / k.u;
  for (int q = 0; q < 3; ++q) k.K[q] = k.e[q] * iu;
  for (int q = 0; q < 3; ++q) k.H[q] = P0[q] + k.K[q];
  k.g1 = sqrt(1. + k.H[0] * k.H[0] + k.H[1] * k.H[1] + k.H[2] * k.H[2]);
  const double gr = 1. / k.g1;
  k.beta = gr * q2m * iu * (0.5 * hx);
  for (int q = 0; q < 3; ++q) k.b[q] = k.fB[q] * k.beta;
  const double *b = k.b, *H = k.H;
  k.co = 2. / (1. + (b[0] * b[0]) + (b[1] * b[1]) + (b[2] * b[2]));
  k.Q[0] = H[1] * b[2] - H[2] * b[1] + H[0];
  k.Q[1] = H[2] * b[0] - H[0] * b[2] + H[1];
  k.Q[2] = H[0] * b[1] - H[1] * b[0] + H[2];
  for (int q = 0; q < 3; ++q) k.p3[q] = k.Q[q] * k.co;
  k.R[0] = k.p3[1] * b[2] - k.p3[2] * b[1];
  k.R[1] = k.p3[2] * b[0] - k.p3[0] * b[2];
  k.R[2] = k.p3[0] * b[1] - k.p3[1] * b[0];
  for (int q = 0; q < 3; ++q) k.P1[q] = H[q] + (k.K[q] + k.R[q]);
  k.g2 = sqrt(1. + k.P1[0] * k.P1[0] + k.P1[1] * k.P1[1] + k.P1[2] * k.P1[2]);
  const double ig2 = 1. / k.g2;
  for (int q = 0; q < 3; ++q) k.Vm[q] = k.P1[q] * ig2;
  for (int q = 0; q < 3; ++q) k.Vb[q] = 0.5 * (k.Vm[q] + k.Vp[q]);
  k.v = 1. - k.Vb[0];
  const double iv = 1. / k.v;
  const double wdj = c.weight * kDj0;
  for (int q = 0; q < 3; ++q) k.dj[q] = wdj * k.Vm[q] * iv;
  k.dj[3] = c.weight * kDrho0 * iv;
  double dy = k.Vb[1] * c.hxy * iv, dz = k.Vb[2] * c.hxz * iv;
  double step = 1., partdx = 0.;
  if (fabs(dy) > 1. || fabs(dz) > 1.) {
    if (fabs(dy) > fabs(dz)) step = partdx = fabs(dy);
    else step = partdx = fabs(dz);
  }
  if (partdx < 1.) partdx = step = 1.;
  k.frozen = !(step < (double)(kMaxSub - 1));
  if (k.frozen) { step = 1.; dy = 0.; dz = 0.; }
  k.s = step;
  k.d[0] = dy;
  k.d[1] = dz;
  // count the sub-steps exactly as the push loop does
  int n = 0;
  double pd = step;
  while (pd > 0.) {
    if (pd > 1.) pd -= 1.;
    else pd = 0.;
    ++n;
  }
  k.nsub = n;
}

// One sub-step position update (the push kernel's wrap and floor).
QSS_HD inline void substep(const Consts &c, int j, int k0, double dy, double dz, double part,
                           double &yt, double &zt, int &jj, int &kk) {
  yt += dy * part + j;
  zt += dz * part + k0;
  while (yt > c.ny) yt -= c.ny;
  while (yt < 0) yt += c.ny;
  while (zt > c.nz) zt -= c.nz;
  while (zt < 0) zt += c.nz;
  jj = (int)yt;
  kk = (int)zt;
  yt -= jj;
  zt -= kk;
  if (jj >= c.ny) jj -= c.ny;
  if (kk >= c.nz) kk -= c.nz;
}

// Sub-step parts in order (replays the push loop's partdx bookkeeping).
struct Parts {
  double step, pd;
  QSS_HD Parts(double s) : step(s), pd(s) {}
  QSS_HD double next() {
    double part;
    if (pd > 1.) { pd -= 1.; part = 1. / step; }
    else { part = pd / step; pd = 0.; }
    return part;
  }
};

struct AdjOut {
  double y0, z0;   // adjoint of the start fractions (= of the global position)
  double P0[3];    // adjoint of the start momentum
  double f[6];     // adjoint of the gathered Ex,Ey,Ez,Bx,By,Bz
  double w[4];     // CIC weights at the start position (for the scatter)
};

// x unchanged, but opaque to the device compiler: values derived from it are
// recomputed rather than kept in registers from an earlier identical
// computation (the adjoint recomputes the kinematics after the sub-step
// replay instead of holding ~35 doubles across it)
QSS_HD inline double fresh(double x) {
#ifdef __HIP_DEVICE_COMPILE__
  __asm__ volatile("" : "+v"(x));
#endif
  return x;
}
QSS_HD inline int fresh(int x) {
#ifdef __HIP_DEVICE_COMPILE__
  __asm__ volatile("" : "+v"(x));
#endif
  return x;
}

// a x b
QSS_HD inline void cross(const double a[3], const double b[3], double r[3]) {
  r[0] = a[1] * b[2] - a[2] * b[1];
  r[1] = a[2] * b[0] - a[0] * b[2];
  r[2] = a[0] * b[1] - a[1] * b[0];
}

// Gathered fields (and, with gy/gz, their derivatives along y and z) at the
// start position; the corner values are not kept.
template <class FN>
QSS_HD inline void gather6(FN F, int o0, const double w[4], const double *wy, const double *wz,
                           double f[6], double *gy, double *gz) {
  for (int q = 0; q < 6; ++q) {
    double Fc[4];
    for (int cn = 0; cn < 4; ++cn) Fc[cn] = F(q, (cn & 1) + o0, (cn >> 1) + o0);
    f[q] = w[0] * Fc[0] + w[1] * Fc[1] + w[2] * Fc[2] + w[3] * Fc[3];
    if (gy) {
      gy[q] = 0;
      gz[q] = 0;
      for (int cn = 0; cn < 4; ++cn) {
        gy[q] += Fc[cn] * wy[cn];
        gz[q] += Fc[cn] * wz[cn];
      }
    }
  }
}

// Reverse of one particle's push.  F(q, dj, dk) returns the gathered field q
// at corner (j+dj, k+dk) of the start cell; JB(q, nj, nk) the deposit adjoint
// of component q (0..3 = Jx, Jy, Jz, rho) at node (nj, nk), indices taken
// periodically by the callee.  full: the new position/momentum adjoints are
// given by NB(i) (i = 0, 1: Ybar, Zbar; 2..4: Pbar), read where they are
// used.  Returns false if the particle's displacement was frozen by the
// sub-step guard (then only the frozen path's derivative is returned).
//
// Registers: the gathers and the kinematics are computed again where they
// are needed (from the same inputs, so bitwise the same values) rather than
// held across the sub-step replay.
template <class FN, class JBN, class NBN>
QSS_HD inline bool push_adjoint_cb(const Consts &c, int j, int k, double y0, double z0,
                                   const double P0[3], FN F, JBN JB, bool full, NBN NB,
                                   AdjOut &o) {
  // what the sub-step replay needs of the kinematics
  int n;
  double dy, dz, s_div, djk[4];
  bool frozen;
  {
    double w[4], wy[4], wz[4], f[6];
    cic_weights(y0, z0, w, wy, wz);
    gather6(F, 0, w, nullptr, nullptr, f, nullptr, nullptr);
    Kin k1;
    kinematics(c, P0, f, k1);
    n = k1.nsub;
    dy = k1.d[0];
    dz = k1.d[1];
    s_div = k1.s;
    frozen = k1.frozen;
    for (int q = 0; q < 4; ++q) djk[q] = k1.dj[q];
  }

  // --- sub-step replay: deposit adjoint and the position chain ---------
  double a_dj[4] = {0, 0, 0, 0};
  double A_y = 0, A_z = 0, PA_y = 0, PA_z = 0, Wy = 0, Wz = 0;
  double pdep_lt = 0, pdep_n = 0, afy_n = 0, afz_n = 0;
  {
    Parts pts(s_div);
    double yt = y0, zt = z0, Pcum = 0;
    int jj = j, kk = k;
    for (int i = 1; i <= n; ++i) {
      const double part = pts.next();
      Pcum += part;
      substep(c, j, k, dy, dz, part, yt, zt, jj, kk);
      double sw[4], swy[4], swz[4];
      cic_weights(yt, zt, sw, swy, swz);
      double a_p = 0, afy = 0, afz = 0;
      for (int q = 0; q < 4; ++q) {
        double G = 0, Gy = 0, Gz = 0;
        for (int cn = 0; cn < 4; ++cn) {
          const double jb = JB(q, jj + (cn & 1), kk + (cn >> 1));
          G += jb * sw[cn];
          Gy += jb * swy[cn];
          Gz += jb * swz[cn];
        }
        a_dj[q] += part * G;
        a_p += djk[q] * G;
        afy += djk[q] * part * Gy;
        afz += djk[q] * part * Gz;
      }
      if (full && i == n) {
        afy += NB(0);
        afz += NB(1);
      }
      A_y += afy;
      A_z += afz;
      PA_y += Pcum * afy;
      PA_z += Pcum * afz;
      const double mi = (double)(i < n - 1 ? i : n - 1);   // min(i, n-1)
      Wy += mi * afy;
      Wz += mi * afz;
      if (i < n) pdep_lt += a_p;
      else { pdep_n = a_p; afy_n = afy; afz_n = afz; }
    }
  }
  double a_d[2] = {PA_y, PA_z};
  if (s_div > 1.) {
    const double s2 = s_div * s_div;
    const double a_s = -(pdep_lt + dy * Wy + dz * Wz) / s2 +
                       (double)(n - 1) / s2 * (pdep_n + dy * afy_n + dz * afz_n);
    if (fabs(dy) > fabs(dz)) a_d[0] += (dy > 0 ? 1. : -1.) * a_s;
    else a_d[1] += (dz > 0 ? 1. : -1.) * a_s;
  }
  if (frozen) { a_d[0] = 0.; a_d[1] = 0.; }

  // --- kinematics reverse -------------------------------------------------
  // the same kinematics again, from inputs the compiler cannot match with the
  // first call's (bitwise the same values)
  Kin kn;
  {
    const double yr = fresh(y0), zr = fresh(z0);
    double Pr[3], w[4], wy[4], wz[4], f[6];
    for (int q = 0; q < 3; ++q) Pr[q] = fresh(P0[q]);
    cic_weights(yr, zr, w, wy, wz);
    gather6(F, fresh(0), w, nullptr, nullptr, f, nullptr, nullptr);
    kinematics(c, Pr, f, kn);
  }
  const double cy = c.hx / c.hy, cz = c.hx / c.hz, v = kn.v;
  const double wdj = c.weight * kDj0;
  double a_v = 0;
  if (!kn.frozen) a_v -= (kn.d[0] * a_d[0] + kn.d[1] * a_d[1]) / v;
  for (int q = 0; q < 4; ++q) a_v -= kn.dj[q] * a_dj[q] / v;
  double a_Vm[3], a_Vb[3], a_Vp[3];
  for (int q = 0; q < 3; ++q) a_Vm[q] = wdj / v * a_dj[q];
  a_Vb[0] = -a_v;
  a_Vb[1] = kn.frozen ? 0. : cy / v * a_d[0];
  a_Vb[2] = kn.frozen ? 0. : cz / v * a_d[1];
  for (int q = 0; q < 3; ++q) {
    a_Vm[q] += 0.5 * a_Vb[q];
    a_Vp[q] = 0.5 * a_Vb[q];
  }
  double a_P1[3];
  {
    const double *P1 = kn.P1, g2 = kn.g2;
    const double dot = P1[0] * a_Vm[0] + P1[1] * a_Vm[1] + P1[2] * a_Vm[2];
    for (int q = 0; q < 3; ++q)
      a_P1[q] = a_Vm[q] / g2 - P1[q] * dot / (g2 * g2 * g2) + (full ? NB(2 + q) : 0.);
  }
  double a_H[3], a_K[3], a_R[3], a_p3[3], a_b[3], a_Q[3];
  for (int q = 0; q < 3; ++q) { a_H[q] = a_P1[q]; a_K[q] = a_P1[q]; a_R[q] = a_P1[q]; }
  cross(kn.b, a_R, a_p3);           // R = p3 x b
  cross(a_R, kn.p3, a_b);
  const double a_co = kn.Q[0] * a_p3[0] + kn.Q[1] * a_p3[1] + kn.Q[2] * a_p3[2];
  for (int q = 0; q < 3; ++q) a_Q[q] = kn.co * a_p3[q];
  {
    double t1[3], t2[3];
    cross(kn.b, a_Q, t1);           // Q = H + H x b
    cross(a_Q, kn.H, t2);
    for (int q = 0; q < 3; ++q) {
      a_H[q] += a_Q[q] + t1[q];
      a_b[q] += t2[q] - kn.co * kn.co * kn.b[q] * a_co;
    }
  }
  double a_fB[3], a_beta = 0;
  for (int q = 0; q < 3; ++q) {
    a_fB[q] = kn.beta * a_b[q];
    a_beta += kn.fB[q] * a_b[q];
  }
  const double a_g1 = -kn.beta / kn.g1 * a_beta;
  double a_u = -kn.beta / kn.u * a_beta;
  for (int q = 0; q < 3; ++q) a_H[q] += kn.H[q] / kn.g1 * a_g1;
  double a_P0[3];
  for (int q = 0; q < 3; ++q) { a_P0[q] = a_H[q]; a_K[q] += a_H[q]; }
  double a_e[3];
  for (int q = 0; q < 3; ++q) {
    a_e[q] = a_K[q] / kn.u;
    a_u -= kn.e[q] * a_K[q] / (kn.u * kn.u);
  }
  a_Vp[0] -= a_u;
  {
    const double g0 = kn.g0;
    const double dot = P0[0] * a_Vp[0] + P0[1] * a_Vp[1] + P0[2] * a_Vp[2];
    for (int q = 0; q < 3; ++q) a_P0[q] += a_Vp[q] / g0 - P0[q] * dot / (g0 * g0 * g0);
  }
  const double c1 = kQ2m * c.hx / 2.;
  for (int q = 0; q < 3; ++q) {
    o.f[q] = c1 * a_e[q];
    o.f[3 + q] = a_fB[q];
  }
  // gather weights
  double ay = A_y, az = A_z;
  {
    const double yr = fresh(y0), zr = fresh(z0);
    double w[4], wy[4], wz[4], f[6], gy[6], gz[6];
    cic_weights(yr, zr, w, wy, wz);
    gather6(F, fresh(0), w, wy, wz, f, gy, gz);
    for (int q = 0; q < 6; ++q) {
      ay += o.f[q] * gy[q];
      az += o.f[q] * gz[q];
    }
    for (int cn = 0; cn < 4; ++cn) o.w[cn] = w[cn];
  }
  o.y0 = ay;
  o.z0 = az;
  for (int q = 0; q < 3; ++q) o.P0[q] = a_P0[q];
  return !kn.frozen;
}

// The same with the final-state adjoints given as values.
template <class FN, class JBN>
QSS_HD inline bool push_adjoint(const Consts &c, int j, int k, double y0, double z0,
                                const double P0[3], FN F, JBN JB, bool full, double Ybar,
                                double Zbar, const double Pbar[3], AdjOut &o) {
  auto nb = [&](int i) { return i == 0 ? Ybar : i == 1 ? Zbar : Pbar[i - 2]; };
  return push_adjoint_cb(c, j, k, y0, z0, P0, F, JB, full, nb, o);
}

}  // namespace qss_adj
