// Finite-difference check of qss_adj::push_adjoint (tertia_amd/csrc/push_adj.h)
// on the CPU: random particles on a small periodic grid, random field planes F
// and random deposit adjoints Jbar; the scalar
//   L = sum_events sum_q Jbar_q . deposit_q  (+ Ybar Y_new + Zbar Z_new + Pbar.P1)
// is differentiated by central differences in every input (start fractions,
// momentum, the 24 gathered corner field values) and compared with the
// hand-written reverse.  Prints the worst relative error and exits 1 if any
// accepted case exceeds the tolerance.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "push_adj.h"

using namespace qss_adj;

namespace {

constexpr int NY = 8, NZ = 8;

struct Case {
  Consts c;
  int j, k;
  double y0, z0, P0[3];
  double Fc[4][6];   // corner fields (0,0),(1,0),(0,1),(1,1)
  bool full;
  double Ybar, Zbar, Pbar[3];
};

std::vector<double> JB(4 * NY * NZ);

double jb(int q, int nj, int nk) {
  nj = ((nj % NY) + NY) % NY;
  nk = ((nk % NZ) + NZ) % NZ;
  return JB[(q * NZ + nk) * NY + nj];
}

// forward scalar: replay of the push (kinematics + sub-steps)
double loss(const Case &cs, double y0, double z0, const double P0[3], const double Fc[4][6]) {
  double w[4], wy[4], wz[4];
  cic_weights(y0, z0, w, wy, wz);
  double f[6];
  for (int q = 0; q < 6; ++q)
    f[q] = w[0] * Fc[0][q] + w[1] * Fc[1][q] + w[2] * Fc[2][q] + w[3] * Fc[3][q];
  Kin kn;
  kinematics(cs.c, P0, f, kn);
  Parts pts(kn.s);
  double yt = y0, zt = z0, L = 0;
  int jj = cs.j, kk = cs.k;
  for (int i = 1; i <= kn.nsub; ++i) {
    const double part = pts.next();
    substep(cs.c, cs.j, cs.k, kn.d[0], kn.d[1], part, yt, zt, jj, kk);
    double sw[4], a[4], b[4];
    cic_weights(yt, zt, sw, a, b);
    for (int q = 0; q < 4; ++q)
      for (int cn = 0; cn < 4; ++cn)
        L += jb(q, jj + (cn & 1), kk + (cn >> 1)) * kn.dj[q] * part * sw[cn];
  }
  if (cs.full) {
    L += cs.Ybar * (jj + yt) + cs.Zbar * (kk + zt);
    for (int q = 0; q < 3; ++q) L += cs.Pbar[q] * kn.P1[q];
  }
  return L;
}

}  // namespace

int main(int argc, char **argv) {
  const int ntrial = argc > 1 ? atoi(argv[1]) : 400;
  std::mt19937_64 rng(20261202);
  std::uniform_real_distribution<double> U(-1.0, 1.0), U01(0.0, 1.0);
  for (auto &v : JB) v = U(rng);
  double worst = 0;
  int accepted = 0, multi = 0, rejected = 0;
  for (int t = 0; t < ntrial; ++t) {
    Case cs;
    cs.c.hx = 0.05;
    cs.c.hy = cs.c.hz = (t % 3 == 0) ? 0.01 : 0.1;   // small cells: sub-cycling
    cs.c.weight = 0.25;
    cs.c.hxy = cs.c.hx / cs.c.hy;
    cs.c.hxz = cs.c.hx / cs.c.hz;
    for (int q = 0; q < 3; ++q) cs.c.bext[q] = (t % 5 == 2) ? 1.5 * U(rng) : 0.0;
    cs.c.ny = NY;
    cs.c.nz = NZ;
    cs.j = 2 + (int)(U01(rng) * 4);
    cs.k = 2 + (int)(U01(rng) * 4);
    cs.y0 = 0.05 + 0.9 * U01(rng);
    cs.z0 = 0.05 + 0.9 * U01(rng);
    const double pm = (t % 2) ? 0.3 : 3.0;
    for (int q = 0; q < 3; ++q) cs.P0[q] = pm * U(rng);
    for (int cn = 0; cn < 4; ++cn)
      for (int q = 0; q < 6; ++q) cs.Fc[cn][q] = 2.0 * U(rng);
    cs.full = t % 4 != 1;
    cs.Ybar = U(rng);
    cs.Zbar = U(rng);
    for (int q = 0; q < 3; ++q) cs.Pbar[q] = U(rng);

    AdjOut o;
    auto Fcb = [&](int q, int dj, int dk) { return cs.Fc[dj + 2 * dk][q]; };
    auto JBf = [&](int q, int nj, int nk) { return jb(q, nj, nk); };
    push_adjoint(cs.c, cs.j, cs.k, cs.y0, cs.z0, cs.P0, Fcb, JBf, cs.full, cs.Ybar, cs.Zbar,
                 cs.Pbar, o);
    // analytic gradient vector: y0, z0, P0[3], Fc[4][6] (= w_c fbar_q)
    double ga[5 + 24];
    ga[0] = o.y0; ga[1] = o.z0;
    for (int q = 0; q < 3; ++q) ga[2 + q] = o.P0[q];
    for (int cn = 0; cn < 4; ++cn)
      for (int q = 0; q < 6; ++q) ga[5 + cn * 6 + q] = o.w[cn] * o.f[q];
    // central differences at two step sizes; a case whose two estimates
    // disagree straddles a sub-step cell crossing (the restart quirk makes
    // the path discontinuous there) and is not used
    auto fd = [&](int idx, double h) {
      Case p = cs, m = cs;
      auto bump = [&](Case &x, double s) {
        if (idx == 0) x.y0 += s;
        else if (idx == 1) x.z0 += s;
        else if (idx < 5) x.P0[idx - 2] += s;
        else x.Fc[(idx - 5) / 6][(idx - 5) % 6] += s;
      };
      bump(p, h);
      bump(m, -h);
      return (loss(p, p.y0, p.z0, p.P0, p.Fc) - loss(m, m.y0, m.z0, m.P0, m.Fc)) / (2 * h);
    };
    double gmax = 1e-12;
    for (double g : ga) gmax = fmax(gmax, fabs(g));
    double err = 0;
    bool ok = true;
    for (int idx = 0; idx < 29; ++idx) {
      const double a = fd(idx, 1e-6), b = fd(idx, 5e-7);
      if (fabs(a - b) > 1e-5 * gmax) { ok = false; break; }
      err = fmax(err, fabs(b - ga[idx]) / gmax);
    }
    double f[6];
    for (int q = 0; q < 6; ++q)
      f[q] = o.w[0] * cs.Fc[0][q] + o.w[1] * cs.Fc[1][q] + o.w[2] * cs.Fc[2][q] +
             o.w[3] * cs.Fc[3][q];
    Kin kn;
    kinematics(cs.c, cs.P0, f, kn);
    if (!ok) { ++rejected; continue; }
    ++accepted;
    if (kn.nsub > 1) ++multi;
    worst = fmax(worst, err);
  }
  printf("accepted %d (multi-sub-step %d) rejected %d worst rel err %.3e\n", accepted, multi,
         rejected, worst);
  return (worst < 1e-6 && accepted > ntrial / 2 && multi > 10) ? 0 : 1;
}
