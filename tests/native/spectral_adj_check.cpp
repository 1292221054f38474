// Dot-product check of the spectral-update transposes in
// tertia_amd/csrc/spectral_mode.h: for random modes and random vectors x, y,
//   <M x, y> == <x, M^T y>
// for the Guess and the Iterate maps (M linear), plus a finite-difference
// check that M is linear (M(x + h e_i) - M(x) = h M e_i), which a wrong
// forward restatement of a nonlinear term would fail.  Exits 1 on failure.
#include <cmath>
#include <cstdio>
#include <random>

#include "spectral_mode.h"

using namespace qss_spec;

template <int NI, int NO, class Fwd, class Adj>
double dot_test(const Mode &m, Fwd fwd, Adj adj, std::mt19937_64 &rng) {
  std::uniform_real_distribution<double> U(-1, 1);
  double x[NI], y[NO], Mx[NO], Mty[NI];
  for (auto &v : x) v = U(rng);
  for (auto &v : y) v = U(rng);
  fwd(m, x, Mx);
  adj(m, y, Mty);
  double a = 0, b = 0, s = 0;
  for (int i = 0; i < NO; ++i) { a += Mx[i] * y[i]; s += fabs(Mx[i] * y[i]); }
  for (int i = 0; i < NI; ++i) { b += x[i] * Mty[i]; s += fabs(x[i] * Mty[i]); }
  // linearity: M(2x) = 2 M(x)
  double x2[NI], M2x[NO];
  for (int i = 0; i < NI; ++i) x2[i] = 2 * x[i];
  fwd(m, x2, M2x);
  double lin = 0, ls = 0;
  for (int i = 0; i < NO; ++i) { lin = fmax(lin, fabs(M2x[i] - 2 * Mx[i])); ls = fmax(ls, fabs(Mx[i])); }
  const double e1 = fabs(a - b) / s, e2 = lin / (ls + 1e-300);
  return fmax(e1, e2);
}

int main() {
  std::mt19937_64 rng(7);
  std::uniform_real_distribution<double> U01(0, 1);
  double worst_g = 0, worst_i = 0;
  for (int t = 0; t < 2000; ++t) {
    Mode m;
    const double ly = 4 + 20 * U01(rng), lz = 4 + 20 * U01(rng);
    const int j = (int)(U01(rng) * 512), k = (int)(U01(rng) * 512);
    m.aky = M_PI / ly * (j + 0.5);
    m.akz = M_PI / lz * (k + 0.5);
    m.ak2 = m.aky * m.aky + m.akz * m.akz;
    m.dens = 0.5 + U01(rng);
    m.hx = 0.005 + 0.1 * U01(rng);
    worst_g = fmax(worst_g, dot_test<GI_N, GO_N>(m, guess, guess_adj, rng));
    worst_i = fmax(worst_i, dot_test<II_N, IO_N>(m, iterate, iterate_adj, rng));
  }
  printf("guess %.3e iterate %.3e\n", worst_g, worst_i);
  return (worst_g < 1e-13 && worst_i < 1e-13) ? 0 : 1;
}
