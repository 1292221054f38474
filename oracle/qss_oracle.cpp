// qss_oracle.cpp — FP64 CPU restatement of the reference quasi-static
// xi-slice sweep (snytav/tertia CPU/FFTW path).
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: it is loaded by the
// tests, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg.  The
// product (tertia_amd/libqss.so) never links or calls it.
//
// What it restates (reference file:line, paths relative to the reference):
//   sweep driver          movepart_splitlayer.cpp:25-558   (oracle_sweep_range)
//   one layer step        movepart_splitlayer.cpp:561-778  (oracle_layer_step)
//   Guess (predictor)     movefieldshydrolinlayersplit.cpp:222-1169
//   Iterate (corrector)   movefieldshydrolinlayersplit.cpp:1173-1671
//   transforms            FFTW REDFT11/RODFT11 r2r 2-D, plans :389-427
//   push + deposit        movepart_splitlayer.cpp:781-1544, 1547-1655
//   CIC gather / deposit  movepart_layer_CIC.cpp:166-203, 449-496
//   ghost exchange        movepart_splitlayer.cpp:1659-1711 (periodic)
//   plasma seeding        mesh.cpp:98-105, 413-548; elements.cpp:97-107
//   beam seeding/deposit  mesh.cpp:108-276; plasma.cpp:350-377;
//                         movebeam.cpp:509-512, 662-770
//   beam interpolation    movepart_splitlayer.cpp:330-348
//   evolving driver       mesh.cpp:43-64 (seeding range), movebeam.cpp:14-660
//                         (push), 662-770 (deposit), step.cpp:63,105 (step)
//
// Parity unpinned: the absolute prefactor of the driver's radiation reaction
// (movebeam.cpp:448-501: 4/9/137, 1.2398e-4/Wavelength, 3/(4 PI), 5.11e5);
// its direction (along -p), its absence for dp parallel to p and its
// scaling with dp_perp^2, 1+p^2, 1/ts and 1/Wavelength are pinned
// (tests/test_oracle_driver.py), the constant itself only by this text.
//
// Build conventions that differ from the reference are listed at the
// function that makes them.  Compile with -ffp-contract=off: the oracle's
// push/deposit arithmetic is evaluated in the reference's operation order.

#include <cmath>
#include <complex>
#include <cstdint>
#include <cstring>
#include <vector>
#include <algorithm>
#include <numeric>
#include <string>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/qss.h"

namespace {

typedef std::complex<double> cplx;

// ------------------------------------------------------------------------
// Counter-based RNG (splitmix64) for jitter and beam noise.  Identical
// integer arithmetic is used on the GPU, so both draw the same numbers.
// ------------------------------------------------------------------------
inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
// uniform in [-1, 1): 53 random bits
inline double unif_pm1(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed ^ mix64(a ^ mix64(b)));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// ------------------------------------------------------------------------
// 1-D half-integer transforms (FFTW REDFT11 / RODFT11 definitions):
//   REDFT11: Y_k = 2 sum_j X_j cos(pi (j+1/2)(k+1/2) / N)
//   RODFT11: Y_k = 2 sum_j X_j sin(pi (j+1/2)(k+1/2) / N)
// Two implementations: the dense matrix (the reference's `HalfInteger`
// table form, CUDA_WRAP/fft_matrix.cu:10-24) and an N/2-point complex FFT
// form used for the CPU speed baseline.
// ------------------------------------------------------------------------
struct Plan1D {
  int n = 0;
  std::vector<double> tc, ts;         // n x n tables 2cos / 2sin
  // FFT form (n even, n/2 power of two)
  bool fft_ok = false;
  int m = 0;
  std::vector<cplx> pre, post, tw;    // twiddles
  std::vector<int> rev;

  void init(int n_) {
    n = n_;
    tc.assign((size_t)n * n, 0.0);
    ts.assign((size_t)n * n, 0.0);
    // angle pi(j+1/2)(k+1/2)/n = 2*pi*m/(8n) with m = (2j+1)(2k+1) mod 8n,
    // reduced exactly in integers so every entry is accurate to ~1 ulp
    for (int j = 0; j < n; ++j)
      for (int k = 0; k < n; ++k) {
        long m = ((long)(2 * j + 1) * (2 * k + 1)) % (8L * n);
        double a = M_PI * (double)m / (4.0 * n);
        tc[(size_t)j * n + k] = 2.0 * std::cos(a);
        ts[(size_t)j * n + k] = 2.0 * std::sin(a);
      }
    fft_ok = false;
    if (n >= 4 && n % 2 == 0) {
      m = n / 2;
      if ((m & (m - 1)) == 0) {
        fft_ok = true;
        pre.resize(m); post.resize(m); tw.resize(m / 2 > 0 ? m / 2 : 1);
        for (int q = 0; q < m; ++q) {
          pre[q] = std::polar(1.0, -M_PI * q / n);
          post[q] = std::polar(1.0, -M_PI * (4.0 * q + 1.0) / (4.0 * n));
        }
        for (int q = 0; q < m / 2; ++q) tw[q] = std::polar(1.0, -2.0 * M_PI * q / m);
        rev.resize(m);
        int lg = 0; while ((1 << lg) < m) ++lg;
        for (int q = 0; q < m; ++q) {
          int r = 0;
          for (int b = 0; b < lg; ++b) if (q & (1 << b)) r |= 1 << (lg - 1 - b);
          rev[q] = r;
        }
      }
    }
  }

  // in-place radix-2 DIT forward FFT of length m
  void fft(cplx *a) const {
    for (int q = 0; q < m; ++q) if (q < rev[q]) std::swap(a[q], a[rev[q]]);
    for (int len = 2; len <= m; len <<= 1) {
      int half = len >> 1, step = m / len;
      for (int s = 0; s < m; s += len)
        for (int q = 0; q < half; ++q) {
          cplx w = tw[q * step];
          cplx u = a[s + q], v = a[s + q + half] * w;
          a[s + q] = u + v;
          a[s + q + half] = u - v;
        }
    }
  }

  // strided 1-D transform: in[i*stride], out[i*stride]
  void apply(int kind, const double *in, long istride, double *out,
             long ostride, bool use_fft, cplx *work, double *xbuf) const {
    if (use_fft && fft_ok) {
      // DST-IV(x)_k = (-1)^k DCT-IV(reverse x)_k
      for (int j = 0; j < n; ++j)
        xbuf[j] = (kind == QSS_KIND_C) ? in[j * istride] : in[(n - 1 - j) * istride];
      // DCT-IV via N/2 complex FFT:
      //   c_q = (x_{2q} + i x_{N-1-2q}) e^{-i pi q/N};  Z = FFT_M(c);
      //   Z_m *= e^{-i pi (4m+1)/(4N)};  Y_{2m} = Re Z_m, Y_{N-1-2m} = -Im Z_m
      for (int q = 0; q < m; ++q)
        work[q] = cplx(xbuf[2 * q], xbuf[n - 1 - 2 * q]) * pre[q];
      fft(work);
      for (int q = 0; q < m; ++q) {
        cplx z = work[q] * post[q];
        xbuf[2 * q] = 2.0 * z.real();
        xbuf[n - 1 - 2 * q] = -2.0 * z.imag();
      }
      for (int k = 0; k < n; ++k) {
        double v = xbuf[k];
        if (kind == QSS_KIND_S && (k & 1)) v = -v;
        out[k * ostride] = v;
      }
      return;
    }
    const double *t = (kind == QSS_KIND_C) ? tc.data() : ts.data();
    for (int k = 0; k < n; ++k) {
      double s = 0.0;
      for (int j = 0; j < n; ++j) s += in[j * istride] * t[(size_t)j * n + k];
      xbuf[k] = s;
    }
    for (int k = 0; k < n; ++k) out[k * ostride] = xbuf[k];
  }
};

// 2-D: out[k][j] = scale * sum T_kz(k,k') T_ky(j,j') in[k'][j'], Nz x Ny,
// y fastest (fftw_plan_r2r_2d(nz, ny, ..., kind_z, kind_y)).
void transform2d(const Plan1D &py, const Plan1D &pz, int kz, int ky,
                 const double *in, double *out, double scale, bool use_fft,
                 int nthreads) {
  const int ny = py.n, nz = pz.n;
  std::vector<double> tmp((size_t)ny * nz);
  int nmax = std::max(ny, nz);
#pragma omp parallel num_threads(nthreads)
  {
    std::vector<cplx> work(nmax);
    std::vector<double> xbuf(nmax);
#pragma omp for schedule(static)
    for (int k = 0; k < nz; ++k)
      py.apply(ky, in + (size_t)k * ny, 1, tmp.data() + (size_t)k * ny, 1,
               use_fft, work.data(), xbuf.data());
#pragma omp for schedule(static)
    for (int j = 0; j < ny; ++j)
      pz.apply(kz, tmp.data() + j, ny, out + j, ny, use_fft, work.data(),
               xbuf.data());
  }
  if (scale != 1.0)
    for (size_t i = 0; i < (size_t)ny * nz; ++i) out[i] *= scale;
}

enum { SP_RHO = 0, SP_EX, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ, SP_JX, SP_JY,
       SP_JZ, SP_RHOP, SP_JXP, SP_JYP, SP_JZP, SP_EXP, SP_EYP, SP_EZP,
       SP_RHOB, SP_JXB, SP_RHOBP, SP_JXBP, SP_N };

}  // namespace

struct oracle {
  qss_params p;
  std::vector<double> bins;
  int ny = 0, nz = 0;
  long ncell = 0;
  int nthreads = 1;
  bool use_fft = false;
  Plan1D ply, plz;
  double hy = 0, hz = 0;
  double pweight = 0;                  // plasma macro-particle weight
  // layers: C (current) and P (previous), fields QSS_EX..QSS_JXBEAM
  std::vector<double> C[QSS_NFIELD], P[QSS_NFIELD];
  std::vector<double> S[SP_N];         // spectral state (fft_of_*)
  // plasma particles (SoA), sorted by (cell, previous slot)
  std::vector<double> y, z, px, py, pz;
  std::vector<int> j, k;
  std::vector<int64_t> id;
  // history: field f of slice s
  uint32_t hmask = 0;
  std::vector<double> hist[QSS_NFIELD];
  double last_residual = 0;
  // evolving driver: beam particles in the 3-D cell grid and the 3-D
  // deposit of nodes (layer 0..L, j, k)
  bool driver = false;
  qss_driver drv{};
  std::vector<int> bi, bj, bk, bsp;
  std::vector<double> bx, by, bz, bpx, bpy, bpz, bw;
  std::vector<double> brho3, bjx3;
};

namespace {

inline double &at(std::vector<double> &v, long n) { return v[n]; }

// --------------------------------------------------------------------------
// Plasma seeding: Mesh::SeedParticles for the front layer (mesh.cpp:413-548)
// with ppc = ppc_y*ppc_z.  ppc 2x2 gives the reference's nseed==4 lattice
// y,z in {1/3,2/3} (mesh.cpp:455-459); weight = (1/ppc)*n0 (elements.cpp:102
// and mesh.cpp:468); id = cell*ppc + iseed with iy = iseed%ppc_y.  Build
// addition: +-jitter uniform position noise (cell units) from the seed.
// --------------------------------------------------------------------------
void seed_plasma(oracle *o) {
  const qss_params &p = o->p;
  int ppc = p.ppc_y * p.ppc_z;
  long np = o->ncell * ppc;
  o->y.resize(np); o->z.resize(np); o->px.assign(np, 0.0); o->py.assign(np, 0.0);
  o->pz.assign(np, 0.0); o->j.resize(np); o->k.resize(np); o->id.resize(np);
  o->pweight = (1.0 / ppc) * p.dens;
  long n = 0;
  for (int kk = 0; kk < o->nz; ++kk)
    for (int jj = 0; jj < o->ny; ++jj)
      for (int is = 0; is < ppc; ++is, ++n) {
        int iy = is % p.ppc_y, iz = is / p.ppc_y;
        int64_t pid = ((int64_t)kk * o->ny + jj) * ppc + is;
        double yy = (iy + 1.0) / (p.ppc_y + 1.0);
        double zz = (iz + 1.0) / (p.ppc_z + 1.0);
        if (p.jitter != 0.0) {
          yy += p.jitter * unif_pm1(p.seed, (uint64_t)pid, 1);
          zz += p.jitter * unif_pm1(p.seed, (uint64_t)pid, 2);
        }
        o->y[n] = yy; o->z[n] = zz; o->j[n] = jj; o->k[n] = kk; o->id[n] = pid;
      }
}

// --------------------------------------------------------------------------
// Beam density of one bunch at a seeding point, Specie::Density case 1
// (plasma.cpp:350-377): nb*exp(-arg), arg = dx^2/Rx^2 + dy^2/R^2 + dz^2/R^2,
// zero beyond arg > 16, R = sqrt(2)*sigma.  dx is measured in xi.  With the
// piecewise profile (bins) the xi factor of bunch 0 is the bin value and the
// cut applies to the transverse part.
// --------------------------------------------------------------------------
double bunch_density(const oracle *o, int b, double xi, double yc, double zc) {
  const qss_bunch &B = o->p.bunch[b];
  double Ry = std::sqrt(2.0) * B.sigma_r;
  double dy = yc - B.y0, dz = zc - B.z0;
  if (b == 0 && o->p.nbins > 0) {
    double arg = dy * dy / Ry / Ry + dz * dz / Ry / Ry;
    if (arg > 16) return 0.0;
    double t = (xi - o->p.bin_xi0) / (o->p.bin_xi1 - o->p.bin_xi0) * o->p.nbins;
    if (!(t >= 0.0) || t >= o->p.nbins) return 0.0;
    int ib = (int)t;
    return o->bins[ib] * std::exp(-arg);
  }
  double Rx = std::sqrt(2.0) * B.sigma_xi;
  double dx = xi - B.xi0;
  double arg = dx * dx / Rx / Rx + dy * dy / Ry / Ry + dz * dz / Ry / Ry;
  if (arg > 16) return 0.0;
  return B.nb * std::exp(-arg);
}

// --------------------------------------------------------------------------
// Beam plane of reference layer i (0..L, L = nslices = front layer): frozen
// beam macro-particles, 4 per 3-D cell at (x=1/2, {1/3,2/3}^2) seeded from
// the density at the cell corner (mesh.cpp:108-163, xco = X(i), yco =
// Y(j)-Ylength/2), deposited trilinearly (movebeam.cpp:662-770).  Layer i
// receives axc=1/2 from cell i and axp=1/2 from cell i-1.  Cell ci lies at
// xi = (L - ci)*hx behind the front.  Nodes outside [0,N) are dropped (the
// beam is not ghost-folded in the reference).
// --------------------------------------------------------------------------
void beam_plane(oracle *o, int i, double *rho, double *jx) {
  const qss_params &p = o->p;
  const int ny = o->ny, nz = o->nz, L = p.nslices;
  std::fill(rho, rho + o->ncell, 0.0);
  std::fill(jx, jx + o->ncell, 0.0);
  for (int side = 0; side < 2; ++side) {
    int ci = (side == 0) ? i : i - 1;
    if (ci < 0 || ci > L) continue;
    double ax = 0.5;  // 1 - 1/2 (axc) or 1/2 (axp)
    double xi = (double)(L - ci) * p.hx;
    for (int b = 0; b < p.nbunch; ++b) {
      const qss_bunch &B = p.bunch[b];
      for (int ck = 0; ck < nz; ++ck)
        for (int cj = 0; cj < ny; ++cj) {
          double yc = o->hy * cj - p.ly / 2.0;
          double zc = o->hz * ck - p.lz / 2.0;
          double dn = bunch_density(o, b, xi, yc, zc);
          if (!(dn > 0.0)) continue;
          for (int is = 0; is < 4; ++is) {
            double yq = 1.0 / 3.0 * (1.0 + is % 2);
            double zq = 1.0 / 3.0 * (1.0 + is / 2);
            double w = 0.25 * dn;
            if (B.noise != 0.0) {
              uint64_t cell = ((uint64_t)(ci + 1) * (uint64_t)nz + ck) * (uint64_t)ny + cj;
              w *= 1.0 + B.noise * unif_pm1(p.seed ^ 0xB3A1ull, cell * 4 + is, (uint64_t)b + 7);
            }
            double drho = w * B.polarity;
            double djx = w * B.polarity * B.vx;
            for (int dz_ = 0; dz_ < 2; ++dz_) {
              int nk = ck + dz_;
              if (nk >= nz) continue;
              double az = dz_ ? zq : 1.0 - zq;
              for (int dy_ = 0; dy_ < 2; ++dy_) {
                int nj = cj + dy_;
                if (nj >= ny) continue;
                double ay = dy_ ? yq : 1.0 - yq;
                long n = nj + (long)ny * nk;
                jx[n] += djx * ax * ay * az;
                rho[n] += drho * ax * ay * az;
              }
            }
          }
        }
    }
  }
}

inline void wavenumbers(const oracle *o, int jj, int kk, double &aky,
                        double &akz, double &ak2) {
  akz = M_PI / o->p.lz * (kk + 0.5);
  aky = M_PI / o->p.ly * (jj + 0.5);
  ak2 = aky * aky + akz * akz;
}

void fwd(oracle *o, int kz, int ky, const std::vector<double> &in,
         std::vector<double> &out) {
  transform2d(o->ply, o->plz, kz, ky, in.data(), out.data(), 1.0, o->use_fft,
              o->nthreads);
}
void inv(oracle *o, int kz, int ky, const std::vector<double> &in,
         std::vector<double> &out) {
  // normalisation / (4 ny nz): movefieldshydrolinlayersplit.cpp:1008-1023
  std::vector<double> tmp(o->ncell);
  transform2d(o->ply, o->plz, kz, ky, in.data(), tmp.data(), 1.0, o->use_fft,
              o->nthreads);
  double d = 4.0 * o->ny * o->nz;
  for (long n = 0; n < o->ncell; ++n) out[n] = tmp[n] / d;
}

const int KC = QSS_KIND_C, KS = QSS_KIND_S;

// Guess: movefieldshydrolinlayersplit.cpp:222-1169 (literal, 19 transforms)
void guess(oracle *o, int first_layer) {
  std::vector<double> *S = o->S;
  const double dens = o->p.dens;
  const double hx = o->p.hx / o->p.nx_split;
  if (first_layer) {  // :479-494 zero all spectra at the first layer
    for (int s = 0; s < SP_N; ++s) std::fill(S[s].begin(), S[s].end(), 0.0);
  }
  // :637-646 beam P spectrum = previous beam spectrum
  S[SP_RHOBP] = S[SP_RHOB];
  S[SP_JXBP] = S[SP_JXB];
  // :647-700 real-space sources: P-layer rho,E,J; C-layer beam
  fwd(o, KC, KC, o->C[QSS_RHOBEAM], S[SP_RHOB]);
  fwd(o, KC, KC, o->C[QSS_JXBEAM], S[SP_JXB]);
  fwd(o, KC, KC, o->P[QSS_RHO], S[SP_RHO]);
  fwd(o, KC, KC, o->P[QSS_EX], S[SP_EX]);
  fwd(o, KC, KS, o->P[QSS_EY], S[SP_EY]);
  fwd(o, KS, KC, o->P[QSS_EZ], S[SP_EZ]);
  fwd(o, KC, KC, o->P[QSS_JX], S[SP_JX]);
  fwd(o, KC, KS, o->P[QSS_JY], S[SP_JY]);
  fwd(o, KS, KC, o->P[QSS_JZ], S[SP_JZ]);
  // :807-901 per-mode predictor
  for (int kk = 0; kk < o->nz; ++kk)
    for (int jj = 0; jj < o->ny; ++jj) {
      double aky, akz, ak2;
      wavenumbers(o, jj, kk, aky, akz, ak2);
      long n1 = jj + (long)o->ny * kk;
      double rp = S[SP_RHOP][n1] = S[SP_RHO][n1];
      double jxp = S[SP_JXP][n1] = S[SP_JX][n1];
      double jyp = S[SP_JYP][n1] = S[SP_JY][n1];
      double jzp = S[SP_JZP][n1] = S[SP_JZ][n1];
      double exp_ = S[SP_EXP][n1] = S[SP_EX][n1];
      double eyp = S[SP_EYP][n1] = S[SP_EY][n1];
      double ezp = S[SP_EZP][n1] = S[SP_EZ][n1];
      double rb = S[SP_RHOB][n1];
      double rbp = S[SP_RHOBP][n1];
      double propagator = (4. - dens * hx * hx) / (4. + dens * hx * hx);
      double denominator = 4. + dens * hx * hx;
      double rho = propagator * rp - (rb + rbp) * dens * hx * hx / denominator +
                   4. * exp_ * hx * (ak2 + dens) / denominator;
      S[SP_RHO][n1] = rho;
      double ex = propagator * exp_ -
                  2. * (2. * rp + rb + rbp) * dens * hx / ((ak2 + dens) * denominator);
      S[SP_EX][n1] = ex;
      double ey = -eyp + aky * (rb + rbp + rp + rho) / (dens + ak2);
      double ez = -ezp + akz * (rb + rbp + rp + rho) / (dens + ak2);
      S[SP_EY][n1] = ey;
      S[SP_EZ][n1] = ez;
      double jy = jyp + hx * dens * (ey + eyp) / 2.;
      double jz = jzp + hx * dens * (ez + ezp) / 2.;
      double jx = jxp + hx * dens * (ex + exp_) / 2.;
      S[SP_JY][n1] = jy; S[SP_JZ][n1] = jz; S[SP_JX][n1] = jx;
      S[SP_BX][n1] = -aky / ak2 * jz + akz / ak2 * jy;
      S[SP_BY][n1] = (-akz * (jx + S[SP_JXB][n1]) + dens * ez) / ak2;
      S[SP_BZ][n1] = (aky * (jx + S[SP_JXB][n1]) - dens * ey) / ak2;
    }
  // :946-1062 ten inverse transforms + normalisation -> C layer
  inv(o, KC, KC, S[SP_RHO], o->C[QSS_RHO]);
  inv(o, KC, KC, S[SP_EX], o->C[QSS_EX]);
  inv(o, KC, KS, S[SP_EY], o->C[QSS_EY]);
  inv(o, KS, KC, S[SP_EZ], o->C[QSS_EZ]);
  inv(o, KC, KC, S[SP_JX], o->C[QSS_JX]);
  inv(o, KC, KS, S[SP_JY], o->C[QSS_JY]);
  inv(o, KS, KC, S[SP_JZ], o->C[QSS_JZ]);
  inv(o, KS, KS, S[SP_BX], o->C[QSS_BX]);
  inv(o, KS, KC, S[SP_BY], o->C[QSS_BY]);
  inv(o, KC, KS, S[SP_BZ], o->C[QSS_BZ]);
}

// Iterate: movefieldshydrolinlayersplit.cpp:1173-1671 (literal, 10
// transforms).  Returns Eerr (:1584-1621, :1641-1643).  final_pass adds the
// build's psi plane (not in the reference).
double iterate(oracle *o, int final_pass) {
  std::vector<double> *S = o->S;
  const double dens = o->p.dens;
  const double hx = o->p.hx / o->p.nx_split;
  // :1252-1345 forward transforms of the fresh deposit (rRho unused)
  fwd(o, KC, KC, o->C[QSS_JX], S[SP_JX]);
  fwd(o, KC, KS, o->C[QSS_JY], S[SP_JY]);
  fwd(o, KS, KC, o->C[QSS_JZ], S[SP_JZ]);
  for (int kk = 0; kk < o->nz; ++kk)
    for (int jj = 0; jj < o->ny; ++jj) {
      double aky, akz, ak2;
      wavenumbers(o, jj, kk, aky, akz, ak2);
      long n1 = jj + (long)o->ny * kk;
      double rp = S[SP_RHOP][n1];
      double jy = S[SP_JY][n1];
      double jz = S[SP_JZ][n1];
      // :1390-1392 continuity
      double rho = rp + S[SP_JX][n1] - S[SP_JXP][n1] -
                   hx * (aky * (jy + S[SP_JYP][n1]) + akz * (jz + S[SP_JZP][n1])) / 2.;
      double rb = S[SP_RHOB][n1];
      double rbp = S[SP_RHOBP][n1];
      double eyp = S[SP_EYP][n1];
      double ezp = S[SP_EZP][n1];
      double newEx = -(aky * jy + akz * jz) / ak2;
      // :1423-1424 uses rho from before this update
      double newEy = -eyp + aky * (rb + rbp + rp + S[SP_RHO][n1]) / (dens + ak2);
      double newEz = -ezp + akz * (rb + rbp + rp + S[SP_RHO][n1]) / (dens + ak2);
      S[SP_EY][n1] = newEy;
      S[SP_EZ][n1] = newEz;
      S[SP_RHO][n1] = rho;
      S[SP_EX][n1] = newEx;
      S[SP_BX][n1] = -aky / ak2 * S[SP_JZ][n1] + akz / ak2 * S[SP_JY][n1];
      S[SP_BY][n1] = (-akz * (S[SP_JX][n1] + S[SP_JXB][n1]) + dens * newEz) / ak2;
      S[SP_BZ][n1] = (aky * (S[SP_JX][n1] + S[SP_JXB][n1]) - dens * newEy) / ak2;
    }
  std::vector<double> nE[3];
  for (int c = 0; c < 3; ++c) nE[c].resize(o->ncell);
  inv(o, KC, KC, S[SP_RHO], o->C[QSS_RHO]);
  inv(o, KC, KC, S[SP_EX], nE[0]);
  inv(o, KC, KS, S[SP_EY], nE[1]);
  inv(o, KS, KC, S[SP_EZ], nE[2]);
  inv(o, KS, KS, S[SP_BX], o->C[QSS_BX]);
  inv(o, KS, KC, S[SP_BY], o->C[QSS_BY]);
  inv(o, KC, KS, S[SP_BZ], o->C[QSS_BZ]);
  // residual :1584-1621: maxima floored at 1e-5
  double err = 0.0;
  for (int c = 0; c < 3; ++c) {
    double mx = 1e-5, dif = 0.0;
    std::vector<double> &old = o->C[QSS_EX + c];
    for (long n = 0; n < o->ncell; ++n) {
      double a = std::fabs(nE[c][n]);
      if (a > mx) mx = a;
      double d = std::fabs(old[n] - nE[c][n]);
      if (d > dif) dif = d;
    }
    err = std::max(err, dif / mx);
    old.swap(nE[c]);
  }
  if (final_pass) {
    // psi (build definition): psi_hat = (rho_hat - Jx_hat)/k_perp^2 in the
    // cos-cos basis, so that d(psi)/d(xi) = Ex holds through continuity.
    std::vector<double> ps(o->ncell);
    for (int kk = 0; kk < o->nz; ++kk)
      for (int jj = 0; jj < o->ny; ++jj) {
        double aky, akz, ak2;
        wavenumbers(o, jj, kk, aky, akz, ak2);
        long n1 = jj + (long)o->ny * kk;
        ps[n1] = (S[SP_RHO][n1] - S[SP_JX][n1]) / ak2;
      }
    inv(o, KC, KC, ps, o->C[QSS_PSI]);
  }
  o->last_residual = err;
  return err;
}

// One particle deposit event (a sub-step of the move)
struct Event {
  int cj, ck;
  double y, z;
  double w[4];  // djx, djy, djz, drho already times part_step
};

// Push of one particle: MoveParticlesLayerSplit body
// (movepart_splitlayer.cpp:880-1512), gather per `shape`.
// Returns the number of events written to ev (>= 1).
int push_one(const oracle *o, const std::vector<double> *F, long n,
             std::vector<Event> &ev, double &npx, double &npy, double &npz,
             int &status) {
  const qss_params &p = o->p;
  const int ny = o->ny, nz = o->nz;
  const double hx = p.hx / p.nx_split, hy = o->hy, hz = o->hz;
  const double q2m = -1.0, djx0 = -1.0, djy0 = -1.0, djz0 = -1.0, drho0 = -1.0;
  const double weight = o->pweight;
  int j = o->j[n], k = o->k[n];
  double yp = o->y[n], zp = o->z[n];
  if (!(yp >= 0 && yp <= 1 && zp >= 0 && zp <= 1)) status = QSS_E_PARTICLE;  // :945-951
  double px = o->px[n], py = o->py[n], pz = o->pz[n];
  double gammap = std::sqrt(1. + px * px + py * py + pz * pz);
  double Vx = px / gammap;
  double Vxp = Vx;
  double Vyp = py / gammap;
  double Vzp = pz / gammap;
  // fields: (P + C)/2 at the start cell (NGP_REF, :1101-1217) or its
  // bilinear interpolation at (j+yp, k+zp) (CIC, movepart_layer_CIC.cpp:166-203)
  double f[6];
  long n00 = j + (long)ny * k;
  if (p.shape == QSS_SHAPE_NGP_REF) {
    for (int c = 0; c < 6; ++c) f[c] = F[c][n00];
  } else {
    int j1 = (j + 1 == ny) ? 0 : j + 1;   // ghost = periodic copy
    int k1 = (k + 1 == nz) ? 0 : k + 1;
    long n10 = j1 + (long)ny * k, n01 = j + (long)ny * k1, n11 = j1 + (long)ny * k1;
    double ayc = 1. - yp, ayp = yp, azc = 1. - zp, azp = zp;
    double acc = ayc * azc, apc = ayp * azc, acp = ayc * azp, app = ayp * azp;
    for (int c = 0; c < 6; ++c)
      f[c] = acc * F[c][n00] + apc * F[c][n10] + acp * F[c][n01] + app * F[c][n11];
  }
  double ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
  ex *= q2m * hx / 2.;
  ey *= q2m * hx / 2.;
  ez *= q2m * hx / 2.;
  px += ex / (1. - Vx);
  py += ey / (1. - Vx);
  pz += ez / (1. - Vx);
  double gamma = std::sqrt(1. + px * px + py * py + pz * pz);
  double gamma_r = 1. / gamma;
  bx += p.b_ext[0];   // bXext..bZext, :867-869, added at :1271-1273
  by += p.b_ext[1];
  bz += p.b_ext[2];
  bx = bx * gamma_r * q2m / (1. - Vx) * hx / 2.;
  by = by * gamma_r * q2m / (1. - Vx) * hx / 2.;
  bz = bz * gamma_r * q2m / (1. - Vx) * hx / 2.;
  double co = 2. / (1. + (bx * bx) + (by * by) + (bz * bz));
  double p3x = py * bz - pz * by + px;
  double p3y = pz * bx - px * bz + py;
  double p3z = px * by - py * bx + pz;
  p3x *= co;
  p3y *= co;
  p3z *= co;
  double px_new = p3y * bz - p3z * by;
  double py_new = p3z * bx - p3x * bz;
  double pz_new = p3x * by - p3y * bx;
  px += ex / (1. - Vx) + px_new;
  py += ey / (1. - Vx) + py_new;
  pz += ez / (1. - Vx) + pz_new;
  gamma = std::sqrt(1. + px * px + py * py + pz * pz);
  double Vxm = px / gamma;
  double Vym = py / gamma;
  double Vzm = pz / gamma;
  Vx = 0.5 * (Vxm + Vxp);
  double Vy = 0.5 * (Vym + Vyp);
  double Vz = 0.5 * (Vzm + Vzp);
  double djx = weight * djx0 * Vxm;
  double djy = weight * djy0 * Vym;
  double djz = weight * djz0 * Vzm;
  double drho = weight * drho0;
  djx *= 1. / (1. - Vx);
  djy *= 1. / (1. - Vx);
  djz *= 1. / (1. - Vx);
  drho *= 1. / (1. - Vx);
  double dy = Vy * hx / hy;
  double dz = Vz * hx / hz;
  dy = dy / (1. - Vx);
  dz = dz / (1. - Vx);
  // sub-cycling :1380-1466
  double partdx = 0., step = 1.;
  if (std::fabs(dy) > 1. || std::fabs(dz) > 1.) {
    if (std::fabs(dy) > std::fabs(dz)) step = partdx = std::fabs(dy);
    else step = partdx = std::fabs(dz);
  }
  if (partdx < 1.) partdx = step = 1.;
  // build guard: more than 62 sub-steps or a non-finite displacement is an
  // error (the reference would loop for ever on inf)
  if (!(step < 63.0)) {
    status = QSS_E_PARTICLE;
    partdx = step = 1.;
    dy = 0.;
    dz = 0.;
  }
  double ytmp = yp, ztmp = zp;
  int jcur = j, kcur = k;
  int nev = 0;
  while (partdx > 0.) {
    double part_step;
    if (partdx > 1.) { partdx -= 1.; part_step = 1. / step; }
    else { part_step = partdx / step; partdx = 0.; }
    // the reference restarts every sub-step from the start cell (its inner
    // j_jump/k_jump shadow the outer ones, :1380-1381 vs :1434-1435)
    int jb = p.fix_subcycle ? jcur : j;
    int kb = p.fix_subcycle ? kcur : k;
    ytmp += dy * part_step + jb;
    ztmp += dz * part_step + kb;
    if (!(ytmp > -2.0 * ny && ytmp < 3.0 * ny && ztmp > -2.0 * nz && ztmp < 3.0 * nz)) {
      status = QSS_E_PARTICLE;
      ytmp = 0.0;
      ztmp = 0.0;
    }
    while (ytmp > ny) ytmp -= ny;
    while (ytmp < 0) ytmp += ny;
    while (ztmp > nz) ztmp -= nz;
    while (ztmp < 0) ztmp += nz;
    int jj = (int)ytmp;
    int kk = (int)ztmp;
    ytmp -= jj;
    ztmp -= kk;
    // build convention: a particle landing exactly on y = Ny (z = Nz) is
    // wrapped to cell 0 instead of being parked in the ghost cell
    if (jj >= ny) jj -= ny;
    if (kk >= nz) kk -= nz;
    Event e;
    e.cj = jj; e.ck = kk; e.y = ytmp; e.z = ztmp;
    e.w[0] = djx * part_step; e.w[1] = djy * part_step;
    e.w[2] = djz * part_step; e.w[3] = drho * part_step;
    ev.push_back(e);
    ++nev;
    jcur = jj; kcur = kk;
  }
  npx = px; npy = py; npz = pz;
  return nev;
}

// MoveParticlesLayerSplit + DepositCurrentsInCellSplit + ExchangeCurrentsSplit
// Events are deposited in slot order (particle, then sub-step) straight into
// the nodes, like DepositCurrentsInCellSplit (:1612-1615) / MoveInCell
// (movepart_layer_CIC.cpp:470-495), on planes with one ghost node per axis;
// the ghosts are then folded periodically, y first, then z, as
// ExchangeCurrentsSplit does (:1685-1711).
int push_deposit(oracle *o, int full_step) {
  const int ny = o->ny, nz = o->nz;
  const long ncell = o->ncell;
  const long np = (long)o->y.size();
  // ClearCurrentsSplit / ClearRhoSplit (mesh.cpp:692-708)
  for (int c = QSS_JX; c <= QSS_RHO; ++c) std::fill(o->C[c].begin(), o->C[c].end(), 0.0);
  std::vector<double> F[6];
  for (int c = 0; c < 6; ++c) {
    F[c].resize(ncell);
    for (long n = 0; n < ncell; ++n) F[c][n] = 0.5 * (o->P[c][n] + o->C[c][n]);
  }
  // per-particle events (parallel, deterministic)
  std::vector<std::vector<Event>> evs(np);
  std::vector<double> npx(np), npy(np), npz(np);
  int status = QSS_OK;
#pragma omp parallel for num_threads(o->nthreads) schedule(static) reduction(min:status)
  for (long n = 0; n < np; ++n) {
    int st = QSS_OK;
    push_one(o, F, n, evs[n], npx[n], npy[n], npz[n], st);
    status = std::min(status, st);
  }
  // deposit in slot order onto (ny+1) x (nz+1) planes (one ghost per axis)
  const int ey = ny + 1, ez = nz + 1;
  std::vector<double> J[4];
  for (int q = 0; q < 4; ++q) J[q].assign((size_t)ey * ez, 0.0);
  const bool cic = (o->p.shape == QSS_SHAPE_CIC);
  for (long n = 0; n < np; ++n)
    for (const Event &e : evs[n]) {
      const long ccc = e.cj + (long)ey * e.ck;
      if (cic) {
        const double ayc = 1. - e.y, ayp = e.y, azc = 1. - e.z, azp = e.z;
        const long cpc = ccc + 1, ccp = ccc + ey, cpp = ccp + 1;
        for (int q = 0; q < 4; ++q) {
          J[q][ccc] += e.w[q] * ayc * azc;
          J[q][cpc] += e.w[q] * ayp * azc;
          J[q][ccp] += e.w[q] * ayc * azp;
          J[q][cpp] += e.w[q] * ayp * azp;
        }
      } else {
        for (int q = 0; q < 4; ++q) J[q][ccc] += e.w[q];
      }
    }
  for (int q = 0; q < 4; ++q) {
    std::vector<double> &A = J[q];
    for (int kk = 0; kk <= nz; ++kk) {        // ghost column j = ny onto j = 0
      A[0 + (long)ey * kk] += A[ny + (long)ey * kk];
      A[ny + (long)ey * kk] = 0.;
    }
    for (int jj = 0; jj <= ny; ++jj) {        // ghost row k = nz onto k = 0
      A[jj] += A[jj + (long)ey * nz];
      A[jj + (long)ey * nz] = 0.;
    }
    for (int kk = 0; kk < nz; ++kk)
      for (int jj = 0; jj < ny; ++jj)
        o->C[QSS_JX + q][jj + (long)ny * kk] = A[jj + (long)ey * kk];
  }
  if (full_step) {
    // particles move to the C layer: last sub-step position, new momenta;
    // storage re-sorted by (cell, previous slot) -- stable.
    std::vector<long> key(np);
    for (long n = 0; n < np; ++n) {
      const Event &e = evs[n].back();
      key[n] = e.cj + (long)ny * e.ck;
    }
    std::vector<long> order(np);
    std::iota(order.begin(), order.end(), 0L);
    std::stable_sort(order.begin(), order.end(),
                     [&](long a, long b) { return key[a] < key[b]; });
    std::vector<double> ny_(np), nz_(np), nx1(np), ny1(np), nz1(np);
    std::vector<int> nj(np), nk(np);
    std::vector<int64_t> nid(np);
    for (long q = 0; q < np; ++q) {
      long n = order[q];
      const Event &e = evs[n].back();
      ny_[q] = e.y; nz_[q] = e.z; nj[q] = e.cj; nk[q] = e.ck;
      nx1[q] = npx[n]; ny1[q] = npy[n]; nz1[q] = npz[n]; nid[q] = o->id[n];
    }
    o->y.swap(ny_); o->z.swap(nz_); o->j.swap(nj); o->k.swap(nk);
    o->px.swap(nx1); o->py.swap(ny1); o->pz.swap(nz1); o->id.swap(nid);
  }
  return status;
}

// Beam planes of the C layer for (slice, isplit): movepart_splitlayer.cpp:
// 330-348 interpolation between reference layers i and i+1.
void beam_layer(oracle *o, int slice, int isplit) {
  const int L = o->p.nslices, S = o->p.nx_split;
  int i = L - 1 - slice;
  std::vector<double> rc(o->ncell), jc(o->ncell);
  if (o->driver) {   // the evolving driver's 3-D deposit, nodes of layer i
    std::copy(o->brho3.begin() + (size_t)i * o->ncell,
              o->brho3.begin() + (size_t)(i + 1) * o->ncell, rc.begin());
    std::copy(o->bjx3.begin() + (size_t)i * o->ncell,
              o->bjx3.begin() + (size_t)(i + 1) * o->ncell, jc.begin());
  } else {
    beam_plane(o, i, rc.data(), jc.data());
  }
  if (S == 1) {
    o->C[QSS_RHOBEAM] = rc; o->C[QSS_JXBEAM] = jc;
    return;
  }
  std::vector<double> rp(o->ncell), jp(o->ncell);
  if (o->driver) {
    std::copy(o->brho3.begin() + (size_t)(i + 1) * o->ncell,
              o->brho3.begin() + (size_t)(i + 2) * o->ncell, rp.begin());
    std::copy(o->bjx3.begin() + (size_t)(i + 1) * o->ncell,
              o->bjx3.begin() + (size_t)(i + 2) * o->ncell, jp.begin());
  } else {
    beam_plane(o, i + 1, rp.data(), jp.data());
  }
  for (long n = 0; n < o->ncell; ++n) {
    o->C[QSS_RHOBEAM][n] = (rc[n] * (S - isplit) + rp[n] * isplit) / S;
    o->C[QSS_JXBEAM][n] = (jc[n] * (S - isplit) + jp[n] * isplit) / S;
  }
}

// --------------------------------------------------------------------------
// Evolving driver (test restatement of Mesh::SeedBeamParticles,
// Mesh::MoveBeamParticles and Mesh::MoveBeamInCell).
// --------------------------------------------------------------------------

// Mesh::SeedBeamParticles over the cells of one rank (mesh.cpp:43-64: i in
// [0, l_Mx-1], j, k interior) for the nseed == 4 lattice (mesh.cpp:148-152:
// x = 1/2, y,z in {1/3,2/3}), weight = GetWeight()*dens = dens/4
// (mesh.cpp:165), momentum (Px0, Py0, Pz0) with zero spreads.  Build
// conventions: particles are stored in (i, k, j, bunch, iseed) order (the
// reference keeps per-cell lists); the bunch noise factor and the density
// at the cell corner are those of the frozen lattice (beam_plane).
void driver_seed(oracle *o, const qss_driver &d) {
  const qss_params &p = o->p;
  const int ny = o->ny, nz = o->nz, L = p.nslices;
  o->drv = d;
  o->driver = true;
  o->bi.clear(); o->bj.clear(); o->bk.clear(); o->bsp.clear();
  o->bx.clear(); o->by.clear(); o->bz.clear();
  o->bpx.clear(); o->bpy.clear(); o->bpz.clear(); o->bw.clear();
  for (int ci = 0; ci < L; ++ci) {
    const double xi = (double)(L - ci) * p.hx;
    for (int ck = 0; ck < nz; ++ck)
      for (int cj = 0; cj < ny; ++cj)
        for (int b = 0; b < p.nbunch; ++b) {
          const qss_bunch &B = p.bunch[b];
          const double yc = o->hy * cj - p.ly / 2.0;
          const double zc = o->hz * ck - p.lz / 2.0;
          const double dn = bunch_density(o, b, xi, yc, zc);
          if (!(dn > 0.0)) continue;
          for (int is = 0; is < 4; ++is) {
            double w = 0.25 * dn;
            if (B.noise != 0.0) {
              uint64_t cell = ((uint64_t)(ci + 1) * (uint64_t)nz + ck) * (uint64_t)ny + cj;
              w *= 1.0 + B.noise * unif_pm1(p.seed ^ 0xB3A1ull, cell * 4 + is, (uint64_t)b + 7);
            }
            o->bi.push_back(ci); o->bj.push_back(cj); o->bk.push_back(ck);
            o->bsp.push_back(b);
            o->bx.push_back(0.5);
            o->by.push_back(1.0 / 3.0 * (1.0 + is % 2));
            o->bz.push_back(1.0 / 3.0 * (1.0 + is / 2));
            o->bpx.push_back(d.p0[b][0]); o->bpy.push_back(d.p0[b][1]);
            o->bpz.push_back(d.p0[b][2]);
            o->bw.push_back(w);
          }
        }
  }
  o->brho3.assign((size_t)(L + 1) * o->ncell, 0.0);
  o->bjx3.assign((size_t)(L + 1) * o->ncell, 0.0);
}

// Field f of the stored wake at 3-D cell (i, j, k): layer i < L is history
// slice L-1-i (the layer stored at movepart_splitlayer.cpp:467-473), layer L
// is the unperturbed front (zero); the stored transverse ghosts j = Ny,
// k = Nz are the periodic copies of ExchangeFieldsSplit
// (movepart_splitlayer.cpp:1659-1682).
inline double wake_field(const oracle *o, int f, int i, int j, int k) {
  const int L = o->p.nslices;
  if (i < 0 || i >= L) return 0.0;
  if (j == o->ny) j = 0;
  if (k == o->nz) k = 0;
  return o->hist[f][(size_t)(L - 1 - i) * o->ncell + j + (long)o->ny * k];
}

// Mesh::MoveBeamParticles (movebeam.cpp:14-660), one rank, no scattering
// (ifscatter = 0), in the reference's operation order.  Particles outside
// the interior cells are not visited (the loops of movebeam.cpp:63-69), so
// they are neither pushed nor deposited.  The E kick is the interpolated
// field itself: its q2m*ts/2 scaling is commented out in the reference
// (movebeam.cpp:283-302).  Returns QSS_E_PARTICLE for an in-cell fraction
// outside [0,1] (the reference's exit(-12), movebeam.cpp:200-206).
int driver_push(oracle *o) {
  const qss_params &p = o->p;
  const qss_driver &d = o->drv;
  const int ny = o->ny, nz = o->nz, L = p.nslices;
  const double ts = d.ts, hx = p.hx, hy = o->hy, hz = o->hz;
  const double PI_REF = 3.1415927;                       // domain.h:10
  const double ElaserPhoton = 1.2398e-4 / d.wavelength;  // movebeam.cpp:19
  std::fill(o->brho3.begin(), o->brho3.end(), 0.0);     // ClearBeamCurrents
  std::fill(o->bjx3.begin(), o->bjx3.end(), 0.0);
  const long np = (long)o->bx.size();
  for (long n = 0; n < np; ++n) {
    const int i = o->bi[n], j = o->bj[n], k = o->bk[n];
    if (i < 0 || i >= L || j < 0 || j >= ny || k < 0 || k >= nz) continue;
    const qss_bunch &B = p.bunch[o->bsp[n]];
    const double weight = o->bw[n];
    const double x = o->bx[n], y = o->by[n], z = o->bz[n];
    double px = o->bpx[n], py = o->bpy[n], pz = o->bpz[n];
    const double pxn = px, pyn = py, pzn = pz;
    const double q2m = d.q2m[o->bsp[n]];
    if (x < 0 || x > 1 || y < 0 || y > 1 || z < 0 || z > 1) return QSS_E_PARTICLE;
    const double axc = 1. - x, axp = x, ayc = 1. - y, ayp = y, azc = 1. - z, azp = z;
    const double accc = axc * ayc * azc, acpc = axc * ayp * azc;
    const double accp = axc * ayc * azp, acpp = axc * ayp * azp;
    const double apcc = axp * ayc * azc, appc = axp * ayp * azc;
    const double apcp = axp * ayc * azp, appp = axp * ayp * azp;
    double g[6];
    for (int f = 0; f < 6; ++f)   // movebeam.cpp:232-272 term order
      g[f] = accc * wake_field(o, f, i, j, k) + acpc * wake_field(o, f, i, j + 1, k) +
             accp * wake_field(o, f, i, j, k + 1) + acpp * wake_field(o, f, i, j + 1, k + 1) +
             apcc * wake_field(o, f, i + 1, j, k) + appc * wake_field(o, f, i + 1, j + 1, k) +
             apcp * wake_field(o, f, i + 1, j, k + 1) +
             appp * wake_field(o, f, i + 1, j + 1, k + 1);
    const double ex = g[0], ey = g[1], ez = g[2];
    double bx = g[3], by = g[4], bz = g[5];
    bx += d.b_ext[0];                                    // movebeam.cpp:311-313
    by += d.b_ext[1];
    bz += d.b_ext[2];
    px += ex;
    py += ey;
    pz += ez;
    double gamma = sqrt(1. + px * px + py * py + pz * pz);
    const double gamma_r = 1. / gamma;
    bx = bx * gamma_r * q2m * ts / 2.;                   // movebeam.cpp:379-381
    by = by * gamma_r * q2m * ts / 2.;
    bz = bz * gamma_r * q2m * ts / 2.;
    const double co = 2. / (1. + (bx * bx) + (by * by) + (bz * bz));
    double p3x = py * bz - pz * by + px;
    double p3y = pz * bx - px * bz + py;
    double p3z = px * by - py * bx + pz;
    p3x *= co;
    p3y *= co;
    p3z *= co;
    const double px_new = p3y * bz - p3z * by;
    const double py_new = p3z * bx - p3x * bz;
    const double pz_new = p3x * by - p3y * bx;
    px += (ex + px_new);
    py += (ey + py_new);
    pz += (ez + pz_new);
    // radiation reaction (movebeam.cpp:448-501); the photon histogram
    // (Synchrotron::AddParticle) is a diagnostic and not restated
    const double p2 = px * px + py * py + pz * pz;
    if (1. + p2 > d.syn_min * d.syn_min) {
      double dpx = px - pxn, dpy = py - pyn, dpz = pz - pzn;
      const double pdp = px * dpx + py * dpy + pz * dpz;
      dpx -= px * (pdp / p2);
      dpy -= py * (pdp / p2);
      dpz -= pz * (pdp / p2);
      const double p2absrev = 1.f / sqrt(p2);
      const double nx = px * p2absrev, ny_ = py * p2absrev, nz_ = pz * p2absrev;
      const double dpabs = sqrt(dpx * dpx + dpy * dpy + dpz * dpz);
      const double Gamma2 = 1. + p2;
      const double omega_c_omega_laser = 3. * Gamma2 * dpabs / (4. * PI_REF * ts);
      const double Ephoton = ElaserPhoton * omega_c_omega_laser;
      const double alfa = 1. / 137.;
      const double nph = 4. / 9. * alfa * dpabs;
      px -= (nph * Ephoton * nx / 5.11e5f);
      py -= (nph * Ephoton * ny_ / 5.11e5f);
      pz -= (nph * Ephoton * nz_ / 5.11e5f);
    }
    gamma = 1. / sqrt(1. + px * px + py * py + pz * pz);
    const double Vx = px * gamma, Vy = py * gamma, Vz = pz * gamma;
    o->bpx[n] = px; o->bpy[n] = py; o->bpz[n] = pz;
    const double polarity = B.polarity;
    const double djx = weight * polarity * Vx;
    const double drho = weight * polarity;
    // move (movebeam.cpp:520-583)
    const double dx = (px * gamma - 1.) * ts / hx;
    const double dy = py * gamma * ts / hy;
    const double dz = pz * gamma * ts / hz;
    // a non-finite or huge move would never leave the wrap loops below
    if (!(std::fabs(dx) < 1e9 && std::fabs(dy) < 1e9 && std::fabs(dz) < 1e9))
      return QSS_E_PARTICLE;
    double xtmp = x + dx, ytmp = y + dy, ztmp = z + dz;
    int i_jump = (int)xtmp, j_jump = (int)ytmp, k_jump = (int)ztmp;
    if (xtmp < 0.) i_jump--;
    if (ytmp < 0.) j_jump--;
    if (ztmp < 0.) k_jump--;
    int itmp = i + i_jump, jtmp = j + j_jump, ktmp = k + k_jump;
    if (itmp < 0) itmp = -1;
    if (itmp > L - 1) itmp = L;
    if (jtmp < 0) jtmp = -1;
    if (jtmp > ny - 1) jtmp = ny;
    if (ktmp < 0) ktmp = -1;
    if (ktmp > nz - 1) ktmp = nz;
    xtmp -= i_jump;
    ytmp -= j_jump;
    ztmp -= k_jump;
    while (xtmp < 0.) xtmp += 1.;
    while (xtmp > 1.) xtmp -= 1.;
    while (ytmp < 0.) ytmp += 1.;
    while (ytmp > 1.) ytmp -= 1.;
    while (ztmp < 0.) ztmp += 1.;
    while (ztmp > 1.) ztmp -= 1.;
    o->bi[n] = itmp; o->bj[n] = jtmp; o->bk[n] = ktmp;
    o->bx[n] = xtmp; o->by[n] = ytmp; o->bz[n] = ztmp;
    // Mesh::MoveBeamInCell (movebeam.cpp:662-770): rho_b and Jx_b on the 8
    // nodes of the new cell; nodes outside layers [0, L] and the interior
    // transverse range are dropped (beam currents are not ghost-folded)
    const double ax[2] = {1. - xtmp, xtmp}, ay[2] = {1. - ytmp, ytmp},
                 az[2] = {1. - ztmp, ztmp};
    for (int a = 0; a < 2; ++a)
      for (int c = 0; c < 2; ++c)
        for (int b = 0; b < 2; ++b) {
          const int ni = itmp + a, nj = jtmp + b, nk = ktmp + c;
          if (ni < 0 || ni > L || nj < 0 || nj >= ny || nk < 0 || nk >= nz) continue;
          const size_t m = (size_t)ni * o->ncell + nj + (size_t)ny * nk;
          o->bjx3[m] += djx * ax[a] * ay[b] * az[c];
          o->brho3[m] += drho * ax[a] * ay[b] * az[c];
        }
  }
  return QSS_OK;
}

void copy_c_to_p(oracle *o) {
  // Cell::operator= (cells.cpp:25-47): fields, currents and density
  for (int f = QSS_EX; f <= QSS_RHO; ++f) o->P[f] = o->C[f];
  o->P[QSS_PSI] = o->C[QSS_PSI];
}

int layer_step(oracle *o, int slice, int isplit, double *residual) {
  beam_layer(o, slice, isplit);
  guess(o, slice == 0);
  int st = QSS_OK;
  for (int it = 0; it < o->p.n_iter; ++it) {
    st = std::min(st, push_deposit(o, 0));
    iterate(o, 0);
  }
  st = std::min(st, push_deposit(o, 1));
  double err = iterate(o, 1);
  if (residual) *residual = err;
  copy_c_to_p(o);
  if (!std::isfinite(err)) st = std::min(st, QSS_E_NONFINITE);
  return st;
}

void reset(oracle *o) {
  for (int f = 0; f < QSS_NFIELD; ++f) {
    std::fill(o->C[f].begin(), o->C[f].end(), 0.0);
    std::fill(o->P[f].begin(), o->P[f].end(), 0.0);
  }
  for (int s = 0; s < SP_N; ++s) std::fill(o->S[s].begin(), o->S[s].end(), 0.0);
  seed_plasma(o);
}

}  // namespace

extern "C" {

oracle *oracle_create(const qss_params *p, int use_fft, int nthreads,
                      uint32_t history_mask) {
  if (!p || p->ny < 2 || p->nz < 2 || p->nslices < 1 || p->nx_split < 1 ||
      p->ppc_y < 1 || p->ppc_z < 1 || p->nbunch < 0 || p->nbunch > QSS_MAX_BUNCH)
    return nullptr;
  oracle *o = new oracle();
  o->p = *p;
  if (p->nbins > 0 && p->bins) o->bins.assign(p->bins, p->bins + p->nbins);
  o->p.bins = nullptr;
  o->ny = p->ny; o->nz = p->nz;
  o->ncell = (long)p->ny * p->nz;
  o->hy = p->ly / p->ny;
  o->hz = p->lz / p->nz;
  o->use_fft = use_fft != 0;
  o->nthreads = nthreads > 0 ? nthreads : 1;
  o->ply.init(p->ny);
  o->plz.init(p->nz);
  for (int f = 0; f < QSS_NFIELD; ++f) {
    o->C[f].assign(o->ncell, 0.0);
    o->P[f].assign(o->ncell, 0.0);
  }
  for (int s = 0; s < SP_N; ++s) o->S[s].assign(o->ncell, 0.0);
  o->hmask = history_mask;
  for (int f = 0; f < QSS_NFIELD; ++f)
    if (history_mask & (1u << f))
      o->hist[f].assign((size_t)o->ncell * p->nslices, 0.0);
  seed_plasma(o);
  return o;
}

void oracle_destroy(oracle *o) { delete o; }

// The reference's layer record of the P layer (PackLayer, para.cpp:94-160):
// planes Bx, By, Bz, Ex, Ey, Ez, Jx, Jy, Jz, Rho, fftJxBeamHydro,
// fftRhoBeamHydro, then per particle (attribute-major) Px, Py, Pz, X, Y, Z,
// Q2m, Weight, i_X, i_Y, i_Z, isort, particles in storage order.
long oracle_layer_record_doubles(const oracle *o) {
  return 12 * o->ncell + 12 * (long)o->y.size();
}

int oracle_pack_layer_record(oracle *o, int iLayer, double *dst) {
  const long nc = o->ncell, np = (long)o->y.size();
  const int f[10] = {QSS_BX, QSS_BY, QSS_BZ, QSS_EX, QSS_EY, QSS_EZ,
                     QSS_JX, QSS_JY, QSS_JZ, QSS_RHO};
  for (int q = 0; q < 10; ++q) std::copy(o->P[f[q]].begin(), o->P[f[q]].end(), dst + q * nc);
  std::copy(o->S[SP_JXB].begin(), o->S[SP_JXB].end(), dst + 10 * nc);
  std::copy(o->S[SP_RHOB].begin(), o->S[SP_RHOB].end(), dst + 11 * nc);
  double *r = dst + 12 * nc;
  for (long i = 0; i < np; ++i) {
    r[i] = o->px[i]; r[i + np] = o->py[i]; r[i + 2 * np] = o->pz[i];
    r[i + 3 * np] = 0.0; r[i + 4 * np] = o->y[i]; r[i + 5 * np] = o->z[i];
    r[i + 6 * np] = -1.0; r[i + 7 * np] = o->pweight;
    r[i + 8 * np] = iLayer; r[i + 9 * np] = o->j[i]; r[i + 10 * np] = o->k[i];
    r[i + 11 * np] = 0.0;
  }
  return QSS_OK;
}

// UnpackLayer (para.cpp:167-280) into the P layer: Np from the length
// (ReceiveLayer, para.cpp:385); ids = record positions.  The oracle's Guess
// re-transforms the P layer at every layer, like the reference, so the
// planes, the beam spectra and the particles are its whole state.
int oracle_unpack_layer_record(oracle *o, const double *src, long n) {
  const long nc = o->ncell;
  if (n < 12 * nc || (n - 12 * nc) % 12 != 0) return QSS_E_ARG;
  const long np = (n - 12 * nc) / 12;
  const int f[10] = {QSS_BX, QSS_BY, QSS_BZ, QSS_EX, QSS_EY, QSS_EZ,
                     QSS_JX, QSS_JY, QSS_JZ, QSS_RHO};
  for (int q = 0; q < 10; ++q) std::copy(src + q * nc, src + (q + 1) * nc, o->P[f[q]].begin());
  std::copy(src + 10 * nc, src + 11 * nc, o->S[SP_JXB].begin());
  std::copy(src + 11 * nc, src + 12 * nc, o->S[SP_RHOB].begin());
  const double *r = src + 12 * nc;
  o->y.resize(np); o->z.resize(np); o->px.resize(np); o->py.resize(np); o->pz.resize(np);
  o->j.resize(np); o->k.resize(np); o->id.resize(np);
  for (long i = 0; i < np; ++i) {
    o->px[i] = r[i]; o->py[i] = r[i + np]; o->pz[i] = r[i + 2 * np];
    o->y[i] = r[i + 4 * np]; o->z[i] = r[i + 5 * np];
    o->j[i] = (int)r[i + 9 * np]; o->k[i] = (int)r[i + 10 * np];
    o->id[i] = i;
  }
  return QSS_OK;
}

int oracle_reset(oracle *o) { reset(o); return QSS_OK; }

int oracle_n_particles(const oracle *o) { return (int)o->y.size(); }

int oracle_layer_step(oracle *o, int slice, int isplit, double *residual) {
  return layer_step(o, slice, isplit, residual);
}

// movepart_splitlayer.cpp:165-478: per layer P <- previous layer, splits,
// store the layer (history).
int oracle_sweep_range(oracle *o, int first, int count, double *residuals) {
  int st = QSS_OK;
  for (int s = first; s < first + count && s < o->p.nslices; ++s) {
    for (int is = 0; is < o->p.nx_split; ++is) {
      double r = 0;
      st = std::min(st, layer_step(o, s, is, &r));
      if (residuals) residuals[(s - first) * o->p.nx_split + is] = r;
    }
    // the layer's fields are in P after copy_c_to_p; its beam planes in C
    for (int f = 0; f < QSS_NFIELD; ++f)
      if (!o->hist[f].empty()) {
        const std::vector<double> &src = f == QSS_RHOBEAM || f == QSS_JXBEAM ? o->C[f] : o->P[f];
        std::copy(src.begin(), src.end(), o->hist[f].begin() + (size_t)s * o->ncell);
      }
  }
  return st;
}

int oracle_sweep(oracle *o, double *residuals) {
  reset(o);
  return oracle_sweep_range(o, 0, o->p.nslices, residuals);
}

int oracle_beam_layer(oracle *o, int slice, int isplit) {
  beam_layer(o, slice, isplit); return QSS_OK;
}
int oracle_beam_plane(oracle *o, int layer, double *rho, double *jx) {
  beam_plane(o, layer, rho, jx); return QSS_OK;
}
int oracle_guess(oracle *o, int first_layer) { guess(o, first_layer); return QSS_OK; }
int oracle_push_deposit(oracle *o, int full_step) { return push_deposit(o, full_step); }
int oracle_iterate(oracle *o, int final_pass, double *residual) {
  double r = iterate(o, final_pass);
  if (residual) *residual = r;
  return QSS_OK;
}
int oracle_copy_c_to_p(oracle *o) { copy_c_to_p(o); return QSS_OK; }

int oracle_get_layer_field(oracle *o, int which, int field, double *out) {
  if (field < 0 || field >= QSS_NFIELD) return QSS_E_ARG;
  std::vector<double> &v = which ? o->P[field] : o->C[field];
  std::copy(v.begin(), v.end(), out);
  return QSS_OK;
}
int oracle_set_layer_field(oracle *o, int which, int field, const double *in) {
  if (field < 0 || field >= QSS_NFIELD) return QSS_E_ARG;
  std::vector<double> &v = which ? o->P[field] : o->C[field];
  std::copy(in, in + o->ncell, v.begin());
  return QSS_OK;
}
int oracle_get_spectral(oracle *o, int sid, double *out) {
  if (sid < 0 || sid >= SP_N) return QSS_E_ARG;
  std::copy(o->S[sid].begin(), o->S[sid].end(), out);
  return QSS_OK;
}
int oracle_set_spectral(oracle *o, int sid, const double *in) {
  if (sid < 0 || sid >= SP_N) return QSS_E_ARG;
  std::copy(in, in + o->ncell, o->S[sid].begin());
  return QSS_OK;
}
int oracle_get_history(oracle *o, int field, int slice, double *out) {
  if (field < 0 || field >= QSS_NFIELD || o->hist[field].empty() || slice < 0 ||
      slice >= o->p.nslices)
    return QSS_E_ARG;
  std::copy(o->hist[field].begin() + (size_t)slice * o->ncell,
            o->hist[field].begin() + (size_t)(slice + 1) * o->ncell, out);
  return QSS_OK;
}
// particles: y,z as global cell coordinates j+y, k+z
int oracle_get_particles(oracle *o, double *y, double *z, double *px, double *py,
                         double *pz, int64_t *id, int32_t *cj, int32_t *ck,
                         double *fy, double *fz) {
  long np = (long)o->y.size();
  for (long n = 0; n < np; ++n) {
    if (y) y[n] = o->j[n] + o->y[n];
    if (z) z[n] = o->k[n] + o->z[n];
    if (px) px[n] = o->px[n];
    if (py) py[n] = o->py[n];
    if (pz) pz[n] = o->pz[n];
    if (id) id[n] = o->id[n];
    if (cj) cj[n] = o->j[n];
    if (ck) ck[n] = o->k[n];
    if (fy) fy[n] = o->y[n];
    if (fz) fz[n] = o->z[n];
  }
  return QSS_OK;
}
// set particles from cell index + in-cell fraction (exact state)
int oracle_set_particles(oracle *o, long np, const int32_t *cj, const int32_t *ck,
                         const double *fy, const double *fz, const double *px,
                         const double *py, const double *pz, const int64_t *id) {
  o->y.assign(fy, fy + np); o->z.assign(fz, fz + np);
  o->j.assign(cj, cj + np); o->k.assign(ck, ck + np);
  o->px.assign(px, px + np); o->py.assign(py, py + np); o->pz.assign(pz, pz + np);
  o->id.assign(id, id + np);
  return QSS_OK;
}

int oracle_set_history(oracle *o, int field, int slice, const double *in) {
  if (field < 0 || field >= QSS_NFIELD || o->hist[field].empty() || slice < 0 ||
      slice >= o->p.nslices)
    return QSS_E_ARG;
  std::copy(in, in + o->ncell, o->hist[field].begin() + (size_t)slice * o->ncell);
  return QSS_OK;
}

// evolving driver
int oracle_driver_seed(oracle *o, const qss_driver *d, int64_t *n) {
  for (int f = QSS_EX; f <= QSS_BZ; ++f)
    if (o->hist[f].empty()) return QSS_E_STATE;
  driver_seed(o, *d);
  if (n) *n = (int64_t)o->bx.size();
  return QSS_OK;
}
int oracle_driver_push(oracle *o) { return o->driver ? driver_push(o) : QSS_E_STATE; }
// Domain::Step (step.cpp:63,105): beam push, then the sweep
int oracle_time_step(oracle *o, double *residuals) {
  if (!o->driver) return QSS_E_STATE;
  int st = driver_push(o);
  if (st != QSS_OK) return st;
  reset(o);
  return oracle_sweep_range(o, 0, o->p.nslices, residuals);
}
int64_t oracle_driver_count(const oracle *o) { return (int64_t)o->bx.size(); }
int oracle_driver_get(oracle *o, int32_t *cell, double *frac, double *mom, double *weight,
                      int32_t *bunch) {
  const long np = (long)o->bx.size();
  for (long n = 0; n < np; ++n) {
    if (cell) { cell[3 * n] = o->bi[n]; cell[3 * n + 1] = o->bj[n]; cell[3 * n + 2] = o->bk[n]; }
    if (frac) { frac[3 * n] = o->bx[n]; frac[3 * n + 1] = o->by[n]; frac[3 * n + 2] = o->bz[n]; }
    if (mom) { mom[3 * n] = o->bpx[n]; mom[3 * n + 1] = o->bpy[n]; mom[3 * n + 2] = o->bpz[n]; }
    if (weight) weight[n] = o->bw[n];
    if (bunch) bunch[n] = o->bsp[n];
  }
  return QSS_OK;
}
int oracle_driver_set(oracle *o, int64_t n, const int32_t *cell, const double *frac,
                      const double *mom, const double *weight, const int32_t *bunch) {
  if (!o->driver || n < 0) return QSS_E_STATE;
  o->bi.resize(n); o->bj.resize(n); o->bk.resize(n); o->bsp.resize(n);
  o->bx.resize(n); o->by.resize(n); o->bz.resize(n);
  o->bpx.resize(n); o->bpy.resize(n); o->bpz.resize(n); o->bw.resize(n);
  for (int64_t q = 0; q < n; ++q) {
    o->bi[q] = cell[3 * q]; o->bj[q] = cell[3 * q + 1]; o->bk[q] = cell[3 * q + 2];
    o->bx[q] = frac[3 * q]; o->by[q] = frac[3 * q + 1]; o->bz[q] = frac[3 * q + 2];
    o->bpx[q] = mom[3 * q]; o->bpy[q] = mom[3 * q + 1]; o->bpz[q] = mom[3 * q + 2];
    o->bw[q] = weight[q]; o->bsp[q] = bunch[q];
  }
  return QSS_OK;
}
int oracle_driver_density(oracle *o, int layer, double *rho, double *jx) {
  if (!o->driver || layer < 0 || layer > o->p.nslices) return QSS_E_ARG;
  const size_t off = (size_t)layer * o->ncell;
  if (rho) std::copy(o->brho3.begin() + off, o->brho3.begin() + off + o->ncell, rho);
  if (jx) std::copy(o->bjx3.begin() + off, o->bjx3.begin() + off + o->ncell, jx);
  return QSS_OK;
}

// standalone transforms (for fixtures / transform parity)
void oracle_transform1d(int n, int kind, const double *in, double *out, int use_fft) {
  Plan1D pl; pl.init(n);
  std::vector<cplx> work(n); std::vector<double> xbuf(n);
  pl.apply(kind, in, 1, out, 1, use_fft != 0, work.data(), xbuf.data());
}
void oracle_transform2d(int ny, int nz, int kind_z, int kind_y, const double *in,
                        double *out, double scale, int use_fft, int nthreads) {
  Plan1D py_, pz_; py_.init(ny); pz_.init(nz);
  transform2d(py_, pz_, kind_z, kind_y, in, out, scale, use_fft != 0,
              nthreads > 0 ? nthreads : 1);
}

}  // extern "C"
