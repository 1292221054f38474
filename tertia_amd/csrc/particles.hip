// particles.hip — plasma seeding, predictor-corrector push and deterministic
// current/charge deposit.
//
// Push: Mesh::MoveParticlesLayerSplit (movepart_splitlayer.cpp:880-1512;
// GPU twin cuMoveSplitParticlesKernel, CUDA_WRAP/plasma_particles.cu:72-657).
// One thread per particle, SoA state in HBM sorted by cell so that the field
// gathers of a wavefront hit neighbouring nodes.  E,B = (P+C)/2 come from the
// interleaved push field written by the inverse-transform epilogue: taken at
// the start cell (NGP_REF, :1101-1217) or bilinearly at (j+y, k+z) (CIC,
// movepart_layer_CIC.cpp:166-203).
//
// Deposit: instead of the reference's CAS-loop double atomics
// (plasma_particles.cu:28-39, 705-715), every push emits deposit events
// (target cell, in-cell position, dJx,dJy,dJz,dRho x part_step), one per
// sub-step (:1391-1466).  Events are counting-sorted by target cell
// (integer atomics only give counts and slots), and the tile kernel sums each
// cell's events in key order (particle slot, sub-step) into four CIC corner
// sums held in LDS, then combines nodes as c00+c10+c01+c11 with periodic
// wrap, which is the ghost fold of ExchangeCurrentsSplit (:1685-1711).  The
// result is bitwise independent of thread scheduling.
//
// Compiled with -ffp-contract=off: the arithmetic is the reference's, in its
// operation order.

#include "qss_internal.h"

namespace {

constexpr int NT = 256;
constexpr int TILE = 16;          // node tile edge of the reduce kernel
constexpr int HC = TILE + 1;      // cells per tile edge incl. the low halo
constexpr int SCAN_B = 1024;      // cells per scan block
constexpr int kMaxSub = 64;       // sub-steps per push encodable in a key

enum { ST_PARTICLE = 1, ST_OVERFLOW = 2, ST_SUBSTEPS = 4 };

__device__ inline uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
__device__ inline double unif_pm1(uint64_t seed, uint64_t a, uint64_t b) {
  uint64_t h = mix64(seed ^ mix64(a ^ mix64(b)));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

// ---------------------------------------------------------------------------
// Seeding: Mesh::SeedFrontParticles -> SeedParticles (mesh.cpp:98-105,
// 413-548): ppc_y x ppc_z lattice at (i+1)/(ppc+1), weight (1/ppc)*n0,
// q2m = -1, at rest; +-jitter from the seed.  Slot order = (cell, seed).
// ---------------------------------------------------------------------------
__global__ void k_seed(double *y, double *z, double *px, double *py, double *pz,
                       int32_t *cell, int64_t *pid, long np, int ppc_y, int ppc_z,
                       double jitter, uint64_t seed) {
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  if (n >= np) return;
  const int ppc = ppc_y * ppc_z;
  const long c = n / ppc;
  const int is = (int)(n - c * ppc);
  const int iy = is % ppc_y, iz = is / ppc_y;
  double yy = (iy + 1.0) / (ppc_y + 1.0);
  double zz = (iz + 1.0) / (ppc_z + 1.0);
  if (jitter != 0.0) {
    yy += jitter * unif_pm1(seed, (uint64_t)n, 1);
    zz += jitter * unif_pm1(seed, (uint64_t)n, 2);
  }
  y[n] = yy;
  z[n] = zz;
  px[n] = 0.0;
  py[n] = 0.0;
  pz[n] = 0.0;
  cell[n] = (int32_t)c;
  pid[n] = n;
}

struct PushArgs {
  const double *y, *z, *px, *py, *pz;
  const int32_t *cell;
  long np;
  const double *F;
  int ny, nz;
  double hx, hy, hz, weight;
  double hxy, hxz;  // hx / hy, hx / hz
  double bext[3];   // uniform external B (movepart_splitlayer.cpp:1271-1273)
  int shape, fix_sub, full;
  int32_t *ev_cell, *ev_rank, *ev_sub;
  double *ev_y, *ev_z, *ev_w, *ev_p;
  int32_t *ov_cell, *ov_rank;
  double *ov_y, *ov_z, *ov_w;
  uint64_t *ov_key;
  long ov_cap;
  int32_t *counters;
  int32_t *cnt, *cnt_main;
};

__device__ inline void gather6(const double *F, long plane, long n, double *f) {
#pragma unroll
  for (int q = 0; q < kPushComp; ++q) f[q] = F[q * plane + n];
}

__global__ __launch_bounds__(NT) void k_push(PushArgs a) {
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  if (n >= a.np) return;
  const int ny = a.ny, nz = a.nz;
  const double hx = a.hx, hy = a.hy, hz = a.hz;
  const double q2m = -1.0, djx0 = -1.0, djy0 = -1.0, djz0 = -1.0, drho0 = -1.0;
  const double weight = a.weight;
  const int32_t c0 = a.cell[n];
  const int j = c0 % ny, k = c0 / ny;
  const double yp = a.y[n], zp = a.z[n];
  // reference exit(-212) (:945-951); NaN also fails the test
  if (!(yp >= 0 && yp <= 1 && zp >= 0 && zp <= 1)) atomicOr(&a.counters[1], ST_PARTICLE);
  double px = a.px[n], py = a.py[n], pz = a.pz[n];
  const double gammap = sqrt(1. + px * px + py * py + pz * pz);
  // one reciprocal per denominator, then products (the tiled push's arithmetic)
  const double igp = 1. / gammap;
  double Vx = px * igp;
  const double Vxp = Vx;
  const double Vyp = py * igp;
  const double Vzp = pz * igp;
  double f[6];
  if (a.shape == QSS_SHAPE_NGP_REF) {
    gather6(a.F, (long)ny * nz, c0, f);
  } else {
    const int j1 = (j + 1 == ny) ? 0 : j + 1;
    const int k1 = (k + 1 == nz) ? 0 : k + 1;
    double f00[6], f10[6], f01[6], f11[6];
    gather6(a.F, (long)ny * nz, c0, f00);
    gather6(a.F, (long)ny * nz, j1 + (long)ny * k, f10);
    gather6(a.F, (long)ny * nz, j + (long)ny * k1, f01);
    gather6(a.F, (long)ny * nz, j1 + (long)ny * k1, f11);
    const double ayc = 1. - yp, ayp = yp, azc = 1. - zp, azp = zp;
    const double acc = ayc * azc, apc = ayp * azc, acp = ayc * azp, app = ayp * azp;
#pragma unroll
    for (int q = 0; q < 6; ++q)
      f[q] = acc * f00[q] + apc * f10[q] + acp * f01[q] + app * f11[q];
  }
  double ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
  ex *= q2m * hx / 2.;
  ey *= q2m * hx / 2.;
  ez *= q2m * hx / 2.;
  const double iu = 1. / (1. - Vx);
  const double kx = ex * iu, ky = ey * iu, kz = ez * iu;
  px += kx;
  py += ky;
  pz += kz;
  double gamma = sqrt(1. + px * px + py * py + pz * pz);
  const double gamma_r = 1. / gamma;
  bx += a.bext[0];
  by += a.bext[1];
  bz += a.bext[2];
  const double beta = gamma_r * q2m * iu * (0.5 * hx);
  bx = bx * beta;
  by = by * beta;
  bz = bz * beta;
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
  px += kx + px_new;
  py += ky + py_new;
  pz += kz + pz_new;
  gamma = sqrt(1. + px * px + py * py + pz * pz);
  const double ig = 1. / gamma;
  const double Vxm = px * ig;
  const double Vym = py * ig;
  const double Vzm = pz * ig;
  Vx = 0.5 * (Vxm + Vxp);
  const double Vy = 0.5 * (Vym + Vyp);
  const double Vz = 0.5 * (Vzm + Vzp);
  const double iv = 1. / (1. - Vx);
  double djx = weight * djx0 * Vxm * iv;
  double djy = weight * djy0 * Vym * iv;
  double djz = weight * djz0 * Vzm * iv;
  double drho = weight * drho0 * iv;
  double dy = Vy * a.hxy * iv;
  double dz = Vz * a.hxz * iv;
  // sub-cycling, movepart_splitlayer.cpp:1380-1466
  double partdx = 0., step = 1.;
  if (fabs(dy) > 1. || fabs(dz) > 1.) {
    if (fabs(dy) > fabs(dz)) step = partdx = fabs(dy);
    else step = partdx = fabs(dz);
  }
  if (partdx < 1.) partdx = step = 1.;
  // bound the loops below: more than kMaxSub-1 sub-steps (or a non-finite
  // displacement) is reported as an error instead of being followed
  if (!(step < (double)(kMaxSub - 1))) {
    atomicOr(&a.counters[1], ST_SUBSTEPS);
    partdx = step = 1.;
    dy = 0.;
    dz = 0.;
  }
  double ytmp = yp, ztmp = zp;
  int jcur = j, kcur = k;
  int sub = 0;
  while (partdx > 0.) {
    double part_step;
    bool last;
    if (partdx > 1.) { partdx -= 1.; part_step = 1. / step; last = false; }
    else { part_step = partdx / step; partdx = 0.; last = true; }
    const int jb = a.fix_sub ? jcur : j;   // quirk: restart from the start cell
    const int kb = a.fix_sub ? kcur : k;
    ytmp += dy * part_step + jb;
    ztmp += dz * part_step + kb;
    if (!(ytmp > -2.0 * ny && ytmp < 3.0 * ny && ztmp > -2.0 * nz && ztmp < 3.0 * nz)) {
      atomicOr(&a.counters[1], ST_PARTICLE);   // non-finite: keep the wrap bounded
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
    if (jj >= ny) jj -= ny;   // y == Ny exactly: cell 0 (not the ghost)
    if (kk >= nz) kk -= nz;
    const int32_t tc = jj + ny * kk;
    const double w0 = djx * part_step, w1 = djy * part_step, w2 = djz * part_step,
                 w3 = drho * part_step;
    if (last) {
      const int rank = atomicAdd(&a.cnt[tc], 1);
      a.ev_cell[n] = tc;
      a.ev_rank[n] = rank;
      a.ev_sub[n] = sub;
      a.ev_y[n] = ytmp;
      a.ev_z[n] = ztmp;
      reinterpret_cast<double4 *>(a.ev_w)[n] = make_double4(w0, w1, w2, w3);
      if (a.full) {
        a.ev_p[3 * n + 0] = px;
        a.ev_p[3 * n + 1] = py;
        a.ev_p[3 * n + 2] = pz;
        atomicAdd(&a.cnt_main[tc], 1);
      }
    } else {
      // an overflow slot is taken first; only a written event is counted, so
      // every counted (cell, rank) slot of the list gets filled
      const int e = atomicAdd(&a.counters[0], 1);
      if (e < a.ov_cap) {
        const int rank = atomicAdd(&a.cnt[tc], 1);
        a.ov_cell[e] = tc;
        a.ov_rank[e] = rank;
        a.ov_y[e] = ytmp;
        a.ov_z[e] = ztmp;
        reinterpret_cast<double4 *>(a.ov_w)[e] = make_double4(w0, w1, w2, w3);
        a.ov_key[e] = (uint64_t)n * kMaxSub + (uint64_t)sub;
      } else {
        atomicOr(&a.counters[1], ST_OVERFLOW);
      }
    }
    jcur = jj;
    kcur = kk;
    ++sub;
  }
}

// ---------------------------------------------------------------------------
// exclusive scan of the per-cell counts (two levels: 1024-cell blocks, then
// the block totals in one workgroup)
// ---------------------------------------------------------------------------
__device__ inline int wave_incl_scan(int v, int lane) {
  for (int o = 1; o < 64; o <<= 1) {
    int u = __shfl_up(v, o, 64);
    if (lane >= o) v += u;
  }
  return v;
}

// block-wide exclusive scan of 4 values per thread; returns the block total
__device__ int block_scan4(int v[4], int *sh) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int s = v[0] + v[1] + v[2] + v[3];
  int inc = wave_incl_scan(s, lane);
  if (lane == 63) sh[w] = inc;
  __syncthreads();
  int wo = 0, tot = 0;
  for (int q = 0; q < NT / 64; ++q) {
    if (q < w) wo += sh[q];
    tot += sh[q];
  }
  int ex = wo + inc - s;
  int r0 = v[0], r1 = v[1], r2 = v[2];
  v[0] = ex;
  v[1] = ex + r0;
  v[2] = ex + r0 + r1;
  v[3] = ex + r0 + r1 + r2;
  __syncthreads();
  return tot;
}

__global__ __launch_bounds__(NT) void k_scan_block(const int32_t *cnt, int32_t *off,
                                                   int32_t *blk, const int32_t *cnt2,
                                                   int32_t *off2, int32_t *blk2,
                                                   long ncell) {
  __shared__ int sh[NT / 64];
  const long base = (long)blockIdx.x * SCAN_B + threadIdx.x * 4;
  int v[4];
  for (int q = 0; q < 4; ++q) v[q] = (base + q < ncell) ? cnt[base + q] : 0;
  int tot = block_scan4(v, sh);
  for (int q = 0; q < 4; ++q)
    if (base + q < ncell) off[base + q] = v[q];
  if (threadIdx.x == 0) blk[blockIdx.x] = tot;
  if (cnt2) {
    for (int q = 0; q < 4; ++q) v[q] = (base + q < ncell) ? cnt2[base + q] : 0;
    tot = block_scan4(v, sh);
    for (int q = 0; q < 4; ++q)
      if (base + q < ncell) off2[base + q] = v[q];
    if (threadIdx.x == 0) blk2[blockIdx.x] = tot;
  }
}

__global__ __launch_bounds__(NT) void k_scan_top(int32_t *blk, int32_t *blk2, int nblk) {
  __shared__ int sh[NT / 64];
  for (int which = 0; which < 2; ++which) {
    int32_t *b = which ? blk2 : blk;
    if (!b) continue;
    int carry = 0;
    for (int base = 0; base < nblk; base += SCAN_B) {
      int v[4];
      const int i0 = base + threadIdx.x * 4;
      for (int q = 0; q < 4; ++q) v[q] = (i0 + q < nblk) ? b[i0 + q] : 0;
      int tot = block_scan4(v, sh);
      for (int q = 0; q < 4; ++q)
        if (i0 + q < nblk) b[i0 + q] = v[q] + carry;
      carry += tot;
      __syncthreads();
    }
  }
}

__device__ inline int cell_off(const int32_t *off, const int32_t *blk, long c) {
  return off[c] + blk[c / SCAN_B];
}

__global__ __launch_bounds__(NT) void k_scatter(const int32_t *ev_cell,
                                                const int32_t *ev_rank,
                                                const int32_t *ov_cell,
                                                const int32_t *ov_rank,
                                                const int32_t *counters, long np,
                                                long ov_cap, const int32_t *off,
                                                const int32_t *blk, int32_t *list) {
  const long e = (long)blockIdx.x * NT + threadIdx.x;
  long nov = counters[0];
  if (nov > ov_cap) nov = ov_cap;
  if (e >= np + nov) return;
  int32_t c, r;
  if (e < np) { c = ev_cell[e]; r = ev_rank[e]; }
  else { c = ov_cell[e - np]; r = ov_rank[e - np]; }
  list[cell_off(off, blk, c) + r] = (int32_t)e;
}

struct ReduceArgs {
  int ny, nz, shape, full;
  long np;
  const int32_t *cnt, *off, *blk, *off_main, *blk_main, *list;
  const int32_t *ev_sub;
  const double *ev_y, *ev_z, *ev_w, *ev_p;
  const uint64_t *ov_key;
  const double *ov_y, *ov_z, *ov_w;
  double *J[4];  // C-layer Jx, Jy, Jz, Dens
  // full step: reordered particle state
  const int64_t *pid_in;
  double *y, *z, *px, *py, *pz;
  int32_t *cell;
  int64_t *pid;
};

__device__ inline uint64_t ev_key(const ReduceArgs &a, int32_t e) {
  return e < a.np ? (uint64_t)e * kMaxSub + (uint64_t)a.ev_sub[e]
                  : a.ov_key[e - a.np];
}

__global__ __launch_bounds__(NT) void k_tile_reduce(ReduceArgs a) {
  __shared__ double cs[HC * HC][16];
  const int j0 = blockIdx.x * TILE, k0 = blockIdx.y * TILE;
  const int ny = a.ny, nz = a.nz;
  const bool cic = a.shape == QSS_SHAPE_CIC;
  for (int ci = threadIdx.x; ci < HC * HC; ci += NT) {
    const int lj = ci % HC, lk = ci / HC;
    int gj = j0 - 1 + lj, gk = k0 - 1 + lk;
    if (gj < 0) gj += ny;
    if (gk < 0) gk += nz;
    double s[16];
#pragma unroll
    for (int q = 0; q < 16; ++q) s[q] = 0.0;
    if (gj < ny && gk < nz) {
      const long gc = gj + (long)ny * gk;
      const int count = a.cnt[gc];
      const int o = cell_off(a.off, a.blk, gc);
      const bool interior = a.full && lj >= 1 && lk >= 1;
      int mslot = interior ? cell_off(a.off_main, a.blk_main, gc) : 0;
      uint64_t last = 0;
      for (int r = 0; r < count; ++r) {
        // next event in key order (selection over the short cell list)
        uint64_t best = ~0ull;
        int32_t be = -1;
        for (int q = 0; q < count; ++q) {
          const int32_t e = a.list[o + q];
          const uint64_t key = ev_key(a, e);
          if ((r == 0 || key > last) && key < best) { best = key; be = e; }
        }
        if (be < 0) break;   // keys are unique by construction; never index -1
        last = best;
        double ey_, ez_;
        double4 w;
        if (be < a.np) {
          ey_ = a.ev_y[be]; ez_ = a.ev_z[be];
          w = reinterpret_cast<const double4 *>(a.ev_w)[be];
        } else {
          const long oe = be - a.np;
          ey_ = a.ov_y[oe]; ez_ = a.ov_z[oe];
          w = reinterpret_cast<const double4 *>(a.ov_w)[oe];
        }
        const double wq[4] = {w.x, w.y, w.z, w.w};
        if (cic) {
          const double ayp = ey_, ayc = 1. - ayp, azp = ez_, azc = 1. - azp;
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            s[0 + q] += wq[q] * ayc * azc;
            s[4 + q] += wq[q] * ayp * azc;
            s[8 + q] += wq[q] * ayc * azp;
            s[12 + q] += wq[q] * ayp * azp;
          }
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) s[q] += wq[q];
        }
        if (interior && be < a.np) {
          // particle moves to the C layer: storage sorted by (cell, key)
          a.y[mslot] = ey_;
          a.z[mslot] = ez_;
          a.px[mslot] = a.ev_p[3 * (long)be + 0];
          a.py[mslot] = a.ev_p[3 * (long)be + 1];
          a.pz[mslot] = a.ev_p[3 * (long)be + 2];
          a.cell[mslot] = (int32_t)gc;
          a.pid[mslot] = a.pid_in[be];
          ++mslot;
        }
      }
    }
#pragma unroll
    for (int q = 0; q < 16; ++q) cs[ci][q] = s[q];
  }
  __syncthreads();
  const int tj = threadIdx.x % TILE, tk = threadIdx.x / TILE;
  const int nj = j0 + tj, nk = k0 + tk;
  if (nj < ny && nk < nz) {
    const int l00 = (tk + 1) * HC + (tj + 1);
    const int l10 = (tk + 1) * HC + tj;
    const int l01 = tk * HC + (tj + 1);
    const int l11 = tk * HC + tj;
    const long n = nj + (long)ny * nk;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      a.J[q][n] = ((cs[l00][q] + cs[l10][4 + q]) + cs[l01][8 + q]) + cs[l11][12 + q];
  }
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }

}  // namespace

hipError_t launch_seed_plasma(qss_ctx *c) {
  const int b = c->cur;
  hipLaunchKernelGGL(k_seed, grid1(c->np), dim3(NT), 0, c->stream, c->py_[b], c->pz_[b],
                     c->ppx[b], c->ppy[b], c->ppz[b], c->pcell[b], c->pid[b], c->np,
                     c->p.ppc_y, c->p.ppc_z, c->p.jitter, (uint64_t)c->p.seed);
  return hipGetLastError();
}

hipError_t launch_push(qss_ctx *c, int full_step) {
  const int b = c->cur;
  PushArgs a;
  a.y = c->py_[b]; a.z = c->pz_[b]; a.px = c->ppx[b]; a.py = c->ppy[b]; a.pz = c->ppz[b];
  a.cell = c->pcell[b];
  a.np = c->np;
  a.F = c->F;
  a.ny = c->ny; a.nz = c->nz;
  a.hx = c->hxs; a.hy = c->hy; a.hz = c->hz; a.weight = c->pweight;
  a.hxy = c->hxs / c->hy; a.hxz = c->hxs / c->hz;
  for (int q = 0; q < 3; ++q) a.bext[q] = c->p.b_ext[q];
  a.shape = c->p.shape; a.fix_sub = c->p.fix_subcycle; a.full = full_step;
  a.ev_cell = c->ev_cell; a.ev_rank = c->ev_rank; a.ev_sub = c->ev_sub;
  a.ev_y = c->ev_y; a.ev_z = c->ev_z; a.ev_w = c->ev_w; a.ev_p = c->ev_p;
  a.ov_cell = c->ov_cell; a.ov_rank = c->ov_rank; a.ov_y = c->ov_y; a.ov_z = c->ov_z;
  a.ov_w = c->ov_w; a.ov_key = c->ov_key; a.ov_cap = c->ov_cap;
  a.counters = c->d_counters;
  a.cnt = c->cnt; a.cnt_main = c->cnt_main;
  hipLaunchKernelGGL(k_push, grid1(c->np), dim3(NT), 0, c->stream, a);
  return hipGetLastError();
}

hipError_t launch_sort_scan(qss_ctx *c, int full_step) {
  hipLaunchKernelGGL(k_scan_block, dim3(c->nblk), dim3(NT), 0, c->stream, c->cnt, c->off,
                     c->blk, full_step ? c->cnt_main : nullptr,
                     full_step ? c->off_main : nullptr, full_step ? c->blk_main : nullptr,
                     c->ncell);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(NT), 0, c->stream, c->blk,
                     full_step ? c->blk_main : nullptr, c->nblk);
  return hipGetLastError();
}

hipError_t launch_scatter(qss_ctx *c) {
  hipLaunchKernelGGL(k_scatter, grid1(c->np + c->ov_cap), dim3(NT), 0, c->stream,
                     c->ev_cell, c->ev_rank, c->ov_cell, c->ov_rank, c->d_counters, c->np,
                     c->ov_cap, c->off, c->blk, c->list);
  return hipGetLastError();
}

hipError_t launch_tile_reduce(qss_ctx *c, int full_step) {
  ReduceArgs a;
  a.ny = c->ny; a.nz = c->nz; a.shape = c->p.shape; a.full = full_step;
  a.np = c->np;
  a.cnt = c->cnt; a.off = c->off; a.blk = c->blk;
  a.off_main = c->off_main; a.blk_main = c->blk_main;
  a.list = c->list;
  a.ev_sub = c->ev_sub; a.ev_y = c->ev_y; a.ev_z = c->ev_z; a.ev_w = c->ev_w;
  a.ev_p = c->ev_p;
  a.ov_key = c->ov_key; a.ov_y = c->ov_y; a.ov_z = c->ov_z; a.ov_w = c->ov_w;
  a.J[0] = c->C[QSS_JX]; a.J[1] = c->C[QSS_JY]; a.J[2] = c->C[QSS_JZ];
  a.J[3] = c->C[QSS_RHO];
  const int b = c->cur, nb = c->cur ^ 1;
  a.pid_in = c->pid[b];
  a.y = c->py_[nb]; a.z = c->pz_[nb]; a.px = c->ppx[nb]; a.py = c->ppy[nb];
  a.pz = c->ppz[nb]; a.cell = c->pcell[nb]; a.pid = c->pid[nb];
  dim3 grid((c->ny + TILE - 1) / TILE, (c->nz + TILE - 1) / TILE);
  hipLaunchKernelGGL(k_tile_reduce, grid, dim3(NT), 0, c->stream, a);
  return hipGetLastError();
}
