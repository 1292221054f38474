// push_tile.hip — tiled predictor-corrector push with a deterministic
// LDS deposit (the default particle path).
//
// Particles are stored sorted by 16x16-cell tile (tile_start[] gives each
// tile's slot range).  One work-group owns one tile for a whole pass:
//   1. it stages the tile's (P+C)/2 field window (17x17 nodes) in LDS, so
//      the CIC/NGP gathers of MoveParticlesLayerSplit
//      (movepart_splitlayer.cpp:1101-1217, movepart_layer_CIC.cpp:166-203)
//      never leave the CU;
//   2. each thread pushes one particle per chunk of 256 (same arithmetic as
//      the reference, -ffp-contract=off);
//   3. every deposit event (one per sub-step, :1391-1466) is keyed by its
//      window cell; the chunk's keys are bitonic-sorted in LDS (register
//      shuffles inside a wave, LDS across waves) and every window node sums
//      the events of its four CIC corner cells in sorted order — a segmented
//      reduction with no atomics (DepositCurrentsInCellSplit :1547-1655 /
//      MoveInCell movepart_layer_CIC.cpp:449-496);
//   4. the 21x21 node window (tile + 2-cell halo + 1) goes to HBM and
//      k_combine adds the windows overlapping each node in a fixed order,
//      which is the periodic ghost fold of ExchangeCurrentsSplit (:1685-1711).
// With the reference's sub-cycle restart every sub-step lands within one
// cell of the start cell, so every event falls inside the window.
//
// Full step: the pushed particle is staged with its destination tile
// direction and its rank among the tile's particles going that way (slot
// order); k_tile_scan / k_permute rebuild the tile-sorted storage: each
// destination tile holds the groups from its 9 source directions in a fixed
// order, each group in source slot order.  The storage order, and so every
// summation order, is a pure function of the input.

#include <hip/hip_ext.h>

#include "qss_internal.h"
#include "split_bf16.h"
#include "tile_sort.h"

namespace {

constexpr int NT = 256;
constexpr int T = kTile;          // cells per tile edge
constexpr int H = 2;              // halo cells
constexpr int W = T + 2 * H + 1;  // window nodes per edge (21)
constexpr int NWIN = W * W;       // 441
constexpr int FW = T + 1;         // gathered field nodes per edge (17)
constexpr int NF = FW * FW;       // 289
constexpr int kMaxSub = 64;
constexpr int kInvalid = 0x7fffffff;

enum { ST_PARTICLE = 1, ST_OVERFLOW = 2, ST_SUBSTEPS = 4, ST_FAR = 8 };

struct TileArgs {
  const double *y, *z, *px, *py, *pz;
  const int32_t *cell;
  const int64_t *pid;
  const int32_t *tile_start;
  int ntj, ntk, ny, nz;
  const double *F;
  const float *Ff;        // fp32 push field (mixed modes)
  double hx, hy, hz, weight;
  double hxy, hxz;        // hx / hy, hx / hz
  double bext[3];         // uniform external B (movepart_splitlayer.cpp:1271-1273)
  int shape, full;
  double *win;            // ntiles x NWIN x 4
  // full step staging (indexed by the old slot)
  double *sy, *sz, *spx, *spy, *spz;
  int32_t *scell, *sdest;  // sdest = direction | rank << 4
  int32_t *tcount;         // ntiles x nsplit x 9
  int32_t *counters;
  long npw;                // particle slots per wake (batched contexts)
  int nsplit;              // work-groups per tile (each takes every nsplit-th chunk)
  uint64_t ny_magic;       // ceil(2^44 / ny): cell / ny = (cell * ny_magic) >> 44
};

// cell -> (j, k) without a division: for 0 <= c < 2^28 and ny <= 2^14,
// c * ceil(2^44/ny) / 2^44 exceeds c/ny by less than 2^-16 < 1/ny, so the
// floor is exact
__device__ inline void cell_jk(int32_t c, int ny, uint64_t magic, int &j, int &k) {
  k = (int)(((uint64_t)(uint32_t)c * magic) >> 44);
  j = c - k * ny;
}

__device__ inline int wrapi(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

// the arguments of wake w of a batch: every per-wake array is stored
// wake-major with a fixed stride
__host__ __device__ inline TileArgs wake_view(const TileArgs &a0, int w) {
  TileArgs a = a0;
  const long pw = (long)w * a0.npw, nt = (long)a0.ntj * a0.ntk;
  const long plane = (long)a0.ny * a0.nz;
  a.y += pw; a.z += pw; a.px += pw; a.py += pw; a.pz += pw;
  a.cell += pw; a.pid += pw;
  a.tile_start += w * (nt + 1);
  if (a.F) a.F += w * kPushComp * plane;
  if (a.Ff) a.Ff += w * kPushComp * plane;
  a.win += w * nt * a0.nsplit * NWIN * 4;
  a.sy += pw; a.sz += pw; a.spx += pw; a.spy += pw; a.spz += pw;
  a.scell += pw; a.sdest += pw;
  a.tcount += w * nt * a0.nsplit * 9;
  return a;
}

using qss_tile::sort256;

// R: arithmetic of the push (double: the reference's; float: mixed modes).
// FT: storage of the push field.  Positions and momenta are carried in
// double in both: the float form computes each slice's kick, rotation and
// displacement in float and adds them to the double state, and the deposit
// weights are summed in double.  With R = double every expression below
// evaluates exactly as the reference's (no contraction, same association).
//
// A tile's particles may be shared by nsplit work-groups (part p takes
// chunks p, p + nsplit, ...): each part writes its own window, direction
// counts and ranks, and the combine / reorder kernels visit the parts in
// order, so results are a pure function of the input and nsplit.
// 128 VGPRs at most: four waves per SIMD, so all the tiles of a 512^2 grid
// (4 work-groups per CU) are resident at once.
// RHO = false: the deposit sums Jx, Jy, Jz only and the windows' density
// component stays 0 (inside a layer step nothing reads the deposited density:
// the corrector takes rho from continuity in spectral space).
template <typename R, typename FT, bool RHO>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
void k_push_tile(TileArgs a0) {
  const TileArgs a = wake_view(a0, blockIdx.y);   // wake of a batch
  const long pw = (long)blockIdx.y * a0.npw;
  __shared__ FT Fw[NF * 6];
  __shared__ int keys_sh[NT];
  __shared__ int sorted[NT];
  __shared__ double ev_y[NT], ev_z[NT];
  __shared__ double4 ev_w[NT];
  __shared__ short cstart[NWIN], cend[NWIN];
  __shared__ int dir_total[9];
  __shared__ int wave_cnt[NT / 64][9];
  __shared__ int flag_sh;

  const int t = threadIdx.x;
  const int tile = blockIdx.x / a.nsplit, part = blockIdx.x - tile * a.nsplit;
  const int tj = tile % a.ntj, tk = tile / a.ntj;
  const int j0 = tj * T, k0 = tk * T;
  const int ny = a.ny, nz = a.nz;
  const double hx = a.hx, hy = a.hy, hz = a.hz;
  const double q2m = -1.0, djx0 = -1.0, djy0 = -1.0, djz0 = -1.0, drho0 = -1.0;
  const double weight = a.weight;
  const bool cic = a.shape == QSS_SHAPE_CIC;
  int status = 0;   // ST_* bits met by this thread, reported once at the end

  // 1. field window: nodes (j0..j0+T, k0..k0+T), periodic
  for (int i = t; i < NF; i += NT) {
    const int lj = i % FW, lk = i / FW;
    const long g = wrapi(j0 + lj, ny) + (long)ny * wrapi(k0 + lk, nz);
    const long plane = (long)ny * nz;
#pragma unroll
    for (int q = 0; q < kPushComp; ++q) {
      if constexpr (sizeof(FT) == 8) Fw[i * 6 + q] = a.F[q * plane + g];
      else Fw[i * 6 + q] = a.Ff[q * plane + g];
    }
  }
  if (t < 9) dir_total[t] = 0;
  double acc0[4] = {0, 0, 0, 0}, acc1[4] = {0, 0, 0, 0};
  const int n0 = t, n1 = t + NT;   // owned window nodes
  __syncthreads();

  const int base = a.tile_start[tile];
  const int np = a.tile_start[tile + 1] - base;
  for (int c0 = part * NT; c0 < np; c0 += a.nsplit * NT) {
    const int slot = base + c0 + t;
    const long gs = pw + slot;   // the wake's particle arrays, unmoved pointers
    const bool active = c0 + t < np;
    int j = 0, k = 0;
    double yp = 0, zp = 0, px = 0, py = 0, pz = 0;
    R djx = 0, djy = 0, djz = 0, drho = 0, dy = 0, dz = 0, partdx = 0, step = 1;
    if (active) {
      const int32_t c = a0.cell[gs];
      cell_jk(c, ny, a.ny_magic, j, k);
      yp = a0.y[gs];
      zp = a0.z[gs];
      if (!(yp >= 0 && yp <= 1 && zp >= 0 && zp <= 1)) status |= (ST_PARTICLE);
      px = a0.px[gs]; py = a0.py[gs]; pz = a0.pz[gs];
      const R rpx = (R)px, rpy = (R)py, rpz = (R)pz;
      const R gammap = sqrt(R(1) + rpx * rpx + rpy * rpy + rpz * rpz);
      // one correctly rounded reciprocal per denominator, then products: the
      // reference divides by the same gamma and 1 - Vx many times over
      // (movepart_splitlayer.cpp:1239-1330); results agree to an ulp per step
      const R igp = R(1) / gammap;
      R Vx = rpx * igp;
      const R Vxp = Vx, Vyp = rpy * igp, Vzp = rpz * igp;
      // gather from the staged window: local node (j - j0, k - k0)
      int lj = j - j0, lk = k - k0;
      if (lj < 0 || lj >= T || lk < 0 || lk >= T) {   // not in this tile: never expected
        status |= (ST_FAR);
        lj = min(max(lj, 0), T - 1);
        lk = min(max(lk, 0), T - 1);
      }
      R f[6];
      const FT *f00 = &Fw[(lj + FW * lk) * 6];
      if (!cic) {
#pragma unroll
        for (int q = 0; q < 6; ++q) f[q] = (R)f00[q];
      } else {
        const FT *f10 = f00 + 6, *f01 = f00 + 6 * FW, *f11 = f01 + 6;
        const R ayc = R(1. - yp), ayp = R(yp), azc = R(1. - zp), azp = R(zp);
        const R acc = ayc * azc, apc = ayp * azc, acp = ayc * azp, app = ayp * azp;
#pragma unroll
        for (int q = 0; q < 6; ++q)
          f[q] = acc * (R)f00[q] + apc * (R)f10[q] + acp * (R)f01[q] + app * (R)f11[q];
      }
      R ex = f[0], ey = f[1], ez = f[2], bx = f[3], by = f[4], bz = f[5];
      ex *= R(q2m * hx / 2.);
      ey *= R(q2m * hx / 2.);
      ez *= R(q2m * hx / 2.);
      // half kick, carried in the double state
      const R iu = R(1) / (R(1) - Vx);
      const R kx = ex * iu, ky = ey * iu, kz = ez * iu;
      px += (double)kx;
      py += (double)ky;
      pz += (double)kz;
      const R hpx = (R)px, hpy = (R)py, hpz = (R)pz;
      R gamma = sqrt(R(1) + hpx * hpx + hpy * hpy + hpz * hpz);
      const R gamma_r = R(1) / gamma;
      bx += R(a.bext[0]);
      by += R(a.bext[1]);
      bz += R(a.bext[2]);
      const R beta = gamma_r * R(q2m) * iu * R(0.5 * hx);
      bx = bx * beta;
      by = by * beta;
      bz = bz * beta;
      const R co = R(2) / (R(1) + (bx * bx) + (by * by) + (bz * bz));
      R p3x = hpy * bz - hpz * by + hpx;
      R p3y = hpz * bx - hpx * bz + hpy;
      R p3z = hpx * by - hpy * bx + hpz;
      p3x *= co;
      p3y *= co;
      p3z *= co;
      const R px_new = p3y * bz - p3z * by;
      const R py_new = p3z * bx - p3x * bz;
      const R pz_new = p3x * by - p3y * bx;
      px += (double)(kx + px_new);
      py += (double)(ky + py_new);
      pz += (double)(kz + pz_new);
      const R npx = (R)px, npy = (R)py, npz = (R)pz;
      gamma = sqrt(R(1) + npx * npx + npy * npy + npz * npz);
      const R ig = R(1) / gamma;
      const R Vxm = npx * ig, Vym = npy * ig, Vzm = npz * ig;
      Vx = R(0.5) * (Vxm + Vxp);
      const R Vy = R(0.5) * (Vym + Vyp), Vz = R(0.5) * (Vzm + Vzp);
      const R iv = R(1) / (R(1) - Vx);
      djx = R(weight * djx0) * Vxm * iv;
      djy = R(weight * djy0) * Vym * iv;
      djz = R(weight * djz0) * Vzm * iv;
      drho = R(weight * drho0) * iv;
      dy = Vy * R(a.hxy) * iv;
      dz = Vz * R(a.hxz) * iv;
      if (fabs(dy) > R(1) || fabs(dz) > R(1)) {
        if (fabs(dy) > fabs(dz)) step = partdx = fabs(dy);
        else step = partdx = fabs(dz);
      }
      if (partdx < R(1)) partdx = step = R(1);
      if (!(step < (R)(kMaxSub - 1))) {
        status |= (ST_SUBSTEPS);
        partdx = step = R(1);
        dy = R(0);
        dz = R(0);
      }
    }
    double ytmp = yp, ztmp = zp;
    int jj = j, kk = k;
    // sub-step rounds: round r deposits every particle's r-th sub-step
    for (;;) {
      const bool have = active && partdx > 0.;
      int key = kInvalid;
      if (have) {
        R part_step;
        if (partdx > R(1)) { partdx -= R(1); part_step = R(1) / step; }
        else { part_step = partdx / step; partdx = R(0); }
        // the reference restarts every sub-step from the start cell
        ytmp += (double)(dy * part_step) + j;
        ztmp += (double)(dz * part_step) + k;
        if (!(ytmp > -2.0 * ny && ytmp < 3.0 * ny && ztmp > -2.0 * nz && ztmp < 3.0 * nz)) {
          status |= (ST_PARTICLE);
          ytmp = 0.0;
          ztmp = 0.0;
        }
        while (ytmp > ny) ytmp -= ny;
        while (ytmp < 0) ytmp += ny;
        while (ztmp > nz) ztmp -= nz;
        while (ztmp < 0) ztmp += nz;
        jj = (int)ytmp;
        kk = (int)ztmp;
        ytmp -= jj;
        ztmp -= kk;
        if (jj >= ny) jj -= ny;
        if (kk >= nz) kk -= nz;
        // window cell: any periodic image with both CIC nodes inside
        int dj = jj - j0, dk = kk - k0;
        if (dj + H < 0) dj += ny;
        else if (dj + H > W - 2) dj -= ny;
        if (dk + H < 0) dk += nz;
        else if (dk + H > W - 2) dk -= nz;
        const int lj = dj + H, lk = dk + H;
        if (lj >= 0 && lj <= W - 2 && lk >= 0 && lk <= W - 2) {
          key = ((lk * W + lj) << 8) | t;
          ev_y[t] = ytmp;
          ev_z[t] = ztmp;
          ev_w[t] = make_double4((double)(djx * part_step), (double)(djy * part_step),
                                 (double)(djz * part_step),
                                 RHO ? (double)(drho * part_step) : 0.0);
        } else {
          status |= (ST_FAR);
        }
      }
      // segmented reduction of this round's events
      for (int i = t; i < NWIN; i += NT) { cstart[i] = 0; cend[i] = 0; }
      const int sk = sort256(key, keys_sh);
      sorted[t] = sk;
      // the event at sorted position t, moved there in place: gathered before
      // the barrier (the sort's barriers made every event visible), stored after
      // it, so the node sums below read events without the sorted[] indirection
      double gy = 0, gz = 0;
      double4 gw = make_double4(0.0, 0.0, 0.0, 0.0);
      if (sk != kInvalid) {
        const int e = sk & 255;
        gy = ev_y[e];
        gz = ev_z[e];
        gw = ev_w[e];
      }
      __syncthreads();
      if (sk != kInvalid) {
        ev_y[t] = gy;
        ev_z[t] = gz;
        ev_w[t] = gw;
        const int cellw = sk >> 8;
        const int prev = t > 0 ? sorted[t - 1] : kInvalid;
        const int next = t + 1 < NT ? sorted[t + 1] : kInvalid;
        if (t == 0 || (prev >> 8) != cellw) cstart[cellw] = (short)t;
        if (next == kInvalid || (next >> 8) != cellw) cend[cellw] = (short)(t + 1);
      }
      __syncthreads();
#pragma unroll
      for (int o = 0; o < 2; ++o) {
        const int n = o == 0 ? n0 : n1;
        if (n >= NWIN) continue;
        double *accn = o == 0 ? acc0 : acc1;
        const int nl_j = n % W, nl_k = n / W;
        // corners: own cell (1-y)(1-z); left y(1-z); below (1-y)z; both yz
#pragma unroll
        for (int cn = 0; cn < 4; ++cn) {
          if (!cic && cn > 0) break;
          const int cj = nl_j - (cn & 1), ck = nl_k - (cn >> 1);
          if (cj < 0 || ck < 0) continue;
          const int cw = ck * W + cj;
          const int e1 = cend[cw];
          for (int i = cstart[cw]; i < e1; ++i) {
            const int e = i;   // events are stored in sorted order
            const double4 wv = ev_w[e];
            const double wq[4] = {wv.x, wv.y, wv.z, wv.w};
            if (cic) {
              const double ayp = ev_y[e], ayc = 1. - ayp, azp = ev_z[e], azc = 1. - azp;
              const double wy = (cn & 1) ? ayp : ayc;
              const double wz = (cn >> 1) ? azp : azc;
#pragma unroll
              for (int q = 0; q < (RHO ? 4 : 3); ++q) accn[q] += wq[q] * wy * wz;
            } else {
#pragma unroll
              for (int q = 0; q < (RHO ? 4 : 3); ++q) accn[q] += wq[q];
            }
          }
        }
      }
      if (t == 0) flag_sh = 0;
      __syncthreads();
      if (active && partdx > 0.) flag_sh = 1;
      __syncthreads();
      const int more = flag_sh;
      __syncthreads();
      if (!more) break;
    }
    if (a.full) {
      // destination tile direction (canonical for 1 or 2 tiles per axis)
      int dtj = 0, dtk = 0;
      if (active) {
        const int ntj_ = a.ntj, ntk_ = a.ntk;
        // |jj/T - tj| < ntj: one correction wraps it
        int d1 = jj / T - tj, d2 = kk / T - tk;
        d1 += d1 < 0 ? ntj_ : 0;
        d2 += d2 < 0 ? ntk_ : 0;
        dtj = d1 == 0 ? 0 : (d1 == ntj_ - 1 && ntj_ > 2 ? -1 : 1);
        dtk = d2 == 0 ? 0 : (d2 == ntk_ - 1 && ntk_ > 2 ? -1 : 1);
        if (d1 > 1 && d1 < ntj_ - 1) status |= (ST_FAR);
        if (d2 > 1 && d2 < ntk_ - 1) status |= (ST_FAR);
      }
      const int dir = (dtj + 1) + 3 * (dtk + 1);
      // rank among this tile's particles with the same direction, slot order
      const int lane = t & 63, wv = t >> 6;
      const unsigned long long below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
      int myrank_w = 0;
      // kept rolled: nine live ballot masks would spill scalar registers
#pragma unroll 1
      for (int d = 0; d < 9; ++d) {
        const unsigned long long m = __ballot(active && dir == d);
        if (lane == 0) wave_cnt[wv][d] = __popcll(m);
        if (active && dir == d) myrank_w = __popcll(m & below);
      }
      __syncthreads();
      int off = dir_total[dir];
      for (int q = 0; q < wv; ++q) off += wave_cnt[q][dir];
      if (active) {
        a0.sy[gs] = ytmp;
        a0.sz[gs] = ztmp;
        a0.spx[gs] = px;
        a0.spy[gs] = py;
        a0.spz[gs] = pz;
        a0.scell[gs] = jj + ny * kk;
        a0.sdest[gs] = dir | (part << 4) | ((off + myrank_w) << 8);
      }
      __syncthreads();
      if (t < 9) {
        int s = 0;
        for (int q = 0; q < NT / 64; ++q) s += wave_cnt[q][t];
        dir_total[t] += s;
      }
      __syncthreads();
    }
  }
  if (status) atomicOr(&a.counters[1], status);
  // 4. window to HBM
  double *wout = a.win + (long)blockIdx.x * NWIN * 4;
  reinterpret_cast<double4 *>(wout)[n0] = make_double4(acc0[0], acc0[1], acc0[2], acc0[3]);
  if (n1 < NWIN)
    reinterpret_cast<double4 *>(wout)[n1] = make_double4(acc1[0], acc1[1], acc1[2], acc1[3]);
  if (a.full && t < 9) a.tcount[blockIdx.x * 9 + t] = dir_total[t];
}

// sum of the windows covering each node, fixed (dtk, dtj, part) order.  Per
// axis a node lies in its own tile's window (dt = 0) and in at most one
// neighbour's (dt = -1 within W - T - H nodes of the tile's low edge, +1 within
// H of its high edge; the two ranges are disjoint), so at most 4 x NS windows
// cover it: all their loads are issued before the first add.  An absent window
// reads the own tile's entry and adds +0.0 instead, which leaves the sum bitwise
// unchanged (the sum starts at +0.0, so it is never -0.0).
// xs0..xs2 (split-bf16 modes, one wake): Jx, Jy, Jz also written as the split
// operands of the forward transform that follows.
// grid: (node blocks of one plane, wakes); 32-bit node index (ny, nz <= 2^14)
static_assert(W - T - H <= T - H, "at most two tiles per axis cover a node");
// RHO = false: the density plane J3 is neither summed nor written
template <int NS, bool RHO>
__global__ __launch_bounds__(NT) void k_combine(const double *win, int ntj, int ntk, int ny,
                                                int nz, uint64_t ny_magic, double *J0,
                                                double *J1, double *J2, double *J3,
                                                uint16_t *xs0, uint16_t *xs1, uint16_t *xs2,
                                                int npc) {
  const int plane = ny * nz;
  const int n1 = blockIdx.x * NT + threadIdx.x;   // node of the wake's plane
  if (n1 >= plane) return;
  const double4 *w4 = reinterpret_cast<const double4 *>(win) +
                      (long)blockIdx.y * ntj * ntk * NS * NWIN;   // the wake's windows
  const long n = (long)blockIdx.y * plane + n1;                   // wake-major node index
  int nj, nk;
  cell_jk(n1, ny, ny_magic, nj, nk);
  const int tj = nj / T, tk = nk / T;
  const int rj = nj - tj * T, rk = nk - tk * T;   // position in the tile, [0, T)
  // covering tiles per axis, in increasing order: (d0, d0 + 1), the second
  // present when the node is near either edge
  const int dk0 = rk < W - T - H ? -1 : 0, dj0 = rj < W - T - H ? -1 : 0;
  const bool hk = rk < W - T - H || rk >= T - H, hj = rj < W - T - H || rj >= T - H;
  long idx[2][2];
#pragma unroll
  for (int a = 0; a < 2; ++a) {
    const int utk = tk + dk0 + a;                 // unwrapped tile row, in [-1, ntk]
    const int wtk = utk < 0 ? utk + ntk : (utk >= ntk ? utk - ntk : utk);
    const int lk = nk - (utk * T - H);
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      const int utj = tj + dj0 + b;
      const int wtj = utj < 0 ? utj + ntj : (utj >= ntj ? utj - ntj : utj);
      const int lj = nj - (utj * T - H);
      idx[a][b] = (long)(wtk * ntj + wtj) * NS * NWIN + lk * W + lj;
    }
  }
  const bool ok[2][2] = {{true, hj}, {hk, hk && hj}};
  double4 v[2][2][NS];
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < NS; ++q)
        v[a][b][q] = w4[(ok[a][b] ? idx[a][b] : idx[0][0]) + (long)q * NWIN];
  double s[4] = {0, 0, 0, 0};
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b)
#pragma unroll
      for (int q = 0; q < NS; ++q) {
        const double4 u = ok[a][b] ? v[a][b][q] : make_double4(0.0, 0.0, 0.0, 0.0);
        s[0] += u.x; s[1] += u.y; s[2] += u.z;
        if (RHO) s[3] += u.w;
      }
  J0[n] = s[0]; J1[n] = s[1]; J2[n] = s[2];
  if (RHO) J3[n] = s[3];
  if (npc) {
    qss_split::store(xs0, plane, n, s[0], npc);
    qss_split::store(xs1, plane, n, s[1], npc);
    qss_split::store(xs2, plane, n, s[2], npc);
  }
}

// new tile starts and group offsets: destination tile u receives the groups
// of source tiles u - dir(d) for d = 0..8 in that order, each as the groups
// of the source's parts 0..nsplit-1.  One 1024-thread work-group per wake:
// a thread holds one destination tile's 9 x nsplit counts in registers.
constexpr int NTS = 1024;
__global__ __launch_bounds__(NTS) void k_tile_scan(const int32_t *tcount, int ntj, int ntk,
                                                   int nsplit, int32_t *tile_start,
                                                   int32_t *goff) {
  __shared__ int wsum[NTS / 64];
  const int nt = ntj * ntk;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int w = blockIdx.x;   // one work-group per wake
  tcount += (long)w * nt * nsplit * 9;
  goff += (long)w * nt * nsplit * 9;
  tile_start += (long)w * (nt + 1);
  int carry = 0;
  for (int b0 = 0; b0 < nt; b0 += NTS) {
    const int u = b0 + threadIdx.x;
    int tot = 0;
    int g[9][kMaxSplit];
    if (u < nt) {
      const int uj = u % ntj, uk = u / ntj;
#pragma unroll
      for (int d = 0; d < 9; ++d) {
        const int dj = d % 3 - 1, dk = d / 3 - 1;
        const int src = wrapi(uk - dk, ntk) * ntj + wrapi(uj - dj, ntj);
#pragma unroll
        for (int q = 0; q < kMaxSplit; ++q) {
          g[d][q] = q < nsplit ? tcount[(src * nsplit + q) * 9 + d] : 0;
          tot += g[d][q];
        }
      }
    }
    // block exclusive scan of tot: inclusive scans within each wave, then
    // the totals of the waves before this one (integer sums: exact)
    int incl = tot;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const int v = __shfl_up(incl, o, 64);
      if (lane >= o) incl += v;
    }
    if (lane == 63) wsum[wv] = incl;
    __syncthreads();
    int before = 0, blk_tot = 0;
#pragma unroll
    for (int q = 0; q < NTS / 64; ++q) {
      const int v = wsum[q];
      before += q < wv ? v : 0;
      blk_tot += v;
    }
    const int start = carry + before + incl - tot;
    if (u < nt) {
      tile_start[u] = start;
      int o = start;
#pragma unroll
      for (int d = 0; d < 9; ++d)
#pragma unroll
        for (int q = 0; q < kMaxSplit; ++q)
          if (q < nsplit) {
            goff[(u * 9 + d) * nsplit + q] = o;
            o += g[d][q];
          }
    }
    __syncthreads();   // wsum is rewritten by the next block of tiles
    carry += blk_tot;
  }
  if (threadIdx.x == 0) tile_start[nt] = carry;
}

__global__ __launch_bounds__(NT) void k_permute(TileArgs a0, const int32_t *goff, long np,
                                                int nw, double *y, double *z, double *px,
                                                double *py, double *pz, int32_t *cell,
                                                int64_t *pid) {
  const long sg = (long)blockIdx.x * NT + threadIdx.x;   // wake-major slot
  if (sg >= np * nw) return;
  const int w = (int)(sg / np);
  const long s = sg - (long)w * np;
  const TileArgs a = wake_view(a0, w);
  goff += (long)w * a.ntj * a.ntk * a.nsplit * 9;
  const long pw = (long)w * a0.npw;
  y += pw; z += pw; px += pw; py += pw; pz += pw; cell += pw; pid += pw;
  const int32_t c = a.scell[s];
  int cj, ck;
  cell_jk(c, a.ny, a.ny_magic, cj, ck);
  const int dst_tile = (ck / T) * a.ntj + cj / T;
  const int sd = a.sdest[s];   // direction | part << 4 | rank << 8
  const long o = goff[(dst_tile * 9 + (sd & 15)) * a.nsplit + ((sd >> 4) & 15)] + (sd >> 8);
  y[o] = a.sy[s];
  z[o] = a.sz[s];
  px[o] = a.spx[s];
  py[o] = a.spy[s];
  pz[o] = a.spz[s];
  cell[o] = c;
  pid[o] = a.pid[s];   // ids are not staged: the old storage still holds them
}

// seeding in tile-major slot order: tile, local cell (j fastest), seed
__global__ void k_seed_tiled(double *y, double *z, double *px, double *py, double *pz,
                             int32_t *cell, int64_t *pid, long np, int nw, int ny, int ntj,
                             int ppc_y, int ppc_z, double jitter, uint64_t seed) {
  const long ng = (long)blockIdx.x * NT + threadIdx.x;   // wake-major slot
  if (ng >= np * nw) return;
  // every wake of a batch starts from the same plasma
  const long n = ng % np;
  y += ng - n; z += ng - n; px += ng - n; py += ng - n; pz += ng - n;
  cell += ng - n; pid += ng - n;
  const int ppc = ppc_y * ppc_z;
  const long per_tile = (long)T * T * ppc;
  const long tile = n / per_tile;
  const long r = n - tile * per_tile;
  const int lc = (int)(r / ppc), is = (int)(r - (long)lc * ppc);
  const int gj = (int)(tile % ntj) * T + lc % T;
  const int gk = (int)(tile / ntj) * T + lc / T;
  const long c = gj + (long)ny * gk;
  const long id = c * ppc + is;
  const int iy = is % ppc_y, iz = is / ppc_y;
  double yy = (iy + 1.0) / (ppc_y + 1.0);
  double zz = (iz + 1.0) / (ppc_z + 1.0);
  if (jitter != 0.0) {
    // the same counter-based draws as the cell-ordered seeding (by id)
    auto mix64 = [](uint64_t x) {
      x += 0x9E3779B97F4A7C15ull;
      x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
      x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
      return x ^ (x >> 31);
    };
    auto unif = [&](uint64_t a2, uint64_t b2) {
      const uint64_t h = mix64(seed ^ mix64(a2 ^ mix64(b2)));
      return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
    };
    yy += jitter * unif((uint64_t)id, 1);
    zz += jitter * unif((uint64_t)id, 2);
  }
  y[n] = yy;
  z[n] = zz;
  px[n] = 0.0;
  py[n] = 0.0;
  pz[n] = 0.0;
  cell[n] = (int32_t)c;
  pid[n] = id;
}

__global__ void k_uniform_tile_start(int32_t *ts, int nt, int nw, int per_tile) {
  const int i = blockIdx.x * NT + threadIdx.x;   // wake-major
  if (i < (nt + 1) * nw) ts[i] = (i % (nt + 1)) * per_tile;
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }

TileArgs make_args(qss_ctx *c, int full) {
  TileArgs a;
  const int b = c->cur;
  a.y = c->py_[b]; a.z = c->pz_[b]; a.px = c->ppx[b]; a.py = c->ppy[b]; a.pz = c->ppz[b];
  a.cell = c->pcell[b]; a.pid = c->pid[b];
  a.tile_start = c->tile_start[b];
  a.ntj = c->ntj; a.ntk = c->ntk; a.ny = c->ny; a.nz = c->nz;
  a.F = c->F;
  a.Ff = c->Ff;
  a.hx = c->hxs; a.hy = c->hy; a.hz = c->hz; a.weight = c->pweight;
  a.hxy = c->hxs / c->hy; a.hxz = c->hxs / c->hz;
  for (int q = 0; q < 3; ++q) a.bext[q] = c->p.b_ext[q];
  a.shape = c->p.shape; a.full = full;
  a.win = c->win;
  const long npall = c->np_cap * c->nw;
  a.sy = c->ev_y; a.sz = c->ev_z; a.spx = c->ev_p; a.spy = c->ev_p + npall;
  a.spz = c->ev_p + 2 * npall;
  a.scell = c->ev_cell; a.sdest = c->ev_rank;
  a.tcount = c->tcount;
  a.counters = c->d_counters;
  a.npw = c->np_cap;
  a.nsplit = c->nsplit;
  a.ny_magic = ((1ull << 44) + (uint64_t)c->ny - 1) / (uint64_t)c->ny;
  return a;
}

}  // namespace

hipError_t launch_push_tile(qss_ctx *c, int full, bool rho, hipEvent_t ev0, hipEvent_t ev1) {
  const dim3 grid(c->ntj * c->ntk * c->nsplit, c->nw), block(NT);
  auto fn = c->mixed ? (rho ? k_push_tile<float, float, true> : k_push_tile<float, float, false>)
                     : (rho ? k_push_tile<double, double, true>
                            : k_push_tile<double, double, false>);
  if (ev0 || ev1)
    hipExtLaunchKernelGGL(fn, grid, block, 0, c->stream, ev0, ev1, 0, make_args(c, full));
  else
    hipLaunchKernelGGL(fn, grid, block, 0, c->stream, make_args(c, full));
  return hipGetLastError();
}

hipError_t launch_combine(qss_ctx *c, bool rho, double *const *out) {
  double *const J[4] = {c->C[QSS_JX], c->C[QSS_JY], c->C[QSS_JZ], c->C[QSS_RHO]};
  const bool split = c->npc && !out;   // sweep deposits only (not the tuning runs)
  if (!out) out = J;
  if (split)
    for (int b = 0; b < 3; ++b) c->xs_src[b] = J[b];
  const uint64_t ny_magic = ((1ull << 44) + (uint64_t)c->ny - 1) / (uint64_t)c->ny;
  auto fn = rho ? (c->nsplit == 1   ? k_combine<1, true>
                   : c->nsplit == 2 ? k_combine<2, true>
                                    : k_combine<4, true>)
                : (c->nsplit == 1   ? k_combine<1, false>
                   : c->nsplit == 2 ? k_combine<2, false>
                                    : k_combine<4, false>);
  if (c->nsplit != 1 && c->nsplit != 2 && c->nsplit != 4) return hipErrorInvalidValue;
  hipLaunchKernelGGL(fn, dim3(grid1(c->ncell).x, c->nw), dim3(NT), 0, c->stream, c->win,
                     c->ntj, c->ntk, c->ny, c->nz, ny_magic, out[0], out[1], out[2], out[3],
                     c->xs[0], c->xs[1], c->xs[2], split ? c->npc : 0);
  return hipGetLastError();
}

hipError_t launch_tile_reorder(qss_ctx *c) {
  const int nb = c->cur ^ 1;
  hipLaunchKernelGGL(k_tile_scan, dim3(c->nw), dim3(NTS), 0, c->stream, c->tcount, c->ntj,
                     c->ntk, c->nsplit, c->tile_start[nb], c->goff);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_permute, grid1(c->np * c->nw), dim3(NT), 0, c->stream, make_args(c, 1),
                     c->goff, c->np, c->nw, c->py_[nb], c->pz_[nb], c->ppx[nb], c->ppy[nb],
                     c->ppz[nb], c->pcell[nb], c->pid[nb]);
  return hipGetLastError();
}

hipError_t launch_seed_tiled(qss_ctx *c) {
  const int b = c->cur;
  hipLaunchKernelGGL(k_seed_tiled, grid1(c->np * c->nw), dim3(NT), 0, c->stream, c->py_[b],
                     c->pz_[b], c->ppx[b], c->ppy[b], c->ppz[b], c->pcell[b], c->pid[b], c->np,
                     c->nw, c->ny, c->ntj, c->p.ppc_y, c->p.ppc_z, c->p.jitter,
                     (uint64_t)c->p.seed);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  const int nt = c->ntj * c->ntk;
  hipLaunchKernelGGL(k_uniform_tile_start, grid1((long)(nt + 1) * c->nw), dim3(NT), 0,
                     c->stream, c->tile_start[b], nt, c->nw, T * T * c->p.ppc_y * c->p.ppc_z);
  return hipGetLastError();
}
