// beam.hip — frozen driver-beam deposit for one layer (rho_b, Jx_b planes).
//
// The reference seeds beam macro-particles 4 per 3-D cell at
// (x=1/2, {1/3,2/3}^2) with weight = Density(cell corner)/4
// (Mesh::SeedBeamParticles, mesh.cpp:108-163; Specie::Density case 1,
// plasma.cpp:350-377) and deposits them trilinearly into 3-D arrays
// (Mesh::MoveBeamInCell, movebeam.cpp:662-770; GPU twin cuMoveBeamInCell,
// CUDA_WRAP/assign_currents.cu:24-173, which races on Jy/Jz).  For a frozen
// driver the 3-D array is never needed: layer i only receives from x-cells i
// and i-1 (weight 1/2 each), so this kernel gathers, per node, the <= 32
// lattice particles that touch it, in a fixed order.  No atomics, bitwise
// deterministic.  The C-layer beam for split s of NxSplit interpolates layers
// i and i+1 (movepart_splitlayer.cpp:330-348, CUDA_WRAP/cuLayers.cu:416-435).

#include "qss_internal.h"
#include "split_bf16.h"

namespace {

constexpr int NT = 256;

struct BeamArgs {
  int ny, nz, nslices, nbunch, nbins;
  double hx, hy, hz, ly, lz, bin_xi0, bin_xi1;
  uint64_t seed;
  qss_bunch b[QSS_MAX_BUNCH];
  const qss_bunch *wb;   // batched wakes: QSS_MAX_BUNCH bunches per wake (or null)
  const double *bins;
  double *rho, *jx;   // C-layer beam planes
  int layer;          // reference layer index i = nslices-1-slice
  int nsplit, isplit;
  const int *slice_dev;   // graph replay: the slice is read from device memory
  uint16_t *xs0, *xs1;    // split-bf16 modes: the planes as split transform operands
  int npc;
};

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

// Each work-group covers a 16 x 16 node tile.  A node (nj, nk) receives from
// the cells (nj-1 .. nj, nk-1 .. nk), so the tile reads the 17 x 17 cells from
// (j0-1, k0-1): their densities are evaluated once per x-cell and bunch into
// LDS and each node sums its four cells from there.
//
// Density of bunch b at cell centre (xi, yc, zc) (Specie::Density case 1,
// plasma.cpp:350-377, with R = sqrt(2) sigma):
//   arg = dx*dx/Rx/Rx + dy*dy/Ry/Ry + dz*dz/Ry/Ry  (the profile-bin bunch: the
//   last two terms), density nb exp(-arg) (bins[t] exp(-arg)), 0 if arg > 16.
// The three terms depend on the x-cell, the column and the row alone, and the
// sum is evaluated left to right, so each is computed once (per work-group:
// the 17 column and 17 row terms per bunch) and a cell adds them in that
// order: the same values as evaluating the expression per cell (this file is
// compiled without FMA contraction, so no product is fused into a sum).
constexpr int BT = 16;            // nodes per tile edge (NT = BT * BT)
constexpr int BC = BT + 1;        // cells per tile edge
constexpr int NBC = BC * BC;      // cells per tile
static_assert(NT == BT * BT, "one thread per node of a tile");
static_assert(3 * QSS_MAX_BUNCH <= 32, "skip bits");

// x-cell ci lies outside bunch b (exact early out: arg = dx^2/Rx^2 +
// (non-negative terms) exceeds 16 as soon as its first term does, rounded sums
// of non-negative terms do not decrease), or outside the x range: its cells
// add nothing
__device__ inline bool xcell_out(const BeamArgs &a, const qss_bunch *bun, int b, int ci) {
  if (ci < 0 || ci > a.nslices) return true;
  if (b == 0 && a.nbins > 0) return false;
  const qss_bunch &B = bun[b];
  const double xi = (double)(a.nslices - ci) * a.hx;
  const double Rx = sqrt(2.0) * B.sigma_xi;
  const double dx = xi - B.xi0;
  return dx * dx / Rx / Rx > 16;
}

// dn[(x * nbunch + b) * NBC + cell] for x-cells ci0 + x, x < nx; bit x *
// nbunch + b of the result is set when that x-cell and bunch add nothing
// (its entries are then not written).  tyz: 2 * BC terms per bunch (scratch).
__device__ unsigned cell_densities(const BeamArgs &a, const qss_bunch *bun, int ci0, int nx,
                                   int j0, int k0, double *dn, double *tyz) {
  // column (dy) and row (dz) terms of every bunch
  for (int idx = threadIdx.x; idx < a.nbunch * 2 * BC; idx += NT) {
    const int b = idx / (2 * BC), r = idx - b * 2 * BC;
    const qss_bunch &B = bun[b];
    const double Ry = sqrt(2.0) * B.sigma_r;
    double d;
    if (r < BC) {
      const double yc = a.hy * (j0 - 1 + r) - a.ly / 2.0;
      d = yc - B.y0;
    } else {
      const double zc = a.hz * (k0 - 1 + r - BC) - a.lz / 2.0;
      d = zc - B.z0;
    }
    tyz[idx] = d * d / Ry / Ry;
  }
  __syncthreads();
  unsigned skip = 0;
  for (int x = 0; x < nx; ++x)
    for (int b = 0; b < a.nbunch; ++b) {
      const int ci = ci0 + x;
      if (xcell_out(a, bun, b, ci)) {   // uniform over the work-group
        skip |= 1u << (x * a.nbunch + b);
        continue;
      }
      const qss_bunch &B = bun[b];
      const double xi = (double)(a.nslices - ci) * a.hx;
      const bool binned = b == 0 && a.nbins > 0;
      double ax = 0.0, amp = B.nb;   // the x term; the amplitude
      if (binned) {
        const double t = (xi - a.bin_xi0) / (a.bin_xi1 - a.bin_xi0) * a.nbins;
        amp = !(t >= 0.0) || t >= a.nbins ? 0.0 : a.bins[(int)t];
      } else {
        const double Rx = sqrt(2.0) * B.sigma_xi;
        const double dx = xi - B.xi0;
        ax = dx * dx / Rx / Rx;
      }
      const double *ty = tyz + b * 2 * BC, *tz = ty + BC;
      double *d = dn + (x * a.nbunch + b) * NBC;
      for (int cell = threadIdx.x; cell < NBC; cell += NT) {
        const int lj = cell % BC, lk = cell / BC;
        const int cj = j0 - 1 + lj, ck = k0 - 1 + lk;
        double v = 0.0;
        if (cj >= 0 && ck >= 0 && cj < a.ny && ck < a.nz) {
          const double arg = binned ? ty[lj] + tz[lk] : ax + ty[lj] + tz[lk];
          if (!(arg > 16) && amp != 0.0) v = amp * exp(-arg);
        }
        d[cell] = v;
      }
    }
  return skip;
}

// value of reference layer i at node (nj, nk); sums in the order
// x-cell (i, i-1), bunch, ck, cj, seed
__device__ void node_value(const BeamArgs &a, const qss_bunch *bun, int i, int nj, int nk,
                           int ci0, int j0, int k0, const double *dn, unsigned skip,
                           double &rho, double &jx) {
  rho = 0.0;
  jx = 0.0;
  for (int side = 0; side < 2; ++side) {
    const int ci = side == 0 ? i : i - 1;
    const double ax = 0.5;
    for (int b = 0; b < a.nbunch; ++b) {
      const int xb = (ci - ci0) * a.nbunch + b;
      if (skip >> xb & 1u) continue;
      const qss_bunch &B = bun[b];
      const double *dnb = dn + xb * NBC;
      for (int ck = nk - 1; ck <= nk; ++ck) {
        if (ck < 0) continue;
        const int dz_ = nk - ck;
        for (int cj = nj - 1; cj <= nj; ++cj) {
          if (cj < 0) continue;
          const int dy_ = nj - cj;
          const double dnc = dnb[(ck - k0 + 1) * BC + (cj - j0 + 1)];
          if (!(dnc > 0.0)) continue;
          for (int is = 0; is < 4; ++is) {
            const double yq = 1.0 / 3.0 * (1.0 + is % 2);
            const double zq = 1.0 / 3.0 * (1.0 + is / 2);
            double w = 0.25 * dnc;
            if (B.noise != 0.0) {
              const uint64_t cell =
                  ((uint64_t)(ci + 1) * (uint64_t)a.nz + ck) * (uint64_t)a.ny + cj;
              w *= 1.0 + B.noise * unif_pm1(a.seed ^ 0xB3A1ull, cell * 4 + is,
                                            (uint64_t)b + 7);
            }
            const double drho = w * B.polarity;
            const double djx = w * B.polarity * B.vx;
            const double az = dz_ ? zq : 1.0 - zq;
            const double ay = dy_ ? yq : 1.0 - yq;
            jx += djx * ax * ay * az;
            rho += drho * ax * ay * az;
          }
        }
      }
    }
  }
}

// grid: (node tiles of the plane, wakes); dynamic LDS: (nx * NBC + 2 * BC) *
// nbunch doubles
__global__ __launch_bounds__(NT) void k_beam_plane(BeamArgs a) {
  extern __shared__ double dn[];
  const int ntj = (a.ny + BT - 1) / BT;
  const int j0 = (blockIdx.x % ntj) * BT, k0 = (blockIdx.x / ntj) * BT;
  const int nj = j0 + threadIdx.x % BT, nk = k0 + threadIdx.x / BT;
  const long nc = (long)a.ny * a.nz;
  const int w = blockIdx.y;   // wake of a batch
  const qss_bunch *bun = a.wb ? a.wb + (long)w * QSS_MAX_BUNCH : a.b;
  const int layer = a.slice_dev ? a.nslices - 1 - *a.slice_dev : a.layer;
  // x-cells layer - 1 .. layer (+ layer + 1 for the split interpolation)
  const int ci0 = layer - 1;
  const int nx = a.nsplit > 1 ? 3 : 2;
  const unsigned skip =
      cell_densities(a, bun, ci0, nx, j0, k0, dn, dn + nx * a.nbunch * NBC);
  __syncthreads();
  if (nj >= a.ny || nk >= a.nz) return;
  const long n = (long)nk * a.ny + nj;
  double *rho = a.rho + w * nc, *jx = a.jx + w * nc;
  double rc, jc;
  node_value(a, bun, layer, nj, nk, ci0, j0, k0, dn, skip, rc, jc);
  if (a.nsplit > 1) {
    double rp, jp;
    node_value(a, bun, layer + 1, nj, nk, ci0, j0, k0, dn, skip, rp, jp);
    const int S = a.nsplit, s = a.isplit;
    rc = (rc * (S - s) + rp * s) / S;
    jc = (jc * (S - s) + jp * s) / S;
  }
  rho[n] = rc;
  jx[n] = jc;
  if (a.npc) {   // single wake (split modes)
    qss_split::store(a.xs0, nc, n, rc, a.npc);
    qss_split::store(a.xs1, nc, n, jc, a.npc);
  }
}

}  // namespace

hipError_t launch_beam_plane(qss_ctx *c, int slice, int isplit) {
  if (c->drv.on) return launch_driver_plane(c, slice, isplit);   // evolving driver
  BeamArgs a;
  a.ny = c->ny; a.nz = c->nz; a.nslices = c->p.nslices; a.nbunch = c->p.nbunch;
  a.nbins = c->p.nbins;
  a.hx = c->p.hx; a.hy = c->hy; a.hz = c->hz; a.ly = c->p.ly; a.lz = c->p.lz;
  a.bin_xi0 = c->p.bin_xi0; a.bin_xi1 = c->p.bin_xi1;
  a.seed = c->p.seed;
  for (int b = 0; b < QSS_MAX_BUNCH; ++b) a.b[b] = c->p.bunch[b];
  a.wb = c->nw > 1 ? c->d_wbunch : nullptr;
  a.bins = c->d_bins;
  a.rho = c->C[QSS_RHOBEAM];
  a.jx = c->C[QSS_JXBEAM];
  a.layer = c->p.nslices - 1 - slice;
  a.nsplit = c->p.nx_split;
  a.isplit = isplit;
  a.slice_dev = c->capturing ? c->d_slice : nullptr;
  a.xs0 = c->xs[0];
  a.xs1 = c->xs[1];
  a.npc = c->npc;
  if (c->npc) {
    c->xs_src[0] = a.rho;
    c->xs_src[1] = a.jx;
  }
  const unsigned tiles = (unsigned)(((c->ny + BT - 1) / BT) * ((c->nz + BT - 1) / BT));
  const size_t lds = sizeof(double) * (size_t)(a.nbunch > 0 ? a.nbunch : 1) *
                     ((a.nsplit > 1 ? 3 : 2) * NBC + 2 * BC);
  hipLaunchKernelGGL(k_beam_plane, dim3(tiles, c->nw), dim3(NT), lds, c->stream, a);
  return hipGetLastError();
}
