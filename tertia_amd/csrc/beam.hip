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

__device__ double bunch_density(const BeamArgs &a, const qss_bunch *bun, int b, double xi,
                                double yc, double zc) {
  const qss_bunch &B = bun[b];
  const double Ry = sqrt(2.0) * B.sigma_r;
  const double dy = yc - B.y0, dz = zc - B.z0;
  if (b == 0 && a.nbins > 0) {
    const double arg = dy * dy / Ry / Ry + dz * dz / Ry / Ry;
    if (arg > 16) return 0.0;
    const double t = (xi - a.bin_xi0) / (a.bin_xi1 - a.bin_xi0) * a.nbins;
    if (!(t >= 0.0) || t >= a.nbins) return 0.0;
    return a.bins[(int)t] * exp(-arg);
  }
  const double Rx = sqrt(2.0) * B.sigma_xi;
  const double dx = xi - B.xi0;
  const double arg = dx * dx / Rx / Rx + dy * dy / Ry / Ry + dz * dz / Ry / Ry;
  if (arg > 16) return 0.0;
  return B.nb * exp(-arg);
}

// value of reference layer i at node (nj, nk); sums in the order
// x-cell (i, i-1), bunch, ck, cj, seed
__device__ void node_value(const BeamArgs &a, const qss_bunch *bun, int i, int nj, int nk,
                           double &rho, double &jx) {
  rho = 0.0;
  jx = 0.0;
  for (int side = 0; side < 2; ++side) {
    const int ci = side == 0 ? i : i - 1;
    if (ci < 0 || ci > a.nslices) continue;
    const double ax = 0.5;
    const double xi = (double)(a.nslices - ci) * a.hx;
    for (int b = 0; b < a.nbunch; ++b) {
      const qss_bunch &B = bun[b];
      // exact early out: arg = dx^2/Rx^2 + (non-negative terms) exceeds 16 as
      // soon as its first term does (rounded sums of non-negative terms do
      // not decrease), so a layer outside the bunch skips its cells
      if (!(b == 0 && a.nbins > 0)) {
        const double Rx = sqrt(2.0) * B.sigma_xi;
        const double dx = xi - B.xi0;
        if (dx * dx / Rx / Rx > 16) continue;
      }
      for (int ck = nk - 1; ck <= nk; ++ck) {
        if (ck < 0) continue;
        const int dz_ = nk - ck;
        for (int cj = nj - 1; cj <= nj; ++cj) {
          if (cj < 0) continue;
          const int dy_ = nj - cj;
          const double yc = a.hy * cj - a.ly / 2.0;
          const double zc = a.hz * ck - a.lz / 2.0;
          const double dn = bunch_density(a, bun, b, xi, yc, zc);
          if (!(dn > 0.0)) continue;
          for (int is = 0; is < 4; ++is) {
            const double yq = 1.0 / 3.0 * (1.0 + is % 2);
            const double zq = 1.0 / 3.0 * (1.0 + is / 2);
            double w = 0.25 * dn;
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

__global__ __launch_bounds__(NT) void k_beam_plane(BeamArgs a) {
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  const long nc = (long)a.ny * a.nz;
  if (n >= nc) return;
  const int w = blockIdx.y;   // wake of a batch
  const qss_bunch *bun = a.wb ? a.wb + (long)w * QSS_MAX_BUNCH : a.b;
  double *rho = a.rho + w * nc, *jx = a.jx + w * nc;
  const int nj = (int)(n % a.ny), nk = (int)(n / a.ny);
  const int layer = a.slice_dev ? a.nslices - 1 - *a.slice_dev : a.layer;
  double rc, jc;
  node_value(a, bun, layer, nj, nk, rc, jc);
  if (a.nsplit > 1) {
    double rp, jp;
    node_value(a, bun, layer + 1, nj, nk, rp, jp);
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
  hipLaunchKernelGGL(k_beam_plane, dim3((unsigned)((c->ncell + NT - 1) / NT), c->nw), dim3(NT),
                     0, c->stream, a);
  return hipGetLastError();
}
