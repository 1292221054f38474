// adjoint.hip — kernels of the reverse (adjoint) xi sweep.
//
// The reference has no derivative of its sweep; these kernels are the
// transposes of the forward kernels, driven by qss_sweep_vjp in qss_api.cpp:
//   k_push_tile_adj  reverse of the tiled push + CIC deposit of one pass
//                    (movepart_splitlayer.cpp:932-1512 with the CIC shape of
//                    movepart_layer_CIC.cpp:166-203,449-496).  One work-group
//                    per 16x16-cell tile, like the forward: it stages the
//                    tile's push field (17x17 nodes) and the deposit adjoint
//                    (21x21 nodes) in LDS, replays every particle's push and
//                    runs its reverse (push_adj.h), writes the particle-state
//                    adjoint by particle id, and sums the field adjoint of
//                    the gather (the transposed CIC gather) per node in a
//                    fixed order: bitonic sort of the chunk's start cells,
//                    then each node adds the events of its four cells.
//   k_fbar_combine   sums the overlapping 17x17 windows per node (fixed
//                    order) into the six field-adjoint planes.
//   k_beam_adj       derivative of one layer's beam planes (beam.hip) with
//                    respect to the piecewise-constant profile value of the
//                    layer's two x-cells, contracted with the beam-plane
//                    adjoints; one work-group, fixed-order reduction.

#include "qss_internal.h"
#include "push_adj.h"
#include "tile_sort.h"

namespace {

constexpr int NT = 256;
constexpr int T = kTile;
constexpr int H = 2;
constexpr int W = T + 2 * H + 1;   // 21: deposit window nodes per edge
constexpr int NWIN = W * W;
constexpr int FW = T + 1;          // 17: gather window nodes per edge
constexpr int NF = FW * FW;
constexpr int kInvalid = 0x7fffffff;
enum { ST_PARTICLE = 1, ST_FAR = 8 };

__device__ inline int wrapi(int a, int n) {
  a %= n;
  return a < 0 ? a + n : a;
}

__global__ __launch_bounds__(NT) void k_push_tile_adj(AdjTileArgs a) {
  __shared__ double Fw[NF * 6];
  __shared__ double Jw[NWIN * 3];
  __shared__ int keys_sh[NT];
  __shared__ int sorted[NT];
  __shared__ double ev_f[NT * 6];
  __shared__ double ev_w[NT * 4];
  __shared__ short cstart[T * T], cend[T * T];

  const int t = threadIdx.x;
  // a tile's particles shared by nsplit work-groups, as in the forward push
  const int tile = blockIdx.x / a.nsplit, part = blockIdx.x - tile * a.nsplit;
  const int tj = tile % a.ntj, tk = tile / a.ntj;
  const int j0 = tj * T, k0 = tk * T;
  const int ny = a.ny, nz = a.nz;
  const long plane = (long)ny * nz;

  for (int i = t; i < NF; i += NT) {
    const int lj = i % FW, lk = i / FW;
    const long g = wrapi(j0 + lj, ny) + (long)ny * wrapi(k0 + lk, nz);
#pragma unroll
    for (int q = 0; q < 6; ++q) Fw[i * 6 + q] = a.F[q * plane + g];
  }
  for (int i = t; i < NWIN; i += NT) {
    const int lj = i % W, lk = i / W;
    const long g = wrapi(j0 - H + lj, ny) + (long)ny * wrapi(k0 - H + lk, nz);
#pragma unroll
    for (int q = 0; q < 3; ++q) Jw[i * 3 + q] = a.Jb[q * plane + g];
  }
  // this work-group's window: chunk sums written by the first chunk, added
  // by later ones (each thread owns nodes t and t + NT)
  double *wout = a.fwin + (long)blockIdx.x * NF * 6;
  bool first = true;
  __syncthreads();

  qss_adj::Consts cs;
  cs.hx = a.hx; cs.hy = a.hy; cs.hz = a.hz; cs.weight = a.weight; cs.ny = ny; cs.nz = nz;
  cs.hxy = a.hx / a.hy;   // the forward push's hx / hy, hx / hz
  cs.hxz = a.hx / a.hz;
  for (int q = 0; q < 3; ++q) cs.bext[q] = a.bext[q];
  const long np = a.np;
  const int base = a.tile_start[tile];
  const int npt = a.tile_start[tile + 1] - base;
  for (int c0 = part * NT; c0 < npt; c0 += a.nsplit * NT) {
    const int slot = base + c0 + t;
    int key = kInvalid;
    if (c0 + t < npt) {
      const int32_t c = a.cell[slot];
      const int j = c % ny, k = c / ny;
      int lj = j - j0, lk = k - k0;
      if (lj < 0 || lj >= T || lk < 0 || lk >= T) {
        atomicOr(&a.counters[1], ST_FAR);
        lj = min(max(lj, 0), T - 1);
        lk = min(max(lk, 0), T - 1);
      }
      const double y0 = a.y[slot], z0 = a.z[slot];
      const double P0[3] = {a.px[slot], a.py[slot], a.pz[slot]};
      const long id = a.pid[slot];
      // adjoint of the new position (0, 1) and momentum (2..4), full step only
      auto NB = [&](int i) { return a.pin[i * np + id]; };
      auto Fn = [&](int q, int dj, int dk) { return Fw[((lj + dj) + FW * (lk + dk)) * 6 + q]; };
      bool far = false;
      auto JBn = [&](int q, int nj, int nk) -> double {
        if (q > 2) return 0.0;   // the deposited density is not transformed
        int dj = nj - j0, dk = nk - k0;
        if (dj + H < 0) dj += ny;
        else if (dj + H > W - 1) dj -= ny;
        if (dk + H < 0) dk += nz;
        else if (dk + H > W - 1) dk -= nz;
        if (dj + H < 0 || dj + H > W - 1 || dk + H < 0 || dk + H > W - 1) {
          far = true;
          return 0.0;
        }
        return Jw[((dk + H) * W + dj + H) * 3 + q];
      };
      qss_adj::AdjOut o;
      qss_adj::push_adjoint_cb(cs, j, k, y0, z0, P0, Fn, JBn, a.full != 0, NB, o);
      if (far) atomicOr(&a.counters[1], ST_FAR);
      if (a.full) {
        a.pout[id] = o.y0;
        a.pout[np + id] = o.z0;
        a.pout[2 * np + id] = o.P0[0];
        a.pout[3 * np + id] = o.P0[1];
        a.pout[4 * np + id] = o.P0[2];
      } else {
        a.pout[id] += o.y0;
        a.pout[np + id] += o.z0;
        a.pout[2 * np + id] += o.P0[0];
        a.pout[3 * np + id] += o.P0[1];
        a.pout[4 * np + id] += o.P0[2];
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) ev_f[t * 6 + q] = o.f[q];
#pragma unroll
      for (int cn = 0; cn < 4; ++cn) ev_w[t * 4 + cn] = o.w[cn];
      key = ((lk * T + lj) << 8) | t;
    }
    // transposed gather: every node sums the events of its four cells
    for (int i = t; i < T * T; i += NT) { cstart[i] = 0; cend[i] = 0; }
    const int sk = qss_tile::sort256(key, keys_sh);
    sorted[t] = sk;
    // the event at sorted position t, moved there in place: gathered before the
    // barrier (the sort's barriers made every event visible), stored after it
    double gf[6], gw[4];
    if (sk != kInvalid) {
      const int e = sk & 255;
#pragma unroll
      for (int q = 0; q < 6; ++q) gf[q] = ev_f[e * 6 + q];
#pragma unroll
      for (int cn = 0; cn < 4; ++cn) gw[cn] = ev_w[e * 4 + cn];
    }
    __syncthreads();
    if (sk != kInvalid) {
#pragma unroll
      for (int q = 0; q < 6; ++q) ev_f[t * 6 + q] = gf[q];
#pragma unroll
      for (int cn = 0; cn < 4; ++cn) ev_w[t * 4 + cn] = gw[cn];
      const int cl = sk >> 8;
      const int prev = t > 0 ? sorted[t - 1] : kInvalid;
      const int next = t + 1 < NT ? sorted[t + 1] : kInvalid;
      if (t == 0 || (prev >> 8) != cl) cstart[cl] = (short)t;
      if (next == kInvalid || (next >> 8) != cl) cend[cl] = (short)(t + 1);
    }
    __syncthreads();
#pragma unroll
    for (int oi = 0; oi < 2; ++oi) {
      const int n = t + oi * NT;
      if (n >= NF) continue;
      double accn[6] = {0, 0, 0, 0, 0, 0};
      const int nl_j = n % FW, nl_k = n / FW;
#pragma unroll
      for (int cn = 0; cn < 4; ++cn) {
        const int cj = nl_j - (cn & 1), ck = nl_k - (cn >> 1);
        if (cj < 0 || ck < 0 || cj >= T || ck >= T) continue;
        const int cl = ck * T + cj;
        const int e1 = cend[cl];
        for (int i = cstart[cl]; i < e1; ++i) {
          const int e = i;   // events are stored in sorted order
          const double wv = ev_w[e * 4 + cn];
#pragma unroll
          for (int q = 0; q < 6; ++q) accn[q] += wv * ev_f[e * 6 + q];
        }
      }
#pragma unroll
      for (int q = 0; q < 6; ++q) wout[n * 6 + q] = first ? accn[q] : wout[n * 6 + q] + accn[q];
    }
    first = false;
    __syncthreads();
  }
  if (first)   // no particles for this work-group
    for (int n = t; n < NF; n += NT)
#pragma unroll
      for (int q = 0; q < 6; ++q) wout[n * 6 + q] = 0.0;
}

// out[q][node] = sum of the 17x17 windows holding the node, fixed order
// pshare (optional): the P-layer share of the field adjoint, 1/2 of each
// gathered field's adjoint, assigned (pmode 1) or accumulated (pmode 2)
__global__ __launch_bounds__(NT) void k_fbar_combine(const double *fwin, int ntj, int ntk, int ny,
                                                     int nz, int nsplit, double *out,
                                                     double *pshare, int pmode) {
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  const long plane = (long)ny * nz;
  if (n >= plane) return;
  const int nj = (int)(n % ny), nk = (int)(n / ny);
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int b = 0; b < 2; ++b) {
    if (b == 1 && nk % T != 0) continue;
    const int utk = nk / T - b, lk = nk - utk * T;
    const int wtk = wrapi(utk, ntk);
    for (int c = 0; c < 2; ++c) {
      if (c == 1 && nj % T != 0) continue;
      const int utj = nj / T - c, lj = nj - utj * T;
      const int wtj = wrapi(utj, ntj);
      for (int p = 0; p < nsplit; ++p) {
        const double *v = fwin + (((long)(wtk * ntj + wtj) * nsplit + p) * NF + lk * FW + lj) * 6;
#pragma unroll
        for (int q = 0; q < 6; ++q) s[q] += v[q];
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) out[q * plane + n] = s[q];
  if (pmode == 1) {
#pragma unroll
    for (int q = 0; q < 6; ++q) pshare[q * plane + n] = 0.5 * s[q];
  } else if (pmode == 2) {
#pragma unroll
    for (int q = 0; q < 6; ++q) pshare[q * plane + n] += 0.5 * s[q];
  }
}

// ---- beam profile adjoint ----------------------------------------------------
struct BeamAdjArgs {
  int ny, nz, nslices;
  double hx, hy, hz, ly, lz, bin_xi0, bin_xi1;
  int nbins;
  uint64_t seed;
  qss_bunch b0;
  const double *rbar, *jbar;   // adjoints of the layer's beam rho, Jx planes
  double *q;                   // q[2*slice + side]
  int slice;
  const int *slice_dev;        // graph replay: the slice is read from device memory
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

// kBeamAdjBlocks work-groups, each reducing its share of the cells in a fixed
// order; partial sums land in q[(2*slice + side) * kBeamAdjBlocks + block]
// and the host adds them in block order
constexpr int NB = 256;
__global__ __launch_bounds__(NB) void k_beam_adj(BeamAdjArgs a) {
  __shared__ double red[NB];
  const int t = threadIdx.x;
  const int slice = a.slice_dev ? *a.slice_dev : a.slice;
  const int i = a.nslices - 1 - slice;   // reference layer index
  const qss_bunch &B = a.b0;
  const double Ry = sqrt(2.0) * B.sigma_r;
  const long ncell = (long)a.ny * a.nz;
  const long stride = (long)NB * kBeamAdjBlocks;
  for (int side = 0; side < 2; ++side) {
    const int ci = side == 0 ? i : i - 1;
    double part = 0.0;
    if (ci >= 0 && ci <= a.nslices) {
      for (long c = (long)blockIdx.x * NB + t; c < ncell; c += stride) {
        const int cj = (int)(c % a.ny), ck = (int)(c / a.ny);
        const double dy = a.hy * cj - a.ly / 2.0 - B.y0;
        const double dz = a.hz * ck - a.lz / 2.0 - B.z0;
        const double arg = dy * dy / Ry / Ry + dz * dz / Ry / Ry;
        if (arg > 16) continue;
        const double e = exp(-arg);
        for (int is = 0; is < 4; ++is) {
          const double yq = 1.0 / 3.0 * (1.0 + is % 2);
          const double zq = 1.0 / 3.0 * (1.0 + is / 2);
          double w = 0.25 * e;
          if (B.noise != 0.0) {
            const uint64_t cell =
                ((uint64_t)(ci + 1) * (uint64_t)a.nz + ck) * (uint64_t)a.ny + cj;
            w *= 1.0 + B.noise * unif_pm1(a.seed ^ 0xB3A1ull, cell * 4 + is, 7);
          }
          for (int cn = 0; cn < 4; ++cn) {
            const int nj = cj + (cn & 1), nk = ck + (cn >> 1);
            if (nj >= a.ny || nk >= a.nz) continue;
            const double ay = (cn & 1) ? yq : 1.0 - yq;
            const double az = (cn >> 1) ? zq : 1.0 - zq;
            const long n = nj + (long)a.ny * nk;
            part += w * B.polarity * 0.5 * ay * az * (a.rbar[n] + B.vx * a.jbar[n]);
          }
        }
      }
    }
    red[t] = part;
    __syncthreads();
    for (int o = NB / 2; o > 0; o >>= 1) {
      if (t < o) red[t] += red[t + o];
      __syncthreads();
    }
    if (t == 0) a.q[(2L * slice + side) * kBeamAdjBlocks + blockIdx.x] = red[0];
    __syncthreads();
  }
}

inline dim3 grid1(long n) { return dim3((unsigned)((n + NT - 1) / NT)); }

}  // namespace

hipError_t launch_push_tile_adj(qss_ctx *c, const AdjTileArgs &a) {
  hipLaunchKernelGGL(k_push_tile_adj, dim3(c->ntj * c->ntk * a.nsplit), dim3(NT), 0, c->stream,
                     a);
  return hipGetLastError();
}

hipError_t launch_fbar_combine(qss_ctx *c, const double *fwin, double *out, double *pshare,
                               int pmode) {
  hipLaunchKernelGGL(k_fbar_combine, grid1(c->ncell), dim3(NT), 0, c->stream, fwin, c->ntj,
                     c->ntk, c->ny, c->nz, c->nsplit, out, pshare, pmode);
  return hipGetLastError();
}

hipError_t launch_beam_adj(qss_ctx *c, int slice, const double *rbar, const double *jbar,
                           double *q) {
  BeamAdjArgs a;
  a.ny = c->ny; a.nz = c->nz; a.nslices = c->p.nslices;
  a.hx = c->p.hx; a.hy = c->hy; a.hz = c->hz; a.ly = c->p.ly; a.lz = c->p.lz;
  a.bin_xi0 = c->p.bin_xi0; a.bin_xi1 = c->p.bin_xi1; a.nbins = c->p.nbins;
  a.seed = c->p.seed;
  a.b0 = c->p.bunch[0];
  a.rbar = rbar; a.jbar = jbar; a.q = q; a.slice = slice;
  a.slice_dev = c->capturing ? c->d_slice : nullptr;
  hipLaunchKernelGGL(k_beam_adj, dim3(kBeamAdjBlocks), dim3(NB), 0, c->stream, a);
  return hipGetLastError();
}
