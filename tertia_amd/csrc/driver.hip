// driver.hip — evolving driver: 3-D beam push through the stored wake and
// the trilinear deposit that the next sweep reads.
//
// k_driver_push restates Mesh::MoveBeamParticles (movebeam.cpp:14-660, the
// host path the reference links, run_control.h:22) per particle: trilinear
// gather of E, B from the 8 corners of the particle's 3-D cell
// (movebeam.cpp:209-272), kick by the interpolated E (its q2m*ts/2 factor is
// commented out in the reference, movebeam.cpp:283-302), Boris rotation
// with b = B q2m ts/(2 gamma), second kick, radiation reaction above
// SynMin (movebeam.cpp:448-501), move by ((Vx-1) ts/hx, Vy ts/hy, Vz ts/hz)
// with the cell clamping of movebeam.cpp:520-583, and deposit of
// w*polarity and w*polarity*Vx on the 8 nodes of the new cell
// (Mesh::MoveBeamInCell, movebeam.cpp:662-770).  Only particles in
// interior cells are visited (the loops of movebeam.cpp:63-69).  The file
// is compiled without FMA contraction so the push is the reference's
// operation order.
//
// Deposit: 64-bit fixed point, value * 2^e rounded to nearest, accumulated
// with integer atomics.  Integer addition is associative, so the 3-D
// deposit is bitwise independent of the order in which particles arrive
// (the reference's CUDA twin races on Jy/Jz, assign_currents.cu:92-104).
// Lanes of a wave whose particles share a cell add their contributions
// first, so one atomic per node and run of lanes reaches L2.  Each
// work-group also reduces the bounding box of the nodes it touched; the
// next push clears only that box instead of the whole 3-D array.

#include <hip/hip_ext.h>

#include <climits>

#include "qss_internal.h"
#include "split_bf16.h"

namespace {

constexpr int NT = 256;

struct DriverArgs {
  long np;
  int L, ny, nz;
  long ncell;
  double ts, hx, hy, hz;
  double bext[3];
  double syn_min2;     // SynMin^2
  double elaser;       // 1.2398e-4 / Wavelength
  double scale;        // 2^e of the fixed-point deposit
  double q2m[QSS_MAX_BUNCH], pol[QSS_MAX_BUNCH];
  const double *hist[6];   // Ex..Bz history (slice-major planes)
  int32_t *ci, *cj, *ck;
  const int32_t *bunch;
  double *x, *y, *z, *px, *py, *pz;
  const double *w;
  unsigned long long *acc_rho, *acc_jx;   // (L+1) x ncell nodes
  int32_t *box;        // min/max of deposited node i, j, k
  int32_t *status;     // bit 0: in-cell fraction outside [0,1]
};

// field of the stored wake at 3-D cell (i, j, k): layer i < L is history
// slice L-1-i, layer L is the unperturbed front; the stored transverse
// ghosts j = Ny, k = Nz are periodic copies (movepart_splitlayer.cpp:
// 1659-1682).  Callers pass 0 <= i <= L, 0 <= j <= Ny, 0 <= k <= Nz.
__device__ inline double wake(const double *h, const DriverArgs &a, int i, int j, int k) {
  if (i >= a.L) return 0.0;
  if (j == a.ny) j = 0;
  if (k == a.nz) k = 0;
  return h[(long)(a.L - 1 - i) * a.ncell + j + (long)a.ny * k];
}

__global__ __launch_bounds__(NT) void k_driver_push(DriverArgs a) {
#pragma clang fp contract(off)
  __shared__ int sbox[6];
  if (threadIdx.x < 6) sbox[threadIdx.x] = (threadIdx.x & 1) ? INT_MIN : INT_MAX;
  __syncthreads();
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  const int i = n < a.np ? a.ci[n] : -1;
  const int j = n < a.np ? a.cj[n] : -1;
  const int k = n < a.np ? a.ck[n] : -1;
  const bool interior = i >= 0 && i < a.L && j >= 0 && j < a.ny && k >= 0 && k < a.nz;
  double x = 0, y = 0, z = 0;
  if (interior) {
    x = a.x[n];
    y = a.y[n];
    z = a.z[n];
  }
  const bool bad = interior && (x < 0 || x > 1 || y < 0 || y > 1 || z < 0 || z > 1);
  if (bad) atomicOr(a.status, 1);
  // deposit data of this lane (dep: the particle was moved)
  int it = 0, jt = 0, kt = 0;
  double xt = 0, yt = 0, zt = 0, djx = 0, drho = 0;
  bool dep = false;
  if (interior && !bad) {
    const int b = a.bunch[n];
    const double weight = a.w[n];
    double px = a.px[n], py = a.py[n], pz = a.pz[n];
    const double pxn = px, pyn = py, pzn = pz;
    const double q2m = a.q2m[b];
    const double axc = 1. - x, axp = x, ayc = 1. - y, ayp = y, azc = 1. - z, azp = z;
    const double accc = axc * ayc * azc, acpc = axc * ayp * azc;
    const double accp = axc * ayc * azp, acpp = axc * ayp * azp;
    const double apcc = axp * ayc * azc, appc = axp * ayp * azc;
    const double apcp = axp * ayc * azp, appp = axp * ayp * azp;
    double g[6];
#pragma unroll
    for (int f = 0; f < 6; ++f) {
      const double *h = a.hist[f];
      g[f] = accc * wake(h, a, i, j, k) + acpc * wake(h, a, i, j + 1, k) +
             accp * wake(h, a, i, j, k + 1) + acpp * wake(h, a, i, j + 1, k + 1) +
             apcc * wake(h, a, i + 1, j, k) + appc * wake(h, a, i + 1, j + 1, k) +
             apcp * wake(h, a, i + 1, j, k + 1) + appp * wake(h, a, i + 1, j + 1, k + 1);
    }
    const double ex = g[0], ey = g[1], ez = g[2];
    double bx = g[3] + a.bext[0], by = g[4] + a.bext[1], bz = g[5] + a.bext[2];
    px += ex;
    py += ey;
    pz += ez;
    double gamma = sqrt(1. + px * px + py * py + pz * pz);
    const double gamma_r = 1. / gamma;
    bx = bx * gamma_r * q2m * a.ts / 2.;
    by = by * gamma_r * q2m * a.ts / 2.;
    bz = bz * gamma_r * q2m * a.ts / 2.;
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
    const double p2 = px * px + py * py + pz * pz;
    if (1. + p2 > a.syn_min2) {
      double dpx = px - pxn, dpy = py - pyn, dpz = pz - pzn;
      const double pdp = px * dpx + py * dpy + pz * dpz;
      dpx -= px * (pdp / p2);
      dpy -= py * (pdp / p2);
      dpz -= pz * (pdp / p2);
      const double p2absrev = 1. / sqrt(p2);
      const double nx = px * p2absrev, ny = py * p2absrev, nz = pz * p2absrev;
      const double dpabs = sqrt(dpx * dpx + dpy * dpy + dpz * dpz);
      const double Gamma2 = 1. + p2;
      const double omega_c = 3. * Gamma2 * dpabs / (4. * 3.1415927 * a.ts);
      const double Ephoton = a.elaser * omega_c;
      const double alfa = 1. / 137.;
      const double nph = 4. / 9. * alfa * dpabs;
      px -= (nph * Ephoton * nx / 511000.);
      py -= (nph * Ephoton * ny / 511000.);
      pz -= (nph * Ephoton * nz / 511000.);
    }
    gamma = 1. / sqrt(1. + px * px + py * py + pz * pz);
    const double Vx = px * gamma;
    const double pol = a.pol[b];
    djx = weight * pol * Vx;
    drho = weight * pol;
    const double dx = (px * gamma - 1.) * a.ts / a.hx;
    const double dy = py * gamma * a.ts / a.hy;
    const double dz = pz * gamma * a.ts / a.hz;
    // the wrap loops below run at most twice for |d| < 1e9 cells; a larger
    // or non-finite move (where the reference's loops would not end) is an
    // error and the particle stays
    const bool far = !(fabs(dx) < 1e9 && fabs(dy) < 1e9 && fabs(dz) < 1e9);
    if (far) atomicOr(a.status, 2);
    xt = far ? x : x + dx;
    yt = far ? y : y + dy;
    zt = far ? z : z + dz;
    int ij = (int)xt, jj = (int)yt, kj = (int)zt;
    if (xt < 0.) ij--;
    if (yt < 0.) jj--;
    if (zt < 0.) kj--;
    it = i + ij;
    jt = j + jj;
    kt = k + kj;
    it = it < 0 ? -1 : (it > a.L - 1 ? a.L : it);
    jt = jt < 0 ? -1 : (jt > a.ny - 1 ? a.ny : jt);
    kt = kt < 0 ? -1 : (kt > a.nz - 1 ? a.nz : kt);
    xt -= ij;
    yt -= jj;
    zt -= kj;
    while (xt < 0.) xt += 1.;
    while (xt > 1.) xt -= 1.;
    while (yt < 0.) yt += 1.;
    while (yt > 1.) yt -= 1.;
    while (zt < 0.) zt += 1.;
    while (zt > 1.) zt -= 1.;
    a.px[n] = px;
    a.py[n] = py;
    a.pz[n] = pz;
    a.ci[n] = it;
    a.cj[n] = jt;
    a.ck[n] = kt;
    a.x[n] = xt;
    a.y[n] = yt;
    a.z[n] = zt;
    dep = !far;
  }
  // Deposit, with every lane of the wave active.  Lanes whose particles
  // ended in the same cell form runs (particles are stored cell by cell and
  // mostly stay); all 8 node slots of a lane share its cell, so one head
  // mask serves every slot.  Each slot's fixed-point values are summed over
  // the run by a segmented suffix scan and the run's first lane issues the
  // atomics: integer sums, so the result does not depend on the grouping.
  const int lane = threadIdx.x & 63;
  const long key = dep ? ((long)(it + 1) * (a.ny + 2) + (jt + 1)) * (long)(a.nz + 2) + (kt + 1)
                       : -1;
  const long prev = __shfl_up(key, 1);
  const bool head = lane == 0 || prev != key;
  const unsigned long long hm = __ballot(head);
  // cont[t]: lanes l whose lanes l+1 .. l+2^t are in l's run
  const unsigned long long nonhead = ~hm;
  const double ax[2] = {1. - xt, xt}, ay[2] = {1. - yt, yt}, az[2] = {1. - zt, zt};
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const int da = q >> 2, dc = (q >> 1) & 1, db = q & 1;
    const int ni = it + da, nj = jt + db, nk = kt + dc;
    const bool valid =
        dep && ni >= 0 && ni <= a.L && nj >= 0 && nj < a.ny && nk >= 0 && nk < a.nz;
    // djx*axc*ayc*azc, ... (movebeam.cpp:697-764), then fixed point
    long long vj = valid ? (long long)rint(djx * ax[da] * ay[db] * az[dc] * a.scale) : 0;
    long long vr = valid ? (long long)rint(drho * ax[da] * ay[db] * az[dc] * a.scale) : 0;
    unsigned long long y = nonhead >> 1;   // bit l: lane l+1 continues l's run
    for (int d = 1; d < 64 && y; d <<= 1) {
      const long long oj = __shfl_down(vj, d), orr = __shfl_down(vr, d);
      if ((y >> lane) & 1ull) {   // lanes l+1 .. l+d are in l's run
        vj += oj;
        vr += orr;
      }
      y &= y >> d;                // now: lanes l+1 .. l+2d
    }
    if (head && valid) {
      const long m = (long)ni * a.ncell + nj + (long)a.ny * nk;
      atomicAdd(a.acc_jx + m, (unsigned long long)vj);
      atomicAdd(a.acc_rho + m, (unsigned long long)vr);
    }
  }
  // bounding box of the touched nodes: wave min/max, then LDS, then global
  int lo[3], hi[3];
  lo[0] = max(it, 0); hi[0] = min(it + 1, a.L);
  lo[1] = max(jt, 0); hi[1] = min(jt + 1, a.ny - 1);
  lo[2] = max(kt, 0); hi[2] = min(kt + 1, a.nz - 1);
  const bool any = dep && lo[0] <= hi[0] && lo[1] <= hi[1] && lo[2] <= hi[2];
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    if (!any) { lo[d] = INT_MAX; hi[d] = INT_MIN; }
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
      lo[d] = min(lo[d], __shfl_xor(lo[d], o));
      hi[d] = max(hi[d], __shfl_xor(hi[d], o));
    }
  }
  if (lane == 0 && lo[0] != INT_MAX) {
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      atomicMin(&sbox[2 * d], lo[d]);
      atomicMax(&sbox[2 * d + 1], hi[d]);
    }
  }
  __syncthreads();
  if (threadIdx.x < 6 && sbox[0] != INT_MAX) {
    if (threadIdx.x & 1) atomicMax(a.box + threadIdx.x, sbox[threadIdx.x]);
    else atomicMin(a.box + threadIdx.x, sbox[threadIdx.x]);
  }
}

// zero the nodes [i0,i1] x [j0,j1] x [k0,k1] of both deposit arrays
__global__ __launch_bounds__(NT) void k_driver_clear(unsigned long long *r,
                                                     unsigned long long *jx, int ny,
                                                     long ncell, int i0, int j0, int k0,
                                                     int nj, int nk, long total) {
  const long t = (long)blockIdx.x * NT + threadIdx.x;
  if (t >= total) return;
  const int jj = (int)(t % nj);
  const long rest = t / nj;
  const int kk = (int)(rest % nk);
  const int ii = (int)(rest / nk);
  const long m = (long)(i0 + ii) * ncell + (j0 + jj) + (long)ny * (k0 + kk);
  r[m] = 0ull;
  jx[m] = 0ull;
}

struct PlaneArgs {
  const unsigned long long *acc_rho, *acc_jx;
  double inv;          // 2^-e
  long ncell;
  int L, layer, nsplit, isplit;
  const int *slice_dev;
  double *rho, *jx;
  uint16_t *xs0, *xs1;
  int npc;
};

// C-layer beam planes of (layer, split) from the 3-D deposit: nodes of
// layer i, interpolated with layer i+1 for split s of S
// (movepart_splitlayer.cpp:330-348)
__global__ __launch_bounds__(NT) void k_driver_plane(PlaneArgs a) {
  const long n = (long)blockIdx.x * NT + threadIdx.x;
  if (n >= a.ncell) return;
  const int layer = a.slice_dev ? a.L - 1 - *a.slice_dev : a.layer;
  const long m = (long)layer * a.ncell + n;
  double rc = (double)(long long)a.acc_rho[m] * a.inv;
  double jc = (double)(long long)a.acc_jx[m] * a.inv;
  if (a.nsplit > 1) {
    const double rp = (double)(long long)a.acc_rho[m + a.ncell] * a.inv;
    const double jp = (double)(long long)a.acc_jx[m + a.ncell] * a.inv;
    const int S = a.nsplit, s = a.isplit;
    rc = (rc * (S - s) + rp * s) / S;
    jc = (jc * (S - s) + jp * s) / S;
  }
  a.rho[n] = rc;
  a.jx[n] = jc;
  if (a.npc) {
    qss_split::store(a.xs0, a.ncell, n, rc, a.npc);
    qss_split::store(a.xs1, a.ncell, n, jc, a.npc);
  }
}

}  // namespace

hipError_t launch_driver_clear(qss_ctx *c) {
  DriverState &d = c->drv;
  if (d.box_lo[0] > d.box_hi[0]) return hipSuccess;   // nothing deposited
  const int nj = d.box_hi[1] - d.box_lo[1] + 1, nk = d.box_hi[2] - d.box_lo[2] + 1;
  const long total = (long)(d.box_hi[0] - d.box_lo[0] + 1) * nj * nk;
  hipLaunchKernelGGL(k_driver_clear, dim3((unsigned)((total + NT - 1) / NT)), dim3(NT), 0,
                     c->stream, d.acc_rho, d.acc_jx, c->ny, c->ncell, d.box_lo[0],
                     d.box_lo[1], d.box_lo[2], nj, nk, total);
  return hipGetLastError();
}

hipError_t launch_driver_push(qss_ctx *c, hipEvent_t ev0, hipEvent_t ev1) {
  DriverState &d = c->drv;
  DriverArgs a;
  a.np = d.np;
  a.L = c->p.nslices;
  a.ny = c->ny;
  a.nz = c->nz;
  a.ncell = c->ncell;
  a.ts = d.par.ts;
  a.hx = c->p.hx;
  a.hy = c->hy;
  a.hz = c->hz;
  for (int q = 0; q < 3; ++q) a.bext[q] = d.par.b_ext[q];
  a.syn_min2 = d.par.syn_min * d.par.syn_min;
  a.elaser = 1.2398e-4 / d.par.wavelength;
  a.scale = d.scale;
  for (int b = 0; b < QSS_MAX_BUNCH; ++b) {
    a.q2m[b] = d.par.q2m[b];
    a.pol[b] = c->p.bunch[b].polarity;
  }
  for (int f = 0; f < 6; ++f) a.hist[f] = c->hist[QSS_EX + f];
  a.ci = d.ci; a.cj = d.cj; a.ck = d.ck; a.bunch = d.bunch;
  a.x = d.x; a.y = d.y; a.z = d.z;
  a.px = d.px; a.py = d.py; a.pz = d.pz;
  a.w = d.w;
  a.acc_rho = d.acc_rho;
  a.acc_jx = d.acc_jx;
  a.box = d.box;
  a.status = d.status;
  const unsigned nb = (unsigned)((d.np + NT - 1) / NT);
  if (nb == 0) return hipSuccess;
  hipExtLaunchKernelGGL(k_driver_push, dim3(nb), dim3(NT), 0, c->stream, ev0, ev1, 0, a);
  return hipGetLastError();
}

hipError_t launch_driver_plane(qss_ctx *c, int slice, int isplit) {
  DriverState &d = c->drv;
  PlaneArgs a;
  a.acc_rho = d.acc_rho;
  a.acc_jx = d.acc_jx;
  a.inv = 1.0 / d.scale;
  a.ncell = c->ncell;
  a.L = c->p.nslices;
  a.layer = c->p.nslices - 1 - slice;
  a.nsplit = c->p.nx_split;
  a.isplit = isplit;
  a.slice_dev = c->capturing ? c->d_slice : nullptr;
  a.rho = c->C[QSS_RHOBEAM];
  a.jx = c->C[QSS_JXBEAM];
  a.xs0 = c->xs[0];
  a.xs1 = c->xs[1];
  a.npc = c->npc;
  if (c->npc) {
    c->xs_src[0] = a.rho;
    c->xs_src[1] = a.jx;
  }
  hipLaunchKernelGGL(k_driver_plane, dim3((unsigned)((c->ncell + NT - 1) / NT)), dim3(NT), 0,
                     c->stream, a);
  return hipGetLastError();
}
