// qss_api.cpp — C ABI and host-side sweep driver of libqss.
//
// The host code mirrors the reference's orchestration:
//   qss_sweep_range  <-> Mesh::MoveAllSplitLayers  (movepart_splitlayer.cpp:25-558)
//   qss_layer_step   <-> Mesh::MoveSplitLayer      (movepart_splitlayer.cpp:561-778)
//   qss_guess        <-> GuessFieldsHydroLinLayerSplit (movefieldshydrolinlayersplit.cpp:222)
//   qss_push_deposit <-> ClearCurrentsSplit + MoveParticlesLayerSplit +
//                        ExchangeCurrentsSplit (mesh.cpp:692, movepart_splitlayer.cpp:781,1685)
//   qss_iterate      <-> IterateFieldsHydroLinLayerSplit + ExchangeFieldsSplit (:1173)
// but all data stays in HBM: the reference's Cell copies P<-layer, C->P and
// layer<-P (:170-187, :423-428, :467-473) become pointer swaps, and the
// spectral P state is kept instead of being re-transformed.

#include "qss_internal.h"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>
#include <string>

namespace {

thread_local std::string g_err;

constexpr int kNClass = QSS_NKCLASS;

// ---- timing helpers --------------------------------------------------------
// Timed(c, cls): event markers around the launches of the scope.
// Timed(c, cls, true): an event pair for ONE launch that takes them itself
// (hipExtLaunchKernel stamps them with the dispatch's start and end), so the
// measured time is the kernel's own duration.
struct Timed {
  qss_ctx *c;
  int cls;
  bool ext;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  Timed(qss_ctx *c_, int cls_, bool ext_ = false) : c(c_), cls(cls_), ext(ext_) {
    if (!c->timing) return;
    auto &pool = c->ev_pool[cls];
    size_t &used = c->ev_used[cls];
    if (used == pool.size()) {
      hipEvent_t a, b;
      if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) {
        c->timing = false;
        return;
      }
      pool.emplace_back(a, b);
    }
    e0 = pool[used].first;
    e1 = pool[used].second;
    ++used;
    if (!ext) (void)hipEventRecord(e0, c->stream);
  }
  ~Timed() {
    if (e1 && !ext) (void)hipEventRecord(e1, c->stream);
  }
};

template <class T>
int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) {
    qss_set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return QSS_E_NOMEM;
  }
  e = hipMemset(*p, 0, n * sizeof(T));
  if (e != hipSuccess) {
    qss_set_error(std::string("hipMemset: ") + hipGetErrorString(e));
    return QSS_E_HIP;
  }
  return QSS_OK;
}

template <class T>
void dfree(T *&p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// dense half-integer tables, exact integer argument reduction:
// angle = pi (2j+1)(2k+1)/(4n) reduced mod 2 pi via (2j+1)(2k+1) mod 8n
void make_tables(int n, std::vector<double> &tc, std::vector<double> &ts) {
  tc.resize((size_t)n * n);
  ts.resize((size_t)n * n);
  for (int j = 0; j < n; ++j)
    for (int k = 0; k < n; ++k) {
      long m = ((long)(2 * j + 1) * (2 * k + 1)) % (8L * n);
      double a = M_PI * (double)m / (4.0 * n);
      tc[(size_t)j * n + k] = 2.0 * std::cos(a);
      ts[(size_t)j * n + k] = 2.0 * std::sin(a);
    }
}

int validate(const qss_params *p) {
  if (!p) { qss_set_error("null params"); return QSS_E_ARG; }
  if (p->ny < 2 || p->nz < 2 || p->ny > 16384 || p->nz > 16384) {
    qss_set_error("ny, nz must be in [2, 16384]");
    return QSS_E_ARG;
  }
  if (!(p->ly > 0) || !(p->lz > 0) || !(p->hx > 0) || !(p->dens > 0)) {
    qss_set_error("ly, lz, hx, dens must be positive");
    return QSS_E_ARG;
  }
  if (p->nslices < 1 || p->nx_split < 1 || p->n_iter < 0 || p->ppc_y < 1 ||
      p->ppc_z < 1 || p->nbunch < 0 || p->nbunch > QSS_MAX_BUNCH) {
    qss_set_error("bad nslices / nx_split / n_iter / ppc / nbunch");
    return QSS_E_ARG;
  }
  if (p->shape != QSS_SHAPE_NGP_REF && p->shape != QSS_SHAPE_CIC) {
    qss_set_error("unknown shape");
    return QSS_E_ARG;
  }
  if (p->precision < QSS_PREC_FP64 || p->precision > QSS_PREC_MIXED_BF16X6) {
    qss_set_error("unknown precision mode");
    return QSS_E_ARG;
  }
  if (p->precision != QSS_PREC_FP64 &&
      (p->ny % kTile != 0 || p->nz % kTile != 0 || p->fix_subcycle)) {
    qss_set_error("mixed precision needs ny, nz multiples of 16 and fix_subcycle = 0");
    return QSS_E_ARG;
  }
  if (p->nbins < 0 || (p->nbins > 0 && (!p->bins || !(p->bin_xi1 > p->bin_xi0)))) {
    qss_set_error("bad piecewise profile");
    return QSS_E_ARG;
  }
  if (p->nwake < 0 || p->nwake > 1024) {
    qss_set_error("nwake must be in [0, 1024]");
    return QSS_E_ARG;
  }
  for (int q = 0; q < 3; ++q)
    if (!std::isfinite(p->b_ext[q])) {
      qss_set_error("b_ext must be finite");
      return QSS_E_ARG;
    }
  if (p->nwake > 1 && (p->ny % kTile != 0 || p->nz % kTile != 0 || p->fix_subcycle ||
                       (p->precision != QSS_PREC_FP64 && p->precision != QSS_PREC_MIXED32))) {
    qss_set_error("a batch of wakes needs ny, nz multiples of 16, fix_subcycle = 0 and "
                  "precision FP64 or MIXED32");
    return QSS_E_ARG;
  }
  long long np = (long long)p->ny * p->nz * p->ppc_y * p->ppc_z * (p->nwake > 1 ? p->nwake : 1);
  if (np > (1LL << 30)) {
    qss_set_error("too many particles per layer");
    return QSS_E_ARG;
  }
  return QSS_OK;
}

// ---- transforms ------------------------------------------------------------
// the recorded split forms no longer describe their planes
void forget_splits(qss_ctx *c) {
  for (auto &p : c->xs_src) p = nullptr;
}

inline int gemm_pick(const qss_ctx *c, int pass, int n) {
  return c->gemm_force >= 0 ? c->gemm_force : c->gemm_var[pass][n];
}

// The factorized form of transform_batch (fact.hip): pre-twiddle, GEMM with
// G_y, butterflies + z pre-twiddle, GEMM with G_z, butterflies + epilogue.
int transform_fact(qss_ctx *c, int n, const int *kz, const int *ky, const double *const *in,
                   double *const *out, double alpha, const int *resid, const int *push,
                   const double *const *pplane, double beta, bool one_wake) {
  const long nc = c->ncell;
  if (c->ffused_ok) {
    // twiddles inside the GEMMs: y pass in -> tmp, z pass tmp -> out with the
    // epilogue
    FusedArgs f{};
    f.ny = c->ny; f.nz = c->nz; f.nbatch = n;
    f.nw = one_wake ? 1 : c->nw;
    f.w_in = nc; f.w_out = nc; f.w_p = nc; f.w_f = nc * kPushComp; f.w_r = kResidPerWake;
    f.alpha = alpha; f.beta = beta;
    f.F = c->F; f.Ff = c->Ff; f.rslots = c->d_resid;
    for (int b = 0; b < n; ++b) {
      f.in[b] = in[b]; f.out[b] = c->tmp[b];
      f.P[b] = pplane ? pplane[b] : nullptr;
      f.kz[b] = kz[b]; f.ky[b] = ky[b];
      f.resid_c[b] = resid ? resid[b] : -1;
      f.push[b] = push ? push[b] : -1;
    }
    f.G = c->d_Gf[0]; f.Gd = c->d_Gfd[0]; f.tw = c->d_twf[0];
    {
      Timed t(c, KC_GEMM, true);
      HIPCHECK(launch_fact_fused(f, 0, c->ffv[0][n], c->stream, t.e0, t.e1));
    }
    if (c->timing) c->gemm_flops += 2.0 * (2.0 * c->nz) * (c->ny / 2) * (c->ny / 2) * n * f.nw;
    for (int b = 0; b < n; ++b) { f.in[b] = c->tmp[b]; f.out[b] = out[b]; }
    f.G = c->d_Gf[1]; f.Gd = c->d_Gfd[1]; f.tw = c->d_twf[1];
    {
      Timed t(c, KC_GEMM, true);
      HIPCHECK(launch_fact_fused(f, 1, c->ffv[1][n], c->stream, t.e0, t.e1));
    }
    if (c->timing) c->gemm_flops += 2.0 * (c->nz / 2) * (2.0 * c->ny) * (c->nz / 2) * n * f.nw;
    return QSS_OK;
  }
  FactArgs f{};
  f.ny = c->ny; f.nz = c->nz; f.nbatch = n;
  f.nw = one_wake ? 1 : c->nw;
  f.w_in = nc; f.w_out = nc; f.w_p = nc; f.w_f = nc * kPushComp; f.w_r = kResidPerWake;
  f.twy = c->d_tw[0]; f.twz = c->d_tw[1];
  f.alpha = alpha; f.beta = beta;
  f.F = c->F; f.Ff = c->Ff; f.rslots = c->d_resid;
  for (int b = 0; b < n; ++b) {
    f.in[b] = in[b]; f.out[b] = out[b];
    f.P[b] = pplane ? pplane[b] : nullptr;
    f.kz[b] = kz[b]; f.ky[b] = ky[b];
    f.resid_c[b] = resid ? resid[b] : -1;
    f.push[b] = push ? push[b] : -1;
    f.t0[b] = c->tmp[b]; f.t1[b] = c->tmp2[b];
  }
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_fact(f, 0, c->stream));
  }
  GemmArgs g{};
  g.nbatch = n;
  g.nw = f.nw;
  g.alpha = 1.0;
  g.wA = nc; g.wB = 0; g.wC = nc;
  g.M = 2 * c->nz; g.N = c->ny / 2; g.K = c->ny / 2;
  for (int b = 0; b < n; ++b) {
    g.b[b].A = c->tmp[b]; g.b[b].B = c->d_G[0]; g.b[b].C = c->tmp2[b];
    g.b[b].resid = -1; g.b[b].push = -1;
  }
  {
    Timed t(c, KC_GEMM, true);
    HIPCHECK(launch_gemm(g, c->stream, QSS_PREC_FP64,
                           c->gemm_force >= 0 ? c->gemm_force : c->fact_var[0][n], t.e0, t.e1));
  }
  if (c->timing) c->gemm_flops += 2.0 * g.M * g.N * g.K * n * g.nw;
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_fact(f, 1, c->stream));
  }
  g.wA = 0; g.wB = nc; g.wC = nc;
  g.M = c->nz / 2; g.N = 2 * c->ny; g.K = c->nz / 2;
  for (int b = 0; b < n; ++b) {
    g.b[b].A = c->d_G[1]; g.b[b].B = c->tmp[b]; g.b[b].C = c->tmp2[b];
  }
  {
    Timed t(c, KC_GEMM, true);
    HIPCHECK(launch_gemm(g, c->stream, QSS_PREC_FP64,
                           c->gemm_force >= 0 ? c->gemm_force : c->fact_var[1][n], t.e0, t.e1));
  }
  if (c->timing) c->gemm_flops += 2.0 * g.M * g.N * g.K * n * g.nw;
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_fact(f, 2, c->stream));
  }
  return QSS_OK;
}

// out_b = alpha * T_kz[b] . in_b . T_ky[b]: y-pass into tmp, then z-pass
// reuse: the split-bf16 operands may come from the producer of the inputs
// (c->xs_src records which planes xs holds); false for caller-owned planes
int transform_batch(qss_ctx *c, int n, const int *kz, const int *ky,
                    const double *const *in, double *const *out, double alpha,
                    const int *resid, const int *push, const double *const *pplane,
                    double beta = 0.0, bool one_wake = false, bool reuse = false) {
  if (n < 1 || n > kGemmMaxBatch) {
    qss_set_error("transform batch size");
    return QSS_E_ARG;
  }
  if (c->fact_use && !c->npc && !c->mixed)
    return transform_fact(c, n, kz, ky, in, out, alpha, resid, push, pplane, beta, one_wake);
  GemmArgs g{};
  g.M = c->nz; g.N = c->ny; g.K = c->ny;
  g.alpha = 1.0;
  g.nbatch = n;
  // batched wakes: every plane operand repeats per wake with the plane stride
  const long nc = c->ncell;
  g.nw = one_wake ? 1 : c->nw;
  g.wA = nc; g.wB = 0; g.wC = nc; g.wP = 0; g.wF = 0; g.wR = 0;
  g.F = c->F;
  g.Ff = c->Ff;
  g.resid = c->d_resid;
  if (c->npc) {
    // split-bf16: the inputs become npc bf16 planes (unless their producer
    // wrote them so); the y pass leaves its result transposed and split for
    // the z pass
    bool have = reuse;
    for (int b = 0; b < n; ++b) have = have && c->xs_src[b] == in[b];
    if (!have) {
      Timed t(c, KC_GEMM, true);
      HIPCHECK(launch_split_bf16(in, c->xs, n, c->ncell, c->npc, c->stream, t.e0, t.e1));
      for (int b = 0; b < n; ++b) c->xs_src[b] = reuse ? in[b] : nullptr;
    }
    g.tout = 1;
  }
  for (int b = 0; b < n; ++b) {
    g.b[b].A = in[b];
    g.b[b].B = c->d_T[ky[b] == QSS_KIND_C ? 0 : 1];
    g.b[b].C = c->tmp[b];
    g.b[b].P = nullptr;
    g.b[b].As = c->xs[b];
    g.b[b].Bts = c->Ts[ky[b] == QSS_KIND_C ? 0 : 1];   // T_y is symmetric
    g.b[b].Cts = c->ts[b];
    g.b[b].resid = -1;
    g.b[b].push = -1;
  }
  {
    Timed t(c, KC_GEMM, true);
    HIPCHECK(launch_gemm(g, c->stream, c->p.precision, gemm_pick(c, 0, n), t.e0, t.e1));
  }
  if (c->timing) c->gemm_flops += 2.0 * g.M * g.N * g.K * n * g.nw;
  g.M = c->nz; g.N = c->ny; g.K = c->nz;
  g.alpha = alpha;
  g.beta = beta;
  g.tout = 0;
  g.wA = 0; g.wB = nc; g.wC = nc; g.wP = nc; g.wF = nc * kPushComp; g.wR = kResidPerWake;
  for (int b = 0; b < n; ++b) {
    g.b[b].A = c->d_T[kz[b] == QSS_KIND_C ? 2 : 3];
    g.b[b].B = c->tmp[b];
    g.b[b].As = c->Ts[kz[b] == QSS_KIND_C ? 2 : 3];
    g.b[b].Bts = c->ts[b];
    g.b[b].Cts = nullptr;
    g.b[b].C = out[b];
    g.b[b].resid = resid ? resid[b] : -1;
    g.b[b].push = push ? push[b] : -1;
    g.b[b].P = pplane ? pplane[b] : nullptr;
  }
  {
    Timed t(c, KC_GEMM, true);
    HIPCHECK(launch_gemm(g, c->stream, c->p.precision, gemm_pick(c, 1, n), t.e0, t.e1));
  }
  // the outputs changed: split forms recorded for them are stale
  for (int b = 0; b < n; ++b)
    for (int q = 0; q < kGemmMaxBatch; ++q)
      if (c->xs_src[q] == out[b]) c->xs_src[q] = nullptr;
  if (c->timing) c->gemm_flops += 2.0 * g.M * g.N * g.K * n * g.nw;
  return QSS_OK;
}

// push field (P+C)/2 from the current planes (after host-side edits)
int refresh_push_field(qss_ctx *c) {
  HIPCHECK(launch_push_field(c));
  return QSS_OK;
}

// ---- the pieces of a layer step ---------------------------------------------
int do_beam(qss_ctx *c, int slice, int isplit) {
  Timed t(c, KC_BEAM);
  HIPCHECK(launch_beam_plane(c, slice, isplit));
  return QSS_OK;
}

int do_guess(qss_ctx *c, int first) {
  if (first) {
    // The reference zeroes every spectrum at the first layer (:479-494) but
    // then re-transforms the P layer, so only the beam spectrum (which
    // becomes the beam P spectrum below) really restarts from zero.  The P
    // spectra here are the context's state: zero after a reset, the split-0
    // result for later splits of the first layer.
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_fill_f64(c->S[SP_RHOB], c->ncell * c->nw, 0.0, c->stream));
    HIPCHECK(launch_fill_f64(c->S[SP_JXB], c->ncell * c->nw, 0.0, c->stream));
  }
  // beam P spectrum = previous beam spectrum (:637-646): swap buffers
  std::swap(c->S[SP_RHOBP], c->S[SP_RHOB]);
  std::swap(c->S[SP_JXBP], c->S[SP_JXB]);
  {
    const int kz[2] = {QSS_KIND_C, QSS_KIND_C}, ky[2] = {QSS_KIND_C, QSS_KIND_C};
    const double *in[2] = {c->C[QSS_RHOBEAM], c->C[QSS_JXBEAM]};
    double *out[2] = {c->S[SP_RHOB], c->S[SP_JXB]};
    CHECK(transform_batch(c, 2, kz, ky, in, out, 1.0, nullptr, nullptr, nullptr, 0.0, false,
                          true));
  }
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_guess_update(c, first));
  }
  // six inverse transforms -> C-layer E,B, with the push-field epilogue
  const int kz[6] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S, QSS_KIND_S, QSS_KIND_S, QSS_KIND_C};
  const int ky[6] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S};
  const double *in[6] = {c->S[SP_EX], c->S[SP_EY], c->S[SP_EZ],
                         c->S[SP_BX], c->S[SP_BY], c->S[SP_BZ]};
  double *out[6] = {c->C[QSS_EX], c->C[QSS_EY], c->C[QSS_EZ],
                    c->C[QSS_BX], c->C[QSS_BY], c->C[QSS_BZ]};
  const int push[6] = {0, 1, 2, 3, 4, 5};
  const double *pp[6] = {c->P[QSS_EX], c->P[QSS_EY], c->P[QSS_EZ],
                         c->P[QSS_BX], c->P[QSS_BY], c->P[QSS_BZ]};
  const double alpha = 1.0 / (4.0 * c->ny * c->nz);
  return transform_batch(c, 6, kz, ky, in, out, alpha, nullptr, push, pp, 0.0, false, true);
}

int do_push_deposit(qss_ctx *c, int full) {
  if (c->tiled) {
    {
      Timed t(c, KC_PUSH, true);
      HIPCHECK(launch_push_tile(c, full, c->deposit_rho, t.e0, t.e1));
    }
    {
      Timed t(c, KC_REDUCE);
      HIPCHECK(launch_combine(c, c->deposit_rho));
      if (full) HIPCHECK(launch_tile_reorder(c));
    }
    if (full) c->cur ^= 1;
    return QSS_OK;
  }
  forget_splits(c);   // the global deposit writes J without its split form
  {
    Timed t(c, KC_SORT);
    HIPCHECK(launch_fill_i32(c->cnt, c->ncell, 0, c->stream));
    if (full) HIPCHECK(launch_fill_i32(c->cnt_main, c->ncell, 0, c->stream));
    HIPCHECK(launch_fill_i32(c->d_counters, 1, 0, c->stream));
  }
  {
    Timed t(c, KC_PUSH);
    HIPCHECK(launch_push(c, full));
  }
  {
    Timed t(c, KC_SORT);
    HIPCHECK(launch_sort_scan(c, full));
    HIPCHECK(launch_scatter(c));
  }
  {
    Timed t(c, KC_REDUCE);
    HIPCHECK(launch_tile_reduce(c, full));
  }
  if (full) c->cur ^= 1;
  return QSS_OK;
}

int do_iterate(qss_ctx *c, int final_pass, double *res_dev) {
  {
    const int kz[3] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S};
    const int ky[3] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C};
    const double *in[3] = {c->C[QSS_JX], c->C[QSS_JY], c->C[QSS_JZ]};
    double *out[3] = {c->S[SP_JX], c->S[SP_JY], c->S[SP_JZ]};
    CHECK(transform_batch(c, 3, kz, ky, in, out, 1.0, nullptr, nullptr, nullptr, 0.0, false,
                          true));
  }
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_iterate_update(c, final_pass));   // also clears the residual slots
  }
  // E (residual + push field), B (push field); final pass: rho and psi too
  int kz[8] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S, QSS_KIND_S, QSS_KIND_S, QSS_KIND_C,
               QSS_KIND_C, QSS_KIND_C};
  int ky[8] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S,
               QSS_KIND_C, QSS_KIND_C};
  const double *in[8] = {c->S[SP_EX], c->S[SP_EY], c->S[SP_EZ], c->S[SP_BX],
                         c->S[SP_BY], c->S[SP_BZ], c->S[SP_RHO], c->S[SP_PSI]};
  double *out[8] = {c->C[QSS_EX], c->C[QSS_EY], c->C[QSS_EZ], c->C[QSS_BX],
                    c->C[QSS_BY], c->C[QSS_BZ], c->C[QSS_RHO], c->C[QSS_PSI]};
  int resid[8] = {0, 1, 2, -1, -1, -1, -1, -1};
  int push[8] = {0, 1, 2, 3, 4, 5, -1, -1};
  if (final_pass)
    for (int b = 0; b < 6; ++b) push[b] = -1;  // the next Guess rebuilds it
  const double *pp[8] = {c->P[QSS_EX], c->P[QSS_EY], c->P[QSS_EZ], c->P[QSS_BX],
                         c->P[QSS_BY], c->P[QSS_BZ], nullptr, nullptr};
  const double alpha = 1.0 / (4.0 * c->ny * c->nz);
  // rho and psi are outputs only (the next layer starts from the kept
  // spectra); a sweep that stores neither leaves them to its end
  const int nb = final_pass && !c->skip_rho_psi ? 8 : 6;
  CHECK(transform_batch(c, nb, kz, ky, in, out, alpha, resid, push, pp, 0.0, false, true));
  if (res_dev) {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_residual_final(c, res_dev));
  }
  return QSS_OK;
}

void swap_c_p(qss_ctx *c) {
  // Cell::operator= moves fields, currents, density (cells.cpp:25-47)
  for (int f = QSS_EX; f <= QSS_RHO; ++f) std::swap(c->C[f], c->P[f]);
  std::swap(c->C[QSS_PSI], c->P[QSS_PSI]);
}

// ---- particle snapshots (checkpoints and the adjoint tape) -------------------
inline size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

struct PartView {
  double *y, *z, *px, *py, *pz;
  int32_t *cell;
  int64_t *pid;
  int32_t *tstart;
};

size_t part_bytes(const qss_ctx *c) {
  const size_t np = c->np_cap, nt = (size_t)c->ntj * c->ntk + 1;
  return 5 * align256(np * 8) + align256(np * 4) + align256(np * 8) + align256(nt * 4);
}

PartView part_view(const qss_ctx *c, char *base) {
  const size_t np = c->np_cap;
  PartView v;
  size_t o = 0;
  double **d[5] = {&v.y, &v.z, &v.px, &v.py, &v.pz};
  for (auto p : d) { *p = reinterpret_cast<double *>(base + o); o += align256(np * 8); }
  v.cell = reinterpret_cast<int32_t *>(base + o); o += align256(np * 4);
  v.pid = reinterpret_cast<int64_t *>(base + o); o += align256(np * 8);
  v.tstart = reinterpret_cast<int32_t *>(base + o);
  return v;
}

// to = 0: context -> snapshot, 1: snapshot -> context
int copy_particles(qss_ctx *c, const PartView &v, int to_ctx) {
  const int b = c->cur;
  const size_t n = c->np_cap, nt = (size_t)c->ntj * c->ntk + 1;
  void *ctx[8] = {c->py_[b], c->pz_[b], c->ppx[b], c->ppy[b], c->ppz[b], c->pcell[b],
                  c->pid[b], c->tile_start[b]};
  void *snap[8] = {v.y, v.z, v.px, v.py, v.pz, v.cell, v.pid, v.tstart};
  const size_t bytes[8] = {n * 8, n * 8, n * 8, n * 8, n * 8, n * 4, n * 8, nt * 4};
  // hipMemcpyDefault: the snapshot may be device memory (checkpoints, a
  // layer sent to another chunk) or host memory
  for (int i = 0; i < 8; ++i)
    if (ctx[i])   // no tile starts without the tiled particle path
      HIPCHECK(hipMemcpyAsync(to_ctx ? ctx[i] : snap[i], to_ctx ? snap[i] : ctx[i], bytes[i],
                              hipMemcpyDefault, c->stream));
  return QSS_OK;
}

// the layer-start particles into a device tape slot, all arrays in one launch
int tape_particles(qss_ctx *c, const PartView &v) {
  const int b = c->cur;
  const long n = c->np_cap, nt = (long)c->ntj * c->ntk + 1;
  CopyBufs cb{{c->py_[b], c->pz_[b], c->ppx[b], c->ppy[b], c->ppz[b], c->pcell[b], c->pid[b],
               c->tile_start[b]},
              {v.y, v.z, v.px, v.py, v.pz, v.cell, v.pid, v.tstart},
              {n * 8, n * 8, n * 8, n * 8, n * 8, n * 4, n * 8, nt * 4}};
  if (!cb.src[7]) cb.nbytes[7] = 0;   // no tile starts without the tiled particle path
  HIPCHECK(launch_copy_bufs(cb, c->stream));
  return QSS_OK;
}

// tape slot of one layer: particles at the layer start + the push field of
// each of the n_iter+1 passes
size_t tape_bytes(const qss_ctx *c) {
  return part_bytes(c) + (size_t)(c->p.n_iter + 1) * kPushComp * align256(c->ncell * 8);
}
double *tape_field(const qss_ctx *c, char *slot, int pass) {
  return reinterpret_cast<double *>(slot + part_bytes(c) +
                                    (size_t)pass * kPushComp * align256(c->ncell * 8));
}

// tape != null: record the layer's start particles and every pass's push
// field (the reverse sweep replays the pushes from them)
int do_layer_step(qss_ctx *c, int slice, int isplit, double *res_dev, char *tape = nullptr) {
  if (tape) CHECK(tape_particles(c, part_view(c, tape)));
  // recording: the inverse transforms' epilogue writes each pass's push field
  // straight into its tape slot, which that pass's push then reads
  double *const F0 = c->F;
  auto field_of = [&](int pass) {
    if (tape) c->F = tape_field(c, tape, pass);
  };
  // the corrector takes rho from continuity in spectral space and the final
  // pass (or the sweep's end) rebuilds the C-layer density from it, so the
  // step's pushes deposit Jx, Jy, Jz only
  struct Restore {
    qss_ctx *c;
    double *F;
    bool rho;
    ~Restore() {
      c->F = F;
      c->deposit_rho = rho;
    }
  } restore{c, F0, c->deposit_rho};
  c->deposit_rho = false;
  CHECK(do_beam(c, slice, isplit));
  field_of(0);
  CHECK(do_guess(c, slice == 0));
  for (int it = 0; it < c->p.n_iter; ++it) {
    CHECK(do_push_deposit(c, 0));
    field_of(it + 1);
    CHECK(do_iterate(c, 0, nullptr));
  }
  CHECK(do_push_deposit(c, 1));
  c->F = F0;   // the final corrector's push field is not replayed
  CHECK(do_iterate(c, 1, res_dev));
  swap_c_p(c);
  return QSS_OK;
}

// checkpoint: particles + P planes + spectral state at a layer start
// header of a packed layer state (qss_pack_layer): magic "QSSLAYR1", format
// version, grid, tiling and capacity, padded to 256 bytes so the arrays after
// it keep their 256-byte alignment
constexpr size_t kStateHeader = 256;
struct StateHeader {
  uint64_t magic;
  int32_t version, ny, nz, ntj, ntk, tiled;
  int64_t np_cap;
  uint64_t payload;
  char pad[kStateHeader - 48];
};
static_assert(sizeof(StateHeader) == kStateHeader, "header size");

constexpr int kRecPlanes = 12, kRecAttr = 12;
// record planes 0..9 (para.cpp:109-118)
constexpr int kRecField[10] = {QSS_BX, QSS_BY, QSS_BZ, QSS_EX, QSS_EY, QSS_EZ,
                               QSS_JX, QSS_JY, QSS_JZ, QSS_RHO};

size_t chk_bytes(const qss_ctx *c) {
  return part_bytes(c) + (size_t)(QSS_NFIELD + SP_N) * align256(c->ncell * 8);
}
StateHeader state_header(const qss_ctx *c) {
  StateHeader h;
  std::memset(&h, 0, sizeof h);
  h.magic = 0x315259414C535351ull;   // "QSSLAYR1" little-endian
  h.version = 1;
  h.ny = c->ny; h.nz = c->nz; h.ntj = c->ntj; h.ntk = c->ntk; h.tiled = c->tiled ? 1 : 0;
  h.np_cap = c->np_cap;
  h.payload = chk_bytes(c);
  return h;
}
int checkpoint(qss_ctx *c, char *base, int restore) {
  CHECK(copy_particles(c, part_view(c, base), restore));
  const size_t pl = align256(c->ncell * 8), bytes = c->ncell * sizeof(double);
  char *p = base + part_bytes(c);
  for (int f = 0; f < QSS_NFIELD; ++f, p += pl)
    HIPCHECK(hipMemcpyAsync(restore ? (void *)c->P[f] : (void *)p,
                            restore ? (const void *)p : (const void *)c->P[f], bytes,
                            hipMemcpyDefault, c->stream));
  for (int s = 0; s < SP_N; ++s, p += pl)
    HIPCHECK(hipMemcpyAsync(restore ? (void *)c->S[s] : (void *)p,
                            restore ? (const void *)p : (const void *)c->S[s], bytes,
                            hipMemcpyDefault, c->stream));
  return QSS_OK;
}

void adj_graph_free(qss_ctx *c) {
  if (c->adj_gexec) (void)hipGraphExecDestroy(c->adj_gexec);
  if (c->adj_graph) (void)hipGraphDestroy(c->adj_graph);
  c->adj_gexec = nullptr;
  c->adj_graph = nullptr;
}

void adj_free(qss_ctx *c) {
  adj_graph_free(c);   // the graph holds the buffers freed here
  AdjState &a = c->adj;
  for (auto &p : a.A) dfree(p);
  dfree(a.pbarEB); dfree(a.fbar); dfree(a.jbar); dfree(a.bbar);
  dfree(a.pbar[0]); dfree(a.pbar[1]); dfree(a.fwin); dfree(a.qbeam);
  dfree(a.chk); dfree(a.tape); dfree(a.exbar);
  a = AdjState();
}

int adj_check(const qss_ctx *c) {
  if (!c->tiled || c->p.precision != QSS_PREC_FP64 || c->p.shape != QSS_SHAPE_CIC ||
      c->p.nx_split != 1 || c->nw != 1) {
    qss_set_error("the adjoint sweep needs one wake, FP64, CIC, nx_split = 1 and ny, nz "
                  "multiples of 16");
    return QSS_E_ARG;
  }
  return QSS_OK;
}

int adj_alloc(qss_ctx *c, int nrun, int K) {
  AdjState &a = c->adj;
  const int nchk = (nrun + K - 1) / K;
  if (a.chk && a.K == K && a.nrun == nrun) return QSS_OK;
  adj_free(c);
  const long nc = c->ncell, np = c->np_cap, nt = (long)c->ntj * c->ntk;
  int r;
  for (int s = 0; s < SP_N; ++s)
    if ((r = dalloc(&a.A[s], nc)) != QSS_OK) return r;
  a.chk_stride = chk_bytes(c);
  a.tape_stride = tape_bytes(c);
  if ((r = dalloc(&a.pbarEB, nc * kPushComp)) || (r = dalloc(&a.fbar, nc * kPushComp)) ||
      (r = dalloc(&a.jbar, nc * 3)) || (r = dalloc(&a.bbar, nc * 2)) ||
      (r = dalloc(&a.pbar[0], np * 5)) || (r = dalloc(&a.pbar[1], np * 5)) ||
      (r = dalloc(&a.fwin, nt * kMaxSplit * 289 * 6)) || (r = dalloc(&a.qbeam, (long)c->p.nslices * 2 * kBeamAdjBlocks)) ||
      (r = dalloc(&a.chk, a.chk_stride * nchk)) || (r = dalloc(&a.tape, a.tape_stride * K)) ||
      (r = dalloc(&a.exbar, (size_t)nrun * nc)))
    return r;
  // the zero fills ran on the null stream
  HIPCHECK(hipDeviceSynchronize());
  a.K = K;
  a.nrun = nrun;
  a.nchk = nchk;
  return QSS_OK;
}

// One push pass of the reverse sweep: adjoint push, windows -> field adjoint,
// P-layer share of it, and the C-layer share into the spectral E,B adjoints.
int adj_push_pass(qss_ctx *c, char *slot, int pass, int full, int cur_adj) {
  AdjState &s = c->adj;
  const PartView v = part_view(c, slot);
  AdjTileArgs g;
  g.y = v.y; g.z = v.z; g.px = v.px; g.py = v.py; g.pz = v.pz;
  g.cell = v.cell; g.pid = v.pid; g.tile_start = v.tstart;
  g.ntj = c->ntj; g.ntk = c->ntk; g.ny = c->ny; g.nz = c->nz;
  g.hx = c->hxs; g.hy = c->hy; g.hz = c->hz; g.weight = c->pweight;
  for (int q = 0; q < 3; ++q) g.bext[q] = c->p.b_ext[q];
  g.F = tape_field(c, slot, pass);
  g.Jb = s.jbar;
  g.full = full;
  g.pin = s.pbar[cur_adj];
  g.pout = s.pbar[cur_adj ^ 1];
  g.np = c->np_cap;
  g.fwin = s.fwin;
  g.nsplit = c->nsplit;
  g.counters = c->d_counters;
  {
    Timed t(c, KC_PUSH);
    HIPCHECK(launch_push_tile_adj(c, g));
  }
  {
    Timed t(c, KC_REDUCE);
    // the layer's first reverse pass (the full step) starts the P-layer share
    HIPCHECK(launch_fbar_combine(c, s.fwin, s.fbar, s.pbarEB, full ? 1 : 2));
  }
  const int kz[6] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S, QSS_KIND_S, QSS_KIND_S, QSS_KIND_C};
  const int ky[6] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S};
  const double *in[6];
  double *out[6];
  for (int q = 0; q < 6; ++q) {
    in[q] = s.fbar + q * c->ncell;
    out[q] = s.A[SP_EX + q];
  }
  return transform_batch(c, 6, kz, ky, in, out, 0.5 / (4.0 * c->ny * c->nz), nullptr,
                         nullptr, nullptr, 1.0);
}

// spectral J adjoints -> real-space deposit adjoints; the J spectra were
// overwritten by the forward transform, so their adjoints restart at zero
int adj_deposit_transform(qss_ctx *c) {
  AdjState &s = c->adj;
  const int kz[3] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S};
  const int ky[3] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C};
  const double *in[3] = {s.A[SP_JX], s.A[SP_JY], s.A[SP_JZ]};
  double *out[3] = {s.jbar, s.jbar + c->ncell, s.jbar + 2 * c->ncell};
  CHECK(transform_batch(c, 3, kz, ky, in, out, 1.0, nullptr, nullptr, nullptr));
  Timed t(c, KC_SPEC);
  HIPCHECK(launch_zero_bufs(ZeroBufs{{s.A[SP_JX], s.A[SP_JY], s.A[SP_JZ], nullptr},
                                      {c->ncell, c->ncell, c->ncell, 0}},
                             c->stream));
  return QSS_OK;
}

// Reverse of do_layer_step for layer `slice` (tape slot), the loss adjoint of
// the layer's final Ex plane being plane `slice` of exbar (device, may be
// null).  Capturing: the slice is read from d_slice at replay.
int adj_layer_step(qss_ctx *c, int slice, char *slot, const double *exbar, int &cur_adj) {
  AdjState &s = c->adj;
  const long nc = c->ncell;
  const double sinv = 1.0 / (4.0 * c->ny * c->nz);
  // final fields of the layer: loss + the next layer's P-layer gathers
  if (exbar)
    HIPCHECK(c->capturing ? launch_axpy(s.pbarEB, exbar, 1.0, nc, c->stream, c->d_slice)
                          : launch_axpy(s.pbarEB, exbar + (size_t)slice * nc, 1.0, nc, c->stream));
  {
    const int kz[6] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S, QSS_KIND_S, QSS_KIND_S, QSS_KIND_C};
    const int ky[6] = {QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S, QSS_KIND_C, QSS_KIND_S};
    const double *in[6];
    double *out[6];
    for (int q = 0; q < 6; ++q) {
      in[q] = s.pbarEB + q * nc;
      out[q] = s.A[SP_EX + q];
    }
    // pbarEB is reassigned by the full step's field-adjoint combine below
    CHECK(transform_batch(c, 6, kz, ky, in, out, sinv, nullptr, nullptr, nullptr, 1.0));
  }
  // final Iterate, full-step push
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_iterate_adj(c, s.A, 1));
  }
  CHECK(adj_deposit_transform(c));
  CHECK(adj_push_pass(c, slot, c->p.n_iter, 1, cur_adj));
  // corrector passes in reverse
  for (int it = c->p.n_iter - 1; it >= 0; --it) {
    {
      Timed t(c, KC_SPEC);
      HIPCHECK(launch_iterate_adj(c, s.A, 0));
    }
    CHECK(adj_deposit_transform(c));
    CHECK(adj_push_pass(c, slot, it, 0, cur_adj));
  }
  {
    Timed t(c, KC_SPEC);
    HIPCHECK(launch_guess_adj(c, s.A, slice == 0));
  }
  // beam spectra -> beam planes -> profile derivative
  {
    const int kz[2] = {QSS_KIND_C, QSS_KIND_C}, ky[2] = {QSS_KIND_C, QSS_KIND_C};
    const double *in[2] = {s.A[SP_RHOB], s.A[SP_JXB]};
    double *out[2] = {s.bbar, s.bbar + nc};
    CHECK(transform_batch(c, 2, kz, ky, in, out, 1.0, nullptr, nullptr, nullptr));
    Timed t(c, KC_BEAM);
    HIPCHECK(launch_beam_adj(c, slice, s.bbar, s.bbar + nc, s.qbeam));
    HIPCHECK(launch_zero_bufs(ZeroBufs{{s.A[SP_RHOB], s.A[SP_JXB], nullptr, nullptr},
                                        {nc, nc, 0, 0}},
                               c->stream));
  }
  // this layer's beam P spectrum is the previous layer's beam spectrum
  std::swap(s.A[SP_RHOB], s.A[SP_RHOBP]);
  cur_adj ^= 1;
  return QSS_OK;
}

// the finished layer into the history (movepart_splitlayer.cpp:467-473);
// wake w of a batch at hist + (w * nslices + s) * ncell.  The layer's fields
// are in P after the closing swap; its beam planes stay in C (the swap
// leaves them, the next layer overwrites them).
const double *history_source(const qss_ctx *c, int f) {
  return f == QSS_RHOBEAM || f == QSS_JXBEAM ? c->C[f] : c->P[f];
}

int store_history(qss_ctx *c, int s) {
  const size_t plane = c->ncell * sizeof(double);
  for (int f = 0; f < QSS_NFIELD; ++f)
    if (c->hist[f])
      HIPCHECK(hipMemcpy2DAsync(c->hist[f] + (size_t)s * c->ncell, plane * c->p.nslices,
                                history_source(c, f), plane, plane, c->nw,
                                hipMemcpyDeviceToDevice, c->stream));
  return QSS_OK;
}

void graph_free(qss_ctx *c) {
  if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
  if (c->graph) (void)hipGraphDestroy(c->graph);
  c->gexec = nullptr;
  c->graph = nullptr;
  adj_graph_free(c);
}

// Layers s and s+1 by one graph launch.  The graph holds two layer steps
// captured from the current pointer state; the layer index is read from
// d_slice by the beam, residual and history kernels and advanced inside the
// graph.  A replay from the other pointer parity runs one layer eagerly
// first (returns 1 then: one layer done).
int sweep_graph_pair(qss_ctx *c, int s, int *done) {
  const bool match = c->gexec && c->g_key_c0 == c->C[0] && c->g_key_cur == c->cur &&
                     c->g_key_rb == c->S[SP_RHOB] && c->g_key_skip == c->skip_rho_psi;
  if (c->gexec && !match) {
    CHECK(do_layer_step(c, s, 0, c->d_res_hist + s));
    CHECK(store_history(c, s));
    *done = 1;
    return QSS_OK;
  }
  if (!c->gexec) {
    c->g_key_c0 = c->C[0];
    c->g_key_cur = c->cur;
    c->g_key_rb = c->S[SP_RHOB];
    c->g_key_skip = c->skip_rho_psi;
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int r = QSS_OK;
    for (int q = 0; q < 2 && r == QSS_OK; ++q) {
      // not the first layer; residual slot of layer 0 (+ d_slice at replay)
      r = do_layer_step(c, 1, 0, c->d_res_hist);
      for (int f = 0; f < QSS_NFIELD && r == QSS_OK; ++f)
        if (c->hist[f] && launch_store_hist(c, history_source(c, f), c->hist[f]) != hipSuccess)
          r = QSS_E_HIP;
      if (r == QSS_OK && launch_add_i32(c->d_slice, 1, c->stream) != hipSuccess) r = QSS_E_HIP;
    }
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != QSS_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (r == QSS_OK) qss_set_error(std::string("graph capture: ") + hipGetErrorString(e));
      return r != QSS_OK ? r : QSS_E_HIP;
    }
    c->graph = g;
    HIPCHECK(hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0));
  }
  HIPCHECK(launch_fill_i32(c->d_slice, 1, s, c->stream));
  HIPCHECK(hipGraphLaunch(c->gexec, c->stream));
  *done = 2;
  return QSS_OK;
}

// One full checkpoint segment of the reverse sweep, layers s0 .. s0+K-1
// (recomputation with the tape, then the reverse layers), by one graph
// launch.  The graph is captured at the first full segment; the slice index
// of every slice-dependent launch is read from d_slice, which the graph
// advances through the recomputation and steps back through the reverse.  A
// segment met in another pointer state (odd K, or a changed loss adjoint
// array) returns *done = false and runs eagerly.  The checkpoint restore
// before it stays outside the graph.
int adj_segment_graph(qss_ctx *c, int s0, const double *exbar, int &cur_adj, bool *done) {
  AdjState &a = c->adj;
  *done = false;
  const void *key[6] = {c->C[0], c->S[SP_RHOB], a.A[SP_RHOB], exbar, a.tape, a.pbar[0]};
  const int keyi[4] = {c->cur, cur_adj, a.K, c->skip_rho_psi ? 1 : 0};
  bool match = c->adj_gexec != nullptr;
  for (int k = 0; k < 6 && match; ++k) match = c->adj_key[k] == key[k];
  for (int k = 0; k < 4 && match; ++k) match = c->adj_key_i[k] == keyi[k];
  if (c->adj_gexec && !match) return QSS_OK;
  if (!c->adj_gexec) {
    for (int k = 0; k < 6; ++k) c->adj_key[k] = key[k];
    for (int k = 0; k < 4; ++k) c->adj_key_i[k] = keyi[k];
    HIPCHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int r = QSS_OK;
    // the captures run no work but advance the host-side pointer state (C/P
    // and beam-spectrum swaps, adjoint parity) by one segment, as the replay
    // does on the device
    for (int l = 0; l < a.K && r == QSS_OK; ++l) {
      // not the first layer; residual slot of layer 0 (+ d_slice at replay)
      r = do_layer_step(c, 1, 0, c->d_res_hist, a.tape + (size_t)l * a.tape_stride);
      if (r == QSS_OK && launch_add_i32(c->d_slice, 1, c->stream) != hipSuccess) r = QSS_E_HIP;
    }
    for (int l = a.K - 1; l >= 0 && r == QSS_OK; --l) {
      if (launch_add_i32(c->d_slice, -1, c->stream) != hipSuccess) {
        r = QSS_E_HIP;
        break;
      }
      r = adj_layer_step(c, 1, a.tape + (size_t)l * a.tape_stride, exbar, cur_adj);
    }
    c->capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (r != QSS_OK || e != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      if (r == QSS_OK) qss_set_error(std::string("graph capture: ") + hipGetErrorString(e));
      return r != QSS_OK ? r : QSS_E_HIP;
    }
    c->adj_graph = g;
    HIPCHECK(hipGraphInstantiate(&c->adj_gexec, g, nullptr, nullptr, 0));
  }
  // a replay advances the device state only; the host-side state of K layers
  // (K even) is the state it started from: every swap and the parity come back
  HIPCHECK(launch_fill_i32(c->d_slice, 1, s0, c->stream));
  HIPCHECK(hipGraphLaunch(c->adj_gexec, c->stream));
  *done = true;
  return QSS_OK;
}

int check_status(qss_ctx *c) {
  int32_t h[2] = {0, 0};
  HIPCHECK(hipMemcpyAsync(h, c->d_counters, sizeof(h), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (h[1] & 2) { qss_set_error("sub-cycle overflow buffer exhausted"); return QSS_E_OVERFLOW; }
  if (h[1] & 4) { qss_set_error("displacement needs more than 62 sub-steps"); return QSS_E_PARTICLE; }
  if (h[1] & 8) { qss_set_error("deposit outside the tile window"); return QSS_E_OVERFLOW; }
  if (h[1] & 1) { qss_set_error("particle outside its cell or non-finite"); return QSS_E_PARTICLE; }
  return QSS_OK;
}

// ---- evolving driver: host side -------------------------------------------
// splitmix64 draws of the frozen lattice's weight noise (the same integer
// arithmetic as the beam kernels, so both lattices carry the same weights)
uint64_t host_mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}
double host_unif_pm1(uint64_t seed, uint64_t a, uint64_t b) {
  const uint64_t h = host_mix64(seed ^ host_mix64(a ^ host_mix64(b)));
  return (double)(h >> 11) * (1.0 / 9007199254740992.0) * 2.0 - 1.0;
}

void driver_free(DriverState &d) {
  dfree(d.ci); dfree(d.cj); dfree(d.ck); dfree(d.bunch);
  dfree(d.x); dfree(d.y); dfree(d.z); dfree(d.px); dfree(d.py); dfree(d.pz); dfree(d.w);
  d.np = 0;
}

// fixed-point scale 2^e of the deposit: |any node sum| <= W = sum |w*pol|
// (trilinear weights in [0,1] summing to 1, |Vx| < 1), so W * 2^e < 2^61
// leaves room for the per-contribution rounding in a signed 64-bit sum
double driver_scale(const std::vector<double> &w, const std::vector<int32_t> &b,
                    const qss_params &p) {
  double W = 0.0;
  for (size_t n = 0; n < w.size(); ++n) W += std::fabs(w[n] * p.bunch[b[n]].polarity);
  if (!(W > 0.0)) return 1.0;
  int ex = 0;
  (void)std::frexp(W, &ex);   // W < 2^ex
  return std::ldexp(1.0, 61 - ex);
}

// particle arrays of n entries from host vectors
int driver_upload(qss_ctx *c, long n, const std::vector<int32_t> &cell,
                  const std::vector<double> &frac, const std::vector<double> &mom,
                  const std::vector<double> &w, const std::vector<int32_t> &bunch) {
  DriverState &d = c->drv;
  driver_free(d);
  std::vector<int32_t> ci(n), cj(n), ck(n);
  std::vector<double> x(n), y(n), z(n), px(n), py(n), pz(n);
  for (long q = 0; q < n; ++q) {
    ci[q] = cell[3 * q]; cj[q] = cell[3 * q + 1]; ck[q] = cell[3 * q + 2];
    x[q] = frac[3 * q]; y[q] = frac[3 * q + 1]; z[q] = frac[3 * q + 2];
    px[q] = mom[3 * q]; py[q] = mom[3 * q + 1]; pz[q] = mom[3 * q + 2];
  }
  int32_t **ip[4] = {&d.ci, &d.cj, &d.ck, &d.bunch};
  const std::vector<int32_t> *iv[4] = {&ci, &cj, &ck, &bunch};
  for (int a = 0; a < 4; ++a) {
    CHECK(dalloc(ip[a], (size_t)n));
    if (n) HIPCHECK(hipMemcpy(*ip[a], iv[a]->data(), n * sizeof(int32_t), hipMemcpyHostToDevice));
  }
  double **dp[7] = {&d.x, &d.y, &d.z, &d.px, &d.py, &d.pz, &d.w};
  const std::vector<double> *dv[7] = {&x, &y, &z, &px, &py, &pz, &w};
  for (int a = 0; a < 7; ++a) {
    CHECK(dalloc(dp[a], (size_t)n));
    if (n) HIPCHECK(hipMemcpy(*dp[a], dv[a]->data(), n * sizeof(double), hipMemcpyHostToDevice));
  }
  d.np = n;
  d.scale = driver_scale(w, bunch, c->p);
  graph_free(c);   // a captured beam-plane launch holds the old scale
  return QSS_OK;
}

int do_reset(qss_ctx *c) {
  forget_splits(c);
  for (int f = 0; f < QSS_NFIELD; ++f) {
    HIPCHECK(launch_fill_f64(c->C[f], c->ncell * c->nw, 0.0, c->stream));
    HIPCHECK(launch_fill_f64(c->P[f], c->ncell * c->nw, 0.0, c->stream));
  }
  for (int s = 0; s < SP_N; ++s)
    HIPCHECK(launch_fill_f64(c->S[s], c->ncell * c->nw, 0.0, c->stream));
  HIPCHECK(launch_fill_f64(c->F, c->ncell * kPushComp * c->nw, 0.0, c->stream));
  if (c->Ff)   // +0.0f is the all-zero bit pattern
    HIPCHECK(launch_fill_i32(reinterpret_cast<int32_t *>(c->Ff), c->ncell * kPushComp * c->nw, 0,
                             c->stream));
  HIPCHECK(launch_fill_i32(c->d_counters, 2, 0, c->stream));
  c->cur = 0;
  c->np = c->np_cap;
  if (c->tiled) HIPCHECK(launch_seed_tiled(c));
  else HIPCHECK(launch_seed_plasma(c));
  return QSS_OK;
}

}  // namespace

void qss_set_error(const std::string &m) { g_err = m; }

extern "C" {

const char *qss_last_error(void) { return g_err.c_str(); }

int qss_create(const qss_params *p, int device, uint32_t history_mask, qss_ctx **out) {
  if (!out) { qss_set_error("null out"); return QSS_E_ARG; }
  *out = nullptr;
  CHECK(validate(p));
  int ndev = 0;
  HIPCHECK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) { qss_set_error("bad device"); return QSS_E_ARG; }
  HIPCHECK(hipSetDevice(device));
  qss_ctx *c = new qss_ctx();
  c->p = *p;
  if (p->nbins > 0) c->bins.assign(p->bins, p->bins + p->nbins);
  c->p.bins = nullptr;
  c->nw = p->nwake > 1 ? p->nwake : 1;
  c->wbunch.resize((size_t)c->nw * QSS_MAX_BUNCH);
  for (int w = 0; w < c->nw; ++w)
    for (int b = 0; b < QSS_MAX_BUNCH; ++b)
      c->wbunch[(size_t)w * QSS_MAX_BUNCH + b] =
          (p->wake_bunch && b < p->nbunch) ? p->wake_bunch[(size_t)w * p->nbunch + b]
                                            : p->bunch[b];
  c->p.wake_bunch = nullptr;
  c->device = device;
  c->ny = p->ny; c->nz = p->nz;
  c->ncell = (long)p->ny * p->nz;
  c->np_cap = c->ncell * p->ppc_y * p->ppc_z;
  c->np = c->np_cap;
  c->ov_cap = std::max<long>(c->np_cap, 4096);
  c->hxs = p->hx / p->nx_split;
  c->hy = p->ly / p->ny;
  c->hz = p->lz / p->nz;
  c->pweight = (1.0 / (p->ppc_y * p->ppc_z)) * p->dens;
  c->nblk = (int)((c->ncell + 1023) / 1024);
  c->mixed = p->precision != QSS_PREC_FP64;
  c->tiled = p->ny % kTile == 0 && p->nz % kTile == 0 && !p->fix_subcycle &&
             (c->mixed || std::getenv("QSS_GLOBAL_DEPOSIT") == nullptr);
  c->ntj = p->ny / kTile;
  c->ntk = p->nz / kTile;
  // default work-groups per tile: about a thousand push work-groups per wake
  // (a function of the grid only, so every wake of a batch sums alike);
  // qss_autotune / qss_set_particle_split override it
  {
    const long nt = (long)c->ntj * c->ntk;
    c->nsplit = nt >= 1024 ? 1 : nt >= 512 ? 2 : 4;
  }
  int r = QSS_OK;
  auto fail = [&](int code) { qss_destroy(c); return code; };
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    qss_set_error("hipStreamCreate failed");
    return fail(QSS_E_HIP);
  }
  const long nc = c->ncell, nw = c->nw, ncw = nc * nw;
  for (int i = 0; i < 2; ++i) {
    const int n = i == 0 ? c->ny : c->nz;
    if ((r = dalloc(&c->d_T[2 * i], (size_t)n * n)) != QSS_OK) return fail(r);
    if ((r = dalloc(&c->d_T[2 * i + 1], (size_t)n * n)) != QSS_OK) return fail(r);
    std::vector<double> tc, ts;
    make_tables(n, tc, ts);
    if (hipMemcpy(c->d_T[2 * i], tc.data(), tc.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->d_T[2 * i + 1], ts.data(), ts.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
      qss_set_error("table upload failed");
      return fail(QSS_E_HIP);
    }
  }
  for (int f = 0; f < QSS_NFIELD; ++f) {
    if ((r = dalloc(&c->C[f], ncw)) != QSS_OK) return fail(r);
    if ((r = dalloc(&c->P[f], ncw)) != QSS_OK) return fail(r);
  }
  for (int s = 0; s < SP_N; ++s)
    if ((r = dalloc(&c->S[s], ncw)) != QSS_OK) return fail(r);
  for (int b = 0; b < kGemmMaxBatch; ++b)
    if ((r = dalloc(&c->tmp[b], ncw)) != QSS_OK) return fail(r);
  // factorized transforms (fact.hip) for large FP64 grids
  c->fact_ok = p->precision == QSS_PREC_FP64 && c->ny % 8 == 0 && c->nz % 8 == 0 &&
               c->ny >= kFactMinN && c->nz >= kFactMinN;
  // dense by default where the twiddles need kernels of their own: at 512^2
  // those three kernels per transform cost more than the halved GEMMs save
  // (0.718 against 0.705 ms per layer step); the fused form (below) is the
  // default where it applies
  c->fact_use = false;
  if (c->fact_ok) {
    for (int b = 0; b < kGemmMaxBatch; ++b)
      if ((r = dalloc(&c->tmp2[b], ncw)) != QSS_OK) return fail(r);
    for (int i = 0; i < 2; ++i) {
      const int n = i == 0 ? c->ny : c->nz;
      std::vector<double> g, tw;
      fact_matrix(n, i == 1, g);
      fact_twiddles(n, tw);
      double *twd = nullptr;
      if ((r = dalloc(&c->d_G[i], g.size())) != QSS_OK || (r = dalloc(&twd, tw.size())) != QSS_OK)
        return fail(r);
      c->d_tw[i] = reinterpret_cast<double2 *>(twd);
      if (hipMemcpy(c->d_G[i], g.data(), g.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
          hipMemcpy(twd, tw.data(), tw.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
        qss_set_error("factorized transform table upload failed");
        return fail(QSS_E_HIP);
      }
    }
    // the fused form where its tile shapes fit (sizes multiples of 128)
    int v0 = -1, v1 = -1;
    for (int v = fact_fused_variants(0) - 1; v >= 0; --v)
      if (fact_fused_shape_ok(0, v, c->ny, c->nz)) v0 = v;
    for (int v = fact_fused_variants(1) - 1; v >= 0; --v)
      if (fact_fused_shape_ok(1, v, c->ny, c->nz)) v1 = v;
    if (v0 >= 0 && v1 >= 0) {
      for (int i = 0; i < 2; ++i) {
        std::vector<double> g, tw;
        fact_fused_matrix(i == 0 ? c->ny : c->nz, i, g);
        fact_fused_twiddles(i == 0 ? c->ny : c->nz, tw);
        double *twd = nullptr;
        if ((r = dalloc(&c->d_Gf[i], g.size())) != QSS_OK || (r = dalloc(&twd, tw.size())) != QSS_OK)
          return fail(r);
        c->d_twf[i] = reinterpret_cast<double2 *>(twd);
        if (hipMemcpy(c->d_Gf[i], g.data(), g.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(twd, tw.data(), tw.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
          qss_set_error("fused transform table upload failed");
          return fail(QSS_E_HIP);
        }
      }
      for (int i = 0; i < 2; ++i) {
        std::vector<double> g;
        fact_fused_matrix(i == 0 ? c->ny : c->nz, 3 + i, g);
        if ((r = dalloc(&c->d_Gfd[i], g.size())) != QSS_OK) return fail(r);
        if (hipMemcpy(c->d_Gfd[i], g.data(), g.size() * 8, hipMemcpyHostToDevice) != hipSuccess) {
          qss_set_error("fused transform table upload failed");
          return fail(QSS_E_HIP);
        }
      }
      for (int b = 0; b <= kGemmMaxBatch; ++b) { c->ffv[0][b] = v0; c->ffv[1][b] = v1; }
      c->ffused_ok = true;
      // faster than the dense tables at 256^2 (0.293 against 0.300 ms per
      // layer step), 384^2 (0.416 / 0.457) and 512^2 (0.591 / 0.695)
      c->fact_use = true;
    }
  }
  if ((r = dalloc(&c->F, ncw * kPushComp)) != QSS_OK) return fail(r);
  if (c->mixed && (r = dalloc(&c->Ff, ncw * kPushComp)) != QSS_OK) return fail(r);
  if (nw > 1) {
    if ((r = dalloc(&c->d_wbunch, c->wbunch.size())) != QSS_OK) return fail(r);
    if (hipMemcpy(c->d_wbunch, c->wbunch.data(), c->wbunch.size() * sizeof(qss_bunch),
                  hipMemcpyHostToDevice) != hipSuccess) {
      qss_set_error("bunch upload failed");
      return fail(QSS_E_HIP);
    }
  }
  c->npc = gemm_pieces(p->precision);
  if (c->npc) {
    for (int i = 0; i < 4; ++i) {
      const long n = i < 2 ? c->ny : c->nz;
      if ((r = dalloc(&c->Ts[i], (size_t)c->npc * n * n)) != QSS_OK) return fail(r);
    }
    for (int b = 0; b < kGemmMaxBatch; ++b)
      if ((r = dalloc(&c->xs[b], (size_t)c->npc * nc)) != QSS_OK ||
          (r = dalloc(&c->ts[b], (size_t)c->npc * nc)) != QSS_OK)
        return fail(r);
  }
  if (!c->bins.empty()) {
    if ((r = dalloc(&c->d_bins, c->bins.size())) != QSS_OK) return fail(r);
    if (hipMemcpy(c->d_bins, c->bins.data(), c->bins.size() * 8, hipMemcpyHostToDevice) !=
        hipSuccess) {
      qss_set_error("bins upload failed");
      return fail(QSS_E_HIP);
    }
  }
  const long np = c->np_cap * nw, ov = c->ov_cap * nw;
  for (int b = 0; b < 2; ++b) {
    if ((r = dalloc(&c->py_[b], np)) || (r = dalloc(&c->pz_[b], np)) ||
        (r = dalloc(&c->ppx[b], np)) || (r = dalloc(&c->ppy[b], np)) ||
        (r = dalloc(&c->ppz[b], np)) || (r = dalloc(&c->pcell[b], np)) ||
        (r = dalloc(&c->pid[b], np)))
      return fail(r);
  }
  if ((r = dalloc(&c->ev_cell, np)) || (r = dalloc(&c->ev_rank, np)) ||
      (r = dalloc(&c->ev_sub, np)) || (r = dalloc(&c->ev_y, np)) ||
      (r = dalloc(&c->ev_z, np)) || (r = dalloc(&c->ev_w, np * 4)) ||
      (r = dalloc(&c->ev_p, np * 3)) || (r = dalloc(&c->ov_cell, ov)) ||
      (r = dalloc(&c->ov_rank, ov)) || (r = dalloc(&c->ov_y, ov)) ||
      (r = dalloc(&c->ov_z, ov)) || (r = dalloc(&c->ov_w, ov * 4)) ||
      (r = dalloc(&c->ov_key, ov)) || (r = dalloc(&c->d_counters, 2)) ||
      (r = dalloc(&c->cnt, nc)) || (r = dalloc(&c->cnt_main, nc)) ||
      (r = dalloc(&c->off, nc)) || (r = dalloc(&c->off_main, nc)) ||
      (r = dalloc(&c->blk, c->nblk)) || (r = dalloc(&c->blk_main, c->nblk)) ||
      (r = dalloc(&c->list, np + ov)) || (r = dalloc(&c->d_resid, (long)kResidPerWake * nw)) ||
      (r = dalloc(&c->d_slice, 1)) ||
      (r = dalloc(&c->d_res_hist, (size_t)p->nslices * p->nx_split * nw)))
    return fail(r);
  if (c->tiled) {
    const long nt = (long)c->ntj * c->ntk;
    if ((r = dalloc(&c->tile_start[0], (nt + 1) * nw)) ||
        (r = dalloc(&c->tile_start[1], (nt + 1) * nw)) ||
        (r = dalloc(&c->win, nt * kMaxSplit * 441 * 4 * nw)) ||
        (r = dalloc(&c->tcount, nt * kMaxSplit * 9 * nw)) ||
        (r = dalloc(&c->goff, nt * kMaxSplit * 9 * nw)))
      return fail(r);
  }
  c->hmask = history_mask;
  for (int f = 0; f < QSS_NFIELD; ++f)
    if (history_mask & (1u << f))
      if ((r = dalloc(&c->hist[f], (size_t)ncw * p->nslices)) != QSS_OK) return fail(r);
  // the zero fills of the allocations ran on the null stream, which does not
  // order with the context's non-blocking stream: wait for them
  if (hipDeviceSynchronize() != hipSuccess) {
    qss_set_error("allocation fill failed");
    return fail(QSS_E_HIP);
  }
  for (int i = 0; c->npc && i < 4; ++i) {
    const long n = i < 2 ? c->ny : c->nz;
    const double *src = c->d_T[i];
    if (launch_split_bf16(&src, &c->Ts[i], 1, n * n, c->npc, c->stream) != hipSuccess) {
      qss_set_error("table split failed");
      return fail(QSS_E_HIP);
    }
  }
  if ((r = do_reset(c)) != QSS_OK) return fail(r);
  if (hipStreamSynchronize(c->stream) != hipSuccess) {
    qss_set_error("seeding failed");
    return fail(QSS_E_HIP);
  }
  *out = c;
  return QSS_OK;
}

int qss_destroy(qss_ctx *c) {
  if (!c) return QSS_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  graph_free(c);
  dfree(c->d_slice);
  adj_free(c);
  driver_free(c->drv);
  dfree(c->drv.acc_rho); dfree(c->drv.acc_jx); dfree(c->drv.box); dfree(c->drv.status);
  for (auto &e : c->drv.ev)
    if (e) (void)hipEventDestroy(e);
  for (auto &p : c->d_T) dfree(p);
  for (int f = 0; f < QSS_NFIELD; ++f) { dfree(c->C[f]); dfree(c->P[f]); dfree(c->hist[f]); }
  for (auto &p : c->S) dfree(p);
  for (auto &p : c->tmp) dfree(p);
  for (auto &p : c->tmp2) dfree(p);
  dfree(c->d_G[0]); dfree(c->d_G[1]); dfree(c->d_tw[0]); dfree(c->d_tw[1]);
  dfree(c->d_Gf[0]); dfree(c->d_Gf[1]); dfree(c->d_twf[0]); dfree(c->d_twf[1]);
  dfree(c->d_Gfd[0]); dfree(c->d_Gfd[1]);
  dfree(c->F); dfree(c->Ff); dfree(c->d_bins); dfree(c->d_wbunch);
  for (auto &p : c->Ts) dfree(p);
  for (auto &p : c->xs) dfree(p);
  for (auto &p : c->ts) dfree(p);
  for (int b = 0; b < 2; ++b) {
    dfree(c->py_[b]); dfree(c->pz_[b]); dfree(c->ppx[b]); dfree(c->ppy[b]);
    dfree(c->ppz[b]); dfree(c->pcell[b]); dfree(c->pid[b]);
  }
  dfree(c->ev_cell); dfree(c->ev_rank); dfree(c->ev_sub); dfree(c->ev_y); dfree(c->ev_z);
  dfree(c->ev_w); dfree(c->ev_p); dfree(c->ov_cell); dfree(c->ov_rank); dfree(c->ov_y);
  dfree(c->ov_z); dfree(c->ov_w); dfree(c->ov_key); dfree(c->d_counters);
  dfree(c->cnt); dfree(c->cnt_main); dfree(c->off); dfree(c->off_main);
  dfree(c->blk); dfree(c->blk_main); dfree(c->list); dfree(c->d_resid);
  dfree(c->d_res_hist);
  dfree(c->tile_start[0]); dfree(c->tile_start[1]); dfree(c->win); dfree(c->tcount);
  dfree(c->goff);
  for (auto &pool : c->ev_pool)
    for (auto &pr : pool) { (void)hipEventDestroy(pr.first); (void)hipEventDestroy(pr.second); }
  if (c->t0) (void)hipEventDestroy(c->t0);
  if (c->t1) (void)hipEventDestroy(c->t1);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return QSS_OK;
}

int qss_info(const qss_ctx *c, int64_t *n_particles, int64_t *n_cells) {
  if (!c) return QSS_E_ARG;
  if (n_particles) *n_particles = c->np;
  if (n_cells) *n_cells = c->ncell;
  return QSS_OK;
}

void *qss_stream(qss_ctx *c) { return c ? (void *)c->stream : nullptr; }

int qss_select_wake(qss_ctx *c, int wake) {
  if (!c || wake < 0 || wake >= c->nw) {
    qss_set_error("wake index out of range");
    return QSS_E_ARG;
  }
  c->sel = wake;
  return QSS_OK;
}

int qss_reset(qss_ctx *c) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  CHECK(do_reset(c));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

// inside a sweep whose history holds neither rho nor psi, the final passes
// leave the layers' rho and psi planes out (only the last layer's are made)
struct SkipRhoPsi {
  qss_ctx *c;
  explicit SkipRhoPsi(qss_ctx *c_) : c(c_) {
    const uint32_t rho_psi = (1u << QSS_RHO) | (1u << QSS_PSI);
    c->skip_rho_psi = (c->hmask & rho_psi) == 0;
  }
  ~SkipRhoPsi() { c->skip_rho_psi = false; }   // also on an early error return
};

// the last layer's rho and psi planes (P after the swap) from its final
// spectra, after a sweep that skipped them
int finish_rho_psi(qss_ctx *c) {
  const int kz[2] = {QSS_KIND_C, QSS_KIND_C}, ky[2] = {QSS_KIND_C, QSS_KIND_C};
  const double *in[2] = {c->S[SP_RHO], c->S[SP_PSI]};
  double *out[2] = {c->P[QSS_RHO], c->P[QSS_PSI]};
  c->skip_rho_psi = false;
  return transform_batch(c, 2, kz, ky, in, out, 1.0 / (4.0 * c->ny * c->nz), nullptr, nullptr,
                         nullptr);
}

int qss_sweep_range(qss_ctx *c, int first, int count, double *residuals) {
  if (!c || first < 0 || count < 0 || first + count > c->p.nslices) {
    qss_set_error("bad sweep range");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  for (auto &u : c->ev_used) u = 0;
  c->gemm_flops = 0;
  if (!c->t0) { HIPCHECK(hipEventCreate(&c->t0)); HIPCHECK(hipEventCreate(&c->t1)); }
  HIPCHECK(hipEventRecord(c->t0, c->stream));
  const int S = c->p.nx_split;
  SkipRhoPsi skip(c);
  for (int s = first; s < first + count;) {
    if (c->use_graph && !c->timing && S == 1 && s >= 1 && first + count - s >= 2) {
      int done = 0;
      CHECK(sweep_graph_pair(c, s, &done));
      s += done;
      continue;
    }
    for (int is = 0; is < S; ++is)
      CHECK(do_layer_step(c, s, is, c->d_res_hist + (size_t)s * S + is));
    CHECK(store_history(c, s));
    ++s;
  }
  if (c->skip_rho_psi && count > 0) CHECK(finish_rho_psi(c));
  HIPCHECK(hipEventRecord(c->t1, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  // wake-major: wake w's residuals at [w*count*S, (w+1)*count*S)
  std::vector<double> res((size_t)count * S * c->nw);
  const size_t row = (size_t)count * S * sizeof(double);
  if (count > 0)
    HIPCHECK(hipMemcpy2D(res.data(), row, c->d_res_hist + (size_t)first * S,
                         (size_t)c->p.nslices * S * sizeof(double), row, c->nw,
                         hipMemcpyDeviceToHost));
  if (residuals) std::copy(res.begin(), res.end(), residuals);
  CHECK(check_status(c));
  for (double v : res)
    if (!std::isfinite(v)) { qss_set_error("non-finite residual"); return QSS_E_NONFINITE; }
  return QSS_OK;
}

int qss_sweep(qss_ctx *c, double *residuals) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  CHECK(do_reset(c));
  return qss_sweep_range(c, 0, c->p.nslices, residuals);
}

int qss_set_bins(qss_ctx *c, const double *bins, int nbins) {
  if (!c || !bins || nbins <= 0 || nbins != c->p.nbins) {
    qss_set_error("bins: the count must equal the nbins given at create");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->bins.assign(bins, bins + nbins);
  HIPCHECK(hipMemcpy(c->d_bins, bins, (size_t)nbins * sizeof(double), hipMemcpyHostToDevice));
  return QSS_OK;
}

int qss_copy_history(qss_ctx *c, int field, int first, int count, double *dst_dev) {
  if (!c || field < 0 || field >= QSS_NFIELD || !c->hist[field] || first < 0 || count < 0 ||
      first + count > c->p.nslices || (!dst_dev && count > 0)) {
    qss_set_error("copy_history: bad field/range or field not stored");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpyAsync(
      dst_dev, c->hist[field] + ((size_t)c->sel * c->p.nslices + first) * c->ncell,
      (size_t)count * c->ncell * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_device_memory(int device, int64_t *used, int64_t *total) {
  int cur = 0;
  HIPCHECK(hipGetDevice(&cur));
  HIPCHECK(hipSetDevice(device));
  size_t fr = 0, tot = 0;
  const hipError_t e = hipMemGetInfo(&fr, &tot);
  (void)hipSetDevice(cur);
  HIPCHECK(e);
  if (used) *used = (int64_t)(tot - fr);
  if (total) *total = (int64_t)tot;
  return QSS_OK;
}

int qss_sweep_checkpointed(qss_ctx *c, int nrun, int every, double *residuals) {
  if (!c || nrun < 1 || nrun > c->p.nslices || every < 0) {
    qss_set_error("checkpointed sweep: bad layer count or interval");
    return QSS_E_ARG;
  }
  CHECK(adj_check(c));
  HIPCHECK(hipSetDevice(c->device));
  // every = 0: one segment of all nrun layers, its tape recorded here
  const bool taped = every == 0;
  CHECK(adj_alloc(c, nrun, taped ? nrun : every));
  c->adj.taped = false;
  for (auto &u : c->ev_used) u = 0;
  c->gemm_flops = 0;
  if (!c->t0) { HIPCHECK(hipEventCreate(&c->t0)); HIPCHECK(hipEventCreate(&c->t1)); }
  HIPCHECK(hipEventRecord(c->t0, c->stream));
  CHECK(do_reset(c));
  AdjState &a = c->adj;
  SkipRhoPsi skip(c);   // as the plain sweep, so that both take the same steps
  if (taped) {
    CHECK(checkpoint(c, a.chk, 0));
    for (int s = 0; s < nrun; ++s) {
      CHECK(do_layer_step(c, s, 0, c->d_res_hist + s, a.tape + (size_t)s * a.tape_stride));
      CHECK(store_history(c, s));
    }
    a.taped = true;
    every = nrun;   // skips the loop below
  }
  for (int s = taped ? nrun : 0; s < nrun;) {
    if (s % every == 0) CHECK(checkpoint(c, a.chk + (size_t)(s / every) * a.chk_stride, 0));
    // pairs of layers by graph launches up to the next checkpoint
    const int lim = std::min(nrun, (s / every + 1) * every);
    if (c->use_graph && !c->timing && s >= 1 && lim - s >= 2) {
      int done = 0;
      CHECK(sweep_graph_pair(c, s, &done));
      s += done;
      continue;
    }
    CHECK(do_layer_step(c, s, 0, c->d_res_hist + s));
    CHECK(store_history(c, s));
    ++s;
  }
  if (c->skip_rho_psi) CHECK(finish_rho_psi(c));
  HIPCHECK(hipEventRecord(c->t1, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (residuals)
    HIPCHECK(hipMemcpy(residuals, c->d_res_hist, (size_t)nrun * sizeof(double),
                       hipMemcpyDeviceToHost));
  return check_status(c);
}

int qss_sweep_vjp(qss_ctx *c, const double *ex_bar, int on_host, double *grad_bins) {
  if (!c || !grad_bins) return QSS_E_ARG;
  AdjState &a = c->adj;
  if (!a.chk || a.nrun < 1) {
    qss_set_error("sweep_vjp needs a preceding qss_sweep_checkpointed");
    return QSS_E_STATE;
  }
  const double *ex_bar_dev = ex_bar;
  if (ex_bar && on_host) {
    HIPCHECK(hipMemcpy(a.exbar, ex_bar, (size_t)a.nrun * c->ncell * sizeof(double),
                       hipMemcpyHostToDevice));
    ex_bar_dev = a.exbar;
  }
  if (c->p.nbins < 1) {
    qss_set_error("sweep_vjp: the driver has no piecewise-constant profile (nbins = 0)");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  for (auto &u : c->ev_used) u = 0;
  c->gemm_flops = 0;
  if (!c->t0) { HIPCHECK(hipEventCreate(&c->t0)); HIPCHECK(hipEventCreate(&c->t1)); }
  HIPCHECK(hipEventRecord(c->t0, c->stream));
  const long nc = c->ncell;
  for (int s = 0; s < SP_N; ++s) HIPCHECK(launch_fill_f64(a.A[s], nc, 0.0, c->stream));
  HIPCHECK(launch_fill_f64(a.pbarEB, nc * kPushComp, 0.0, c->stream));
  HIPCHECK(launch_fill_f64(a.pbar[0], c->np_cap * 5, 0.0, c->stream));
  HIPCHECK(launch_fill_f64(a.pbar[1], c->np_cap * 5, 0.0, c->stream));
  HIPCHECK(launch_fill_f64(a.qbeam, (long)c->p.nslices * 2 * kBeamAdjBlocks, 0.0, c->stream));
  int cur_adj = 0;
  SkipRhoPsi skip(c);   // the recomputation takes the checkpointed forward's steps
  if (a.taped)          // the forward recorded every layer: reverse only
    for (int s = a.nrun - 1; s >= 0; --s)
      CHECK(adj_layer_step(c, s, a.tape + (size_t)s * a.tape_stride, ex_bar_dev, cur_adj));
  for (int seg = a.taped ? -1 : a.nchk - 1; seg >= 0; --seg) {
    const int s0 = seg * a.K, s1 = std::min(a.nrun, s0 + a.K);
    // recompute the segment from its checkpoint, recording the tape
    CHECK(checkpoint(c, a.chk + (size_t)seg * a.chk_stride, 1));
    // full segments after the first layer's by one graph launch (even K:
    // every such segment starts in the same pointer state)
    if (c->use_graph && !c->timing && s0 > 0 && s1 - s0 == a.K && a.K % 2 == 0) {
      bool done = false;
      CHECK(adj_segment_graph(c, s0, ex_bar_dev, cur_adj, &done));
      if (done) continue;
    }
    for (int s = s0; s < s1; ++s)
      CHECK(do_layer_step(c, s, 0, c->d_res_hist + s, a.tape + (size_t)(s - s0) * a.tape_stride));
    for (int s = s1 - 1; s >= s0; --s)
      CHECK(adj_layer_step(c, s, a.tape + (size_t)(s - s0) * a.tape_stride, ex_bar_dev, cur_adj));
  }
  HIPCHECK(hipEventRecord(c->t1, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  CHECK(check_status(c));
  std::vector<double> qp((size_t)a.nrun * 2 * kBeamAdjBlocks), q((size_t)a.nrun * 2, 0.0);
  HIPCHECK(hipMemcpy(qp.data(), a.qbeam, qp.size() * sizeof(double), hipMemcpyDeviceToHost));
  for (size_t i = 0; i < q.size(); ++i)   // work-group partials in a fixed order
    for (int b = 0; b < kBeamAdjBlocks; ++b) q[i] += qp[i * kBeamAdjBlocks + b];
  // layer s receives the profile at xi = (s+1) hx and (s+2) hx (beam.hip)
  const int nb = c->p.nbins, L = c->p.nslices;
  for (int b = 0; b < nb; ++b) grad_bins[b] = 0.0;
  for (int s = 0; s < a.nrun; ++s)
    for (int side = 0; side < 2; ++side) {
      const int ci = (L - 1 - s) - side;
      if (ci < 0 || ci > L) continue;
      const double xi = (double)(L - ci) * c->p.hx;
      const double t = (xi - c->p.bin_xi0) / (c->p.bin_xi1 - c->p.bin_xi0) * nb;
      if (!(t >= 0.0) || t >= nb) continue;
      grad_bins[(int)t] += q[(size_t)s * 2 + side];
    }
  return QSS_OK;
}

int qss_layer_step(qss_ctx *c, int slice, int isplit, double *residual) {
  if (!c || slice < 0 || slice >= c->p.nslices || isplit < 0 || isplit >= c->p.nx_split) {
    qss_set_error("bad slice / split");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  double *dst = c->d_res_hist + (size_t)slice * c->p.nx_split + isplit;
  CHECK(do_layer_step(c, slice, isplit, dst));
  double r = 0;
  HIPCHECK(hipMemcpyAsync(&r, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (residual) *residual = r;
  CHECK(check_status(c));
  if (!std::isfinite(r)) { qss_set_error("non-finite residual"); return QSS_E_NONFINITE; }
  return QSS_OK;
}

int qss_beam_layer(qss_ctx *c, int slice, int isplit) {
  if (!c || slice < 0 || slice >= c->p.nslices || isplit < 0 || isplit >= c->p.nx_split)
    return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  CHECK(do_beam(c, slice, isplit));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_guess(qss_ctx *c, int first_layer) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  CHECK(do_guess(c, first_layer));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->push_dirty = false;
  return QSS_OK;
}

int qss_push_deposit(qss_ctx *c, int full_step) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  if (c->push_dirty) {
    CHECK(refresh_push_field(c));
    c->push_dirty = false;
  }
  CHECK(do_push_deposit(c, full_step));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return check_status(c);
}

int qss_iterate(qss_ctx *c, int final_pass, double *residual) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  double *dst = c->d_res_hist;
  CHECK(do_iterate(c, final_pass, dst));
  double r = 0;
  HIPCHECK(hipMemcpyAsync(&r, dst, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (residual) *residual = r;
  c->push_dirty = false;
  return QSS_OK;
}

int qss_copy_c_to_p(qss_ctx *c) {
  if (!c) return QSS_E_ARG;
  forget_splits(c);
  swap_c_p(c);
  c->push_dirty = true;
  return QSS_OK;
}

int qss_transform2d(qss_ctx *c, int kind_z, int kind_y, const double *in_dev,
                    double *out_dev, double scale) {
  return qss_transform2d_batched(c, 1, &kind_z, &kind_y, &in_dev, &out_dev, scale);
}

int qss_transform2d_batched(qss_ctx *c, int n, const int *kind_z, const int *kind_y,
                            const double *const *in_dev, double *const *out_dev,
                            double scale) {
  if (!c || !kind_z || !kind_y || !in_dev || !out_dev || n < 1 || n > kGemmMaxBatch)
    return QSS_E_ARG;
  for (int b = 0; b < n; ++b)
    if ((kind_z[b] != 0 && kind_z[b] != 1) || (kind_y[b] != 0 && kind_y[b] != 1) ||
        !in_dev[b] || !out_dev[b] || ((uintptr_t)in_dev[b] & 15) ||
        ((uintptr_t)out_dev[b] & 15)) {
      qss_set_error("transform planes must be non-null and 16-byte aligned");
      return QSS_E_ARG;
    }
  HIPCHECK(hipSetDevice(c->device));
  CHECK(transform_batch(c, n, kind_z, kind_y, in_dev, out_dev, scale, nullptr, nullptr,
                        nullptr, 0.0, true));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_get_layer_field(qss_ctx *c, int which, int field, double *host) {
  if (!c || !host || field < 0 || field >= QSS_NFIELD || (which != 0 && which != 1))
    return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  double *src = (which ? c->P[field] : c->C[field]) + (size_t)c->sel * c->ncell;
  HIPCHECK(hipMemcpyAsync(host, src, c->ncell * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_set_layer_field(qss_ctx *c, int which, int field, const double *host) {
  if (!c || !host || field < 0 || field >= QSS_NFIELD || (which != 0 && which != 1))
    return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  forget_splits(c);
  double *dst = (which ? c->P[field] : c->C[field]) + (size_t)c->sel * c->ncell;
  HIPCHECK(hipMemcpyAsync(dst, host, c->ncell * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  if (field <= QSS_BZ) c->push_dirty = true;
  return QSS_OK;
}

int qss_get_history(qss_ctx *c, int field, int slice, double *host) {
  if (!c || !host || field < 0 || field >= QSS_NFIELD || !c->hist[field] || slice < 0 ||
      slice >= c->p.nslices)
    return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpy(host,
                     c->hist[field] + ((size_t)c->sel * c->p.nslices + slice) * c->ncell,
                     c->ncell * sizeof(double), hipMemcpyDeviceToHost));
  return QSS_OK;
}

int qss_layer_ptr(qss_ctx *c, int which, int field, double **dev) {
  if (!c || !dev || field < 0 || field >= QSS_NFIELD || (which != 0 && which != 1))
    return QSS_E_ARG;
  *dev = (which ? c->P[field] : c->C[field]) + (size_t)c->sel * c->ncell;
  forget_splits(c);   // the caller may write the plane
  return QSS_OK;
}

int qss_history_ptr(qss_ctx *c, int field, double **dev) {
  if (!c || !dev || field < 0 || field >= QSS_NFIELD || !c->hist[field]) return QSS_E_ARG;
  *dev = c->hist[field] + (size_t)c->sel * c->p.nslices * c->ncell;
  return QSS_OK;
}

int qss_get_spectral(qss_ctx *c, int id, double *host) {
  if (!c || !host || id < 0 || id >= QSS_NSPECTRAL) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpyAsync(host, c->S[id] + (size_t)c->sel * c->ncell,
                          c->ncell * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_set_spectral(qss_ctx *c, int id, const double *host) {
  if (!c || !host || id < 0 || id >= QSS_NSPECTRAL) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  forget_splits(c);
  HIPCHECK(hipMemcpyAsync(c->S[id] + (size_t)c->sel * c->ncell, host,
                          c->ncell * sizeof(double), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

int qss_get_particles(qss_ctx *c, int32_t *cell_j, int32_t *cell_k, double *fy, double *fz,
                      double *px, double *py, double *pz, int64_t *id) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  const int b = c->cur;
  const size_t n = (size_t)c->np, o = (size_t)c->sel * c->np_cap;
  auto get = [&](void *dst, const void *src, size_t bytes) -> int {
    if (!dst) return QSS_OK;
    HIPCHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    return QSS_OK;
  };
  std::vector<int32_t> cell(n);
  CHECK(get(cell.data(), c->pcell[b] + o, n * 4));
  CHECK(get(fy, c->py_[b] + o, n * 8));
  CHECK(get(fz, c->pz_[b] + o, n * 8));
  CHECK(get(px, c->ppx[b] + o, n * 8));
  CHECK(get(py, c->ppy[b] + o, n * 8));
  CHECK(get(pz, c->ppz[b] + o, n * 8));
  CHECK(get(id, c->pid[b] + o, n * 8));
  HIPCHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < n; ++i) {
    if (cell_j) cell_j[i] = cell[i] % c->ny;
    if (cell_k) cell_k[i] = cell[i] / c->ny;
  }
  return QSS_OK;
}

int qss_set_particles(qss_ctx *c, int64_t n, const int32_t *cell_j, const int32_t *cell_k,
                      const double *fy, const double *fz, const double *px, const double *py,
                      const double *pz, const int64_t *id) {
  if (!c || n < 0 || n > c->np_cap || !cell_j || !cell_k || !fy || !fz || !px || !py ||
      !pz || !id)
    return QSS_E_ARG;
  if (c->nw > 1 && n != c->np) {
    qss_set_error("a wake of a batch keeps its particle count");
    return QSS_E_ARG;
  }
  graph_free(c);   // launch sizes follow the particle count
  std::vector<int32_t> cell((size_t)n);
  for (int64_t i = 0; i < n; ++i) {
    if (cell_j[i] < 0 || cell_j[i] >= c->ny || cell_k[i] < 0 || cell_k[i] >= c->nz) {
      qss_set_error("particle cell out of range");
      return QSS_E_ARG;
    }
    cell[i] = cell_j[i] + c->ny * cell_k[i];
  }
  // the tiled path needs slot order grouped by tile (stable: the given order
  // is kept inside a tile)
  std::vector<int64_t> order((size_t)n);
  std::iota(order.begin(), order.end(), (int64_t)0);
  std::vector<int32_t> tstart;
  if (c->tiled) {
    auto tile_of = [&](int64_t i) {
      return (cell_k[i] / kTile) * c->ntj + cell_j[i] / kTile;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](int64_t a, int64_t b) { return tile_of(a) < tile_of(b); });
    const int nt = c->ntj * c->ntk;
    tstart.assign(nt + 1, 0);
    for (int64_t i = 0; i < n; ++i) ++tstart[tile_of(i) + 1];
    for (int q = 0; q < nt; ++q) tstart[q + 1] += tstart[q];
  }
  std::vector<int32_t> ocell((size_t)n);
  std::vector<double> ofy((size_t)n), ofz((size_t)n), opx((size_t)n), opy((size_t)n),
      opz((size_t)n);
  std::vector<int64_t> oid((size_t)n);
  for (int64_t q = 0; q < n; ++q) {
    const int64_t i = order[q];
    ocell[q] = cell[i]; ofy[q] = fy[i]; ofz[q] = fz[i];
    opx[q] = px[i]; opy[q] = py[i]; opz[q] = pz[i]; oid[q] = id[i];
  }
  HIPCHECK(hipSetDevice(c->device));
  const int b = c->cur;
  const size_t o = (size_t)c->sel * c->np_cap;
  HIPCHECK(hipMemcpy(c->pcell[b] + o, ocell.data(), n * 4, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->py_[b] + o, ofy.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->pz_[b] + o, ofz.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->ppx[b] + o, opx.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->ppy[b] + o, opy.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->ppz[b] + o, opz.data(), n * 8, hipMemcpyHostToDevice));
  HIPCHECK(hipMemcpy(c->pid[b] + o, oid.data(), n * 8, hipMemcpyHostToDevice));
  if (c->tiled)
    HIPCHECK(hipMemcpy(c->tile_start[b] + (size_t)c->sel * (c->ntj * c->ntk + 1),
                       tstart.data(), tstart.size() * 4, hipMemcpyHostToDevice));
  c->np = n;
  return QSS_OK;
}

int qss_set_transform_method(qss_ctx *c, int method) {
  if (!c || (method != QSS_TRANSFORM_DENSE && method != QSS_TRANSFORM_FACTORED)) {
    qss_set_error("unknown transform method");
    return QSS_E_ARG;
  }
  if (method == QSS_TRANSFORM_FACTORED && !c->fact_ok) {
    qss_set_error("the factored transform needs FP64 and ny, nz multiples of 8, at least " +
                  std::to_string(kFactMinN));
    return QSS_E_ARG;
  }
  graph_free(c);   // a captured graph holds the other method's launches
  c->fact_use = method == QSS_TRANSFORM_FACTORED;
  return QSS_OK;
}

int qss_transform_method(const qss_ctx *c) {
  if (!c) return QSS_E_ARG;
  return c->fact_use ? QSS_TRANSFORM_FACTORED : QSS_TRANSFORM_DENSE;
}

int qss_gemm_variants(int precision) {
  if (precision < QSS_PREC_FP64 || precision > QSS_PREC_MIXED_BF16X6) return QSS_E_ARG;
  return gemm_num_variants(precision);
}

int qss_set_gemm_variant(qss_ctx *c, int variant) {
  if (!c || variant < -1 || variant >= gemm_num_variants(c->p.precision)) return QSS_E_ARG;
  graph_free(c);
  c->gemm_force = variant;
  return QSS_OK;
}

int qss_gemm_choice(const qss_ctx *c, int *chosen) {
  if (!c || !chosen) return QSS_E_ARG;
  for (int p = 0; p < 2; ++p)
    for (int b = 0; b <= kGemmMaxBatch; ++b)
      chosen[p * (kGemmMaxBatch + 1) + b] = c->gemm_var[p][b];
  return QSS_OK;
}

int qss_fact_choice(const qss_ctx *c, int *chosen) {
  if (!c || !chosen) return QSS_E_ARG;
  constexpr int S = kGemmMaxBatch + 1;
  for (int p = 0; p < 2; ++p)
    for (int b = 0; b < S; ++b) {
      chosen[p * S + b] = c->fact_var[p][b];
      chosen[(2 + p) * S + b] = c->ffused_ok ? c->ffv[p][b] : -1;
    }
  return QSS_OK;
}

int qss_set_fact_choice(qss_ctx *c, const int *chosen) {
  if (!c || !chosen) return QSS_E_ARG;
  constexpr int S = kGemmMaxBatch + 1;
  const int nv = gemm_num_variants(QSS_PREC_FP64);
  for (int i = 0; i < 2 * S; ++i)
    if (chosen[i] < 0 || chosen[i] >= nv) {
      qss_set_error("factored gemm choice out of range");
      return QSS_E_ARG;
    }
  for (int p = 0; p < 2 && c->ffused_ok; ++p)
    for (int b = 0; b < S; ++b)
      if (!fact_fused_shape_ok(p, chosen[(2 + p) * S + b], c->ny, c->nz)) {
        qss_set_error("fused transform choice out of range for this grid");
        return QSS_E_ARG;
      }
  graph_free(c);
  for (int p = 0; p < 2; ++p)
    for (int b = 0; b < S; ++b) {
      c->fact_var[p][b] = chosen[p * S + b];
      if (c->ffused_ok) c->ffv[p][b] = chosen[(2 + p) * S + b];
    }
  return QSS_OK;
}

int qss_enable_graph(qss_ctx *c, int on) {
  if (!c) return QSS_E_ARG;
  c->use_graph = on != 0;
  if (!c->use_graph) graph_free(c);
  return QSS_OK;
}

int qss_set_gemm_choice(qss_ctx *c, const int *chosen) {
  if (!c || !chosen) return QSS_E_ARG;
  graph_free(c);
  const int nv = gemm_num_variants(c->p.precision);
  for (int i = 0; i < 2 * (kGemmMaxBatch + 1); ++i)
    if (chosen[i] < 0 || chosen[i] >= nv) {
      qss_set_error("gemm choice out of range");
      return QSS_E_ARG;
    }
  for (int p = 0; p < 2; ++p)
    for (int b = 0; b <= kGemmMaxBatch; ++b)
      c->gemm_var[p][b] = chosen[p * (kGemmMaxBatch + 1) + b];
  return QSS_OK;
}

int qss_autotune(qss_ctx *c, int reps, float *table_ms) {
  if (!c || reps < 1) return QSS_E_ARG;
  graph_free(c);
  forget_splits(c);   // the timing runs use the split buffers as scratch
  HIPCHECK(hipSetDevice(c->device));
  hipEvent_t e0, e1;
  HIPCHECK(hipEventCreate(&e0));
  HIPCHECK(hipEventCreate(&e1));
  int rc = QSS_OK;
  const int prec = c->p.precision;
  const int nv = gemm_num_variants(prec);
  for (int pass = 0; pass < 2 && rc == QSS_OK; ++pass)
    for (int b = 1; b <= kGemmMaxBatch && rc == QSS_OK; ++b) {
      GemmArgs g{};
      g.M = c->nz; g.N = c->ny; g.K = pass == 0 ? c->ny : c->nz;
      g.alpha = 1.0; g.nbatch = b; g.F = c->F; g.Ff = c->Ff; g.resid = c->d_resid;
      // time the launch shape the sweep uses: all wakes of the batch
      g.nw = c->nw;
      g.wA = pass == 0 ? c->ncell : 0;
      g.wB = pass == 0 ? 0 : c->ncell;
      g.wC = c->ncell;
      g.wP = pass == 0 ? 0 : c->ncell;
      g.wF = pass == 0 ? 0 : c->ncell * kPushComp;
      g.wR = pass == 0 ? 0 : kResidPerWake;
      g.tout = (c->npc && pass == 0) ? 1 : 0;
      // the z passes of 6 and 8 fields carry the corrector's epilogue in the
      // sweep (push field from the P plane, residual of E): time them with it.
      // The push field and residual slots are scratch between layer steps.
      const bool epi = pass == 1 && (b == 6 || b == 8);
      for (int q = 0; q < b; ++q) {
        // operands: any plane of the right size; outputs: the scratch planes
        g.b[q].A = pass == 0 ? c->S[q] : c->d_T[2];
        g.b[q].B = pass == 0 ? c->d_T[0] : c->S[q];
        g.b[q].C = c->tmp[q];
        g.b[q].As = pass == 0 ? c->xs[q] : c->Ts[2];
        g.b[q].Bts = pass == 0 ? c->Ts[0] : c->ts[q];
        g.b[q].Cts = c->ts[q];
        g.b[q].P = epi && q < 6 ? c->P[q] : nullptr;
        g.b[q].resid = epi && q < 3 ? q : -1;
        g.b[q].push = epi && q < 6 ? q : -1;
      }
      float best = 1e30f;
      int bestv = 0;
      for (int v = 0; v < nv; ++v) {
        if (launch_gemm(g, c->stream, prec, v) != hipSuccess) { rc = QSS_E_HIP; break; }
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) (void)launch_gemm(g, c->stream, prec, v);
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = QSS_E_HIP; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (table_ms) table_ms[(pass * kGemmMaxBatch + (b - 1)) * nv + v] = ms / reps;
        if (ms < best) { best = ms; bestv = v; }
      }
      c->gemm_var[pass][b] = bestv;
    }
  // the factorized transform's two GEMM shapes (no epilogue: it runs in the
  // post kernel), per pass and batch size
  for (int pass = 0; pass < 2 && rc == QSS_OK && c->fact_ok; ++pass)
    for (int b = 1; b <= kGemmMaxBatch && rc == QSS_OK; ++b) {
      GemmArgs g{};
      g.alpha = 1.0; g.nbatch = b; g.nw = c->nw;
      if (pass == 0) {
        g.M = 2 * c->nz; g.N = c->ny / 2; g.K = c->ny / 2;
        g.wA = c->ncell; g.wB = 0;
      } else {
        g.M = c->nz / 2; g.N = 2 * c->ny; g.K = c->nz / 2;
        g.wA = 0; g.wB = c->ncell;
      }
      g.wC = c->ncell;
      for (int q = 0; q < b; ++q) {
        g.b[q].A = pass == 0 ? c->S[q] : c->d_G[1];
        g.b[q].B = pass == 0 ? c->d_G[0] : c->S[q];
        g.b[q].C = c->tmp2[q];
        g.b[q].resid = -1;
        g.b[q].push = -1;
      }
      float best = 1e30f;
      int bestv = 0;
      const int nv64 = gemm_num_variants(QSS_PREC_FP64);
      for (int v = 0; v < nv64; ++v) {
        if (launch_gemm(g, c->stream, QSS_PREC_FP64, v) != hipSuccess) { rc = QSS_E_HIP; break; }
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) (void)launch_gemm(g, c->stream, QSS_PREC_FP64, v);
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = QSS_E_HIP; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; bestv = v; }
      }
      c->fact_var[pass][b] = bestv;
    }
  // the fused form's tile shapes, per pass and batch size; the z pass timed
  // with the corrector's epilogue where the sweep runs it (6 and 8 fields)
  for (int pass = 0; pass < 2 && rc == QSS_OK && c->ffused_ok; ++pass)
    for (int b = 1; b <= kGemmMaxBatch && rc == QSS_OK; ++b) {
      FusedArgs f{};
      f.ny = c->ny; f.nz = c->nz; f.nbatch = b; f.nw = c->nw;
      f.w_in = c->ncell; f.w_out = c->ncell; f.w_p = c->ncell;
      f.w_f = c->ncell * kPushComp; f.w_r = kResidPerWake;
      f.alpha = 1.0; f.F = c->F; f.Ff = c->Ff; f.rslots = c->d_resid;
      f.G = c->d_Gf[pass]; f.Gd = c->d_Gfd[pass]; f.tw = c->d_twf[pass];
      const bool epi = pass == 1 && (b == 6 || b == 8);
      for (int q = 0; q < b; ++q) {
        f.in[q] = pass == 0 ? c->S[q] : c->tmp2[q];
        f.out[q] = pass == 0 ? c->tmp2[q] : c->tmp[q];
        f.P[q] = epi && q < 6 ? c->P[q] : nullptr;
        f.kz[q] = f.ky[q] = QSS_KIND_C;
        f.resid_c[q] = epi && q < 3 ? q : -1;
        f.push[q] = epi && q < 6 ? q : -1;
      }
      float best = 1e30f;
      int bestv = c->ffv[pass][b];
      for (int v = 0; v < fact_fused_variants(pass); ++v) {
        if (!fact_fused_shape_ok(pass, v, c->ny, c->nz)) continue;
        if (launch_fact_fused(f, pass, v, c->stream) != hipSuccess) { rc = QSS_E_HIP; break; }
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < reps; ++r) (void)launch_fact_fused(f, pass, v, c->stream);
        (void)hipEventRecord(e1, c->stream);
        if (hipEventSynchronize(e1) != hipSuccess) { rc = QSS_E_HIP; break; }
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best) { best = ms; bestv = v; }
      }
      c->ffv[pass][b] = bestv;
    }
  // work-groups per tile of the push: one predictor pass (state unchanged)
  // plus the window combine into scratch planes, per candidate
  if (rc == QSS_OK && c->tiled && c->np > 0) {
    double *const outs[4] = {c->tmp[0], c->tmp[1], c->tmp[2], c->tmp[3]};
    float best = 1e30f;
    int bests = c->nsplit;
    for (int sp = 1; sp <= kMaxSplit && rc == QSS_OK; sp *= 2) {
      c->nsplit = sp;
      // the form a layer step runs (no density deposit)
      if (launch_push_tile(c, 0, false) != hipSuccess ||
          launch_combine(c, false, outs) != hipSuccess) {
        rc = QSS_E_HIP;
        break;
      }
      (void)hipEventRecord(e0, c->stream);
      for (int r = 0; r < reps; ++r) {
        (void)launch_push_tile(c, 0, false);
        (void)launch_combine(c, false, outs);
      }
      (void)hipEventRecord(e1, c->stream);
      if (hipEventSynchronize(e1) != hipSuccess) { rc = QSS_E_HIP; break; }
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      if (ms < best) { best = ms; bests = sp; }
    }
    c->nsplit = bests;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != QSS_OK) qss_set_error("autotune launch failed");
  return rc;
}

int qss_set_particle_split(qss_ctx *c, int nsplit) {
  if (!c || (nsplit != 1 && nsplit != 2 && nsplit != 4)) {
    qss_set_error("particle split must be 1, 2 or 4");
    return QSS_E_ARG;
  }
  graph_free(c);
  c->nsplit = nsplit;
  return QSS_OK;
}

int qss_particle_split(const qss_ctx *c) { return c ? c->nsplit : QSS_E_ARG; }

int qss_enable_timing(qss_ctx *c, int on) {
  if (!c) return QSS_E_ARG;
  c->timing = on != 0;
  return QSS_OK;
}

int qss_timing(qss_ctx *c, double *ms_total, double *ms_kernel, int64_t *launches,
               double *gemm_flops) {
  if (!c) return QSS_E_ARG;
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  float ms = 0.f;
  if (c->t0 && c->t1) HIPCHECK(hipEventElapsedTime(&ms, c->t0, c->t1));
  if (ms_total) *ms_total = ms;
  for (int k = 0; k < kNClass; ++k) {
    double sum = 0;
    for (size_t i = 0; i < c->ev_used[k]; ++i) {
      float e = 0.f;
      HIPCHECK(hipEventElapsedTime(&e, c->ev_pool[k][i].first, c->ev_pool[k][i].second));
      sum += e;
    }
    if (ms_kernel) ms_kernel[k] = sum;
    if (launches) launches[k] = (int64_t)c->ev_used[k];
  }
  if (gemm_flops) *gemm_flops = c->gemm_flops;
  return QSS_OK;
}

// ---- layer hand-off between xi chunks ----------------------------------------
// The layer state a sweep continues from: a 256-byte header, then the
// particles (current buffer and tile starts), the P-layer planes and the
// spectral state in the layout of the adjoint's checkpoints.  The header
// names the grid the state belongs to; unpacking checks it, the size, and the
// particle cells and tile starts before anything reaches the context.
int qss_layer_state_bytes(const qss_ctx *c, size_t *bytes) {
  if (!c || !bytes) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer hand-off: one wake"); return QSS_E_ARG; }
  *bytes = kStateHeader + chk_bytes(c);
  return QSS_OK;
}

int qss_pack_layer(qss_ctx *c, void *dst) {
  if (!c || !dst) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer hand-off: one wake"); return QSS_E_ARG; }
  HIPCHECK(hipSetDevice(c->device));
  const StateHeader h = state_header(c);
  char *base = static_cast<char *>(dst);
  HIPCHECK(hipMemcpyAsync(base, &h, sizeof h, hipMemcpyDefault, c->stream));
  CHECK(checkpoint(c, base + kStateHeader, 0));
  HIPCHECK(hipStreamSynchronize(c->stream));   // h lives on this stack frame
  return QSS_OK;
}

int qss_unpack_layer(qss_ctx *c, const void *src, size_t bytes) {
  if (!c || !src) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer hand-off: one wake"); return QSS_E_ARG; }
  HIPCHECK(hipSetDevice(c->device));
  if (bytes != kStateHeader + chk_bytes(c)) {
    qss_set_error("layer hand-off: the buffer has " + std::to_string(bytes) +
                  " bytes, this context's state " +
                  std::to_string(kStateHeader + chk_bytes(c)));
    return QSS_E_ARG;
  }
  const char *base = static_cast<const char *>(src);
  StateHeader h{}, want = state_header(c);
  HIPCHECK(hipMemcpy(&h, base, sizeof h, hipMemcpyDefault));
  if (std::memcmp(&h, &want, sizeof h) != 0) {
    qss_set_error("layer hand-off: the buffer holds the state of another grid or format "
                  "(header mismatch)");
    return QSS_E_ARG;
  }
  // the cells and tile starts drive the push kernels' indexing: check them
  const PartView v = part_view(c, const_cast<char *>(base + kStateHeader));
  std::vector<int32_t> cell(c->np_cap);
  HIPCHECK(hipMemcpy(cell.data(), v.cell, cell.size() * sizeof(int32_t), hipMemcpyDefault));
  for (int32_t q : cell)
    if (q < 0 || q >= c->ncell) {
      qss_set_error("layer hand-off: a particle cell is outside the grid");
      return QSS_E_ARG;
    }
  if (c->tiled) {
    const long nt = (long)c->ntj * c->ntk;
    std::vector<int32_t> ts(nt + 1);
    HIPCHECK(hipMemcpy(ts.data(), v.tstart, ts.size() * sizeof(int32_t), hipMemcpyDefault));
    bool ok = ts[0] == 0 && ts[nt] == c->np_cap;
    for (long t = 0; ok && t < nt; ++t) {
      ok = ts[t] <= ts[t + 1] && ts[t + 1] <= c->np_cap;
      for (long q = ts[t]; ok && q < ts[t + 1]; ++q) {   // each particle in its own tile
        const int j = cell[q] % c->ny, k = cell[q] / c->ny;
        ok = (k / kTile) * c->ntj + j / kTile == t;
      }
    }
    if (!ok) {
      qss_set_error("layer hand-off: the tile starts do not describe the particles");
      return QSS_E_ARG;
    }
  }
  HIPCHECK(hipStreamSynchronize(c->stream));
  forget_splits(c);   // the planes change under any recorded split operands
  CHECK(checkpoint(c, const_cast<char *>(base + kStateHeader), 1));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->np = c->np_cap;
  c->push_dirty = true;
  return QSS_OK;
}

// ---- the reference's layer record (PackLayer / UnpackLayer, para.cpp:94-160, 167-280)
// Doubles: 12 planes of Ny*Nz (y fastest) in the order Bx, By, Bz, Ex, Ey, Ez,
// Jx, Jy, Jz, Rho, fftJxBeamHydro, fftRhoBeamHydro (para.cpp:109-120), then
// 12 attributes of Np particles, attribute-major: Px, Py, Pz, X, Y, Z, Q2m,
// Weight, i_X, i_Y, i_Z, isort (para.cpp:138-149; integers stored as doubles).
// The planes are the P layer (the last layer swept); the two beam planes are
// the beam spectra the next layer's Guess takes as its previous beam
// spectrum (CUDA_WRAP_getBeamFFT, paraCPUlayers.cu:56).  Np comes from the
// record's length as in ReceiveLayer (para.cpp:385).
int qss_layer_record_doubles(const qss_ctx *c, size_t *n) {
  if (!c || !n) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer record: one wake"); return QSS_E_ARG; }
  *n = (size_t)kRecPlanes * c->ncell + (size_t)kRecAttr * c->np;
  return QSS_OK;
}

int qss_pack_layer_record(qss_ctx *c, int iLayer, double *dst) {
  if (!c || !dst) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer record: one wake"); return QSS_E_ARG; }
  HIPCHECK(hipSetDevice(c->device));
  const long nc = c->ncell, np = c->np;
  for (int q = 0; q < 10; ++q)
    HIPCHECK(hipMemcpyAsync(dst + q * nc, c->P[kRecField[q]], nc * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(dst + 10 * nc, c->S[SP_JXB], nc * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(dst + 11 * nc, c->S[SP_RHOB], nc * sizeof(double),
                          hipMemcpyDeviceToHost, c->stream));
  const int b = c->cur;
  std::vector<double> a(5 * np);
  std::vector<int32_t> cell(np);
  double *const from[5] = {c->ppx[b], c->ppy[b], c->ppz[b], c->py_[b], c->pz_[b]};
  for (int q = 0; q < 5; ++q)
    HIPCHECK(hipMemcpyAsync(a.data() + q * np, from[q], np * sizeof(double),
                            hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(cell.data(), c->pcell[b], np * sizeof(int32_t), hipMemcpyDeviceToHost,
                          c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  double *r = dst + kRecPlanes * nc;
  for (long i = 0; i < np; ++i) {   // storage (slot) order
    r[i] = a[i];                            // Px
    r[i + np] = a[i + np];                  // Py
    r[i + 2 * np] = a[i + 2 * np];          // Pz
    r[i + 3 * np] = 0.0;                    // X: layer particles sit at x = 0
    r[i + 4 * np] = a[i + 3 * np];          // Y (in-cell fraction)
    r[i + 5 * np] = a[i + 4 * np];          // Z
    r[i + 6 * np] = -1.0;                   // Q2m of the electrons
    r[i + 7 * np] = c->pweight;             // Weight
    r[i + 8 * np] = (double)iLayer;         // i_X
    r[i + 9 * np] = (double)(cell[i] % c->ny);   // i_Y
    r[i + 10 * np] = (double)(cell[i] / c->ny);  // i_Z
    r[i + 11 * np] = 0.0;                   // isort: species 0
  }
  return QSS_OK;
}

int qss_unpack_layer_record(qss_ctx *c, const double *src, size_t n) {
  if (!c || !src) return QSS_E_ARG;
  if (c->nw != 1) { qss_set_error("layer record: one wake"); return QSS_E_ARG; }
  const long nc = c->ncell;
  if (n < (size_t)kRecPlanes * nc || (n - (size_t)kRecPlanes * nc) % kRecAttr != 0) {
    qss_set_error("layer record: the length is not 12 Ny*Nz + 12 Np doubles");
    return QSS_E_ARG;
  }
  const long np = (long)((n - (size_t)kRecPlanes * nc) / kRecAttr);
  if (np != c->np_cap) {
    qss_set_error("layer record: " + std::to_string(np) + " particles, the sweep carries " +
                  std::to_string(c->np_cap) + " (ppc per cell)");
    return QSS_E_ARG;
  }
  const double *r = src + kRecPlanes * nc;
  std::vector<int32_t> cell(np);
  for (long i = 0; i < np; ++i) {
    const double jy = r[i + 9 * np], kz = r[i + 10 * np];
    const double y = r[i + 4 * np], z = r[i + 5 * np];
    const bool ok = jy >= 0 && jy < c->ny && jy == std::floor(jy) && kz >= 0 && kz < c->nz &&
                    kz == std::floor(kz) && y >= 0 && y <= 1 && z >= 0 && z <= 1 &&
                    r[i + 6 * np] == -1.0 && r[i + 7 * np] == c->pweight &&
                    r[i + 11 * np] == 0.0 && std::isfinite(r[i]) &&
                    std::isfinite(r[i + np]) && std::isfinite(r[i + 2 * np]);
    if (!ok) {
      qss_set_error("layer record: particle " + std::to_string(i) +
                    " is not a plasma electron of this sweep (cell, in-cell position, Q2m, "
                    "Weight, isort or a momentum)");
      return QSS_E_ARG;
    }
    cell[i] = (int32_t)jy + c->ny * (int32_t)kz;
  }
  // storage order: tile-major (stable, so a record written by
  // qss_pack_layer_record comes back in the order it was written)
  std::vector<long> order(np);
  std::iota(order.begin(), order.end(), 0L);
  std::vector<int32_t> ts;
  if (c->tiled) {
    const long nt = (long)c->ntj * c->ntk;
    auto tile = [&](long i) {
      return (cell[i] / c->ny / kTile) * c->ntj + (cell[i] % c->ny) / kTile;
    };
    std::stable_sort(order.begin(), order.end(),
                     [&](long x, long y) { return tile(x) < tile(y); });
    ts.assign(nt + 1, 0);
    for (long i = 0; i < np; ++i) ++ts[tile(i) + 1];
    for (long t = 0; t < nt; ++t) ts[t + 1] += ts[t];
  }
  std::vector<double> a(5 * np);
  std::vector<int32_t> sc(np);
  std::vector<int64_t> id(np);
  const int src_attr[5] = {4, 5, 0, 1, 2};   // y, z, px, py, pz
  for (long q = 0; q < np; ++q) {
    const long i = order[q];
    for (int m = 0; m < 5; ++m) a[m * np + q] = r[i + src_attr[m] * np];
    sc[q] = cell[i];
    id[q] = i;   // the record's position labels the particle
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  forget_splits(c);
  const int b = c->cur;
  double *const to[5] = {c->py_[b], c->pz_[b], c->ppx[b], c->ppy[b], c->ppz[b]};
  for (int m = 0; m < 5; ++m)
    HIPCHECK(hipMemcpyAsync(to[m], a.data() + m * np, np * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->pcell[b], sc.data(), np * sizeof(int32_t), hipMemcpyHostToDevice,
                          c->stream));
  HIPCHECK(hipMemcpyAsync(c->pid[b], id.data(), np * sizeof(int64_t), hipMemcpyHostToDevice,
                          c->stream));
  if (c->tiled)
    HIPCHECK(hipMemcpyAsync(c->tile_start[b], ts.data(), ts.size() * sizeof(int32_t),
                            hipMemcpyHostToDevice, c->stream));
  for (int q = 0; q < 10; ++q)
    HIPCHECK(hipMemcpyAsync(c->P[kRecField[q]], src + q * nc, nc * sizeof(double),
                            hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->S[SP_JXB], src + 10 * nc, nc * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipMemcpyAsync(c->S[SP_RHOB], src + 11 * nc, nc * sizeof(double),
                          hipMemcpyHostToDevice, c->stream));
  for (int f : {QSS_RHOBEAM, QSS_JXBEAM, QSS_PSI})
    HIPCHECK(launch_fill_f64(c->P[f], nc, 0.0, c->stream));
  // the spectra the next Guess reads as its P-layer spectra: the reference
  // re-transforms the P layer at every layer (movefieldshydrolinlayersplit.cpp:
  // 569-623); this context keeps them, so a record supplies them this way
  const int kz[7] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_C, QSS_KIND_S,
                     QSS_KIND_C, QSS_KIND_C, QSS_KIND_S};
  const int ky[7] = {QSS_KIND_C, QSS_KIND_C, QSS_KIND_S, QSS_KIND_C,
                     QSS_KIND_C, QSS_KIND_S, QSS_KIND_C};
  const double *in[7] = {c->P[QSS_RHO], c->P[QSS_EX], c->P[QSS_EY], c->P[QSS_EZ],
                         c->P[QSS_JX], c->P[QSS_JY], c->P[QSS_JZ]};
  double *out[7] = {c->S[SP_RHO], c->S[SP_EX], c->S[SP_EY], c->S[SP_EZ],
                    c->S[SP_JX], c->S[SP_JY], c->S[SP_JZ]};
  CHECK(transform_batch(c, 7, kz, ky, in, out, 1.0, nullptr, nullptr, nullptr));
  HIPCHECK(hipStreamSynchronize(c->stream));
  c->np = c->np_cap;
  c->push_dirty = true;
  return QSS_OK;
}

int qss_set_history(qss_ctx *c, int field, int slice, const double *host) {
  if (!c || !host || field < 0 || field >= QSS_NFIELD || !c->hist[field] || slice < 0 ||
      slice >= c->p.nslices) {
    qss_set_error("set_history: no history for this field or slice out of range");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipMemcpyAsync(c->hist[field] + ((size_t)c->sel * c->p.nslices + slice) * c->ncell,
                          host, c->ncell * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  return QSS_OK;
}

// Mesh::SeedBeamParticles over cells i in [0, nslices) (mesh.cpp:43-64),
// nseed == 4 lattice (mesh.cpp:148-152), weight = profile/4 (mesh.cpp:165).
// Particles are stored in (i, k, j, bunch, iseed) order.  Cells outside a
// bunch's support are skipped by exact early outs: the profile is zero when
// arg = dx^2/Rx^2 + dy^2/R^2 + dz^2/R^2 > 16, and a rounded sum of
// non-negative terms is at least each of its terms.
int qss_driver_seed(qss_ctx *c, const qss_driver *d, int64_t *n_particles) {
  if (!c || !d) return QSS_E_ARG;
  if (c->nw != 1 || c->p.nbins > 0) {
    qss_set_error("driver: one wake with Gaussian bunches (no profile bins)");
    return QSS_E_ARG;
  }
  for (int f = QSS_EX; f <= QSS_BZ; ++f)
    if (!c->hist[f]) {
      qss_set_error("driver: needs the history of Ex, Ey, Ez, Bx, By, Bz");
      return QSS_E_STATE;
    }
  if (!(std::isfinite(d->ts) && d->wavelength > 0.0)) {
    qss_set_error("driver: ts must be finite and wavelength positive");
    return QSS_E_ARG;
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  const qss_params &p = c->p;
  const int L = p.nslices, ny = c->ny, nz = c->nz;
  std::vector<int32_t> cell, bunch;
  std::vector<double> frac, mom, w;
  for (int ci = 0; ci < L; ++ci) {
    const double xi = (double)(L - ci) * p.hx;
    bool any_x = false;
    for (int b = 0; b < p.nbunch; ++b) {
      const double Rx = std::sqrt(2.0) * p.bunch[b].sigma_xi, dx = xi - p.bunch[b].xi0;
      any_x = any_x || !(dx * dx / Rx / Rx > 16);
    }
    if (!any_x) continue;
    for (int ck = 0; ck < nz; ++ck) {
      const double zc = c->hz * ck - p.lz / 2.0;
      // candidate cells of this row: the union over bunches of the cells
      // within 4 R of the bunch centre plus 2 cells (outside it
      // dy^2/R^2 > 16, so arg > 16); every candidate gets the exact test
      int jlo = ny, jhi = -1;
      for (int b = 0; b < p.nbunch; ++b) {
        const double Ry = std::sqrt(2.0) * p.bunch[b].sigma_r;
        const double ctr = (p.bunch[b].y0 + p.ly / 2.0) / c->hy, half = 4.0 * Ry / c->hy + 2.0;
        jlo = std::min(jlo, (int)std::max(0.0, std::floor(ctr - half)));
        jhi = std::max(jhi, (int)std::min((double)(ny - 1), std::ceil(ctr + half)));
      }
      for (int cj = jlo; cj <= jhi; ++cj) {
        const double yc = c->hy * cj - p.ly / 2.0;
        for (int b = 0; b < p.nbunch; ++b) {
          const qss_bunch &B = p.bunch[b];
          const double Rx = std::sqrt(2.0) * B.sigma_xi, Ry = std::sqrt(2.0) * B.sigma_r;
          const double dx = xi - B.xi0, dy = yc - B.y0, dz = zc - B.z0;
          if (dx * dx / Rx / Rx > 16) continue;
          if (dz * dz / Ry / Ry > 16) continue;
          const double arg = dx * dx / Rx / Rx + dy * dy / Ry / Ry + dz * dz / Ry / Ry;
          if (arg > 16) continue;
          const double dn = B.nb * std::exp(-arg);
          if (!(dn > 0.0)) continue;
          for (int is = 0; is < 4; ++is) {
            double wt = 0.25 * dn;
            if (B.noise != 0.0) {
              const uint64_t cl = ((uint64_t)(ci + 1) * (uint64_t)nz + ck) * (uint64_t)ny + cj;
              wt *= 1.0 + B.noise * host_unif_pm1(p.seed ^ 0xB3A1ull, cl * 4 + is,
                                                   (uint64_t)b + 7);
            }
            cell.insert(cell.end(), {ci, cj, ck});
            frac.insert(frac.end(), {0.5, 1.0 / 3.0 * (1.0 + is % 2), 1.0 / 3.0 * (1.0 + is / 2)});
            mom.insert(mom.end(), {d->p0[b][0], d->p0[b][1], d->p0[b][2]});
            w.push_back(wt);
            bunch.push_back(b);
          }
        }
      }
    }
  }
  DriverState &ds = c->drv;
  ds.par = *d;
  CHECK(driver_upload(c, (long)w.size(), cell, frac, mom, w, bunch));
  const size_t nodes = (size_t)(L + 1) * c->ncell;
  if (!ds.acc_rho) CHECK(dalloc(&ds.acc_rho, nodes));
  else HIPCHECK(hipMemset(ds.acc_rho, 0, nodes * sizeof(unsigned long long)));
  if (!ds.acc_jx) CHECK(dalloc(&ds.acc_jx, nodes));
  else HIPCHECK(hipMemset(ds.acc_jx, 0, nodes * sizeof(unsigned long long)));
  if (!ds.box) CHECK(dalloc(&ds.box, 6));
  if (!ds.status) CHECK(dalloc(&ds.status, 1));
  for (auto &e : ds.ev)
    if (!e) HIPCHECK(hipEventCreate(&e));
  ds.box_lo[0] = 1; ds.box_hi[0] = 0;   // nothing deposited yet
  ds.on = true;
  graph_free(c);   // a captured graph holds the frozen lattice's kernel
  if (n_particles) *n_particles = (int64_t)ds.np;
  return QSS_OK;
}

int qss_driver_push(qss_ctx *c) {
  if (!c) return QSS_E_ARG;
  DriverState &d = c->drv;
  if (!d.on) { qss_set_error("driver: not seeded"); return QSS_E_STATE; }
  HIPCHECK(hipSetDevice(c->device));
  // Mesh::ClearBeamCurrents: only the nodes the previous push touched are
  // non-zero
  const bool clear = d.box_lo[0] <= d.box_hi[0];
  if (clear) HIPCHECK(hipEventRecord(d.ev[0], c->stream));
  HIPCHECK(launch_driver_clear(c));
  if (clear) HIPCHECK(hipEventRecord(d.ev[1], c->stream));
  const int32_t empty[6] = {INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN, INT32_MAX, INT32_MIN};
  HIPCHECK(hipMemcpyAsync(d.box, empty, sizeof(empty), hipMemcpyHostToDevice, c->stream));
  HIPCHECK(launch_fill_i32(d.status, 1, 0, c->stream));
  HIPCHECK(launch_driver_push(c, d.ev[2], d.ev[3]));
  int32_t box[6], st = 0;
  HIPCHECK(hipMemcpyAsync(box, d.box, sizeof(box), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipMemcpyAsync(&st, d.status, sizeof(st), hipMemcpyDeviceToHost, c->stream));
  HIPCHECK(hipStreamSynchronize(c->stream));
  d.ms_clear = 0.f;
  if (clear) HIPCHECK(hipEventElapsedTime(&d.ms_clear, d.ev[0], d.ev[1]));
  d.ms_push = 0.f;
  if (d.np) HIPCHECK(hipEventElapsedTime(&d.ms_push, d.ev[2], d.ev[3]));
  if (box[0] <= box[1]) {
    for (int q = 0; q < 3; ++q) {
      const int lim = q == 0 ? c->p.nslices : (q == 1 ? c->ny - 1 : c->nz - 1);
      d.box_lo[q] = std::max(0, box[2 * q]);
      d.box_hi[q] = std::min(lim, box[2 * q + 1]);
    }
  } else {
    d.box_lo[0] = 1; d.box_hi[0] = 0;
  }
  if (st & 2) { qss_set_error("driver: non-finite or unbounded move"); return QSS_E_PARTICLE; }
  if (st & 1) {
    qss_set_error("driver: in-cell fraction outside [0,1] (reference exit(-12))");
    return QSS_E_PARTICLE;
  }
  return QSS_OK;
}

int qss_time_step(qss_ctx *c, double *residuals) {
  CHECK(qss_driver_push(c));
  return qss_sweep(c, residuals);
}

int qss_driver_get(qss_ctx *c, int32_t *cell, double *frac, double *mom, double *weight,
                   int32_t *bunch) {
  if (!c) return QSS_E_ARG;
  DriverState &d = c->drv;
  if (!d.on) { qss_set_error("driver: not seeded"); return QSS_E_STATE; }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  const long n = d.np;
  if (n == 0) return QSS_OK;
  std::vector<int32_t> iv(n);
  std::vector<double> dv(n);
  int32_t *ia[3] = {d.ci, d.cj, d.ck};
  double *fa[3] = {d.x, d.y, d.z}, *ma[3] = {d.px, d.py, d.pz};
  for (int a = 0; a < 3; ++a) {
    if (cell) {
      HIPCHECK(hipMemcpy(iv.data(), ia[a], n * sizeof(int32_t), hipMemcpyDeviceToHost));
      for (long q = 0; q < n; ++q) cell[3 * q + a] = iv[q];
    }
    if (frac) {
      HIPCHECK(hipMemcpy(dv.data(), fa[a], n * sizeof(double), hipMemcpyDeviceToHost));
      for (long q = 0; q < n; ++q) frac[3 * q + a] = dv[q];
    }
    if (mom) {
      HIPCHECK(hipMemcpy(dv.data(), ma[a], n * sizeof(double), hipMemcpyDeviceToHost));
      for (long q = 0; q < n; ++q) mom[3 * q + a] = dv[q];
    }
  }
  if (weight) HIPCHECK(hipMemcpy(weight, d.w, n * sizeof(double), hipMemcpyDeviceToHost));
  if (bunch) HIPCHECK(hipMemcpy(bunch, d.bunch, n * sizeof(int32_t), hipMemcpyDeviceToHost));
  return QSS_OK;
}

int qss_driver_set(qss_ctx *c, int64_t n, const int32_t *cell, const double *frac,
                   const double *mom, const double *weight, const int32_t *bunch) {
  if (!c || n < 0 || (n > 0 && (!cell || !frac || !mom || !weight || !bunch)))
    return QSS_E_ARG;
  DriverState &d = c->drv;
  if (!d.on) { qss_set_error("driver: not seeded"); return QSS_E_STATE; }
  const int lim[3] = {c->p.nslices, c->ny, c->nz};
  for (int64_t q = 0; q < n; ++q) {
    if (bunch[q] < 0 || bunch[q] >= c->p.nbunch) {
      qss_set_error("driver_set: bunch index out of range");
      return QSS_E_ARG;
    }
    for (int a = 0; a < 3; ++a)
      if (cell[3 * q + a] < -1 || cell[3 * q + a] > lim[a]) {
        qss_set_error("driver_set: cell index outside [-1, N]");
        return QSS_E_ARG;
      }
  }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<int32_t> cv(cell, cell + 3 * n), bv(bunch, bunch + n);
  std::vector<double> fv(frac, frac + 3 * n), mv(mom, mom + 3 * n), wv(weight, weight + n);
  return driver_upload(c, (long)n, cv, fv, mv, wv, bv);
}

int qss_driver_density(qss_ctx *c, int layer, double *rho, double *jx) {
  if (!c) return QSS_E_ARG;
  DriverState &d = c->drv;
  if (!d.on) { qss_set_error("driver: not seeded"); return QSS_E_STATE; }
  if (layer < 0 || layer > c->p.nslices) { qss_set_error("driver: layer out of range"); return QSS_E_ARG; }
  HIPCHECK(hipSetDevice(c->device));
  HIPCHECK(hipStreamSynchronize(c->stream));
  std::vector<unsigned long long> v(c->ncell);
  const double inv = 1.0 / d.scale;
  unsigned long long *src[2] = {d.acc_rho, d.acc_jx};
  double *dst[2] = {rho, jx};
  for (int a = 0; a < 2; ++a) {
    if (!dst[a]) continue;
    HIPCHECK(hipMemcpy(v.data(), src[a] + (size_t)layer * c->ncell,
                       c->ncell * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    for (long n = 0; n < c->ncell; ++n) dst[a][n] = (double)(long long)v[n] * inv;
  }
  return QSS_OK;
}

int qss_driver_timing(qss_ctx *c, double *ms_clear, double *ms_push) {
  if (!c) return QSS_E_ARG;
  if (ms_clear) *ms_clear = c->drv.ms_clear;
  if (ms_push) *ms_push = c->drv.ms_push;
  return QSS_OK;
}

}  // extern "C"
