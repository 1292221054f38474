// qss_internal.h — device-side layout and kernel launchers of libqss.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/qss.h"

// spectral state ids (fft_of_* arrays, movefieldshydrolinlayersplit.cpp:267-313)
enum {
  SP_RHO = 0, SP_EX, SP_EY, SP_EZ, SP_BX, SP_BY, SP_BZ, SP_JX, SP_JY, SP_JZ,
  SP_RHOP, SP_JXP, SP_JYP, SP_JZP, SP_EXP, SP_EYP, SP_EZP,
  SP_RHOB, SP_JXB, SP_RHOBP, SP_JXBP, SP_PSI, SP_N
};

// push field: (P+C)/2 of Ex,Ey,Ez,Bx,By,Bz as 6 consecutive Ny*Nz planes
constexpr int kPushComp = 6;
// cells per tile edge of the tiled particle path
constexpr int kTile = 16;
// work-groups per tile of the tiled push (1, 2 or 4; chosen per context)
constexpr int kMaxSplit = 4;

// ---------------------------------------------------------------------------
// transform GEMM (tertia_amd/csrc/gemm.hip)
// C_b = alpha * A_b . B_b, all row-major, M x K . K x N, per-batch epilogue.
// Split-bf16 modes take their operands pre-split into npc bf16 pieces (bit
// patterns, piece p of an operand at +p*plane): A as [M][K], B transposed as
// Bt[N][K], both k-contiguous; with tout the result is stored the same way,
// transposed and split (Ct[N][M]), ready to be the next pass's Bt.
// ---------------------------------------------------------------------------
constexpr int kGemmMaxBatch = 8;
// residual maxima of the corrector: each wave adds its maxima (atomicMax on
// the bit patterns) into slot blockIdx % kResidSlots, 8 u64 per slot (dif[3],
// max[3]); spreading the atomics over slots keeps thousands of waves from
// serialising on one address.  The finaliser takes the max over the slots.
constexpr int kResidSlots = 128;
constexpr int kResidPerWake = kResidSlots * 8;
constexpr int kMaxPieces = 3;
struct GemmBatch {
  const double *A;
  const double *B;
  double *C;
  const double *P;   // P-layer plane for the push-field epilogue (or null)
  const uint16_t *As;   // split operands (bf16 modes)
  const uint16_t *Bts;
  uint16_t *Cts;        // tout: transposed split result
  int resid;         // -1, or component 0..2: residual max|new-old|, max|new|
  int push;          // -1, or component 0..5 of the push field
};
struct GemmArgs {
  int M, N, K;
  double alpha;
  double beta;                  // C = alpha A.B + beta C (fp64 store)
  int nbatch;
  int tout;                     // bf16 modes: store Cts instead of C
  double *F;                    // push field planes (kPushComp x M*N)
  float *Ff;                    // its fp32 form (mixed modes; F unused when set)
  unsigned long long *resid;    // kResidPerWake per wake: slots of dif[3], max[3] (bits)
  GemmBatch b[kGemmMaxBatch];
  // batched wakes (dense kernels): the launch covers nw copies of every
  // field; copy w uses operand + w * stride (stride 0 = shared operand)
  int nw;
  long wA, wB, wC, wP, wF, wR;
};
// prec = QSS_PREC_*: selects the MFMA type and the variant table
// ev0/ev1 (optional): events that receive the kernel's own start/end
// timestamps (hipExtLaunchKernel), no marker packets around the dispatch
hipError_t launch_gemm(const GemmArgs &a, hipStream_t s, int prec, int variant,
                       hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
int gemm_num_variants(int prec);
const char *gemm_variant_name(int prec, int v);
int gemm_pieces(int prec);   // bf16 pieces per operand element (0: not split)
// out[b] (npc planes of n bf16) = pieces of in[b] (n doubles), b < nb
hipError_t launch_split_bf16(const double *const *in, uint16_t *const *out, int nb, long n,
                             int npc, hipStream_t s, hipEvent_t ev0 = nullptr,
                             hipEvent_t ev1 = nullptr);

// ---------------------------------------------------------------------------
// factorized transform (tertia_amd/csrc/fact.hip): one radix-2 step of the
// DCT-IV's half-length DFT, two half-size GEMMs per 2-D transform
// ---------------------------------------------------------------------------
struct FactArgs {
  int ny, nz, nbatch, nw;
  const double *in[kGemmMaxBatch];
  double *out[kGemmMaxBatch];
  const double *P[kGemmMaxBatch];    // P-layer planes of the push-field epilogue
  long w_in, w_out, w_p;             // wake strides (0: shared)
  int kz[kGemmMaxBatch], ky[kGemmMaxBatch];
  int resid_c[kGemmMaxBatch], push[kGemmMaxBatch];   // -1: none
  double *t0[kGemmMaxBatch], *t1[kGemmMaxBatch];    // intermediates (per wake: ny*nz)
  const double2 *twy, *twz;          // twiddles of the y and z axes
  double alpha, beta;
  double *F;
  float *Ff;
  long w_f;
  unsigned long long *rslots;
  long w_r;
};
// stage 0: pre-twiddle along y; 1: butterflies along y + pre-twiddle along z;
// 2: butterflies along z + epilogue
hipError_t launch_fact(const FactArgs &a, int stage, hipStream_t s, hipEvent_t ev0 = nullptr,
                       hipEvent_t ev1 = nullptr);
void fact_twiddles(int n, std::vector<double> &tw);
void fact_matrix(int n, bool transpose, std::vector<double> &g);
// grids with both sizes multiples of 8 and at least this large use it in FP64
constexpr int kFactMinN = 256;

// the same with the twiddles inside the GEMMs (fact_fused.hip): pass 0 reads
// X and writes the y-transformed plane, pass 1 reads that and writes the
// result through the dense path's epilogue.  Sizes multiples of 128, <= 1024.
struct FusedArgs {
  int ny, nz, nbatch, nw;
  const double *in[kGemmMaxBatch];
  double *out[kGemmMaxBatch];
  const double *P[kGemmMaxBatch];
  long w_in, w_out, w_p;
  int kz[kGemmMaxBatch], ky[kGemmMaxBatch];
  int resid_c[kGemmMaxBatch], push[kGemmMaxBatch];
  const double *G;                   // the pass's permuted table (h x h), pre folded in
  const double *Gd;                  // the pass's table for the direct kernels (their K
                                     // order; k-major for the z pass)
  const double2 *tw;                 // the pass's butterfly twiddles (mid', post)
  double alpha, beta;
  double *F;
  float *Ff;
  long w_f;
  unsigned long long *rslots;
  long w_r;
};
int fact_fused_variants(int pass);
bool fact_fused_shape_ok(int pass, int variant, int ny, int nz);
hipError_t launch_fact_fused(const FusedArgs &a, int pass, int variant, hipStream_t s,
                             hipEvent_t ev0 = nullptr, hipEvent_t ev1 = nullptr);
void fact_fused_matrix(int n, int pass, std::vector<double> &g);
void fact_fused_twiddles(int n, std::vector<double> &tw);

// ---------------------------------------------------------------------------
// adjoint sweep (tertia_amd/csrc/adjoint.hip, qss_sweep_vjp)
// ---------------------------------------------------------------------------
struct AdjTileArgs {
  const double *y, *z, *px, *py, *pz;   // particles at the layer start (taped)
  const int32_t *cell;
  const int64_t *pid;
  const int32_t *tile_start;
  int ntj, ntk, ny, nz;
  double hx, hy, hz, weight;
  double bext[3];       // uniform external B
  const double *F;      // the pass's push field (6 planes, taped)
  const double *Jb;     // deposit adjoint planes of Jx, Jy, Jz
  int full;
  const double *pin;    // full step: adjoint of the new state by id (5 x np)
  double *pout;         // adjoint of the start state by id (5 x np); full: =, else +=
  long np;
  double *fwin;         // per tile and part: 17x17 nodes x 6 field adjoints
  int nsplit;           // work-groups per tile
  int32_t *counters;
};

// buffers of the checkpointed forward and of the reverse sweep
struct AdjState {
  int K = 0, nrun = 0, nchk = 0;
  bool taped = false;          // the forward recorded every layer's tape (every = 0)
  double *A[SP_N] = {};        // adjoint of the spectral state
  double *pbarEB = nullptr;    // 6 planes: adjoint of the final E,B of a layer
  double *fbar = nullptr;      // 6 planes: adjoint of a push field
  double *jbar = nullptr;      // 3 planes: adjoint of the deposited Jx,Jy,Jz
  double *bbar = nullptr;      // 2 planes: adjoint of the beam rho, Jx planes
  double *pbar[2] = {};        // particle-state adjoints by id (5 x np_cap)
  double *fwin = nullptr;      // ntiles x nsplit x 289 x 6
  double *qbeam = nullptr;     // per layer and x-cell side: d(loss)/d(profile)
  char *chk = nullptr;         // checkpoints (state at layer starts)
  size_t chk_stride = 0;
  char *tape = nullptr;        // per layer of a segment (all layers when taped):
                               // particles + push fields
  size_t tape_stride = 0;
  double *exbar = nullptr;     // loss adjoint of the Ex history given on the host
};

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
// evolving driver (qss_driver_*): beam particles in the 3-D cell grid and
// the fixed-point 3-D deposit the sweeps read
struct DriverState {
  bool on = false;
  qss_driver par{};
  long np = 0;
  int32_t *ci = nullptr, *cj = nullptr, *ck = nullptr, *bunch = nullptr;
  double *x = nullptr, *y = nullptr, *z = nullptr;
  double *px = nullptr, *py = nullptr, *pz = nullptr, *w = nullptr;
  unsigned long long *acc_rho = nullptr, *acc_jx = nullptr;   // (nslices+1) x ncell
  double scale = 1.0;            // 2^e: fixed-point resolution 2^-e
  int32_t *box = nullptr;        // device: min/max node i, j, k of the last push
  int32_t *status = nullptr;     // device status bits of the last push
  int box_lo[3] = {1, 0, 0}, box_hi[3] = {0, -1, -1};   // host copy (empty)
  hipEvent_t ev[4] = {};         // dispatch-stamped clear and push
  float ms_clear = 0.f, ms_push = 0.f;
};

struct qss_ctx {
  qss_params p{};
  std::vector<double> bins;
  std::vector<qss_bunch> wbunch;   // nw * QSS_MAX_BUNCH (per-wake bunches)
  qss_bunch *d_wbunch = nullptr;
  int nw = 1;                      // wakes in the batch; per-wake arrays are
                                   // stored wake-major with a fixed stride
  int sel = 0;                     // wake of the data-access calls
  int device = 0;
  hipStream_t stream = nullptr;
  int ny = 0, nz = 0;
  long ncell = 0;
  long np = 0;            // plasma particles in the layer
  long np_cap = 0;        // particle capacity (ny*nz*ppc)
  bool push_dirty = false;   // C/P E,B edited from the host: rebuild F
  double gemm_flops = 0;     // algorithmic flops of timed GEMM launches
  long ov_cap = 0;        // capacity of the sub-cycle overflow event buffer
  double hxs = 0, hy = 0, hz = 0, pweight = 0;

  double *d_T[4] = {};    // Cy, Sy, Cz, Sz transform matrices
  double *C[QSS_NFIELD] = {};
  double *P[QSS_NFIELD] = {};
  double *S[SP_N] = {};
  double *tmp[kGemmMaxBatch] = {};   // transform intermediates
  // factorized transform: eligible, in use, G matrices (y, z), twiddles,
  // second intermediate per batch slot, GEMM variants per pass and batch size
  bool fact_ok = false, fact_use = false;
  // fused form (fact_fused.hip): eligible, permuted tables (y, z), their
  // butterfly twiddles, variants
  bool ffused_ok = false;
  double *d_Gf[2] = {};
  double *d_Gfd[2] = {};             // the tables of the direct kernels
  double2 *d_twf[2] = {};
  int ffv[2][kGemmMaxBatch + 1] = {};
  double *d_G[2] = {};
  double2 *d_tw[2] = {};
  double *tmp2[kGemmMaxBatch] = {};
  int fact_var[2][kGemmMaxBatch + 1] = {};
  double *F = nullptr;               // push field
  float *Ff = nullptr;               // push field, fp32 (mixed modes)
  bool mixed = false;                // precision != QSS_PREC_FP64
  int npc = 0;                       // bf16 pieces per element (split modes)
  uint16_t *Ts[4] = {};              // split tables (npc planes each)
  uint16_t *xs[kGemmMaxBatch] = {};  // split transform inputs
  uint16_t *ts[kGemmMaxBatch] = {};  // split, transposed y-pass results
  // the plane whose split form xs[b] holds (null: none).  Kernels that
  // produce transform inputs write the split form too and record it here;
  // the next transform then skips its split pass.  Any other write to such a
  // plane (host edits, transform outputs) clears the record.
  const double *xs_src[kGemmMaxBatch] = {};
  double *d_bins = nullptr;

  // particles (SoA), ping-pong for the full-step reorder
  double *py_[2] = {}, *pz_[2] = {}, *ppx[2] = {}, *ppy[2] = {}, *ppz[2] = {};
  int32_t *pcell[2] = {};
  int64_t *pid[2] = {};
  int cur = 0;

  // deposit events: main (one per particle) + overflow (extra sub-steps)
  int32_t *ev_cell = nullptr, *ev_rank = nullptr, *ev_sub = nullptr;
  double *ev_y = nullptr, *ev_z = nullptr, *ev_w = nullptr, *ev_p = nullptr;
  int32_t *ov_cell = nullptr, *ov_rank = nullptr;
  double *ov_y = nullptr, *ov_z = nullptr, *ov_w = nullptr;
  uint64_t *ov_key = nullptr;
  int32_t *d_counters = nullptr;   // [0] overflow count, [1] status flags
  // counting sort by target cell
  int32_t *cnt = nullptr, *cnt_main = nullptr;
  int32_t *off = nullptr, *off_main = nullptr;
  int32_t *blk = nullptr, *blk_main = nullptr;
  int nblk = 0;
  int32_t *list = nullptr;

  // tiled particle path (default when ny, nz are multiples of kTile and the
  // sub-cycle restart quirk is kept): particles sorted by tile
  bool tiled = false;
  int ntj = 0, ntk = 0;
  int nsplit = 1;                // work-groups per tile
  bool deposit_rho = true;       // false inside a layer step (see launch_push_tile)
  bool skip_rho_psi = false;     // inside a sweep: final passes leave rho, psi
  int32_t *tile_start[2] = {};   // ping-pong with the particle arrays
  double *win = nullptr;         // ntiles x nsplit x 441 x 4 deposit windows
  int32_t *tcount = nullptr;     // ntiles x nsplit x 9 full-step direction counts
  int32_t *goff = nullptr;       // ntiles x 9 x nsplit group offsets

  unsigned long long *d_resid = nullptr;   // kResidPerWake per wake and pass
  double *d_res_hist = nullptr;            // residual of each layer step
  int res_count = 0;

  uint32_t hmask = 0;
  double *hist[QSS_NFIELD] = {};

  AdjState adj;
  DriverState drv;

  // hipGraph replay of two layer steps (the C/P and particle ping-pong
  // pointers return to the same state every two layers); the layer index
  // of a replay is read by the kernels from d_slice
  bool use_graph = false;
  bool capturing = false;
  int *d_slice = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  // one checkpoint segment of the reverse sweep (recomputation + reverse),
  // replayed for every full segment captured in the same pointer state
  hipGraph_t adj_graph = nullptr;
  hipGraphExec_t adj_gexec = nullptr;
  const void *adj_key[6] = {};
  int adj_key_i[4] = {-1, -1, -1, -1};
  const double *g_key_c0 = nullptr;   // pointer state the graph was captured in
  int g_key_cur = -1;
  const double *g_key_rb = nullptr;
  bool g_key_skip = false;

  // transform GEMM tile shape per (pass, batch): pass 0 = X.T_y, 1 = T_z.Y
  int gemm_var[2][kGemmMaxBatch + 1] = {};
  int gemm_force = -1;    // >= 0: use this variant for every launch

  // timing
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pool[QSS_NKCLASS];
  size_t ev_used[QSS_NKCLASS] = {};
  hipEvent_t t0 = nullptr, t1 = nullptr;
};

// kernel classes for timing (QSS_NKCLASS of them)
enum { KC_PUSH = 0, KC_GEMM = 1, KC_SPEC = 2, KC_BEAM = 3, KC_SORT = 4, KC_REDUCE = 5 };

// launchers (tertia_amd/csrc/*.hip); all enqueue on c->stream
hipError_t launch_fill_i32(int32_t *p, long n, int32_t v, hipStream_t s);
hipError_t launch_fill_f64(double *p, long n, double v, hipStream_t s);
// device-to-device copies of up to kCopyBufs buffers in one launch (unused
// slots: nbytes = 0); 16-byte aligned buffers, sizes multiples of 4 bytes
constexpr int kCopyBufs = 8;
struct CopyBufs {
  const void *src[kCopyBufs];
  void *dst[kCopyBufs];
  long nbytes[kCopyBufs];
};
hipError_t launch_copy_bufs(const CopyBufs &a, hipStream_t s);
// zero fills of up to kFillBufs double buffers in one launch (unused slots: n = 0)
constexpr int kFillBufs = 4;
struct ZeroBufs {
  double *p[kFillBufs];
  long n[kFillBufs];
};
hipError_t launch_zero_bufs(const ZeroBufs &a, hipStream_t s);
hipError_t launch_fill_u64(unsigned long long *p, long n, unsigned long long v, hipStream_t s);
hipError_t launch_seed_plasma(qss_ctx *c);
hipError_t launch_beam_plane(qss_ctx *c, int slice, int isplit);
hipError_t launch_driver_clear(qss_ctx *c);
hipError_t launch_driver_push(qss_ctx *c, hipEvent_t ev0, hipEvent_t ev1);
hipError_t launch_driver_plane(qss_ctx *c, int slice, int isplit);
hipError_t launch_guess_update(qss_ctx *c, int first_layer);
hipError_t launch_iterate_update(qss_ctx *c, int final_pass);
hipError_t launch_push(qss_ctx *c, int full_step);
hipError_t launch_sort_scan(qss_ctx *c, int full_step);
hipError_t launch_scatter(qss_ctx *c);
hipError_t launch_tile_reduce(qss_ctx *c, int full_step);
hipError_t launch_residual_final(qss_ctx *c, double *dst);
hipError_t launch_push_field(qss_ctx *c);
// rho = false: the deposit leaves the density out (a layer step, whose
// corrector takes rho from continuity)
hipError_t launch_push_tile(qss_ctx *c, int full, bool rho, hipEvent_t ev0 = nullptr,
                            hipEvent_t ev1 = nullptr);
// out (optional): 4 destination planes instead of the C-layer Jx, Jy, Jz, Rho
hipError_t launch_combine(qss_ctx *c, bool rho, double *const *out = nullptr);
hipError_t launch_tile_reorder(qss_ctx *c);
hipError_t launch_seed_tiled(qss_ctx *c);
hipError_t launch_axpy(double *y, const double *x, double a, long n, hipStream_t s,
                       const int *slice_dev = nullptr);
// graph replay helpers: history store at the device-side slice, slice += v
hipError_t launch_store_hist(qss_ctx *c, const double *src, double *hist);
hipError_t launch_add_i32(int *p, int v, hipStream_t s);
// adjoints of the spectral updates on the adjoint state A (flag: first layer
// for the Guess, final pass for the Iterate)
hipError_t launch_guess_adj(qss_ctx *c, double *const *A, int first_layer);
hipError_t launch_iterate_adj(qss_ctx *c, double *const *A, int final_pass);
hipError_t launch_push_tile_adj(qss_ctx *c, const AdjTileArgs &a);
// out = the six field-adjoint planes; pshare (pmode 1: =, 2: +=) their halves
hipError_t launch_fbar_combine(qss_ctx *c, const double *fwin, double *out, double *pshare,
                               int pmode);
// q: per layer, side and work-group partial sums (kBeamAdjBlocks per pair)
constexpr int kBeamAdjBlocks = 64;
hipError_t launch_beam_adj(qss_ctx *c, int slice, const double *rbar, const double *jbar,
                           double *q);

// host helpers
void qss_set_error(const std::string &m);

// early returns of the C-ABI functions: a HIP error (message recorded) or a
// non-OK status of a nested call
#define HIPCHECK(x)                                                              \
  do {                                                                           \
    hipError_t e_ = (x);                                                         \
    if (e_ != hipSuccess) {                                                      \
      qss_set_error(std::string(#x) + ": " + hipGetErrorString(e_));             \
      return QSS_E_HIP;                                                          \
    }                                                                            \
  } while (0)

#define CHECK(x)                                                                 \
  do {                                                                           \
    int r_ = (x);                                                                \
    if (r_ != QSS_OK) return r_;                                                 \
  } while (0)

