/*
 * qss.h — C ABI of the MI355X-native quasi-static xi-slice wake sweep.
 *
 * The reference (snytav/tertia) drives one layer step through C++ member
 * functions of `Mesh` that in turn call a C-style `CUDA_WRAP_*` API with
 * file-static device buffers.  This header replaces that surface with an
 * explicit context: every entry point takes a `qss_ctx*`, returns an int
 * status (QSS_OK or a negative QSS_E_* code, never exit()), and launches all
 * device work on the context's HIP stream.  Plain C types only.
 *
 * Coordinates follow the reference: x = xi (propagation, layers processed
 * from the front layer l_Mx-1 down to 0), y (index j, fastest) and z
 * (index k).  Every real-space and spectral plane is Nz x Ny doubles,
 * element (j,k) at offset j + Ny*k (reference `n = j + ny*k`,
 * movefieldshydrolinlayersplit.cpp:597).
 *
 * Thread-compatibility: one context per host thread; contexts are
 * independent.  Device buffers are owned by the context.
 */
#ifndef QSS_H
#define QSS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ----------------------------------------------------- */
#define QSS_OK              0
#define QSS_E_ARG          -1   /* bad argument / unsupported size          */
#define QSS_E_HIP          -2   /* HIP runtime error (message via last_error)*/
#define QSS_E_NOMEM        -3   /* device allocation failed                 */
#define QSS_E_STATE        -4   /* call out of order (e.g. no particles)    */
#define QSS_E_OVERFLOW     -5   /* sub-cycle event buffer too small         */
#define QSS_E_NONFINITE    -6   /* non-finite residual / field detected     */
#define QSS_E_PARTICLE     -7   /* particle outside its cell (ref exit(-212))*/

/* ---- enums ------------------------------------------------------------ */
/* transform kinds along one axis: FFTW REDFT11 (cos) / RODFT11 (sin),
 * reference movefieldshydrolinlayersplit.cpp:389-427                      */
#define QSS_KIND_C 0
#define QSS_KIND_S 1

/* particle shape: NGP_REF reproduces the linked split path
 * (movepart_splitlayer.cpp:1101-1217 gather at the start cell,
 * :1547-1655 single-cell deposit); CIC uses the bilinear gather/deposit of
 * movepart_layer_CIC.cpp:166-203,449-496 inside the same predictor-corrector */
#define QSS_SHAPE_NGP_REF 0
#define QSS_SHAPE_CIC     1

/* precision of the GPU path (the oracle is always FP64).  In every mixed
 * mode the plasma push runs in FP32 on an FP32 copy of the push field,
 * while particle positions and momenta are accumulated across slices in
 * FP64, deposits are summed in FP64, and the spectral state, the transform
 * operands/results in HBM and the corrector residual stay FP64.  Mixed modes
 * need ny, nz multiples of 16 and fix_subcycle = 0 (the tiled particle path). */
#define QSS_PREC_FP64         0   /* FP64 push, FP64 MFMA transforms          */
#define QSS_PREC_MIXED32      1   /* transforms on FP32 MFMA (fp32 operands)  */
#define QSS_PREC_MIXED_BF16X3 2   /* transforms on BF16 MFMA, operands split
                                     into 2 bf16 pieces, 3 products          */
#define QSS_PREC_MIXED_BF16X6 3   /* 3 bf16 pieces, 6 products (~fp32)       */

/* field ids for history / layer access */
enum qss_field {
  QSS_EX = 0, QSS_EY, QSS_EZ, QSS_BX, QSS_BY, QSS_BZ,   /* fields  (f_Fields)  */
  QSS_JX, QSS_JY, QSS_JZ, QSS_RHO,                      /* currents, f_Dens    */
  QSS_RHOBEAM, QSS_JXBEAM,                              /* beam sources        */
  QSS_PSI,                                              /* wake potential      */
  QSS_NFIELD
};

/* ---- parameters ------------------------------------------------------- */
#define QSS_MAX_BUNCH 4

/* One frozen bi-Gaussian driver/witness bunch.  Profile follows
 * Specie::Density case 1 (plasma.cpp:350-377):
 *   n = nb * exp(-dx^2/Rx^2 - dy^2/R^2 - dz^2/R^2), zero where arg > 16,
 * with Rx = sqrt(2)*sigma_xi and R = sqrt(2)*sigma_r.  Macro-particles are
 * seeded 4 per 3-D cell at (x=1/2, {1/3,2/3}^2) (mesh.cpp:140-163), weight =
 * profile(cell corner)/4 * (1 + noise*u), u ~ U(-1,1) from the seed, and are
 * deposited trilinearly (movebeam.cpp:662-770) with dRho = w*polarity,
 * dJx = w*polarity*vx (movebeam.cpp:509-512).                               */
typedef struct qss_bunch {
  double nb;        /* peak density / n0                                    */
  double sigma_r;   /* transverse rms size (k_p^-1)                         */
  double sigma_xi;  /* longitudinal rms size (k_p^-1)                       */
  double xi0;       /* centre, distance behind the front layer (k_p^-1)     */
  double y0, z0;    /* transverse centre relative to the box centre         */
  double polarity;  /* charge in electron units (-1 for electrons)          */
  double vx;        /* frozen longitudinal velocity (1)                     */
  double noise;     /* relative macro-particle weight noise amplitude       */
} qss_bunch;

typedef struct qss_params {
  int32_t ny, nz;        /* transverse cells (l_My, l_Mz)                    */
  double  ly, lz;        /* transverse box (Ylength, Zlength)                */
  double  hx;            /* layer thickness Hx                               */
  int32_t nx_split;      /* NxSplit (HxSplit = hx / nx_split)                */
  int32_t nslices;       /* number of layers swept (l_Mx)                    */
  int32_t n_iter;        /* predictor-corrector iterations nIter             */
  double  dens;          /* plasma density n0 (units k_p^2)                  */
  int32_t ppc_y, ppc_z;  /* plasma particles per cell along y, z (2,2)       */
  double  jitter;        /* +- uniform position jitter, cell units           */
  uint64_t seed;         /* seeds jitter and beam noise                      */
  int32_t shape;         /* QSS_SHAPE_*                                      */
  int32_t fix_subcycle;  /* 0: reproduce the sub-cycle restart quirk         */
  int32_t precision;     /* QSS_PREC_*                                       */
  int32_t nbunch;        /* bunches used                                     */
  qss_bunch bunch[QSS_MAX_BUNCH];
  /* optional piecewise-constant longitudinal profile replacing the
   * Gaussian xi factor of bunch[0]: nbins values over [bin_xi0, bin_xi1) */
  int32_t nbins;
  double  bin_xi0, bin_xi1;
  const double *bins;    /* host pointer, nbins values (copied at create)   */
  /* batch of independent wakes on this grid, advanced together: every
   * kernel launch of a layer step covers all of them (0 or 1: one wake).
   * wake_bunch (host, nwake*nbunch, copied at create; NULL = every wake uses
   * `bunch`) gives wake w its bunches at [w*nbunch, (w+1)*nbunch).  The
   * plasma, profile bins and grid are shared.  nwake > 1 needs the tiled
   * particle path (ny, nz multiples of 16, fix_subcycle = 0) and precision
   * FP64 or MIXED32; the adjoint works on single-wake contexts.            */
  int32_t nwake;
  const qss_bunch *wake_bunch;
  /* uniform external magnetic field felt by the plasma electrons, added to
   * the gathered B before the Boris rotation (bXext..bZext =
   * BxExternal..BzExternal / sf_DimFields, movepart_splitlayer.cpp:867-869,
   * 1271-1273; sf_DimFields = 1, domain.cpp:204).  Zero-initialise the
   * struct and the field is off.                                           */
  double  b_ext[3];
} qss_params;

/* ---- context ---------------------------------------------------------- */
typedef struct qss_ctx qss_ctx;

/* Creates a context on HIP device `device`, allocates the layer buffers
 * (SoA plasma particles, C/P field planes, spectral state, transform
 * matrices) and seeds the front layer (Mesh::SeedFrontParticles,
 * mesh.cpp:98-105).  history_mask: bit f set => store field f of every swept
 * layer in device memory (nslices planes each). */
int  qss_create(const qss_params *p, int device, uint32_t history_mask,
                qss_ctx **out);
int  qss_destroy(qss_ctx *c);
const char *qss_last_error(void);
/* Returns the compiled-in sizes: particles per layer etc. (per wake) */
int  qss_info(const qss_ctx *c, int64_t *n_particles, int64_t *n_cells);
/* Batched contexts: the wake that the data-access calls below (layer
 * fields, spectra, history, particles, pointers) refer to; default 0.
 * Sweep residuals of a batch are returned wake-major.                      */
int  qss_select_wake(qss_ctx *c, int wake);
/* The HIP stream all work of this context runs on (hipStream_t). */
void *qss_stream(qss_ctx *c);

/* ---- whole sweep: Mesh::MoveAllSplitLayers (movepart_splitlayer.cpp:25) */
/* Re-seeds the front layer, zeroes the spectral state and sweeps
 * `nslices` layers front to back.  residuals (host, may be NULL) receives
 * the final-pass residual of every layer step (nslices*nx_split). */
int  qss_sweep(qss_ctx *c, double *residuals);
/* Sweeps only layers [first, first+count) of the slice order (0 = first
 * layer behind the front); used for timing windows and checkpoints.       */
int  qss_sweep_range(qss_ctx *c, int first, int count, double *residuals);
int  qss_reset(qss_ctx *c);   /* re-seed particles, zero fields/spectra    */

/* on != 0: qss_sweep/qss_sweep_range replay pairs of layer steps from one
 * captured hipGraph (layers >= 1, nx_split = 1, timing off) instead of
 * launching every kernel from the host.  Results are bitwise those of the
 * eager launches.  Off by default.                                         */
int  qss_enable_graph(qss_ctx *c, int on);

/* ---- one layer step: Mesh::MoveSplitLayer (movepart_splitlayer.cpp:561) */
/* slice: 0-based position behind the front; isplit < nx_split.
 * Sets the C/P beam planes (interpolation, movepart_splitlayer.cpp:330-348),
 * runs Guess, nIter x (deposit, Iterate), the full-step push and the final
 * Iterate, then copies C -> P.  residual (host, may be NULL).              */
int  qss_layer_step(qss_ctx *c, int slice, int isplit, double *residual);

/* ---- the pieces of a layer step ---------------------------------------- */
/* CUDA_WRAP_interpolateLayers (cuLayers.cu:437) + beam deposit of layers
 * i and i+1 (CUDA_WRAP_beam_move deposit part, cuBeam.cu:739): fills the
 * C and P beam planes for (slice, isplit).                                  */
int  qss_beam_layer(qss_ctx *c, int slice, int isplit);
/* Mesh::GuessFieldsHydroLinLayerSplit (movefieldshydrolinlayersplit.cpp:222):
 * beam forward transforms, spectral predictor, inverse -> C-layer E, B.    */
int  qss_guess(qss_ctx *c, int first_layer);
/* Mesh::ClearCurrentsSplit + MoveParticlesLayerSplit + ExchangeCurrentsSplit
 * (mesh.cpp:692, movepart_splitlayer.cpp:781,1685): push every plasma
 * particle of P through E,B = (P+C)/2 and deposit Jx,Jy,Jz,Dens into C.
 * full_step != 0 also stores the pushed particles as the C layer.          */
int  qss_push_deposit(qss_ctx *c, int full_step);
/* Mesh::IterateFieldsHydroLinLayerSplit (movefieldshydrolinlayersplit.cpp:
 * 1173) + ExchangeFieldsSplit: corrector; residual = Eerr.  final != 0 also
 * produces the real-space rho and psi planes.                               */
int  qss_iterate(qss_ctx *c, int final_pass, double *residual);
/* C -> P at the end of a split step (movepart_splitlayer.cpp:423-428).     */
int  qss_copy_c_to_p(qss_ctx *c);

/* ---- transforms: CUDA_WRAP_fourierHalfInteger2D (half_integer2D.cu:583) */
/* out = scale * T_kz . in . T_ky on device pointers (Nz x Ny planes),
 * T^c_{jk} = 2cos(pi(j+1/2)(k+1/2)/N), T^s = 2sin(...).  scale = 1 for the
 * forward transform, 1/(4 Ny Nz) for the inverse (normalization.cu:241).
 * `in` and `out` may not alias.                                             */
int  qss_transform2d(qss_ctx *c, int kind_z, int kind_y, const double *in_dev,
                     double *out_dev, double scale);
/* Batched form: n transforms in one launch pair. */
int  qss_transform2d_batched(qss_ctx *c, int n, const int *kind_z,
                             const int *kind_y, const double *const *in_dev,
                             double *const *out_dev, double scale);

/* ---- transform tile shapes --------------------------------------------- */
/* The reference plans its transforms with FFTW_ESTIMATE
 * (movefieldshydrolinlayersplit.cpp:389-427).  Here the transform is a
 * batched MFMA GEMM compiled for several tile shapes; qss_autotune times
 * every shape on this context's buffers for every launch shape the sweep
 * uses (1..8 fields, both passes, the 6- and 8-field inverse passes with
 * their corrector epilogue) and keeps the fastest.  It also picks the
 * particles per work-group of the tiled push (qss_set_particle_split) by
 * timing a predictor pass.  It overwrites only scratch state (transform
 * intermediates, the push field and residual slots that the next layer step
 * rebuilds).  reps >= 1.                                                   */
int  qss_autotune(qss_ctx *c, int reps, float *table_ms);
/* table_ms (may be NULL): 2*8*qss_gemm_variants(precision) floats, the mean
 * time of one launch for [pass][nbatch-1][variant].                         */
/* variant >= 0 forces one tile shape for every launch (testing); -1 returns
 * to the tuned/default choice.  qss_gemm_variants returns the number of tile
 * shapes compiled for a precision mode (QSS_PREC_*).                        */
int  qss_set_gemm_variant(qss_ctx *c, int variant);
int  qss_gemm_variants(int precision);
/* chosen[pass*9 + nbatch] for pass 0 (X.T_y), 1 (T_z.Y), nbatch 1..8.      */
int  qss_gemm_choice(const qss_ctx *c, int *chosen);
/* Installs a choice table (same layout), e.g. one tuned on another context
 * of the same grid and precision.                                           */
int  qss_set_gemm_choice(qss_ctx *c, const int *chosen);
/* Particles per work-group of the tiled push: a 16x16-cell tile's particles
 * are shared by nsplit = 1, 2 or 4 work-groups (each takes every nsplit-th
 * chunk of 256).  No reference counterpart (the reference launches one
 * thread per particle, plasma_particles.cu:761-863).  The default depends on
 * the grid only (1 from 1024 tiles, 2 from 512, else 4); results are
 * bitwise reproducible for a given nsplit and differ across nsplit by
 * summation order only.  qss_particle_split returns the current value.    */
int  qss_set_particle_split(qss_ctx *c, int nsplit);
int  qss_particle_split(const qss_ctx *c);

/* ---- data access (host buffers, Ny*Nz doubles, synchronous) ------------ */
/* The reference keeps the layer fields in the Cell members of its C and P
 * layer arrays (p_CellLayerC / p_CellLayerP, mesh.h:30-31; Cell, cells.h:30),
 * stores each finished layer into the 3-D cell array (p_CellArray, mesh.h:29;
 * movepart_splitlayer.cpp:467-473), and keeps the plasma particles in per-cell
 * lists (Cell::p_Particles, cells.h:80; Particle, particles.h:12).  These calls
 * read and write the same quantities.
 * which: 0 = C layer, 1 = P layer.  field: QSS_EX..QSS_PSI.  A layer step
 * ends by trading the C and P planes, so after qss_layer_step or a sweep the
 * finished layer is the P layer and the C planes are the next step's scratch
 * (inside a sweep its density plane is not written).                        */
int  qss_get_layer_field(qss_ctx *c, int which, int field, double *host);
int  qss_set_layer_field(qss_ctx *c, int which, int field, const double *host);
/* Device pointer of a layer plane; valid until the next layer step (the C
 * and P planes trade places at the end of every step).                      */
int  qss_layer_ptr(qss_ctx *c, int which, int field, double **dev);
/* history of field f for slice s (requires history_mask bit). */
int  qss_get_history(qss_ctx *c, int field, int slice, double *host);
/* Device pointer of the history block of field f (nslices*Ny*Nz doubles). */
int  qss_history_ptr(qss_ctx *c, int field, double **dev);
/* Plasma particles in slot order (sorted by cell after every full step).
 * State is exact: cell indices (j,k) plus in-cell fractions y,z in [0,1]
 * (the reference's Particle f_Y/f_Z inside its Cell); id = seeding index.
 * Any pointer may be NULL.  set: n <= the capacity fixed at create.        */
int  qss_get_particles(qss_ctx *c, int32_t *cell_j, int32_t *cell_k, double *fy,
                       double *fz, double *px, double *py, double *pz,
                       int64_t *id);
int  qss_set_particles(qss_ctx *c, int64_t n, const int32_t *cell_j,
                       const int32_t *cell_k, const double *fy, const double *fz,
                       const double *px, const double *py, const double *pz,
                       const int64_t *id);
/* Spectral state (fft_of_* arrays, movefieldshydrolinlayersplit.cpp:267-313):
 * id 0..20 = Rho,Ex,Ey,Ez,Bx,By,Bz,Jx,Jy,Jz,RhoP,JxP,JyP,JzP,ExP,EyP,EzP,
 * RhoBeam,JxBeam,RhoBeamP,JxBeamP.                                          */
#define QSS_NSPECTRAL 21
int  qss_get_spectral(qss_ctx *c, int id, double *host);
int  qss_set_spectral(qss_ctx *c, int id, const double *host);

/* ---- differentiation: adjoint of the sweep ------------------------------ */
/* The reference has no derivative of its sweep.  These entry points give the
 * gradient of a scalar loss of the Ex planes (the longitudinal field, history
 * of QSS_EX) with respect to the driver's piecewise-constant longitudinal
 * profile (qss_params.nbins > 0), by a hand-written adjoint: reverse push
 * with the transposed CIC gather and deposit, transposed spectral updates and
 * transposed transform GEMMs, recomputing each segment of `every` layers from
 * a checkpoint.  FP64, CIC, nx_split = 1, ny and nz multiples of 16.        */
/* Replaces the profile values (nbins = the count given at create).          */
int  qss_set_bins(qss_ctx *c, const double *bins, int nbins);
/* Re-seeds and sweeps layers [0, nrun), storing the state at every
 * `every`-th layer start (device memory, kept in the context) and the
 * history planes of history_mask.  every = 0: the forward records the
 * reverse sweep's tape of every layer instead (particles at the layer start
 * and each pass's push field, 2 x 8 x (3 + ppc x 3) doubles per cell and
 * layer for nrun layers, ~25 MB per layer at 256^2 x 4 ppc), and
 * qss_sweep_vjp recomputes nothing.  residuals: host, nrun values or NULL. */
int  qss_sweep_checkpointed(qss_ctx *c, int nrun, int every, double *residuals);
/* Vector-Jacobian product of the preceding checkpointed sweep: ex_bar =
 * dLoss/dEx of layers [0, nrun) (nrun*Ny*Nz doubles, layer-major; device
 * memory, or host memory if on_host != 0; NULL = 0); grad_bins (host,
 * nbins) = dLoss/dbins.  Leaves the layer state unspecified (call qss_reset
 * or a sweep next).                                                          */
int  qss_sweep_vjp(qss_ctx *c, const double *ex_bar, int on_host, double *grad_bins);
/* Device-to-device copy of history planes [first, first+count) of a field
 * (no counterpart in the reference; the adjoint's loss reads them).         */
int  qss_copy_history(qss_ctx *c, int field, int first, int count, double *dst_dev);
/* Bytes of device memory in use on `device` (all processes) and in total
 * (no counterpart in the reference).                                        */
int  qss_device_memory(int device, int64_t *used, int64_t *total);

/* ---- layer hand-off between xi chunks ----------------------------------- */
/* The reference pipelines a sweep over MPI ranks that own consecutive xi
 * chunks: a rank receives the boundary layer from the rank in front, sweeps
 * its chunk and sends its last layer on (PackLayer / SendLayer /
 * ReceiveLayer, para.cpp:94-412; movepart_splitlayer.cpp:125,545).
 *
 * Two forms of that hand-off:
 *
 * 1. The state a sweep continues from, bitwise: a 256-byte header (format
 *    magic and version, grid, tiling, particle capacity), the particles with
 *    their tile starts, the P-layer planes and the spectral state, as one
 *    flat buffer of qss_layer_state_bytes bytes in device memory (what a
 *    device-to-device send carries) or host memory.  qss_unpack_layer takes
 *    the buffer's size and returns QSS_E_ARG, leaving the context untouched,
 *    when the size or header does not match this context or a particle cell
 *    or tile start is out of place.  A context that unpacks the state of a
 *    context of the same parameters after layer s-1 and sweeps layers
 *    [s, ...) with qss_sweep_range gives the layers of one unbroken sweep
 *    bitwise.  Single-wake contexts.                                         */
int  qss_layer_state_bytes(const qss_ctx *c, size_t *bytes);
int  qss_pack_layer(qss_ctx *c, void *dst);
int  qss_unpack_layer(qss_ctx *c, const void *src, size_t bytes);
/* 2. The reference's own record (PackLayer para.cpp:94-160, UnpackLayer
 *    :167-280), host memory, qss_layer_record_doubles doubles: 12 planes of
 *    Ny*Nz (y fastest) Bx, By, Bz, Ex, Ey, Ez, Jx, Jy, Jz, Rho,
 *    fftJxBeamHydro, fftRhoBeamHydro (the beam spectra of the last layer),
 *    then 12 attribute arrays of Np particles: Px, Py, Pz, X, Y, Z, Q2m,
 *    Weight, i_X, i_Y, i_Z, isort.  qss_pack_layer_record writes the P layer
 *    (the last layer swept) as reference layer iLayer (i_X); unpack takes Np
 *    from the length as ReceiveLayer does (para.cpp:385), requires every
 *    particle to be a plasma electron of this sweep (Q2m = -1, the plasma
 *    weight, isort 0, cell inside the grid, fractions in [0, 1]) and Np =
 *    ppc * Ny * Nz, and rebuilds the P-layer spectra from the planes as the
 *    reference's Guess does at every layer; the continuation then agrees
 *    with an unbroken sweep to rounding.  Single-wake contexts.            */
int  qss_layer_record_doubles(const qss_ctx *c, size_t *n);
int  qss_pack_layer_record(qss_ctx *c, int iLayer, double *dst);
int  qss_unpack_layer_record(qss_ctx *c, const double *src, size_t n);

/* ---- transform method ------------------------------------------------------ */
/* No counterpart in the reference (whose FFTW plans pick their own
 * algorithm).  DENSE: each 1-D pass is one GEMM with the N x N half-integer
 * table (the reference's HalfInteger matrix form, fft_matrix.cu:10-24).
 * FACTORED: one radix-2 step of the DCT-IV's half-length complex DFT turns
 * each pass into an (N/2 x N/2) GEMM applied to both halves of every row,
 * half the multiply-adds (FP64 contexts with ny, nz multiples of 8 and at
 * least 256).  Where ny and nz are multiples of 128 (at most 1024) the
 * twiddles run inside the two GEMMs; elsewhere in twiddle kernels between
 * them.  Contexts start FACTORED where the twiddles run inside the GEMMs
 * (measured faster at 256^2, 384^2 and 512^2) and DENSE otherwise.  Results
 * agree to rounding.  Setting a method drops a captured graph.              */
#define QSS_TRANSFORM_DENSE    0
#define QSS_TRANSFORM_FACTORED 1
int  qss_set_transform_method(qss_ctx *c, int method);
int  qss_transform_method(const qss_ctx *c);
/* The FACTORED method's tile shapes (qss_autotune tunes them), 4 blocks of
 * 9 ints, each indexed by batch size 0..8 like qss_gemm_choice: the GEMMs of
 * the twiddle-kernel form (y pass, z pass; FP64 variants), then the fused
 * kernels (y pass, z pass; their own variant numbers, -1 where the context
 * has no fused form).  Copying a choice between contexts of the same grid
 * makes their transforms bitwise alike.                                    */
#define QSS_FACT_CHOICE_INTS (4 * 9)
int  qss_fact_choice(const qss_ctx *c, int *chosen);
int  qss_set_fact_choice(qss_ctx *c, const int *chosen);

/* ---- allocation-free transform ------------------------------------------ */
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): qss_transform2d without a context.  out = scale *
 * T_kind_z . in . T_kind_y for one Ny x Nz plane (y fastest) in device
 * memory, FP64 MFMA, enqueued on `stream` (a hipStream_t of the current
 * device; NULL = its null stream): no allocation and no synchronisation.
 * The transform tables and the intermediate plane live in a fixed per-device
 * scratch array of 3 x 512 x 512 doubles (6 MiB, held by each device the
 * library runs on until the process ends), so ny, nz <= 512 (QSS_E_ARG
 * beyond).  Every call on a device shares that scratch: calls that overlap
 * on one device race (issue them on one stream).                           */
int  qss_transform2d_noalloc(int ny, int nz, int kind_z, int kind_y, const double *in_dev,
                             double *out_dev, double scale, void *stream);

/* Writes history plane `slice` of field f (requires the history_mask bit);
 * the wake an evolving driver is pushed through (qss_driver_push).         */
int  qss_set_history(qss_ctx *c, int field, int slice, const double *host);

/* ---- evolving driver: Domain::Step (step.cpp:23-105) -------------------- */
/* The bunches above are a frozen driver: every sweep deposits their seeding
 * lattice.  An evolving driver keeps the beam macro-particles in the 3-D
 * cell grid (layer i = x-cell, transverse cell j, k, in-cell fractions
 * x, y, z), pushes them between sweeps through the stored wake
 * (Mesh::MoveBeamParticles, movebeam.cpp:14-660, the host path the
 * reference links, run_control.h:22) and deposits rho_b and Jx_b
 * trilinearly into 3-D node arrays (Mesh::MoveBeamInCell,
 * movebeam.cpp:662-770) that the following sweeps interpolate per layer
 * (movepart_splitlayer.cpp:330-348) instead of the frozen lattice.         */
typedef struct qss_driver {
  double ts;                    /* time step Ts (&Domain Ts)                  */
  double b_ext[3];              /* external B (BxExternal.. / sf_DimFields)   */
  double syn_min;               /* &Synchrotron SynMin: radiation reaction of
                                   movebeam.cpp:448-501 where 1+p^2 > SynMin^2
                                   (reference default 100)                    */
  double wavelength;            /* &Domain Wavelength: photon energy scale
                                   1.2398e-4 / Wavelength (default 0.001)     */
  double p0[QSS_MAX_BUNCH][3];  /* seeded momentum Px0, Py0, Pz0 per bunch    */
  double q2m[QSS_MAX_BUNCH];    /* charge/mass of each bunch's particles      */
} qss_driver;

/* Mesh::SeedBeamParticles (mesh.cpp:43-64, 108-276) for every bunch: 4
 * particles per 3-D cell at (x = 1/2, {1/3,2/3}^2) of the cells i in
 * [0, nslices), j, k with a positive profile at the cell corner, weight =
 * profile/4 (times the bunch's noise factor), momentum p0.  From now on the
 * sweeps read the driver's 3-D deposit, which is zero until the first
 * qss_driver_push.  Needs history of QSS_EX..QSS_BZ, one wake, no profile
 * bins.  n_particles (may be NULL) receives the count.                      */
int  qss_driver_seed(qss_ctx *c, const qss_driver *d, int64_t *n_particles);
/* Mesh::MoveBeamParticles: clears the 3-D deposit (Mesh::ClearBeamCurrents,
 * mesh.cpp:648-662), then every particle in an interior cell (0 <= i <
 * nslices, 0 <= j < Ny, 0 <= k < Nz) gathers E, B trilinearly from the
 * stored wake (layer i = history slice nslices-1-i, layer nslices = the
 * unperturbed front, transverse ghosts periodic), gets the reference's
 * kick / Boris rotation / kick and radiation reaction, moves by
 * ((Vx-1) ts/hx, Vy ts/hy, Vz ts/hz) with the reference's cell clamping
 * to [-1, N], and deposits w*polarity and w*polarity*Vx on the 8 nodes of
 * its new cell.  Particles in clamped (ghost) cells stay where they are.
 * Deposits are summed in 64-bit fixed point (order-independent, so the
 * result is bitwise reproducible): resolution 2^-e with e fixed at seeding
 * from the total |w*polarity|, at most 2^-(e+1) rounding per contribution. */
int  qss_driver_push(qss_ctx *c);
/* Domain::Step (step.cpp:63,105): qss_driver_push, then qss_sweep.          */
int  qss_time_step(qss_ctx *c, double *residuals);
/* Driver particles (host arrays, any may be NULL): cell = (i, j, k) per
 * particle, frac = in-cell (x, y, z) in [0,1], mom = (px, py, pz), weight,
 * bunch index.                                                              */
int  qss_driver_get(qss_ctx *c, int32_t *cell, double *frac, double *mom, double *weight,
                    int32_t *bunch);
/* Replaces the driver particles (n of them, same layout; needs a seeded
 * driver for the push parameters).  Recomputes the fixed-point scale.       */
int  qss_driver_set(qss_ctx *c, int64_t n, const int32_t *cell, const double *frac,
                    const double *mom, const double *weight, const int32_t *bunch);
/* 3-D deposit of reference layer `layer` in [0, nslices] (Ny*Nz doubles
 * each, host; either may be NULL).                                           */
int  qss_driver_density(qss_ctx *c, int layer, double *rho, double *jx);
/* Kernel times of the last qss_driver_push (dispatch-stamped HIP events):
 * the clear of the previous deposit's box and the push + deposit.  No
 * counterpart in the reference.                                             */
int  qss_driver_timing(qss_ctx *c, double *ms_clear, double *ms_push);

/* ---- timing of the last qss_sweep/qss_sweep_range (HIP events) --------- */
/* No counterpart in the reference's layer call surface.
 * Kernel classes: 0 push, 1 transform GEMM, 2 spectral update, 3 beam,
 * 4 counting sort (fill+scan+scatter), 5 deposit tile reduce.              */
#define QSS_NKCLASS 6
/* on != 0: record a HIP event pair around every launch of the next sweeps.
 * ms_total: wall time of the last sweep window on the stream; ms_kernel[i],
 * launches[i]: summed event time / launch count per class; gemm_flops:
 * algorithmic flops (2*M*N*K per transform pass and field) of the GEMMs.   */
int  qss_enable_timing(qss_ctx *c, int on);
int  qss_timing(qss_ctx *c, double *ms_total, double *ms_kernel,
                int64_t *launches, double *gemm_flops);

/* ---- HDF5 output and restart files: Mesh::Save / Mesh::Load ------------ */
/* Self-contained (no libhdf5): files of the HDF5 format the reference's
 * library defaults produce (superblock 0, symbol-table root group, contiguous
 * datasets), readable by any HDF5 reader; the reader takes such files from
 * any writer.  Host-only: these calls need no GPU except qss_save/qss_load.
 *
 * Particle records as stored (SavedParticle / SavedIon, particles.h:111-153;
 * compound members registered at save.cpp:225-251): */
typedef struct qss_saved_particle {
  int32_t cell;                 /* ghosted cell number GetN(i,j,k)          */
  int32_t pad_;
  double  x, y, z;              /* position (cell + in-cell fraction) * h   */
  double  px, py, pz, weight, q2m;
} qss_saved_particle;           /* 72 bytes                                 */
typedef struct qss_saved_ion {
  int32_t cell;
  int32_t pad_;
  double  x, y, z, px, py, pz, weight, q2m;
  int32_t i_z, sort;            /* charge state, species                     */
} qss_saved_ion;                /* 80 bytes                                 */

#define QSS_H5_F64       0      /* H5T_IEEE_F64LE                            */
#define QSS_H5_I32       1      /* H5T_STD_I32LE                             */
#define QSS_H5_I64       2      /* H5T_STD_I64LE                             */
#define QSS_H5_PARTICLE  3      /* qss_saved_particle records               */
#define QSS_H5_ION       4      /* qss_saved_ion records                    */
typedef struct qss_h5_set {
  const char *name;             /* dataset name in the root group           */
  int32_t     kind;             /* QSS_H5_*                                  */
  int32_t     rank;             /* 1..3                                      */
  int64_t     dims[3];          /* slowest first                             */
  const void *data;             /* host, product(dims) elements              */
} qss_h5_set;
/* WriteHDF5Recorddouble / RecordInt / RecordLong (myhdfshell.cpp) for a
 * whole file: creates `path` (truncating) with n datasets.                  */
int  qss_h5_write(const char *path, int n, const qss_h5_set *sets);
/* Shape and kind of a dataset (kind: QSS_H5_* of the stored type, -1 for
 * other types); dims[3] filled slowest first, unused entries 0.  A helper of
 * this library's readers; the reference's readers know their datasets.      */
int  qss_h5_info(const char *path, const char *name, int32_t *kind, int32_t *rank,
                 int64_t *dims);
/* ReadHDF5Record*: elements [first, first+count) of a dataset, converted to
 * `kind` (integer and float widths; compound records member by member,
 * matched by name).                                                        */
int  qss_h5_read(const char *path, const char *name, int32_t kind, int64_t first,
                 int64_t count, void *dst);
/* Dataset names of the file, one per line, NUL-terminated (a helper, as
 * qss_h5_info).                                                             */
int  qss_h5_list(const char *path, char *names, size_t cap);

/* Mesh::Save (save.cpp:178-835) of PE 0: writes <dir>/vs<isave>_3d_fields.h5
 * (meshinfo, phase, X Y Z axes with the window shift, ex ey ez bx by bz jx jy
 * jz dens Dens<sort> as [Mz][My][Mx] from the history of the selected wake;
 * needs history of QSS_EX..QSS_JZ, and QSS_RHOBEAM without an evolving
 * driver) and <dir>/vs<isave>_3d_particles.h5 (the evolving driver's
 * particles as Ions_of_sort_<1+bunch>_at_PE_0 records, Nparticles_at_PE_0).
 * Sort 0 is the plasma, which has no particles in the 3-D mesh.            */
int  qss_save(qss_ctx *c, const char *dir, int isave, double phase, int shift);
/* Mesh::Load (load.cpp:285-1080): checks meshinfo against the context,
 * loads every field of ex..jz that the context keeps history of into that
 * history (selected wake) and, with a seeded evolving driver, replaces its
 * particles by the file's records of sorts 1..nbunch (qss_driver_set; the
 * in-cell fraction is x/hx - i, clamped to [0, 1]).  phase may be NULL.     */
int  qss_load(qss_ctx *c, const char *dir, int isave, double *phase);

#ifdef __cplusplus
}
#endif
#endif /* QSS_H */
