/*
 * crava_hip.h -- C ABI of the MI355X (gfx950) implementation of CRAVA's
 * Fourier-domain Bayesian AVO inversion hot path.
 *
 * Plain C: opaque handles, plain pointers and sizes.  Every entry point
 * returns CRV_OK (0) or a negative CRV_E_* code; the text of the last error
 * on a context is returned by crv_last_error().  Numerical failures of a
 * single wavenumber (a Cholesky factorisation that does not succeed) are not
 * errors: they reproduce the reference's fallback and are counted in
 * crv_post_stats.
 *
 * Grid storage follows FFTGrid exactly (reference src/fftgrid.h:284-322,
 * src/fftgrid.cpp:40-70): one buffer of rsize = rnxp*nyp*nzp reals with
 * rnxp = 2*(nxp/2+1), x fastest; real index i + rnxp*j + rnxp*nyp*k; after the
 * forward transform the same buffer holds cnxp*nyp*nzp complex values,
 * complex index i + cnxp*j + cnxp*nyp*k (src/fftgrid.cpp:1544,1677).
 * Host buffers passed to upload/download use that same layout, pad columns
 * included; they are borrowed only for the duration of the call.
 *
 * Threading: one host thread per crv_ctx; all device work of a context is
 * issued on the context's HIP stream in call order.
 */
#ifndef CRAVA_HIP_H
#define CRAVA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- return codes ------------------------------------------------------ */
#define CRV_OK              0
#define CRV_E_INVALID      -1  /* bad argument / inconsistent dimensions     */
#define CRV_E_STATE        -2  /* grid in the wrong (transformed) state      */
#define CRV_E_HIP          -3  /* HIP runtime error                          */
#define CRV_E_NOMEM        -4  /* device allocation failed                   */
#define CRV_E_UNSUPPORTED  -5  /* size or option not supported               */

/* ---- precision --------------------------------------------------------- */
#define CRV_FP32 32
#define CRV_FP64 64

/* ---- grid types: same values as FFTGrid::gridTypes (src/fftgrid.h:193) - */
#define CRV_CTMISSING  0
#define CRV_DATA       1
#define CRV_PARAMETER  2
#define CRV_COVARIANCE 3
#define CRV_VELOCITY   4
#define CRV_OPERATOR   5

typedef struct crv_ctx  crv_ctx;
typedef struct crv_grid crv_grid;

typedef struct crv_grid_info {
  int nx, ny, nz;        /* un-padded sizes                                  */
  int nxp, nyp, nzp;     /* padded sizes                                     */
  int cnxp, rnxp;        /* nxp/2+1 and 2*cnxp                               */
  int precision;         /* CRV_FP32 | CRV_FP64                              */
  int cubetype;          /* CRV_DATA ...                                     */
  int is_transformed;    /* FFTGrid::istransformed_                          */
  size_t rsize;          /* rnxp*nyp*nzp reals                               */
  size_t csize;          /* cnxp*nyp*nzp complex values                      */
} crv_grid_info;

/* ---- context ------------------------------------------------------------
 * Replaces the global FFTW plan creation of FFTGrid::fftInPlace
 * (src/fftgrid.cpp:1799-1804): a context owns the device, the stream and the
 * cached transform plans (twiddle tables) for every grid size it has seen.
 * Grids, posterior plans, graphs and networks made on a context are
 * destroyed before it; after crv_ctx_destroy their handles are invalid.   */
int         crv_ctx_create(int device, crv_ctx** out);
int         crv_ctx_destroy(crv_ctx* ctx);
const char* crv_last_error(const crv_ctx* ctx);
int         crv_ctx_synchronize(crv_ctx* ctx);
/* Use an external HIP stream (hipStream_t passed as void*; NULL = own). */
int         crv_ctx_set_stream(crv_ctx* ctx, void* hip_stream);
/* Library version string, e.g. "crava_hip 1.0 gfx950". */
const char* crv_version(void);
/* Kernel timing with HIP events on the context's stream (no counterpart in
 * the reference, whose phase timers are src/timings.h:11-81).  While enabled
 * every kernel launch is bracketed by a pair of events.  The report
 * synchronises, then writes one line per kernel name,
 * "name launches total_ms min_ms max_ms", clears the record and returns the
 * number of bytes needed (buf may be NULL to query).                        */
int crv_ctx_profile(crv_ctx* ctx, int enable);
int crv_ctx_profile_report(crv_ctx* ctx, char* buf, size_t len);
/* Launch-shape tuning of the transforms (no counterpart in the reference,
 * which plans FFTW_ESTIMATE per call, src/fftgrid.cpp:1801-1804).  Applies to
 * every plan of the context, present and future; 0 restores the default.
 *   CRV_TUNE_CHUNK_MB  interleave the x and y passes per chunk of z planes of
 *                      about this many MB, so the second pass reads the first
 *                      pass's output from the memory-side cache (0 = off)
 *   CRV_TUNE_TX_Y/Z    x columns per y / z pencil tile (power of two)
 *   CRV_TUNE_ROWS      maximum x rows per block (power of two)            */
#define CRV_TUNE_CHUNK_MB 1
#define CRV_TUNE_TX_Y     2
#define CRV_TUNE_TX_Z     3
#define CRV_TUNE_ROWS     4
/*   CRV_TUNE_POST_BLOCKS  posterior blocks (256 wavenumber-threads) per CU,
 *                      0 = as many as fit                                  */
#define CRV_TUNE_POST_BLOCKS 5
/*   CRV_TUNE_BATCH_MB  batched transforms (crv_fft_*_batch) launch sub-batches
 *                      of about this many MB of grids, so that each pass's
 *                      output is still in the memory-side cache when the
 *                      next pass reads it; grids larger than twice this go
 *                      in one launch per pass (0 = 256, the cache's size) */
#define CRV_TUNE_BATCH_MB 6
/*   CRV_TUNE_ORDER     1 + a bit mask of the passes that walk their tiles
 *                      backwards (1 x R2C, 2 y forward, 4 z forward, 8 z
 *                      inverse, 16 y inverse, 32 x C2R): a pass that starts
 *                      on the tiles its predecessor wrote last finds them in
 *                      the memory-side cache (0 = the library's default)   */
#define CRV_TUNE_ORDER    7
/*   CRV_TUNE_ZPOST     crv_invert's z pass with the posterior of the means
 *                      inside it (FP32 refine plans whose wavenumbers all
 *                      stay in FP32; results bit-equal to the separate z
 *                      forward pass, posterior and z inverse pass): 0 = off
 *                      (the default), 1 = on with the default tile width,
 *                      2, 4, ... = on with that many x columns per tile    */
#define CRV_TUNE_ZPOST    8
/*   CRV_TUNE_ZINV      z inverse passes of FP32 two-stage z lengths as
 *                      persistent blocks that load a tile's inputs while
 *                      the previous tile is transformed (results bit-equal):
 *                      0 = for the covariances' generating z inverse of
 *                      crv_invert (the default), 1 = for every such z
 *                      inverse, 2 = off                                     */
#define CRV_TUNE_ZINV     9
int crv_ctx_tune(crv_ctx* ctx, int key, int value);

/* Autotuner for one grid size: times the candidate launch shapes of each
 * pass (x rows per tile, y and z pencil tile widths, and the order of the
 * radix stages where a length has no two-stage kernel) with the transforms
 * themselves on scratch memory, and keeps the fastest in the size's plan.
 * shape (may be NULL) receives the choice.  A later crv_ctx_tune call
 * resets every plan to the defaults.                                        */
typedef struct crv_fft_shape {
  int rows_per_tile, tx_y, tx_z;
  int radix_order[3];            /* x, y, z: 0 default, 1 reversed           */
  double ms_r2c, ms_c2r;         /* one transform with the chosen shapes    */
} crv_fft_shape;
int crv_fft_autotune(crv_ctx* ctx, int nxp, int nyp, int nzp, int precision, int reps, crv_fft_shape* shape);

/* ---- batches and graphs --------------------------------------------------
 * count grids of one size in one allocation, grids[i] i rsize elements
 * after grids[0] (each handle is destroyed on its own; the memory goes with
 * the last one).  Batched entry points run one launch per pass over the
 * whole batch.                                                              */
int crv_grid_create_batch(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision,
                          int cubetype, int count, crv_grid** grids);
/* SeismicParametersHolder::FFTCovGrids / invFFTCovGrids
 * (seismicparametersholder.cpp:273-322) and the means'/data's transforms
 * over a batch: every grid of one contiguous batch (same size, type and
 * state) transformed by sub-batch launches.  _from writes the spectra of
 * the src batch into the dst batch (src unchanged).                         */
int crv_fft_r2c_batch(crv_grid** grids, int count);
int crv_fft_r2c_batch_from(crv_grid** dst, crv_grid** src, int count);
int crv_fft_c2r_batch(crv_grid** grids, int count);
/* HIP-graph capture of the context's stream: calls between begin and end
 * are recorded (kernel launches only: plans must exist, see
 * crv_posterior_exec and the batched simulation), then replayed with
 * crv_graph_launch.                                                         */
/* nreal realizations of one posterior with the counter noise stream (seed,
 * realization0 + r): out is a contiguous batch of 3 nreal grids, out[3r+g]
 * the realization r of parameter g (posterior means added when post_mean is
 * given); one launch for the noise and scaling of all realizations, then
 * the batched inverse transforms.  Launches only: capturable.              */
int crv_simulate_batch(crv_ctx* ctx, crv_grid* cov[6], crv_grid* post_mean[3], unsigned long long seed,
                       unsigned long long realization0, int nreal, int fp64_chol, crv_grid** out);
typedef struct crv_graph crv_graph;
int crv_graph_begin(crv_ctx* ctx);
int crv_graph_end(crv_ctx* ctx, crv_graph** out);
int crv_graph_launch(crv_graph* graph);
int crv_graph_destroy(crv_graph* graph);

/* ---- grids --------------------------------------------------------------
 * crv_grid_create replaces FFTGrid::FFTGrid + createRealGrid
 * (src/fftgrid.cpp:40-70,1243-1295).  The buffer is uninitialised.          */
int crv_grid_create(crv_ctx* ctx, int nx, int ny, int nz,
                    int nxp, int nyp, int nzp,
                    int precision, int cubetype, crv_grid** out);
int crv_grid_destroy(crv_grid* g);
int crv_grid_info_get(const crv_grid* g, crv_grid_info* info);
int crv_grid_set_cubetype(crv_grid* g, int cubetype);
/* FFTGrid::setTransformedStatus (src/fftgrid.h:185). */
int crv_grid_set_transformed(crv_grid* g, int is_transformed);
/* Device address of the grid buffer (hipDeviceptr as void*). */
int crv_grid_device_ptr(const crv_grid* g, void** dptr);

/* Host <-> device copies of the whole rsize buffer (real view) or csize
 * buffer (complex view, interleaved re/im).  The element type is float for
 * CRV_FP32 grids and double for CRV_FP64 grids.  nelem must equal rsize
 * (real) or 2*csize (complex).  upload_real clears the transformed flag,
 * upload_complex sets it; download copies whatever the buffer holds.        */
int crv_grid_upload_real(crv_grid* g, const void* host, size_t nelem);
int crv_grid_upload_complex(crv_grid* g, const void* host, size_t nelem);
int crv_grid_download(crv_grid* g, void* host, size_t nelem);
/* Device-to-device copy of buffer, flag and type (same sizes); when the
 * precisions differ every element is converted (rounded to nearest). */
int crv_grid_copy(crv_grid* dst, const crv_grid* src);
/* FFTGrid::add (src/fftgrid.cpp:1870-1891): dst += src over the whole real
 * buffer (untransformed) or complex buffer (transformed). */
int crv_grid_add(crv_grid* dst, const crv_grid* src);
/* FFTGrid::multiplyByScalar (src/fftgrid.cpp:1985-1995): all rsize reals. */
int crv_grid_multiply_by_scalar(crv_grid* g, double scalar);

/* ---- transforms ---------------------------------------------------------
 * crv_fft_r2c replaces FFTGrid::fftInPlace (src/fftgrid.cpp:1783-1808):
 * requires !is_transformed and cubetype != CTMISSING; scales by
 * 1/sqrt(nxp*nyp*nzp) unless COVARIANCE or OPERATOR; unnormalised forward
 * real-to-complex 3D DFT with sign -1 (FFTW_FORWARD) over (nzp, nyp, nxp),
 * x last; sets is_transformed.
 * crv_fft_c2r replaces FFTGrid::invFFTInPlace (src/fftgrid.cpp:1811-1841):
 * requires is_transformed; unnormalised inverse (sign +1) complex-to-real,
 * then scales by 1/N (COVARIANCE, OPERATOR) or 1/sqrt(N) (others); the
 * imaginary parts of the kx = 0 and kx = nxp/2 planes are ignored, as a
 * half-complex inverse does.  Pad columns i >= nxp hold unspecified values
 * afterwards, as in the reference.                                          */
int crv_fft_r2c(crv_grid* g);
int crv_fft_c2r(crv_grid* g);
/* Out-of-place forms: src is left untouched, dst (same sizes, precision)
 * receives the result and src's cubetype.  Same arithmetic as in place.     */
int crv_fft_r2c_from(crv_grid* dst, const crv_grid* src);
/* C2R of src into dst followed by dst += add_real (untransformed grid of the
 * same size): FFTGrid::invFFTInPlace + FFTGrid::add fused (crava.cpp:1672-1708).
 * add_real may be NULL.  src is consumed (holds unspecified values after). */
int crv_fft_c2r_add(crv_grid* dst, crv_grid* src, const crv_grid* add_real);

/* ---- posterior ----------------------------------------------------------
 * crv_posterior replaces Crava::computePostMeanResidAndFFTCov
 * (src/crava.cpp:975-1285; signature src/crava.h:43-64) for the
 * constant-thickness branch.
 *
 * Inputs/outputs, as in the reference:
 *   mean[3]   in: prior means (real, untransformed, PARAMETER) -- they are
 *             transformed here if not yet transformed (crava.cpp:177-184);
 *             out: posterior means, inverse transformed (crava.cpp:1262-1264)
 *   data[n]   in: seismic data spectra or real data (transformed here if
 *             needed); out: residual spectra (crava.cpp:1193-1195)
 *   err_corr  in: real errCorr grid (COVARIANCE), transformed here
 *             (crava.cpp:1062); may be NULL when separable != 0
 *   cov[6]    order aa, bb, rr, ab, ar, br.  in: prior covariance grids
 *             (real, COVARIANCE; transformed here if not yet transformed,
 *             seismicparametersholder.cpp:299-322) or ignored when
 *             separable != 0; out: posterior covariance spectra
 *             (crava.cpp:1196-1205).
 * The per-kz tables follow Crava::fillErrMultVectors (crava.cpp:1336-1364)
 * for constant thickness: kwd[k][l] = conj(W_l(k)) * D1(k) (complex),
 * em1[k][l] = conj(W_l(k))/|W_l|, em2[k][l] = conj(E3_l(k))/|E3_l|,
 * em3[k][l] = 1 -- all interleaved re/im doubles, nzp*n complex each.       */

#define CRV_MODE_PLAIN        0   /* arithmetic in the grids' precision      */
#define CRV_MODE_FP32_REFINE  1   /* FP32 grids; FP32 solve with FP32
                                     accumulation; wavenumbers rejected in
                                     FP32, with condition estimate above
                                     cond_threshold or |k| below
                                     lowk_fraction of Nyquist are re-solved
                                     in FP64 with one refinement step        */

typedef struct crv_post_params {
  int           ntheta;        /* n, 1..8                                    */
  int           nz;            /* un-padded nz (frequency axis scaling)      */
  double        lz;            /* simbox lz (ms), frequency = nz*1000/(lz*nzp)*min(k,nzp-k) */
  double        min_rel_thick; /* simbox->getMinRelThick()                   */
  double        low_cut;       /* Hz                                         */
  double        high_cut;      /* Hz                                         */
  double        wnc;           /* white-noise component                      */
  const float*  A;             /* ntheta x 3, row-major (Aki-Richards rows)  */
  const double* err_theta_cov; /* ntheta x ntheta                            */
  const double* kwd;           /* nzp x ntheta complex                       */
  const double* em1;           /* nzp x ntheta complex                       */
  const double* em2;           /* nzp x ntheta complex                       */
  const double* em3;           /* nzp x ntheta complex                       */
  int           mode;          /* CRV_MODE_*                                 */
  double        cond_threshold;/* FP64 re-solve above this (default 1e5)     */
  double        lowk_fraction; /* FP64 re-solve for |k|/Nyquist below this   */
  /* Separable prior (gradI = gradJ = 0): prior covariance spectrum
   * Sigma0[a][b] * |Re(cxy[kx,ky] * ct[kz])| and error spectrum
   * |Re(cxy[kx,ky])|, with cxy the 2D spectrum of the lateral correlation
   * (cnxp x nyp reals) and ct the 1D spectrum of the temporal correlation
   * (nzp reals).  When separable == 0 these are ignored and cov/err_corr
   * grids are read as the reference does.                                   */
  int           separable;
  const double* sigma0;        /* 3x3 row-major                              */
  const double* cxy;           /* cnxp*nyp                                   */
  const double* ct;            /* nzp                                        */
  int           skip_mean_inverse; /* 1: leave posterior means as spectra     */
} crv_post_params;

typedef struct crv_post_stats {
  long long n_wavenumbers;     /* cnxp*nyp*nzp                               */
  long long n_inverted;        /* wavenumbers inside the frequency band      */
  long long n_chol_fail;       /* fell back to the prior (crava.cpp:1421)    */
  long long n_fp64;            /* re-solved in FP64 (CRV_MODE_FP32_REFINE)   */
  double    max_cond;          /* largest condition estimate seen            */
} crv_post_stats;

int crv_posterior(crv_ctx* ctx, const crv_post_params* p,
                  crv_grid* mean[3], crv_grid** data, crv_grid* err_corr,
                  crv_grid* cov[6], crv_post_stats* stats);

/* The same update split for batches and HIP-graph capture: a plan holds the
 * per-kz tables on the device for one grid size and precision;
 * crv_posterior_exec only launches kernels (no allocation, no host
 * synchronisation) on inputs that are already transformed and leaves the
 * posterior means as spectra; crv_post_plan_stats reads the counters of the
 * last exec (synchronises).  Execs of one plan must not overlap.          */
typedef struct crv_post_plan crv_post_plan;
int crv_post_plan_create(crv_ctx* ctx, const crv_post_params* p, int nxp, int nyp, int nzp, int precision,
                         crv_post_plan** out);
int crv_post_plan_destroy(crv_post_plan* plan);
int crv_posterior_exec(crv_post_plan* plan, crv_grid* mean[3], crv_grid** data, crv_grid* err_corr,
                       crv_grid* cov[6]);
int crv_post_plan_stats(crv_post_plan* plan, crv_post_stats* stats);
/* The inversion sequence of the Crava constructor (src/crava.cpp:205-300):
 * the FFTs of the data and the prior means (fftInPlace, crava.cpp:263-271),
 * computePostMeanResidAndFFTCov (crava.cpp:975-1285), nreal realizations of
 * simulate with the counter noise stream (crava.cpp:1550-1734) and
 * invFFTCovGrids (seismicparametersholder.cpp:310-322), in one call on a
 * separable-prior plan.  In: mean[3] prior means and data[ntheta] cubes,
 * real, or, when mean_src / data_src are given, those grids (left
 * unchanged; mean and data then only receive the results).  Out: mean[3]
 * posterior means (real), data the residual spectra,
 * cov[6] the posterior covariance (real), sims (a contiguous batch of
 * 3 nreal grids, may be NULL when nreal = 0) the realizations with the
 * posterior means added.  The results equal crv_fft_r2c of the inputs,
 * crv_posterior_exec, crv_fft_c2r of the means, crv_simulate_batch and
 * crv_fft_c2r of the covariances, bit for bit, which is also the order of
 * the work when chunk_rows <= 0; with 0 < chunk_rows < nyp each chunk of
 * chunk_rows y rows goes through the z transforms, the posterior and the
 * noise in turn (smaller working sets, but smaller launches).  One
 * exception: with chunk_rows <= 0 on an FP32 refine plan (separable prior,
 * up to 6 angles), the realizations outside the inverted band scale the noise by
 * sqrt(lam) times the plane's factor of the prior, where
 * crv_simulate_batch factors the stored FP32 product lam S: those values
 * differ by rounding (about 1e-7 relative).  And on such plans whose nyp has
 * a two-stage kernel the covariances' inverse transform runs its y pass on
 * the inverted band's planes first and its z pass after it, generating the
 * other planes from the y transform of |cxy| (the sequence does z, then y):
 * the covariances then differ from the sequence's by rounding (below 1e-6
 * relative RMS).  Launches only (capturable); statistics through
 * crv_post_plan_stats.                                                      */
int crv_invert(crv_post_plan* plan, crv_grid* mean[3], crv_grid** data, crv_grid* const* mean_src,
               crv_grid* const* data_src, crv_grid* cov[6], unsigned long long seed,
               unsigned long long realization0, int nreal, crv_grid** sims, int chunk_rows);

/* ---- simulation ---------------------------------------------------------
 * crv_simulate replaces one realization of Crava::simulate
 * (src/crava.cpp:1550-1734, loop body 1599-1708): complex white noise as
 * FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1191-1240) for seed0, seed1,
 * seed2 in that order; per wavenumber the 3x3 posterior covariance (upper
 * triangle from cov[], lower = conjugate) is Cholesky factored
 * (lib_matrCholCpx) and the noise multiplied by L (lib_matrProdCholVec);
 * on failure the noise is zeroed; the three grids are inverse transformed
 * with PARAMETER scaling and the real posterior means added.
 *
 * Noise sources:
 *   CRV_NOISE_REFERENCE  the reference RandomGen stream (32-bit LCG
 *                        69069*s+1 and Marsaglia-Bray normals,
 *                        libs/lib/random.cpp:38-119), continued from
 *                        *lcg_state, which is updated on return;
 *   CRV_NOISE_COUNTER    a counter-based generator (Philox-4x32-10 with
 *                        Box-Muller), keyed by (seed, realization); same
 *                        Hermitian structure as the reference fill;
 *   CRV_NOISE_HOST       three complex noise grids supplied by the caller
 *                        (host arrays, csize complex each, grid precision). */
#define CRV_NOISE_REFERENCE 0
#define CRV_NOISE_COUNTER   1
#define CRV_NOISE_HOST      2

typedef struct crv_noise_spec {
  int                mode;
  unsigned int*      lcg_state;    /* CRV_NOISE_REFERENCE (in/out)            */
  unsigned long long seed;         /* CRV_NOISE_COUNTER                       */
  unsigned long long realization;  /* CRV_NOISE_COUNTER                       */
  const void*        host_noise[3];/* CRV_NOISE_HOST                          */
} crv_noise_spec;

/* crv_invert with the realizations' noise chosen by noise (crv_noise_spec
 * above): CRV_NOISE_COUNTER (seed, realization: crv_invert itself) or
 * CRV_NOISE_REFERENCE, the reference's RandomGen stream from
 * *noise->lcg_state (updated), drawn as Crava::simulate draws it
 * (crava.cpp:1598-1603: seed0, seed1, seed2 of realization 0, then of
 * realization 1, ...).  With the reference stream the 3 nreal grids'
 * normals are decoded first (crv_fill_complex_noise's decode; it
 * synchronises the stream, so this form is not capturable), and the
 * realizations equal crv_simulate's with the same state on the same
 * covariances and means (fp64_chol 2 on FP32 refine plans, 0 otherwise),
 * bit for bit, but for the out-of-band rounding exception stated for
 * crv_invert.  CRV_NOISE_HOST is not accepted here.                       */
int crv_invert_noise(crv_post_plan* plan, crv_grid* mean[3], crv_grid** data, crv_grid* const* mean_src,
                     crv_grid* const* data_src, crv_grid* cov[6], const crv_noise_spec* noise, int nreal,
                     crv_grid** sims, int chunk_rows);

/* cov: posterior covariance spectra (aa, bb, rr, ab, ar, br), transformed.
 * post_mean: posterior means in real space (untransformed) or NULL.
 * out[3]: destination grids (alpha, beta, rho realization), real on return.
 * fp64_chol (FP32 grids): 0 = factor the 3x3 matrices in float with
 * lib_matrCholCpx's rules, as the reference does; 1 = in FP64 with one
 * Newton refinement step; 2 = in float, and in FP64 with refinement only
 * where the float factorisation is rejected or its pivot ratio squared
 * exceeds 1e3 (the precision of crv_invert's realizations on refine plans). */
int crv_simulate(crv_ctx* ctx, crv_grid* cov[6], crv_grid* post_mean[3],
                 const crv_noise_spec* noise, int fp64_chol, crv_grid* out[3]);

/* FFTGrid::fillInComplexNoise alone (src/fftgrid.cpp:1191-1240): fills g
 * (becomes transformed, PARAMETER) with the reference noise stream from
 * *lcg_state (updated).  Used to check the stream against the reference.   */
int crv_fill_complex_noise(crv_grid* g, unsigned int* lcg_state);

/* Number of N(0,1) draws FFTGrid::fillInComplexNoise makes on one grid of
 * this padded size (a function of nxp, nyp, nzp only). */
long long crv_noise_draws(int nxp, int nyp, int nzp);

/* The reference stream's decode (used by crv_fill_complex_noise,
 * crv_simulate and crv_invert_noise) evaluates RandomGen::rnorm01
 * (libs/lib/random.cpp:38-111) on the device; a normal whose accept/reject
 * decision (the exp of g(x), the log of the polar branch) or float-rounded
 * value lies within a guard margin of where the device and the host libm
 * could disagree is evaluated again on the host, so the decoded stream
 * equals the host's sequential one.  No reference counterpart: a test hook.
 * scale > 0 multiplies the margins (1 by default; larger sends more normals
 * to the host), 0 keeps the current scale; force_table != 0 makes the next
 * decode put every normal near a margin into its host table and decode
 * again (the path a differing decision takes).  info (may be NULL)
 * receives the last decode's fix passes, host-evaluated normals, values
 * patched from the host and decode runs. */
int crv_noise_guard(crv_ctx* ctx, double scale, int force_table, long long info[4]);

/* ---- inputs of the path generated on the GPU ----------------------------
 * crv_fill_prior_spectra: the six prior covariance spectra that
 * SeismicParametersHolder::FFTCovGrids + getParameterCovariance yield for
 * zero correlation gradients (seismicparametersholder.cpp:198-222, 299-322,
 * 408-426): cov[q] = Sigma0[a][b] |cxy[kx,ky] ct[kz]| (im 0), transformed,
 * COVARIANCE.  cxy is cnxp*nyp doubles (row ky), ct is nzp doubles.        */
int crv_fill_prior_spectra(crv_ctx* ctx, const double* sigma0, const double* cxy, const double* ct,
                           crv_grid* cov[6]);
/* crv_fill_prior_spectra without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  cxy and ct are staged in a fixed device array of 4 MiB per
 * GPU (kept until the process ends) instead of a per-call allocation; a grid
 * with cnxp*nyp + nzp above 524288 is refused with CRV_E_UNSUPPORTED.  The
 * results equal crv_fill_prior_spectra's bit for bit.  Every call on one GPU
 * shares that array: calls on different streams of the same GPU that
 * overlap in time race (calls on one context's stream are ordered).        */
int crv_fill_prior_spectra_noalloc(crv_ctx* ctx, const double* sigma0, const double* cxy, const double* ct,
                                   crv_grid* cov[6]);
/* Real grid set to value in columns i < nxp, 0 in the pad (untransformed).
 * FFTGrid::fillInConstant (fftgrid.cpp:875-892) with a zero pad.           */
int crv_grid_fill(crv_grid* g, double value);
/* Crava::computeSyntSeismic in the Fourier domain (crava.cpp:1767-1894):
 * data[l](k) = kwd[kz][l] * sum_m A[l][m] model[m](k), with kwd as in
 * crv_post_params (kD * conj(W_l)); model transformed; data become
 * transformed DATA grids.                                                   */
int crv_forward_model(crv_ctx* ctx, int ntheta, const float* A, const double* kwd, crv_grid* model[3],
                      crv_grid** data);
/* Adds N(0, sigma^2) white noise (counter stream keyed by seed, stream) to
 * the columns i < nxp of a real grid (the seismic noise of the synthetic
 * surveys).                                                                 */
int crv_add_white_noise(crv_grid* g, double sigma, unsigned long long seed, unsigned long long stream);
/* Sum, sum of squares and count of the non-zero samples of the un-padded box
 * (i < nx, j < ny, k < nz): Crava::computeDataVariance (crava.cpp:407-449)
 * is sum2 / count.                                                          */
int crv_grid_moments(crv_grid* g, double* sum, double* sum2, long long* count);

/* ---- rock physics (the model behind the learned surrogate) ---------------
 * crv_rock_forward replaces, per sample, the rplib chain used to generate
 * facies-conditional elastic samples (DistributionsRockDEM ->
 * DEMTools::CalcEffectiveBulkAndShearModulus -> DEM::CalcEffectiveModulus ->
 * OrdDiffEqSolver::Ode45, then RockGassmann::ComputeSeismicAndElasticParams;
 * rplib/dem.cpp:39-189, rplib/orddiffeqsolver.cpp:17-167,
 * rplib/demmodelling.cpp:29-55,178-236,274-354, rplib/rockgassmann.cpp:79-97):
 * Hill-mixed quartz/clay mineral, DEM dry frame with empty pores, Batzle-Wang
 * brine and gas (Reuss mix by water saturation), Gassmann.  One thread per
 * sample in FP64.  d_in: n x 5 doubles (porosity, clay fraction, water
 * saturation, pore aspect ratio, effective pressure MPa); d_out: n x 3
 * (ln vp, ln vs in m/s, ln rho in kg/m3) as float (out_fp32) or double;
 * d_status (optional): 0 ok, 1 DEM integration failed.  Device pointers.    */
typedef struct crv_rock_params {
  double k_quartz, g_quartz, rho_quartz;   /* kPa, kPa, g/cm3 */
  double k_clay, g_clay, rho_clay;
  double temperature;                      /* deg C */
  double salinity;                         /* weight fraction */
  double gas_gravity;
} crv_rock_params;

int crv_rock_forward(crv_ctx* ctx, const crv_rock_params* p, long long n, const double* d_in, void* d_out,
                     int out_fp32, int* d_status);

/* ---- rock-physics surrogate ---------------------------------------------
 * A 5 -> 128 -> 128 -> 128 -> 3 tanh MLP standing in for crv_rock_forward
 * when facies-conditional elastic samples are generated in bulk (the role
 * rplib plays for faciesprob: rplib/distributionsrockdem.cpp:125-156,
 * src/faciesprob.cpp:2494-2559).  Training and inference run on BF16 matrix
 * cores with FP32 accumulation; FP32 master weights and Adam moments.
 * Inputs are crv_rock_forward's (porosity, clay fraction, Sw, aspect ratio,
 * effective pressure MPa) as float; features with bit f of in_log_mask set
 * are log10'd; then x~ = (x - in_lo) * in_scale - 1.  Outputs are the
 * normalised (ln vp, ln vs, ln rho): y = y~ * out_std + out_mean.  The loss
 * is the mean squared error in normalised units.
 * Parameters (crv_mlp_param_count() floats): W1a [128][16] (bias in column
 * 5), W2a and W3a [128][144] (bias in column 128), W4a [32][144] (rows 0..2
 * used, bias in column 128), row-major, concatenated; padding stays zero.
 * Device pointers: d_x n x 5, d_t / d_y n x 3 floats.                      */
typedef struct crv_mlp crv_mlp;
typedef struct crv_mlp_norm {
  float in_lo[5], in_scale[5];
  int   in_log_mask;
  float out_mean[3], out_std[3];
} crv_mlp_norm;
long long crv_mlp_param_count(void);
int crv_mlp_create(crv_ctx* ctx, const crv_mlp_norm* norm, long long max_batch, crv_mlp** out);
int crv_mlp_destroy(crv_mlp* h);
/* host params -> device (Adam moments reset) / device -> host */
int crv_mlp_set_params(crv_mlp* h, const float* params);
int crv_mlp_get_params(crv_mlp* h, float* params);
/* gradient of the loss at the current parameters (host copy; either
 * output may be NULL); no update */
int crv_mlp_gradients(crv_mlp* h, const float* d_x, const float* d_t, long long n, float* grad, float* loss);
/* one Adam step (beta 0.9 / 0.999, eps 1e-8) on a batch; step counts from 1 */
int crv_mlp_train_step(crv_mlp* h, const float* d_x, const float* d_t, long long n, float lr, int step,
                       float* loss);
int crv_mlp_predict(crv_mlp* h, const float* d_x, long long n, float* d_y);
/* matrix-core flops of the last forward and backward launches (padded shapes) */
int crv_mlp_flops(crv_mlp* h, double* fwd, double* bwd);
/* Error gate of the surrogate with rplib as fallback: every sample whose
 * cell of the error table predicts an error above thresh is recomputed with
 * the crv_rock_forward chain and its outputs overwritten (NaN when the DEM
 * integration fails).  The table has nb^5 cells over the normalised inputs
 * (cell index sum_f b_f nb^(4-f), b_f = floor(nb (x~_f + 1) / 2) clamped),
 * a cell's value being the largest error seen there (+inf when never
 * seen).  d_x n x 5, d_y n x 3 floats, d_table nb^5 floats, device memory. */
int crv_rock_refine(crv_ctx* ctx, const crv_rock_params* p, long long n, const float* d_x, float* d_y,
                    const crv_mlp_norm* norm, const float* d_table, int nb, float thresh, long long* n_refined);

/* ---- facies probabilities -------------------------------------------------
 * FaciesProb's density estimate and classification (src/faciesprob.cpp:
 * 544-930): crv_facies_density bins n facies-conditional samples (d_samples
 * n x 3 floats: ln vp, ln vs, ln rho) on the box, smooths the histogram by
 * a Gaussian of covariance smooth_cov (3 x 3, log units: posterior
 * uncertainty plus kernel bandwidth) through the library's transforms, and
 * normalises it to a density on `density` (an FP32 grid; x bins = nxp along
 * ln vp, y along ln vs, z along ln rho; the transform is circular, so the
 * box should extend beyond the samples by a few smoothing widths).
 * crv_facies_prob gives every voxel of the real posterior means (FP32)
 * P(f | m) = prior_f p_f(m) / sum_g prior_g p_g(m), p_f by trilinear
 * interpolation of the nf (<= 8) densities.                                */
typedef struct crv_facies_box {
  double lo[3], hi[3];
} crv_facies_box;
int crv_facies_density(crv_ctx* ctx, const float* d_samples, long long n, const crv_facies_box* box,
                       const double* smooth_cov, crv_grid* density);
int crv_facies_prob(crv_ctx* ctx, int nf, crv_grid* const* density, const double* prior,
                    const crv_facies_box* box, crv_grid* mean[3], crv_grid** prob);
/* crv_facies_prob with FaciesProb's undefined class (calculateFaciesProb,
 * src/faciesprob.cpp:2008-2056, p_undefined = ModelSettings p_undef_, 0.01
 * by default, modelsettings.cpp:119): the denominator gains
 * p_undefined / (number of bins) for densities that sum to 1 over the box
 * (the reference's PosteriorElasticPDF3D mass per bin; in general
 * p_undefined times the mean bin value of density[0], the same share for
 * unit-integral densities), so voxels far from every facies get
 * probabilities near 0 instead of ratios of vanishing densities.
 * prob_undef (may be NULL) receives the undefined class's probability
 * (faciesProbUndef_).  p_undefined = 0 is crv_facies_prob.                 */
int crv_facies_prob_undef(crv_ctx* ctx, int nf, crv_grid* const* density, const double* prior,
                          double p_undefined, const crv_facies_box* box, crv_grid* mean[3], crv_grid** prob,
                          crv_grid* prob_undef);
/* The density with a choice of smoothing kernel:
 *   CRV_SMOOTH_FOURIER  crv_facies_density (the Gaussian's Fourier transform,
 *                       clamped at 0, normalised to unit integral);
 *   CRV_SMOOTH_SAMPLED  PosteriorElasticPDF3D's estimate (reference
 *                       src/posteriorelasticpdf3d.cpp:12-131, 369-420): the
 *                       histogram divided by n, circularly convolved with the
 *                       Gaussian exp(-x^T S^-1 x / 2) sampled at the bins'
 *                       cyclic lags and normalised to sum 1 -- a probability
 *                       mass per bin that sums to 1 (densities of one box
 *                       differ from it by the bin volume only, so
 *                       crv_facies_prob takes either).  Samples outside the
 *                       box are skipped (the reference does not check).  */
#define CRV_SMOOTH_FOURIER 0
#define CRV_SMOOTH_SAMPLED 1
int crv_facies_density_kernel(crv_ctx* ctx, const float* d_samples, long long n, const crv_facies_box* box,
                              const double* smooth_cov, int kernel, crv_grid* density);

/* ---- padding (src/fftgrid.cpp:2857-2885, src/modelgeneral.cpp:616-622) -- */
/* Smallest 2^a 3^b 5^c 7^d >= leastint (FFTGrid::findClosestFactorableNumber). */
int crv_find_closest_factorable(int leastint);
/* ModelGeneral::setPaddingSize: findClosestFactorable(ceil(n*(1+pad_fac))). */
int crv_padding_size(int n, double pad_fac);

/* ---- prior temporal correlation ------------------------------------------
 * SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef
 * (seismicparametersholder.cpp:430-498): the prior temporal correlation
 * corr (ncorr lags, corr[0] != 0) laid out circularly on nzp points (lag
 * min(k, nzp - k), 0 beyond ncorr), its spectrum replaced by its modulus and
 * set to 0 for frequency indices <= low_int_cut (ModelGeneral's lowIntCut =
 * floor(lowCut nzp dt / 1000), modelgeneral.cpp:4081), transformed back and
 * divided by its lag-0 value.  out: nzp floats.  CRV_E_INVALID when the
 * unnormalised lag-0 value is not above 1e-5 (the reference aborts: too few
 * layers).  Host only (no device work).                                  */
int crv_circ_corr_t(const float* corr, int ncorr, int nzp, int low_int_cut, float* out);

/* ---- output (next to the path) ------------------------------------------
 * Derived parameters of ParameterOutput::writeParameters
 * (src/parameteroutput.cpp:14-420) from the real log-parameter grids
 * alpha = ln vp, beta = ln vs, rho = ln rho, evaluated in double and stored
 * in out's precision (out may alias an input):
 *   CRV_OUT_EXP          exp(alpha), RMISSING kept (FFTGrid::expTransf,
 *                        src/fftgrid.cpp:1726-1743; vp, vs or rho output)
 *   CRV_OUT_AI / SI      exp(alpha + rho) / exp(beta + rho)
 *   CRV_OUT_VPVS         exp(alpha - beta)
 *   CRV_OUT_POISSON      0.5 (q - 2) / (q - 1), q = exp(2 (alpha - beta))
 *   CRV_OUT_LAME_MU      exp(rho + 2 beta - 13.81551)
 *   CRV_OUT_LAME_LAMBDA  exp(rho) (exp(2 alpha - 13.81551) - 2 exp(2 beta - 13.81551))
 *   CRV_OUT_MU_RHO       exp(2 (beta + rho) - 13.81551)
 *   CRV_OUT_LAMBDA_RHO   exp(2 (alpha + rho) - 13.81551) - 2 exp(2 (beta + rho) - 13.81551) */
#define CRV_OUT_EXP         0
#define CRV_OUT_AI          1
#define CRV_OUT_SI          2
#define CRV_OUT_VPVS        3
#define CRV_OUT_POISSON     4
#define CRV_OUT_LAME_MU     5
#define CRV_OUT_LAME_LAMBDA 6
#define CRV_OUT_MU_RHO      7
#define CRV_OUT_LAMBDA_RHO  8
int crv_grid_derived(crv_ctx* ctx, int kind, crv_grid* alpha, crv_grid* beta, crv_grid* rho, crv_grid* out);
/* CRAVA's binary grid format (FFTGrid::writeCravaFile / readCravaFile,
 * src/fftgrid.cpp:2411-2498): the line "crava_fftgrid_binary", the simbox
 * geometry below in this order, rnxp, nyp, nzp, then the rsize floats of the
 * padded real grid; numbers big-endian IEEE.  Write and read move the grid
 * through pinned staging with the copies overlapping the file I/O.  FP32
 * real grids; read checks the dimensions and returns the geometry in sb
 * (may be NULL).  crv_crava_header encodes the header into buf (NULL: the
 * length only) and returns its length.                                    */
typedef struct {
  double x0, y0, dx, dy;
  int nx, ny;
  double il0, xl0, il_step_x, il_step_y, xl_step_x, xl_step_y, angle;
} crv_simbox_info;
int crv_crava_header(const crv_simbox_info* sb, int rnxp, int nyp, int nzp, unsigned char* buf, size_t len);
int crv_grid_write_crava(crv_grid* g, const char* path, const crv_simbox_info* sb);
int crv_grid_read_crava(crv_grid* g, const char* path, crv_simbox_info* sb);
/* STORM grid files (FFTGrid::writeStormFile, src/fftgrid.cpp:2117-2176; the
 * header is Simbox::getStormHeader, src/simbox.cpp:408-458):
 *   storm_petro_binary            (storm_petro_ascii for text files)
 *   0 <cubetype> -99999.000000
 *   FFTGrid
 *   <xmin> <lx> <ymin> <ly> <top_name> <bot_name> 0.0 0.0
 *                                 (flat: <xmin> <lx> <ymin> <ly> 0.0 <lz> 0.0 0.0)
 *   <lz> <angle in degrees>
 *   (empty line)
 *   <nx> <ny> <nz>
 * reals in fixed notation with 6 decimals; then the nx ny nz values of the
 * real grid, x fastest (the padded nxp nyp nzp with padding != 0), as
 * big-endian floats, or as text (6 decimals fixed, or scientific), one line
 * per x row; then the line "0".  FP32 real grids.  The caller appends the
 * reference's suffix (".storm" binary, ".dat" text) to path.
 * crv_storm_header encodes the header into buf (NULL: the length only) and
 * returns its length.                                                      */
typedef struct {
  double xmin, lx, ymin, ly, lz, angle;  /* angle in radians */
  const char* top_name;                  /* surfaces named in the header (flat = 0) */
  const char* bot_name;
  int flat;
} crv_storm_info;
int crv_storm_header(const crv_storm_info* si, int cubetype, int nx, int ny, int nz, int ascii, char* buf,
                     size_t len);
int crv_grid_write_storm(crv_grid* g, const char* path, const crv_storm_info* si, int ascii, int padding,
                         int scientific);

/* ---- prior and error correlation grids (general prior) ------------------
 * SeismicParametersHolder::initializeCorrelations (src/seismicparameters-
 * holder.cpp:197-216): the six real prior covariance grids cov[6] (aa, bb,
 * rr, ab, ar, br) as FFTGrid::fillInParamCorr (src/fftgrid.cpp:1072-1117)
 * fills them -- value(i, j, k) = corrXY(i, j) times the temporal
 * correlation circ_corr_t (nzp values, crv_circ_corr_t) at the cyclic,
 * linearly interpolated shift k + ci grad_i + cj grad_j (ci, cj the lateral
 * offsets nearest zero), RMISSING in the padding -- each multiplied by
 * float(sigma0[q]).  corr_xy: the lateral correlation surface, nxp * nyp
 * values (i + nxp j).  Float arithmetic as the reference (bit for bit).  */
int crv_fill_param_corr(crv_grid* cov[6], const double* corr_xy, const float* circ_corr_t, float grad_i,
                        float grad_j, const double sigma0[6]);
/* FFTGrid::fillInErrCorr (src/fftgrid.cpp:1018-1069): the real error
 * correlation grid, corrXY(i, j) times the weight of plane 0 at the shifted
 * position s = k + ci grad_i + cj grad_j: within one sample of plane 0 (or
 * of plane nzp, cyclically) 1 - (s minus s truncated toward zero), as the
 * reference's int() gives it (so 1 + |s| for s in (-1, 0)); 0 elsewhere;
 * RMISSING in the padding.                                                */
int crv_fill_err_corr(crv_grid* err, const double* corr_xy, float grad_i, float grad_j);

/* ---- prediction kriging to wells ---------------------------------------
 * Crava::doPostKriging (src/crava.cpp:1736-1765): the posterior means
 * mean[3] (real, updated in place) kriged to well data with the separable
 * covariances of the six real posterior covariance grids cov[6] (aa, bb,
 * rr, ab, ar, br), as CKrigingAdmin::KrigAll does it (src/krigingadmin.cpp:
 * block size and data neighbourhood from the covariance ranges, per block
 * a kriging system over up to data_target (+-10 %) data, every cell
 * updated with k(cell) . x).  pts: the blocked-log points, log Vp, log Vs,
 * log rho per cell of the simbox (0 <= i < nx ...), CRV_WELL_MISSING where
 * a value is absent; points of one cell are averaged.  The block systems
 * are solved on the host (FP64 Cholesky, where the reference calls LAPACK
 * posv); the cells are updated on the device.  smooth != 0 (the
 * smooth-kriged-parameters option, off by default in the reference) is not
 * built and returns CRV_E_UNSUPPORTED.  Synchronises.                     */
#define CRV_WELL_MISSING (-99999.0f)
typedef struct crv_well_point {
  int   i, j, k;
  float alpha, beta, rho;
} crv_well_point;
typedef struct crv_krig_stats {
  int       block[3];          /* dxBlock, dyBlock, dzBlock               */
  int       block_ext[3];      /* data neighbourhood beyond the block     */
  int       range[3];          /* covariance ranges used (cells)          */
  int       data_points;       /* after merging points of one cell        */
  int       blocks;            /* blocks per parameter                    */
  int       empty_blocks[3];   /* blocks without data, per parameter      */
  int       max_block_data[3]; /* largest system, per parameter           */
  long long cells_updated[3];
  long long cells_missing[3];  /* trend RMISSING: left as it is           */
} crv_krig_stats;
int crv_krig_post(crv_grid* mean[3], crv_grid* cov[6], const crv_well_point* pts, int npts, int data_target,
                  int smooth, crv_krig_stats* stats);

#ifdef __cplusplus
}
#endif
#endif /* CRAVA_HIP_H */
