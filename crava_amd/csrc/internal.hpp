// Internal declarations shared by the HIP translation units of libcrava_hip.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstddef>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/crava_hip.h"

// ---------------------------------------------------------------------------
// complex helpers (float2 / double2 as interleaved re, im)
template <typename T> struct cplx_of;
template <> struct cplx_of<float>  { using type = float2; };
template <> struct cplx_of<double> { using type = double2; };
template <typename T> using cplx_t = typename cplx_of<T>::type;

template <typename C> __host__ __device__ inline C cmake(decltype(C::x) a, decltype(C::x) b) { C r; r.x = a; r.y = b; return r; }
template <typename C> __host__ __device__ inline C cadd(C a, C b) { return cmake<C>(a.x + b.x, a.y + b.y); }
template <typename C> __host__ __device__ inline C csub(C a, C b) { return cmake<C>(a.x - b.x, a.y - b.y); }
template <typename C> __host__ __device__ inline C cmul(C a, C b) { return cmake<C>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <typename C> __host__ __device__ inline C cmulc(C a, C b) { return cmake<C>(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
template <typename C> __host__ __device__ inline C cconj(C a) { return cmake<C>(a.x, -a.y); }
template <typename C> __host__ __device__ inline C cscale(C a, decltype(C::x) s) { return cmake<C>(a.x * s, a.y * s); }

// ---------------------------------------------------------------------------
// Unsigned division by a runtime constant d (numerators below 2^31):
// q = (umulhi(n, m) + n) >> s with s = ceil(log2 d), m = 2^32 (2^s - d) / d + 1.
struct FastDiv {
  unsigned m; int s; int d;
};
inline FastDiv make_fastdiv(int d) {
  FastDiv f; f.d = d; f.s = 0;
  while ((1u << f.s) < (unsigned)d) ++f.s;
  unsigned long long num = (1ull << 32) * (unsigned long long)((1ull << f.s) - d);
  f.m = (unsigned)(num / (unsigned)d + 1);
  return f;
}
__device__ __forceinline__ int fdiv(int n, FastDiv f) {
  return (int)((__umulhi((unsigned)n, f.m) + (unsigned)n) >> f.s);
}

// ---------------------------------------------------------------------------
// 1D transform plan: Stockham stages over LDS; the twiddles of all stages
// of one axis are a contiguous range [tw0, tw0 + twn) of the plan's table,
// copied to LDS by each block.
constexpr int kMaxStages = 16;
struct Fft1D {
  int L;                 // transform length
  int nst;               // number of radix stages
  int R[kMaxStages];     // radix of each stage
  int Ns[kMaxStages];    // product of the radices before the stage
  int tw[kMaxStages];    // offset of the stage's twiddles in the table
  int tw0, twn;          // the axis's range of the table
};

struct FftPlan {
  int nxp = 0, nyp = 0, nzp = 0, prec = 0;
  bool xeven = true;     // nxp even: x transform via a length nxp/2 complex FFT
  Fft1D fx{}, fy{}, fz{};
  int post_off = 0;      // offset of exp(-2 pi i k / nxp), k = 0..nxp/2
  void* tw = nullptr;    // device twiddle table (float2 or double2)
  // launch shapes (chosen by crv_plan_tune or defaults)
  int rows_per_block = 0;
  int tx_y = 0, tx_z = 0;
  int chunk_mb = 0;      // x/y pass interleave chunk (MB of grid), 0 = off
  int batch_mb = 0;      // batched transforms: MB of grids per launch, 0 = 256
  int order = 0;         // CRV_TUNE_ORDER bits: passes that walk their tiles backwards
  bool half[3] = {false, false, false};  // x, y, z: general kernel with half the elements per thread
  bool wide[3] = {false, false, false};  // y, z (index 1, 2): 512-thread column kernel
  int rev[3] = {0, 0, 0};  // radix order per axis (0 default, 1 reversed)
  // reference noise: per-row prefix of draws (nyp*nzp+1 int64), lazily made
  long long* noise_prefix = nullptr;
  long long noise_draws = 0;
};

struct ProfEvent {
  const char* name;
  hipEvent_t a, b;
};

struct crv_ctx {
  int tune[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};   // CRV_TUNE_* overrides (0 = default)
  bool prof_on = false;
  std::vector<ProfEvent> prof;
  int device = 0;
  int num_cu = 256;      // compute units of the device (persistent grids)
  int post_blocks_per_cu = 0;  // posterior grid size (0: occupancy limit)
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::string last_error;
  std::map<std::tuple<int, int, int, int>, FftPlan*> plans;
  // small device scratch: statistics accumulators
  void* dstats = nullptr;
  // reference-noise scratch (grown on demand)
  void*  lcg_scratch = nullptr;
  size_t lcg_scratch_bytes = 0;
  double* normals = nullptr;
  size_t  normals_cap = 0;
  // reference-noise decode: scale of the guard margins (1 by default) and
  // what the last decode did: fix passes, normals resolved on the host
  // (decisions, values), decode runs
  double noise_guard = 1.0;
  bool noise_force_table = false;
  long long noise_info[4] = {0, 0, 0, 0};
};

struct crv_grid {
  crv_ctx* ctx = nullptr;
  int nx = 0, ny = 0, nz = 0, nxp = 0, nyp = 0, nzp = 0, cnxp = 0, rnxp = 0;
  int precision = CRV_FP32;
  int cubetype = CRV_CTMISSING;
  bool transformed = false;
  size_t rsize = 0, csize = 0;
  void* d = nullptr;
  int* shared = nullptr;   // batch grids: count of live grids in the allocation
  void* base = nullptr;    // batch grids: start of the allocation
  size_t elem_bytes() const { return precision == CRV_FP64 ? 8 : 4; }
};

// error helpers -------------------------------------------------------------
int crv_set_error(crv_ctx* ctx, int code, const std::string& msg);
#define CRV_HIP_CHECK(ctx, expr)                                                     \
  do {                                                                               \
    hipError_t e__ = (expr);                                                         \
    if (e__ != hipSuccess)                                                           \
      return crv_set_error((ctx), CRV_E_HIP, std::string(#expr) + ": " +             \
                                                 hipGetErrorString(e__));            \
  } while (0)

// Brackets the launches of its scope with events when profiling is on.
struct ProfScope {
  crv_ctx* c;
  const char* name;
  hipEvent_t a = nullptr, b = nullptr;
  ProfScope(crv_ctx* ctx, const char* n) : c(ctx), name(n) {
    if (c->prof_on && hipEventCreate(&a) == hipSuccess && hipEventCreate(&b) == hipSuccess)
      (void)hipEventRecord(a, c->stream);
  }
  ~ProfScope() {
    if (a && b) {
      (void)hipEventRecord(b, c->stream);
      c->prof.push_back(ProfEvent{name, a, b});
    }
  }
};

// plan cache (fft.hip) -------------------------------------------------------
int crv_get_plan(crv_ctx* ctx, int nxp, int nyp, int nzp, int prec, FftPlan** out);
void crv_free_plan(FftPlan* p);
// (re)select the launch shapes of a plan from the context's tuning values
int crv_plan_shapes(crv_ctx* ctx, FftPlan* p);

// transform launchers (fft.hip) ------------------------------------------------
// Forward R2C of src (real view) into dst (may alias), scale applied on load.
int crv_launch_r2c(crv_ctx* ctx, FftPlan* plan, void* dst, const void* src, double scale);
// Inverse C2R of buf in place (complex view in, real view out); result written
// to dst (may alias buf) times scale, plus add (real grid) if non-null.
int crv_launch_c2r(crv_ctx* ctx, FftPlan* plan, void* buf, void* dst, double scale, const void* add);
int crv_launch_c2r_batch(crv_ctx* ctx, FftPlan* plan, void* buf, int nbatch, long long bstride, double scale,
                         const void* const* add, int nadd);
int crv_launch_r2c_batch(crv_ctx* ctx, FftPlan* plan, void* dst, const void* src, int nbatch, long long bstride,
                         double scale);
// grids per launch of a batched transform of nbatch grids (CRV_TUNE_BATCH_MB)
int crv_sub_batch(const FftPlan* p, int nbatch);
// Covariance spectra that the z inverse pass generates instead of reading
// (FP32, separable prior, crv_invert): the grids of a launch are the
// covariances q0 + b.  Outside the band (band[k] == 0) the element of column
// i, y row j is lam[j cnxp + i] S[6 k + q].  With rows (eigen-basis plans)
// the band's planes are generated too: plane k's row bidx[k] of covariance
// q, rows + (q nband + bidx[k]) rw (rw = 4 + 4 n), holds S_q, cond(D),
// l_max, l_min, then for each of the n eigenvalues l_p, Re G_pq, Im G_pq, 0,
// and the element is lam S_q - lam^2 sum_p w_p G_pq unless the posterior
// queued the wavenumber for its FP64 solve (same rule), whose value is read;
// noqueue: the plan has shown that no wavenumber can be queued.
struct CovGen {
  const float* lam = nullptr;
  const float* S = nullptr;
  const int* band = nullptr;
  int q0 = 0;
  const int* bidx = nullptr;
  const float* rows = nullptr;
  int nband = 0, rw = 0, n = 0;
  float rx = 0, ry = 0, rz = 0, lowk2 = 0, cond_thr = 0;
  int noqueue = 0;
  // the band planes (nband ints; crv_launch_cov_band_y)
  const int* bplanes = nullptr;
  // with lamy (complex, [nyp][cnxp]) the z inverse's input outside the band
  // is lamy[o cnxp + i] S[6 k + q], for covariances whose y inverse has been
  // done already: lamy is the y inverse transform of lam (the separable
  // prior's plane after its y pass)
  const float2* lamy = nullptr;
};
// The y inverse pass of the band planes only of nbatch covariance grids
// (FP32, eigen-basis records, covariances q0 + b): the elements are generated
// as the eigen-basis posterior's entries (the queued ones read from the grid),
// transformed along y and stored in place; the other planes are left as they
// are.  Needs a two-stage kernel for nyp.
int crv_launch_cov_band_y(crv_ctx* ctx, FftPlan* plan, void* buf, int nbatch, long long bstride, const CovGen& g);
bool crv_cov_band_y_fits(const FftPlan* p);
// LDS of the generating z inverse of a plan with nband band planes of rw
// floats each fits the column kernels' 64 KB
bool crv_cov_gen_fits(const FftPlan* p, int nband, int rw);

// The z pass of crv_invert with the eigen-basis posterior of the means in
// between (FP32 plans with eigen-basis records where no wavenumber can be
// queued for the FP64 solve): data[l], after their x and y forward passes,
// become the residual spectra; mean[m], likewise, the z inverse transforms
// of the posterior mean spectra.  stats[1] receives the largest condition
// estimate (atomic max of the double's bits).
struct ZPost {
  void* data[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  void* mean[3] = {nullptr, nullptr, nullptr};
  int n = 0, nband = 0;
  const float* lam = nullptr;     // [nyp][cnxp] |cxy|
  const float* erec = nullptr;    // eigen-basis plane records
  const int* bidx = nullptr;      // [nzp] row among the band planes, -1 outside
  const int* bplanes = nullptr;   // [nband] the band planes
  unsigned long long* stats = nullptr;
};
// tile width of the fused z pass for a plan with nband band planes, 0 when
// it does not apply (or CRV_TUNE_ZPOST is 1)
int crv_zpost_tx(crv_ctx* ctx, const FftPlan* p, int nband);
int crv_launch_zpost(crv_ctx* ctx, FftPlan* p, const ZPost& z, int tx);

// One pass of a transform over part of a batch of nbatch grids bstride reals
// apart: pass 0 x R2C (rows [o0, o1), row = j + nyp k; reads src, scales),
// 1 y forward (z planes [o0, o1)), 2 z forward (y rows [o0, o1)), 3 z
// inverse, 4 y inverse, 5 x C2R (rows, in place, scales, adds add[b % nadd]).
// gen (pass 3, FP32): the grids are covariances whose spectra are generated
// where CovGen says so.
int crv_launch_pass(crv_ctx* ctx, FftPlan* plan, int pass, void* buf, const void* src, int nbatch, long long bstride,
                    int o0, int o1, double scale, const void* const* add, int nadd, const CovGen* gen = nullptr);
// Noise and Cholesky scaling of nreal realizations over y rows [j0, j1):
// realization r, parameter g written at out0 + (3 r + g) gstride complex
// elements (simulate.hip).  Noise: the counter stream (seed, real0), or with
// normals the decoded reference stream, grid g of realization r from draw
// (3 r + g) noise_draws (the plan's reference-stream prefix must exist).
int crv_launch_sim_rows(crv_ctx* ctx, FftPlan* plan, const void* const cov[6], void* out0, long long gstride,
                        int nreal, unsigned long long seed, unsigned long long real0, int fp64_chol, int j0, int j1,
                        const double* normals = nullptr);

// elementwise (simulate.hip) ---------------------------------------------------
int crv_launch_axpy(crv_ctx* ctx, int prec, void* dst, const void* src, size_t n);
int crv_launch_scale(crv_ctx* ctx, int prec, void* dst, double s, size_t n);
int crv_launch_convert(crv_ctx* ctx, int dprec, void* dst, int sprec, const void* src, size_t n);

// noise (noise.hip) ------------------------------------------------------------
int crv_noise_prefix(crv_ctx* ctx, FftPlan* plan);
// Decode nnormals reference normals from *state into ctx->normals; updates *state.
int crv_reference_normals(crv_ctx* ctx, unsigned int* state, long long nnormals);
// Fill one grid from ctx->normals starting at draw offset `first`.
int crv_launch_fill_noise(crv_ctx* ctx, FftPlan* plan, crv_grid* g, long long first);

// Per-device address of the staging array of crv_fill_prior_spectra_noalloc
// (synth.hip); called by crv_ctx_create with the device current.
int crv_noalloc_setup(int device);

// FFT scale factors by grid type and precision (capi.hip)
double crv_scale_forward(const crv_grid* g);
double crv_scale_inverse(const crv_grid* g);

// Host-side LCG skip-ahead: state after `steps` applications of s <- 69069 s + 1.
unsigned int crv_lcg_skip(unsigned int s, unsigned long long steps);
