// Inputs of the hot path generated on the GPU: prior covariance spectra,
// constant prior means, the linearised forward model and white noise, and
// the data-variance reduction.  They stand in for the SEG-Y cubes and the
// model set-up that CRAVA reads or builds before the inversion:
//  * prior spectra: FFTCovGrids + getParameterCovariance for zero correlation
//    gradients, where the 3D spectrum of corrT(k) corrXY(i, j) Sigma0[a][b]
//    is Sigma0[a][b] cxy(kx, ky) ct(kz) (reference
//    src/seismicparametersholder.cpp:198-222, 299-322, 339-426);
//  * forward model: Crava::computeSyntSeismic in the Fourier domain, data_l =
//    conj(W_l) D1 (A_l . m) (src/crava.cpp:1767-1894);
//  * data variance: Crava::computeDataVariance, mean of d^2 over the
//    un-padded box without exact zeros, float squares summed in double
//    (src/crava.cpp:407-449).
#include "internal.hpp"
#include "noise.hpp"

#include <algorithm>
#include <cstring>
#include <mutex>
#include <vector>

namespace {

constexpr int kThreads = 256;

int nblocks(long long n) { return (int)std::min<long long>((n + kThreads - 1) / kThreads, 65536); }

template <typename T>
__global__ void k_prior_spectra(cplx_t<T>* c0, cplx_t<T>* c1, cplx_t<T>* c2, cplx_t<T>* c3, cplx_t<T>* c4,
                                cplx_t<T>* c5, const double* cxy, const double* ct, double s0, double s1, double s2,
                                double s3, double s4, double s5, int cnxp, int nyp, long long nc) {
  using C = cplx_t<T>;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const long long plane = (long long)cnxp * nyp;
    const int k = (int)(i / plane);
    const int rem = (int)(i - (long long)k * plane);
    const double sc = fabs(cxy[rem] * ct[k]);
    c0[i] = cmake<C>((T)(s0 * sc), (T)0);
    c1[i] = cmake<C>((T)(s1 * sc), (T)0);
    c2[i] = cmake<C>((T)(s2 * sc), (T)0);
    c3[i] = cmake<C>((T)(s3 * sc), (T)0);
    c4[i] = cmake<C>((T)(s4 * sc), (T)0);
    c5[i] = cmake<C>((T)(s5 * sc), (T)0);
  }
}

template <typename T>
__global__ void k_fill_real(T* g, T v, int nxp, int rnxp, long long rsize) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < rsize; i += (long long)gridDim.x * blockDim.x)
    g[i] = ((int)(i % rnxp) < nxp) ? v : (T)0;
}

struct FwdArgs {
  const void* m[3];
  void* d[8];
  const double2* K;   // [nzp][n][3] = kwd[k][l] * A[l][m]
  int n, cnxp, nyp;
  long long nc;
};

template <typename T>
__global__ void k_forward(FwdArgs a) {
  using C = cplx_t<T>;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < a.nc; i += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(i / ((long long)a.cnxp * a.nyp));
    const C m0 = ((const C*)a.m[0])[i], m1 = ((const C*)a.m[1])[i], m2 = ((const C*)a.m[2])[i];
    for (int l = 0; l < a.n; ++l) {
      const double2* K = a.K + ((size_t)k * a.n + l) * 3;
      double re = K[0].x * m0.x - K[0].y * m0.y + K[1].x * m1.x - K[1].y * m1.y + K[2].x * m2.x - K[2].y * m2.y;
      double im = K[0].x * m0.y + K[0].y * m0.x + K[1].x * m1.y + K[1].y * m1.x + K[2].x * m2.y + K[2].y * m2.x;
      ((C*)a.d[l])[i] = cmake<C>((T)re, (T)im);
    }
  }
}

template <typename T>
__global__ void k_white_noise(T* g, double sigma, unsigned long long seed, unsigned long long stream, int nxp, int rnxp,
                              long long rsize) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < rsize; i += (long long)gridDim.x * blockDim.x) {
    if ((int)(i % rnxp) >= nxp) continue;
    double n0, n1;
    counter_pair(seed, stream, 7, i, n0, n1);
    g[i] = (T)((double)g[i] + sigma * n0);
  }
}

// per-block partial sums of d^2 (float products, as the reference) and of
// the count of non-zero samples inside the un-padded box
template <typename T>
__global__ void k_moments(const T* g, int nx, int ny, int nz, int rnxp, int nyp, double* psum2, double* psum,
                          unsigned long long* pcnt) {
  __shared__ double s2[kThreads], s1[kThreads];
  __shared__ unsigned long long sc[kThreads];
  double a2 = 0, a1 = 0;
  unsigned long long c = 0;
  const long long n = (long long)nx * ny * nz;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % nx);
    const long long jk = t / nx;
    const int j = (int)(jk % ny), k = (int)(jk / ny);
    const T v = g[(size_t)i + (size_t)rnxp * j + (size_t)rnxp * nyp * k];
    if (v != (T)0) {
      const float vf = (float)v;
      a2 += (sizeof(T) == 4) ? (double)(vf * vf) : (double)v * (double)v;
      a1 += (double)v;
      ++c;
    }
  }
  s2[threadIdx.x] = a2; s1[threadIdx.x] = a1; sc[threadIdx.x] = c;
  __syncthreads();
  for (int o = kThreads / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s2[threadIdx.x] += s2[threadIdx.x + o];
      s1[threadIdx.x] += s1[threadIdx.x + o];
      sc[threadIdx.x] += sc[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    psum2[blockIdx.x] = s2[0];
    psum[blockIdx.x] = s1[0];
    pcnt[blockIdx.x] = sc[0];
  }
}

// Argument checks shared by the two prior-spectra entry points.
// The lateral index of column i as the reference's fills take it: i, or
// i - n for the upper half (the cyclic offset nearest zero, ties to i).
__device__ __forceinline__ int cycle_of(int i, int n) {
  const int c = i - n;
  return (i < -c) ? i : c;
}

constexpr float kRMissing = -99999.0f;  // RMISSING (src/definitions.h)

// SeismicParametersHolder::initializeCorrelations (seismicparametersholder.
// cpp:197-216): FFTGrid::fillInParamCorr (fftgrid.cpp:1072-1117) -- the
// temporal correlation shifted by the lateral offset times the correlation
// gradient, interpolated linearly and cyclically, times the lateral
// correlation -- then each grid scaled by its prior covariance (float
// multiplyByScalar over every element, padding included).  Float
// arithmetic in the reference's order, contraction off.
struct ParamCorrArgs {
  void* g[6];
  float s[6];
  const float* cxy;   // float(priorCorrXY(i + nxp j))
  const float* ct;    // circCorrT[k]
  float gi, gj;
  int nxp, nyp, nzp, rnxp;
  long long rsize;
};

template <typename T>
__global__ void k_param_corr(ParamCorrArgs a) {
#pragma clang fp contract(off)
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < a.rsize;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % a.rnxp);
    const long long r = t / a.rnxp;
    const int j = (int)(r % a.nyp), k = (int)(r / a.nyp);
    float value;
    if (i < a.nxp) {
      const int ci = cycle_of(i, a.nxp), cj = cycle_of(j, a.nyp);
      float subK = (float)k + (float)ci * a.gi;
      subK = subK + (float)cj * a.gj;
      int baseK = (int)floorf(subK);
      subK -= (float)baseK;
      while (baseK < 0) baseK += a.nzp;
      while (baseK >= a.nzp) baseK -= a.nzp;
      value = (1.0f - subK) * a.ct[baseK];
      value += subK * a.ct[baseK != a.nzp - 1 ? baseK + 1 : 0];
      value *= a.cxy[i + a.nxp * j];
    } else {
      value = kRMissing;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) ((T*)a.g[q])[t] = (T)(value * a.s[q]);
  }
}

// FFTGrid::fillInErrCorr (fftgrid.cpp:1018-1069): the lateral correlation
// times the weight of plane 0 at the shifted position s: within one sample
// of plane 0 or nzp, 1 - (s minus its truncation toward zero), moved into
// [-1, 1) cyclically (a negative base weighs by the other side); 0 elsewhere.
template <typename T>
__global__ void k_err_corr(T* g, const float* cxy, float gi, float gj, int nxp, int nyp, int nzp, int rnxp,
                           long long rsize) {
#pragma clang fp contract(off)
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < rsize;
       t += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(t % rnxp);
    const long long r = t / rnxp;
    const int j = (int)(r % nyp), k = (int)(r / nyp);
    const int ci = cycle_of(i, nxp), cj = cycle_of(j, nyp);
    float subK = (float)k + (float)ci * gi;
    subK = subK + (float)cj * gj;
    float value;
    if (i < nxp) {
      if (fabsf(subK) < 1.0f || fabsf(subK - (float)nzp) < 1.0f) {
        int baseK = (int)subK;
        subK -= (float)baseK;
        while (baseK < -1) baseK += nzp;
        while (baseK >= 1) baseK -= nzp;
        if (baseK < 0) subK = 1.0f - subK;
        value = (1.0f - subK) * cxy[i + nxp * j];
      } else {
        value = 0.0f;
      }
    } else {
      value = kRMissing;
    }
    g[t] = (T)value;
  }
}

int check_prior_args(crv_ctx* ctx, const double* sigma0, const double* cxy, const double* ct, crv_grid* cov[6]) {
  if (!sigma0 || !cxy || !ct || !cov) return crv_set_error(ctx, CRV_E_INVALID, "NULL argument");
  crv_grid* g0 = cov[0];
  for (int q = 0; q < 6; ++q)
    if (!cov[q] || cov[q]->nxp != g0->nxp || cov[q]->nyp != g0->nyp || cov[q]->nzp != g0->nzp ||
        cov[q]->precision != g0->precision)
      return crv_set_error(ctx, CRV_E_INVALID, "cov grids inconsistent");
  return CRV_OK;
}

// Copies cxy and ct to d (device, cnxp*nyp + nzp doubles) and fills the six
// spectra from there; stream-ordered, so d may be released or reused by a
// later call on the same stream.
int launch_prior_spectra(crv_ctx* ctx, const double* sigma0, const double* cxy, const double* ct, crv_grid* cov[6],
                         double* d) {
  crv_grid* g0 = cov[0];
  const size_t nxy = (size_t)g0->cnxp * g0->nyp;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(d, cxy, nxy * 8, hipMemcpyHostToDevice, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(d + nxy, ct, (size_t)g0->nzp * 8, hipMemcpyHostToDevice, ctx->stream));
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  double s[6];
  for (int q = 0; q < 6; ++q) s[q] = sigma0[pa[q] * 3 + pb[q]];
  const long long nc = (long long)g0->csize;
  ProfScope ps(ctx, "prior_spectra");
  if (g0->precision == CRV_FP64)
    hipLaunchKernelGGL(k_prior_spectra<double>, dim3(nblocks(nc)), dim3(kThreads), 0, ctx->stream,
                       (double2*)cov[0]->d, (double2*)cov[1]->d, (double2*)cov[2]->d, (double2*)cov[3]->d,
                       (double2*)cov[4]->d, (double2*)cov[5]->d, d, d + nxy, s[0], s[1], s[2], s[3], s[4], s[5],
                       g0->cnxp, g0->nyp, nc);
  else
    hipLaunchKernelGGL(k_prior_spectra<float>, dim3(nblocks(nc)), dim3(kThreads), 0, ctx->stream,
                       (float2*)cov[0]->d, (float2*)cov[1]->d, (float2*)cov[2]->d, (float2*)cov[3]->d,
                       (float2*)cov[4]->d, (float2*)cov[5]->d, d, d + nxy, s[0], s[1], s[2], s[3], s[4], s[5],
                       g0->cnxp, g0->nyp, nc);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// Staging area of crv_fill_prior_spectra_noalloc: a file-scope device array
// (each GPU holds its own copy until the process ends).  In host code its
// name is only a host-side stand-in; host code uses the per-device address
// that hipGetSymbolAddress returns (crv_noalloc_setup) instead.
constexpr size_t kPriorScratchDoubles = (size_t)1 << 19;   // 4 MiB
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};
int g_scratch_state[kMaxDevices] = {};   // 0 not looked up, 1 ok, -1 lookup failed

}  // namespace

static __device__ __align__(256) double scratch[kPriorScratchDoubles];

// Called by crv_ctx_create with `device` current: looks up this device's
// address of the staging array once per device and keeps it.
int crv_noalloc_setup(int device) {
  if (device < 0 || device >= kMaxDevices) return CRV_E_UNSUPPORTED;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_state[device] == 0) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e == hipSuccess && p) {
      g_scratch_ptr[device] = (double*)p;
      g_scratch_state[device] = 1;
    } else {
      g_scratch_state[device] = -1;
    }
  }
  return g_scratch_state[device] == 1 ? CRV_OK : CRV_E_HIP;
}

extern "C" int crv_fill_prior_spectra(crv_ctx* ctx, const double* sigma0, const double* cxy, const double* ct,
                                      crv_grid* cov[6]) {
  if (!ctx) return CRV_E_INVALID;
  int rc = check_prior_args(ctx, sigma0, cxy, ct, cov);
  if (rc != CRV_OK) return rc;
  crv_grid* g0 = cov[0];
  double* d = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&d, ((size_t)g0->cnxp * g0->nyp + g0->nzp) * 8, ctx->stream));
  rc = launch_prior_spectra(ctx, sigma0, cxy, ct, cov, d);
  CRV_HIP_CHECK(ctx, hipFreeAsync(d, ctx->stream));
  if (rc != CRV_OK) return rc;
  for (int q = 0; q < 6; ++q) { cov[q]->transformed = true; cov[q]->cubetype = CRV_COVARIANCE; }
  return CRV_OK;
}

extern "C" int crv_fill_prior_spectra_noalloc(crv_ctx* ctx, const double* sigma0, const double* cxy,
                                              const double* ct, crv_grid* cov[6]) {
  if (!ctx) return CRV_E_INVALID;
  int rc = check_prior_args(ctx, sigma0, cxy, ct, cov);
  if (rc != CRV_OK) return rc;
  crv_grid* g0 = cov[0];
  const size_t need = (size_t)g0->cnxp * g0->nyp + g0->nzp;
  if (need > kPriorScratchDoubles)
    return crv_set_error(ctx, CRV_E_UNSUPPORTED,
                         "fill_prior_spectra_noalloc: cnxp*nyp + nzp = " + std::to_string(need) +
                             " doubles exceeds the staging area of " + std::to_string(kPriorScratchDoubles));
  double* d = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (ctx->device >= 0 && ctx->device < kMaxDevices && g_scratch_state[ctx->device] == 1)
      d = g_scratch_ptr[ctx->device];
  }
  if (!d) return crv_set_error(ctx, CRV_E_HIP, "fill_prior_spectra_noalloc: staging area address unavailable");
  rc = launch_prior_spectra(ctx, sigma0, cxy, ct, cov, d);
  if (rc != CRV_OK) return rc;
  for (int q = 0; q < 6; ++q) { cov[q]->transformed = true; cov[q]->cubetype = CRV_COVARIANCE; }
  return CRV_OK;
}

// float(priorCorrXY) on the device (nxp nyp values)
static int upload_corr_xy(crv_ctx* ctx, const double* corr_xy, int nxp, int nyp, float** out) {
  std::vector<float> h((size_t)nxp * nyp);
  for (size_t q = 0; q < h.size(); ++q) h[q] = (float)corr_xy[q];
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)out, h.size() * sizeof(float), ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(*out, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  // the host array goes out of scope: finish the copy first
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return CRV_OK;
}

extern "C" int crv_fill_param_corr(crv_grid* cov[6], const double* corr_xy, const float* circ_corr_t, float grad_i,
                                   float grad_j, const double sigma0[6]) {
  if (!cov || !cov[0]) return CRV_E_INVALID;
  crv_ctx* ctx = cov[0]->ctx;
  if (!corr_xy || !circ_corr_t || !sigma0) return crv_set_error(ctx, CRV_E_INVALID, "param corr: NULL argument");
  crv_grid* g0 = cov[0];
  for (int q = 0; q < 6; ++q)
    if (!cov[q] || cov[q]->nxp != g0->nxp || cov[q]->nyp != g0->nyp || cov[q]->nzp != g0->nzp ||
        cov[q]->precision != g0->precision)
      return crv_set_error(ctx, CRV_E_INVALID, "param corr: grids inconsistent");
  if (!std::isfinite(grad_i) || !std::isfinite(grad_j) || std::fabs(grad_i) * g0->nxp > 1e6f ||
      std::fabs(grad_j) * g0->nyp > 1e6f)
    return crv_set_error(ctx, CRV_E_INVALID, "param corr: correlation gradient out of range");
  ParamCorrArgs a;
  for (int q = 0; q < 6; ++q) {
    a.g[q] = cov[q]->d;
    a.s[q] = (float)sigma0[q];
  }
  float* dxy = nullptr;
  int rc = upload_corr_xy(ctx, corr_xy, g0->nxp, g0->nyp, &dxy);
  if (rc != CRV_OK) return rc;
  float* dct = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dct, g0->nzp * sizeof(float), ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(dct, circ_corr_t, g0->nzp * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  a.cxy = dxy; a.ct = dct; a.gi = grad_i; a.gj = grad_j;
  a.nxp = g0->nxp; a.nyp = g0->nyp; a.nzp = g0->nzp; a.rnxp = g0->rnxp; a.rsize = (long long)g0->rsize;
  if (g0->precision == CRV_FP64)
    hipLaunchKernelGGL(k_param_corr<double>, dim3(nblocks(a.rsize)), dim3(kThreads), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL(k_param_corr<float>, dim3(nblocks(a.rsize)), dim3(kThreads), 0, ctx->stream, a);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  CRV_HIP_CHECK(ctx, hipFreeAsync(dxy, ctx->stream));
  CRV_HIP_CHECK(ctx, hipFreeAsync(dct, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  for (int q = 0; q < 6; ++q) {
    cov[q]->transformed = false;
    cov[q]->cubetype = CRV_COVARIANCE;
  }
  return CRV_OK;
}

extern "C" int crv_fill_err_corr(crv_grid* err, const double* corr_xy, float grad_i, float grad_j) {
  if (!err) return CRV_E_INVALID;
  crv_ctx* ctx = err->ctx;
  if (!corr_xy) return crv_set_error(ctx, CRV_E_INVALID, "err corr: NULL argument");
  if (!std::isfinite(grad_i) || !std::isfinite(grad_j))
    return crv_set_error(ctx, CRV_E_INVALID, "err corr: correlation gradient not finite");
  float* dxy = nullptr;
  int rc = upload_corr_xy(ctx, corr_xy, err->nxp, err->nyp, &dxy);
  if (rc != CRV_OK) return rc;
  const long long rs = (long long)err->rsize;
  if (err->precision == CRV_FP64)
    hipLaunchKernelGGL(k_err_corr<double>, dim3(nblocks(rs)), dim3(kThreads), 0, ctx->stream, (double*)err->d, dxy,
                       grad_i, grad_j, err->nxp, err->nyp, err->nzp, err->rnxp, rs);
  else
    hipLaunchKernelGGL(k_err_corr<float>, dim3(nblocks(rs)), dim3(kThreads), 0, ctx->stream, (float*)err->d, dxy,
                       grad_i, grad_j, err->nxp, err->nyp, err->nzp, err->rnxp, rs);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  CRV_HIP_CHECK(ctx, hipFreeAsync(dxy, ctx->stream));
  err->transformed = false;
  err->cubetype = CRV_COVARIANCE;
  return CRV_OK;
}

extern "C" int crv_grid_fill(crv_grid* g, double value) {
  if (!g) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  ProfScope ps(ctx, "fill");
  if (g->precision == CRV_FP64)
    hipLaunchKernelGGL(k_fill_real<double>, dim3(nblocks(g->rsize)), dim3(kThreads), 0, ctx->stream, (double*)g->d,
                       value, g->nxp, g->rnxp, (long long)g->rsize);
  else
    hipLaunchKernelGGL(k_fill_real<float>, dim3(nblocks(g->rsize)), dim3(kThreads), 0, ctx->stream, (float*)g->d,
                       (float)value, g->nxp, g->rnxp, (long long)g->rsize);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  g->transformed = false;
  return CRV_OK;
}

extern "C" int crv_forward_model(crv_ctx* ctx, int ntheta, const float* A, const double* kwd, crv_grid* model[3],
                                 crv_grid** data) {
  if (!ctx) return CRV_E_INVALID;
  if (ntheta < 1 || ntheta > 8 || !A || !kwd || !model || !data)
    return crv_set_error(ctx, CRV_E_INVALID, "forward model: bad arguments");
  crv_grid* g0 = model[0];
  for (int m = 0; m < 3; ++m)
    if (!model[m] || !model[m]->transformed || model[m]->nxp != g0->nxp || model[m]->nyp != g0->nyp ||
        model[m]->nzp != g0->nzp || model[m]->precision != g0->precision)
      return crv_set_error(ctx, CRV_E_INVALID, "forward model: model grids must be transformed and consistent");
  for (int l = 0; l < ntheta; ++l)
    if (!data[l] || data[l]->nxp != g0->nxp || data[l]->nyp != g0->nyp || data[l]->nzp != g0->nzp ||
        data[l]->precision != g0->precision)
      return crv_set_error(ctx, CRV_E_INVALID, "forward model: data grids inconsistent");
  const int nzp = g0->nzp;
  std::vector<double> K((size_t)nzp * ntheta * 3 * 2);
  for (int k = 0; k < nzp; ++k)
    for (int l = 0; l < ntheta; ++l)
      for (int m = 0; m < 3; ++m) {
        const size_t t = ((size_t)k * ntheta + l) * 2;
        K[(((size_t)k * ntheta + l) * 3 + m) * 2] = kwd[t] * A[l * 3 + m];
        K[(((size_t)k * ntheta + l) * 3 + m) * 2 + 1] = kwd[t + 1] * A[l * 3 + m];
      }
  double2* dK = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dK, K.size() * 8, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(dK, K.data(), K.size() * 8, hipMemcpyHostToDevice, ctx->stream));
  FwdArgs a{};
  for (int m = 0; m < 3; ++m) a.m[m] = model[m]->d;
  for (int l = 0; l < ntheta; ++l) a.d[l] = data[l]->d;
  a.K = dK; a.n = ntheta; a.cnxp = g0->cnxp; a.nyp = g0->nyp; a.nc = (long long)g0->csize;
  {
    ProfScope ps(ctx, "forward_model");
    if (g0->precision == CRV_FP64)
      hipLaunchKernelGGL(k_forward<double>, dim3(nblocks(a.nc)), dim3(kThreads), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL(k_forward<float>, dim3(nblocks(a.nc)), dim3(kThreads), 0, ctx->stream, a);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  CRV_HIP_CHECK(ctx, hipFreeAsync(dK, ctx->stream));
  for (int l = 0; l < ntheta; ++l) { data[l]->transformed = true; data[l]->cubetype = CRV_DATA; }
  return CRV_OK;
}

extern "C" int crv_add_white_noise(crv_grid* g, double sigma, unsigned long long seed, unsigned long long stream) {
  if (!g) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "add_white_noise: grid must be real");
  ProfScope ps(ctx, "white_noise");
  if (g->precision == CRV_FP64)
    hipLaunchKernelGGL(k_white_noise<double>, dim3(nblocks(g->rsize)), dim3(kThreads), 0, ctx->stream, (double*)g->d,
                       sigma, seed, stream, g->nxp, g->rnxp, (long long)g->rsize);
  else
    hipLaunchKernelGGL(k_white_noise<float>, dim3(nblocks(g->rsize)), dim3(kThreads), 0, ctx->stream, (float*)g->d,
                       sigma, seed, stream, g->nxp, g->rnxp, (long long)g->rsize);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

extern "C" int crv_grid_moments(crv_grid* g, double* sum, double* sum2, long long* count) {
  if (!g || !sum || !sum2 || !count) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "moments: grid must be real");
  const int nb = 1024;
  unsigned char* d = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&d, (size_t)nb * 24, ctx->stream));
  {
    ProfScope ps(ctx, "moments");
    if (g->precision == CRV_FP64)
      hipLaunchKernelGGL(k_moments<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const double*)g->d, g->nx, g->ny,
                         g->nz, g->rnxp, g->nyp, (double*)d, (double*)(d + nb * 8), (unsigned long long*)(d + nb * 16));
    else
      hipLaunchKernelGGL(k_moments<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, (const float*)g->d, g->nx, g->ny,
                         g->nz, g->rnxp, g->nyp, (double*)d, (double*)(d + nb * 8), (unsigned long long*)(d + nb * 16));
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  std::vector<unsigned char> h((size_t)nb * 24);
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(h.data(), d, h.size(), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  CRV_HIP_CHECK(ctx, hipFreeAsync(d, ctx->stream));
  double s2 = 0, s1 = 0;
  unsigned long long c = 0;
  for (int b = 0; b < nb; ++b) {
    double x;
    unsigned long long y;
    memcpy(&x, h.data() + b * 8, 8); s2 += x;
    memcpy(&x, h.data() + nb * 8 + b * 8, 8); s1 += x;
    memcpy(&y, h.data() + nb * 16 + b * 8, 8); c += y;
  }
  *sum = s1; *sum2 = s2; *count = (long long)c;
  return CRV_OK;
}
