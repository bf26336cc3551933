// C ABI: contexts, grids, transforms and padding helpers.
#include "internal.hpp"

#include <cmath>
#include <cstring>
#include <new>
#include <algorithm>
#include <cstdio>

int crv_set_error(crv_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->last_error = msg;
  return code;
}

// Scale factors of FFTGrid::fftInPlace / invFFTInPlace (fftgrid.cpp:1796-1797,
// 1827-1830).  FP32 grids use the reference's float expressions; FP64 grids
// the exact values.
double crv_scale_forward(const crv_grid* g) {
  if (g->cubetype == CRV_COVARIANCE || g->cubetype == CRV_OPERATOR) return 1.0;
  const long long n = (long long)g->nxp * g->nyp * g->nzp;
  if (g->precision == CRV_FP64) return 1.0 / std::sqrt((double)n);
  return (double)(1.0f / std::sqrt((float)n));
}
double crv_scale_inverse(const crv_grid* g) {
  const long long n = (long long)g->nxp * g->nyp * g->nzp;
  if (g->cubetype == CRV_COVARIANCE || g->cubetype == CRV_OPERATOR)
    return g->precision == CRV_FP64 ? 1.0 / (double)n : (double)(float)(1.0 / (double)n);
  if (g->precision == CRV_FP64) return 1.0 / std::sqrt((double)n);
  return (double)(float)(1.0 / (double)std::sqrt((float)n));
}

extern "C" {

const char* crv_version(void) { return "crava_hip 1.0 gfx950"; }

int crv_ctx_create(int device, crv_ctx** out) {
  if (!out) return CRV_E_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return CRV_E_HIP;
  if (hipSetDevice(device) != hipSuccess) return CRV_E_HIP;
  crv_ctx* c = new (std::nothrow) crv_ctx();
  if (!c) return CRV_E_NOMEM;
  c->device = device;
  int ncu = 0;
  if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && ncu > 0)
    c->num_cu = ncu;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) { delete c; return CRV_E_HIP; }
  c->own_stream = true;
  // a failed lookup is reported by the allocation-free entry point itself
  (void)crv_noalloc_setup(device);
  *out = c;
  return CRV_OK;
}

int crv_ctx_destroy(crv_ctx* ctx) {
  if (!ctx) return CRV_E_INVALID;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  for (auto& kv : ctx->plans) crv_free_plan(kv.second);
  if (ctx->lcg_scratch) (void)hipFree(ctx->lcg_scratch);
  if (ctx->normals) (void)hipFree(ctx->normals);
  if (ctx->own_stream && ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return CRV_OK;
}

const char* crv_last_error(const crv_ctx* ctx) { return ctx ? ctx->last_error.c_str() : "NULL context"; }

int crv_ctx_synchronize(crv_ctx* ctx) {
  if (!ctx) return CRV_E_INVALID;
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return CRV_OK;
}

int crv_ctx_set_stream(crv_ctx* ctx, void* s) {
  if (!ctx) return CRV_E_INVALID;
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (s) {
    if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
    ctx->stream = (hipStream_t)s;
    ctx->own_stream = false;
  } else if (!ctx->own_stream) {
    CRV_HIP_CHECK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    ctx->own_stream = true;
  }
  return CRV_OK;
}

int crv_grid_create(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision, int cubetype,
                    crv_grid** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!out) return crv_set_error(ctx, CRV_E_INVALID, "out is NULL");
  *out = nullptr;
  if (nxp < 1 || nyp < 1 || nzp < 1 || nx < 0 || ny < 0 || nz < 0 || nx > nxp || ny > nyp || nz > nzp)
    return crv_set_error(ctx, CRV_E_INVALID, "bad grid dimensions");
  if (precision != CRV_FP32 && precision != CRV_FP64) return crv_set_error(ctx, CRV_E_INVALID, "bad precision");
  if (cubetype < CRV_CTMISSING || cubetype > CRV_OPERATOR) return crv_set_error(ctx, CRV_E_INVALID, "bad cubetype");
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, nxp, nyp, nzp, precision, &plan);
  if (rc != CRV_OK) return rc;
  crv_grid* g = new (std::nothrow) crv_grid();
  if (!g) return crv_set_error(ctx, CRV_E_NOMEM, "host allocation");
  g->ctx = ctx;
  g->nx = nx; g->ny = ny; g->nz = nz; g->nxp = nxp; g->nyp = nyp; g->nzp = nzp;
  g->cnxp = nxp / 2 + 1; g->rnxp = 2 * g->cnxp;
  g->precision = precision; g->cubetype = cubetype; g->transformed = false;
  g->rsize = (size_t)g->rnxp * nyp * nzp;
  g->csize = (size_t)g->cnxp * nyp * nzp;
  const hipError_t e = hipMalloc(&g->d, g->rsize * g->elem_bytes());
  if (e != hipSuccess) {
    delete g;
    return crv_set_error(ctx, CRV_E_NOMEM, std::string("grid allocation: ") + hipGetErrorString(e));
  }
  *out = g;
  return CRV_OK;
}

int crv_grid_create_batch(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision,
                          int cubetype, int count, crv_grid** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!out || count < 1) return crv_set_error(ctx, CRV_E_INVALID, "grid batch: bad count or out");
  if (nxp < 1 || nyp < 1 || nzp < 1 || nx < 0 || ny < 0 || nz < 0 || nx > nxp || ny > nyp || nz > nzp)
    return crv_set_error(ctx, CRV_E_INVALID, "bad grid dimensions");
  if (precision != CRV_FP32 && precision != CRV_FP64) return crv_set_error(ctx, CRV_E_INVALID, "bad precision");
  if (cubetype < CRV_CTMISSING || cubetype > CRV_OPERATOR) return crv_set_error(ctx, CRV_E_INVALID, "bad cubetype");
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, nxp, nyp, nzp, precision, &plan);
  if (rc != CRV_OK) return rc;
  const int cnxp = nxp / 2 + 1;
  const size_t rsize = (size_t)2 * cnxp * nyp * nzp;
  const size_t eb = precision == CRV_FP64 ? 8 : 4;
  void* base = nullptr;
  const hipError_t e = hipMalloc(&base, rsize * eb * (size_t)count);
  if (e != hipSuccess) return crv_set_error(ctx, CRV_E_NOMEM, std::string("grid batch allocation: ") + hipGetErrorString(e));
  int* shared = new (std::nothrow) int(count);
  if (!shared) { (void)hipFree(base); return CRV_E_NOMEM; }
  for (int i = 0; i < count; ++i) {
    crv_grid* g = new (std::nothrow) crv_grid();
    if (!g) {
      for (int q = 0; q < i; ++q) delete out[q];
      delete shared;
      (void)hipFree(base);
      return crv_set_error(ctx, CRV_E_NOMEM, "host allocation");
    }
    g->ctx = ctx;
    g->nx = nx; g->ny = ny; g->nz = nz; g->nxp = nxp; g->nyp = nyp; g->nzp = nzp;
    g->cnxp = cnxp; g->rnxp = 2 * cnxp;
    g->precision = precision; g->cubetype = cubetype; g->transformed = false;
    g->rsize = rsize;
    g->csize = (size_t)cnxp * nyp * nzp;
    g->d = (char*)base + (size_t)i * rsize * eb;
    g->shared = shared;
    g->base = base;
    out[i] = g;
  }
  return CRV_OK;
}

int crv_grid_destroy(crv_grid* g) {
  if (!g) return CRV_E_INVALID;
  if (g->shared) {
    if (--*g->shared == 0) {
      (void)hipStreamSynchronize(g->ctx->stream);
      (void)hipFree(g->base);
      delete g->shared;
    }
  } else if (g->d) {
    (void)hipStreamSynchronize(g->ctx->stream);
    (void)hipFree(g->d);
  }
  delete g;
  return CRV_OK;
}

int crv_grid_info_get(const crv_grid* g, crv_grid_info* info) {
  if (!g || !info) return CRV_E_INVALID;
  info->nx = g->nx; info->ny = g->ny; info->nz = g->nz;
  info->nxp = g->nxp; info->nyp = g->nyp; info->nzp = g->nzp;
  info->cnxp = g->cnxp; info->rnxp = g->rnxp;
  info->precision = g->precision; info->cubetype = g->cubetype;
  info->is_transformed = g->transformed ? 1 : 0;
  info->rsize = g->rsize; info->csize = g->csize;
  return CRV_OK;
}

int crv_grid_set_cubetype(crv_grid* g, int cubetype) {
  if (!g) return CRV_E_INVALID;
  if (cubetype < CRV_CTMISSING || cubetype > CRV_OPERATOR) return crv_set_error(g->ctx, CRV_E_INVALID, "bad cubetype");
  g->cubetype = cubetype;
  return CRV_OK;
}

int crv_grid_set_transformed(crv_grid* g, int t) {
  if (!g) return CRV_E_INVALID;
  g->transformed = t != 0;
  return CRV_OK;
}

int crv_grid_device_ptr(const crv_grid* g, void** dptr) {
  if (!g || !dptr) return CRV_E_INVALID;
  *dptr = g->d;
  return CRV_OK;
}

int crv_grid_upload_real(crv_grid* g, const void* host, size_t nelem) {
  if (!g || !host) return CRV_E_INVALID;
  if (nelem != g->rsize) return crv_set_error(g->ctx, CRV_E_INVALID, "upload_real: nelem != rsize");
  CRV_HIP_CHECK(g->ctx, hipMemcpyAsync(g->d, host, nelem * g->elem_bytes(), hipMemcpyHostToDevice, g->ctx->stream));
  CRV_HIP_CHECK(g->ctx, hipStreamSynchronize(g->ctx->stream));
  g->transformed = false;
  return CRV_OK;
}

int crv_grid_upload_complex(crv_grid* g, const void* host, size_t nelem) {
  if (!g || !host) return CRV_E_INVALID;
  if (nelem != 2 * g->csize) return crv_set_error(g->ctx, CRV_E_INVALID, "upload_complex: nelem != 2*csize");
  CRV_HIP_CHECK(g->ctx, hipMemcpyAsync(g->d, host, nelem * g->elem_bytes(), hipMemcpyHostToDevice, g->ctx->stream));
  CRV_HIP_CHECK(g->ctx, hipStreamSynchronize(g->ctx->stream));
  g->transformed = true;
  return CRV_OK;
}

int crv_grid_download(crv_grid* g, void* host, size_t nelem) {
  if (!g || !host) return CRV_E_INVALID;
  if (nelem != g->rsize && nelem != 2 * g->csize)
    return crv_set_error(g->ctx, CRV_E_INVALID, "download: nelem must be rsize or 2*csize");
  CRV_HIP_CHECK(g->ctx, hipMemcpyAsync(host, g->d, nelem * g->elem_bytes(), hipMemcpyDeviceToHost, g->ctx->stream));
  CRV_HIP_CHECK(g->ctx, hipStreamSynchronize(g->ctx->stream));
  return CRV_OK;
}

int crv_grid_copy(crv_grid* dst, const crv_grid* src) {
  if (!dst || !src) return CRV_E_INVALID;
  if (dst->nxp != src->nxp || dst->nyp != src->nyp || dst->nzp != src->nzp)
    return crv_set_error(dst->ctx, CRV_E_INVALID, "copy: size mismatch");
  if (dst->precision != src->precision) {
    const int rc = crv_launch_convert(dst->ctx, dst->precision, dst->d, src->precision, src->d, src->rsize);
    if (rc != CRV_OK) return rc;
  } else {
    CRV_HIP_CHECK(dst->ctx, hipMemcpyAsync(dst->d, src->d, src->rsize * src->elem_bytes(), hipMemcpyDeviceToDevice,
                                           dst->ctx->stream));
  }
  dst->transformed = src->transformed;
  dst->cubetype = src->cubetype;
  return CRV_OK;
}

int crv_grid_add(crv_grid* dst, const crv_grid* src) {
  if (!dst || !src) return CRV_E_INVALID;
  if (dst->nxp != src->nxp || dst->nyp != src->nyp || dst->nzp != src->nzp || dst->precision != src->precision)
    return crv_set_error(dst->ctx, CRV_E_INVALID, "add: size/precision mismatch");
  // FFTGrid::add adds over csize complex (transformed) or rsize reals; both
  // cover the same buffer.
  const size_t n = dst->transformed ? 2 * dst->csize : dst->rsize;
  return crv_launch_axpy(dst->ctx, dst->precision, dst->d, src->d, n);
}

int crv_grid_multiply_by_scalar(crv_grid* g, double s) {
  if (!g) return CRV_E_INVALID;
  return crv_launch_scale(g->ctx, g->precision, g->d, s, g->rsize);
}

int crv_fft_r2c(crv_grid* g) {
  if (!g) return CRV_E_INVALID;
  return crv_fft_r2c_from(g, g);
}

int crv_fft_r2c_from(crv_grid* dst, const crv_grid* src) {
  if (!dst || !src) return CRV_E_INVALID;
  crv_ctx* ctx = dst->ctx;
  if (src->transformed) return crv_set_error(ctx, CRV_E_STATE, "fftInPlace: grid already transformed");
  if (src->cubetype == CRV_CTMISSING) return crv_set_error(ctx, CRV_E_STATE, "fftInPlace: cubetype missing");
  if (dst->nxp != src->nxp || dst->nyp != src->nyp || dst->nzp != src->nzp || dst->precision != src->precision)
    return crv_set_error(ctx, CRV_E_INVALID, "r2c: size/precision mismatch");
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, src->nxp, src->nyp, src->nzp, src->precision, &plan);
  if (rc != CRV_OK) return rc;
  rc = crv_launch_r2c(ctx, plan, dst->d, src->d, crv_scale_forward(src));
  if (rc != CRV_OK) return rc;
  dst->cubetype = src->cubetype;
  dst->transformed = true;
  return CRV_OK;
}

int crv_fft_c2r(crv_grid* g) {
  if (!g) return CRV_E_INVALID;
  return crv_fft_c2r_add(g, g, nullptr);
}

// grids[0..count) one contiguous batch of one size, type and state
static int check_batch(crv_grid** g, int count, bool transformed, const char* what) {
  if (!g || count < 1 || !g[0]) return CRV_E_INVALID;
  crv_ctx* ctx = g[0]->ctx;
  const size_t bytes = g[0]->rsize * g[0]->elem_bytes();
  for (int i = 0; i < count; ++i) {
    const crv_grid* x = g[i];
    if (!x || x->nxp != g[0]->nxp || x->nyp != g[0]->nyp || x->nzp != g[0]->nzp || x->precision != g[0]->precision ||
        x->cubetype != g[0]->cubetype)
      return crv_set_error(ctx, CRV_E_INVALID, std::string(what) + ": grids differ in size, precision or type");
    if (x->cubetype == CRV_CTMISSING) return crv_set_error(ctx, CRV_E_STATE, std::string(what) + ": cubetype missing");
    if (x->transformed != transformed)
      return crv_set_error(ctx, CRV_E_STATE, std::string(what) + ": wrong transformed state");
    if ((const char*)x->d != (const char*)g[0]->d + (size_t)i * bytes)
      return crv_set_error(ctx, CRV_E_INVALID, std::string(what) + ": grids are not one contiguous batch");
  }
  return CRV_OK;
}

int crv_fft_r2c_batch_from(crv_grid** dst, crv_grid** src, int count) {
  int rc = check_batch(src, count, false, "fft_r2c_batch");
  if (rc != CRV_OK) return rc;
  crv_ctx* ctx = src[0]->ctx;
  if (dst != src) {
    for (int i = 0; i < count; ++i) {
      if (!dst || !dst[i] || dst[i]->nxp != src[0]->nxp || dst[i]->nyp != src[0]->nyp ||
          dst[i]->nzp != src[0]->nzp || dst[i]->precision != src[0]->precision)
        return crv_set_error(ctx, CRV_E_INVALID, "fft_r2c_batch: destination grids differ");
      if ((const char*)dst[i]->d != (const char*)dst[0]->d + (size_t)i * src[0]->rsize * src[0]->elem_bytes())
        return crv_set_error(ctx, CRV_E_INVALID, "fft_r2c_batch: destination is not one contiguous batch");
    }
  }
  FftPlan* plan = nullptr;
  if ((rc = crv_get_plan(ctx, src[0]->nxp, src[0]->nyp, src[0]->nzp, src[0]->precision, &plan)) != CRV_OK) return rc;
  rc = crv_launch_r2c_batch(ctx, plan, dst[0]->d, src[0]->d, count, (long long)src[0]->rsize,
                            crv_scale_forward(src[0]));
  if (rc != CRV_OK) return rc;
  for (int i = 0; i < count; ++i) { dst[i]->cubetype = src[0]->cubetype; dst[i]->transformed = true; }
  return CRV_OK;
}

int crv_fft_r2c_batch(crv_grid** grids, int count) { return crv_fft_r2c_batch_from(grids, grids, count); }

int crv_fft_c2r_batch(crv_grid** grids, int count) {
  int rc = check_batch(grids, count, true, "fft_c2r_batch");
  if (rc != CRV_OK) return rc;
  crv_ctx* ctx = grids[0]->ctx;
  FftPlan* plan = nullptr;
  if ((rc = crv_get_plan(ctx, grids[0]->nxp, grids[0]->nyp, grids[0]->nzp, grids[0]->precision, &plan)) != CRV_OK)
    return rc;
  rc = crv_launch_c2r_batch(ctx, plan, grids[0]->d, count, (long long)grids[0]->rsize, crv_scale_inverse(grids[0]),
                            nullptr, 0);
  if (rc != CRV_OK) return rc;
  for (int i = 0; i < count; ++i) grids[i]->transformed = false;
  return CRV_OK;
}

int crv_fft_c2r_add(crv_grid* dst, crv_grid* src, const crv_grid* add) {
  if (!dst || !src) return CRV_E_INVALID;
  crv_ctx* ctx = dst->ctx;
  if (!src->transformed) return crv_set_error(ctx, CRV_E_STATE, "invFFTInPlace: grid not transformed");
  if (src->cubetype == CRV_CTMISSING) return crv_set_error(ctx, CRV_E_STATE, "invFFTInPlace: cubetype missing");
  if (dst->nxp != src->nxp || dst->nyp != src->nyp || dst->nzp != src->nzp || dst->precision != src->precision)
    return crv_set_error(ctx, CRV_E_INVALID, "c2r: size/precision mismatch");
  if (add) {
    if (add->nxp != src->nxp || add->nyp != src->nyp || add->nzp != src->nzp || add->precision != src->precision)
      return crv_set_error(ctx, CRV_E_INVALID, "c2r: add grid size/precision mismatch");
    if (add->transformed) return crv_set_error(ctx, CRV_E_STATE, "c2r: add grid must be real");
  }
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, src->nxp, src->nyp, src->nzp, src->precision, &plan);
  if (rc != CRV_OK) return rc;
  rc = crv_launch_c2r(ctx, plan, src->d, dst->d, crv_scale_inverse(src), add ? add->d : nullptr);
  if (rc != CRV_OK) return rc;
  dst->cubetype = src->cubetype;
  dst->transformed = false;
  return CRV_OK;
}

// FFTGrid::findClosestFactorableNumber (fftgrid.cpp:2857-2885): the bounds
// are the reference's float-log expressions; the search is exact.
int crv_find_closest_factorable(int leastint) {
  if (leastint < 1) return -1;
  const int maxant2 = (int)std::ceil((double)std::log((float)leastint) / std::log(2.0f));
  const int maxant3 = (int)std::ceil((double)std::log((float)leastint) / std::log(3.0f));
  const int maxant5 = (int)std::ceil((double)std::log((float)leastint) / std::log(5.0f));
  const int maxant7 = (int)std::ceil((double)std::log((float)leastint) / std::log(7.0f));
  int closest = (int)std::pow(2.0, maxant2);
  // products are exact integers in double; the reference's int cast of a
  // product beyond INT_MAX yields a negative number, which never qualifies
  for (int i = 0; i <= maxant2; ++i)
    for (int j = 0; j <= maxant3; ++j)
      for (int k = 0; k <= maxant5; ++k)
        for (int l = 0; l <= maxant7; ++l) {
          const double f = std::pow(2.0, i) * std::pow(3.0, j) * std::pow(5.0, k) * std::pow(7.0, l);
          if (f > 2147483647.0) continue;
          if ((int)f >= leastint && (int)f < closest) closest = (int)f;
        }
  return closest;
}

// ModelGeneral::setPaddingSize (modelgeneral.cpp:616-622)
int crv_padding_size(int n, double pad_fac) {
  const int leastint = (int)std::ceil(n * (1.0f + pad_fac));
  return crv_find_closest_factorable(leastint);
}

// Circular temporal correlation made positive definite
// (SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef,
// seismicparametersholder.cpp:430-498).  The transforms are direct DFTs in
// double (nzp is a trace length); the reference's unnormalised transforms
// and its final division by the lag-0 value give the same result up to
// rounding.
int crv_circ_corr_t(const float* corr, int ncorr, int nzp, int low_int_cut, float* out) {
  if (!corr || !out || ncorr < 1 || nzp < 1 || corr[0] == 0.0f) return CRV_E_INVALID;
  std::vector<double> c(nzp);
  for (int k = 0; k < nzp; ++k) {
    const int refk = k < nzp / 2 + 1 ? k : nzp - k;
    c[k] = refk < ncorr ? (double)corr[refk] : 0.0;
  }
  // half spectrum of the even sequence; its modulus, zero up to the cut
  const int nh = nzp / 2 + 1;
  std::vector<double> a(nh);
  const double w = 2.0 * M_PI / nzp;
  for (int f = 0; f < nh; ++f) {
    double re = 0.0, im = 0.0;
    for (int t = 0; t < nzp; ++t) {
      const long long r = ((long long)f * t) % nzp;
      re += c[t] * std::cos(w * (double)r);
      im -= c[t] * std::sin(w * (double)r);
    }
    a[f] = f <= low_int_cut ? 0.0 : std::sqrt(re * re + im * im);
  }
  // inverse real transform of the (real, even) half spectrum, then the
  // division by the value at lag 0
  std::vector<double> x(nzp);
  for (int t = 0; t < nzp; ++t) {
    double v = a[0];
    for (int f = 1; f < nh; ++f) {
      const long long r = ((long long)f * t) % nzp;
      const double m = (2 * f == nzp) ? 1.0 : 2.0;
      v += m * a[f] * std::cos(w * (double)r);
    }
    x[t] = v / nzp;
  }
  // the reference stops when the lag-0 value (of its unnormalised inverse,
  // nzp times this one) is not above 1e-5: too few layers for the correlation
  if (x[0] * nzp <= 1e-5) return CRV_E_INVALID;
  for (int t = 0; t < nzp; ++t) out[t] = (float)(x[t] / x[0]);
  return CRV_OK;
}

}  // extern "C"

extern "C" int crv_ctx_profile(crv_ctx* ctx, int enable) {
  if (!ctx) return CRV_E_INVALID;
  ctx->prof_on = enable != 0;
  return CRV_OK;
}

extern "C" int crv_ctx_profile_report(crv_ctx* ctx, char* buf, size_t len) {
  if (!ctx) return CRV_E_INVALID;
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  struct Acc { long long n = 0; double tot = 0, mn = 1e30, mx = 0; };
  std::map<std::string, Acc> acc;
  std::vector<std::string> order;
  for (auto& e : ctx->prof) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, e.a, e.b) == hipSuccess) {
      if (!acc.count(e.name)) order.push_back(e.name);
      Acc& a = acc[e.name];
      a.n++; a.tot += ms; a.mn = std::min(a.mn, (double)ms); a.mx = std::max(a.mx, (double)ms);
    }
    (void)hipEventDestroy(e.a);
    (void)hipEventDestroy(e.b);
  }
  ctx->prof.clear();
  std::string out;
  char line[256];
  for (auto& nm : order) {
    const Acc& a = acc[nm];
    snprintf(line, sizeof(line), "%s %lld %.6f %.6f %.6f\n", nm.c_str(), a.n, a.tot, a.mn, a.mx);
    out += line;
  }
  if (buf && len) {
    const size_t k = std::min(len - 1, out.size());
    memcpy(buf, out.data(), k);
    buf[k] = 0;
  }
  return (int)out.size() + 1;
}

extern "C" int crv_ctx_tune(crv_ctx* ctx, int key, int value) {
  if (!ctx) return CRV_E_INVALID;
  if (key < CRV_TUNE_CHUNK_MB || key > CRV_TUNE_ZINV || value < 0 || (key == CRV_TUNE_ORDER && value > 64))
    return crv_set_error(ctx, CRV_E_INVALID, "crv_ctx_tune: bad key or value");
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (key == CRV_TUNE_POST_BLOCKS) {
    ctx->post_blocks_per_cu = value;
    return CRV_OK;
  }
  const int old = ctx->tune[key];
  ctx->tune[key] = value;
  for (auto& kv : ctx->plans) {
    const int rc = crv_plan_shapes(ctx, kv.second);
    if (rc != CRV_OK) {
      ctx->tune[key] = old;
      for (auto& kv2 : ctx->plans) (void)crv_plan_shapes(ctx, kv2.second);
      return rc;
    }
  }
  return CRV_OK;
}

// ---- HIP graphs ---------------------------------------------------------------
struct crv_graph {
  crv_ctx* ctx = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

extern "C" int crv_graph_begin(crv_ctx* ctx) {
  if (!ctx) return CRV_E_INVALID;
  CRV_HIP_CHECK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  return CRV_OK;
}

extern "C" int crv_graph_end(crv_ctx* ctx, crv_graph** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!out) return crv_set_error(ctx, CRV_E_INVALID, "graph_end: out is NULL");
  *out = nullptr;
  hipGraph_t g = nullptr;
  CRV_HIP_CHECK(ctx, hipStreamEndCapture(ctx->stream, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t ie = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  if (ie != hipSuccess) {
    (void)hipGraphDestroy(g);
    return crv_set_error(ctx, CRV_E_HIP, std::string("graph instantiate: ") + hipGetErrorString(ie));
  }
  crv_graph* gr = new (std::nothrow) crv_graph();
  if (!gr) { (void)hipGraphExecDestroy(e); (void)hipGraphDestroy(g); return CRV_E_NOMEM; }
  gr->ctx = ctx; gr->graph = g; gr->exec = e;
  *out = gr;
  return CRV_OK;
}

extern "C" int crv_graph_launch(crv_graph* gr) {
  if (!gr) return CRV_E_INVALID;
  CRV_HIP_CHECK(gr->ctx, hipGraphLaunch(gr->exec, gr->ctx->stream));
  return CRV_OK;
}

extern "C" int crv_graph_destroy(crv_graph* gr) {
  if (!gr) return CRV_E_INVALID;
  (void)hipStreamSynchronize(gr->ctx->stream);
  if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
  if (gr->graph) (void)hipGraphDestroy(gr->graph);
  delete gr;
  return CRV_OK;
}
