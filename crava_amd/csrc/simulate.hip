// Posterior simulation and grid arithmetic on gfx950.
//
// One realization of Crava::simulate (reference src/crava.cpp:1599-1708):
// noise for seed0, seed1, seed2 is generated where it is used (counter
// stream) or read from the decoded reference stream; per wavenumber the
// 3 x 3 posterior covariance is factored with lib_matrCholCpx's rules and the
// noise vector multiplied by L (lib_matrProdCholVec, lib_matr.c:453-470), or
// zeroed when the factorisation is rejected (crava.cpp:1639-1651); the three
// inverse transforms then add the posterior means in their last pass
// (FFTGrid::add, fftgrid.cpp:1870-1891).
#include "internal.hpp"
#include "linalg.hpp"
#include "noise.hpp"

#include <algorithm>
#include <cmath>

namespace {

constexpr int kThreads = 256;

template <typename T>
struct SimArgs {
  const cplx_t<T>* cov[6];   // aa bb rr ab ar br
  cplx_t<T>* out[3];
  NoiseGeom geo;
  int mode;                  // CRV_NOISE_*
  const double* normals;     // reference: 3 consecutive grids of D draws each
  long long draws;           // D
  const cplx_t<T>* host[3];  // host-supplied noise (uploaded)
  unsigned long long seed, real;
  long long nc;
};

template <typename T, typename U>
__global__ __launch_bounds__(kThreads) void k_simulate(SimArgs<T> a) {
  using C = cplx_t<T>;
  using CU = cplx_t<U>;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long i = (long long)blockIdx.x * kThreads + threadIdx.x; i < a.nc; i += stride) {
    CU P[6];
    const C c0 = a.cov[0][i], c1 = a.cov[1][i], c2 = a.cov[2][i];
    const C c3 = a.cov[3][i], c4 = a.cov[4][i], c5 = a.cov[5][i];
    // lower triangle packed: P00, P10, P11, P20, P21, P22 (lower = conj(upper))
    P[0] = cmake<CU>((U)c0.x, (U)c0.y);
    P[1] = cmake<CU>((U)c3.x, -(U)c3.y);
    P[2] = cmake<CU>((U)c1.x, (U)c1.y);
    P[3] = cmake<CU>((U)c4.x, -(U)c4.y);
    P[4] = cmake<CU>((U)c5.x, -(U)c5.y);
    P[5] = cmake<CU>((U)c2.x, (U)c2.y);
    U dmin, dmax;
    const bool ok = chol_ref<U, 3>(P, dmin, dmax);
    CU z[3];
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      double re = 0.0, im = 0.0;
      if (ok) {
        if (a.mode == CRV_NOISE_COUNTER) {
          noise_counter_value(a.geo, a.seed, a.real, g, i, re, im);
        } else if (a.mode == CRV_NOISE_REFERENCE) {
          noise_ref_value(a.geo, a.normals + g * a.draws, i, re, im);
        } else {
          const C h = a.host[g][i];
          re = (double)h.x; im = (double)h.y;
        }
      }
      z[g] = cmake<CU>((U)re, (U)im);
    }
    CU o[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      CU x = cmake<CU>((U)0, (U)0);
      if (ok) {
#pragma unroll
        for (int c = 0; c <= r; ++c) x = cadd(x, cmul(P[r * (r + 1) / 2 + c], z[c]));
      }
      o[r] = x;
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) a.out[g][i] = cmake<C>((T)o[g].x, (T)o[g].y);
  }
}

template <typename T>
__global__ void k_axpy(T* __restrict__ d, const T* __restrict__ s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] += s[i];
}
template <typename T>
__global__ void k_scale(T* d, T s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) d[i] *= s;
}

template <typename S, typename D>
__global__ void k_convert(D* __restrict__ d, const S* __restrict__ s, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    d[i] = (D)s[i];
}

int nblocks_for(size_t n) { return (int)std::min<size_t>((n + kThreads - 1) / kThreads, 65536); }

}  // namespace

int crv_launch_convert(crv_ctx* ctx, int dprec, void* dst, int sprec, const void* src, size_t n) {
  ProfScope ps(ctx, "convert");
  if (dprec == CRV_FP64 && sprec == CRV_FP32)
    hipLaunchKernelGGL((k_convert<float, double>), dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (double*)dst,
                       (const float*)src, n);
  else if (dprec == CRV_FP32 && sprec == CRV_FP64)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (float*)dst,
                       (const double*)src, n);
  else
    return crv_set_error(ctx, CRV_E_INVALID, "convert: unsupported precision pair");
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

int crv_launch_axpy(crv_ctx* ctx, int prec, void* dst, const void* src, size_t n) {
  ProfScope ps(ctx, "axpy");
  if (prec == CRV_FP64)
    hipLaunchKernelGGL(k_axpy<double>, dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (double*)dst, (const double*)src, n);
  else
    hipLaunchKernelGGL(k_axpy<float>, dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (float*)dst, (const float*)src, n);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

int crv_launch_scale(crv_ctx* ctx, int prec, void* dst, double s, size_t n) {
  ProfScope ps(ctx, "scale");
  if (prec == CRV_FP64)
    hipLaunchKernelGGL(k_scale<double>, dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (double*)dst, s, n);
  else
    hipLaunchKernelGGL(k_scale<float>, dim3(nblocks_for(n)), dim3(kThreads), 0, ctx->stream, (float*)dst, (float)s, n);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

extern "C" int crv_simulate(crv_ctx* ctx, crv_grid* cov[6], crv_grid* post_mean[3], const crv_noise_spec* noise,
                            int fp64_chol, crv_grid* out[3]) {
  if (!ctx) return CRV_E_INVALID;
  if (!cov || !out || !noise) return crv_set_error(ctx, CRV_E_INVALID, "NULL argument");
  crv_grid* g0 = cov[0];
  if (!g0) return crv_set_error(ctx, CRV_E_INVALID, "cov[0] is NULL");
  auto same = [&](const crv_grid* g) {
    return g && g->nxp == g0->nxp && g->nyp == g0->nyp && g->nzp == g0->nzp && g->precision == g0->precision;
  };
  for (int q = 0; q < 6; ++q) {
    if (!same(cov[q])) return crv_set_error(ctx, CRV_E_INVALID, "cov grids inconsistent");
    if (!cov[q]->transformed) return crv_set_error(ctx, CRV_E_STATE, "cov grid not transformed");
  }
  for (int g = 0; g < 3; ++g) {
    if (!same(out[g])) return crv_set_error(ctx, CRV_E_INVALID, "out grids inconsistent");
    for (int q = 0; q < 6; ++q)
      if (out[g] == cov[q]) return crv_set_error(ctx, CRV_E_INVALID, "out grid aliases a cov grid");
    if (post_mean && post_mean[g]) {
      if (!same(post_mean[g])) return crv_set_error(ctx, CRV_E_INVALID, "post_mean grids inconsistent");
      if (post_mean[g]->transformed) return crv_set_error(ctx, CRV_E_STATE, "post_mean must be real (untransformed)");
    }
  }
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, g0->nxp, g0->nyp, g0->nzp, g0->precision, &plan);
  if (rc != CRV_OK) return rc;
  const long long nc = (long long)g0->csize;
  const bool f64 = g0->precision == CRV_FP64;
  void* hostbuf = nullptr;
  long long draws = 0;
  if (noise->mode == CRV_NOISE_REFERENCE) {
    if (!noise->lcg_state) return crv_set_error(ctx, CRV_E_INVALID, "reference noise needs lcg_state");
    if ((rc = crv_noise_prefix(ctx, plan)) != CRV_OK) return rc;
    draws = plan->noise_draws;
    if ((rc = crv_reference_normals(ctx, noise->lcg_state, 3 * draws)) != CRV_OK) return rc;
  } else if (noise->mode == CRV_NOISE_HOST) {
    const size_t gb = (size_t)nc * 2 * g0->elem_bytes();
    for (int g = 0; g < 3; ++g)
      if (!noise->host_noise[g]) return crv_set_error(ctx, CRV_E_INVALID, "host noise missing");
    CRV_HIP_CHECK(ctx, hipMallocAsync(&hostbuf, 3 * gb, ctx->stream));
    for (int g = 0; g < 3; ++g)
      CRV_HIP_CHECK(ctx, hipMemcpyAsync((char*)hostbuf + g * gb, noise->host_noise[g], gb, hipMemcpyHostToDevice, ctx->stream));
  } else if (noise->mode != CRV_NOISE_COUNTER) {
    return crv_set_error(ctx, CRV_E_INVALID, "unknown noise mode");
  }
  const int nb = (int)std::min<long long>((nc + kThreads - 1) / kThreads, 256LL * 64);
  const size_t gb = (size_t)nc * 2 * g0->elem_bytes();
  {
  ProfScope ps(ctx, "simulate_scale");
  if (f64) {
    SimArgs<double> a{};
    for (int q = 0; q < 6; ++q) a.cov[q] = (const double2*)cov[q]->d;
    for (int g = 0; g < 3; ++g) a.out[g] = (double2*)out[g]->d;
    a.geo = make_noise_geom(plan); a.mode = noise->mode; a.normals = ctx->normals; a.draws = draws;
    for (int g = 0; g < 3; ++g) a.host[g] = hostbuf ? (const double2*)((char*)hostbuf + g * gb) : nullptr;
    a.seed = noise->seed; a.real = noise->realization; a.nc = nc;
    hipLaunchKernelGGL((k_simulate<double, double>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  } else {
    SimArgs<float> a{};
    for (int q = 0; q < 6; ++q) a.cov[q] = (const float2*)cov[q]->d;
    for (int g = 0; g < 3; ++g) a.out[g] = (float2*)out[g]->d;
    a.geo = make_noise_geom(plan); a.mode = noise->mode; a.normals = ctx->normals; a.draws = draws;
    for (int g = 0; g < 3; ++g) a.host[g] = hostbuf ? (const float2*)((char*)hostbuf + g * gb) : nullptr;
    a.seed = noise->seed; a.real = noise->realization; a.nc = nc;
    if (fp64_chol)
      hipLaunchKernelGGL((k_simulate<float, double>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((k_simulate<float, float>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  }
  }
  CRV_HIP_CHECK(ctx, hipGetLastError());
  if (hostbuf) CRV_HIP_CHECK(ctx, hipFreeAsync(hostbuf, ctx->stream));
  // seed grids: PARAMETER type, transformed (fftgrid.cpp:1194-1196)
  for (int g = 0; g < 3; ++g) {
    out[g]->cubetype = CRV_PARAMETER;
    const double sc = crv_scale_inverse(out[g]);
    const void* add = (post_mean && post_mean[g]) ? post_mean[g]->d : nullptr;
    if ((rc = crv_launch_c2r(ctx, plan, out[g]->d, out[g]->d, sc, add)) != CRV_OK) return rc;
    out[g]->transformed = false;
  }
  return CRV_OK;
}

extern "C" int crv_fill_complex_noise(crv_grid* g, unsigned int* lcg_state) {
  if (!g || !lcg_state) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, g->nxp, g->nyp, g->nzp, g->precision, &plan);
  if (rc != CRV_OK) return rc;
  if ((rc = crv_noise_prefix(ctx, plan)) != CRV_OK) return rc;
  if ((rc = crv_reference_normals(ctx, lcg_state, plan->noise_draws)) != CRV_OK) return rc;
  if ((rc = crv_launch_fill_noise(ctx, plan, g, 0)) != CRV_OK) return rc;
  g->transformed = true;
  g->cubetype = CRV_PARAMETER;
  return CRV_OK;
}
