// Posterior simulation and grid arithmetic on gfx950.
//
// One realization of Crava::simulate (reference src/crava.cpp:1599-1708):
// noise for seed0, seed1, seed2 is generated where it is used (counter
// stream) or read from the decoded reference stream; per wavenumber the
// 3 x 3 posterior covariance is factored with lib_matrCholCpx's rules and the
// noise vector multiplied by L (lib_matrProdCholVec, lib_matr.c:453-470), or
// zeroed when the factorisation is rejected (crava.cpp:1639-1651).  With
// fp64_chol 1 the factor of FP32 covariances is computed in FP64 and refined
// by one Newton step before it scales the noise; with 2 it is made in float
// as the reference does and again in FP64 only where the float one is
// rejected or ill-conditioned.  The three
// inverse transforms then add the posterior means in their last pass
// (FFTGrid::add, fftgrid.cpp:1870-1891).
#include "internal.hpp"
#include "linalg.hpp"
#include "noise.hpp"
#include "simkern.hpp"

#include <algorithm>
#include <type_traits>
#include <cmath>

namespace {

constexpr int kThreads = 256;

template <typename T>
struct SimArgs {
  const cplx_t<T>* cov[6];   // aa bb rr ab ar br
  cplx_t<T>* out[3];
  NoiseGeom geo;
  int mode;                  // CRV_NOISE_*
  const double* normals;     // reference: grid g of realization r at (3 r + g) D
  long long draws;           // D
  const cplx_t<T>* host[3];  // host-supplied noise (uploaded)
  unsigned long long seed, real;
  long long nc;
  // realizations real, real + 1, ..., real + nreal - 1 (counter stream;
  // 1 otherwise), written rstride complex elements apart: each wavenumber's
  // factor is made once for all of them
  int nreal;
  long long rstride;
  // wavenumbers covered: y rows [j0, j0 + nyr) of every kz plane, ncr of
  // them (the whole grid: j0 = 0, nyr = nyp, ncr = nc); dplr divides by
  // cnxp nyr, dcx by cnxp (indices below 2^31)
  int cnxp, nyp, j0, nyr;
  long long ncr;
  FastDiv dplr, dcx;
};

template <typename T>
void sim_rows(SimArgs<T>& a, int cnxp, int nyp, int nzp, int j0, int j1) {
  a.cnxp = cnxp; a.nyp = nyp; a.j0 = j0; a.nyr = j1 - j0;
  a.ncr = (long long)cnxp * a.nyr * nzp;
  a.dplr = make_fastdiv(cnxp * a.nyr);
  a.dcx = make_fastdiv(cnxp);
}

// FM (factor mode): 0 the reference's factorisation in the grids' precision
// (without divisions for FP64 grids); 1 in FP64 with one refinement step
// (FP32 grids); 2 in float with the reference's rules, and in FP64 with
// refinement where that is rejected or ill-conditioned (FP32 grids).
template <typename T, int FM>
__global__ __launch_bounds__(kThreads) void k_simulate(SimArgs<T> a) {
  using C = cplx_t<T>;
  constexpr bool F64 = FM == 1;
  using U = typename std::conditional<F64, double, T>::type;
  using CU = cplx_t<U>;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < a.ncr; t += stride) {
    const int k = fdiv((int)t, a.dplr);
    const int rem = (int)t - k * a.dplr.d;
    const int jl = fdiv(rem, a.dcx);
    const long long i = (long long)(rem - jl * a.cnxp) + (long long)a.cnxp * ((long long)(a.j0 + jl) + (long long)a.nyp * k);
    const C cv[6] = {a.cov[0][i], a.cov[1][i], a.cov[2][i], a.cov[3][i], a.cov[4][i], a.cov[5][i]};
    CU Lf[6];
    bool ok;
    if constexpr (FM == 2) ok = sim_factor_mixed(cv, Lf);
    else ok = sim_factor<T, F64>(cv, Lf);
    for (int r = 0; r < a.nreal; ++r) {
      C o[3] = {cmake<C>((T)0, (T)0), cmake<C>((T)0, (T)0), cmake<C>((T)0, (T)0)};
      if (ok) {
        T re[3], im[3];
        if (a.mode == CRV_NOISE_COUNTER) {
          sim_noise3<T>(a.geo, a.seed, a.real + (unsigned long long)r, i, re, im);
        } else if (a.mode == CRV_NOISE_REFERENCE) {
          ref_noise3<T>(a.geo, a.normals, a.draws, r, i, re, im);
        } else {
#pragma unroll
          for (int g = 0; g < 3; ++g) {
            const C h = a.host[g][i];
            re[g] = h.x; im[g] = h.y;
          }
        }
        sim_apply<T, U>(Lf, re, im, o);
      }
#pragma unroll
      for (int g = 0; g < 3; ++g) a.out[g][(long long)r * a.rstride + i] = o[g];
    }
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
  if (nc >= (1LL << 31)) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "simulate: grid of 2^31 or more wavenumbers");
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
    a.seed = noise->seed; a.real = noise->realization; a.nc = nc; a.nreal = 1;
    sim_rows(a, plan->nxp / 2 + 1, plan->nyp, plan->nzp, 0, plan->nyp);
    hipLaunchKernelGGL((k_simulate<double, 0>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  } else {
    SimArgs<float> a{};
    for (int q = 0; q < 6; ++q) a.cov[q] = (const float2*)cov[q]->d;
    for (int g = 0; g < 3; ++g) a.out[g] = (float2*)out[g]->d;
    a.geo = make_noise_geom(plan); a.mode = noise->mode; a.normals = ctx->normals; a.draws = draws;
    for (int g = 0; g < 3; ++g) a.host[g] = hostbuf ? (const float2*)((char*)hostbuf + g * gb) : nullptr;
    a.seed = noise->seed; a.real = noise->realization; a.nc = nc; a.nreal = 1;
    sim_rows(a, plan->nxp / 2 + 1, plan->nyp, plan->nzp, 0, plan->nyp);
    if (fp64_chol == 2)
      hipLaunchKernelGGL((k_simulate<float, 2>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
    else if (fp64_chol)
      hipLaunchKernelGGL((k_simulate<float, 1>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((k_simulate<float, 0>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
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

// Noise and Cholesky scaling of nreal counter-stream realizations over y
// rows [j0, j1) of every kz plane: realization r, parameter g at
// out0 + (3 r + g) gstride complex elements (see internal.hpp).
int crv_launch_sim_rows(crv_ctx* ctx, FftPlan* plan, const void* const cov[6], void* out0, long long gstride,
                        int nreal, unsigned long long seed, unsigned long long real0, int fp64_chol, int j0, int j1,
                        const double* normals) {
  const int cnxp = plan->nxp / 2 + 1;
  const long long nc = (long long)cnxp * plan->nyp * plan->nzp;
  if (nc >= (1LL << 31)) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "simulate: grid of 2^31 or more wavenumbers");
  if (j0 < 0 || j1 > plan->nyp || j1 <= j0 || nreal < 1 || nreal > 65535)
    return crv_set_error(ctx, CRV_E_INVALID, "simulate: bad rows or realizations");
  const long long ncr = (long long)cnxp * (j1 - j0) * plan->nzp;
  const int nb = (int)std::min<long long>((ncr + kThreads - 1) / kThreads, 2048);
  ProfScope ps(ctx, "simulate_scale_batch");
  if (plan->prec == CRV_FP64) {
    SimArgs<double> a{};
    for (int q = 0; q < 6; ++q) a.cov[q] = (const double2*)cov[q];
    for (int g = 0; g < 3; ++g) a.out[g] = (double2*)out0 + g * gstride;
    a.geo = make_noise_geom(plan); a.mode = normals ? CRV_NOISE_REFERENCE : CRV_NOISE_COUNTER;
    a.normals = normals; a.draws = plan->noise_draws;
    a.seed = seed; a.real = real0; a.nc = nc; a.nreal = nreal; a.rstride = 3 * gstride;
    sim_rows(a, cnxp, plan->nyp, plan->nzp, j0, j1);
    hipLaunchKernelGGL((k_simulate<double, 0>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  } else {
    SimArgs<float> a{};
    for (int q = 0; q < 6; ++q) a.cov[q] = (const float2*)cov[q];
    for (int g = 0; g < 3; ++g) a.out[g] = (float2*)out0 + g * gstride;
    a.geo = make_noise_geom(plan); a.mode = normals ? CRV_NOISE_REFERENCE : CRV_NOISE_COUNTER;
    a.normals = normals; a.draws = plan->noise_draws;
    a.seed = seed; a.real = real0; a.nc = nc; a.nreal = nreal; a.rstride = 3 * gstride;
    sim_rows(a, cnxp, plan->nyp, plan->nzp, j0, j1);
    if (fp64_chol == 2)
      hipLaunchKernelGGL((k_simulate<float, 2>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
    else if (fp64_chol)
      hipLaunchKernelGGL((k_simulate<float, 1>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
    else
      hipLaunchKernelGGL((k_simulate<float, 0>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  }
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

extern "C" int crv_simulate_batch(crv_ctx* ctx, crv_grid* cov[6], crv_grid* post_mean[3], unsigned long long seed,
                                  unsigned long long realization0, int nreal, int fp64_chol, crv_grid** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!cov || !out || nreal < 1 || nreal > 65535) return crv_set_error(ctx, CRV_E_INVALID, "simulate_batch: bad args");
  crv_grid* g0 = cov[0];
  if (!g0) return crv_set_error(ctx, CRV_E_INVALID, "cov[0] is NULL");
  auto same = [&](const crv_grid* g) {
    return g && g->nxp == g0->nxp && g->nyp == g0->nyp && g->nzp == g0->nzp && g->precision == g0->precision;
  };
  for (int q = 0; q < 6; ++q) {
    if (!same(cov[q])) return crv_set_error(ctx, CRV_E_INVALID, "cov grids inconsistent");
    if (!cov[q]->transformed) return crv_set_error(ctx, CRV_E_STATE, "cov grid not transformed");
  }
  const size_t eb = g0->elem_bytes();
  const int nout = 3 * nreal;
  for (int b = 0; b < nout; ++b) {
    if (!same(out[b])) return crv_set_error(ctx, CRV_E_INVALID, "out grids inconsistent");
    if ((char*)out[b]->d != (char*)out[0]->d + (size_t)b * out[0]->rsize * eb)
      return crv_set_error(ctx, CRV_E_INVALID, "simulate_batch: out must be one contiguous grid batch");
  }
  const void* add[3] = {nullptr, nullptr, nullptr};
  int nadd = 0;
  if (post_mean) {
    for (int g = 0; g < 3; ++g) {
      if (!post_mean[g]) { nadd = 0; break; }
      if (!same(post_mean[g])) return crv_set_error(ctx, CRV_E_INVALID, "post_mean grids inconsistent");
      if (post_mean[g]->transformed) return crv_set_error(ctx, CRV_E_STATE, "post_mean must be real (untransformed)");
      add[g] = post_mean[g]->d;
      nadd = 3;
    }
  }
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, g0->nxp, g0->nyp, g0->nzp, g0->precision, &plan);
  if (rc != CRV_OK) return rc;
  const long long gstride = (long long)out[0]->rsize / 2;  // complex elements per grid
  const void* cv[6];
  for (int q = 0; q < 6; ++q) cv[q] = cov[q]->d;
  if ((rc = crv_launch_sim_rows(ctx, plan, cv, out[0]->d, gstride, nreal, seed, realization0, fp64_chol, 0,
                                g0->nyp)) != CRV_OK)
    return rc;
  for (int b = 0; b < nout; ++b) { out[b]->cubetype = CRV_PARAMETER; out[b]->transformed = true; }
  const double sc = crv_scale_inverse(out[0]);
  if ((rc = crv_launch_c2r_batch(ctx, plan, out[0]->d, nout, (long long)out[0]->rsize, sc, add, nadd)) != CRV_OK)
    return rc;
  for (int b = 0; b < nout; ++b) out[b]->transformed = false;
  return CRV_OK;
}
