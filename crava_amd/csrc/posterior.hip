// Per-wavenumber Bayesian posterior update on gfx950.
//
// Replaces the k, j, i loop of Crava::computePostMeanResidAndFFTCov
// (reference src/crava.cpp:1084-1247) together with invertSeismicData
// (1388-1438), getNextErrorVariance (1441-1469),
// SeismicParametersHolder::getParameterCovariance (seismicparametersholder.cpp:
// 339-426) and lib_matrCholCpx (libs/lib/lib_matr.c:78-167).
//
// One thread owns one wavenumber (i, j, k) of the half spectrum; it gathers
// 3 means, n data, the error and prior spectra, factors the n x n Hermitian
// system, and scatters the posterior mean, the n residuals and the 6 entries
// of the 3 x 3 posterior covariance.  The kz-only quantities (forward matrix
// K = conj(W_l) D1 A, error shape, white-noise diagonal, band mask) are
// tabulated once per call on the host, so a wavenumber costs one pass over
// its own grid entries: the kernel is bound by HBM bandwidth.
//
// Algebra.  With M = K Sm K^H + Se = L L^H, Y = L^-1 K Sm and w = L^-1 (d - K mu):
//   mu_post = mu + Y^H w,  S_post = Sm - Y^H Y,  res = d - K mu_post,
// which equals the reference's KScc (M^-1 KS) and mu + (M^-1 KS)^H (d - K mu).
#include "internal.hpp"
#include "linalg.hpp"

#include <cmath>
#include <cstring>

namespace {

constexpr int kThreads = 256;

template <typename T>
struct PostArgs {
  cplx_t<T>* mean[3];
  cplx_t<T>* data[8];
  const cplx_t<T>* err;       // errCorr spectrum (non-separable)
  cplx_t<T>* cov[6];          // aa bb rr ab ar br
  const cplx_t<T>* Kt;        // [nzp][n][3]
  const cplx_t<T>* Et;        // [nzp][n][n]
  const T* Dt;                // [nzp][n]
  const int* inv;             // [nzp]
  // FP64 copies of the tables (refine mode fallback)
  const double2* Kt64;
  const double2* Et64;
  const double* Dt64;
  const double* cxy;          // separable prior: [nyp][cnxp]
  const double* ct;           // [nzp]
  double sig0[6];
  int separable;
  int nxp, cnxp, nyp, nzp;
  long long nc;
  unsigned long long* stats;  // [0] inverted [1] chol fail [2] fp64 [3] max cond (bits)
  unsigned int* list;         // refine mode: indices to redo in FP64
  unsigned int* list_count;
  double cond_thr;
  double lowk2;
};

// Gather the cleaned prior (all entries real: getParameterCovariance keeps
// Sigma0[a][b] * |Re(c)/Sigma0[a][b]|) and the error-spectrum multiplier.
template <typename T>
__device__ __forceinline__ void load_prior(const PostArgs<T>& a, long long idx, int i, int j, int k,
                                           double s[6], double& lam) {
  if (a.separable) {
    const double cx = a.cxy[(size_t)j * a.cnxp + i];
    const double sc = fabs(cx * a.ct[k]);
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = a.sig0[q] * sc;
    lam = fabs(cx);
  } else {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double re = (double)a.cov[q][idx].x;
      const double sg = a.sig0[q];
      s[q] = (sg != 0.0) ? sg * fabs(re / sg) : 0.0;
    }
    lam = fabs((double)a.err[idx].x);
  }
}

template <typename T>
__device__ __forceinline__ void store_cov_real(const PostArgs<T>& a, long long idx, const double s[6]) {
  using C = cplx_t<T>;
#pragma unroll
  for (int q = 0; q < 6; ++q) a.cov[q][idx] = cmake<C>((T)s[q], (T)0);
}

__device__ __forceinline__ void stat_add(unsigned long long* p, unsigned long long v) {
  // one atomic per wavefront
  unsigned long long s = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
  if ((threadIdx.x & 63) == 0 && s) atomicAdd(p, s);
}
__device__ __forceinline__ void stat_max(unsigned long long* p, double v) {
  double m = v;
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmax(m, __shfl_down(m, off, 64));
  if ((threadIdx.x & 63) == 0 && m > 0.0) atomicMax(p, (unsigned long long)__double_as_longlong(m));
}

// Full solve for one wavenumber in precision U (storage T).  Returns false
// when the factorisation is rejected (posterior = prior).
template <typename T, typename U, int N>
__device__ __forceinline__ bool solve_one(const PostArgs<T>& a, long long idx, int k, const double s[6],
                                          double lam, const cplx_t<U>* Kt, const cplx_t<U>* Et, const U* Dt,
                                          double& cond) {
  using C = cplx_t<U>;
  using CT = cplx_t<T>;
  C K[N][3];
#pragma unroll
  for (int l = 0; l < N; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m) K[l][m] = Kt[((size_t)k * N + l) * 3 + m];
  const U S[3][3] = {{(U)s[0], (U)s[3], (U)s[4]}, {(U)s[3], (U)s[1], (U)s[5]}, {(U)s[4], (U)s[5], (U)s[2]}};
  C KS[N][4];  // columns 0..2: K Sm ; column 3: d - K mu
#pragma unroll
  for (int l = 0; l < N; ++l)
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      C x = cmake<C>((U)0, (U)0);
#pragma unroll
      for (int q = 0; q < 3; ++q) x = cadd(x, cscale(K[l][q], S[q][m]));
      KS[l][m] = x;
    }
  C A[N * (N + 1) / 2];
  const U lm = (U)lam;
#pragma unroll
  for (int l = 0; l < N; ++l)
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      C x = cscale(Et[((size_t)k * N + l) * N + m], lm);
      if (l == m) x.x += Dt[(size_t)k * N + l];
#pragma unroll
      for (int q = 0; q < 3; ++q) x = cadd(x, cmulc(KS[l][q], K[m][q]));
      A[l * (l + 1) / 2 + m] = x;
    }
  U dmin, dmax;
  if (!chol_ref<U, N>(A, dmin, dmax)) return false;
  cond = (double)(dmax / dmin) * (double)(dmax / dmin);
  C mu[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) { CT v = a.mean[m][idx]; mu[m] = cmake<C>((U)v.x, (U)v.y); }
  C d[N];
#pragma unroll
  for (int l = 0; l < N; ++l) {
    CT v = a.data[l][idx];
    d[l] = cmake<C>((U)v.x, (U)v.y);
    C x = d[l];
#pragma unroll
    for (int m = 0; m < 3; ++m) x = csub(x, cmul(K[l][m], mu[m]));
    KS[l][3] = x;
  }
  lsolve<U, N, 4>(A, KS);
  // mu_post = mu + Y^H w ; S_post = Sm - Y^H Y
  C mp[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    C x = mu[m];
#pragma unroll
    for (int l = 0; l < N; ++l) x = cadd(x, cmulc(KS[l][3], KS[l][m]));
    mp[m] = x;
    a.mean[m][idx] = cmake<CT>((T)x.x, (T)x.y);
  }
#pragma unroll
  for (int l = 0; l < N; ++l) {
    C x = d[l];
#pragma unroll
    for (int m = 0; m < 3; ++m) x = csub(x, cmul(K[l][m], mp[m]));
    a.data[l][idx] = cmake<CT>((T)x.x, (T)x.y);
  }
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    C x = cmake<C>((U)s[q], (U)0);
#pragma unroll
    for (int l = 0; l < N; ++l) x = csub(x, cmulc(KS[l][pb[q]], KS[l][pa[q]]));
    a.cov[q][idx] = cmake<CT>((T)x.x, (T)x.y);
  }
  return true;
}

__device__ __forceinline__ void decode(long long idx, int cnxp, int nyp, int& i, int& j, int& k) {
  const long long plane = (long long)cnxp * nyp;
  k = (int)(idx / plane);
  const int rem = (int)(idx - (long long)k * plane);
  j = rem / cnxp;
  i = rem - j * cnxp;
}

// Plain mode: everything in the grids' precision T.
template <typename T, int N>
__global__ __launch_bounds__(kThreads) void k_posterior(PostArgs<T> a) {
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long base = (long long)blockIdx.x * kThreads; base < a.nc; base += stride) {
    const long long idx = base + threadIdx.x;
    unsigned long long ninv = 0, nfail = 0;
    double cond = 0.0;
    if (idx < a.nc) {
      int i, j, k;
      decode(idx, a.cnxp, a.nyp, i, j, k);
      double s[6], lam;
      load_prior(a, idx, i, j, k, s, lam);
      bool done = false;
      if (a.inv[k]) {
        ninv = 1;
        done = solve_one<T, T, N>(a, idx, k, s, lam, a.Kt, a.Et, a.Dt, cond);
        if (!done) nfail = 1;
      }
      if (!done) store_cov_real(a, idx, s);
    }
    stat_add(&a.stats[0], ninv);
    stat_add(&a.stats[1], nfail);
    stat_max(&a.stats[3], cond);
  }
}

// FP32 storage, FP32 factorisation + one step of iterative refinement with
// FP64 residuals; wavenumbers that are ill-conditioned (cond > cond_thr),
// near |k| = 0, or rejected in FP32 are queued for an FP64 solve.
template <int N>
__global__ __launch_bounds__(kThreads) void k_posterior_refine(PostArgs<float> a) {
  using C = float2;
  using D = double2;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long base = (long long)blockIdx.x * kThreads; base < a.nc; base += stride) {
    const long long idx = base + threadIdx.x;
    unsigned long long ninv = 0;
    double cond = 0.0;
    if (idx < a.nc) {
      int i, j, k;
      decode(idx, a.cnxp, a.nyp, i, j, k);
      double s[6], lam;
      load_prior(a, idx, i, j, k, s, lam);
      bool done = false;
      if (a.inv[k]) {
        ninv = 1;
        const int ky = min(j, a.nyp - j), kz = min(k, a.nzp - k);
        const double fx = (double)i / (0.5 * a.nxp), fy = (double)ky / (0.5 * a.nyp), fz = (double)kz / (0.5 * a.nzp);
        const bool lowk = (fx * fx + fy * fy + fz * fz) < a.lowk2;
        C K[N][3];
#pragma unroll
        for (int l = 0; l < N; ++l)
#pragma unroll
          for (int m = 0; m < 3; ++m) K[l][m] = a.Kt[((size_t)k * N + l) * 3 + m];
        const float S[3][3] = {{(float)s[0], (float)s[3], (float)s[4]},
                               {(float)s[3], (float)s[1], (float)s[5]},
                               {(float)s[4], (float)s[5], (float)s[2]}};
        C B[N][4];
#pragma unroll
        for (int l = 0; l < N; ++l)
#pragma unroll
          for (int m = 0; m < 3; ++m) {
            C x = cmake<C>(0.f, 0.f);
#pragma unroll
            for (int q = 0; q < 3; ++q) x = cadd(x, cscale(K[l][q], S[q][m]));
            B[l][m] = x;
          }
        C A[N * (N + 1) / 2];
#pragma unroll
        for (int l = 0; l < N; ++l)
#pragma unroll
          for (int m = 0; m <= l; ++m) {
            C x = cscale(a.Et[((size_t)k * N + l) * N + m], (float)lam);
            if (l == m) x.x += a.Dt[(size_t)k * N + l];
#pragma unroll
            for (int q = 0; q < 3; ++q) x = cadd(x, cmulc(B[l][q], K[m][q]));
            A[l * (l + 1) / 2 + m] = x;
          }
        C M[N * (N + 1) / 2];
#pragma unroll
        for (int t = 0; t < N * (N + 1) / 2; ++t) M[t] = A[t];
        float dmin, dmax;
        const bool ok = chol_ref<float, N>(A, dmin, dmax);
        if (ok) cond = (double)(dmax / dmin) * (double)(dmax / dmin);
        if (ok && !lowk && cond <= a.cond_thr) {
          C mu[3];
#pragma unroll
          for (int m = 0; m < 3; ++m) mu[m] = a.mean[m][idx];
          C d[N];
#pragma unroll
          for (int l = 0; l < N; ++l) {
            d[l] = a.data[l][idx];
            C x = d[l];
#pragma unroll
            for (int m = 0; m < 3; ++m) x = csub(x, cmul(K[l][m], mu[m]));
            B[l][3] = x;
          }
          // One right-hand side at a time (columns of K Sm, then d - K mu):
          // x = M^-1 b in FP32, residual b - M x in FP64, x += M^-1 residual.
          // Each refined column is folded into the outputs at once, so only
          // one solution vector is live.
          C cv[6];
          C mp[3];
#pragma unroll
          for (int c = 0; c < 4; ++c) {
            C x[N][1], y[N][1];
#pragma unroll
            for (int l = 0; l < N; ++l) x[l][0] = B[l][c];
            lsolve<float, N, 1>(A, x);
            lhsolve<float, N, 1>(A, x);
#pragma unroll
            for (int l = 0; l < N; ++l) {
              D r = cmake<D>((double)B[l][c].x, (double)B[l][c].y);
#pragma unroll
              for (int m = 0; m < N; ++m) {
                const C mm = (m <= l) ? M[l * (l + 1) / 2 + m] : cconj(M[m * (m + 1) / 2 + l]);
                r = csub(r, cmul(cmake<D>((double)mm.x, (double)mm.y), cmake<D>((double)x[m][0].x, (double)x[m][0].y)));
              }
              y[l][0] = cmake<C>((float)r.x, (float)r.y);
            }
            lsolve<float, N, 1>(A, y);
            lhsolve<float, N, 1>(A, y);
#pragma unroll
            for (int l = 0; l < N; ++l) x[l][0] = cadd(x[l][0], y[l][0]);
            if (c < 3) {
              // S_post[a][c] = Sm[a][c] - sum_l conj(KS[l][a]) x[l], a <= c
#pragma unroll
              for (int aa = 0; aa <= c; ++aa) {
                C acc = cmake<C>(0.f, 0.f);
#pragma unroll
                for (int l = 0; l < N; ++l) acc = cadd(acc, cmulc(x[l][0], B[l][aa]));
                const int q = (aa == c) ? c : (c == 1 ? 3 : (aa == 0 ? 4 : 5));
                cv[q] = acc;
              }
            } else {
#pragma unroll
              for (int m = 0; m < 3; ++m) {
                C acc = mu[m];
#pragma unroll
                for (int l = 0; l < N; ++l) acc = cadd(acc, cmulc(x[l][0], B[l][m]));
                mp[m] = acc;
              }
            }
          }
#pragma unroll
          for (int m = 0; m < 3; ++m) a.mean[m][idx] = mp[m];
#pragma unroll
          for (int l = 0; l < N; ++l) {
            C x = d[l];
#pragma unroll
            for (int m = 0; m < 3; ++m) x = csub(x, cmul(K[l][m], mp[m]));
            a.data[l][idx] = x;
          }
#pragma unroll
          for (int q = 0; q < 6; ++q)
            a.cov[q][idx] = cmake<C>((float)(s[q] - (double)cv[q].x), (float)(-(double)cv[q].y));
          done = true;
        } else {
          const unsigned int slot = atomicAdd(a.list_count, 1u);
          a.list[slot] = (unsigned int)idx;
          done = true;  // finished by k_posterior_list64
        }
      }
      if (!done) store_cov_real(a, idx, s);
    }
    stat_add(&a.stats[0], ninv);
    stat_max(&a.stats[3], cond);
  }
}

// FP64 solve for the queued wavenumbers (FP32 storage).
template <int N>
__global__ __launch_bounds__(kThreads) void k_posterior_list64(PostArgs<float> a) {
  const unsigned int n = *a.list_count;
  for (unsigned int t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
    const long long idx = a.list[t];
    int i, j, k;
    decode(idx, a.cnxp, a.nyp, i, j, k);
    double s[6], lam;
    load_prior(a, idx, i, j, k, s, lam);
    double cond = 0.0;
    const bool ok = solve_one<float, double, N>(a, idx, k, s, lam, a.Kt64, a.Et64, a.Dt64, cond);
    if (!ok) {
      store_cov_real(a, idx, s);
      atomicAdd(&a.stats[1], 1ull);
    }
    atomicAdd(&a.stats[2], 1ull);
  }
}

template <typename T>
int launch_plain(crv_ctx* ctx, int n, const PostArgs<T>& a, int nblocks) {
  ProfScope ps(ctx, "posterior");
  switch (n) {
#define CASE(NN) case NN: hipLaunchKernelGGL((k_posterior<T, NN>), dim3(nblocks), dim3(kThreads), 0, ctx->stream, a); break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return crv_set_error(ctx, CRV_E_UNSUPPORTED, "ntheta must be 1..8");
  }
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

int launch_refine(crv_ctx* ctx, int n, const PostArgs<float>& a, int nblocks) {
  switch (n) {
#define CASE(NN)                                                                                   \
  case NN:                                                                                         \
    {                                                                                              \
      ProfScope p1(ctx, "posterior_refine");                                                       \
      hipLaunchKernelGGL((k_posterior_refine<NN>), dim3(nblocks), dim3(kThreads), 0, ctx->stream, a); \
    }                                                                                              \
    {                                                                                              \
      ProfScope p2(ctx, "posterior_fp64_list");                                                    \
      hipLaunchKernelGGL((k_posterior_list64<NN>), dim3(1024), dim3(kThreads), 0, ctx->stream, a);  \
    }                                                                                              \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return crv_set_error(ctx, CRV_E_UNSUPPORTED, "ntheta must be 1..8");
  }
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

int check_grid(crv_ctx* ctx, const crv_grid* g, const crv_grid* ref, const char* what) {
  if (!g) return crv_set_error(ctx, CRV_E_INVALID, std::string(what) + " is NULL");
  if (g->nxp != ref->nxp || g->nyp != ref->nyp || g->nzp != ref->nzp || g->precision != ref->precision)
    return crv_set_error(ctx, CRV_E_INVALID, std::string(what) + ": inconsistent size or precision");
  return CRV_OK;
}

// make sure a grid is in the Fourier domain (FFTCovGrids / fftInPlace semantics)
int ensure_transformed(crv_grid* g) {
  if (g->transformed) return CRV_OK;
  return crv_fft_r2c(g);
}

}  // namespace

extern "C" int crv_posterior(crv_ctx* ctx, const crv_post_params* p, crv_grid* mean[3], crv_grid** data,
                             crv_grid* err_corr, crv_grid* cov[6], crv_post_stats* stats) {
  if (!ctx) return CRV_E_INVALID;
  if (!p || !mean || !data || !cov) return crv_set_error(ctx, CRV_E_INVALID, "NULL argument");
  const int n = p->ntheta;
  if (n < 1 || n > 8) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "ntheta must be 1..8");
  crv_grid* g0 = mean[0];
  if (!g0) return crv_set_error(ctx, CRV_E_INVALID, "mean[0] is NULL");
  int rc;
  for (int m = 0; m < 3; ++m) if ((rc = check_grid(ctx, mean[m], g0, "mean")) != CRV_OK) return rc;
  for (int l = 0; l < n; ++l) if ((rc = check_grid(ctx, data[l], g0, "data")) != CRV_OK) return rc;
  for (int q = 0; q < 6; ++q) if ((rc = check_grid(ctx, cov[q], g0, "cov")) != CRV_OK) return rc;
  if (!p->separable && (rc = check_grid(ctx, err_corr, g0, "err_corr")) != CRV_OK) return rc;
  if (!p->A || !p->err_theta_cov || !p->kwd || !p->em1 || !p->em2 || !p->em3 || !p->sigma0)
    return crv_set_error(ctx, CRV_E_INVALID, "missing table");
  if (p->separable && (!p->cxy || !p->ct)) return crv_set_error(ctx, CRV_E_INVALID, "separable prior without cxy/ct");
  const bool refine = (p->mode == CRV_MODE_FP32_REFINE);
  if (refine && g0->precision != CRV_FP32)
    return crv_set_error(ctx, CRV_E_INVALID, "CRV_MODE_FP32_REFINE needs FP32 grids");

  const int nxp = g0->nxp, nyp = g0->nyp, nzp = g0->nzp, cnxp = g0->cnxp;
  // Forward transforms that the reference performs inside the call or just
  // before it (crava.cpp:177-184, 1039, 1062).
  for (int m = 0; m < 3; ++m) if ((rc = ensure_transformed(mean[m])) != CRV_OK) return rc;
  for (int l = 0; l < n; ++l) if ((rc = ensure_transformed(data[l])) != CRV_OK) return rc;
  if (!p->separable) {
    for (int q = 0; q < 6; ++q) if ((rc = ensure_transformed(cov[q])) != CRV_OK) return rc;
    if ((rc = ensure_transformed(err_corr)) != CRV_OK) return rc;
  }

  // ---- kz tables (crava.cpp:1104-1121, 1336-1364, 1441-1469) -------------
  std::vector<double> Kt((size_t)nzp * n * 3 * 2), Et((size_t)nzp * n * n * 2), Dt((size_t)nzp * n);
  std::vector<int> inv(nzp);
  const double wnc = p->wnc;
  for (int k = 0; k < nzp; ++k) {
    // crava.cpp:1104-1105: float numerator over a double product, compared
    // with lowCut * minRelThick in double and with highCut in float
    const float fr = (float)((double)(p->nz * 1000.0f) / (p->lz * nzp) * (double)std::min(k, nzp - k));
    inv[k] = ((double)fr > p->low_cut * p->min_rel_thick && fr < (float)p->high_cut) ? 1 : 0;
    for (int l = 0; l < n; ++l) {
      const double wr = p->kwd[((size_t)k * n + l) * 2], wi = p->kwd[((size_t)k * n + l) * 2 + 1];
      for (int m = 0; m < 3; ++m) {
        const double am = p->A[l * 3 + m];
        Kt[(((size_t)k * n + l) * 3 + m) * 2] = wr * am;
        Kt[(((size_t)k * n + l) * 3 + m) * 2 + 1] = wi * am;
      }
      const double* e1 = p->em1 + ((size_t)k * n) * 2;
      const double* e2 = p->em2 + ((size_t)k * n) * 2;
      const double* e3 = p->em3 + ((size_t)k * n) * 2;
      for (int m = 0; m < n; ++m) {
        const double th = p->err_theta_cov[l * n + m];
        const double c = 0.5 * (1.0 - wnc) * th;
        // e_l * conj(e_m), summed over the two error shapes
        const double re = e1[2 * l] * e1[2 * m] + e1[2 * l + 1] * e1[2 * m + 1] + e2[2 * l] * e2[2 * m] +
                          e2[2 * l + 1] * e2[2 * m + 1];
        const double im = -e1[2 * l] * e1[2 * m + 1] + e1[2 * l + 1] * e1[2 * m] - e2[2 * l] * e2[2 * m + 1] +
                          e2[2 * l + 1] * e2[2 * m];
        Et[(((size_t)k * n + l) * n + m) * 2] = c * re;
        Et[(((size_t)k * n + l) * n + m) * 2 + 1] = (l == m) ? 0.0 : c * im;
      }
      Dt[(size_t)k * n + l] = wnc * p->err_theta_cov[l * n + l] * e3[2 * l] * e3[2 * l];
    }
  }
  const bool f64 = (g0->precision == CRV_FP64);
  const size_t esz = f64 ? 8 : 4;
  // device table layout: [Kt T][Et T][Dt T][inv int][Kt64][Et64][Dt64][cxy][ct][stats][list count]
  auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
  const size_t bK = al(Kt.size() * esz), bE = al(Et.size() * esz), bD = al(Dt.size() * esz), bI = al(nzp * 4);
  const size_t bK64 = refine ? al(Kt.size() * 8) : 0, bE64 = refine ? al(Et.size() * 8) : 0,
               bD64 = refine ? al(Dt.size() * 8) : 0;
  const size_t bX = p->separable ? al((size_t)cnxp * nyp * 8) : 0, bT = p->separable ? al((size_t)nzp * 8) : 0;
  const size_t bS = 256;
  const size_t total = bK + bE + bD + bI + bK64 + bE64 + bD64 + bX + bT + bS;
  std::vector<unsigned char> h(total, 0);
  size_t off = 0;
  auto put = [&](const void* src, size_t bytes, size_t slot) {
    if (bytes) memcpy(h.data() + off, src, bytes);
    const size_t o = off;
    off += slot;
    return o;
  };
  size_t oK, oE, oD;
  if (f64) {
    oK = put(Kt.data(), Kt.size() * 8, bK);
    oE = put(Et.data(), Et.size() * 8, bE);
    oD = put(Dt.data(), Dt.size() * 8, bD);
  } else {
    std::vector<float> fK(Kt.begin(), Kt.end()), fE(Et.begin(), Et.end()), fD(Dt.begin(), Dt.end());
    oK = put(fK.data(), fK.size() * 4, bK);
    oE = put(fE.data(), fE.size() * 4, bE);
    oD = put(fD.data(), fD.size() * 4, bD);
  }
  const size_t oI = put(inv.data(), (size_t)nzp * 4, bI);
  size_t oK64 = 0, oE64 = 0, oD64 = 0;
  if (refine) {
    oK64 = put(Kt.data(), Kt.size() * 8, bK64);
    oE64 = put(Et.data(), Et.size() * 8, bE64);
    oD64 = put(Dt.data(), Dt.size() * 8, bD64);
  }
  size_t oX = 0, oT = 0;
  if (p->separable) {
    oX = put(p->cxy, (size_t)cnxp * nyp * 8, bX);
    oT = put(p->ct, (size_t)nzp * 8, bT);
  }
  const size_t oS = put(nullptr, 0, bS);
  unsigned char* dtab = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dtab, total, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(dtab, h.data(), total, hipMemcpyHostToDevice, ctx->stream));
  unsigned int* dlist = nullptr;
  const long long nc = (long long)cnxp * nyp * nzp;
  if (refine) CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dlist, (size_t)nc * 4, ctx->stream));

  const int nblocks = (int)std::min<long long>((nc + kThreads - 1) / kThreads, 256LL * 64);
  if (f64) {
    PostArgs<double> a{};
    for (int m = 0; m < 3; ++m) a.mean[m] = (double2*)mean[m]->d;
    for (int l = 0; l < n; ++l) a.data[l] = (double2*)data[l]->d;
    a.err = p->separable ? nullptr : (const double2*)err_corr->d;
    for (int q = 0; q < 6; ++q) a.cov[q] = (double2*)cov[q]->d;
    a.Kt = (const double2*)(dtab + oK); a.Et = (const double2*)(dtab + oE); a.Dt = (const double*)(dtab + oD);
    a.inv = (const int*)(dtab + oI);
    a.cxy = p->separable ? (const double*)(dtab + oX) : nullptr;
    a.ct = p->separable ? (const double*)(dtab + oT) : nullptr;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) a.sig0[q] = p->sigma0[pa[q] * 3 + pb[q]];
    a.separable = p->separable; a.nxp = nxp; a.cnxp = cnxp; a.nyp = nyp; a.nzp = nzp; a.nc = nc;
    a.stats = (unsigned long long*)(dtab + oS);
    rc = launch_plain<double>(ctx, n, a, nblocks);
  } else {
    PostArgs<float> a{};
    for (int m = 0; m < 3; ++m) a.mean[m] = (float2*)mean[m]->d;
    for (int l = 0; l < n; ++l) a.data[l] = (float2*)data[l]->d;
    a.err = p->separable ? nullptr : (const float2*)err_corr->d;
    for (int q = 0; q < 6; ++q) a.cov[q] = (float2*)cov[q]->d;
    a.Kt = (const float2*)(dtab + oK); a.Et = (const float2*)(dtab + oE); a.Dt = (const float*)(dtab + oD);
    a.inv = (const int*)(dtab + oI);
    a.cxy = p->separable ? (const double*)(dtab + oX) : nullptr;
    a.ct = p->separable ? (const double*)(dtab + oT) : nullptr;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; ++q) a.sig0[q] = p->sigma0[pa[q] * 3 + pb[q]];
    a.separable = p->separable; a.nxp = nxp; a.cnxp = cnxp; a.nyp = nyp; a.nzp = nzp; a.nc = nc;
    a.stats = (unsigned long long*)(dtab + oS);
    if (refine) {
      a.Kt64 = (const double2*)(dtab + oK64); a.Et64 = (const double2*)(dtab + oE64);
      a.Dt64 = (const double*)(dtab + oD64);
      a.list = dlist;
      a.list_count = (unsigned int*)(dtab + oS + 64);
      a.cond_thr = p->cond_threshold > 0 ? p->cond_threshold : 1e5;
      a.lowk2 = p->lowk_fraction * p->lowk_fraction;
      rc = launch_refine(ctx, n, a, nblocks);
    } else {
      rc = launch_plain<float>(ctx, n, a, nblocks);
    }
  }
  if (rc != CRV_OK) { (void)hipFreeAsync(dtab, ctx->stream); if (dlist) (void)hipFreeAsync(dlist, ctx->stream); return rc; }
  for (int l = 0; l < n; ++l) data[l]->transformed = true;
  for (int q = 0; q < 6; ++q) { cov[q]->transformed = true; cov[q]->cubetype = CRV_COVARIANCE; }

  if (stats) {
    unsigned long long hs[4];
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(hs, dtab + oS, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    stats->n_wavenumbers = nc;
    stats->n_inverted = (long long)hs[0];
    stats->n_chol_fail = (long long)hs[1];
    stats->n_fp64 = (long long)hs[2];
    double mc; memcpy(&mc, &hs[3], 8);
    stats->max_cond = mc;
  }
  CRV_HIP_CHECK(ctx, hipFreeAsync(dtab, ctx->stream));
  if (dlist) CRV_HIP_CHECK(ctx, hipFreeAsync(dlist, ctx->stream));

  if (!p->skip_mean_inverse) {
    for (int m = 0; m < 3; ++m) if ((rc = crv_fft_c2r(mean[m])) != CRV_OK) return rc;
  }
  return CRV_OK;
}
