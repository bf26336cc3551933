// Per-wavenumber Bayesian posterior update on gfx950.
//
// Replaces the k, j, i loop of Crava::computePostMeanResidAndFFTCov
// (reference src/crava.cpp:1084-1247) together with invertSeismicData
// (1388-1438), getNextErrorVariance (1441-1469),
// SeismicParametersHolder::getParameterCovariance (seismicparametersholder.cpp:
// 339-426) and lib_matrCholCpx (libs/lib/lib_matr.c:78-167).
//
// One thread owns one wavenumber (i, j, k) of the half spectrum; it gathers
// 3 means, n data and the prior, factors the n x n Hermitian system, and
// scatters the posterior mean, the n residuals and the 6 entries of the
// 3 x 3 posterior covariance.  Everything that depends on kz only is
// tabulated once per call on the host in FP64: the forward matrix
// K = conj(W_l) D1 A (wavelet times Aki-Richards), the error shape E, the
// white-noise diagonal D and the band mask.  For the separable prior the
// prior and error spectra at (i, j, k) are |cxy(i,j)| * |ct(k)| * Sigma0 and
// |cxy(i,j)|, so with lam = |cxy(i,j)|
//   M = K S K^H + lam E + D = lam (|ct| K Sigma0 K^H + E) + D,  K S = lam |ct| K Sigma0,
// and the bracketed per-kz matrices are tabulated too: a wavenumber then
// costs one Cholesky factorisation and its solves, no matrix products.
//
// Algebra.  With M = L L^H, Y = L^-1 K S and w = L^-1 (d - K mu):
//   mu_post = mu + Y^H w,  S_post = S - Y^H Y,  res = d - K mu_post,
// which equals the reference's KScc (M^-1 K S) and mu + (M^-1 K S)^H (d - K mu).
//
// Precision.  CRV_MODE_PLAIN computes in the grids' precision.
// CRV_MODE_FP32_REFINE (FP32 grids) solves in FP32 with FP32 accumulation
// and sends a wavenumber to an FP64 solve with one step of iterative
// refinement when the FP32 factorisation is rejected, when its condition
// estimate (largest over smallest Cholesky pivot) exceeds the threshold, or
// when |k| is below the low-wavenumber fraction of Nyquist.
#include "internal.hpp"
#include "eiggen.hpp"
#include "linalg.hpp"
#include "noise.hpp"
#include "simkern.hpp"

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <type_traits>

namespace {

constexpr int kThreads = 256;

// Per-kz tables in precision U; P = n (n + 1) / 2 packed lower triangle.
template <typename U>
struct KzTab {
  const cplx_t<U>* K;   // [nzp][n][3]  forward matrix
  const cplx_t<U>* E;   // [nzp][P]     error shape; separable: |ct| K Sigma0 K^H + E
  const cplx_t<U>* B;   // [nzp][n][3]  separable: |ct| K Sigma0
  const U* D;           // [nzp][n]     white-noise diagonal
  const U* S;           // [nzp][6]     separable: |ct| Sigma0 (aa bb rr ab ar br)
  const U* lam;         // [nyp][cnxp]  separable: |cxy|
};

template <typename T>
struct PostArgs {
  cplx_t<T>* mean[3];
  cplx_t<T>* data[8];
  const cplx_t<T>* err;   // errCorr spectrum (general prior)
  cplx_t<T>* cov[6];      // aa bb rr ab ar br
  KzTab<T> t;
  KzTab<double> t64;      // FP64 copies (refine mode)
  const int* inv;         // [nzp] band mask
  T sig0[6];
  double sig0d[6];
  int nxp, cnxp, nyp, nzp;
  long long nc;
  unsigned long long* stats;  // [0] rejected factorisations [1] max condition (double bits)
  unsigned int* list;         // refine mode: wavenumbers for the FP64 solve
  unsigned int* list_count;
  float cond_thr;
  float lowk2;
  float rx, ry, rz;           // 2 / nxp, 2 / nyp, 2 / nzp
  FastDiv dplane, dcx;        // by cnxp nyp, cnxp (indices below 2^31)
  // the main pass covers y rows [j0, j0 + nyr) of every kz plane:
  // ncr = cnxp nyr nzp wavenumbers, dplr divides by cnxp nyr
  int j0, nyr;
  long long ncr;
  FastDiv dplr;
  // FP32 separable plans with n <= kRecMaxN: the per-kz tables of plane k as
  // one record of 128 floats at rec + 64 k (see k_posterior_rec); cpp =
  // chunks of 64 wavenumbers per plane of the row range
  const float2* rec;
  int cpp;
  // FP32 refine mode, separable prior, n <= kRecMaxN, D > 0 on the band's
  // planes: the plane's eigen-basis record of kEigRec floats at erec +
  // kEigRec k (see k_posterior_eig); nullptr: none
  const float* erec;
  // simulation fused into the FP32 passes (crv_invert): nsim realizations of
  // the counter noise stream, realization r of parameter g written at
  // sim + (3 r + g) sstride complex elements
  float2* sim;
  long long sstride;
  int nsim;
  unsigned long long seed, real0;
  NoiseGeom geo;
  // with normals: the decoded reference stream instead of the counter one
  // (grid g of realization r from draw (3 r + g) draws; geo has the prefix)
  const double* normals;
  long long draws;
  // the covariances of planes outside the band are left unwritten (the z
  // inverse pass generates them from the same tables, crv_invert)
  int skip_prior_cov;
  // eigen-basis kernel: the covariances inside the band are left unwritten
  // too, but for the wavenumbers of the FP64 list (generated likewise)
  int skip_band_cov;
  // eigen-basis kernel: the realizations only, the means and residuals
  // being updated by the z pass (crv_invert's fused z pass)
  int sim_only;
};

// the noise of realization r at wavenumber idx: the counter stream, or the
// decoded reference stream
__device__ __forceinline__ void fused_noise(const PostArgs<float>& a, int r, long long idx, float re[3],
                                            float im[3]) {
  if (a.normals) ref_noise3<float>(a.geo, a.normals, a.draws, r, idx, re, im);
  else sim_noise3<float>(a.geo, a.seed, a.real0 + (unsigned long long)r, idx, re, im);
}

// Simulation of one wavenumber from its covariance c as stored (FP32):
// SIM 1 factors in float with the reference's rules; SIM 2 likewise, and in
// FP64 with one refinement step where the float factorisation is rejected
// or ill-conditioned (sim_factor_mixed).
template <int SIM>
__device__ __forceinline__ void fused_sim(const PostArgs<float>& a, long long idx, const float2 c[6]) {
  if constexpr (SIM > 0) {
    float2 Lf[6];
    const bool ok = (SIM == 2) ? sim_factor_mixed(c, Lf) : sim_factor<float, false>(c, Lf);
    for (int r = 0; r < a.nsim; ++r) {
      float2 o[3] = {make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f)};
      if (ok) {
        float re[3], im[3];
        fused_noise(a, r, idx, re, im);
        sim_apply<float, float>(Lf, re, im, o);
      }
#pragma unroll
      for (int g = 0; g < 3; ++g) a.sim[(size_t)(3 * r + g) * a.sstride + idx] = o[g];
    }
  }
}

// Simulation of a wavenumber outside the inverted band, whose covariance is
// its plane's prior lam S with S real: the factor is sqrt(lam) L, L the
// plane's factor of S (real, packed lower 00 10 11 20 21 22; zero where S is
// rejected), made once per plane on the host.  o = sqrt(lam) L z in FP32.
__device__ __forceinline__ void fused_sim_prior(const PostArgs<float>& a, long long idx, float lam,
                                                const float L[6]) {
  const bool ok = L[0] > 0.0f && lam > 0.0f;
  const float sl = sqrtf(lam);
  float l[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) l[q] = sl * L[q];
  for (int r = 0; r < a.nsim; ++r) {
    float2 o[3] = {make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f), make_float2(0.0f, 0.0f)};
    if (ok) {
      float re[3], im[3];
      fused_noise(a, r, idx, re, im);
      o[0] = make_float2(l[0] * re[0], l[0] * im[0]);
      o[1] = make_float2(l[1] * re[0] + l[2] * re[1], l[1] * im[0] + l[2] * im[1]);
      o[2] = make_float2(l[3] * re[0] + l[4] * re[1] + l[5] * re[2], l[3] * im[0] + l[4] * im[1] + l[5] * im[2]);
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) a.sim[(size_t)(3 * r + g) * a.sstride + idx] = o[g];
  }
}

constexpr int kRecMaxN = 6;

// wavenumber t of the main pass's row range -> (i, j, k) and its index
template <typename T>
__device__ __forceinline__ long long decode_rows(const PostArgs<T>& a, long long t, int& i, int& j, int& k) {
  k = fdiv((int)t, a.dplr);
  const int rem = (int)t - k * a.dplr.d;
  const int jl = fdiv(rem, a.dcx);
  i = rem - jl * a.cnxp;
  j = a.j0 + jl;
  return (long long)i + (long long)a.cnxp * ((long long)j + (long long)a.nyp * k);
}

template <typename T>
__device__ __forceinline__ void decode(const PostArgs<T>& a, long long idx, int& i, int& j, int& k) {
  k = fdiv((int)idx, a.dplane);
  const int rem = (int)idx - k * a.dplane.d;
  j = fdiv(rem, a.dcx);
  i = rem - j * a.cnxp;
}

// Cleaned prior from the covariance grids (getParameterCovariance keeps
// Sigma0[a][b] * |Re(c) / Sigma0[a][b]|, a real value) and the error
// multiplier |Re errCorr|.
template <typename T, typename U>
__device__ __forceinline__ void gather_general(const PostArgs<T>& a, long long idx, const U* sig0, U s[6], U& lam) {
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const U re = (U)a.cov[q][idx].x;
    const U sg = sig0[q];
    s[q] = (sg != (U)0) ? sg * fabs(re / sg) : (U)0;
  }
  lam = fabs((U)a.err[idx].x);
}

template <typename T, typename U>
__device__ __forceinline__ void store_prior(const PostArgs<T>& a, long long idx, const U s[6]) {
#pragma unroll
  for (int q = 0; q < 6; ++q) a.cov[q][idx] = cmake<cplx_t<T>>((T)s[q], (T)0);
}

// M (packed lower) and the right-hand sides B = K S (n x 3), in precision U.
// Separable: B holds |ct| K Sigma0 (not yet multiplied by lam).
template <typename U, int N, bool SEP>
__device__ __forceinline__ void build_system(const KzTab<U>& t, int k, U lam, const U s[6], cplx_t<U>* M,
                                             cplx_t<U> (*B)[4]) {
  using C = cplx_t<U>;
  constexpr int P = N * (N + 1) / 2;
  if constexpr (SEP) {
#pragma unroll
    for (int q = 0; q < P; ++q) M[q] = cscale(t.E[(size_t)k * P + q], lam);
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
      for (int m = 0; m < 3; ++m) B[l][m] = t.B[((size_t)k * N + l) * 3 + m];
  } else {
    const C* Kk = t.K + (size_t)k * N * 3;
    const U S[3][3] = {{s[0], s[3], s[4]}, {s[3], s[1], s[5]}, {s[4], s[5], s[2]}};
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        C x = cmake<C>((U)0, (U)0);
#pragma unroll
        for (int q = 0; q < 3; ++q) x = cadd(x, cscale(Kk[l * 3 + q], S[q][m]));
        B[l][m] = x;
      }
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
      for (int m = 0; m <= l; ++m) {
        C x = cscale(t.E[(size_t)k * P + l * (l + 1) / 2 + m], lam);
#pragma unroll
        for (int q = 0; q < 3; ++q) x = cadd(x, cmulc(B[l][q], Kk[m * 3 + q]));
        M[l * (l + 1) / 2 + m] = x;
      }
  }
#pragma unroll
  for (int l = 0; l < N; ++l) M[l * (l + 1) / 2 + l].x += t.D[(size_t)k * N + l];
}

// One reduction per block for the statistics.
__device__ __forceinline__ void block_stats(unsigned long long* st, unsigned long long nfail, double cmax) {
  __shared__ unsigned long long s_f[kThreads / 64];
  __shared__ double s_c[kThreads / 64];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    nfail += __shfl_down(nfail, off, 64);
    cmax = fmax(cmax, __shfl_down(cmax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) { s_f[threadIdx.x >> 6] = nfail; s_c[threadIdx.x >> 6] = cmax; }
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long f = 0;
    double c = 0.0;
#pragma unroll
    for (int w = 0; w < kThreads / 64; ++w) { f += s_f[w]; c = fmax(c, s_c[w]); }
    if (f) atomicAdd(&st[0], f);
    if (c > 0.0) atomicMax(&st[1], (unsigned long long)__double_as_longlong(c));
  }
}

// Main pass: arithmetic in the grids' precision T.  LIST (FP32 refine mode)
// queues rejected, ill-conditioned and low-|k| wavenumbers for the FP64 pass
// instead of finishing them.
template <typename T, int N, bool SEP, bool LIST>
__global__ __launch_bounds__(kThreads) void k_posterior(PostArgs<T> a) {
  using C = cplx_t<T>;
  constexpr int P = N * (N + 1) / 2;
  unsigned long long nfail = 0;
  double cmax = 0.0;
  const long long stride = (long long)gridDim.x * kThreads;
  for (long long t = (long long)blockIdx.x * kThreads + threadIdx.x; t < a.ncr; t += stride) {
    int i, j, k;
    const long long idx = decode_rows(a, t, i, j, k);
    T lam, s[6];
    if constexpr (SEP) {
      lam = a.t.lam[(size_t)j * a.cnxp + i];
#pragma unroll
      for (int q = 0; q < 6; ++q) s[q] = lam * a.t.S[(size_t)k * 6 + q];
    } else {
      gather_general<T, T>(a, idx, a.sig0, s, lam);
    }
    if (!a.inv[k]) { store_prior(a, idx, s); continue; }
    C A[P];
    C B[N][4];  // columns 0..2: K S (separable: |ct| K Sigma0), column 3: d - K mu
    build_system<T, N, SEP>(a.t, k, lam, s, A, B);
    T rd[N], pmin = (T)0, pmax = (T)0;
    const bool ok = chol_rcp<T, N>(A, rd, pmin, pmax);
    const T cond = ok ? pmax / pmin : (T)0;
    if constexpr (LIST) {
      const int ky = min(j, a.nyp - j), kz = min(k, a.nzp - k);
      const float fx = (float)i * a.rx, fy = (float)ky * a.ry, fz = (float)kz * a.rz;
      const bool lowk = (fx * fx + fy * fy + fz * fz) < a.lowk2;
      if (!ok || lowk || cond > a.cond_thr) {
        const unsigned int slot = atomicAdd(a.list_count, 1u);
        a.list[slot] = (unsigned int)idx;
        continue;
      }
    } else {
      if (!ok) { ++nfail; store_prior(a, idx, s); continue; }
    }
    cmax = fmax(cmax, (double)cond);
    const C* Kk = a.t.K + (size_t)k * N * 3;
    C mu[3], d[N];
#pragma unroll
    for (int m = 0; m < 3; ++m) mu[m] = a.mean[m][idx];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      d[l] = a.data[l][idx];
      C x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mu[m]));
      B[l][3] = x;
    }
    lsolve_r<T, N, 4>(A, rd, B);
    const T sb = SEP ? lam : (T)1;  // K S = sb * B
    C mp[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      C x = cmake<C>((T)0, (T)0);
#pragma unroll
      for (int l = 0; l < N; ++l) x = cadd(x, cmulc(B[l][3], B[l][m]));
      mp[m] = cadd(mu[m], cscale(x, sb));
      a.mean[m][idx] = mp[m];
    }
#pragma unroll
    for (int l = 0; l < N; ++l) {
      C x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mp[m]));
      a.data[l][idx] = x;
    }
    const T sb2 = sb * sb;
    constexpr int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      C x = cmake<C>((T)0, (T)0);
#pragma unroll
      for (int l = 0; l < N; ++l) x = cadd(x, cmulc(B[l][pb[q]], B[l][pa[q]]));
      a.cov[q][idx] = cmake<C>(s[q] - sb2 * x.x, -sb2 * x.y);
    }
  }
  block_stats(a.stats, nfail, cmax);
}

// Main pass of FP32 separable plans (n <= kRecMaxN), same arithmetic as
// k_posterior<float, N, true, LIST>.  Each wave takes 64 consecutive
// wavenumbers of one kz plane, so everything that depends on kz is uniform
// across the wave: the plane's record (E P complex, B 3n complex, K 3n
// complex, D n, S 6, the band flag at float 127) is read with one float2
// per lane and broadcast with readlane instead of per-lane table loads.
// Where the record has room (n <= 5) it also holds the plane's factor of S
// (OL), which the FP64-refine simulation (SIM 2) uses outside the band.
template <int N, bool LIST, int SIM>
__global__ __launch_bounds__(kThreads) void k_posterior_rec(PostArgs<float> a) {
  using C = float2;
  constexpr int P = N * (N + 1) / 2;
  constexpr int OE = 0, OB = 2 * P, OK = OB + 6 * N, OD = OK + 6 * N, OS = OD + N, OL = OS + 6;
  static_assert(OS + 6 <= 127, "record overflow");
  constexpr bool PF = SIM == 2 && OL + 6 <= 127;
  unsigned long long nfail = 0;
  double cmax = 0.0;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwv = (int)gridDim.x * (kThreads / 64);
  const int plane = a.dplr.d, cpp = a.cpp, nch = cpp * a.nzp;
  for (int ch = wv; ch < nch; ch += nwv) {
    const int k = ch / cpp;
    const int p0 = (ch - k * cpp) * 64;
    const float2 rv = a.rec[(size_t)k * 64 + lane];
    auto R = [&](int m) -> float {
      const float v = (m & 1) ? rv.y : rv.x;
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), m >> 1));
    };
    // lanes past the plane's end repeat its last wavenumber and store nothing
    const bool live = p0 + lane < plane;
    const int pp = live ? p0 + lane : plane - 1;
    const int jl = fdiv(pp, a.dcx);
    const int i = pp - jl * a.cnxp, j = a.j0 + jl;
    const long long idx = (long long)i + (long long)a.cnxp * ((long long)j + (long long)a.nyp * k);
    const float lam = a.t.lam[(size_t)j * a.cnxp + i];
    float s[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = lam * R(OS + q);
    if (R(127) == 0.0f) {
      if (live) {
        if (!a.skip_prior_cov) store_prior(a, idx, s);
        if constexpr (PF) {
          float Lp[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) Lp[q] = R(OL + q);
          fused_sim_prior(a, idx, lam, Lp);
        } else if constexpr (SIM > 0) {
          float2 c[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) c[q] = make_float2(s[q], 0.0f);
          fused_sim<SIM>(a, idx, c);
        }
      }
      continue;
    }
    C A[P];
    C B[N][4];
#pragma unroll
    for (int q = 0; q < P; ++q) A[q] = cscale(cmake<C>(R(OE + 2 * q), R(OE + 2 * q + 1)), lam);
#pragma unroll
    for (int l = 0; l < N; ++l)
#pragma unroll
      for (int m = 0; m < 3; ++m) B[l][m] = cmake<C>(R(OB + 2 * (l * 3 + m)), R(OB + 2 * (l * 3 + m) + 1));
#pragma unroll
    for (int l = 0; l < N; ++l) A[l * (l + 1) / 2 + l].x += R(OD + l);
    // the forward matrix, taken while the whole wave is active
    C Kk[N * 3];
#pragma unroll
    for (int q = 0; q < N * 3; ++q) Kk[q] = cmake<C>(R(OK + 2 * q), R(OK + 2 * q + 1));
    float rd[N], pmin = 0.0f, pmax = 0.0f;
    const bool ok = chol_rcp<float, N>(A, rd, pmin, pmax);
    const float cond = ok ? pmax / pmin : 0.0f;
    if constexpr (LIST) {
      const int ky = min(j, a.nyp - j), kz = min(k, a.nzp - k);
      const float fx = (float)i * a.rx, fy = (float)ky * a.ry, fz = (float)kz * a.rz;
      const bool lowk = (fx * fx + fy * fy + fz * fz) < a.lowk2;
      if (!ok || lowk || cond > a.cond_thr) {
        if (live) {
          const unsigned int slot = atomicAdd(a.list_count, 1u);
          a.list[slot] = (unsigned int)idx;
        }
        continue;
      }
    } else {
      if (!ok) {
        if (live) {
          ++nfail;
          store_prior(a, idx, s);
          if constexpr (SIM > 0) {
            float2 c[6];
#pragma unroll
            for (int q = 0; q < 6; ++q) c[q] = make_float2(s[q], 0.0f);
            fused_sim<SIM>(a, idx, c);
          }
        }
        continue;
      }
    }
    if (!live) continue;
    cmax = fmax(cmax, (double)cond);
    C mu[3], d[N];
#pragma unroll
    for (int m = 0; m < 3; ++m) mu[m] = a.mean[m][idx];
#pragma unroll
    for (int l = 0; l < N; ++l) {
      d[l] = a.data[l][idx];
      C x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mu[m]));
      B[l][3] = x;
    }
    lsolve_r<float, N, 4>(A, rd, B);
    const float sb = lam;  // K S = lam * B
    C mp[3];
#pragma unroll
    for (int m = 0; m < 3; ++m) {
      C x = cmake<C>(0.0f, 0.0f);
#pragma unroll
      for (int l = 0; l < N; ++l) x = cadd(x, cmulc(B[l][3], B[l][m]));
      mp[m] = cadd(mu[m], cscale(x, sb));
      a.mean[m][idx] = mp[m];
    }
#pragma unroll
    for (int l = 0; l < N; ++l) {
      C x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mp[m]));
      a.data[l][idx] = x;
    }
    const float sb2 = sb * sb;
    constexpr int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    C cq[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      C x = cmake<C>(0.0f, 0.0f);
#pragma unroll
      for (int l = 0; l < N; ++l) x = cadd(x, cmulc(B[l][pb[q]], B[l][pa[q]]));
      cq[q] = cmake<C>(s[q] - sb2 * x.x, -sb2 * x.y);
      a.cov[q][idx] = cq[q];
    }
    fused_sim<SIM>(a, idx, cq);
  }
  block_stats(a.stats, nfail, cmax);
}

// ---- eigen-basis posterior (FP32 refine mode, separable prior) -------------
// On a plane k of the band M(lam) = lam E + D with D diagonal and positive.
// With F = D^-1/2 E D^-1/2 = U diag(l) U^H (made on the host in FP64),
// T = U^H D^-1/2 and V = T B (B = |ct| K Sigma0): M^-1 = T^H diag(w) T with
// w_p = 1 / (lam l_p + 1), so with r = d - K mu and z = T r
//   mu_post = mu + lam V^H diag(w) z,
//   S_post  = lam S - lam^2 V^H diag(w) V = lam S - lam^2 sum_p w_p G_p,
//   res     = r - K (mu_post - mu),
// where G_p = conj(V_pa) V_pb per covariance q = (a, b).  The solve is
// diagonal in that basis.  The condition estimate is the bound
// cond(D) (lam l_max + 1) / (lam l_min + 1) >= cond(M).
// the record of plane k as scalar (wave-uniform) loads
typedef const __attribute__((address_space(4))) float* EigRecPtr;

template <int N, int SIM>
__global__ __launch_bounds__(kThreads) void k_posterior_eig(PostArgs<float> a) {
  using C = float2;
  using Lo = EigLayout<N>;
  double cmax = 0.0;
  const int lane = threadIdx.x & 63;
  const int wv = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (kThreads / 64) + (threadIdx.x >> 6)));
  const int nwv = (int)gridDim.x * (kThreads / 64);
  const int plane = a.dplr.d, cpp = a.cpp, nch = cpp * a.nzp;
  for (int ch = wv; ch < nch; ch += nwv) {
    const int k = ch / cpp;
    const int p0 = (ch - k * cpp) * 64;
    const EigRecPtr er = (EigRecPtr)(a.erec + (size_t)k * kEigRec);
    // lanes past the plane's end repeat its last wavenumber and store nothing
    const bool live = p0 + lane < plane;
    const int pp = live ? p0 + lane : plane - 1;
    const int jl = fdiv(pp, a.dcx);
    const int i = pp - jl * a.cnxp, j = a.j0 + jl;
    const long long idx = (long long)i + (long long)a.cnxp * ((long long)j + (long long)a.nyp * k);
    const float lam = a.t.lam[(size_t)j * a.cnxp + i];
    float s[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) s[q] = lam * er[Lo::OS + q];
    if (er[Lo::OBAND] == 0.0f) {
      if (live) {
        if (!a.skip_prior_cov) store_prior(a, idx, s);
        if constexpr (SIM == 2) {
          float Lp[6];
#pragma unroll
          for (int q = 0; q < 6; ++q) Lp[q] = er[Lo::OF + q];
          fused_sim_prior(a, idx, lam, Lp);
        }
      }
      continue;
    }
    const EigQueueGeo qg{a.rx, a.ry, a.rz, a.lowk2, a.cond_thr, a.nyp, a.nzp};
    if (eig_queued(qg, lam, er[Lo::OC], er[Lo::OC + 1], er[Lo::OC + 2], i, j, k)) {
      if (live) {
        const unsigned int slot = atomicAdd(a.list_count, 1u);
        a.list[slot] = (unsigned int)idx;
      }
      continue;
    }
    if (!live) continue;
    cmax = fmax(cmax, (double)eig_cond(lam, er[Lo::OC], er[Lo::OC + 1], er[Lo::OC + 2]));
    if (!a.sim_only) {
      C mu[3], d[N], mp[3], res[N];
#pragma unroll
      for (int m = 0; m < 3; ++m) mu[m] = a.mean[m][idx];
#pragma unroll
      for (int l = 0; l < N; ++l) d[l] = a.data[l][idx];
      eig_mean_update<N>(N, er, lam, mu, d, mp, res);
#pragma unroll
      for (int m = 0; m < 3; ++m) a.mean[m][idx] = mp[m];
#pragma unroll
      for (int l = 0; l < N; ++l) a.data[l][idx] = res[l];
    }
    float w[N];
#pragma unroll
    for (int p = 0; p < N; ++p) w[p] = eig_weight(lam, er[Lo::OW + p]);
    // covariances lam S - lam^2 sum_p w_p G_p
    const float l2 = lam * lam;
    C cq[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      float gr[N], gi[N];
#pragma unroll
      for (int p = 0; p < N; ++p) {
        gr[p] = er[Lo::OG + 9 * p + (q < 3 ? q : 3 + 2 * (q - 3))];
        gi[p] = q < 3 ? 0.0f : er[Lo::OG + 9 * p + 4 + 2 * (q - 3)];
      }
      cq[q] = eig_cov_entry<N>(N, q < 3, s[q], l2, w, gr, gi);
    }
    if (!a.skip_band_cov) {
#pragma unroll
      for (int q = 0; q < 6; ++q) a.cov[q][idx] = cq[q];
    }
    fused_sim<SIM>(a, idx, cq);
  }
  block_stats(a.stats, 0ull, cmax);
}

// FP64 solve with one step of iterative refinement for the queued
// wavenumbers (FP32 storage): per right-hand side b (columns of K S, then
// d - K mu), x = M^-1 b, r = b - M x, x += M^-1 r, with M built in FP64 from
// the FP64 tables.
template <int N, bool SEP, int SIM = 0>
__global__ __launch_bounds__(kThreads) void k_posterior_list64(PostArgs<float> a) {
  using D = double2;
  constexpr int P = N * (N + 1) / 2;
  unsigned long long nfail = 0;
  double cmax = 0.0;
  const unsigned int n = *a.list_count;
  for (unsigned int t = blockIdx.x * kThreads + threadIdx.x; t < n; t += gridDim.x * kThreads) {
    const long long idx = a.list[t];
    int i, j, k;
    decode(a, idx, i, j, k);
    double lam, s[6];
    if constexpr (SEP) {
      lam = a.t64.lam[(size_t)j * a.cnxp + i];
#pragma unroll
      for (int q = 0; q < 6; ++q) s[q] = lam * a.t64.S[(size_t)k * 6 + q];
    } else {
      gather_general<float, double>(a, idx, a.sig0d, s, lam);
    }
    D M[P], B[N][4];
    build_system<double, N, SEP>(a.t64, k, lam, s, M, B);
    if constexpr (SEP) {
#pragma unroll
      for (int l = 0; l < N; ++l)
#pragma unroll
        for (int m = 0; m < 3; ++m) B[l][m] = cscale(B[l][m], lam);
    }
    D A[P];
#pragma unroll
    for (int q = 0; q < P; ++q) A[q] = M[q];
    double rd[N], pmin = 0.0, pmax = 0.0;
    if (!chol_rcp<double, N>(A, rd, pmin, pmax)) {
      ++nfail;
      store_prior(a, idx, s);
      if constexpr (SIM > 0) {
        float2 c[6];
#pragma unroll
        for (int q = 0; q < 6; ++q) c[q] = make_float2((float)s[q], 0.0f);
        fused_sim<SIM>(a, idx, c);
      }
      continue;
    }
    cmax = fmax(cmax, pmax / pmin);
    const D* Kk = a.t64.K + (size_t)k * N * 3;
    D mu[3], d[N];
#pragma unroll
    for (int m = 0; m < 3; ++m) { const float2 v = a.mean[m][idx]; mu[m] = cmake<D>(v.x, v.y); }
#pragma unroll
    for (int l = 0; l < N; ++l) {
      const float2 v = a.data[l][idx];
      d[l] = cmake<D>(v.x, v.y);
      D x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mu[m]));
      B[l][3] = x;
    }
    D cv[6], mp[3];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      D x[N][1], y[N][1];
#pragma unroll
      for (int l = 0; l < N; ++l) x[l][0] = B[l][c];
      lsolve_r<double, N, 1>(A, rd, x);
      lhsolve_r<double, N, 1>(A, rd, x);
#pragma unroll
      for (int l = 0; l < N; ++l) {
        D r = B[l][c];
#pragma unroll
        for (int m = 0; m < N; ++m) {
          const D mm = (m <= l) ? M[l * (l + 1) / 2 + m] : cconj(M[m * (m + 1) / 2 + l]);
          r = csub(r, cmul(mm, x[m][0]));
        }
        y[l][0] = r;
      }
      lsolve_r<double, N, 1>(A, rd, y);
      lhsolve_r<double, N, 1>(A, rd, y);
#pragma unroll
      for (int l = 0; l < N; ++l) x[l][0] = cadd(x[l][0], y[l][0]);
      if (c < 3) {
        // (B^H M^-1 B)[aa][c] for aa <= c
#pragma unroll
        for (int aa = 0; aa <= c; ++aa) {
          D acc = cmake<D>(0.0, 0.0);
#pragma unroll
          for (int l = 0; l < N; ++l) acc = cadd(acc, cmulc(x[l][0], B[l][aa]));
          const int q = (aa == c) ? c : (c == 1 ? 3 : (aa == 0 ? 4 : 5));
          cv[q] = acc;
        }
      } else {
#pragma unroll
        for (int m = 0; m < 3; ++m) {
          D acc = mu[m];
#pragma unroll
          for (int l = 0; l < N; ++l) acc = cadd(acc, cmulc(x[l][0], B[l][m]));
          mp[m] = acc;
        }
      }
    }
#pragma unroll
    for (int m = 0; m < 3; ++m) a.mean[m][idx] = make_float2((float)mp[m].x, (float)mp[m].y);
#pragma unroll
    for (int l = 0; l < N; ++l) {
      D x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) x = csub(x, cmul(Kk[l * 3 + m], mp[m]));
      a.data[l][idx] = make_float2((float)x.x, (float)x.y);
    }
    float2 cq[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      cq[q] = make_float2((float)(s[q] - cv[q].x), (float)(-cv[q].y));
      a.cov[q][idx] = cq[q];
    }
    fused_sim<SIM>(a, idx, cq);
  }
  block_stats(a.stats, nfail, cmax);
}

// Persistent grid: as many blocks as fit on the device at once.
template <typename F>
int persistent_blocks(crv_ctx* ctx, F kernel, long long work) {
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kThreads, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  if (ctx->post_blocks_per_cu > 0) per_cu = ctx->post_blocks_per_cu;
  const long long want = (work + kThreads - 1) / kThreads;
  return (int)std::max<long long>(1, std::min<long long>(want, (long long)per_cu * ctx->num_cu));
}

template <typename T, int N, bool SEP, bool LIST>
int launch_main(crv_ctx* ctx, const PostArgs<T>& a) {
  if constexpr (std::is_same<T, float>::value && SEP && LIST && N <= kRecMaxN) {
    if (a.erec) {
      auto kern = a.nsim ? k_posterior_eig<N, 2> : k_posterior_eig<N, 0>;
      const int nb = persistent_blocks(ctx, kern, a.ncr);
      ProfScope ps(ctx, a.nsim ? "posterior_sim_fp32" : "posterior_fp32");
      hipLaunchKernelGGL(kern, dim3(nb), dim3(kThreads), 0, ctx->stream, a);
      CRV_HIP_CHECK(ctx, hipGetLastError());
      return CRV_OK;
    }
  }
  if constexpr (std::is_same<T, float>::value && SEP && N <= kRecMaxN) {
    if (a.rec) {
      auto go = [&](auto kern) {
        const int nb = persistent_blocks(ctx, kern, a.ncr);
        ProfScope ps(ctx, a.nsim ? (LIST ? "posterior_sim_fp32" : "posterior_sim")
                                 : (LIST ? "posterior_fp32" : "posterior"));
        hipLaunchKernelGGL(kern, dim3(nb), dim3(kThreads), 0, ctx->stream, a);
        return hipGetLastError();
      };
      hipError_t e;
      if (a.nsim == 0) e = go(k_posterior_rec<N, LIST, 0>);
      else if (LIST) e = go(k_posterior_rec<N, LIST, 2>);
      else e = go(k_posterior_rec<N, LIST, 1>);
      if (e != hipSuccess) return crv_set_error(ctx, CRV_E_HIP, std::string("posterior: ") + hipGetErrorString(e));
      return CRV_OK;
    }
  }
  const int nb = persistent_blocks(ctx, k_posterior<T, N, SEP, LIST>, a.ncr);
  ProfScope ps(ctx, LIST ? "posterior_fp32" : "posterior");
  hipLaunchKernelGGL((k_posterior<T, N, SEP, LIST>), dim3(nb), dim3(kThreads), 0, ctx->stream, a);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

template <int N, bool SEP>
int launch_list(crv_ctx* ctx, const PostArgs<float>& a) {
  ProfScope ps(ctx, "posterior_fp64_list");
  if (a.nsim)
    hipLaunchKernelGGL((k_posterior_list64<N, SEP, 2>), dim3(256), dim3(kThreads), 0, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_posterior_list64<N, SEP>), dim3(256), dim3(kThreads), 0, ctx->stream, a);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

template <typename T, bool SEP, bool LIST>
int dispatch(crv_ctx* ctx, int n, const PostArgs<T>& a) {
  int rc = CRV_OK;
  switch (n) {
#define CASE(NN)                                                                          \
  case NN:                                                                                \
    rc = launch_main<T, NN, SEP, LIST>(ctx, a);                                           \
    if constexpr (LIST && std::is_same<T, float>::value)                                  \
      if (rc == CRV_OK) rc = launch_list<NN, SEP>(ctx, a);                                \
    break;
    CASE(1) CASE(2) CASE(3) CASE(4) CASE(5) CASE(6) CASE(7) CASE(8)
#undef CASE
    default: return crv_set_error(ctx, CRV_E_UNSUPPORTED, "ntheta must be 1..8");
  }
  return rc;
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

// Host-side FP64 tables, interleaved re/im for complex entries.
struct HostTables {
  std::vector<double> K, E, B, D, S, lam;
  std::vector<int> inv;
};

// Serialises the tables in precision U into blob, recording offsets.
struct TabOffsets { size_t K = 0, E = 0, B = 0, D = 0, S = 0, lam = 0; };

template <typename U>
TabOffsets put_tables(std::vector<unsigned char>& blob, const HostTables& h) {
  TabOffsets o;
  auto put = [&](const std::vector<double>& v) {
    const size_t at = blob.size();
    const size_t bytes = (v.size() * sizeof(U) + 255) & ~(size_t)255;
    blob.resize(at + std::max<size_t>(bytes, 256), 0);
    U* dst = (U*)(blob.data() + at);
    for (size_t q = 0; q < v.size(); ++q) dst[q] = (U)v[q];
    return at;
  };
  o.K = put(h.K); o.E = put(h.E); o.B = put(h.B); o.D = put(h.D); o.S = put(h.S); o.lam = put(h.lam);
  return o;
}

template <typename U>
KzTab<U> tab_at(unsigned char* base, const TabOffsets& o) {
  KzTab<U> t;
  t.K = (const cplx_t<U>*)(base + o.K); t.E = (const cplx_t<U>*)(base + o.E); t.B = (const cplx_t<U>*)(base + o.B);
  t.D = (const U*)(base + o.D); t.S = (const U*)(base + o.S); t.lam = (const U*)(base + o.lam);
  return t;
}

__global__ void k_real_to_complex(const float* __restrict__ x, float2* __restrict__ z, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
    z[i] = make_float2(x[i], 0.0f);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = 0ull;
}

int validate_params(crv_ctx* ctx, const crv_post_params* p) {
  if (!p) return crv_set_error(ctx, CRV_E_INVALID, "NULL parameters");
  if (p->ntheta < 1 || p->ntheta > 8) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "ntheta must be 1..8");
  if (!p->A || !p->err_theta_cov || !p->kwd || !p->em1 || !p->em2 || !p->em3 || !p->sigma0)
    return crv_set_error(ctx, CRV_E_INVALID, "missing table");
  if (p->separable && (!p->cxy || !p->ct)) return crv_set_error(ctx, CRV_E_INVALID, "separable prior without cxy/ct");
  return CRV_OK;
}

// Factor L (packed lower 00 10 11 20 21 22) of a plane's prior S as the
// FP32 record holds it (aa bb rr ab ar br), in extended precision with
// lib_matrCholCpx's acceptance rule (lib_matr.c:78-167, the rule of
// chol3_f64); all zero where S is rejected.
void prior_factor(const float* S, float* L) {
  using E = long double;
  for (int q = 0; q < 6; ++q) L[q] = 0.0f;
  const E p00 = S[0], p10 = S[3], p11 = S[1], p20 = S[4], p21 = S[5], p22 = S[2];
  if (p00 <= 0) return;
  const E tol = (E)kCholTol * p00;
  if (p00 <= tol) return;
  const E l00 = sqrtl(p00), l10 = p10 / l00, l20 = p20 / l00;
  const E d1 = p11 - l10 * l10;
  if (d1 <= tol) return;
  const E l11 = sqrtl(d1), l21 = (p21 - l20 * l10) / l11;
  const E d2 = p22 - l20 * l20 - l21 * l21;
  if (d2 <= tol) return;
  const E l22 = sqrtl(d2);
  const E v[6] = {l00, l10, l11, l20, l21, l22};
  for (int q = 0; q < 6; ++q) L[q] = (float)v[q];
}

// kz tables (crava.cpp:1104-1121, 1336-1364, 1441-1469) in FP64
HostTables build_tables(const crv_post_params* p, int nyp, int nzp, int cnxp, long long& ninv_planes) {
  const int n = p->ntheta;
  const int P = n * (n + 1) / 2;
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  HostTables h;
  h.K.assign((size_t)nzp * n * 3 * 2, 0.0);
  h.E.assign((size_t)nzp * P * 2, 0.0);
  h.D.assign((size_t)nzp * n, 0.0);
  h.inv.assign(nzp, 0);
  if (p->separable) {
    h.B.assign((size_t)nzp * n * 3 * 2, 0.0);
    h.S.assign((size_t)nzp * 6, 0.0);
    h.lam.resize((size_t)cnxp * nyp);
    for (size_t q = 0; q < h.lam.size(); ++q) h.lam[q] = std::fabs(p->cxy[q]);
  }
  const double wnc = p->wnc;
  ninv_planes = 0;
  for (int k = 0; k < nzp; ++k) {
    // crava.cpp:1104-1105: float numerator over a double product, compared
    // with lowCut * minRelThick in double and with highCut in float
    const float fr = (float)((double)(p->nz * 1000.0f) / (p->lz * nzp) * (double)std::min(k, nzp - k));
    h.inv[k] = ((double)fr > p->low_cut * p->min_rel_thick && fr < (float)p->high_cut) ? 1 : 0;
    ninv_planes += h.inv[k];
    for (int l = 0; l < n; ++l) {
      const double wr = p->kwd[((size_t)k * n + l) * 2], wi = p->kwd[((size_t)k * n + l) * 2 + 1];
      for (int m = 0; m < 3; ++m) {
        const double am = p->A[l * 3 + m];
        h.K[(((size_t)k * n + l) * 3 + m) * 2] = wr * am;
        h.K[(((size_t)k * n + l) * 3 + m) * 2 + 1] = wi * am;
      }
      const double* e1 = p->em1 + ((size_t)k * n) * 2;
      const double* e2 = p->em2 + ((size_t)k * n) * 2;
      const double* e3 = p->em3 + ((size_t)k * n) * 2;
      for (int m = 0; m <= l; ++m) {
        const double th = p->err_theta_cov[l * n + m];
        const double c = 0.5 * (1.0 - wnc) * th;
        // e_l * conj(e_m), summed over the two error shapes
        const double re = e1[2 * l] * e1[2 * m] + e1[2 * l + 1] * e1[2 * m + 1] + e2[2 * l] * e2[2 * m] +
                          e2[2 * l + 1] * e2[2 * m + 1];
        const double im = -e1[2 * l] * e1[2 * m + 1] + e1[2 * l + 1] * e1[2 * m] - e2[2 * l] * e2[2 * m + 1] +
                          e2[2 * l + 1] * e2[2 * m];
        const size_t o = ((size_t)k * P + l * (l + 1) / 2 + m) * 2;
        h.E[o] = c * re;
        h.E[o + 1] = (l == m) ? 0.0 : c * im;
      }
      h.D[(size_t)k * n + l] = wnc * p->err_theta_cov[l * n + l] * e3[2 * l] * e3[2 * l];
    }
    if (p->separable) {
      // |ct| Sigma0, |ct| K Sigma0 and |ct| K Sigma0 K^H + E
      const double act = std::fabs(p->ct[k]);
      for (int q = 0; q < 6; ++q) h.S[(size_t)k * 6 + q] = act * p->sigma0[pa[q] * 3 + pb[q]];
      const double* Kk = h.K.data() + (size_t)k * n * 3 * 2;
      double* Bk = h.B.data() + (size_t)k * n * 3 * 2;
      for (int l = 0; l < n; ++l)
        for (int m = 0; m < 3; ++m) {
          double xr = 0.0, xi = 0.0;
          for (int q = 0; q < 3; ++q) {
            const double sv = act * p->sigma0[q * 3 + m];
            xr += Kk[(l * 3 + q) * 2] * sv;
            xi += Kk[(l * 3 + q) * 2 + 1] * sv;
          }
          Bk[(l * 3 + m) * 2] = xr;
          Bk[(l * 3 + m) * 2 + 1] = xi;
        }
      for (int l = 0; l < n; ++l)
        for (int m = 0; m <= l; ++m) {
          double xr = 0.0, xi = 0.0;  // sum_q B[l][q] conj(K[m][q])
          for (int q = 0; q < 3; ++q) {
            const double br = Bk[(l * 3 + q) * 2], bi = Bk[(l * 3 + q) * 2 + 1];
            const double kr = Kk[(m * 3 + q) * 2], ki = Kk[(m * 3 + q) * 2 + 1];
            xr += br * kr + bi * ki;
            xi += bi * kr - br * ki;
          }
          const size_t o = ((size_t)k * P + l * (l + 1) / 2 + m) * 2;
          h.E[o] += xr;
          h.E[o + 1] = (l == m) ? 0.0 : h.E[o + 1] + xi;
        }
    }
  }
  return h;
}

}  // namespace

// Eigen-decomposition A = U diag(l) U^H of a Hermitian n x n matrix (row
// major, overwritten) by cyclic Jacobi rotations; false if it does not
// converge (non-finite input).
bool herm_eig(int n, std::vector<std::complex<double>>& A, std::vector<double>& l,
              std::vector<std::complex<double>>& U) {
  using Cx = std::complex<double>;
  U.assign((size_t)n * n, Cx(0.0, 0.0));
  for (int p = 0; p < n; ++p) U[(size_t)p * n + p] = 1.0;
  bool done = false;
  for (int sweep = 0; sweep < 60 && !done; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int p = 0; p < n; ++p)
      for (int q = 0; q < n; ++q) (p == q ? dia : off) += std::norm(A[(size_t)p * n + q]);
    if (!std::isfinite(off) || !std::isfinite(dia)) return false;
    if (off <= 1e-30 * dia) { done = true; break; }
    for (int p = 0; p < n; ++p)
      for (int q = p + 1; q < n; ++q) {
        const Cx b = A[(size_t)p * n + q];
        const double ab = std::abs(b);
        if (ab == 0.0) continue;
        // J = diag(1, conj(ph)) [[c, s], [-s, c]] on (p, q) makes A_pq zero
        const Cx ph = b / ab;
        const double tau = (A[(size_t)q * n + q].real() - A[(size_t)p * n + p].real()) / (2.0 * ab);
        const double t = (tau >= 0.0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        auto cols = [&](std::vector<Cx>& X) {
          for (int r = 0; r < n; ++r) {
            const Cx x = X[(size_t)r * n + p], y = X[(size_t)r * n + q];
            X[(size_t)r * n + p] = c * x - sn * std::conj(ph) * y;
            X[(size_t)r * n + q] = sn * x + c * std::conj(ph) * y;
          }
        };
        cols(A);
        for (int r = 0; r < n; ++r) {
          const Cx x = A[(size_t)p * n + r], y = A[(size_t)q * n + r];
          A[(size_t)p * n + r] = c * x - sn * ph * y;
          A[(size_t)q * n + r] = sn * x + c * ph * y;
        }
        cols(U);
      }
  }
  if (!done) return false;
  l.resize(n);
  for (int p = 0; p < n; ++p) l[p] = A[(size_t)p * n + p].real();
  return true;
}

// Eigen-basis records (k_posterior_eig's EigLayout) of every plane: S, its
// factor and K on all planes, T, V, G, l and the condition data on the
// band's.  false (no records) where a band plane's noise diagonal D is not
// positive or its decomposition fails.
bool eig_records(const HostTables& h, int n, int nzp, std::vector<float>& out) {
  using Cx = std::complex<double>;
  const int P = n * (n + 1) / 2;
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  const int OT = 0, OV = OT + 2 * n * n, OK = OV + 6 * n, OG = OK + 6 * n, OW = OG + 9 * n, OS = OW + n,
            OF = OS + 6, OC = OF + 6;
  if (OC + 3 > kEigRec - 1) return false;
  out.assign((size_t)nzp * kEigRec, 0.0f);
  std::vector<double> dh(n), lv;
  std::vector<Cx> F((size_t)n * n), U, T((size_t)n * n), V((size_t)n * 3);
  for (int k = 0; k < nzp; ++k) {
    float* rk = out.data() + (size_t)k * kEigRec;
    for (int q = 0; q < 6; ++q) rk[OS + q] = (float)h.S[(size_t)k * 6 + q];
    prior_factor(rk + OS, rk + OF);
    for (int q = 0; q < 6 * n; ++q) rk[OK + q] = (float)h.K[(size_t)k * 6 * n + q];
    if (!h.inv[k]) continue;
    rk[kEigRec - 1] = 1.0f;
    double dmin = 0.0, dmax = 0.0;
    for (int l = 0; l < n; ++l) {
      const double d = h.D[(size_t)k * n + l];
      if (!(d > 0.0) || !std::isfinite(d)) return false;
      dh[l] = 1.0 / std::sqrt(d);
      dmin = l ? std::min(dmin, d) : d;
      dmax = l ? std::max(dmax, d) : d;
    }
    for (int l = 0; l < n; ++l)
      for (int m = 0; m <= l; ++m) {
        const size_t o = ((size_t)k * P + l * (l + 1) / 2 + m) * 2;
        const Cx f = dh[l] * Cx(h.E[o], l == m ? 0.0 : h.E[o + 1]) * dh[m];
        F[(size_t)l * n + m] = f;
        F[(size_t)m * n + l] = std::conj(f);
      }
    if (!herm_eig(n, F, lv, U)) return false;
    for (int p = 0; p < n; ++p)
      for (int l = 0; l < n; ++l) T[(size_t)p * n + l] = std::conj(U[(size_t)l * n + p]) * dh[l];
    for (int p = 0; p < n; ++p)
      for (int m = 0; m < 3; ++m) {
        Cx v = 0.0;
        for (int l = 0; l < n; ++l) {
          const size_t o = (((size_t)k * n + l) * 3 + m) * 2;
          v += T[(size_t)p * n + l] * Cx(h.B[o], h.B[o + 1]);
        }
        V[(size_t)p * 3 + m] = v;
      }
    double lmin = 0.0, lmax = 0.0;
    for (int p = 0; p < n; ++p) {
      const double lp = std::max(lv[p], 0.0);
      lmin = p ? std::min(lmin, lp) : lp;
      lmax = p ? std::max(lmax, lp) : lp;
      rk[OW + p] = (float)lp;
      for (int l = 0; l < n; ++l) {
        rk[OT + 2 * (p * n + l)] = (float)T[(size_t)p * n + l].real();
        rk[OT + 2 * (p * n + l) + 1] = (float)T[(size_t)p * n + l].imag();
      }
      for (int m = 0; m < 3; ++m) {
        rk[OV + 2 * (p * 3 + m)] = (float)V[(size_t)p * 3 + m].real();
        rk[OV + 2 * (p * 3 + m) + 1] = (float)V[(size_t)p * 3 + m].imag();
      }
      for (int q = 0; q < 3; ++q) rk[OG + 9 * p + q] = (float)std::norm(V[(size_t)p * 3 + q]);
      for (int q = 3; q < 6; ++q) {
        const Cx g = std::conj(V[(size_t)p * 3 + pa[q]]) * V[(size_t)p * 3 + pb[q]];
        rk[OG + 9 * p + 3 + 2 * (q - 3)] = (float)g.real();
        rk[OG + 9 * p + 4 + 2 * (q - 3)] = (float)g.imag();
      }
    }
    rk[OC] = (float)(dmax / dmin);
    rk[OC + 1] = (float)lmax;
    rk[OC + 2] = (float)lmin;
  }
  return true;
}

// True when the eigen-basis posterior can queue no wavenumber of the grid
// for the FP64 solve (eig_queued false everywhere), shown on the host from
// the tables the kernels read: the low-|k| test is smallest at i = 0, ky = 0
// and the plane's smallest |kz|, where the kernel's r2 is exactly fz * fz;
// the condition bound cond(D) (lam l_max + 1) / (lam l_min + 1) grows with
// lam (l_max >= l_min >= 0), so it is largest at the table's largest lam,
// where it must stay a relative 1e-5 below the threshold (far more than the
// few float roundings of the kernel's test).
bool eig_never_queued(const std::vector<float>& er, const std::vector<int>& inv, const std::vector<double>& lam,
                      int n, int nzp, float lowk2, float thr) {
  const int OC = 2 * n * n + 22 * n + 12;
  float lmax_tab = 0.0f;
  for (double v : lam) {
    const float f = (float)v;
    if (!std::isfinite(f) || f < 0.0f) return false;
    lmax_tab = std::max(lmax_tab, f);
  }
  const float rz = 2.0f / (float)nzp;
  for (int k = 0; k < nzp; ++k) {
    if (!inv[k]) continue;
    const int kz = std::min(k, nzp - k);
    const float fz = (float)kz * rz;
    if (fz * fz < lowk2) return false;
    const float* rk = er.data() + (size_t)k * kEigRec;
    const double cd = rk[OC], lmx = rk[OC + 1], lmn = rk[OC + 2];
    if (!(lmn >= 0.0) || !(lmx >= lmn) || !std::isfinite(cd)) return false;
    const double bound = cd * ((double)lmax_tab * lmx + 1.0) / ((double)lmax_tab * lmn + 1.0);
    if (!(bound <= (double)thr * (1.0 - 1e-5))) return false;
  }
  return true;
}

struct crv_post_plan {
  crv_ctx* ctx = nullptr;
  int n = 0, nxp = 0, nyp = 0, nzp = 0, cnxp = 0, precision = 0;
  bool separable = false, refine = false, async = false;
  long long nc = 0, ninv_planes = 0;
  unsigned char* dtab = nullptr;  // [tables T][tables FP64 (refine)][inv][stats 256 B]
  TabOffsets oT, o64;
  size_t oI = 0, oS = 0, oR = 0;   // oR: FP32 separable records (0: none)
  size_t oE = 0;                   // eigen-basis records (0: none)
  // with eigen-basis records: the band planes' tables of the generating z
  // inverse (CovGen): row index per plane, then 6 x nband rows of gw floats
  size_t oGI = 0, oGR = 0;
  int nband = 0, gw = 0;
  // with eigen-basis records: the band planes (nband ints), and whether no
  // wavenumber of the grid can be queued for the FP64 solve (zpost_ok), in
  // which case crv_invert may run the posterior of the means inside its z
  // pass (crv_launch_zpost)
  size_t oBP = 0;
  bool zpost_ok = false;
  // with eigen-basis records (plans made by crv_post_plan_create): the y
  // inverse transform of lam as a complex plane (nyp x cnxp), from which the
  // covariances' z inverse generates the planes outside the band after the
  // band's y pass (crv_launch_cov_band_y)
  float2* dlamy = nullptr;
  unsigned int* dlist = nullptr;
  double sig0[6] = {0, 0, 0, 0, 0, 0};
  float cond_thr = 1e5f;
  double lowk2 = 0.0;
};

namespace {

int plan_make(crv_ctx* ctx, const crv_post_params* p, int nxp, int nyp, int nzp, int precision, bool async,
              crv_post_plan** out) {
  int rc = validate_params(ctx, p);
  if (rc != CRV_OK) return rc;
  if (precision != CRV_FP32 && precision != CRV_FP64) return crv_set_error(ctx, CRV_E_INVALID, "bad precision");
  const bool refine = (p->mode == CRV_MODE_FP32_REFINE);
  if (refine && precision != CRV_FP32) return crv_set_error(ctx, CRV_E_INVALID, "CRV_MODE_FP32_REFINE needs FP32 grids");
  crv_post_plan* pl = new (std::nothrow) crv_post_plan();
  if (!pl) return CRV_E_NOMEM;
  pl->ctx = ctx; pl->n = p->ntheta; pl->nxp = nxp; pl->nyp = nyp; pl->nzp = nzp; pl->cnxp = nxp / 2 + 1;
  pl->precision = precision; pl->separable = p->separable != 0; pl->refine = refine; pl->async = async;
  pl->nc = (long long)pl->cnxp * nyp * nzp;
  if (pl->nc >= (1LL << 31)) {
    delete pl;
    return crv_set_error(ctx, CRV_E_UNSUPPORTED, "posterior: grid of 2^31 or more wavenumbers");
  }
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) pl->sig0[q] = p->sigma0[pa[q] * 3 + pb[q]];
  pl->cond_thr = (float)(p->cond_threshold > 0 ? p->cond_threshold : 1e5);
  pl->lowk2 = p->lowk_fraction * p->lowk_fraction;
  const HostTables h = build_tables(p, nyp, nzp, pl->cnxp, pl->ninv_planes);
  std::vector<unsigned char> blob;
  pl->oT = (precision == CRV_FP64) ? put_tables<double>(blob, h) : put_tables<float>(blob, h);
  if (refine) pl->o64 = put_tables<double>(blob, h);
  if (precision == CRV_FP32 && pl->separable && p->ntheta <= kRecMaxN) {
    // one 128-float record per kz: E, B, K, D, S (the FP32 tables' values)
    // and the band flag at float 127
    const int n = p->ntheta, P = n * (n + 1) / 2;
    pl->oR = blob.size();
    blob.resize(pl->oR + (size_t)nzp * 128 * sizeof(float), 0);
    float* r = (float*)(blob.data() + pl->oR);
    for (int k = 0; k < nzp; ++k) {
      float* rk = r + (size_t)k * 128;
      int o = 0;
      for (int q = 0; q < 2 * P; ++q) rk[o++] = (float)h.E[(size_t)k * 2 * P + q];
      for (int q = 0; q < 6 * n; ++q) rk[o++] = (float)h.B[(size_t)k * 6 * n + q];
      for (int q = 0; q < 6 * n; ++q) rk[o++] = (float)h.K[(size_t)k * 6 * n + q];
      for (int q = 0; q < n; ++q) rk[o++] = (float)h.D[(size_t)k * n + q];
      for (int q = 0; q < 6; ++q) rk[o++] = (float)h.S[(size_t)k * 6 + q];
      if (o + 6 <= 127) prior_factor(rk + o - 6, rk + o);
      rk[127] = h.inv[k] ? 1.0f : 0.0f;
    }
  }
  if (precision == CRV_FP32 && pl->separable && refine && p->ntheta <= kRecMaxN) {
    std::vector<float> er;
    if (eig_records(h, p->ntheta, nzp, er)) {
      pl->oE = (blob.size() + 255) & ~(size_t)255;
      blob.resize(pl->oE + er.size() * sizeof(float), 0);
      memcpy(blob.data() + pl->oE, er.data(), er.size() * sizeof(float));
      // the generating z inverse's copy of the band planes' values: the same
      // floats as the records, per covariance q: S_q, cond(D), l_max, l_min,
      // then per eigenvalue p: l_p, Re G_pq, Im G_pq, 0 (CovGen; groups of
      // four floats, read as float4)
      const int n = p->ntheta;
      const int OG = 2 * n * n + 12 * n, OW = OG + 9 * n, OS = OW + n, OC = OS + 12;
      std::vector<int> bidx(nzp, -1);
      int nb = 0;
      for (int k = 0; k < nzp; ++k)
        if (h.inv[k]) bidx[k] = nb++;
      const int gw = 4 + 4 * n;
      std::vector<float> rows((size_t)6 * nb * gw, 0.0f);
      for (int k = 0; k < nzp; ++k) {
        if (bidx[k] < 0) continue;
        const float* rk = er.data() + (size_t)k * kEigRec;
        for (int q = 0; q < 6; ++q) {
          float* t = rows.data() + ((size_t)q * nb + bidx[k]) * gw;
          t[0] = rk[OS + q];
          t[1] = rk[OC]; t[2] = rk[OC + 1]; t[3] = rk[OC + 2];
          for (int pp = 0; pp < n; ++pp) {
            t[4 + 4 * pp] = rk[OW + pp];
            t[5 + 4 * pp] = rk[OG + 9 * pp + (q < 3 ? q : 3 + 2 * (q - 3))];
            t[6 + 4 * pp] = q < 3 ? 0.0f : rk[OG + 9 * pp + 4 + 2 * (q - 3)];
          }
        }
      }
      pl->nband = nb; pl->gw = gw;
      pl->oGI = (blob.size() + 255) & ~(size_t)255;
      blob.resize(pl->oGI + (size_t)nzp * sizeof(int), 0);
      memcpy(blob.data() + pl->oGI, bidx.data(), (size_t)nzp * sizeof(int));
      pl->oGR = (blob.size() + 255) & ~(size_t)255;
      blob.resize(pl->oGR + rows.size() * sizeof(float), 0);
      memcpy(blob.data() + pl->oGR, rows.data(), rows.size() * sizeof(float));
      std::vector<int> bplanes;
      for (int k = 0; k < nzp; ++k)
        if (h.inv[k]) bplanes.push_back(k);
      pl->oBP = (blob.size() + 255) & ~(size_t)255;
      blob.resize(pl->oBP + std::max<size_t>(1, bplanes.size()) * sizeof(int), 0);
      if (!bplanes.empty()) memcpy(blob.data() + pl->oBP, bplanes.data(), bplanes.size() * sizeof(int));
      pl->zpost_ok = nb > 0 && eig_never_queued(er, h.inv, h.lam, n, nzp, (float)pl->lowk2, pl->cond_thr);
    }
  }
  pl->oI = blob.size();
  blob.resize(pl->oI + (((size_t)nzp * 4 + 255) & ~(size_t)255), 0);
  memcpy(blob.data() + pl->oI, h.inv.data(), (size_t)nzp * 4);
  pl->oS = blob.size();
  blob.resize(pl->oS + 256, 0);
  hipError_t e;
  if (async) {
    e = hipMallocAsync((void**)&pl->dtab, blob.size(), ctx->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(pl->dtab, blob.data(), blob.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess && refine) e = hipMallocAsync((void**)&pl->dlist, (size_t)pl->nc * 4, ctx->stream);
  } else {
    e = hipMalloc((void**)&pl->dtab, blob.size());
    if (e == hipSuccess) e = hipMemcpy(pl->dtab, blob.data(), blob.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess && refine) e = hipMalloc((void**)&pl->dlist, (size_t)pl->nc * 4);
  }
  if (e != hipSuccess) {
    crv_post_plan_destroy(pl);
    return crv_set_error(ctx, CRV_E_HIP, std::string("posterior plan: ") + hipGetErrorString(e));
  }
  if (!async && pl->oE && pl->oGR && pl->oBP && precision == CRV_FP32) {
    // lam's y inverse transform, with the transforms' own y pass
    FftPlan* fp = nullptr;
    const size_t nl = (size_t)pl->cnxp * nyp;
    if ((rc = crv_get_plan(ctx, nxp, nyp, nzp, CRV_FP32, &fp)) == CRV_OK && crv_cov_band_y_fits(fp)) {
      e = hipMalloc((void**)&pl->dlamy, nl * sizeof(float2));
      if (e == hipSuccess) {
        hipLaunchKernelGGL(k_real_to_complex, dim3((unsigned)std::min<size_t>((nl + 255) / 256, 4096)), dim3(256), 0,
                           ctx->stream, tab_at<float>(pl->dtab, pl->oT).lam, pl->dlamy, (long long)nl);
        e = hipGetLastError();
      }
      if (e == hipSuccess) rc = crv_launch_pass(ctx, fp, 4, pl->dlamy, pl->dlamy, 1, 0, 0, 1, 1.0, nullptr, 0);
      if (e == hipSuccess && rc == CRV_OK) e = hipStreamSynchronize(ctx->stream);
      if (e != hipSuccess || rc != CRV_OK) {
        crv_post_plan_destroy(pl);
        return rc != CRV_OK ? rc : crv_set_error(ctx, CRV_E_HIP, std::string("posterior plan: ") + hipGetErrorString(e));
      }
    }
  }
  *out = pl;
  return CRV_OK;
}

}  // namespace

extern "C" int crv_post_plan_create(crv_ctx* ctx, const crv_post_params* p, int nxp, int nyp, int nzp,
                                    int precision, crv_post_plan** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!out || nxp < 1 || nyp < 1 || nzp < 1) return crv_set_error(ctx, CRV_E_INVALID, "post_plan_create: bad args");
  *out = nullptr;
  return plan_make(ctx, p, nxp, nyp, nzp, precision, false, out);
}

extern "C" int crv_post_plan_destroy(crv_post_plan* pl) {
  if (!pl) return CRV_E_INVALID;
  crv_ctx* ctx = pl->ctx;
  if (pl->async) {
    if (pl->dtab) (void)hipFreeAsync(pl->dtab, ctx->stream);
    if (pl->dlist) (void)hipFreeAsync(pl->dlist, ctx->stream);
  } else {
    (void)hipStreamSynchronize(ctx->stream);
    if (pl->dtab) (void)hipFree(pl->dtab);
    if (pl->dlist) (void)hipFree(pl->dlist);
    if (pl->dlamy) (void)hipFree(pl->dlamy);
  }
  delete pl;
  return CRV_OK;
}

namespace {

// after a row range's FP64 pass: count its queued wavenumbers into stats[9]
// and empty the list for the next range
__global__ void k_list_done(unsigned long long* stats) {
  unsigned int* cnt = (unsigned int*)(stats + 8);
  stats[9] += *cnt;
  *cnt = 0u;
}

}  // namespace

// The posterior over y rows [j0, j1) of every kz plane (launches only).
// reset clears the statistics first (the first range of an exec).
// sim (FP32 plans with records only): the realizations' spectra, written by
// the posterior passes themselves (see fused_sim)
struct FusedSim {
  float2* sim = nullptr;
  long long sstride = 0;
  int nsim = 0;
  unsigned long long seed = 0, real0 = 0;
  // reference stream: the decoded normals, draws per grid, row prefix
  const double* normals = nullptr;
  long long draws = 0;
  const long long* prefix = nullptr;
  bool skip_prior_cov = false;
  bool skip_band_cov = false;
  bool sim_only = false;
};

int exec_rows(crv_post_plan* pl, crv_grid* const* mean, crv_grid* const* data, crv_grid* err_corr,
              crv_grid* const* cov, int j0, int j1, bool reset, const FusedSim& fs = FusedSim()) {
  crv_ctx* ctx = pl->ctx;
  const int n = pl->n;
  unsigned long long* dstats = (unsigned long long*)(pl->dtab + pl->oS);
  if (reset) {
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, dstats, 32);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  const int nyr = j1 - j0;
  int rc;
  if (pl->precision == CRV_FP64) {
    PostArgs<double> a{};
    for (int m = 0; m < 3; ++m) a.mean[m] = (double2*)mean[m]->d;
    for (int l = 0; l < n; ++l) a.data[l] = (double2*)data[l]->d;
    a.err = pl->separable ? nullptr : (const double2*)err_corr->d;
    for (int q = 0; q < 6; ++q) a.cov[q] = (double2*)cov[q]->d;
    a.t = tab_at<double>(pl->dtab, pl->oT);
    a.inv = (const int*)(pl->dtab + pl->oI);
    for (int q = 0; q < 6; ++q) a.sig0[q] = a.sig0d[q] = pl->sig0[q];
    a.nxp = pl->nxp; a.cnxp = pl->cnxp; a.nyp = pl->nyp; a.nzp = pl->nzp; a.nc = pl->nc;
    a.dplane = make_fastdiv(pl->cnxp * pl->nyp); a.dcx = make_fastdiv(pl->cnxp);
    a.j0 = j0; a.nyr = nyr; a.ncr = (long long)pl->cnxp * nyr * pl->nzp; a.dplr = make_fastdiv(pl->cnxp * nyr);
    a.stats = dstats;
    rc = pl->separable ? dispatch<double, true, false>(ctx, n, a) : dispatch<double, false, false>(ctx, n, a);
  } else {
    PostArgs<float> a{};
    for (int m = 0; m < 3; ++m) a.mean[m] = (float2*)mean[m]->d;
    for (int l = 0; l < n; ++l) a.data[l] = (float2*)data[l]->d;
    a.err = pl->separable ? nullptr : (const float2*)err_corr->d;
    for (int q = 0; q < 6; ++q) a.cov[q] = (float2*)cov[q]->d;
    a.t = tab_at<float>(pl->dtab, pl->oT);
    a.inv = (const int*)(pl->dtab + pl->oI);
    for (int q = 0; q < 6; ++q) { a.sig0d[q] = pl->sig0[q]; a.sig0[q] = (float)pl->sig0[q]; }
    a.nxp = pl->nxp; a.cnxp = pl->cnxp; a.nyp = pl->nyp; a.nzp = pl->nzp; a.nc = pl->nc;
    a.dplane = make_fastdiv(pl->cnxp * pl->nyp); a.dcx = make_fastdiv(pl->cnxp);
    a.j0 = j0; a.nyr = nyr; a.ncr = (long long)pl->cnxp * nyr * pl->nzp; a.dplr = make_fastdiv(pl->cnxp * nyr);
    a.stats = dstats;
    if (pl->oE) a.erec = (const float*)(pl->dtab + pl->oE);
    if (pl->oR) {
      a.rec = (const float2*)(pl->dtab + pl->oR);
      a.cpp = (pl->cnxp * nyr + 63) / 64;
      a.skip_prior_cov = fs.skip_prior_cov ? 1 : 0;
      a.skip_band_cov = (fs.skip_band_cov && pl->oE) ? 1 : 0;
      a.sim_only = (fs.sim_only && pl->oE) ? 1 : 0;
      if (fs.nsim > 0) {
        a.sim = fs.sim; a.sstride = fs.sstride; a.nsim = fs.nsim; a.seed = fs.seed; a.real0 = fs.real0;
        a.geo = make_noise_geom_dims(pl->cnxp, pl->nyp, pl->nzp);
        a.normals = fs.normals; a.draws = fs.draws; a.geo.prefix = fs.prefix;
      }
    }
    if (pl->refine) {
      a.t64 = tab_at<double>(pl->dtab, pl->o64);
      a.list = pl->dlist;
      a.list_count = (unsigned int*)(pl->dtab + pl->oS + 64);
      a.cond_thr = pl->cond_thr;
      a.lowk2 = (float)pl->lowk2;
      a.rx = 2.0f / (float)pl->nxp; a.ry = 2.0f / (float)pl->nyp; a.rz = 2.0f / (float)pl->nzp;
      rc = pl->separable ? dispatch<float, true, true>(ctx, n, a) : dispatch<float, false, true>(ctx, n, a);
      if (rc == CRV_OK) {
        hipLaunchKernelGGL(k_list_done, dim3(1), dim3(1), 0, ctx->stream, dstats);
        CRV_HIP_CHECK(ctx, hipGetLastError());
      }
    } else {
      rc = pl->separable ? dispatch<float, true, false>(ctx, n, a) : dispatch<float, false, false>(ctx, n, a);
    }
  }
  return rc;
}

extern "C" int crv_posterior_exec(crv_post_plan* pl, crv_grid* mean[3], crv_grid** data, crv_grid* err_corr,
                                  crv_grid* cov[6]) {
  if (!pl) return CRV_E_INVALID;
  crv_ctx* ctx = pl->ctx;
  if (!mean || !data || !cov) return crv_set_error(ctx, CRV_E_INVALID, "NULL argument");
  const int n = pl->n;
  auto fits = [&](const crv_grid* g, bool need_transformed) {
    return g && g->nxp == pl->nxp && g->nyp == pl->nyp && g->nzp == pl->nzp && g->precision == pl->precision &&
           (!need_transformed || g->transformed);
  };
  for (int m = 0; m < 3; ++m)
    if (!fits(mean[m], true)) return crv_set_error(ctx, CRV_E_STATE, "mean: wrong size/precision or not transformed");
  for (int l = 0; l < n; ++l)
    if (!fits(data[l], true)) return crv_set_error(ctx, CRV_E_STATE, "data: wrong size/precision or not transformed");
  for (int q = 0; q < 6; ++q)
    if (!fits(cov[q], !pl->separable)) return crv_set_error(ctx, CRV_E_STATE, "cov: wrong size/precision or state");
  if (!pl->separable && !fits(err_corr, true))
    return crv_set_error(ctx, CRV_E_STATE, "err_corr: wrong size/precision or not transformed");
  int rc = exec_rows(pl, mean, data, err_corr, cov, 0, pl->nyp, true);
  if (rc != CRV_OK) return rc;
  for (int l = 0; l < n; ++l) data[l]->transformed = true;
  for (int q = 0; q < 6; ++q) { cov[q]->transformed = true; cov[q]->cubetype = CRV_COVARIANCE; }
  return CRV_OK;
}

// ---- the whole inversion, pipelined through the memory-side cache ----------
namespace {

// grids of one role: one launch per pass when they form one contiguous batch
struct Family {
  crv_grid* const* g = nullptr;
  int n = 0;
  bool contiguous = false;
};

Family make_family(crv_grid* const* g, int n) {
  Family f;
  f.g = g; f.n = n; f.contiguous = true;
  for (int b = 1; b < n; ++b)
    if ((char*)g[b]->d != (char*)g[0]->d + (size_t)b * g[0]->rsize * g[0]->elem_bytes()) f.contiguous = false;
  return f;
}

int family_pass(crv_ctx* ctx, FftPlan* fp, const Family& f, int pass, int o0, int o1, double scale,
                const void* const* add, int nadd) {
  if (f.n == 0) return CRV_OK;
  if (f.contiguous)
    return crv_launch_pass(ctx, fp, pass, f.g[0]->d, f.g[0]->d, f.n, (long long)f.g[0]->rsize, o0, o1, scale, add,
                           nadd);
  for (int b = 0; b < f.n; ++b) {
    const void* a1[1] = {nadd ? add[b % nadd] : nullptr};
    const int rc = crv_launch_pass(ctx, fp, pass, f.g[b]->d, f.g[b]->d, 1, 0, o0, o1, scale, a1, nadd ? 1 : 0);
    if (rc != CRV_OK) return rc;
  }
  return CRV_OK;
}

// x and y passes (forward: x then y; inverse: y then x), and with with_z the
// z pass too (forward last, inverse first), over a family in sub-batches
// that fit the memory-side cache, so each pass of a sub-batch reads the
// previous pass's output from there; cs: the family is the covariances, whose
// spectra the z inverse generates where CovGen says so
int family_xy(crv_ctx* ctx, FftPlan* fp, const Family& f, bool inverse, double scale, const void* const* add,
              int nadd, crv_grid* const* src = nullptr, bool with_z = false, const CovGen* cs = nullptr) {
  const int nrows = fp->nyp * fp->nzp;
  // a forward x pass may read other grids (src): one launch per sub-batch
  // when they form a batch with the same stride, else one per grid
  const bool src_batch = src && make_family(src, f.n).contiguous && src[0]->rsize == f.g[0]->rsize;
  const int sb = (f.contiguous && (!src || src_batch)) ? crv_sub_batch(fp, f.n) : 1;
  for (int b0 = 0; b0 < f.n; b0 += sb) {
    Family s;
    s.g = f.g + b0; s.n = std::min(sb, f.n - b0); s.contiguous = f.contiguous;
    const void* ad[3] = {nullptr, nullptr, nullptr};
    for (int q = 0; q < nadd; ++q) ad[q] = add[(b0 + q) % nadd];
    int rc;
    if (!inverse) {
      if (src) {
        if ((rc = crv_launch_pass(ctx, fp, 0, s.g[0]->d, src[b0]->d, s.n, (long long)s.g[0]->rsize, 0, nrows, scale,
                                  nullptr, 0)) != CRV_OK)
          return rc;
      } else if ((rc = family_pass(ctx, fp, s, 0, 0, nrows, scale, nullptr, 0)) != CRV_OK) {
        return rc;
      }
      if ((rc = family_pass(ctx, fp, s, 1, 0, fp->nzp, 1.0, nullptr, 0)) != CRV_OK) return rc;
      if (with_z && (rc = family_pass(ctx, fp, s, 2, 0, fp->nyp, 1.0, nullptr, 0)) != CRV_OK) return rc;
    } else {
      if (with_z && cs) {
        // covariances b0.. of the family: q = b0 + b
        CovGen g = *cs;
        if (s.contiguous) {
          g.q0 = b0;
          rc = crv_launch_pass(ctx, fp, 3, s.g[0]->d, s.g[0]->d, s.n, (long long)s.g[0]->rsize, 0, fp->nyp, 1.0,
                               nullptr, 0, &g);
        } else {
          rc = CRV_OK;
          for (int b = 0; b < s.n && rc == CRV_OK; ++b) {
            g.q0 = b0 + b;
            rc = crv_launch_pass(ctx, fp, 3, s.g[b]->d, s.g[b]->d, 1, 0, 0, fp->nyp, 1.0, nullptr, 0, &g);
          }
        }
        if (rc != CRV_OK) return rc;
      } else if (with_z && (rc = family_pass(ctx, fp, s, 3, 0, fp->nyp, 1.0, nullptr, 0)) != CRV_OK) {
        return rc;
      }
      if ((rc = family_pass(ctx, fp, s, 4, 0, fp->nzp, 1.0, nullptr, 0)) != CRV_OK) return rc;
      if ((rc = family_pass(ctx, fp, s, 5, 0, nrows, scale, ad, nadd)) != CRV_OK) return rc;
    }
  }
  return CRV_OK;
}

}  // namespace

extern "C" int crv_invert_noise(crv_post_plan* pl, crv_grid* mean[3], crv_grid** data, crv_grid* const* mean_src,
                                crv_grid* const* data_src, crv_grid* cov[6], const crv_noise_spec* noise, int nreal,
                                crv_grid** sims, int chunk_rows) {
  if (!pl) return CRV_E_INVALID;
  crv_ctx* ctx = pl->ctx;
  if (!mean || !data || !cov || nreal < 0 || nreal > 21845 || (nreal > 0 && !sims) || !noise)
    return crv_set_error(ctx, CRV_E_INVALID, "crv_invert: bad arguments");
  if (noise->mode != CRV_NOISE_COUNTER && noise->mode != CRV_NOISE_REFERENCE)
    return crv_set_error(ctx, CRV_E_INVALID, "crv_invert: noise must be the counter or the reference stream");
  if (noise->mode == CRV_NOISE_REFERENCE && !noise->lcg_state)
    return crv_set_error(ctx, CRV_E_INVALID, "reference noise needs lcg_state");
  const bool refnoise = noise->mode == CRV_NOISE_REFERENCE;
  const unsigned long long seed = noise->seed, realization0 = noise->realization;
  if (!pl->separable) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "crv_invert needs a separable-prior plan");
  const int n = pl->n;
  auto fits = [&](const crv_grid* g) {
    return g && g->nxp == pl->nxp && g->nyp == pl->nyp && g->nzp == pl->nzp && g->precision == pl->precision;
  };
  for (int m = 0; m < 3; ++m) {
    const crv_grid* in = mean_src ? mean_src[m] : mean[m];
    if (!fits(mean[m]) || !fits(in) || in->transformed)
      return crv_set_error(ctx, CRV_E_STATE, "mean: wrong size or not real");
  }
  for (int l = 0; l < n; ++l) {
    const crv_grid* in = data_src ? data_src[l] : data[l];
    if (!fits(data[l]) || !fits(in) || in->transformed)
      return crv_set_error(ctx, CRV_E_STATE, "data: wrong size or not real");
  }
  if (mean_src)
    for (int m = 0; m < 3; ++m) mean[m]->cubetype = mean_src[m]->cubetype;
  if (data_src)
    for (int l = 0; l < n; ++l) data[l]->cubetype = data_src[l]->cubetype;
  for (int q = 0; q < 6; ++q)
    if (!fits(cov[q])) return crv_set_error(ctx, CRV_E_INVALID, "cov: wrong size");
  const int nsim = 3 * nreal;
  for (int b = 0; b < nsim; ++b) {
    if (!fits(sims[b])) return crv_set_error(ctx, CRV_E_INVALID, "sims: wrong size");
    if ((char*)sims[b]->d != (char*)sims[0]->d + (size_t)b * sims[0]->rsize * sims[0]->elem_bytes())
      return crv_set_error(ctx, CRV_E_INVALID, "sims must be one contiguous grid batch");
  }
  FftPlan* fp = nullptr;
  int rc = crv_get_plan(ctx, pl->nxp, pl->nyp, pl->nzp, pl->precision, &fp);
  if (rc != CRV_OK) return rc;
  // reference stream: the 3 nreal grids' normals decoded up front, in the
  // order Crava::simulate draws them (crava.cpp:1598-1603)
  const double* normals = nullptr;
  if (refnoise && nsim > 0) {
    if ((rc = crv_noise_prefix(ctx, fp)) != CRV_OK) return rc;
    if ((rc = crv_reference_normals(ctx, noise->lcg_state, (long long)nsim * fp->noise_draws)) != CRV_OK) return rc;
    normals = ctx->normals;
  }
  const int nyp = pl->nyp;
  const Family fd = make_family(data, n), fm = make_family(mean, 3), fc = make_family(cov, 6);
  const Family fs = nsim ? make_family(sims, nsim) : Family();
  // chunk_rows <= 0 or >= nyp: whole transforms, each grid's passes back to
  // back (the order of the one-call sequence); otherwise chunks of y rows
  const bool chunked = chunk_rows > 0 && chunk_rows < nyp;
  if (!chunked) chunk_rows = nyp;
  // whole grids on an FP32 plan with per-plane records: the posterior
  // passes write the realizations themselves (no re-read of the covariances)
  const bool fused = !chunked && nsim > 0 && pl->oR != 0;
  // whole grids on an FP32 plan with records: the covariances of planes
  // outside the band are the cleaned prior lam(i,j) S(k); the posterior
  // leaves them unwritten and the z inverse pass generates them
  const bool synth = !chunked && pl->oR != 0;
  // with eigen-basis records the band's planes as well (but for the
  // wavenumbers of the FP64 list), when their tables fit the z inverse's LDS
  const bool synth_band = synth && pl->oE != 0 && pl->oGR != 0 && crv_cov_gen_fits(fp, pl->nband, pl->gw);
  // whole grids on a plan where no wavenumber can need the FP64 solve: the
  // posterior of the means and the residuals runs inside the z pass
  // (crv_launch_zpost), so their spectra never make the HBM round trip
  // between the z forward pass, the posterior and the z inverse pass
  const int ztx = (!chunked && synth && pl->oE != 0 && pl->zpost_ok && n <= 6) ? crv_zpost_tx(ctx, fp, pl->nband) : 0;
  const bool zfused = ztx > 0;
  // forward x and y passes of the data and the means (and their z passes
  // unless the fused z pass does them)
  const bool fwd_z = !chunked && !zfused;
  if ((rc = family_xy(ctx, fp, fd, false, crv_scale_forward(data[0]), nullptr, 0, data_src, fwd_z)) != CRV_OK)
    return rc;
  if ((rc = family_xy(ctx, fp, fm, false, crv_scale_forward(mean[0]), nullptr, 0, mean_src, fwd_z)) != CRV_OK)
    return rc;
  for (int q = 0; q < 6; ++q) cov[q]->cubetype = CRV_COVARIANCE;
  const void* cv[6];
  for (int q = 0; q < 6; ++q) cv[q] = cov[q]->d;
  if (zfused) {
    unsigned long long* dstats = (unsigned long long*)(pl->dtab + pl->oS);
    hipLaunchKernelGGL(k_zero_u64, dim3(1), dim3(64), 0, ctx->stream, dstats, 32);
    CRV_HIP_CHECK(ctx, hipGetLastError());
    ZPost z;
    for (int l = 0; l < n; ++l) z.data[l] = data[l]->d;
    for (int m = 0; m < 3; ++m) z.mean[m] = mean[m]->d;
    z.n = n; z.nband = pl->nband;
    z.lam = tab_at<float>(pl->dtab, pl->oT).lam;
    z.erec = (const float*)(pl->dtab + pl->oE);
    z.bidx = (const int*)(pl->dtab + pl->oGI);
    z.bplanes = (const int*)(pl->dtab + pl->oBP);
    z.stats = dstats;
    if ((rc = crv_launch_zpost(ctx, fp, z, ztx)) != CRV_OK) return rc;
    // the realizations' spectra (and the band's covariances where the z
    // inverse does not generate them) from the posterior kernel alone
    if (nsim > 0 || !synth_band) {
      FusedSim fsim;
      if (nsim > 0) {
        fsim.sim = (float2*)sims[0]->d; fsim.sstride = (long long)sims[0]->rsize / 2; fsim.nsim = nreal;
        fsim.seed = seed; fsim.real0 = realization0;
        fsim.normals = normals; fsim.draws = fp->noise_draws; fsim.prefix = fp->noise_prefix;
      }
      fsim.skip_prior_cov = synth;
      fsim.skip_band_cov = synth_band;
      fsim.sim_only = true;
      if ((rc = exec_rows(pl, mean, data, nullptr, cov, 0, nyp, false, fsim)) != CRV_OK) return rc;
    }
  }
  // per chunk of y rows: z forward, posterior, simulation, z inverse
  for (int j0 = 0; j0 < nyp && !zfused; j0 += chunk_rows) {
    const int j1 = std::min(nyp, j0 + chunk_rows);
    if (chunked) {
      if ((rc = family_pass(ctx, fp, fd, 2, j0, j1, 1.0, nullptr, 0)) != CRV_OK) return rc;
      if ((rc = family_pass(ctx, fp, fm, 2, j0, j1, 1.0, nullptr, 0)) != CRV_OK) return rc;
    }
    FusedSim fsim;
    if (fused) {
      fsim.sim = (float2*)sims[0]->d; fsim.sstride = (long long)sims[0]->rsize / 2; fsim.nsim = nreal;
      fsim.seed = seed; fsim.real0 = realization0;
      fsim.normals = normals; fsim.draws = fp->noise_draws; fsim.prefix = fp->noise_prefix;
    }
    fsim.skip_prior_cov = synth;
    fsim.skip_band_cov = synth_band;
    if ((rc = exec_rows(pl, mean, data, nullptr, cov, j0, j1, j0 == 0, fsim)) != CRV_OK) return rc;
    if (chunked) {
      if (nsim && (rc = crv_launch_sim_rows(ctx, fp, cv, sims[0]->d, (long long)sims[0]->rsize / 2, nreal, seed,
                                            realization0, pl->refine ? 2 : 0, j0, j1, normals)) != CRV_OK)
        return rc;
      if ((rc = family_pass(ctx, fp, fm, 3, j0, j1, 1.0, nullptr, 0)) != CRV_OK) return rc;
      if ((rc = family_pass(ctx, fp, fc, 3, j0, j1, 1.0, nullptr, 0)) != CRV_OK) return rc;
      if ((rc = family_pass(ctx, fp, fs, 3, j0, j1, 1.0, nullptr, 0)) != CRV_OK) return rc;
    }
  }
  // inverse (y and x, after the z pass when not chunked): the means first,
  // whose real values the realizations add in their last pass; the
  // realizations' noise when not chunked after the covariances' spectra are
  // complete and before their inverse transforms
  if ((rc = family_xy(ctx, fp, fm, true, crv_scale_inverse(mean[0]), nullptr, 0, nullptr, fwd_z)) != CRV_OK)
    return rc;
  if (!chunked && nsim && !fused &&
      (rc = crv_launch_sim_rows(ctx, fp, cv, sims[0]->d, (long long)sims[0]->rsize / 2, nreal, seed, realization0,
                                pl->refine ? 2 : 0, 0, nyp, normals)) != CRV_OK)
    return rc;
  if (nsim) {
    for (int b = 0; b < nsim; ++b) sims[b]->cubetype = CRV_PARAMETER;
    const void* add[3] = {mean[0]->d, mean[1]->d, mean[2]->d};
    if ((rc = family_xy(ctx, fp, fs, true, crv_scale_inverse(sims[0]), add, 3, nullptr, !chunked)) != CRV_OK)
      return rc;
  }
  CovGen cs;
  if (synth) {
    const KzTab<float> t = tab_at<float>(pl->dtab, pl->oT);
    cs.lam = t.lam; cs.S = t.S; cs.band = (const int*)(pl->dtab + pl->oI);
    if (synth_band) {
      cs.bidx = (const int*)(pl->dtab + pl->oGI);
      cs.rows = (const float*)(pl->dtab + pl->oGR);
      cs.nband = pl->nband; cs.rw = pl->gw; cs.n = pl->n;
      cs.rx = 2.0f / (float)pl->nxp; cs.ry = 2.0f / (float)pl->nyp; cs.rz = 2.0f / (float)pl->nzp;
      cs.lowk2 = (float)pl->lowk2; cs.cond_thr = pl->cond_thr;
      cs.noqueue = pl->zpost_ok ? 1 : 0;
    }
  }
  if (synth_band && pl->dlamy && crv_cov_band_y_fits(fp)) {
    // per covariance: the y inverse of the band's planes, generated there;
    // the z inverse, generating the planes outside the band from the y
    // transform of lam and reading the band's; the x C2R pass
    cs.bplanes = (const int*)(pl->dtab + pl->oBP);
    // in sub-batches of contiguous grids sized to the memory-side cache (one
    // grid at a time otherwise), each pass reading its predecessor's output
    const int sb = fc.contiguous ? crv_sub_batch(fp, 6) : 1;
    const long long rs = (long long)cov[0]->rsize;
    for (int q0 = 0; q0 < 6; q0 += sb) {
      const int nb = std::min(sb, 6 - q0);
      CovGen g = cs;
      g.q0 = q0;
      if ((rc = crv_launch_cov_band_y(ctx, fp, cov[q0]->d, nb, rs, g)) != CRV_OK) return rc;
      CovGen gz = cs;
      gz.q0 = q0;
      gz.rows = nullptr;
      gz.lamy = pl->dlamy;
      if ((rc = crv_launch_pass(ctx, fp, 3, cov[q0]->d, cov[q0]->d, nb, rs, 0, nyp, 1.0, nullptr, 0, &gz)) != CRV_OK)
        return rc;
      if ((rc = crv_launch_pass(ctx, fp, 5, cov[q0]->d, cov[q0]->d, nb, rs, 0, nyp * pl->nzp,
                                crv_scale_inverse(cov[q0]), nullptr, 0)) != CRV_OK)
        return rc;
    }
  } else if ((rc = family_xy(ctx, fp, fc, true, crv_scale_inverse(cov[0]), nullptr, 0, nullptr, !chunked,
                             synth ? &cs : nullptr)) != CRV_OK) {
    return rc;
  }
  for (int l = 0; l < n; ++l) data[l]->transformed = true;
  for (int m = 0; m < 3; ++m) mean[m]->transformed = false;
  for (int q = 0; q < 6; ++q) cov[q]->transformed = false;
  for (int b = 0; b < nsim; ++b) sims[b]->transformed = false;
  return CRV_OK;
}

extern "C" int crv_invert(crv_post_plan* pl, crv_grid* mean[3], crv_grid** data, crv_grid* const* mean_src,
                          crv_grid* const* data_src, crv_grid* cov[6], unsigned long long seed,
                          unsigned long long realization0, int nreal, crv_grid** sims, int chunk_rows) {
  crv_noise_spec ns;
  memset(&ns, 0, sizeof(ns));
  ns.mode = CRV_NOISE_COUNTER;
  ns.seed = seed;
  ns.realization = realization0;
  return crv_invert_noise(pl, mean, data, mean_src, data_src, cov, &ns, nreal, sims, chunk_rows);
}

extern "C" int crv_post_plan_stats(crv_post_plan* pl, crv_post_stats* stats) {
  if (!pl || !stats) return CRV_E_INVALID;
  crv_ctx* ctx = pl->ctx;
  unsigned long long hs[16];
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(hs, pl->dtab + pl->oS, sizeof(hs), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  stats->n_wavenumbers = pl->nc;
  stats->n_inverted = pl->ninv_planes * pl->cnxp * pl->nyp;
  stats->n_chol_fail = (long long)hs[0];
  stats->n_fp64 = pl->refine ? (long long)hs[9] : 0;
  double mc;
  memcpy(&mc, &hs[1], 8);
  stats->max_cond = mc;
  return CRV_OK;
}

extern "C" int crv_posterior(crv_ctx* ctx, const crv_post_params* p, crv_grid* mean[3], crv_grid** data,
                             crv_grid* err_corr, crv_grid* cov[6], crv_post_stats* stats) {
  if (!ctx) return CRV_E_INVALID;
  if (!mean || !data || !cov) return crv_set_error(ctx, CRV_E_INVALID, "NULL argument");
  int rc = validate_params(ctx, p);
  if (rc != CRV_OK) return rc;
  const int n = p->ntheta;
  crv_grid* g0 = mean[0];
  if (!g0) return crv_set_error(ctx, CRV_E_INVALID, "mean[0] is NULL");
  for (int m = 0; m < 3; ++m) if ((rc = check_grid(ctx, mean[m], g0, "mean")) != CRV_OK) return rc;
  for (int l = 0; l < n; ++l) if ((rc = check_grid(ctx, data[l], g0, "data")) != CRV_OK) return rc;
  for (int q = 0; q < 6; ++q) if ((rc = check_grid(ctx, cov[q], g0, "cov")) != CRV_OK) return rc;
  if (!p->separable && (rc = check_grid(ctx, err_corr, g0, "err_corr")) != CRV_OK) return rc;
  if (p->mode == CRV_MODE_FP32_REFINE && g0->precision != CRV_FP32)
    return crv_set_error(ctx, CRV_E_INVALID, "CRV_MODE_FP32_REFINE needs FP32 grids");
  // Forward transforms that the reference performs inside the call or just
  // before it (crava.cpp:177-184, 1039, 1062).
  for (int m = 0; m < 3; ++m) if ((rc = ensure_transformed(mean[m])) != CRV_OK) return rc;
  for (int l = 0; l < n; ++l) if ((rc = ensure_transformed(data[l])) != CRV_OK) return rc;
  if (!p->separable) {
    for (int q = 0; q < 6; ++q) if ((rc = ensure_transformed(cov[q])) != CRV_OK) return rc;
    if ((rc = ensure_transformed(err_corr)) != CRV_OK) return rc;
  }
  crv_post_plan* pl = nullptr;
  if ((rc = plan_make(ctx, p, g0->nxp, g0->nyp, g0->nzp, g0->precision, true, &pl)) != CRV_OK) return rc;
  rc = crv_posterior_exec(pl, mean, data, err_corr, cov);
  if (rc == CRV_OK && stats) rc = crv_post_plan_stats(pl, stats);
  crv_post_plan_destroy(pl);
  if (rc != CRV_OK) return rc;
  if (!p->skip_mean_inverse) {
    for (int m = 0; m < 3; ++m) if ((rc = crv_fft_c2r(mean[m])) != CRV_OK) return rc;
  }
  return CRV_OK;
}
