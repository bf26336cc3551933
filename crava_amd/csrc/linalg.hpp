// Small dense complex linear algebra shared by the posterior and simulation
// kernels (one thread per matrix, everything in registers).
#pragma once
#include "internal.hpp"

constexpr double kCholTol = 1e-20;  // TOL of lib_matr.c:20

// Scaled complex Cholesky with the reference's acceptance rule
// (lib_matr.c:78-167): divide by Re A00 (fail if <= 0), reject a row whose
// scaled diagonal or Schur pivot is <= 1e-20, multiply L by sqrt(A00).
// Only the lower triangle (row-major packed A[i*(i+1)/2 + j]) is used.
template <typename U, int N>
__device__ __forceinline__ bool chol_ref(cplx_t<U>* A, U& dmin, U& dmax) {
  using C = cplx_t<U>;
  const U factor = A[0].x;
  if (factor <= (U)0) return false;
#pragma unroll
  for (int t = 0; t < N * (N + 1) / 2; ++t) A[t] = cmake<C>(A[t].x / factor, A[t].y / factor);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int ri = i * (i + 1) / 2;
    if (A[ri + i].x <= (U)kCholTol) return false;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const int rj = j * (j + 1) / 2;
      C lr = cmake<C>((U)0, (U)0);
#pragma unroll
      for (int q = 0; q < j; ++q) lr = cadd(lr, cmulc(A[ri + q], A[rj + q]));
      const U d = A[rj + j].x;
      const U help = d * d;
      const C x = csub(A[ri + j], lr);
      A[ri + j] = cmake<C>((x.x * d) / help, (x.y * d) / help);
    }
    U lr = (U)0;
#pragma unroll
    for (int q = 0; q < i; ++q) lr += A[ri + q].x * A[ri + q].x + A[ri + q].y * A[ri + q].y;
    const U dd = A[ri + i].x - lr;
    if (dd <= (U)kCholTol) return false;
    A[ri + i] = cmake<C>(sqrt(dd), (U)0);
  }
  const U sf = sqrt(factor);
  dmin = A[0].x; dmax = A[0].x;
#pragma unroll
  for (int t = 0; t < N * (N + 1) / 2; ++t) A[t] = cmake<C>(A[t].x * sf, A[t].y * sf);
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const U d = A[i * (i + 1) / 2 + i].x;
    dmin = fmin(dmin, d); dmax = fmax(dmax, d);
  }
  return true;
}

__device__ __forceinline__ float rsqrt_any(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_any(double x) { return 1.0 / sqrt(x); }

// The same factorisation and acceptance rule without the scaling pass and
// without divisions.  Scaling A by 1/A00 only moves the thresholds:
// a scaled pivot p / A00 <= 1e-20 is the unscaled pivot p <= 1e-20 * A00, so
// the rule is applied to unscaled pivots and the factor keeps its reciprocal
// diagonal rd[i] = 1 / L[i][i].  Off-diagonal L entries are stored in A;
// the diagonal of A is left as it was.  pmin / pmax are the smallest and
// largest pivots L[i][i]^2, so (dmax/dmin)^2 of the diagonal is pmax / pmin.
template <typename U, int N>
__device__ __forceinline__ bool chol_rcp(cplx_t<U>* A, U* rd, U& pmin, U& pmax) {
  using C = cplx_t<U>;
  const U factor = A[0].x;
  if (factor <= (U)0) return false;
  const U tol = (U)kCholTol * factor;
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int ri = i * (i + 1) / 2;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const int rj = j * (j + 1) / 2;
      C x = A[ri + j];
#pragma unroll
      for (int q = 0; q < j; ++q) x = csub(x, cmulc(A[ri + q], A[rj + q]));
      A[ri + j] = cscale(x, rd[j]);
    }
    U p = A[ri + i].x;
#pragma unroll
    for (int q = 0; q < i; ++q) p -= A[ri + q].x * A[ri + q].x + A[ri + q].y * A[ri + q].y;
    if (p <= tol) return false;
    rd[i] = rsqrt_any(p);
    if (i == 0) { pmin = p; pmax = p; } else { pmin = fmin(pmin, p); pmax = fmax(pmax, p); }
  }
  return true;
}

// L y = b with L from chol_rcp (strict lower part in L, 1/diag in rd)
template <typename U, int N, int NC>
__device__ __forceinline__ void lsolve_r(const cplx_t<U>* L, const U* rd, cplx_t<U> (*b)[NC]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int ri = i * (i + 1) / 2;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx_t<U> x = b[i][c];
#pragma unroll
      for (int j = 0; j < i; ++j) x = csub(x, cmul(b[j][c], L[ri + j]));
      b[i][c] = cscale(x, rd[i]);
    }
  }
}

// L^H x = y with L from chol_rcp
template <typename U, int N, int NC>
__device__ __forceinline__ void lhsolve_r(const cplx_t<U>* L, const U* rd, cplx_t<U> (*b)[NC]) {
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx_t<U> x = b[i][c];
#pragma unroll
      for (int j = N - 1; j > i; --j) x = csub(x, cmulc(b[j][c], L[j * (j + 1) / 2 + i]));
      b[i][c] = cscale(x, rd[i]);
    }
  }
}

// forward substitution L y = b (b overwritten), N x NC right-hand sides
template <typename U, int N, int NC>
__device__ __forceinline__ void lsolve(const cplx_t<U>* L, cplx_t<U> (*b)[NC]) {
#pragma unroll
  for (int i = 0; i < N; ++i) {
    const int ri = i * (i + 1) / 2;
    const U d = L[ri + i].x;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx_t<U> x = b[i][c];
#pragma unroll
      for (int j = 0; j < i; ++j) x = csub(x, cmul(b[j][c], L[ri + j]));
      b[i][c] = cmake<cplx_t<U>>(x.x / d, x.y / d);
    }
  }
}

// backward substitution L^H x = y (y overwritten)
template <typename U, int N, int NC>
__device__ __forceinline__ void lhsolve(const cplx_t<U>* L, cplx_t<U> (*b)[NC]) {
#pragma unroll
  for (int i = N - 1; i >= 0; --i) {
    const U d = L[i * (i + 1) / 2 + i].x;
#pragma unroll
    for (int c = 0; c < NC; ++c) {
      cplx_t<U> x = b[i][c];
#pragma unroll
      for (int j = N - 1; j > i; --j) x = csub(x, cmulc(b[j][c], L[j * (j + 1) / 2 + i]));
      b[i][c] = cmake<cplx_t<U>>(x.x / d, x.y / d);
    }
  }
}

