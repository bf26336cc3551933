// Per-wavenumber pieces of posterior simulation (Crava::simulate, reference
// src/crava.cpp:1599-1708), shared by the simulation kernel and the
// posterior kernel that fuses it: the 3 x 3 factorisation of the posterior
// covariance with lib_matrCholCpx's acceptance rule (libs/lib/lib_matr.c:
// 78-167; optionally in FP64 with one Newton refinement step for FP32
// grids), and L z (lib_matrProdCholVec, lib_matr.c:453-470).
#pragma once
#include "internal.hpp"
#include "linalg.hpp"

#include <type_traits>

__device__ __forceinline__ double rsqrt_f64(double x) {
  // hardware estimate and two Newton steps
  double y = __builtin_amdgcn_rsq(x);
  y = y * (1.5 - 0.5 * x * y * y);
  return y * (1.5 - 0.5 * x * y * y);
}

// 3 x 3 Cholesky with lib_matrCholCpx's acceptance rule (factor A00 > 0,
// scaled pivots > 1e-20) without divisions; P and L packed lower
// (00, 10, 11, 20, 21, 22).
__device__ __forceinline__ bool chol3_f64(const double2* P, double2* L) {
  using D = double2;
  const double factor = P[0].x;
  if (factor <= 0.0) return false;
  const double tol = kCholTol * factor;
  const double p0 = P[0].x;
  if (p0 <= tol) return false;
  const double r0 = rsqrt_f64(p0);
  L[0] = cmake<D>(p0 * r0, 0.0);
  L[1] = cscale(P[1], r0);
  const double p1 = P[2].x - (L[1].x * L[1].x + L[1].y * L[1].y);
  if (p1 <= tol) return false;
  const double r1 = rsqrt_f64(p1);
  L[2] = cmake<D>(p1 * r1, 0.0);
  L[3] = cscale(P[3], r0);
  L[4] = cscale(csub(P[4], cmulc(L[3], L[1])), r1);
  const double p2 = P[5].x - (L[3].x * L[3].x + L[3].y * L[3].y) - (L[4].x * L[4].x + L[4].y * L[4].y);
  if (p2 <= tol) return false;
  L[5] = cmake<D>(p2 * rsqrt_f64(p2), 0.0);
  return true;
}

// One step of iterative refinement of a Cholesky factor (a Newton step for
// L L^H = P): E = P - L L^H, X = L^-1 E L^-H, L <- L + L Phi(X), where
// Phi keeps the strict lower triangle of X and half its diagonal.
__device__ __forceinline__ void chol3_refine(const double2* P, double2* L) {
  using D = double2;
  auto at = [&](int i, int j) -> D { return L[i * (i + 1) / 2 + j]; };
  // E, full Hermitian 3 x 3 from the lower part
  D E[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      D x = P[i * (i + 1) / 2 + j];
#pragma unroll
      for (int k = 0; k <= j; ++k) x = csub(x, cmulc(at(i, k), at(j, k)));
      if (i == j) x.y = 0.0;
      E[i][j] = x;
      E[j][i] = cconj(x);
    }
  const double r[3] = {1.0 / L[0].x, 1.0 / L[2].x, 1.0 / L[5].x};
  // Y = L^-1 E (columns), then X = L^-1 Y^H
  D Y[3][3], X[3][3];
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      D x = E[i][c];
#pragma unroll
      for (int k = 0; k < i; ++k) x = csub(x, cmul(at(i, k), Y[k][c]));
      Y[i][c] = cscale(x, r[i]);
    }
#pragma unroll
  for (int c = 0; c < 3; ++c)
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      D x = cconj(Y[c][i]);
#pragma unroll
      for (int k = 0; k < i; ++k) x = csub(x, cmul(at(i, k), X[k][c]));
      X[i][c] = cscale(x, r[i]);
    }
  D dL[6];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j <= i; ++j) {
      D x = cmake<D>(0.0, 0.0);
#pragma unroll
      for (int k = j; k <= i; ++k) {
        const D phi = (k == j) ? cmake<D>(0.5 * X[j][j].x, 0.0) : X[k][j];
        x = cadd(x, cmul(at(i, k), phi));
      }
      dL[i * (i + 1) / 2 + j] = x;
    }
#pragma unroll
  for (int q = 0; q < 6; ++q) L[q] = cadd(L[q], dL[q]);
  L[0].y = 0.0; L[2].y = 0.0; L[5].y = 0.0;
}

// The factor of one wavenumber's covariance c = (aa, bb, rr, ab, ar, br) as
// stored in the grids: in FP64 with a refinement step (F64), or in T with
// the reference's rules.  false: rejected (the realization is zero there).
template <typename T, bool F64>
__device__ __forceinline__ bool sim_factor(const cplx_t<T> c[6],
                                           cplx_t<typename std::conditional<F64, double, T>::type> Lf[6]) {
  using U = typename std::conditional<F64, double, T>::type;
  using CU = cplx_t<U>;
  // lower triangle packed: P00, P10, P11, P20, P21, P22 (lower = conj(upper))
  CU P[6];
  P[0] = cmake<CU>((U)c[0].x, (U)c[0].y);
  P[1] = cmake<CU>((U)c[3].x, -(U)c[3].y);
  P[2] = cmake<CU>((U)c[1].x, (U)c[1].y);
  P[3] = cmake<CU>((U)c[4].x, -(U)c[4].y);
  P[4] = cmake<CU>((U)c[5].x, -(U)c[5].y);
  P[5] = cmake<CU>((U)c[2].x, (U)c[2].y);
  bool ok;
  if constexpr (std::is_same<U, double>::value) {
    ok = chol3_f64(P, Lf);
    if (F64 && ok) chol3_refine(P, Lf);
  } else {
    U dmin, dmax;
#pragma unroll
    for (int q = 0; q < 6; ++q) Lf[q] = P[q];
    ok = chol_ref<U, 3>(Lf, dmin, dmax);
  }
  return ok;
}

// The factor of an FP32 covariance c = (aa, bb, rr, ab, ar, br) in float
// without divisions (reciprocal square roots), with lib_matrCholCpx's
// acceptance rule applied to the unscaled pivots (p <= 1e-20 A00, the form of
// its test on the matrix divided by A00); where a pivot is rejected or the
// pivot ratio max p / min p (the squared ratio of the factor's diagonal)
// exceeds kSimCondF32 the factor is made again in FP64 with one refinement
// step and rounded to float.  false: rejected in FP64 as well (the
// realization is zero there).
constexpr float kSimCondF32 = 1e3f;
__device__ __forceinline__ bool sim_factor_mixed(const float2 c[6], float2 L[6]) {
  const float a00 = c[0].x;
  const float tol = (float)kCholTol * a00;
  const float r0 = __builtin_amdgcn_rsqf(a00);
  L[0] = make_float2(a00 * r0, 0.0f);
  L[1] = make_float2(c[3].x * r0, -c[3].y * r0);
  const float p1 = c[1].x - (L[1].x * L[1].x + L[1].y * L[1].y);
  const float r1 = __builtin_amdgcn_rsqf(p1);
  L[2] = make_float2(p1 * r1, 0.0f);
  L[3] = make_float2(c[4].x * r0, -c[4].y * r0);
  // P21 - L20 conj(L10), with P21 = conj(br)
  const float2 x = make_float2(c[5].x - (L[3].x * L[1].x + L[3].y * L[1].y),
                               -c[5].y - (L[3].y * L[1].x - L[3].x * L[1].y));
  L[4] = make_float2(x.x * r1, x.y * r1);
  const float p2 = c[2].x - (L[3].x * L[3].x + L[3].y * L[3].y) - (L[4].x * L[4].x + L[4].y * L[4].y);
  L[5] = make_float2(p2 * __builtin_amdgcn_rsqf(p2), 0.0f);
  const float pmin = fminf(a00, fminf(p1, p2)), pmax = fmaxf(a00, fmaxf(p1, p2));
  const bool ok = a00 > 0.0f && a00 > tol && p1 > tol && p2 > tol;
  if (ok && pmax <= kSimCondF32 * pmin) return true;
  double2 Ld[6];
  if (!sim_factor<float, true>(c, Ld)) return false;
#pragma unroll
  for (int q = 0; q < 6; ++q) L[q] = make_float2((float)Ld[q].x, (float)Ld[q].y);
  return true;
}

// o = L z for the noise z = re + i im, rounded to T
template <typename T, typename U>
__device__ __forceinline__ void sim_apply(const cplx_t<U> Lf[6], const T re[3], const T im[3], cplx_t<T> o[3]) {
  using CU = cplx_t<U>;
  CU z[3];
#pragma unroll
  for (int g = 0; g < 3; ++g) z[g] = cmake<CU>((U)re[g], (U)im[g]);
#pragma unroll
  for (int r = 0; r < 3; ++r) {
    CU x = cmake<CU>((U)0, (U)0);
#pragma unroll
    for (int c = 0; c <= r; ++c) x = cadd(x, cmul(Lf[r * (r + 1) / 2 + c], z[c]));
    o[r] = cmake<cplx_t<T>>((T)x.x, (T)x.y);
  }
}
