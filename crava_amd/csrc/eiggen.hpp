// Pieces of the eigen-basis posterior (FP32 refine mode, separable prior)
// that two kernels evaluate on the same values and must agree on bit for
// bit: the posterior kernel itself, and the z inverse pass that generates the
// covariance planes from the per-plane tables instead of reading them.
//   * the queue rule: a wavenumber goes to the FP64 solve when |k| is below
//     the low-wavenumber limit or its condition bound
//     cond(D) (lam l_max + 1) / (lam l_min + 1) exceeds the threshold;
//   * entry q of S_post = lam S - lam^2 sum_p w_p G_p, w_p = 1 / (lam l_p + 1).
// Every multiply-add is written as fmaf so that both kernels round alike.
#pragma once
#include "internal.hpp"

struct EigQueueGeo {
  float rx, ry, rz;   // 2 / nxp, 2 / nyp, 2 / nzp
  float lowk2;        // (low-wavenumber fraction of Nyquist)^2
  float cond_thr;
  int nyp, nzp;
};

// condition bound of wavenumber (i, j, k) from the plane's cond(D), l_max,
// l_min: cond(D) (lam l_max + 1) / (lam l_min + 1)
__device__ __forceinline__ float eig_cond(float lam, float cd, float lmax, float lmin) {
  return cd * fmaf(lam, lmax, 1.0f) / fmaf(lam, lmin, 1.0f);
}

// true when wavenumber (i, j, k) is queued for the FP64 solve: |k| below the
// limit, or the condition bound above the threshold (compared without the
// division: the denominator lam l_min + 1 is at least 1), or not finite
__device__ __forceinline__ bool eig_queued(const EigQueueGeo& g, float lam, float cd, float lmax, float lmin, int i,
                                           int j, int k) {
  const float num = cd * fmaf(lam, lmax, 1.0f), den = fmaf(lam, lmin, 1.0f);
  const int ky = min(j, g.nyp - j), kz = min(k, g.nzp - k);
  const float fx = (float)i * g.rx, fy = (float)ky * g.ry, fz = (float)kz * g.rz;
  const float r2 = fmaf(fx, fx, fmaf(fy, fy, fz * fz));
  return r2 < g.lowk2 || !(num <= g.cond_thr * den);
}

__device__ __forceinline__ float eig_weight(float lam, float lp) { return __builtin_amdgcn_rcpf(fmaf(lam, lp, 1.0f)); }

// entry q of the posterior covariance: s = lam S_q, l2 = lam^2, w the
// weights, gr / gi the real and imaginary parts of G_pq (gi unused for the
// diagonal entries, whose imaginary part is 0)
template <int NMAX>
__device__ __forceinline__ float2 eig_cov_entry(int n, bool diag, float s, float l2, const float* w, const float* gr,
                                                const float* gi) {
  float xr = 0.0f, xi = 0.0f;
#pragma unroll
  for (int p = 0; p < NMAX; ++p)
    if (p < n) {
      xr = fmaf(w[p], gr[p], xr);
      if (!diag) xi = fmaf(w[p], gi[p], xi);
    }
  return diag ? make_float2(fmaf(-l2, xr, s), 0.0f) : make_float2(fmaf(-l2, xr, s), -l2 * xi);
}
