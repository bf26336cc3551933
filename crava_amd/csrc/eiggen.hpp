// Pieces of the eigen-basis posterior (FP32 refine mode, separable prior)
// that several kernels evaluate on the same values and must agree on bit for
// bit: the posterior kernel itself, the z inverse pass that generates the
// covariance planes from the per-plane tables instead of reading them, and
// the z pass that runs the posterior of the means between its forward and
// inverse transforms.
//   * the plane record (EigLayout);
//   * the queue rule: a wavenumber goes to the FP64 solve when |k| is below
//     the low-wavenumber limit or its condition bound
//     cond(D) (lam l_max + 1) / (lam l_min + 1) exceeds the threshold;
//   * entry q of S_post = lam S - lam^2 sum_p w_p G_p, w_p = 1 / (lam l_p + 1);
//   * the update of the means and the residuals (eig_mean_update).
// Every multiply-add is written as fmaf so that the kernels round alike.
#pragma once
#include "internal.hpp"

// Plane record of the eigen-basis posterior, kEigRec floats per kz plane.
// On a plane k of the band M(lam) = lam E + D with D diagonal and positive;
// with F = D^-1/2 E D^-1/2 = U diag(l) U^H (made on the host in FP64),
// T = U^H D^-1/2 and V = T B (B = |ct| K Sigma0), M^-1 = T^H diag(w) T.
constexpr int kEigRec = 256;  // floats per plane record
constexpr int kEigMaxN = 6;   // angles the record has room for
template <int N>
struct EigLayout {
  static constexpr int OT = 0;              // T, N x N complex [p][l]
  static constexpr int OV = OT + 2 * N * N; // V, N x 3 complex [p][m]
  static constexpr int OK = OV + 6 * N;     // K, N x 3 complex [l][m]
  static constexpr int OG = OK + 6 * N;     // G_p: 3 real, then 3 complex (ab, ar, br)
  static constexpr int OW = OG + 9 * N;     // l_p
  static constexpr int OS = OW + N;         // |ct| Sigma0 (aa bb rr ab ar br)
  static constexpr int OF = OS + 6;         // factor of S (fused_sim_prior)
  static constexpr int OC = OF + 6;         // cond(D), l_max, l_min
  static constexpr int OBAND = kEigRec - 1; // 1: inverted plane
  static_assert(OC + 3 <= OBAND, "record overflow");
};

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
// diagonal entries, whose imaginary part is 0).  Terms p >= n are computed
// and discarded by a select (their inputs may hold anything): a branch on
// the run-time n would split the caller's code into blocks that the
// compiler schedules one after another.
template <int NMAX>
__device__ __forceinline__ float2 eig_cov_entry(int n, bool diag, float s, float l2, const float* w, const float* gr,
                                                const float* gi) {
  float xr = 0.0f, xi = 0.0f;
#pragma unroll
  for (int p = 0; p < NMAX; ++p) {
    const float tr = fmaf(w[p], gr[p], xr);
    xr = p < n ? tr : xr;
    if (!diag) {
      const float ti = fmaf(w[p], gi[p], xi);
      xi = p < n ? ti : xi;
    }
  }
  return diag ? make_float2(fmaf(-l2, xr, s), 0.0f) : make_float2(fmaf(-l2, xr, s), -l2 * xi);
}

// Posterior mean and residuals of one wavenumber of a band plane from its
// record er (any pointer type: scalar loads where the plane is uniform over
// the wave, vector loads otherwise), lam = |cxy(i,j)|, the prior means mu
// and the data d (n <= NMAX angles; NMAX = n when n is a compile-time
// constant).  With r = d - K mu and z = T r:
//   mp = mu + dm,  dm = lam V^H diag(w) z,  res = r - K dm.
template <int NMAX, typename RecP>
__device__ __forceinline__ void eig_mean_update(int n, RecP er, float lam, const float2 mu[3], const float2 d[NMAX],
                                                float2 mp[3], float2 res[NMAX]) {
  const int OV = 2 * n * n, OK = OV + 6 * n, OW = OK + 15 * n;
  float2 r[NMAX];
#pragma unroll
  for (int l = 0; l < NMAX; ++l) {
    if (l < n) {
      float2 x = d[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const float kr = er[OK + 2 * (l * 3 + m)], ki = er[OK + 2 * (l * 3 + m) + 1];
        x.x = fmaf(-kr, mu[m].x, fmaf(ki, mu[m].y, x.x));
        x.y = fmaf(-kr, mu[m].y, fmaf(-ki, mu[m].x, x.y));
      }
      r[l] = x;
    }
  }
  // y = lam diag(w) T r
  float2 y[NMAX];
#pragma unroll
  for (int p = 0; p < NMAX; ++p) {
    if (p < n) {
      float2 z = make_float2(0.0f, 0.0f);
#pragma unroll
      for (int l = 0; l < NMAX; ++l) {
        if (l < n) {
          const float tr = er[2 * (p * n + l)], ti = er[2 * (p * n + l) + 1];
          z.x = fmaf(tr, r[l].x, fmaf(-ti, r[l].y, z.x));
          z.y = fmaf(tr, r[l].y, fmaf(ti, r[l].x, z.y));
        }
      }
      const float g = lam * eig_weight(lam, er[OW + p]);
      y[p] = make_float2(g * z.x, g * z.y);
    }
  }
  // mean update V^H y, residual r - K dm
  float2 dm[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    float2 x = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int p = 0; p < NMAX; ++p) {
      if (p < n) {
        const float vr = er[OV + 2 * (p * 3 + m)], vi = er[OV + 2 * (p * 3 + m) + 1];
        x.x = fmaf(vr, y[p].x, fmaf(vi, y[p].y, x.x));
        x.y = fmaf(vr, y[p].y, fmaf(-vi, y[p].x, x.y));
      }
    }
    dm[m] = x;
    mp[m] = make_float2(mu[m].x + x.x, mu[m].y + x.y);
  }
#pragma unroll
  for (int l = 0; l < NMAX; ++l) {
    if (l < n) {
      float2 x = r[l];
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const float kr = er[OK + 2 * (l * 3 + m)], ki = er[OK + 2 * (l * 3 + m) + 1];
        x.x = fmaf(-kr, dm[m].x, fmaf(ki, dm[m].y, x.x));
        x.y = fmaf(-kr, dm[m].y, fmaf(-ki, dm[m].x, x.y));
      }
      res[l] = x;
    }
  }
}
