// Facies probabilities from facies-conditional elastic samples (gfx950).
//
// The role of FaciesProb's density estimation and classification
// (reference src/faciesprob.cpp:544-930, posteriorelasticpdf3d.cpp): for
// each facies, the (ln vp, ln vs, ln rho) samples generated from its rock
// model (rplib or its surrogate) are binned on a regular box, the
// histogram is smoothed by a Gaussian whose covariance is the posterior
// uncertainty plus a kernel bandwidth (a product in the Fourier domain,
// using the library's own transforms), and normalised to a density; a
// voxel with posterior mean m then gets
//   P(f | m) = pi_f p_f(m) / sum_g pi_g p_g(m)
// with p_f read by trilinear interpolation.
//
// Density grids are FP32 FFT grids of the box: x = ln vp (nxp bins),
// y = ln vs, z = ln rho; bin centres lo + (i + 0.5) * (hi - lo) / nbins.
#include "internal.hpp"

#include <algorithm>
#include <cmath>

namespace {

constexpr int kThreads = 256;

struct Box {
  float lo[3], inv_w[3];  // inverse bin width
  int nb[3];
  int rnx;                // row length of the real grid
};

Box make_box(const crv_facies_box* b, const crv_grid* g) {
  Box x;
  x.nb[0] = g->nxp; x.nb[1] = g->nyp; x.nb[2] = g->nzp;
  for (int d = 0; d < 3; ++d) {
    x.lo[d] = (float)b->lo[d];
    x.inv_w[d] = (float)(x.nb[d] / (b->hi[d] - b->lo[d]));
  }
  x.rnx = g->rnxp;
  return x;
}

// histogram: one count per sample inside the box (float atomics)
__global__ void k_hist(const float* __restrict__ s, long long n, Box b, float* __restrict__ h) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    int c[3];
    bool in = true;
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float v = s[3 * i + d];
      const float t = (v - b.lo[d]) * b.inv_w[d];
      c[d] = (int)floorf(t);
      in = in && (t == t) && c[d] >= 0 && c[d] < b.nb[d];
    }
    if (in) atomicAdd(&h[(size_t)c[0] + (size_t)b.rnx * (c[1] + (size_t)b.nb[1] * c[2])], 1.0f);
  }
}

// multiply the half spectrum by the Gaussian exp(-2 pi^2 f^T S f), f in
// cycles per unit of (ln vp, ln vs, ln rho); S given in those units
__global__ void k_gauss(float2* __restrict__ X, int nx, int ny, int nz, double3 L, double S00, double S11,
                        double S22, double S01, double S02, double S12) {
  const int cnx = nx / 2 + 1;
  const long long nc = (long long)cnx * ny * nz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const int kx = (int)(i % cnx);
    const long long r = i / cnx;
    const int ky = (int)(r % ny), kz = (int)(r / ny);
    const double fx = kx / L.x;
    const double fy = (ky <= ny / 2 ? ky : ky - ny) / L.y;
    const double fz = (kz <= nz / 2 ? kz : kz - nz) / L.z;
    const double q = S00 * fx * fx + S11 * fy * fy + S22 * fz * fz + 2.0 * (S01 * fx * fy + S02 * fx * fz + S12 * fy * fz);
    const float g = (float)exp(-2.0 * 9.8696044010893586188 * q);
    X[i] = make_float2(X[i].x * g, X[i].y * g);
  }
}

// clamp to >= 0 (smoothing ringing) and scale to a density
__global__ void k_normalise(float* __restrict__ d, int nx, int ny, int nz, int rnx, float scale) {
  const long long n = (long long)nx * ny * nz;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % nx);
    const long long r = t / nx;
    const size_t i = (size_t)x + (size_t)rnx * r;
    d[i] = fmaxf(d[i], 0.0f) * scale;
  }
}

__global__ void k_sum(const float* __restrict__ d, int nx, long long rows, int rnx, double* __restrict__ out) {
  __shared__ double sh[kThreads];
  double acc = 0.0;
  const long long n = (long long)nx * rows;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % nx);
    const long long r = t / nx;
    acc += fmax((double)d[(size_t)x + (size_t)rnx * r], 0.0);
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = kThreads / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(out, sh[0]);
}

// PosteriorElasticPDF3D::SetupSmoothingGaussian3D (reference
// src/posteriorelasticpdf3d.cpp:369-420): the Gaussian exp(-x^T Si x / 2)
// sampled at the cyclic bin lags (j for j <= n/2, j - n above), stored in
// float; the sum is taken separately and the kernel divided by it.
__global__ void k_zero_d(double* p) { *p = 0.0; }

__global__ void k_smoother(float* __restrict__ s, int nx, int ny, int nz, int rnx, double dx, double dy, double dz,
                           double i00, double i11, double i22, double i10, double i20, double i21) {
  const long long n = (long long)rnx * ny * nz;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(t % rnx);
    const long long r = t / rnx;
    const int k = (int)(r % ny), l = (int)(r / ny);
    if (j >= nx) { s[t] = 0.0f; continue; }
    const double x = (j <= nx / 2 ? j : j - nx) * dx;
    const double y = (k <= ny / 2 ? k : k - ny) * dy;
    const double z = (l <= nz / 2 ? l : l - nz) * dz;
    const double q = x * x * i00 + y * y * i11 + z * z * i22 + 2.0 * (x * y * i10 + x * z * i20 + y * z * i21);
    s[t] = (float)exp(-0.5 * q);
  }
}

// spectrum of the histogram times spectrum of the kernel, the kernel's
// normalisation folded in
__global__ void k_cmul(float2* __restrict__ a, const float2* __restrict__ b, long long nc, float scale) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nc; i += (long long)gridDim.x * blockDim.x) {
    const float2 x = a[i], y = b[i];
    a[i] = make_float2((y.x * x.x - y.y * x.y) * scale, (y.y * x.x + y.x * x.y) * scale);
  }
}

__device__ __forceinline__ float trilinear(const float* __restrict__ d, const Box& b, float v0, float v1, float v2) {
  float t[3] = {(v0 - b.lo[0]) * b.inv_w[0] - 0.5f, (v1 - b.lo[1]) * b.inv_w[1] - 0.5f,
                (v2 - b.lo[2]) * b.inv_w[2] - 0.5f};
  int i0[3];
  float w[3];
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    // outside the box the density is that of the nearest edge bin
    const float tc = fminf(fmaxf(t[k], 0.0f), (float)(b.nb[k] - 1));
    i0[k] = min((int)floorf(tc), b.nb[k] - 2 < 0 ? 0 : b.nb[k] - 2);
    w[k] = tc - (float)i0[k];
  }
  float acc = 0.0f;
#pragma unroll
  for (int c = 0; c < 8; ++c) {
    const int dx = c & 1, dy = (c >> 1) & 1, dz = c >> 2;
    const int x = min(i0[0] + dx, b.nb[0] - 1), y = min(i0[1] + dy, b.nb[1] - 1), z = min(i0[2] + dz, b.nb[2] - 1);
    const float wt = (dx ? w[0] : 1.0f - w[0]) * (dy ? w[1] : 1.0f - w[1]) * (dz ? w[2] : 1.0f - w[2]);
    acc += wt * d[(size_t)x + (size_t)b.rnx * (y + (size_t)b.nb[1] * z)];
  }
  return acc;
}

struct ProbArgs {
  const float* dens[8];
  float prior[8];
  int nf;
  Box b;
  const float* m[3];   // posterior means (real grids, FP32)
  float* p[8];         // outputs (real grids)
  float* pu;           // undefined class output (may be null)
  // undefined class (FaciesProb::calculateFaciesProb, faciesprob.cpp:2008,
  // 2020-2050): p_undef times the mean bin value of the first density --
  // p_undef / (number of bins) for densities that sum to 1 over the box, as
  // the reference's, and the same uniform share for unit-integral ones
  float p_undef;
  const double* dsum;  // sum of the first density over its bins (p_undef > 0)
  double nbins;
  int nx, ny, nz, rnx; // voxel grid (padded, row length)
};

__global__ void k_prob(ProbArgs a) {
  const long long n = (long long)a.nx * a.ny * a.nz;
  const float undef = a.p_undef > 0.0f ? (float)((double)a.p_undef * (*a.dsum / a.nbins)) : 0.0f;
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(t % a.nx);
    const long long r = t / a.nx;
    const size_t i = (size_t)x + (size_t)a.rnx * r;
    const float v0 = a.m[0][i], v1 = a.m[1][i], v2 = a.m[2][i];
    float q[8], s = undef;
    for (int f = 0; f < a.nf; ++f) {
      q[f] = a.prior[f] * trilinear(a.dens[f], a.b, v0, v1, v2);
      s += q[f];
    }
    const float inv = s > 0.0f ? 1.0f / s : 0.0f;
    for (int f = 0; f < a.nf; ++f) a.p[f][i] = s > 0.0f ? q[f] * inv : 1.0f / (float)a.nf;
    if (a.pu) a.pu[i] = s > 0.0f ? undef * inv : 0.0f;
  }
}

int nblocks(long long n) { return (int)std::min<long long>((n + kThreads - 1) / kThreads, 8192); }

}  // namespace

extern "C" int crv_facies_density(crv_ctx* ctx, const float* d_samples, long long n, const crv_facies_box* box,
                                  const double* smooth_cov, crv_grid* density) {
  if (!ctx) return CRV_E_INVALID;
  if (!box || !density || !smooth_cov || n < 0 || (n > 0 && !d_samples))
    return crv_set_error(ctx, CRV_E_INVALID, "facies_density: bad args");
  if (density->precision != CRV_FP32) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: FP32 grid needed");
  for (int d = 0; d < 3; ++d)
    if (!(box->hi[d] > box->lo[d])) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: empty box");
  const Box b = make_box(box, density);
  float* g = (float*)density->d;
  CRV_HIP_CHECK(ctx, hipMemsetAsync(g, 0, density->rsize * sizeof(float), ctx->stream));
  {
    ProfScope ps(ctx, "facies_hist");
    if (n > 0) hipLaunchKernelGGL(k_hist, dim3(nblocks(n)), dim3(kThreads), 0, ctx->stream, d_samples, n, b, g);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  // smoothing: unscaled forward and inverse transforms of the histogram
  density->cubetype = CRV_OPERATOR;  // OPERATOR: no 1/sqrt(N) on the forward pass, 1/N on the inverse
  density->transformed = false;
  int rc = crv_fft_r2c(density);
  if (rc != CRV_OK) return rc;
  const double3 L = make_double3(box->hi[0] - box->lo[0], box->hi[1] - box->lo[1], box->hi[2] - box->lo[2]);
  const long long nc = (long long)density->csize;
  {
    ProfScope ps(ctx, "facies_smooth");
    hipLaunchKernelGGL(k_gauss, dim3(nblocks(nc)), dim3(kThreads), 0, ctx->stream, (float2*)density->d,
                       density->nxp, density->nyp, density->nzp, L, smooth_cov[0], smooth_cov[4], smooth_cov[8],
                       smooth_cov[1], smooth_cov[2], smooth_cov[5]);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  if ((rc = crv_fft_c2r(density)) != CRV_OK) return rc;
  // normalise: integral over the box = 1
  double* dsum = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dsum, sizeof(double), ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(dsum, 0, sizeof(double), ctx->stream));
  hipLaunchKernelGGL(k_sum, dim3(256), dim3(kThreads), 0, ctx->stream, (const float*)g, density->nxp,
                     (long long)density->nyp * density->nzp, density->rnxp, dsum);
  double hs = 0.0;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(&hs, dsum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipFreeAsync(dsum, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  const double cell = L.x * L.y * L.z / ((double)density->nxp * density->nyp * density->nzp);
  const float scale = hs > 0.0 ? (float)(1.0 / (hs * cell)) : 0.0f;
  hipLaunchKernelGGL(k_normalise, dim3(nblocks((long long)density->nxp * density->nyp * density->nzp)),
                     dim3(kThreads), 0, ctx->stream, g, density->nxp, density->nyp, density->nzp, density->rnxp,
                     scale);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// PosteriorElasticPDF3D's density (reference src/posteriorelasticpdf3d.cpp:
// 12-131): the histogram of the samples divided by their number, circularly
// convolved with the sampled Gaussian kernel normalised to sum 1 (products of
// unscaled transforms, 1/N on the inverse): a probability mass per bin that
// sums to 1 over the box.  The kernel's sum is taken in double on the device.
static int facies_density_sampled(crv_ctx* ctx, const float* d_samples, long long n, const crv_facies_box* box,
                                  const double* sigma, crv_grid* density) {
  // inverse of the 3x3 kernel covariance (InvertSquareMatrix), by cofactors
  const double* a = sigma;
  const double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
  const double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
  if (!(det > 0.0)) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: kernel covariance not positive definite");
  double inv[9];
  inv[0] = c00 / det; inv[1] = (a[2] * a[7] - a[1] * a[8]) / det; inv[2] = (a[1] * a[5] - a[2] * a[4]) / det;
  inv[3] = c01 / det; inv[4] = (a[0] * a[8] - a[2] * a[6]) / det; inv[5] = (a[2] * a[3] - a[0] * a[5]) / det;
  inv[6] = c02 / det; inv[7] = (a[1] * a[6] - a[0] * a[7]) / det; inv[8] = (a[0] * a[4] - a[1] * a[3]) / det;
  const Box b = make_box(box, density);
  const int nx = density->nxp, ny = density->nyp, nz = density->nzp;
  float* g = (float*)density->d;
  FftPlan* plan = nullptr;
  int rc = crv_get_plan(ctx, nx, ny, nz, CRV_FP32, &plan);
  if (rc != CRV_OK) return rc;
  const size_t rbytes = density->rsize * sizeof(float);
  float* sm = nullptr;
  double* dsum = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&sm, rbytes, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dsum, sizeof(double), ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(g, 0, rbytes, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(dsum, 0, sizeof(double), ctx->stream));
  {
    ProfScope ps(ctx, "facies_hist");
    if (n > 0) hipLaunchKernelGGL(k_hist, dim3(nblocks(n)), dim3(kThreads), 0, ctx->stream, d_samples, n, b, g);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  const double dx = (box->hi[0] - box->lo[0]) / nx, dy = (box->hi[1] - box->lo[1]) / ny,
               dz = (box->hi[2] - box->lo[2]) / nz;
  {
    ProfScope ps(ctx, "facies_kernel");
    hipLaunchKernelGGL(k_smoother, dim3(nblocks((long long)density->rsize)), dim3(kThreads), 0, ctx->stream, sm, nx,
                       ny, nz, density->rnxp, dx, dy, dz, inv[0], inv[4], inv[8], inv[3], inv[6], inv[7]);
    hipLaunchKernelGGL(k_sum, dim3(256), dim3(kThreads), 0, ctx->stream, (const float*)sm, nx, (long long)ny * nz,
                       density->rnxp, dsum);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  double ksum = 0.0;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(&ksum, dsum, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (!(ksum > 0.0)) {
    (void)hipFreeAsync(sm, ctx->stream);
    (void)hipFreeAsync(dsum, ctx->stream);
    return crv_set_error(ctx, CRV_E_INVALID, "facies_density: empty smoothing kernel");
  }
  // unscaled forward transforms; the product carries 1/(n sum N): the
  // histogram's 1/n, the kernel's 1/sum and the inverse transform's 1/N
  rc = crv_launch_r2c(ctx, plan, g, g, 1.0);
  if (rc == CRV_OK) rc = crv_launch_r2c(ctx, plan, sm, sm, 1.0);
  if (rc == CRV_OK) {
    ProfScope ps(ctx, "facies_smooth");
    const double N = (double)nx * ny * nz;
    const float scale = (float)(1.0 / ((n > 0 ? (double)n : 1.0) * ksum * N));
    hipLaunchKernelGGL(k_cmul, dim3(nblocks((long long)density->csize)), dim3(kThreads), 0, ctx->stream, (float2*)g,
                       (const float2*)sm, (long long)density->csize, scale);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) rc = crv_set_error(ctx, CRV_E_HIP, hipGetErrorString(e));
  }
  if (rc == CRV_OK) rc = crv_launch_c2r(ctx, plan, g, g, 1.0, nullptr);
  CRV_HIP_CHECK(ctx, hipFreeAsync(sm, ctx->stream));
  CRV_HIP_CHECK(ctx, hipFreeAsync(dsum, ctx->stream));
  if (rc != CRV_OK) return rc;
  density->cubetype = CRV_PARAMETER;
  density->transformed = false;
  return CRV_OK;
}

extern "C" int crv_facies_density_kernel(crv_ctx* ctx, const float* d_samples, long long n,
                                         const crv_facies_box* box, const double* smooth_cov, int kernel,
                                         crv_grid* density) {
  if (kernel == CRV_SMOOTH_FOURIER) return crv_facies_density(ctx, d_samples, n, box, smooth_cov, density);
  if (!ctx) return CRV_E_INVALID;
  if (kernel != CRV_SMOOTH_SAMPLED) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: unknown kernel");
  if (!box || !density || !smooth_cov || n < 0 || (n > 0 && !d_samples))
    return crv_set_error(ctx, CRV_E_INVALID, "facies_density: bad args");
  if (density->precision != CRV_FP32) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: FP32 grid needed");
  for (int d = 0; d < 3; ++d)
    if (!(box->hi[d] > box->lo[d])) return crv_set_error(ctx, CRV_E_INVALID, "facies_density: empty box");
  return facies_density_sampled(ctx, d_samples, n, box, smooth_cov, density);
}

extern "C" int crv_facies_prob_undef(crv_ctx* ctx, int nf, crv_grid* const* density, const double* prior,
                                     double p_undefined, const crv_facies_box* box, crv_grid* mean[3],
                                     crv_grid** prob, crv_grid* prob_undef) {
  if (!ctx) return CRV_E_INVALID;
  if (nf < 1 || nf > 8 || !density || !prior || !box || !mean || !prob || !(p_undefined >= 0.0))
    return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: bad args");
  const crv_grid* d0 = density[0];
  const crv_grid* m0 = mean[0];
  if (!d0 || !m0) return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: NULL grid");
  auto same_out = [&](const crv_grid* p) {
    return p && p->nxp == m0->nxp && p->nyp == m0->nyp && p->nzp == m0->nzp && p->precision == CRV_FP32;
  };
  for (int f = 0; f < nf; ++f) {
    const crv_grid* d = density[f];
    if (!d || d->nxp != d0->nxp || d->nyp != d0->nyp || d->nzp != d0->nzp || d->precision != CRV_FP32 ||
        d->transformed)
      return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: densities must be real FP32 grids of one size");
    if (!same_out(prob[f])) return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: output grids inconsistent");
  }
  if (prob_undef && !same_out(prob_undef))
    return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: undefined-class grid inconsistent");
  for (int g = 0; g < 3; ++g) {
    const crv_grid* m = mean[g];
    if (!m || m->nxp != m0->nxp || m->nyp != m0->nyp || m->nzp != m0->nzp || m->precision != CRV_FP32 ||
        m->transformed)
      return crv_set_error(ctx, CRV_E_INVALID, "facies_prob: means must be real FP32 grids of one size");
  }
  ProbArgs a{};
  a.nf = nf;
  a.b = make_box(box, d0);
  double ps = 0.0;
  for (int f = 0; f < nf; ++f) ps += prior[f];
  for (int f = 0; f < nf; ++f) {
    a.dens[f] = (const float*)density[f]->d;
    a.prior[f] = (float)(ps > 0.0 ? prior[f] / ps : 1.0 / nf);
    a.p[f] = (float*)prob[f]->d;
    prob[f]->transformed = false;
  }
  if (prob_undef) {
    a.pu = (float*)prob_undef->d;
    prob_undef->transformed = false;
  }
  for (int g = 0; g < 3; ++g) a.m[g] = (const float*)mean[g]->d;
  a.nx = m0->nxp; a.ny = m0->nyp; a.nz = m0->nzp; a.rnx = m0->rnxp;
  a.p_undef = (float)p_undefined;
  a.nbins = (double)d0->nxp * d0->nyp * d0->nzp;
  double* dsum = nullptr;
  if (p_undefined > 0.0) {
    CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&dsum, sizeof(double), ctx->stream));
    // a kernel clears the accumulator (no memset node in a captured graph)
    hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(1), 0, ctx->stream, dsum);
    hipLaunchKernelGGL(k_sum, dim3(256), dim3(kThreads), 0, ctx->stream, (const float*)d0->d, d0->nxp,
                       (long long)d0->nyp * d0->nzp, d0->rnxp, dsum);
    CRV_HIP_CHECK(ctx, hipGetLastError());
    a.dsum = dsum;
  }
  {
    ProfScope ps_(ctx, "facies_prob");
    hipLaunchKernelGGL(k_prob, dim3(nblocks((long long)a.nx * a.ny * a.nz)), dim3(kThreads), 0, ctx->stream, a);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  if (dsum) CRV_HIP_CHECK(ctx, hipFreeAsync(dsum, ctx->stream));
  return CRV_OK;
}

extern "C" int crv_facies_prob(crv_ctx* ctx, int nf, crv_grid* const* density, const double* prior,
                               const crv_facies_box* box, crv_grid* mean[3], crv_grid** prob) {
  return crv_facies_prob_undef(ctx, nf, density, prior, 0.0, box, mean, prob, nullptr);
}
