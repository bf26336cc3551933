// Rock-physics forward model on gfx950: one sample per thread, in FP64.
//
// The model that the learned surrogate stands in for (rplib):
//  * mineral: Hill average of quartz and clay (rplib/demmodelling.cpp:178-236);
//  * dry frame: DEM with empty spheroidal pores, Berryman's P and Q factors
//    (rplib/dem.cpp:39-189), integrated with adaptive Runge-Kutta-Fehlberg
//    4(5) and the reference's step control (rplib/orddiffeqsolver.cpp:17-167);
//  * fluid: Batzle-Wang brine (rplib/demmodelling.cpp:29-55, 274-336) mixed
//    with Batzle-Wang gas (published 1992 equations; rplib has no gas model)
//    by a Reuss average, density linear;
//  * Gassmann substitution in the reference's form (rplib/rockgassmann.cpp:
//    79-97) and vp, vs (rplib/demmodelling.cpp:338-354).
// Units as rplib: moduli in kPa, density g/cm3, velocity m/s, T in deg C,
// P in MPa.  Outputs ln vp, ln vs, ln rho (rho in kg/m3).
#include "internal.hpp"

#include <algorithm>

namespace {

constexpr int kThreads = 128;

__device__ void dem_rhs(double asp, double conc, double ka, double mua, const double y[2], double t, double out[2]) {
  const double k = y[0], mu = y[1];
  if (asp == 1.0) asp = 0.99;
  const double theta = (asp / pow(1 - asp * asp, 1.5)) * (acos(asp) - asp * sqrt(1 - asp * asp));
  const double fn = ((asp * asp) / (1 - asp * asp)) * (3 * theta - 2);
  const double nu = (3 * k - 2 * mu) / (2 * (3 * k + mu));
  const double r = (1 - 2 * nu) / (2 * (1 - nu));
  const double a = mua / mu - 1;
  const double b = (1.0 / 3.0) * (ka / k - mua / mu);
  const double f1 = 1 + a * (1.5 * (fn + theta) - r * (1.5 * fn + 2.5 * theta - 4.0 / 3.0));
  double f2 = 1 + a * (1 + 1.5 * (fn + theta) - (r / 2) * (3 * fn + 5 * theta)) + b * (3 - 4 * r);
  f2 += (a / 2) * (a + 3 * b) * (3 - 4 * r) * (fn + theta - r * (fn - theta + 2 * theta * theta));
  const double f3 = 1 + a * (1 - (fn + 1.5 * theta) + r * (fn + theta));
  const double f4 = 1 + (a / 4) * (fn + 3 * theta - r * (fn - theta));
  const double f5 = a * (-fn + r * (fn + theta - 4.0 / 3.0)) + b * theta * (3 - 4 * r);
  const double f6 = 1 + a * (1 + fn - r * (fn + theta)) + b * (1 - theta) * (3 - 4 * r);
  const double f7 = 2 + (a / 4) * (3 * fn + 9 * theta - r * (3 * fn + 5 * theta)) + b * theta * (3 - 4 * r);
  const double f8 = a * (1 - 2 * r + (fn / 2) * (r - 1) + (theta / 2) * (5 * r - 3)) + b * (1 - theta) * (3 - 4 * r);
  const double f9 = a * ((r - 1) * fn - r * theta) + b * theta * (3 - 4 * r);
  const double p = f1 / f2;
  const double q = ((2 / f3) + (1 / f4) + ((f4 * f5 + f6 * f7 - f8 * f9) / (f2 * f4))) / 5.0;
  out[0] = conc * (ka - k) * p / (1 - t);
  out[1] = conc * (mua - mu) * q / (1 - t);
}

// returns 0 on success, 1 when the step size collapsed or the budget ran out
__device__ int dem_rkf45(double asp, double tfinal, double ka, double mua, double y[2], int max_steps) {
  const double al[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
  const double be[5][5] = {{1.0 / 4.0, 0, 0, 0, 0},
                           {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0},
                           {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0},
                           {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0, 0},
                           {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0,
                            -5643.0 / 20520.0}};
  const double g0[6] = {902880.0 / 7618050.0, 0, 3953664.0 / 7618050.0, 3855735.0 / 7618050.0,
                        -1371249.0 / 7618050.0, 277020.0 / 7618050.0};
  const double g1[6] = {-2090.0 / 752400.0, 0, 22528.0 / 752400.0, 21970.0 / 752400.0, -15048.0 / 752400.0,
                        -27360.0 / 752400.0};
  const double tol = 1e-5;
  double t = 0.0;
  const double hmax = tfinal / 16.0;
  double h = hmax / 8.0;
  double f[6][2];
  int iters = 0;
  while (t < tfinal && (t + h) > t) {
    if (++iters > max_steps) return 1;
    if (t + h > tfinal) h = tfinal - t;
    dem_rhs(asp, 1.0, ka, mua, y, t, f[0]);
#pragma unroll
    for (int j = 0; j < 5; ++j) {
      double y1[2] = {y[0], y[1]};
#pragma unroll
      for (int m = 0; m <= j; ++m) {
        y1[0] += h * be[j][m] * f[m][0];
        y1[1] += h * be[j][m] * f[m][1];
      }
      dem_rhs(asp, 1.0, ka, mua, y1, t + al[j] * h, f[j + 1]);
    }
    double e0 = 0, e1 = 0;
#pragma unroll
    for (int m = 0; m < 6; ++m) { e0 += h * g1[m] * f[m][0]; e1 += h * g1[m] * f[m][1]; }
    const double delta = fmax(fabs(e0), fabs(e1));
    const double tau = fmax(fmax(fabs(y[0]), fabs(y[1])), 1.0) * tol;
    if (delta <= tau) {
      t += h;
#pragma unroll
      for (int m = 0; m < 6; ++m) { y[0] += h * g0[m] * f[m][0]; y[1] += h * g0[m] * f[m][1]; }
    }
    if (delta != 0.0) {
      h = 0.8 * h * pow(tau / delta, 0.2);
      if (hmax < h) h = hmax;
    }
  }
  return (t < tfinal) ? 1 : 0;
}

__device__ double water_velocity(double T, double P) {
  const double w[5][4] = {{1402.85, 1.524, 3.437E-3, -1.197E-5},
                          {4.871, -0.0111, 1.739E-4, -1.628E-6},
                          {-0.04783, 2.747E-4, -2.135E-6, 1.237E-8},
                          {1.487E-4, -6.503E-7, -1.455E-8, 1.327E-10},
                          {-2.197E-7, 7.987E-10, 5.230E-11, -4.614E-13}};
  double s = 0.0, ti = 1.0;
  for (int i = 0; i < 5; ++i) {
    double pj = 1.0;
    for (int j = 0; j < 4; ++j) { s += w[i][j] * ti * pj; pj *= P; }
    ti *= T;
  }
  return s;
}

__device__ double water_density(double T, double P) {
  return 1.0 + 1E-6 * (-80 * T - 3.3 * (T * T) + 0.00175 * (T * T * T) + 489 * P - 2 * T * P + 0.016 * (T * T) * P -
                       1.3E-5 * (T * T * T) * P - 0.333 * (P * P) - 0.002 * T * (P * P));
}

__device__ void brine(double T, double P, double S, double& k_kpa, double& rho) {
  const double v = water_velocity(T, P) +
                   S * (1170.0 - 9.6 * T + 0.055 * (T * T) - 8.5E-5 * (T * T * T) + 2.6 * P - 0.0029 * T * P -
                        0.0476 * (P * P)) +
                   pow(S, 1.5) * (780.0 - 10.0 * P + 0.16 * (P * P)) - 1820 * (S * S);
  rho = water_density(T, P) +
        S * (0.668 + 0.44 * S + 1E-6 * (300 * P - 2400 * P * S + T * (80 + 3 * T - 3300 * S - 13 * P + 47 * P * S)));
  k_kpa = rho * v * v;  // (rho v^2 / 1e3 MPa) * 1e3
}

// Batzle & Wang (1992) gas, gravity G: density g/cm3 and modulus kPa
__device__ void gas(double T, double P, double G, double& k_kpa, double& rho) {
  const double Ta = T + 273.15;
  const double Tpr = Ta / (94.72 + 170.75 * G);
  const double Ppr = P / (4.892 - 0.4048 * G);
  const double c = 0.45 + 8.0 * (0.56 - 1.0 / Tpr) * (0.56 - 1.0 / Tpr);
  const double E = 0.109 * (3.85 - Tpr) * (3.85 - Tpr) * exp(-c * pow(Ppr, 1.2) / Tpr);
  const double a3 = 0.03 + 0.00527 * (3.5 - Tpr) * (3.5 - Tpr) * (3.5 - Tpr);
  const double Z = a3 * Ppr + (0.642 * Tpr - 0.007 * Tpr * Tpr * Tpr * Tpr - 0.52) + E;
  const double dZ = a3 + E * (-c * 1.2 * pow(Ppr, 0.2) / Tpr);
  rho = 28.8 * G * P / (Z * 8.31441 * Ta);
  const double g0 = 0.85 + 5.6 / (Ppr + 2) + 27.1 / ((Ppr + 3.5) * (Ppr + 3.5)) - 8.7 * exp(-0.65 * (Ppr + 1));
  k_kpa = P * g0 / (1.0 - Ppr / Z * dZ) * 1E3;
}

// one sample: in = (porosity, clay, Sw, aspect, Pe MPa) -> ln vp, ln vs, ln rho
__device__ __forceinline__ int rock_sample(const crv_rock_params& p, double phi, double clay, double sw, double asp,
                                           double pe, int max_steps, double out[3]) {
  const double vq = 1.0 - clay, vc = clay;
  const double kmin = 0.5 * ((vq * p.k_quartz + vc * p.k_clay) + 1.0 / (vq / p.k_quartz + vc / p.k_clay));
  const double gmin = 0.5 * ((vq * p.g_quartz + vc * p.g_clay) + 1.0 / (vq / p.g_quartz + vc / p.g_clay));
  const double rhomin = vq * p.rho_quartz + vc * p.rho_clay;
  double y[2] = {kmin, gmin};
  int st = 0;
  if (phi != 1.0) st = dem_rkf45(asp, phi, 0.0, 0.0, y, max_steps);
  const double kd = y[0], gd = y[1];
  double kb, rb, kg, rg;
  brine(p.temperature, pe, p.salinity, kb, rb);
  gas(p.temperature, pe, p.gas_gravity, kg, rg);
  const double kfl = 1.0 / (sw / kb + (1.0 - sw) / kg);
  const double rfl = sw * rb + (1.0 - sw) * rg;
  const double t1 = (1 - kd / kmin) * (1 - kd / kmin);
  const double k = kd + t1 / (phi / kfl + (1 - phi) / kmin + kd / (kmin * kmin));
  const double rho = (1 - phi) * rhomin + phi * rfl;
  out[0] = log(sqrt((k + 4.0 / 3.0 * gd) / rho));
  out[1] = log(sqrt(gd / rho));
  out[2] = log(rho * 1000.0);
  return st;
}

template <typename TO>
__global__ __launch_bounds__(kThreads) void k_rock(crv_rock_params p, long long n, const double* __restrict__ in,
                                                   TO* __restrict__ out, int* __restrict__ status, int max_steps) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double o[3];
    const int st = rock_sample(p, in[5 * i], in[5 * i + 1], in[5 * i + 2], in[5 * i + 3], in[5 * i + 4], max_steps, o);
    out[3 * i] = (TO)o[0];
    out[3 * i + 1] = (TO)o[1];
    out[3 * i + 2] = (TO)o[2];
    if (status) status[i] = st;
  }
}

// Error gate: list the samples whose error-table cell predicts more than
// thresh (one atomic per wavefront for the list slots).
__global__ __launch_bounds__(kThreads) void k_gate(long long n, const float* __restrict__ x, crv_mlp_norm nm,
                                                   const float* __restrict__ table, int nb, float thresh,
                                                   unsigned int* __restrict__ list, unsigned int* __restrict__ count) {
  for (long long base = (long long)blockIdx.x * blockDim.x; base < n; base += (long long)gridDim.x * blockDim.x) {
    const long long i = base + threadIdx.x;
    bool flag = false;
    if (i < n) {
      int cell = 0;
#pragma unroll
      for (int f = 0; f < 5; ++f) {
        float v = x[5 * i + f];
        if (nm.in_log_mask & (1 << f)) v = log10f(v);
        const float t = ((v - nm.in_lo[f]) * nm.in_scale[f]) * 0.5f;  // [0, 1] inside the box
        const int b = min(nb - 1, max(0, (int)floorf(t * (float)nb)));
        cell = cell * nb + b;
      }
      flag = !(table[cell] <= thresh);
    }
    const unsigned long long mask = __ballot(flag);
    const int lane = threadIdx.x & 63;
    unsigned int slot0 = 0;
    if (lane == 0 && mask) slot0 = atomicAdd(count, (unsigned int)__popcll(mask));
    slot0 = __shfl(slot0, 0, 64);
    if (flag) list[slot0 + __popcll(mask & ((1ull << lane) - 1ull))] = (unsigned int)i;
  }
}

// rplib chain for the listed samples (float in/out)
__global__ __launch_bounds__(kThreads) void k_rock_list(crv_rock_params p, const unsigned int* __restrict__ count,
                                                        const unsigned int* __restrict__ list,
                                                        const float* __restrict__ x, float* __restrict__ y,
                                                        int max_steps) {
  const unsigned int m = *count;
  for (unsigned int t = blockIdx.x * blockDim.x + threadIdx.x; t < m; t += gridDim.x * blockDim.x) {
    const long long i = list[t];
    double o[3];
    const int st = rock_sample(p, x[5 * i], x[5 * i + 1], x[5 * i + 2], x[5 * i + 3], x[5 * i + 4], max_steps, o);
    const float nan = __builtin_nanf("");
    y[3 * i] = st ? nan : (float)o[0];
    y[3 * i + 1] = st ? nan : (float)o[1];
    y[3 * i + 2] = st ? nan : (float)o[2];
  }
}

}  // namespace

extern "C" int crv_rock_forward(crv_ctx* ctx, const crv_rock_params* p, long long n, const double* d_in, void* d_out,
                                int out_fp32, int* d_status) {
  if (!ctx) return CRV_E_INVALID;
  if (!p || n < 0 || (n > 0 && (!d_in || !d_out))) return crv_set_error(ctx, CRV_E_INVALID, "rock_forward: bad args");
  if (n == 0) return CRV_OK;
  const int nb = (int)std::min<long long>((n + kThreads - 1) / kThreads, 1 << 20);
  ProfScope ps(ctx, "rock_forward");
  if (out_fp32)
    hipLaunchKernelGGL(k_rock<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, *p, n, d_in, (float*)d_out, d_status,
                       4096);
  else
    hipLaunchKernelGGL(k_rock<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, *p, n, d_in, (double*)d_out,
                       d_status, 4096);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

extern "C" int crv_rock_refine(crv_ctx* ctx, const crv_rock_params* p, long long n, const float* d_x, float* d_y,
                               const crv_mlp_norm* norm, const float* d_table, int nb, float thresh,
                               long long* n_refined) {
  if (!ctx) return CRV_E_INVALID;
  if (!p || !norm || n < 0 || nb < 1 || (n > 0 && (!d_x || !d_y || !d_table)))
    return crv_set_error(ctx, CRV_E_INVALID, "rock_refine: bad args");
  if (n > 0xffffffffLL) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "rock_refine: more than 2^32 samples");
  if (n_refined) *n_refined = 0;
  if (n == 0) return CRV_OK;
  unsigned int* buf = nullptr;
  CRV_HIP_CHECK(ctx, hipMallocAsync((void**)&buf, ((size_t)n + 64) * sizeof(unsigned int), ctx->stream));
  unsigned int* count = buf;
  unsigned int* list = buf + 64;
  CRV_HIP_CHECK(ctx, hipMemsetAsync(count, 0, sizeof(unsigned int), ctx->stream));
  {
    ProfScope ps(ctx, "rock_gate");
    const int nb_g = (int)std::min<long long>((n + kThreads - 1) / kThreads, 1 << 16);
    hipLaunchKernelGGL(k_gate, dim3(nb_g), dim3(kThreads), 0, ctx->stream, n, d_x, *norm, d_table, nb, thresh, list,
                       count);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  {
    ProfScope ps(ctx, "rock_refine");
    const int nb_r = (int)std::min<long long>((n + kThreads - 1) / kThreads, 4096);
    hipLaunchKernelGGL(k_rock_list, dim3(nb_r), dim3(kThreads), 0, ctx->stream, *p, (const unsigned int*)count,
                       (const unsigned int*)list, d_x, d_y, 4096);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  unsigned int hc = 0;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(&hc, count, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipFreeAsync(buf, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (n_refined) *n_refined = hc;
  return CRV_OK;
}
