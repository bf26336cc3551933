// Complex white noise for posterior simulation on gfx950.
//
// Two sources:
//  * the reference stream: RandomGen (reference libs/lib/random.cpp:38-119,
//    libs/lib/random.h:11-14) -- a 32-bit LCG s <- 69069 s + 1 feeding
//    Marsaglia-Bray normals that consume a data-dependent number of uniforms.
//    It is decoded in parallel: the draw sequence is cut into chunks; each
//    chunk first walks its normals from its own first draw and marks the
//    starts; walks from any other entry point join that path within a few
//    normals, so a second pass started at the true entry (the previous
//    chunk's exit) finds the join, and passes repeat until no exit changes.
//    An exclusive scan of per-chunk counts places every normal.
//  * a counter-based stream (Philox-4x32-10, Salmon et al., SC'11, with a
//    Box-Muller transform) keyed by (seed, realization, grid, index), which
//    needs no decode and is computed where it is used.
// Either stream is laid onto the half spectrum with the reference's
// Hermitian rules (FFTGrid::fillInComplexNoise, src/fftgrid.cpp:1191-1240).
#include "internal.hpp"
#include "noise.hpp"

#include <algorithm>
#include <cstring>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 2048;           // draws per chunk
constexpr int kWords = kChunk / 32;    // bitmap words per chunk

struct Affine { unsigned a, c; };      // s -> a s + c

struct SkipTable { Affine p[32]; };    // LCG^(2^i)

SkipTable make_skip() {
  SkipTable t;
  Affine f{69069u, 1u};
  for (int i = 0; i < 32; ++i) {
    t.p[i] = f;
    // f o f
    f = Affine{f.a * f.a, f.a * f.c + f.c};
  }
  return t;
}

__device__ __forceinline__ unsigned skip_dev(unsigned s, unsigned long long n, const SkipTable& t) {
  for (int i = 0; i < 32 && n; ++i, n >>= 1)
    if (n & 1) s = t.p[i].a * s + t.p[i].c;
  return s;
}

// One reference normal starting from state s (the state before its first
// draw).  Returns the value; cnt receives the number of draws, s the state
// after them.  Arithmetic is kept in the order of random.cpp:38-111 with
// contraction off so every double operation rounds as on the host.
__device__ double ref_normal(unsigned& s, int& cnt) {
#pragma clang fp contract(off)
  const double invmod = 1.0 / 4294967296.0;
  auto unif = [&]() { s = 69069u * s + 1u; ++cnt; return (double)s * invmod; };
  cnt = 0;
  double x;
  const double u = unif();
  if (u < 0.8638) {
    const double u1 = unif(), u2 = unif(), u3 = unif();
    x = 2.0 * (u1 + u2 + u3) - 3.0;
  } else if (u < 0.9745) {
    const double u1 = unif(), u2 = unif();
    x = 1.5 * (u1 + u2 - 1.0);
  } else if (u < 0.9973002039) {
    double u2;
    do {
      const double u1 = unif();
      u2 = unif();
      x = 6.0 * u1 - 3.0;
      // g(x) of random.cpp:96-111
      const double ax = fabs(x);
      double g = 17.49731196 * exp(-x * x / 2.0);
      if (ax < 1.0)
        g -= 2.0 * 2.36785163 * (3.0 - x * x) + 2.15787544 * (1.5 - ax);
      else if (ax < 1.5)
        g -= 2.36785163 * (3.0 - ax) * (3.0 - ax) + 2.15787544 * (1.5 - ax);
      else
        g -= 2.36785163 * (3.0 - ax) * (3.0 - ax);
      if (!(0.358 * u2 > g)) break;
    } while (true);
  } else {
    double sv, tv;
    do {
      double v1, v2, w;
      do {
        const double u1 = unif(), u2 = unif();
        v1 = 2.0 * u1 - 1.0;
        v2 = 2.0 * u2 - 1.0;
        w = v1 * v1 + v2 * v2;
      } while (w >= 1.0);
      const double c = sqrt((9.0 - 2.0 * log(w)) / w);
      sv = c * v1;
      tv = c * v2;
    } while (fabs(sv) <= 3.0 && fabs(tv) <= 3.0);
    x = (fabs(sv) > 3.0) ? sv : tv;
  }
  return x;
}

// Pass 1: walk each chunk from its first draw; mark starts; record exit.
__global__ void k_lcg_spec(unsigned s0, long long nchunks, SkipTable tab, unsigned* bitmap, int* exit0,
                           int* count0) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nchunks) return;
  unsigned words[kWords];
#pragma unroll
  for (int w = 0; w < kWords; ++w) words[w] = 0u;
  unsigned s = skip_dev(s0, (unsigned long long)b * kChunk, tab);
  int q = 0, n = 0;
  while (q < kChunk) {
    words[q >> 5] |= 1u << (q & 31);
    int c;
    (void)ref_normal(s, c);
    q += c;
    ++n;
  }
#pragma unroll
  for (int w = 0; w < kWords; ++w) bitmap[b * kWords + w] = words[w];
  exit0[b] = q - kChunk;
  count0[b] = n;
}

// Pass 2: walk from the true entry until the walk lands on a marked start.
__global__ void k_lcg_fix(unsigned s0, long long nchunks, SkipTable tab, const unsigned* bitmap,
                          const int* exit0, const int* count0, const int* exit_cur, int* exit_next,
                          int* count, unsigned* changed) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nchunks) return;
  const int entry = (b == 0) ? 0 : exit_cur[b - 1];
  int q = entry, n = 0, ex;
  if (q >= kChunk) {
    ex = q - kChunk;
  } else {
    unsigned s = skip_dev(s0, (unsigned long long)b * kChunk + q, tab);
    const unsigned* bm = bitmap + b * kWords;
    bool merged = false;
    while (q < kChunk) {
      if ((bm[q >> 5] >> (q & 31)) & 1u) { merged = true; break; }
      int c;
      (void)ref_normal(s, c);
      q += c;
      ++n;
    }
    if (merged) {
      // starts of the first walk at or after q
      int rest = __popc(bm[q >> 5] >> (q & 31));
      for (int w = (q >> 5) + 1; w < kWords; ++w) rest += __popc(bm[w]);
      n += rest;
      ex = exit0[b];
    } else {
      ex = q - kChunk;
    }
  }
  exit_next[b] = ex;
  count[b] = n;
  if (ex != exit_cur[b]) atomicAdd(changed, 1u);
}

// Exclusive scan of counts (single block), base[b] = sum_{c<b} count[c].
__global__ void k_scan(const int* count, long long n, long long* base, long long* total) {
  __shared__ long long part[1024];
  const int t = threadIdx.x;
  const long long per = (n + blockDim.x - 1) / blockDim.x;
  const long long lo = t * per, hi = min(n, lo + per);
  long long s = 0;
  for (long long i = lo; i < hi; ++i) s += count[i];
  part[t] = s;
  __syncthreads();
  if (t == 0) {
    long long run = 0;
    for (int i = 0; i < (int)blockDim.x; ++i) { long long v = part[i]; part[i] = run; run += v; }
    *total = run;
  }
  __syncthreads();
  long long run = part[t];
  for (long long i = lo; i < hi; ++i) { base[i] = run; run += count[i]; }
}

// Pass 3: emit normals along the true path.
__global__ void k_lcg_emit(unsigned s0, long long nchunks, SkipTable tab, const int* exit_fin,
                           const long long* base, long long nneed, double* out, unsigned* end_state) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nchunks) return;
  int q = (b == 0) ? 0 : exit_fin[b - 1];
  if (q >= kChunk) return;
  long long m = base[b];
  if (m >= nneed) return;
  unsigned s = skip_dev(s0, (unsigned long long)b * kChunk + q, tab);
  while (q < kChunk && m < nneed) {
    int c;
    const double v = ref_normal(s, c);
    out[m] = v;
    if (m == nneed - 1) *end_state = s;
    q += c;
    ++m;
  }
}

// Lay the decoded normals onto one grid (reference Hermitian rules).
template <typename T>
__global__ void k_fill(cplx_t<T>* g, NoiseGeom geo, const double* normals, long long first) {
  using C = cplx_t<T>;
  const long long n = (long long)geo.cnxp * geo.nyp * geo.nzp;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x) {
    double re, im;
    noise_ref_value(geo, normals + first, i, re, im);
    g[i] = cmake<C>((T)re, (T)im);
  }
}

}  // namespace

unsigned int crv_lcg_skip(unsigned int s, unsigned long long steps) {
  static const SkipTable t = make_skip();
  for (int i = 0; i < 32 && steps; ++i, steps >>= 1)
    if (steps & 1) s = t.p[i].a * s + t.p[i].c;
  return s;
}

// Draws per row of the half spectrum and their prefix (fftgrid.cpp:1197-1238).
int crv_noise_prefix(crv_ctx* ctx, FftPlan* plan) {
  if (plan->noise_prefix) return CRV_OK;
  const int nxp = plan->nxp, nyp = plan->nyp, nzp = plan->nzp, cnxp = nxp / 2 + 1;
  const long long nrows = (long long)nyp * nzp;
  std::vector<long long> pre(nrows + 1);
  long long run = 0;
  for (long long jk = 0; jk < nrows; ++jk) {
    pre[jk] = run;
    const int j = (int)(jk % nyp), k = (int)(jk / nyp);
    const int jcc = (j == 0) ? 0 : nyp - j, kcc = (k == 0) ? 0 : nzp - k;
    const long long jkcc = jcc + (long long)kcc * nyp;
    auto special_draws = [&]() -> int { return jkcc == jk ? 1 : (jkcc > jk ? 2 : 0); };
    for (int x = 0; x < cnxp; ++x) {
      const long long i = jk * cnxp + x;
      const bool special = (x == 0) || (x == cnxp - 1 && (i % 2) == 1);
      run += special ? special_draws() : 2;
    }
  }
  pre[nrows] = run;
  CRV_HIP_CHECK(ctx, hipMalloc(&plan->noise_prefix, (nrows + 1) * sizeof(long long)));
  CRV_HIP_CHECK(ctx, hipMemcpy(plan->noise_prefix, pre.data(), (nrows + 1) * sizeof(long long), hipMemcpyHostToDevice));
  plan->noise_draws = run;
  return CRV_OK;
}

int crv_reference_normals(crv_ctx* ctx, unsigned int* state, long long nneed) {
  if (nneed <= 0) return CRV_OK;
  const SkipTable tab = make_skip();
  // draws available: ~3.87 per normal on average; take 4.0 plus margin
  long long nchunks = (long long)((4.0 * (double)nneed) / kChunk) + 64;
  for (int attempt = 0; attempt < 4; ++attempt, nchunks *= 2) {
    const size_t bmb = (size_t)nchunks * kWords * 4;
    const size_t ib = (size_t)nchunks * 4;
    const size_t need = bmb + 5 * ib + (size_t)nchunks * 8 + 64;
    if (ctx->lcg_scratch_bytes < need) {
      if (ctx->lcg_scratch) CRV_HIP_CHECK(ctx, hipFree(ctx->lcg_scratch));
      ctx->lcg_scratch = nullptr; ctx->lcg_scratch_bytes = 0;
      CRV_HIP_CHECK(ctx, hipMalloc(&ctx->lcg_scratch, need));
      ctx->lcg_scratch_bytes = need;
    }
    if (ctx->normals_cap < (size_t)nneed) {
      if (ctx->normals) CRV_HIP_CHECK(ctx, hipFree(ctx->normals));
      ctx->normals = nullptr; ctx->normals_cap = 0;
      CRV_HIP_CHECK(ctx, hipMalloc(&ctx->normals, (size_t)nneed * sizeof(double)));
      ctx->normals_cap = (size_t)nneed;
    }
    unsigned char* p = (unsigned char*)ctx->lcg_scratch;
    unsigned* bitmap = (unsigned*)p; p += bmb;
    int* exit0 = (int*)p; p += ib;
    int* count0 = (int*)p; p += ib;
    int* exa = (int*)p; p += ib;
    int* exb = (int*)p; p += ib;
    int* count = (int*)p; p += ib;
    long long* base = (long long*)p; p += (size_t)nchunks * 8;
    unsigned* misc = (unsigned*)p;  // [0] changed, [1] end state, [2..3] total (as long long at +8)
    long long* total = (long long*)(p + 8);
    const int nb = (int)((nchunks + kThreads - 1) / kThreads);
    hipLaunchKernelGGL(k_lcg_spec, dim3(nb), dim3(kThreads), 0, ctx->stream, *state, nchunks, tab, bitmap, exit0, count0);
    CRV_HIP_CHECK(ctx, hipGetLastError());
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(exa, exit0, ib, hipMemcpyDeviceToDevice, ctx->stream));
    int* cur = exa; int* nxt = exb;
    for (long long it = 0;; ++it) {
      CRV_HIP_CHECK(ctx, hipMemsetAsync(misc, 0, 4, ctx->stream));
      hipLaunchKernelGGL(k_lcg_fix, dim3(nb), dim3(kThreads), 0, ctx->stream, *state, nchunks, tab, bitmap, exit0,
                         count0, cur, nxt, count, misc);
      CRV_HIP_CHECK(ctx, hipGetLastError());
      unsigned changed = 0;
      CRV_HIP_CHECK(ctx, hipMemcpyAsync(&changed, misc, 4, hipMemcpyDeviceToHost, ctx->stream));
      CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      std::swap(cur, nxt);
      if (changed == 0) break;
      if (it > nchunks + 2) return crv_set_error(ctx, CRV_E_INVALID, "reference noise decode did not settle");
    }
    hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, ctx->stream, count, nchunks, base, total);
    CRV_HIP_CHECK(ctx, hipGetLastError());
    long long have = 0;
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(&have, total, 8, hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (have < nneed) continue;  // not enough draws decoded: widen and redo
    hipLaunchKernelGGL(k_lcg_emit, dim3(nb), dim3(kThreads), 0, ctx->stream, *state, nchunks, tab, cur, base, nneed,
                       ctx->normals, misc + 1);
    CRV_HIP_CHECK(ctx, hipGetLastError());
    unsigned endst = 0;
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(&endst, misc + 1, 4, hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *state = endst;
    return CRV_OK;
  }
  return crv_set_error(ctx, CRV_E_INVALID, "reference noise decode: stream too short");
}

int crv_launch_fill_noise(crv_ctx* ctx, FftPlan* plan, crv_grid* g, long long first) {
  NoiseGeom geo = make_noise_geom(plan);
  const long long n = (long long)geo.cnxp * geo.nyp * geo.nzp;
  const int nb = (int)std::min<long long>((n + kThreads - 1) / kThreads, 65536);
  ProfScope ps(ctx, "noise_fill");
  if (g->precision == CRV_FP64)
    hipLaunchKernelGGL(k_fill<double>, dim3(nb), dim3(kThreads), 0, ctx->stream, (double2*)g->d, geo, ctx->normals, first);
  else
    hipLaunchKernelGGL(k_fill<float>, dim3(nb), dim3(kThreads), 0, ctx->stream, (float2*)g->d, geo, ctx->normals, first);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

extern "C" long long crv_noise_draws(int nxp, int nyp, int nzp) {
  if (nxp < 1 || nyp < 1 || nzp < 1) return -1;
  const int cnxp = nxp / 2 + 1;
  long long run = 0;
  const long long nrows = (long long)nyp * nzp;
  for (long long jk = 0; jk < nrows; ++jk) {
    const int j = (int)(jk % nyp), k = (int)(jk / nyp);
    const int jcc = (j == 0) ? 0 : nyp - j, kcc = (k == 0) ? 0 : nzp - k;
    const long long jkcc = jcc + (long long)kcc * nyp;
    const int sd = jkcc == jk ? 1 : (jkcc > jk ? 2 : 0);
    for (int x = 0; x < cnxp; ++x) {
      const long long i = jk * cnxp + x;
      const bool special = (x == 0) || (x == cnxp - 1 && (i % 2) == 1);
      run += special ? sd : 2;
    }
  }
  return run;
}
