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
#include <cmath>
#include <cstring>
#include <vector>

namespace {

constexpr int kThreads = 256;
constexpr int kChunk = 1024;           // draws per chunk
constexpr int kWords = kChunk / 32;    // bitmap words per chunk
constexpr int kHostCap = 16384;        // normals resolved on the host per run
constexpr int kScanTile = 4096;        // counts per block of the scan
constexpr int kScanTopMax = 1024 * 16; // tiles the top-level scan handles
constexpr int kFixBatch = 3;           // fix passes between convergence checks

// Guard margins.  The device and the host evaluate the reference's normal
// with the same IEEE double operations in the same order (contraction off),
// so they agree bit for bit except through exp() (the g(x) test of the
// third branch) and log() (the polar branch), whose device and host results
// may differ in their last bits.  Bounds, for libraries within 2 ulp of
// each other: g(x) = 17.49731196 exp(-x^2/2) - t(x) with |t| <= 21 differs by
// at most 17.5 * 2 * 2^-52 + 2 ulp(21) < 2e-14; the polar value c v1,
// c = sqrt((9 - 2 ln w) / w), by at most 5 * 2^-52 relative.  A decision
// whose operands are closer than kGMargin (absolute, third branch) or
// kRelMargin (relative, |s| <= 3 and |t| <= 3 of the polar branch) and a
// polar value within kRelMargin of a float rounding boundary (of the value
// or of its product with float(1/sqrt 2), the two roundings the grids use)
// are evaluated on the host instead, with the host's libm: the decode then
// equals the host's sequential stream exactly, at the rate of about one
// host evaluation per 1e6 polar normals.
constexpr double kGMargin = 1e-12;
constexpr double kRelMargin = 1e-14;

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

// First-uniform thresholds: u = s 2^-32 is exact, so u < c exactly when
// s < ceil(c 2^32) (c the double nearest the reference's literal).
constexpr unsigned kT1 = 3709992751u;  // 0.8638
constexpr unsigned kT2 = 4185445630u;  // 0.9745
constexpr unsigned kT3 = 4283371761u;  // 0.9973002039
// The third branch's test first in FP32: |g32(x) - g(x)| stays below 3.1e-5
// (x rounded to float: 1.9e-6 through the exponential term and 2.1e-6
// through the polynomial; the exponent's two roundings 9.4e-6; __expf 7e-6;
// the constants and products 2e-6; the polynomial's roundings 7.5e-6; the
// difference 1.3e-6), so a decision with |0.358 u2 - g32| above kG32Margin
// is the exact one, whatever libm evaluates g; the rest are decided in FP64.
constexpr double kG32Margin = 2e-4;

__device__ __forceinline__ float g_fast(float xf, double ax) {
  const float axf = fabsf(xf);
  const float e = __expf(-0.5f * xf * xf);
  float t;
  if (ax < 1.0)
    t = 2.0f * 2.36785163f * (3.0f - xf * xf) + 2.15787544f * (1.5f - axf);
  else if (ax < 1.5)
    t = 2.36785163f * (3.0f - axf) * (3.0f - axf) + 2.15787544f * (1.5f - axf);
  else
    t = 2.36785163f * (3.0f - axf) * (3.0f - axf);
  return 17.49731196f * e - t;
}

__host__ __device__ inline unsigned lcg_next(unsigned s) { return 69069u * s + 1u; }

// One reference normal from state s (the state before its first draw):
// the method of random.cpp:38-111 (Marsaglia-Bray, the branches, constants
// and g(x) of the reference).  The uniforms are s 2^-32, so the branch
// tests are integer compares and the sums 2 (u1 + u2 + u3) - 3,
// 1.5 (u1 + u2 - 1), 6 u1 - 3, 2 u1 - 1 are exact in double whichever way
// they are formed: they are formed from the integers; every rounded
// operation is done as the reference does it, in its order, contraction
// off.  Every branch begins with two draws after u, made before branching.
// cnt receives the number of draws and s the state after them; near
// (margins scaled by gs) is set when a decision or the value lies within
// the guard margins above.  FAST (device): the third branch's test first in
// FP32.
template <bool FAST>
__host__ __device__ inline double ref_normal_core(unsigned& s, int& cnt, bool& near, double gs) {
#pragma clang fp contract(off)
  const double invmod = 1.0 / 4294967296.0;
  const double gm = kGMargin * gs, rm = kRelMargin * gs;
  near = false;
  const unsigned su = lcg_next(s);
  unsigned p = lcg_next(su);
  unsigned q = lcg_next(p);
  cnt = 3;
  double x;
  if (su < kT1) {
    const unsigned r = lcg_next(q);
    cnt = 4;
    s = r;
    const unsigned long long S = (unsigned long long)p + q + r;
    x = (double)(2 * (int)(S >> 32) - 3) + (double)(unsigned)S * 0x1p-31;
  } else if (su < kT2) {
    s = q;
    const long long S = (long long)p + (long long)q - 4294967296LL;
    x = (double)(3 * S) * 0x1p-33;
  } else if (su < kT3) {
    for (;;) {
      x = (double)(6LL * (long long)p - 3LL * 4294967296LL) * 0x1p-32;
      const double ax = fabs(x);
      const double lhs = 0.358 * ((double)q * invmod);
      bool reject = false, decided = false;
      if constexpr (FAST) {
        const double d = lhs - (double)g_fast((float)x, ax);
        if (fabs(d) > kG32Margin) {
          reject = d > 0.0;
          decided = true;
        }
      }
      if (!decided) {
        // g(x) of random.cpp:96-111
        double g = 17.49731196 * exp(-x * x / 2.0);
        if (ax < 1.0)
          g -= 2.0 * 2.36785163 * (3.0 - x * x) + 2.15787544 * (1.5 - ax);
        else if (ax < 1.5)
          g -= 2.36785163 * (3.0 - ax) * (3.0 - ax) + 2.15787544 * (1.5 - ax);
        else
          g -= 2.36785163 * (3.0 - ax) * (3.0 - ax);
        if (fabs(lhs - g) <= gm) near = true;
        reject = lhs > g;
      }
      if (!reject) break;
      p = lcg_next(q);
      q = lcg_next(p);
      cnt += 2;
    }
    s = q;
  } else {
    double sv, tv;
    for (;;) {
      double v1, v2, w;
      for (;;) {
        v1 = (double)(2LL * (long long)p - 4294967296LL) * 0x1p-32;
        v2 = (double)(2LL * (long long)q - 4294967296LL) * 0x1p-32;
        w = v1 * v1 + v2 * v2;
        if (!(w >= 1.0)) break;
        p = lcg_next(q);
        q = lcg_next(p);
        cnt += 2;
      }
      const double c = sqrt((9.0 - 2.0 * log(w)) / w);
      sv = c * v1;
      tv = c * v2;
      if (fabs(fabs(sv) - 3.0) <= 3.0 * rm || fabs(fabs(tv) - 3.0) <= 3.0 * rm) near = true;
      if (!(fabs(sv) <= 3.0 && fabs(tv) <= 3.0)) break;
      p = lcg_next(q);
      q = lcg_next(p);
      cnt += 2;
    }
    s = q;
    x = (fabs(sv) > 3.0) ? sv : tv;
    // |x| > 3: the value and its float(1/sqrt 2) multiple against the
    // midpoints between adjacent floats
    const double sd = (double)(float)0.70710678118654752440;
    const double vs[2] = {x, sd * x};
    for (int e = 0; e < 2; ++e) {
      const double v = vs[e];
      const float f = (float)v;
      if ((double)f == v) continue;
      const unsigned fb = __builtin_bit_cast(unsigned, f);
      const float fn = __builtin_bit_cast(float, fabs(v) > fabs((double)f) ? fb + 1u : fb - 1u);
      const double mid = 0.5 * ((double)f + (double)fn);
      if (fabs(v - mid) <= rm * fabs(v)) near = true;
    }
  }
  return x;
}

// A normal resolved on the host: its starting state, draws and value
// (sorted by state for the device's binary search).
struct Resolved { unsigned s; int cnt; double v; };

// A normal near a margin without a table entry: its starting state, the
// device's draws for it and where its value went: m >= 0 the output index
// (emit pass), m <= -2 the staging index -2 - m (spec pass), -1 nowhere
// (fix pass).
struct Listed { unsigned s; int cnt; long long m; };

struct Guard {
  const Resolved* tab;
  int ntab;
  double scale;
  Listed* list;
  unsigned* nlist;
};

__device__ __forceinline__ bool guard_lookup(const Guard& G, unsigned s0, int& cnt, double& v) {
  int lo = 0, hi = G.ntab;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (G.tab[mid].s < s0) lo = mid + 1; else hi = mid;
  }
  if (lo < G.ntab && G.tab[lo].s == s0) {
    cnt = G.tab[lo].cnt;
    v = G.tab[lo].v;
    return true;
  }
  return false;
}

// The normal from state s, taking the host's value and draws for states in
// the table; an unresolved normal near a margin is listed for the host.
__device__ __forceinline__ double ref_normal(unsigned& s, int& cnt, const Guard& G, long long m) {
  const unsigned s0 = s;
  bool near;
  double x = ref_normal_core<true>(s, cnt, near, G.scale);
  if (near) {
    double v;
    int c;
    if (guard_lookup(G, s0, c, v)) {
      s = s0;
      for (int k = 0; k < c; ++k) s = lcg_next(s);
      cnt = c;
      x = v;
    } else {
      const unsigned slot = atomicAdd(G.nlist, 1u);
      if (slot < (unsigned)kHostCap) G.list[slot] = Listed{s0, cnt, m};
    }
  }
  return x;
}

// normals that start in one chunk of 1024 draws: at most 342 (each takes 3
// or more); staged in 44 tiles of 8 per chunk
constexpr int kCap = 352;

// staged normal k of chunk b: each chunk's normals in order (kCap slots)
__device__ __forceinline__ long long stage_at(int k, long long b, long long nchunks) {
  return b * kCap + k;
}

// Pass 1: walk each chunk from its first draw; mark the starts, stage the
// normals (through 8 LDS slots per lane, written 64 bytes at a time),
// record exit and count.
__global__ __launch_bounds__(kThreads) void k_lcg_spec(unsigned s0, long long nchunks, SkipTable tab, Guard G,
                                                       unsigned* bitmap, int* exit0, int* count0, double* stage) {
  __shared__ double st[8 * kThreads];
  const int t = threadIdx.x;
  const long long b = (long long)blockIdx.x * blockDim.x + t;
  if (b >= nchunks) return;
  unsigned* bm = bitmap + b * kWords;
  unsigned s = skip_dev(s0, (unsigned long long)b * kChunk, tab);
  int q = 0, n = 0, cw = 0;
  unsigned word = 0u;
  auto tile_out = [&](int k0) {
    double2* dst = (double2*)(stage + stage_at(k0, b, nchunks));
#pragma unroll
    for (int j = 0; j < 4; ++j) dst[j] = make_double2(st[(2 * j) * kThreads + t], st[(2 * j + 1) * kThreads + t]);
  };
  while (q < kChunk) {
    while ((q >> 5) != cw) {
      bm[cw++] = word;
      word = 0u;
    }
    word |= 1u << (q & 31);
    int c;
    st[(n & 7) * kThreads + t] = ref_normal(s, c, G, -2 - stage_at(n, b, nchunks));
    q += c;
    ++n;
    if ((n & 7) == 0) tile_out(n - 8);
  }
  if (n & 7) tile_out(n & ~7);
  for (; cw < kWords; ++cw) {
    bm[cw] = word;
    word = 0u;
  }
  exit0[b] = q - kChunk;
  count0[b] = n;
}

// Pass 2: walk from the true entry until the walk lands on a marked start
// (recorded in merge_q; -1 when it does not).  A pass whose predecessor
// changed nothing only carries the exits over.
__global__ __launch_bounds__(kThreads) void k_lcg_fix(unsigned s0, long long nchunks, SkipTable tab, Guard G,
                                                      const unsigned* bitmap, const int* exit0, const int* count0,
                                                      const int* exit_cur, int* exit_next, int* count, int* merge_q,
                                                      const unsigned* changed_prev, unsigned* changed) {
  const long long b = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nchunks) return;
  if (changed_prev && *changed_prev == 0u) {
    exit_next[b] = exit_cur[b];
    return;
  }
  const int entry = (b == 0) ? 0 : exit_cur[b - 1];
  int q = entry, n = 0, ex, mq = -1;
  if (q >= kChunk) {
    ex = q - kChunk;
  } else {
    unsigned s = skip_dev(s0, (unsigned long long)b * kChunk + q, tab);
    const unsigned* bm = bitmap + b * kWords;
    bool merged = false;
    while (q < kChunk) {
      if ((bm[q >> 5] >> (q & 31)) & 1u) { merged = true; break; }
      int c;
      (void)ref_normal(s, c, G, -1);
      q += c;
      ++n;
    }
    if (merged) {
      // starts of the first walk at or after q
      int rest = __popc(bm[q >> 5] >> (q & 31));
      for (int w = (q >> 5) + 1; w < kWords; ++w) rest += __popc(bm[w]);
      n += rest;
      ex = exit0[b];
      mq = q;
    } else {
      ex = q - kChunk;
    }
  }
  exit_next[b] = ex;
  count[b] = n;
  merge_q[b] = mq;
  if (ex != exit_cur[b]) atomicAdd(changed, 1u);
}

// exclusive scan of the 256 values v of a block (one per thread); the
// block's total in *tot
__device__ __forceinline__ long long block_scan_excl(long long v, long long* sh, long long* tot) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  long long x = v;
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const long long y = __shfl_up(x, d, 64);
    if (lane >= d) x += y;
  }
  if (lane == 63) sh[w] = x;
  __syncthreads();
  long long off = 0;
  for (int k = 0; k < w; ++k) off += sh[k];
  *tot = sh[0] + sh[1] + sh[2] + sh[3];
  __syncthreads();
  return off + x - v;
}

// Scan of the chunk counts, three levels: tile sums, the scan of the tile
// sums (one block), then each tile's scan from its offset.
__global__ __launch_bounds__(256) void k_scan_tiles(const int* count, long long n, long long* tile_sum) {
  __shared__ long long sh[4];
  const long long t0 = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * 16;
  long long s = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e)
    if (t0 + e < n) s += count[t0 + e];
  long long tot;
  (void)block_scan_excl(s, sh, &tot);
  if (threadIdx.x == 0) tile_sum[blockIdx.x] = tot;
}

__global__ __launch_bounds__(256) void k_scan_top(long long* tile_sum, int ntiles, long long* total) {
  __shared__ long long sh[4];
  const int per = (ntiles + 255) / 256;
  const int lo = threadIdx.x * per, hi = min(ntiles, lo + per);
  long long s = 0;
  for (int i = lo; i < hi; ++i) s += tile_sum[i];
  long long tot;
  long long run = block_scan_excl(s, sh, &tot);
  for (int i = lo; i < hi; ++i) {
    const long long v = tile_sum[i];
    tile_sum[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) *total = tot;
}

__global__ __launch_bounds__(256) void k_scan_apply(const int* count, long long n, const long long* tile_off,
                                                    long long* base) {
  __shared__ long long sh[4];
  const long long t0 = (long long)blockIdx.x * kScanTile + (long long)threadIdx.x * 16;
  int c[16];
  long long s = 0;
#pragma unroll
  for (int e = 0; e < 16; ++e) {
    c[e] = (t0 + e < n) ? count[t0 + e] : 0;
    s += c[e];
  }
  long long tot;
  long long run = tile_off[blockIdx.x] + block_scan_excl(s, sh, &tot);
#pragma unroll
  for (int e = 0; e < 16; ++e)
    if (t0 + e < n) {
      base[t0 + e] = run;
      run += c[e];
    }
}

// starts of the first walk before position q of a chunk
__device__ __forceinline__ int bitmap_rank(const unsigned* bm, int q) {
  int r = 0;
  for (int w = 0; w < (q >> 5); ++w) r += __popc(bm[w]);
  if (q & 31) r += __popc(bm[q >> 5] & ((1u << (q & 31)) - 1u));
  return r;
}

// position of the first walk's k-th start (0-based) in a chunk; kChunk +
// exit when there is none
__device__ __forceinline__ int bitmap_select(const unsigned* bm, int k, int exit) {
  for (int w = 0; w < kWords; ++w) {
    unsigned v = bm[w];
    const int c = __popc(v);
    if (k < c) {
      for (; k > 0; --k) v &= v - 1u;
      return (w << 5) + __builtin_ctz(v);
    }
    k -= c;
  }
  return kChunk + exit;
}

// Pass 3: the normals along the true path.  Each chunk walks from its true
// entry to the merge with the first walk (a few normals; to its end when
// they do not merge), then the wave copies every chunk's staged normals
// from the merge on.  The lane whose range holds the last normal needed
// records the state after it.
__global__ __launch_bounds__(64) void k_lcg_emit(unsigned s0, long long nchunks, SkipTable tab, Guard G,
                                                 const int* exit_fin, const int* merge_q, const unsigned* bitmap,
                                                 const int* exit0, const int* count0, const double* stage,
                                                 const long long* base, long long nneed, double* out,
                                                 unsigned* end_state) {
  const int lane = threadIdx.x;
  const long long b = (long long)blockIdx.x * 64 + lane;
  int r0 = 0, ncopy = 0;
  long long m = nneed;
  if (b < nchunks) {
    int q = (b == 0) ? 0 : exit_fin[b - 1];
    m = base[b];
    const int mq = merge_q[b];
    if (q < kChunk && m < nneed) {
      const int stop = mq >= 0 ? mq : kChunk;
      if (q < stop) {
        unsigned s = skip_dev(s0, (unsigned long long)b * kChunk + q, tab);
        while (q < stop && m < nneed) {
          int c;
          out[m] = ref_normal(s, c, G, m);
          if (m == nneed - 1) *end_state = s;
          q += c;
          ++m;
        }
      }
      if (mq >= 0 && m < nneed) {
        const unsigned* bm = bitmap + b * kWords;
        r0 = bitmap_rank(bm, mq);
        ncopy = (int)min((long long)(count0[b] - r0), nneed - m);
        if (m + ncopy == nneed && ncopy > 0) {
          // the last needed normal is the first walk's (r0 + ncopy - 1)-th
          const int pos = bitmap_select(bm, r0 + ncopy, exit0[b]);
          *end_state = skip_dev(s0, (unsigned long long)b * kChunk + pos, tab);
        }
      }
    }
  }
  // the copy: the wave's 64 runs as one sequence of values, 64 consecutive
  // ones per step (contiguous loads and stores), eight steps in flight
  __shared__ int pre[65];
  __shared__ int r0s[64];
  __shared__ long long dsts[64];
  int inc = ncopy;
  for (int d = 1; d < 64; d <<= 1) {
    const int y = __shfl_up(inc, d, 64);
    if (lane >= d) inc += y;
  }
  pre[lane + 1] = inc;
  if (lane == 0) pre[0] = 0;
  r0s[lane] = r0;
  dsts[lane] = m;
  __syncthreads();
  const int W = pre[64];
  const long long b0 = (long long)blockIdx.x * 64;
  // the lane's cursor: run L, its first sequence index pL, the next run's
  // pN, its staged and output positions
  int L = 0, pL = 0, pN = pre[1];
  long long sL = stage_at(r0s[0], b0, nchunks), dL = dsts[0];
  constexpr int kDepth = 8;
  for (int t0 = 0; t0 < W; t0 += kDepth * 64) {
    long long src[kDepth], dst[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) {
      const int t = t0 + u * 64 + lane;
      src[u] = -1;
      if (t < W) {
        while (t >= pN) {
          ++L;
          pL = pN;
          pN = pre[L + 1];
          sL = stage_at(r0s[L], b0 + L, nchunks);
          dL = dsts[L];
        }
        src[u] = sL + (t - pL);
        dst[u] = dL + (t - pL);
      }
    }
    double v[kDepth];
#pragma unroll
    for (int u = 0; u < kDepth; ++u) v[u] = src[u] >= 0 ? stage[src[u]] : 0.0;
#pragma unroll
    for (int u = 0; u < kDepth; ++u)
      if (src[u] >= 0) out[dst[u]] = v[u];
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

// host values written where the listed normals went (output or staging)
__global__ void k_patch(double* out, double* stage, const Listed* list, const double* v, int n) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const long long m = list[t].m;
  if (m >= 0) out[m] = v[t];
  else if (m <= -2) stage[-2 - m] = v[t];
}

// the host's normal from state s (its libm): value and draws
Resolved host_normal(unsigned s) {
  Resolved r;
  r.s = s;
  bool near;
  unsigned t = s;
  r.v = ref_normal_core<false>(t, r.cnt, near, 1.0);
  return r;
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

namespace {

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// device scratch of one decode of nchunks chunks
struct DecodeBufs {
  unsigned* bitmap;
  int *exit0, *count0, *exa, *exb, *count, *merge;
  long long *base, *tiles;
  unsigned* misc;      // [0, kFixBatch) changed, [4] end state, [5] listed
  long long* total;    // misc + 32 bytes
  Listed* list;
  Resolved* tab;
  double* vals;
  double* stage;
};

constexpr int kTabCap = 4 * kHostCap;

size_t decode_bytes(long long nchunks, DecodeBufs* d, void* p0) {
  const long long ntiles = (nchunks + kScanTile - 1) / kScanTile;
  size_t o = 0;
  auto take = [&](size_t b) { const size_t at = o; o += align256(b); return (char*)p0 + at; };
  DecodeBufs x;
  x.bitmap = (unsigned*)take((size_t)nchunks * kWords * 4);
  x.exit0 = (int*)take((size_t)nchunks * 4);
  x.count0 = (int*)take((size_t)nchunks * 4);
  x.exa = (int*)take((size_t)nchunks * 4);
  x.exb = (int*)take((size_t)nchunks * 4);
  x.count = (int*)take((size_t)nchunks * 4);
  x.merge = (int*)take((size_t)nchunks * 4);
  x.base = (long long*)take((size_t)nchunks * 8);
  x.tiles = (long long*)take((size_t)ntiles * 8);
  x.misc = (unsigned*)take(64);
  x.total = (long long*)((char*)x.misc + 32);
  x.list = (Listed*)take((size_t)kHostCap * sizeof(Listed));
  x.tab = (Resolved*)take((size_t)kTabCap * sizeof(Resolved));
  x.vals = (double*)take((size_t)kHostCap * 8);
  x.stage = (double*)take((size_t)nchunks * kCap * 8);
  if (d) *d = x;
  return o;
}

// What the host does with the normals listed near a margin: false when one
// needs a new run (a count that differs from the device's, a full list, the
// test hook), after adding them to the table; else their host values are
// written where the device put its own.
bool resolve_listed(crv_ctx* ctx, const DecodeBufs& d, unsigned nl, bool& force, std::vector<Resolved>& table,
                    long long info[4], int& rc) {
  rc = CRV_OK;
  const int nlist = (int)std::min<unsigned>(nl, (unsigned)kHostCap);
  if (nlist == 0) return true;
  std::vector<Listed> list(nlist);
  if (hipMemcpy(list.data(), d.list, nlist * sizeof(Listed), hipMemcpyDeviceToHost) != hipSuccess) {
    rc = crv_set_error(ctx, CRV_E_HIP, "reference noise decode: list download failed");
    return false;
  }
  std::vector<double> vals(nlist);
  bool rerun = nl > (unsigned)kHostCap || force;
  for (int e = 0; e < nlist; ++e) {
    const Resolved h = host_normal(list[e].s);
    vals[e] = h.v;
    if (h.cnt != list[e].cnt) rerun = true;
  }
  info[1] += nlist;
  if (rerun) {
    force = false;
    for (const Listed& e : list) table.push_back(host_normal(e.s));
    std::sort(table.begin(), table.end(), [](const Resolved& a, const Resolved& b) { return a.s < b.s; });
    table.erase(std::unique(table.begin(), table.end(), [](const Resolved& a, const Resolved& b) { return a.s == b.s; }),
                table.end());
    if (table.size() > (size_t)kTabCap)
      rc = crv_set_error(ctx, CRV_E_UNSUPPORTED, "reference noise decode: too many normals near a margin");
    return false;
  }
  if (hipMemcpyAsync(d.vals, vals.data(), nlist * 8, hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    rc = crv_set_error(ctx, CRV_E_HIP, "reference noise decode: upload failed");
    return false;
  }
  hipLaunchKernelGGL(k_patch, dim3((nlist + 255) / 256), dim3(256), 0, ctx->stream, ctx->normals, d.stage, d.list,
                     d.vals, nlist);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(ctx->stream) != hipSuccess) {
    rc = crv_set_error(ctx, CRV_E_HIP, "reference noise decode: patch failed");
    return false;
  }
  for (const Listed& e : list) info[2] += e.m != -1;
  return true;
}

}  // namespace

int crv_reference_normals(crv_ctx* ctx, unsigned int* state, long long nneed) {
  if (nneed <= 0) return CRV_OK;
  const SkipTable tab = make_skip();
  // draws available: about 3.87 per normal on average; 4.0 plus a margin
  // exceeds the need by far more than its spread
  long long nchunks = (long long)((4.0 * (double)nneed) / kChunk) + 64;
  // earlier work on the stream may still read the normals or the scratch
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (ctx->normals_cap < (size_t)nneed) {
    if (ctx->normals) CRV_HIP_CHECK(ctx, hipFree(ctx->normals));
    ctx->normals = nullptr; ctx->normals_cap = 0;
    CRV_HIP_CHECK(ctx, hipMalloc(&ctx->normals, (size_t)nneed * sizeof(double)));
    ctx->normals_cap = (size_t)nneed;
  }
  std::vector<Resolved> table;
  bool force = ctx->noise_force_table;
  ctx->noise_force_table = false;
  long long info[4] = {0, 0, 0, 0};
  for (int run = 0; run < 16; ++run) {
    if ((nchunks + kScanTile - 1) / kScanTile > kScanTopMax)
      return crv_set_error(ctx, CRV_E_UNSUPPORTED, "reference noise decode: stream too long");
    const size_t need = decode_bytes(nchunks, nullptr, nullptr);
    if (ctx->lcg_scratch_bytes < need) {
      if (ctx->lcg_scratch) CRV_HIP_CHECK(ctx, hipFree(ctx->lcg_scratch));
      ctx->lcg_scratch = nullptr; ctx->lcg_scratch_bytes = 0;
      CRV_HIP_CHECK(ctx, hipMalloc(&ctx->lcg_scratch, need));
      ctx->lcg_scratch_bytes = need;
    }
    DecodeBufs d;
    decode_bytes(nchunks, &d, ctx->lcg_scratch);
    info[3] = run + 1;
    Guard G;
    G.tab = d.tab; G.ntab = (int)table.size(); G.scale = ctx->noise_guard; G.list = d.list; G.nlist = d.misc + 5;
    if (!table.empty())
      CRV_HIP_CHECK(ctx, hipMemcpyAsync(d.tab, table.data(), table.size() * sizeof(Resolved), hipMemcpyHostToDevice,
                                        ctx->stream));
    CRV_HIP_CHECK(ctx, hipMemsetAsync(d.misc, 0, 64, ctx->stream));
    const int nbw = (int)((nchunks + 63) / 64);
    {
    ProfScope ps(ctx, "lcg_spec");
    hipLaunchKernelGGL(k_lcg_spec, dim3((nchunks + kThreads - 1) / kThreads), dim3(kThreads), 0, ctx->stream, *state, nchunks, tab, G, d.bitmap, d.exit0,
                       d.count0, d.stage);
    }
    CRV_HIP_CHECK(ctx, hipGetLastError());
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(d.exa, d.exit0, (size_t)nchunks * 4, hipMemcpyDeviceToDevice, ctx->stream));
    int* cur = d.exa;
    int* nxt = d.exb;
    const int nb = (int)((nchunks + kThreads - 1) / kThreads);
    unsigned misc[8];
    for (long long it = 0;; it += kFixBatch) {
      CRV_HIP_CHECK(ctx, hipMemsetAsync(d.misc, 0, 4 * kFixBatch, ctx->stream));
      for (int p = 0; p < kFixBatch; ++p) {
        ProfScope ps(ctx, "lcg_fix");
        hipLaunchKernelGGL(k_lcg_fix, dim3(nb), dim3(kThreads), 0, ctx->stream, *state, nchunks, tab, G, d.bitmap,
                           d.exit0, d.count0, cur, nxt, d.count, d.merge, p ? d.misc + p - 1 : nullptr, d.misc + p);
        CRV_HIP_CHECK(ctx, hipGetLastError());
        std::swap(cur, nxt);
      }
      CRV_HIP_CHECK(ctx, hipMemcpyAsync(misc, d.misc, 32, hipMemcpyDeviceToHost, ctx->stream));
      CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
      info[0] += kFixBatch;
      if (misc[kFixBatch - 1] == 0) break;
      if (it > nchunks + 2) return crv_set_error(ctx, CRV_E_INVALID, "reference noise decode did not settle");
    }
    // normals listed by the first two passes: staged values patched, or a
    // new run
    int rc = CRV_OK;
    if (!resolve_listed(ctx, d, misc[5], force, table, info, rc)) {
      if (rc != CRV_OK) return rc;
      continue;
    }
    CRV_HIP_CHECK(ctx, hipMemsetAsync(d.misc + 5, 0, 4, ctx->stream));
    const int ntiles = (int)((nchunks + kScanTile - 1) / kScanTile);
    {
    ProfScope ps(ctx, "lcg_scan");
    hipLaunchKernelGGL(k_scan_tiles, dim3(ntiles), dim3(256), 0, ctx->stream, d.count, nchunks, d.tiles);
    hipLaunchKernelGGL(k_scan_top, dim3(1), dim3(256), 0, ctx->stream, d.tiles, ntiles, d.total);
    hipLaunchKernelGGL(k_scan_apply, dim3(ntiles), dim3(256), 0, ctx->stream, d.count, nchunks, d.tiles, d.base);
    }
    CRV_HIP_CHECK(ctx, hipGetLastError());
    long long have = 0;
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(&have, d.total, 8, hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (have < nneed) {  // not enough draws decoded: widen and redo
      nchunks *= 2;
      continue;
    }
    {
    ProfScope ps(ctx, "lcg_emit");
    hipLaunchKernelGGL(k_lcg_emit, dim3(nbw), dim3(64), 0, ctx->stream, *state, nchunks, tab, G, cur, d.merge,
                       d.bitmap, d.exit0, d.count0, d.stage, d.base, nneed, ctx->normals, d.misc + 4);
    }
    CRV_HIP_CHECK(ctx, hipGetLastError());
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(misc, d.misc, 32, hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    if (!resolve_listed(ctx, d, misc[5], force, table, info, rc)) {
      if (rc != CRV_OK) return rc;
      continue;
    }
    *state = misc[4];
    memcpy(ctx->noise_info, info, sizeof(info));
    return CRV_OK;
  }
  return crv_set_error(ctx, CRV_E_INVALID, "reference noise decode did not complete");
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

extern "C" int crv_noise_guard(crv_ctx* ctx, double scale, int force_table, long long info[4]) {
  if (!ctx) return CRV_E_INVALID;
  if (!(scale >= 0.0) || scale > 1e12) return crv_set_error(ctx, CRV_E_INVALID, "noise guard: scale out of range");
  if (scale > 0.0) ctx->noise_guard = scale;
  ctx->noise_force_table = force_table != 0;
  if (info) memcpy(info, ctx->noise_info, sizeof(ctx->noise_info));
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
