// Device-side noise values on the half spectrum (shared by the fill and the
// simulation kernels).  The Hermitian structure is that of
// FFTGrid::fillInComplexNoise (reference src/fftgrid.cpp:1191-1240):
//  * element i = x + cnxp*(j + nyp*k) is "special" when x == 0, or when
//    x == cnxp-1 and i is odd (the reference tests the parity of the global
//    index, not of nxp);
//  * a special element whose conjugate partner (nyp-j, nzp-k) is itself gets
//    a real N(0,1); one whose partner comes later gets (N, N)/sqrt(2); one
//    whose partner came earlier copies the partner's conjugate;
//  * every other element gets (N, N)/sqrt(2).
#pragma once
#include "internal.hpp"

struct NoiseGeom {
  int cnxp, nyp, nzp;
  const long long* prefix;  // draws before each row (reference stream only)
  FastDiv dc, dy;           // by cnxp, nyp
  int small;                // cnxp nyp nzp < 2^31: 32-bit index arithmetic
};

// the geometry of a grid of half-spectrum size cnxp x nyp x nzp without a
// reference-stream prefix (counter noise only)
inline NoiseGeom make_noise_geom_dims(int cnxp, int nyp, int nzp) {
  NoiseGeom g;
  g.cnxp = cnxp; g.nyp = nyp; g.nzp = nzp; g.prefix = nullptr;
  g.dc = make_fastdiv(cnxp); g.dy = make_fastdiv(nyp);
  g.small = ((long long)cnxp * nyp * nzp) < (1LL << 31);
  return g;
}

inline NoiseGeom make_noise_geom(const FftPlan* p) {
  NoiseGeom g;
  g.cnxp = p->nxp / 2 + 1; g.nyp = p->nyp; g.nzp = p->nzp; g.prefix = p->noise_prefix;
  g.dc = make_fastdiv(g.cnxp); g.dy = make_fastdiv(g.nyp);
  g.small = ((long long)g.cnxp * g.nyp * g.nzp) < (1LL << 31);
  return g;
}

// classification of element i: 0 ordinary, 1 self-conjugate, 2 owns draws
// (partner later), 3 copies the partner (returned in partner, its row in
// pjk); jk and x: the row and column of element i
__device__ __forceinline__ int noise_kind(const NoiseGeom& g, long long i, long long& partner, long long& jk, int& x,
                                          long long& pjk) {
  if (g.small) {
    const int q = fdiv((int)i, g.dc);
    x = (int)i - q * g.cnxp;
    jk = q;
  } else {
    x = (int)(i % g.cnxp);
    jk = i / g.cnxp;
  }
  const bool special = (x == 0) || (x == g.cnxp - 1 && (i & 1));
  if (!special) return 0;
  int j, k;
  if (g.small) {
    k = fdiv((int)jk, g.dy);
    j = (int)jk - k * g.nyp;
  } else {
    j = (int)(jk % g.nyp);
    k = (int)(jk / g.nyp);
  }
  const int jcc = (j == 0) ? 0 : g.nyp - j, kcc = (k == 0) ? 0 : g.nzp - k;
  const long long jkcc = jcc + (long long)kcc * g.nyp;
  if (jkcc == jk) return 1;
  if (jkcc > jk) return 2;
  pjk = jkcc;
  partner = jkcc * g.cnxp + x;
  return 3;
}

// float(1/sqrt(2.0)) of fftgrid.cpp:1195, applied as float*double -> float
__device__ __forceinline__ double noise_std() { return (double)(float)0.70710678118654752440; }

// draw offset of element (jk, x) inside the reference stream of one grid
__device__ __forceinline__ long long noise_ref_offset(const NoiseGeom& g, long long jk, int x) {
  const long long r = g.prefix[jk];
  if (x == 0) return r;
  // draws of element 0 of this row
  int j, k;
  if (g.small) {
    k = fdiv((int)jk, g.dy);
    j = (int)jk - k * g.nyp;
  } else {
    j = (int)(jk % g.nyp);
    k = (int)(jk / g.nyp);
  }
  const int jcc = (j == 0) ? 0 : g.nyp - j, kcc = (k == 0) ? 0 : g.nzp - k;
  const long long jkcc = jcc + (long long)kcc * g.nyp;
  const int e0 = jkcc == jk ? 1 : (jkcc > jk ? 2 : 0);
  return r + e0 + 2LL * (x - 1);
}

// reference-stream value (re, im) of element i; normals indexed from the
// grid's first draw
__device__ __forceinline__ void noise_ref_value(const NoiseGeom& g, const double* normals, long long i, double& re,
                                                double& im) {
  long long partner = 0, jk, pjk = 0;
  int x;
  const int kind = noise_kind(g, i, partner, jk, x, pjk);
  const double sd = noise_std();
  if (kind == 1) {
    re = (double)(float)normals[noise_ref_offset(g, jk, x)];
    im = 0.0;
  } else {
    const long long o = noise_ref_offset(g, kind == 3 ? pjk : jk, x);
    re = (double)(float)(sd * normals[o]);
    im = (double)(float)(sd * normals[o + 1]);
    if (kind == 3) im = -im;
  }
}

// reference-stream noise of the three grids of realization r at element i:
// grid q of realization r starts at draw (3 r + q) draws of the decoded
// normals (Crava::simulate fills seed0, seed1, seed2 of each realization in
// turn from one generator, crava.cpp:1598-1603); values rounded to float as
// the reference's grids hold them
template <typename U>
__device__ __forceinline__ void ref_noise3(const NoiseGeom& g, const double* normals, long long draws, int r,
                                           long long i, U re[3], U im[3]) {
  long long partner = 0, jk, pjk = 0;
  int x;
  const int kind = noise_kind(g, i, partner, jk, x, pjk);
  const long long o = noise_ref_offset(g, kind == 3 ? pjk : jk, x);
  const double sd = noise_std();
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const double* nq = normals + (long long)(3 * r + q) * draws + o;
    if (kind == 1) {
      re[q] = (U)(float)nq[0];
      im[q] = (U)0;
    } else {
      re[q] = (U)(float)(sd * nq[0]);
      const float v = (float)(sd * nq[1]);
      im[q] = (U)(kind == 3 ? -v : v);
    }
  }
}

// ---- counter-based stream ---------------------------------------------------
// Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
// easy as 1, 2, 3", SC 2011): multipliers and Weyl key increments as
// published there.
__host__ __device__ inline void philox10(unsigned c[4], unsigned k0, unsigned k1) {
  for (int r = 0; r < 10; ++r) {
    const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
    const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
    const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
    const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
    const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
    k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
  }
}

// two N(0,1) for (seed, realization, grid, element): Box-Muller on the first
// two Philox words, uniforms (w + 0.5) 2^-32 in (0, 1)
__device__ __forceinline__ void counter_pair(unsigned long long seed, unsigned long long real, int grid, long long i,
                                             double& n0, double& n1) {
  unsigned c[4] = {(unsigned)i, (unsigned)((unsigned long long)i >> 32), (unsigned)real,
                   ((unsigned)(real >> 32) & 0x00FFFFFFu) | ((unsigned)grid << 24)};
  philox10(c, (unsigned)seed, (unsigned)(seed >> 32));
  const double u1 = ((double)c[0] + 0.5) * (1.0 / 4294967296.0);
  const double u2 = ((double)c[1] + 0.5) * (1.0 / 4294967296.0);
  const double r = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  n0 = r * cs;
  n1 = r * sn;
}

// Simulation noise of the three grids at one element.  One Philox call per
// (element, realization) gives four 32-bit words, cut into six 21-bit
// fields: the top 21 bits of each word, f0..f3, and from their low 11 bits
// f4 = low0 * 2^10 + low1 / 2 and f5 = low2 * 2^10 + low3 / 2.  Each field
// f gives the uniform (2 f + 1) 2^-22 in (0, 1), exact in FP32; grid g takes
// the Box-Muller pair of fields (2g, 2g + 1): r cos(2 pi u2), r sin(2 pi u2)
// with r = sqrt(-2 ln u1), evaluated in the precision U of the grids.  The
// radius field limits |N| to sqrt(2 ln 2^22) = 5.5, a tail of probability
// 4e-8.  An element that copies its Hermitian partner draws with the
// partner's index.
__device__ __forceinline__ void sim_words(unsigned long long seed, unsigned long long real, long long s,
                                          unsigned w[4]) {
  w[0] = (unsigned)s;
  w[1] = (unsigned)((unsigned long long)s >> 32);
  w[2] = (unsigned)real;
  w[3] = (unsigned)(real >> 32) & 0x00FFFFFFu;
  philox10(w, (unsigned)seed, (unsigned)(seed >> 32));
}

__device__ __forceinline__ void sim_fields(const unsigned w[4], unsigned f[6]) {
#pragma unroll
  for (int q = 0; q < 4; ++q) f[q] = w[q] >> 11;
  f[4] = ((w[0] & 0x7FFu) << 10) | ((w[1] & 0x7FFu) >> 1);
  f[5] = ((w[2] & 0x7FFu) << 10) | ((w[3] & 0x7FFu) >> 1);
}

// sin(2 pi u), cos(2 pi u) for u = (2m + 1) 2^-22 in (0, 1).  With
// q = rint(4u), v = u - q/4 is exact and within [-1/8, 1/8], so
// t = 2 pi v is within pi/4, where the Taylor series to t^9 (sin) and t^10
// (cos) stay within 2e-9 of the functions; q mod 4 rotates the pair by
// quarter turns.
__device__ __forceinline__ void sincos_2pi_unit(float u, float& sn, float& cs) {
  const float q = rintf(4.0f * u);
  const float t = 6.28318530717958647692f * fmaf(-0.25f, q, u);
  const float t2 = t * t;
  float ps = fmaf(t2, 2.7557319e-6f, -1.9841270e-4f);  // 1/9!, -1/7!
  ps = fmaf(t2, ps, 8.3333333e-3f);                    // 1/5!
  ps = fmaf(t2, ps, -1.6666667e-1f);                   // -1/3!
  const float s = fmaf(t * t2, ps, t);
  float pc = fmaf(t2, -2.7557319e-7f, 2.4801587e-5f);  // -1/10!, 1/8!
  pc = fmaf(t2, pc, -1.3888889e-3f);                   // -1/6!
  pc = fmaf(t2, pc, 4.1666667e-2f);                    // 1/4!
  pc = fmaf(t2, pc, -0.5f);
  const float c = fmaf(t2, pc, 1.0f);
  // quarter turn k = q mod 4: (sin, cos) = (s, c), (c, -s), (-s, -c), (-c, s),
  // by a swap on odd k and sign flips (k & 2 for sin, (k + 1) & 2 for cos)
  // rather than branches on k, which differs between the lanes of a wave
  const unsigned k = (unsigned)(int)q;
  const float a = (k & 1u) ? c : s, b = (k & 1u) ? s : c;
  sn = __uint_as_float(__float_as_uint(a) ^ ((k & 2u) << 30));
  cs = __uint_as_float(__float_as_uint(b) ^ (((k + 1u) & 2u) << 30));
}

// Box-Muller on two 21-bit fields
__device__ __forceinline__ void sim_box_muller(unsigned a, unsigned b, float& n0, float& n1) {
  const float u1 = (float)((a << 1) | 1u) * 0x1p-22f;
  const float u2 = (float)((b << 1) | 1u) * 0x1p-22f;
  // -2 ln u1 >= 4.7e-7: a normal number, where the hardware square root is
  // within 1 ulp
  const float r = __builtin_amdgcn_sqrtf(-2.0f * logf(u1));
  float sn, cs;
  sincos_2pi_unit(u2, sn, cs);
  n0 = r * cs;
  n1 = r * sn;
}
__device__ __forceinline__ void sim_box_muller(unsigned a, unsigned b, double& n0, double& n1) {
  const double u1 = (double)((a << 1) | 1u) * 0x1p-22;
  const double u2 = (double)((b << 1) | 1u) * 0x1p-22;
  const double r = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincospi(2.0 * u2, &sn, &cs);
  n0 = r * cs;
  n1 = r * sn;
}

// (re, im) of grids 0..2 at element i, Hermitian rules of fillInComplexNoise,
// values rounded to float as the reference's float grids hold them
template <typename U>
__device__ __forceinline__ void sim_noise3(const NoiseGeom& g, unsigned long long seed, unsigned long long real,
                                           long long i, U re[3], U im[3]) {
  long long partner = 0, jk, pjk = 0;
  int x;
  const int kind = noise_kind(g, i, partner, jk, x, pjk);
  const long long s = (kind == 3) ? partner : i;
  unsigned w[4], f[6];
  sim_words(seed, real, s, w);
  sim_fields(w, f);
  U n[3][2];
#pragma unroll
  for (int g = 0; g < 3; ++g) sim_box_muller(f[2 * g], f[2 * g + 1], n[g][0], n[g][1]);
  const U sd = (U)noise_std();
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    if (kind == 1) {
      re[q] = (U)(float)n[q][0];
      im[q] = (U)0;
    } else {
      re[q] = (U)(float)(sd * n[q][0]);
      im[q] = (U)(float)(sd * n[q][1]);
      if (kind == 3) im[q] = -im[q];
    }
  }
}
