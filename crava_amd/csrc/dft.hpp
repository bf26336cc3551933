// Small DFTs in registers for the mixed-radix transforms (gfx950).
//
// Dft<T, R, INV>::run(v) replaces v[0..R) by its DFT (sign -1 forward, +1
// inverse, no scaling), natural order in and out.  Radices 2, 3, 4, 8 are
// direct butterflies, 5 and 7 use the symmetric split of an odd prime
// length, and composite radices (6, 9, 10, 12, 15, 16, 18, 20, 25) are
// two-level transforms whose inner twiddles are compile-time constants.
#pragma once
#include "internal.hpp"

#include <type_traits>

namespace dft {

// ---- compile-time twiddles ---------------------------------------------------
constexpr double kPi = 3.141592653589793238462643383279502884;

__host__ __device__ constexpr double taylor_sin(double x) {
  double x2 = x * x, t = x, s = x;
  for (int i = 1; i <= 13; ++i) { t *= -x2 / ((2.0 * i) * (2.0 * i + 1.0)); s += t; }
  return s;
}
__host__ __device__ constexpr double taylor_cos(double x) {
  double x2 = x * x, t = 1.0, s = 1.0;
  for (int i = 1; i <= 13; ++i) { t *= -x2 / ((2.0 * i - 1.0) * (2.0 * i)); s += t; }
  return s;
}
struct CosSin { double c, s; };
// cos and sin of 2 pi m / n: the octant is found in integers, the series is
// evaluated on [0, pi/4]
__host__ __device__ constexpr CosSin cossin2pi(long long m, long long n) {
  m %= n;
  if (m < 0) m += n;
  const long long q8 = 8 * m, o = q8 / n, r = q8 - o * n;
  const double a = kPi / 4.0 * (double)r / (double)n;        // alpha
  const double b = kPi / 4.0 * (double)(n - r) / (double)n;  // pi/4 - alpha
  switch (o) {
    case 0: return {taylor_cos(a), taylor_sin(a)};
    case 1: return {taylor_sin(b), taylor_cos(b)};
    case 2: return {-taylor_sin(a), taylor_cos(a)};
    case 3: return {-taylor_cos(b), taylor_sin(b)};
    case 4: return {-taylor_cos(a), -taylor_sin(a)};
    case 5: return {-taylor_sin(b), -taylor_cos(b)};
    case 6: return {taylor_sin(a), -taylor_cos(a)};
    default: return {taylor_cos(b), -taylor_sin(b)};
  }
}

template <int B, int E, typename F>
__device__ __forceinline__ void sfor(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    sfor<B + 1, E>(f);
  }
}

// multiply by -i (forward) or +i (inverse)
template <typename C, bool INV> __device__ __forceinline__ C rot(C a) {
  return INV ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// a * exp(-+ 2 pi i e / N) with e, N known at compile time
template <typename T, int E, int N, bool INV>
__device__ __forceinline__ cplx_t<T> twiddle_ct(cplx_t<T> a) {
  using C = cplx_t<T>;
  constexpr int e = E % N;
  if constexpr (e == 0) {
    return a;
  } else if constexpr ((4 * e) % N == 0) {
    constexpr int q = 4 * e / N;  // quarter turns
    if constexpr (q == 1) return rot<C, INV>(a);
    else if constexpr (q == 2) return cmake<C>(-a.x, -a.y);
    else return rot<C, !INV>(a);
  } else {
    constexpr CosSin w = cossin2pi(e, N);
    constexpr T wr = (T)w.c, wi = (T)(INV ? w.s : -w.s);
    return cmake<C>(a.x * wr - a.y * wi, a.x * wi + a.y * wr);
  }
}

// ---- small DFTs (sign -1 forward, +1 inverse), in place ------------------
template <typename T, int R, bool INV> struct Dft;

template <typename T, bool INV> struct Dft<T, 1, INV> {
  __device__ __forceinline__ static void run(cplx_t<T>*) {}
};

template <typename T, bool INV> struct Dft<T, 2, INV> {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    C a = v[0];
    v[0] = cadd(a, v[1]);
    v[1] = csub(a, v[1]);
  }
};

template <typename T, bool INV> struct Dft<T, 4, INV> {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = rot<C, INV>(csub(v[1], v[3]));
    v[0] = cadd(t0, t2); v[2] = csub(t0, t2);
    v[1] = cadd(t1, t3); v[3] = csub(t1, t3);
  }
};

template <typename T, bool INV> struct Dft<T, 8, INV> {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    const T h = (T)0.70710678118654752440;  // sqrt(1/2)
    C e[4] = {v[0], v[2], v[4], v[6]};
    C o[4] = {v[1], v[3], v[5], v[7]};
    Dft<T, 4, INV>::run(e);
    Dft<T, 4, INV>::run(o);
    // o[m] *= exp(-+ 2 pi i m / 8)
    C o1, o3;
    if (!INV) {
      o1 = cmake<C>(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
      o3 = cmake<C>(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    } else {
      o1 = cmake<C>(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y));
      o3 = cmake<C>(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
    }
    C o2 = rot<C, INV>(o[2]);
    v[0] = cadd(e[0], o[0]); v[4] = csub(e[0], o[0]);
    v[1] = cadd(e[1], o1);   v[5] = csub(e[1], o1);
    v[2] = cadd(e[2], o2);   v[6] = csub(e[2], o2);
    v[3] = cadd(e[3], o3);   v[7] = csub(e[3], o3);
  }
};

template <typename T, bool INV> struct Dft<T, 3, INV> {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    const T s3 = (T)0.86602540378443864676;  // sin(2 pi / 3)
    C t = cadd(v[1], v[2]);
    C m = cmake<C>(v[0].x - (T)0.5 * t.x, v[0].y - (T)0.5 * t.y);
    C d = rot<C, INV>(csub(v[1], v[2]));
    v[0] = cadd(v[0], t);
    v[1] = cmake<C>(m.x + s3 * d.x, m.y + s3 * d.y);
    v[2] = cmake<C>(m.x - s3 * d.x, m.y - s3 * d.y);
  }
};

// Odd prime radix by the symmetric split: a_q = v_q + v_{R-q},
// b_q = v_q - v_{R-q}; X_m = v0 + sum cos(2 pi q m/R) a_q + sum sin(2 pi q m/R) rot(b_q),
// X_{R-m} = v0 + sum cos(...) a_q - sum sin(...) rot(b_q).
template <typename T, int R, bool INV> struct DftOdd {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    constexpr int H = (R - 1) / 2;
    C a[H], b[H];
#pragma unroll
    for (int q = 1; q <= H; ++q) {
      a[q - 1] = cadd(v[q], v[R - q]);
      b[q - 1] = rot<C, INV>(csub(v[q], v[R - q]));
    }
    C x0 = v[0];
    C s0 = x0;
#pragma unroll
    for (int q = 0; q < H; ++q) s0 = cadd(s0, a[q]);
    sfor<1, H + 1>([&](auto mc) {
      constexpr int m = decltype(mc)::value;
      C cs = x0, sn = cmake<C>((T)0, (T)0);
      sfor<1, H + 1>([&](auto qc) {
        constexpr int q = decltype(qc)::value;
        constexpr CosSin w = cossin2pi(q * m, R);
        constexpr T cj = (T)w.c, sj = (T)w.s;
        cs = cmake<C>(cs.x + cj * a[q - 1].x, cs.y + cj * a[q - 1].y);
        sn = cmake<C>(sn.x + sj * b[q - 1].x, sn.y + sj * b[q - 1].y);
      });
      v[m] = cadd(cs, sn);
      v[R - m] = csub(cs, sn);
    });
    v[0] = s0;
  }
};
template <typename T, bool INV> struct Dft<T, 5, INV> : DftOdd<T, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 7, INV> : DftOdd<T, 7, INV> {};

// Two-level DFT of length N = A B in registers: n = A n2 + n1, k = k1 + B k2;
// B-point DFTs over n2, twiddle exp(-+ 2 pi i n1 k1 / N), A-point DFTs over n1.
template <typename T, int A, int B, bool INV> struct DftAB {
  using C = cplx_t<T>;
  __device__ __forceinline__ static void run(C* v) {
    constexpr int N = A * B;
    C y[A][B];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1)
#pragma unroll
      for (int n2 = 0; n2 < B; ++n2) y[n1][n2] = v[A * n2 + n1];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) Dft<T, B, INV>::run(y[n1]);
    sfor<1, A>([&](auto n1c) {
      constexpr int n1 = decltype(n1c)::value;
      sfor<1, B>([&](auto k1c) {
        constexpr int k1 = decltype(k1c)::value;
        y[n1][k1] = twiddle_ct<T, n1 * k1, N, INV>(y[n1][k1]);
      });
    });
#pragma unroll
    for (int k1 = 0; k1 < B; ++k1) {
      C z[A];
#pragma unroll
      for (int n1 = 0; n1 < A; ++n1) z[n1] = y[n1][k1];
      Dft<T, A, INV>::run(z);
#pragma unroll
      for (int k2 = 0; k2 < A; ++k2) v[k1 + B * k2] = z[k2];
    }
  }
};
template <typename T, bool INV> struct Dft<T, 6, INV> : DftAB<T, 2, 3, INV> {};
template <typename T, bool INV> struct Dft<T, 9, INV> : DftAB<T, 3, 3, INV> {};
template <typename T, bool INV> struct Dft<T, 10, INV> : DftAB<T, 2, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 12, INV> : DftAB<T, 4, 3, INV> {};
template <typename T, bool INV> struct Dft<T, 14, INV> : DftAB<T, 2, 7, INV> {};
template <typename T, bool INV> struct Dft<T, 15, INV> : DftAB<T, 3, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 16, INV> : DftAB<T, 4, 4, INV> {};
template <typename T, bool INV> struct Dft<T, 18, INV> : DftAB<T, 2, 9, INV> {};
template <typename T, bool INV> struct Dft<T, 20, INV> : DftAB<T, 4, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 21, INV> : DftAB<T, 3, 7, INV> {};
template <typename T, bool INV> struct Dft<T, 25, INV> : DftAB<T, 5, 5, INV> {};

}  // namespace dft
