// 3D real<->complex transforms over CRAVA's padded in-place grids on gfx950.
//
// Replaces FFTGrid::fftInPlace / invFFTInPlace (reference src/fftgrid.cpp:
// 1783-1841), which call FFTW 2.1.2 rfftw3d(nzp, nyp, nxp) in place.  The
// transform is split into three passes, each one read and one write of the
// grid through HBM:
//   rows : x, contiguous rows of rnxp reals; a real transform of length nxp
//          (via a complex transform of length nxp/2 when nxp is even);
//   cols : y, pencils of stride cnxp complex, tiles of TX consecutive x;
//   cols : z, pencils of stride cnxp*nyp complex, tiles of TX consecutive x.
// Each pencil tile is staged in LDS and transformed with self-sorting
// (Stockham) mixed-radix stages of radix 2, 3, 4, 5, 7 and 8; twiddles come
// from a per-size table computed on the host in long double.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace {

// 256-thread blocks with tiles of at most 4096 (FP32) / 2048 (FP64) complex
// values (32 KB of LDS): three blocks share a CU, so one block's global
// loads overlap another's LDS stages.
constexpr int kThreads = 256;

template <typename T> struct Ept;
template <> struct Ept<float>  { static constexpr int value = 16; };  // 4096 complex / block
template <> struct Ept<double> { static constexpr int value = 8; };   // 2048 complex / block

// Unsigned division by a runtime constant (n < 2^31): q = (umulhi(n,m)+n)>>s.
struct FastDiv {
  unsigned m; int s; int d;
};
FastDiv make_fastdiv(int d) {
  FastDiv f; f.d = d; f.s = 0;
  while ((1u << f.s) < (unsigned)d) ++f.s;
  unsigned long long num = (1ull << 32) * (unsigned long long)((1ull << f.s) - d);
  f.m = (unsigned)(num / (unsigned)d + 1);
  return f;
}
__device__ __forceinline__ int fdiv(int n, FastDiv f) {
  return (int)((__umulhi((unsigned)n, f.m) + (unsigned)n) >> f.s);
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of tiles so that neighbouring tiles (which share cache lines at
// their edges) meet in the same L2.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks) {
  const int x = bid & 7, i = bid >> 3;
  const int q = nblocks >> 3, r = nblocks & 7;
  return x * q + min(x, r) + i;
}

// multiply by -i (forward) or +i (inverse)
template <typename C, bool INV> __device__ __forceinline__ C rot(C a) {
  return INV ? cmake<C>(-a.y, a.x) : cmake<C>(a.y, -a.x);
}

// ---- small DFTs (sign -1 forward, +1 inverse), in place ------------------
template <typename T, int R, bool INV> struct Dft;

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

// cos(2 pi j / R), sin(2 pi j / R) for j = 1 .. (R-1)/2
template <int R> struct OddTab;
template <> struct OddTab<5> {
  __device__ static constexpr double c(int j) { return j == 1 ? 0.30901699437494742410 : -0.80901699437494742410; }
  __device__ static constexpr double s(int j) { return j == 1 ? 0.95105651629515357212 : 0.58778525229247312917; }
};
template <> struct OddTab<7> {
  __device__ static constexpr double c(int j) {
    return j == 1 ? 0.62348980185873353053 : (j == 2 ? -0.22252093395631440429 : -0.90096886790241912624);
  }
  __device__ static constexpr double s(int j) {
    return j == 1 ? 0.78183148246802980871 : (j == 2 ? 0.97492791218182360702 : 0.43388373911755812048);
  }
};

// Odd radix by the symmetric split: a_q = v_q + v_{R-q}, b_q = v_q - v_{R-q};
// X_m = v0 + sum cos(2 pi q m/R) a_q + sum sin(2 pi q m/R) rot(b_q),
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
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      C cs = x0, sn = cmake<C>((T)0, (T)0);
#pragma unroll
      for (int q = 1; q <= H; ++q) {
        const int j = (q * m) % R;
        // cos/sin of 2 pi j / R with j possibly > R/2
        const T cj = (T)(j <= H ? OddTab<R>::c(j) : OddTab<R>::c(R - j));
        const T sj = (T)(j <= H ? OddTab<R>::s(j) : -OddTab<R>::s(R - j));
        cs = cmake<C>(cs.x + cj * a[q - 1].x, cs.y + cj * a[q - 1].y);
        sn = cmake<C>(sn.x + sj * b[q - 1].x, sn.y + sj * b[q - 1].y);
      }
      v[m] = cadd(cs, sn);
      v[R - m] = csub(cs, sn);
    }
    v[0] = s0;
  }
};
template <typename T, bool INV> struct Dft<T, 5, INV> : DftOdd<T, 5, INV> {};
template <typename T, bool INV> struct Dft<T, 7, INV> : DftOdd<T, 7, INV> {};

// ---- one Stockham stage over an LDS tile ---------------------------------
// Tile element (e, c) lives at lds[e*S + c]; C = 1 << cshift pencils.
// Butterfly b -> pencil c = b % C, index j = b / C; inputs j + r*L/R;
// outputs (j - k)*R + k + r*Ns with k = j % Ns.
template <typename T, int R, bool INV>
__device__ __forceinline__ void stage(cplx_t<T>* lds, int S, int cshift, int L, int Ns,
                                      const cplx_t<T>* __restrict__ tw) {
  using C = cplx_t<T>;
  constexpr int MAXB = (Ept<T>::value + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
  const int cm = (1 << cshift) - 1;
  const int tid = threadIdx.x;
  C v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = tid + q * kThreads;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = lds[(j + r * Lr) * S + c];
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = tid + q * kThreads;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
      const int k = j % Ns;
      const C* w = tw + k * (R - 1);
#pragma unroll
      for (int r = 1; r < R; ++r) v[q][r] = INV ? cmulc(v[q][r], w[r - 1]) : cmul(v[q][r], w[r - 1]);
      Dft<T, R, INV>::run(v[q]);
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(base + r * Ns) * S + c] = v[q][r];
    }
  }
  __syncthreads();
}

template <typename T, bool INV>
__device__ __forceinline__ void run_stages(cplx_t<T>* lds, int S, int cshift, const Fft1D& f,
                                           const cplx_t<T>* __restrict__ tw) {
  for (int s = 0; s < f.nst; ++s) {
    const cplx_t<T>* t = tw + f.tw[s];
    switch (f.R[s]) {
      case 2: stage<T, 2, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
      case 3: stage<T, 3, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
      case 4: stage<T, 4, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
      case 5: stage<T, 5, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
      case 7: stage<T, 7, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
      default: stage<T, 8, INV>(lds, S, cshift, f.L, f.Ns[s], t); break;
    }
  }
}

// ---- x pass: rows ----------------------------------------------------------
template <typename T>
struct RowArgs {
  T* dst;              // real view (output of C2R / row base of R2C output)
  const T* src;        // real view input (R2C) or complex input buffer (C2R)
  const T* add;        // C2R: real grid added to the output (may be null)
  Fft1D f;             // length M = nxp/2 (even) or nxp (odd)
  const cplx_t<T>* tw;
  const cplx_t<T>* wpost;  // exp(-2 pi i k / nxp), k = 0..nxp/2 (even only)
  int nrows, rnxp, nxp, cnxp;
  int xeven;
  int pshift;          // rows per block = 1 << pshift
  FastDiv divL;        // by f.L
  FastDiv divL1;       // by f.L + 1 (even) or cnxp (odd)
  T scale;
};

template <typename T>
__global__ __launch_bounds__(kThreads) void k_rows_r2c(RowArgs<T> a) {
  using C = cplx_t<T>;
  extern __shared__ unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int P = 1 << a.pshift, S = P + 1;
  const int row0 = xcd_remap(blockIdx.x, gridDim.x) << a.pshift;
  const int nr = min(P, a.nrows - row0);
  const int L = a.f.L;
  const int tid = threadIdx.x;
  // load: every load of the thread is issued before the first LDS store, so
  // a wave keeps Ept loads in flight
  constexpr int EPT = Ept<T>::value;
  C buf[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * kThreads;
    buf[q] = cmake<C>((T)0, (T)0);
    if (idx < P * L) {
      const int c = fdiv(idx, a.divL), e = idx - c * L;
      if (c < nr) {
        const T* row = a.src + (size_t)(row0 + c) * a.rnxp;
        buf[q] = a.xeven ? *reinterpret_cast<const C*>(row + 2 * e) : cmake<C>(row[e], (T)0);
      }
    }
  }
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * kThreads;
    if (idx < P * L) {
      const int c = fdiv(idx, a.divL), e = idx - c * L;
      lds[e * S + c] = cscale(buf[q], a.scale);
    }
  }
  __syncthreads();
  run_stages<T, false>(lds, S, a.pshift, a.f, a.tw);
  // post-process and store cnxp complex per row
  const int K = a.cnxp;
  for (int idx = tid; idx < P * K; idx += kThreads) {
    const int c = fdiv(idx, a.divL1), k = idx - c * K;
    if (c >= nr) continue;
    C X;
    if (a.xeven) {
      const int M = L;
      C za = lds[(k == M ? 0 : k) * S + c];
      C zb = cconj(lds[(k == 0 ? 0 : M - k) * S + c]);
      C d = csub(za, zb);
      C t = cmul(a.wpost[k], d);
      // X = 0.5 (za + zb) - 0.5 i t
      X = cmake<C>((T)0.5 * (za.x + zb.x + t.y), (T)0.5 * (za.y + zb.y - t.x));
    } else {
      X = lds[k * S + c];
    }
    C* out = reinterpret_cast<C*>(a.dst + (size_t)(row0 + c) * a.rnxp);
    out[k] = X;
  }
}

template <typename T>
__global__ __launch_bounds__(kThreads) void k_rows_c2r(RowArgs<T> a) {
  using C = cplx_t<T>;
  extern __shared__ unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int P = 1 << a.pshift, S = P + 1;
  const int row0 = xcd_remap(blockIdx.x, gridDim.x) << a.pshift;
  const int nr = min(P, a.nrows - row0);
  const int L = a.f.L;
  const int tid = threadIdx.x;
  constexpr int EPT = Ept<T>::value;
  C ba[EPT], bb[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * kThreads;
    ba[q] = cmake<C>((T)0, (T)0);
    bb[q] = ba[q];
    if (idx < P * L) {
      const int c = fdiv(idx, a.divL), k = idx - c * L;
      if (c < nr) {
        const C* X = reinterpret_cast<const C*>(a.src + (size_t)(row0 + c) * a.rnxp);
        if (a.xeven) {
          ba[q] = X[k];
          bb[q] = X[L - k];
        } else {
          ba[q] = (k < a.cnxp) ? X[k] : cconj(X[L - k]);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = tid + q * kThreads;
    if (idx < P * L) {
      const int c = fdiv(idx, a.divL), k = idx - c * L;
      C z;
      if (a.xeven) {
        C xa = ba[q], xb = bb[q];
        if (k == 0) { xa.y = (T)0; xb.y = (T)0; }
        xb = cconj(xb);
        C d = csub(xa, xb);
        C t = cmulc(d, a.wpost[k]);  // d * exp(+2 pi i k / nxp)
        // Z = (xa + xb) + i t
        z = cmake<C>(xa.x + xb.x - t.y, xa.y + xb.y + t.x);
      } else {
        z = ba[q];
        if (k == 0) z.y = (T)0;
      }
      lds[k * S + c] = z;
    }
  }
  __syncthreads();
  run_stages<T, true>(lds, S, a.pshift, a.f, a.tw);
  for (int idx = tid; idx < P * L; idx += kThreads) {
    const int c = fdiv(idx, a.divL), n = idx - c * L;
    if (c >= nr) continue;
    const size_t rbase = (size_t)(row0 + c) * a.rnxp;
    C z = lds[n * S + c];
    if (a.xeven) {
      C o = cscale(z, a.scale);
      if (a.add) {
        C ad = *reinterpret_cast<const C*>(a.add + rbase + 2 * n);
        o = cadd(o, ad);
      }
      *reinterpret_cast<C*>(a.dst + rbase + 2 * n) = o;
    } else {
      T o = z.x * a.scale;
      if (a.add) o += a.add[rbase + n];
      a.dst[rbase + n] = o;
    }
  }
}

// ---- y / z passes: strided complex pencils ---------------------------------
template <typename T>
struct ColArgs {
  cplx_t<T>* data;
  Fft1D f;
  const cplx_t<T>* tw;
  long long estride;   // element stride of a pencil (complex units)
  long long ostride;   // stride of the outer index
  int ncols;           // number of x columns (cnxp)
  int ntx;             // x tiles per outer index
  int cshift;          // TX = 1 << cshift
};

template <typename T, bool INV>
__global__ __launch_bounds__(kThreads) void k_cols(ColArgs<T> a) {
  using C = cplx_t<T>;
  extern __shared__ unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int tile = xcd_remap(blockIdx.x, gridDim.x);
  const int o = tile / a.ntx, t = tile - o * a.ntx;
  const int TX = 1 << a.cshift, cm = TX - 1;
  const int x0 = t * TX;
  const int nc = min(TX, a.ncols - x0);
  C* base = a.data + (size_t)o * a.ostride + x0;
  const int n = a.f.L << a.cshift;
  constexpr int EPT = Ept<T>::value;
  C buf[EPT];
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    const int c = idx & cm, e = idx >> a.cshift;
    buf[q] = (idx < n && c < nc) ? base[(size_t)e * a.estride + c] : cmake<C>((T)0, (T)0);
  }
#pragma unroll
  for (int q = 0; q < EPT; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    if (idx < n) lds[idx] = buf[q];
  }
  __syncthreads();
  run_stages<T, INV>(lds, TX, a.cshift, a.f, a.tw);
  for (int idx = threadIdx.x; idx < n; idx += kThreads) {
    const int c = idx & cm, e = idx >> a.cshift;
    if (c < nc) base[(size_t)e * a.estride + c] = lds[idx];
  }
}

// ---- planning ----------------------------------------------------------------
std::vector<int> factorize(int L) {
  std::vector<int> r;
  int m = L;
  while (m % 8 == 0) { r.push_back(8); m /= 8; }
  if (m % 4 == 0) { r.push_back(4); m /= 4; }
  if (m % 2 == 0) { r.push_back(2); m /= 2; }
  for (int p : {5, 3, 7}) while (m % p == 0) { r.push_back(p); m /= p; }
  if (m != 1) r.clear();
  return r;
}

bool plan_1d(int L, Fft1D& f, std::vector<long double>& twr, std::vector<long double>& twi) {
  f.L = L;
  std::vector<int> rs = (L == 1) ? std::vector<int>{} : factorize(L);
  if (L != 1 && rs.empty()) return false;
  if ((int)rs.size() > kMaxStages) return false;
  f.nst = (int)rs.size();
  int Ns = 1;
  const long double twopi = 6.283185307179586476925286766559L;
  for (int s = 0; s < f.nst; ++s) {
    const int R = rs[s];
    f.R[s] = R; f.Ns[s] = Ns; f.tw[s] = (int)twr.size();
    const long long den = (long long)Ns * R;
    for (int k = 0; k < Ns; ++k)
      for (int r = 1; r < R; ++r) {
        const long long num = ((long long)k * r) % den;
        const long double ang = -twopi * (long double)num / (long double)den;
        twr.push_back(cosl(ang));
        twi.push_back(sinl(ang));
      }
    Ns *= R;
  }
  return true;
}

int pow2_floor(int v) { int p = 1; while (p * 2 <= v) p *= 2; return p; }
int ilog2(int v) { int s = 0; while ((1 << s) < v) ++s; return s; }

template <typename T>
size_t row_lds_bytes(const FftPlan* p) { return (size_t)((p->rows_per_block + 1) * p->fx.L) * sizeof(cplx_t<T>); }

template <typename T>
int launch_rows(crv_ctx* ctx, FftPlan* p, bool inv, T* dst, const T* src, const T* add, int row_begin,
                int row_end, double scale) {
  const int cnxp = p->nxp / 2 + 1, rnxp = 2 * cnxp;
  const size_t off = (size_t)row_begin * rnxp;
  RowArgs<T> a{};
  a.dst = dst + off; a.src = src + off; a.add = add ? add + off : nullptr;
  a.f = p->fx; a.tw = reinterpret_cast<const cplx_t<T>*>(p->tw);
  a.wpost = a.tw + p->post_off;
  a.nrows = row_end - row_begin; a.rnxp = rnxp; a.nxp = p->nxp; a.cnxp = cnxp; a.xeven = p->xeven ? 1 : 0;
  a.pshift = ilog2(p->rows_per_block);
  a.divL = make_fastdiv(p->fx.L); a.divL1 = make_fastdiv(cnxp);
  a.scale = (T)scale;
  const int nb = (a.nrows + p->rows_per_block - 1) / p->rows_per_block;
  if (inv) {
    ProfScope ps(ctx, "fft_x_c2r");
    hipLaunchKernelGGL(k_rows_c2r<T>, dim3(nb), dim3(kThreads), row_lds_bytes<T>(p), ctx->stream, a);
  } else {
    ProfScope ps(ctx, "fft_x_r2c");
    hipLaunchKernelGGL(k_rows_r2c<T>, dim3(nb), dim3(kThreads), row_lds_bytes<T>(p), ctx->stream, a);
  }
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// y pass over z planes [k0, k1) or z pass over y rows [j0, j1)
template <typename T>
int launch_cols(crv_ctx* ctx, FftPlan* p, bool inv, bool zaxis, cplx_t<T>* data, int o0, int o1) {
  using C = cplx_t<T>;
  const int cnxp = p->nxp / 2 + 1;
  ColArgs<T> a{};
  a.tw = reinterpret_cast<const C*>(p->tw);
  a.ncols = cnxp;
  int tx;
  if (zaxis) {
    a.f = p->fz; a.estride = (long long)cnxp * p->nyp; a.ostride = cnxp; tx = p->tx_z;
  } else {
    a.f = p->fy; a.estride = cnxp; a.ostride = (long long)cnxp * p->nyp; tx = p->tx_y;
  }
  a.data = data + (size_t)o0 * a.ostride;
  a.cshift = ilog2(tx);
  a.ntx = (cnxp + tx - 1) / tx;
  const int nb = a.ntx * (o1 - o0);
  const size_t lds = (size_t)a.f.L * tx * sizeof(C);
  static const char* names[4] = {"fft_y_fwd", "fft_y_inv", "fft_z_fwd", "fft_z_inv"};
  ProfScope ps(ctx, names[(zaxis ? 2 : 0) + (inv ? 1 : 0)]);
  if (inv)
    hipLaunchKernelGGL((k_cols<T, true>), dim3(nb), dim3(kThreads), lds, ctx->stream, a);
  else
    hipLaunchKernelGGL((k_cols<T, false>), dim3(nb), dim3(kThreads), lds, ctx->stream, a);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// z planes per chunk when the x and y passes are interleaved so that the
// second pass of a chunk finds the first pass's output still in the
// memory-side cache; 0 = whole grid per pass
int chunk_planes(const FftPlan* p, size_t elem) {
  if (p->chunk_mb <= 0) return p->nzp;
  const size_t plane = (size_t)(p->nxp / 2 + 1) * 2 * p->nyp * elem;
  const size_t want = (size_t)p->chunk_mb << 20;
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)p->nzp, want / std::max<size_t>(plane, 1)));
}

template <typename T>
int launch_r2c(crv_ctx* ctx, FftPlan* p, T* dst, const T* src, double scale) {
  using C = cplx_t<T>;
  C* cd = reinterpret_cast<C*>(dst);
  const int cp = chunk_planes(p, sizeof(T));
  int rc;
  for (int k0 = 0; k0 < p->nzp; k0 += cp) {
    const int k1 = std::min(p->nzp, k0 + cp);
    if ((rc = launch_rows<T>(ctx, p, false, dst, src, nullptr, k0 * p->nyp, k1 * p->nyp, scale)) != CRV_OK) return rc;
    if (p->nyp > 1 && (rc = launch_cols<T>(ctx, p, false, false, cd, k0, k1)) != CRV_OK) return rc;
  }
  if (p->nzp > 1 && (rc = launch_cols<T>(ctx, p, false, true, cd, 0, p->nyp)) != CRV_OK) return rc;
  return CRV_OK;
}

template <typename T>
int launch_c2r(crv_ctx* ctx, FftPlan* p, T* buf, T* dst, double scale, const T* add) {
  using C = cplx_t<T>;
  C* cd = reinterpret_cast<C*>(buf);
  int rc;
  if (p->nzp > 1 && (rc = launch_cols<T>(ctx, p, true, true, cd, 0, p->nyp)) != CRV_OK) return rc;
  const int cp = chunk_planes(p, sizeof(T));
  for (int k0 = 0; k0 < p->nzp; k0 += cp) {
    const int k1 = std::min(p->nzp, k0 + cp);
    if (p->nyp > 1 && (rc = launch_cols<T>(ctx, p, true, false, cd, k0, k1)) != CRV_OK) return rc;
    if ((rc = launch_rows<T>(ctx, p, true, dst, buf, add, k0 * p->nyp, k1 * p->nyp, scale)) != CRV_OK) return rc;
  }
  return CRV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
void crv_free_plan(FftPlan* p) {
  if (!p) return;
  if (p->tw) (void)hipFree(p->tw);
  if (p->noise_prefix) (void)hipFree(p->noise_prefix);
  delete p;
}

int crv_get_plan(crv_ctx* ctx, int nxp, int nyp, int nzp, int prec, FftPlan** out) {
  auto key = std::make_tuple(nxp, nyp, nzp, prec);
  auto it = ctx->plans.find(key);
  if (it != ctx->plans.end()) { *out = it->second; return CRV_OK; }
  FftPlan* p = new FftPlan();
  p->nxp = nxp; p->nyp = nyp; p->nzp = nzp; p->prec = prec;
  p->xeven = (nxp % 2 == 0);
  std::vector<long double> tr, ti;
  const int Lx = p->xeven ? nxp / 2 : nxp;
  if (!plan_1d(Lx, p->fx, tr, ti) || !plan_1d(nyp, p->fy, tr, ti) || !plan_1d(nzp, p->fz, tr, ti)) {
    delete p;
    return crv_set_error(ctx, CRV_E_UNSUPPORTED, "grid size is not a product of 2, 3, 5, 7");
  }
  p->post_off = (int)tr.size();
  const long double twopi = 6.283185307179586476925286766559L;
  for (int k = 0; k <= nxp / 2; ++k) {
    const long double ang = -twopi * (long double)k / (long double)nxp;
    tr.push_back(cosl(ang));
    ti.push_back(sinl(ang));
  }
  {
    const int rc = crv_plan_shapes(ctx, p);
    if (rc != CRV_OK) { delete p; return rc; }
  }
  const size_t n = tr.size();
  hipError_t e;
  if (prec == CRV_FP64) {
    std::vector<double2> h(n);
    for (size_t i = 0; i < n; ++i) { h[i].x = (double)tr[i]; h[i].y = (double)ti[i]; }
    e = hipMalloc(&p->tw, n * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpy(p->tw, h.data(), n * sizeof(double2), hipMemcpyHostToDevice);
  } else {
    std::vector<float2> h(n);
    for (size_t i = 0; i < n; ++i) { h[i].x = (float)tr[i]; h[i].y = (float)ti[i]; }
    e = hipMalloc(&p->tw, n * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(p->tw, h.data(), n * sizeof(float2), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    crv_free_plan(p);
    return crv_set_error(ctx, CRV_E_HIP, std::string("plan twiddles: ") + hipGetErrorString(e));
  }
  ctx->plans[key] = p;
  *out = p;
  return CRV_OK;
}

// Launch shapes: a tile holds at most 4096 (FP32) / 2048 (FP64) complex
// values (32 KB of LDS); x rows per block P with (P+1) L <= that; y/z tiles
// of TX consecutive x columns, 16 (FP32) or 8 (FP64) at most, halved until
// the pencil tile fits.  CRV_TUNE_* values override, within those limits.
int crv_plan_shapes(crv_ctx* ctx, FftPlan* p) {
  const int maxe = kThreads * ((p->prec == CRV_FP64) ? Ept<double>::value : Ept<float>::value);
  const int txdef = (p->prec == CRV_FP64) ? 8 : 16;
  const int Lx = p->fx.L;
  int P = 1;
  const int pmax = ctx->tune[CRV_TUNE_ROWS] > 0 ? ctx->tune[CRV_TUNE_ROWS] : 64;
  while (2 * P <= pmax && (2 * P + 1) * Lx <= maxe) P *= 2;
  if ((P + 1) * Lx > maxe) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "nxp too large");
  p->rows_per_block = P;
  auto pick_tx = [&](int L, int want) {
    int t = want > 0 ? pow2_floor(want) : txdef;
    while (t > 1 && L * t > maxe) t /= 2;
    return (L * t <= maxe) ? t : 0;
  };
  p->tx_y = pick_tx(p->nyp, ctx->tune[CRV_TUNE_TX_Y]);
  p->tx_z = pick_tx(p->nzp, ctx->tune[CRV_TUNE_TX_Z]);
  if (!p->tx_y || !p->tx_z) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "nyp or nzp too large");
  p->chunk_mb = ctx->tune[CRV_TUNE_CHUNK_MB];
  return CRV_OK;
}

int crv_launch_r2c(crv_ctx* ctx, FftPlan* plan, void* dst, const void* src, double scale) {
  if (plan->prec == CRV_FP64)
    return launch_r2c<double>(ctx, plan, (double*)dst, (const double*)src, scale);
  return launch_r2c<float>(ctx, plan, (float*)dst, (const float*)src, scale);
}

int crv_launch_c2r(crv_ctx* ctx, FftPlan* plan, void* buf, void* dst, double scale, const void* add) {
  if (plan->prec == CRV_FP64)
    return launch_c2r<double>(ctx, plan, (double*)buf, (double*)dst, scale, (const double*)add);
  return launch_c2r<float>(ctx, plan, (float*)buf, (float*)dst, scale, (const float*)add);
}
