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
//
// Each tile goes through self-sorting (Stockham) mixed-radix stages.  In
// the column passes the first stage reads its butterflies straight from HBM
// into registers and writes LDS, middle stages go LDS -> registers -> LDS,
// and the last stage writes HBM straight from registers.  Row tiles move
// between HBM and LDS with the lanes along the rows (coalesced) and every
// stage runs LDS -> LDS; the real-to-complex post-processing (R2C) and
// pre-processing (C2R) work on the staged rows.  Lengths that are a product
// of two radices from a fixed list of pairs up to 25 x 18 (composite radices
// done in registers, dft.hpp) -- the production grid's 225, 450 and 320
// among them -- get kernels compiled for their pair: two stages, i.e. one
// LDS exchange and one barrier per tile.  Other lengths run through radices
// 2..8 with the stage radix dispatched at run time.
// Twiddles of all stages come from a per-size table computed on the host in
// long double and copied to LDS by each block; the index arithmetic uses
// multiply-shift division, no runtime modulo.
// Floating-point contraction only within an expression (fused multiply-add
// where the source writes a * b + c): the rounding of every butterfly then
// follows from the source alone, not from the kernel it is inlined into, so
// different kernels that run the same stages (the z passes and the fused z
// pass of crv_invert) produce the same bits.
#pragma clang fp contract(on)
#include "internal.hpp"
#include "dft.hpp"
#include "eiggen.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <unordered_map>

namespace {

using dft::Dft;

// 256-thread blocks with tiles of at most 4096 (FP32) / 2048 (FP64) complex
// values (32 KB of LDS).
constexpr int kThreads = 256;

// passes that walk their tiles backwards by default (CRV_TUNE_ORDER bits):
// the y forward pass (after an x pass that wrote the grid plane by plane)
// and the x C2R pass (after the y inverse pass), which then start on the
// planes still in the memory-side cache (measured at 450x450x320 FP32 in
// the inversion step: 8.72 -> 8.54 ms), and the x R2C pass, which with its
// streaming input loads leaves the y pass a little more of its output when
// it walks backwards too (7.28 -> 7.24 ms)
constexpr int kDefaultOrder = (1 << 0) | (1 << 1) | (1 << 5);

template <typename T> struct Ept;
template <> struct Ept<float>  { static constexpr int value = 16; };  // 4096 complex / block
template <> struct Ept<double> { static constexpr int value = 8; };   // 2048 complex / block

// Elements per thread of a kernel: the two-stage (radix-pair) kernels and
// the general kernel for long pencils hold Ept; the general kernel whose
// tile fits half of that (H) holds Ept / 2 -- its stage loop dispatches the
// radix at run time and would otherwise need about twice the registers.
template <typename T, int RF, bool H>
constexpr int ept() { return (RF == 0 && H) ? Ept<T>::value / 2 : Ept<T>::value; }


// per-stage divisors: L / R_s (butterflies per pencil) and Ns
struct StageDiv {
  FastDiv lr[kMaxStages];
  FastDiv ns[kMaxStages];
};

StageDiv make_stage_div(const Fft1D& f) {
  StageDiv d{};
  for (int s = 0; s < f.nst; ++s) {
    d.lr[s] = make_fastdiv(f.L / f.R[s]);
    d.ns[s] = make_fastdiv(f.Ns[s]);
  }
  return d;
}

// Blocks are dealt round-robin to the 8 XCDs; give each XCD a contiguous
// range of tiles so that neighbouring tiles (which share cache lines at
// their edges) meet in the same L2.  rev walks each range backwards: a pass
// that follows another over the same grid then starts on the tiles written
// last, which the memory-side cache still holds.
__device__ __forceinline__ int xcd_remap(int bid, int nblocks, int rev) {
  const int x = bid & 7, i = bid >> 3;
  const int q = nblocks >> 3, r = nblocks & 7;
  const int cnt = q + (x < r ? 1 : 0);
  return x * q + min(x, r) + (rev ? cnt - 1 - i : i);
}

// Two consecutive reals (p aligned to their pair) in one streaming load: two
// scalar loads of p and p + 1 in a kernel that also reads single reals on
// another branch are not merged by the compiler.
template <typename T>
__device__ __forceinline__ cplx_t<T> load_pair_nt(const T* p) {
  typedef T pair_t __attribute__((ext_vector_type(2)));
  const pair_t r = __builtin_nontemporal_load(reinterpret_cast<const pair_t*>(p));
  return cmake<cplx_t<T>>(r.x, r.y);
}

template <typename T>
__device__ __forceinline__ cplx_t<T> twiddle(cplx_t<T> a, cplx_t<T> w, bool inv) {
  return inv ? cmulc(a, w) : cmul(a, w);
}

// Butterfly b of a stage -> (pencil c, butterfly index j).  Pencil-fast
// (c = b mod C) keeps LDS accesses consecutive across lanes and, for the
// column passes, HBM accesses coalesced over the tile's x columns;
// index-fast (j = b mod Lr) makes the HBM accesses of a row pass run along
// the row.
template <bool JFAST>
__device__ __forceinline__ void bmap(int b, int cshift, FastDiv dlr, int& c, int& j) {
  if (JFAST) {
    c = fdiv(b, dlr);
    j = b - c * dlr.d;
  } else {
    c = b & ((1 << cshift) - 1);
    j = b >> cshift;
  }
}

// ---- stage pieces -----------------------------------------------------------
// Tile element (e, c) lives at lds[e*S + c]; C = 1 << cshift pencils.
// Stage s: butterfly (c, j) reads e = j + r*Lr, twiddles by tw[k*(R-1) + r-1]
// with k = j mod Ns, and writes e' = (j - k)*R + k + r*Ns.

// first stage (Ns = 1, no twiddles): inputs from load(c, e)
template <typename T, int R, bool INV, bool JFAST, int E, typename Load>
__device__ __forceinline__ void stage_first(cplx_t<T>* lds, int S, int cshift, int L, FastDiv dlr, Load load) {
  using C = cplx_t<T>;
  constexpr int MAXB = (E + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
  C v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      int c, j;
      bmap<JFAST>(b, cshift, dlr, c, j);
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = load(c, j + r * Lr);
    }
  }
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      int c, j;
      bmap<JFAST>(b, cshift, dlr, c, j);
      Dft<T, R, INV>::run(v[q]);
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(j * R + r) * S + c] = v[q][r];
    }
  }
}

// first stage whose inputs are made in two steps: pre(c, e) for every
// element of the thread's butterflies (the loads, all in flight together),
// then gen(c, e, p) from what pre returned.  A generator whose value needs a
// load and then a branch on it would otherwise wait for each load in turn.
template <typename T, int R, bool INV, int E, typename Pre, typename Gen>
__device__ __forceinline__ void stage_first_pre(cplx_t<T>* lds, int S, int cshift, int L, Pre pre, Gen gen) {
  using C = cplx_t<T>;
  using P = decltype(pre(0, 0));
  constexpr int MAXB = (E + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
  const int cm = (1 << cshift) - 1;
  P pv[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
#pragma unroll
      for (int r = 0; r < R; ++r) pv[q][r] = pre(c, j + r * Lr);
    }
  }
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
      C v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = gen(c, j + r * Lr, pv[q][r]);
      Dft<T, R, INV>::run(v);
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(j * R + r) * S + c] = v[r];
    }
  }
}

// middle stage: LDS -> LDS (pencil-fast); ends with the tile written.
// TW = false: a first stage (Ns = 1), whose twiddles are all 1.
template <typename T, int R, bool INV, bool TW, int E>
__device__ __forceinline__ void stage_mid(cplx_t<T>* lds, int S, int cshift, int L, int Ns, FastDiv dns,
                                          const cplx_t<T>* tw) {
  using C = cplx_t<T>;
  constexpr int MAXB = (E + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
  const int cm = (1 << cshift) - 1;
  C v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
      v[q][0] = lds[j * S + c];
      if constexpr (TW) {
        const int k = j - fdiv(j, dns) * Ns;
        const C* w = tw + k * (R - 1);
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = twiddle<T>(lds[(j + r * Lr) * S + c], w[r - 1], INV);
      } else {
#pragma unroll
        for (int r = 1; r < R; ++r) v[q][r] = lds[(j + r * Lr) * S + c];
      }
      Dft<T, R, INV>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
      const int k = TW ? j - fdiv(j, dns) * Ns : 0;
      const int base = (j - k) * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(base + r * Ns) * S + c] = v[q][r];
    }
  }
  __syncthreads();
}

// first stage (Ns = 1, no twiddles) whose inputs come from load(c, e) --
// values derived from LDS in another layout; ends with the tile written in
// the lds[e*S + c] layout over that LDS (all reads finish before the writes)
template <typename T, int R, bool INV, int E, typename Load>
__device__ __forceinline__ void stage_first_lds(cplx_t<T>* lds, int S, int cshift, int L, Load load) {
  using C = cplx_t<T>;
  constexpr int MAXB = (E + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
  const int cm = (1 << cshift) - 1;
  C v[MAXB][R];
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
#pragma unroll
      for (int r = 0; r < R; ++r) v[q][r] = load(c, j + r * Lr);
      Dft<T, R, INV>::run(v[q]);
    }
  }
  __syncthreads();
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      const int c = b & cm, j = b >> cshift;
#pragma unroll
      for (int r = 0; r < R; ++r) lds[(j * R + r) * S + c] = v[q][r];
    }
  }
  __syncthreads();
}

// last stage (Ns = L/R, so k = j): LDS -> store(c, e)
template <typename T, int R, bool INV, bool JFAST, int E, typename Store>
__device__ __forceinline__ void stage_last(const cplx_t<T>* lds, int S, int cshift, int L, FastDiv dlr,
                                           const cplx_t<T>* tw, Store store) {
  using C = cplx_t<T>;
  constexpr int MAXB = (E + R - 1) / R;
  const int Lr = L / R;
  const int total = Lr << cshift;
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int b = threadIdx.x + q * (int)blockDim.x;
    if (b < total) {
      int c, j;
      bmap<JFAST>(b, cshift, dlr, c, j);
      const C* w = tw + j * (R - 1);
      C v[R];
      v[0] = lds[j * S + c];
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = twiddle<T>(lds[(j + r * Lr) * S + c], w[r - 1], INV);
      Dft<T, R, INV>::run(v);
#pragma unroll
      for (int r = 0; r < R; ++r) store(c, j + r * Lr, v[r]);
    }
  }
}

// a one-stage transform: load -> DFT -> store, no LDS
template <typename T, int R, bool INV, int E, typename Load, typename Store>
__device__ __forceinline__ void stage_single(int cshift, Load load, Store store) {
  using C = cplx_t<T>;
  constexpr int MAXB = (E + R - 1) / R;
  const int total = 1 << cshift;  // one butterfly per pencil
#pragma unroll
  for (int q = 0; q < MAXB; ++q) {
    const int c = threadIdx.x + q * (int)blockDim.x;
    if (c < total) {
      C v[R];
#pragma unroll
      for (int r = 0; r < R; ++r) v[r] = load(c, r);
      Dft<T, R, INV>::run(v);
#pragma unroll
      for (int r = 0; r < R; ++r) store(c, r, v[r]);
    }
  }
}

#define CRV_RADIX_SWITCH(RR, CALL)                     \
  switch (RR) {                                        \
    case 2: { constexpr int R_ = 2; CALL; } break;     \
    case 3: { constexpr int R_ = 3; CALL; } break;     \
    case 4: { constexpr int R_ = 4; CALL; } break;     \
    case 5: { constexpr int R_ = 5; CALL; } break;     \
    case 7: { constexpr int R_ = 7; CALL; } break;     \
    default: { constexpr int R_ = 8; CALL; } break;    \
  }

// Radix pairs of the two-stage kernels (first stage, last stage), and the
// radices of the general kernel, whose stage loop dispatches at run time.
#define CRV_FFT_PAIRS(X)                                                                              \
  X(25, 18) X(15, 15) X(20, 16) X(16, 16) X(16, 8) X(20, 20) X(25, 16) X(18, 16) X(16, 12) X(25, 8)   \
  X(18, 15) X(18, 12) X(15, 9) X(12, 9) X(25, 10) X(25, 15) X(20, 9) X(15, 10) X(12, 7) X(15, 8)      \
  X(15, 14) X(21, 15) X(12, 5) X(7, 6) X(15, 5)
constexpr int kSmallRadices[] = {8, 4, 2, 5, 3, 7};

// Whole transform of a tile.  load/store move HBM data; with LAST_TO_LDS the
// last stage also writes LDS (natural order) and store is not used.
template <typename T, bool INV, bool JFAST, bool LAST_TO_LDS, int RF, int RL, bool H, typename Load, typename Store>
__device__ __forceinline__ void transform_tile(cplx_t<T>* lds, int S, int cshift, const Fft1D& f, const StageDiv& sd,
                                               const cplx_t<T>* twl, Load load, Store store) {
  constexpr int E = ept<T, RF, H>();
  const int L = f.L, nst = f.nst;
  if constexpr (RF > 0) {
    // two stages of radices RF, RL known at compile time
    stage_first<T, RF, INV, JFAST, E>(lds, S, cshift, L, sd.lr[0], load);
    __syncthreads();
    const cplx_t<T>* t = twl + (f.tw[1] - f.tw0);
    if constexpr (LAST_TO_LDS)
      stage_mid<T, RL, INV, true, E>(lds, S, cshift, L, f.Ns[1], sd.ns[1], t);
    else
      stage_last<T, RL, INV, JFAST, E>(lds, S, cshift, L, sd.lr[1], t, store);
    return;
  }
  if (nst == 0) {
    // length 1: the tile is its own transform
    const int total = 1 << cshift;
    for (int b = threadIdx.x; b < total; b += (int)blockDim.x) {
      if (LAST_TO_LDS) lds[b] = load(b, 0);  // element 0 of pencil b
      else store(b, 0, load(b, 0));
    }
    __syncthreads();
    return;
  }
  if (nst == 1 && !LAST_TO_LDS) {
    CRV_RADIX_SWITCH(f.R[0], (stage_single<T, R_, INV, E>(cshift, load, store)));
    return;
  }
  CRV_RADIX_SWITCH(f.R[0], (stage_first<T, R_, INV, JFAST, E>(lds, S, cshift, L, sd.lr[0], load)));
  __syncthreads();
  const int nmid_end = LAST_TO_LDS ? nst : nst - 1;
  for (int s = 1; s < nmid_end; ++s) {
    const cplx_t<T>* t = twl + (f.tw[s] - f.tw0);
    CRV_RADIX_SWITCH(f.R[s], (stage_mid<T, R_, INV, true, E>(lds, S, cshift, L, f.Ns[s], sd.ns[s], t)));
  }
  if (!LAST_TO_LDS) {
    const cplx_t<T>* t = twl + (f.tw[nst - 1] - f.tw0);
    CRV_RADIX_SWITCH(f.R[nst - 1], (stage_last<T, R_, INV, JFAST, E>(lds, S, cshift, L, sd.lr[nst - 1], t, store)));
  }
}

// Whole two-stage transform of a pencil-fast tile whose first stage makes
// its inputs with pre / gen (stage_first_pre) and whose last stage stores.
template <typename T, bool INV, int RF, int RL, typename Pre, typename Gen, typename Store>
__device__ __forceinline__ void transform_tile_pre(cplx_t<T>* lds, int S, int cshift, const Fft1D& f,
                                                   const StageDiv& sd, const cplx_t<T>* twl, Pre pre, Gen gen,
                                                   Store store) {
  static_assert(RF > 0, "two-stage lengths only");
  constexpr int E = ept<T, RF, false>();
  stage_first_pre<T, RF, INV, E>(lds, S, cshift, f.L, pre, gen);
  __syncthreads();
  stage_last<T, RL, INV, false, E>(lds, S, cshift, f.L, sd.lr[1], twl + (f.tw[1] - f.tw0), store);
}

// Whole transform of a tile already in LDS in natural order (lds[e*S + c]),
// left there in natural order; every stage runs LDS -> LDS.  The caller
// synchronises before; the tile is complete on return.
template <typename T, bool INV, int RF, int RL, bool H>
__device__ __forceinline__ void transform_tile_lds(cplx_t<T>* lds, int S, int cshift, const Fft1D& f,
                                                   const StageDiv& sd, const cplx_t<T>* twl) {
  constexpr int E = ept<T, RF, H>();
  const int L = f.L;
  if constexpr (RF > 0) {
    stage_mid<T, RF, INV, false, E>(lds, S, cshift, L, 1, sd.ns[0], twl);
    stage_mid<T, RL, INV, true, E>(lds, S, cshift, L, f.Ns[1], sd.ns[1], twl + (f.tw[1] - f.tw0));
    return;
  }
  // general lengths: every stage with its twiddles (those of the first stage
  // are all 1), one radix dispatch per stage
  for (int s = 0; s < f.nst; ++s) {
    const cplx_t<T>* t = twl + (f.tw[s] - f.tw0);
    CRV_RADIX_SWITCH(f.R[s], (stage_mid<T, R_, INV, true, E>(lds, S, cshift, L, f.Ns[s], sd.ns[s], t)));
  }
}

// p (inside the dynamic LDS block base) rounded up to a 16-byte boundary,
// kept an LDS pointer: a round trip through an integer would make it a
// generic one, whose loads are flat loads that the compiler waits for
// together with the global loads in flight
template <typename U, typename V>
__device__ __forceinline__ U* lds_align16(unsigned char* base, V* p) {
  const unsigned off = (unsigned)(reinterpret_cast<unsigned char*>(p) - base);
  return reinterpret_cast<U*>(base + ((off + 15u) & ~15u));
}

template <typename T>
__device__ __forceinline__ void copy_to_lds(cplx_t<T>* dst, const cplx_t<T>* src, int n) {
  for (int i = threadIdx.x; i < n; i += (int)blockDim.x) dst[i] = src[i];
}

// ---- x pass: rows ----------------------------------------------------------
template <typename T>
struct RowArgs {
  T* dst;              // real view (output of C2R / row base of R2C output)
  const T* src;        // real view input (R2C) or complex input buffer (C2R)
  const T* add[3];     // C2R: real grid added to the output of batch entry b: add[b % nadd]
  int nadd;
  long long bstride;   // elements between the grids of a batch (blockIdx.y)
  Fft1D f;             // length M = nxp/2 (even) or nxp (odd)
  StageDiv sd;
  const cplx_t<T>* tw;
  const cplx_t<T>* wpost;  // exp(-2 pi i k / nxp), k = 0..nxp/2 (even only)
  int nrows, rnxp, nxp, cnxp;
  int xeven;
  int pshift;          // rows per block = 1 << pshift
  int tile;            // complex elements of the tile region
  int rev;             // walk the tiles backwards (xcd_remap)
  FastDiv divK;        // by cnxp
  FastDiv divL;        // by f.L
  T scale;
};

// Row tiles move between HBM and LDS with lanes running along the rows
// (fully coalesced), and every stage of the transform runs LDS -> LDS.  The
// tile region holds max((P+1) L, P cnxp) complex: the transform's
// lds[e*S + c] layout, or, for a C2R, the staged half-spectrum rows.
template <typename T, int RF, int RL, bool H>
__global__ __launch_bounds__(kThreads) void k_rows_r2c(RowArgs<T> a) {
  using C = cplx_t<T>;
  constexpr int E = ept<T, RF, H>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int P = 1 << a.pshift, S = P + 1;
  const int L = a.f.L, K = a.cnxp;
  C* twl = lds + a.tile;
  C* wpl = twl + a.f.twn;
  const int row0 = xcd_remap(blockIdx.x, gridDim.x, a.rev) << a.pshift;
  const int nr = min(P, a.nrows - row0);
  const size_t boff = (size_t)blockIdx.y * a.bstride;
  const T* src0 = a.src + boff + (size_t)row0 * a.rnxp;
  T* dst0 = a.dst + boff;
  // rows -> registers -> tile (natural order), scaled
  const int PL = P * L;
  C v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    v[q] = cmake<C>((T)0, (T)0);
    if (idx < PL) {
      const int c = fdiv(idx, a.divL), e = idx - c * L;
      if (c < nr) {
        const T* row = src0 + (size_t)c * a.rnxp;
        // the input cube is read once: streaming loads keep it from
        // displacing the spectrum this pass writes (the next pass's input)
        // in the memory-side cache
        if (a.xeven) {
          v[q] = load_pair_nt(row + 2 * e);
        } else {
          v[q] = cmake<C>(__builtin_nontemporal_load(row + e), (T)0);
        }
      }
    }
  }
  // the twiddles go to LDS while the rows are in flight
  copy_to_lds<T>(twl, a.tw + a.f.tw0, a.f.twn);
  if (a.xeven) copy_to_lds<T>(wpl, a.wpost, K);
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    if (idx < PL) {
      const int c = fdiv(idx, a.divL), e = idx - c * L;
      lds[e * S + c] = cscale(v[q], a.scale);
    }
  }
  __syncthreads();
  transform_tile_lds<T, false, RF, RL, H>(lds, S, a.pshift, a.f, a.sd, twl);
  __syncthreads();
  // post-process and store cnxp complex per row.  Even lengths: each thread
  // reads the LDS values of four of its elements before combining and
  // storing them (clamped indexes, which stay inside the tile, keep the
  // reads free of branches so that they are in flight together).
  const int PK = P * K;
  if (a.xeven) {
    const int M = L;
    constexpr int U = 4;
    for (int i0 = threadIdx.x; i0 < PK; i0 += U * kThreads) {
      C za[U], zb[U], w[U];
      int cu[U], ku[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int idx = min(i0 + u * kThreads, PK - 1);
        const int c = fdiv(idx, a.divK), k = idx - c * K;
        cu[u] = c;
        ku[u] = k;
        za[u] = lds[(k == M ? 0 : k) * S + c];
        zb[u] = lds[(k == 0 ? 0 : M - k) * S + c];
        w[u] = wpl[k];
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        if (i0 + u * kThreads >= PK || cu[u] >= nr) continue;
        const C zc = cconj(zb[u]);
        const C d = csub(za[u], zc);
        const C t = cmul(w[u], d);
        // X = 0.5 (za + zb) - 0.5 i t
        const C X = cmake<C>((T)0.5 * (za[u].x + zc.x + t.y), (T)0.5 * (za[u].y + zc.y - t.x));
        reinterpret_cast<C*>(dst0 + (size_t)(row0 + cu[u]) * a.rnxp)[ku[u]] = X;
      }
    }
  } else {
    for (int idx = threadIdx.x; idx < PK; idx += kThreads) {
      const int c = fdiv(idx, a.divK), k = idx - c * K;
      if (c >= nr) continue;
      reinterpret_cast<C*>(dst0 + (size_t)(row0 + c) * a.rnxp)[k] = lds[k * S + c];
    }
  }
}

template <typename T, int RF, int RL, bool H>
__global__ __launch_bounds__(kThreads) void k_rows_c2r(RowArgs<T> a) {
  using C = cplx_t<T>;
  constexpr int E = ept<T, RF, H>();
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int P = 1 << a.pshift, S = P + 1;
  const int L = a.f.L, K = a.cnxp;
  C* twl = lds + a.tile;
  C* wpl = twl + a.f.twn;
  copy_to_lds<T>(twl, a.tw + a.f.tw0, a.f.twn);
  if (a.xeven) copy_to_lds<T>(wpl, a.wpost, K);
  const int row0 = xcd_remap(blockIdx.x, gridDim.x, a.rev) << a.pshift;
  const int nr = min(P, a.nrows - row0);
  const size_t boff = (size_t)blockIdx.y * a.bstride;
  const T* src0 = a.src + boff + (size_t)row0 * a.rnxp;
  T* dst0 = a.dst + boff + (size_t)row0 * a.rnxp;
  // the added grids are separate allocations, not members of the batch
  const T* add = a.nadd ? a.add[blockIdx.y % a.nadd] + (size_t)row0 * a.rnxp : nullptr;
  // half-spectrum rows -> registers -> staging rows lds[c*KS + k]; the
  // two-stage kernels read the rows across lanes (one row per lane), so
  // their stride is odd
  const int KS = (RF > 0) ? (K | 1) : K;
  const int PK = P * K, PL = P * L;
  C v[E];
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    if (idx < PK) {
      const int c = fdiv(idx, a.divK), k = idx - c * K;
      v[q] = (c < nr) ? reinterpret_cast<const C*>(src0 + (size_t)c * a.rnxp)[k] : cmake<C>((T)0, (T)0);
    }
  }
  // rows of the added grid (even lengths)
  C w[E];
  auto load_add = [&]() {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      w[q] = cmake<C>((T)0, (T)0);
      if (a.xeven && add && idx < PL) {
        const int c = fdiv(idx, a.divL), n = idx - c * L;
        if (c < nr) {
          const T* p2 = add + (size_t)c * a.rnxp + 2 * n;
          w[q] = cmake<C>(__builtin_nontemporal_load(p2), __builtin_nontemporal_load(p2 + 1));
        }
      }
    }
  };
  // FP32: loaded before the transform, so that their latency hides behind
  // it; the kernel stays within 128 registers, at the four waves per SIMD
  // its LDS allows.  FP64 rows take twice the registers and load after the
  // transform, which measures faster for them.
  constexpr bool early_add = sizeof(T) == 4;
  if constexpr (early_add) load_add();
#pragma unroll
  for (int q = 0; q < E; ++q) {
    const int idx = threadIdx.x + q * kThreads;
    if (idx < PK) {
      if constexpr (RF > 0) {
        const int c = fdiv(idx, a.divK), k = idx - c * K;
        lds[c * KS + k] = v[q];
      } else {
        lds[idx] = v[q];
      }
    }
  }
  __syncthreads();
  // Z[k] of the half-length complex transform (or the full Hermitian row)
  // from staged row c
  auto zval = [&](int c, int k) -> C {
    const C* X = lds + c * KS;
    if (a.xeven) {
      C xa = X[k], xb = X[L - k];
      if (k == 0) { xa.y = (T)0; xb.y = (T)0; }
      xb = cconj(xb);
      const C d = csub(xa, xb);
      const C t = cmulc(d, wpl[k]);  // d * exp(+2 pi i k / nxp)
      // Z = (xa + xb) + i t
      return cmake<C>(xa.x + xb.x - t.y, xa.y + xb.y + t.x);
    }
    C z = (k < K) ? X[k] : cconj(X[L - k]);
    if (k == 0) z.y = (T)0;
    return z;
  };
  if constexpr (RF > 0) {
    // the first stage reads Z from the staged rows; its output replaces them
    stage_first_lds<T, RF, true, E>(lds, S, a.pshift, L, zval);
    stage_mid<T, RL, true, true, E>(lds, S, a.pshift, L, a.f.Ns[1], a.sd.ns[1], twl + (a.f.tw[1] - a.f.tw0));
  } else {
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      if (idx < PL) {
        const int c = fdiv(idx, a.divL), k = idx - c * L;
        v[q] = zval(c, k);
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      if (idx < PL) {
        const int c = fdiv(idx, a.divL), k = idx - c * L;
        lds[k * S + c] = v[q];
      }
    }
    __syncthreads();
    transform_tile_lds<T, true, RF, RL, H>(lds, S, a.pshift, a.f, a.sd, twl);
  }
  __syncthreads();
  // tile -> rows, scaled, plus the added grid
  if (a.xeven) {
    if constexpr (!early_add) load_add();
    // the real grid is the transform's final output, not read by the next
    // pass: streaming stores leave the memory-side cache to the passes
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      if (idx < PL) {
        const int c = fdiv(idx, a.divL), n = idx - c * L;
        if (c < nr) {
          const C o = cadd(cscale(lds[n * S + c], a.scale), w[q]);
          T* p2 = dst0 + (size_t)c * a.rnxp + 2 * n;
          __builtin_nontemporal_store(o.x, p2);
          __builtin_nontemporal_store(o.y, p2 + 1);
        }
      }
    }
  } else {
    T wr[E];
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      wr[q] = (T)0;
      if (add && idx < PL) {
        const int c = fdiv(idx, a.divL), n = idx - c * L;
        if (c < nr) wr[q] = __builtin_nontemporal_load(add + (size_t)c * a.rnxp + n);
      }
    }
#pragma unroll
    for (int q = 0; q < E; ++q) {
      const int idx = threadIdx.x + q * kThreads;
      if (idx < PL) {
        const int c = fdiv(idx, a.divL), n = idx - c * L;
        if (c < nr) __builtin_nontemporal_store(lds[n * S + c].x * a.scale + wr[q], dst0 + (size_t)c * a.rnxp + n);
      }
    }
  }
}

// ---- y / z passes: strided complex pencils ---------------------------------
template <typename T>
struct ColArgs {
  cplx_t<T>* data;
  Fft1D f;
  StageDiv sd;
  const cplx_t<T>* tw;
  long long estride;   // element stride of a pencil (complex units)
  long long ostride;   // stride of the outer index
  int ncols;           // number of x columns (cnxp)
  int ntx;             // x tiles per outer index
  int cshift;          // TX = 1 << cshift
  long long bstride;   // elements between the grids of a batch (blockIdx.y)
  int rev;             // walk the tiles backwards (xcd_remap)
  // SYN (z inverse of separable covariance spectra whose prior planes the
  // posterior did not write): where band[e] is 0 the element of column x,
  // y row o is lam[o cnxp + x] S[6 e + q], covariance q = q0 + blockIdx.y;
  // with syn_rows the band's planes are generated as well (CovGen)
  const float* syn_lam;
  const float* syn_S;
  const int* syn_band;
  int syn_q0;
  const int* syn_bidx;
  const float* syn_rows;
  int syn_nband, syn_rw, syn_n, syn_j0, syn_noqueue;
  int syn_nrows;       // y rows of the launch (persistent SYN blocks walk ntx x nrows tiles)
  EigQueueGeo syn_geo;
  // complex per-column factor outside the band in place of syn_lam (the y
  // inverse transform of lam, [y row][column]; CovGen::lamy)
  const float2* syn_lamy;
};

// W: 512-thread blocks (tiles of up to 8192 complex FP32 values), so that
// the pencil tiles of the two-stage lengths can be 16 columns (128 B) wide.
// The generating z inverse (SYN) with its plane tables takes up to 64 KB of
// LDS, so two of its blocks share a CU: it may hold 128 VGPRs (4 waves per
// SIMD) without spilling.
constexpr int cols_waves_per_eu(bool wide, bool syn, int L) { return wide && syn ? 4 : 1; }

template <typename T, bool INV, int RF, int RL, bool H, bool SYN = false, bool W = false>
__global__ __attribute__((amdgpu_flat_work_group_size(1, W ? 2 * kThreads : kThreads),
                          amdgpu_waves_per_eu(cols_waves_per_eu(W, SYN, RF * RL)))) void k_cols(ColArgs<T> a) {
  using C = cplx_t<T>;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int TX = 1 << a.cshift;
  C* twl = lds + a.f.L * TX;
  copy_to_lds<T>(twl, a.tw + a.f.tw0, a.f.twn);
  if constexpr (SYN) {
    // Persistent blocks: the tables of this covariance (S, NaN inside the
    // band; with the band generated too, each plane's row index and the band
    // planes' rows) are staged in LDS once, then the block walks the tiles
    // vb = blockIdx.x + i gridDim.x (gridDim.x a multiple of the 8 XCDs, so
    // that xcd_remap keeps each XCD on a contiguous range).  Each tile goes
    // through the transform as a z inverse pass's does, its first stage
    // generating the elements where that one loads them (the lanes of a wave
    // share a few planes) and its last stage storing to HBM.
    float* synS = reinterpret_cast<float*>(twl + a.f.twn);
    int* synB = reinterpret_cast<int*>(synS + a.f.L);
    // band rows at a 16-byte boundary (read as float4)
    float* synT = lds_align16<float>(smem_raw, synB + a.f.L);
    const int q = a.syn_q0 + blockIdx.y;
    for (int e = threadIdx.x; e < a.f.L; e += (int)blockDim.x)
      synS[e] = a.syn_band[e] ? __builtin_nanf("") : a.syn_S[6 * e + q];
    if (a.syn_rows) {
      for (int e = threadIdx.x; e < a.f.L; e += (int)blockDim.x) synB[e] = a.syn_bidx[e];
      const float* rq = a.syn_rows + (size_t)q * a.syn_nband * a.syn_rw;
      for (int t = threadIdx.x; t < a.syn_nband * a.syn_rw; t += (int)blockDim.x) synT[t] = rq[t];
    }
    const int n = a.syn_n;
    const int ntiles = a.ntx * a.syn_nrows;
    for (int vb = blockIdx.x; vb < ntiles; vb += gridDim.x) {
      // the column shift, opaque in each tile: otherwise the element and LDS
      // offsets, the same for every tile, are hoisted out of the loop and
      // held in registers across it
      int csh = a.cshift;
      asm volatile("" : "+v"(csh));
      const int txs = 1 << csh;
      const int tile = xcd_remap(vb, ntiles, a.rev);
      const int o = tile / a.ntx, tt = tile - o * a.ntx;
      const int x0 = tt * txs;
      const int nc = min(txs, a.ncols - x0);
      C* base = a.data + (size_t)blockIdx.y * a.bstride + (size_t)o * a.ostride + x0;
      // where the block size is a multiple of TX, the first stage gives each
      // thread butterflies of one column, c = threadIdx.x mod TX: its lam and
      // lam's y transform are loaded into registers with the band reads
      // (otherwise each element loads them)
      const bool percol = txs <= (int)blockDim.x;
      const size_t lrow = (size_t)o * a.ncols + x0;
      const size_t li = lrow + min((int)threadIdx.x & (txs - 1), nc - 1);
      const float cl0 = a.syn_lam[li];
      const float2 cly0 = a.syn_lamy ? a.syn_lamy[li] : make_float2(0.0f, 0.0f);
      auto col_lam = [&](int c) { return percol ? cl0 : a.syn_lam[lrow + c]; };
      auto col_lamy = [&](int c) { return percol ? cly0 : a.syn_lamy[lrow + c]; };
      // the previous tile's last stage has read the LDS tile
      __syncthreads();
      // an element of a band plane that the generator does not make itself
      // (no tables here, or the wavenumber went to the FP64 solve) is read
      auto band_read = [&](int c, int e) -> bool {
        if (!a.syn_rows) return true;
        if (a.syn_noqueue) return false;
        const float4 h = reinterpret_cast<const float4*>(synT + synB[e] * a.syn_rw)[0];
        return eig_queued(a.syn_geo, col_lam(c), h.y, h.z, h.w, x0 + c, a.syn_j0 + o, e);
      };
      // the reads first, all in flight together; then the generated values
      auto pre = [&](int c, int e) -> C {
        if (c < nc && __builtin_isnan(synS[e]) && band_read(c, e)) return base[(size_t)e * a.estride + c];
        return cmake<C>((T)0, (T)0);
      };
      auto gen = [&](int c, int e, C v) -> C {
        if (c >= nc) return cmake<C>((T)0, (T)0);
        const float sv = synS[e];
        if (!__builtin_isnan(sv)) {
          if (a.syn_lamy) {
            const float2 ly = col_lamy(c);
            return cmake<C>((T)(ly.x * sv), (T)(ly.y * sv));
          }
          return cmake<C>((T)(col_lam(c) * sv), (T)0);
        }
        if (band_read(c, e)) return v;
        // a band plane: the eigen-basis posterior's covariance entry (the
        // rows are padded to kEigMaxN angles; terms p >= n are discarded)
        const float lam = col_lam(c);
        const float4* tb = reinterpret_cast<const float4*>(synT + synB[e] * a.syn_rw);
        const float4 h = tb[0];   // S_q, cond(D), l_max, l_min
        float w[kEigMaxN], gr[kEigMaxN], gi[kEigMaxN];
#pragma unroll
        for (int p = 0; p < kEigMaxN; ++p) {
          const float4 g = p < n ? tb[1 + p] : make_float4(0.0f, 0.0f, 0.0f, 0.0f);
          w[p] = eig_weight(lam, g.x);
          gr[p] = g.y;
          gi[p] = g.z;
        }
        const float2 g = eig_cov_entry<kEigMaxN>(n, q < 3, lam * h.x, lam * lam, w, gr, gi);
        return cmake<C>((T)g.x, (T)g.y);
      };
      auto store = [&](int c, int e, C v) {
        if (c < nc) base[(size_t)e * a.estride + c] = v;
      };
      if constexpr (RF > 0) {
        transform_tile_pre<T, INV, RF, RL>(lds, txs, csh, a.f, a.sd, twl, pre, gen, store);
      } else {
        auto load = [&](int c, int e) -> C { return gen(c, e, pre(c, e)); };
        transform_tile<T, INV, false, false, RF, RL, H>(lds, txs, csh, a.f, a.sd, twl, load, store);
      }
    }
    return;
  }
  const int tile = xcd_remap(blockIdx.x, gridDim.x, a.rev);
  const int o = tile / a.ntx, t = tile - o * a.ntx;
  const int x0 = t * TX;
  const int nc = min(TX, a.ncols - x0);
  C* base = a.data + (size_t)blockIdx.y * a.bstride + (size_t)o * a.ostride + x0;
  auto load = [&](int c, int e) -> C {
    if (c >= nc) return cmake<C>((T)0, (T)0);
    return base[(size_t)e * a.estride + c];
  };
  auto store = [&](int c, int e, C v) {
    if (c < nc) base[(size_t)e * a.estride + c] = v;
  };
  transform_tile<T, INV, false, false, RF, RL, H>(lds, TX, a.cshift, a.f, a.sd, twl, load, store);
}

// ---- y inverse of the covariances' band planes (crv_invert) ------------------
// The covariance spectra of the inverted band are the eigen-basis posterior's
// entries lam S_q - lam^2 sum_p w_p G_pq, generated from the plane's row of
// tables; outside the band they are the separable prior lam(i, j) |ct(k)|
// Sigma0[q].  The y and z inverse transforms commute, so the y pass runs on
// the band's planes only (generating them in its first stage), and the z
// inverse pass then generates the other planes from the y transform of lam
// (CovGen::lamy): 40 % of a y pass per covariance instead of a whole one.
struct BandYArgs {
  float2* data;        // covariance grid of batch entry 0 (blockIdx.y: q = q0 + blockIdx.y)
  long long bstride;   // complex elements between the grids of the batch
  Fft1D f;
  StageDiv sd;
  const float2* tw;
  long long pstride;   // cnxp nyp
  int ncols, ntx, cshift, rev;
  int nband, rw, n, q0, noqueue;
  const int* bplanes;
  const float* lam;    // [nyp][cnxp]
  const float* rows;   // rows + (q nband + b) rw: S_q, cond(D), l_max, l_min, then (l_p, Re G, Im G, 0)
  EigQueueGeo geo;
};

template <int RF, int RL, bool W>
__global__ __attribute__((amdgpu_flat_work_group_size(1, W ? 2 * kThreads : kThreads))) void k_cov_band_y(
    BandYArgs a) {
  using C = float2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int TX = 1 << a.cshift;
  C* twl = lds + a.f.L * TX;
  float4* row = lds_align16<float4>(smem_raw, twl + a.f.twn);
  copy_to_lds<float>(twl, a.tw + a.f.tw0, a.f.twn);
  const int tile = xcd_remap(blockIdx.x, gridDim.x, a.rev);
  const int b = tile / a.ntx, t = tile - b * a.ntx;
  const int k = a.bplanes[b];
  const int q = a.q0 + blockIdx.y;
  const float* rq = a.rows + ((size_t)q * a.nband + b) * a.rw;
  // the row, padded to kEigMaxN angles (the terms p >= n are discarded)
  for (int i = threadIdx.x; i < 4 + 4 * kEigMaxN; i += (int)blockDim.x)
    reinterpret_cast<float*>(row)[i] = i < a.rw ? rq[i] : 0.0f;
  __syncthreads();
  const int x0 = t * TX, nc = min(TX, a.ncols - x0);
  C* base = a.data + (size_t)blockIdx.y * a.bstride + (size_t)k * a.pstride + x0;
  // the plane's row, the same for the whole block, in scalar registers
  auto uni = [](float x) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(x))); };
  const float4 h4 = row[0];
  const float4 h = make_float4(uni(h4.x), uni(h4.y), uni(h4.z), uni(h4.w));   // S_q, cond(D), l_max, l_min
  float lp[kEigMaxN], gr[kEigMaxN], gi[kEigMaxN];
#pragma unroll
  for (int p = 0; p < kEigMaxN; ++p) {
    const float4 g = row[1 + p];
    lp[p] = uni(g.x);
    gr[p] = uni(g.y);
    gi[p] = uni(g.z);
  }
  const int n = a.n;
  // lam of an element (columns past the grid's read the last column's and
  // are zeroed below), then its entry
  auto pre = [&](int c, int j) -> float { return a.lam[(size_t)j * a.ncols + x0 + min(c, nc - 1)]; };
  auto gen = [&](int c, int j, float lam) -> C {
    if (c < nc && !a.noqueue && eig_queued(a.geo, lam, h.y, h.z, h.w, x0 + c, j, k))
      return base[(size_t)j * a.ncols + c];   // the FP64 solve's value
    float w[kEigMaxN];
#pragma unroll
    for (int p = 0; p < kEigMaxN; ++p) w[p] = eig_weight(lam, lp[p]);
    const float2 v = eig_cov_entry<kEigMaxN>(n, q < 3, lam * h.x, lam * lam, w, gr, gi);
    return c < nc ? v : make_float2(0.0f, 0.0f);
  };
  auto store = [&](int c, int j, C v) {
    if (c < nc) base[(size_t)j * a.ncols + c] = v;
  };
  // Every thread generates elements i = j TX + c (natural order) into LDS,
  // all their lam loads first; the first stage then runs LDS -> LDS.  The
  // first stage's own mapping would leave the generation, the kernel's
  // main arithmetic, to the threads that hold butterflies (288 of 512 for
  // a first radix of 25): 0.062 -> 0.058 ms at 450 x 450 x 320.
  const int lt = a.f.L << a.cshift, cm = TX - 1;
  constexpr int NB = Ept<float>::value;   // tiles hold at most NB blockDim elements
  float lv[NB];
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int i = (int)threadIdx.x + u * (int)blockDim.x;
    if (i < lt) lv[u] = pre(i & cm, i >> a.cshift);
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int i = (int)threadIdx.x + u * (int)blockDim.x;
    if (i < lt) lds[i] = gen(i & cm, i >> a.cshift, lv[u]);
  }
  __syncthreads();
  constexpr int E = ept<float, RF, false>();
  stage_mid<float, RF, true, false, E>(lds, TX, a.cshift, a.f.L, 1, a.sd.ns[0], twl);
  stage_last<float, RL, true, false, E>(lds, TX, a.cshift, a.f.L, a.sd.lr[1], twl + (a.f.tw[1] - a.f.tw0), store);
}

// the element off bytes after p
__device__ __forceinline__ float2& at_off(float2* p, unsigned off) {
  return *reinterpret_cast<float2*>(reinterpret_cast<char*>(p) + off);
}

// ---- z inverse pass with the next tile's loads in flight ---------------------
// FP32 two-stage z lengths.  Persistent blocks walk the tiles vb =
// blockIdx.x + i gridDim.x (gridDim.x a multiple of the 8 XCDs).  The
// elements of a block's next tile are loaded into registers in natural
// order (element i = e TX + c of the tile in thread i mod blockDim) while
// its current tile goes through both stages and its stores, so that their
// latency hides behind that work instead of being waited for once per tile;
// at the next tile they are written to LDS in the lds[e TX + c] layout and
// the first stage runs LDS -> LDS.  Each element goes through the same
// butterflies as in k_cols, so the results are bit-equal.  SYN: the
// covariances' z inverse after the band's y pass (no band rows): the planes
// outside the band (S finite) are generated as f(column) S[e], f = lam's y
// transform (syn_lamy) or lam, and only the band's are read.
// elements per thread held for the next tile: a 16-column tile over the
// block size the plan gives it (512 threads above 4096 values); the
// launcher uses the kernel only for tiles within that
constexpr int zinv_nat(int L) { return L * 16 > 4096 ? (L * 16 + 511) / 512 : (L * 16 + 255) / 256; }

template <int RF, int RL, bool SYN>
__global__ __attribute__((amdgpu_flat_work_group_size(1, 2 * kThreads), amdgpu_waves_per_eu(4))) void k_zinv_pf(
    ColArgs<float> a) {
  using C = float2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  constexpr int L = RF * RL;
  C* twl = lds + L * (1 << a.cshift);
  float* synS = reinterpret_cast<float*>(twl + a.f.twn);
  copy_to_lds<float>(twl, a.tw + a.f.tw0, a.f.twn);
  if constexpr (SYN) {
    const int q = a.syn_q0 + blockIdx.y;
    for (int e = threadIdx.x; e < L; e += (int)blockDim.x)
      synS[e] = a.syn_band[e] ? __builtin_nanf("") : a.syn_S[6 * e + q];
  }
  __syncthreads();
  constexpr int E = ept<float, RF, false>();
  const int NT = (int)blockDim.x;
  const int ntiles = a.ntx * a.syn_nrows;
  C* const gbase = a.data + (size_t)blockIdx.y * a.bstride;
  // byte offsets from the tile's first column fit 32 bits (the launcher
  // checks the grid's size), so each address is one register over the
  // block-uniform base
  const unsigned es8 = (unsigned)a.estride * 8u;
  struct Tile { C* base; size_t frow; int x0, nc; };
  auto tile_of = [&](int vb, int csh) {
    const int tile = xcd_remap(vb, ntiles, a.rev);
    const int o = tile / a.ntx;
    Tile t;
    t.x0 = (tile - o * a.ntx) << csh;
    t.nc = min(1 << csh, a.ncols - t.x0);
    t.base = gbase + (size_t)o * a.ostride + t.x0;
    t.frow = (size_t)o * a.ncols + t.x0;
    return t;
  };
  // the next tile's elements (zero where generated or past the grid's
  // columns) and, for SYN, the factor of this thread's column c =
  // threadIdx.x mod TX (the block size is a multiple of TX)
  constexpr int NAT = zinv_nat(L);
  C pv[NAT];
  float2 fv = make_float2(0.0f, 0.0f);
  auto issue = [&](const Tile& t, int csh) {
    const int cm = (1 << csh) - 1, lt = L << csh;
#pragma unroll
    for (int u = 0; u < NAT; ++u) {
      const int i = (int)threadIdx.x + u * NT;
      if (i < lt) {
        const int e = i >> csh, c = i & cm;
        const bool rd = c < t.nc && (!SYN || __builtin_isnan(synS[e]));
        pv[u] = rd ? at_off(t.base, (unsigned)e * es8 + (unsigned)c * 8u) : make_float2(0.0f, 0.0f);
      }
    }
    if constexpr (SYN) {
      const size_t fi = t.frow + min((int)threadIdx.x & cm, t.nc - 1);
      fv = a.syn_lamy ? a.syn_lamy[fi] : make_float2(a.syn_lam[fi], 0.0f);
    }
  };
  int vb = blockIdx.x;
  if (vb < ntiles) issue(tile_of(vb, a.cshift), a.cshift);
  for (; vb < ntiles; vb += (int)gridDim.x) {
    // the column shift, opaque in each tile (otherwise the offsets, the same
    // for every tile, are hoisted out of the loop and held in registers)
    int csh = a.cshift;
    asm volatile("" : "+v"(csh));
    const int txs = 1 << csh, cm = txs - 1, lt = L << csh;
    const Tile t = tile_of(vb, csh);
    // the previous tile's last stage has read the LDS tile
    __syncthreads();
#pragma unroll
    for (int u = 0; u < NAT; ++u) {
      const int i = (int)threadIdx.x + u * NT;
      if (i < lt) {
        C v = pv[u];
        if constexpr (SYN) {
          const float sv = synS[i >> csh];
          if ((i & cm) < t.nc && !__builtin_isnan(sv))
            v = a.syn_lamy ? make_float2(fv.x * sv, fv.y * sv) : make_float2(fv.x * sv, 0.0f);
        }
        lds[i] = v;
      }
    }
    if (vb + (int)gridDim.x < ntiles) issue(tile_of(vb + (int)gridDim.x, csh), csh);
    __syncthreads();
    stage_mid<float, RF, true, false, E>(lds, txs, csh, L, 1, a.sd.ns[0], twl);
    auto store = [&](int c, int e, C v) {
      if (c < t.nc) at_off(t.base, (unsigned)e * es8 + (unsigned)c * 8u) = v;
    };
    stage_last<float, RL, true, false, E>(lds, txs, csh, L, a.sd.lr[1], twl + (a.f.tw[1] - a.f.tw0), store);
  }
}

// ---- z pass with the posterior of the means (crv_invert) ---------------------
// One tile is TX consecutive x columns of one y row, whole z pencils, for
// the n data cubes and the 3 means (FP32, eigen-basis records).  The pencil
// sets of a tile go through LDS one after another: all threads load a set
// (lanes along the columns) and write it to LDS in natural order; the loads
// of the next two sets are in flight meanwhile.  A data set's elements
// outside the inverted band are their own residual and go straight back to
// HBM, the band's stay in registers; the mean sets are transformed into three
// slots.  The band's wavenumbers then get eig_mean_update (posterior means
// into the slots, residuals to HBM), and the slots are inverse transformed
// and stored.  The blocks are persistent: the loads of a block's next tile
// are issued before the posterior and the inverse transforms of its current
// one.  Each element sees the arithmetic of the separate z forward pass,
// k_posterior_eig and z inverse pass, so the results are bit-equal; the
// spectra of the means and the data never make the HBM round trip between
// them.
constexpr int kZThreads = 512;
constexpr int kZQ = 2;    // band wavenumbers per thread at most
constexpr int kZPF = 5;   // elements per thread of a pencil set at most

struct ZPostArgs {
  float2* data[kEigMaxN];
  float2* mean[3];
  Fft1D f;
  StageDiv sd;
  const float2* tw;
  int estride;         // cnxp nyp
  int ncols, ntx, ntiles, cshift, rev;
  int n, nband;
  const float* lam;    // [nyp][cnxp]
  const float* erec;   // plane records, kEigRec floats per plane
  const int* bidx;     // [nzp] row among the band planes, -1 outside the band
  const int* bplanes;  // [nband] band planes
  unsigned long long* stats;
};

template <int RF, int RL>
__global__ __attribute__((amdgpu_flat_work_group_size(1, kZThreads), amdgpu_waves_per_eu(4))) void k_zpost(
    ZPostArgs a) {
  using C = float2;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* lds = reinterpret_cast<C*>(smem_raw);
  const int TX = 1 << a.cshift, L = a.f.L, LT = L * TX;
  C* twl = lds + 3 * LT;
  int* sbidx = reinterpret_cast<int*>(twl + a.f.twn);
  int* sbpl = sbidx + L;
  copy_to_lds<float>(twl, a.tw + a.f.tw0, a.f.twn);
  for (int e = threadIdx.x; e < L; e += kZThreads) sbidx[e] = a.bidx[e];
  for (int e = threadIdx.x; e < a.nband; e += kZThreads) sbpl[e] = a.bplanes[e];
  const int nbw = a.nband * TX;
  const int es = a.estride, cm = TX - 1;
  const int nset = a.n + 3;
  const int n = a.n;
  const int OC = 2 * n * n + 22 * n + 12;   // EigLayout<n>::OC
  // the tile's y row and first column, its column count and element offset
  struct Tile { int o, x0, nc; size_t off; };
  auto tile_of = [&](int vb) {
    const int tile = xcd_remap(vb, a.ntiles, a.rev);
    Tile T;
    T.o = tile / a.ntx;
    T.x0 = (tile - T.o * a.ntx) * TX;
    T.nc = min(TX, a.ncols - T.x0);
    T.off = (size_t)T.o * a.ncols + T.x0;
    return T;
  };
  auto set_ptr = [&](const Tile& T, int s) -> C* { return (s < n ? a.data[s] : a.mean[s - n]) + T.off; };
  // element u of this thread in a pencil set: t = threadIdx.x + u kZThreads,
  // column t & cm, plane t >> cshift.  csh: the tile's column shift, made
  // opaque where it is used in a loop: otherwise the element and LDS
  // offsets, the same for every set, are hoisted out of the loops and held
  // in registers across them
  C pa[kZPF], pb[kZPF];
  auto prefetch = [&](C (&pf)[kZPF], const Tile& T, int s, int csh) {
    const C* base = set_ptr(T, s);
    const int cmk = (1 << csh) - 1;
#pragma unroll
    for (int u = 0; u < kZPF; ++u) {
      const int tt = threadIdx.x + u * kZThreads;
      const int c = tt & cmk, e = tt >> csh;
      pf[u] = (tt < LT && c < T.nc) ? base[e * es + c] : make_float2(0.0f, 0.0f);
    }
  };
  C d[kZQ][kEigMaxN];
  auto step = [&](C (&pf)[kZPF], const Tile& T, int s) {
    int csh = a.cshift;
    asm volatile("" : "+v"(csh));
    const int txs = 1 << csh, cmk = txs - 1;
    C* slot = lds + (s < n ? 0 : s - n) * LT;
#pragma unroll
    for (int u = 0; u < kZPF; ++u) {
      const int tt = threadIdx.x + u * kZThreads;
      if (tt < LT) slot[tt] = pf[u];
    }
    __syncthreads();
    if (s + 2 < nset) prefetch(pf, T, s + 2, csh);
    transform_tile_lds<float, false, RF, RL, false>(slot, txs, csh, a.f, a.sd, twl);
    if (s < n) {
      // outside the band the spectrum is the residual
      C* base = set_ptr(T, s);
      for (int tt = threadIdx.x; tt < LT; tt += kZThreads) {
        const int c = tt & cmk, e = tt >> csh;
        if (c < T.nc && sbidx[e] < 0) base[e * es + c] = slot[tt];
      }
#pragma unroll
      for (int q = 0; q < kZQ; ++q) {
        const int b = threadIdx.x + q * kZThreads;
        const C v = b < nbw ? slot[sbpl[b >> csh] * txs + (b & cmk)] : make_float2(0.0f, 0.0f);
#pragma unroll
        for (int ll = 0; ll < kEigMaxN; ++ll)
          if (ll == s) d[q][ll] = v;
      }
      __syncthreads();
    }
  };
  double cmax = 0.0;
  int vb = blockIdx.x;
  Tile T = tile_of(vb);
  prefetch(pa, T, 0, a.cshift);
  prefetch(pb, T, 1, a.cshift);   // nset >= 4
  while (vb < a.ntiles) {
    for (int s = 0; s < nset; s += 2) {
      step(pa, T, s);
      if (s + 1 < nset) step(pb, T, s + 1);
    }
    // posterior of the band's wavenumbers, one after the other
    int pcs = a.cshift;
    asm volatile("" : "+v"(pcs));
#pragma unroll
    for (int q = 0; q < kZQ; ++q) {
      const int b = threadIdx.x + q * kZThreads;
      const int c = b & ((1 << pcs) - 1);
      if (b < nbw && c < T.nc) {
        const int e = sbpl[b >> pcs];
        const float* er = a.erec + (size_t)e * kEigRec;
        const float lam = a.lam[T.off + c];
        C mu[3], mp[3], res[kEigMaxN];
#pragma unroll
        for (int m = 0; m < 3; ++m) mu[m] = lds[m * LT + (e << pcs) + c];
        eig_mean_update<kEigMaxN>(n, er, lam, mu, d[q], mp, res);
#pragma unroll
        for (int m = 0; m < 3; ++m) lds[m * LT + (e << pcs) + c] = mp[m];
#pragma unroll
        for (int l = 0; l < kEigMaxN; ++l)
          if (l < n) a.data[l][T.off + (size_t)e * es + c] = res[l];
        cmax = fmax(cmax, (double)eig_cond(lam, er[OC], er[OC + 1], er[OC + 2]));
      }
      // keep the two wavenumbers' record loads apart (register pressure)
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
    // the next tile's first two sets load while this one's means are
    // inverse transformed and stored
    const int vn = vb + (int)gridDim.x;
    const Tile Tn = tile_of(vn < a.ntiles ? vn : vb);
    if (vn < a.ntiles) {
      int csh = a.cshift;
      asm volatile("" : "+v"(csh));
      prefetch(pa, Tn, 0, csh);
      prefetch(pb, Tn, 1, csh);
    }
    // means: inverse transform of the slots
#pragma unroll 1
    for (int m = 0; m < 3; ++m) {
      int csh = a.cshift;
      asm volatile("" : "+v"(csh));
      const int txs = 1 << csh, cmk = txs - 1;
      C* slot = lds + m * LT;
      transform_tile_lds<float, true, RF, RL, false>(slot, txs, csh, a.f, a.sd, twl);
      C* base = a.mean[m] + T.off;
      for (int tt = threadIdx.x; tt < LT; tt += kZThreads) {
        const int c = tt & cmk, e = tt >> csh;
        if (c < T.nc) base[e * es + c] = slot[tt];
      }
    }
    // the slots' stores above read LDS that the next tile's first set writes
    __syncthreads();
    vb = vn;
    T = Tn;
  }
  // largest condition estimate of the block
  __shared__ double s_c[kZThreads / 64];
#pragma unroll
  for (int sh = 32; sh > 0; sh >>= 1) cmax = fmax(cmax, __shfl_down(cmax, sh, 64));
  if ((threadIdx.x & 63) == 0) s_c[threadIdx.x >> 6] = cmax;
  __syncthreads();
  if (threadIdx.x == 0) {
    double cm2 = 0.0;
#pragma unroll
    for (int w = 0; w < kZThreads / 64; ++w) cm2 = fmax(cm2, s_c[w]);
    if (cm2 > 0.0) atomicMax(&a.stats[1], (unsigned long long)__double_as_longlong(cm2));
  }
}

// ---- planning ----------------------------------------------------------------
// A length that is one of the two-stage radix pairs uses it; any other
// 2^a 3^b 5^c 7^d length is split into radices 8, 4, 2, 5, 3, 7.
std::vector<int> factorize(int L) {
#define CRV_PAIR_MATCH(A, B) if (L == (A) * (B)) return {A, B};
  CRV_FFT_PAIRS(CRV_PAIR_MATCH)
#undef CRV_PAIR_MATCH
  std::vector<int> r;
  int m = L;
  while (m % 8 == 0) { r.push_back(8); m /= 8; }
  if (m % 4 == 0) { r.push_back(4); m /= 4; }
  if (m % 2 == 0) { r.push_back(2); m /= 2; }
  for (int p : {5, 3, 7}) while (m % p == 0) { r.push_back(p); m /= p; }
  if (m != 1) r.clear();
  return r;
}

bool is_pair(const std::vector<int>& rs) {
#define CRV_PAIR_IS(A, B) if (rs.size() == 2 && rs[0] == (A) && rs[1] == (B)) return true;
  CRV_FFT_PAIRS(CRV_PAIR_IS)
#undef CRV_PAIR_IS
  return false;
}

// the axis runs in a two-stage kernel (the selection of rows_kernel / cols_kernel)
bool axis_has_pair(const Fft1D& f) { return f.nst == 2 && is_pair(std::vector<int>(f.R, f.R + 2)); }

// reverse: the stages of a length without a two-stage kernel run in the
// opposite order (an autotuning choice)
bool plan_1d(int L, Fft1D& f, std::vector<long double>& twr, std::vector<long double>& twi, bool reverse = false) {
  f.L = L;
  f.tw0 = (int)twr.size();
  std::vector<int> rs = (L == 1) ? std::vector<int>{} : factorize(L);
  if (L != 1 && rs.empty()) return false;
  if (reverse && !is_pair(rs)) std::reverse(rs.begin(), rs.end());
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
  f.twn = (int)twr.size() - f.tw0;
  return true;
}

int pow2_floor(int v) { int p = 1; while (p * 2 <= v) p *= 2; return p; }

// complex FP32 values of a 512-thread column tile: twice the 256-thread
// budget, less what keeps the tile, the axis's twiddles and the generated
// z-inverse tables (L + TX floats) within 64 KB of LDS
int wide_budget(const Fft1D& f, int maxe) {
  return std::min(2 * maxe, 65536 / (int)sizeof(float2) - f.twn - (f.L + 64) / 2);
}
int ilog2(int v) { int s = 0; while ((1 << s) < v) ++s; return s; }

// complex elements of a row tile: the transform's (P+1) L, or the staged
// half-spectrum rows P cnxp of a C2R
int row_tile_elems(const FftPlan* p) {
  const int cnxp = p->nxp / 2 + 1;
  return std::max((p->rows_per_block + 1) * p->fx.L, p->rows_per_block * (cnxp | 1));
}

template <typename T>
size_t row_lds_bytes(const FftPlan* p) {
  const int cnxp = p->nxp / 2 + 1;
  return (size_t)(row_tile_elems(p) + p->fx.twn + cnxp) * sizeof(cplx_t<T>);
}

// kernel for an axis plan: the two-stage kernel of its radix pair, or the
// general one
template <typename T>
const void* rows_kernel(const Fft1D& f, bool inv, bool half) {
  if (f.nst == 2) {
#define CRV_PAIR_ROWS(A, B) \
    if (f.R[0] == (A) && f.R[1] == (B)) \
      return inv ? (const void*)k_rows_c2r<T, A, B, false> : (const void*)k_rows_r2c<T, A, B, false>;
    CRV_FFT_PAIRS(CRV_PAIR_ROWS)
#undef CRV_PAIR_ROWS
  }
  if (half) return inv ? (const void*)k_rows_c2r<T, 0, 0, true> : (const void*)k_rows_r2c<T, 0, 0, true>;
  return inv ? (const void*)k_rows_c2r<T, 0, 0, false> : (const void*)k_rows_r2c<T, 0, 0, false>;
}

// z inverse with the prior planes generated (float grids); wide: 512 threads
const void* cols_kernel_syn(const Fft1D& f, bool half, bool wide) {
  if (f.nst == 2) {
#define CRV_PAIR_SYN(A, B)                                                                     \
    if (f.R[0] == (A) && f.R[1] == (B))                                                        \
      return wide ? (const void*)k_cols<float, true, A, B, false, true, true>                  \
                  : (const void*)k_cols<float, true, A, B, false, true>;
    CRV_FFT_PAIRS(CRV_PAIR_SYN)
#undef CRV_PAIR_SYN
  }
  return half ? (const void*)k_cols<float, true, 0, 0, true, true> : (const void*)k_cols<float, true, 0, 0, false, true>;
}

// the prefetching z inverse kernel of a length, or nullptr where there is
// none or where it spills registers to scratch (its registers would then
// wait for the loads it means to keep in flight)
const void* zinv_pf_kernel(const Fft1D& f, bool syn) {
  if (f.nst != 2) return nullptr;
  const void* k = nullptr;
#define CRV_PAIR_ZPF(A, B)                                                                              \
  if (!k && f.R[0] == (A) && f.R[1] == (B))                                                              \
    k = syn ? (const void*)k_zinv_pf<A, B, true> : (const void*)k_zinv_pf<A, B, false>;
  CRV_FFT_PAIRS(CRV_PAIR_ZPF)
#undef CRV_PAIR_ZPF
  if (!k) return nullptr;
  static std::mutex mu;
  static std::unordered_map<const void*, bool> spills;
  std::lock_guard<std::mutex> lock(mu);
  auto it = spills.find(k);
  if (it == spills.end()) {
    hipFuncAttributes at{};
    const bool sp = hipFuncGetAttributes(&at, k) != hipSuccess || at.localSizeBytes > 0;
    it = spills.emplace(k, sp).first;
  }
  return it->second ? nullptr : k;
}

template <typename T, bool INV>
const void* cols_kernel(const Fft1D& f, bool half, bool wide) {
  if (f.nst == 2) {
    if constexpr (std::is_same<T, float>::value) {
#define CRV_PAIR_COLS(A, B)                                                                    \
      if (f.R[0] == (A) && f.R[1] == (B))                                                      \
        return wide ? (const void*)k_cols<T, INV, A, B, false, false, true>                    \
                    : (const void*)k_cols<T, INV, A, B, false>;
      CRV_FFT_PAIRS(CRV_PAIR_COLS)
#undef CRV_PAIR_COLS
    } else {
#define CRV_PAIR_COLS(A, B) \
      if (f.R[0] == (A) && f.R[1] == (B)) return (const void*)k_cols<T, INV, A, B, false>;
      CRV_FFT_PAIRS(CRV_PAIR_COLS)
#undef CRV_PAIR_COLS
    }
  }
  return half ? (const void*)k_cols<T, INV, 0, 0, true> : (const void*)k_cols<T, INV, 0, 0, false>;
}

// Batches: nbatch grids bstride elements apart (real units), one launch.
struct Batch {
  int n = 1;
  long long stride = 0;
  int nadd = 0;
  const void* add[3] = {nullptr, nullptr, nullptr};
};

template <typename T>
int launch_rows(crv_ctx* ctx, FftPlan* p, bool inv, T* dst, const T* src, const Batch& bt, int row_begin,
                int row_end, double scale) {
  const int cnxp = p->nxp / 2 + 1, rnxp = 2 * cnxp;
  const size_t off = (size_t)row_begin * rnxp;
  RowArgs<T> a{};
  a.dst = dst + off; a.src = src + off;
  a.nadd = bt.nadd;
  for (int q = 0; q < bt.nadd; ++q) a.add[q] = (const T*)bt.add[q] + off;
  a.bstride = bt.stride;
  a.f = p->fx; a.sd = make_stage_div(p->fx); a.tw = reinterpret_cast<const cplx_t<T>*>(p->tw);
  a.wpost = a.tw + p->post_off;
  a.nrows = row_end - row_begin; a.rnxp = rnxp; a.nxp = p->nxp; a.cnxp = cnxp; a.xeven = p->xeven ? 1 : 0;
  a.pshift = ilog2(p->rows_per_block);
  a.tile = row_tile_elems(p);
  a.rev = (p->order >> (inv ? 5 : 0)) & 1;
  a.divK = make_fastdiv(cnxp);
  a.divL = make_fastdiv(p->fx.L);
  a.scale = (T)scale;
  const int nb = (a.nrows + p->rows_per_block - 1) / p->rows_per_block;
  void* args[] = {&a};
  ProfScope ps(ctx, inv ? "fft_x_c2r" : "fft_x_r2c");
  CRV_HIP_CHECK(ctx, hipLaunchKernel(rows_kernel<T>(p->fx, inv, p->half[0]), dim3(nb, bt.n), dim3(kThreads), args,
                                     row_lds_bytes<T>(p), ctx->stream));
  return CRV_OK;
}

// LDS of a column tile of TX columns, with the generating z inverse's
// tables (gen != nullptr: S and the planes' row indices; rows: band rows)
size_t cols_lds_bytes(const Fft1D& f, int tx, size_t elem, const CovGen* gen) {
  size_t lds = ((size_t)f.L * tx + f.twn) * elem;
  if (gen) {
    lds += (size_t)f.L * (sizeof(float) + sizeof(int)) + 16;
    if (gen->rows) lds += (size_t)gen->nband * gen->rw * sizeof(float);
  }
  return lds;
}

// y pass over z planes [k0, k1) or z pass over y rows [j0, j1)
template <typename T>
int launch_cols(crv_ctx* ctx, FftPlan* p, bool inv, bool zaxis, cplx_t<T>* data, const Batch& bt, int o0, int o1,
                const CovGen* zs = nullptr) {
  using C = cplx_t<T>;
  const int cnxp = p->nxp / 2 + 1;
  ColArgs<T> a{};
  a.bstride = bt.stride / 2;  // complex units
  a.tw = reinterpret_cast<const C*>(p->tw);
  a.ncols = cnxp;
  int tx;
  if (zaxis) {
    a.f = p->fz; a.estride = (long long)cnxp * p->nyp; a.ostride = cnxp; tx = p->tx_z;
  } else {
    a.f = p->fy; a.estride = cnxp; a.ostride = (long long)cnxp * p->nyp; tx = p->tx_y;
  }
  a.sd = make_stage_div(a.f);
  a.rev = (p->order >> (zaxis ? (inv ? 3 : 2) : (inv ? 4 : 1))) & 1;
  a.data = data + (size_t)o0 * a.ostride;
  a.cshift = ilog2(tx);
  a.ntx = (cnxp + tx - 1) / tx;
  const int nb = a.ntx * (o1 - o0);
  const bool syn = zs && zaxis && inv && std::is_same<T, float>::value;
  const size_t lds = cols_lds_bytes(a.f, tx, sizeof(C), syn ? zs : nullptr);
  static const char* names[4] = {"fft_y_fwd", "fft_y_inv", "fft_z_fwd", "fft_z_inv"};
  ProfScope ps(ctx, syn ? "fft_z_inv_prior" : names[(zaxis ? 2 : 0) + (inv ? 1 : 0)]);
  void* args[] = {&a};
  const bool half = p->half[zaxis ? 2 : 1];
  const bool wide = p->wide[zaxis ? 2 : 1];
  const void* kern = inv ? cols_kernel<T, true>(a.f, half, wide) : cols_kernel<T, false>(a.f, half, wide);
  const int threads = wide ? 2 * kThreads : kThreads;
  a.syn_nrows = o1 - o0;
  if (syn) {
    // the generated elements of row o0 + o use the tables of rows from o0
    a.syn_lam = zs->lam + (size_t)o0 * cnxp;
    a.syn_S = zs->S;
    a.syn_band = zs->band;
    a.syn_q0 = zs->q0;
    a.syn_bidx = zs->bidx;
    a.syn_rows = zs->rows;
    a.syn_nband = zs->nband; a.syn_rw = zs->rw; a.syn_n = zs->n; a.syn_j0 = o0; a.syn_noqueue = zs->noqueue;
    a.syn_lamy = zs->lamy ? zs->lamy + (size_t)o0 * cnxp : nullptr;
    a.syn_geo = EigQueueGeo{zs->rx, zs->ry, zs->rz, zs->lowk2, zs->cond_thr, p->nyp, p->nzp};
  }
  // persistent launches: as many blocks as fit at once (a multiple of 8, the
  // XCDs), at most one per tile
  auto persistent = [&](const void* k, size_t bytes) {
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, threads, bytes) != hipSuccess || per_cu < 1)
      per_cu = 1;
    const int want = std::max(8, (per_cu * ctx->num_cu / std::max(1, bt.n) + 7) / 8 * 8);
    return hipLaunchKernel(k, dim3(std::min(nb, want), bt.n), dim3(threads), args, bytes, ctx->stream);
  };
  // FP32 two-stage z inverse: the kernel with the next tile's loads in
  // flight for the covariances' generating pass (without band rows), and
  // on request for the others, whose reads it measured slower
  if constexpr (std::is_same<T, float>::value) {
    const int zt = ctx->tune[CRV_TUNE_ZINV];
    const size_t lpf = ((size_t)a.f.L * tx + a.f.twn) * sizeof(C) + (syn ? (size_t)a.f.L * sizeof(float) : 0);
    const void* pf = (zaxis && inv && (zt == 1 || (zt == 0 && syn)) && tx <= threads && lpf <= 65536 &&
                      (size_t)a.f.L * tx <= (size_t)zinv_nat(a.f.L) * threads &&
                      (!syn || !zs->rows) && (size_t)a.f.L * a.estride * sizeof(C) < (1ull << 32))
                         ? zinv_pf_kernel(a.f, syn)
                         : nullptr;
    if (pf) {
      CRV_HIP_CHECK(ctx, persistent(pf, lpf));
      return CRV_OK;
    }
  }
  if (syn) {
    CRV_HIP_CHECK(ctx, persistent(cols_kernel_syn(a.f, half, wide), lds));
    return CRV_OK;
  }
  CRV_HIP_CHECK(ctx, hipLaunchKernel(kern, dim3(nb, bt.n), dim3(wide ? 2 * kThreads : kThreads), args, lds,
                                     ctx->stream));
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
int launch_r2c(crv_ctx* ctx, FftPlan* p, T* dst, const T* src, double scale, const Batch& bt) {
  using C = cplx_t<T>;
  C* cd = reinterpret_cast<C*>(dst);
  const int cp = chunk_planes(p, sizeof(T));
  int rc;
  for (int k0 = 0; k0 < p->nzp; k0 += cp) {
    const int k1 = std::min(p->nzp, k0 + cp);
    if ((rc = launch_rows<T>(ctx, p, false, dst, src, bt, k0 * p->nyp, k1 * p->nyp, scale)) != CRV_OK) return rc;
    if (p->nyp > 1 && (rc = launch_cols<T>(ctx, p, false, false, cd, bt, k0, k1)) != CRV_OK) return rc;
  }
  if (p->nzp > 1 && (rc = launch_cols<T>(ctx, p, false, true, cd, bt, 0, p->nyp)) != CRV_OK) return rc;
  return CRV_OK;
}

template <typename T>
int launch_c2r(crv_ctx* ctx, FftPlan* p, T* buf, T* dst, double scale, const Batch& bt) {
  using C = cplx_t<T>;
  C* cd = reinterpret_cast<C*>(buf);
  int rc;
  if (p->nzp > 1 && (rc = launch_cols<T>(ctx, p, true, true, cd, bt, 0, p->nyp)) != CRV_OK) return rc;
  const int cp = chunk_planes(p, sizeof(T));
  for (int k0 = 0; k0 < p->nzp; k0 += cp) {
    const int k1 = std::min(p->nzp, k0 + cp);
    if (p->nyp > 1 && (rc = launch_cols<T>(ctx, p, true, false, cd, bt, k0, k1)) != CRV_OK) return rc;
    if ((rc = launch_rows<T>(ctx, p, true, dst, buf, bt, k0 * p->nyp, k1 * p->nyp, scale)) != CRV_OK) return rc;
  }
  return CRV_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
namespace {

// the three axis plans and their twiddle table (replaces an existing one)
int build_twiddles(crv_ctx* ctx, FftPlan* p, int rev_x, int rev_y, int rev_z) {
  std::vector<long double> tr, ti;
  const int nxp = p->nxp;
  const int Lx = p->xeven ? nxp / 2 : nxp;
  if (!plan_1d(Lx, p->fx, tr, ti, rev_x) || !plan_1d(p->nyp, p->fy, tr, ti, rev_y) ||
      !plan_1d(p->nzp, p->fz, tr, ti, rev_z))
    return crv_set_error(ctx, CRV_E_UNSUPPORTED, "grid size is not a product of 2, 3, 5, 7");
  p->post_off = (int)tr.size();
  const long double twopi = 6.283185307179586476925286766559L;
  for (int k = 0; k <= nxp / 2; ++k) {
    const long double ang = -twopi * (long double)k / (long double)nxp;
    tr.push_back(cosl(ang));
    ti.push_back(sinl(ang));
  }
  const size_t n = tr.size();
  void* tw = nullptr;
  hipError_t e;
  if (p->prec == CRV_FP64) {
    std::vector<double2> h(n);
    for (size_t i = 0; i < n; ++i) { h[i].x = (double)tr[i]; h[i].y = (double)ti[i]; }
    e = hipMalloc(&tw, n * sizeof(double2));
    if (e == hipSuccess) e = hipMemcpy(tw, h.data(), n * sizeof(double2), hipMemcpyHostToDevice);
  } else {
    std::vector<float2> h(n);
    for (size_t i = 0; i < n; ++i) { h[i].x = (float)tr[i]; h[i].y = (float)ti[i]; }
    e = hipMalloc(&tw, n * sizeof(float2));
    if (e == hipSuccess) e = hipMemcpy(tw, h.data(), n * sizeof(float2), hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    if (tw) (void)hipFree(tw);
    return crv_set_error(ctx, CRV_E_HIP, std::string("plan twiddles: ") + hipGetErrorString(e));
  }
  if (p->tw) {
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(p->tw);
  }
  p->tw = tw;
  p->rev[0] = rev_x; p->rev[1] = rev_y; p->rev[2] = rev_z;
  return CRV_OK;
}

}  // namespace

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
  int rc = build_twiddles(ctx, p, 0, 0, 0);
  if (rc == CRV_OK) rc = crv_plan_shapes(ctx, p);
  if (rc != CRV_OK) {
    crv_free_plan(p);
    return rc;
  }
  ctx->plans[key] = p;
  *out = p;
  return CRV_OK;
}

// Launch shapes: a tile holds at most 4096 (FP32) / 2048 (FP64) complex
// values (32 KB of LDS); x rows per block P with (P+1) L and P cnxp <= that; y/z tiles
// of TX consecutive x columns, 16 (FP32) or 8 (FP64) at most, halved until
// the pencil tile fits.  FP32 two-stage y/z axes whose 16-column tile does
// not fit 4096 values take the 512-thread kernel (up to 8192 values, 64 KB).
// CRV_TUNE_* values override, within those limits.
int crv_plan_shapes(crv_ctx* ctx, FftPlan* p) {
  const int maxe = kThreads * ((p->prec == CRV_FP64) ? Ept<double>::value : Ept<float>::value);
  const int txdef = (p->prec == CRV_FP64) ? 8 : 16;
  const int Lx = p->fx.L;
  const int pmax = ctx->tune[CRV_TUNE_ROWS] > 0 ? ctx->tune[CRV_TUNE_ROWS] : 64;
  const int cnx = p->nxp / 2 + 1;
  // x rows per block within a budget of complex values per tile; axes
  // without a two-stage kernel use the half budget (and the kernel with half
  // the registers) when their tiles fit it
  const int cns = cnx | 1;  // staged C2R rows (odd stride)
  auto pick_rows = [&](int budget) {
    int P = 1;
    while (2 * P <= pmax && (2 * P + 1) * Lx <= budget && 2 * P * cns <= budget) P *= 2;
    return ((P + 1) * Lx <= budget && P * cns <= budget) ? P : 0;
  };
  p->half[0] = !axis_has_pair(p->fx) && pick_rows(maxe / 2) > 0;
  const int P = pick_rows(p->half[0] ? maxe / 2 : maxe);
  if (!P) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "nxp too large");
  p->rows_per_block = P;
  auto pick_tx = [&](int L, int want, int budget) {
    int t = want > 0 ? pow2_floor(want) : txdef;
    while (t > 1 && L * t > budget) t /= 2;
    return (L * t <= budget) ? t : 0;
  };
  p->half[1] = !axis_has_pair(p->fy) && pick_tx(p->nyp, ctx->tune[CRV_TUNE_TX_Y], maxe / 2) > 0;
  p->half[2] = !axis_has_pair(p->fz) && pick_tx(p->nzp, ctx->tune[CRV_TUNE_TX_Z], maxe / 2) > 0;
  // FP32 two-stage axes may use 512-thread blocks (twice the budget), so
  // that the tile is 16 columns (128 B per pencil row) wide
  const bool wy = p->prec == CRV_FP32 && axis_has_pair(p->fy);
  const bool wz = p->prec == CRV_FP32 && axis_has_pair(p->fz);
  p->tx_y = pick_tx(p->nyp, ctx->tune[CRV_TUNE_TX_Y], p->half[1] ? maxe / 2 : (wy ? wide_budget(p->fy, maxe) : maxe));
  p->tx_z = pick_tx(p->nzp, ctx->tune[CRV_TUNE_TX_Z], p->half[2] ? maxe / 2 : (wz ? wide_budget(p->fz, maxe) : maxe));
  p->wide[1] = wy && p->tx_y * p->nyp > maxe;
  p->wide[2] = wz && p->tx_z * p->nzp > maxe;
  if (!p->tx_y || !p->tx_z) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "nyp or nzp too large");
  p->chunk_mb = ctx->tune[CRV_TUNE_CHUNK_MB];
  p->batch_mb = ctx->tune[CRV_TUNE_BATCH_MB];
  p->order = ctx->tune[CRV_TUNE_ORDER] > 0 ? ctx->tune[CRV_TUNE_ORDER] - 1 : kDefaultOrder;
  return CRV_OK;
}

int crv_launch_r2c(crv_ctx* ctx, FftPlan* plan, void* dst, const void* src, double scale) {
  Batch bt;
  if (plan->prec == CRV_FP64)
    return launch_r2c<double>(ctx, plan, (double*)dst, (const double*)src, scale, bt);
  return launch_r2c<float>(ctx, plan, (float*)dst, (const float*)src, scale, bt);
}

int crv_launch_c2r(crv_ctx* ctx, FftPlan* plan, void* buf, void* dst, double scale, const void* add) {
  Batch bt;
  if (add) { bt.nadd = 1; bt.add[0] = add; }
  if (plan->prec == CRV_FP64)
    return launch_c2r<double>(ctx, plan, (double*)buf, (double*)dst, scale, bt);
  return launch_c2r<float>(ctx, plan, (float*)buf, (float*)dst, scale, bt);
}

// Grids per launch of a batched transform.  A pass over a grid that fits
// the memory-side cache leaves its output there for the next pass, so a
// batch of such grids goes in sub-batches of about batch_mb (default 256 MB,
// the cache's size); grids more than twice that size get no reuse anyway and
// the whole batch goes in one launch per pass.
int sub_batch(const FftPlan* p, int nbatch) {
  const size_t eb = p->prec == CRV_FP64 ? 8 : 4;
  const size_t grid = (size_t)2 * (p->nxp / 2 + 1) * p->nyp * p->nzp * eb;
  const size_t mall = (size_t)(p->batch_mb > 0 ? p->batch_mb : 256) << 20;
  if (grid > 2 * mall) return nbatch;
  return (int)std::max<size_t>(1, std::min<size_t>((size_t)nbatch, mall / grid));
}

int crv_sub_batch(const FftPlan* p, int nbatch) { return sub_batch(p, nbatch); }

// nbatch grids of this plan, bstride reals apart, transformed in place; the
// output of batch entry b gets add[b % nadd] added
int crv_launch_c2r_batch(crv_ctx* ctx, FftPlan* plan, void* buf, int nbatch, long long bstride, double scale,
                         const void* const* add, int nadd) {
  if (nbatch < 1 || nbatch > 65535 || nadd < 0 || nadd > 3) return crv_set_error(ctx, CRV_E_INVALID, "bad batch");
  const int sb = sub_batch(plan, nbatch);
  const size_t eb = plan->prec == CRV_FP64 ? 8 : 4;
  for (int b0 = 0; b0 < nbatch; b0 += sb) {
    Batch bt;
    bt.n = std::min(sb, nbatch - b0); bt.stride = bstride; bt.nadd = nadd;
    for (int q = 0; q < nadd; ++q) bt.add[q] = add[(b0 + q) % nadd];
    char* g = (char*)buf + (size_t)b0 * bstride * eb;
    const int rc = plan->prec == CRV_FP64
                       ? launch_c2r<double>(ctx, plan, (double*)g, (double*)g, scale, bt)
                       : launch_c2r<float>(ctx, plan, (float*)g, (float*)g, scale, bt);
    if (rc != CRV_OK) return rc;
  }
  return CRV_OK;
}

int crv_launch_r2c_batch(crv_ctx* ctx, FftPlan* plan, void* dst, const void* src, int nbatch, long long bstride,
                         double scale) {
  if (nbatch < 1 || nbatch > 65535) return crv_set_error(ctx, CRV_E_INVALID, "bad batch");
  const int sb = sub_batch(plan, nbatch);
  const size_t eb = plan->prec == CRV_FP64 ? 8 : 4;
  for (int b0 = 0; b0 < nbatch; b0 += sb) {
    Batch bt;
    bt.n = std::min(sb, nbatch - b0); bt.stride = bstride;
    const size_t off = (size_t)b0 * bstride * eb;
    char* d = (char*)dst + off;
    const char* sp = (const char*)src + off;
    const int rc = plan->prec == CRV_FP64
                       ? launch_r2c<double>(ctx, plan, (double*)d, (const double*)sp, scale, bt)
                       : launch_r2c<float>(ctx, plan, (float*)d, (const float*)sp, scale, bt);
    if (rc != CRV_OK) return rc;
  }
  return CRV_OK;
}

bool crv_cov_gen_fits(const FftPlan* p, int nband, int rw) {
  CovGen g;
  g.rows = (const float*)p;  // any non-null pointer: count the band tables
  g.nband = nband; g.rw = rw;
  return cols_lds_bytes(p->fz, p->tx_z, sizeof(float2), &g) <= 65536;
}

// One pass of a transform over part of a batch (see internal.hpp).
// ---- y inverse of the covariances' band planes -------------------------------
const void* cov_band_y_kernel(const Fft1D& f, bool wide) {
  if (f.nst != 2) return nullptr;
#define CRV_PAIR_BANDY(A, B)                                                                       \
  if (f.R[0] == (A) && f.R[1] == (B))                                                              \
    return wide ? (const void*)k_cov_band_y<A, B, true> : (const void*)k_cov_band_y<A, B, false>;
  CRV_FFT_PAIRS(CRV_PAIR_BANDY)
#undef CRV_PAIR_BANDY
  return nullptr;
}

bool crv_cov_band_y_fits(const FftPlan* p) {
  return p->prec == CRV_FP32 && p->nyp > 1 && p->nzp > 1 && cov_band_y_kernel(p->fy, p->wide[1]) != nullptr;
}

int crv_launch_cov_band_y(crv_ctx* ctx, FftPlan* p, void* buf, int nbatch, long long bstride, const CovGen& g) {
  const void* kern = crv_cov_band_y_fits(p) ? cov_band_y_kernel(p->fy, p->wide[1]) : nullptr;
  if (!kern || !g.rows || !g.bplanes || g.nband < 1 || nbatch < 1 || nbatch > 6 || g.rw % 4 != 0 ||
      g.rw > 4 + 4 * kEigMaxN)
    return crv_set_error(ctx, CRV_E_UNSUPPORTED, "covariance band y pass: unsupported plan");
  BandYArgs a{};
  const int cnxp = p->nxp / 2 + 1;
  a.data = (float2*)buf; a.bstride = bstride / 2;
  a.f = p->fy; a.sd = make_stage_div(p->fy); a.tw = (const float2*)p->tw;
  a.pstride = (long long)cnxp * p->nyp;
  a.ncols = cnxp; a.cshift = ilog2(p->tx_y); a.ntx = (cnxp + p->tx_y - 1) / p->tx_y;
  a.rev = (p->order >> 4) & 1;
  a.nband = g.nband; a.rw = g.rw; a.n = g.n; a.q0 = g.q0; a.noqueue = g.noqueue;
  a.bplanes = g.bplanes; a.lam = g.lam; a.rows = g.rows;
  a.geo = EigQueueGeo{g.rx, g.ry, g.rz, g.lowk2, g.cond_thr, p->nyp, p->nzp};
  const size_t lds = ((size_t)p->fy.L * p->tx_y + p->fy.twn) * sizeof(float2) + 16 + (4 + 4 * kEigMaxN) * sizeof(float);
  void* args[] = {&a};
  ProfScope ps(ctx, "fft_y_inv_band");
  CRV_HIP_CHECK(ctx, hipLaunchKernel(kern, dim3(a.ntx * g.nband, nbatch), dim3(p->wide[1] ? 2 * kThreads : kThreads),
                                     args, lds, ctx->stream));
  return CRV_OK;
}

// ---- fused z pass of crv_invert -----------------------------------------------
size_t zpost_lds_bytes(const FftPlan* p, int nband, int tx) {
  return ((size_t)3 * p->fz.L * tx + p->fz.twn) * sizeof(float2) + (size_t)(p->fz.L + nband) * sizeof(int);
}

const void* zpost_kernel(const Fft1D& f) {
  if (f.nst != 2) return nullptr;
#define CRV_PAIR_ZPOST(A, B) \
  if (f.R[0] == (A) && f.R[1] == (B)) return (const void*)k_zpost<A, B>;
  CRV_FFT_PAIRS(CRV_PAIR_ZPOST)
#undef CRV_PAIR_ZPOST
  return nullptr;
}

int crv_zpost_tx(crv_ctx* ctx, const FftPlan* p, int nband) {
  const int want = ctx->tune[CRV_TUNE_ZPOST];
  if (want == 0 || p->prec != CRV_FP32 || nband < 1 || !zpost_kernel(p->fz)) return 0;
  const int L = p->fz.L;
  // the widest tile (16 columns at most, or the tuned width) whose band
  // wavenumbers the threads can hold, whose stages the block covers and
  // whose LDS leaves room for a second block per CU (the tuned width may
  // take the whole 160 KB)
  if ((long long)(p->nxp / 2 + 1) * p->nyp * p->nzp >= (1LL << 31)) return 0;
  for (int t = want >= 2 ? pow2_floor(want) : 16; t >= 1; t /= 2) {
    if ((long long)nband * t > (long long)kZQ * kZThreads || L * t > kZPF * kZThreads) continue;
    bool ok = true;
    for (int st = 0; st < 2; ++st) {
      const int R = p->fz.R[st];
      if ((L / R) * t > ((Ept<float>::value + R - 1) / R) * kZThreads) ok = false;
    }
    if (ok && zpost_lds_bytes(p, nband, t) <= (want >= 2 ? 160u : 80u) * 1024u) return t;
  }
  return 0;
}

int crv_launch_zpost(crv_ctx* ctx, FftPlan* p, const ZPost& z, int tx) {
  const void* kern = zpost_kernel(p->fz);
  if (!kern || tx < 1 || z.n < 1 || z.n > kEigMaxN)
    return crv_set_error(ctx, CRV_E_UNSUPPORTED, "fused z pass: unsupported shape");
  ZPostArgs a{};
  for (int l = 0; l < z.n; ++l) a.data[l] = (float2*)z.data[l];
  for (int m = 0; m < 3; ++m) a.mean[m] = (float2*)z.mean[m];
  a.f = p->fz; a.sd = make_stage_div(p->fz); a.tw = (const float2*)p->tw;
  const int cnxp = p->nxp / 2 + 1;
  a.estride = cnxp * p->nyp; a.ncols = cnxp;
  a.cshift = ilog2(tx); a.ntx = (cnxp + tx - 1) / tx;
  a.rev = (p->order >> 2) & 1;
  a.n = z.n; a.nband = z.nband; a.lam = z.lam; a.erec = z.erec; a.bidx = z.bidx; a.bplanes = z.bplanes;
  a.stats = z.stats;
  const size_t lds = zpost_lds_bytes(p, z.nband, tx);
  if (lds > 64u * 1024u)
    CRV_HIP_CHECK(ctx, hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  a.ntiles = a.ntx * p->nyp;
  // persistent: as many blocks as fit at once (a multiple of the 8 XCDs)
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kZThreads, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const int grid = std::min(a.ntiles, std::max(8, per_cu * ctx->num_cu / 8 * 8));
  void* args[] = {&a};
  ProfScope ps(ctx, "fft_z_post");
  CRV_HIP_CHECK(ctx, hipLaunchKernel(kern, dim3(grid), dim3(kZThreads), args, lds, ctx->stream));
  return CRV_OK;
}

int crv_launch_pass(crv_ctx* ctx, FftPlan* plan, int pass, void* buf, const void* src, int nbatch, long long bstride,
                    int o0, int o1, double scale, const void* const* add, int nadd, const CovGen* gen) {
  if (nbatch < 1 || nbatch > 65535 || nadd < 0 || nadd > 3 || pass < 0 || pass > 5 || o0 < 0 || o1 < o0)
    return crv_set_error(ctx, CRV_E_INVALID, "bad pass");
  if (o1 == o0) return CRV_OK;
  Batch bt;
  bt.n = nbatch; bt.stride = bstride; bt.nadd = nadd;
  for (int q = 0; q < nadd; ++q) bt.add[q] = add[q];
  const bool f64 = plan->prec == CRV_FP64;
  switch (pass) {
    case 0:
      return f64 ? launch_rows<double>(ctx, plan, false, (double*)buf, (const double*)src, bt, o0, o1, scale)
                 : launch_rows<float>(ctx, plan, false, (float*)buf, (const float*)src, bt, o0, o1, scale);
    case 5:
      return f64 ? launch_rows<double>(ctx, plan, true, (double*)buf, (const double*)buf, bt, o0, o1, scale)
                 : launch_rows<float>(ctx, plan, true, (float*)buf, (const float*)buf, bt, o0, o1, scale);
    default: {
      const bool inv = pass >= 3, zaxis = (pass == 2 || pass == 3);
      if ((zaxis ? plan->nzp : plan->nyp) == 1) return CRV_OK;
      const CovGen* zp = (pass == 3 && gen && gen->lam && !f64) ? gen : nullptr;
      return f64 ? launch_cols<double>(ctx, plan, inv, zaxis, (double2*)buf, bt, o0, o1)
                 : launch_cols<float>(ctx, plan, inv, zaxis, (float2*)buf, bt, o0, o1, zp);
    }
  }
}

// ---- autotuner ---------------------------------------------------------------
extern "C" int crv_fft_autotune(crv_ctx* ctx, int nxp, int nyp, int nzp, int precision, int reps,
                                crv_fft_shape* shape) {
  if (!ctx) return CRV_E_INVALID;
  if (precision != CRV_FP32 && precision != CRV_FP64) return crv_set_error(ctx, CRV_E_INVALID, "bad precision");
  if (reps < 1) reps = 3;
  FftPlan* p = nullptr;
  int rc = crv_get_plan(ctx, nxp, nyp, nzp, precision, &p);
  if (rc != CRV_OK) return rc;
  const size_t eb = precision == CRV_FP64 ? 8 : 4;
  const size_t rsize = (size_t)2 * (nxp / 2 + 1) * nyp * nzp;
  void* buf = nullptr;
  CRV_HIP_CHECK(ctx, hipMalloc(&buf, rsize * eb));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto cleanup = [&]() {
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(buf);
  };
  if (hipMemset(buf, 0, rsize * eb) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess) {
    cleanup();
    return crv_set_error(ctx, CRV_E_HIP, "autotune: scratch set-up failed");
  }
  const bool prof = ctx->prof_on;
  ctx->prof_on = false;
  Batch bt;
  const bool f64 = precision == CRV_FP64;
  // milliseconds per call of fn, median of reps after a warm-up call
  auto timed = [&](auto fn, float& ms) -> int {
    int r = fn();
    if (r != CRV_OK) return r;
    std::vector<float> t;
    for (int i = 0; i < reps; ++i) {
      CRV_HIP_CHECK(ctx, hipEventRecord(e0, ctx->stream));
      if ((r = fn()) != CRV_OK) return r;
      CRV_HIP_CHECK(ctx, hipEventRecord(e1, ctx->stream));
      CRV_HIP_CHECK(ctx, hipEventSynchronize(e1));
      float x = 0.0f;
      CRV_HIP_CHECK(ctx, hipEventElapsedTime(&x, e0, e1));
      t.push_back(x);
    }
    std::sort(t.begin(), t.end());
    ms = t[t.size() / 2];
    return CRV_OK;
  };
  auto rows = [&]() -> int {
    const int nr = nyp * nzp;
    int r = f64 ? launch_rows<double>(ctx, p, false, (double*)buf, (const double*)buf, bt, 0, nr, 1.0)
                : launch_rows<float>(ctx, p, false, (float*)buf, (const float*)buf, bt, 0, nr, 1.0);
    if (r != CRV_OK) return r;
    return f64 ? launch_rows<double>(ctx, p, true, (double*)buf, (const double*)buf, bt, 0, nr, 1.0)
               : launch_rows<float>(ctx, p, true, (float*)buf, (const float*)buf, bt, 0, nr, 1.0);
  };
  auto cols = [&](bool z) -> int {
    const int o1 = z ? nyp : nzp;
    int r = f64 ? launch_cols<double>(ctx, p, false, z, (double2*)buf, bt, 0, o1)
                : launch_cols<float>(ctx, p, false, z, (float2*)buf, bt, 0, o1);
    if (r != CRV_OK) return r;
    return f64 ? launch_cols<double>(ctx, p, true, z, (double2*)buf, bt, 0, o1)
               : launch_cols<float>(ctx, p, true, z, (float2*)buf, bt, 0, o1);
  };
  const int maxe = kThreads * (f64 ? Ept<double>::value : Ept<float>::value);
  rc = CRV_OK;
  // radix order of each axis (only lengths without a two-stage kernel)
  int rev[3] = {0, 0, 0};
  for (int ax = 0; ax < 3 && rc == CRV_OK; ++ax) {
    const Fft1D& f = ax == 0 ? p->fx : (ax == 1 ? p->fy : p->fz);
    if (f.nst < 2 || is_pair(std::vector<int>(f.R, f.R + f.nst))) continue;
    float best = 1e30f;
    for (int r = 0; r < 2 && rc == CRV_OK; ++r) {
      int tr[3] = {rev[0], rev[1], rev[2]};
      tr[ax] = r;
      if ((rc = build_twiddles(ctx, p, tr[0], tr[1], tr[2])) != CRV_OK) break;
      float ms = 0.0f;
      rc = ax == 0 ? timed(rows, ms) : timed([&]() { return cols(ax == 2); }, ms);
      if (rc == CRV_OK && ms < best) { best = ms; rev[ax] = r; }
    }
  }
  if (rc == CRV_OK) rc = build_twiddles(ctx, p, rev[0], rev[1], rev[2]);
  // x rows per tile
  if (rc == CRV_OK) {
    float best = 1e30f;
    int bestP = p->rows_per_block;
    const int bx = p->half[0] ? maxe / 2 : maxe;
    for (int P = 1; P <= 64 && rc == CRV_OK; P *= 2) {
      if ((P + 1) * p->fx.L > bx || P * ((p->nxp / 2 + 1) | 1) > bx) break;
      p->rows_per_block = P;
      float ms = 0.0f;
      if ((rc = timed(rows, ms)) == CRV_OK && ms < best) { best = ms; bestP = P; }
    }
    p->rows_per_block = bestP;
  }
  // pencil tile widths
  for (int ax = 1; ax <= 2 && rc == CRV_OK; ++ax) {
    const int L = ax == 1 ? nyp : nzp;
    if (L < 2) continue;
    int& tx = ax == 1 ? p->tx_y : p->tx_z;
    float best = 1e30f;
    int bestT = tx;
    const Fft1D& fa = ax == 1 ? p->fy : p->fz;
    const bool wok = !f64 && axis_has_pair(fa);
    const int bt_ = p->half[ax] ? maxe / 2 : (wok ? wide_budget(fa, maxe) : maxe);
    const bool wide0 = p->wide[ax];
    for (int t = 1; t <= 64 && rc == CRV_OK; t *= 2) {
      if (L * t > bt_) break;
      tx = t;
      p->wide[ax] = wok && L * t > maxe;
      float ms = 0.0f;
      if ((rc = timed([&]() { return cols(ax == 2); }, ms)) == CRV_OK && ms < best) { best = ms; bestT = t; }
    }
    tx = bestT;
    p->wide[ax] = (rc == CRV_OK) ? (wok && L * bestT > maxe) : wide0;
  }
  if (rc == CRV_OK && shape) {
    shape->rows_per_tile = p->rows_per_block;
    shape->tx_y = p->tx_y;
    shape->tx_z = p->tx_z;
    for (int ax = 0; ax < 3; ++ax) shape->radix_order[ax] = p->rev[ax];
    float a = 0.0f, b = 0.0f;
    rc = timed([&]() {
      return f64 ? launch_r2c<double>(ctx, p, (double*)buf, (const double*)buf, 1.0, bt)
                 : launch_r2c<float>(ctx, p, (float*)buf, (const float*)buf, 1.0, bt);
    }, a);
    if (rc == CRV_OK)
      rc = timed([&]() {
        return f64 ? launch_c2r<double>(ctx, p, (double*)buf, (double*)buf, 1.0, bt)
                   : launch_c2r<float>(ctx, p, (float*)buf, (float*)buf, 1.0, bt);
      }, b);
    shape->ms_r2c = a;
    shape->ms_c2r = b;
  }
  ctx->prof_on = prof;
  cleanup();
  return rc;
}
