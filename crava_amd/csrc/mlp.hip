// Rock-physics surrogate: a 5 -> 128 -> 128 -> 128 -> 3 tanh MLP trained and
// evaluated with BF16 matrix-core (MFMA) kernels and FP32 accumulation.
//
// It stands in for the per-sample rplib chain (DEM inclusion ODE + Gassmann
// with Batzle-Wang fluids; reference rplib/dem.cpp:39-189,
// rplib/orddiffeqsolver.cpp:17-167, rplib/rockgassmann.cpp:79-97) when
// facies-conditional elastic samples are generated for faciesprob
// (reference rplib/distributionsrockdem.cpp:125-156,
// src/faciesprob.cpp:2494-2559); crv_rock_forward stays the fallback and the
// measure of its error.
//
// Parameters live in augmented matrices whose last used column holds the
// bias, so every layer is one product with a constant-1 input column:
//   W1a [128][16]  (5 inputs, bias at column 5, zero pad)
//   W2a [128][144], W3a [128][144]  (128 inputs, bias at column 128)
//   W4a [32][144]  (3 outputs used, rows 3..31 zero)
// FP32 master copies and Adam moments; the kernels read a BF16 copy.
//
// Kernels (gfx950, wave64, v_mfma_f32_32x32x16_bf16):
//  * k_mlp_forward: persistent blocks of 512 threads; all weights in LDS;
//    tiles of 128 samples go through the four layers on chip (activations
//    in LDS as BF16).  Training mode also stores the activations and the
//    loss gradient for the backward products.
//  * k_gemm: tiled 64 x 64 BF16 product with LDS staging, used for the
//    backward pass: dZ = (dZ' W) * (1 - H^2) and split-K weight gradients
//    dW = dZ'^T H (partial sums per K slice).
//  * k_reduce / k_reduce_adam: sum the slices (and with it the Adam step and
//    the BF16 refresh of a training step).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

struct crv_mlp {
  crv_ctx* ctx = nullptr;
  crv_mlp_norm norm{};
  long long max_batch = 0;
  // parameters
  float* w = nullptr;       // FP32 master, kParams
  float* m = nullptr;       // Adam first moment
  float* v = nullptr;       // Adam second moment
  float* g = nullptr;       // reduced gradient
  __bf16* wb = nullptr;     // BF16 copy read by the kernels
  // per-batch scratch
  __bf16* xa = nullptr;     // [B][16]
  __bf16* h[3] = {nullptr, nullptr, nullptr};  // [B][144]
  __bf16* dy = nullptr;     // [B][32]
  __bf16* dz[2] = {nullptr, nullptr};          // [B][128]
  float* part = nullptr;    // [slices][kParams]
  float* loss = nullptr;    // device scalar
  int slices = 0;
  double flops_fwd = 0, flops_bwd = 0;  // MFMA flops of the last call
};

namespace {

constexpr int kT = 512;          // threads of the forward block (8 waves)
constexpr int kTile = 128;       // samples per forward tile
constexpr int kHid = 128;
constexpr int kKa = 144;         // augmented K of layers 2..4
constexpr int kK1 = 16;          // augmented K of layer 1
constexpr int kNo = 32;          // padded output rows
constexpr int kLd = kKa + 8;     // LDS row stride (bf16) of activations and W2..W4
static_assert(kK1 == 16 && (kLd * 2) % 16 == 0, "tile rows start at 16-byte boundaries; layer 1 reads 16 columns");
constexpr int kLd1 = kK1 + 8;    // LDS row stride of W1
constexpr int kOffW1 = 0, kOffW2 = kHid * kK1, kOffW3 = kOffW2 + kHid * kKa, kOffW4 = kOffW3 + kHid * kKa;
constexpr int kParams = kOffW4 + kNo * kKa;  // 43520
constexpr int kKChunk = 256;     // rows per split-K slice of the weight gradients (enough blocks to fill the chip)

__device__ __forceinline__ float tanh_fast(float x) {
  const float e = __expf(-2.0f * fabsf(x));
  const float t = __fdividef(1.0f - e, 1.0f + e);
  return copysignf(t, x);
}

struct FwdArgs {
  const float* x;       // [n][5]
  const float* t;       // [n][3] targets (training) or null
  float* y;             // [n][3] predictions (inference) or null
  __bf16* xa;           // training: stored inputs / activations / output gradient
  __bf16* h1;
  __bf16* h2;
  __bf16* h3;
  __bf16* dy;
  float* loss;
  const __bf16* w;
  long long n;
  crv_mlp_norm nm;
  float grad_scale;
  int train;
};

// one layer over the tile: act (rows x K) -> act (rows x N) through W
// (N rows of stride ldw); 8 waves = 4 row groups x 2 column halves.
template <int KST, int NT, bool TANH>
__device__ __forceinline__ void mlp_layer(__bf16* act, const __bf16* wl, int ldw, floatx16* acc_out) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int rg = wave & 3, nh = wave >> 2;
  const int r = 32 * rg + (lane & 31), kq = 8 * (lane >> 5);
  bf16x8 a[KST];
#pragma unroll
  for (int s = 0; s < KST; ++s) a[s] = *reinterpret_cast<const bf16x8*>(act + r * kLd + 16 * s + kq);
  __syncthreads();
  floatx16 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = floatx16{};
  const bool active = (NT == 2) || (nh == 0);  // the 32-wide output layer uses half the waves
  if (active) {
#pragma unroll
    for (int s = 0; s < KST; ++s)
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int n = 64 * nh + 32 * t + (lane & 31);
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(wl + n * ldw + 16 * s + kq);
        acc[t] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[s], b, acc[t], 0, 0, 0);
      }
  }
  if constexpr (TANH) {
#pragma unroll
    for (int t = 0; t < NT; ++t)
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int row = 32 * rg + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const int col = 64 * nh + 32 * t + (lane & 31);
        act[row * kLd + col] = (__bf16)tanh_fast(acc[t][i]);
      }
    __syncthreads();
  } else {
#pragma unroll
    for (int t = 0; t < NT; ++t) acc_out[t] = acc[t];
  }
}

__device__ __forceinline__ void store_rows(__bf16* dst, const __bf16* act, long long row0, long long n, int cols) {
  // cols is a multiple of 8; 16-byte chunks
  const int cpr = cols / 8;
  for (int c = threadIdx.x; c < kTile * cpr; c += kT) {
    const int r = c / cpr, q = c - r * cpr;
    if (row0 + r < n)
      *reinterpret_cast<bf16x8*>(dst + (row0 + r) * cols + 8 * q) = *reinterpret_cast<const bf16x8*>(act + r * kLd + 8 * q);
  }
}

__global__ __launch_bounds__(kT) void k_mlp_forward(FwdArgs a) {
  extern __shared__ unsigned char smem_raw[];
  __bf16* w1 = reinterpret_cast<__bf16*>(smem_raw);
  __bf16* w2 = w1 + kHid * kLd1;
  __bf16* w3 = w2 + kHid * kLd;
  __bf16* w4 = w3 + kHid * kLd;
  __bf16* act = w4 + kNo * kLd;
  // weights -> LDS (16-byte chunks)
  auto load_w = [&](__bf16* dst, int ld, const __bf16* src, int rows, int cols) {
    const int cpr = cols / 8;
    for (int c = threadIdx.x; c < rows * cpr; c += kT) {
      const int r = c / cpr, q = c - r * cpr;
      *reinterpret_cast<bf16x8*>(dst + r * ld + 8 * q) = *reinterpret_cast<const bf16x8*>(src + r * cols + 8 * q);
    }
  };
  load_w(w1, kLd1, a.w + kOffW1, kHid, kK1);
  load_w(w2, kLd, a.w + kOffW2, kHid, kKa);
  load_w(w3, kLd, a.w + kOffW3, kHid, kKa);
  load_w(w4, kLd, a.w + kOffW4, kNo, kKa);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  float lsum = 0.0f;
  const long long ntiles = (a.n + kTile - 1) / kTile;
  // the inputs of sample threadIdx.x of a tile (threads below kTile), loaded
  // one tile ahead so that their latency hides behind the current tile's
  // layers
  float xin[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  auto load_in = [&](long long tile) {
    const long long gr = tile * kTile + threadIdx.x;
    if (threadIdx.x < kTile && tile < ntiles && gr < a.n) {
#pragma unroll
      for (int f = 0; f < 5; ++f) xin[f] = a.x[gr * 5 + f];
    }
  };
  load_in(blockIdx.x);
  // columns kHid..kKa-1 of the tile rows: the bias column of layers 2..4 (1)
  // and zero padding.  No layer writes them, so they are set once; layer 1
  // reads columns 0..kK1-1 only, set for each tile below, and writes 0..kHid-1
  for (int e = threadIdx.x; e < kTile * (kKa - kHid); e += kT) {
    const int r = e / (kKa - kHid), c = kHid + e % (kKa - kHid);
    act[r * kLd + c] = (__bf16)(c == kHid ? 1.0f : 0.0f);
  }
  for (long long tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const long long row0 = tile * kTile;
    __syncthreads();
    if (threadIdx.x < kTile) {
      const int r = threadIdx.x;
      const long long gr = row0 + r;
      // the 5 scaled inputs, the layer-1 bias column (1) and zero padding
      bf16x8 v0, v1;
#pragma unroll
      for (int c = 0; c < 8; ++c) { v0[c] = (__bf16)0.0f; v1[c] = (__bf16)0.0f; }
      if (gr < a.n) {
#pragma unroll
        for (int f = 0; f < 5; ++f) {
          float v = xin[f];
          if (a.nm.in_log_mask & (1 << f)) v = log10f(v);
          v0[f] = (__bf16)((v - a.nm.in_lo[f]) * a.nm.in_scale[f] - 1.0f);
        }
      }
      v0[5] = (__bf16)1.0f;
      bf16x8* ar = reinterpret_cast<bf16x8*>(act + r * kLd);   // rows are 16-byte aligned (kLd = 152)
      ar[0] = v0;
      ar[1] = v1;
    }
    load_in(tile + gridDim.x);
    __syncthreads();
    if (a.train) store_rows(a.xa, act, row0, a.n, kK1);
    mlp_layer<1, 2, true>(act, w1, kLd1, nullptr);
    // layer 1 overwrote columns 0..127; column 128 still holds 1
    if (a.train) store_rows(a.h1, act, row0, a.n, kKa);
    mlp_layer<9, 2, true>(act, w2, kLd, nullptr);
    if (a.train) store_rows(a.h2, act, row0, a.n, kKa);
    mlp_layer<9, 2, true>(act, w3, kLd, nullptr);
    if (a.train) store_rows(a.h3, act, row0, a.n, kKa);
    floatx16 out[1];
    mlp_layer<9, 1, false>(act, w4, kLd, out);
    if ((wave >> 2) == 0) {
      const int rg = wave & 3, col = lane & 31;
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int row = 32 * rg + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
        const long long gr = row0 + row;
        if (gr >= a.n) continue;
        if (a.train) {
          float gv = 0.0f;
          if (col < 3) {
            const float tn = (a.t[gr * 3 + col] - a.nm.out_mean[col]) / a.nm.out_std[col];
            const float e = out[0][i] - tn;
            lsum += e * e;
            gv = a.grad_scale * e;
          }
          a.dy[gr * kNo + col] = (__bf16)gv;
        } else if (col < 3) {
          a.y[gr * 3 + col] = out[0][i] * a.nm.out_std[col] + a.nm.out_mean[col];
        }
      }
    }
  }
  if (a.train) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) lsum += __shfl_down(lsum, off, 64);
    if (lane == 0 && lsum != 0.0f) atomicAdd(a.loss, lsum);
  }
}

// ---- backward products ------------------------------------------------------
struct GemmArgs {
  const __bf16* A; int lda; int a_trans;   // A(m,k) = a_trans ? A[k*lda+m] : A[m*lda+k]
  const __bf16* B; int ldb; int b_rowk;    // B(k,n) = b_rowk ? B[k*ldb+n] : B[n*ldb+k]
  int M, N;
  long long K;
  int kchunk;                              // split-K slice length (mode 0)
  int mode;                                // 0: fp32 partial P[slice][M][N]; 1: bf16 C = acc*(1-H^2)
  float* P;
  __bf16* C; int ldc;
  const __bf16* H; int ldh;
  int vec;                                 // A, B 16-byte aligned, lda, ldb multiples of 8: 8-wide loads
};

constexpr int kGT = 256;   // 4 waves, 64 x 64 tile
constexpr int kGK = 128;   // K per LDS tile: 8 loads of 16 B in flight per thread, 8 MFMA steps per wave
constexpr int kGLd = kGK + 8;

// one 64 x 64 output tile (bx, by) of K slice bz
__device__ __forceinline__ void gemm_block(const GemmArgs& g, int bx, int by, int bz) {
  __shared__ __bf16 As[64 * kGLd];
  __shared__ __bf16 Bs[64 * kGLd];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int m0 = bx * 64, n0 = by * 64;
  const long long kb = (long long)bz * g.kchunk;
  const long long ke = (g.K < kb + g.kchunk) ? g.K : kb + g.kchunk;
  const int wm = wave & 1, wn = wave >> 1;
  floatx16 acc = floatx16{};
  for (long long k0 = kb; k0 < ke; k0 += kGK) {
    // chunks of 32 and MFMA steps of 16 past the end of K are skipped
    const int kt = (int)((ke - k0 < kGK) ? ke - k0 : kGK);
    // A tile 64 x kGK -> As[m][k], 32 values per thread as 4 chunks of 8
    bf16x8 va[kGK / 32], vb[kGK / 32];
#pragma unroll
    for (int c = 0; c < kGK / 32; ++c) {
      bf16x8 v = bf16x8{};
      if (32 * c >= kt) { va[c] = v; continue; }
      if (!g.a_trans) {
        // m-major: 8 consecutive k of one m
        const int m = tid >> 2, k8 = 8 * (tid & 3) + 32 * c;
        const long long gm = m0 + m, gk = k0 + k8;
        if (gm < g.M) {
          if (g.vec && gk + 8 <= ke) v = *reinterpret_cast<const bf16x8*>(g.A + gm * g.lda + gk);
          else
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (gk + j < ke) ? g.A[gm * g.lda + gk + j] : (__bf16)0.0f;
        }
      } else {
        // k-major: 8 consecutive m of one k
        const int k = (tid >> 3) + 32 * c, m8 = 8 * (tid & 7);
        const long long gk = k0 + k, gm = m0 + m8;
        if (gk < ke) {
          if (g.vec && gm + 8 <= g.M) v = *reinterpret_cast<const bf16x8*>(g.A + gk * g.lda + gm);
          else
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (gm + j < g.M) ? g.A[gk * g.lda + gm + j] : (__bf16)0.0f;
        }
      }
      va[c] = v;
    }
    // B tile kGK x 64 -> Bs[n][k]
#pragma unroll
    for (int c = 0; c < kGK / 32; ++c) {
      bf16x8 v = bf16x8{};
      if (32 * c >= kt) { vb[c] = v; continue; }
      if (g.b_rowk) {
        const int k = (tid >> 3) + 32 * c, n8 = 8 * (tid & 7);
        const long long gk = k0 + k, gn = n0 + n8;
        if (gk < ke) {
          if (g.vec && gn + 8 <= g.N) v = *reinterpret_cast<const bf16x8*>(g.B + gk * g.ldb + gn);
          else
#pragma unroll
            for (int j = 0; j < 8; ++j) v[j] = (gn + j < g.N) ? g.B[gk * g.ldb + gn + j] : (__bf16)0.0f;
        }
      } else {
        const int n = tid >> 2, k8 = 8 * (tid & 3) + 32 * c;
        const long long gn = n0 + n, gk = k0 + k8;
        if (gn < g.N) {
#pragma unroll
          for (int j = 0; j < 8; ++j) v[j] = (gk + j < ke) ? g.B[gn * g.ldb + gk + j] : (__bf16)0.0f;
        }
      }
      vb[c] = v;
    }
#pragma unroll
    for (int c = 0; c < kGK / 32; ++c) {
      if (32 * c >= kt) continue;
      if (!g.a_trans) {
        *reinterpret_cast<bf16x8*>(As + (tid >> 2) * kGLd + 8 * (tid & 3) + 32 * c) = va[c];
      } else {
        const int k = (tid >> 3) + 32 * c, m8 = 8 * (tid & 7);
#pragma unroll
        for (int j = 0; j < 8; ++j) As[(m8 + j) * kGLd + k] = va[c][j];
      }
      if (g.b_rowk) {
        const int k = (tid >> 3) + 32 * c, n8 = 8 * (tid & 7);
#pragma unroll
        for (int j = 0; j < 8; ++j) Bs[(n8 + j) * kGLd + k] = vb[c][j];
      } else {
        *reinterpret_cast<bf16x8*>(Bs + (tid >> 2) * kGLd + 8 * (tid & 3) + 32 * c) = vb[c];
      }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kGK / 16; ++s) {
      if (16 * s >= kt) break;
      const bf16x8 av = *reinterpret_cast<const bf16x8*>(As + (32 * wm + (lane & 31)) * kGLd + 16 * s + 8 * (lane >> 5));
      const bf16x8 bv = *reinterpret_cast<const bf16x8*>(Bs + (32 * wn + (lane & 31)) * kGLd + 16 * s + 8 * (lane >> 5));
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(av, bv, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int col = n0 + 32 * wn + (lane & 31);
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = m0 + 32 * wm + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    if (row >= g.M || col >= g.N) continue;
    if (g.mode == 0) {
      g.P[((size_t)bz * g.M + row) * g.N + col] = acc[i];
    } else {
      const float h = (float)g.H[(size_t)row * g.ldh + col];
      g.C[(size_t)row * g.ldc + col] = (__bf16)(acc[i] * (1.0f - h * h));
    }
  }
}

__global__ __launch_bounds__(kGT) void k_gemm(GemmArgs g) { gemm_block(g, blockIdx.x, blockIdx.y, blockIdx.z); }

// Two independent products in one launch: blocks below n0 = gx0 gy0 gz0
// take tiles of g0, the others tiles of g1 (gx1 x gy1 tiles per K slice)
__global__ __launch_bounds__(kGT) void k_gemm_pair(GemmArgs g0, int gx0, int gy0, int gz0, GemmArgs g1, int gx1,
                                                   int gy1) {
  const int n0 = gx0 * gy0 * gz0;
  int b = blockIdx.x;
  if (b < n0) {
    gemm_block(g0, b % gx0, (b / gx0) % gy0, b / (gx0 * gy0));
  } else {
    b -= n0;
    gemm_block(g1, b % gx1, (b / gx1) % gy1, b / (gx1 * gy1));
  }
}

struct Seg { int off, size, slices; const float* P; };
struct ReduceArgs { Seg s[4]; float* g; };

__global__ void k_reduce(ReduceArgs a) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kParams; p += gridDim.x * blockDim.x) {
    int l = 0;
    while (l < 3 && p >= a.s[l + 1].off) ++l;
    const int q = p - a.s[l].off;
    // eight independent partial sums keep eight loads in flight
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const int ns = a.s[l].slices;
    int z = 0;
    for (; z + 8 <= ns; z += 8)
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += a.s[l].P[(size_t)(z + u) * a.s[l].size + q];
    for (; z < ns; ++z) acc[0] += a.s[l].P[(size_t)z * a.s[l].size + q];
    a.g[p] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
  }
}

// the reduction of k_reduce and the Adam step (m, v, the bias-corrected
// step lr_t and the BF16 refresh) in one pass over the parameters
__global__ void k_reduce_adam(ReduceArgs a, float* w, float* m, float* v, __bf16* wb, float lr_t, float b1, float b2,
                              float eps) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < kParams; p += gridDim.x * blockDim.x) {
    int l = 0;
    while (l < 3 && p >= a.s[l + 1].off) ++l;
    const int q = p - a.s[l].off;
    float acc[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const int ns = a.s[l].slices;
    int z = 0;
    for (; z + 8 <= ns; z += 8)
#pragma unroll
      for (int u = 0; u < 8; ++u) acc[u] += a.s[l].P[(size_t)(z + u) * a.s[l].size + q];
    for (; z < ns; ++z) acc[0] += a.s[l].P[(size_t)z * a.s[l].size + q];
    const float gi = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    a.g[p] = gi;
    const float mi = b1 * m[p] + (1.0f - b1) * gi;
    const float vi = b2 * v[p] + (1.0f - b2) * gi * gi;
    m[p] = mi;
    v[p] = vi;
    const float wi = w[p] - lr_t * mi / (sqrtf(vi) + eps);
    w[p] = wi;
    wb[p] = (__bf16)wi;
  }
}

__global__ void k_to_bf16(const float* w, __bf16* wb, int n) {
  for (int p = blockIdx.x * blockDim.x + threadIdx.x; p < n; p += gridDim.x * blockDim.x) wb[p] = (__bf16)w[p];
}

size_t fwd_lds_bytes() { return (size_t)(kHid * kLd1 + 2 * kHid * kLd + kNo * kLd + kTile * kLd) * sizeof(__bf16); }

int launch_forward(crv_mlp* h, const float* x, const float* t, float* y, long long n, bool train) {
  crv_ctx* ctx = h->ctx;
  FwdArgs a{};
  a.x = x; a.t = t; a.y = y; a.xa = h->xa; a.h1 = h->h[0]; a.h2 = h->h[1]; a.h3 = h->h[2]; a.dy = h->dy;
  a.loss = h->loss; a.w = h->wb; a.n = n; a.nm = h->norm; a.train = train ? 1 : 0;
  a.grad_scale = (float)(2.0 / (3.0 * (double)n));
  const long long ntiles = (n + kTile - 1) / kTile;
  const int nb = (int)std::max<long long>(1, std::min<long long>(ntiles, ctx->num_cu));
  const size_t lds = fwd_lds_bytes();
  static bool attr_set = false;
  if (!attr_set) {
    CRV_HIP_CHECK(ctx, hipFuncSetAttribute((const void*)k_mlp_forward, hipFuncAttributeMaxDynamicSharedMemorySize,
                                           (int)lds));
    attr_set = true;
  }
  ProfScope ps(ctx, train ? "mlp_forward_train" : "mlp_forward");
  hipLaunchKernelGGL(k_mlp_forward, dim3(nb), dim3(kT), lds, ctx->stream, a);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  // MFMA work as issued (padded shapes): 2 * rows * (16*128 + 2*144*128 + 144*32)
  h->flops_fwd = 2.0 * (double)(ntiles * kTile) * (kK1 * kHid + 2.0 * kKa * kHid + kKa * kNo);
  return CRV_OK;
}

// a product with its grid: 64 x 64 output tiles, `slices` K slices
struct GemmLaunch {
  GemmArgs g;
  int gx, gy, gz;
};

GemmLaunch make_gemm(GemmArgs g, int slices) {
  auto al16 = [](const void* p) { return ((uintptr_t)p & 15) == 0; };
  g.vec = al16(g.A) && al16(g.B) && g.lda % 8 == 0 && g.ldb % 8 == 0;
  return GemmLaunch{g, (g.M + 63) / 64, (g.N + 63) / 64, slices};
}

int launch_gemm(crv_ctx* ctx, const GemmLaunch& l, const char* name) {
  ProfScope ps(ctx, name);
  hipLaunchKernelGGL(k_gemm, dim3(l.gx, l.gy, l.gz), dim3(kGT), 0, ctx->stream, l.g);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// two independent products in one launch
int launch_gemm_pair(crv_ctx* ctx, const GemmLaunch& a, const GemmLaunch& b, const char* name) {
  const int nb = a.gx * a.gy * a.gz + b.gx * b.gy * b.gz;
  ProfScope ps(ctx, name);
  hipLaunchKernelGGL(k_gemm_pair, dim3(nb), dim3(kGT), 0, ctx->stream, a.g, a.gx, a.gy, a.gz, b.g, b.gx, b.gy);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  return CRV_OK;
}

// dW[M=out][N=in] partials = dZ^T (rows x out) . H (rows x in)
GemmLaunch weight_grad(const __bf16* dz, int ldz, int nout, const __bf16* hin, int ldh, int nin, long long n,
                       float* P) {
  GemmArgs g{};
  g.A = dz; g.lda = ldz; g.a_trans = 1;
  g.B = hin; g.ldb = ldh; g.b_rowk = 1;
  g.M = nout; g.N = nin; g.K = n; g.kchunk = kKChunk; g.mode = 0; g.P = P;
  return make_gemm(g, (int)((n + kKChunk - 1) / kKChunk));
}

// dZ_prev[rows][128] = (dZ . W[:, :128]) * (1 - H_prev^2)
GemmLaunch back_prop(const __bf16* dz, int ldz, int kdim, const __bf16* w, int ldw, const __bf16* hprev, __bf16* out,
                     long long n) {
  GemmArgs g{};
  g.A = dz; g.lda = ldz; g.a_trans = 0;
  g.B = w; g.ldb = ldw; g.b_rowk = 1;
  g.M = (int)n; g.N = kHid; g.K = kdim; g.kchunk = kdim; g.mode = 1;
  g.C = out; g.ldc = kHid; g.H = hprev; g.ldh = kKa;
  return make_gemm(g, 1);
}

// forward (training) + backward products: the per-slice partial gradients,
// and in r what sums them.  Each layer's weight gradient and the next
// layer's back-propagation share a launch (they read the same dZ and write
// different arrays), so the chain is the forward pass, three paired
// launches and the first layer's weight gradient.
int partials(crv_mlp* h, const float* x, const float* t, long long n, ReduceArgs& r) {
  crv_ctx* ctx = h->ctx;
  if (n < 1 || n > h->max_batch) return crv_set_error(ctx, CRV_E_INVALID, "mlp: batch size out of range");
  CRV_HIP_CHECK(ctx, hipMemsetAsync(h->loss, 0, sizeof(float), ctx->stream));
  int rc;
  if ((rc = launch_forward(h, x, t, nullptr, n, true)) != CRV_OK) return rc;
  const int slices = (int)((n + kKChunk - 1) / kKChunk);
  float* P4 = h->part;
  float* P3 = P4 + (size_t)slices * kNo * kKa;
  float* P2 = P3 + (size_t)slices * kHid * kKa;
  float* P1 = P2 + (size_t)slices * kHid * kKa;
  const __bf16* W2 = h->wb + kOffW2;
  const __bf16* W3 = h->wb + kOffW3;
  const __bf16* W4 = h->wb + kOffW4;
  if ((rc = launch_gemm_pair(ctx, weight_grad(h->dy, kNo, kNo, h->h[2], kKa, kKa, n, P4),
                              back_prop(h->dy, kNo, kNo, W4, kKa, h->h[2], h->dz[0], n), "mlp_dW4_dZ3")) != CRV_OK)
    return rc;
  if ((rc = launch_gemm_pair(ctx, weight_grad(h->dz[0], kHid, kHid, h->h[1], kKa, kKa, n, P3),
                              back_prop(h->dz[0], kHid, kHid, W3, kKa, h->h[1], h->dz[1], n), "mlp_dW3_dZ2")) !=
      CRV_OK)
    return rc;
  if ((rc = launch_gemm_pair(ctx, weight_grad(h->dz[1], kHid, kHid, h->h[0], kKa, kKa, n, P2),
                              back_prop(h->dz[1], kHid, kHid, W2, kKa, h->h[0], h->dz[0], n), "mlp_dW2_dZ1")) !=
      CRV_OK)
    return rc;
  if ((rc = launch_gemm(ctx, weight_grad(h->dz[0], kHid, kHid, h->xa, kK1, kK1, n, P1), "mlp_dW1")) != CRV_OK)
    return rc;
  r = ReduceArgs{};
  r.s[0] = Seg{kOffW1, kHid * kK1, slices, P1};
  r.s[1] = Seg{kOffW2, kHid * kKa, slices, P2};
  r.s[2] = Seg{kOffW3, kHid * kKa, slices, P3};
  r.s[3] = Seg{kOffW4, kNo * kKa, slices, P4};
  r.g = h->g;
  const double rows = (double)slices * kKChunk;
  // padded MFMA work: weight gradients over every row, back-propagation
  // products over 64-row tiles
  h->flops_bwd = 2.0 * rows * ((double)kNo * kKa + 2.0 * kHid * kKa + kHid * kK1) +
                 2.0 * (double)((n + 63) / 64 * 64) * ((double)kNo * kHid + 2.0 * kHid * kHid);
  return CRV_OK;
}

// the summed gradient in h->g
int gradients(crv_mlp* h, const float* x, const float* t, long long n) {
  ReduceArgs r;
  int rc = partials(h, x, t, n, r);
  if (rc != CRV_OK) return rc;
  ProfScope ps(h->ctx, "mlp_reduce");
  hipLaunchKernelGGL(k_reduce, dim3(85), dim3(512), 0, h->ctx->stream, r);
  CRV_HIP_CHECK(h->ctx, hipGetLastError());
  return CRV_OK;
}

}  // namespace

extern "C" {

long long crv_mlp_param_count(void) { return kParams; }

int crv_mlp_create(crv_ctx* ctx, const crv_mlp_norm* norm, long long max_batch, crv_mlp** out) {
  if (!ctx) return CRV_E_INVALID;
  if (!norm || !out || max_batch < 1) return crv_set_error(ctx, CRV_E_INVALID, "mlp_create: bad argument");
  *out = nullptr;
  crv_mlp* h = new (std::nothrow) crv_mlp();
  if (!h) return CRV_E_NOMEM;
  h->ctx = ctx;
  h->norm = *norm;
  h->max_batch = max_batch;
  h->slices = (int)((max_batch + kKChunk - 1) / kKChunk);
  const size_t B = (size_t)max_batch;
  const size_t pbytes = (size_t)kParams * sizeof(float);
  const size_t part = (size_t)h->slices * kParams * sizeof(float);
  bool ok = hipMalloc(&h->w, pbytes) == hipSuccess && hipMalloc(&h->m, pbytes) == hipSuccess &&
            hipMalloc(&h->v, pbytes) == hipSuccess && hipMalloc(&h->g, pbytes) == hipSuccess &&
            hipMalloc(&h->wb, (size_t)kParams * 2) == hipSuccess && hipMalloc(&h->xa, B * kK1 * 2) == hipSuccess &&
            hipMalloc(&h->dy, B * kNo * 2) == hipSuccess && hipMalloc(&h->part, part) == hipSuccess &&
            hipMalloc(&h->loss, 256) == hipSuccess;
  for (int l = 0; l < 3 && ok; ++l) ok = hipMalloc(&h->h[l], B * kKa * 2) == hipSuccess;
  for (int l = 0; l < 2 && ok; ++l) ok = hipMalloc(&h->dz[l], B * kHid * 2) == hipSuccess;
  if (!ok) {
    crv_mlp_destroy(h);
    return crv_set_error(ctx, CRV_E_NOMEM, "mlp_create: device allocation failed");
  }
  if (hipMemset(h->w, 0, pbytes) != hipSuccess || hipMemset(h->m, 0, pbytes) != hipSuccess ||
      hipMemset(h->v, 0, pbytes) != hipSuccess || hipMemset(h->wb, 0, (size_t)kParams * 2) != hipSuccess) {
    crv_mlp_destroy(h);
    return crv_set_error(ctx, CRV_E_HIP, "mlp_create: memset failed");
  }
  *out = h;
  return CRV_OK;
}

int crv_mlp_destroy(crv_mlp* h) {
  if (!h) return CRV_E_INVALID;
  if (h->ctx && h->ctx->stream) (void)hipStreamSynchronize(h->ctx->stream);
  void* bufs[] = {h->w, h->m, h->v, h->g, h->wb, h->xa, h->h[0], h->h[1], h->h[2], h->dy, h->dz[0], h->dz[1],
                  h->part, h->loss};
  for (void* b : bufs)
    if (b) (void)hipFree(b);
  delete h;
  return CRV_OK;
}

int crv_mlp_set_params(crv_mlp* h, const float* params) {
  if (!h || !params) return CRV_E_INVALID;
  crv_ctx* ctx = h->ctx;
  const size_t pbytes = (size_t)kParams * sizeof(float);
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(h->w, params, pbytes, hipMemcpyHostToDevice, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(h->m, 0, pbytes, ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(h->v, 0, pbytes, ctx->stream));
  hipLaunchKernelGGL(k_to_bf16, dim3(85), dim3(512), 0, ctx->stream, (const float*)h->w, h->wb, kParams);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return CRV_OK;
}

int crv_mlp_get_params(crv_mlp* h, float* params) {
  if (!h || !params) return CRV_E_INVALID;
  crv_ctx* ctx = h->ctx;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(params, h->w, (size_t)kParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return CRV_OK;
}

int crv_mlp_gradients(crv_mlp* h, const float* d_x, const float* d_t, long long n, float* grad, float* loss) {
  if (!h || !d_x || !d_t) return CRV_E_INVALID;
  crv_ctx* ctx = h->ctx;
  int rc = gradients(h, d_x, d_t, n);
  if (rc != CRV_OK) return rc;
  if (grad)
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(grad, h->g, (size_t)kParams * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  float l = 0.0f;
  CRV_HIP_CHECK(ctx, hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  if (loss) *loss = l / (3.0f * (float)n);
  return CRV_OK;
}

int crv_mlp_train_step(crv_mlp* h, const float* d_x, const float* d_t, long long n, float lr, int step,
                       float* loss) {
  if (!h || !d_x || !d_t || step < 1) return CRV_E_INVALID;
  crv_ctx* ctx = h->ctx;
  ReduceArgs r;
  int rc = partials(h, d_x, d_t, n, r);
  if (rc != CRV_OK) return rc;
  const double b1 = 0.9, b2 = 0.999;
  const float lr_t = (float)(lr * std::sqrt(1.0 - std::pow(b2, step)) / (1.0 - std::pow(b1, step)));
  {
    ProfScope ps(ctx, "mlp_reduce_adam");
    hipLaunchKernelGGL(k_reduce_adam, dim3(85), dim3(512), 0, ctx->stream, r, h->w, h->m, h->v, h->wb, lr_t,
                       (float)b1, (float)b2, 1e-8f);
    CRV_HIP_CHECK(ctx, hipGetLastError());
  }
  if (loss) {
    float l = 0.0f;
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    *loss = l / (3.0f * (float)n);
  }
  return CRV_OK;
}

int crv_mlp_predict(crv_mlp* h, const float* d_x, long long n, float* d_y) {
  if (!h || !d_x || !d_y || n < 0) return CRV_E_INVALID;
  if (n == 0) return CRV_OK;
  return launch_forward(h, d_x, nullptr, d_y, n, false);
}

int crv_mlp_flops(crv_mlp* h, double* fwd, double* bwd) {
  if (!h) return CRV_E_INVALID;
  if (fwd) *fwd = h->flops_fwd;
  if (bwd) *bwd = h->flops_bwd;
  return CRV_OK;
}

}  // extern "C"
