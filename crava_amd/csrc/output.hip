// Output of inverted grids: derived elastic parameters and CRAVA's binary
// grid format.
//
// crv_grid_derived replaces ParameterOutput's per-voxel loops (reference
// src/parameteroutput.cpp:14-420: expTransf of the log parameters,
// acoustic and shear impedance, vp/vs, Poisson's ratio, Lame parameters,
// mu rho and lambda rho) with one elementwise kernel.  The reference
// evaluates each formula in double from the float log values and stores a
// float; the kernel does the same.  RMISSING stays RMISSING in the
// exponential of one grid (FFTGrid::expTransf, src/fftgrid.cpp:1726-1743).
//
// crv_grid_write_crava / crv_grid_read_crava write and read the format of
// FFTGrid::writeCravaFile / readCravaFile (src/fftgrid.cpp:2411-2498): the
// line "crava_fftgrid_binary", the simbox geometry (4 doubles, 2 ints, 7
// doubles, as listed in crv_simbox_info), rnxp, nyp, nzp and the rsize
// floats of the padded real grid, every number big-endian IEEE (NRLib's
// default, libs/nrlib/iotools/fileio.hpp:266-288).  The grid moves through
// two pinned staging buffers: the copy of one chunk overlaps the byte
// swapping and file I/O of the other.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace {

constexpr int kThreads = 256;
constexpr double kRMissing = -99999.0;

template <typename T>
__global__ void k_derived(int kind, const T* __restrict__ a, const T* __restrict__ b, const T* __restrict__ r,
                          T* out, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const double A = a ? (double)a[i] : 0.0, B = b ? (double)b[i] : 0.0, R = r ? (double)r[i] : 0.0;
    double v;
    switch (kind) {
      case CRV_OUT_EXP:        v = (A == kRMissing) ? kRMissing : exp(A); break;
      case CRV_OUT_AI:         v = exp(A + R); break;
      case CRV_OUT_SI:         v = exp(B + R); break;
      case CRV_OUT_VPVS:       v = exp(A - B); break;
      case CRV_OUT_POISSON: {
        const double q = exp(2.0 * (A - B));
        v = 0.5 * (q - 2.0) / (q - 1.0);
        break;
      }
      case CRV_OUT_LAME_MU:    v = exp(R + 2.0 * B - 13.81551); break;
      case CRV_OUT_LAME_LAMBDA: v = exp(R) * (exp(2.0 * A - 13.81551) - 2.0 * exp(2.0 * B - 13.81551)); break;
      case CRV_OUT_MU_RHO:     v = exp(2.0 * (B + R) - 13.81551); break;
      default:                 v = exp(2.0 * (A + R) - 13.81551) - 2.0 * exp(2.0 * (B + R) - 13.81551); break;
    }
    out[i] = (T)v;
  }
}

// big-endian encoders of the format's numbers
void put_be(unsigned char* p, const void* v, int nbytes) {
  const unsigned char* s = (const unsigned char*)v;
  for (int i = 0; i < nbytes; ++i) p[i] = s[nbytes - 1 - i];
}
void get_be(const unsigned char* p, void* v, int nbytes) {
  unsigned char* d = (unsigned char*)v;
  for (int i = 0; i < nbytes; ++i) d[i] = p[nbytes - 1 - i];
}

const char kMagic[] = "crava_fftgrid_binary\n";
constexpr size_t kMagicLen = sizeof(kMagic) - 1;

// pinned staging: two buffers of kChunk bytes
constexpr size_t kChunk = (size_t)16 << 20;

struct Staging {
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  ~Staging() {
    for (int i = 0; i < 2; ++i) {
      if (buf[i]) (void)hipHostFree(buf[i]);
      if (ev[i]) (void)hipEventDestroy(ev[i]);
    }
  }
  hipError_t init() {
    for (int i = 0; i < 2; ++i) {
      hipError_t e = hipHostMalloc(&buf[i], kChunk, hipHostMallocDefault);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&ev[i], hipEventDisableTiming);
      if (e != hipSuccess) return e;
    }
    return hipSuccess;
  }
};

}  // namespace

extern "C" int crv_crava_header(const crv_simbox_info* sb, int rnxp, int nyp, int nzp, unsigned char* buf,
                                size_t len) {
  if (!sb) return CRV_E_INVALID;
  const size_t need = kMagicLen + 11 * 8 + 2 * 4 + 3 * 4;
  if (!buf) return (int)need;
  if (len < need) return CRV_E_INVALID;
  unsigned char* p = buf;
  memcpy(p, kMagic, kMagicLen); p += kMagicLen;
  auto d = [&](double v) { put_be(p, &v, 8); p += 8; };
  auto i4 = [&](int v) { put_be(p, &v, 4); p += 4; };
  d(sb->x0); d(sb->y0); d(sb->dx); d(sb->dy);
  i4(sb->nx); i4(sb->ny);
  d(sb->il0); d(sb->xl0); d(sb->il_step_x); d(sb->il_step_y); d(sb->xl_step_x); d(sb->xl_step_y); d(sb->angle);
  i4(rnxp); i4(nyp); i4(nzp);
  return (int)need;
}

extern "C" int crv_grid_derived(crv_ctx* ctx, int kind, crv_grid* alpha, crv_grid* beta, crv_grid* rho,
                                crv_grid* out) {
  if (!ctx || !out) return CRV_E_INVALID;
  if (kind < CRV_OUT_EXP || kind > CRV_OUT_LAMBDA_RHO) return crv_set_error(ctx, CRV_E_INVALID, "derived: bad kind");
  const bool na = kind == CRV_OUT_EXP || kind == CRV_OUT_AI || kind == CRV_OUT_VPVS || kind == CRV_OUT_POISSON ||
                  kind == CRV_OUT_LAME_LAMBDA || kind == CRV_OUT_LAMBDA_RHO;
  const bool nb = kind == CRV_OUT_SI || kind == CRV_OUT_VPVS || kind == CRV_OUT_POISSON || kind == CRV_OUT_LAME_MU ||
                  kind == CRV_OUT_LAME_LAMBDA || kind == CRV_OUT_MU_RHO || kind == CRV_OUT_LAMBDA_RHO;
  const bool nr = kind == CRV_OUT_AI || kind == CRV_OUT_SI || kind == CRV_OUT_LAME_MU || kind == CRV_OUT_LAME_LAMBDA ||
                  kind == CRV_OUT_MU_RHO || kind == CRV_OUT_LAMBDA_RHO;
  auto ok = [&](const crv_grid* g, bool need) {
    if (!need) return true;
    return g && g->nxp == out->nxp && g->nyp == out->nyp && g->nzp == out->nzp && g->precision == out->precision &&
           !g->transformed;
  };
  if (!ok(alpha, na) || !ok(beta, nb) || !ok(rho, nr))
    return crv_set_error(ctx, CRV_E_STATE, "derived: inputs missing, of another size or not real");
  const size_t n = out->rsize;
  const int nblk = (int)std::min<size_t>((n + kThreads - 1) / kThreads, 65536);
  ProfScope ps(ctx, "derived");
  if (out->precision == CRV_FP64)
    hipLaunchKernelGGL(k_derived<double>, dim3(nblk), dim3(kThreads), 0, ctx->stream, kind,
                       na ? (const double*)alpha->d : nullptr, nb ? (const double*)beta->d : nullptr,
                       nr ? (const double*)rho->d : nullptr, (double*)out->d, n);
  else
    hipLaunchKernelGGL(k_derived<float>, dim3(nblk), dim3(kThreads), 0, ctx->stream, kind,
                       na ? (const float*)alpha->d : nullptr, nb ? (const float*)beta->d : nullptr,
                       nr ? (const float*)rho->d : nullptr, (float*)out->d, n);
  CRV_HIP_CHECK(ctx, hipGetLastError());
  out->transformed = false;
  out->cubetype = CRV_PARAMETER;
  return CRV_OK;
}

extern "C" int crv_grid_write_crava(crv_grid* g, const char* path, const crv_simbox_info* sb) {
  if (!g || !path || !sb) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  if (g->precision != CRV_FP32) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "write_crava: FP32 grids only");
  if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "write_crava: grid is transformed");
  const int rnxp = 2 * (g->nxp / 2 + 1);
  unsigned char head[256];
  const int hl = crv_crava_header(sb, rnxp, g->nyp, g->nzp, head, sizeof(head));
  if (hl < 0) return crv_set_error(ctx, CRV_E_INVALID, "write_crava: header");
  FILE* f = fopen(path, "wb");
  if (!f) return crv_set_error(ctx, CRV_E_INVALID, std::string("write_crava: cannot open ") + path);
  Staging st;
  hipError_t e = st.init();
  bool io_ok = fwrite(head, 1, (size_t)hl, f) == (size_t)hl;
  const size_t total = g->rsize * sizeof(float);
  const size_t nchunk = (total + kChunk - 1) / kChunk;
  // chunk c goes to buffer c % 2; its copy is queued before chunk c - 1 is
  // swapped and written
  auto queue = [&](size_t c) {
    const size_t off = c * kChunk, len = std::min(kChunk, total - off);
    hipError_t r = hipMemcpyAsync(st.buf[c & 1], (const char*)g->d + off, len, hipMemcpyDeviceToHost, ctx->stream);
    if (r == hipSuccess) r = hipEventRecord(st.ev[c & 1], ctx->stream);
    return r;
  };
  if (e == hipSuccess && nchunk) e = queue(0);
  for (size_t c = 0; c < nchunk && e == hipSuccess && io_ok; ++c) {
    if (c + 1 < nchunk && (e = queue(c + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(st.ev[c & 1])) != hipSuccess) break;
    const size_t len = std::min(kChunk, total - c * kChunk);
    unsigned char* b = (unsigned char*)st.buf[c & 1];
    for (size_t i = 0; i + 4 <= len; i += 4) {
      std::swap(b[i], b[i + 3]);
      std::swap(b[i + 1], b[i + 2]);
    }
    io_ok = fwrite(b, 1, len, f) == len;
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  io_ok = (fclose(f) == 0) && io_ok;
  if (e != hipSuccess) return crv_set_error(ctx, CRV_E_HIP, std::string("write_crava: ") + hipGetErrorString(e));
  if (!io_ok) return crv_set_error(ctx, CRV_E_INVALID, std::string("write_crava: write failed: ") + path);
  return CRV_OK;
}

extern "C" int crv_grid_read_crava(crv_grid* g, const char* path, crv_simbox_info* sb) {
  if (!g || !path) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  if (g->precision != CRV_FP32) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "read_crava: FP32 grids only");
  FILE* f = fopen(path, "rb");
  if (!f) return crv_set_error(ctx, CRV_E_INVALID, std::string("read_crava: cannot open ") + path);
  unsigned char head[256];
  const size_t hl = kMagicLen + 11 * 8 + 2 * 4 + 3 * 4;
  if (fread(head, 1, hl, f) != hl || memcmp(head, kMagic, kMagicLen) != 0) {
    fclose(f);
    return crv_set_error(ctx, CRV_E_INVALID, std::string("read_crava: not a CRAVA binary grid: ") + path);
  }
  const unsigned char* p = head + kMagicLen;
  crv_simbox_info s{};
  auto d = [&](double& v) { get_be(p, &v, 8); p += 8; };
  auto i4 = [&](int& v) { get_be(p, &v, 4); p += 4; };
  d(s.x0); d(s.y0); d(s.dx); d(s.dy); i4(s.nx); i4(s.ny);
  d(s.il0); d(s.xl0); d(s.il_step_x); d(s.il_step_y); d(s.xl_step_x); d(s.xl_step_y); d(s.angle);
  int rnxp, nyp, nzp;
  i4(rnxp); i4(nyp); i4(nzp);
  if (rnxp != 2 * (g->nxp / 2 + 1) || nyp != g->nyp || nzp != g->nzp) {
    fclose(f);
    return crv_set_error(ctx, CRV_E_INVALID, "read_crava: the grid on file has other dimensions");
  }
  if (sb) *sb = s;
  Staging st;
  hipError_t e = st.init();
  bool io_ok = true;
  const size_t total = g->rsize * sizeof(float);
  const size_t nchunk = (total + kChunk - 1) / kChunk;
  for (size_t c = 0; c < nchunk && e == hipSuccess; ++c) {
    // buffer c % 2 is free once its previous copy (chunk c - 2) is done
    if (c >= 2 && (e = hipEventSynchronize(st.ev[c & 1])) != hipSuccess) break;
    const size_t off = c * kChunk, len = std::min(kChunk, total - off);
    unsigned char* b = (unsigned char*)st.buf[c & 1];
    if (fread(b, 1, len, f) != len) { io_ok = false; break; }
    for (size_t i = 0; i + 4 <= len; i += 4) {
      std::swap(b[i], b[i + 3]);
      std::swap(b[i + 1], b[i + 2]);
    }
    e = hipMemcpyAsync((char*)g->d + off, b, len, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(st.ev[c & 1], ctx->stream);
  }
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  fclose(f);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) return crv_set_error(ctx, CRV_E_HIP, std::string("read_crava: ") + hipGetErrorString(e));
  if (!io_ok) return crv_set_error(ctx, CRV_E_INVALID, std::string("read_crava: file too short: ") + path);
  g->transformed = false;
  return CRV_OK;
}
