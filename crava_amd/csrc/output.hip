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
//
// crv_grid_write_storm writes FFTGrid::writeStormFile's STORM files (binary
// or text) through the same staging.
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

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

// ---- STORM grid files --------------------------------------------------------
// FFTGrid::writeStormFile (src/fftgrid.cpp:2117-2176) with the header of
// Simbox::getStormHeader (src/simbox.cpp:408-458): reals as NRLib::ToString
// (value, 6) prints them (fixed notation, 6 decimals), the values as
// big-endian floats or as text (std::fixed with 6 decimals, or
// std::scientific), one text line per x row, then the line "0".

namespace {

std::string fixed6(double v) {
  char b[64];
  snprintf(b, sizeof(b), "%.6f", v);
  return b;
}

std::string storm_header(const crv_storm_info* si, int cubetype, int nx, int ny, int nz, bool ascii) {
  const double pi = 3.14159265358979323846;
  std::string h = ascii ? "storm_petro_ascii\n" : "storm_petro_binary\n";
  h += "0 " + std::to_string(cubetype) + " " + fixed6(kRMissing) + "\n";
  h += "FFTGrid\n";
  const std::string xy = fixed6(si->xmin) + " " + fixed6(si->lx) + " " + fixed6(si->ymin) + " " + fixed6(si->ly);
  if (si->flat)
    h += xy + " 0.0 " + fixed6(si->lz) + " 0.0 0.0\n";
  else
    h += xy + " " + si->top_name + " " + si->bot_name + " 0.0 0.0\n";
  h += fixed6(si->lz) + " " + fixed6(si->angle * 180 / pi) + "\n\n";
  h += std::to_string(nx) + " " + std::to_string(ny) + " " + std::to_string(nz) + "\n";
  return h;
}

}  // namespace

extern "C" int crv_storm_header(const crv_storm_info* si, int cubetype, int nx, int ny, int nz, int ascii, char* buf,
                                size_t len) {
  if (!si || (!si->flat && (!si->top_name || !si->bot_name)) || nx < 1 || ny < 1 || nz < 1) return CRV_E_INVALID;
  const std::string h = storm_header(si, cubetype, nx, ny, nz, ascii != 0);
  if (!buf) return (int)h.size();
  if (len < h.size()) return CRV_E_INVALID;
  memcpy(buf, h.data(), h.size());
  return (int)h.size();
}

extern "C" int crv_grid_write_storm(crv_grid* g, const char* path, const crv_storm_info* si, int ascii, int padding,
                                    int scientific) {
  if (!g || !path || !si) return CRV_E_INVALID;
  crv_ctx* ctx = g->ctx;
  if (!si->flat && (!si->top_name || !si->bot_name))
    return crv_set_error(ctx, CRV_E_INVALID, "write_storm: top and base surface names needed (flat = 0)");
  if (g->precision != CRV_FP32) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "write_storm: FP32 grids only");
  if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "write_storm: grid is transformed");
  const int nx = padding ? g->nxp : g->nx, ny = padding ? g->nyp : g->ny, nz = padding ? g->nzp : g->nz;
  const size_t rnxp = 2 * (size_t)(g->nxp / 2 + 1);
  const std::string head = storm_header(si, g->cubetype, nx, ny, nz, ascii != 0);
  FILE* f = fopen(path, ascii ? "w" : "wb");
  if (!f) return crv_set_error(ctx, CRV_E_INVALID, std::string("write_storm: cannot open ") + path);
  Staging st;
  hipError_t e = st.init();
  bool io_ok = fwrite(head.data(), 1, head.size(), f) == head.size();
  // rows r = j + nyp k of the planes k < nz, kChunk bytes of whole rows per
  // chunk; rows j >= ny of a plane are staged and skipped
  const size_t row_bytes = rnxp * sizeof(float);
  const size_t nrows = (size_t)g->nyp * nz, rows_per = std::max<size_t>(1, kChunk / row_bytes);
  const size_t nchunk = (nrows + rows_per - 1) / rows_per;
  if (row_bytes > kChunk) e = hipErrorInvalidValue;
  auto queue = [&](size_t c) {
    const size_t r0 = c * rows_per, n = std::min(rows_per, nrows - r0);
    hipError_t r = hipMemcpyAsync(st.buf[c & 1], (const char*)g->d + r0 * row_bytes, n * row_bytes,
                                  hipMemcpyDeviceToHost, ctx->stream);
    if (r == hipSuccess) r = hipEventRecord(st.ev[c & 1], ctx->stream);
    return r;
  };
  std::vector<unsigned char> out;
  if (e == hipSuccess && nchunk) e = queue(0);
  for (size_t c = 0; c < nchunk && e == hipSuccess && io_ok; ++c) {
    if (c + 1 < nchunk && (e = queue(c + 1)) != hipSuccess) break;
    if ((e = hipEventSynchronize(st.ev[c & 1])) != hipSuccess) break;
    const size_t r0 = c * rows_per, n = std::min(rows_per, nrows - r0);
    const float* b = (const float*)st.buf[c & 1];
    out.clear();
    char txt[64];
    for (size_t q = 0; q < n; ++q) {
      if ((int)((r0 + q) % (size_t)g->nyp) >= ny) continue;
      const float* row = b + q * rnxp;
      for (int i = 0; i < nx; ++i) {
        if (ascii) {
          const int m = snprintf(txt, sizeof(txt), scientific ? "%e" : "%.6f", (double)row[i]);
          out.insert(out.end(), txt, txt + m);
          out.push_back(i + 1 < nx ? ' ' : '\n');
        } else {
          unsigned char v[4];
          put_be(v, &row[i], 4);
          out.insert(out.end(), v, v + 4);
        }
      }
    }
    io_ok = fwrite(out.data(), 1, out.size(), f) == out.size();
  }
  if (io_ok) io_ok = fwrite("0\n", 1, 2, f) == 2;
  const hipError_t e2 = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = e2;
  io_ok = (fclose(f) == 0) && io_ok;
  if (e != hipSuccess) return crv_set_error(ctx, CRV_E_HIP, std::string("write_storm: ") + hipGetErrorString(e));
  if (!io_ok) return crv_set_error(ctx, CRV_E_INVALID, std::string("write_storm: write failed: ") + path);
  return CRV_OK;
}
