// Prediction kriging of the posterior means to well data on gfx950
// (Crava::doPostKriging, reference src/crava.cpp:1736-1765).
//
// Host side, as CKrigingAdmin plans it (src/krigingadmin.cpp):
//  * one separable covariance table per posterior covariance grid
//    (CovGridSeparated(const FFTGrid&), src/covgridseparated.cpp:16-56): the
//    plane gamma_xy(i, j) = C(i, j, 0) and the column
//    gamma_z(k) = C(0, 0, k) / C(0, 0, k0), k0 where |C(0, 0, k)| stops rising;
//  * tapering of all six by Gaussians whose ranges come from the three
//    auto-covariances (findTaperRanges / performTapering,
//    covgridseparated.cpp:223-328), then integer ranges where |gamma|
//    falls below 5 % (EstimateRange*, covgridseparated.cpp:146-221);
//  * the block size and data neighbourhood (EstimateSizeOfBlock,
//    krigingadmin.cpp:768-939), and per block the data box grown or shrunk
//    by bisection until it holds the data target within 10 %
//    (FindDataInDataBlockLoop, krigingadmin.cpp:441-575);
//  * per block the kriging system K x = r over the box's data (SetMatrix,
//    krigingadmin.cpp:592-703, upper triangle), solved by Cholesky in FP64
//    (the reference calls LAPACK posv through NRLib::CholeskySolve).
// Device side: every cell of every block adds k(cell) . x to its trend,
// k(cell) the covariances between the cell and the block's data
// (SetKrigVector, krigingadmin.cpp:705-766; KrigBlock, 515-556).
#include "internal.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

constexpr float kMissing = -99999.0f;  // RMISSING (src/definitions.h)
constexpr int kTolPercent = 10;        // maxDataTolerance_ (krigingadmin.h:97)

// ---- separable covariance (CovGridSeparated) --------------------------------

struct CovSep {
  int nxp = 0, nyp = 0, nzp = 0;
  std::vector<float> xy, z;
  float g2(int i1, int j1, int k1, int i2, int j2, int k2) const {  // GetGamma2
    int di = i2 - i1, dj = j2 - j1, dk = k2 - k1;
    if (std::abs(di) >= nxp / 2 || std::abs(dj) >= nyp / 2 || std::abs(dk) >= nzp / 2) return 0.0f;
    if (di < 0) di += nxp;
    if (dj < 0) dj += nyp;
    if (dk < 0) dk += nzp;
    return xy[di + dj * nxp] * z[dk];
  }
};

// the table of one real covariance grid (values read from the device)
int cov_sep_from_grid(crv_grid* g, CovSep& c) {
  crv_ctx* ctx = g->ctx;
  c.nxp = g->nxp; c.nyp = g->nyp; c.nzp = g->nzp;
  c.xy.assign((size_t)c.nxp * c.nyp, 0.0f);
  c.z.assign(c.nzp, 0.0f);
  const size_t eb = g->elem_bytes();
  std::vector<char> plane((size_t)g->rnxp * g->nyp * eb);
  CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  CRV_HIP_CHECK(ctx, hipMemcpy(plane.data(), g->d, plane.size(), hipMemcpyDeviceToHost));
  auto val = [&](const char* base, size_t idx) -> float {
    return eb == 8 ? (float)((const double*)base)[idx] : ((const float*)base)[idx];
  };
  for (int j = 0; j < c.nyp; ++j)
    for (int i = 0; i < c.nxp; ++i) c.xy[i + (size_t)j * c.nxp] = val(plane.data(), i + (size_t)g->rnxp * j);
  std::vector<float> col(c.nzp);
  const size_t kstride = (size_t)g->rnxp * g->nyp * eb;
  std::vector<char> colv((size_t)c.nzp * eb);
  CRV_HIP_CHECK(ctx, hipMemcpy2D(colv.data(), eb, g->d, kstride, eb, c.nzp, hipMemcpyDeviceToHost));
  for (int k = 0; k < c.nzp; ++k) col[k] = val(colv.data(), k);
  // the first local maximum of |C(0, 0, k)| from k = 0 (its signed value
  // normalises the column)
  float maxval = std::fabs(col[0]);
  int i = 1;
  float value = c.nzp > 1 ? std::fabs(col[1]) : 0.0f;
  while (value > maxval && i < c.nzp - 1) {
    maxval = std::fabs(col[i]);
    ++i;
    value = std::fabs(col[i]);
  }
  maxval = col[i - 1];
  if (maxval == 0.0f) return crv_set_error(ctx, CRV_E_INVALID, "kriging: covariance grid is zero at the origin");
  for (int k = 0; k < c.nzp; ++k) c.z[k] = col[k] / maxval;
  return CRV_OK;
}

void estimate_ranges(const CovSep& c, int& rx, int& ry, int& rz) {
  rx = ry = rz = 0;
  const float f = 0.05f;
  const float sill = f * c.xy[0];
  const int nx2 = c.nxp / 2, ny2 = c.nyp / 2;
  auto at = [&](int i, int j) { return std::fabs(c.xy[i + (size_t)j * c.nxp]); };
  [&] {
    for (int i = nx2 - 1; i >= 0; --i) {
      for (int j = ny2 - 1; j >= 0; --j)
        if (at(i, j) >= sill) { rx = i; return; }
      for (int j = ny2; j < c.nyp; ++j)
        if (at(i, j) >= sill) { rx = i; return; }
    }
  }();
  [&] {
    for (int j = ny2 - 1; j >= 0; --j) {
      for (int i = nx2 - 1; i >= 0; --i)
        if (at(i, j) >= sill) { ry = j; return; }
      for (int i = nx2; i < c.nxp; ++i)
        if (at(i, j) >= sill) { ry = j; return; }
    }
  }();
  for (int k = 0; k <= c.nzp / 2; ++k)
    if (std::fabs(c.z[k]) <= 0.05f) { rz = k - 1; break; }
}

// sum of gamma_xy along a ray at angle a, relative to gamma_xy(0, 0)
// (angleSum, covgridseparated.cpp:244-278, its stepping kept as it is)
float angle_sum(const CovSep& c, float angle) {
  const float cosang = std::cos(angle), sinang = std::sin(angle);
  float di, dj;
  if (std::fabs(cosang) > std::fabs(sinang)) {
    di = 1.0f;
    dj = sinang / cosang;
  } else {
    dj = cosang / sinang;
    di = 1.0f;
  }
  const double maxi = c.nxp / 2 + 1, maxj = c.nyp / 2 + 1;
  float fi = 0.0f, fj = 0.0f, sum = 0.0f;
  while (std::fabs(fi) < maxi && std::fabs(fj) < maxj) {
    const int i = fi >= 0 ? (int)std::floor(fi) : c.nxp + (int)std::floor(fi);
    const int j = fj >= 0 ? (int)std::floor(fj) : c.nyp + (int)std::floor(fj);
    sum += c.xy[i + (size_t)j * c.nxp];
    fi += di;
    fj += dj;
  }
  return sum / c.xy[0];
}

void taper_ranges(const CovSep& c, float& rx, float& ry, float& rz) {
  float sum = 0.0f;
  for (int k = 0; k < c.nzp / 2; ++k) sum += std::fabs(c.z[k]);
  rz = 2.5f * sum;
  const double pi = 3.14159265358979323846;
  float angle = 0.0f;
  const float da = (float)pi / 20.0f;
  rx = ry = 0.0f;
  while (angle < pi) {
    sum = 2.5f * angle_sum(c, angle);
    if (sum * std::cos(angle) > rx) rx = sum * std::cos(angle);
    if (sum * std::sin(angle) > ry) ry = sum * std::sin(angle);
    angle += da;
  }
}

void taper(CovSep& c, float rx, float ry, float rz) {
  const float sq3 = std::sqrt(3.0f);
  auto w = [&](int n, float r) { return (float)std::exp(-std::pow((float)n / r * sq3, 2)); };
  const int nz2 = (int)std::ceil((float)c.nzp / 2.0f);
  for (int k = 0; k < nz2; ++k) {
    c.z[k] *= w(k, rz);
    c.z[c.nzp - 1 - k] *= w(k + 1, rz);
  }
  const int nx2 = (int)std::ceil((float)c.nxp / 2.0f), ny2 = (int)std::ceil((float)c.nyp / 2.0f);
  for (int j = 0; j < c.nyp; ++j)
    for (int i = 0; i < nx2; ++i) {
      c.xy[i + (size_t)j * c.nxp] *= w(i, rx);
      c.xy[(c.nxp - i - 1) + (size_t)j * c.nxp] *= w(i + 1, rx);
    }
  for (int i = 0; i < c.nxp; ++i)
    for (int j = 0; j < ny2; ++j) {
      c.xy[i + (size_t)j * c.nxp] *= w(j, ry);
      c.xy[i + (size_t)(c.nyp - j - 1) * c.nxp] *= w(j + 1, ry);
    }
}

// ---- boxes (CBox, src/box.cpp) ----------------------------------------------

struct Box {
  int i0 = 0, j0 = 0, k0 = 0, i1 = 0, j1 = 0, k1 = 0;
  bool inside(int i, int j, int k) const { return i0 <= i && i <= i1 && j0 <= j && j <= j1 && k0 <= k && k <= k1; }
  bool operator==(const Box& b) const {
    return i0 == b.i0 && j0 == b.j0 && k0 == b.k0 && i1 == b.i1 && j1 == b.j1 && k1 == b.k1;
  }
};

Box make_box(int a0, int b0, int c0, int a1, int b1, int c1, const int* dims) {
  Box x;
  x.i0 = std::min(a0, a1); x.i1 = std::max(a0, a1);
  x.j0 = std::min(b0, b1); x.j1 = std::max(b0, b1);
  x.k0 = std::min(c0, c1); x.k1 = std::max(c0, c1);
  if (dims) {
    auto tr = [&](int& i, int& j, int& k) {
      i = std::min(std::max(i, 0), dims[0] - 1);
      j = std::min(std::max(j, 0), dims[1] - 1);
      k = std::min(std::max(k, 0), dims[2] - 1);
    };
    tr(x.i0, x.j0, x.k0);
    tr(x.i1, x.j1, x.k1);
  }
  return x;
}

// ModifyBox without ranges or truncation: the box halfway to s
void move_half(Box& b, const Box& s) {
  b.i1 = (int)((b.i1 + s.i1) / 2.0f); b.j1 = (int)((b.j1 + s.j1) / 2.0f); b.k1 = (int)((b.k1 + s.k1) / 2.0f);
  b.i0 = (int)((b.i0 + s.i0) / 2.0f); b.j0 = (int)((b.j0 + s.j0) / 2.0f); b.k0 = (int)((b.k0 + s.k0) / 2.0f);
}

// ---- well data (KrigingData3D / CBWellPt) -----------------------------------

struct WellPt {
  int i, j, k;
  float v[3];   // log Vp, log Vs, log rho (trend subtracted before kriging)
  int nv[3];    // observations merged into the cell
};

struct Plan {
  int nx, ny, nz;
  CovSep cov[6];  // aa bb rr ab ar br
  std::vector<WellPt> pts;
  std::vector<int> by_i, ival;  // point indices sorted by i (and index); their i
  int nvalid = 0;
  float rangeX = 0, rangeY = 0, rangeZ = 0;
  int rA[3], rB[3], rR[3];
  int dxB = 0, dyB = 0, dzB = 0, dxE = 0, dyE = 0, dzE = 0;
  int target = 0;
};

float cpu_time(const Plan& p, float dxB, float dyE, float& nd, bool& rapid) {  // CalcCPUTime
  const float tBigK = 1.50e-6f, tChol = 1.90e-8f, tSolve = 7.00e-7f, tSmallK = 1.64e-6f;
  const float dxE = (dyE * p.rangeX) / p.rangeY, dzE = (dyE * p.rangeZ) / p.rangeY;
  const float dyB = (dxB * p.rangeY) / p.rangeX, dzB = (dxB * p.rangeZ) / p.rangeX;
  const float V = (float)(p.nx * p.ny * p.nz);
  const float v = dxB * dyB * dzB;
  const float Nss = V / v;
  nd = (dxB + 2 * dxE) * (dyB + 2 * dyE) * (dzB + 2 * dzE) * p.nvalid / V;
  const float TBigK = tBigK * Nss * nd * nd, TChol = tChol * Nss * nd * nd * nd, TSolve = tSolve * Nss * nd * nd;
  const float TSmall = tSmallK * V * nd;
  rapid = (TChol >= 0.5 * TSmall);
  return TBigK + TChol + TSolve + TSmall;
}

void block_size2(Plan& p) {  // EstimateSizeOfBlock2
  p.dyB = (int)std::ceil((p.dxB * p.rangeY) / p.rangeX);
  p.dzB = (int)std::ceil((p.dxB * p.rangeZ) / p.rangeX);
  if (p.dyE > (int)p.rangeY) p.dyE = (int)p.rangeY;
  p.dxE = (int)std::ceil((p.dyE * p.rangeX) / p.rangeY);
  p.dzE = (int)std::ceil((p.dyE * p.rangeZ) / p.rangeY);
  if (p.dxE > (int)p.rangeX) p.dxE = (int)p.rangeX;
  if (p.dzE > (int)p.rangeZ) p.dzE = (int)p.rangeZ;
}

void block_size(Plan& p) {  // EstimateSizeOfBlock, krigingadmin.cpp:768-923
  float tr[3][3];
  for (int a = 0; a < 3; ++a) taper_ranges(p.cov[a], tr[a][0], tr[a][1], tr[a][2]);
  const float rx = std::max(std::max(tr[0][0], tr[1][0]), tr[2][0]);
  const float ry = std::max(std::max(tr[0][1], tr[1][1]), tr[2][1]);
  const float rz = std::max(std::max(tr[0][2], tr[1][2]), tr[2][2]);
  for (int q = 0; q < 6; ++q) taper(p.cov[q], rx, ry, rz);
  estimate_ranges(p.cov[0], p.rA[0], p.rA[1], p.rA[2]);
  estimate_ranges(p.cov[1], p.rB[0], p.rB[1], p.rB[2]);
  estimate_ranges(p.cov[2], p.rR[0], p.rR[1], p.rR[2]);
  p.rangeX = (float)std::max(p.rA[0], p.rB[0]); p.rangeX = std::max((float)p.rR[0], p.rangeX);
  p.rangeY = (float)std::max(p.rA[1], p.rB[1]); p.rangeY = std::max((float)p.rR[1], p.rangeY);
  p.rangeZ = (float)std::max(p.rA[2], p.rB[2]); p.rangeZ = std::max((float)p.rR[2], p.rangeZ);
  const int ndata = (int)p.pts.size();
  if (ndata <= p.target) {
    p.dxB = p.nx; p.dyB = p.ny; p.dzB = p.nz;
    p.dxE = (int)p.rangeX; p.dyE = (int)p.rangeY; p.dzE = (int)p.rangeZ;
    return;
  }
  if (p.rangeX == 0.0f && p.rangeY == 0.0f) {
    p.dxB = 1; p.dyB = 1; p.dxE = 0; p.dyE = 0;
    p.dzB = (int)std::ceil(p.rangeZ); p.dzE = p.dzB;
    return;
  } else if (p.rangeX == 0.0f) {
    p.dxB = 1; p.dxE = 0;
    p.dyB = (int)std::ceil(p.rangeY); p.dyE = p.dyB;
    p.dzB = (int)std::ceil(p.rangeZ); p.dzE = p.dzB;
    return;
  } else if (p.rangeY == 0.0f) {
    p.dyB = 1; p.dyE = 0;
    p.dxB = (int)std::ceil(p.rangeX); p.dxE = p.dxB;
    p.dzB = (int)std::ceil(p.rangeZ); p.dzE = p.dzB;
    return;
  }
  float nd;
  bool rapid = false;
  int dxT = 0;
  const int dyE = (int)p.rangeY;
  float tMin = cpu_time(p, 1, (float)dyE, nd, rapid);
  float ndT = nd;
  for (int dxB = 2; dxB <= p.nx; ++dxB) {
    const float t1 = cpu_time(p, (float)dxB, (float)dyE, nd, rapid);
    if (t1 < tMin) { tMin = t1; dxT = dxB; ndT = nd; }
  }
  if (ndT < p.target) {
    p.dxB = dxT;
    p.dyE = dyE;
    block_size2(p);
    return;
  }
  bool test = false;
  float sumdy = 0.0f, sumdx = 0.0f;
  int count = 0;
  cpu_time(p, 1.0f, 1.0f, nd, rapid);
  cpu_time(p, (float)p.nx, p.rangeY, nd, rapid);
  const int nd2 = p.target;
  const float V = (float)(p.nx * p.ny * p.nz), R = p.rangeX * p.rangeY * p.rangeZ;
  const float a = (float)std::pow(nd2 * V / (p.nvalid * R), 0.333333333f);
  const int maxdy = (int)(0.5f * (a - 1.0f / p.rangeX) * p.rangeY);
  for (int dy = 1; dy <= maxdy; ++dy) {
    const float dx2 = p.rangeX * (a - 2 * dy / p.rangeY);
    if (dx2 < 1) continue;
    if (rapid) {
      test = true;
      sumdy += dy;
      sumdx += dx2;
      ++count;
      break;
    }
  }
  if (test) {
    p.dxB = (int)std::ceil(sumdx / count);
    p.dyE = (int)std::ceil(sumdy / count);
    block_size2(p);
    return;
  }
  p.dxB = p.nx; p.dyB = p.ny; p.dzB = p.nz;
  p.dxE = (int)p.rangeX; p.dyE = (int)p.rangeY; p.dzE = (int)p.rangeZ;
}

// the data of a box for kriging parameter g (FindDataInDataBlock): lists of
// point indices per parameter; a point whose own parameter g is valid adds
// only that one, else its other valid parameters
enum Fit { kRight, kSmall, kBig };
Fit find_data(const Plan& p, int g, const Box& box, std::vector<int> lists[3], std::vector<int>& hit) {
  for (int a = 0; a < 3; ++a) lists[a].clear();
  // the points inside, in their order (through the index by i)
  hit.clear();
  const auto lo = std::lower_bound(p.ival.begin(), p.ival.end(), box.i0);
  const auto hi = std::upper_bound(p.ival.begin(), p.ival.end(), box.i1);
  for (auto it = lo; it != hi; ++it) {
    const int m = p.by_i[it - p.ival.begin()];
    const WellPt& w = p.pts[m];
    if (box.inside(w.i, w.j, w.k)) hit.push_back(m);
  }
  std::sort(hit.begin(), hit.end());
  int total = 0;
  const int cmin = (int)(p.target * (1.0f - kTolPercent / 100.0f));
  const int cmax = (int)(p.target * (1.0f + kTolPercent / 100.0f));
  for (const int m : hit) {
    const WellPt& w = p.pts[m];
    if (w.nv[g] > 0) {
      lists[g].push_back(m);
      ++total;
    } else {
      for (int a = 0; a < 3; ++a)
        if (a != g && w.nv[a] > 0) {
          lists[a].push_back(m);
          ++total;
        }
    }
  }
  if (total <= cmax && total >= cmin) return kRight;
  return total < cmin ? kSmall : kBig;
}

// FindDataInDataBlockLoop, krigingadmin.cpp:519-575: the lists of the last
// search made
void find_data_loop(const Plan& p, int g, const Box& block, Box data, std::vector<int> lists[3],
                    std::vector<int>& hit) {
  Fit cur = find_data(p, g, data, lists, hit);
  Box minB = data;
  Box maxB = make_box(block.i0 - (int)p.rangeX, block.j0 - (int)p.rangeY, block.k0 - (int)p.rangeZ,
                      block.i1 + (int)p.rangeX, block.j1 + (int)p.rangeY, block.k1 + (int)p.rangeZ, nullptr);
  if (cur == kSmall) {
    if (find_data(p, g, maxB, lists, hit) != kBig) {
      data = maxB;
      cur = kRight;
    }
  } else if (cur == kBig) {
    if (find_data(p, g, minB, lists, hit) != kSmall) {
      data = minB;
      cur = kRight;
    }
  }
  while (cur != kRight) {
    if (cur == kSmall) {
      minB = data;
      move_half(data, maxB);
    } else {
      maxB = data;
      move_half(data, minB);
    }
    if (data == maxB || data == minB) break;
    cur = find_data(p, g, data, lists, hit);
  }
}

// FP64 Cholesky solve of the symmetric system whose upper triangle is K
bool chol_solve(std::vector<double>& K, int n, std::vector<double>& x) {
  // factor U^T U = K in place (upper triangle, row-major: K[i n + j], i <= j)
  for (int j = 0; j < n; ++j) {
    double d = K[(size_t)j * n + j];
    for (int k = 0; k < j; ++k) d -= K[(size_t)k * n + j] * K[(size_t)k * n + j];
    if (!(d > 0.0)) return false;
    const double ujj = std::sqrt(d);
    K[(size_t)j * n + j] = ujj;
    for (int i = j + 1; i < n; ++i) {
      double s = K[(size_t)j * n + i];
      for (int k = 0; k < j; ++k) s -= K[(size_t)k * n + j] * K[(size_t)k * n + i];
      K[(size_t)j * n + i] = s / ujj;
    }
  }
  for (int i = 0; i < n; ++i) {  // U^T y = r
    double s = x[i];
    for (int k = 0; k < i; ++k) s -= K[(size_t)k * n + i] * x[k];
    x[i] = s / K[(size_t)i * n + i];
  }
  for (int i = n - 1; i >= 0; --i) {  // U x = y
    double s = x[i];
    for (int k = i + 1; k < n; ++k) s -= K[(size_t)i * n + k] * x[k];
    x[i] = s / K[(size_t)i * n + i];
  }
  return true;
}

// ---- device ---------------------------------------------------------------

struct KBlock { int off, n; };
struct KTabs {
  const float* xy[3];   // per data type (alpha, beta, rho data): the table used
  const float* z[3];
  int flip[3];          // 1: covariance from the datum to the cell
};
struct KGeo {
  int nx, ny, nz;       // simbox
  int dxB, dyB, dzB, nbx, nby;
  int nxp, nyp, nzp;    // padded (the covariance window)
  long long sj, sk;     // grid strides of y and z
};

constexpr int kKrigThreads = 256;

// every cell of the simbox: result = trend + sum_d gamma_d(cell) x_d over
// the data of the cell's block, in the system's order (alpha, beta, rho
// data), the sum in FP64; cells of blocks without data are left alone
template <typename T>
__global__ __launch_bounds__(kKrigThreads) void k_krig_cells(T* trend, KGeo g, const KBlock* blocks,
                                                             const int4* dpos, const double* dx, KTabs tabs,
                                                             unsigned long long* counts) {
  const long long ncell = (long long)g.nx * g.ny * g.nz;
  const long long c = (long long)blockIdx.x * kKrigThreads + threadIdx.x;
  bool updated = false, missing = false;
  if (c < ncell) {
    const int ci = (int)(c % g.nx);
    const long long r = c / g.nx;
    const int cj = (int)(r % g.ny), ck = (int)(r / g.ny);
    const KBlock bk = blocks[ci / g.dxB + g.nbx * (cj / g.dyB + g.nby * (ck / g.dzB))];
    if (bk.n > 0) {
      double s = 0.0;
      for (int e = 0; e < bk.n; ++e) {
        const int4 d = dpos[bk.off + e];
        const int ty = d.w;
        const int sg = tabs.flip[ty] ? -1 : 1;
        int di = sg * (d.x - ci), dj = sg * (d.y - cj), dk = sg * (d.z - ck);
        float gv = 0.0f;
        if (abs(di) < g.nxp / 2 && abs(dj) < g.nyp / 2 && abs(dk) < g.nzp / 2) {
          if (di < 0) di += g.nxp;
          if (dj < 0) dj += g.nyp;
          if (dk < 0) dk += g.nzp;
          gv = tabs.xy[ty][di + dj * g.nxp] * tabs.z[ty][dk];
        }
        s = fma((double)gv, dx[bk.off + e], s);
      }
      T* cell = trend + ci + g.sj * cj + g.sk * ck;
      const T v = *cell;
      if (v == (T)kMissing) {
        missing = true;
      } else {
        if constexpr (sizeof(T) == 4) *cell = v + (float)s;  // float result + float(k . x), as the reference
        else *cell = v + s;
        updated = true;
      }
    }
  }
  // the counts, one atomic per wave
  const unsigned long long bu = __ballot(updated), bm = __ballot(missing);
  if ((threadIdx.x & 63) == 0) {
    if (bu) atomicAdd(&counts[0], (unsigned long long)__popcll(bu));
    if (bm) atomicAdd(&counts[1], (unsigned long long)__popcll(bm));
  }
}

// the trend at each well point (SubtractTrends, krigingadmin.cpp:400-439)
template <typename T>
__global__ void k_gather(const T* g, long long sj, long long sk, const int4* pos, int n, float* out) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) out[t] = (float)g[pos[t].x + sj * pos[t].y + sk * pos[t].z];
}

}  // namespace

extern "C" int crv_krig_post(crv_grid* mean[3], crv_grid* cov[6], const crv_well_point* pts, int npts,
                             int data_target, int smooth, crv_krig_stats* stats) {
  if (!mean || !mean[0]) return CRV_E_INVALID;
  crv_ctx* ctx = mean[0]->ctx;
  if (!cov || !pts || npts < 0 || data_target < 1)
    return crv_set_error(ctx, CRV_E_INVALID, "kriging: bad arguments");
  if (smooth)
    return crv_set_error(ctx, CRV_E_UNSUPPORTED,
                         "kriging: the smoothing of kriged values (smooth-kriged-parameters) is not built");
  crv_grid* g0 = mean[0];
  for (int a = 0; a < 3; ++a) {
    crv_grid* g = mean[a];
    if (!g || g->nxp != g0->nxp || g->nyp != g0->nyp || g->nzp != g0->nzp || g->precision != g0->precision ||
        g->nx != g0->nx || g->ny != g0->ny || g->nz != g0->nz)
      return crv_set_error(ctx, CRV_E_INVALID, "kriging: mean grids inconsistent");
    if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "kriging: mean grids must be real");
  }
  for (int q = 0; q < 6; ++q) {
    crv_grid* g = cov[q];
    if (!g || g->nxp != g0->nxp || g->nyp != g0->nyp || g->nzp != g0->nzp)
      return crv_set_error(ctx, CRV_E_INVALID, "kriging: covariance grids inconsistent");
    if (g->transformed) return crv_set_error(ctx, CRV_E_STATE, "kriging: covariance grids must be real");
  }
  Plan p;
  p.nx = g0->nx; p.ny = g0->ny; p.nz = g0->nz;
  p.target = data_target;
  const int dims[3] = {p.nx, p.ny, p.nz};
  // data: points of one cell merged (their means), points without any value
  // dropped (KrigingData3D::addData / divide)
  for (int m = 0; m < npts; ++m) {
    const crv_well_point& w = pts[m];
    if (w.i < 0 || w.i >= p.nx || w.j < 0 || w.j >= p.ny || w.k < 0 || w.k >= p.nz)
      return crv_set_error(ctx, CRV_E_INVALID, "kriging: well point outside the grid");
    const float v[3] = {w.alpha, w.beta, w.rho};
    if (v[0] == kMissing && v[1] == kMissing && v[2] == kMissing) continue;
    int at = -1;
    for (int b = 0; b < (int)p.pts.size(); ++b)
      if (p.pts[b].i == w.i && p.pts[b].j == w.j && p.pts[b].k == w.k) { at = b; break; }
    if (at < 0) {
      p.pts.push_back(WellPt{w.i, w.j, w.k, {0.0f, 0.0f, 0.0f}, {0, 0, 0}});
      at = (int)p.pts.size() - 1;
    }
    for (int a = 0; a < 3; ++a)
      if (v[a] != kMissing) {
        p.pts[at].v[a] += v[a];
        p.pts[at].nv[a]++;
      }
  }
  for (WellPt& w : p.pts)
    for (int a = 0; a < 3; ++a) {
      if (w.nv[a] > 0) w.v[a] /= (float)w.nv[a];
      p.nvalid += w.nv[a] > 0;
    }
  crv_krig_stats st;
  memset(&st, 0, sizeof(st));
  st.data_points = (int)p.pts.size();
  int rc;
  for (int q = 0; q < 6; ++q)
    if ((rc = cov_sep_from_grid(cov[q], p.cov[q])) != CRV_OK) return rc;
  block_size(p);
  p.dxB = std::max(p.dxB, 1); p.dyB = std::max(p.dyB, 1); p.dzB = std::max(p.dzB, 1);
  if (!(p.dxE <= p.rangeX && p.dyE <= p.rangeY && p.dzE <= p.rangeZ))
    return crv_set_error(ctx, CRV_E_INVALID, "kriging: data neighbourhood beyond the covariance ranges");
  st.block[0] = p.dxB; st.block[1] = p.dyB; st.block[2] = p.dzB;
  st.block_ext[0] = p.dxE; st.block_ext[1] = p.dyE; st.block_ext[2] = p.dzE;
  st.range[0] = (int)p.rangeX; st.range[1] = (int)p.rangeY; st.range[2] = (int)p.rangeZ;
  const int np = (int)p.pts.size();
  if (np == 0) {
    if (stats) *stats = st;
    return CRV_OK;
  }
  // residuals: data minus the trends at the points
  {
    std::vector<int4> pos(np);
    for (int m = 0; m < np; ++m) pos[m] = make_int4(p.pts[m].i, p.pts[m].j, p.pts[m].k, 0);
    int4* dpos = nullptr;
    float* dval = nullptr;
    CRV_HIP_CHECK(ctx, hipMalloc(&dpos, np * sizeof(int4)));
    CRV_HIP_CHECK(ctx, hipMalloc(&dval, 3 * np * sizeof(float)));
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(dpos, pos.data(), np * sizeof(int4), hipMemcpyHostToDevice, ctx->stream));
    const long long sj = g0->rnxp, sk = (long long)g0->rnxp * g0->nyp;
    for (int a = 0; a < 3; ++a) {
      if (g0->precision == CRV_FP64)
        hipLaunchKernelGGL(k_gather<double>, dim3((np + 255) / 256), dim3(256), 0, ctx->stream,
                           (const double*)mean[a]->d, sj, sk, dpos, np, dval + a * np);
      else
        hipLaunchKernelGGL(k_gather<float>, dim3((np + 255) / 256), dim3(256), 0, ctx->stream,
                           (const float*)mean[a]->d, sj, sk, dpos, np, dval + a * np);
    }
    std::vector<float> tr(3 * np);
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(tr.data(), dval, 3 * np * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
    CRV_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
    CRV_HIP_CHECK(ctx, hipFree(dpos));
    CRV_HIP_CHECK(ctx, hipFree(dval));
    for (int a = 0; a < 3; ++a)
      for (int m = 0; m < np; ++m)
        if (p.pts[m].nv[a] > 0 && tr[a * np + m] != kMissing) p.pts[m].v[a] -= tr[a * np + m];
  }
  // per parameter: the blocks' systems (host threads over the blocks), then
  // every cell on the device
  const int nbx = (p.nx + p.dxB - 1) / p.dxB, nby = (p.ny + p.dyB - 1) / p.dyB, nbz = (p.nz + p.dzB - 1) / p.dzB;
  const long long nblocks = (long long)nbx * nby * nbz;
  if (nblocks > (1LL << 30)) return crv_set_error(ctx, CRV_E_UNSUPPORTED, "kriging: too many blocks");
  st.blocks = (int)nblocks;
  p.by_i.resize(np);
  for (int m = 0; m < np; ++m) p.by_i[m] = m;
  std::sort(p.by_i.begin(), p.by_i.end(), [&](int x, int y) {
    return p.pts[x].i != p.pts[y].i ? p.pts[x].i < p.pts[y].i : x < y;
  });
  p.ival.resize(np);
  for (int m = 0; m < np; ++m) p.ival[m] = p.pts[p.by_i[m]].i;
  // covariance table per (kriged parameter, data type) and its direction
  // (SetKrigVector: alpha uses aa, ab, ar; beta ab~, bb~, br; rho ar~, br~,
  // rr~); the system's upper triangle uses table tab_of[row type][column
  // type] from the row datum to the column datum (SetMatrix)
  const int tab_of[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};
  const int flip_of[3][3] = {{0, 0, 0}, {1, 1, 0}, {1, 1, 1}};
  float* dtab = nullptr;
  const size_t nxy = (size_t)p.cov[0].nxp * p.cov[0].nyp, nzp = p.cov[0].nzp;
  CRV_HIP_CHECK(ctx, hipMalloc(&dtab, 6 * (nxy + nzp) * sizeof(float)));
  for (int q = 0; q < 6; ++q) {
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(dtab + q * (nxy + nzp), p.cov[q].xy.data(), nxy * sizeof(float),
                                      hipMemcpyHostToDevice, ctx->stream));
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(dtab + q * (nxy + nzp) + nxy, p.cov[q].z.data(), nzp * sizeof(float),
                                      hipMemcpyHostToDevice, ctx->stream));
  }
  unsigned long long* dcount = nullptr;
  CRV_HIP_CHECK(ctx, hipMalloc(&dcount, 6 * sizeof(unsigned long long)));
  CRV_HIP_CHECK(ctx, hipMemsetAsync(dcount, 0, 6 * sizeof(unsigned long long), ctx->stream));
  std::vector<void*> keep;
  auto cleanup = [&]() {
    (void)hipStreamSynchronize(ctx->stream);
    for (void* v : keep) (void)hipFree(v);
    (void)hipFree(dtab);
    (void)hipFree(dcount);
  };
  const int nthr = (int)std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
  for (int g = 0; g < 3; ++g) {
    // thread t plans blocks [t nblocks / nthr, (t + 1) nblocks / nthr)
    struct Part {
      std::vector<int4> pos;
      std::vector<double> x;
      int empty = 0, maxn = 0;
      bool fail = false;
    };
    std::vector<Part> part(nthr);
    std::vector<KBlock> blocks(nblocks);
    auto plan_range = [&](int t) {
      Part& P = part[t];
      std::vector<int> lists[3], hit, idx, ty;
      std::vector<double> K, x;
      const long long b0 = nblocks * t / nthr, b1 = nblocks * (t + 1) / nthr;
      for (long long b = b0; b < b1 && !P.fail; ++b) {
        const int bi = (int)(b % nbx), bj = (int)((b / nbx) % nby), bk = (int)(b / ((long long)nbx * nby));
        const int i1 = bi * p.dxB, j1 = bj * p.dyB, k1 = bk * p.dzB;
        const Box block = make_box(i1, j1, k1, i1 + p.dxB - 1, j1 + p.dyB - 1, k1 + p.dzB - 1, dims);
        const Box data = make_box(i1 - p.dxE, j1 - p.dyE, k1 - p.dzE, i1 + p.dxB + p.dxE - 1, j1 + p.dyB + p.dyE - 1,
                                  k1 + p.dzB + p.dzE - 1, dims);
        find_data_loop(p, g, block, data, lists, hit);
        const int n = (int)(lists[0].size() + lists[1].size() + lists[2].size());
        blocks[b] = KBlock{(int)P.pos.size(), n};
        if (n == 0) {
          ++P.empty;
          continue;
        }
        P.maxn = std::max(P.maxn, n);
        // the order of the system: alpha data, beta data, rho data
        idx.clear();
        ty.clear();
        for (int a = 0; a < 3; ++a)
          for (int m : lists[a]) {
            idx.push_back(m);
            ty.push_back(a);
          }
        K.assign((size_t)n * n, 0.0);
        x.assign(n, 0.0);
        for (int r = 0; r < n; ++r) {
          const WellPt& wr = p.pts[idx[r]];
          x[r] = wr.v[ty[r]];
          for (int c = r; c < n; ++c) {
            const WellPt& wc = p.pts[idx[c]];
            K[(size_t)r * n + c] = p.cov[tab_of[ty[r]][ty[c]]].g2(wr.i, wr.j, wr.k, wc.i, wc.j, wc.k);
          }
        }
        if (!chol_solve(K, n, x)) {
          P.fail = true;
          break;
        }
        for (int r = 0; r < n; ++r) {
          const WellPt& w = p.pts[idx[r]];
          P.pos.push_back(make_int4(w.i, w.j, w.k, ty[r]));
          P.x.push_back(x[r]);
        }
      }
    };
    std::vector<std::thread> pool;
    for (int t = 1; t < nthr; ++t) pool.emplace_back(plan_range, t);
    plan_range(0);
    for (auto& th : pool) th.join();
    std::vector<int4> dposv;
    std::vector<double> dxv;
    std::vector<long long> base(nthr, 0);
    for (int t = 0; t < nthr; ++t) {
      if (part[t].fail) {
        cleanup();
        return crv_set_error(ctx, CRV_E_INVALID, "kriging: block system not positive definite");
      }
      base[t] = (long long)dposv.size();
      dposv.insert(dposv.end(), part[t].pos.begin(), part[t].pos.end());
      dxv.insert(dxv.end(), part[t].x.begin(), part[t].x.end());
      st.empty_blocks[g] += part[t].empty;
      st.max_block_data[g] = std::max(st.max_block_data[g], part[t].maxn);
    }
    if (dposv.size() > (size_t)INT32_MAX) {
      cleanup();
      return crv_set_error(ctx, CRV_E_UNSUPPORTED, "kriging: too many block data");
    }
    for (int t = 0; t < nthr; ++t) {
      const long long b0 = nblocks * t / nthr, b1 = nblocks * (t + 1) / nthr;
      for (long long b = b0; b < b1; ++b) blocks[b].off += (int)base[t];
    }
    if (dposv.empty()) continue;
    KBlock* db = nullptr;
    int4* dp = nullptr;
    double* dxp = nullptr;
    CRV_HIP_CHECK(ctx, hipMalloc(&db, blocks.size() * sizeof(KBlock)));
    keep.push_back(db);
    CRV_HIP_CHECK(ctx, hipMalloc(&dp, dposv.size() * sizeof(int4)));
    keep.push_back(dp);
    CRV_HIP_CHECK(ctx, hipMalloc(&dxp, dxv.size() * sizeof(double)));
    keep.push_back(dxp);
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(db, blocks.data(), blocks.size() * sizeof(KBlock), hipMemcpyHostToDevice, ctx->stream));
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(dp, dposv.data(), dposv.size() * sizeof(int4), hipMemcpyHostToDevice, ctx->stream));
    CRV_HIP_CHECK(ctx, hipMemcpyAsync(dxp, dxv.data(), dxv.size() * sizeof(double), hipMemcpyHostToDevice, ctx->stream));
    KTabs tb;
    for (int a = 0; a < 3; ++a) {
      const int q = tab_of[g][a];
      tb.xy[a] = dtab + q * (nxy + nzp);
      tb.z[a] = dtab + q * (nxy + nzp) + nxy;
      tb.flip[a] = flip_of[g][a];
    }
    KGeo geo;
    geo.nx = p.nx; geo.ny = p.ny; geo.nz = p.nz;
    geo.dxB = p.dxB; geo.dyB = p.dyB; geo.dzB = p.dzB; geo.nbx = nbx; geo.nby = nby;
    geo.nxp = g0->nxp; geo.nyp = g0->nyp; geo.nzp = g0->nzp;
    geo.sj = g0->rnxp; geo.sk = (long long)g0->rnxp * g0->nyp;
    const long long ncell = (long long)p.nx * p.ny * p.nz;
    const unsigned nb = (unsigned)((ncell + kKrigThreads - 1) / kKrigThreads);
    {
      ProfScope ps(ctx, "krig_cells");
      if (g0->precision == CRV_FP64)
        hipLaunchKernelGGL(k_krig_cells<double>, dim3(nb), dim3(kKrigThreads), 0, ctx->stream, (double*)mean[g]->d,
                           geo, db, dp, dxp, tb, dcount + 2 * g);
      else
        hipLaunchKernelGGL(k_krig_cells<float>, dim3(nb), dim3(kKrigThreads), 0, ctx->stream, (float*)mean[g]->d,
                           geo, db, dp, dxp, tb, dcount + 2 * g);
    }
    if (hipGetLastError() != hipSuccess) {
      cleanup();
      return crv_set_error(ctx, CRV_E_HIP, "kriging: launch failed");
    }
  }
  unsigned long long hc[6];
  const hipError_t e = hipMemcpyAsync(hc, dcount, sizeof(hc), hipMemcpyDeviceToHost, ctx->stream);
  cleanup();
  if (e != hipSuccess) return crv_set_error(ctx, CRV_E_HIP, "kriging: counter download failed");
  for (int g = 0; g < 3; ++g) {
    st.cells_updated[g] = (long long)hc[2 * g];
    st.cells_missing[g] = (long long)hc[2 * g + 1];
  }
  if (stats) *stats = st;
  return CRV_OK;
}
