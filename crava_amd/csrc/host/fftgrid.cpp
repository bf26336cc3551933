#include "fftgrid.h"

#include <cstring>

namespace crava_hip {

FFTGrid::FFTGrid(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision)
    : ctx_(ctx), grid_(nullptr), nx_(nx), ny_(ny), nz_(nz), nxp_(nxp), nyp_(nyp), nzp_(nzp), precision_(precision) {
  check(crv_grid_create(ctx, nx, ny, nz, nxp, nyp, nzp, precision, CRV_CTMISSING, &grid_), "FFTGrid");
}

FFTGrid::~FFTGrid() {
  if (grid_) crv_grid_destroy(grid_);
}

void FFTGrid::check(int rc, const char* what) const {
  if (rc != CRV_OK) throw HipError(std::string(what) + ": " + crv_last_error(ctx_));
}

void FFTGrid::createRealGrid() { check(crv_grid_set_transformed(grid_, 0), "createRealGrid"); }

void FFTGrid::createComplexGrid() { check(crv_grid_set_transformed(grid_, 1), "createComplexGrid"); }

void FFTGrid::setType(int cubeType) {
  check(crv_grid_set_cubetype(grid_, cubeType), "setType");
}

int FFTGrid::getType() const {
  crv_grid_info i;
  check(crv_grid_info_get(grid_, &i), "getType");
  return i.cubetype;
}

bool FFTGrid::getIsTransformed() const {
  crv_grid_info i;
  check(crv_grid_info_get(grid_, &i), "getIsTransformed");
  return i.is_transformed != 0;
}

void FFTGrid::setTransformedStatus(bool status) { check(crv_grid_set_transformed(grid_, status ? 1 : 0), "setTransformedStatus"); }

void FFTGrid::fillInConstant(float value) { check(crv_grid_fill(grid_, value), "fillInConstant"); }

void FFTGrid::fillInComplexNoise(RandomGen* ranGen) {
  if (!ranGen) throw HipError("fillInComplexNoise: NULL RandomGen");
  check(crv_fill_complex_noise(grid_, ranGen->state()), "fillInComplexNoise");
}

void FFTGrid::fftInPlace() {
  if (getIsTransformed()) throw HipError("fftInPlace: grid already transformed");
  if (getType() == CTMISSING) throw HipError("fftInPlace: grid type missing");
  check(crv_fft_r2c(grid_), "fftInPlace");
}

void FFTGrid::invFFTInPlace() {
  if (!getIsTransformed()) throw HipError("invFFTInPlace: grid not transformed");
  if (getType() == CTMISSING) throw HipError("invFFTInPlace: grid type missing");
  check(crv_fft_c2r(grid_), "invFFTInPlace");
}

void FFTGrid::add(FFTGrid* g) {
  if (!g || g->getNxp() != nxp_) throw HipError("add: grids differ");
  check(crv_grid_add(grid_, g->grid_), "add");
}

void FFTGrid::multiplyByScalar(float scalar) { check(crv_grid_multiply_by_scalar(grid_, scalar), "multiplyByScalar"); }

int FFTGrid::expTransf() {
  check(crv_grid_derived(ctx_, CRV_OUT_EXP, grid_, nullptr, nullptr, grid_), "expTransf");
  return 0;
}

// the reference appends IO::SuffixCrava() (".crava") to the name it is given
void FFTGrid::writeCravaFile(const std::string& fileName, const crv_simbox_info& simbox) {
  check(crv_grid_write_crava(grid_, (fileName + ".crava").c_str(), &simbox), "writeCravaFile");
}

// the reference appends IO::SuffixStormBinary() (".storm") or IO::SuffixGeneralData() (".dat")
void FFTGrid::writeStormFile(const std::string& fileName, const crv_storm_info& storm, bool ascii, bool padding,
                             bool flat, bool scientific_format) {
  crv_storm_info si = storm;
  si.flat = flat ? 1 : 0;
  const std::string path = fileName + (ascii ? ".dat" : ".storm");
  check(crv_grid_write_storm(grid_, path.c_str(), &si, ascii ? 1 : 0, padding ? 1 : 0, scientific_format ? 1 : 0),
        "writeStormFile");
}

void FFTGrid::readCravaFile(const std::string& fileName, std::string& errText) {
  if (crv_grid_read_crava(grid_, fileName.c_str(), nullptr) != CRV_OK)
    errText += std::string("Error: ") + crv_last_error(ctx_) + "\n";
}

void FFTGrid::setRealValues(const std::vector<float>& r) {
  if ((long long)r.size() != getrsize()) throw HipError("setRealValues: size != rsize");
  if (precision_ == CRV_FP32) {
    check(crv_grid_upload_real(grid_, r.data(), r.size()), "setRealValues");
  } else {
    std::vector<double> d(r.begin(), r.end());
    check(crv_grid_upload_real(grid_, d.data(), d.size()), "setRealValues");
  }
}

void FFTGrid::setComplexValues(const std::vector<float>& c) {
  if ((long long)c.size() != 2 * getcsize()) throw HipError("setComplexValues: size != 2*csize");
  if (precision_ == CRV_FP32) {
    check(crv_grid_upload_complex(grid_, c.data(), c.size()), "setComplexValues");
  } else {
    std::vector<double> d(c.begin(), c.end());
    check(crv_grid_upload_complex(grid_, d.data(), d.size()), "setComplexValues");
  }
}

std::vector<float> FFTGrid::getRealValues() const {
  const size_t n = (size_t)getrsize();
  if (precision_ == CRV_FP32) {
    std::vector<float> r(n);
    check(crv_grid_download(grid_, r.data(), n), "getRealValues");
    return r;
  }
  std::vector<double> d(n);
  check(crv_grid_download(grid_, d.data(), n), "getRealValues");
  return std::vector<float>(d.begin(), d.end());
}

std::vector<float> FFTGrid::getComplexValues() const {
  const size_t n = (size_t)(2 * getcsize());
  if (precision_ == CRV_FP32) {
    std::vector<float> r(n);
    check(crv_grid_download(grid_, r.data(), n), "getComplexValues");
    return r;
  }
  std::vector<double> d(n);
  check(crv_grid_download(grid_, d.data(), n), "getComplexValues");
  return std::vector<float>(d.begin(), d.end());
}

}  // namespace crava_hip
