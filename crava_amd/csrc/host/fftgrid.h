// Host C++ mirror of the reference's FFTGrid and RandomGen for the hot path
// (reference src/fftgrid.h:22-323, libs/lib/random.h:16-34).  Storage lives
// on the GPU behind the C ABI (include/crava_hip.h); the member functions
// keep the reference's names, argument meaning and state rules (asserts
// become exceptions).  A caller of the reference switches by constructing
// these objects where it constructed FFTGrid.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "crava_hip.h"

namespace crava_hip {

class HipError : public std::runtime_error {
 public:
  explicit HipError(const std::string& m) : std::runtime_error(m) {}
};

// libs/lib/random.h: the generator's whole state is the 32-bit LCG seed.
class RandomGen {
 public:
  explicit RandomGen(unsigned int seed) : seed_(seed) {}
  unsigned int* state() { return &seed_; }
  unsigned int seed() const { return seed_; }

 private:
  unsigned int seed_;
};

class FFTGrid {
 public:
  enum gridTypes { CTMISSING, DATA, PARAMETER, COVARIANCE, VELOCITY, OPERATOR };

  // FFTGrid(nx, ny, nz, nxp, nyp, nzp) (fftgrid.cpp:40-70), on a context
  FFTGrid(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision = CRV_FP32);
  virtual ~FFTGrid();
  FFTGrid(const FFTGrid&) = delete;
  FFTGrid& operator=(const FFTGrid&) = delete;

  void createRealGrid();      // fftgrid.cpp:1243-1259 (buffer exists; state = real)
  void createComplexGrid();   // state = transformed
  void fillInConstant(float value);                        // pad columns set to 0
  void fillInComplexNoise(RandomGen* ranGen);              // fftgrid.cpp:1191-1240

  virtual void fftInPlace();                               // fftgrid.cpp:1783-1808
  virtual void invFFTInPlace();                            // fftgrid.cpp:1811-1841
  virtual void add(FFTGrid* fftGrid);                      // fftgrid.cpp:1870-1891
  virtual void multiplyByScalar(float scalar);             // fftgrid.cpp:1985-1995
  int  expTransf();                                         // fftgrid.cpp:1726-1743
  // fftgrid.cpp:2411-2498: CRAVA's binary grid format with the simbox geometry
  void writeCravaFile(const std::string& fileName, const crv_simbox_info& simbox);
  void readCravaFile(const std::string& fileName, std::string& errText);
  // fftgrid.cpp:2117-2176: appends ".storm" (binary) or ".dat" (text) to fileName;
  // the simbox's geometry and surface names come in storm (its flat field is set from flat)
  void writeStormFile(const std::string& fileName, const crv_storm_info& storm, bool ascii, bool padding = false,
                      bool flat = false, bool scientific_format = false);

  // host copies in the reference layout (rvalue_ with pads / cvalue_)
  void setRealValues(const std::vector<float>& r);
  void setComplexValues(const std::vector<float>& c);      // interleaved re, im
  std::vector<float> getRealValues() const;
  std::vector<float> getComplexValues() const;

  int  getType() const;
  void setType(int cubeType);
  bool getIsTransformed() const;
  void setTransformedStatus(bool status);
  int  getNx() const { return nx_; }
  int  getNy() const { return ny_; }
  int  getNz() const { return nz_; }
  int  getNxp() const { return nxp_; }
  int  getNyp() const { return nyp_; }
  int  getNzp() const { return nzp_; }
  int  getCNxp() const { return nxp_ / 2 + 1; }
  int  getRNxp() const { return 2 * (nxp_ / 2 + 1); }
  long long getrsize() const { return (long long)getRNxp() * nyp_ * nzp_; }
  long long getcsize() const { return (long long)getCNxp() * nyp_ * nzp_; }

  crv_grid* handle() const { return grid_; }
  crv_ctx* context() const { return ctx_; }

  static int findClosestFactorableNumber(int leastint) { return crv_find_closest_factorable(leastint); }

 protected:
  void check(int rc, const char* what) const;

  crv_ctx* ctx_;
  crv_grid* grid_;
  int nx_, ny_, nz_, nxp_, nyp_, nzp_;
  int precision_;
};

}  // namespace crava_hip
