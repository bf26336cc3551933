// Host C++ mirror of the reference's Crava inversion members on the hot path
// and of SeismicParametersHolder's covariance grids (reference
// src/crava.h:43-64, 154; src/seismicparametersholder.h).  The arithmetic
// runs on the GPU through the C ABI; these classes keep the reference's call
// shapes so that its driver code maps one to one.
#pragma once

#include <complex>
#include <vector>

#include "fftgrid.h"

namespace crava_hip {

// SeismicParametersHolder: prior covariance grids (real, COVARIANCE) that
// computePostMeanResidAndFFTCov transforms and overwrites with the posterior
// covariance spectra, and the 3x3 prior parameter covariance priorVar0_.
class SeismicParametersHolder {
 public:
  SeismicParametersHolder(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, int precision,
                          const double priorVar0[9]);
  ~SeismicParametersHolder();
  FFTGrid* GetCovAlpha() { return cov_[0]; }
  FFTGrid* GetCovBeta() { return cov_[1]; }
  FFTGrid* GetCovRho() { return cov_[2]; }
  FFTGrid* GetCrCovAlphaBeta() { return cov_[3]; }
  FFTGrid* GetCrCovAlphaRho() { return cov_[4]; }
  FFTGrid* GetCrCovBetaRho() { return cov_[5]; }
  FFTGrid** grids() { return cov_; }
  const double* priorVar0() const { return priorVar0_; }
  void FFTCovGrids();     // seismicparametersholder.cpp:299-322
  void invFFTCovGrids();  // seismicparametersholder.cpp:273-296
  // computeCircCorrT + makeCircCorrTPosDef (seismicparametersholder.cpp:
  // 430-498): nzp values; throws when the correlation is undefined.
  static std::vector<float> computeCircCorrT(const std::vector<float>& priorCorrT, int minIntFq, int nzp);

 private:
  FFTGrid* cov_[6];
  double priorVar0_[9];
};

// Per-kz spectra that the reference reads from its Wavelet objects in
// fillErrMultVectors (crava.cpp:1336-1364): kD * conj(W_l), conj(W_l)/|W_l|,
// conj(E3_l)/|E3_l| and 1, each nzp x ntheta.
struct KzTables {
  int nzp = 0, ntheta = 0;
  std::vector<std::complex<double>> kwd, em1, em2, em3;
};

struct Precision {
  int mode = CRV_MODE_PLAIN;       // or CRV_MODE_FP32_REFINE
  double cond_threshold = 1e5;
  double lowk_fraction = 0.02;
};

class Crava {
 public:
  // Crava::computePostMeanResidAndFFTCov (crava.cpp:975-1285) for constant
  // thickness: postAlpha/Beta/Rho in = prior means (real), out = posterior
  // means (real); seisData in = data, out = residual spectra; errCorr is
  // transformed here; the covariance grids of seismicParameters go from
  // prior (real) to posterior spectra.  nThreads and fileGrid are accepted
  // for signature parity and unused (one GPU stream; no disk paging).
  static int computePostMeanResidAndFFTCov(SeismicParametersHolder& seismicParameters, FFTGrid*& postAlpha,
                                           FFTGrid*& postBeta, FFTGrid*& postRho, FFTGrid*& errCorr,
                                           FFTGrid** seisData, const KzTables& tables, double** errThetaCov,
                                           float** A, int ntheta, float lowCut, float highCut, float wnc, int nz,
                                           int nzp, int nyp, int nxp, double lz, double minRelThick,
                                           int nThreads, bool fileGrid, const Precision& precision,
                                           crv_post_stats* stats);

  // Crava::simulate (crava.cpp:1550-1734): nSim realizations from the
  // reference noise stream of ranGen; realization r is written to
  // out[3*r .. 3*r+2] (alpha, beta, rho, real, posterior mean added).
  static int simulate(SeismicParametersHolder& seismicParameters, FFTGrid* postMean[3], RandomGen* ranGen,
                      int nSim, FFTGrid** out, bool fp64Cholesky);
};

// Model set-up of the configurations (restating the reference's wavelet,
// reflection-matrix and error-covariance construction; see model.cpp).
KzTables makeRickerTables(int nzp, double dt, const std::vector<double>& peakHz);
std::vector<float> reflectionMatrix(const std::vector<double>& anglesDeg, double vsvp);
std::vector<double> errThetaCov(const std::vector<double>& dataVar, const std::vector<double>& sn,
                                const std::vector<double>& anglesDeg, double angularRangeDeg);

}  // namespace crava_hip
