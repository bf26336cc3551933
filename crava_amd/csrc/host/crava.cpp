#include "crava.h"

#include <cmath>
#include <cstring>

namespace crava_hip {

SeismicParametersHolder::SeismicParametersHolder(crv_ctx* ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                                 int precision, const double priorVar0[9]) {
  for (int q = 0; q < 6; ++q) {
    cov_[q] = new FFTGrid(ctx, nx, ny, nz, nxp, nyp, nzp, precision);
    cov_[q]->setType(FFTGrid::COVARIANCE);
  }
  std::memcpy(priorVar0_, priorVar0, sizeof(priorVar0_));
}

SeismicParametersHolder::~SeismicParametersHolder() {
  for (auto* g : cov_) delete g;
}

void SeismicParametersHolder::FFTCovGrids() {
  for (auto* g : cov_)
    if (!g->getIsTransformed()) g->fftInPlace();
}

void SeismicParametersHolder::invFFTCovGrids() {
  for (auto* g : cov_)
    if (g->getIsTransformed()) g->invFFTInPlace();
}

std::vector<float> SeismicParametersHolder::computeCircCorrT(const std::vector<float>& priorCorrT, int minIntFq,
                                                            int nzp) {
  std::vector<float> out(nzp);
  if (priorCorrT.empty() ||
      crv_circ_corr_t(priorCorrT.data(), (int)priorCorrT.size(), nzp, minIntFq, out.data()) != CRV_OK)
    throw HipError("computeCircCorrT: the circular temporal correlation is undefined (too few layers)");
  return out;
}

int Crava::computePostMeanResidAndFFTCov(SeismicParametersHolder& sp, FFTGrid*& postAlpha, FFTGrid*& postBeta,
                                         FFTGrid*& postRho, FFTGrid*& errCorr, FFTGrid** seisData,
                                         const KzTables& t, double** errThetaCov, float** A, int ntheta,
                                         float lowCut, float highCut, float wnc, int nz, int nzp, int nyp, int nxp,
                                         double lz, double minRelThick, int /*nThreads*/, bool /*fileGrid*/,
                                         const Precision& prec, crv_post_stats* stats) {
  if (ntheta < 1 || ntheta > 8 || t.ntheta != ntheta || t.nzp != nzp) throw HipError("computePostMeanResidAndFFTCov: tables");
  if (postAlpha->getNxp() != nxp || postAlpha->getNyp() != nyp || postAlpha->getNzp() != nzp)
    throw HipError("computePostMeanResidAndFFTCov: grid sizes");
  std::vector<float> Af((size_t)ntheta * 3);
  std::vector<double> th((size_t)ntheta * ntheta);
  for (int l = 0; l < ntheta; ++l) {
    for (int m = 0; m < 3; ++m) Af[l * 3 + m] = A[l][m];
    for (int m = 0; m < ntheta; ++m) th[l * ntheta + m] = errThetaCov[l][m];
  }
  crv_post_params p{};
  p.ntheta = ntheta;
  p.nz = nz;
  p.lz = lz;
  p.min_rel_thick = minRelThick;
  p.low_cut = lowCut;
  p.high_cut = highCut;
  p.wnc = wnc;
  p.A = Af.data();
  p.err_theta_cov = th.data();
  p.kwd = reinterpret_cast<const double*>(t.kwd.data());
  p.em1 = reinterpret_cast<const double*>(t.em1.data());
  p.em2 = reinterpret_cast<const double*>(t.em2.data());
  p.em3 = reinterpret_cast<const double*>(t.em3.data());
  p.mode = prec.mode;
  p.cond_threshold = prec.cond_threshold;
  p.lowk_fraction = prec.lowk_fraction;
  p.separable = 0;
  p.sigma0 = sp.priorVar0();
  crv_grid* mean[3] = {postAlpha->handle(), postBeta->handle(), postRho->handle()};
  std::vector<crv_grid*> data(ntheta);
  for (int l = 0; l < ntheta; ++l) data[l] = seisData[l]->handle();
  crv_grid* cov[6];
  for (int q = 0; q < 6; ++q) cov[q] = sp.grids()[q]->handle();
  const int rc = crv_posterior(postAlpha->context(), &p, mean, data.data(), errCorr->handle(), cov, stats);
  if (rc != CRV_OK) throw HipError(std::string("computePostMeanResidAndFFTCov: ") + crv_last_error(postAlpha->context()));
  return 0;
}

int Crava::simulate(SeismicParametersHolder& sp, FFTGrid* postMean[3], RandomGen* ranGen, int nSim, FFTGrid** out,
                    bool fp64Cholesky) {
  crv_grid* cov[6];
  for (int q = 0; q < 6; ++q) cov[q] = sp.grids()[q]->handle();
  crv_grid* pm[3] = {postMean[0]->handle(), postMean[1]->handle(), postMean[2]->handle()};
  for (int r = 0; r < nSim; ++r) {
    crv_grid* o[3] = {out[3 * r]->handle(), out[3 * r + 1]->handle(), out[3 * r + 2]->handle()};
    crv_noise_spec ns{};
    ns.mode = CRV_NOISE_REFERENCE;
    ns.lcg_state = ranGen->state();
    const int rc = crv_simulate(out[0]->context(), cov, pm, &ns, fp64Cholesky ? 1 : 0, o);
    if (rc != CRV_OK) throw HipError(std::string("simulate: ") + crv_last_error(out[0]->context()));
  }
  return 0;
}

}  // namespace crava_hip
