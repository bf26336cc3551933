// C entry point over the host C++ classes, used to drive them from tests:
// one inversion of a survey given as host arrays in the reference layout,
// through FFTGrid / SeismicParametersHolder / Crava exactly as the
// reference's Crava constructor sequences them (src/crava.cpp:156-300):
// FFTs of the data and means, computePostMeanResidAndFFTCov, simulate (one
// realization from RandomGen(seed)), invFFTCovGrids.
#include <cstdio>
#include <cstring>
#include <memory>

#include "crava.h"

using namespace crava_hip;

extern "C" int crvh_invert(int nxp, int nyp, int nzp, int ntheta, const float* const* mean_real,
                           const float* const* data_real, const float* const* cov_real, const float* err_real,
                           const double* sigma0, const double* angles_deg, const double* peak_hz, const double* sn,
                           double dt, double vsvp, double low_cut, double high_cut, double wnc, int mode,
                           unsigned int seed, float* const* post_mean_out, float* const* cov_spec_out,
                           float* const* real_out, char* err, int errlen) {
  crv_ctx* ctx = nullptr;
  if (crv_ctx_create(0, &ctx) != CRV_OK) {
    if (err && errlen) snprintf(err, errlen, "no GPU");
    return CRV_E_HIP;
  }
  int rc = 0;
  try {
    const int nx = nxp, ny = nyp, nz = nzp;
    const long long rsize = (long long)2 * (nxp / 2 + 1) * nyp * nzp;
    auto mk = [&](int type) {
      auto g = std::make_unique<FFTGrid>(ctx, nx, ny, nz, nxp, nyp, nzp, CRV_FP32);
      g->setType(type);
      return g;
    };
    std::unique_ptr<FFTGrid> mean[3], dataG[8];
    for (int m = 0; m < 3; ++m) {
      mean[m] = mk(FFTGrid::PARAMETER);
      mean[m]->setRealValues(std::vector<float>(mean_real[m], mean_real[m] + rsize));
    }
    std::vector<double> dataVar(ntheta);
    for (int l = 0; l < ntheta; ++l) {
      dataG[l] = mk(FFTGrid::DATA);
      dataG[l]->setRealValues(std::vector<float>(data_real[l], data_real[l] + rsize));
      double s, s2;
      long long c;
      crv_grid_moments(dataG[l]->handle(), &s, &s2, &c);
      dataVar[l] = (double)(float)(s2 / (double)(c ? c : 1));
    }
    SeismicParametersHolder sp(ctx, nx, ny, nz, nxp, nyp, nzp, CRV_FP32, sigma0);
    for (int q = 0; q < 6; ++q) sp.grids()[q]->setRealValues(std::vector<float>(cov_real[q], cov_real[q] + rsize));
    auto errCorr = mk(FFTGrid::COVARIANCE);
    errCorr->setRealValues(std::vector<float>(err_real, err_real + rsize));
    // crava.cpp:177-184: transform data and prior means before the inversion
    for (int l = 0; l < ntheta; ++l) dataG[l]->fftInPlace();
    for (int m = 0; m < 3; ++m) mean[m]->fftInPlace();
    const std::vector<double> ang(angles_deg, angles_deg + ntheta), pk(peak_hz, peak_hz + ntheta),
        snv(sn, sn + ntheta);
    const KzTables tab = makeRickerTables(nzp, dt, pk);
    std::vector<float> A = reflectionMatrix(ang, vsvp);
    std::vector<double> th = errThetaCov(dataVar, snv, ang, 10.0);
    float* Arow[8];
    double* Trow[8];
    for (int l = 0; l < ntheta; ++l) { Arow[l] = &A[l * 3]; Trow[l] = &th[(size_t)l * ntheta]; }
    FFTGrid* pa = mean[0].get();
    FFTGrid* pb = mean[1].get();
    FFTGrid* pr = mean[2].get();
    FFTGrid* ec = errCorr.get();
    FFTGrid* seis[8];
    for (int l = 0; l < ntheta; ++l) seis[l] = dataG[l].get();
    Precision prec;
    prec.mode = mode;
    Crava::computePostMeanResidAndFFTCov(sp, pa, pb, pr, ec, seis, tab, Trow, Arow, ntheta, (float)low_cut,
                                         (float)high_cut, (float)wnc, nz, nzp, nyp, nxp, nz * dt, 1.0, 1, false,
                                         prec, nullptr);
    std::unique_ptr<FFTGrid> sim[3];
    FFTGrid* simp[3];
    for (int a = 0; a < 3; ++a) { sim[a] = mk(FFTGrid::PARAMETER); simp[a] = sim[a].get(); }
    FFTGrid* pm[3] = {pa, pb, pr};
    RandomGen rg(seed);
    Crava::simulate(sp, pm, &rg, 1, simp, mode == CRV_MODE_FP32_REFINE);
    for (int q = 0; q < 6; ++q) {
      auto c = sp.grids()[q]->getComplexValues();
      std::memcpy(cov_spec_out[q], c.data(), c.size() * sizeof(float));
    }
    sp.invFFTCovGrids();
    for (int a = 0; a < 3; ++a) {
      auto r = mean[a]->getRealValues();
      std::memcpy(post_mean_out[a], r.data(), r.size() * sizeof(float));
      auto s = sim[a]->getRealValues();
      std::memcpy(real_out[a], s.data(), s.size() * sizeof(float));
    }
  } catch (const std::exception& e) {
    if (err && errlen) snprintf(err, errlen, "%s", e.what());
    rc = CRV_E_INVALID;
  }
  crv_ctx_destroy(ctx);
  return rc;
}

// The output path of the host FFTGrid (ParameterOutput's vp output:
// expTransf, then FFTGrid::writeCravaFile; reference src/parameteroutput.cpp:
// 84-93, src/fftgrid.cpp:1726-1743,2411-2445), and readCravaFile of the file
// written (src/fftgrid.cpp:2448-2498) into a second grid: log_real (rsize
// floats) in, the exponentiated grid read back in exp_out.  The exponentiated
// grid is also written with writeStormFile (binary, padded, flat simbox of
// 25 x 25 x 4 cells from the origin) to base_name + ".storm".
extern "C" int crvh_write_read_crava(int nxp, int nyp, int nzp, const float* log_real, const char* base_name,
                                     float* exp_out, char* err, int errlen) {
  crv_ctx* ctx = nullptr;
  if (crv_ctx_create(0, &ctx) != CRV_OK) {
    if (err && errlen) snprintf(err, errlen, "no GPU");
    return CRV_E_HIP;
  }
  int rc = 0;
  try {
    const long long rsize = (long long)2 * (nxp / 2 + 1) * nyp * nzp;
    crv_simbox_info sb{};
    sb.dx = 25.0; sb.dy = 25.0; sb.nx = nxp; sb.ny = nyp;
    {
      FFTGrid g(ctx, nxp, nyp, nzp, nxp, nyp, nzp, CRV_FP32);
      g.setType(FFTGrid::PARAMETER);
      g.setRealValues(std::vector<float>(log_real, log_real + rsize));
      g.expTransf();
      g.writeCravaFile(base_name, sb);
      // the padded grid as a flat STORM binary file (fftgrid.cpp:2117-2150)
      crv_storm_info si{};
      si.lx = 25.0 * nxp; si.ly = 25.0 * nyp; si.lz = 4.0 * nzp;
      g.writeStormFile(base_name, si, false, true, true);
    }
    FFTGrid h(ctx, nxp, nyp, nzp, nxp, nyp, nzp, CRV_FP32);
    std::string errText;
    h.readCravaFile(std::string(base_name) + ".crava", errText);
    if (!errText.empty()) throw HipError(errText);
    auto r = h.getRealValues();
    std::memcpy(exp_out, r.data(), r.size() * sizeof(float));
  } catch (const std::exception& e) {
    if (err && errlen) snprintf(err, errlen, "%s", e.what());
    rc = CRV_E_INVALID;
  }
  crv_ctx_destroy(ctx);
  return rc;
}
