// Model set-up restated from the reference for the synthetic configurations:
//  * Ricker wavelet (src/wavelet.cpp:1354-1358) sampled at dt in FFT order,
//    zero beyond its +-119 ms support (src/wavelet.cpp:130-186, 416-463);
//  * difference operators and their spectra (src/wavelet1D.cpp:367-478),
//    time-domain norms (src/wavelet.cpp:790-808);
//  * the per-kz tables of Crava::fillErrMultVectors (src/crava.cpp:1336-1364);
//  * PP reflection rows (src/modelavodynamic.cpp:877-902);
//  * errThetaCov with the generalised-exponential angular correlation
//    (src/crava.cpp:451-487, src/vario.cpp:81-94).
#include <cmath>

#include "crava.h"

namespace crava_hip {

namespace {

const double kPi = 3.14159265358979323846;

std::vector<std::complex<double>> dft(const std::vector<float>& x) {
  const int n = (int)x.size();
  std::vector<std::complex<double>> X(n);
  for (int k = 0; k < n; ++k) {
    std::complex<double> s = 0;
    for (int t = 0; t < n; ++t) {
      const long long r = ((long long)k * t) % n;
      const double a = -2.0 * kPi * (double)r / n;
      s += (double)x[t] * std::complex<double>(std::cos(a), std::sin(a));
    }
    X[k] = s;
  }
  return X;
}

float norm_time(const std::vector<float>& w) {
  double s = 0;
  for (float v : w) s += (double)(v * v);
  return (float)std::sqrt(s);
}

}  // namespace

KzTables makeRickerTables(int nzp, double dt, const std::vector<double>& peakHz) {
  KzTables t;
  const int n = (int)peakHz.size();
  t.nzp = nzp;
  t.ntheta = n;
  t.kwd.resize((size_t)nzp * n);
  t.em1.resize((size_t)nzp * n);
  t.em2.resize((size_t)nzp * n);
  t.em3.assign((size_t)nzp * n, 1.0);
  std::vector<float> d1(nzp, 0.f);
  d1[0] = -1.f;
  d1[nzp - 1] = 1.f;
  const auto D1 = dft(d1);
  for (int l = 0; l < n; ++l) {
    std::vector<float> w(nzp);
    for (int k = 0; k < nzp; ++k) {
      const double tms = (double)(float)(k < nzp / 2 + 1 ? dt * k : dt * (k - nzp));
      const double c = kPi * kPi * peakHz[l] * peakHz[l] * tms * tms * 1e-6;
      w[k] = std::fabs(tms) > 119.0 ? 0.f : (float)((1.0 - 2.0 * c) * std::exp(-c));
    }
    // central(backward(forward(w))), cyclic
    std::vector<float> f(nzp), b(nzp), e3(nzp);
    for (int i = 0; i < nzp; ++i) f[i] = w[(i + 1) % nzp] - w[i];
    for (int i = 0; i < nzp; ++i) b[i] = f[i] - f[(i - 1 + nzp) % nzp];
    for (int i = 0; i < nzp; ++i) e3[i] = (float)(0.5 * ((double)b[(i + 1) % nzp] - b[(i - 1 + nzp) % nzp]));
    const auto W = dft(w);
    const auto E3 = dft(e3);
    const double nw = norm_time(w), n3 = norm_time(e3);
    for (int k = 0; k < nzp; ++k) {
      t.kwd[(size_t)k * n + l] = D1[k] * std::conj(W[k]);
      t.em1[(size_t)k * n + l] = std::conj(W[k]) / nw;
      t.em2[(size_t)k * n + l] = std::conj(E3[k]) / n3;
    }
  }
  return t;
}

std::vector<float> reflectionMatrix(const std::vector<double>& anglesDeg, double vsvp) {
  std::vector<float> A(anglesDeg.size() * 3);
  const double g2 = vsvp * vsvp;
  for (size_t i = 0; i < anglesDeg.size(); ++i) {
    const double th = anglesDeg[i] * kPi / 180.0;
    const double s2 = std::sin(th) * std::sin(th), t2 = std::tan(th) * std::tan(th);
    A[i * 3 + 0] = (float)((1.0 + t2) / 2.0);
    A[i * 3 + 1] = (float)(-4 * g2 * s2);
    A[i * 3 + 2] = (float)((1.0 - 4.0 * g2 * s2) / 2.0);
  }
  return A;
}

std::vector<double> errThetaCov(const std::vector<double>& dataVar, const std::vector<double>& sn,
                                const std::vector<double>& anglesDeg, double angularRangeDeg) {
  const size_t n = dataVar.size();
  std::vector<double> th(n * n);
  const double range = angularRangeDeg * kPi / 180.0;
  for (size_t i = 0; i < n; ++i)
    for (size_t j = 0; j < n; ++j) {
      const float evi = (float)dataVar[i] / (float)sn[i], evj = (float)dataVar[j] / (float)sn[j];
      const float dth = (float)(anglesDeg[i] * kPi / 180.0) - (float)(anglesDeg[j] * kPi / 180.0);
      const float corr = (float)std::exp(-3.0 * std::fabs((double)dth) / range);
      th[i * n + j] = (double)(float)(std::sqrt(evi) * std::sqrt(evj) * corr);
    }
  return th;
}

}  // namespace crava_hip
