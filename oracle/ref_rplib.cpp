// ORACLE (test infrastructure only).  C entry points over the reference's
// rplib DEM and fluid routines, compiled from rplib/dem.cpp,
// rplib/orddiffeqsolver.cpp and rplib/demmodelling.cpp into
// oracle/_ref/libcravaref.so by oracle/Makefile.  No reference code here.
#include <vector>

#include "rplib/demmodelling.h"

extern "C" {

// DEMTools::CalcEffectiveBulkAndShearModulus (demmodelling.cpp:253-271)
int ref_dem_effective(int n, const double* k, const double* mu, const double* asp, const double* conc, double kbg,
                      double mubg, double* keff, double* mueff) {
  std::vector<double> K(k, k + n), M(mu, mu + n), A(asp, asp + n), C(conc, conc + n);
  try {
    DEMTools::CalcEffectiveBulkAndShearModulus(K, M, A, C, kbg, mubg, *keff, *mueff);
  } catch (...) {
    return 1;
  }
  return 0;
}

double ref_brine_velocity(double t, double p, double s) { return DEMTools::CalcVelocityOfBrineFromTPS(t, p, s); }
double ref_brine_density(double t, double p, double s) { return DEMTools::CalcDensityOfBrineFromTPS(t, p, s); }
double ref_brine_modulus(double t, double p, double s) { return DEMTools::CalcBulkModulusOfBrineFromTPS(t, p, s); }
double ref_water_velocity(double t, double p) { return DEMTools::CalcVelocityOfWaterFromTP(t, p); }
double ref_water_density(double t, double p) { return DEMTools::CalcDensityOfWaterFromTP(t, p); }

double ref_hill(int n, const double* prop, const double* frac) {
  std::vector<double> P(prop, prop + n), F(frac, frac + n);
  return DEMTools::CalcEffectiveElasticModuliUsingHill(P, F);
}

void ref_seismic_from_elastic(double k, double mu, double rho, double* vp, double* vs) {
  DEMTools::CalcSeismicParamsFromElasticParams(k, mu, rho, *vp, *vs);
}

}  // extern "C"
