// ORACLE (test infrastructure only).  C entry points over the reference's
// rplib DEM, fluid and Gassmann routines, compiled from rplib/dem.cpp,
// rplib/orddiffeqsolver.cpp, rplib/demmodelling.cpp, rplib/rockgassmann.cpp
// and the rock/dry-rock/fluid classes it is built from (rplib/rock.cpp,
// dryrock.cpp, fluid.cpp, dryrocktabulatedmodulus.cpp,
// fluidtabulatedmodulus.cpp) into oracle/_ref/libcravaref.so by
// oracle/Makefile.  No reference code here.
#include <vector>

#include "rplib/demmodelling.h"
#include "rplib/dryrocktabulatedmodulus.h"
#include "rplib/fluidtabulatedmodulus.h"
#include "rplib/rockgassmann.h"

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

// RockGassmann(fluid, dryrock) -> ComputeSeismicAndElasticParams
// (rockgassmann.cpp:8-17,79-97) on a dry rock and a fluid of given moduli,
// as the reference builds a Gassmann rock from DistributionsRockGassmann.
// Units as rplib: moduli kPa, density g/cm3, velocities m/s.
void ref_gassmann(double k_dry, double mu_dry, double rho_dry, double porosity, double k_mineral, double k_fluid,
                  double rho_fluid, double* out /* k, mu, rho, vp, vs */) {
  const std::vector<double> u;
  DryRockTabulatedModulus dry(k_dry, mu_dry, rho_dry, porosity, k_mineral, u);
  FluidTabulatedModulus fluid(k_fluid, rho_fluid, u);
  RockGassmann rock(&fluid, &dry);
  double k, mu, rho, vp, vs, rho2;
  rock.GetElasticParams(k, mu, rho);
  rock.GetSeismicParams(vp, vs, rho2);
  out[0] = k; out[1] = mu; out[2] = rho; out[3] = vp; out[4] = vs;
}

}  // extern "C"
