/*
 * ORACLE -- test infrastructure, not product code.
 *
 * Restatement of the rplib rock-physics forward model that the learned
 * surrogate replaces:
 *   - DEM inclusion model, Berryman's spheroid P and Q factors
 *     (rplib/dem.cpp:39-189; J.G. Berryman, J. Acoust. Soc. Am. 68 (1980));
 *   - adaptive Runge-Kutta-Fehlberg 4(5) integration with the reference's
 *     step control (rplib/orddiffeqsolver.cpp:17-167);
 *   - Hill / Reuss / Voigt mixing (rplib/demmodelling.cpp:178-236);
 *   - Batzle-Wang water and brine (rplib/demmodelling.cpp:29-55, 274-336;
 *     M. Batzle and Z. Wang, Geophysics 57 (1992));
 *   - Gassmann substitution in the reference's form (rplib/rockgassmann.cpp:
 *     79-97);
 *   - vp, vs from moduli (rplib/demmodelling.cpp:338-354).
 * rplib has no Batzle-Wang gas; the gas of the gas-sand facies follows the
 * published Batzle-Wang (1992) gas equations (eqs. 9-11), gas gravity 0.6.
 *
 * Units as rplib: moduli in kPa, density in g/cm3, velocity in m/s,
 * temperature in deg C, pressure in MPa.
 */
#include <math.h>
#include <string.h>

/* ------------------------------------------------------------- DEM RHS */
typedef struct {
  int n;
  double k[4], mu[4], asp[4], conc[4];
} or_dem;

static void dem_rhs(const or_dem* d, const double y[2], double t, double out[2]) {
  double sum_conc = 0.0;
  for (int i = 0; i < d->n; i++) sum_conc += d->conc[i];
  double krhs = 0.0, murhs = 0.0;
  for (int i = 0; i < d->n; i++) {
    const double ka = d->k[i], mua = d->mu[i];
    double asp = d->asp[i];
    const double conc = d->conc[i] / sum_conc;
    const double k = y[0], mu = y[1];
    if (asp == 1.0) asp = 0.99;
    const double theta = (asp / pow(1 - asp * asp, 1.5)) * (acos(asp) - asp * sqrt(1 - asp * asp));
    const double fn = ((asp * asp) / (1 - asp * asp)) * (3 * theta - 2);
    const double nu = (3 * k - 2 * mu) / (2 * (3 * k + mu));
    const double r = (1 - 2 * nu) / (2 * (1 - nu));
    const double a = mua / mu - 1;
    const double b = (1.0 / 3.0) * (ka / k - mua / mu);
    const double f1 = 1 + a * (1.5 * (fn + theta) - r * (1.5 * fn + 2.5 * theta - 4.0 / 3.0));
    double f2 = 1 + a * (1 + 1.5 * (fn + theta) - (r / 2) * (3 * fn + 5 * theta)) + b * (3 - 4 * r);
    f2 += (a / 2) * (a + 3 * b) * (3 - 4 * r) * (fn + theta - r * (fn - theta + 2 * theta * theta));
    const double f3 = 1 + a * (1 - (fn + 1.5 * theta) + r * (fn + theta));
    const double f4 = 1 + (a / 4) * (fn + 3 * theta - r * (fn - theta));
    const double f5 = a * (-fn + r * (fn + theta - 4.0 / 3.0)) + b * theta * (3 - 4 * r);
    const double f6 = 1 + a * (1 + fn - r * (fn + theta)) + b * (1 - theta) * (3 - 4 * r);
    const double f7 = 2 + (a / 4) * (3 * fn + 9 * theta - r * (3 * fn + 5 * theta)) + b * theta * (3 - 4 * r);
    const double f8 = a * (1 - 2 * r + (fn / 2) * (r - 1) + (theta / 2) * (5 * r - 3)) + b * (1 - theta) * (3 - 4 * r);
    const double f9 = a * ((r - 1) * fn - r * theta) + b * theta * (3 - 4 * r);
    const double p = (3 * f1 / f2) / 3.0;
    const double q = ((2 / f3) + (1 / f4) + ((f4 * f5 + f6 * f7 - f8 * f9) / (f2 * f4))) / 5.0;
    krhs += conc * (ka - k) * p;
    murhs += conc * (mua - mu) * q;
  }
  out[0] = krhs / (1 - t);
  out[1] = murhs / (1 - t);
}

/* ------------------------------------------------ RKF45 (reference control) */
/* returns 0 on success, 1 when the step size collapsed ("Singularity likely") */
static int rkf45(const or_dem* d, double t0, double tfinal, double y[2], double tol, int* nsteps) {
  static const double al[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
  static const double be[5][6] = {
      {1.0 / 4.0, 0, 0, 0, 0, 0},
      {3.0 / 32.0, 9.0 / 32.0, 0, 0, 0, 0},
      {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0, 0, 0, 0},
      {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0, 0, 0},
      {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0, -5643.0 / 20520.0, 0}};
  static const double ga[2][6] = {
      {902880.0 / 7618050.0, 0, 3953664.0 / 7618050.0, 3855735.0 / 7618050.0, -1371249.0 / 7618050.0,
       277020.0 / 7618050.0},
      {-2090.0 / 752400.0, 0, 22528.0 / 752400.0, 21970.0 / 752400.0, -15048.0 / 752400.0, -27360.0 / 752400.0}};
  double t = t0;
  const double hmax = (tfinal - t) / 16.0;
  double h = hmax / 8.0;
  double f[6][2];
  int steps = 0;
  while (t < tfinal && (t + h) > t) {
    if (t + h > tfinal) h = tfinal - t;
    dem_rhs(d, y, t, f[0]);
    for (int j = 0; j < 5; j++) {
      double y1[2] = {y[0], y[1]};
      for (int i = 0; i < 2; i++)
        for (int m = 0; m < 6; m++) y1[i] += h * be[j][m] * f[m][i];
      dem_rhs(d, y1, t + al[j] * h, f[j + 1]);
    }
    double e[2] = {0, 0};
    for (int i = 0; i < 2; i++)
      for (int m = 0; m < 6; m++) e[i] += h * ga[1][m] * f[m][i];
    double delta = fabs(e[0]);
    if (fabs(e[1]) > delta) delta = fabs(e[1]);
    double tau = fabs(y[0]);
    if (fabs(y[1]) > tau) tau = fabs(y[1]);
    if (1.0 > tau) tau = 1.0;
    tau *= tol;
    if (delta <= tau) {
      t += h;
      for (int i = 0; i < 2; i++)
        for (int m = 0; m < 6; m++) y[i] += h * ga[0][m] * f[m][i];
      steps++;
    }
    if (delta != 0.0) {
      h = 0.8 * h * pow(tau / delta, 0.2);
      if (hmax < h) h = hmax;
    }
  }
  if (nsteps) *nsteps = steps;
  return (t < tfinal) ? 1 : 0;
}

/* DEM::CalcEffectiveModulus (dem.cpp:39-101) for n inclusion types */
int or_dem_effective(int n, const double* k, const double* mu, const double* asp, const double* conc, double kbg,
                     double mubg, double* keff, double* mueff, int* nsteps) {
  or_dem d;
  d.n = n;
  double sum = 0.0;
  for (int i = 0; i < n; i++) {
    d.k[i] = k[i]; d.mu[i] = mu[i]; d.asp[i] = asp[i]; d.conc[i] = conc[i];
    sum += conc[i];
  }
  if (n > 1 && sum == 1.0) {
    for (int i = 0; i < n; i++) d.conc[i] *= 0.999999;
    sum *= 0.999999;
  }
  if (sum == 1.0) {
    *keff = k[n - 1];
    *mueff = mu[n - 1];
    return 0;
  }
  double y[2] = {kbg, mubg};
  const int rc = rkf45(&d, 0.0, sum, y, 1e-5, nsteps);
  *keff = y[0];
  *mueff = y[1];
  return rc;
}

/* ------------------------------------------------------- fluids (B&W) */
double or_water_velocity(double T, double P) {
  static const double w[5][4] = {{1402.85, 1.524, 3.437E-3, -1.197E-5},
                                 {4.871, -0.0111, 1.739E-4, -1.628E-6},
                                 {-0.04783, 2.747E-4, -2.135E-6, 1.237E-8},
                                 {1.487E-4, -6.503E-7, -1.455E-8, 1.327E-10},
                                 {-2.197E-7, 7.987E-10, 5.230E-11, -4.614E-13}};
  double s = 0.0;
  for (int i = 0; i < 5; i++)
    for (int j = 0; j < 4; j++) s += w[i][j] * pow(T, i) * pow(P, j);
  return s;
}

double or_water_density(double T, double P) {
  return 1.0 + 1E-6 * (-80 * T - 3.3 * (T * T) + 0.00175 * (T * T * T) + 489 * P - 2 * T * P + 0.016 * (T * T) * P -
                       1.3E-5 * (T * T * T) * P - 0.333 * (P * P) - 0.002 * T * (P * P));
}

double or_brine_velocity(double T, double P, double S) {
  return or_water_velocity(T, P) +
         S * (1170.0 - 9.6 * T + 0.055 * (T * T) - 8.5E-5 * (T * T * T) + 2.6 * P - 0.0029 * T * P - 0.0476 * (P * P)) +
         pow(S, 1.5) * (780.0 - 10.0 * P + 0.16 * (P * P)) - 1820 * (S * S);
}

double or_brine_density(double T, double P, double S) {
  return or_water_density(T, P) +
         S * (0.668 + 0.44 * S + 1E-6 * (300 * P - 2400 * P * S + T * (80 + 3 * T - 3300 * S - 13 * P + 47 * P * S)));
}

/* brine bulk modulus in kPa (rplib: rho v^2 / 1e3 in MPa) */
double or_brine_modulus_kpa(double T, double P, double S) {
  const double v = or_brine_velocity(T, P, S);
  return or_brine_density(T, P, S) * v * v / 1E3 * 1E3;
}

/* Batzle & Wang (1992) gas: density (g/cm3) and adiabatic modulus (kPa) */
void or_gas_props(double T, double P, double G, double* rho, double* k_kpa) {
  const double Ta = T + 273.15;
  const double Tpr = Ta / (94.72 + 170.75 * G);
  const double Ppr = P / (4.892 - 0.4048 * G);
  const double ex = (0.45 + 8.0 * (0.56 - 1.0 / Tpr) * (0.56 - 1.0 / Tpr)) * pow(Ppr, 1.2) / Tpr;
  const double E = 0.109 * (3.85 - Tpr) * (3.85 - Tpr) * exp(-ex);
  const double Z = (0.03 + 0.00527 * pow(3.5 - Tpr, 3)) * Ppr + (0.642 * Tpr - 0.007 * pow(Tpr, 4) - 0.52) + E;
  const double dEdP = E * (-(0.45 + 8.0 * (0.56 - 1.0 / Tpr) * (0.56 - 1.0 / Tpr)) * 1.2 * pow(Ppr, 0.2) / Tpr);
  const double dZdP = (0.03 + 0.00527 * pow(3.5 - Tpr, 3)) + dEdP;
  const double R = 8.31441;
  *rho = 28.8 * G * P / (Z * R * Ta);
  const double g0 = 0.85 + 5.6 / (Ppr + 2) + 27.1 / ((Ppr + 3.5) * (Ppr + 3.5)) - 8.7 * exp(-0.65 * (Ppr + 1));
  const double k_mpa = P * g0 / (1.0 - Ppr / Z * dZdP);
  *k_kpa = k_mpa * 1E3;
}

/* ---------------------------------------------- mixing and substitution */
/* Hill average of n constituents (rplib/demmodelling.cpp:178-236): the mean
   of the Voigt (arithmetic) and Reuss (harmonic) averages by volume fraction. */
double or_hill(int n, const double* prop, const double* frac) {
  double v = 0.0, r = 0.0;
  int zero = 0;
  for (int i = 0; i < n; i++) {
    v += frac[i] * prop[i];
    if (prop[i] == 0.0) zero = 1;
  }
  /* the Reuss average is 0 when a constituent has zero modulus */
  if (!zero) {
    for (int i = 0; i < n; i++) r += frac[i] / prop[i];
    r = 1.0 / r;
  }
  return 0.5 * (r + v);
}

/* RockGassmann::ComputeSeismicAndElasticParams in the reference's form
   (rplib/rockgassmann.cpp:79-97), including its "+ k_dry / K_min^2" term in
   the denominator; shear modulus unchanged, density rho_dry + phi rho_fl,
   then vp, vs (rplib/demmodelling.cpp:338-354).  out: k, mu, rho, vp, vs. */
void or_gassmann(double k_dry, double mu_dry, double rho_dry, double phi, double k_min, double k_fl, double rho_fl,
                 double* out) {
  const double t1 = (1 - k_dry / k_min) * (1 - k_dry / k_min);
  const double k = k_dry + t1 / (phi / k_fl + (1 - phi) / k_min + k_dry / (k_min * k_min));
  const double rho = rho_dry + phi * rho_fl;
  out[0] = k;
  out[1] = mu_dry;
  out[2] = rho;
  out[3] = sqrt((k + 4.0 / 3.0 * mu_dry) / rho);
  out[4] = sqrt(mu_dry / rho);
}

/* ------------------------------------------------------ one rock sample */
typedef struct {
  double kq, gq, rhoq;     /* quartz: kPa, kPa, g/cm3 */
  double kc, gc, rhoc;     /* clay */
  double temp_c;           /* fluid temperature */
  double salinity;         /* brine salinity (weight fraction) */
  double gas_gravity;
  int gas;                 /* 0: brine + gas mixture by Sw; facies rules applied by the caller */
} or_rock_params;

/* inputs: porosity, clay fraction, water saturation, aspect ratio, effective
   pressure (MPa).  outputs: ln vp, ln vs (m/s), ln rho (kg/m3).  Returns 0
   on success, 1 when the DEM integration failed. */
int or_rock_forward(const or_rock_params* p, const double in[5], double out[3]) {
  const double phi = in[0], clay = in[1], sw = in[2], asp = in[3], pe = in[4];
  /* mineral: Hill average of quartz and clay */
  const double vq = 1.0 - clay, vc = clay;
  const double fr[2] = {vq, vc}, ks[2] = {p->kq, p->kc}, gs[2] = {p->gq, p->gc};
  const double kmin = or_hill(2, ks, fr), gmin = or_hill(2, gs, fr);
  const double rhomin = vq * p->rhoq + vc * p->rhoc;
  /* dry frame: DEM with empty spheroidal pores of aspect asp */
  double k0 = 0.0, g0 = 0.0, a = asp, c = phi, kd, gd;
  if (or_dem_effective(1, &k0, &g0, &a, &c, kmin, gmin, &kd, &gd, 0)) return 1;
  /* fluid: brine and gas, Reuss mix of moduli, linear density */
  const double kb = or_brine_modulus_kpa(p->temp_c, pe, p->salinity);
  const double rb = or_brine_density(p->temp_c, pe, p->salinity);
  double rg, kg;
  or_gas_props(p->temp_c, pe, p->gas_gravity, &rg, &kg);
  const double kfl = 1.0 / (sw / kb + (1.0 - sw) / kg);
  const double rfl = sw * rb + (1.0 - sw) * rg;
  /* Gassmann in the reference's form (rockgassmann.cpp:90); the dry frame
     of empty pores has density (1 - phi) rho_mineral */
  double gm[5];
  or_gassmann(kd, gd, (1 - phi) * rhomin, phi, kmin, kfl, rfl, gm);
  out[0] = log(gm[3]);
  out[1] = log(gm[4]);
  out[2] = log(gm[2] * 1000.0);
  return 0;
}

void or_rock_forward_batch(const or_rock_params* p, long n, const double* in, double* out, int* status) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 256)
#endif
  for (long i = 0; i < n; i++) status[i] = or_rock_forward(p, in + 5 * i, out + 3 * i);
}
