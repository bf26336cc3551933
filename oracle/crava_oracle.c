/*
 * ORACLE -- test infrastructure, not product code.  Only tests/, the smoke
 * check and bench.py's CPU baseline may load this library, and only to check
 * or to time the reference algorithm on the CPU.
 *
 * Plain-C restatement of the reference algorithm on the hot path:
 *   - RandomGen (libs/lib/random.cpp:38-119, random.h:11-14): 32-bit LCG
 *     s <- 69069 s + 1, u = s 2^-32, Marsaglia-Bray normals;
 *   - FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1191-1240);
 *   - lib_matr complex routines used by invertSeismicData and simulate
 *     (libs/lib/lib_matr.c:78-167 Cholesky, 330-365 AXeqB, 453-470
 *     ProdCholVec, 525-566 products, 674-737 mat-vec, 745-800 add/sub/adjoint);
 *   - the per-wavenumber posterior loop (post_loop.inc), in float (the
 *     reference precision) and in double (the FP64 oracle);
 *   - the simulate scaling loop (src/crava.cpp:1617-1656);
 *   - FFTGrid::findClosestFactorableNumber (src/fftgrid.cpp:2857-2885).
 *
 * Pinning (tests/test_oracle.py): the RNG, Cholesky and posterior loop are
 * checked bit-for-bit against the reference's own sources compiled into
 * oracle/_ref (libs/lib/random.cpp, libs/lib/lib_matr.c) and against golden
 * vectors written from them (tests/golden/); the posterior additionally
 * against the closed-form Gaussian update; the padding search against brute
 * force.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float re, im; } ocf;
typedef struct { double re, im; } ocd;

/* ------------------------------------------------------------------ RNG */
static double or_unif01(uint32_t* s) {
  *s = 69069u * *s + 1u;
  return (double)*s * ((1.0 / 1073741824.0) / 4.0);
}

static double or_g(double x) {
  const double a = 17.49731196, b = 2.36785163, c = 2.15787544;
  const double ax = fabs(x);
  double r = a * exp(-x * x / 2.0);
  if (ax < 1.0) r -= 2.0 * b * (3.0 - x * x) + c * (1.5 - ax);
  else if (ax < 1.5) r -= b * (3.0 - ax) * (3.0 - ax) + c * (1.5 - ax);
  else r -= b * (3.0 - ax) * (3.0 - ax);
  return r;
}

double or_rnorm01(uint32_t* s) {
  double u, u1, u2, u3, x;
  u = or_unif01(s);
  if (u < 0.8638) {
    u1 = or_unif01(s); u2 = or_unif01(s); u3 = or_unif01(s);
    x = 2.0 * (u1 + u2 + u3) - 3.0;
  } else if (u < 0.9745) {
    u1 = or_unif01(s); u2 = or_unif01(s);
    x = 1.5 * (u1 + u2 - 1.0);
  } else if (u < 0.9973002039) {
    do {
      u1 = or_unif01(s); u2 = or_unif01(s);
      x = 6.0 * u1 - 3.0;
    } while (0.358 * u2 > or_g(x));
  } else {
    double v1, v2, w, cc, sv, tv;
    do {
      do {
        u1 = or_unif01(s); u2 = or_unif01(s);
        v1 = 2.0 * u1 - 1.0; v2 = 2.0 * u2 - 1.0;
        w = v1 * v1 + v2 * v2;
      } while (w >= 1.0);
      cc = sqrt((9.0 - 2.0 * log(w)) / w);
      sv = cc * v1; tv = cc * v2;
    } while (fabs(sv) <= 3.0 && fabs(tv) <= 3.0);
    x = (fabs(sv) > 3.0) ? sv : tv;
  }
  return x;
}

/* n normals from *state; state updated */
void or_rnorm_seq(uint32_t* state, long n, double* out) {
  for (long i = 0; i < n; i++) out[i] = or_rnorm01(state);
}

/* --------------------------------------------- FFTGrid::fillInComplexNoise */
/* out: 2*csize floats (or doubles when out_d != NULL) */
static void fill_noise_impl(int nxp, int nyp, int nzp, uint32_t* state, float* out, double* out_d) {
  const int cnxp = nxp / 2 + 1;
  const long csize = (long)cnxp * nyp * nzp;
  const float sd = (float)(1 / sqrt(2.0));
  for (long i = 0; i < csize; i++) {
    double re, im;
    if (((i % cnxp) == 0) || (((i % cnxp) == cnxp - 1) && ((i % 2) == 1))) {
      const long xshift = i % cnxp;
      const long jkind = (i - xshift) / cnxp;
      const long jind = jkind % nyp, kind = jkind / nyp;
      const long jcc = (jind == 0) ? 0 : nyp - jind, kcc = (kind == 0) ? 0 : nzp - kind;
      const long jkcc = jcc + kcc * nyp;
      if (jkcc == jkind) {
        re = (float)or_rnorm01(state);
        im = 0;
      } else if (jkcc > jkind) {
        re = (float)(sd * or_rnorm01(state));
        im = (float)(sd * or_rnorm01(state));
      } else {
        const long cci = jkcc * cnxp + xshift;
        if (out) { re = out[2 * cci]; im = -out[2 * cci + 1]; }
        else { re = out_d[2 * cci]; im = -out_d[2 * cci + 1]; }
      }
    } else {
      re = (float)(sd * or_rnorm01(state));
      im = (float)(sd * or_rnorm01(state));
    }
    if (out) { out[2 * i] = (float)re; out[2 * i + 1] = (float)im; }
    else { out_d[2 * i] = re; out_d[2 * i + 1] = im; }
  }
}
void or_fill_complex_noise(int nxp, int nyp, int nzp, uint32_t* state, float* out) {
  fill_noise_impl(nxp, nyp, nzp, state, out, NULL);
}
void or_fill_complex_noise_d(int nxp, int nyp, int nzp, uint32_t* state, double* out) {
  fill_noise_impl(nxp, nyp, nzp, state, NULL, out);
}

long or_noise_draws(int nxp, int nyp, int nzp) {
  const int cnxp = nxp / 2 + 1;
  long draws = 0;
  for (long jk = 0; jk < (long)nyp * nzp; jk++) {
    const long j = jk % nyp, k = jk / nyp;
    const long jcc = (j == 0) ? 0 : nyp - j, kcc = (k == 0) ? 0 : nzp - k;
    const long jkcc = jcc + kcc * nyp;
    for (int x = 0; x < cnxp; x++) {
      const long i = jk * cnxp + x;
      if (x == 0 || (x == cnxp - 1 && (i % 2) == 1)) draws += (jkcc == jk) ? 1 : (jkcc > jk ? 2 : 0);
      else draws += 2;
    }
  }
  return draws;
}

/* --------------------------------------------------- lib_matr restatement */
#define OR_TOL 1e-20

#define DEFINE_MATR(SUF, R, C)                                                                       \
  int or_chol##SUF(int n, C** x) {                                                                   \
    C lr;                                                                                            \
    R help;                                                                                          \
    R factor = x[0][0].re;                                                                           \
    if (factor <= 0) return 1;                                                                       \
    for (int i = 0; i < n; i++)                                                                      \
      for (int j = 0; j < n; j++) { x[i][j].re = x[i][j].re / factor; x[i][j].im = x[i][j].im / factor; } \
    for (int i = 0; i < n; i++) {                                                                    \
      if (x[i][i].re <= OR_TOL) return 1;                                                            \
      for (int j = 0; j < i; j++) {                                                                  \
        lr.re = 0; lr.im = 0;                                                                        \
        for (int k = 0; k < j; k++) {                                                                \
          lr.re += (x[i][k].re * x[j][k].re) + (x[i][k].im * x[j][k].im);                            \
          lr.im += -(x[i][k].re * x[j][k].im) + (x[i][k].im * x[j][k].re);                           \
        }                                                                                            \
        help = x[j][j].re * x[j][j].re + x[j][j].im * x[j][j].im;                                    \
        const R are = x[i][j].re, aim = x[i][j].im;                                                  \
        x[i][j].re = ((are - lr.re) * x[j][j].re + (aim - lr.im) * x[j][j].im) / help;               \
        x[i][j].im = ((aim - lr.im) * x[j][j].re - (are - lr.re) * x[j][j].im) / help;               \
      }                                                                                              \
      lr.re = 0;                                                                                     \
      for (int k = 0; k < i; k++) lr.re += (x[i][k].re * x[i][k].re) + (x[i][k].im * x[i][k].im);    \
      lr.re = x[i][i].re - lr.re;                                                                    \
      if (lr.re <= OR_TOL) return 1;                                                                 \
      x[i][i].re = (R)sqrt(lr.re);                                                                   \
      x[i][i].im = 0;                                                                                \
    }                                                                                                \
    for (int i = 0; i < n; i++)                                                                      \
      for (int j = 0; j <= i; j++) { x[i][j].re *= (R)sqrt(factor); x[i][j].im *= (R)sqrt(factor); } \
    return 0;                                                                                        \
  }                                                                                                  \
  void or_axeqb##SUF(int n, C** L, C** X, int ncol) {                                                \
    C lx;                                                                                            \
    R help;                                                                                          \
    for (int c = 0; c < ncol; c++) {                                                                 \
      for (int i = 0; i < n; i++) {                                                                  \
        lx = X[i][c];                                                                                \
        for (int j = 0; j < i; j++) {                                                                \
          lx.re -= (X[j][c].re * L[i][j].re - X[j][c].im * L[i][j].im);                              \
          lx.im -= (X[j][c].im * L[i][j].re + X[j][c].re * L[i][j].im);                              \
        }                                                                                            \
        help = L[i][i].re * L[i][i].re + L[i][i].im * L[i][i].im;                                    \
        X[i][c].re = (lx.re * L[i][i].re + lx.im * L[i][i].im) / help;                               \
        X[i][c].im = (lx.im * L[i][i].re - lx.re * L[i][i].im) / help;                               \
      }                                                                                              \
      for (int i = n - 1; i >= 0; i--) {                                                             \
        lx = X[i][c];                                                                                \
        for (int j = n - 1; j > i; j--) {                                                            \
          lx.re = lx.re - (X[j][c].re * L[j][i].re + X[j][c].im * L[j][i].im);                       \
          lx.im = lx.im - (-X[j][c].re * L[j][i].im + X[j][c].im * L[j][i].re);                      \
        }                                                                                            \
        help = L[i][i].re * L[i][i].re + L[i][i].im * L[i][i].im;                                    \
        X[i][c].re = (lx.re * L[i][i].re - lx.im * L[i][i].im) / help;                               \
        X[i][c].im = (lx.im * L[i][i].re + lx.re * L[i][i].im) / help;                               \
      }                                                                                              \
    }                                                                                                \
  }                                                                                                  \
  static void or_prod##SUF(C** a, C** b, int n1, int n2, int n3, C** o) {                            \
    for (int i = 0; i < n1; i++)                                                                     \
      for (int j = 0; j < n3; j++) {                                                                 \
        C x; x.re = 0; x.im = 0;                                                                     \
        for (int k = 0; k < n2; k++) {                                                               \
          x.re += a[i][k].re * b[k][j].re - a[i][k].im * b[k][j].im;                                 \
          x.im += a[i][k].im * b[k][j].re + a[i][k].re * b[k][j].im;                                 \
        }                                                                                            \
        o[i][j] = x;                                                                                 \
      }                                                                                              \
  }                                                                                                  \
  static void or_prodadj##SUF(C** a, C** b, int n1, int n2, int n3, C** o) {                         \
    for (int i = 0; i < n1; i++)                                                                     \
      for (int j = 0; j < n3; j++) {                                                                 \
        C x; x.re = 0; x.im = 0;                                                                     \
        for (int k = 0; k < n2; k++) {                                                               \
          x.re += a[i][k].re * b[j][k].re + a[i][k].im * b[j][k].im;                                 \
          x.im += a[i][k].im * b[j][k].re - a[i][k].re * b[j][k].im;                                 \
        }                                                                                            \
        o[i][j] = x;                                                                                 \
      }                                                                                              \
  }                                                                                                  \
  static void or_addmat##SUF(C** x, int n1, int n2, C** y) {                                         \
    for (int i = 0; i < n1; i++)                                                                     \
      for (int j = 0; j < n2; j++) { y[i][j].re += x[i][j].re; y[i][j].im += x[i][j].im; }           \
  }                                                                                                  \
  static void or_submat##SUF(C** x, int n1, int n2, C** y) {                                         \
    for (int i = 0; i < n1; i++)                                                                     \
      for (int j = 0; j < n2; j++) { y[i][j].re -= x[i][j].re; y[i][j].im -= x[i][j].im; }           \
  }                                                                                                  \
  static void or_adjoint##SUF(C** m, int n1, int n2, C** o) {                                        \
    for (int i = 0; i < n1; i++)                                                                     \
      for (int j = 0; j < n2; j++) { o[j][i].re = m[i][j].re; o[j][i].im = -m[i][j].im; }            \
  }                                                                                                  \
  static void or_matvec##SUF(C** m, C* v, int n1, int n2, C* o) {                                    \
    for (int i = 0; i < n1; i++) {                                                                   \
      C x; x.re = 0; x.im = 0;                                                                       \
      for (int j = 0; j < n2; j++) {                                                                 \
        x.re += m[i][j].re * v[j].re - m[i][j].im * v[j].im;                                         \
        x.im += m[i][j].im * v[j].re + m[i][j].re * v[j].im;                                         \
      }                                                                                              \
      o[i] = x;                                                                                      \
    }                                                                                                \
  }                                                                                                  \
  static void or_adjmatvec##SUF(C** m, C* v, int n1, int n2, C* o) {                                 \
    for (int i = 0; i < n1; i++) {                                                                   \
      C x; x.re = 0; x.im = 0;                                                                       \
      for (int j = 0; j < n2; j++) {                                                                 \
        x.re += m[j][i].re * v[j].re + m[j][i].im * v[j].im;                                         \
        x.im += -m[j][i].im * v[j].re + m[j][i].re * v[j].im;                                        \
      }                                                                                              \
      o[i] = x;                                                                                      \
    }                                                                                                \
  }                                                                                                  \
  static void or_addvec##SUF(C* x, int n, C* y) {                                                    \
    for (int i = 0; i < n; i++) { y[i].re += x[i].re; y[i].im += x[i].im; }                          \
  }                                                                                                  \
  static void or_subvec##SUF(C* x, int n, C* y) {                                                    \
    for (int i = 0; i < n; i++) { y[i].re -= x[i].re; y[i].im -= x[i].im; }                          \
  }                                                                                                  \
  void or_prodcholvec##SUF(int n, C** m, C* v) {                                                     \
    C in[8];                                                                                         \
    for (int i = 0; i < n; i++) { in[i] = v[i]; v[i].re = 0; v[i].im = 0; }                          \
    for (int i = 0; i < n; i++)                                                                      \
      for (int j = 0; j < i + 1; j++) {                                                              \
        v[i].re += m[i][j].re * in[j].re - m[i][j].im * in[j].im;                                    \
        v[i].im += m[i][j].re * in[j].im + m[i][j].im * in[j].re;                                    \
      }                                                                                              \
  }

DEFINE_MATR(_f, float, ocf)
DEFINE_MATR(_d, double, ocd)

/* flat wrappers (row-major n x n complex, interleaved) for tests */
int or_chol_flat_f(int n, float* a) {
  ocf* rows[8];
  for (int i = 0; i < n; i++) rows[i] = (ocf*)(a + 2 * n * i);
  return or_chol_f(n, rows);
}
int or_chol_flat_d(int n, double* a) {
  ocd* rows[8];
  for (int i = 0; i < n; i++) rows[i] = (ocd*)(a + 2 * n * i);
  return or_chol_d(n, rows);
}
/* L (n x n, from chol) and B (n x ncol): solves L L^H X = B in place */
void or_axeqb_flat_f(int n, float* L, float* B, int ncol) {
  ocf *rl[8], *rb[8];
  for (int i = 0; i < n; i++) { rl[i] = (ocf*)(L + 2 * n * i); rb[i] = (ocf*)(B + 2 * ncol * i); }
  or_axeqb_f(n, rl, rb, ncol);
}

/* ------------------------------------------------------ posterior loops */
#define PL_NAME or_posterior_f
#define PL_ERRVAR or_errvar_f
#define PL_PARVAR or_parvar_f
#define PL_REAL float
#define PL_CPX ocf
#define PL_PROD or_prod_f
#define PL_PRODADJ or_prodadj_f
#define PL_ADDMAT or_addmat_f
#define PL_CHOL or_chol_f
#define PL_ADJOINT or_adjoint_f
#define PL_AXEQB or_axeqb_f
#define PL_SUBMAT or_submat_f
#define PL_MATVEC or_matvec_f
#define PL_SUBVEC or_subvec_f
#define PL_ADJMATVEC or_adjmatvec_f
#define PL_ADDVEC or_addvec_f
#include "post_loop.inc"
#undef PL_NAME
#undef PL_ERRVAR
#undef PL_PARVAR
#undef PL_REAL
#undef PL_CPX
#undef PL_PROD
#undef PL_PRODADJ
#undef PL_ADDMAT
#undef PL_CHOL
#undef PL_ADJOINT
#undef PL_AXEQB
#undef PL_SUBMAT
#undef PL_MATVEC
#undef PL_SUBVEC
#undef PL_ADJMATVEC
#undef PL_ADDVEC

#define PL_NAME or_posterior_d
#define PL_ERRVAR or_errvar_d
#define PL_PARVAR or_parvar_d
#define PL_REAL double
#define PL_CPX ocd
#define PL_PROD or_prod_d
#define PL_PRODADJ or_prodadj_d
#define PL_ADDMAT or_addmat_d
#define PL_CHOL or_chol_d
#define PL_ADJOINT or_adjoint_d
#define PL_AXEQB or_axeqb_d
#define PL_SUBMAT or_submat_d
#define PL_MATVEC or_matvec_d
#define PL_SUBVEC or_subvec_d
#define PL_ADJMATVEC or_adjmatvec_d
#define PL_ADDVEC or_addvec_d
#include "post_loop.inc"

/* The posterior loop's own error-variance and parameter-covariance helpers,
   exposed for direct checks.  e1, e2, e3: n complex (interleaved), theta
   n x n, out: n x n complex (interleaved, row-major); covre: the real parts
   of the six covariance spectra, out 3 x 3 complex. */
void or_error_variance_d(int n, const double* e1, const double* e2, const double* e3, const double* theta,
                         double lam, float wnc, double* out) {
  ocd *rows[8];
  for (int l = 0; l < n; l++) rows[l] = (ocd*)(out + 2 * n * l);
  or_errvar_d(n, (const ocd*)e1, (const ocd*)e2, (const ocd*)e3, theta, fabs(lam), wnc, rows);
}
void or_error_variance_f(int n, const float* e1, const float* e2, const float* e3, const double* theta, float lam,
                         float wnc, float* out) {
  ocf *rows[8];
  for (int l = 0; l < n; l++) rows[l] = (ocf*)(out + 2 * n * l);
  or_errvar_f(n, (const ocf*)e1, (const ocf*)e2, (const ocf*)e3, theta, (float)sqrt(lam * lam), wnc, rows);
}
void or_parameter_covariance_d(const double* covre, const double* sigma0, double* out) {
  ocd *rows[3];
  for (int a = 0; a < 3; a++) rows[a] = (ocd*)(out + 6 * a);
  or_parvar_d(covre, sigma0, rows);
}

/* ------------------------------------------ simulate scaling (crava.cpp:1617-1656) */
#define DEFINE_SIMSCALE(SUF, R, C)                                                                    \
  long or_simulate_scale##SUF(long csize, R** cov, R** seed) {                                        \
    long nfail = 0;                                                                                   \
    C Pb[3][3];                                                                                       \
    C* P[3] = {Pb[0], Pb[1], Pb[2]};                                                                  \
    C z[3];                                                                                           \
    for (long i = 0; i < csize; i++) {                                                                \
      P[0][0].re = cov[0][2 * i]; P[0][0].im = cov[0][2 * i + 1];                                     \
      P[1][1].re = cov[1][2 * i]; P[1][1].im = cov[1][2 * i + 1];                                     \
      P[2][2].re = cov[2][2 * i]; P[2][2].im = cov[2][2 * i + 1];                                     \
      P[0][1].re = cov[3][2 * i]; P[0][1].im = cov[3][2 * i + 1];                                     \
      P[0][2].re = cov[4][2 * i]; P[0][2].im = cov[4][2 * i + 1];                                     \
      P[1][2].re = cov[5][2 * i]; P[1][2].im = cov[5][2 * i + 1];                                     \
      P[1][0].re = P[0][1].re; P[1][0].im = -P[0][1].im;                                              \
      P[2][0].re = P[0][2].re; P[2][0].im = -P[0][2].im;                                              \
      P[2][1].re = P[1][2].re; P[2][1].im = -P[1][2].im;                                              \
      for (int g = 0; g < 3; g++) { z[g].re = seed[g][2 * i]; z[g].im = seed[g][2 * i + 1]; }          \
      if (or_chol##SUF(3, P) == 0) {                                                                  \
        or_prodcholvec##SUF(3, P, z);                                                                 \
      } else {                                                                                        \
        nfail++;                                                                                      \
        for (int g = 0; g < 3; g++) { z[g].re = 0; z[g].im = 0; }                                     \
      }                                                                                               \
      for (int g = 0; g < 3; g++) { seed[g][2 * i] = z[g].re; seed[g][2 * i + 1] = z[g].im; }          \
    }                                                                                                 \
    return nfail;                                                                                     \
  }
DEFINE_SIMSCALE(_f, float, ocf)
DEFINE_SIMSCALE(_d, double, ocd)

/* ------------------------------------------ padding (fftgrid.cpp:2857-2885) */
int or_find_closest_factorable(int leastint) {
  const int m2 = (int)ceil((double)log((float)leastint) / log(2.0f));
  const int m3 = (int)ceil((double)log((float)leastint) / log(3.0f));
  const int m5 = (int)ceil((double)log((float)leastint) / log(5.0f));
  const int m7 = (int)ceil((double)log((float)leastint) / log(7.0f));
  int closest = (int)pow(2.0, m2);
  for (int i = 0; i <= m2; i++)
    for (int j = 0; j <= m3; j++)
      for (int k = 0; k <= m5; k++)
        for (int l = 0; l <= m7; l++) {
          const double f = pow(2.0, i) * pow(3.0, j) * pow(5.0, k) * pow(7.0, l);
          if (f > 2147483647.0) continue;
          if ((int)f >= leastint && (int)f < closest) closest = (int)f;
        }
  return closest;
}
