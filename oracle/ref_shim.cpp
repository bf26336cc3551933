// ORACLE (test infrastructure only).  C entry points over the reference's
// own numerics, compiled from the sources under the reference tree into
// oracle/_ref/libcravaref.so by oracle/Makefile:
//   libs/fft (FFTW 2.1.2, single precision), libs/lib/lib_matr.c,
//   libs/lib/random.cpp.
// This file holds no reference code: it calls those routines in the order
// the reference's callers do (src/fftgrid.cpp:1783-1841 for the transforms,
// src/fftgrid.cpp:1191-1240 for the noise fill, src/crava.cpp:1076-1247 and
// 1617-1656 for the loops via post_loop.inc).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <unistd.h>

#include "fftw.h"
#include "rfftw.h"
#include "lib/lib_matr.h"
#include "lib/random.h"

extern "C" {

// ---------------------------------------------------------------- RandomGen
static unsigned read_state(const RandomGen& g) {
  char name[] = "/tmp/crava_ref_seedXXXXXX";
  const int fd = mkstemp(name);
  if (fd < 0) return 0;
  close(fd);
  g.writeSeedFile(name);
  FILE* f = fopen(name, "r");
  long long v = 0;
  if (f) { if (fscanf(f, "%lld", &v) != 1) v = 0; fclose(f); }
  unlink(name);
  return (unsigned)v;
}

// n normals from a generator seeded with `seed`; returns the final state
unsigned ref_rnorm_seq(unsigned seed, long n, double* out) {
  RandomGen g(seed);
  for (long i = 0; i < n; ++i) out[i] = RandomGen::rnorm01();
  return read_state(g);
}

void ref_unif_seq(unsigned seed, long n, double* out) {
  RandomGen g(seed);
  for (long i = 0; i < n; ++i) out[i] = RandomGen::unif01();
}

// FFTGrid::fillInComplexNoise with the reference generator; returns state
unsigned ref_fill_complex_noise(int nxp, int nyp, int nzp, unsigned seed, float* out) {
  RandomGen g(seed);
  const int cnxp = nxp / 2 + 1;
  const long csize = (long)cnxp * nyp * nzp;
  fftw_complex* cv = reinterpret_cast<fftw_complex*>(out);
  const float sd = float(1 / sqrt(2.0));
  for (long i = 0; i < csize; ++i) {
    if (((i % cnxp) == 0) || (((i % cnxp) == cnxp - 1) && ((i % 2) == 1))) {
      const long xshift = i % cnxp, jkind = (i - xshift) / cnxp;
      const long jind = jkind % nyp, kind = jkind / nyp;
      const long jcc = (jind == 0) ? 0 : nyp - jind, kcc = (kind == 0) ? 0 : nzp - kind;
      const long jkcc = jcc + kcc * nyp;
      if (jkcc == jkind) {
        cv[i].re = float(RandomGen::rnorm01());
        cv[i].im = 0;
      } else if (jkcc > jkind) {
        cv[i].re = float(sd * RandomGen::rnorm01());
        cv[i].im = float(sd * RandomGen::rnorm01());
      } else {
        const long cci = jkcc * cnxp + xshift;
        cv[i].re = cv[cci].re;
        cv[i].im = -cv[cci].im;
      }
    } else {
      cv[i].re = float(sd * RandomGen::rnorm01());
      cv[i].im = float(sd * RandomGen::rnorm01());
    }
  }
  return read_state(g);
}

// ---------------------------------------------------------------- lib_matr
int ref_chol_flat(int n, float* a) {
  fftw_complex* rows[8];
  for (int i = 0; i < n; ++i) rows[i] = reinterpret_cast<fftw_complex*>(a + 2 * n * i);
  return lib_matrCholCpx(n, rows);
}

void ref_axeqb_flat(int n, float* L, float* B, int ncol) {
  fftw_complex *rl[8], *rb[8];
  for (int i = 0; i < n; ++i) {
    rl[i] = reinterpret_cast<fftw_complex*>(L + 2 * n * i);
    rb[i] = reinterpret_cast<fftw_complex*>(B + 2 * ncol * i);
  }
  lib_matrAXeqBMatCpx(n, rl, rb, ncol);
}

#define PL_NAME ref_posterior
#define PL_ERRVAR ref_errvar
#define PL_PARVAR ref_parvar
#define PL_REAL float
#define PL_CPX fftw_complex
#define PL_PROD lib_matrProdCpx
#define PL_PRODADJ lib_matrProdAdjointCpx
#define PL_ADDMAT lib_matrAddMatCpx
#define PL_CHOL lib_matrCholCpx
#define PL_ADJOINT lib_matrAdjoint
#define PL_AXEQB lib_matrAXeqBMatCpx
#define PL_SUBMAT lib_matrSubtMatCpx
#define PL_MATVEC lib_matrProdMatVecCpx
#define PL_SUBVEC lib_matrSubtVecCpx
#define PL_ADJMATVEC lib_matrProdAdjointMatVecCpx
#define PL_ADDVEC lib_matrAddVecCpx
#include "post_loop.inc"

// Crava::simulate's per-wavenumber loop (crava.cpp:1617-1656)
long ref_simulate_scale(long csize, float** cov, float** seed) {
  long nfail = 0;
  fftw_complex Pb[3][3];
  fftw_complex* P[3] = {Pb[0], Pb[1], Pb[2]};
  fftw_complex z[3];
  for (long i = 0; i < csize; ++i) {
    P[0][0].re = cov[0][2 * i]; P[0][0].im = cov[0][2 * i + 1];
    P[1][1].re = cov[1][2 * i]; P[1][1].im = cov[1][2 * i + 1];
    P[2][2].re = cov[2][2 * i]; P[2][2].im = cov[2][2 * i + 1];
    P[0][1].re = cov[3][2 * i]; P[0][1].im = cov[3][2 * i + 1];
    P[0][2].re = cov[4][2 * i]; P[0][2].im = cov[4][2 * i + 1];
    P[1][2].re = cov[5][2 * i]; P[1][2].im = cov[5][2 * i + 1];
    P[1][0].re = P[0][1].re; P[1][0].im = -P[0][1].im;
    P[2][0].re = P[0][2].re; P[2][0].im = -P[0][2].im;
    P[2][1].re = P[1][2].re; P[2][1].im = -P[1][2].im;
    for (int g = 0; g < 3; ++g) { z[g].re = seed[g][2 * i]; z[g].im = seed[g][2 * i + 1]; }
    if (lib_matrCholCpx(3, P) == 0) {
      lib_matrProdCholVec(3, P, z);
    } else {
      ++nfail;
      for (int g = 0; g < 3; ++g) { z[g].re = 0; z[g].im = 0; }
    }
    for (int g = 0; g < 3; ++g) { seed[g][2 * i] = z[g].re; seed[g][2 * i + 1] = z[g].im; }
  }
  return nfail;
}

// ------------------------------------------------------------- transforms
// FFTGrid::fftInPlace: scale all rsize reals by 1/sqrt(N) unless COVARIANCE
// (3) or OPERATOR (5), then rfftw3d(nzp, nyp, nxp) real-to-complex in place.
void ref_fft3d_r2c(int nxp, int nyp, int nzp, float* data, int cubetype) {
  const long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  if (cubetype != 3 && cubetype != 5) {
    const float s = 1.0f / std::sqrt(static_cast<float>(nxp * nyp * nzp));
    for (long i = 0; i < rsize; ++i) data[i] *= s;
  }
  rfftwnd_plan plan = rfftw3d_create_plan(nzp, nyp, nxp, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_real_to_complex(plan, data, reinterpret_cast<fftw_complex*>(data));
  fftwnd_destroy_plan(plan);
}

// FFTGrid::invFFTInPlace: complex-to-real in place, then scale by 1/N
// (COVARIANCE, OPERATOR) or 1/sqrt(N).
void ref_fft3d_c2r(int nxp, int nyp, int nzp, float* data, int cubetype) {
  const long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  float scale;
  if (cubetype == 3 || cubetype == 5) scale = float(1.0 / (nxp * nyp * nzp));
  else scale = float(1.0 / std::sqrt(float(nxp * nyp * nzp)));
  rfftwnd_plan plan = rfftw3d_create_plan(nzp, nyp, nxp, FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_complex_to_real(plan, reinterpret_cast<fftw_complex*>(data), data);
  fftwnd_destroy_plan(plan);
  for (long i = 0; i < rsize; ++i) data[i] *= scale;
}

}  // extern "C"
