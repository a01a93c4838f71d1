/*
 * ref_harness.cpp -- drives the reference's own compiled code (oracle/_ref, test
 * infrastructure only).  Linked against, and compiled from their places in the CRAVA tree:
 *   libs/fft/fftw/*.c, libs/fft/rfftw/*.c   FFTW 2.1.2, float (libs/fft/include/config.h:90)
 *   libs/lib/lib_matr.c                     small complex linear algebra, symmetric eigen
 *   libs/lib/random.cpp                     RandomGen (LCG + Marsaglia-Bray)
 * The harness restates only the call sequences (FFTGrid::fftInPlace/invFFTInPlace,
 * the per-wavenumber loop of avoinversion.cpp:1016-1083, simulate's colouring 1421-1460,
 * fillInComplexNoise fftgrid.cpp:997-1045); the arithmetic is the reference's.
 * fftgrid.cpp/avoinversion.cpp themselves need a cblas.h the image lacks and are not built.
 */
#include <math.h>
#include <stdint.h>

#include <mutex>
#include <vector>
#include <stdlib.h>
#include <string.h>
#include "fftw.h"
#include "rfftw.h"
#include "lib/lib_matr.h"
#include "lib/random.h"

/* FFTW 2's planner keeps process-wide state (the twiddle cache): plans are created and destroyed
 * under one lock, so that timings on several threads at once are safe; executing a plan is
 * reentrant. */
static std::mutex plan_mu;
static rfftwnd_plan make_plan(int nzp, int nyp, int nxp, fftw_direction dir)
{
  std::lock_guard<std::mutex> lock(plan_mu);
  return rfftw3d_create_plan(nzp, nyp, nxp, dir, FFTW_ESTIMATE | FFTW_IN_PLACE);
}
static void drop_plan(rfftwnd_plan p)
{
  std::lock_guard<std::mutex> lock(plan_mu);
  rfftwnd_destroy_plan(p);
}

extern "C" {

/* FFTGrid::fftInPlace, fftgrid.cpp:1587-1612 */
void ref_grid_fft(int nxp, int nyp, int nzp, int cubetype, float *rvalue)
{
  size_t rsize = (size_t)2 * (nxp / 2 + 1) * nyp * nzp;
  if (cubetype != 3) {
    float s = 1.0f / sqrtf((float)((size_t)nxp * nyp * nzp));
    for (size_t i = 0; i < rsize; i++) rvalue[i] *= s;
  }
  rfftwnd_plan plan = make_plan(nzp, nyp, nxp, FFTW_REAL_TO_COMPLEX);
  rfftwnd_one_real_to_complex(plan, rvalue, (fftw_complex *)rvalue);
  drop_plan(plan);
}

/* FFTGrid::invFFTInPlace, fftgrid.cpp:1615-1645 */
void ref_grid_ifft(int nxp, int nyp, int nzp, int cubetype, float *rvalue)
{
  size_t n = (size_t)nxp * nyp * nzp;
  size_t rsize = (size_t)2 * (nxp / 2 + 1) * nyp * nzp;
  float scale = (cubetype == 3) ? float(1.0 / n) : float(1.0 / sqrtf(float(n)));
  rfftwnd_plan plan = make_plan(nzp, nyp, nxp, FFTW_COMPLEX_TO_REAL);
  rfftwnd_one_complex_to_real(plan, (fftw_complex *)rvalue, rvalue);
  drop_plan(plan);
  for (size_t i = 0; i < rsize; i++) rvalue[i] *= scale;
}

/* 1D in-place r2c of length n (Wavelet::fft1DInPlace, wavelet.cpp:255-270) */
void ref_fft1d_r2c(int n, float *buf)
{
  rfftwnd_plan plan = rfftwnd_create_plan(1, &n, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_real_to_complex(plan, buf, (fftw_complex *)buf);
  rfftwnd_destroy_plan(plan);
}

/* CommonData::ResampleTrace (commondata.h:136-190) with the reference's FFTW 2.1.2 plans as
 * FillInData makes them (commondata.cpp:6531-6538): the n_data samples zero-padded to nt, r2c,
 * the cnt bins into a zero spectrum of cmt bins, c2r of length mt = 10 nt, scaled by 1 / nt.
 * fine receives the rmt = 2 (mt / 2 + 1) floats of the in-place buffer (the reference reads up
 * to index rmt - 1). */
void ref_resample_trace(int n_data, const float *data, int nt, float *fine)
{
  const int mt = 10 * nt, cnt = nt / 2 + 1, rnt = 2 * cnt, cmt = mt / 2 + 1, rmt = 2 * cmt;
  std::vector<float> amp(rnt), ampf(rmt);
  rfftwnd_plan p1, p2;
  {
    std::lock_guard<std::mutex> lock(plan_mu);
    p1 = rfftwnd_create_plan(1, &nt, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
    p2 = rfftwnd_create_plan(1, &mt, FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE | FFTW_IN_PLACE);
  }
  for (int i = 0; i < rnt; i++) amp[i] = i < n_data ? data[i] : 0.0f;
  rfftwnd_one_real_to_complex(p1, amp.data(), (fftw_complex *)amp.data());
  for (int i = 0; i < rmt; i++) ampf[i] = i < 2 * cnt ? amp[i] : 0.0f;
  rfftwnd_one_complex_to_real(p2, (fftw_complex *)ampf.data(), ampf.data());
  const float scale = 1 / static_cast<float>(nt);
  for (int i = 0; i < rmt; i++) fine[i] = scale * ampf[i];
  drop_plan(p1);
  drop_plan(p2);
}

static fftw_complex **cmat(int n1, int n2)
{
  fftw_complex **m = new fftw_complex *[n1];
  for (int i = 0; i < n1; i++) m[i] = new fftw_complex[n2];
  return m;
}
static void cfree(fftw_complex **m, int n1)
{
  for (int i = 0; i < n1; i++) delete[] m[i];
  delete[] m;
}

/* The posterior update in avoinversion.cpp's call order, with lib_matr from the reference.
 * Arguments as orc_posterior_f32 (packed grids), flattened to plain pointers. */
long ref_posterior(int nxp, int nyp, int nzp, int nt, const float *Kt, const float *e1t,
                   const float *e2t, const float *e3t, const int *invert, const double *s0e,
                   float wnc, const double *pv, int prior_mode, const float *rxy, const float *rt,
                   const float *lam_grid, const float *lam2d, float *mean, float *data,
                   float *cov, uint8_t *flags)
{
  const int cnxp = nxp / 2 + 1;
  const size_t csize = (size_t)cnxp * nyp * nzp;
  fftw_complex **K = cmat(nt, 3), **KS = cmat(nt, 3), **KScc = cmat(3, nt), **parVar = cmat(3, 3);
  fftw_complex **margVar = cmat(nt, nt), **errVar = cmat(nt, nt), **reduceVar = cmat(3, 3);
  fftw_complex *e1 = new fftw_complex[nt], *e2 = new fftw_complex[nt], *e3 = new fftw_complex[nt];
  fftw_complex *ijkData = new fftw_complex[nt], *ijkDataMean = new fftw_complex[nt], *ijkRes = new fftw_complex[nt];
  fftw_complex ijkMean[3], ijkAns[3];
  const double s6[6] = {pv[0], pv[4], pv[8], pv[1], pv[2], pv[5]};
  long nfail = 0;
  for (int k = 0; k < nzp; k++) {
    for (int l = 0; l < nt; l++) {
      for (int j = 0; j < 3; j++) {
        K[l][j].re = Kt[2 * ((k * nt + l) * 3 + j)];
        K[l][j].im = Kt[2 * ((k * nt + l) * 3 + j) + 1];
      }
      e1[l].re = e1t[2 * (k * nt + l)]; e1[l].im = e1t[2 * (k * nt + l) + 1];
      e2[l].re = e2t[2 * (k * nt + l)]; e2[l].im = e2t[2 * (k * nt + l) + 1];
      e3[l].re = e3t[2 * (k * nt + l)]; e3[l].im = e3t[2 * (k * nt + l) + 1];
    }
    bool invert_frequency = invert[k] != 0;
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < cnxp; i++) {
        size_t idx = i + (size_t)cnxp * (j + (size_t)nyp * k);
        for (int g = 0; g < 3; g++) {
          ijkMean[g].re = mean[2 * (g * csize + idx)];
          ijkMean[g].im = mean[2 * (g * csize + idx) + 1];
        }
        for (int l = 0; l < nt; l++) {
          ijkData[l].re = data[2 * (l * csize + idx)];
          ijkData[l].im = data[2 * (l * csize + idx) + 1];
          ijkRes[l] = ijkData[l];
        }
        /* getNextParameterCovariance, seismicparametersholder.cpp:485-556 */
        float a6[6];
        for (int g = 0; g < 6; g++) {
          float t;
          if (prior_mode == 0) {
            t = cov[2 * (g * csize + idx)];
            if (s6[g] != 0) t = t / static_cast<float>(s6[g]);
          } else {
            t = rxy[(size_t)j * cnxp + i] * rt[k];
          }
          a6[g] = float(sqrt(t * t)) * static_cast<float>(s6[g]);
        }
        parVar[0][0].re = a6[0]; parVar[1][1].re = a6[1]; parVar[2][2].re = a6[2];
        parVar[0][1].re = parVar[1][0].re = a6[3];
        parVar[0][2].re = parVar[2][0].re = a6[4];
        parVar[1][2].re = parVar[2][1].re = a6[5];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) parVar[a][b].im = 0.0f;
        /* getNextErrorVariance, avoinversion.cpp:1236-1272 */
        float lt = lam_grid ? lam_grid[2 * idx] : lam2d[(size_t)j * cnxp + i];
        float lam = float(sqrt(lt * lt));
        if (invert_frequency)
          for (int l = 0; l < nt; l++)
            for (int m = 0; m < nt; m++) {
              double sl = s0e[l * nt + m];
              errVar[l][m].re = float(0.5 * (1.0 - wnc) * sl * lam * (e1[l].re * e1[m].re + e1[l].im * e1[m].im));
              errVar[l][m].re += float(0.5 * (1.0 - wnc) * sl * lam * (e2[l].re * e2[m].re + e2[l].im * e2[m].im));
              if (l == m) {
                errVar[l][m].re += float(wnc * sl * e3[l].re * e3[l].re);
                errVar[l][m].im = 0.0;
              } else {
                errVar[l][m].im = float(0.5 * (1.0 - wnc) * sl * lam * (-e1[l].re * e1[m].im + e1[l].im * e1[m].re));
                errVar[l][m].im += float(0.5 * (1.0 - wnc) * sl * lam * (-e2[l].re * e2[m].im + e2[l].im * e2[m].re));
              }
            }
        uint8_t flag = 0;
        if (invert_frequency) {
          flag = 1;
          lib_matrProdCpx(K, parVar, nt, 3, 3, KS);
          lib_matrProdAdjointCpx(KS, K, nt, 3, nt, margVar);
          lib_matrAddMatCpx(errVar, nt, nt, margVar);
          if (lib_matrCholCpx(nt, margVar) == 0) {
            lib_matrAdjoint(KS, nt, 3, KScc);
            lib_matrAXeqBMatCpx(nt, margVar, KS, 3);
            lib_matrProdCpx(KScc, KS, 3, nt, 3, reduceVar);
            lib_matrSubtMatCpx(reduceVar, 3, 3, parVar);
            lib_matrProdMatVecCpx(K, ijkMean, nt, 3, ijkDataMean);
            lib_matrSubtVecCpx(ijkDataMean, nt, ijkData);
            lib_matrProdAdjointMatVecCpx(KS, ijkData, 3, nt, ijkAns);
            lib_matrAddVecCpx(ijkAns, 3, ijkMean);
            lib_matrProdMatVecCpx(K, ijkMean, nt, 3, ijkData);
            lib_matrSubtVecCpx(ijkData, nt, ijkRes);
          } else {
            flag = 2;
            nfail++;
          }
        }
        if (flags) flags[idx] = flag;
        for (int g = 0; g < 3; g++) {
          mean[2 * (g * csize + idx)] = ijkMean[g].re;
          mean[2 * (g * csize + idx) + 1] = ijkMean[g].im;
        }
        fftw_complex o6[6] = {parVar[0][0], parVar[1][1], parVar[2][2], parVar[0][1], parVar[0][2], parVar[1][2]};
        for (int g = 0; g < 6; g++) {
          cov[2 * (g * csize + idx)] = o6[g].re;
          cov[2 * (g * csize + idx) + 1] = o6[g].im;
        }
        for (int l = 0; l < nt; l++) {
          data[2 * (l * csize + idx)] = ijkRes[l].re;
          data[2 * (l * csize + idx) + 1] = ijkRes[l].im;
        }
      }
  }
  cfree(K, nt); cfree(KS, nt); cfree(KScc, 3); cfree(parVar, 3);
  cfree(margVar, nt); cfree(errVar, nt); cfree(reduceVar, 3);
  delete[] e1; delete[] e2; delete[] e3;
  delete[] ijkData; delete[] ijkDataMean; delete[] ijkRes;
  return nfail;
}

/* simulate's colouring, avoinversion.cpp:1421-1460 */
long ref_colour(long csize, const float *cov, float *seed)
{
  fftw_complex **P = cmat(3, 3);
  fftw_complex z[3];
  long nfail = 0;
  for (long idx = 0; idx < csize; idx++) {
    P[0][0].re = cov[2 * (0 * csize + idx)]; P[0][0].im = cov[2 * (0 * csize + idx) + 1];
    P[1][1].re = cov[2 * (1 * csize + idx)]; P[1][1].im = cov[2 * (1 * csize + idx) + 1];
    P[2][2].re = cov[2 * (2 * csize + idx)]; P[2][2].im = cov[2 * (2 * csize + idx) + 1];
    P[0][1].re = cov[2 * (3 * csize + idx)]; P[0][1].im = cov[2 * (3 * csize + idx) + 1];
    P[0][2].re = cov[2 * (4 * csize + idx)]; P[0][2].im = cov[2 * (4 * csize + idx) + 1];
    P[1][2].re = cov[2 * (5 * csize + idx)]; P[1][2].im = cov[2 * (5 * csize + idx) + 1];
    P[1][0].re = P[0][1].re; P[1][0].im = -P[0][1].im;
    P[2][0].re = P[0][2].re; P[2][0].im = -P[0][2].im;
    P[2][1].re = P[1][2].re; P[2][1].im = -P[1][2].im;
    for (int g = 0; g < 3; g++) {
      z[g].re = seed[2 * (g * csize + idx)];
      z[g].im = seed[2 * (g * csize + idx) + 1];
    }
    if (lib_matrCholCpx(3, P) == 0) {
      lib_matrProdCholVec(3, P, z);
    } else {
      for (int g = 0; g < 3; g++) z[g].re = z[g].im = 0.0;
      nfail++;
    }
    for (int g = 0; g < 3; g++) {
      seed[2 * (g * csize + idx)] = z[g].re;
      seed[2 * (g * csize + idx) + 1] = z[g].im;
    }
  }
  cfree(P, 3);
  return nfail;
}

/* RandomGen: the stream state is the reference's process-wide static (random.cpp:136). */
void ref_rng_seed(unsigned int seed) { RandomGen g(seed); (void)g; }
double ref_unif01(void) { return RandomGen::unif01(); }
double ref_rnorm01(void) { return RandomGen::rnorm01(); }
void ref_rnorm01_n(double *out, long n) { for (long i = 0; i < n; i++) out[i] = RandomGen::rnorm01(); }

/* fftgrid.cpp:997-1045 call sequence, drawing from the reference RandomGen */
void ref_fill_complex_noise(int nxp, int nyp, int nzp, float *cv)
{
  const size_t cnxp = nxp / 2 + 1;
  const size_t csize = cnxp * nyp * nzp;
  float sd = float(1 / sqrt(2.0));
  for (size_t i = 0; i < csize; i++) {
    if (((i % cnxp) == 0) || (((i % cnxp) == cnxp - 1) && ((i % 2) == 1))) {
      int xshift = i % cnxp;
      int jkind = (i - xshift) / cnxp;
      int jind = jkind % nyp, kind = jkind / nyp;
      int jccind = (jind == 0) ? 0 : nyp - jind;
      int kccind = (kind == 0) ? 0 : nzp - kind;
      int jkccind = jccind + kccind * nyp;
      if (jkccind == jkind) {
        cv[2 * i] = float(RandomGen::rnorm01());
        cv[2 * i + 1] = 0;
      } else if (jkccind > jkind) {
        cv[2 * i] = float(sd * RandomGen::rnorm01());
        cv[2 * i + 1] = float(sd * RandomGen::rnorm01());
      } else {
        size_t cci = jkccind * cnxp + xshift;
        cv[2 * i] = cv[2 * cci];
        cv[2 * i + 1] = -cv[2 * cci + 1];
      }
    } else {
      cv[2 * i] = float(sd * RandomGen::rnorm01());
      cv[2 * i + 1] = float(sd * RandomGen::rnorm01());
    }
  }
}

/* lib_matr_eigen (lib_matr.c:920-1000, tred2 + tqli): a is n x n row-major, vec receives the
 * eigenvectors as columns (row-major), val the eigenvalues; returns the routine's error flag. */
int ref_matr_eigen(int n, const double *a, double *vec, double *val)
{
  double **m = (double **)malloc(sizeof(double *) * n);
  double **v = (double **)malloc(sizeof(double *) * n);
  for (int i = 0; i < n; i++) {
    m[i] = (double *)malloc(sizeof(double) * n);
    v[i] = (double *)malloc(sizeof(double) * n);
    for (int j = 0; j < n; j++) m[i][j] = a[i * n + j];
  }
  int err = 0;
  lib_matr_eigen(m, n, v, val, &err);
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < n; j++) vec[i * n + j] = v[i][j];
    free(m[i]);
    free(v[i]);
  }
  free(m);
  free(v);
  return err;
}

} /* extern "C" */
