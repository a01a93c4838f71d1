/*
 * orc_setup.c -- host-side setup restated from CRAVA (test infrastructure only; see oracle.h).
 * These produce the small per-angle / per-k_z tables and correlation factors that the
 * per-wavenumber update consumes.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* wavelet.cpp:1484-1488 (t in ms, peak frequency in Hz) */
double orc_ricker(double t_ms, double peak_hz)
{
  double c = M_PI * M_PI * peak_hz * peak_hz * t_ms * t_ms * 1e-6;
  return (1 - 2 * c) * exp(-c);
}

/* commondata.cpp:2938-2942, PP row of the Aki-Richards matrix (angles in radians) */
void orc_aki_richards(double theta, double vsvp, double *row)
{
  double vsvp2 = vsvp * vsvp;
  double sint = sin(theta), sint2 = sint * sint;
  double tan2t = tan(theta) * tan(theta);
  row[0] = (float)((1.0 + tan2t) / 2.0);
  row[1] = (float)(-4 * vsvp2 * sint2);
  row[2] = (float)((1.0 - 4.0 * vsvp2 * sint2) / 2.0);
}

/* vario.cpp:35-44 (findDist) and 66-74 (SphericalVario::corr) */
float orc_spherical_corr(float dx, float dy, float range1, float range2, float angle)
{
  float txx = (float)(cos(angle) * cos(angle) / (range1 * range1) + sin(angle) * sin(angle) / (range2 * range2));
  float tyy = (float)(sin(angle) * sin(angle) / (range1 * range1) + cos(angle) * cos(angle) / (range2 * range2));
  float txy = (float)(2 * cos(angle) * sin(angle) * (1.0 / (range1 * range1) - 1.0 / (range2 * range2)));
  float dist = (float)sqrt(txx * dx * dx + tyy * dy * dy + txy * dx * dy);
  if (dist >= 1.0) return 0.0f;
  return (float)(1.0 + dist * dist * dist / 2.0 - 1.5 * dist);
}

/* commondata.cpp:9180-9218: (*grid)(j*nx+i) = corr(refi*dx, refj*dy) with wrap-around lags */
void orc_corr_xy(int nxp, int nyp, float dx, float dy, float range1, float range2, float angle, float *out)
{
  for (int j = 0; j < nyp; j++)
    for (int i = 0; i < nxp; i++) {
      int refi = (i < nxp / 2 + 1) ? i : i - nxp;
      int refj = (j < nyp / 2 + 1) ? j : j - nyp;
      out[(long)j * nxp + i] = orc_spherical_corr(refi * dx, refj * dy, range1, range2, angle);
    }
}

/* seismicparametersholder.cpp:624-656 (computeCircCorrT) and 805-837 (makeCircCorrTPosDef):
 * mirror the correlation onto the circle, take |FFT| with frequencies <= min_int_fq zeroed,
 * transform back (unnormalised) and rescale so that the value at lag 0 is 1. */
int orc_circ_corr_t(const double *corr_t, int n, int min_int_fq, int nzp, float *out)
{
  double *c = (double *)calloc(2 * nzp, sizeof(double));
  for (int k = 0; k < nzp; k++) {
    int refk = (k < nzp / 2 + 1) ? k : nzp - k;
    c[2 * k] = (refk < n) ? (float)corr_t[refk] : 0.0;
  }
  orc_cfft(nzp, -1, c, 1, 1, nzp);
  for (int k = 0; k < nzp; k++) {
    int refk = (k < nzp / 2 + 1) ? k : nzp - k;
    double a = (refk <= min_int_fq) ? 0.0 : sqrt(c[2 * refk] * c[2 * refk] + c[2 * refk + 1] * c[2 * refk + 1]);
    c[2 * k] = a;
    c[2 * k + 1] = 0.0;
  }
  orc_cfft(nzp, +1, c, 1, 1, nzp);
  int err = 0;
  double scale = 0.0;
  if (c[0] > 1.e-5f)
    scale = (float)(1.0f / c[0]);
  else
    err = 1;
  for (int k = 0; k < nzp; k++) out[k] = (float)(c[2 * k] * (float)scale);
  free(c);
  return err;
}

/* real DFT of a length-n real sequence, bin k (full-spectrum value), sign -1, unnormalised */
static void dft_bin(const double *x, int n, int k, double *re, double *im)
{
  double sr = 0.0, si = 0.0;
  for (int t = 0; t < n; t++) {
    long e = ((long)k * t) % n;
    double a = 2.0 * M_PI * (double)e / (double)n;
    sr += x[t] * cos(a);
    si -= x[t] * sin(a);
  }
  *re = sr;
  *im = si;
}

/* wavelet.cpp:882-900 findNorm on a real array: sqrt(sum of float squares) in double */
static float find_norm(const float *r, int n)
{
  double norm2 = 0.0;
  for (int i = 0; i < n; i++) norm2 += (double)(r[i] * r[i]);
  return (float)sqrt(norm2);
}

void orc_kz_tables(int nz, int nzp, double dt_ms, int nt, const float *wav, const float *A,
                   float low_cut, float high_cut, float min_rel_thick,
                   float *K, float *e1, float *e2, float *e3, int *invert)
{
  double *w = (double *)malloc(sizeof(double) * nzp);
  double *s = (double *)malloc(sizeof(double) * nzp);
  float *fd = (float *)malloc(sizeof(float) * nzp);
  float *bd = (float *)malloc(sizeof(float) * nzp);
  float *cd = (float *)malloc(sizeof(float) * nzp);
  float *wn = (float *)malloc(sizeof(float) * nt);
  float *sn = (float *)malloc(sizeof(float) * nt);
  double *diff = (double *)calloc(nzp, sizeof(double));
  diff[0] = -1.0; /* wavelet1D.cpp:411-415, FIRSTORDERFORWARDDIFF */
  diff[nzp - 1] = 1.0;
  double lz = nz * dt_ms;
  for (int l = 0; l < nt; l++) {
    const float *wl = wav + (long)l * nzp;
    wn[l] = find_norm(wl, nzp);
    /* errorSmooth3 = central(backward(forward(w))), wavelet1D.cpp:441-500 */
    for (int i = 0; i < nzp; i++) fd[i] = (i == nzp - 1) ? wl[0] - wl[i] : wl[i + 1] - wl[i];
    for (int i = 0; i < nzp; i++) bd[i] = (i == 0) ? fd[0] - fd[nzp - 1] : fd[i] - fd[i - 1];
    for (int i = 0; i < nzp; i++) {
      if (i == 0) cd[i] = (float)(0.5 * (bd[i + 1] - bd[nzp - 1]));
      else if (i == nzp - 1) cd[i] = (float)(0.5 * (bd[0] - bd[i - 1]));
      else cd[i] = (float)(0.5 * (bd[i + 1] - bd[i - 1]));
    }
    sn[l] = find_norm(cd, nzp);
  }
  for (int k = 0; k < nzp; k++) {
    float rf = (float)((nz * 1000.0f) / (lz * nzp) * (k < nzp - k ? k : nzp - k));
    invert[k] = (rf > low_cut * (double)min_rel_thick) && (rf < high_cut);
    double kdr, kdi;
    dft_bin(diff, nzp, k, &kdr, &kdi);
    float kDr = (float)kdr, kDi = (float)kdi;
    for (int l = 0; l < nt; l++) {
      const float *wl = wav + (long)l * nzp;
      for (int i = 0; i < nzp; i++) w[i] = wl[i];
      double wr, wi;
      dft_bin(w, nzp, k, &wr, &wi);
      /* fillkW: conj(W); lib_matrProdScalVecCpx(kD, kW) */
      float kwr = (float)wr, kwi = -(float)wi;
      float pr = kDr * kwr - kDi * kwi;
      float pi = kDr * kwi + kDi * kwr;
      for (int j = 0; j < 3; j++) { /* lib_matrProdDiagCpxR */
        K[2 * ((k * nt + l) * 3 + j)] = pr * A[l * 3 + j];
        K[2 * ((k * nt + l) * 3 + j) + 1] = pi * A[l * 3 + j];
      }
      /* fillkWNorm(wavelet) */
      e1[2 * (k * nt + l)] = (float)wr / wn[l];
      e1[2 * (k * nt + l) + 1] = -((float)wi / wn[l]);
      /* fillkWNorm(errorSmooth3) */
      for (int i = 0; i < nzp; i++) fd[i] = wl[i];
      for (int i = 0; i < nzp; i++) bd[i] = (i == nzp - 1) ? fd[0] - fd[i] : fd[i + 1] - fd[i];
      for (int i = 0; i < nzp; i++) fd[i] = (i == 0) ? bd[0] - bd[nzp - 1] : bd[i] - bd[i - 1];
      for (int i = 0; i < nzp; i++) {
        if (i == 0) cd[i] = (float)(0.5 * (fd[i + 1] - fd[nzp - 1]));
        else if (i == nzp - 1) cd[i] = (float)(0.5 * (fd[0] - fd[i - 1]));
        else cd[i] = (float)(0.5 * (fd[i + 1] - fd[i - 1]));
      }
      for (int i = 0; i < nzp; i++) s[i] = cd[i];
      double sr, si;
      dft_bin(s, nzp, k, &sr, &si);
      e2[2 * (k * nt + l)] = (float)sr / sn[l];
      e2[2 * (k * nt + l) + 1] = -((float)si / sn[l]);
      e3[2 * (k * nt + l)] = 1.0f; /* lib_matrFillOnesVecCpx */
      e3[2 * (k * nt + l) + 1] = 0.0f;
    }
  }
  free(w); free(s); free(fd); free(bd); free(cd); free(wn); free(sn); free(diff);
}

/* ModelAVODynamic::ComputeDataVariance (modelavodynamic.cpp:575-620) for one stack, in the
 * reference's order: the sequential sweep k, j, i over the padded real grid rv (rnxp floats per
 * row), float products of the non-zero box samples summed in double per row, rows per plane,
 * planes in total; returns float(total / count), 0 when every sample is zero. */
float orc_data_variance(int nx, int ny, int nz, int nxp, int nyp, int nzp, const float *rv, long *count)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  double totvar = 0;
  long ndata = 0;
  for (int k = 0; k < nzp; k++) {
    double tmpvar1 = 0;
    for (int j = 0; j < nyp; j++) {
      double tmpvar2 = 0;
      for (int i = 0; i < rnxp; i++) {
        float tmp = rv[(long)rnxp * (j + (long)nyp * k) + i];
        if (k < nz && j < ny && i < nx && tmp != 0.0) {
          tmpvar2 += (double)(tmp * tmp);
          ndata++;
        }
      }
      tmpvar1 += tmpvar2;
    }
    totvar += tmpvar1;
  }
  if (count) *count = ndata;
  return ndata == 0 ? 0.0f : (float)(totvar / (double)ndata);
}
