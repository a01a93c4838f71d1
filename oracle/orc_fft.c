/*
 * orc_fft.c -- FFT part of the oracle (test infrastructure only; see oracle.h).
 *
 * Restates the transforms CRAVA obtains from the vendored FFTW 2.1.2 (float):
 *   FFTGrid::fftInPlace     src/fftgrid.cpp:1587-1612  (rfftw3d_create_plan(nzp,nyp,nxp,...))
 *   FFTGrid::invFFTInPlace  src/fftgrid.cpp:1615-1645
 *   rfftwnd_real2c_aux / rfftwnd_c2real_aux  libs/fft/rfftw/rfftwnd.c:207-288 (pass order)
 * Conventions: forward sign -1 (libs/fft/include/fftw.h:61), unnormalised; the in-place
 * layout has row pitch rnxp = 2*(nxp/2+1) reals (fftgrid.cpp:54-58); the x axis is the
 * half-spectrum axis.  Evaluated in double with exact-angle twiddles.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

/* smallest 2^a 3^b 5^c 7^d >= leastint.  fftgrid.cpp:2630-2658 enumerates exponents up to
 * ceil(log_p(leastint)) and keeps the smallest product >= leastint, which is this number. */
int orc_factorable(int leastint)
{
  if (leastint <= 1) return 1;
  int best = 1;
  while (best < leastint) best *= 2;
  for (long a = 1; a <= 2L * best; a *= 2)
    for (long b = a; b <= 2L * best; b *= 3)
      for (long c = b; c <= 2L * best; c *= 5)
        for (long d = c; d <= 2L * best; d *= 7)
          if (d >= leastint && d < best) best = (int)d;
  return best;
}

/* multiintervalgrid.cpp:922-928: ceil(n*(1+pad)) then the closest factorable number. */
int orc_padding_size(int n, double pad_fraction)
{
  int leastint = (int)ceil(n * (1.0f + pad_fraction));
  return orc_factorable(leastint);
}

static void twiddle(long n, long e, int sign, double *re, double *im)
{
  e %= n;
  if (e < 0) e += n;
  double a = 2.0 * M_PI * (double)e / (double)n;
  *re = cos(a);
  *im = sign * sin(a);
}

void orc_cdft_naive(int n, int sign, const double *in, double *out)
{
  for (int k = 0; k < n; ++k) {
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < n; ++t) {
      double wr, wi;
      twiddle(n, (long)k * t, sign, &wr, &wi);
      sr += in[2 * t] * wr - in[2 * t + 1] * wi;
      si += in[2 * t] * wi + in[2 * t + 1] * wr;
    }
    out[2 * k] = sr;
    out[2 * k + 1] = si;
  }
}

static int smallest_factor(int n)
{
  if (n % 4 == 0 && n != 4) return 4;
  for (int p = 2; p * p <= n; ++p)
    if (n % p == 0) return p;
  return n;
}

/* decimation in time, recursive; in: strided, out: contiguous (length n). */
static void fft_rec(int n, int sign, const double *in, long istride, double *out, double *tmp)
{
  if (n == 1) {
    out[0] = in[0];
    out[1] = in[1];
    return;
  }
  int p = smallest_factor(n);
  int m = n / p;
  for (int r = 0; r < p; ++r)
    fft_rec(m, sign, in + 2 * r * istride, istride * p, out + 2 * r * m, tmp);
  for (int k = 0; k < m; ++k) {
    for (int r = 0; r < p; ++r) {
      double wr, wi;
      twiddle(n, (long)r * k, sign, &wr, &wi);
      double xr = out[2 * (r * m + k)], xi = out[2 * (r * m + k) + 1];
      tmp[2 * r] = xr * wr - xi * wi;
      tmp[2 * r + 1] = xr * wi + xi * wr;
    }
    for (int q = 0; q < p; ++q) {
      double sr = 0.0, si = 0.0;
      for (int r = 0; r < p; ++r) {
        double wr, wi;
        twiddle(p, (long)r * q, sign, &wr, &wi);
        sr += tmp[2 * r] * wr - tmp[2 * r + 1] * wi;
        si += tmp[2 * r] * wi + tmp[2 * r + 1] * wr;
      }
      out[2 * (q * m + k)] = sr;
      out[2 * (q * m + k) + 1] = si;
    }
  }
}

void orc_cfft(int n, int sign, double *data, long stride, long howmany, long dist)
{
  double *line = (double *)malloc(sizeof(double) * 2 * n);
  double *res = (double *)malloc(sizeof(double) * 2 * n);
  double *tmp = (double *)malloc(sizeof(double) * 2 * (n + 8));
  for (long h = 0; h < howmany; ++h) {
    double *base = data + 2 * h * dist;
    for (int e = 0; e < n; ++e) {
      line[2 * e] = base[2 * e * stride];
      line[2 * e + 1] = base[2 * e * stride + 1];
    }
    fft_rec(n, sign, line, 1, res, tmp);
    for (int e = 0; e < n; ++e) {
      base[2 * e * stride] = res[2 * e];
      base[2 * e * stride + 1] = res[2 * e + 1];
    }
  }
  free(line);
  free(res);
  free(tmp);
}

void orc_rfft3d_r2c(int nxp, int nyp, int nzp, double *buf)
{
  const int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  double *line = (double *)malloc(sizeof(double) * 2 * nxp);
  double *res = (double *)malloc(sizeof(double) * 2 * nxp);
  double *tmp = (double *)malloc(sizeof(double) * 2 * (nxp + 8));
  /* x: real rows -> half spectra (rfftw_real2c_aux) */
  for (long row = 0; row < (long)nyp * nzp; ++row) {
    double *r = buf + row * rnxp;
    for (int i = 0; i < nxp; ++i) {
      line[2 * i] = r[i];
      line[2 * i + 1] = 0.0;
    }
    fft_rec(nxp, -1, line, 1, res, tmp);
    memcpy(r, res, sizeof(double) * 2 * cnxp);
  }
  free(line);
  free(res);
  free(tmp);
  /* y within each z plane, then z */
  for (int k = 0; k < nzp; ++k)
    orc_cfft(nyp, -1, buf + 2L * cnxp * nyp * k, cnxp, cnxp, 1);
  orc_cfft(nzp, -1, buf, (long)cnxp * nyp, (long)cnxp * nyp, 1);
}

void orc_rfft3d_c2r(int nxp, int nyp, int nzp, double *buf)
{
  const int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  orc_cfft(nzp, +1, buf, (long)cnxp * nyp, (long)cnxp * nyp, 1);
  for (int k = 0; k < nzp; ++k)
    orc_cfft(nyp, +1, buf + 2L * cnxp * nyp * k, cnxp, cnxp, 1);
  double *line = (double *)malloc(sizeof(double) * 2 * nxp);
  double *res = (double *)malloc(sizeof(double) * 2 * nxp);
  double *tmp = (double *)malloc(sizeof(double) * 2 * (nxp + 8));
  for (long row = 0; row < (long)nyp * nzp; ++row) {
    double *r = buf + row * rnxp;
    /* Hermitian extension; FFTW's hc2r reads only Re of the DC and (even n) Nyquist bins */
    line[0] = r[0];
    line[1] = 0.0;
    for (int kx = 1; kx < cnxp; ++kx) {
      line[2 * kx] = r[2 * kx];
      line[2 * kx + 1] = r[2 * kx + 1];
      line[2 * (nxp - kx)] = r[2 * kx];
      line[2 * (nxp - kx) + 1] = -r[2 * kx + 1];
    }
    if (nxp % 2 == 0) {
      line[nxp] = r[nxp];
      line[nxp + 1] = 0.0;
    }
    fft_rec(nxp, +1, line, 1, res, tmp);
    for (int i = 0; i < nxp; ++i) r[i] = res[2 * i];
    for (int i = nxp; i < rnxp; ++i) r[i] = 0.0;
  }
  free(line);
  free(res);
  free(tmp);
}

static void scale_real(double *buf, long n, double s)
{
  for (long i = 0; i < n; ++i) buf[i] *= s;
}

/* fftInPlace: DATA/PARAMETER are multiplied by 1/sqrt(N) (computed in float) first. */
void orc_grid_fft(int nxp, int nyp, int nzp, int cubetype, double *buf)
{
  long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  if (cubetype != ORC_COVARIANCE)
    scale_real(buf, rsize, 1.0f / sqrtf((float)((long)nxp * nyp * nzp)));
  orc_rfft3d_r2c(nxp, nyp, nzp, buf);
}

/* invFFTInPlace: x 1/N for COVARIANCE, x 1/sqrt(N) otherwise, after the transform. */
void orc_grid_ifft(int nxp, int nyp, int nzp, int cubetype, double *buf)
{
  long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  double n = (double)((long)nxp * nyp * nzp);
  /* fftgrid.cpp:1628-1631: float(1.0/N) or float(1.0/sqrt(float(N))) (float sqrt overload) */
  float s = (cubetype == ORC_COVARIANCE) ? (float)(1.0 / n) : (float)(1.0 / (double)sqrtf((float)n));
  orc_rfft3d_c2r(nxp, nyp, nzp, buf);
  scale_real(buf, rsize, s);
}

void orc_grid_fft_f32(int nxp, int nyp, int nzp, int cubetype, float *buf)
{
  long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  double *d = (double *)malloc(sizeof(double) * rsize);
  for (long i = 0; i < rsize; ++i) d[i] = buf[i];
  orc_grid_fft(nxp, nyp, nzp, cubetype, d);
  for (long i = 0; i < rsize; ++i) buf[i] = (float)d[i];
  free(d);
}

void orc_grid_ifft_f32(int nxp, int nyp, int nzp, int cubetype, float *buf)
{
  long rsize = 2L * (nxp / 2 + 1) * nyp * nzp;
  double *d = (double *)malloc(sizeof(double) * rsize);
  for (long i = 0; i < rsize; ++i) d[i] = buf[i];
  orc_grid_ifft(nxp, nyp, nzp, cubetype, d);
  for (long i = 0; i < rsize; ++i) buf[i] = (float)d[i];
  free(d);
}
