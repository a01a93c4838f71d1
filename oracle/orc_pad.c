/*
 * orc_pad.c -- padding fills on the input side of the path (test infrastructure only).
 *
 * Sequential restatements, in the reference's loop order, of how FFTGrid fills the grids that
 * enter the path: the background model's linear edge interpolation, the seismic data's trace
 * interpolation and tapered extrapolation, and the prior and error correlation grids.  Float
 * arithmetic as the reference's x86-64 build does it (no FMA; the Makefile passes
 * -ffp-contract=off).
 */
#include <math.h>
#include <stdlib.h>

#include "oracle.h"

#define ORC_RMISSING -99999.0f
#define ORC_IMISSING -99999

static size_t ridx(int rnxp, int nyp, int i, int j, int k)
{
  return (size_t)i + (size_t)rnxp * (size_t)j + (size_t)rnxp * (size_t)nyp * (size_t)k;
}

/* FFTGrid::FFTGrid(const NRLib::Grid<float> *, nxp, nyp, nzp), fftgrid.cpp:125-213: the cube
 * (i fastest, NRLib Grid::GetIndex, grid.hpp:231-236) is copied and the padding filled in
 * loop order k, j, i; every padding value is a two-point blend of values written earlier. */
void orc_fill_padded(int nx, int ny, int nz, int nxp, int nyp, int nzp, const float *cube, float *rv)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  const float sum_c = 1.0f / (float)(nzp - nz + 1);
  const float sum_b = 1.0f / (float)(nyp - ny + 1);
  const float sum_a = 1.0f / (float)(nxp - nx + 1);
  for (int k = 0; k < nzp; ++k)
    for (int j = 0; j < nyp; ++j)
      for (int i = 0; i < rnxp; ++i) {
        float v;
        if (i < nx && j < ny && k < nz) {
          v = cube[(size_t)i + (size_t)nx * j + (size_t)nx * ny * k];
        } else if (i >= nxp) {
          v = 0.0f;
        } else if (k >= nz) {
          const float lo = rv[ridx(rnxp, nyp, i, j, 0)], hi = rv[ridx(rnxp, nyp, i, j, nz - 1)];
          v = lo * (sum_c * (float)(k - nz + 1)) + hi * (sum_c * (float)(nzp - k));
        } else if (j >= ny) {
          const float lo = rv[ridx(rnxp, nyp, i, 0, k)], hi = rv[ridx(rnxp, nyp, i, ny - 1, k)];
          v = lo * (sum_b * (float)(j - ny + 1)) + hi * (sum_b * (float)(nyp - j));
        } else {
          const float lo = rv[ridx(rnxp, nyp, 0, j, k)], hi = rv[ridx(rnxp, nyp, nx - 1, j, k)];
          v = lo * (sum_a * (float)(i - nx + 1)) + hi * (sum_a * (float)(nxp - i));
        }
        rv[ridx(rnxp, nyp, i, j, k)] = v;
      }
}

/* FFTGrid::getFillNumber, fftgrid.cpp:1103-1140: the in-box index whose trace a padding index
 * copies (mirror towards the nearer end, clamped at n/2). */
int orc_fill_number(int i, int n, int np)
{
  if (i >= np) return ORC_IMISSING;
  if (i < n) return i;
  const int to_np = np - i - 1, past_n = i - n + 1;
  if (past_n < to_np) return (n - past_n) > n / 2 ? (n - past_n) : n / 2;
  return to_np < n / 2 ? to_np : n / 2;
}

/* FFTGrid::getDistToBoundary, fftgrid.cpp:1157-1200: distance into the padding in units of
 * the taper length min(n, np - n) / 2.1. */
float orc_dist_to_boundary(int i, int n, int np)
{
  if (i >= np) return ORC_RMISSING;
  if (i < n) return 0.0f;
  const int m = n < np - n ? n : np - n;
  const float taper = (float)(m / 2.1);
  const int to_np = np - i, past_n = i - (n - 1);
  return past_n < to_np ? (float)past_n / taper : (float)to_np / taper;
}

/* FFTGrid::extrapolateSeismic, fftgrid.cpp:2484-2518: traces outside the rectangle of good
 * traces copy the clamped fill-number trace, tapered by max(1 - dx^2 - dy^2 - dz^2, 0)^3. */
void orc_extrapolate_seismic(int nx, int ny, int nz, int nxp, int nyp, int nzp, float *rv, int imin, int imax,
                             int jmin, int jmax)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  for (int j = 0; j < nyp; ++j) {
    int rj = orc_fill_number(j, ny, nyp);
    rj = rj < jmin ? jmin : (rj > jmax ? jmax : rj);
    for (int i = 0; i < nxp; ++i) {
      if (!(i < imin || i > imax || j < jmin || j > jmax)) continue;
      int ri = orc_fill_number(i, nx, nxp);
      ri = ri < imin ? imin : (ri > imax ? imax : ri);
      for (int k = 0; k < nzp; ++k) {
        const float v = rv[ridx(rnxp, nyp, ri, rj, k)];
        const float dx = orc_dist_to_boundary(i, nx, nxp), dy = orc_dist_to_boundary(j, ny, nyp);
        const float dz = orc_dist_to_boundary(k, nz, nzp);
        double t = 1.0 - dx * dx - dy * dy - dz * dz;
        if (t < 0.0) t = 0.0;
        const float mult = (float)pow(t, 3);
        rv[ridx(rnxp, nyp, i, j, k)] = mult * v;
      }
    }
  }
}

/* FFTGrid::interpolateTrace, fftgrid.cpp:2440-2481: a bad trace (odd flag) with at least two
 * good (even-flag) 4-neighbours becomes their mean over the whole padded column. */
static int interp_trace(int nx, int ny, int nzp, int rnxp, int nyp, float *rv, short *flags, int idx, int i,
                        int j)
{
  const int l = i > 0 && flags[idx - 1] % 2 == 0, r = i < nx - 1 && flags[idx + 1] % 2 == 0;
  const int u = j > 0 && flags[idx - nx] % 2 == 0, d = j < ny - 1 && flags[idx + nx] % 2 == 0;
  const int nt = l + r + u + d;
  if (nt > 1) {
    for (int k = 0; k < nzp; ++k) {
      float s = l ? rv[ridx(rnxp, nyp, i - 1, j, k)] : 0.0f;
      if (r) s += rv[ridx(rnxp, nyp, i + 1, j, k)];
      if (u) s += rv[ridx(rnxp, nyp, i, j - 1, k)];
      if (d) s += rv[ridx(rnxp, nyp, i, j + 1, k)];
      rv[ridx(rnxp, nyp, i, j, k)] = s / (float)nt;
    }
    flags[idx]--;
  }
  return flags[idx];
}

/* FFTGrid::interpolateSeismic, fftgrid.cpp:2349-2437: traces whose energy over the nz box
 * samples is at most threshold * (total energy) / (nx ny) are interpolated from neighbours
 * (flag bit 0: bad, bit 1: a bad trace came earlier; a success re-walks the pending earlier
 * traces), then the padding is extrapolated from the rectangle of good traces.  Returns the
 * number of bad traces; *imin..*jmax receive the rectangle. */
int orc_interpolate_seismic(int nx, int ny, int nz, int nxp, int nyp, int nzp, float *rv, float threshold,
                            int *rect)
{
  const int rnxp = 2 * (nxp / 2 + 1), ntr = nx * ny;
  short *flags = (short *)malloc(sizeof(short) * (size_t)ntr);
  float *energy = (float *)malloc(sizeof(float) * (size_t)ntr);
  float total = 0.0f;
  for (int j = 0, t = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i, ++t) {
      float e = 0.0f;
      for (int k = 0; k < nz; ++k) {
        const float v = rv[ridx(rnxp, nyp, i, j, k)];
        e += v * v;
      }
      energy[t] = e;
      total += e;
    }
  const float limit = threshold * total / (float)ntr;
  int imin = nx, imax = 0, jmin = ny, jmax = 0, nbad = 0, seen_bad = 0;
  for (int j = 0, t = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i, ++t) {
      const int bad = energy[t] <= limit;
      nbad += bad;
      flags[t] = (short)(seen_bad + bad);
      if (bad) {
        seen_bad = 2;
      } else {
        if (i < imin) imin = i;
        if (i > imax) imax = i;
        if (j < jmin) jmin = j;
        if (j > jmax) jmax = j;
      }
    }
  for (int cur = 0; cur < ntr; ++cur) {
    if (flags[cur] % 2 != 1) continue;
    if (interp_trace(nx, ny, nzp, rnxp, nyp, rv, flags, cur, cur % nx, cur / nx) % 2 != 0) continue;
    int t = cur - 1;
    for (; t >= 0 && flags[t] > 1; --t)
      if (flags[t] % 2 == 1) interp_trace(nx, ny, nzp, rnxp, nyp, rv, flags, t, t % nx, t / nx);
    if (t < 0) t = 0;
    for (++t; flags[t - 1] == 0 && t <= cur; ++t)
      if (flags[t] > 1) flags[t] -= 2;
  }
  orc_extrapolate_seismic(nx, ny, nz, nxp, nyp, nzp, rv, imin, imax, jmin, jmax);
  free(flags);
  free(energy);
  if (rect) {
    rect[0] = imin;
    rect[1] = imax;
    rect[2] = jmin;
    rect[3] = jmax;
  }
  return nbad;
}

/* Signed cyclic offset of index i on a lattice of n (FFTGrid's cycleI / cycleJ). */
static int cyc(int i, int n)
{
  int c = i - n;
  if (i < -c) c = i;
  return c;
}

/* FFTGrid::fillInParamCorr / FillInLateralCorr (fftgrid.cpp:830-923), followed by
 * multiplyByScalar(scale) (seismicparametersholder.cpp:336-348 scales by priorVar0 after the
 * fill): the circular temporal correlation interpolated at k + cycleI gradI + cycleJ gradJ
 * (cyclically) times the lateral correlation corr_xy[i + nxp j]; RMISSING on the row padding. */
void orc_fill_param_corr(int nxp, int nyp, int nzp, const float *corr_xy, const float *corr_t, float grad_i,
                         float grad_j, float scale, float *rv)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  for (int k = 0; k < nzp; ++k)
    for (int j = 0; j < nyp; ++j)
      for (int i = 0; i < rnxp; ++i) {
        float v;
        if (i < nxp) {
          float sub = k + cyc(i, nxp) * grad_i + cyc(j, nyp) * grad_j;
          int base = (int)floorf(sub);
          sub -= base;
          while (base < 0) base += nzp;
          while (base >= nzp) base -= nzp;
          v = (1 - sub) * corr_t[base];
          v += sub * corr_t[base != nzp - 1 ? base + 1 : 0];
          v *= corr_xy[i + nxp * j];
        } else {
          v = ORC_RMISSING;
        }
        rv[ridx(rnxp, nyp, i, j, k)] = v * scale;
      }
}

/* FFTGrid::fillInErrCorr (fftgrid.cpp:777-828): the lateral correlation, weighted by the
 * distance of k + cycleI gradI + cycleJ gradJ from the k = 0 plane (cyclically, within one
 * sample), zero elsewhere; RMISSING on the row padding. */
void orc_fill_err_corr(int nxp, int nyp, int nzp, const float *corr_xy, float grad_i, float grad_j, float *rv)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  for (int k = 0; k < nzp; ++k)
    for (int j = 0; j < nyp; ++j)
      for (int i = 0; i < rnxp; ++i) {
        float v;
        if (i < nxp) {
          float sub = k + cyc(i, nxp) * grad_i + cyc(j, nyp) * grad_j;
          if (fabsf(sub) < 1.0f || fabsf(sub - nzp) < 1.0f) {
            int base = (int)sub;
            sub -= base;
            while (base < -1) base += nzp;
            while (base >= 1) base -= nzp;
            if (base < 0) sub = 1 - sub;
            v = (1 - sub) * corr_xy[i + nxp * j];
          } else {
            v = 0.0f;
          }
        } else {
          v = ORC_RMISSING;
        }
        rv[ridx(rnxp, nyp, i, j, k)] = v;
      }
}
