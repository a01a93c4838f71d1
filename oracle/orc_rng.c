/*
 * orc_rng.c -- random numbers and complex noise.  Test infrastructure only; see oracle.h.
 *
 * Reference stream: libs/lib/random.{h,cpp} -- seed = 69069*seed + 1 (mod 2^32)
 * (random.h:10-13, random.cpp:113-118), u = seed * 2^-32, normals by Marsaglia-Bray
 * (random.cpp:38-111).  The reference keeps the seed in a process-wide static
 * (random.cpp:136); here it is an explicit state word.
 *
 * Counter-based stream: the generator the GPU uses where the reference stream is not
 * requested.  Philox-4x32-10 (Salmon et al., SC'11) keyed by the 64-bit seed, counter =
 * (cell index lo, cell index hi, stream, 0); the four 32-bit outputs become two normals by
 * the Box-Muller transform in double precision, rounded to float.
 */
#include <math.h>
#include <stdint.h>
#include "oracle.h"

double orc_unif01(uint32_t *state)
{
  *state = 69069u * *state + 1u;
  return (double)(*state) * (1.0 / 4294967296.0);
}

/* Ripley p.84, random.cpp:96-111 */
static double mb_g(double x)
{
  const double a = 17.49731196, b = 2.36785163, c = 2.15787544;
  double absx = fabs(x);
  double result = a * exp(-x * x / 2.0);
  if (absx < 1.0)
    result -= 2.0 * b * (3.0 - x * x) + c * (1.5 - absx);
  else if (absx < 1.5)
    result -= b * (3.0 - absx) * (3.0 - absx) + c * (1.5 - absx);
  else
    result -= b * (3.0 - absx) * (3.0 - absx);
  return result;
}

double orc_rnorm01(uint32_t *st)
{
  double u = orc_unif01(st), u1, u2, u3, x;
  if (u < 0.8638) {
    u1 = orc_unif01(st);
    u2 = orc_unif01(st);
    u3 = orc_unif01(st);
    x = 2.0 * (u1 + u2 + u3) - 3.0;
  } else if (u < 0.9745) {
    u1 = orc_unif01(st);
    u2 = orc_unif01(st);
    x = 1.5 * (u1 + u2 - 1.0);
  } else if (u < 0.9973002039) {
    do {
      u1 = orc_unif01(st);
      u2 = orc_unif01(st);
      x = 6.0 * u1 - 3.0;
    } while (0.358 * u2 > mb_g(x));
  } else {
    double s, t;
    do {
      double v1, v2, w;
      do {
        u1 = orc_unif01(st);
        u2 = orc_unif01(st);
        v1 = 2.0 * u1 - 1.0;
        v2 = 2.0 * u2 - 1.0;
        w = v1 * v1 + v2 * v2;
      } while (w >= 1.0);
      double c = sqrt((9.0 - 2.0 * log(w)) / w);
      s = c * v1;
      t = c * v2;
    } while ((fabs(s) <= 3.0) && (fabs(t) <= 3.0));
    x = (fabs(s) > 3.0) ? s : t;
  }
  return x;
}

void orc_rnorm01_n(uint32_t *state, double *out, long n)
{
  for (long i = 0; i < n; i++) out[i] = orc_rnorm01(state);
}

/* fftgrid.cpp:997-1045.  The Hermitian treatment applies on the x = 0 column and, by the
 * reference's test of the parity of the linear index (not of nxp), on the x = cnxp-1 column
 * where the linear index is odd. */
void orc_fill_complex_noise(int nxp, int nyp, int nzp, uint32_t *st, float *cv)
{
  const long cnxp = nxp / 2 + 1;
  const long csize = cnxp * nyp * nzp;
  const float sd = (float)(1 / sqrt(2.0));
  for (long i = 0; i < csize; i++) {
    long xs = i % cnxp;
    if (xs == 0 || (xs == cnxp - 1 && (i % 2) == 1)) {
      long jk = (i - xs) / cnxp;
      long jind = jk % nyp, kind = jk / nyp;
      long jcc = (jind == 0) ? 0 : nyp - jind;
      long kcc = (kind == 0) ? 0 : nzp - kind;
      long jkcc = jcc + kcc * nyp;
      if (jkcc == jk) {
        cv[2 * i] = (float)orc_rnorm01(st);
        cv[2 * i + 1] = 0;
      } else if (jkcc > jk) {
        cv[2 * i] = (float)(sd * orc_rnorm01(st));
        cv[2 * i + 1] = (float)(sd * orc_rnorm01(st));
      } else {
        long cci = jkcc * cnxp + xs;
        cv[2 * i] = cv[2 * cci];
        cv[2 * i + 1] = -cv[2 * cci + 1];
      }
    } else {
      cv[2 * i] = (float)(sd * orc_rnorm01(st));
      cv[2 * i + 1] = (float)(sd * orc_rnorm01(st));
    }
  }
}

/* ---- Philox-4x32-10 ---- */
static void philox(uint32_t ctr[4], const uint32_t key0[2])
{
  uint32_t k0 = key0[0], k1 = key0[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)0xD2511F53u * ctr[0];
    uint64_t p1 = (uint64_t)0xCD9E8D57u * ctr[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ ctr[1] ^ k0;
    uint32_t n1 = lo1;
    uint32_t n2 = hi0 ^ ctr[3] ^ k1;
    uint32_t n3 = lo0;
    ctr[0] = n0; ctr[1] = n1; ctr[2] = n2; ctr[3] = n3;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

/* Two standard normals for (seed, stream, idx): the Philox block of counter
 * (idx, stream >> 1, 0), key seed; Box-Muller on words (w0, w1) = block words 0-1 for an even
 * stream, 2-3 for an odd one: u1 = ((w0 >> 8) + 1) 2^-24 in (0, 1], u2 = (w1 >> 8) 2^-24,
 * evaluated in double here (the device evaluates the same formula in float). */
void orc_cb_normal2(uint64_t seed, uint32_t stream, uint64_t idx, float *z0, float *z1)
{
  uint32_t c[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), stream >> 1, 0u};
  uint32_t k[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  philox(c, k);
  const uint32_t w0 = (stream & 1u) ? c[2] : c[0], w1 = (stream & 1u) ? c[3] : c[1];
  double u1 = (double)((w0 >> 8) + 1u) * (1.0 / 16777216.0);
  double u2 = (double)(w1 >> 8) * (1.0 / 16777216.0);
  double r = sqrt(-2.0 * log(u1));
  double th = 2.0 * M_PI * u2;
  *z0 = (float)(r * cos(th));
  *z1 = (float)(r * sin(th));
}

/* Raw Philox-4x32-10 block (known-answer tests). */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
  uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
  philox(c, key);
  for (int i = 0; i < 4; i++) out[i] = c[i];
}

/* Same Hermitian pattern as the reference fill; value of cell i from its own counter, the
 * conjugate-mirror cells copy their partner's value. */
static void cb_cell(int nxp, int nyp, int nzp, uint64_t seed, uint32_t stream, long i, float *out)
{
  const long cnxp = nxp / 2 + 1;
  const double sd = 1 / sqrt(2.0);
  long xs = i % cnxp;
  float z0, z1;
  if (xs == 0 || (xs == cnxp - 1 && (i % 2) == 1)) {
    long jk = (i - xs) / cnxp;
    long jind = jk % nyp, kind = jk / nyp;
    long jcc = (jind == 0) ? 0 : nyp - jind;
    long kcc = (kind == 0) ? 0 : nzp - kind;
    long jkcc = jcc + kcc * nyp;
    if (jkcc == jk) {
      orc_cb_normal2(seed, stream, (uint64_t)i, &z0, &z1);
      out[0] = z0;
      out[1] = 0;
    } else if (jkcc > jk) {
      orc_cb_normal2(seed, stream, (uint64_t)i, &z0, &z1);
      out[0] = (float)(sd * z0);
      out[1] = (float)(sd * z1);
    } else {
      long cci = jkcc * cnxp + xs;
      orc_cb_normal2(seed, stream, (uint64_t)cci, &z0, &z1);
      out[0] = (float)(sd * z0);
      out[1] = -(float)(sd * z1);
    }
  } else {
    orc_cb_normal2(seed, stream, (uint64_t)i, &z0, &z1);
    out[0] = (float)(sd * z0);
    out[1] = (float)(sd * z1);
  }
}

void orc_fill_complex_noise_cb(int nxp, int nyp, int nzp, uint64_t seed, uint32_t stream, float *cv)
{
  const long csize = (long)(nxp / 2 + 1) * nyp * nzp;
  for (long i = 0; i < csize; i++) cb_cell(nxp, nyp, nzp, seed, stream, i, cv + 2 * i);
}

/* The same noise on the k planes ks[0..nk) only: cv holds nk planes of (nxp/2+1) * nyp cells. */
void orc_fill_complex_noise_cb_planes(int nxp, int nyp, int nzp, uint64_t seed, uint32_t stream, const int *ks,
                                      int nk, float *cv)
{
  const long plane = (long)(nxp / 2 + 1) * nyp;
  for (int p = 0; p < nk; p++)
    for (long c = 0; c < plane; c++)
      cb_cell(nxp, nyp, nzp, seed, stream, (long)ks[p] * plane + c, cv + 2 * ((long)p * plane + c));
}
