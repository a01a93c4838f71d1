// cg_dft.h -- complex helpers and the DFT butterflies of the LDS FFT engines: radix 2/3/4/5/7/8,
// and composite sizes up to 32 computed as two smaller DFTs in registers.
// X[q] = sum_r v[r] w^(r q), w = exp(DIR 2 pi i / p); DIR = -1 forward (FFTW sign), +1 inverse.
#pragma once

#include <hip/hip_runtime.h>

namespace cgdft {

__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// multiply by DIR * i
template <int DIR>
__device__ __forceinline__ float2 mul_di(float2 a)
{
  return DIR > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// ---------------------------------------------------------------- small DFTs
// X[q] = sum_r v[r] w^(r q), w = exp(DIR 2 pi i / p)
template <int DIR>
__device__ __forceinline__ void dft2(float2 *v)
{
  float2 a = v[0], b = v[1];
  v[0] = cadd(a, b);
  v[1] = csub(a, b);
}

template <int DIR>
__device__ __forceinline__ void dft3(float2 *v)
{
  const float s = 0.86602540378443864676f;
  float2 t1 = cadd(v[1], v[2]);
  float2 t2 = make_float2(v[0].x - 0.5f * t1.x, v[0].y - 0.5f * t1.y);
  float2 t3 = mul_di<DIR>(cscale(csub(v[1], v[2]), s));
  v[0] = cadd(v[0], t1);
  v[1] = cadd(t2, t3);
  v[2] = csub(t2, t3);
}

template <int DIR>
__device__ __forceinline__ void dft4(float2 *v)
{
  float2 a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
  float2 c = cadd(v[1], v[3]), d = mul_di<DIR>(csub(v[1], v[3]));
  v[0] = cadd(a, c);
  v[2] = csub(a, c);
  v[1] = cadd(b, d);
  v[3] = csub(b, d);
}

template <int DIR>
__device__ __forceinline__ void dft5(float2 *v)
{
  const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
  const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
  float2 a1 = cadd(v[1], v[4]), b1 = csub(v[1], v[4]);
  float2 a2 = cadd(v[2], v[3]), b2 = csub(v[2], v[3]);
  float2 x0 = v[0];
  float2 r1 = make_float2(x0.x + c1 * a1.x + c2 * a2.x, x0.y + c1 * a1.y + c2 * a2.y);
  float2 r2 = make_float2(x0.x + c2 * a1.x + c1 * a2.x, x0.y + c2 * a1.y + c1 * a2.y);
  float2 i1 = mul_di<DIR>(make_float2(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
  float2 i2 = mul_di<DIR>(make_float2(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
  v[0] = cadd(x0, cadd(a1, a2));
  v[1] = cadd(r1, i1);
  v[4] = csub(r1, i1);
  v[2] = cadd(r2, i2);
  v[3] = csub(r2, i2);
}

template <int DIR>
__device__ __forceinline__ void dft7(float2 *v)
{
  // cos(2 pi m / 7), sin(2 pi m / 7), m = 1..3
  const float C[3] = {0.62348980185873353053f, -0.22252093395631440429f, -0.90096886790241912624f};
  const float S[3] = {0.78183148246802980871f, 0.97492791218182360702f, 0.43388373911755812048f};
  float2 a[3], b[3];
#pragma unroll
  for (int m = 0; m < 3; ++m) {
    a[m] = cadd(v[m + 1], v[6 - m]);
    b[m] = csub(v[m + 1], v[6 - m]);
  }
  float2 x0 = v[0];
  v[0] = cadd(x0, cadd(a[0], cadd(a[1], a[2])));
#pragma unroll
  for (int q = 1; q <= 3; ++q) {
    float2 re = x0, im = make_float2(0.f, 0.f);
#pragma unroll
    for (int m = 1; m <= 3; ++m) {
      const int e = (m * q) % 7;  // exponent
      // cos/sin(2 pi e / 7) from the m = 1..3 tables with symmetry
      const float c = (e <= 3) ? C[e - 1] : C[7 - e - 1];
      const float s = (e <= 3) ? S[e - 1] : -S[7 - e - 1];
      re = make_float2(re.x + c * a[m - 1].x, re.y + c * a[m - 1].y);
      im = make_float2(im.x + s * b[m - 1].x, im.y + s * b[m - 1].y);
    }
    float2 t = mul_di<DIR>(im);
    v[q] = cadd(re, t);
    v[7 - q] = csub(re, t);
  }
}

template <int DIR>
__device__ __forceinline__ void dft8(float2 *v)
{
  const float h = 0.70710678118654752440f;
  float2 e[4] = {v[0], v[2], v[4], v[6]};
  float2 o[4] = {v[1], v[3], v[5], v[7]};
  dft4<DIR>(e);
  dft4<DIR>(o);
  // w^1 = (1 + DIR i)/sqrt2, w^2 = DIR i, w^3 = (-1 + DIR i)/sqrt2
  float2 o1 = cscale(cadd(o[1], mul_di<DIR>(o[1])), h);
  float2 o2 = mul_di<DIR>(o[2]);
  float2 o3 = cscale(csub(mul_di<DIR>(o[3]), o[3]), h);
  v[0] = cadd(e[0], o[0]);
  v[4] = csub(e[0], o[0]);
  v[1] = cadd(e[1], o1);
  v[5] = csub(e[1], o1);
  v[2] = cadd(e[2], o2);
  v[6] = csub(e[2], o2);
  v[3] = cadd(e[3], o3);
  v[7] = csub(e[3], o3);
}

// ---------------------------------------------------------------- composite sizes
// P = A B (A the largest base radix dividing P with B = P / A > 1), in registers:
//   X[k1 + A k2] = sum_n2 w_B^(n2 k2) [ w_P^(n2 k1) sum_n1 v[B n1 + n2] w_A^(n1 k1) ]
// with compile-time twiddles w_P^m (constexpr tables in double, rounded to float).
constexpr double kTwoPi = 6.283185307179586476925286766559;

constexpr double cx_sin_taylor(double x)  // |x| <= pi / 2: 14 terms reach double precision
{
  double term = x, sum = x;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cx_cos_taylor(double x)
{
  double term = 1.0, sum = 1.0;
  for (int k = 1; k < 14; ++k) {
    term *= -x * x / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
// cos and sin of 2 pi m / p: quadrant q = floor(4 m / p) and residual angle
// (pi / 2) (4 m - q p) / p in [0, pi / 2), both exact in integers.
constexpr void cx_sincos(long m, long p, double &c, double &s)
{
  m %= p;
  if (m < 0) m += p;
  const long q = (4 * m) / p;
  const double phi = (kTwoPi / 4.0) * (double)(4 * m - q * p) / (double)p;
  const double cp = cx_cos_taylor(phi), sp = cx_sin_taylor(phi);
  switch (q) {
    case 0: c = cp; s = sp; break;
    case 1: c = -sp; s = cp; break;
    case 2: c = -cp; s = -sp; break;
    default: c = sp; s = -cp; break;
  }
}

template <int P>
struct TwTable {
  float c[P];
  float s[P];
};
template <int P>
constexpr TwTable<P> make_tw()
{
  TwTable<P> t{};
  for (int m = 0; m < P; ++m) {
    double c = 0, s = 0;
    cx_sincos(m, P, c, s);
    t.c[m] = (float)c;
    t.s[m] = (float)s;
  }
  return t;
}
template <int P>
struct Tw {
  static constexpr TwTable<P> t = make_tw<P>();
};

constexpr int base_radix(int p)
{
  return p % 8 == 0 && p > 8 ? 8 : p % 7 == 0 && p > 7 ? 7 : p % 5 == 0 && p > 5 ? 5 : p % 4 == 0 && p > 4 ? 4
         : p % 3 == 0 && p > 3 ? 3 : p % 2 == 0 && p > 2 ? 2 : p;
}

template <int DIR, int P>
__device__ __forceinline__ void dft(float2 *v);

template <int DIR, int P>
__device__ __forceinline__ void dft_composite(float2 *v)
{
  constexpr int A = base_radix(P);
  constexpr int B = P / A;
  float2 y[P];
#pragma unroll
  for (int n2 = 0; n2 < B; ++n2) {
    float2 t[A];
#pragma unroll
    for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
    dft<DIR, A>(t);
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
      const int m = (n2 * k1) % P;
      float2 w = make_float2(Tw<P>::t.c[m], Tw<P>::t.s[m]);  // exp(+i 2 pi m / P)
      if (DIR < 0) w.y = -w.y;
      y[n2 * A + k1] = (n2 == 0 || k1 == 0) ? t[k1] : cmul(t[k1], w);
    }
  }
#pragma unroll
  for (int k1 = 0; k1 < A; ++k1) {
    float2 t[B];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) t[n2] = y[n2 * A + k1];
    dft<DIR, B>(t);
#pragma unroll
    for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = t[k2];
  }
}

template <int DIR, int P>
__device__ __forceinline__ void dft(float2 *v)
{
  if constexpr (P == 1) {
  } else if constexpr (P == 2) {
    dft2<DIR>(v);
  } else if constexpr (P == 3) {
    dft3<DIR>(v);
  } else if constexpr (P == 4) {
    dft4<DIR>(v);
  } else if constexpr (P == 5) {
    dft5<DIR>(v);
  } else if constexpr (P == 7) {
    dft7<DIR>(v);
  } else if constexpr (P == 8) {
    dft8<DIR>(v);
  } else {
    dft_composite<DIR, P>(v);
  }
}

}  // namespace cgdft
