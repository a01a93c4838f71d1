// cg_dft.h -- complex helpers and the radix-2/3/4/5/7/8 DFT butterflies of the LDS FFT engines.
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

template <int DIR, int P>
__device__ __forceinline__ void dft(float2 *v)
{
  if (P == 2) dft2<DIR>(v);
  if (P == 3) dft3<DIR>(v);
  if (P == 4) dft4<DIR>(v);
  if (P == 5) dft5<DIR>(v);
  if (P == 7) dft7<DIR>(v);
  if (P == 8) dft8<DIR>(v);
}

}  // namespace cgdft
