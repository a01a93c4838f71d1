// cg_colour.h -- per-cell pieces of the realisation kernels, shared by the colouring kernel
// (cg_noise.hip) and the colouring fused into the inverse z pass (cg_fft_fixed.h):
// counter-based noise in the FFTGrid::fillInComplexNoise layout, lib_matrCholCpx(3) and the
// product L z of AVOInversion::simulate (avoinversion.cpp:1421-1460).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cg_noise.h"

namespace cgcol {

// Philox4x32-10 (Salmon et al., SC'11).  Each round's two 32 x 32 -> 64-bit products are taken
// whole (one v_mad_u64_u32 each) rather than as separate low and high halves.
__device__ __forceinline__ void philox(uint32_t (&c)[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t lo0 = (uint32_t)p0, hi0 = (uint32_t)(p0 >> 32);
    const uint32_t lo1 = (uint32_t)p1, hi1 = (uint32_t)(p1 >> 32);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
}

// Two N(0,1) for (seed, stream, cell): u1 = (c0>>8 + 1) 2^-24 in (0, 1], u2 = (c1>>8) 2^-24.
__device__ __forceinline__ float2 cb_normal2(uint64_t seed, uint32_t stream, uint64_t cell)
{
  uint32_t c[4] = {(uint32_t)cell, (uint32_t)(cell >> 32), stream, 0u};
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float u1 = (float)((c[0] >> 8) + 1u) * 5.9604644775390625e-08f;
  const float u2 = (float)(c[1] >> 8) * 5.9604644775390625e-08f;
  const float r = sqrtf(-2.0f * logf(u1));
  float s, co;
  sincospif(2.0f * u2, &s, &co);
  return make_float2(r * co, r * s);
}

// Noise value of cell idx in the fillInComplexNoise layout.
__device__ __forceinline__ float2 noise_cell(uint64_t seed, uint32_t stream, long idx, int cnxp, int nyp, int nzp)
{
  const float sd = 0.70710678118654752440f;  // float(1/sqrt(2)), fftgrid.cpp:1002
  const int xs = (int)(idx % cnxp);
  if (xs == 0 || (xs == cnxp - 1 && (idx & 1))) {
    const long jk = idx / cnxp;
    const long jind = jk % nyp, kind = jk / nyp;
    const long jcc = jind == 0 ? 0 : nyp - jind;
    const long kcc = kind == 0 ? 0 : nzp - kind;
    const long jkcc = jcc + kcc * nyp;
    if (jkcc == jk) {
      const float2 z = cb_normal2(seed, stream, (uint64_t)idx);
      return make_float2(z.x, 0.f);
    }
    if (jkcc > jk) {
      const float2 z = cb_normal2(seed, stream, (uint64_t)idx);
      return make_float2(sd * z.x, sd * z.y);
    }
    const float2 z = cb_normal2(seed, stream, (uint64_t)(jkcc * cnxp + xs));
    return make_float2(sd * z.x, -(sd * z.y));
  }
  const float2 z = cb_normal2(seed, stream, (uint64_t)idx);
  return make_float2(sd * z.x, sd * z.y);
}

struct c32 {
  float re, im;
};

// lib_matrCholCpx(3, P) on the lower triangle; returns 0 on success.
__device__ __forceinline__ int chol3(c32 (&m)[6])
{
  // t(i,j) = i(i+1)/2 + j: 0:(0,0) 1:(1,0) 2:(1,1) 3:(2,0) 4:(2,1) 5:(2,2)
  const float factor = m[0].re;
  if (!(factor > 0.f)) return 1;
  const float inv = 1.0f / factor;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    m[t].re *= inv;
    m[t].im *= inv;
  }
  const double tol = 1e-20;
  // row 0
  if (!((double)m[0].re > tol)) return 1;
  m[0].re = sqrtf(m[0].re);
  m[0].im = 0.f;
  // row 1
  if (!((double)m[2].re > tol)) return 1;
  m[1].re = m[1].re / m[0].re;
  m[1].im = m[1].im / m[0].re;
  float piv = m[2].re - (m[1].re * m[1].re + m[1].im * m[1].im);
  if (!((double)piv > tol)) return 1;
  m[2].re = sqrtf(piv);
  m[2].im = 0.f;
  // row 2
  if (!((double)m[5].re > tol)) return 1;
  m[3].re = m[3].re / m[0].re;
  m[3].im = m[3].im / m[0].re;
  {
    const float lr = m[3].re * m[1].re + m[3].im * m[1].im;
    const float li = -(m[3].re * m[1].im) + m[3].im * m[1].re;
    m[4].re = (m[4].re - lr) / m[2].re;
    m[4].im = (m[4].im - li) / m[2].re;
  }
  piv = m[5].re - (m[3].re * m[3].re + m[3].im * m[3].im + m[4].re * m[4].re + m[4].im * m[4].im);
  if (!((double)piv > tol)) return 1;
  m[5].re = sqrtf(piv);
  m[5].im = 0.f;
  const float sf = sqrtf(factor);
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    m[t].re *= sf;
    m[t].im *= sf;
  }
  return 0;
}

// One cell of the colouring: out[g] = L z_g for the posterior covariance at idx (z = 0 where the
// factorisation fails, which the return value reports).  Noise from the counter generator
// (streams stream0 .. stream0+2), or `z` as given when from_grids.
__device__ __forceinline__ bool colour_cell(const cg::ColourParams &p, long idx, float2 (&z)[3], float2 (&o)[3])
{
  const float2 c00 = p.cov[0][idx], c11 = p.cov[1][idx], c22 = p.cov[2][idx];
  const float2 c01 = p.cov[3][idx], c02 = p.cov[4][idx], c12 = p.cov[5][idx];
  // Hermitian completion of the six stored entries (avoinversion.cpp:1425-1437)
  c32 m[6] = {{c00.x, c00.y}, {c01.x, -c01.y}, {c11.x, c11.y}, {c02.x, -c02.y}, {c12.x, -c12.y}, {c22.x, c22.y}};
  if (!p.from_grids) {
#pragma unroll
    for (int g = 0; g < 3; ++g) z[g] = noise_cell(p.seed, p.stream0 + g, idx, p.cnxp, p.nyp, p.nzp);
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) o[g] = make_float2(0.f, 0.f);
  const bool failed = chol3(m) != 0;
  if (!failed) {
    const int row[3] = {0, 1, 3};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) {
        const c32 l = m[row[a] + b];
        o[a].x += l.re * z[b].x - l.im * z[b].y;
        o[a].y += l.re * z[b].y + l.im * z[b].x;
      }
  }
  return failed;
}

}  // namespace cgcol
