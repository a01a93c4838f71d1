// cg_colour.h -- per-cell pieces of the realisation kernels, shared by the colouring kernel
// (cg_noise.hip) and the colouring fused into the inverse z pass (cg_fft_fixed.h):
// counter-based noise in the FFTGrid::fillInComplexNoise layout, lib_matrCholCpx(3) and the
// product L z of AVOInversion::simulate (avoinversion.cpp:1421-1460).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

#include "cg_mem.h"
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

// Box-Muller on two 32-bit words: u1 = (a>>8 + 1) 2^-24 in (0, 1], u2 = (b>>8) 2^-24;
// (r cos 2 pi u2, r sin 2 pi u2) with r = sqrt(-2 ln u1), from the hardware's log2, square root,
// sine and cosine (the latter two take turns, so 2 pi u2 needs no range reduction).  u1 >= 2^-24
// keeps every operand normal.
__device__ __forceinline__ float2 box_muller(uint32_t a, uint32_t b)
{
#pragma clang fp contract(off)
  const float u1 = (float)((a >> 8) + 1u) * 5.9604644775390625e-08f;
  const float u2 = (float)(b >> 8) * 5.9604644775390625e-08f;
  const float r = __builtin_amdgcn_sqrtf(-1.38629436111989061883f * __builtin_amdgcn_logf(u1));  // -2 ln 2 log2 u1
  return make_float2(r * __builtin_amdgcn_cosf(u2), r * __builtin_amdgcn_sinf(u2));
}

// One Philox block per (seed, cell, stream pair): counter (cell, stream >> 1, 0), key seed.
__device__ __forceinline__ void cb_block(uint64_t seed, uint32_t pair, uint64_t cell, uint32_t (&c)[4])
{
  c[0] = (uint32_t)cell;
  c[1] = (uint32_t)(cell >> 32);
  c[2] = pair;
  c[3] = 0u;
  philox(c, (uint32_t)seed, (uint32_t)(seed >> 32));
}

// Two N(0,1) for (seed, stream, cell): words 0-1 of the stream pair's block for an even stream,
// words 2-3 for an odd one.
__device__ __forceinline__ float2 cb_normal2(uint64_t seed, uint32_t stream, uint64_t cell)
{
  uint32_t c[4];
  cb_block(seed, stream >> 1, cell, c);
  return (stream & 1u) ? box_muller(c[2], c[3]) : box_muller(c[0], c[1]);
}

// The normal pairs of streams s0, s0 + 1, s0 + 2 at one cell: two Philox blocks.
__device__ __forceinline__ void cb_normal2x3(uint64_t seed, uint32_t s0, uint64_t cell, float2 (&z)[3])
{
  uint32_t a[4], b[4];
  cb_block(seed, s0 >> 1, cell, a);
  cb_block(seed, (s0 >> 1) + 1u, cell, b);
  if (s0 & 1u) {
    z[0] = box_muller(a[2], a[3]);
    z[1] = box_muller(b[0], b[1]);
    z[2] = box_muller(b[2], b[3]);
  } else {
    z[0] = box_muller(a[0], a[1]);
    z[1] = box_muller(a[2], a[3]);
    z[2] = box_muller(b[0], b[1]);
  }
}

// The fillInComplexNoise layout: cell idx takes its value from the normal pair (z0, z1) drawn at
// cell *src: kind 0: (z0, z1) / sqrt 2; kind 1 (self-conjugate cell): (z0, 0); kind 2 (the
// conjugate mirror of an earlier cell): (z0, -z1) / sqrt 2.
__device__ __forceinline__ int noise_source(long idx, int cnxp, int nyp, int nzp, long *src)
{
  const int xs = (int)(idx % cnxp);
  *src = idx;
  if (xs == 0 || (xs == cnxp - 1 && (idx & 1))) {
    const long jk = idx / cnxp;
    const long jind = jk % nyp, kind = jk / nyp;
    const long jcc = jind == 0 ? 0 : nyp - jind;
    const long kcc = kind == 0 ? 0 : nzp - kind;
    const long jkcc = jcc + kcc * nyp;
    if (jkcc == jk) return 1;
    if (jkcc > jk) return 0;
    *src = jkcc * cnxp + xs;
    return 2;
  }
  return 0;
}

// noise_source for 32-bit cell indices (grids of fewer than 2^32 complex cells), with the
// divisions by cnxp and nyp done by multiplication.
__device__ __forceinline__ int noise_source32(uint32_t idx, const cg::ColourParams &p, uint32_t *src)
{
  const uint32_t jk = cg::fast_div(idx, p.by_cnxp);
  const uint32_t xs = idx - jk * (uint32_t)p.cnxp;
  *src = idx;
  if (xs == 0 || (xs == (uint32_t)p.cnxp - 1 && (idx & 1))) {
    const uint32_t kind = cg::fast_div(jk, p.by_nyp), jind = jk - kind * (uint32_t)p.nyp;
    const uint32_t jcc = jind == 0 ? 0 : p.nyp - jind;
    const uint32_t kcc = kind == 0 ? 0 : p.nzp - kind;
    const uint32_t jkcc = jcc + kcc * (uint32_t)p.nyp;
    if (jkcc == jk) return 1;
    if (jkcc > jk) return 0;
    *src = jkcc * (uint32_t)p.cnxp + xs;
    return 2;
  }
  return 0;
}

__device__ __forceinline__ float2 noise_value(int kind, float2 z)
{
  const float sd = 0.70710678118654752440f;  // float(1/sqrt(2)), fftgrid.cpp:1002
  if (kind == 1) return make_float2(z.x, 0.f);
  return make_float2(sd * z.x, kind == 2 ? -(sd * z.y) : sd * z.y);
}

// Noise value of cell idx of one stream.
__device__ __forceinline__ float2 noise_cell(uint64_t seed, uint32_t stream, long idx, int cnxp, int nyp, int nzp)
{
  long src;
  const int kind = noise_source(idx, cnxp, nyp, nzp, &src);
  return noise_value(kind, cb_normal2(seed, stream, (uint64_t)src));
}

struct c32 {
  float re, im;
};

// lib_matrCholCpx(3, P) on the lower triangle; returns 0 on success.  The pivot tests are the
// reference's (value > 1e-20 in double), as float comparisons with the smallest float above
// 1e-20; divisions by a pivot are products with its reciprocal.
__device__ __forceinline__ int chol3(c32 (&m)[6])
{
#pragma clang fp contract(off)
  // t(i,j) = i(i+1)/2 + j: 0:(0,0) 1:(1,0) 2:(1,1) 3:(2,0) 4:(2,1) 5:(2,2)
  const float factor = m[0].re;
  if (!(factor > 0.f)) return 1;
  const float inv = 1.0f / factor;
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    m[t].re *= inv;
    m[t].im *= inv;
  }
  const float tol = __uint_as_float(0x1e3ce509u);  // 1.00000005e-20f: x > 1e-20 (double) <=> x >= tol
  // row 0
  if (!(m[0].re >= tol)) return 1;
  m[0].re = __builtin_amdgcn_sqrtf(m[0].re);
  m[0].im = 0.f;
  const float r0 = __builtin_amdgcn_rcpf(m[0].re);
  // row 1
  if (!(m[2].re >= tol)) return 1;
  m[1].re = m[1].re * r0;
  m[1].im = m[1].im * r0;
  float piv = m[2].re - (m[1].re * m[1].re + m[1].im * m[1].im);
  if (!(piv >= tol)) return 1;
  m[2].re = __builtin_amdgcn_sqrtf(piv);
  m[2].im = 0.f;
  // row 2
  if (!(m[5].re >= tol)) return 1;
  m[3].re = m[3].re * r0;
  m[3].im = m[3].im * r0;
  {
    const float r2 = __builtin_amdgcn_rcpf(m[2].re);
    const float lr = m[3].re * m[1].re + m[3].im * m[1].im;
    const float li = -(m[3].re * m[1].im) + m[3].im * m[1].re;
    m[4].re = (m[4].re - lr) * r2;
    m[4].im = (m[4].im - li) * r2;
  }
  piv = m[5].re - (m[3].re * m[3].re + m[3].im * m[3].im + m[4].re * m[4].re + m[4].im * m[4].im);
  if (!(piv >= tol)) return 1;
  m[5].re = __builtin_amdgcn_sqrtf(piv);
  m[5].im = 0.f;
  const float sf = sqrtf(factor);
#pragma unroll
  for (int t = 0; t < 6; ++t) {
    m[t].re *= sf;
    m[t].im *= sf;
  }
  return 0;
}

// o + l z with explicitly placed fused multiply-adds (the per-cell functions compile with
// contraction off, so a single and a paired cell round alike whatever kernel inlines them).
__device__ __forceinline__ float2 lz(const c32 &l, const float2 &z, float2 o)
{
  o.x = __builtin_fmaf(l.re, z.x, o.x);
  o.x = __builtin_fmaf(-l.im, z.y, o.x);
  o.y = __builtin_fmaf(l.re, z.y, o.y);
  o.y = __builtin_fmaf(l.im, z.x, o.y);
  return o;
}

// One cell of the colouring: out[g] = L z_g (+ the mean spectrum, when given) for the posterior
// covariance at idx (L z = 0 where the factorisation fails, which the return value reports).  Noise from the counter generator
// (streams stream0 .. stream0+2), or `z` as given when from_grids.
__device__ __forceinline__ bool colour_cell(const cg::ColourParams &p, long idx, float2 (&z)[3], float2 (&o)[3])
{
#pragma clang fp contract(off)
  const float2 c00 = cg::gld(p.cov[0] + idx), c11 = cg::gld(p.cov[1] + idx), c22 = cg::gld(p.cov[2] + idx);
  const float2 c01 = cg::gld(p.cov[3] + idx), c02 = cg::gld(p.cov[4] + idx), c12 = cg::gld(p.cov[5] + idx);
  float2 mv[3];  // loaded with the covariance, before the arithmetic that hides their latency
#pragma unroll
  for (int g = 0; g < 3; ++g) mv[g] = p.mean[g] ? cg::gld(p.mean[g] + idx) : make_float2(0.f, 0.f);
  // Hermitian completion of the six stored entries (avoinversion.cpp:1425-1437)
  c32 m[6] = {{c00.x, c00.y}, {c01.x, -c01.y}, {c11.x, c11.y}, {c02.x, -c02.y}, {c12.x, -c12.y}, {c22.x, c22.y}};
  if (!p.from_grids) {
    long src;
    int kind;
    if (p.csize <= 0xFFFFFFFFl) {  // workgroup-uniform
      uint32_t s32;
      kind = noise_source32((uint32_t)idx, p, &s32);
      src = s32;
    } else {
      kind = noise_source(idx, p.cnxp, p.nyp, p.nzp, &src);
    }
    cb_normal2x3(p.seed, p.stream0, (uint64_t)src, z);
#pragma unroll
    for (int g = 0; g < 3; ++g) z[g] = noise_value(kind, z[g]);
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
        o[a] = lz(l, z[b], o[a]);
      }
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    o[g].x += mv[g].x;
    o[g].y += mv[g].y;
  }
  return failed;
}

// Whether a pair's six streams (stream0 .. stream0 + 5) are three whole Philox blocks.
__device__ __host__ __forceinline__ bool pair_in_three_blocks(const cg::ColourParams &p)
{
  return p.stream1 == p.stream0 + 3u && !(p.stream0 & 1u);
}

// colour_cell for two realisations of one posterior (counter noise only): the covariance and
// mean are read and the 3x3 factor computed once; each realisation's arithmetic is colour_cell's,
// operation for operation, so the pair equals two single cells bit for bit.  THREE: the streams
// are three whole Philox blocks (pair_in_three_blocks), drawn once each.
template <bool THREE>
__device__ __forceinline__ bool colour_cell2t(const cg::ColourParams &p, long idx, float2 (&o)[3], float2 (&o2)[3])
{
#pragma clang fp contract(off)
  const float2 c00 = cg::gld(p.cov[0] + idx), c11 = cg::gld(p.cov[1] + idx), c22 = cg::gld(p.cov[2] + idx);
  const float2 c01 = cg::gld(p.cov[3] + idx), c02 = cg::gld(p.cov[4] + idx), c12 = cg::gld(p.cov[5] + idx);
  float2 mv[3];
#pragma unroll
  for (int g = 0; g < 3; ++g) mv[g] = p.mean[g] ? cg::gld(p.mean[g] + idx) : make_float2(0.f, 0.f);
  c32 m[6] = {{c00.x, c00.y}, {c01.x, -c01.y}, {c11.x, c11.y}, {c02.x, -c02.y}, {c12.x, -c12.y}, {c22.x, c22.y}};
  long src;
  int kind;
  if (p.csize <= 0xFFFFFFFFl) {
    uint32_t s32;
    kind = noise_source32((uint32_t)idx, p, &s32);
    src = s32;
  } else {
    kind = noise_source(idx, p.cnxp, p.nyp, p.nzp, &src);
  }
  float2 z[3], z2[3];
  if (THREE) {
    // six consecutive streams from an even one: exactly three Philox blocks, the same words
    // cb_normal2x3 takes for each realisation
    uint32_t a[4], b[4], d[4];
    const uint32_t q = p.stream0 >> 1;
    cb_block(p.seed, q, (uint64_t)src, a);
    cb_block(p.seed, q + 1u, (uint64_t)src, b);
    cb_block(p.seed, q + 2u, (uint64_t)src, d);
    z[0] = box_muller(a[0], a[1]);
    z[1] = box_muller(a[2], a[3]);
    z[2] = box_muller(b[0], b[1]);
    z2[0] = box_muller(b[2], b[3]);
    z2[1] = box_muller(d[0], d[1]);
    z2[2] = box_muller(d[2], d[3]);
  } else {
    cb_normal2x3(p.seed, p.stream0, (uint64_t)src, z);
    cb_normal2x3(p.seed, p.stream1, (uint64_t)src, z2);
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    z[g] = noise_value(kind, z[g]);
    z2[g] = noise_value(kind, z2[g]);
    o[g] = make_float2(0.f, 0.f);
    o2[g] = make_float2(0.f, 0.f);
  }
  const bool failed = chol3(m) != 0;
  if (!failed) {
    const int row[3] = {0, 1, 3};
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
      for (int b = 0; b <= a; ++b) {
        const c32 l = m[row[a] + b];
        o[a] = lz(l, z[b], o[a]);
        o2[a] = lz(l, z2[b], o2[a]);
      }
  }
#pragma unroll
  for (int g = 0; g < 3; ++g) {
    o[g].x += mv[g].x;
    o[g].y += mv[g].y;
    o2[g].x += mv[g].x;
    o2[g].y += mv[g].y;
  }
  return failed;
}

__device__ __forceinline__ bool colour_cell2(const cg::ColourParams &p, long idx, float2 (&o)[3], float2 (&o2)[3])
{
  return pair_in_three_blocks(p) ? colour_cell2t<true>(p, idx, o, o2) : colour_cell2t<false>(p, idx, o, o2);
}

}  // namespace cgcol
