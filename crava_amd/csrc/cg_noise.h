// cg_noise.h -- parameters of the realisation kernels (cg_noise.hip).
#pragma once

#include "cg_internal.h"

namespace cg {

// floor(n / d) for 32-bit n by a multiply-high, an add and a shift (Granlund & Montgomery,
// "Division by invariant integers using multiplication", PLDI 1994, fig. 4.1):
//   l = ceil(log2 d), m = floor(2^32 (2^l - d) / d) + 1, q = (umulhi(m, n) + n) >> l.
struct FastDiv {
  uint32_t m, l;
};
inline FastDiv make_fastdiv(uint32_t d)
{
  uint32_t l = 0;
  while ((1ull << l) < d) ++l;
  return FastDiv{(uint32_t)(((1ull << 32) * ((1ull << l) - d)) / d + 1), l};
}
__device__ __forceinline__ uint32_t fast_div(uint32_t n, FastDiv f)
{
  return (uint32_t)(((uint64_t)__umulhi(n, f.m) + n) >> f.l);
}

struct ColourParams {
  long csize;
  int cnxp, nyp, nzp;
  FastDiv by_cnxp, by_nyp;  // set by colour_dims
  const float2 *cov[6];   // posterior covariance spectra 00 11 22 01 02 12
  float2 *out[3];         // coloured spectra (noise in, if from_grids)
  const float2 *mean[3];  // posterior mean spectra added to L z (null: the mean is added later)
  int from_grids;
  uint64_t seed;
  uint32_t stream0;       // streams stream0, stream0+1, stream0+2
  unsigned long long *fails;
  // the second realisation of a pair (k_colour_zinv2): outputs and streams stream1 .. stream1+2
  float2 *out2[3];
  uint32_t stream1;
};

struct ForwardParams {
  long plane;
  int nzp, nt;
  const float2 *K;  // [nzp][nt][3], device
  const float2 *m[3];
  float2 *d[8];
};

int forward_model(cg_ctx *ctx, const ForwardParams &p);
int add_white_noise(cg_ctx *ctx, float *x, int nxp, int rnxp, long rows, float sigma, uint64_t seed, uint32_t stream);
// Sum of squares over the nx * ny * nz box of a real grid in a fixed order (rows in i order, rows
// of a plane in j order, planes in k order); ref: the data variance's definition (float products,
// zero samples skipped and counted in *dcount).  work: box_sum_squares_work(ny, nz) bytes of
// device memory; dout / dcount device memory.
int box_sum_squares(cg_ctx *ctx, const float *x, int nx, int ny, int nz, int rnxp, int nyp, bool ref, void *work,
                    double *dout, unsigned long long *dcount);
size_t box_sum_squares_work(int ny, int nz);
int fill_noise(cg_ctx *ctx, float2 *out, long csize, int cnxp, int nyp, int nzp, uint64_t seed, uint32_t stream);
int colour(cg_ctx *ctx, const ColourParams &p);
// Two realisations of one posterior (out2 / stream1 set, counter noise), unfused.
int colour2(cg_ctx *ctx, const ColourParams &p);
// Fills the divisors of p from p.cnxp and p.nyp.
inline void colour_dims(ColourParams &p)
{
  p.by_cnxp = make_fastdiv((uint32_t)p.cnxp);
  p.by_nyp = make_fastdiv((uint32_t)p.nyp);
}

struct ColourItem {
  ColourParams p;
  unsigned block0;
};
int colour_batch(cg_ctx *ctx, const ColourItem *dev_items, int n, unsigned blocks);
int zero_counters(cg_ctx *ctx, unsigned long long *d, int n);

}  // namespace cg
