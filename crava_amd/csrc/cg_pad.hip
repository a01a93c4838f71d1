// cg_pad.hip -- padding fills of the grids that enter the path: the background model's linear
// edge interpolation (fftgrid.cpp:125-213) and the seismic data's trace interpolation and
// tapered extrapolation (fftgrid.cpp:2349-2518).
//
// The reference fills the padding in one sequential sweep whose later values read earlier
// ones.  Here every cell evaluates the chain it depends on directly (z padding from y-padded
// values, those from x-padded values, those from the cube), so the cells are written in any
// order, with the same floats.  Every product and sum is rounded on its own, as the
// reference's x86-64 build computes them (no fused multiply-add).
#include "cg_internal.h"

#pragma clang fp contract(off)

// named (not anonymous) so that profilers show the kernels' names
namespace cgpad {

struct PadGeom {
  int nx, ny, nz, nxp, nyp, nzp, rnxp;
  float sa, sb, sc;  // 1 / (padding + 1) per axis, rounded on the host as the reference does
};

__device__ __forceinline__ float blend(float lo, float wlo, float hi, float whi) { return lo * wlo + hi * whi; }

// The fill runs as four launches over the row classes (box rows, y-padding rows, z-padding
// rows, rows in both paddings), so each is branch-free.  A workgroup takes kRows rows of one
// plane; their source cube rows (row j, or rows 0 and ny - 1 in the y padding, of plane k, or
// planes 0 and nz - 1 in the z padding) and those rows' edge samples are fixed before the cell
// loop, so per cell the x blend is arithmetic and the in-row sample (clamped index) an
// unconditional load: a thread's loads for its cell pair (x even, x + 1; rnxp is even) of every
// row issue before the first blend.  Blends nest x, then y, then z, as the reference's sweep
// composes them.
constexpr int kRows = 4;

template <bool YPAD, bool ZPAD>
__global__ void __launch_bounds__(1024)
k_fill_padded(float *__restrict__ rv, const float *__restrict__ cube, PadGeom g, int j_begin, int j_end, int k_begin)
{
  constexpr int NS = (YPAD ? 2 : 1) * (ZPAD ? 2 : 1);  // source rows per output row
  const int k = k_begin + blockIdx.y;
  const float wz0 = g.sc * (float)(k - g.nz + 1), wz1 = g.sc * (float)(g.nzp - k);
  const long nx = g.nx;
  const float *src[kRows][NS];
  float lo[kRows][NS], hi[kRows][NS], wy0[kRows], wy1[kRows];
#pragma unroll
  for (int r = 0; r < kRows; ++r) {
    const int j = min(j_begin + (int)blockIdx.x * kRows + r, j_end - 1);  // clamped rows are not stored
    wy0[r] = g.sb * (float)(j - g.ny + 1);
    wy1[r] = g.sb * (float)(g.nyp - j);
#pragma unroll
    for (int q = 0; q < NS; ++q) {
      // q: bit 0 selects row ny - 1 (y padding), the next bit plane nz - 1 (z padding)
      const int jq = YPAD ? ((q & 1) ? g.ny - 1 : 0) : j;
      const int kq = ZPAD ? ((q >> (YPAD ? 1 : 0)) & 1 ? g.nz - 1 : 0) : k;
      src[r][q] = cube + ((long)kq * g.ny + jq) * nx;
      lo[r][q] = src[r][q][0];
      hi[r][q] = src[r][q][nx - 1];
    }
  }
  const int pairs = g.rnxp / 2;
  for (int p = threadIdx.x; p < pairs; p += blockDim.x) {
    float2 v[kRows];
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      float c[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * p + h;
        const bool in = i < g.nx;
        const int ic = in ? i : g.nx - 1;
        const float wx0 = g.sa * (float)(i - g.nx + 1), wx1 = g.sa * (float)(g.nxp - i);
        float x[NS];
#pragma unroll
        for (int q = 0; q < NS; ++q) {
          const float s = src[r][q][ic];
          x[q] = in ? s : blend(lo[r][q], wx0, hi[r][q], wx1);
        }
        float val;
        if (YPAD) {
          val = blend(x[0], wy0[r], x[1], wy1[r]);
          if (ZPAD) val = blend(val, wz0, blend(x[2], wy0[r], x[3], wy1[r]), wz1);
        } else {
          val = ZPAD ? blend(x[0], wz0, x[1], wz1) : x[0];
        }
        c[h] = i < g.nxp ? val : 0.f;  // the row padding of the in-place layout
      }
      v[r] = make_float2(c[0], c[1]);
    }
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int j = j_begin + (int)blockIdx.x * kRows + r;
      if (j < j_end) reinterpret_cast<float2 *>(rv + ((long)k * g.nyp + j) * g.rnxp)[p] = v[r];
    }
  }
}

// Energy of every box trace over its nz samples, summed in k order (one thread per trace).
__global__ void __launch_bounds__(cg::kThreads)
k_trace_energy(const float *__restrict__ rv, float *__restrict__ energy, int nx, int ny, int nz, int rnxp, int nyp)
{
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nx * ny) return;
  const int i = t % nx, j = t / nx;
  const long plane = (long)rnxp * nyp;
  const float *p = rv + (long)j * rnxp + i;
  float e = 0.f;
  for (int k = 0; k < nz; ++k) {
    const float v = p[k * plane];
    e += v * v;
  }
  energy[t] = e;
}

// One trace interpolation: op = trace index << 4 | neighbour mask (1 left, 2 right, 4 up,
// 8 down); the column becomes the mean of the masked neighbours, summed in that order.
__device__ __forceinline__ void trace_op(float *__restrict__ rv, unsigned op, int nx, int rnxp, long plane, int nzp)
{
  const unsigned m = op & 15u, t = op >> 4;
  const int i = (int)(t % (unsigned)nx), j = (int)(t / (unsigned)nx);
  const float n = (float)__popc(m);
  float *col = rv + (long)j * rnxp + i;
  for (int k = threadIdx.x; k < nzp; k += blockDim.x) {
    const long z = k * plane;
    float s = (m & 1u) ? col[z - 1] : 0.f;
    if (m & 2u) s += col[z + 1];
    if (m & 4u) s += col[z - rnxp];
    if (m & 8u) s += col[z + rnxp];
    col[z] = s / n;
  }
}

// A run of ops in order (each may read columns written by the ones before): one workgroup,
// the barrier after each op makes its column visible to the next.
__global__ void __launch_bounds__(1024)
k_trace_ops_serial(float *__restrict__ rv, const unsigned *__restrict__ ops, int nops, int nx, int rnxp, int nyp,
                   int nzp)
{
  for (int o = 0; o < nops; ++o) {
    trace_op(rv, ops[o], nx, rnxp, (long)rnxp * nyp, nzp);
    __syncthreads();
  }
}

// Ops of one dependency level (none reads or writes a column another of them writes): one
// workgroup each.
__global__ void __launch_bounds__(cg::kThreads)
k_trace_ops_level(float *__restrict__ rv, const unsigned *__restrict__ ops, int nx, int rnxp, int nyp, int nzp)
{
  trace_op(rv, ops[blockIdx.x], nx, rnxp, (long)rnxp * nyp, nzp);
}

// Cells outside the rectangle [imin, imax] x [jmin, jmax] of good traces: the clamped
// fill-number trace times max(1 - dx^2 - dy^2 - dz^2, 0)^3.  Reads only rectangle traces,
// writes only others, so it runs in place.  grid.x over nxp * nyp, grid.y = k.
__global__ void __launch_bounds__(cg::kThreads)
k_extrapolate(float *__restrict__ rv, const int *__restrict__ ref_i, const float *__restrict__ dist_x,
              const int *__restrict__ ref_j, const float *__restrict__ dist_y, const float *__restrict__ dist_z,
              int nxp, int nyp, int rnxp, int imin, int imax, int jmin, int jmax)
{
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (unsigned)nxp * (unsigned)nyp) return;
  const int i = (int)(c % (unsigned)nxp), j = (int)(c / (unsigned)nxp), k = blockIdx.y;
  if (i >= imin && i <= imax && j >= jmin && j <= jmax) return;
  const long base = (long)k * nyp * rnxp;
  const float v = rv[base + (long)ref_j[j] * rnxp + ref_i[i]];
  const float dx = dist_x[i], dy = dist_y[j], dz = dist_z[k];
  double t = 1.0 - (double)(dx * dx) - (double)(dy * dy) - (double)(dz * dz);
  t = t > 0.0 ? t : 0.0;
  const float mult = (float)(t * t * t);
  rv[base + (long)j * rnxp + i] = mult * v;
}

// Signed cyclic offset of index i on a lattice of n (FFTGrid's cycleI / cycleJ).
__device__ __forceinline__ int cyc(int i, int n)
{
  const int c = i - n;
  return i < -c ? i : c;
}

// FFTGrid::fillInParamCorr / FillInLateralCorr (fftgrid.cpp:830-923) with the multiplyByScalar
// that follows them (seismicparametersholder.cpp:336-348) applied to every cell: the circular
// temporal correlation interpolated at k + cycleI gradI + cycleJ gradJ times the lateral
// correlation, RMISSING on the row padding.  A workgroup takes kRows rows (j) of plane k, a thread
// cell pairs (float2 stores).  Without gradients (GRAD false) the interpolation point is k for the
// whole plane, so the table reads are uniform; the arithmetic is the same either way.
template <bool GRAD>
__global__ void __launch_bounds__(1024)
k_fill_param_corr(float *__restrict__ rv, const float *__restrict__ cxy, const float *__restrict__ ct, int nxp,
                  int nyp, int nzp, int rnxp, float gi, float gj, float scale)
{
  const int k = blockIdx.y;
  // the plane's uniform interpolation (no gradients)
  const float c0 = ct[k], c1 = ct[k != nzp - 1 ? k + 1 : 0];
  for (int p = threadIdx.x; p < rnxp / 2; p += blockDim.x) {
#pragma unroll
    for (int r = 0; r < kRows; ++r) {
      const int j = (int)blockIdx.x * kRows + r;
      if (j >= nyp) break;
      float v[2];
#pragma unroll
      for (int h = 0; h < 2; ++h) {
        const int i = 2 * p + h;
        float x = -99999.0f;
        if (i < nxp) {
          float sub = (float)k + (float)cyc(i, nxp) * gi + (float)cyc(j, nyp) * gj;
          int base = (int)floorf(sub);
          sub -= (float)base;
          float t0 = c0, t1 = c1;
          if (GRAD) {
            base %= nzp;  // the reference's cyclic wrap loops
            if (base < 0) base += nzp;
            t0 = ct[base];
            t1 = ct[base != nzp - 1 ? base + 1 : 0];
          }
          x = (1.f - sub) * t0;
          x += sub * t1;
          x *= cxy[i + nxp * j];
        }
        v[h] = x * scale;
      }
      reinterpret_cast<float2 *>(rv + ((long)k * nyp + j) * rnxp)[p] = make_float2(v[0], v[1]);
    }
  }
}

// FFTGrid::fillInErrCorr (fftgrid.cpp:777-828): the lateral correlation weighted by the distance
// of k + cycleI gradI + cycleJ gradJ from the k = 0 plane (cyclically, within one sample), zero
// elsewhere, RMISSING on the row padding.
__global__ void __launch_bounds__(cg::kThreads)
k_fill_err_corr(float *__restrict__ rv, const float *__restrict__ cxy, int nxp, int nyp, int nzp, int rnxp, float gi,
                float gj)
{
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (unsigned)rnxp * (unsigned)nyp) return;
  const int i = (int)(c % (unsigned)rnxp), j = (int)(c / (unsigned)rnxp), k = blockIdx.y;
  float v = -99999.0f;
  if (i < nxp) {
    float sub = (float)k + (float)cyc(i, nxp) * gi + (float)cyc(j, nyp) * gj;
    v = 0.f;
    if (fabsf(sub) < 1.0f || fabsf(sub - (float)nzp) < 1.0f) {
      int base = (int)sub;  // 0 here, or nzp - 1 or nzp near the far end
      sub -= (float)base;
      if (base >= 1) base -= nzp;
      if (base < 0) sub = 1.f - sub;
      v = (1.f - sub) * cxy[i + nxp * j];
    }
  }
  rv[(long)k * nyp * rnxp + c] = v;
}

// The trace taper of CravaResult::ComputeSyntSeismic (cravaresult.cpp:2338-2346): box traces
// keep their nz samples and run linearly from the last sample back to the first over the z
// padding (fac ((k - nz) r_0 + (nzp - k - 1) r_{nz-1}), fac = 1 / (nzp - nz - 1)); columns
// outside the box and the row padding are 0 (the z-only operators that follow never mix
// columns).  grid.x over a plane's rnxp * nyp cells, grid.y = k.
__global__ void __launch_bounds__(cg::kThreads)
k_fill_trace_taper(float *__restrict__ rv, const float *__restrict__ cube, int nx, int ny, int nz, int nyp, int rnxp,
                   int nzp, float fac)
{
  const unsigned c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= (unsigned)rnxp * (unsigned)nyp) return;
  const int i = (int)(c % (unsigned)rnxp), j = (int)(c / (unsigned)rnxp), k = blockIdx.y;
  float v = 0.f;
  if (i < nx && j < ny) {
    const long plane = (long)nx * ny, col = (long)j * nx + i;
    if (k < nz) {
      v = cube[k * plane + col];
    } else {
      const float r0 = cube[col], r1 = cube[(nz - 1) * plane + col];
      v = fac * ((float)(k - nz) * r0 + (float)(nzp - k - 1) * r1);
    }
  }
  rv[(long)k * nyp * rnxp + c] = v;
}

}  // namespace cgpad

namespace cg {

using namespace cgpad;

int fill_padded(cg_ctx *ctx, float *rv, const float *cube, int nx, int ny, int nz, int nxp, int nyp, int nzp)
{
  PadGeom g{nx, ny, nz, nxp, nyp, nzp, 2 * (nxp / 2 + 1), 0.f, 0.f, 0.f};
  g.sa = 1.0f / (float)(nxp - nx + 1);
  g.sb = 1.0f / (float)(nyp - ny + 1);
  g.sc = 1.0f / (float)(nzp - nz + 1);
  const int pairs = g.rnxp / 2, threads = pairs >= 1024 ? 1024 : (pairs + 63) / 64 * 64;
  auto grid = [](int j0, int j1, int k0, int k1) {
    return dim3((unsigned)((j1 - j0 + kRows - 1) / kRows), (unsigned)(k1 - k0));
  };
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  hipStream_t s = ctx->stream;
  k_fill_padded<false, false><<<grid(0, ny, 0, nz), threads, 0, s>>>(rv, cube, g, 0, ny, 0);
  if (nyp > ny) k_fill_padded<true, false><<<grid(ny, nyp, 0, nz), threads, 0, s>>>(rv, cube, g, ny, nyp, 0);
  if (nzp > nz) k_fill_padded<false, true><<<grid(0, ny, nz, nzp), threads, 0, s>>>(rv, cube, g, 0, ny, nz);
  if (nyp > ny && nzp > nz)
    k_fill_padded<true, true><<<grid(ny, nyp, nz, nzp), threads, 0, s>>>(rv, cube, g, ny, nyp, nz);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fill_padded");
}

int fill_param_corr(cg_ctx *ctx, float *rv, const float *cxy, const float *ct, int nxp, int nyp, int nzp, float gi,
                    float gj, float scale)
{
  const int rnxp = 2 * (nxp / 2 + 1), pairs = rnxp / 2;
  const int threads = pairs >= 1024 ? 1024 : (pairs + 63) / 64 * 64;
  const dim3 grid((unsigned)((nyp + kRows - 1) / kRows), (unsigned)nzp);
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  if (gi == 0.f && gj == 0.f)
    k_fill_param_corr<false><<<grid, threads, 0, ctx->stream>>>(rv, cxy, ct, nxp, nyp, nzp, rnxp, gi, gj, scale);
  else
    k_fill_param_corr<true><<<grid, threads, 0, ctx->stream>>>(rv, cxy, ct, nxp, nyp, nzp, rnxp, gi, gj, scale);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fill_param_corr");
}

int fill_trace_taper(cg_ctx *ctx, float *rv, const float *cube, int nx, int ny, int nz, int nxp, int nyp, int nzp)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  const float fac = 1.0f / (float)(nzp - nz - 1);
  const unsigned cells = (unsigned)rnxp * (unsigned)nyp;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_fill_trace_taper<<<dim3((cells + kThreads - 1) / kThreads, (unsigned)nzp), kThreads, 0, ctx->stream>>>(
      rv, cube, nx, ny, nz, nyp, rnxp, nzp, fac);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fill_trace_taper");
}

int fill_err_corr(cg_ctx *ctx, float *rv, const float *cxy, int nxp, int nyp, int nzp, float gi, float gj)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  const unsigned cells = (unsigned)rnxp * (unsigned)nyp;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_fill_err_corr<<<dim3((cells + kThreads - 1) / kThreads, (unsigned)nzp), kThreads, 0, ctx->stream>>>(
      rv, cxy, nxp, nyp, nzp, rnxp, gi, gj);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fill_err_corr");
}

int trace_energy(cg_ctx *ctx, const float *rv, float *energy, int nx, int ny, int nz, int rnxp, int nyp)
{
  const int n = nx * ny;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_trace_energy<<<(n + kThreads - 1) / kThreads, kThreads, 0, ctx->stream>>>(rv, energy, nx, ny, nz, rnxp, nyp);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_trace_energy");
}

int trace_ops(cg_ctx *ctx, float *rv, const unsigned *ops, int nops, bool one_level, int nx, int rnxp, int nyp,
              int nzp)
{
  if (nops <= 0) return CG_OK;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  if (one_level) {
    k_trace_ops_level<<<(unsigned)nops, kThreads, 0, ctx->stream>>>(rv, ops, nx, rnxp, nyp, nzp);
  } else {
    const int threads = nzp >= 1024 ? 1024 : (nzp + 63) / 64 * 64;
    k_trace_ops_serial<<<1, threads, 0, ctx->stream>>>(rv, ops, nops, nx, rnxp, nyp, nzp);
  }
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_trace_ops");
}

int extrapolate(cg_ctx *ctx, float *rv, const int *ref_i, const float *dist_x, const int *ref_j, const float *dist_y,
                const float *dist_z, int nxp, int nyp, int nzp, int imin, int imax, int jmin, int jmax)
{
  const unsigned cells = (unsigned)nxp * (unsigned)nyp;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  k_extrapolate<<<dim3((cells + kThreads - 1) / kThreads, (unsigned)nzp), kThreads, 0, ctx->stream>>>(
      rv, ref_i, dist_x, ref_j, dist_y, dist_z, nxp, nyp, 2 * (nxp / 2 + 1), imin, imax, jmin, jmax);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_extrapolate");
}

}  // namespace cg
