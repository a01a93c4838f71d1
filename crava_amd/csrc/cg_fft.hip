// cg_fft.hip -- mixed-radix Stockham FFT passes staged in LDS, gfx950.
//
// Replaces the FFTW 2.1.2 transforms behind FFTGrid::fftInPlace / invFFTInPlace
// (fftgrid.cpp:1587-1645; pass order rfftw/rfftwnd.c:207-288):
//   forward : x real->half-complex (one row per LDS sequence), then y, then z complex passes
//   inverse : z, y complex passes, then x half-complex->real ignoring Im of DC and Nyquist
// Each pass reads and writes the grid once.  A workgroup loads a tile (rows for x, or TI
// consecutive columns over the whole axis for y/z, 64-128 B row segments) into LDS, runs the
// radix stages between two LDS buffers and writes the tile back.  Sign -1 forward.
// Radices 2, 3, 4, 5, 7, 8 cover every padded length CRAVA produces (2^a 3^b 5^c 7^d).
#include <cmath>

#include "cg_dft.h"
#include "cg_internal.h"

// named (not anonymous) so that profilers show the kernels' names
namespace cgfft {

using namespace cgdft;

// One Stockham radix-P stage over `nbatch` sequences of length n held in LDS.
// Element e of sequence b lives at buf[b*bstride + e*estride].  Butterfly j reads
// elements j + q*(n/P), applies w_{Ns*P}^{q k} (k = j mod Ns) and writes elements
// (j/Ns)*Ns*P + k + q*Ns.  `bfast` chooses which index runs fastest across threads.
template <int DIR, int P>
__device__ __forceinline__ void stage(const float2 *__restrict__ src, float2 *__restrict__ dst, int n, int Ns,
                                      int estride, int bstride, int nbatch, bool bfast,
                                      const float2 *__restrict__ tw, int tid, int nthr)
{
  const int nb = n / P;
  const int total = nb * nbatch;
  const int twstep = n / (Ns * P);
  for (int w = tid; w < total; w += nthr) {
    int b, j;
    if (bfast) {
      b = w % nbatch;
      j = w / nbatch;
    } else {
      j = w % nb;
      b = w / nb;
    }
    const int k = j % Ns;
    float2 v[P];
    const float2 *s = src + b * bstride;
#pragma unroll
    for (int q = 0; q < P; ++q) v[q] = s[(j + q * nb) * estride];
    if (k != 0) {
#pragma unroll
      for (int q = 1; q < P; ++q) {
        float2 t = tw[q * k * twstep];  // q k twstep < n
        if (DIR > 0) t.y = -t.y;
        v[q] = cmul(v[q], t);
      }
    }
    if (P == 2) dft2<DIR>(v);
    if (P == 3) dft3<DIR>(v);
    if (P == 4) dft4<DIR>(v);
    if (P == 5) dft5<DIR>(v);
    if (P == 7) dft7<DIR>(v);
    if (P == 8) dft8<DIR>(v);
    const int d = (j / Ns) * Ns * P + k;
    float2 *o = dst + b * bstride;
#pragma unroll
    for (int q = 0; q < P; ++q) o[(d + q * Ns) * estride] = v[q];
  }
}

// Runs all stages; data starts in bufA and the final result is returned (bufA or bufB).
template <int DIR>
__device__ float2 *lds_fft(float2 *bufA, float2 *bufB, const cg::FftPlan &plan, int estride, int bstride,
                           int nbatch, bool bfast, const float2 *tw, int tid, int nthr)
{
  float2 *src = bufA, *dst = bufB;
  int Ns = 1;
  for (int s = 0; s < plan.nstages; ++s) {
    const int p = plan.radix[s];
    switch (p) {
      case 2: stage<DIR, 2>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
      case 3: stage<DIR, 3>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
      case 4: stage<DIR, 4>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
      case 5: stage<DIR, 5>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
      case 7: stage<DIR, 7>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
      default: stage<DIR, 8>(src, dst, plan.n, Ns, estride, bstride, nbatch, bfast, tw, tid, nthr); break;
    }
    Ns *= p;
    __syncthreads();
    float2 *t = src;
    src = dst;
    dst = t;
  }
  return src;
}

__device__ __forceinline__ void load_twiddles(float2 *lds_tw, const float2 *gtw, int n, int tid, int nthr)
{
  for (int m = tid; m < n; m += nthr) lds_tw[m] = gtw[m];
}

// ---------------------------------------------------------------- x axis, forward
// rows of nxp reals at pitch rnxp floats -> cnxp complex bins in place.  Even nxp: the row is
// read as nxp/2 complex numbers z[m] = x[2m] + i x[2m+1]; after the half-length FFT,
// X[k] = E[k] + w^k O[k] with E = (Z[k] + conj Z[n2-k]) / 2, O = -i (Z[k] - conj Z[n2-k]) / 2,
// w = exp(-2 pi i / nxp).  Odd nxp: full-length complex FFT of (x, 0).
// `src` and `grid` may alias (in-place): every load of a workgroup's rows precedes the
// barrier and every store follows the last stage.
__device__ __forceinline__ void x_r2c_body(const float *src, float *grid, int nxp, long rows, int rows_per_wg,
                                           const cg::FftPlan &plan, const float2 *__restrict__ tw_fft,
                                           const float2 *__restrict__ tw_row, float scale, unsigned blk)
{
  extern __shared__ float2 smem[];
  const int n = plan.n;  // nxp/2 (even) or nxp (odd)
  const int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  const bool even = (nxp % 2) == 0;
  float2 *tw = smem;
  float2 *A = smem + n;
  float2 *B = A + (size_t)rows_per_wg * n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long row0 = (long)blk * rows_per_wg;
  load_twiddles(tw, tw_fft, n, tid, nthr);
  for (int idx = tid; idx < rows_per_wg * n; idx += nthr) {
    const int r = idx / n, m = idx % n;
    const long row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < rows) {
      const float *p = src + row * rnxp;
      if (even) {
        v = reinterpret_cast<const float2 *>(p)[m];
        v = cscale(v, scale);
      } else {
        v = make_float2(p[m] * scale, 0.f);
      }
    }
    A[idx] = v;
  }
  __syncthreads();
  float2 *Z = lds_fft<-1>(A, B, plan, 1, n, rows_per_wg, false, tw, tid, nthr);
  for (int idx = tid; idx < rows_per_wg * cnxp; idx += nthr) {
    const int r = idx / cnxp, k = idx % cnxp;
    const long row = row0 + r;
    if (row >= rows) continue;
    float2 X;
    if (even) {
      const float2 zk = Z[r * n + (k % n)];
      const float2 zn = cconj(Z[r * n + ((n - k) % n)]);
      const float2 E = cscale(cadd(zk, zn), 0.5f);
      const float2 d = csub(zk, zn);
      const float2 O = make_float2(0.5f * d.y, -0.5f * d.x);  // -i d / 2
      X = cadd(E, cmul(tw_row[k], O));
    } else {
      X = Z[r * n + k];
    }
    reinterpret_cast<float2 *>(grid + row * rnxp)[k] = X;
  }
}

__global__ void __launch_bounds__(cg::kThreads)
k_fft_x_r2c(const float *src, float *grid, int nxp, long rows, int rows_per_wg, cg::FftPlan plan,
            const float2 *__restrict__ tw_fft, const float2 *__restrict__ tw_row, float scale)
{
  x_r2c_body(src, grid, nxp, rows, rows_per_wg, plan, tw_fft, tw_row, scale, cg_xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------- x axis, inverse
// cnxp half-complex bins -> nxp reals, unnormalised, then x scale (+ add[] if given).  The
// imaginary parts of the DC and (even nxp) Nyquist bins are ignored, as FFTW's hc2r does.
// Even nxp: Z[k] = (X[k] + conj X[n2-k]) + i w^-k (X[k] - conj X[n2-k]); z = IFFT_n2(Z);
// x[2m] = Re z[m], x[2m+1] = Im z[m].  Odd nxp: Hermitian extension and full-length IFFT.
__device__ __forceinline__ void x_c2r_body(float *__restrict__ grid, int nxp, long rows, int rows_per_wg,
                                           const cg::FftPlan &plan, const float2 *__restrict__ tw_fft,
                                           const float2 *__restrict__ tw_row, float scale,
                                           const float *__restrict__ add, unsigned blk)
{
  extern __shared__ float2 smem[];
  const int n = plan.n;
  const int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  const bool even = (nxp % 2) == 0;
  float2 *tw = smem;
  float2 *A = smem + n;
  float2 *B = A + (size_t)rows_per_wg * n;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const long row0 = (long)blk * rows_per_wg;
  load_twiddles(tw, tw_fft, n, tid, nthr);
  for (int idx = tid; idx < rows_per_wg * n; idx += nthr) {
    const int r = idx / n, k = idx % n;
    const long row = row0 + r;
    float2 v = make_float2(0.f, 0.f);
    if (row < rows) {
      const float2 *p = reinterpret_cast<const float2 *>(grid + row * rnxp);
      if (even) {
        float2 xk = p[k];
        float2 xn = p[n - k];
        if (k == 0) {
          xk.y = 0.f;  // DC
          xn.y = 0.f;  // Nyquist
        }
        xn = cconj(xn);
        const float2 E = cadd(xk, xn);
        const float2 O = cmul(csub(xk, xn), cconj(tw_row[k]));
        v = make_float2(E.x - O.y, E.y + O.x);  // E + i O
      } else {
        if (k == 0) {
          v = make_float2(p[0].x, 0.f);
        } else if (k < cnxp) {
          v = p[k];
        } else {
          v = cconj(p[n - k]);
        }
      }
    }
    A[idx] = v;
  }
  __syncthreads();
  float2 *Z = lds_fft<+1>(A, B, plan, 1, n, rows_per_wg, false, tw, tid, nthr);
  for (int idx = tid; idx < rows_per_wg * cnxp; idx += nthr) {
    const int r = idx / cnxp, m = idx % cnxp;
    const long row = row0 + r;
    if (row >= rows) continue;
    float *p = grid + row * rnxp;
    const float *a = add ? add + row * rnxp : nullptr;
    if (even) {
      // complex slot m < n holds reals 2m, 2m+1; slot n is the pitch padding
      float2 v = (m < n) ? cscale(Z[r * n + m], scale) : make_float2(0.f, 0.f);
      if (a && m < n) {
        v.x += a[2 * m];
        v.y += a[2 * m + 1];
      }
      reinterpret_cast<float2 *>(p)[m] = v;
    } else {
      // reals 2m, 2m+1 from the real parts of z[2m], z[2m+1] (index nxp is padding)
      const int i0 = 2 * m, i1 = 2 * m + 1;
      float v0 = Z[r * n + i0].x * scale;
      float v1 = (i1 < nxp) ? Z[r * n + i1].x * scale : 0.f;
      if (a) {
        v0 += a[i0];
        if (i1 < nxp) v1 += a[i1];
      }
      reinterpret_cast<float2 *>(p)[m] = make_float2(v0, v1);
    }
  }
}

__global__ void __launch_bounds__(cg::kThreads)
k_fft_x_c2r(float *__restrict__ grid, int nxp, long rows, int rows_per_wg, cg::FftPlan plan,
            const float2 *__restrict__ tw_fft, const float2 *__restrict__ tw_row, float scale,
            const float *__restrict__ add)
{
  x_c2r_body(grid, nxp, rows, rows_per_wg, plan, tw_fft, tw_row, scale, add, cg_xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------- strided complex axis
// Transforms along an axis with element stride S (complex units).  The tile is TI consecutive
// inner columns over all n elements: data[o*OS + e*S + c0 + t].
template <int DIR>
__device__ __forceinline__ void strided_body(float2 *__restrict__ data, const cg::FftPlan &plan, long S, long inner,
                                             long OS, int TI, unsigned tiles_per_outer,
                                             const float2 *__restrict__ tw_fft, float scale, unsigned blk)
{
  extern __shared__ float2 smem[];
  const int n = plan.n;
  float2 *tw = smem;
  float2 *A = smem + n;
  float2 *B = A + (size_t)n * TI;
  const int tid = threadIdx.x, nthr = blockDim.x;
  const unsigned o = blk / tiles_per_outer;
  const long c0 = (long)(blk % tiles_per_outer) * TI;
  float2 *base = data + (long)o * OS + c0;
  const int ncol = (int)((inner - c0) < TI ? (inner - c0) : TI);
  load_twiddles(tw, tw_fft, n, tid, nthr);
  for (int idx = tid; idx < n * TI; idx += nthr) {
    const int e = idx / TI, t = idx % TI;
    A[idx] = (t < ncol) ? base[e * S + t] : make_float2(0.f, 0.f);
  }
  __syncthreads();
  float2 *R = lds_fft<DIR>(A, B, plan, TI, 1, TI, true, tw, tid, nthr);
  for (int idx = tid; idx < n * TI; idx += nthr) {
    const int e = idx / TI, t = idx % TI;
    if (t < ncol) base[e * S + t] = (scale == 1.0f) ? R[idx] : cscale(R[idx], scale);
  }
}

template <int DIR>
__global__ void __launch_bounds__(cg::kThreads)
k_fft_strided(float2 *__restrict__ data, cg::FftPlan plan, long S, long inner, long OS, int TI,
              unsigned tiles_per_outer, const float2 *__restrict__ tw_fft, float scale)
{
  strided_body<DIR>(data, plan, S, inner, OS, TI, tiles_per_outer, tw_fft, scale, cg_xcd_remap(blockIdx.x, gridDim.x));
}

// ---------------------------------------------------------------- batched passes
// One launch transforms many grids of any sizes: item i owns blocks [block0_i, block0_{i+1}).
__device__ __forceinline__ int find_item(const cg::FftItem *items, int nitems, unsigned g)
{
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int KIND, int DIR>
__global__ void __launch_bounds__(cg::kThreads) k_fft_batch(const cg::FftItem *__restrict__ items, int nitems)
{
  const unsigned g = cg_xcd_remap(blockIdx.x, gridDim.x);
  const cg::FftItem &it = items[find_item(items, nitems, g)];
  const unsigned blk = g - it.block0;
  if (KIND == cg::FFT_ITEM_X_R2C)
    x_r2c_body(it.src, it.grid, it.nxp, it.rows, it.tile, it.plan, it.tw, it.twr, it.scale, blk);
  else if (KIND == cg::FFT_ITEM_X_C2R)
    x_c2r_body(it.grid, it.nxp, it.rows, it.tile, it.plan, it.tw, it.twr, it.scale, it.add, blk);
  else
    strided_body<DIR>(it.data, it.plan, it.S, it.inner, it.OS, it.tile, it.tiles_per_outer, it.tw, it.scale, blk);
}

constexpr size_t kLdsBudget = 80 * 1024;  // bytes per workgroup: two workgroups per CU (160 KB LDS)

}  // namespace cgfft

namespace cg {

using namespace cgfft;

bool make_plan(int n, FftPlan *plan)
{
  plan->n = n;
  plan->nstages = 0;
  int m = n;
  // largest radices first: fewer LDS round trips
  const int order[] = {8, 7, 5, 4, 3, 2};
  while (m > 1) {
    bool found = false;
    for (int p : order) {
      if (m % p == 0) {
        if (plan->nstages >= kMaxStages) return false;
        plan->radix[plan->nstages++] = p;
        m /= p;
        found = true;
        break;
      }
    }
    if (!found) return false;
  }
  return true;
}

int fft_x_r2c(cg_ctx *ctx, const float *src, float *grid, int nxp, long rows, float scale)
{
  const bool even = (nxp % 2) == 0;
  int rc = CG_OK;
  hipEvent_t ev;
  if (ctx->fft_engine == CG_FFT_AUTO) {
    launch_begin(ctx, CG_K_FFT_X, &ev);
    const int done = fixed_x(ctx, 1, src, grid, nxp, rows, scale, nullptr, &rc);
    launch_end(ctx, CG_K_FFT_X, ev);
    if (done) return rc;
  }
  FftPlan plan;
  const int n = even ? nxp / 2 : nxp;
  if (!make_plan(n, &plan)) return fail(ctx, CG_EINVAL, "x length has a prime factor above 7");
  const float2 *twf = twiddle_table(ctx, n);
  const float2 *twr = twiddle_table(ctx, nxp);
  if (!twf || !twr) return CG_ENOMEM;
  int rpw = (int)((kLdsBudget / sizeof(float2) - n) / (2 * (size_t)n));
  if (rpw < 1) rpw = 1;
  if (rpw > 16) rpw = 16;
  const size_t lds = sizeof(float2) * ((size_t)n + 2 * (size_t)rpw * n);
  const unsigned nblk = (unsigned)((rows + rpw - 1) / rpw);
  launch_begin(ctx, CG_K_FFT_X, &ev);
  k_fft_x_r2c<<<nblk, kThreads, lds, ctx->stream>>>(src, grid, nxp, rows, rpw, plan, twf, twr, scale);
  launch_end(ctx, CG_K_FFT_X, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fft_x_r2c");
}

int fft_x_c2r(cg_ctx *ctx, float *grid, int nxp, long rows, float scale, const float *add)
{
  const bool even = (nxp % 2) == 0;
  int rc = CG_OK;
  hipEvent_t ev;
  if (ctx->fft_engine == CG_FFT_AUTO) {
    launch_begin(ctx, CG_K_FFT_X, &ev);
    const int done = fixed_x(ctx, 0, nullptr, grid, nxp, rows, scale, add, &rc);
    launch_end(ctx, CG_K_FFT_X, ev);
    if (done) return rc;
  }
  FftPlan plan;
  const int n = even ? nxp / 2 : nxp;
  if (!make_plan(n, &plan)) return fail(ctx, CG_EINVAL, "x length has a prime factor above 7");
  const float2 *twf = twiddle_table(ctx, n);
  const float2 *twr = twiddle_table(ctx, nxp);
  if (!twf || !twr) return CG_ENOMEM;
  int rpw = (int)((kLdsBudget / sizeof(float2) - n) / (2 * (size_t)n));
  if (rpw < 1) rpw = 1;
  if (rpw > 16) rpw = 16;
  const size_t lds = sizeof(float2) * ((size_t)n + 2 * (size_t)rpw * n);
  const unsigned nblk = (unsigned)((rows + rpw - 1) / rpw);
  launch_begin(ctx, CG_K_FFT_X, &ev);
  k_fft_x_c2r<<<nblk, kThreads, lds, ctx->stream>>>(grid, nxp, rows, rpw, plan, twf, twr, scale, add);
  launch_end(ctx, CG_K_FFT_X, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fft_x_c2r");
}

int fft_strided(cg_ctx *ctx, float2 *data, int n, long S, long inner, long outer, long OS, int dir, float scale,
                int kclass)
{
  hipEvent_t ev;
  if (ctx->fft_engine == CG_FFT_AUTO) {
    int rc = CG_OK;
    launch_begin(ctx, kclass, &ev);
    const int done = fixed_strided(ctx, data, n, S, inner, outer, OS, dir, scale, &rc);
    launch_end(ctx, kclass, ev);
    if (done) return rc;
  }
  FftPlan plan;
  if (!make_plan(n, &plan)) return fail(ctx, CG_EINVAL, "axis length has a prime factor above 7");
  const float2 *twf = twiddle_table(ctx, n);
  if (!twf) return CG_ENOMEM;
  int TI = 16;
  while (TI > 1 && sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI) > kLdsBudget) TI /= 2;
  if (sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI) > 160 * 1024)
    return fail(ctx, CG_EINVAL, "axis length too long for the LDS engine");
  const size_t lds = sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI);
  const unsigned tpo = (unsigned)((inner + TI - 1) / TI);
  const unsigned long nblk = (unsigned long)tpo * (unsigned long)outer;
  if (nblk > 0x7fffffffUL) return fail(ctx, CG_EINVAL, "grid too large");
  launch_begin(ctx, kclass, &ev);
  if (dir < 0)
    k_fft_strided<-1><<<(unsigned)nblk, kThreads, lds, ctx->stream>>>(data, plan, S, inner, OS, TI, tpo, twf,
                                                                        scale);
  else
    k_fft_strided<+1><<<(unsigned)nblk, kThreads, lds, ctx->stream>>>(data, plan, S, inner, OS, TI, tpo, twf,
                                                                        scale);
  launch_end(ctx, kclass, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fft_strided");
}

int fft_strided(cg_ctx *ctx, float2 *data, int n, long S, long inner, long outer, long OS, int dir, float scale)
{
  return fft_strided(ctx, data, n, S, inner, outer, OS, dir, scale, CG_K_FFT_Y);
}

int plan_fft_item(cg_ctx *ctx, int kind, FftItem *it, size_t *lds)
{
  if (kind == FFT_ITEM_X_R2C || kind == FFT_ITEM_X_C2R) {
    const bool even = (it->nxp % 2) == 0;
    const int n = even ? it->nxp / 2 : it->nxp;
    if (!make_plan(n, &it->plan)) return fail(ctx, CG_EINVAL, "x length has a prime factor above 7");
    it->n = n;
    it->tw = twiddle_table(ctx, n);
    it->twr = twiddle_table(ctx, it->nxp);
    if (!it->tw || !it->twr) return CG_ENOMEM;
    int rpw = (int)((kLdsBudget / sizeof(float2) - n) / (2 * (size_t)n));
    rpw = rpw < 1 ? 1 : (rpw > 16 ? 16 : rpw);
    it->tile = rpw;
    it->nblocks = (unsigned)((it->rows + rpw - 1) / rpw);
    *lds = sizeof(float2) * ((size_t)n + 2 * (size_t)rpw * n);
    return CG_OK;
  }
  const int n = it->n;
  if (!make_plan(n, &it->plan)) return fail(ctx, CG_EINVAL, "axis length has a prime factor above 7");
  it->tw = twiddle_table(ctx, n);
  if (!it->tw) return CG_ENOMEM;
  int TI = 16;
  while (TI > 1 && sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI) > kLdsBudget) TI /= 2;
  if (sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI) > 160 * 1024)
    return fail(ctx, CG_EINVAL, "axis length too long for the LDS engine");
  it->tile = TI;
  it->tiles_per_outer = (unsigned)((it->inner + TI - 1) / TI);
  it->nblocks = it->tiles_per_outer * (unsigned)it->outer;
  *lds = sizeof(float2) * ((size_t)n + 2 * (size_t)n * TI);
  return CG_OK;
}

int launch_fft_batch(cg_ctx *ctx, int kind, int dir, const FftItem *dev_items, int nitems, unsigned total_blocks,
                     size_t lds, int kclass)
{
  if (nitems <= 0 || total_blocks == 0) return CG_OK;
  hipEvent_t ev;
  launch_begin(ctx, kclass, &ev);
  if (kind == FFT_ITEM_X_R2C)
    k_fft_batch<FFT_ITEM_X_R2C, -1><<<total_blocks, kThreads, lds, ctx->stream>>>(dev_items, nitems);
  else if (kind == FFT_ITEM_X_C2R)
    k_fft_batch<FFT_ITEM_X_C2R, +1><<<total_blocks, kThreads, lds, ctx->stream>>>(dev_items, nitems);
  else if (dir < 0)
    k_fft_batch<FFT_ITEM_STRIDED, -1><<<total_blocks, kThreads, lds, ctx->stream>>>(dev_items, nitems);
  else
    k_fft_batch<FFT_ITEM_STRIDED, +1><<<total_blocks, kThreads, lds, ctx->stream>>>(dev_items, nitems);
  launch_end(ctx, kclass, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fft_batch");
}

// FFTGrid::fftInPlace: pre-scale (fused into the x pass), x r2c, y, z.  With src != null the
// x pass reads the real cube from src (same layout) and leaves it unchanged.
int grid_fft_forward(cg_grid *g, const float *src)
{
  cg_ctx *ctx = g->ctx;
  const long N = (long)g->nxp * g->nyp * g->nzp;
  const float s = (g->cubetype == CG_CT_COVARIANCE) ? 1.0f : 1.0f / sqrtf((float)N);
  int rc = fft_x_r2c(ctx, src ? src : g->d, g->d, g->nxp, (long)g->nyp * g->nzp, s);
  if (rc) return rc;
  float2 *c = reinterpret_cast<float2 *>(g->d);
  const long cnxp = g->cnxp, plane = cnxp * g->nyp;
  if (g->nyp > 1) {
    rc = fft_strided(ctx, c, g->nyp, cnxp, cnxp, g->nzp, plane, -1, 1.0f, CG_K_FFT_Y);
    if (rc) return rc;
  }
  if (g->nzp > 1) {
    rc = fft_strided(ctx, c, g->nzp, plane, plane, 1, 0, -1, 1.0f, CG_K_FFT_Z);
    if (rc) return rc;
  }
  return CG_OK;
}

// FFTGrid::invFFTInPlace: z, y complex, x c2r with the scale (and an optional real add) fused.
static int grid_fft_inverse_impl(cg_grid *g, const float *add_real, bool with_z)
{
  cg_ctx *ctx = g->ctx;
  const size_t N = (size_t)g->nxp * g->nyp * g->nzp;
  const float s = (g->cubetype == CG_CT_COVARIANCE) ? float(1.0 / (double)N) : float(1.0 / sqrtf(float(N)));
  float2 *c = reinterpret_cast<float2 *>(g->d);
  const long cnxp = g->cnxp, plane = cnxp * g->nyp;
  int rc;
  if (with_z && g->nzp > 1) {
    rc = fft_strided(ctx, c, g->nzp, plane, plane, 1, 0, +1, 1.0f, CG_K_FFT_Z);
    if (rc) return rc;
  }
  if (g->nyp > 1) {
    rc = fft_strided(ctx, c, g->nyp, cnxp, cnxp, g->nzp, plane, +1, 1.0f, CG_K_FFT_Y);
    if (rc) return rc;
  }
  return fft_x_c2r(ctx, g->d, g->nxp, (long)g->nyp * g->nzp, s, add_real);
}

int grid_fft_inverse(cg_grid *g, const float *add_real) { return grid_fft_inverse_impl(g, add_real, true); }

int grid_fft_inverse_xy(cg_grid *g, const float *add_real) { return grid_fft_inverse_impl(g, add_real, false); }

}  // namespace cg
