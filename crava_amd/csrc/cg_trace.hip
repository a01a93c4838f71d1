// cg_trace.hip -- per-trace filters along z of real grids: the data rescaling of the
// non-constant-thickness branch, AVOInversion::divideDataByScaleWavelet with
// computeAdjustmentFactor (avoinversion.cpp:485-719) and multiplyDataByScaleWaveletAndWriteToFile
// (:721-790).
//
// The reference transforms every trace on its own (nzp-point R2C, a complex factor per frequency,
// C2R).  Here the whole grid is transformed along z at once: the real grid viewed as complex
// columns holds two neighbouring traces per column (x = 2m in the real part, x = 2m + 1 in the
// imaginary part), so one complex z pass of the FFT engine transforms both.  The filter kernel then
// separates the two spectra at k and nzp - k, X_a = (Z(k) + conj Z(-k)) / 2 and
// X_b = (Z(k) - conj Z(-k)) / 2i, multiplies each by its trace's factor, and packs them back
// (Z'(k) = Y_a + i Y_b, Z'(-k) = conj Y_a + i conj Y_b; at k = 0 and nzp / 2 only the real parts,
// as the reference's half-complex inverse reads them); an inverse z pass returns the traces.
//
// Factors per (trace, k), in float as the reference computes them:
//  * local wavelet (Wavelet1D::shiftAndScale, wavelet1D.cpp:580-602): W_loc(m) = gain e^{-2 pi i m s / nzp} W(m),
//    s the local time shift in samples;
//  * its stretch by sf = relThick * localStretch (Wavelet::getCAmp(k, sf), wavelet.cpp:337-394): linear
//    interpolation of the half spectrum at k / sf, divided by sf;
//  * errorSmooth3 of the local wavelet: D3(m) W_loc(m), D3 the spectrum of the central, backward and
//    forward differences, normalised by its norm within the frequency band (= gain times the global
//    one, the phase having modulus one);
//  * divide: the adjustment factor of computeAdjustmentFactor (avoinversion.cpp:671-714) inside the
//    band lowCut < deltaF k < highCut (deltaF = nz 1000 / (relThick lz nzp)), zero outside;
//  * multiply: conj(W_loc stretched) (avoinversion.cpp:759-767) for box traces.
#include <cmath>

#include "cg_internal.h"
#include "cg_trace.h"

// named (not anonymous) so that profilers show the kernels' names
namespace cgtrace {

struct C2 {
  float re, im;
};

__device__ __forceinline__ C2 ld2(const float2 *p, int m)
{
  const float2 v = p[m];
  return C2{v.x, v.y};
}

// The local wavelet's spectrum at bin m (shiftAndScale: the multiplier rounded to float).
__device__ __forceinline__ C2 wloc_bin(const cg::ScaleWaveletParams &p, int m, float ishift, float gain)
{
  const C2 w = ld2(p.W, m);
  if (ishift == 0.0f) return C2{gain * w.re, gain * w.im};
  float sn, cs;
  // the angle 2 pi ishift m / nzp, as a fraction of a turn reduced to [-0.5, 0.5) before the
  // transcendental (the reference evaluates cos and sin of the angle itself in double)
  float turns = ishift * (float)m / (float)p.nzp;
  turns -= rintf(turns);
  sincospif(2.0f * turns, &sn, &cs);
  const float ar = gain * cs, ai = -gain * sn;
  return C2{ar * w.re - ai * w.im, ai * w.re + ar * w.im};
}

// errorSmooth3 of the local wavelet at bin m, normalised (multiplied by inv_norm).
__device__ __forceinline__ C2 sloc_bin(const cg::ScaleWaveletParams &p, int m, float ishift, float gain,
                                       float inv_norm)
{
  const C2 w = wloc_bin(p, m, ishift, gain);
  const C2 d = ld2(p.D3, m);
  return C2{(d.re * w.re - d.im * w.im) * inv_norm, (d.im * w.re + d.re * w.im) * inv_norm};
}

// Wavelet::getCAmp(k, scale) for k <= nzp / 2 of the spectrum given by bin(m).
template <class Bin>
__device__ __forceinline__ C2 stretched(int cnzp, int k, float scale, Bin bin)
{
  const float omega = (float)k / scale;
  if (omega >= (float)cnzp) return C2{0.0f, 0.0f};
  const int omL = (int)floorf(omega);
  int omU = omL + 1;
  if (omU >= cnzp) omU -= 1;
  const float d = omega - (float)omL;
  const C2 a = bin(omL), b = bin(omU);
  return C2{(a.re * (1.0f - d) + b.re * d) / scale, (a.im * (1.0f - d) + b.im * d) / scale};
}

// The reference's local-wavelet lookup of padded trace (i, j) in divideDataByScaleWavelet
// (avoinversion.cpp:516-533): indices mirrored into the box.
__device__ __forceinline__ int mirror_index(int i, int n, int np)
{
  int ii = i;
  if (ii > 3 * n - 1) ii = 0;
  if (ii > (np + n) / 2) ii = np - ii;
  if (ii >= n) ii = 2 * n - ii - 1;
  return ii;
}

// Factor of real trace x (< nxp), row j, frequency k (<= nzp / 2); `identity` if the trace is not
// filtered.
__device__ __forceinline__ C2 trace_factor(const cg::ScaleWaveletParams &p, int x, int j, int k, bool &identity)
{
  identity = false;
  const bool divide = p.mode == CG_SCALE_WAVELET_DIVIDE;
  int li = x, lj = j;
  if (divide) {
    li = mirror_index(x, p.nx, p.nxp);
    lj = mirror_index(j, p.ny, p.nyp);
  } else if (x >= p.nx || j >= p.ny) {
    identity = true;
    return C2{1.0f, 0.0f};
  }
  // Wavelet::getLocalTimeshift / getLocalGainFactor (wavelet.cpp:1251-1278): defaults outside the grids
  const bool in_box = li >= 0 && li < p.nx && lj >= 0 && lj < p.ny;
  const long lb = (long)lj * p.nx + li;
  const float ishift = (in_box && p.shift) ? p.shift[lb] / p.dz : 0.0f;
  const float gain = (in_box && p.gain) ? p.gain[lb] : 1.0f;
  const float lstretch = (in_box && p.stretch) ? p.stretch[lb] : 1.0f;
  const double rel = p.rel_thick[(long)j * p.nxp + x];
  const float sf = (float)(rel * (double)lstretch);
  const int cnzp = p.nzp / 2 + 1;
  const C2 wl = stretched(cnzp, k, sf, [&](int m) { return wloc_bin(p, m, ishift, gain); });
  if (!divide) return C2{wl.re, -wl.im};
  const double deltaF = (double)p.nz * 1000.0 / (rel * p.lz * (double)p.nzp);
  if (!(deltaF * k < p.high_cut && deltaF * k > p.low_cut)) return C2{0.0f, 0.0f};
  // computeAdjustmentFactor (avoinversion.cpp:671-714)
  const float inv_n3 = 1.0f / (gain * p.norm_f3);
  const float wn2 = p.wnorm * p.wnorm;
  const float wLoc2 = wl.re * wl.re + wl.im * wl.im;
  const C2 wg = ld2(p.W, k);
  const float wGlob2 = wg.re * wg.re + wg.im * wg.im;
  const C2 sl = stretched(cnzp, k, sf, [&](int m) { return sloc_bin(p, m, ishift, gain, inv_n3); });
  const float sizeLoc = wLoc2 / wn2 + (sl.re * sl.re + sl.im * sl.im);
  const C2 wu = wloc_bin(p, k, ishift, gain);
  const C2 su = sloc_bin(p, k, ishift, gain, inv_n3);
  const float sizeGlob = (wu.re * wu.re + wu.im * wu.im) / wn2 + (su.re * su.re + su.im * su.im);
  float sigma2Loc = p.error_var * (p.wnc + (1.0f - p.wnc) * 0.5f * sizeLoc);
  const float sigma2Glob = p.error_var * (p.wnc + (1.0f - p.wnc) * 0.5f * sizeGlob);
  const float tau2 = p.tau2[k];
  const float tol = 0.05f * p.modw;
  float wGlob2Adj = wGlob2, wLoc2Adj = wLoc2;
  if (wGlob2 < tol) wGlob2Adj = 0.5f * (wGlob2 + tol);
  if (wLoc2 < tol) wLoc2Adj = 0.5f * (wLoc2 + tol);
  if (wLoc2Adj < 0.05f * wGlob2Adj) {
    wLoc2Adj = 0.5f * (wLoc2 + 0.05f * wGlob2Adj);
    sigma2Loc *= wLoc2Adj / wLoc2;
  }
  const float den = tau2 * wLoc2Adj + sigma2Loc;
  const float glob = wGlob2Adj / (tau2 * wGlob2 + sigma2Glob);
  return C2{(wl.re / den) / glob, (wl.im / den) / glob};
}

__device__ __forceinline__ C2 cmul(C2 a, C2 b) { return C2{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }

// One packed column q (complex slot m of row j) and the frequency pair (k, nzp - k), k <= nzp / 2.
__global__ void __launch_bounds__(cg::kThreads) k_scale_wavelet(cg::ScaleWaveletParams p)
{
  const long plane = (long)p.cnxp * p.nyp;
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const int half = p.nzp / 2;
  if (t >= plane * (half + 1)) return;
  const int k = (int)(t / plane);
  const long q = t - (long)k * plane;
  const int m = (int)(q % p.cnxp), j = (int)(q / p.cnxp);
  const int kn = k == 0 ? 0 : p.nzp - k;
  float2 *col = p.data + q;
  const float2 z1 = col[(long)k * plane], z2 = col[(long)kn * plane];
  // the two traces' spectra at k: X_a = (Z(k) + conj Z(-k)) / 2, X_b = (Z(k) - conj Z(-k)) / 2i
  const C2 xa{0.5f * (z1.x + z2.x), 0.5f * (z1.y - z2.y)};
  const C2 xb{0.5f * (z1.y + z2.y), -0.5f * (z1.x - z2.x)};
  const int xa_i = 2 * m, xb_i = 2 * m + 1;
  bool ida = true, idb = true;
  C2 fa{1.0f, 0.0f}, fb{1.0f, 0.0f};
  if (xa_i < p.nxp) fa = trace_factor(p, xa_i, j, k, ida);
  if (xb_i < p.nxp) fb = trace_factor(p, xb_i, j, k, idb);
  C2 ya = ida ? xa : cmul(xa, fa);
  C2 yb = idb ? xb : cmul(xb, fb);
  const bool self = k == 0 || 2 * k == p.nzp;  // self-conjugate bins: the real parts (half-complex)
  if (self) {
    if (!ida) ya.im = 0.0f;
    if (!idb) yb.im = 0.0f;
  }
  // Z'(k) = Y_a + i Y_b, Z'(-k) = conj(Y_a) + i conj(Y_b)
  col[(long)k * plane] = make_float2(ya.re - yb.im, ya.im + yb.re);
  if (!self) col[(long)kn * plane] = make_float2(ya.re + yb.im, -ya.im + yb.re);
}

// divideDataByScaleWavelet's taper of the z padding: sample k > nz of every trace times
// max(1 - dist^2, 0) (avoinversion.cpp:540-545), dist = getDistToBoundary(k, nz, nzp).
__global__ void __launch_bounds__(cg::kThreads) k_taper_z(float *rv, long row_floats, int k0, int nk,
                                                          const float *taper)
{
  const long n = row_floats * nk;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long)gridDim.x * blockDim.x) {
    const int kk = (int)(t / row_floats);
    rv[(long)(k0 + kk) * row_floats + (t - (long)kk * row_floats)] *= taper[kk];
  }
}

}  // namespace cgtrace

namespace cg {

using namespace cgtrace;

int scale_wavelet_filter(cg_ctx *ctx, const ScaleWaveletParams &p, const float *taper_dev)
{
  const long plane = (long)p.cnxp * p.nyp;
  const float s = 1.0f / sqrtf((float)p.nzp);  // the reference's rData / sqrt(nzp), before and after
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  int rc = CG_OK;
  if (p.mode == CG_SCALE_WAVELET_DIVIDE && p.nzp > p.nz + 1) {
    const int nk = p.nzp - p.nz - 1;
    const long rowf = 2 * plane;
    size_t b = ((size_t)rowf * nk + kThreads - 1) / kThreads;
    if (b > 65536) b = 65536;
    k_taper_z<<<(unsigned)b, kThreads, 0, ctx->stream>>>(reinterpret_cast<float *>(p.data), rowf, p.nz + 1, nk,
                                                          taper_dev);
    rc = hip_fail(ctx, hipGetLastError(), "k_taper_z");
  }
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  if (rc) return rc;
  rc = fft_strided(ctx, p.data, p.nzp, plane, plane, 1, 0, -1, s, CG_K_FFT_Z);
  if (rc) return rc;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  const long n = plane * (p.nzp / 2 + 1);
  k_scale_wavelet<<<(unsigned)((n + kThreads - 1) / kThreads), kThreads, 0, ctx->stream>>>(p);
  rc = hip_fail(ctx, hipGetLastError(), "k_scale_wavelet");
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  if (rc) return rc;
  return fft_strided(ctx, p.data, p.nzp, plane, plane, 1, 0, +1, s, CG_K_FFT_Z);
}

}  // namespace cg

// ---------------------------------------------------------------- seismic traces onto the grid
namespace cgtrace {

// CommonData::SmoothTraceInGuardZone (commondata.cpp:6847-6885) and the zero padding of
// ResampleTrace (commondata.h:148-160): sample n of trace t, in float as the reference (the
// factors from double angles, then squared in float; the start factor first where both apply).
__global__ void __launch_bounds__(cg::kThreads) k_pack_traces(cg::TraceFillParams p)
{
#pragma clang fp contract(off)
  const long total = (long)p.ntraces * p.rnt;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int t = (int)(idx / p.rnt), n = (int)(idx - (long)t * p.rnt);
    const int len = p.length[t];
    float v = 0.f;
    if (n < len) {
      v = p.samples[(long)t * p.stride + n];
      if (n < p.ns) {
        const float s = (float)sin(1.5707963267948966 * ((double)n / (double)p.ns));
        v *= s * s;
      }
      const int ke = n - (len - p.ns);
      if (ke >= 0) {
        const float c = (float)cos(1.5707963267948966 * ((double)(ke + 1) / (double)p.ns));
        v *= c * c;
      }
    }
    p.a[idx] = v;
  }
}

// FillInData's preparation of a trace for a non-DATA grid (commondata.cpp:6612-6644), one
// workgroup per trace (at most 1024 samples, four per thread): where the first sample is zero,
// every zero sample is dropped (kept samples compacted by a workgroup scan) and, for SEG-Y traces,
// z0 moves down one dz_data per leading zero (added one at a time, as the reference does); the line through the
// first and last kept samples is subtracted, in float: s[m] - (first + m inc),
// inc = (last - first) / (n - 1); zero padding to rnt.  The trend's fine samples
// are first + l inc2, inc2 = (last - first) / (10 (n - 1)), l < 10 (n - 1) + 1 (:6684-6692).
constexpr int kDetrendThreads = 256;
__global__ void __launch_bounds__(kDetrendThreads) k_pack_detrend(cg::TraceFillParams p)
{
#pragma clang fp contract(off)
  __shared__ float kept[1024];
  __shared__ int scan[kDetrendThreads];
  __shared__ int lead;
  const int t = blockIdx.x, tid = threadIdx.x;
  const int len = p.length[t];
  const float *src = p.samples + (long)t * p.stride;
  const bool trim = src[0] == 0.0f;
  float v[4];
  int c = 0;
  if (tid == 0) lead = len;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int n = 4 * tid + q;
    v[q] = n < len ? src[n] : 0.0f;
    if (n < len && v[q] != 0.0f) atomicMin(&lead, n);
    c += (n < len && (!trim || v[q] != 0.0f)) ? 1 : 0;
  }
  scan[tid] = c;
  __syncthreads();
  for (int d = 1; d < kDetrendThreads; d <<= 1) {  // inclusive scan of the kept counts
    const int add = tid >= d ? scan[tid - d] : 0;
    __syncthreads();
    scan[tid] += add;
    __syncthreads();
  }
  int pos = scan[tid] - c;
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int n = 4 * tid + q;
    if (n < len && (!trim || v[q] != 0.0f)) kept[pos++] = v[q];
  }
  __syncthreads();
  const int nk = scan[kDetrendThreads - 1];
  const float first = nk > 0 ? kept[0] : 0.0f, last = nk > 0 ? kept[nk - 1] : 0.0f;
  const float inc = (last - first) / (float)(nk - 1);
  float *dst = p.a + (long)t * p.rnt;
  for (int m = tid; m < p.rnt; m += kDetrendThreads) dst[m] = m < nk ? kept[m] - (first + (float)m * inc) : 0.0f;
  if (tid == 0) {
    float z0 = p.z0_data[t];
    if (trim && p.trim_shift)
      for (int i = 0; i < lead; ++i) z0 = z0 + p.dz_data;
    cg::TraceInfo ti;
    ti.z0 = z0;
    ti.first = first;
    ti.inc2 = (last - first) / (10.0f * (float)(nk - 1));
    ti.n_trend = nk > 0 ? (int)(10.0f * (float)(nk - 1) + 1.0f) : 0;
    p.info[t] = ti;
  }
}

// The cnt = rnt / 2 bins of each trace's spectrum as the lowest bins of a zero spectrum of
// cmt = rmt / 2 bins (ResampleTrace, commondata.h:166-176).
__global__ void __launch_bounds__(cg::kThreads) k_spread_spectra(cg::TraceFillParams p)
{
  const long total = (long)p.ntraces * p.rmt;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int t = (int)(idx / p.rmt), f = (int)(idx - (long)t * p.rmt);
    p.b[idx] = f < p.rnt ? p.a[(long)t * p.rnt + f] : 0.f;
  }
}

constexpr float kMissing = -99999.0f;  // RMISSING

// CommonData::InterpolateGridValues (commondata.cpp:6947-6995) with GetZSimboxIndex (:7052-7062)
// and SetTrace (:7084-7092): real cell (i, j, k) of the padded grid from the fine trace of its
// column's trace (zero for columns without one), in the reference's float arithmetic.  Non-DATA
// grids add the trend interpolated by InterpolateAndShiftTrend (:6997-7050: its end values
// outside it) and take RMISSING in dead columns (:6731-6742).
__global__ void __launch_bounds__(cg::kThreads) k_fill_columns(cg::TraceFillParams p)
{
#pragma clang fp contract(off)
  const long plane = (long)p.nyp * p.rnxp;
  const long total = plane * p.nzp;
  const float inv = 1.0f / p.dz_fine;
  for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += (long)gridDim.x * blockDim.x) {
    const int k = (int)(idx / plane);
    const long col = idx - (long)k * plane;
    const int t = p.column_trace[col];
    float v = 0.f;
    cg::TraceInfo ti{0.0f, 0.0f, 0.0f, 1};
    if (p.detrend && t >= 0) ti = p.info[t];
    if (p.detrend && (t == -2 || (t >= 0 && ti.n_trend == 0))) {
      v = kMissing;
    } else if (t >= 0) {
      const float z0_shift = p.z0_grid[col] - (p.detrend ? ti.z0 : p.z0_data[t]);
      const int refk = k < (p.nz + p.nzp) / 2 ? k : k - p.nzp;
      const float dl = (z0_shift + (float)refk * p.dz_grid[col]) * inv;
      const float fl = floorf(dl), cl = ceilf(dl);
      const int l1 = (int)fl, l2 = (int)cl;
      const float *fine = p.b + (long)t * p.rmt;
      const int n_fine = p.rmt;
      if (l2 < 0 || l1 > n_fine - 1) {
        v = 0.f;
      } else if (l1 < 0) {
        v = fine[l2];
      } else if (l2 > n_fine - 1 || l1 == l2) {
        v = fine[l1];
      } else {
        const float w1 = cl - dl, w2 = dl - fl;
        v = w1 * fine[l1] + w2 * fine[l2];
      }
      if (p.detrend) {
        const int nt = ti.n_trend;
        const auto trend = [&](int l) { return ti.first + (float)l * ti.inc2; };
        float tr;
        if (l2 < 0) tr = trend(0);
        else if (l1 > nt - 1) tr = trend(nt - 1);
        else if (l1 < 0) tr = trend(l2);
        else if (l2 > nt - 1 || l1 == l2) tr = trend(l1);
        else tr = (cl - dl) * trend(l1) + (dl - fl) * trend(l2);
        v = v + tr;
      }
    }
    p.grid[idx] = v;
  }
}

}  // namespace cgtrace

namespace cg {

static unsigned grid_stride_blocks(long total)
{
  long b = (total + kThreads - 1) / kThreads;
  return (unsigned)(b < 8192 ? (b > 0 ? b : 1) : 8192);
}

int pack_traces(cg_ctx *ctx, const TraceFillParams &p)
{
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgtrace::k_pack_traces<<<grid_stride_blocks((long)p.ntraces * p.rnt), kThreads, 0, ctx->stream>>>(p);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_pack_traces");
}

int pack_detrend(cg_ctx *ctx, const TraceFillParams &p)
{
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgtrace::k_pack_detrend<<<(unsigned)p.ntraces, cgtrace::kDetrendThreads, 0, ctx->stream>>>(p);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_pack_detrend");
}

int spread_spectra(cg_ctx *ctx, const TraceFillParams &p)
{
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgtrace::k_spread_spectra<<<grid_stride_blocks((long)p.ntraces * p.rmt), kThreads, 0, ctx->stream>>>(p);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_spread_spectra");
}

int fill_columns(cg_ctx *ctx, const TraceFillParams &p)
{
  hipEvent_t ev;
  launch_begin(ctx, CG_K_ELEMENTWISE, &ev);
  cgtrace::k_fill_columns<<<grid_stride_blocks((long)p.nyp * p.rnxp * p.nzp), kThreads, 0, ctx->stream>>>(p);
  launch_end(ctx, CG_K_ELEMENTWISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_fill_columns");
}

}  // namespace cg
