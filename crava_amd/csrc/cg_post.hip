// cg_post.hip -- per-wavenumber Gaussian conditioning, one wavenumber per lane.
//
// Replaces the loop of AVOInversion::computePostMeanResidAndFFTCov (avoinversion.cpp:944-1083)
// with getNextParameterCovariance (seismicparametersholder.cpp:485-556), getNextErrorVariance
// (avoinversion.cpp:1236-1272) and the lib_matr kernels (lib_matr.c:78-167, 330-365, 524-736).
// For every wavenumber (i, j, k) with lowCut < f(k) < highCut:
//   M      = K Sigma_m K^H + Sigma_e                      (ntheta x ntheta Hermitian)
//   L      = chol(M)  [normalised by M00, TOL 1e-20, prior kept on failure]
//   X      = M^-1 K Sigma_m
//   Sigma' = Sigma_m - (K Sigma_m)^H X
//   m'     = m + X^H (d - K m),   residual = d - K m'
// Outside the band the prior passes through.  The Cholesky factor, right-hand sides and
// products stay in registers; no MFMA (nothing here is a dense contraction).
//
// Structure used for the separable prior Sigma_m(k) = r(k) Sigma0 (r = |rxy(kx,ky) rt(kz)|):
//   K Sigma_m K^H = r G(kz),   G = K Sigma0 K^H,   K Sigma_m = r S(kz),   S = K Sigma0,
//   Sigma_e = lambda Q(kz) + diag(D(kz)),  Q_lm = 0.5 (1-wnc) Sigma0e_lm (e1_l e1_m* + e2_l e2_m*),
//   D_l = wnc Sigma0e_ll |e3_l|^2,
// so each lane assembles M from three per-plane tables (computed once on the host in double).
#include <cmath>
#include <mutex>

#include "cg_internal.h"
#include "cg_mem.h"
#include "cg_post.h"

#include "cg_post_dev.h"

namespace cgpost {

using cg::gld;
using cg::gst;
using cg::kat;

// Loads the prior covariance entries of one wavenumber (getNextParameterCovariance:
// |Re(c_ab)/s_ab| * s_ab with zero imaginary part).
template <bool SEP>
__device__ __forceinline__ float prior6(const cg::PostParams &p, long idx, long c, int k, float (&P6)[6])
{
  float r = 0.f;
  if (SEP) {
    const float t = gld(p.rxy + c) * kat(p.rt, k);
    r = sqrtf(t * t);
#pragma unroll
    for (int g = 0; g < 6; ++g) P6[g] = r * p.s6f[g];
  } else {
#pragma unroll
    for (int g = 0; g < 6; ++g) {
      float t = gld(p.cov[g] + idx).x;
      if (p.s6f[g] != 0.f) t = t / p.s6f[g];
      P6[g] = sqrtf(t * t) * p.s6f[g];
    }
  }
  return r;
}

// FP32 pass.  grid: x over the cells of one k plane, y = k, so the per-plane tables are
// wavefront-uniform loads.  With MIXED, wavenumbers whose FP32 factor fails or whose pivot
// ratio exceeds cond_limit are appended to p.list untouched, for the FP64 pass.
template <int NT, bool SEP, bool MIXED>
__device__ __forceinline__ void post32_body(const cg::PostParams &p, int k, long c)
{
  const long plane = (long)p.cnxp * p.nyp;
  const bool active = c < plane;
  // the lateral tables rxy and lam2d are [nyp][cnxp]: cell c of the plane is their entry c
  const long cc = active ? c : 0;
  const long idx = (long)k * plane + cc;
  const bool inv = kat(p.invert, k) != 0;
  bool did_inv = false, did_fail = false, defer = false;
  if (active) {
    float P6[6];
    const float r = prior6<SEP>(p, idx, cc, k, P6);
    bool write_prior = !inv;
    if (inv) {
      did_inv = true;
      const float lt = p.lam_grid ? gld(p.lam_grid + idx).x : gld(p.lam2d + cc);
      const float lam = sqrtf(lt * lt);
      cpx<float> m3[3], d[NT], cov6[6], res[NT];
#pragma unroll
      for (int a = 0; a < 3; ++a) {
        const float2 v = gld(p.mean[a] + idx);
        m3[a] = mk<float>(v.x, v.y);
      }
#pragma unroll
      for (int l = 0; l < NT; ++l) {
        const float2 v = gld(p.data[l] + idx);
        d[l] = mk<float>(v.x, v.y);
      }
      float cond = 0.f;
      const int fail = update<NT, float, SEP>(kat(p.tab32, k), r, lam, P6, m3, d, cov6, res, p.want_residual != 0, cond);
      if (MIXED && (fail || cond > p.cond_limit)) {
        defer = true;
      } else if (fail) {  // posterior = prior (avoinversion.cpp:1040-1042)
        did_fail = true;
        write_prior = true;
      } else {
#pragma unroll
        for (int a = 0; a < 3; ++a) gst(p.mean[a] + idx, make_float2(m3[a].re, m3[a].im));
#pragma unroll
        for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2(cov6[g].re, cov6[g].im));
        if (p.want_residual) {
#pragma unroll
          for (int l = 0; l < NT; ++l) gst(p.data[l] + idx, make_float2(res[l].re, res[l].im));
        }
      }
    }
    if (write_prior) {
#pragma unroll
      for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2(P6[g], 0.f));
    }
  }
  if (MIXED && p.mask) {  // mark the deferred cell; one counter add per wavefront
    if (defer) atomicOr(p.mask + ((unsigned long long)idx >> 6), 1ull << ((unsigned long long)idx & 63u));
    const unsigned long long m = __ballot(defer);
    if (m && __lane_id() == 0) atomicAdd(p.list_len, (unsigned long long)__popcll(m));
  } else if (MIXED) {  // wave-aggregated append to the FP64 list
    const unsigned long long m = __ballot(defer);
    if (m) {
      unsigned long long base = 0;
      if (__lane_id() == 0) base = atomicAdd(p.list_len, (unsigned long long)__popcll(m));
      base = __shfl(base, 0);
      const unsigned lane = __lane_id();
      const unsigned long long below = lane ? (m & ((~0ull) >> (64 - lane))) : 0ull;
      if (defer) gst(p.list + (base + __popcll(below)), (unsigned)idx);
    }
  }
  if (p.stats) {
    wave_count(p.stats, did_inv);
    wave_count(p.stats + 64, did_fail);
  }
}

template <int NT, bool SEP, bool MIXED>
__global__ void __launch_bounds__(cg::kThreads) k_posterior(cg::PostParams p)
{
  post32_body<NT, SEP, MIXED>(p, blockIdx.y, (long)blockIdx.x * blockDim.x + threadIdx.x);
}

// FP64 re-solve of one deferred wavenumber (its inputs were left intact by the FP32 pass).
template <int NT, bool SEP>
__device__ __forceinline__ void post64_cell(const cg::PostParams &p, long idx)
{
  const long plane = (long)p.cnxp * p.nyp;
  const int k = (int)(idx / plane);
  const long c = idx % plane;
  float P6[6];
  prior6<SEP>(p, idx, c, k, P6);
  double P6d[6], rd = 0.0;
  if (SEP) {
    rd = fabs((double)gld(p.rxy + c) * (double)kat(p.rt, k));
#pragma unroll
    for (int g = 0; g < 6; ++g) P6d[g] = rd * p.s6d[g];
  } else {
#pragma unroll
    for (int g = 0; g < 6; ++g) P6d[g] = (double)P6[g];
  }
  const float lt = p.lam_grid ? gld(p.lam_grid + idx).x : gld(p.lam2d + c);
  const double lam = fabs((double)lt);
  cpx<double> m3[3], d[NT], cov6[6], res[NT];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const float2 v = gld(p.mean[a] + idx);
    m3[a] = mk<double>(v.x, v.y);
  }
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    const float2 v = gld(p.data[l] + idx);
    d[l] = mk<double>(v.x, v.y);
  }
  double cond = 0.0;
  const int fail = update<NT, double, SEP>(kat(p.tab64, k), rd, lam, P6d, m3, d, cov6, res, p.want_residual != 0, cond);
  if (fail) {
    if (p.stats) atomicAdd(p.stats + 64, 1ull);  // rare: the FP64 factor failed too
#pragma unroll
    for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2(P6[g], 0.f));
  } else {
#pragma unroll
    for (int a = 0; a < 3; ++a) gst(p.mean[a] + idx, make_float2((float)m3[a].re, (float)m3[a].im));
#pragma unroll
    for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2((float)cov6[g].re, (float)cov6[g].im));
    if (p.want_residual) {
#pragma unroll
      for (int l = 0; l < NT; ++l) gst(p.data[l] + idx, make_float2((float)res[l].re, (float)res[l].im));
    }
  }
}

// FP64 pass over the deferred wavenumbers.  Fixed grid, every thread strides over the list
// and stops at its length, so every wave exits.
template <int NT, bool SEP>
__device__ __forceinline__ void post64_body(const cg::PostParams &p, unsigned long long e0, unsigned long long stride)
{
  const unsigned long long n = *p.list_len;
  for (unsigned long long e = e0; e < n; e += stride) post64_cell<NT, SEP>(p, (long)gld(p.list + e));
}

// FP64 pass over the deferral bit mask: a fixed grid strides over its words (csize / 64, rounded
// up); each set bit is re-solved and its word cleared, so the mask is all zero again afterwards.
template <int NT, bool SEP>
__global__ void __launch_bounds__(cg::kThreads) k_posterior64_mask(cg::PostParams p)
{
  const unsigned long long nw = ((unsigned long long)p.csize + 63ull) >> 6;
  for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; w < nw;
       w += (unsigned long long)gridDim.x * blockDim.x) {
    unsigned long long bits = p.mask[w];
    if (!bits) continue;
    p.mask[w] = 0ull;
    while (bits) {
      const int b = __ffsll((long long)bits) - 1;
      bits &= bits - 1ull;
      post64_cell<NT, SEP>(p, (long)(w * 64ull + (unsigned long long)b));
    }
  }
}

__global__ void __launch_bounds__(cg::kThreads) k_zero_words(unsigned long long *d, unsigned long long n)
{
  for (unsigned long long w = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; w < n;
       w += (unsigned long long)gridDim.x * blockDim.x)
    d[w] = 0ull;
}

template <int NT, bool SEP>
__global__ void __launch_bounds__(cg::kThreads) k_posterior64(cg::PostParams p)
{
  post64_body<NT, SEP>(p, (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x,
                       (unsigned long long)gridDim.x * blockDim.x);
}

// Batched forms: many inversions of any size (same nt and prior form) in one launch.  Item i
// owns blocks [block0_i, block0_i + nzp_i * bpp_i); block b of an item covers plane b / bpp.
__device__ __forceinline__ int find_post_item(const cg::PostItem *items, int nitems, unsigned g)
{
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

template <int NT, bool SEP, bool MIXED>
__global__ void __launch_bounds__(cg::kThreads) k_posterior_batch(const cg::PostItem *__restrict__ items, int nitems)
{
  const cg::PostItem &it = items[find_post_item(items, nitems, blockIdx.x)];
  const unsigned local = blockIdx.x - it.block0;
  post32_body<NT, SEP, MIXED>(it.p, (int)(local / it.bpp), (long)(local % it.bpp) * blockDim.x + threadIdx.x);
}

template <int NT, bool SEP>
__global__ void __launch_bounds__(cg::kThreads) k_posterior64_batch(const cg::PostItem *__restrict__ items)
{
  post64_body<NT, SEP>(items[blockIdx.y].p, (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x,
                       (unsigned long long)gridDim.x * blockDim.x);
}

template <int NT>
void launch_batch_nt(const cg::PostItem *items, int n, unsigned blocks, hipStream_t s, bool sep, bool mixed)
{
  if (sep) {
    if (mixed) k_posterior_batch<NT, true, true><<<blocks, cg::kThreads, 0, s>>>(items, n);
    else k_posterior_batch<NT, true, false><<<blocks, cg::kThreads, 0, s>>>(items, n);
  } else {
    if (mixed) k_posterior_batch<NT, false, true><<<blocks, cg::kThreads, 0, s>>>(items, n);
    else k_posterior_batch<NT, false, false><<<blocks, cg::kThreads, 0, s>>>(items, n);
  }
  if (mixed) {
    dim3 g64(128, (unsigned)n);
    if (sep) k_posterior64_batch<NT, true><<<g64, cg::kThreads, 0, s>>>(items);
    else k_posterior64_batch<NT, false><<<g64, cg::kThreads, 0, s>>>(items);
  }
}

template <int NT>
void launch_nt(const cg::PostParams &p, dim3 grid, hipStream_t s, bool sep, bool mixed)
{
  if (sep) {
    if (mixed) k_posterior<NT, true, true><<<grid, cg::kThreads, 0, s>>>(p);
    else k_posterior<NT, true, false><<<grid, cg::kThreads, 0, s>>>(p);
  } else {
    if (mixed) k_posterior<NT, false, true><<<grid, cg::kThreads, 0, s>>>(p);
    else k_posterior<NT, false, false><<<grid, cg::kThreads, 0, s>>>(p);
  }
}

template <int NT>
void launch_nt64(const cg::PostParams &p, unsigned blocks, hipStream_t s, bool sep)
{
  if (p.mask) {
    if (sep) k_posterior64_mask<NT, true><<<blocks, cg::kThreads, 0, s>>>(p);
    else k_posterior64_mask<NT, false><<<blocks, cg::kThreads, 0, s>>>(p);
  } else {
    if (sep) k_posterior64<NT, true><<<blocks, cg::kThreads, 0, s>>>(p);
    else k_posterior64<NT, false><<<blocks, cg::kThreads, 0, s>>>(p);
  }
}

// The allocation-free posterior's buffer: device memory of fixed size on every device the library
// runs on, held until the process ends.  Host code never passes this name to a kernel (in host
// code it is only a stand-in for the symbol); it uses the address noalloc_scratch_init looked up.
static __device__ __align__(256) unsigned char scratch[cg::kNoAllocBytes];


}  // namespace cgpost

namespace cg {

using namespace cgpost;

int launch_posterior_batch(cg_ctx *ctx, int nt, bool sep, bool mixed, const PostItem *dev_items, int n,
                           unsigned blocks)
{
  if (n <= 0 || blocks == 0) return CG_OK;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_POSTERIOR, &ev);
  switch (nt) {
    case 1: launch_batch_nt<1>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 2: launch_batch_nt<2>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 3: launch_batch_nt<3>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 4: launch_batch_nt<4>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 5: launch_batch_nt<5>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 6: launch_batch_nt<6>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    case 7: launch_batch_nt<7>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
    default: launch_batch_nt<8>(dev_items, n, blocks, ctx->stream, sep, mixed); break;
  }
  launch_end(ctx, CG_K_POSTERIOR, ev);
  return hip_fail(ctx, hipGetLastError(), "k_posterior_batch");
}

namespace {
constexpr int kMaxDevices = 64;
std::mutex scratch_mu;
void *scratch_ptr[kMaxDevices] = {};
bool scratch_tried[kMaxDevices] = {};
}  // namespace

int noalloc_scratch_init(int device)
{
  if (device < 0 || device >= kMaxDevices) return CG_EINVAL;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_tried[device]) return scratch_ptr[device] ? CG_OK : CG_EHIP;
  scratch_tried[device] = true;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != device) return CG_EHIP;  // the caller makes it current
  void *ptr = nullptr;
  const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !ptr) return CG_EHIP;
  scratch_ptr[device] = ptr;
  return CG_OK;
}

void *noalloc_scratch(int device)
{
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_ptr[device];
}

int zero_words(cg_ctx *ctx, unsigned long long *d, size_t n)
{
  if (n == 0) return CG_OK;
  size_t b = (n + kThreads - 1) / kThreads;
  if (b > 4096) b = 4096;
  k_zero_words<<<(unsigned)b, kThreads, 0, ctx->stream>>>(d, (unsigned long long)n);
  return hip_fail(ctx, hipGetLastError(), "k_zero_words");
}

int launch_posterior(cg_ctx *ctx, const PostParams &p)
{
  const long plane = (long)p.cnxp * p.nyp;
  dim3 grid((unsigned)((plane + kThreads - 1) / kThreads), (unsigned)p.nzp);
  const bool sep = p.prior_mode == CG_PRIOR_SEPARABLE;
  const bool mixed = p.precision == CG_PREC_MIXED;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_POSTERIOR, &ev);
  switch (p.nt) {
    case 1: launch_nt<1>(p, grid, ctx->stream, sep, mixed); break;
    case 2: launch_nt<2>(p, grid, ctx->stream, sep, mixed); break;
    case 3: launch_nt<3>(p, grid, ctx->stream, sep, mixed); break;
    case 4: launch_nt<4>(p, grid, ctx->stream, sep, mixed); break;
    case 5: launch_nt<5>(p, grid, ctx->stream, sep, mixed); break;
    case 6: launch_nt<6>(p, grid, ctx->stream, sep, mixed); break;
    case 7: launch_nt<7>(p, grid, ctx->stream, sep, mixed); break;
    default: launch_nt<8>(p, grid, ctx->stream, sep, mixed); break;
  }
  int rc = hip_fail(ctx, hipGetLastError(), "k_posterior");
  if (rc == CG_OK && mixed) {
    const unsigned blocks = 1024;  // grid-stride over the deferred list
    switch (p.nt) {
      case 1: launch_nt64<1>(p, blocks, ctx->stream, sep); break;
      case 2: launch_nt64<2>(p, blocks, ctx->stream, sep); break;
      case 3: launch_nt64<3>(p, blocks, ctx->stream, sep); break;
      case 4: launch_nt64<4>(p, blocks, ctx->stream, sep); break;
      case 5: launch_nt64<5>(p, blocks, ctx->stream, sep); break;
      case 6: launch_nt64<6>(p, blocks, ctx->stream, sep); break;
      case 7: launch_nt64<7>(p, blocks, ctx->stream, sep); break;
      default: launch_nt64<8>(p, blocks, ctx->stream, sep); break;
    }
    rc = hip_fail(ctx, hipGetLastError(), "k_posterior64");
  }
  launch_end(ctx, CG_K_POSTERIOR, ev);
  return rc;
}

}  // namespace cg
