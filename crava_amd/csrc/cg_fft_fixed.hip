// cg_fft_fixed.hip -- FFT passes with compile-time radix plans for the common padded lengths.
//
// Same passes and conventions as cg_fft.hip (FFTGrid::fftInPlace / invFFTInPlace,
// fftgrid.cpp:1587-1645; FFTW pass order rfftwnd.c:207-288), but the transform length, the
// radix sequence and the thread split are template parameters:
//  * every sequence of length N is worked on by TPS threads holding N/TPS elements each in
//    registers; a stage reads all its butterfly inputs from LDS, waits at a barrier, and
//    writes its outputs back in place (Stockham autosort), so one LDS buffer suffices;
//  * index arithmetic (j mod Ns, j / Ns, strides) folds to shifts and multiplies;
//  * tiles: y/z passes take TI (16 or 8) consecutive columns over the whole axis (128 or 64 B
//    row segments); x passes take RW rows, LDS row pitch N+1 to spread banks.
// Lengths without a plan here use the runtime-plan kernels of cg_fft.hip.
#include <cmath>

#include "cg_dft.h"
#include "cg_internal.h"

namespace {

using namespace cgdft;

template <int N, int DIR, int TPS, int ES, int NS, int... RS>
struct Stages;

template <int N, int DIR, int TPS, int ES, int NS>
struct Stages<N, DIR, TPS, ES, NS> {
  static __device__ __forceinline__ void run(float2 *, int, const float2 *) {}
};

template <int N, int DIR, int TPS, int ES, int NS, int P, int... RS>
struct Stages<N, DIR, TPS, ES, NS, P, RS...> {
  static __device__ __forceinline__ void run(float2 *buf, int t, const float2 *tw)
  {
    constexpr int NB = N / P;
    constexpr int BPT = NB / TPS;
    static_assert(NB % TPS == 0, "threads per sequence must divide every stage");
    constexpr int TWS = N / (NS * P);
    float2 v[BPT][P];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
#pragma unroll
      for (int q = 0; q < P; ++q) v[b][q] = buf[(j + q * NB) * ES];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int q = 1; q < P; ++q) {
          float2 w = tw[q * k * TWS];
          if (DIR > 0) w.y = -w.y;
          v[b][q] = cmul(v[b][q], w);
        }
      }
      dft<DIR, P>(v[b]);
      const int d = (j / NS) * NS * P + k;
#pragma unroll
      for (int q = 0; q < P; ++q) buf[(d + q * NS) * ES] = v[b][q];
    }
    __syncthreads();
    Stages<N, DIR, TPS, ES, NS * P, RS...>::run(buf, t, tw);
  }
};

// Radix plans: TPS divides N/p for every radix p of the plan.
template <int N>
struct Plan;
#define CG_PLAN(n, tps, ...)                                                      \
  template <>                                                                     \
  struct Plan<n> {                                                                \
    static constexpr int TPS = tps;                                               \
    template <int DIR, int ES>                                                    \
    static __device__ __forceinline__ void run(float2 *b, int t, const float2 *w) \
    {                                                                             \
      Stages<n, DIR, tps, ES, 1, __VA_ARGS__>::run(b, t, w);                      \
    }                                                                             \
  };
CG_PLAN(160, 8, 4, 4, 2, 5)
CG_PLAN(192, 8, 8, 8, 3)
CG_PLAN(256, 16, 4, 4, 4, 4)
CG_PLAN(288, 12, 8, 4, 3, 3)
CG_PLAN(320, 16, 4, 4, 4, 5)
CG_PLAN(384, 16, 8, 8, 2, 3)
CG_PLAN(512, 32, 8, 4, 4, 4)
CG_PLAN(576, 24, 8, 8, 3, 3)
CG_PLAN(640, 16, 8, 8, 2, 5)
CG_PLAN(768, 32, 8, 8, 4, 3)
CG_PLAN(1152, 48, 8, 8, 2, 3, 3)
#undef CG_PLAN

__device__ __forceinline__ void load_tw(float2 *tw, const float2 *g, int n, int tid, int nthr)
{
  for (int m = tid; m < n; m += nthr) tw[m] = g[m];
}

// ---------------------------------------------------------------- y / z passes
template <int N, int DIR, int TI>
__global__ void __launch_bounds__(TI *Plan<N>::TPS)
k_ffts(float2 *__restrict__ data, long S, long inner, long OS, unsigned tiles_per_outer,
       const float2 *__restrict__ twg, float scale)
{
  constexpr int NT = TI * Plan<N>::TPS;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  const int tid = threadIdx.x;
  const unsigned blk = cg_xcd_remap(blockIdx.x, gridDim.x);
  const unsigned o = blk / tiles_per_outer;
  const long c0 = (long)(blk % tiles_per_outer) * TI;
  float2 *base = data + (long)o * OS + c0;
  const int ncol = (int)((inner - c0) < TI ? (inner - c0) : TI);
  load_tw(tw, twg, N, tid, NT);
  if (ncol == TI) {
    // all loads issued before the first LDS store: N / TPS requests in flight per thread
    constexpr int PER = N * TI / NT;
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int idx = tid + i * NT;
      v[i] = base[(long)(idx / TI) * S + (idx % TI)];
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) buf[tid + i * NT] = v[i];
  } else {
    for (int idx = tid; idx < N * TI; idx += NT) {
      const int c = idx % TI;
      buf[idx] = c < ncol ? base[(long)(idx / TI) * S + c] : make_float2(0.f, 0.f);
    }
  }
  __syncthreads();
  Plan<N>::template run<DIR, TI>(buf + (tid % TI), tid / TI, tw);
  if (ncol == TI) {
#pragma unroll
    for (int i = 0; i < N * TI / NT; ++i) {
      const int idx = tid + i * NT;
      base[(long)(idx / TI) * S + (idx % TI)] = cscale(buf[idx], scale);
    }
  } else {
    for (int idx = tid; idx < N * TI; idx += NT) {
      const int c = idx % TI;
      if (c < ncol) base[(long)(idx / TI) * S + c] = cscale(buf[idx], scale);
    }
  }
}

// ---------------------------------------------------------------- x pass, even nxp = 2 N
// A block takes RW consecutive rows.  Each row holds N + 1 complex slots (the in-place layout:
// 2(N + 1) floats), so the block's rows are one contiguous span of global memory, copied in and
// out flat (coalesced).  In LDS the rows are interleaved, slot m of row r at m * ES + r with odd
// ES = RW + 1: the FFT stages then address consecutive rows from consecutive lanes (no bank
// conflicts), and the flat copies stride by ES (odd, so also conflict-free).
//
// Forward: z[m] = x[2m] + i x[2m+1] (scaled), Z = FFT_N(z),
//   X[k] = (Z[k] + conj Z[N-k]) / 2 - i w^k (Z[k] - conj Z[N-k]) / 2,  w = exp(-2 pi i / 2N);
// bins k and N - k are formed from the same pair of slots, so each work item rewrites its own
// pair in place.
template <int N, int RW>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftx_r2c(const float *src, float *dst, long rows, const float2 *__restrict__ twg,
           const float2 *__restrict__ tw_row, float scale)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = N + 1;  // complex slots per row
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  const int tid = threadIdx.x;
  const long row0 = (long)cg_xcd_remap(blockIdx.x, gridDim.x) * RW;
  const int nr = (int)((rows - row0) < RW ? (rows - row0) : RW);
  const float2 *in = reinterpret_cast<const float2 *>(src) + row0 * CN;
  float2 *out = reinterpret_cast<float2 *>(dst) + row0 * CN;
  load_tw(tw, twg, N, tid, NT);
  constexpr int PER = (RW * CN + NT - 1) / NT;
  float2 v[PER];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    v[i] = f < nr * CN ? in[f] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < RW * CN) buf[m * ES + r] = cscale(v[i], scale);
  }
  __syncthreads();
  Plan<N>::template run<-1, ES>(buf + tid % RW, tid / RW, tw);
  for (int idx = tid; idx < RW * (N / 2 + 1); idx += NT) {
    const int r = idx % RW, k = idx / RW;
    const int kn = k == 0 ? 0 : N - k;
    const float2 a = buf[k * ES + r], b = buf[kn * ES + r];
    // X[k] from (Z[k], conj Z[N-k]); X[N-k] from (Z[N-k], conj Z[k])
    const float2 ea = cscale(cadd(a, cconj(b)), 0.5f), da = csub(a, cconj(b));
    const float2 eb = cscale(cadd(b, cconj(a)), 0.5f), db = csub(b, cconj(a));
    const float2 xa = cadd(ea, cmul(tw_row[k], make_float2(0.5f * da.y, -0.5f * da.x)));
    const float2 xb = cadd(eb, cmul(tw_row[N - k], make_float2(0.5f * db.y, -0.5f * db.x)));
    buf[k * ES + r] = xa;
    buf[(N - k) * ES + r] = xb;  // k = 0 writes bin N into the spare slot
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) out[f] = buf[m * ES + r];
  }
}

// Inverse: Z[k] = (X[k] + conj X[N-k]) + i w^-k (X[k] - conj X[N-k]) with Im X[0] = Im X[N] = 0,
// z = IFFT_N(z), x[2m] = Re z[m], x[2m+1] = Im z[m]; then x scale (+ add); the row's spare slot
// is written as zero.
template <int N, int RW>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftx_c2r(float *grid, long rows, const float2 *__restrict__ twg, const float2 *__restrict__ tw_row, float scale,
           const float *__restrict__ add)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = N + 1;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  const int tid = threadIdx.x;
  const long row0 = (long)cg_xcd_remap(blockIdx.x, gridDim.x) * RW;
  const int nr = (int)((rows - row0) < RW ? (rows - row0) : RW);
  float2 *io = reinterpret_cast<float2 *>(grid) + row0 * CN;
  const float2 *ad = add ? reinterpret_cast<const float2 *>(add) + row0 * CN : nullptr;
  load_tw(tw, twg, N, tid, NT);
  constexpr int PER = (RW * CN + NT - 1) / NT;
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[i] = f < nr * CN ? io[f] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < RW * CN) buf[m * ES + r] = v[i];
    }
  }
  __syncthreads();
  for (int idx = tid; idx < RW * (N / 2 + 1); idx += NT) {
    const int r = idx % RW, k = idx / RW;
    float2 a = buf[k * ES + r], b = buf[(N - k) * ES + r];  // X[k], X[N-k]
    if (k == 0) {
      a.y = 0.f;
      b.y = 0.f;
    }
    const float2 ea = cadd(a, cconj(b)), oa = cmul(csub(a, cconj(b)), cconj(tw_row[k]));
    buf[k * ES + r] = make_float2(ea.x - oa.y, ea.y + oa.x);
    if (k > 0 && k < N - k) {
      const float2 eb = cadd(b, cconj(a)), ob = cmul(csub(b, cconj(a)), cconj(tw_row[N - k]));
      buf[(N - k) * ES + r] = make_float2(eb.x - ob.y, eb.y + ob.x);
    }
  }
  __syncthreads();
  Plan<N>::template run<+1, ES>(buf + tid % RW, tid / RW, tw);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) {
      float2 v = make_float2(0.f, 0.f);
      if (m < N) {
        v = cscale(buf[m * ES + r], scale);
        if (ad) {
          const float2 q = ad[f];
          v.x += q.x;
          v.y += q.y;
        }
      }
      io[f] = v;
    }
  }
}

template <int N, int DIR, int TI>
int launch_strided(cg_ctx *ctx, float2 *data, long S, long inner, long outer, long OS, float scale, const float2 *tw)
{
  const unsigned tpo = (unsigned)((inner + TI - 1) / TI);
  const unsigned long nblk = (unsigned long)tpo * (unsigned long)outer;
  const size_t lds = sizeof(float2) * (N + (size_t)N * TI);
  k_ffts<N, DIR, TI><<<(unsigned)nblk, TI * Plan<N>::TPS, lds, ctx->stream>>>(data, S, inner, OS, tpo, tw, scale);
  return cg::hip_fail(ctx, hipGetLastError(), "k_ffts");
}

template <int N, int DIR>
int strided_n(cg_ctx *ctx, float2 *data, long S, long inner, long outer, long OS, float scale, const float2 *tw)
{
  // 16 columns (128 B segments) while the tile fits 80 KB, else 8
  if ((size_t)N * 16 * sizeof(float2) <= 80 * 1024 && 16 * Plan<N>::TPS <= 1024)
    return launch_strided<N, DIR, 16>(ctx, data, S, inner, outer, OS, scale, tw);
  return launch_strided<N, DIR, 8>(ctx, data, S, inner, outer, OS, scale, tw);
}

// Rows per block: the most (32, 16, 8) whose LDS (twiddles + (N + 1) (RW + 1) slots) fits 80 KB.
template <int N>
constexpr bool x_fits(int rw)
{
  return (N + (N + 1) * (rw + 1)) * 8 <= 80 * 1024 && rw * Plan<N>::TPS <= 512;
}
template <int N>
constexpr int rows_per_wg()
{
  return x_fits<N>(32) ? 32 : x_fits<N>(16) ? 16 : 8;
}

template <int N>
int x_r2c_n(cg_ctx *ctx, const float *src, float *dst, long rows, float scale, const float2 *tw, const float2 *twr)
{
  constexpr int RW = rows_per_wg<N>();
  const size_t lds = sizeof(float2) * (N + (size_t)(N + 1) * (RW + 1));
  const unsigned nblk = (unsigned)((rows + RW - 1) / RW);
  k_fftx_r2c<N, RW><<<nblk, RW * Plan<N>::TPS, lds, ctx->stream>>>(src, dst, rows, tw, twr, scale);
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftx_r2c");
}

template <int N>
int x_c2r_n(cg_ctx *ctx, float *grid, long rows, float scale, const float *add, const float2 *tw, const float2 *twr)
{
  constexpr int RW = rows_per_wg<N>();
  const size_t lds = sizeof(float2) * (N + (size_t)(N + 1) * (RW + 1));
  const unsigned nblk = (unsigned)((rows + RW - 1) / RW);
  k_fftx_c2r<N, RW><<<nblk, RW * Plan<N>::TPS, lds, ctx->stream>>>(grid, rows, tw, twr, scale, add);
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftx_c2r");
}

#define CG_FIXED_LENGTHS(X) X(160) X(192) X(256) X(288) X(320) X(384) X(512) X(576) X(640) X(768) X(1152)

}  // namespace

namespace cg {

// Returns 1 if a fixed plan handled the pass (rc in *rc), 0 to fall back to the generic engine.
int fixed_strided(cg_ctx *ctx, float2 *data, int n, long S, long inner, long outer, long OS, int dir, float scale,
                  int *rc)
{
  const float2 *tw = nullptr;
  switch (n) {
#define X(L) case L:
    CG_FIXED_LENGTHS(X)
#undef X
    tw = twiddle_table(ctx, n);
    if (!tw) {
      *rc = CG_ENOMEM;
      return 1;
    }
    break;
    default: return 0;
  }
  switch (n) {
#define X(L)                                                                                    \
  case L:                                                                                       \
    *rc = dir < 0 ? strided_n<L, -1>(ctx, data, S, inner, outer, OS, scale, tw)                \
                  : strided_n<L, +1>(ctx, data, S, inner, outer, OS, scale, tw);               \
    return 1;
    CG_FIXED_LENGTHS(X)
#undef X
  }
  return 0;
}

int fixed_x(cg_ctx *ctx, int forward, const float *src, float *grid, int nxp, long rows, float scale, const float *add,
            int *rc)
{
  if (nxp % 2) return 0;
  const int n = nxp / 2;
  switch (n) {
#define X(L) case L:
    CG_FIXED_LENGTHS(X)
#undef X
    break;
    default: return 0;
  }
  const float2 *tw = twiddle_table(ctx, n);
  const float2 *twr = twiddle_table(ctx, nxp);
  if (!tw || !twr) {
    *rc = CG_ENOMEM;
    return 1;
  }
  switch (n) {
#define X(L)                                                                                    \
  case L:                                                                                       \
    *rc = forward ? x_r2c_n<L>(ctx, src, grid, rows, scale, tw, twr)                           \
                  : x_c2r_n<L>(ctx, grid, rows, scale, add, tw, twr);                          \
    return 1;
    CG_FIXED_LENGTHS(X)
#undef X
  }
  return 0;
}

}  // namespace cg
