// cg_fft_fixed.h -- FFT passes with compile-time radix plans (templates; instantiated per length
// by cg_fft_fixed_part.hip, dispatched by cg_fft_fixed.hip).
//
// Same passes and conventions as cg_fft.hip (FFTGrid::fftInPlace / invFFTInPlace,
// fftgrid.cpp:1587-1645; FFTW pass order rfftwnd.c:207-288), but the transform length, the
// radix sequence and the thread split are template parameters:
//  * every sequence of length N is worked on by TPS threads holding the butterfly inputs of a
//    stage in registers; a stage reads all its inputs from LDS, waits at a barrier, and writes
//    its outputs back in place (Stockham autosort), so one LDS buffer suffices;
//  * index arithmetic (j mod Ns, j / Ns, strides) folds to shifts and multiplies;
//  * every load of a tile is issued before the first LDS store (N / TPS requests in flight per
//    thread), which is what keeps HBM busy;
//  * one kernel serves one grid or a list of grids of the same length (descriptor per grid,
//    blocks [block0_i, block0_{i+1}) belong to grid i), so a batch of cubes of mixed sizes
//    needs one launch per (pass, length).
// The plans (n, TPS, radices) come from cg_fft_plans.h.
#pragma once

#include "cg_colour.h"
#include "cg_dft.h"
#include "cg_fft_plans.h"
#include "cg_internal.h"

namespace cgfix {

using namespace cgdft;

// Workgroup barrier for the LDS exchanges of these kernels: each wave's LDS operations are
// complete (lgkmcnt 0) before it, but outstanding global loads and stores are not waited for
// (__syncthreads also waits for them), so stores and prefetches into LDS stay in flight across
// the FFT stages.  No kernel here passes data between threads through global memory.
__device__ __forceinline__ void lds_sync()
{
  __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt and expcnt at their maximum (no wait)
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

// Waits for every outstanding global load and store of the wave (vmcnt 0), including loads
// into LDS.
__device__ __forceinline__ void vm_wait()
{
  __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0); expcnt and lgkmcnt at their maximum
  asm volatile("" ::: "memory");
}

// ---------------------------------------------------------------- radix stages
template <int N, int DIR, int TPS, int ES, int NS, int... RS>
struct Stages;

template <int N, int DIR, int TPS, int ES, int NS>
struct Stages<N, DIR, TPS, ES, NS> {
  static __device__ __forceinline__ void run(float2 *, int, const float2 *) {}
};

// Element e of the sequence lives at buf[e * ES].  Butterfly j (j < N / P) reads elements
// j + q N/P, applies w_{Ns P}^{q k} (k = j mod Ns) and writes (j / Ns) Ns P + k + q Ns.
template <int N, int DIR, int TPS, int ES, int NS, int P, int... RS>
struct Stages<N, DIR, TPS, ES, NS, P, RS...> {
  static __device__ __forceinline__ void run(float2 *buf, int t, const float2 *tw)
  {
    constexpr int NB = N / P;
    constexpr int BPT = (NB + TPS - 1) / TPS;
    constexpr bool EVEN = NB % TPS == 0;  // otherwise the last butterfly of some threads is idle
    constexpr int TWS = N / (NS * P);
    float2 v[BPT][P];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
      if (EVEN || j < NB) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[b][q] = buf[(j + q * NB) * ES];
      }
    }
    lds_sync();
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
      if (EVEN || j < NB) {
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
    }
    lds_sync();
    Stages<N, DIR, TPS, ES, NS * P, RS...>::run(buf, t, tw);
  }
};

// The same stages with the first stage's inputs taken from registers and the last stage's
// outputs handed to a sink instead of LDS (the y/z passes load the first stage's butterfly
// inputs straight from global memory and store the last stage's outputs straight back, which
// saves two LDS round trips and two barriers per tile).
//   First<P1>::Regs<BPT1>: v[b][q] = element (t + b TPS) + q N/P1 of the sequence.
//   sink(e, value) receives element e of the transformed sequence.
template <int N, int DIR, int TPS, int ES, int NS, int... RS>
struct StagesIO;

template <int N, int DIR, int TPS, int ES, int NS, int P, int... RS>
struct StagesIO<N, DIR, TPS, ES, NS, P, RS...> {
  static constexpr int P_ = P;
  static constexpr int NB = N / P;
  static constexpr int BPT = (NB + TPS - 1) / TPS;
  static constexpr bool EVEN = NB % TPS == 0;
  static constexpr bool LAST = sizeof...(RS) == 0;

  // one stage from values already in registers (v), outputs to LDS or, last, to the sink
  template <class Sink>
  static __device__ __forceinline__ void step(float2 (&v)[BPT][P], float2 *buf, int t, const float2 *tw, Sink &sink)
  {
    constexpr int TWS = N / (NS * P);
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
      if (EVEN || j < NB) {
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
        for (int q = 0; q < P; ++q) {
          if constexpr (LAST) sink(d + q * NS, v[b][q]);
          else buf[(d + q * NS) * ES] = v[b][q];
        }
      }
    }
    if constexpr (!LAST) {
      lds_sync();
      StagesIO<N, DIR, TPS, ES, NS * P, RS...>::from_lds(buf, t, tw, sink);
    }
  }

  template <class Sink>
  static __device__ __forceinline__ void from_lds(float2 *buf, int t, const float2 *tw, Sink &sink)
  {
    float2 v[BPT][P];
#pragma unroll
    for (int b = 0; b < BPT; ++b) {
      const int j = t + b * TPS;
      if (EVEN || j < NB) {
#pragma unroll
        for (int q = 0; q < P; ++q) v[b][q] = buf[(j + q * NB) * ES];
      }
    }
    lds_sync();
    step(v, buf, t, tw, sink);
  }
};

template <int N>
struct Plan;
#define CG_PLAN(n, tps, ...)                                                                    \
  template <>                                                                                   \
  struct Plan<n> {                                                                              \
    static constexpr int TPS = tps;                                                             \
    template <int DIR, int ES>                                                                  \
    static __device__ __forceinline__ void run(float2 *b, int t, const float2 *w)               \
    {                                                                                           \
      Stages<n, DIR, tps, ES, 1, __VA_ARGS__>::run(b, t, w);                                    \
    }                                                                                           \
    template <int DIR, int ES>                                                                  \
    using IO = StagesIO<n, DIR, tps, ES, 1, __VA_ARGS__>;                                       \
  };
CG_FFT_PLANS(CG_PLAN)
#undef CG_PLAN

constexpr size_t kLds = 80 * 1024;  // bytes per workgroup aimed for: two workgroups per CU

constexpr size_t kLdsMax = 160 * 1024;  // LDS of one CU

// Tile variants.  Every length is compiled with two tile sizes per pass (where two exist); the
// autotuner (cg_autotune) times both on a grid's geometry and keeps the faster, otherwise
// variant 0 runs.
// y/z passes, variant 0: 16 columns (128 B row segments) while the tile fits kLds, else 8; more
// for short lengths so that a workgroup has at least two wavefronts.  Variant 1: twice as wide
// (segments that straddle fewer cache lines when the row pitch is not a multiple of 128 B) if
// that fits one CU's LDS, else half as wide.
template <int N>
constexpr int strided_tile0()
{
  constexpr int T = Plan<N>::TPS;
  int ti = ((size_t)N * 17 * 8 <= kLds) ? 16 : 8;
  while (ti < 64 && ti * T < 128 && (size_t)N * (2 * ti + 1) * 8 <= kLds) ti *= 2;
  return ti;
}
template <int N, int V>
constexpr int strided_tile()
{
  constexpr int t0 = strided_tile0<N>();
  if (V == 0) return t0;
  if ((size_t)N * (2 * t0 + 1) * 8 <= kLdsMax && 2 * t0 * Plan<N>::TPS <= 1024) return 2 * t0;
  return t0 >= 8 ? t0 / 2 : t0;
}

// x pass, variant 0: the most rows (up to 128) whose LDS (N twiddles + (N + 1)(RW + 1) slots)
// fits kLds with at most 512 threads; variant 1: half as many rows (more workgroups per CU).
// Even nxp, where x_rows_dma allows it: variant 0 is the double-buffered kernel, variant 1 the
// single-buffered one with x_rows<N, 0>() rows.
template <int N, int V>
constexpr int x_rows()
{
  constexpr int T = Plan<N>::TPS;
  int rw = 128;
  while (rw > 1 && ((size_t)(N + (N + 1) * (rw + 1)) * 8 > kLds || rw * T > 512)) rw /= 2;
  return (V == 1 && rw > 1) ? rw / 2 : rw;
}
template <int N>
constexpr int x_rows_dma();
// Rows per tile of variant V of the even-nxp x pass.
template <int N, int V>
constexpr int xe_rows()
{
  return x_rows_dma<N>() > 0 ? (V == 0 ? x_rows_dma<N>() : x_rows<N, 0>()) : x_rows<N, V>();
}

__device__ __forceinline__ void load_tw(float2 *tw, const float2 *g, int n, int tid, int nthr)
{
  for (int m = tid; m < n; m += nthr) tw[m] = g[m];
}

template <typename A>
__device__ __forceinline__ int find_item(const A *items, int nitems, unsigned g)
{
  int lo = 0, hi = nitems - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (items[mid].block0 <= g) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// Resolves the grid of this workgroup: `one` for a single-grid launch, else the item that owns
// the (XCD-remapped) block index.
template <typename A>
__device__ __forceinline__ unsigned resolve(const A &one, const A *items, int nitems, A &a)
{
  const unsigned g = cg_xcd_remap(blockIdx.x, gridDim.x);
  if (!items) {
    a = one;
    return g;
  }
  a = items[find_item(items, nitems, g)];
  return g - a.block0;
}

// As resolve(), with the tiles of each grid taken in descending order (nblk(a): the grid's
// tiles).  The y and z passes run this way and the x passes ascending, so that a pass starts on
// the part of the grid its predecessor wrote last (x forward then y; y inverse then x), which is
// still in the memory-side cache; with the XCD remap each XCD walks its band of the grid.
template <typename A, class NB>
__device__ __forceinline__ unsigned resolve_desc(const A &one, const A *items, int nitems, A &a, NB nblk)
{
  const unsigned g = resolve(one, items, nitems, a);
  return nblk(a) - 1u - g;
}

// ---------------------------------------------------------------- y / z passes
// Columns are numbered q = c + inner * o (c < inner the contiguous index, o < outer the slab) and
// a block takes TI consecutive q, so every tile but the grid's last is full even where inner is
// not a multiple of TI (y pass: inner = nxp/2 + 1); a tile that crosses a slab boundary reads
// two shorter segments.  Element e of column q lives at (q % inner) + (q / inner) OS + e S.
template <int N, int DIR, int TI>
__global__ void __launch_bounds__(TI *Plan<N>::TPS)
k_ffts(cg::StridedArgs one, const cg::StridedArgs *__restrict__ items, int nitems, const float2 *__restrict__ twg)
{
  using IO = typename Plan<N>::template IO<DIR, TI>;
  constexpr int NT = TI * Plan<N>::TPS;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  cg::StridedArgs a;
  const unsigned blk = resolve_desc(one, items, nitems, a, [](const cg::StridedArgs &x) { return (x.ncols + TI - 1) / TI; });
  const int tid = threadIdx.x;
  const int c = tid % TI, t = tid / TI;
  const unsigned q = blk * TI + c;
  const bool live = q < a.ncols;
  float2 *col = a.data + (live ? (long)(q % a.inner) + (long)(q / a.inner) * a.OS : 0);
  load_tw(tw, twg, N, tid, NT);
  // the first stage's butterfly inputs, straight from global memory: lanes of a wavefront
  // cover TI consecutive columns of consecutive rows (TI * 8-byte row segments)
  float2 v[IO::BPT][IO::P_];
#pragma unroll
  for (int b = 0; b < IO::BPT; ++b) {
    const int j = t + b * Plan<N>::TPS;
#pragma unroll
    for (int r = 0; r < IO::P_; ++r)
      v[b][r] = (live && (IO::EVEN || j < IO::NB)) ? cg::gld(col + (long)(j + r * IO::NB) * a.S) : make_float2(0.f, 0.f);
  }
  lds_sync();  // twiddles in LDS
  const float scale = a.scale;
  auto sink = [&](int e, float2 x) {
    if (live) cg::gst(col + (long)e * a.S, cscale(x, scale));
  };
  IO::step(v, buf + c, t, tw, sink);
}

// ---------------------------------------------------------------- x pass
// A block takes RW consecutive rows of the in-place layout (CN complex slots = 2 CN floats per
// row), one contiguous span of global memory copied in and out flat (coalesced).  In LDS the
// rows are interleaved, slot m of row r at m * ES + r with odd ES = RW + 1: the FFT stages
// address consecutive rows from consecutive lanes, and the flat copies stride by ES (odd, so
// free of bank conflicts too).

// Even nxp = 2N.  Forward: z[m] = x[2m] + i x[2m+1] (scaled), Z = FFT_N(z),
//   X[k] = (Z[k] + conj Z[N-k]) / 2 - i w^k (Z[k] - conj Z[N-k]) / 2,  w = exp(-2 pi i / 2N);
// bins k and N - k come from the same pair of slots, so each work item rewrites its own pair.
template <int N, int RW>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftx_r2c(cg::XArgs one, const cg::XArgs *__restrict__ items, int nitems, const float2 *__restrict__ twg,
           const float2 *__restrict__ tw_row)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = N + 1;  // complex slots per row
  constexpr int PER = (RW * CN + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  cg::XArgs a;
  const unsigned blk = resolve(one, items, nitems, a);
  const int tid = threadIdx.x;
  const long row0 = (long)blk * RW;
  const int nr = (int)((a.rows - row0) < RW ? (a.rows - row0) : RW);
  const float2 *in = reinterpret_cast<const float2 *>(a.src) + row0 * CN;
  float2 *out = reinterpret_cast<float2 *>(a.dst) + row0 * CN;
  load_tw(tw, twg, N, tid, NT);
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[i] = f < nr * CN ? cg::gld(in + f) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < RW * CN) buf[m * ES + r] = cscale(v[i], a.scale);
    }
  }
  lds_sync();
  Plan<N>::template run<-1, ES>(buf + tid % RW, tid / RW, tw);
  for (int idx = tid; idx < RW * (N / 2 + 1); idx += NT) {
    const int r = idx % RW, k = idx / RW;
    const int kn = k == 0 ? 0 : N - k;
    const float2 za = buf[k * ES + r], zb = buf[kn * ES + r];
    // X[k] from (Z[k], conj Z[N-k]); X[N-k] from (Z[N-k], conj Z[k])
    const float2 ea = cscale(cadd(za, cconj(zb)), 0.5f), da = csub(za, cconj(zb));
    const float2 eb = cscale(cadd(zb, cconj(za)), 0.5f), db = csub(zb, cconj(za));
    const float2 w = tw_row[k];  // w^(N-k) = -conj(w^k) for w = exp(-i pi / N)
    const float2 xa = cadd(ea, cmul(w, make_float2(0.5f * da.y, -0.5f * da.x)));
    const float2 xb = cadd(eb, cmul(make_float2(-w.x, w.y), make_float2(0.5f * db.y, -0.5f * db.x)));
    buf[k * ES + r] = xa;
    buf[(N - k) * ES + r] = xb;  // k = 0 writes bin N into the spare slot
  }
  lds_sync();
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) cg::gst(out + f, buf[m * ES + r]);
  }
}

// Even nxp, inverse: Z[k] = (X[k] + conj X[N-k]) + i w^-k (X[k] - conj X[N-k]) with
// Im X[0] = Im X[N] = 0, z = IFFT_N(Z), x[2m] = Re z[m], x[2m+1] = Im z[m]; then x scale (+ add);
// the row's spare slot is written as zero.
template <int N, int RW, bool ADD>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftx_c2r(cg::XArgs one, const cg::XArgs *__restrict__ items, int nitems, const float2 *__restrict__ twg,
           const float2 *__restrict__ tw_row)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = N + 1;
  constexpr int PER = (RW * CN + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  cg::XArgs a;
  const unsigned blk = resolve(one, items, nitems, a);
  const int tid = threadIdx.x;
  const long row0 = (long)blk * RW;
  const int nr = (int)((a.rows - row0) < RW ? (a.rows - row0) : RW);
  float2 *io = reinterpret_cast<float2 *>(a.dst) + row0 * CN;
  const float2 *ad = a.add ? reinterpret_cast<const float2 *>(a.add) + row0 * CN : nullptr;
  load_tw(tw, twg, N, tid, NT);
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[i] = f < nr * CN ? cg::gld(io + f) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < RW * CN) buf[m * ES + r] = v[i];
    }
  }
  lds_sync();
  for (int idx = tid; idx < RW * (N / 2 + 1); idx += NT) {
    const int r = idx % RW, k = idx / RW;
    float2 xa = buf[k * ES + r], xb = buf[(N - k) * ES + r];  // X[k], X[N-k]
    if (k == 0) {
      xa.y = 0.f;
      xb.y = 0.f;
    }
    const float2 w = tw_row[k];  // w^(N-k) = -conj(w^k) for w = exp(-i pi / N)
    const float2 ea = cadd(xa, cconj(xb)), oa = cmul(csub(xa, cconj(xb)), cconj(w));
    buf[k * ES + r] = make_float2(ea.x - oa.y, ea.y + oa.x);
    if (k > 0 && k < N - k) {
      const float2 eb = cadd(xb, cconj(xa)), ob = cmul(csub(xb, cconj(xa)), make_float2(-w.x, -w.y));
      buf[(N - k) * ES + r] = make_float2(eb.x - ob.y, eb.y + ob.x);
    }
  }
  lds_sync();
  Plan<N>::template run<+1, ES>(buf + tid % RW, tid / RW, tw);
  float2 v[ADD ? PER : 1];
  if (ADD) {  // every load of the added grid in flight before the first store
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[ADD ? i : 0] = f < nr * CN ? cg::gld(ad + f) : make_float2(0.f, 0.f);
    }
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) {
      float2 o = make_float2(0.f, 0.f);
      if (m < N) {
        o = cscale(buf[m * ES + r], a.scale);
        if (ADD) {
          o.x += v[ADD ? i : 0].x;
          o.y += v[ADD ? i : 0].y;
        }
      }
      cg::gst(io + f, o);
    }
  }
}

// ---------------------------------------------------------------- x pass, double-buffered
// Even nxp, the second tile variant.  Persistent workgroups (a multiple of 8, as many as the CUs
// hold): workgroup b serves XCD b % 8, which walks its eighth of the tiles (a contiguous band)
// in steps of gridDim.x / 8.  Two LDS tile buffers hold RW rows each in the grid's own layout
// (row r at r CN, CN = N + 1 odd for even N, so consecutive rows fall on distinct bank pairs):
// while the FFT works on one buffer, the next tile's rows stream into the other with
// global_load_lds (no registers, no LDS store instructions), and the previous tile's stores
// are in flight.  Iteration: start the next tile's copy; transform; read the tile out to
// registers; wait for the copy (and the previous stores); barrier; store.
struct TileWalk {
  unsigned t, hi, step;
  __device__ __forceinline__ explicit TileWalk(unsigned ntiles)
  {
    const unsigned xcd = blockIdx.x % 8u, q = ntiles / 8u, r = ntiles % 8u;
    const unsigned lo = xcd * q + (xcd < r ? xcd : r);
    hi = lo + q + (xcd < r ? 1u : 0u);
    t = lo + blockIdx.x / 8u;
    step = gridDim.x / 8u;
  }
};

// A tile's rows: source (forward: the real grid, inverse: the grid itself), destination, added
// grid, scale, first row and float2 slots (nr CN).  The descriptor fields are read with scalar
// loads (workgroup-uniform), so that looking up the next tile never waits for vector memory.
struct XTile {
  const float2 *src;
  float2 *dst;
  const float2 *add;
  float scale;
  long row0;
  int n;
};

template <int RW, int CN, bool FWD>
__device__ __forceinline__ XTile x_tile(const XTile &one, long one_rows, const cg::XArgs *items, int nitems,
                                        unsigned tile)
{
  XTile x = one;
  long rows = one_rows;
  unsigned blk = tile;
  if (items) {
    typedef const __attribute__((address_space(4))) cg::XArgs *KArgs;
    const KArgs it = (KArgs)items + find_item(items, nitems, tile);
    x.dst = reinterpret_cast<float2 *>(it->dst);
    x.src = FWD ? reinterpret_cast<const float2 *>(it->src) : x.dst;
    x.add = reinterpret_cast<const float2 *>(it->add);
    x.scale = it->scale;
    rows = it->rows;
    blk = tile - it->block0;
  }
  x.row0 = (long)blk * RW;
  const long left = rows - x.row0;
  x.n = (int)(left < RW ? left : RW) * CN;
  return x;
}

// Descriptor fields are pointers loaded from memory: their accesses go through cg::gld / cg::gst
// (cg_mem.h), which keep them global operations rather than flat ones (flat ones also count
// against the LDS counter that lds_sync waits for).

// One 16-byte global_load_lds per lane: lane l of the wave writes LDS bytes [lds + 16 l, +16).
// Issued as inline assembly so that the compiler does not track it (its own tracking makes
// later LDS reads of the other buffer wait for the copy); the caller waits with vm_wait and a
// barrier before reading the destination.  m0 holds the wave's LDS destination; no other code
// in these kernels uses m0.
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"  // m0 is reserved: declared clobbered on purpose
__device__ __forceinline__ void lds_copy16(const float2 *g, unsigned lds)
{
  asm volatile("s_mov_b32 m0, %1\n\tglobal_load_lds_dwordx4 %0, off" : : "v"(g), "s"(lds) : "memory", "m0");
}
#pragma clang diagnostic pop

// n float2 from global src (16-byte aligned) to LDS dst: 16-byte pieces, one per lane (a wave
// fills 1 KB of consecutive LDS), and an odd last float2 by a plain load.
template <int NT>
__device__ __forceinline__ void copy_to_lds(float2 *dst, const float2 *src, int n, int tid)
{
  const int pieces = n >> 1;
  const int wbase = tid & ~63;
  for (int base = 0; base < pieces; base += NT) {
    const unsigned lds = __builtin_amdgcn_readfirstlane(
        (unsigned)(uintptr_t)(__attribute__((address_space(3))) float2 *)(dst + 2 * (base + wbase)));
    if (base + tid < pieces) lds_copy16(src + 2 * (base + tid), lds);
  }
  if ((n & 1) && tid == 0) dst[n - 1] = cg::gld(src + (n - 1));
}

template <int N, int RW, bool FWD, bool ADD>
struct XDma {
  static constexpr int TPS = Plan<N>::TPS;
  static constexpr int NT = RW * TPS;
  static constexpr int CN = N + 1;
  static constexpr int TILE = RW * CN;
  static constexpr int PER = (TILE + NT - 1) / NT;

  // Transforms tile `cur` (in `buf`), prefetches tile `nxt` (when has_next) into `other`, and
  // stores `cur`.  Everything the transform reads is in LDS (twiddles included): a global load
  // issued after the copy would wait for the copy (loads complete in order).
  //   Forward: the R2C post-pass reads the pairs (k, N - k) of a row from LDS with k on
  //   consecutive lanes and stores X[k] and X[N - k] straight to HBM (both runs coalesced), after
  //   the wait for the copy, so the stores drain during the next tile's transform.
  //   Inverse: pre-pass, transform, read-out to registers (+ the added rows), wait, stores.
  // Ends with a barrier: `buf` is free and `other` holds the next tile for every wave.
  static __device__ __forceinline__ void step(float2 *buf, float2 *other, const float2 *tw, const XTile &cur,
                                              const XTile &nxt, bool has_next, const float2 *tw_row, int tid)
  {
    float2 va[ADD ? PER : 1];
    if (ADD) {  // issued before the copy, used after the transform
      const float2 *ad = cur.add + cur.row0 * CN;
#pragma unroll
      for (int i = 0; i < PER; ++i) {
        const int f = tid + i * NT;
        va[ADD ? i : 0] = f < cur.n ? cg::gld(ad + f) : make_float2(0.f, 0.f);
      }
    }
    if (has_next) copy_to_lds<NT>(other, nxt.src + nxt.row0 * CN, nxt.n, tid);
    const int r = tid % RW, t = tid / RW;
    const float scale = cur.scale;
    float2 *out = cur.dst + cur.row0 * CN;
    if (FWD) {
      Plan<N>::template run<-1, 1>(buf + r * CN, t, tw);
      vm_wait();  // the next tile's rows are in `other`; the previous tile's stores are done
      constexpr int H = N / 2 + 1;
      const int nrows = cur.n / CN;
      for (int idx = tid; idx < nrows * H; idx += NT) {
        const int rr = idx / H, k = idx - rr * H;
        const float2 *row = buf + rr * CN;
        const int kn = k == 0 ? 0 : N - k;
        const float2 za = row[k], zb = row[kn];
        const float2 ea = cscale(cadd(za, cconj(zb)), 0.5f), da = csub(za, cconj(zb));
        const float2 eb = cscale(cadd(zb, cconj(za)), 0.5f), db = csub(zb, cconj(za));
        const float2 wk = tw_row[k];  // w^(N-k) = -conj(w^k) for w = exp(-i pi / N)
        const float2 xa = cadd(ea, cmul(wk, make_float2(0.5f * da.y, -0.5f * da.x)));
        const float2 xb = cadd(eb, cmul(make_float2(-wk.x, wk.y), make_float2(0.5f * db.y, -0.5f * db.x)));
        float2 *orow = out + rr * CN;
        cg::gst(orow + k, cscale(xa, scale));
        if (N - k != k) cg::gst(orow + (N - k), cscale(xb, scale));  // k = 0: bin N in the spare slot
      }
      lds_sync();
      return;
    }
    for (int idx = tid; idx < RW * (N / 2 + 1); idx += NT) {
      const int rr = idx % RW, k = idx / RW;
      float2 *row = buf + rr * CN;
      float2 xa = row[k], xb = row[N - k];  // X[k], X[N-k]
      if (k == 0) {
        xa.y = 0.f;
        xb.y = 0.f;
      }
      const float2 wk = tw_row[k];
      const float2 ea = cadd(xa, cconj(xb)), oa = cmul(csub(xa, cconj(xb)), cconj(wk));
      row[k] = make_float2(ea.x - oa.y, ea.y + oa.x);
      if (k > 0 && k < N - k) {
        const float2 eb = cadd(xb, cconj(xa)), ob = cmul(csub(xb, cconj(xa)), make_float2(-wk.x, -wk.y));
        row[N - k] = make_float2(eb.x - ob.y, eb.y + ob.x);
      }
    }
    lds_sync();
    Plan<N>::template run<+1, 1>(buf + r * CN, t, tw);
    float2 o[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int m = f % CN;
      float2 x = make_float2(0.f, 0.f);
      if (f < TILE && m < N) x = cscale(buf[f], scale);
      if (ADD && m < N) {
        x.x += va[ADD ? i : 0].x;
        x.y += va[ADD ? i : 0].y;
      }
      o[i] = x;
    }
    vm_wait();   // the next tile's rows are in `other`; earlier stores are done
    lds_sync();  // ... for every wave; `buf` is free
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      if (f < cur.n) cg::gst(out + f, o[i]);
    }
  }
};

template <int N, int RW, bool FWD, bool ADD>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftx_dma(cg::XArgs one, const cg::XArgs *__restrict__ items, int nitems, unsigned ntiles,
           const float2 *__restrict__ twg, const float2 *__restrict__ tw_row)
{
  using K = XDma<N, RW, FWD, ADD>;
  __shared__ __align__(16) float2 s_tw[N];
  __shared__ __align__(16) float2 s_twr[N / 2 + 1];
  __shared__ __align__(16) float2 s_b0[K::TILE];
  __shared__ __align__(16) float2 s_b1[K::TILE];
  const int tid = threadIdx.x;
  TileWalk w(ntiles);
  if (w.t >= w.hi) return;  // workgroup-uniform: no tile for this workgroup
  load_tw(s_tw, twg, N, tid, K::NT);
  load_tw(s_twr, tw_row, N / 2 + 1, tid, K::NT);
  XTile o1;  // the single grid, from the kernel arguments
  o1.dst = reinterpret_cast<float2 *>(one.dst);
  o1.src = FWD ? reinterpret_cast<const float2 *>(one.src) : o1.dst;
  o1.add = reinterpret_cast<const float2 *>(one.add);
  o1.scale = one.scale;
  o1.row0 = 0;
  o1.n = 0;
  const long rows1 = one.rows;
  XTile cur = x_tile<RW, K::CN, FWD>(o1, rows1, items, nitems, w.t);
  copy_to_lds<K::NT>(s_b0, cur.src + cur.row0 * K::CN, cur.n, tid);
  vm_wait();
  lds_sync();
  for (;;) {  // two tiles per trip, so that each buffer is a fixed array in each half
    unsigned tn = w.t + w.step;
    bool more = tn < w.hi;
    XTile nxt = more ? x_tile<RW, K::CN, FWD>(o1, rows1, items, nitems, tn) : cur;
    K::step(s_b0, s_b1, s_tw, cur, nxt, more, s_twr, tid);
    if (!more) break;
    w.t = tn;
    cur = nxt;
    tn = w.t + w.step;
    more = tn < w.hi;
    nxt = more ? x_tile<RW, K::CN, FWD>(o1, rows1, items, nitems, tn) : cur;
    K::step(s_b1, s_b0, s_tw, cur, nxt, more, s_twr, tid);
    if (!more) break;
    w.t = tn;
    cur = nxt;
  }
}

// Rows per tile of the double-buffered kernel: the most (up to 32) whose two buffers and the
// twiddles fit one CU's LDS with at most 1024 threads; 0 where that leaves fewer than 2 rows.
template <int N>
constexpr int x_rows_dma()
{
  constexpr int T = Plan<N>::TPS;
  int rw = 16;
  while (rw >= 2 && ((size_t)(N + N / 2 + 1 + 2 * (size_t)(N + 1) * rw) * 8 > kLdsMax || rw * T > 1024)) rw /= 2;
  return rw >= 2 ? rw : 0;
}

// Odd nxp = N: each row holds N reals and one padding float, CN = (N + 1) / 2 float pairs.
// Forward: full-length complex FFT of (x, 0); bins 0..CN-1 are kept.
template <int N, int RW>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftxo_r2c(cg::XArgs one, const cg::XArgs *__restrict__ items, int nitems, const float2 *__restrict__ twg)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = (N + 1) / 2;
  constexpr int PER = (RW * CN + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  cg::XArgs a;
  const unsigned blk = resolve(one, items, nitems, a);
  const int tid = threadIdx.x;
  const long row0 = (long)blk * RW;
  const int nr = (int)((a.rows - row0) < RW ? (a.rows - row0) : RW);
  const float2 *in = reinterpret_cast<const float2 *>(a.src) + row0 * CN;
  float2 *out = reinterpret_cast<float2 *>(a.dst) + row0 * CN;
  load_tw(tw, twg, N, tid, NT);
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[i] = f < nr * CN ? cg::gld(in + f) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < RW * CN) {
        buf[(2 * m) * ES + r] = make_float2(v[i].x * a.scale, 0.f);
        if (2 * m + 1 < N) buf[(2 * m + 1) * ES + r] = make_float2(v[i].y * a.scale, 0.f);
      }
    }
  }
  lds_sync();
  Plan<N>::template run<-1, ES>(buf + tid % RW, tid / RW, tw);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) cg::gst(out + f, buf[m * ES + r]);
  }
}

// Odd nxp, inverse: Hermitian extension Z[0] = Re X[0], Z[k] = X[k], Z[N-k] = conj X[k]
// (1 <= k < CN), z = IFFT_N(Z), x[m] = Re z[m] scale (+ add); the padding float is zero.
template <int N, int RW, bool ADD>
__global__ void __launch_bounds__(RW *Plan<N>::TPS)
k_fftxo_c2r(cg::XArgs one, const cg::XArgs *__restrict__ items, int nitems, const float2 *__restrict__ twg)
{
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = RW * TPS;
  constexpr int ES = RW + 1;
  constexpr int CN = (N + 1) / 2;
  constexpr int PER = (RW * CN + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;
  cg::XArgs a;
  const unsigned blk = resolve(one, items, nitems, a);
  const int tid = threadIdx.x;
  const long row0 = (long)blk * RW;
  const int nr = (int)((a.rows - row0) < RW ? (a.rows - row0) : RW);
  float2 *io = reinterpret_cast<float2 *>(a.dst) + row0 * CN;
  const float2 *ad = a.add ? reinterpret_cast<const float2 *>(a.add) + row0 * CN : nullptr;
  load_tw(tw, twg, N, tid, NT);
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[i] = f < nr * CN ? cg::gld(io + f) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, k = f - r * CN;
      if (f < RW * CN) {
        if (k == 0) {
          buf[r] = make_float2(v[i].x, 0.f);
        } else {
          buf[k * ES + r] = v[i];
          buf[(N - k) * ES + r] = cconj(v[i]);
        }
      }
    }
  }
  lds_sync();
  Plan<N>::template run<+1, ES>(buf + tid % RW, tid / RW, tw);
  float2 v[ADD ? PER : 1];
  if (ADD) {  // every load of the added grid in flight before the first store
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      v[ADD ? i : 0] = f < nr * CN ? cg::gld(ad + f) : make_float2(0.f, 0.f);
    }
  }
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < nr * CN) {
      float2 o;
      o.x = buf[(2 * m) * ES + r].x * a.scale;
      o.y = (2 * m + 1 < N) ? buf[(2 * m + 1) * ES + r].x * a.scale : 0.f;
      if (ADD) {
        o.x += v[ADD ? i : 0].x;
        if (2 * m + 1 < N) o.y += v[ADD ? i : 0].y;
      }
      cg::gst(io + f, o);
    }
  }
}

// ---------------------------------------------------------------- realisation: colour + inverse z
// AVOInversion::simulate colours the noise (avoinversion.cpp:1421-1460) and then transforms the
// three grids back (z, y, x passes).  Here one workgroup colours TI columns over all N planes
// of the three output grids straight into LDS (counter noise, chol(Sigma_post), L z) and runs
// their inverse z transforms before anything reaches HBM: the coloured spectra are never
// written and read back.  The y and x passes follow as usual.
// TI columns: at most 16, within kLds (two workgroups per CU) and 1024 threads.
template <int N>
constexpr int colour_tile()
{
  constexpr int T = Plan<N>::TPS;
  int ti = 16;
  while (ti > 1 && ((size_t)(N + 3 * (size_t)N * ti) * 8 > kLds || 3 * ti * T > 1024)) ti /= 2;
  return ti;
}

template <int N, int TI, bool LIST>
__global__ void __launch_bounds__(3 * TI *Plan<N>::TPS)
k_colour_zinv(cg::ColourParams one, const cg::ColourItem *__restrict__ items, int nitems,
              const float2 *__restrict__ twg)
{
  using IO = typename Plan<N>::template IO<+1, TI>;
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = 3 * TI * TPS;
  constexpr int CELLS = N * TI;
  constexpr int PER = (CELLS + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;  // grid g, plane z, column c at (g N + z) TI + c
  const int tid = threadIdx.x;
  // one realisation (the kernel argument, in scalar registers), or a list of them (item i owns
  // tiles [block0_i, block0_{i+1})), whose workgroup's descriptor is staged in LDS
  __shared__ cg::ColourParams staged;
  unsigned tile = cg_xcd_remap(blockIdx.x, gridDim.x);
  if (LIST) {
    const int it = find_item(items, nitems, tile);
    tile -= items[it].block0;
    if (tid == 0) staged = items[it].p;
  }
  const cg::ColourParams &p = LIST ? staged : one;
  const long q0 = (long)tile * TI;
  load_tw(tw, twg, N, tid, NT);
  if (LIST) lds_sync();  // the staged descriptor (the twiddles are next read after phase 1's barrier)
  const long plane = (long)p.cnxp * p.nyp;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int c = e % TI, z = e / TI;
    const bool live = e < CELLS && q0 + c < plane;
    float2 zn[3], o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    bool failed = false;
    if (live) failed = cgcol::colour_cell(p, q0 + c + (long)z * plane, zn, o);
    if (p.fails) {  // one atomic per wavefront, issued by the lowest failing lane
      const unsigned long long msk = __ballot(failed);
      if (failed && (int)__lane_id() == __ffsll((long long)msk) - 1)
        atomicAdd(p.fails, (unsigned long long)__popcll(msk));
    }
    if (e < CELLS) {
#pragma unroll
      for (int g = 0; g < 3; ++g) buf[(g * N + z) * TI + c] = o[g];
    }
  }
  lds_sync();
  const int g = tid / (TI * TPS), r = tid % (TI * TPS), c = r % TI, t = r / TI;
  const bool live = q0 + c < plane;
  float2 *col = p.out[g] + (live ? q0 + c : 0);
  auto sink = [&](int e, float2 x) {
    if (live) cg::gst(col + (long)e * plane, x);
  };
  IO::from_lds(buf + g * N * TI + c, t, tw, sink);
}

// Two realisations of one posterior per workgroup.  Wide tiles matter most here (the covariance
// and mean columns are read in TI-cell row segments): 16 columns with the second realisation
// held in registers (k_colour_zinv2r, three column sets in LDS) where a thread then keeps at
// most 6 cells, else six column sets (k_colour_zinv2) with the widest tile that fits the CU's
// LDS and 1024 threads.  Measured on the 576 and 1152 cubes against 4-column tiles at two
// workgroups per CU: 2.14 -> 1.72 ms (nzp 320, registers, 16 columns), 13.5 -> 10.3 ms
// (nzp 384, six sets, 8 columns).
template <int N>
constexpr bool colour2_in_registers()
{
  constexpr int T = Plan<N>::TPS;
  return (size_t)(N + 3 * (size_t)N * 16) * 8 <= kLdsMax && 3 * 16 * T <= 1024 && (N + 3 * T - 1) / (3 * T) <= 6;
}

template <int N>
constexpr int colour_tile2()
{
  constexpr int T = Plan<N>::TPS;
  if (colour2_in_registers<N>()) return 16;
  int ti = 16;
  while (ti > 1 && ((size_t)(N + 6 * (size_t)N * ti) * 8 > kLdsMax || 6 * ti * T > 1024)) ti /= 2;
  return ti;
}

// k_colour_zinv for realisations stream0 and stream1 of one posterior: the covariance is read
// and factorised once per cell (colour_cell2), both realisations' columns are transformed from
// LDS; grid g < 3 goes to out[g], g >= 3 to out2[g - 3].  One pair (`one`), or a list of them
// (item i owns tiles [block0_i, block0_{i+1}); its descriptor is staged in LDS).
template <int N, int TI, bool THREE, bool LIST>
__global__ void __launch_bounds__(6 * TI *Plan<N>::TPS)
k_colour_zinv2(cg::ColourParams one, const cg::ColourItem *__restrict__ items, int nitems,
               const float2 *__restrict__ twg)
{
  using IO = typename Plan<N>::template IO<+1, TI>;
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = 6 * TI * TPS;
  constexpr int CELLS = N * TI;
  constexpr int PER = (CELLS + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;  // grid g (0..5), plane z, column c at (g N + z) TI + c
  const int tid = threadIdx.x;
  __shared__ cg::ColourParams staged;
  unsigned tile = cg_xcd_remap(blockIdx.x, gridDim.x);
  if (LIST) {
    const int it = find_item(items, nitems, tile);
    tile -= items[it].block0;
    if (tid == 0) staged = items[it].p;
  }
  const cg::ColourParams &p = LIST ? staged : one;
  const long q0 = (long)tile * TI;
  load_tw(tw, twg, N, tid, NT);
  if (LIST) lds_sync();  // the staged descriptor
  const long plane = (long)p.cnxp * p.nyp;
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int c = e % TI, z = e / TI;
    const bool live = e < CELLS && q0 + c < plane;
    float2 o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    float2 o2[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    bool failed = false;
    if (live) failed = cgcol::colour_cell2t<THREE>(p, q0 + c + (long)z * plane, o, o2);
    if (p.fails) {  // a failed cell zeroes the noise of both realisations: counted twice
      const unsigned long long msk = __ballot(failed);
      if (failed && (int)__lane_id() == __ffsll((long long)msk) - 1)
        atomicAdd(p.fails, 2ull * (unsigned long long)__popcll(msk));
    }
    if (e < CELLS) {
#pragma unroll
      for (int g = 0; g < 3; ++g) {
        buf[(g * N + z) * TI + c] = o[g];
        buf[((g + 3) * N + z) * TI + c] = o2[g];
      }
    }
  }
  lds_sync();
  const int g = tid / (TI * TPS), r = tid % (TI * TPS), c = r % TI, t = r / TI;
  const bool live = q0 + c < plane;
  float2 *col = (g < 3 ? p.out[g] : p.out2[g - 3]) + (live ? q0 + c : 0);
  auto sink = [&](int e, float2 x) {
    if (live) cg::gst(col + (long)e * plane, x);
  };
  IO::from_lds(buf + g * N * TI + c, t, tw, sink);
}

// k_colour_zinv2 with half the LDS per column: the first realisation's three columns go through
// LDS and their inverse z transforms while the second's coloured values wait in registers, then
// take the same LDS and transforms.  Twice the columns per workgroup at the same LDS (wider row
// segments of the covariance reads), the same cells and arithmetic (bit for bit).
// (Register budget: the workgroup's waves spread over the CU's four SIMDs -- a 16-column tile at
// nzp 320 is 960 threads, four waves per SIMD, at most 128 registers per lane.  A 12-column tile
// (three waves per SIMD) fits without spilling but reads 96-byte row segments that straddle
// cache lines: 2.18 ms per pair against 1.72 ms.)
template <int N, int TI, bool THREE, bool LIST>
__global__ void __launch_bounds__(3 * TI *Plan<N>::TPS) __attribute__((amdgpu_waves_per_eu((3 * TI * Plan<N>::TPS + 255) / 256)))
k_colour_zinv2r(cg::ColourParams one, const cg::ColourItem *__restrict__ items, int nitems,
                const float2 *__restrict__ twg)
{
  using IO = typename Plan<N>::template IO<+1, TI>;
  constexpr int TPS = Plan<N>::TPS;
  constexpr int NT = 3 * TI * TPS;
  constexpr int CELLS = N * TI;
  constexpr int PER = (CELLS + NT - 1) / NT;
  extern __shared__ float2 smem[];
  float2 *tw = smem;
  float2 *buf = smem + N;  // grid g (0..2), plane z, column c at (g N + z) TI + c
  const int tid = threadIdx.x;
  __shared__ cg::ColourParams staged;
  unsigned tile = cg_xcd_remap(blockIdx.x, gridDim.x);
  if (LIST) {
    const int it = find_item(items, nitems, tile);
    tile -= items[it].block0;
    if (tid == 0) staged = items[it].p;
  }
  const cg::ColourParams &p = LIST ? staged : one;
  const long q0 = (long)tile * TI;
  load_tw(tw, twg, N, tid, NT);
  if (LIST) lds_sync();  // the staged descriptor
  const long plane = (long)p.cnxp * p.nyp;
  float2 keep[PER][3];
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int c = e % TI, z = e / TI;
    const bool live = e < CELLS && q0 + c < plane;
    float2 o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
#pragma unroll
    for (int g = 0; g < 3; ++g) keep[i][g] = make_float2(0.f, 0.f);
    bool failed = false;
    if (live) failed = cgcol::colour_cell2t<THREE>(p, q0 + c + (long)z * plane, o, keep[i]);
    if (p.fails) {  // a failed cell zeroes the noise of both realisations: counted twice
      const unsigned long long msk = __ballot(failed);
      if (failed && (int)__lane_id() == __ffsll((long long)msk) - 1)
        atomicAdd(p.fails, 2ull * (unsigned long long)__popcll(msk));
    }
    if (e < CELLS) {
#pragma unroll
      for (int g = 0; g < 3; ++g) buf[(g * N + z) * TI + c] = o[g];
    }
    // one cell's loads in flight at a time: the kept values already take 6 PER registers
    __builtin_amdgcn_sched_barrier(0);
  }
  lds_sync();
  const int g = tid / (TI * TPS), r = tid % (TI * TPS), c = r % TI, t = r / TI;
  const bool live = q0 + c < plane;
  const long off = live ? q0 + c : 0;
  {
    float2 *col = p.out[g] + off;
    auto sink = [&](int e, float2 x) {
      if (live) cg::gst(col + (long)e * plane, x);
    };
    IO::from_lds(buf + g * N * TI + c, t, tw, sink);
  }
  // The last stage of from_lds reads its inputs and passes a barrier before it computes and
  // stores, so no thread reads the column sets any more: the second realisation may overwrite them.
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int e = tid + i * NT;
    const int cc = e % TI, z = e / TI;
    if (e < CELLS) {
#pragma unroll
      for (int gg = 0; gg < 3; ++gg) buf[(gg * N + z) * TI + cc] = keep[i][gg];
    }
  }
  lds_sync();
  {
    float2 *col = p.out2[g] + off;
    auto sink = [&](int e, float2 x) {
      if (live) cg::gst(col + (long)e * plane, x);
    };
    IO::from_lds(buf + g * N * TI + c, t, tw, sink);
  }
}

// Wavefronts of the fused kernel one CU holds (LDS-limited).
template <int N>
constexpr int colour_waves()
{
  constexpr int TI = colour_tile<N>();
  constexpr size_t lds = sizeof(float2) * (N + 3 * (size_t)N * TI);
  return (int)(kLdsMax / lds) * ((3 * TI * Plan<N>::TPS + 63) / 64);
}

// ---------------------------------------------------------------- host launchers
// One realisation (`one`; nblocks = its tiles) or a list (`items`, block0 in tiles of
// colour_tile<N>() columns; nblocks = all tiles).
template <int N>
int colour_zinv_launch(cg_ctx *ctx, const cg::ColourParams *one, const cg::ColourItem *items, int nitems,
                       unsigned nblocks, const float2 *tw)
{
  constexpr int TI = colour_tile<N>();
  if (nblocks == 0) return CG_OK;
  const size_t lds = sizeof(float2) * (N + 3 * (size_t)N * TI);
  const cg::ColourParams o = one ? *one : cg::ColourParams{};
  if (items)
    k_colour_zinv<N, TI, true><<<nblocks, 3 * TI * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  else
    k_colour_zinv<N, TI, false><<<nblocks, 3 * TI * Plan<N>::TPS, lds, ctx->stream>>>(o, nullptr, 0, tw);
  return cg::hip_fail(ctx, hipGetLastError(), "k_colour_zinv");
}

// k_colour_zinv2 (REG false) or k_colour_zinv2r (REG true) with TI columns per workgroup: one
// pair (`one`; nblocks = its tiles), or a list of pairs whose six streams are three whole Philox
// blocks each (even first realisations; nblocks = all tiles).
template <int N, int TI, bool REG>
int colour_zinv2_go(cg_ctx *ctx, const cg::ColourParams *one, const cg::ColourItem *items, int nitems,
                    unsigned nblk, const float2 *tw)
{
  constexpr int SETS = REG ? 3 : 6;
  constexpr unsigned NT = SETS * TI * Plan<N>::TPS;
  const size_t lds = sizeof(float2) * (N + SETS * (size_t)N * TI);
  if (nblk == 0) return CG_OK;
  const cg::ColourParams o = one ? *one : cg::ColourParams{};
  if (items) {
    if constexpr (REG) k_colour_zinv2r<N, TI, true, true><<<nblk, NT, lds, ctx->stream>>>(o, items, nitems, tw);
    else k_colour_zinv2<N, TI, true, true><<<nblk, NT, lds, ctx->stream>>>(o, items, nitems, tw);
  } else if (cgcol::pair_in_three_blocks(o)) {
    if constexpr (REG) k_colour_zinv2r<N, TI, true, false><<<nblk, NT, lds, ctx->stream>>>(o, nullptr, 0, tw);
    else k_colour_zinv2<N, TI, true, false><<<nblk, NT, lds, ctx->stream>>>(o, nullptr, 0, tw);
  } else {
    if constexpr (REG) k_colour_zinv2r<N, TI, false, false><<<nblk, NT, lds, ctx->stream>>>(o, nullptr, 0, tw);
    else k_colour_zinv2<N, TI, false, false><<<nblk, NT, lds, ctx->stream>>>(o, nullptr, 0, tw);
  }
  return cg::hip_fail(ctx, hipGetLastError(), REG ? "k_colour_zinv2r" : "k_colour_zinv2");
}

// Two realisations of one posterior (p.out / p.stream0 and p.out2 / p.stream1), or a list of
// such pairs (block0 in tiles of colour_tile2<N>() columns).
template <int N>
int colour_zinv2_launch(cg_ctx *ctx, const cg::ColourParams *one, const cg::ColourItem *items, int nitems,
                        unsigned nblocks, const float2 *tw)
{
  return colour_zinv2_go<N, colour_tile2<N>(), colour2_in_registers<N>()>(ctx, one, items, nitems, nblocks, tw);
}

template <int N, int V>
int strided_launch(cg_ctx *ctx, int dir, const cg::StridedArgs *one, const cg::StridedArgs *items, int nitems,
                   unsigned nblocks, const float2 *tw)
{
  constexpr int TI = strided_tile<N, V>();
  const size_t lds = sizeof(float2) * (N + (size_t)N * TI);
  const cg::StridedArgs o = one ? *one : cg::StridedArgs{};
  if (dir < 0)
    k_ffts<N, -1, TI><<<nblocks, TI * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  else
    k_ffts<N, +1, TI><<<nblocks, TI * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  return cg::hip_fail(ctx, hipGetLastError(), "k_ffts");
}

// Workgroups of a persistent launch over `ntiles` tiles: as many as the CUs hold at once
// (`per_cu`, the kernel's occupancy), a multiple of 8 (one share per XCD), no more than the tiles
// need.
inline unsigned persistent_blocks(const cg_ctx *ctx, unsigned ntiles, int per_cu)
{
  const unsigned cap = (unsigned)(per_cu > 0 ? per_cu : 1) * (unsigned)ctx->num_cu;
  const unsigned want = ntiles < cap ? ntiles : cap;
  return ((want + 7u) / 8u) * 8u;
}

// Workgroups of `kernel` one CU holds with `threads` threads (asked once per kernel: every
// device this library runs on is a gfx950).
template <typename K>
int occupancy(K kernel, int threads)
{
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, kernel, threads, 0) != hipSuccess || n < 1) n = 1;
  return n;
}

template <int N, bool FWD, bool ADD>
int xdma_launch(cg_ctx *ctx, const cg::XArgs *one, const cg::XArgs *items, int nitems, unsigned ntiles,
                const float2 *tw, const float2 *twr)
{
  constexpr int RW = x_rows_dma<N>();
  constexpr int NT = RW * Plan<N>::TPS;
  if (ntiles == 0) return CG_OK;
  static const int occ = occupancy(k_fftx_dma<N, RW, FWD, ADD>, NT);
  const cg::XArgs o = one ? *one : cg::XArgs{};
  k_fftx_dma<N, RW, FWD, ADD><<<persistent_blocks(ctx, ntiles, occ), NT, 0, ctx->stream>>>(o, items, nitems, ntiles,
                                                                                            tw, twr);
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftx_dma");
}

// `add`: the inverse adds a real grid (every descriptor of a list has one, or none has).
// `nblocks`: tiles of xe_rows<N, V>() rows (of the grid, or of all descriptors of a list).
template <int N, int V>
int xe_launch(cg_ctx *ctx, int forward, int add, const cg::XArgs *one, const cg::XArgs *items, int nitems,
              unsigned nblocks, const float2 *tw, const float2 *twr)
{
  if constexpr (V == 0 && x_rows_dma<N>() > 0) {
    if (forward) return xdma_launch<N, true, false>(ctx, one, items, nitems, nblocks, tw, twr);
    if (add) return xdma_launch<N, false, true>(ctx, one, items, nitems, nblocks, tw, twr);
    return xdma_launch<N, false, false>(ctx, one, items, nitems, nblocks, tw, twr);
  }
  constexpr int RW = xe_rows<N, V>();
  const size_t lds = sizeof(float2) * (N + (size_t)(N + 1) * (RW + 1));
  const cg::XArgs o = one ? *one : cg::XArgs{};
  if (forward)
    k_fftx_r2c<N, RW><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw, twr);
  else if (add)
    k_fftx_c2r<N, RW, true><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw, twr);
  else
    k_fftx_c2r<N, RW, false><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw, twr);
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftx");
}

template <int N, int V>
int xo_launch(cg_ctx *ctx, int forward, int add, const cg::XArgs *one, const cg::XArgs *items, int nitems,
              unsigned nblocks, const float2 *tw, const float2 *)
{
  constexpr int RW = x_rows<N, V>();
  const size_t lds = sizeof(float2) * (N + (size_t)(N + 1) * (RW + 1));
  const cg::XArgs o = one ? *one : cg::XArgs{};
  if (forward)
    k_fftxo_r2c<N, RW><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  else if (add)
    k_fftxo_c2r<N, RW, true><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  else
    k_fftxo_c2r<N, RW, false><<<nblocks, RW * Plan<N>::TPS, lds, ctx->stream>>>(o, items, nitems, tw);
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftxo");
}


// ---------------------------------------------------------------- x pass with the first z stage
// The z transform of length nzp = N1 N2 in two stages (n = n1 + N1 n2, k = k2 + N2 k1,
// w = exp(-2 pi i / nzp)):
//   stage 1: for each n1, a length-N2 DFT over n2 (planes n1 + N1 n2), times w^(n1 k2), into
//            plane n1 + N1 k2;
//   stage 2: for each k2, a length-N1 DFT over the N1 consecutive planes n1 + N1 k2, after which
//            plane k1 + N1 k2 holds frequency k2 + N2 k1.
// A tile of this x pass is the N2 rows of one (y, n1), which it transforms along x in LDS as
// k_fftx_r2c does and then along z (stage 1) before anything reaches HBM; stage 2 belongs to a
// pass whose tiles are N1 consecutive planes (the posterior, which also starts the inverse
// transform with the inverse of stage 2).  The y pass commutes with both stages.  Inverse: the
// inverse stage 1 (times w^-(n1 k2), then a length-N2 inverse DFT over k2) and the x c2r.
// Even nxp = 2N only (CN = N + 1 complex slots per row).  Tile t: y = t mod nyp, n1 = t / nyp.
struct XZArgs {
  const float *src;  // forward: the real grid read (may be the grid itself)
  float *dst;        // the grid
  const float *add;  // inverse: real grid added to the output, or null
  int nyp, n1;       // rows per plane; N1
  float scale;
};

// The x transform of an xz tile: the length's plan, or for the split geometries' lengths a plan
// with more threads per row (288: 24 instead of 12; 576: 32 instead of 24), so that the z stage's
// N + 1 bins take one round over the workgroup's threads and more loads are in flight
// (k_fftxz_fwd<288, 20>: 214 -> 199 us per launch at 576 x 576 x 320).
template <int N>
struct XZPlan {
  static constexpr int TPS = Plan<N>::TPS;
  template <int DIR, int ES>
  static __device__ __forceinline__ void run(float2 *b, int t, const float2 *w)
  {
    Plan<N>::template run<DIR, ES>(b, t, w);
  }
};
template <>
struct XZPlan<288> {
  static constexpr int TPS = 24;
  template <int DIR, int ES>
  static __device__ __forceinline__ void run(float2 *b, int t, const float2 *w)
  {
    Stages<288, DIR, 24, ES, 1, 12, 24>::run(b, t, w);
  }
};
template <>
struct XZPlan<576> {
  static constexpr int TPS = 32;
  template <int DIR, int ES>
  static __device__ __forceinline__ void run(float2 *b, int t, const float2 *w)
  {
    Stages<576, DIR, 32, ES, 1, 24, 24>::run(b, t, w);
  }
};

template <int N, int N2>
struct XZ {
  static constexpr int TPS = XZPlan<N>::TPS;
  static constexpr int NT = N2 * TPS;
  static constexpr int ES = N2 + 1;  // odd: slot m of row r at m ES + r
  static constexpr int CN = N + 1;
  static constexpr int PER = (N2 * CN + NT - 1) / NT;
};

template <int N, int N2>
__global__ void __launch_bounds__(N2 * XZPlan<N>::TPS)
k_fftxz_fwd(XZArgs a, const float2 *__restrict__ twg, const float2 *__restrict__ tw_row,
            const float2 *__restrict__ twz)
{
  using K = XZ<N, N2>;
  constexpr int NT = K::NT, ES = K::ES, CN = K::CN, PER = K::PER;
  __shared__ float2 s_tw[N];
  __shared__ float2 s_twr[N / 2 + 1];
  __shared__ float2 s_twz[N2];
  __shared__ float2 buf[CN * ES];
  const int tid = threadIdx.x;
  const unsigned t = cg_xcd_remap(blockIdx.x, gridDim.x);
  const int y = (int)(t % (unsigned)a.nyp), n1 = (int)(t / (unsigned)a.nyp);
  const long pitch = (long)a.n1 * a.nyp * CN;  // float2 from row n2 to row n2 + 1 (N1 planes)
  const long first = ((long)n1 * a.nyp + y) * CN;
  const float2 *in = reinterpret_cast<const float2 *>(a.src) + first;
  float2 *out = reinterpret_cast<float2 *>(a.dst) + first;
  load_tw(s_tw, twg, N, tid, NT);
  load_tw(s_twr, tw_row, N / 2 + 1, tid, NT);
  if (tid < N2) s_twz[tid] = twz[n1 * tid];  // n1 k2 < nzp
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      v[i] = f < N2 * CN ? cg::gld(in + r * pitch + m) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < N2 * CN) buf[m * ES + r] = cscale(v[i], a.scale);
    }
  }
  lds_sync();
  XZPlan<N>::template run<-1, ES>(buf + tid % N2, tid / N2, s_tw);
  for (int idx = tid; idx < N2 * (N / 2 + 1); idx += NT) {  // R2C post-pass, as k_fftx_r2c
    const int r = idx % N2, k = idx / N2;
    const int kn = k == 0 ? 0 : N - k;
    const float2 za = buf[k * ES + r], zb = buf[kn * ES + r];
    const float2 ea = cscale(cadd(za, cconj(zb)), 0.5f), da = csub(za, cconj(zb));
    const float2 eb = cscale(cadd(zb, cconj(za)), 0.5f), db = csub(zb, cconj(za));
    const float2 w = s_twr[k];
    const float2 xa = cadd(ea, cmul(w, make_float2(0.5f * da.y, -0.5f * da.x)));
    const float2 xb = cadd(eb, cmul(make_float2(-w.x, w.y), make_float2(0.5f * db.y, -0.5f * db.x)));
    buf[k * ES + r] = xa;
    buf[(N - k) * ES + r] = xb;
  }
  lds_sync();
  // stage 1 along z: bin m of the N2 rows, one sequence per thread, stored row by row (lanes on
  // consecutive bins: coalesced)
  for (int m = tid; m < CN; m += NT) {
    float2 c[N2];
#pragma unroll
    for (int r = 0; r < N2; ++r) c[r] = buf[m * ES + r];
    dft<-1, N2>(c);
#pragma unroll
    for (int k2 = 0; k2 < N2; ++k2) cg::gst(out + k2 * pitch + m, k2 == 0 ? c[0] : cmul(c[k2], s_twz[k2]));
  }
}

template <int N, int N2, bool ADD>
__global__ void __launch_bounds__(N2 * XZPlan<N>::TPS)
k_fftxz_inv(XZArgs a, const float2 *__restrict__ twg, const float2 *__restrict__ tw_row,
            const float2 *__restrict__ twz)
{
  using K = XZ<N, N2>;
  constexpr int NT = K::NT, ES = K::ES, CN = K::CN, PER = K::PER;
  __shared__ float2 s_tw[N];
  __shared__ float2 s_twr[N / 2 + 1];
  __shared__ float2 s_twz[N2];
  __shared__ float2 buf[CN * ES];
  const int tid = threadIdx.x;
  const unsigned t = cg_xcd_remap(blockIdx.x, gridDim.x);
  const int y = (int)(t % (unsigned)a.nyp), n1 = (int)(t / (unsigned)a.nyp);
  const long pitch = (long)a.n1 * a.nyp * CN;
  const long first = ((long)n1 * a.nyp + y) * CN;
  float2 *io = reinterpret_cast<float2 *>(a.dst) + first;
  const float2 *ad = ADD ? reinterpret_cast<const float2 *>(a.add) + first : nullptr;
  load_tw(s_tw, twg, N, tid, NT);
  load_tw(s_twr, tw_row, N / 2 + 1, tid, NT);
  if (tid < N2) s_twz[tid] = cconj(twz[n1 * tid]);
  {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      v[i] = f < N2 * CN ? cg::gld(io + r * pitch + m) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      if (f < N2 * CN) buf[m * ES + r] = v[i];
    }
  }
  float2 va[ADD ? PER : 1];
  if (ADD) {  // the added rows' loads in flight during the transforms
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      const int f = tid + i * NT;
      const int r = f / CN, m = f - r * CN;
      va[ADD ? i : 0] = f < N2 * CN ? cg::gld(ad + r * pitch + m) : make_float2(0.f, 0.f);
    }
  }
  lds_sync();
  // inverse stage 1: row k2 times w^-(n1 k2), then the inverse DFT over k2 into rows n2
  for (int m = tid; m < CN; m += NT) {
    float2 c[N2];
#pragma unroll
    for (int r = 0; r < N2; ++r) c[r] = r == 0 ? buf[m * ES] : cmul(buf[m * ES + r], s_twz[r]);
    dft<+1, N2>(c);
#pragma unroll
    for (int r = 0; r < N2; ++r) buf[m * ES + r] = c[r];
  }
  lds_sync();
  for (int idx = tid; idx < N2 * (N / 2 + 1); idx += NT) {  // C2R pre-pass, as k_fftx_c2r
    const int r = idx % N2, k = idx / N2;
    float2 xa = buf[k * ES + r], xb = buf[(N - k) * ES + r];
    if (k == 0) {
      xa.y = 0.f;
      xb.y = 0.f;
    }
    const float2 w = s_twr[k];
    const float2 ea = cadd(xa, cconj(xb)), oa = cmul(csub(xa, cconj(xb)), cconj(w));
    buf[k * ES + r] = make_float2(ea.x - oa.y, ea.y + oa.x);
    if (k > 0 && k < N - k) {
      const float2 eb = cadd(xb, cconj(xa)), ob = cmul(csub(xb, cconj(xa)), make_float2(-w.x, -w.y));
      buf[(N - k) * ES + r] = make_float2(eb.x - ob.y, eb.y + ob.x);
    }
  }
  lds_sync();
  XZPlan<N>::template run<+1, ES>(buf + tid % N2, tid / N2, s_tw);
#pragma unroll
  for (int i = 0; i < PER; ++i) {
    const int f = tid + i * NT;
    const int r = f / CN, m = f - r * CN;
    if (f < N2 * CN) {
      float2 o = make_float2(0.f, 0.f);
      if (m < N) {
        o = cscale(buf[m * ES + r], a.scale);
        if (ADD) {
          o.x += va[ADD ? i : 0].x;
          o.y += va[ADD ? i : 0].y;
        }
      }
      cg::gst(io + r * pitch + m, o);
    }
  }
}

template <int N, int N2>
int xz_launch(cg_ctx *ctx, int forward, const XZArgs &a, int nzp, const float2 *tw, const float2 *twr,
              const float2 *twz)
{
  const unsigned nblk = (unsigned)a.nyp * (unsigned)a.n1;
  if (forward)
    k_fftxz_fwd<N, N2><<<nblk, XZ<N, N2>::NT, 0, ctx->stream>>>(a, tw, twr, twz);
  else if (a.add)
    k_fftxz_inv<N, N2, true><<<nblk, XZ<N, N2>::NT, 0, ctx->stream>>>(a, tw, twr, twz);
  else
    k_fftxz_inv<N, N2, false><<<nblk, XZ<N, N2>::NT, 0, ctx->stream>>>(a, tw, twr, twz);
  (void)nzp;
  return cg::hip_fail(ctx, hipGetLastError(), "k_fftxz");
}

}  // namespace cgfix
