// cg_invert.hip -- kernels of the split-z inversion (cg_invert): the posterior means and
// covariance of one inversion with the z transform cut into two stages, so that the posterior
// pass also carries the z transform's last forward stage and the inverse's first stage.
//
// Pass sequence of one inversion (nzp = N1 N2; G = 3 + ntheta input grids):
//   G x  k_fftxz_fwd   x r2c + z stage 1                  read the real cube, write the grid
//   G x  y pass (k_ffts, forward)
//   1 x  k_posterior_z z stage 2 of G grids, posterior, inverse z stage 2 of the 3 means
//   3 x  y pass (inverse)
//   3 x  k_fftxz_inv   inverse z stage 1 + x c2r
// against x, y, z, posterior, z, y, x of the unsplit sequence: the forward z passes (2 G grids
// of traffic), the posterior's round trip of the spectra and the means' inverse z passes
// (6 grids) are replaced by one read of the G grids and one write of the means.
//
// k_posterior_z: tile = N1 consecutive planes (n1 + N1 k2, one k2) x 64 consecutive cells of a
// plane.  The G grids' tile rows go to LDS; one thread per (grid, column) runs the length-N1
// DFT along z in registers; then wave w of W takes planes k1 = w, w + W, ... -- frequency
// k = k2 + N2 k1, one k per wave, so the per-plane tables are scalar loads as in k_posterior --
// and the 64 lanes the 64 columns; the covariance goes straight to HBM at plane k (the natural
// spectrum layout), the means back to LDS, and one thread per (mean, column) runs the inverse
// length-N1 DFT and stores the column's N1 planes.  Per-cell arithmetic is cg_post_dev.h's
// update, as in k_posterior (avoinversion.cpp:944-1083).
//
// FP64 re-solves (CG_PREC_MIXED): a tile with a deferred wavenumber writes the covariance of its
// other cells, leaves the means of the columns holding one unwritten (the prior means stay in
// HBM, and the seismic grids are only ever read) and is appended to a tile list;
// k_posterior_z_redo then runs each listed tile again with the deferred cells in FP64.  Every
// tile of the list is written by exactly one workgroup, so the result does not depend on the
// list's order.
#include "cg_fft_fixed.h"
#include "cg_noise.h"
#include "cg_post_dev.h"

namespace cgpost {

using cgdft::dft;
using cgfix::lds_sync;

constexpr int kZCols = 64;  // cells of a plane per tile (one per lane)

// Waves per workgroup: the smallest divisor of N1 from 8 up (N1 itself below 8), so that every
// wave takes N1 / W planes.
constexpr int z_waves(int n1)
{
  if (n1 <= 8) return n1;
  for (int w = 8; w < n1; ++w)
    if (n1 % w == 0) return w;
  return n1;
}
template <int N1>
constexpr int z_threads()
{
  return 64 * z_waves(N1);
}

// Marks a deferred wavenumber's mean entries in LDS (a NaN with a payload no input carries): the
// inverse stage-2 transform of its column is then not stored.
constexpr unsigned kDeferMark = 0x7fc0defeu;

// One 64-column tile.  buf: [G][N1][64] float2 in LDS.
template <int NT, int N1, bool MIXED, bool REDO>
__device__ __forceinline__ void post_z_tile(const cg::PostZArgs &z, unsigned t, float2 *buf)
{
  constexpr int G = 3 + NT, TC = kZCols, NTH = z_threads<N1>(), W = NTH / 64;
  static_assert((TC * N1) % NTH == 0, "a tile's rows of one grid must be whole rounds of the threads");
  constexpr int PER = G * N1 * TC / NTH;
  const cg::PostParams &p = z.p;
  const int tid = threadIdx.x;
  const long plane = (long)p.cnxp * p.nyp;
  const unsigned k2 = t / z.ctiles, ct = t % z.ctiles;
  const long c0 = (long)ct * TC;
  // 1. the G grids' rows of planes n1 + N1 k2, cells c0 .. c0 + 63 (512-byte segments).  With an
  // even plane size every row starts 16-byte aligned and a thread loads cell pairs (half the load
  // instructions: 1557 -> 1538 us per pass at 576 x 576 x 320).
  constexpr bool kPairs = (TC * N1) % (2 * NTH) == 0;
  if (kPairs && (plane & 1) == 0) {
    if constexpr (kPairs) {
      constexpr int RPG2 = TC * N1 / (2 * NTH), PER2 = G * RPG2;
      typedef float fvec4 __attribute__((ext_vector_type(4)));
      fvec4 v[PER2];
#pragma unroll
      for (int i = 0; i < PER2; ++i) {
        const int g = i / RPG2;
        const int e = 2 * ((i % RPG2) * NTH + tid);  // the pair's first cell within the grid's tile
        const int c = e % TC, n1 = e / TC;
        const float2 *src = g < 3 ? p.mean[g < 3 ? g : 0] : p.data[g >= 3 ? g - 3 : 0];
        v[i] = c0 + c < plane  // c0 + c even and plane even: the pair's second cell is in range too
                   ? *(const __attribute__((address_space(1))) fvec4 *)(src + (long)(n1 + N1 * (long)k2) * plane + c0 + c)
                   : fvec4{0.f, 0.f, 0.f, 0.f};
      }
#pragma unroll
      for (int i = 0; i < PER2; ++i) {
        const int e = 2 * ((i % RPG2) * NTH + tid);
        *reinterpret_cast<fvec4 *>(buf + (i / RPG2) * N1 * TC + e) = v[i];
      }
    }
  } else {
    float2 v[PER];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      constexpr int RPG = TC * N1 / NTH;  // rounds per grid
      const int g = i / RPG;
      const int e = (i % RPG) * NTH + tid;  // within the grid's tile
      const int c = e % TC, n1 = e / TC;
      const float2 *src = g < 3 ? p.mean[g < 3 ? g : 0] : p.data[g >= 3 ? g - 3 : 0];
      v[i] = c0 + c < plane ? cg::gld(src + (long)(n1 + N1 * (long)k2) * plane + c0 + c) : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
      constexpr int RPG = TC * N1 / NTH;
      const int g = i / RPG;
      const int e = (i % RPG) * NTH + tid;
      buf[g * N1 * TC + e] = v[i];
    }
  }
  lds_sync();
  // 2. z stage 2, forward: one (grid, column) per thread
  for (int s = tid; s < G * TC; s += NTH) {
    const int g = s / TC, c = s % TC;
    float2 w[N1];
#pragma unroll
    for (int n = 0; n < N1; ++n) w[n] = buf[(g * N1 + n) * TC + c];
    dft<-1, N1>(w);
#pragma unroll
    for (int n = 0; n < N1; ++n) buf[(g * N1 + n) * TC + c] = w[n];
  }
  lds_sync();
  // 3. the posterior of each cell: wave wv takes planes k1 = wv, wv + 8, ...
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), c = tid & 63;
  const long cc = c0 + c;
  const bool active = cc < plane;
  const long ccs = active ? cc : 0;
#pragma unroll 1
  for (int k1 = wv; k1 < N1; k1 += W) {
    const int k = (int)k2 + z.n2 * k1;  // wavefront-uniform
    const long idx = (long)k * plane + ccs;
    const bool inv = kat(p.invert, k) != 0;
    bool did_inv = false, did_fail = false, defer = false;
    if (active) {
      const float tr = gld(p.rxy + ccs) * kat(p.rt, k);
      const float r = sqrtf(tr * tr);
      float P6[6];
#pragma unroll
      for (int g = 0; g < 6; ++g) P6[g] = r * p.s6f[g];
      bool write_prior = !inv, write_cov = false;
      cpx<float> cov6[6];
      if (inv) {
        did_inv = true;
        const float lt = gld(p.lam2d + ccs);
        const float lam = sqrtf(lt * lt);
        cpx<float> m3[3], d[NT], res[NT];
#pragma unroll
        for (int a = 0; a < 3; ++a) {
          const float2 v = buf[(a * N1 + k1) * TC + c];
          m3[a] = mk<float>(v.x, v.y);
        }
#pragma unroll
        for (int l = 0; l < NT; ++l) {
          const float2 v = buf[((3 + l) * N1 + k1) * TC + c];
          d[l] = mk<float>(v.x, v.y);
        }
        float cond = 0.f;
        const int fail = update<NT, float, true>(kat(p.tab32, k), r, lam, P6, m3, d, cov6, res, false, cond);
        if (MIXED && (fail || cond > p.cond_limit)) {
          defer = true;
        } else if (fail) {  // posterior = prior (avoinversion.cpp:1040-1042)
          did_fail = true;
          write_prior = true;
        } else {
          write_cov = true;
#pragma unroll
          for (int a = 0; a < 3; ++a) buf[(a * N1 + k1) * TC + c] = make_float2(m3[a].re, m3[a].im);
        }
        if (REDO && defer) {  // the FP64 re-solve of k_posterior64 on the same inputs
          const double rd = fabs((double)gld(p.rxy + ccs) * (double)kat(p.rt, k));
          double P6d[6];
#pragma unroll
          for (int g = 0; g < 6; ++g) P6d[g] = rd * p.s6d[g];
          cpx<double> m3d[3], dd[NT], cov6d[6], resd[NT];
#pragma unroll
          for (int a = 0; a < 3; ++a) {
            const float2 v = buf[(a * N1 + k1) * TC + c];
            m3d[a] = mk<double>(v.x, v.y);
          }
#pragma unroll
          for (int l = 0; l < NT; ++l) {
            const float2 v = buf[((3 + l) * N1 + k1) * TC + c];
            dd[l] = mk<double>(v.x, v.y);
          }
          double cond64 = 0.0;
          const int fail64 = update<NT, double, true>(kat(p.tab64, k), rd, fabs((double)lt), P6d, m3d, dd, cov6d, resd,
                                                      false, cond64);
          if (fail64) {
            did_fail = true;  // rare: the FP64 factor failed too
            write_prior = true;
          } else {
#pragma unroll
            for (int g = 0; g < 6; ++g) cov6[g] = mk<float>((float)cov6d[g].re, (float)cov6d[g].im);
            write_cov = true;
#pragma unroll
            for (int a = 0; a < 3; ++a) buf[(a * N1 + k1) * TC + c] = make_float2((float)m3d[a].re, (float)m3d[a].im);
          }
        }
      }
      if (write_cov) {
#pragma unroll
        for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2(cov6[g].re, cov6[g].im));
      } else if (write_prior) {
#pragma unroll
        for (int g = 0; g < 6; ++g) gst(p.cov[g] + idx, make_float2(P6[g], 0.f));
      }
    }
    if (MIXED && !REDO && defer) {
      const float mark = __uint_as_float(kDeferMark);
#pragma unroll
      for (int a = 0; a < 3; ++a) buf[(a * N1 + k1) * TC + c] = make_float2(mark, mark);
    }
    if (p.stats) {
      if (!REDO) wave_count(p.stats, did_inv);
      // REDO counts only the re-solved cells' outcome; their tile's other cells were counted
      // by the first pass
      wave_count(p.stats + 64, REDO ? (defer && did_fail) : did_fail);
    }
    if (REDO) {
      const unsigned long long m = __ballot(defer);
      if (m && __lane_id() == 0) atomicAdd(p.list_len, (unsigned long long)__popcll(m));
    }
  }
  lds_sync();
  // 4. inverse z stage 2 of the means: one (mean, column) per thread, the column's N1 planes
  // stored with lanes on consecutive cells.  A column holding a deferred wavenumber keeps its
  // prior means in HBM; the wave of mean 0 lists the tile for the FP64 pass.
  for (int s = tid; s < 3 * TC; s += NTH) {
    const int a = s / TC, cl = s % TC;
    float2 w[N1];
    bool marked = false;
#pragma unroll
    for (int n = 0; n < N1; ++n) {
      w[n] = buf[(a * N1 + n) * TC + cl];
      if (MIXED && !REDO) marked |= __float_as_uint(w[n].x) == kDeferMark;
    }
    if (MIXED && !REDO) {
      const unsigned long long m = __ballot(marked);
      if (a == 0 && m && __lane_id() == 0) z.redo[atomicAdd(z.redo_len, 1ull)] = t;
      if (marked) continue;
    }
    dft<+1, N1>(w);
    if (c0 + cl < plane) {
      float2 *dst = p.mean[a] + (long)N1 * k2 * plane + c0 + cl;
#pragma unroll
      for (int n = 0; n < N1; ++n) gst(dst + (long)n * plane, w[n]);
    }
  }
}

template <int NT, int N1, bool MIXED>
__global__ void __launch_bounds__(z_threads<N1>()) k_posterior_z(cg::PostZArgs z)
{
  __shared__ __attribute__((aligned(16))) float2 buf[(3 + NT) * N1 * kZCols];
  post_z_tile<NT, N1, MIXED, false>(z, cg_xcd_remap(blockIdx.x, gridDim.x), buf);
}

// The deferred tiles again, FP64 where the first pass deferred (workgroups stride over the list;
// the loop bound is the same for every thread of a workgroup).
template <int NT, int N1>
__global__ void __launch_bounds__(z_threads<N1>()) k_posterior_z_redo(cg::PostZArgs z)
{
  __shared__ __attribute__((aligned(16))) float2 buf[(3 + NT) * N1 * kZCols];
  const unsigned long long n = *z.redo_len;
  for (unsigned long long i = blockIdx.x; i < n; i += gridDim.x) {
    post_z_tile<NT, N1, true, true>(z, z.redo[i], buf);
    lds_sync();  // the tile's LDS is read to the end before the next tile's rows arrive
  }
}

template <int NT, int N1>
int launch_z_nt(cg_ctx *ctx, const cg::PostZArgs &z, unsigned tiles, bool mixed)
{
  if (mixed) {
    k_posterior_z<NT, N1, true><<<tiles, z_threads<N1>(), 0, ctx->stream>>>(z);
    int rc = cg::hip_fail(ctx, hipGetLastError(), "k_posterior_z");
    if (rc) return rc;
    const unsigned blocks = tiles < 1024u ? tiles : 1024u;
    k_posterior_z_redo<NT, N1><<<blocks, z_threads<N1>(), 0, ctx->stream>>>(z);
    return cg::hip_fail(ctx, hipGetLastError(), "k_posterior_z_redo");
  }
  k_posterior_z<NT, N1, false><<<tiles, z_threads<N1>(), 0, ctx->stream>>>(z);
  return cg::hip_fail(ctx, hipGetLastError(), "k_posterior_z");
}

// Realisations with the split z transform (cg_simulate / cg_simulate_pair on such grids):
// colouring of the N1 planes k2 + N2 k1 (one k2) x 64 cells into LDS -- wave w takes
// k1 = w, w + W, ...; the same per-cell functions as the unsplit colouring kernels, so the noise
// and the coloured values are those of any other path -- then the inverse stage 2 over k1 into
// planes n1 + N1 k2 of the output grids.  The y pass and the inverse stage 1 with the x c2r (and
// the real-domain mean) follow.  PAIR: realisations stream0 and stream1, out and out2.
// MEAN: posterior means in the same split layout (the posterior pass's output before the means'
// own inverse y and x passes), mean_split, are added at the store; their tile rows are requested
// at the start and wait in LDS, so the reads overlap the colouring.
struct SplitMeans {
  const float2 *m[3];
};

template <int N1, bool PAIR, bool THREE, bool MEAN>
__global__ void __launch_bounds__(z_threads<N1>())
k_colour_z2(cg::ColourParams p, int n2, unsigned ctiles, SplitMeans mean_split)
{
  constexpr int TC = kZCols, NTH = z_threads<N1>(), W = NTH / 64, S = PAIR ? 6 : 3;
  static_assert((3 * N1 * TC) % NTH == 0, "the means' tile rows must be whole rounds of the threads");
  constexpr int MPER = 3 * N1 * TC / NTH;
  __shared__ float2 buf[S * N1 * TC];
  __shared__ float2 mbuf[MEAN ? 3 * N1 * TC : 1];  // [3][N1][64]: the means' tile rows
  const int tid = threadIdx.x;
  const unsigned t = cg_xcd_remap(blockIdx.x, gridDim.x);
  const unsigned k2 = t / ctiles, ct = t % ctiles;
  const long plane = (long)p.cnxp * p.nyp;
  const long c0 = (long)ct * TC;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6), c = tid & 63;
  const bool active = c0 + c < plane;
  // the means' rows (planes n1 + N1 k2) requested first, in flight during the colouring
  float2 mv[MEAN ? MPER : 1];
  if (MEAN) {
#pragma unroll
    for (int i = 0; i < MPER; ++i) {
      const int e = i * NTH + tid;
      const int g = e / (N1 * TC), n1 = (e / TC) % N1, cl = e % TC;
      mv[MEAN ? i : 0] = c0 + cl < plane ? gld(mean_split.m[g] + (long)(n1 + N1 * (long)k2) * plane + c0 + cl)
                                        : make_float2(0.f, 0.f);
    }
  }
#pragma unroll 1
  for (int k1 = wv; k1 < N1; k1 += W) {
    const long idx = (long)((int)k2 + n2 * k1) * plane + (active ? c0 + c : 0);
    float2 o[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    float2 o2[3] = {make_float2(0.f, 0.f), make_float2(0.f, 0.f), make_float2(0.f, 0.f)};
    bool failed = false;
    if (active) {
      if (PAIR) {
        failed = cgcol::colour_cell2t<THREE>(p, idx, o, o2);
      } else {
        float2 zn[3];
        failed = cgcol::colour_cell(p, idx, zn, o);
      }
    }
    if (p.fails) {  // one atomic per wavefront; a failed cell zeroes the noise of both of a pair
      const unsigned long long msk = __ballot(failed);
      if (failed && (int)__lane_id() == __ffsll((long long)msk) - 1)
        atomicAdd(p.fails, (PAIR ? 2ull : 1ull) * (unsigned long long)__popcll(msk));
    }
#pragma unroll
    for (int g = 0; g < 3; ++g) {
      buf[(g * N1 + k1) * TC + c] = o[g];
      if (PAIR) buf[((g + 3) * N1 + k1) * TC + c] = o2[g];
    }
  }
  if (MEAN) {
#pragma unroll
    for (int i = 0; i < MPER; ++i) mbuf[i * NTH + tid] = mv[MEAN ? i : 0];
  }
  lds_sync();
  // inverse stage 2 of the coloured columns (+ the split-layout means)
  for (int s = tid; s < S * TC; s += NTH) {
    const int g = s / TC, cl = s % TC, gm = g < 3 ? g : g - 3;
    float2 w[N1];
#pragma unroll
    for (int n = 0; n < N1; ++n) w[n] = buf[(g * N1 + n) * TC + cl];
    dft<+1, N1>(w);
    if (c0 + cl < plane) {
      float2 *dst = (g < 3 ? p.out[g] : p.out2[g - 3]) + (long)N1 * k2 * plane + c0 + cl;
#pragma unroll
      for (int n = 0; n < N1; ++n) {
        if (MEAN) {
          const float2 m = mbuf[(gm * N1 + n) * TC + cl];
          w[n].x += m.x;
          w[n].y += m.y;
        }
        gst(dst + (long)n * plane, w[n]);
      }
    }
  }
}

}  // namespace cgpost

namespace {

// Compiled splits of the split-z inversion: x half-length N (nxp = 2N), N1, N2 (nzp = N1 N2).
struct XZEntry {
  int n, n1, n2;
  int (*launch)(cg_ctx *, int, const cgfix::XZArgs &, int, const float2 *, const float2 *, const float2 *);
};
// (nzp 320 was also measured split 20 x 16: the posterior pass at 20 planes per tile holds one
// workgroup per CU and took 1.87 ms against 1.56 ms, the 16-row x passes 224 us against 215 us.)
const XZEntry kXZ[] = {
    {288, 16, 20, &cgfix::xz_launch<288, 20>},
    {576, 16, 24, &cgfix::xz_launch<576, 24>},
};

const XZEntry *xz_find(int nxp, int nzp)
{
  if (nxp % 2) return nullptr;
  for (const XZEntry &e : kXZ)
    if (e.n == nxp / 2 && e.n1 * e.n2 == nzp) return &e;
  return nullptr;
}

}  // namespace

namespace cg {

bool split_z_supported(int nxp, int nzp, int *n1, int *n2)
{
  const XZEntry *e = xz_find(nxp, nzp);
  if (!e) return false;
  if (n1) *n1 = e->n1;
  if (n2) *n2 = e->n2;
  return true;
}

int fft_xz(cg_ctx *ctx, int forward, const float *src, float *grid, const float *add, int nxp, int nyp, int nzp,
           float scale)
{
  const XZEntry *e = xz_find(nxp, nzp);
  if (!e) return fail(ctx, CG_EINVAL, "split-z x pass: no compiled split for this grid");
  const float2 *tw = twiddle_table(ctx, e->n);
  const float2 *twr = twiddle_table(ctx, nxp);
  const float2 *twz = twiddle_table(ctx, nzp);
  if (!tw || !twr || !twz) return CG_ENOMEM;
  cgfix::XZArgs a{forward ? src : nullptr, grid, forward ? nullptr : add, nyp, e->n1, scale};
  hipEvent_t ev;
  launch_begin(ctx, CG_K_FFT_X, &ev);
  const int rc = e->launch(ctx, forward, a, nzp, tw, twr, twz);
  launch_end(ctx, CG_K_FFT_X, ev);
  return rc;
}

template <int N1>
static int launch_z_n1(cg_ctx *ctx, const PostZArgs &z, unsigned tiles, bool mixed)
{
  switch (z.p.nt) {
    case 1: return cgpost::launch_z_nt<1, N1>(ctx, z, tiles, mixed);
    case 2: return cgpost::launch_z_nt<2, N1>(ctx, z, tiles, mixed);
    case 3: return cgpost::launch_z_nt<3, N1>(ctx, z, tiles, mixed);
    case 4: return cgpost::launch_z_nt<4, N1>(ctx, z, tiles, mixed);
    case 5: return cgpost::launch_z_nt<5, N1>(ctx, z, tiles, mixed);
    case 6: return cgpost::launch_z_nt<6, N1>(ctx, z, tiles, mixed);
    case 7: return cgpost::launch_z_nt<7, N1>(ctx, z, tiles, mixed);
    default: return cgpost::launch_z_nt<8, N1>(ctx, z, tiles, mixed);
  }
}

int launch_posterior_z(cg_ctx *ctx, const PostZArgs &z, int n1, bool mixed)
{
  const long plane = (long)z.p.cnxp * z.p.nyp;
  const unsigned long tiles = (unsigned long)z.n2 * (unsigned long)((plane + cgpost::kZCols - 1) / cgpost::kZCols);
  if (tiles >= (1ul << 31) || n1 != 16) return fail(ctx, CG_EINVAL, "split-z posterior: unsupported split");
  hipEvent_t ev;
  launch_begin(ctx, CG_K_POSTERIOR, &ev);
  const int rc = launch_z_n1<16>(ctx, z, (unsigned)tiles, mixed);
  launch_end(ctx, CG_K_POSTERIOR, ev);
  return rc;
}


int colour_split(cg_ctx *ctx, const ColourParams &p, int n1, int n2, bool pair, const float2 *const *mean_split)
{
  cgpost::SplitMeans ms{{nullptr, nullptr, nullptr}};
  for (int g = 0; g < 3 && mean_split; ++g) ms.m[g] = mean_split[g];
  const long plane = (long)p.cnxp * p.nyp;
  const unsigned long ctiles = (unsigned long)((plane + cgpost::kZCols - 1) / cgpost::kZCols);
  const unsigned long tiles = ctiles * (unsigned long)n2;
  if (tiles >= (1ul << 31) || n1 != 16) return fail(ctx, CG_EINVAL, "split-z colouring: unsupported split");
  constexpr int NTH = cgpost::z_threads<16>();
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  const unsigned nb = (unsigned)tiles, ctl = (unsigned)ctiles;
  const bool three = pair && cgcol::pair_in_three_blocks(p);
  if (ms.m[0]) {
    if (!pair) cgpost::k_colour_z2<16, false, false, true><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
    else if (three) cgpost::k_colour_z2<16, true, true, true><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
    else cgpost::k_colour_z2<16, true, false, true><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
  } else {
    if (!pair) cgpost::k_colour_z2<16, false, false, false><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
    else if (three) cgpost::k_colour_z2<16, true, true, false><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
    else cgpost::k_colour_z2<16, true, false, false><<<nb, NTH, 0, ctx->stream>>>(p, n2, ctl, ms);
  }
  launch_end(ctx, CG_K_NOISE, ev);
  return hip_fail(ctx, hipGetLastError(), "k_colour_z2");
}

}  // namespace cg
