// cg_fft_fixed.hip -- dispatch to the compile-time-plan FFT passes (cg_fft_fixed.h), which
// cg_fft_fixed_part.hip instantiates for every length of cg_fft_plans.h.  Lengths without a
// compiled plan fall back to the runtime-plan kernels of cg_fft.hip.
#include <array>
#include <map>
#include <mutex>

#include "cg_fft_plans.h"
#include "cg_internal.h"
#include "cg_noise.h"

namespace cg {

const FixedEntry *fixed_part_0(int *count);
const FixedEntry *fixed_part_1(int *count);
const FixedEntry *fixed_part_2(int *count);
const FixedEntry *fixed_part_3(int *count);
const FixedEntry *fixed_part_4(int *count);
const FixedEntry *fixed_part_5(int *count);
const FixedEntry *fixed_part_6(int *count);
const FixedEntry *fixed_part_7(int *count);
static_assert(CG_FFT_NPARTS == 8, "one declaration per part");

namespace {

struct Registry {
  std::map<int, const FixedEntry *> by[FIX_KINDS];
  Registry()
  {
    const FixedEntry *(*parts[])(int *) = {fixed_part_0, fixed_part_1, fixed_part_2, fixed_part_3,
                                           fixed_part_4, fixed_part_5, fixed_part_6, fixed_part_7};
    for (auto part : parts) {
      int n = 0;
      const FixedEntry *e = part(&n);
      for (int i = 0; i < n; ++i) by[e[i].kind][e[i].n] = &e[i];
    }
  }
};

const Registry &registry()
{
  static const Registry r;  // thread-safe initialisation
  return r;
}

// Autotuned tile variants per (device, kind, length, pass geometry), shared by the contexts of a
// device.
using TuneKey = std::array<long, 6>;
std::mutex tune_mu;
std::map<TuneKey, int> tune_table;

TuneKey tune_key(const cg_ctx *ctx, const FixedEntry *e, long g0, long g1, long g2)
{
  return TuneKey{(long)ctx->device, (long)e->kind, (long)e->n, g0, g1, g2};
}

}  // namespace

int fixed_variant(cg_ctx *ctx, const FixedEntry *e, long g0, long g1, long g2)
{
  if (ctx->fft_variant == 0 || ctx->fft_variant == 1) return ctx->fft_variant;
  std::lock_guard<std::mutex> lock(tune_mu);
  const auto it = tune_table.find(tune_key(ctx, e, g0, g1, g2));
  return it == tune_table.end() ? 0 : it->second;
}

const FixedEntry *fixed_find(int kind, int n)
{
  if (kind < 0 || kind >= FIX_KINDS) return nullptr;
  const auto &m = registry().by[kind];
  const auto it = m.find(n);
  return it == m.end() ? nullptr : it->second;
}

int fixed_strided(cg_ctx *ctx, float2 *data, int n, long S, long inner, long outer, long OS, int dir, float scale,
                  int *rc)
{
  const FixedEntry *e = fixed_find(FIX_STRIDED, n);
  if (!e) return 0;
  const unsigned long ncols = (unsigned long)inner * (unsigned long)outer;
  if (ncols >= (1ul << 31)) {
    *rc = fail(ctx, CG_EINVAL, "strided FFT: too many columns");
    return 1;
  }
  const float2 *tw = twiddle_table(ctx, n);
  if (!tw) {
    *rc = CG_ENOMEM;
    return 1;
  }
  StridedArgs a{data, S, OS, (unsigned)inner, (unsigned)ncols, 0u, scale};
  const int v = fixed_variant(ctx, e, S, inner, (long)ncols);
  const unsigned nblk = (unsigned)((ncols + e->tile[v] - 1) / e->tile[v]);
  *rc = e->s[v](ctx, dir, &a, nullptr, 0, nblk, tw);
  return 1;
}

int fixed_x(cg_ctx *ctx, int forward, const float *src, float *grid, int nxp, long rows, float scale, const float *add,
            int *rc)
{
  const bool even = nxp % 2 == 0;
  const FixedEntry *e = fixed_find(even ? FIX_XE : FIX_XO, even ? nxp / 2 : nxp);
  if (!e) return 0;
  const float2 *tw = twiddle_table(ctx, e->n);
  const float2 *twr = even ? twiddle_table(ctx, nxp) : nullptr;
  if (!tw || (even && !twr)) {
    *rc = CG_ENOMEM;
    return 1;
  }
  const int v = fixed_variant(ctx, e, rows, forward, add != nullptr);
  const unsigned long nblk = ((unsigned long)rows + e->tile[v] - 1) / e->tile[v];
  if (nblk >= (1ul << 31)) {
    *rc = fail(ctx, CG_EINVAL, "x FFT: too many rows");
    return 1;
  }
  XArgs a{forward ? src : nullptr, grid, forward ? nullptr : add, rows, 0u, scale};
  *rc = e->x[v](ctx, forward, !forward && add, &a, nullptr, 0, (unsigned)nblk, tw, twr);
  return 1;
}

int fixed_strided_list(cg_ctx *ctx, const FixedEntry *e, int v, int dir, const StridedArgs *dev_items, int nitems,
                       unsigned nblocks)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  const float2 *tw = twiddle_table(ctx, e->n);
  if (!tw) return CG_ENOMEM;
  return e->s[v](ctx, dir, nullptr, dev_items, nitems, nblocks, tw);
}

int fixed_x_list(cg_ctx *ctx, const FixedEntry *e, int v, int forward, int add, const XArgs *dev_items, int nitems,
                 unsigned nblocks, int nxp)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  const bool even = e->kind == FIX_XE;
  const float2 *tw = twiddle_table(ctx, e->n);
  const float2 *twr = even ? twiddle_table(ctx, nxp) : nullptr;
  if (!tw || (even && !twr)) return CG_ENOMEM;
  return e->x[v](ctx, forward, !forward && add, nullptr, dev_items, nitems, nblocks, tw, twr);
}

// Tune-table key of the fused-or-not choice for a realisation of nzp planes of `plane` cells.
static TuneKey colour_key(const cg_ctx *ctx, int nzp, long plane)
{
  return TuneKey{(long)ctx->device, (long)FIX_KINDS, (long)nzp, plane, 0, 0};
}

const FixedEntry *fixed_colour_entry(cg_ctx *ctx, int nzp, long plane)
{
  const FixedEntry *e = fixed_find(FIX_STRIDED, nzp);
  if (!e || !e->cz || ctx->fused_colour == 0 || ctx->fft_engine != CG_FFT_AUTO || nzp < 2) return nullptr;
  if (ctx->fused_colour != 1) {
    std::lock_guard<std::mutex> lock(tune_mu);
    const auto it = tune_table.find(colour_key(ctx, nzp, plane));
    if (it != tune_table.end() ? it->second == 0 : e->cz_waves < 12) return nullptr;
  }
  return e;
}

int fixed_colour_zinv(cg_ctx *ctx, const ColourParams &p, int nzp, long plane, int *rc)
{
  const FixedEntry *e = p.from_grids ? nullptr : fixed_colour_entry(ctx, nzp, plane);
  if (!e) return 0;
  const float2 *tw = twiddle_table(ctx, nzp);
  if (!tw) {
    *rc = CG_ENOMEM;
    return 1;
  }
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  *rc = e->cz(ctx, &p, nullptr, 0, (unsigned)((plane + e->cz_tile - 1) / e->cz_tile), tw);
  launch_end(ctx, CG_K_NOISE, ev);
  return 1;
}

// Tune-table key of the fused-or-not choice for a pair of realisations.
static TuneKey colour2_key(const cg_ctx *ctx, int nzp, long plane)
{
  return TuneKey{(long)ctx->device, (long)FIX_KINDS + 1, (long)nzp, plane, 0, 0};
}

// The paired colouring is fused where the autotuner timed it faster than the colouring kernel +
// six z passes for this geometry; untuned, where its tile has at least 8 columns (64-byte row
// segments; narrower tiles, at the longest lengths, read the covariance too finely).
const FixedEntry *fixed_colour2_entry(cg_ctx *ctx, int nzp, long plane)
{
  const FixedEntry *e = fixed_find(FIX_STRIDED, nzp);
  if (!e || !e->cz2 || ctx->fused_colour == 0 || ctx->fft_engine != CG_FFT_AUTO || nzp < 2) return nullptr;
  if (ctx->fused_colour != 1) {
    std::lock_guard<std::mutex> lock(tune_mu);
    const auto it = tune_table.find(colour2_key(ctx, nzp, plane));
    if (it != tune_table.end() ? it->second == 0 : e->cz2_tile < 8) return nullptr;
  }
  return e;
}

int fixed_colour_zinv2(cg_ctx *ctx, const ColourParams &p, int nzp, long plane, int *rc)
{
  const FixedEntry *e = p.from_grids ? nullptr : fixed_colour2_entry(ctx, nzp, plane);
  if (!e) return 0;
  const float2 *tw = twiddle_table(ctx, nzp);
  if (!tw) {
    *rc = CG_ENOMEM;
    return 1;
  }
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  *rc = e->cz2(ctx, &p, nullptr, 0, (unsigned)((plane + e->cz2_tile - 1) / e->cz2_tile), tw);
  launch_end(ctx, CG_K_NOISE, ev);
  return 1;
}

int fixed_colour_zinv2_list(cg_ctx *ctx, const FixedEntry *e, const ColourItem *dev_items, int nitems,
                            unsigned nblocks)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  if (!e->cz2) return fail(ctx, CG_EINVAL, "paired colouring: length without a paired kernel");
  const float2 *tw = twiddle_table(ctx, e->n);
  if (!tw) return CG_ENOMEM;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  const int rc = e->cz2(ctx, nullptr, dev_items, nitems, nblocks, tw);
  launch_end(ctx, CG_K_NOISE, ev);
  return rc;
}

int fixed_colour_zinv_list(cg_ctx *ctx, const FixedEntry *e, const ColourItem *dev_items, int nitems,
                           unsigned nblocks)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  const float2 *tw = twiddle_table(ctx, e->n);
  if (!tw) return CG_ENOMEM;
  hipEvent_t ev;
  launch_begin(ctx, CG_K_NOISE, &ev);
  const int rc = e->cz(ctx, nullptr, dev_items, nitems, nblocks, tw);
  launch_end(ctx, CG_K_NOISE, ev);
  return rc;
}

// Times both tile variants of every pass of a grid of padded size nxp x nyp x nzp (x forward
// out of place, x inverse with and without the added grid, y and z forward + inverse) on scratch
// grids and records the faster for that geometry.
int autotune_geometry(cg_ctx *ctx, int nxp, int nyp, int nzp)
{
  const long cnxp = nxp / 2 + 1, plane = cnxp * nyp, csize = plane * nzp;
  float2 *A = nullptr, *B = nullptr;
  hipError_t e = hipMalloc(&A, sizeof(float2) * csize);
  if (e == hipSuccess) e = hipMalloc(&B, sizeof(float2) * csize);
  if (e == hipSuccess) e = hipMemsetAsync(A, 0, sizeof(float2) * csize, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(B, 0, sizeof(float2) * csize, ctx->stream);
  hipEvent_t t0 = nullptr, t1 = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&t0);
  if (e == hipSuccess) e = hipEventCreate(&t1);
  int rc = e == hipSuccess ? CG_OK : hip_fail(ctx, e, "autotune scratch");
  // mean time of `reps` runs of f after one untimed run; negative on failure
  auto time_it = [&](auto f) -> float {
    if (f()) return -1.f;
    const int reps = 3;
    if (hipEventRecord(t0, ctx->stream) != hipSuccess) return -1.f;
    for (int i = 0; i < reps; ++i)
      if (f()) return -1.f;
    if (hipEventRecord(t1, ctx->stream) != hipSuccess || hipEventSynchronize(t1) != hipSuccess) return -1.f;
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    return ms / reps;
  };
  auto record = [&](const FixedEntry *en, long g0, long g1, long g2, const float (&ms)[2]) {
    if (ms[0] < 0.f || ms[1] < 0.f) return;
    std::lock_guard<std::mutex> lock(tune_mu);
    tune_table[tune_key(ctx, en, g0, g1, g2)] = ms[1] < ms[0] ? 1 : 0;
  };
  const bool even = nxp % 2 == 0;
  const FixedEntry *ex = fixed_find(even ? FIX_XE : FIX_XO, even ? nxp / 2 : nxp);
  // (The double-buffered even-nxp kernel is not tuned: on warm scratch grids the single-buffered
  // one times faster, in the transform sequence of an inversion it is slower.)
  if (!rc && ex && ex->tune && ex->tile[0] != ex->tile[1]) {
    const float2 *tw = twiddle_table(ctx, ex->n), *twr = even ? twiddle_table(ctx, nxp) : nullptr;
    const long rows = (long)nyp * nzp;
    for (int mode = 0; mode < 3 && tw; ++mode) {  // forward; inverse; inverse + add
      const int fwd = mode == 0;
      const float *add = mode == 2 ? reinterpret_cast<const float *>(B) : nullptr;
      float ms[2];
      for (int v = 0; v < 2; ++v) {
        XArgs a{fwd ? reinterpret_cast<const float *>(B) : nullptr, reinterpret_cast<float *>(A), add, rows, 0u,
                1.0f};
        const unsigned nblk = (unsigned)((rows + ex->tile[v] - 1) / ex->tile[v]);
        ms[v] = time_it([&] { return ex->x[v](ctx, fwd, add != nullptr, &a, nullptr, 0, nblk, tw, twr); });
      }
      record(ex, rows, fwd, add != nullptr, ms);
    }
  }
  for (int axis = 1; axis <= 2 && !rc; ++axis) {
    const int n = axis == 1 ? nyp : nzp;
    const FixedEntry *es = n > 1 ? fixed_find(FIX_STRIDED, n) : nullptr;
    if (!es || es->tile[0] == es->tile[1]) continue;
    const float2 *tw = twiddle_table(ctx, n);
    if (!tw) continue;
    const long S = axis == 1 ? cnxp : plane, inner = S, OS = axis == 1 ? plane : 0;
    const long ncols = axis == 1 ? cnxp * nzp : plane;
    float ms[2];
    for (int v = 0; v < 2; ++v) {
      StridedArgs a{A, S, OS, (unsigned)inner, (unsigned)ncols, 0u, 1.0f};
      const unsigned nblk = (unsigned)((ncols + es->tile[v] - 1) / es->tile[v]);
      ms[v] = time_it([&] {
        int r = es->s[v](ctx, -1, &a, nullptr, 0, nblk, tw);
        return r ? r : es->s[v](ctx, +1, &a, nullptr, 0, nblk, tw);
      });
    }
    record(es, S, inner, ncols, ms);
  }
  // A realisation's colouring and inverse z pass: fused, or the colouring kernel followed by
  // three z passes.  Posterior covariance spectra: the identity (diagonal A = 1, off-diagonal
  // B = 0), so every cell takes the factorisation's full path; three separate output grids.
  const FixedEntry *ez = nzp > 1 ? fixed_find(FIX_STRIDED, nzp) : nullptr;
  if (!rc && ez && ez->cz) {
    const float2 *tw = twiddle_table(ctx, nzp);
    float2 *out[3] = {nullptr, nullptr, nullptr}, *out2[3] = {nullptr, nullptr, nullptr};
    for (int g = 0; g < 3 && e == hipSuccess; ++g) e = hipMalloc(&out[g], sizeof(float2) * csize);
    if (e == hipSuccess)
      rc = fill_constant(ctx, reinterpret_cast<float *>(A), 2 * (size_t)csize, 1, 1, 1.0f);  // every float
    if (tw && e == hipSuccess && !rc) {
      ColourParams p{};
      p.csize = csize;
      p.cnxp = (int)cnxp;
      p.nyp = nyp;
      p.nzp = nzp;
      colour_dims(p);
      for (int g = 0; g < 6; ++g) p.cov[g] = g < 3 ? A : B;
      for (int g = 0; g < 3; ++g) p.out[g] = out[g];
      p.seed = 1;
      const int v = fixed_variant(ctx, ez, plane, plane, plane);
      StridedArgs sa[3];
      for (int g = 0; g < 3; ++g) sa[g] = StridedArgs{out[g], plane, 0, (unsigned)plane, (unsigned)plane, 0u, 1.0f};
      const unsigned nblk = (unsigned)((plane + ez->tile[v] - 1) / ez->tile[v]);
      float ms[2];
      ms[0] = time_it([&] {
        int r = colour(ctx, p);
        for (int g = 0; g < 3 && !r; ++g) r = ez->s[v](ctx, +1, &sa[g], nullptr, 0, nblk, tw);
        return r;
      });
      const unsigned czb = (unsigned)((plane + ez->cz_tile - 1) / ez->cz_tile);
      ms[1] = time_it([&] { return ez->cz(ctx, &p, nullptr, 0, czb, tw); });
      if (ms[0] >= 0.f && ms[1] >= 0.f) {
        std::lock_guard<std::mutex> lock(tune_mu);
        tune_table[colour_key(ctx, nzp, plane)] = ms[1] < ms[0] ? 1 : 0;
      }
      // two realisations per launch: the pair colouring kernel + six z passes, or fused
      for (int g = 0; g < 3 && e == hipSuccess && ez->cz2; ++g) e = hipMalloc(&out2[g], sizeof(float2) * csize);
      if (e == hipSuccess && ez->cz2) {
        for (int g = 0; g < 3; ++g) p.out2[g] = out2[g];
        p.stream1 = 3;
        StridedArgs sb[3];
        for (int g = 0; g < 3; ++g)
          sb[g] = StridedArgs{out2[g], plane, 0, (unsigned)plane, (unsigned)plane, 0u, 1.0f};
        ms[0] = time_it([&] {
          int r = colour2(ctx, p);
          for (int g = 0; g < 3 && !r; ++g) r = ez->s[v](ctx, +1, &sa[g], nullptr, 0, nblk, tw);
          for (int g = 0; g < 3 && !r; ++g) r = ez->s[v](ctx, +1, &sb[g], nullptr, 0, nblk, tw);
          return r;
        });
        const unsigned cz2b = (unsigned)((plane + ez->cz2_tile - 1) / ez->cz2_tile);
        ms[1] = time_it([&] { return ez->cz2(ctx, &p, nullptr, 0, cz2b, tw); });
        if (ms[0] >= 0.f && ms[1] >= 0.f) {
          std::lock_guard<std::mutex> lock(tune_mu);
          tune_table[colour2_key(ctx, nzp, plane)] = ms[1] < ms[0] ? 1 : 0;
        }
      }
    }
    (void)hipStreamSynchronize(ctx->stream);
    for (int g = 0; g < 3; ++g) {
      if (out[g]) (void)hipFree(out[g]);
      if (out2[g]) (void)hipFree(out2[g]);
    }
    if (e != hipSuccess && !rc) rc = hip_fail(ctx, e, "autotune scratch");
  }
  if (t0) (void)hipEventDestroy(t0);
  if (t1) (void)hipEventDestroy(t1);
  (void)hipStreamSynchronize(ctx->stream);
  if (A) (void)hipFree(A);
  if (B) (void)hipFree(B);
  return rc;
}

}  // namespace cg
