// cg_fft_fixed.hip -- dispatch to the compile-time-plan FFT passes (cg_fft_fixed.h), which
// cg_fft_fixed_part.hip instantiates for every length of cg_fft_plans.h.  Lengths without a
// compiled plan fall back to the runtime-plan kernels of cg_fft.hip.
#include <map>

#include "cg_fft_plans.h"
#include "cg_internal.h"

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

}  // namespace

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
  const unsigned nblk = (unsigned)((ncols + e->tile - 1) / e->tile);
  *rc = e->s(ctx, dir, &a, nullptr, 0, nblk, tw);
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
  const unsigned long nblk = ((unsigned long)rows + e->tile - 1) / e->tile;
  if (nblk >= (1ul << 31)) {
    *rc = fail(ctx, CG_EINVAL, "x FFT: too many rows");
    return 1;
  }
  XArgs a{forward ? src : nullptr, grid, forward ? nullptr : add, rows, 0u, scale};
  *rc = e->x(ctx, forward, &a, nullptr, 0, (unsigned)nblk, tw, twr);
  return 1;
}

int fixed_strided_list(cg_ctx *ctx, const FixedEntry *e, int dir, const StridedArgs *dev_items, int nitems,
                       unsigned nblocks)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  const float2 *tw = twiddle_table(ctx, e->n);
  if (!tw) return CG_ENOMEM;
  return e->s(ctx, dir, nullptr, dev_items, nitems, nblocks, tw);
}

int fixed_x_list(cg_ctx *ctx, const FixedEntry *e, int forward, const XArgs *dev_items, int nitems, unsigned nblocks,
                 int nxp)
{
  if (nitems <= 0 || nblocks == 0) return CG_OK;
  const bool even = e->kind == FIX_XE;
  const float2 *tw = twiddle_table(ctx, e->n);
  const float2 *twr = even ? twiddle_table(ctx, nxp) : nullptr;
  if (!tw || (even && !twr)) return CG_ENOMEM;
  return e->x(ctx, forward, nullptr, dev_items, nitems, nblocks, tw, twr);
}

}  // namespace cg
