// cg_fft_fixed_part.hip -- instantiates the compile-time-plan passes of one part of the length
// list in cg_fft_plans.h and exports them as a table; built once per part with -DCG_PART=<k>.
#include "cg_fft_fixed.h"

#ifndef CG_PART
#error "build with -DCG_PART=<part index>"
#endif

#define CG_CAT2(a, b) a##b
#define CG_CAT(a, b) CG_CAT2(a, b)

namespace {

// two tile variants per length and pass (the second may repeat the first where no other fits)
#define S(n)                                                                                  \
  {cg::FIX_STRIDED, n, {cgfix::strided_tile<n, 0>(), cgfix::strided_tile<n, 1>()},           \
   {&cgfix::strided_launch<n, 0>, &cgfix::strided_launch<n, 1>}, {nullptr, nullptr},         \
   &cgfix::colour_zinv_launch<n>, cgfix::colour_waves<n>(), cgfix::colour_tile<n>(), 1,     \
   &cgfix::colour_zinv2_launch<n>, cgfix::colour_tile2<n>()},
#define XE(n)                                                                        \
  {cg::FIX_XE, n, {cgfix::xe_rows<n, 0>(), cgfix::xe_rows<n, 1>()}, {nullptr, nullptr}, \
   {&cgfix::xe_launch<n, 0>, &cgfix::xe_launch<n, 1>}, nullptr, 0, 0, cgfix::x_rows_dma<n>() > 0 ? 0 : 1, nullptr, 0},
#define XO(n)                                                                        \
  {cg::FIX_XO, n, {cgfix::x_rows<n, 0>(), cgfix::x_rows<n, 1>()}, {nullptr, nullptr}, \
   {&cgfix::xo_launch<n, 0>, &cgfix::xo_launch<n, 1>}, nullptr, 0, 0, 1, nullptr, 0},
const cg::FixedEntry entries[] = {CG_CAT(CG_FFT_PART, CG_PART)(S, XE, XO)};
#undef S
#undef XE
#undef XO

}  // namespace

namespace cg {

const FixedEntry *CG_CAT(fixed_part_, CG_PART)(int *count)
{
  *count = (int)(sizeof(entries) / sizeof(entries[0]));
  return entries;
}

}  // namespace cg
