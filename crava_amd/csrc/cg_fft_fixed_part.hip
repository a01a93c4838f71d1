// cg_fft_fixed_part.hip -- instantiates the compile-time-plan passes of one part of the length
// list in cg_fft_plans.h and exports them as a table; built once per part with -DCG_PART=<k>.
#include "cg_fft_fixed.h"

#ifndef CG_PART
#error "build with -DCG_PART=<part index>"
#endif

#define CG_CAT2(a, b) a##b
#define CG_CAT(a, b) CG_CAT2(a, b)

namespace {

#define S(n) {cg::FIX_STRIDED, n, cgfix::strided_tile<n>(), &cgfix::strided_launch<n>, nullptr},
#define XE(n) {cg::FIX_XE, n, cgfix::x_rows<n>(), nullptr, &cgfix::xe_launch<n>},
#define XO(n) {cg::FIX_XO, n, cgfix::x_rows<n>(), nullptr, &cgfix::xo_launch<n>},
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
