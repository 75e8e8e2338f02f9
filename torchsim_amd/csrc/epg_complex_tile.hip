// One lane count's instantiations of the complex kernel. Built once per lane
// count with TSIM_COMPLEX_L defined, so they compile in parallel.
#include "epg_complex_impl.h"

#ifndef TSIM_COMPLEX_L
#error "build with -DTSIM_COMPLEX_L=<lanes per atom>"
#endif

#define TSIM_CAT2(a, b) a##b
#define TSIM_CAT(a, b) TSIM_CAT2(a, b)

namespace tsim {

ComplexLaunch TSIM_CAT(pick_complex_, TSIM_COMPLEX_L)(int tm, bool exch) {
  return pick_complex_mask<TSIM_COMPLEX_L>(tm, exch);
}

ComplexBatchLaunch TSIM_CAT(pick_complex_batch_, TSIM_COMPLEX_L)(int tm) {
  return pick_complex_batch_mask<TSIM_COMPLEX_L>(tm);
}

}  // namespace tsim
