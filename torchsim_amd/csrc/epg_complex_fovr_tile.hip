// One lane count's instantiations of the complex forward-over-reverse kernel.
// Built once per lane count with TSIM_COMPLEX_FOVR_L defined, so they compile
// in parallel.
#include "epg_complex_fovr_impl.h"

#ifndef TSIM_COMPLEX_FOVR_L
#error "build with -DTSIM_COMPLEX_FOVR_L=<lanes per atom>"
#endif

#define TSIM_CAT2(a, b) a##b
#define TSIM_CAT(a, b) TSIM_CAT2(a, b)

namespace tsim {

ComplexFovrLaunch TSIM_CAT(pick_complex_fovr_, TSIM_COMPLEX_FOVR_L)(bool exch) {
  return exch ? &launch_complex_fovr<TSIM_COMPLEX_FOVR_L, true> : &launch_complex_fovr<TSIM_COMPLEX_FOVR_L, false>;
}

}  // namespace tsim
