// One lane count's instantiations of the semisolid-pool / profiled-pulse
// kernel (every tangent mask, instantaneous or profiled pulses). Built once per lane count with TSIM_SEMISOLID_L defined.
#include "epg_semisolid_impl.h"

#ifndef TSIM_SEMISOLID_L
#error "build with -DTSIM_SEMISOLID_L=<lanes per atom>"
#endif

#define TSIM_CAT2(a, b) a##b
#define TSIM_CAT(a, b) TSIM_CAT2(a, b)

namespace tsim {

SemisolidLaunch TSIM_CAT(pick_semisolid_, TSIM_SEMISOLID_L)(int tm, bool profiled) {
  return pick_semisolid_mask<TSIM_SEMISOLID_L>(tm, profiled);
}

}  // namespace tsim
