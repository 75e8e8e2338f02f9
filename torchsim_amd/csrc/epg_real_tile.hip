// One state tile's instantiations of the real-subspace kernel. Built once per
// tile with TSIM_TILE_ID / TSIM_TILE_SL / TSIM_TILE_L defined, so the tiles
// compile in parallel.
#include "epg_real_impl.h"

#ifndef TSIM_TILE_ID
#error "build with -DTSIM_TILE_ID=<n> -DTSIM_TILE_SL=<orders per lane> -DTSIM_TILE_L=<lanes per atom>"
#endif

#define TSIM_CAT2(a, b) a##b
#define TSIM_CAT(a, b) TSIM_CAT2(a, b)

namespace tsim {

RealLaunch TSIM_CAT(pick_real_tile_, TSIM_TILE_ID)(int tm, int variant) {
  return pick_real_kernel<TSIM_TILE_SL, TSIM_TILE_L>(tm, variant);
}

BatchLaunch TSIM_CAT(pick_batch_tile_, TSIM_TILE_ID)(int tm, int variant) {
  return pick_batch_kernel<TSIM_TILE_SL, TSIM_TILE_L>(tm, variant);
}

}  // namespace tsim
