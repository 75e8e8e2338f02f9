// Complex EPG kernels (placeholder until the complex path lands).
#include "epg_common.h"

extern "C" int tsim_complex_simulate(void* const*, int64_t, int64_t, int64_t, int64_t, int64_t, int,
                                     int, int64_t, void*) {
  tsim::set_error("the complex (off-axis / exchanging-pool) kernel is not built yet");
  return tsim::UNSUPPORTED;
}
