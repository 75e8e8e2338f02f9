// One lane count's instantiations of the complex adjoint and JVP kernels. Built once
// per lane count with TSIM_COMPLEX_ADJOINT_L defined, so they compile in
// parallel.
#include "epg_complex_adjoint_impl.h"

#ifndef TSIM_COMPLEX_ADJOINT_L
#error "build with -DTSIM_COMPLEX_ADJOINT_L=<lanes per atom>"
#endif

#define TSIM_CAT2(a, b) a##b
#define TSIM_CAT(a, b) TSIM_CAT2(a, b)

namespace tsim {

ComplexAdjointLaunch TSIM_CAT(pick_complex_adjoint_, TSIM_COMPLEX_ADJOINT_L)(bool exch) {
  return exch ? &launch_complex_adjoint<TSIM_COMPLEX_ADJOINT_L, true>
              : &launch_complex_adjoint<TSIM_COMPLEX_ADJOINT_L, false>;
}

ComplexJvpLaunch TSIM_CAT(pick_complex_jvp_, TSIM_COMPLEX_ADJOINT_L)(bool exch) {
  return exch ? &launch_complex_jvp<TSIM_COMPLEX_ADJOINT_L, true> : &launch_complex_jvp<TSIM_COMPLEX_ADJOINT_L, false>;
}

}  // namespace tsim
