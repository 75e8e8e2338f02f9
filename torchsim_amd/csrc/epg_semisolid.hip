// Host side of the semisolid-pool kernel: argument checks, lane-count choice
// and the C ABI entry point (declared in include/torchsim_hip.h).
#include "epg_semisolid_adjoint_impl.h"

#include <cstring>

namespace tsim {
namespace {

SemisolidLaunch pick_semisolid(int lanes, int tm, bool profiled) {
  switch (lanes) {
#define TSIM_SEMISOLID_CASE(LANES) \
  case LANES:                      \
    return pick_semisolid_##LANES(tm, profiled);
    TSIM_COMPLEX_LANES(TSIM_SEMISOLID_CASE)
#undef TSIM_SEMISOLID_CASE
    default:
      return nullptr;
  }
}

int semisolid_lanes(int64_t states) {
  for (int l : {4, 8, 10, 12, 16, 20, 32, 64})
    if (states <= l) return l;
  return -1;
}

}  // namespace
}  // namespace tsim


namespace tsim {
namespace {

// Checks and arguments shared by the entry points. `bound` asks for the
// semisolid pool (its buffers and the lineshape required); a profile table
// makes the pulses profiled. Returns the tangent mask in `tm`.
int fill_shaped(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                int64_t output_count, int features, int64_t tissue_stride, const int32_t* columns,
                int64_t column_count, float factor_re, float factor_im, bool bound, const float* lineshape,
                int64_t lineshape_bins, float lineshape_step, const float* profile, int64_t locations,
                int64_t profile_bins, float profile_step, void* signal, void* jacobian, SemisolidArgs& a,
                int& tm) {
  if (packed == nullptr || packed[T_T1] == nullptr || packed[T_T2] == nullptr) {
    set_error("packed buffers with t1 and t2 are required");
    return BAD_ARGUMENT;
  }
  if (atom_count < 0 || event_count < 0 || state_count < 1 || output_count < 0) {
    set_error("invalid EPG buffer dimensions");
    return BAD_ARGUMENT;
  }
  if (state_count > 64) {
    set_error("at most 64 configuration orders are carried, %lld asked", (long long)state_count);
    return UNSUPPORTED;
  }
  if (tissue_stride != 0 && tissue_stride != 1) {
    set_error("tissue_stride must be 0 or 1");
    return BAD_ARGUMENT;
  }
  if (features & FEATURE_MOVING) {
    set_error("the semisolid and profiled kernel carries no flow term");
    return UNSUPPORTED;
  }
  if (packed[T_POOL_B_FRACTION] != nullptr) {
    set_error("an exchanging pool is not built beside a semisolid pool or a profiled pulse");
    return UNSUPPORTED;
  }
  if (bound) {
    if (packed[T_BOUND_FRACTION] == nullptr || packed[T_BOUND_EXCHANGE] == nullptr ||
        packed[T_T1_BOUND] == nullptr) {
      set_error("the semisolid pool needs its fraction, exchange rate and T1 buffers");
      return BAD_ARGUMENT;
    }
    if (lineshape == nullptr || lineshape_bins < 2 || !(lineshape_step > 0.0f)) {
      set_error("the lineshape table needs at least two knots and a positive spacing");
      return BAD_ARGUMENT;
    }
  }
  if (profile != nullptr) {
    if (locations < 1 || profile_bins < 2 || !(profile_step > 0.0f)) {
      set_error("the profile table needs at least one position, two flip knots and a positive spacing");
      return BAD_ARGUMENT;
    }
    if (atom_count % locations != 0) {
      set_error("atoms come one per slice position: %lld atoms for %lld positions", (long long)atom_count,
                (long long)locations);
      return BAD_ARGUMENT;
    }
  }
  if (column_count < 0 || column_count > 4 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 4 Jacobian columns, got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  memset(&a, 0, sizeof(a));
  a.t1 = (const float*)packed[T_T1];
  a.t2 = (const float*)packed[T_T2];
  a.m0 = (const float*)packed[T_M0];
  a.b1 = (const float*)packed[T_B1];
  a.b1_phase = (const float*)packed[T_B1_PHASE];
  a.b0 = (const float*)packed[T_B0];
  a.inv_eff = (const float*)packed[T_INV_EFF];
  a.diffusion = (const float*)packed[T_DIFFUSION];
  if (bound) {
    a.bound_fraction = (const float*)packed[T_BOUND_FRACTION];
    a.bound_exchange = (const float*)packed[T_BOUND_EXCHANGE];
    a.t1_bound = (const float*)packed[T_T1_BOUND];
  }
  a.saturation = (const float*)packed[TISSUE_COUNT + E_SATURATION];
  a.rf_frequency = (const float*)packed[TISSUE_COUNT + E_RF_FREQUENCY];
  if (bound && (a.saturation == nullptr || a.rf_frequency == nullptr)) {
    set_error("the saturation and rf_frequency event buffers are required");
    return BAD_ARGUMENT;
  }
  a.lineshape = lineshape;
  a.lineshape_bins = lineshape_bins;
  a.lineshape_step = lineshape_step;
  a.profile = profile;
  a.profile_bins = profile_bins;
  a.profile_step = profile_step;
  a.locations = profile ? locations : 1;
  a.stride_opt = tissue_stride;
  a.atoms = atom_count;
  a.outputs = output_count;
  a.states = (int)state_count;
  a.density = (features & FEATURE_DENSITY) && a.m0 ? 1 : 0;
  a.inverting = (features & FEATURE_INVERTING) && a.inv_eff ? 1 : 0;
  a.transmit = (features & FEATURE_TRANSMIT) && a.b1 ? 1 : 0;
  a.off_axis = (features & FEATURE_OFF_AXIS) ? 1 : 0;
  a.diffusing = (features & FEATURE_DIFFUSING) && a.diffusion ? 1 : 0;
  a.factor_re = factor_re;
  a.factor_im = factor_im;
  a.col_t1 = a.col_t2 = a.col_b1 = a.col_m0 = -1;
  tm = 0;
  for (int64_t c = 0; c < column_count; ++c) {
    int* slot;
    switch (columns[c]) {
      case T_T1: tm |= CT1; slot = &a.col_t1; break;
      case T_T2: tm |= CT2; slot = &a.col_t2; break;
      case T_B1: tm |= CB1; slot = &a.col_b1; break;
      case T_M0: slot = &a.col_m0; break;
      default:
        set_error("column %lld: property %d carries no tangent here", (long long)c, columns[c]);
        return BAD_ARGUMENT;
    }
    if (*slot >= 0) {
      set_error("property %d appears twice among the Jacobian columns", columns[c]);
      return BAD_ARGUMENT;
    }
    *slot = (int)c;
  }
  if (a.col_b1 >= 0 && !a.transmit) {
    set_error("a B1 column needs the transmit feature and a b1 buffer");
    return BAD_ARGUMENT;
  }
  a.columns = (int)column_count;
  a.signal = (float2*)signal;
  a.jacobian = column_count > 0 ? (float2*)jacobian : nullptr;
  return OK;
}

// Checks and launch of the forward entry points.
int run_shaped(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
               int64_t output_count, int features, int64_t tissue_stride, const int32_t* columns,
               int64_t column_count, float factor_re, float factor_im, bool bound, const float* lineshape,
               int64_t lineshape_bins, float lineshape_step, const float* profile, int64_t locations,
               int64_t profile_bins, float profile_step, void* signal, void* jacobian, void* stream) {
  SemisolidArgs a;
  int tm = 0;
  int st = fill_shaped(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, columns,
                       column_count, factor_re, factor_im, bound, lineshape, lineshape_bins, lineshape_step, profile,
                       locations, profile_bins, profile_step, signal, jacobian, a, tm);
  if (st != OK) return st;
  const bool sized = atom_count > 0 && output_count > 0;
  if (sized && (signal == nullptr || (column_count > 0 && jacobian == nullptr))) {
    set_error("the signal (and, with columns, the Jacobian) output is required");
    return BAD_ARGUMENT;
  }
  if (!sized) return OK;
  hipStream_t s = (hipStream_t)stream;
  const int L = semisolid_lanes(state_count);
  SemisolidLaunch launch = pick_semisolid(L, tm, profile != nullptr);
  if (!launch) {
    set_error("no kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  a.events = compiled.codes;
  a.event_count = event_count;
  const int apw = 64 / L;
  const int64_t waves = (atom_count + apw - 1) / apw;
  const int threads = waves < 4LL * 256 ? 64 : 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * apw;
  const int64_t blocks = (atom_count + atoms_per_block - 1) / atoms_per_block;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  launch(a, dim3((unsigned)blocks), dim3(threads), s);
  return hip_status(hipGetLastError(), "semisolid_kernel launch");
}

}  // namespace
}  // namespace tsim

using namespace tsim;

extern "C" int tsim_semisolid_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                                       int64_t state_count, int64_t output_count, int features,
                                       int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                                       float factor_re, float factor_im, const float* lineshape,
                                       int64_t lineshape_bins, float lineshape_step, void* signal, void* jacobian,
                                       void* stream) {
  return run_shaped(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, columns,
                    column_count, factor_re, factor_im, true, lineshape, lineshape_bins, lineshape_step, nullptr, 1,
                    0, 1.0f, signal, jacobian, stream);
}

extern "C" int tsim_profiled_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                                      int64_t state_count, int64_t output_count, int features, int64_t tissue_stride,
                                      const int32_t* columns, int64_t column_count, float factor_re, float factor_im,
                                      const float* profile, int64_t locations, int64_t profile_bins,
                                      float profile_step, const float* lineshape, int64_t lineshape_bins,
                                      float lineshape_step, void* signal, void* jacobian, void* stream) {
  if (profile == nullptr) {
    set_error("the profile table is required");
    return BAD_ARGUMENT;
  }
  const bool bound = packed != nullptr && packed[T_BOUND_FRACTION] != nullptr;
  return run_shaped(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, columns,
                    column_count, factor_re, factor_im, bound, lineshape, lineshape_bins, lineshape_step, profile,
                    locations, profile_bins, profile_step, signal, jacobian, stream);
}

namespace tsim {
namespace {

using SemisolidVjpLaunch = void (*)(const SemisolidVjpArgs&, dim3, dim3, hipStream_t);

template <int L, bool PROFILED>
void launch_semisolid_vjp(const SemisolidVjpArgs& a, dim3 grid, dim3 block, hipStream_t s) {
  hipLaunchKernelGGL((semisolid_vjp_kernel<L, PROFILED>), grid, block, 0, s, a);
}

SemisolidVjpLaunch pick_semisolid_vjp(int lanes, bool profiled) {
  switch (lanes) {
#define TSIM_SEMISOLID_VJP_CASE(LANES) \
  case LANES:                          \
    return profiled ? &launch_semisolid_vjp<LANES, true> : &launch_semisolid_vjp<LANES, false>;
    TSIM_COMPLEX_LANES(TSIM_SEMISOLID_VJP_CASE)
#undef TSIM_SEMISOLID_VJP_CASE
    default:
      return nullptr;
  }
}

// Bytes of entry states one launch keeps (the launch covers as many atoms:
// 3,570 atoms of an 880-pulse train, 4 events a pulse, at 10 orders).
constexpr int64_t kTrajectoryBytes = 4LL << 30;

}  // namespace
}  // namespace tsim

namespace tsim {
namespace {

int run_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count, int64_t output_count,
            int features, int64_t tissue_stride, float factor_re, float factor_im, bool bound, const float* lineshape,
            int64_t lineshape_bins, float lineshape_step, const float* profile, int64_t locations,
            int64_t profile_bins, float profile_step, const void* cotangent, float* grad_flip, float* grad_phase,
            float* grad_duration, void* stream) {
  SemisolidArgs a;
  int tm = 0;
  int st = fill_shaped(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, nullptr,
                       0, factor_re, factor_im, bound, lineshape, lineshape_bins, lineshape_step, profile, locations,
                       profile_bins, profile_step, nullptr, nullptr, a, tm);
  if (st != OK) return st;
  if (event_count > 0 && (grad_flip == nullptr || grad_phase == nullptr || grad_duration == nullptr)) {
    set_error("the flip, phase and duration gradient outputs are required");
    return BAD_ARGUMENT;
  }
  if (atom_count > 0 && output_count > 0 && cotangent == nullptr) {
    set_error("the signal cotangent is required");
    return BAD_ARGUMENT;
  }
  if (event_count == 0) return OK;
  hipStream_t s = (hipStream_t)stream;
  const unsigned sum_blocks = (unsigned)((3 * event_count + 255) / 256);
  if (atom_count == 0 || output_count == 0) {
    // nothing recorded: every gradient is zero (no waves to sum)
    hipLaunchKernelGGL(semisolid_sum_partials, dim3(sum_blocks), dim3(256), 0, s, (const float*)nullptr,
                       (int64_t)0, event_count, grad_flip, grad_phase, grad_duration, 0);
    return hip_status(hipGetLastError(), "semisolid gradient clear");
  }
  const int L = semisolid_lanes(state_count);
  SemisolidVjpLaunch launch = pick_semisolid_vjp(L, profile != nullptr);
  if (!launch) {
    set_error("no adjoint kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  const int64_t apw = 64 / L;
  // lanes one launch covers: its entry states within the budget, one wave at least
  int64_t wave_cap = kTrajectoryBytes / (8 * (int64_t)sizeof(float) * event_count * 64);
  if (wave_cap < 1) wave_cap = 1;
  const int64_t waves_all = (atom_count + apw - 1) / apw;
  const int64_t waves = waves_all < wave_cap ? waves_all : wave_cap;
  const int64_t threads = waves * 64;
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  a.events = compiled.codes;
  a.event_count = event_count;
  float* trajectory = nullptr;
  float* partial = nullptr;
  st = scratch_alloc((void**)&trajectory, sizeof(float) * 8 * (size_t)event_count * (size_t)threads, s,
                     "semisolid adjoint states");
  if (st != OK) return st;
  st = scratch_alloc((void**)&partial, sizeof(float) * 3 * (size_t)event_count * (size_t)waves, s,
                     "semisolid adjoint partials");
  if (st != OK) {
    (void)hipFreeAsync(trajectory, s);
    return st;
  }
  SemisolidVjpArgs q;
  q.p = a;
  q.cotangent = (const float2*)cotangent;
  q.trajectory = trajectory;
  q.partial = partial;
  q.threads = threads;
  const int64_t atoms_per_launch = waves * apw;
  int first = 1;
  for (int64_t begin = 0; begin < atom_count && st == OK; begin += atoms_per_launch) {
    const int64_t count = atom_count - begin < atoms_per_launch ? atom_count - begin : atoms_per_launch;
    const int64_t launch_waves = (count + apw - 1) / apw;
    q.first_atom = begin;
    q.launch_atoms = count;
    launch(q, dim3((unsigned)launch_waves), dim3(64), s);
    st = hip_status(hipGetLastError(), "semisolid_vjp_kernel launch");
    if (st != OK) break;
    hipLaunchKernelGGL(semisolid_sum_partials, dim3(sum_blocks), dim3(256), 0, s, (const float*)partial,
                       launch_waves, event_count, grad_flip, grad_phase, grad_duration, first ? 0 : 1);
    st = hip_status(hipGetLastError(), "semisolid gradient sum");
    first = 0;
  }
  (void)hipFreeAsync(partial, s);
  (void)hipFreeAsync(trajectory, s);
  return st;
}

}  // namespace
}  // namespace tsim

extern "C" int tsim_semisolid_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                                  int64_t output_count, int features, int64_t tissue_stride, float factor_re,
                                  float factor_im, const float* lineshape, int64_t lineshape_bins,
                                  float lineshape_step, const void* cotangent, float* grad_flip, float* grad_phase,
                                  float* grad_duration, void* stream) {
  return run_vjp(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, factor_re,
                 factor_im, true, lineshape, lineshape_bins, lineshape_step, nullptr, 1, 0, 1.0f, cotangent, grad_flip,
                 grad_phase, grad_duration, stream);
}

extern "C" int tsim_profiled_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                                 int64_t output_count, int features, int64_t tissue_stride, float factor_re,
                                 float factor_im, const float* profile, int64_t locations, int64_t profile_bins,
                                 float profile_step, const float* lineshape, int64_t lineshape_bins,
                                 float lineshape_step, const void* cotangent, float* grad_flip, float* grad_phase,
                                 float* grad_duration, void* stream) {
  if (profile == nullptr) {
    set_error("the profile table is required");
    return BAD_ARGUMENT;
  }
  const bool bound = packed != nullptr && packed[T_BOUND_FRACTION] != nullptr;
  return run_vjp(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, factor_re,
                 factor_im, bound, lineshape, lineshape_bins, lineshape_step, profile, locations, profile_bins,
                 profile_step, cotangent, grad_flip, grad_phase, grad_duration, stream);
}

namespace tsim {
namespace {

using EventJvpLaunch = void (*)(const SemisolidEventJvpArgs&, dim3, dim3, hipStream_t);

template <int L, bool PROFILED>
void launch_event_jvp(const SemisolidEventJvpArgs& a, dim3 grid, dim3 block, hipStream_t s) {
  hipLaunchKernelGGL((semisolid_event_jvp_kernel<L, PROFILED>), grid, block, 0, s, a);
}

EventJvpLaunch pick_event_jvp(int lanes, bool profiled) {
  switch (lanes) {
#define TSIM_EVENT_JVP_CASE(LANES) \
  case LANES:                      \
    return profiled ? &launch_event_jvp<LANES, true> : &launch_event_jvp<LANES, false>;
    TSIM_COMPLEX_LANES(TSIM_EVENT_JVP_CASE)
#undef TSIM_EVENT_JVP_CASE
    default:
      return nullptr;
  }
}

}  // namespace
}  // namespace tsim

extern "C" int tsim_semisolid_event_jvp(void* const* packed, int64_t atom_count, int64_t event_count,
                                        int64_t state_count, int64_t output_count, int features,
                                        int64_t tissue_stride, float factor_re, float factor_im, const float* profile,
                                        int64_t locations, int64_t profile_bins, float profile_step,
                                        const float* lineshape, int64_t lineshape_bins, float lineshape_step,
                                        const float* d_flip, const float* d_phase, const float* d_duration,
                                        void* tangent, void* stream) {
  const bool bound = packed != nullptr && packed[T_BOUND_FRACTION] != nullptr;
  SemisolidArgs a;
  int tm = 0;
  int st = fill_shaped(packed, atom_count, event_count, state_count, output_count, features, tissue_stride, nullptr,
                       0, factor_re, factor_im, bound, lineshape, lineshape_bins, lineshape_step, profile, locations,
                       profile_bins, profile_step, nullptr, nullptr, a, tm);
  if (st != OK) return st;
  if (!bound && profile == nullptr) {
    set_error("the event tangents run with a semisolid pool or a profile table");
    return BAD_ARGUMENT;
  }
  if (atom_count == 0 || output_count == 0) return OK;
  if (tangent == nullptr) {
    set_error("the tangent output is required");
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  const int L = semisolid_lanes(state_count);
  EventJvpLaunch launch = pick_event_jvp(L, profile != nullptr);
  if (!launch) {
    set_error("no kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  a.events = compiled.codes;
  a.event_count = event_count;
  SemisolidEventJvpArgs q;
  q.p = a;
  q.d_flip = d_flip;
  q.d_phase = d_phase;
  q.d_duration = d_duration;
  q.tangent = (float2*)tangent;
  const int64_t apw = 64 / L;
  const int64_t waves = (atom_count + apw - 1) / apw;
  if (waves > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  launch(q, dim3((unsigned)waves), dim3(64), s);
  return hip_status(hipGetLastError(), "semisolid_event_jvp_kernel launch");
}
