// Host side of the complex adjoint: argument checks, the checkpoint and
// scratch workspace, the launch, and the fixed-order reduction of the
// per-wave event partials (declared in include/torchsim_hip.h).
#include "epg_complex_fovr_impl.h"

#include <cmath>
#include <cstring>

namespace tsim {
namespace {

ComplexAdjointLaunch pick_complex_adjoint(int lanes, bool exch) {
  switch (lanes) {
#define TSIM_COMPLEX_ADJOINT_CASE(LANES) \
  case LANES:                            \
    return pick_complex_adjoint_##LANES(exch);
    TSIM_COMPLEX_LANES(TSIM_COMPLEX_ADJOINT_CASE)
#undef TSIM_COMPLEX_ADJOINT_CASE
    default:
      return nullptr;
  }
}

// Fewest lanes covering the orders (the forward complex kernel's rule).
int adjoint_lanes(int64_t states) {
  for (int l : {4, 8, 10, 12, 16, 20, 32, 64})
    if (states <= l) return l;
  return -1;
}

// Per-event partials summed over waves in a fixed order (added to what the
// outputs hold with `accumulate`).
__global__ void reduce_event_partials(const float* __restrict__ partial, int64_t waves, int64_t events,
                                      float* __restrict__ grad_flip, float* __restrict__ grad_phase,
                                      float* __restrict__ grad_duration, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= 3 * events) return;
  float v = 0.0f;
  for (int64_t w = 0; w < waves; ++w) v += partial[w * 3 * events + i];
  const int64_t e = i / 3;
  const int k = (int)(i % 3);
  float* out = k == 0 ? grad_flip : (k == 1 ? grad_phase : grad_duration);
  if (out) out[e] = accumulate ? out[e] + v : v;
}

// The signal cotangent an M0 column adds: its cotangent over M0 (the column
// is the signal over M0), joined to the signal's own (or alone).
__global__ void join_m0_cotangent(const float2* __restrict__ signal_cotangent, const float2* __restrict__ jacobian_cotangent,
                                  int64_t columns, int64_t column, const float* __restrict__ m0, int64_t stride,
                                  int density, int64_t atoms, int64_t outputs, float2* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= atoms * outputs) return;
  const int64_t atom = i / outputs, o = i - atom * outputs;
  const float scale = 1.0f / (density ? m0[atom * stride] : 1.0f);
  const float2 w = jacobian_cotangent[(atom * columns + column) * outputs + o];
  float2 v = make_float2(w.x * scale, w.y * scale);
  if (signal_cotangent) {
    v.x += signal_cotangent[i].x;
    v.y += signal_cotangent[i].y;
  }
  out[i] = v;
}

__global__ void clear_floats(float* __restrict__ a, int64_t count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) a[i] = 0.0f;
}

int clear(float* a, int64_t count, hipStream_t s) {
  if (a == nullptr || count <= 0) return OK;
  hipLaunchKernelGGL(clear_floats, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, a, count);
  return hip_status(hipGetLastError(), "gradient clear");
}

// Checks and fills what the adjoint and the JVP share (the packed buffers,
// sizes and feature flags); the event codes come later.
int fill_common(ComplexAdjointArgs& a, void* const* packed, int64_t atom_count, int64_t event_count,
                int64_t state_count, int64_t output_count, int pool_kind, int features, int64_t tissue_stride,
                float factor_re, float factor_im) {
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
  if (pool_kind != 0 && pool_kind != 2) {
    set_error("pool kind %d is not built (single pool 0 or one exchanging pool 2)", pool_kind);
    return UNSUPPORTED;
  }
  if (tissue_stride != 0 && tissue_stride != 1) {
    set_error("tissue_stride must be 0 or 1");
    return BAD_ARGUMENT;
  }
  if (features & FEATURE_MOVING) {
    set_error("the complex adjoint and JVP carry no flow term");
    return UNSUPPORTED;
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
  a.pool_fraction = (const float*)packed[T_POOL_B_FRACTION];
  a.pool_exchange = (const float*)packed[T_POOL_B_EXCHANGE];
  a.t1_pool = (const float*)packed[T_T1_POOL_B];
  a.t2_pool = (const float*)packed[T_T2_POOL_B];
  a.pool_shift = (const float*)packed[T_POOL_B_SHIFT];
  if (pool_kind == 2 && (!a.pool_fraction || !a.pool_exchange || !a.t1_pool || !a.t2_pool)) {
    set_error("an exchanging pool needs its fraction, exchange rate, T1 and T2 buffers");
    return BAD_ARGUMENT;
  }
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
  a.event_count = event_count;
  return OK;
}

// The CG slot a tissue property's gradient or tangent takes, or -1.
int carried_slot(int tissue) {
  switch (tissue) {
    case T_T1: return CG_T1;
    case T_T2: return CG_T2;
    case T_M0: return CG_M0;
    case T_B1: return CG_B1;
    case T_B1_PHASE: return CG_B1_PHASE;
    case T_B0: return CG_B0;
    case T_INV_EFF: return CG_IE;
    case T_DIFFUSION: return CG_D;
    default: return -1;
  }
}

ComplexFovrLaunch pick_complex_fovr(int lanes, bool exch) {
  switch (lanes) {
#define TSIM_COMPLEX_FOVR_CASE(LANES) \
  case LANES:                         \
    return pick_complex_fovr_##LANES(exch);
    TSIM_COMPLEX_LANES(TSIM_COMPLEX_FOVR_CASE)
#undef TSIM_COMPLEX_FOVR_CASE
    default:
      return nullptr;
  }
}

ComplexJvpLaunch pick_complex_jvp(int lanes, bool exch) {
  switch (lanes) {
#define TSIM_COMPLEX_JVP_CASE(LANES) \
  case LANES:                        \
    return pick_complex_jvp_##LANES(exch);
    TSIM_COMPLEX_LANES(TSIM_COMPLEX_JVP_CASE)
#undef TSIM_COMPLEX_JVP_CASE
    default:
      return nullptr;
  }
}

}  // namespace
}  // namespace tsim

using namespace tsim;

extern "C" int tsim_complex_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                                int64_t output_count, int pool_kind, int features, int64_t tissue_stride,
                                float factor_re, float factor_im, const void* signal_cotangent,
                                int64_t segment_events, float* grad_flip, float* grad_phase, float* grad_duration,
                                float* const* tissue_gradients, int64_t tissue_gradient_count, void* stream) {
  ComplexAdjointArgs a;
  int fill = fill_common(a, packed, atom_count, event_count, state_count, output_count, pool_kind, features,
                         tissue_stride, factor_re, factor_im);
  if (fill != OK) return fill;
  if (segment_events < 0) {
    set_error("segment_events must be positive (or 0 for the library's choice)");
    return BAD_ARGUMENT;
  }
  if (signal_cotangent == nullptr && atom_count > 0 && output_count > 0) {
    set_error("the signal cotangent is required");
    return BAD_ARGUMENT;
  }
  if (tissue_gradient_count != 0 && (tissue_gradients == nullptr || tissue_gradient_count != TISSUE_COUNT)) {
    set_error("tissue_gradients holds %d pointers (tissue order) or tissue_gradient_count is 0", TISSUE_COUNT);
    return BAD_ARGUMENT;
  }
  if (tissue_gradient_count == TISSUE_COUNT) {
    for (int q = 0; q < TISSUE_COUNT; ++q) {
      if (tissue_gradients[q] == nullptr) continue;
      const int slot = carried_slot(q);
      if (slot < 0) {
        set_error("tissue property %d carries no gradient in the complex adjoint", q);
        return UNSUPPORTED;
      }
      a.grad[slot] = tissue_gradients[q];
    }
  }
  a.cotangent = (const float2*)signal_cotangent;

  hipStream_t s = (hipStream_t)stream;
  int st = OK;
  for (float* g : {grad_flip, grad_phase, grad_duration})
    if (st == OK) st = clear(g, event_count, s);
  if (st != OK) return st;
  if (atom_count == 0 || event_count == 0 || output_count == 0) {
    for (int q = 0; q < CG_COUNT && st == OK; ++q) st = clear(a.grad[q], atom_count, s);
    return st;
  }
  const int L = adjoint_lanes(state_count);
  ComplexAdjointLaunch launch = pick_complex_adjoint(L, pool_kind == 2);
  if (!launch) {
    set_error("no complex adjoint kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  // one checkpoint about every sqrt(E) events by default: checkpoints and one
  // segment's entry states then take about the same space
  int64_t segment = segment_events;
  if (segment == 0) segment = (int64_t)std::ceil(std::sqrt((double)event_count));
  if (segment > event_count) segment = event_count;
  const int apw = 64 / L;
  const int64_t waves_needed = (atom_count + apw - 1) / apw;
  const int threads = waves_needed < 4LL * 256 ? 64 : 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * apw;
  const int64_t blocks = (atom_count + atoms_per_block - 1) / atoms_per_block;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  const int64_t total = blocks * threads;
  const int64_t waves = total / 64;
  const int64_t floats = 6LL * (pool_kind == 2 ? 2 : 1);
  const int64_t segments = (event_count + segment - 1) / segment;
  const size_t ck = sizeof(float) * (size_t)(segments * floats * total);
  const size_t sc = sizeof(float) * (size_t)(segment * floats * total);
  const size_t pa = sizeof(float) * (size_t)(waves * event_count * 3);
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  char* work = nullptr;
  st = scratch_alloc((void**)&work, ck + sc + pa, s, "complex adjoint workspace");
  if (st != OK) return st;
  a.events = compiled.codes;
  a.event_count = event_count;
  a.segment = segment;
  a.threads_total = total;
  a.checkpoints = (float*)work;
  a.scratch = (float*)(work + ck);
  a.partial = (float*)(work + ck + sc);
  const size_t lds = sizeof(float) * (size_t)(threads / 64) * 3 * CFLUSH * 64;
  launch(a, dim3((unsigned)blocks), dim3(threads), lds, s);
  st = hip_status(hipGetLastError(), "complex_adjoint_kernel launch");
  if (st == OK) {
    const int64_t rb = (3 * event_count + 255) / 256;
    hipLaunchKernelGGL(reduce_event_partials, dim3((unsigned)rb), dim3(256), 0, s, (const float*)a.partial, waves,
                       event_count, grad_flip, grad_phase, grad_duration, 0);
    st = hip_status(hipGetLastError(), "reduce_event_partials launch");
  }
  (void)hipFreeAsync(work, s);
  return st;
}

extern "C" int tsim_complex_jvp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                                int64_t output_count, int pool_kind, int features, int64_t tissue_stride,
                                float factor_re, float factor_im, const float* const* tangents,
                                int64_t tangent_count, void* signal, void* tangent, void* stream) {
  ComplexJvpArgs q;
  memset(&q, 0, sizeof(q));
  int st = fill_common(q.base, packed, atom_count, event_count, state_count, output_count, pool_kind, features,
                       tissue_stride, factor_re, factor_im);
  if (st != OK) return st;
  if (tangent_count != 0 && (tangents == nullptr || tangent_count != TISSUE_COUNT + 3)) {
    set_error("tangents holds %d pointers (17 tissue, then duration, flip, phase) or tangent_count is 0",
              TISSUE_COUNT + 3);
    return BAD_ARGUMENT;
  }
  if (tangent == nullptr && atom_count > 0 && output_count > 0) {
    set_error("the tangent output is required");
    return BAD_ARGUMENT;
  }
  if (tangent_count > 0) {
    for (int k = 0; k < TISSUE_COUNT; ++k) {
      if (tangents[k] == nullptr) continue;
      const int slot = carried_slot(k);
      if (slot < 0) {
        set_error("tissue property %d carries no tangent in the complex JVP", k);
        return UNSUPPORTED;
      }
      q.tissue_tangent[slot] = tangents[k];
    }
    q.duration_tangent = tangents[TISSUE_COUNT + 0];
    q.flip_tangent = tangents[TISSUE_COUNT + 1];
    q.phase_tangent = tangents[TISSUE_COUNT + 2];
  }
  q.signal = (float2*)signal;
  q.tangent = (float2*)tangent;
  hipStream_t s = (hipStream_t)stream;
  if (atom_count == 0 || output_count == 0) return OK;
  // samples no event records stay zero (a fill kernel, not a memset, so a
  // caller may capture the call in a graph)
  if (q.signal) st = clear((float*)q.signal, 2 * atom_count * output_count, s);
  if (st == OK) st = clear((float*)q.tangent, 2 * atom_count * output_count, s);
  if (st != OK || event_count == 0) return st;
  const int L = adjoint_lanes(state_count);
  ComplexJvpLaunch launch = pick_complex_jvp(L, pool_kind == 2);
  if (!launch) {
    set_error("no complex JVP kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  const int apw = 64 / L;
  const int64_t waves = (atom_count + apw - 1) / apw;
  const int threads = waves < 4LL * 256 ? 64 : 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * apw;
  const int64_t blocks = (atom_count + atoms_per_block - 1) / atoms_per_block;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  q.base.events = compiled.codes;
  launch(q, dim3((unsigned)blocks), dim3(threads), s);
  return hip_status(hipGetLastError(), "complex_jvp_kernel launch");
}

extern "C" int tsim_complex_vjp_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                                         int64_t state_count, int64_t output_count, int pool_kind, int features,
                                         int64_t tissue_stride, float factor_re, float factor_im,
                                         const void* signal_cotangent, const int32_t* columns, int64_t column_count,
                                         const void* jacobian_cotangent, int64_t segment_events, float* grad_flip,
                                         float* grad_phase, float* grad_duration, void* stream) {
  ComplexFovrArgs q;
  memset(&q, 0, sizeof(q));
  int st = fill_common(q.base, packed, atom_count, event_count, state_count, output_count, pool_kind, features,
                       tissue_stride, factor_re, factor_im);
  if (st != OK) return st;
  ComplexAdjointArgs& a = q.base;
  if (column_count < 0 || column_count > 4 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 4 Jacobian columns, got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (segment_events < 0) {
    set_error("segment_events must be positive (or 0 for the library's choice)");
    return BAD_ARGUMENT;
  }
  const bool sized = atom_count > 0 && output_count > 0;
  if (column_count > 0 && jacobian_cotangent == nullptr && sized) {
    set_error("Jacobian columns asked for but the Jacobian cotangent is null");
    return BAD_ARGUMENT;
  }
  int m0_column = -1;
  int tangent[4];
  int tangents = 0;
  int64_t tangent_column[4];
  for (int64_t c = 0; c < column_count; ++c) {
    int kind;
    switch (columns[c]) {
      case T_T1: kind = CG_T1; break;
      case T_T2: kind = CG_T2; break;
      case T_B1: kind = CG_B1; break;
      case T_M0: kind = CG_M0; break;
      default:
        set_error("column %lld: property %d carries no complex-kernel tangent", (long long)c, columns[c]);
        return BAD_ARGUMENT;
    }
    for (int j = 0; j < tangents; ++j)
      if (tangent[j] == kind) {
        set_error("property %d appears twice among the Jacobian columns", columns[c]);
        return BAD_ARGUMENT;
      }
    if (kind == CG_M0) {
      if (m0_column >= 0) {
        set_error("property %d appears twice among the Jacobian columns", columns[c]);
        return BAD_ARGUMENT;
      }
      m0_column = (int)c;
      continue;
    }
    if (kind == CG_B1 && !a.transmit) {
      set_error("a B1 column needs the transmit feature and a b1 buffer");
      return BAD_ARGUMENT;
    }
    tangent_column[tangents] = c;
    tangent[tangents++] = kind;
  }
  hipStream_t s = (hipStream_t)stream;
  for (float* g : {grad_flip, grad_phase, grad_duration})
    if (st == OK) st = clear(g, event_count, s);
  if (st != OK || !sized || event_count == 0) return st;

  // the signal's cotangent (with an M0 column's) through the first-order adjoint
  if (signal_cotangent != nullptr || m0_column >= 0) {
    void* joined = nullptr;
    const void* w = signal_cotangent;
    if (m0_column >= 0) {
      st = scratch_alloc(&joined, sizeof(float2) * (size_t)(atom_count * output_count), s, "M0 cotangent");
      if (st != OK) return st;
      const int64_t n = atom_count * output_count;
      hipLaunchKernelGGL(join_m0_cotangent, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                         (const float2*)signal_cotangent, (const float2*)jacobian_cotangent, column_count,
                         (int64_t)m0_column, a.m0, tissue_stride, a.density, atom_count, output_count,
                         (float2*)joined);
      st = hip_status(hipGetLastError(), "join_m0_cotangent launch");
      w = joined;
    }
    if (st == OK)
      st = tsim_complex_vjp(packed, atom_count, event_count, state_count, output_count, pool_kind, features,
                            tissue_stride, factor_re, factor_im, w, segment_events, grad_flip, grad_phase,
                            grad_duration, nullptr, 0, stream);
    if (joined) (void)hipFreeAsync(joined, s);
    if (st != OK) return st;
  }
  if (tangents == 0) return OK;

  const int L = adjoint_lanes(state_count);
  ComplexFovrLaunch launch = pick_complex_fovr(L, pool_kind == 2);
  if (!launch) {
    set_error("no complex forward-over-reverse kernel for %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  int64_t segment = segment_events;
  if (segment == 0) segment = (int64_t)std::ceil(std::sqrt((double)event_count));
  if (segment > event_count) segment = event_count;
  const int apw = 64 / L;
  const int64_t waves_needed = (atom_count + apw - 1) / apw;
  const int threads = waves_needed < 4LL * 256 ? 64 : 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * apw;
  const int64_t blocks = (atom_count + atoms_per_block - 1) / atoms_per_block;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  const int64_t total = blocks * threads;
  const int64_t waves = total / 64;
  const int64_t floats = 2 * 6LL * (pool_kind == 2 ? 2 : 1);  // the state and its tangent
  const int64_t segments = (event_count + segment - 1) / segment;
  const size_t ck = sizeof(float) * (size_t)(segments * floats * total);
  const size_t sc = sizeof(float) * (size_t)(segment * floats * total);
  const size_t pa = sizeof(float) * (size_t)(waves * event_count * 3);
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  char* work = nullptr;
  st = scratch_alloc((void**)&work, ck + sc + pa, s, "complex forward-over-reverse workspace");
  if (st != OK) return st;
  a.events = compiled.codes;
  a.segment = segment;
  a.threads_total = total;
  a.checkpoints = (float*)work;
  a.scratch = (float*)(work + ck);
  a.partial = (float*)(work + ck + sc);
  q.jacobian_cotangent = (const float2*)jacobian_cotangent;
  q.columns = column_count;
  const size_t lds = sizeof(float) * (size_t)(threads / 64) * 3 * CFLUSH * 64;
  const int64_t rb = (3 * event_count + 255) / 256;
  for (int j = 0; j < tangents && st == OK; ++j) {
    q.tangent = tangent[j];
    q.column = tangent_column[j];
    launch(q, dim3((unsigned)blocks), dim3(threads), lds, s);
    st = hip_status(hipGetLastError(), "complex_fovr_kernel launch");
    if (st == OK) {
      hipLaunchKernelGGL(reduce_event_partials, dim3((unsigned)rb), dim3(256), 0, s, (const float*)a.partial, waves,
                         event_count, grad_flip, grad_phase, grad_duration, 1);
      st = hip_status(hipGetLastError(), "reduce_event_partials launch");
    }
  }
  (void)hipFreeAsync(work, s);
  return st;
}
