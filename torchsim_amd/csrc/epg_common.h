// Shared definitions for the gfx950 EPG kernels: the event bits of the packed
// stream, the compact per-event record the kernels walk, and the status
// plumbing of the C ABI.
//
// Event bits follow the reference's packed action word
// (src/torchsim/_epg_cpu.cpp:36-41, sequence/_accelerators.py _pack_events).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tsim {

constexpr uint32_t PRE_SHIFT = 1;
constexpr uint32_t POST_SHIFT = 2;
constexpr uint32_t INVERSION = 4;
constexpr uint32_t SPOIL_AFTER = 8;
constexpr uint32_t SHIFT_AFTER = 16;
constexpr uint32_t RECORD = 32;

// Event kinds of the packed stream (reference sequence/_description.py:31-36).
constexpr int32_t KIND_WAIT = 0;
constexpr int32_t KIND_RF = 1;
constexpr int32_t KIND_ADC = 2;

// Feature bits, in the order the reference packs them
// (src/torchsim/_epg_cpu.cpp:89-96, sequence/_parameters.py FEATURE_BITS).
constexpr int FEATURE_OFF_AXIS = 1 << 0;
constexpr int FEATURE_MOVING = 1 << 1;
constexpr int FEATURE_DIFFUSING = 1 << 2;
constexpr int FEATURE_TRANSMIT = 1 << 3;
constexpr int FEATURE_DENSITY = 1 << 4;
constexpr int FEATURE_INVERTING = 1 << 5;

// Tissue property order of the packed ABI (reference _epg_cpu.cpp:613-620).
enum TissueIndex : int {
  T_T1 = 0, T_T2, T_M0, T_B1, T_B1_PHASE, T_B0, T_INV_EFF, T_DIFFUSION,
  T_VELOCITY, T_BOUND_FRACTION, T_BOUND_EXCHANGE, T_T1_BOUND,
  T_POOL_B_FRACTION, T_POOL_B_EXCHANGE, T_T1_POOL_B, T_T2_POOL_B,
  T_POOL_B_SHIFT, TISSUE_COUNT
};
// Event buffers after the tissue block (reference _epg_cpu.cpp:622-631).
enum EventIndex : int {
  E_DURATION = 0, E_KIND, E_FLIP, E_PHASE, E_ACTION, E_OUTPUT_INDEX,
  E_SHIM_INDEX, E_SATURATION, E_RF_FREQUENCY, EVENT_COUNT
};
constexpr int PACKED_COUNT = TISSUE_COUNT + EVENT_COUNT;

// Control word of a compiled event: the action bits in the low byte, the
// kind above them.
constexpr uint32_t CTRL_KIND_SHIFT = 8;

// One event as the kernels read it: 32 bytes, a single scalar load. Built on
// the device from the reference-format buffers by a one-thread-per-event pass.
struct alignas(32) EventCode {
  float dt;        // interval before the event, seconds
  float flip;      // radians
  float versine;   // 1 - cos(flip) = 2 sin^2(flip/2), for a launch whose flip carries no B1 map
  float sine;      // sin(flip)
  float phase;     // radians
  uint32_t ctrl;   // action | kind << 8
  int32_t out;     // output index, -1 when not recorded
  int32_t event;   // position in the packed stream (adjoint bookkeeping)
};

// Event codes are never written while a kernel runs. Reading them through the
// constant address space lets the compiler turn a wave-uniform index into
// scalar loads (one s_load per code) instead of vector loads plus readlanes.
__device__ __forceinline__ EventCode load_event(const EventCode* codes, int64_t e) {
  using Word = const __attribute__((address_space(4))) uint32_t;
  Word* w = (Word*)(codes + e);
  EventCode ev;
  ev.dt = __uint_as_float(w[0]);
  ev.flip = __uint_as_float(w[1]);
  ev.versine = __uint_as_float(w[2]);
  ev.sine = __uint_as_float(w[3]);
  ev.phase = __uint_as_float(w[4]);
  ev.ctrl = w[5];
  ev.out = (int32_t)w[6];
  ev.event = (int32_t)w[7];
  return ev;
}

// Status codes every entry point returns. Zero is success; the message of the
// last failure on the calling thread is available from tsim_last_error().
enum Status : int {
  OK = 0,
  BAD_ARGUMENT = 1,
  UNSUPPORTED = 2,
  HIP_FAILURE = 3,
};

struct MacroOp;

// Stream-ordered scratch holding a compiled event stream: allocated and
// freed on the caller's stream, so calls on distinct streams never share it.
struct Compiled {
  EventCode* codes = nullptr;
  MacroOp* ops = nullptr;
  hipStream_t stream = nullptr;
  ~Compiled();
};

// Compile the reference-format event buffers (packed[17..22]) into EventCodes.
int compile_packed(void* const* packed, int64_t event_count, hipStream_t stream, Compiled& out);

// The same into a caller's buffer of `events` EventCodes (stream-ordered).
int compile_stream(const float* duration, const int32_t* kind, const float* flip, const float* phase,
                   const uint8_t* action, const int32_t* output_index, int64_t events, EventCode* out,
                   hipStream_t stream);

// Group compiled events into the real kernel's macro-ops (after compile_packed).
int compile_ops(int64_t event_count, int64_t outputs, Compiled& out);

// The same grouping into caller-provided ops (op_capacity(events) of them);
// with op_count, the number of ops (end marker excluded) is written there.
// zero_mode 1 gives every event with a zero interval a relaxation part at
// dt = 0, 2 each event flagged in zero_mask (DEVICE, one byte per event);
// the adjoint then returns those intervals' duration gradients.
int pack_stream(const EventCode* codes, int64_t events, int64_t outputs, MacroOp* ops, hipStream_t stream,
                int64_t* op_count = nullptr, const uint8_t* zero_mask = nullptr, int zero_mode = 0);

// Copy a host table of `bytes` to `device` on stream `s` through pooled
// pinned staging (no host wait; under graph capture the staging buffer is
// kept for the graph's replays).
int upload_table(const void* host, size_t bytes, void* device, hipStream_t s);

void set_error(const char* fmt, ...);
const char* last_error();
int hip_status(hipError_t err, const char* where);

// Stream-ordered scratch (hipMallocAsync). The first call on a device raises
// its default pool's release threshold, so freed scratch stays mapped for the
// next launch instead of going back to the HIP runtime's allocator at every
// synchronization.
int scratch_alloc(void** ptr, size_t bytes, hipStream_t stream, const char* what);

}  // namespace tsim

extern "C" int tsim_complex_simulate(void* const* pointers, int64_t pointer_count,
                                     int64_t atom_count, int64_t event_count,
                                     int64_t state_count, int64_t output_count, int pool_kind,
                                     int features, int64_t tissue_stride, void* stream);
