// Real-subspace EPG forward fused with forward-mode tangents, for gfx950.
//
// What it computes is the reference's real-subspace state machine
// (src/torchsim/_epg_cpu.cpp:3041-3128 forward, :3136-3285 one-tangent JVP,
// sequence/_epg_triton.py:14221-14620), carrying the value and up to four
// tangent directions (T1, T2, B1, inversion efficiency) through ONE walk of
// the event stream, where the reference runs one forward and one JVP kernel
// per differentiated property. The M0 column needs no recursion: it is the
// recorded a0 itself.
//
// Layout. An atom's S configuration orders are split into L contiguous blocks
// of SL = S/L registers, one block per lane, so an atom occupies L adjacent
// lanes of a wave (64/L atoms per wave). Relaxation and the RF T-matrix are
// per-state register arithmetic; the gradient shift moves registers within a
// lane and crosses block boundaries with one DPP row shift per plane
// (row_shr:1 for F+ moving up, row_shl:1 for F- moving down). Every per-event
// quantity is uniform across the launch and lives in scalar registers; the
// event stream is read one 32-byte record at a time through the scalar cache.
//
// Representation (reference _epg_cpu.cpp:3029-3038): F+ = e^{i phi} i a,
// F- = e^{-i phi} i b, Z = c, all real; the recorded sample is i * m0 * a0.
#include "epg_real_impl.h"

#include <atomic>

#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace tsim {

namespace {

thread_local char g_error[512] = "";

}  // namespace

const char* last_error() { return g_error; }

void set_error(const char* fmt, ...) {
  va_list args;
  va_start(args, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, args);
  va_end(args);
}

int hip_status(hipError_t err, const char* where) {
  if (err == hipSuccess) return OK;
  set_error("%s: %s", where, hipGetErrorString(err));
  return HIP_FAILURE;
}

// ---------------------------------------------------------------------------
// Event compilation: reference-format buffers -> one EventCode per event.
// ---------------------------------------------------------------------------

__global__ void compile_events(const float* __restrict__ duration,
                               const int32_t* __restrict__ kind,
                               const float* __restrict__ flip,
                               const float* __restrict__ phase,
                               const uint8_t* __restrict__ action,
                               const int32_t* __restrict__ output_index,
                               int64_t events, EventCode* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= events) return;
  EventCode code;
  code.dt = duration[e];
  code.flip = flip[e];
  // 1 - cos(flip) from the half angle: formed as 1 - cos it would carry
  // cos's rounding (about 6e-8) into a quantity that is small for small
  // flips, and a train repeats that error on every pulse.
  const float h = sinf(0.5f * code.flip);
  code.versine = 2.0f * (h * h);
  code.sine = sinf(code.flip);
  code.phase = phase ? phase[e] : 0.0f;
  code.ctrl = (uint32_t)action[e] | ((uint32_t)kind[e] << CTRL_KIND_SHIFT);
  code.out = output_index[e];
  code.event = (int32_t)e;
  out[e] = code;
}

int compile_stream(const float* duration, const int32_t* kind, const float* flip,
                   const float* phase, const uint8_t* action,
                   const int32_t* output_index, int64_t events, EventCode* out,
                   hipStream_t stream) {
  if (events <= 0) return OK;
  const int threads = 256;
  const int64_t blocks = (events + threads - 1) / threads;
  hipLaunchKernelGGL(compile_events, dim3((unsigned)blocks), dim3(threads), 0, stream,
                     duration, kind, flip, phase, action, output_index, events, out);
  return hip_status(hipGetLastError(), "compile_events launch");
}

// ---------------------------------------------------------------------------
// Macro-ops: the event stream regrouped for the real kernel.
//
// Each event contributes up to five parts in its own order: a relaxation
// (dt != 0), a pre-shift, its action (rotation or inversion for an RF event,
// a record for a recording ADC event), a post-shift, and a spoil or shift
// after it. A part has a position: relax 0, rotation/inversion/spoil 1
// (the kernel applies all three as one plane transform), record 2, shift 3. An op is a maximal run of parts with increasing positions
// that ends at the first position-3 part, so a new op starts wherever a part's
// position does not exceed its predecessor's, or follows a shift/spoil. This
// cut rule is local, which makes the grouping a scan: one block, each thread
// summarising a contiguous chunk of events, a block scan of the summaries,
// then each thread placing its parts. The parts of one op may come from
// several threads; each writes its own fields, and the control bits are
// or-ed in atomically.
// ---------------------------------------------------------------------------

namespace {

constexpr int kPackThreads = 1024;

struct PartSummary {
  int has;      // any part in the range
  int first;    // position of its first part
  int last;     // position of its last part
  int64_t cuts; // op boundaries strictly inside the range
};

__device__ __forceinline__ bool cut_between(int prev, int next) { return next <= prev || prev == 3; }

__device__ __forceinline__ PartSummary combine(const PartSummary& x, const PartSummary& y) {
  if (!x.has) return y;
  if (!y.has) return x;
  PartSummary r;
  r.has = 1;
  r.first = x.first;
  r.last = y.last;
  r.cuts = x.cuts + y.cuts + (cut_between(x.last, y.first) ? 1 : 0);
  return r;
}

// Call f(part, position) for each part of event `ev` in order; part is one
// of OP_RELAX, OP_SHIFT, OP_ROTATE, OP_INVERT, OP_RECORD, OP_SPOIL. With
// zero_interval an event whose interval is zero still gets its relaxation
// part (an identity at dt = 0), so the adjoint returns that interval's
// duration gradient.
template <typename F>
__device__ __forceinline__ void for_each_part(const EventCode& ev, bool zero_interval, F&& f) {
  const uint32_t act = ev.ctrl & 0xFFu;
  const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
  if (ev.dt != 0.0f || zero_interval) f(OP_RELAX, 0);
  if (act & PRE_SHIFT) f(OP_SHIFT, 3);
  if (kind == KIND_RF) {
    f((act & INVERSION) ? OP_INVERT : OP_ROTATE, 1);
  } else if (kind == KIND_ADC && (act & RECORD)) {
    f(OP_RECORD, 2);
  }
  if (act & POST_SHIFT) f(OP_SHIFT, 3);
  if (act & SPOIL_AFTER) {
    f(OP_SPOIL, 1);
  } else if (act & SHIFT_AFTER) {
    f(OP_SHIFT, 3);
  }
}

// zero_mode: 0 no zero-interval relaxation parts (forward kernels), 1 one
// for every event, 2 one for each event flagged in zero_mask.
__global__ void __launch_bounds__(kPackThreads) pack_ops(const EventCode* __restrict__ codes, int64_t events,
                                                         int64_t outputs, MacroOp* __restrict__ ops,
                                                         int64_t* __restrict__ op_count,
                                                         const uint8_t* __restrict__ zero_mask, int zero_mode) {
  __shared__ int s_has[kPackThreads];
  __shared__ int s_first[kPackThreads];
  __shared__ int s_last[kPackThreads];
  __shared__ int64_t s_cuts[kPackThreads];
  __shared__ int64_t s_recs[kPackThreads];
  __shared__ int s_sequential;
  const int t = (int)threadIdx.x;
  const int64_t chunk = (events + kPackThreads - 1) / kPackThreads;
  const int64_t lo = chunk * t < events ? chunk * t : events;
  const int64_t hi = lo + chunk < events ? lo + chunk : events;
  if (t == 0) s_sequential = 1;

  PartSummary mine{0, 0, 0, 0};
  int64_t my_recs = 0;
  auto zero_interval = [&](int64_t e) -> bool {
    return zero_mode == 1 || (zero_mode == 2 && zero_mask[e] != 0);
  };
  for (int64_t e = lo; e < hi; ++e) {
    for_each_part(codes[e], zero_interval(e), [&](uint32_t part, int pos) {
      if (!mine.has) {
        mine.has = 1;
        mine.first = pos;
      } else if (cut_between(mine.last, pos)) {
        ++mine.cuts;
      }
      mine.last = pos;
      if (part == OP_RECORD) ++my_recs;
    });
  }
  // exclusive scan of the record counts (the ordinal of each thread's first record)
  s_recs[t] = my_recs;
  __syncthreads();
  for (int step = 1; step < kPackThreads; step <<= 1) {
    const int64_t left = t >= step ? s_recs[t - step] : 0;
    __syncthreads();
    s_recs[t] += left;
    __syncthreads();
  }
  int64_t record = s_recs[t] - my_recs;
  // inclusive Hillis-Steele scan of the summaries
  s_has[t] = mine.has;
  s_first[t] = mine.first;
  s_last[t] = mine.last;
  s_cuts[t] = mine.cuts;
  __syncthreads();
  for (int step = 1; step < kPackThreads; step <<= 1) {
    PartSummary left{0, 0, 0, 0};
    if (t >= step) left = PartSummary{s_has[t - step], s_first[t - step], s_last[t - step], s_cuts[t - step]};
    const PartSummary self{s_has[t], s_first[t], s_last[t], s_cuts[t]};
    __syncthreads();
    const PartSummary merged = combine(left, self);
    s_has[t] = merged.has;
    s_first[t] = merged.first;
    s_last[t] = merged.last;
    s_cuts[t] = merged.cuts;
    __syncthreads();
  }
  PartSummary before{0, 0, 0, 0};
  if (t > 0) before = PartSummary{s_has[t - 1], s_first[t - 1], s_last[t - 1], s_cuts[t - 1]};
  const int64_t count = s_has[kPackThreads - 1] ? s_cuts[kPackThreads - 1] + 1 : 0;
  __syncthreads();

  // clear every op and the two end-marker slots (the kernel reads one op
  // past the first), then mark the end
  for (int64_t i = t; i <= count + 1; i += kPackThreads) {
    MacroOp z;
    z.dt = 0.0f;
    z.flip = 0.0f;
    z.versine = 0.0f;
    z.sine = 0.0f;
    z.ctrl = i >= count ? OP_END : 0u;
    z.out = -1;
    z.event = -1;
    z.relax_event = -1;
    ops[i] = z;
  }
  __syncthreads();

  int has = before.has;
  int last = before.last;
  int64_t index = before.has ? before.cuts : 0;
  for (int64_t e = lo; e < hi; ++e) {
    const EventCode ev = codes[e];
    for_each_part(ev, zero_interval(e), [&](uint32_t part, int pos) {
      if (has && cut_between(last, pos)) ++index;
      has = 1;
      last = pos;
      MacroOp* op = ops + index;
      if (part == OP_RELAX) {
        op->dt = ev.dt;
        op->relax_event = ev.event;
      } else if (part == OP_ROTATE) {
        op->flip = ev.flip;
        op->versine = ev.versine;
        op->sine = ev.sine;
        op->event = ev.event;
      } else if (part == OP_RECORD) {
        op->out = ev.out;
        // staged writes need the k-th record to fill output k
        if ((int64_t)ev.out != record || ev.out < 0 || (int64_t)ev.out >= outputs) s_sequential = 0;
        ++record;
      }
      atomicOr(&op->ctrl, part);
    });
  }
  __syncthreads();
  if (t == 0 && count > 0 && s_sequential) atomicOr(&ops[0].ctrl, OPS_SEQUENTIAL);
  if (t == 0 && op_count != nullptr) *op_count = count;
}

}  // namespace

int compile_ops(int64_t event_count, int64_t outputs, Compiled& out) {
  const int64_t capacity = op_capacity(event_count > 0 ? event_count : 0);
  int st = scratch_alloc((void**)&out.ops, sizeof(MacroOp) * (size_t)capacity, out.stream, "macro-ops");
  if (st != OK) return st;
  hipLaunchKernelGGL(pack_ops, dim3(1), dim3(kPackThreads), 0, out.stream, out.codes,
                     event_count > 0 ? event_count : 0, outputs, out.ops, (int64_t*)nullptr,
                     (const uint8_t*)nullptr, 0);
  return hip_status(hipGetLastError(), "pack_ops launch");
}

int pack_stream(const EventCode* codes, int64_t events, int64_t outputs, MacroOp* ops, hipStream_t stream,
                int64_t* op_count, const uint8_t* zero_mask, int zero_mode) {
  hipLaunchKernelGGL(pack_ops, dim3(1), dim3(kPackThreads), 0, stream, codes, events > 0 ? events : 0, outputs,
                     ops, op_count, zero_mask, zero_mode);
  return hip_status(hipGetLastError(), "pack_ops launch");
}

// Output staging for a launch writing complex signal / Jacobian rows: the
// log2 of the samples per flush (K) and of the padded stream count, sized
// so one wave's stage fits STAGE_BYTES_PER_WAVE; K < 2 writes directly.
void set_staging(RealArgs& args, int lanes_per_atom) {
  args.stage_k_log2 = -1;
  args.stage_s_log2 = 0;
  const bool reads = args.normal != nullptr;  // normal equations stage the measured samples
  if (!reads && (args.out_real != nullptr || (args.signal == nullptr && args.jacobian == nullptr))) return;
  const int streams = reads ? 1 : 1 + args.columns;
  int s_log2 = 0;
  while ((1 << s_log2) < streams) ++s_log2;
  const int64_t slot_bytes = (int64_t)(64 / lanes_per_atom) * (1 << s_log2) * (int64_t)sizeof(float2);
  int k_log2 = 0;
  while (k_log2 < 4 && slot_bytes * (2 << k_log2) <= STAGE_BYTES_PER_WAVE) ++k_log2;
  if (k_log2 < 1) return;
  args.stage_k_log2 = k_log2;
  args.stage_s_log2 = s_log2;
}

// Dynamic LDS bytes of a block of `threads` threads for the staging above.
size_t stage_bytes(const RealArgs& args, int lanes_per_atom, int threads = 256) {
  if (args.stage_k_log2 < 1) return 0;
  return (size_t)(threads / 64) * (64 / lanes_per_atom) * ((size_t)1 << (args.stage_s_log2 + args.stage_k_log2)) *
         sizeof(float2);
}

// Threads per block of a real-kernel launch over `lanes` lanes: one-wave
// blocks when four-wave blocks would occupy fewer than all 256 compute units,
// so the waves spread over every unit (each with its own caches and memory
// path); four-wave blocks otherwise. The kernel indexes by global thread id
// and by wave within the block only.
int real_block_threads(int64_t lanes) { return lanes < 256LL * 256 ? 64 : 256; }

// ---------------------------------------------------------------------------
// The real-subspace kernel.
// ---------------------------------------------------------------------------

// ---------------------------------------------------------------------------
// Dispatch over the instantiated tiles.
// ---------------------------------------------------------------------------

// A state tile: SL orders per lane, L lanes per atom.
struct Tile {
  int sl;
  int l;
};

#define TSIM_TILE_ENTRY(ID, SL, L) {SL, L},
constexpr Tile kTiles[] = {TSIM_TILES(TSIM_TILE_ENTRY)};
#undef TSIM_TILE_ENTRY
constexpr int kTileCount = sizeof(kTiles) / sizeof(kTiles[0]);

RealLaunch pick_real(int tile, int tm, int variant) {
  switch (tile) {
#define TSIM_TILE_CASE(ID, SL, L) \
  case ID:                        \
    return pick_real_tile_##ID(tm, variant);
    TSIM_TILES(TSIM_TILE_CASE)
#undef TSIM_TILE_CASE
    default:
      return nullptr;
  }
}

// The variant bits of a launch: what the features ask for, and whether the
// tile carries orders past the state count (which must then be kept zero).
int variant_of(int features, bool accum, int tile, int64_t states) {
  int v = 0;
  if (features & FEATURE_DIFFUSING) v |= V_DIFF;
  if (features & FEATURE_TRANSMIT) v |= V_B1MAP;
  if (accum) v |= V_ACCUM;
  if ((int64_t)kTiles[tile].sl * kTiles[tile].l != states) v |= V_PADDED;
  return v;
}

int default_tile(int states, int tangents, int min_lanes = 1, bool accumulate = false) {
  // Among the tiles covering S orders, weigh the padded orders (work spent on
  // zeros), the in-lane register moves and DPP crossings of a wider atom, and
  // the occupancy the register file leaves: 3 planes * SL orders *
  // (1 + tangents) floats of state per lane, and for the normal equations
  // their FP64 sums and the staged data on every lane (~40 more). A
  // projection needs enough lanes per atom for its basis directions
  // (min_lanes).
  int best = -1;
  float best_cost = 1e30f;
  for (int t = 0; t < kTileCount; ++t) {
    const Tile tile = kTiles[t];
    const int cover = tile.sl * tile.l;
    if (cover < states || tile.l < min_lanes) continue;
    const int regs = 3 * tile.sl * (1 + tangents) + (accumulate ? 40 : 0);
    const float occupancy = regs <= 96 ? 1.0f : (regs <= 150 ? 1.12f : (regs <= 200 ? 1.3f : 1.8f));
    const float cost = (float)cover / (float)states * (1.0f + 0.06f * (float)(tile.l - 1)) * occupancy;
    if (cost < best_cost) {
      best_cost = cost;
      best = t;
    }
  }
  return best;
}

int run_real(RealArgs& args, int tm, int features, int tile, hipStream_t stream) {
  if (tile < 0) tile = default_tile(args.states, popcount4(tangent_class(tm)));
  if (tile < 0 || tile >= kTileCount) {
    set_error("no state tile covers %d orders (at most %d)", args.states,
              kTiles[kTileCount - 1].sl * kTiles[kTileCount - 1].l);
    return UNSUPPORTED;
  }
  if (kTiles[tile].sl * kTiles[tile].l < args.states) {
    set_error("tile %d covers %d orders, fewer than the %d asked for", tile,
              kTiles[tile].sl * kTiles[tile].l, args.states);
    return BAD_ARGUMENT;
  }
  if (args.atoms == 0 || args.outputs == 0) return OK;
  const int L = kTiles[tile].l;
  const int64_t lanes = args.atoms * L;
  const int threads = real_block_threads(lanes);
  const int64_t blocks = (lanes + threads - 1) / threads;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)args.atoms);
    return BAD_ARGUMENT;
  }
  RealLaunch launch = pick_real(tile, tm, variant_of(features, false, tile, args.states));
  if (!launch) {
    set_error("no real kernel for this variant");
    return UNSUPPORTED;
  }
  set_staging(args, L);
  launch(args, dim3((unsigned)blocks), dim3(threads), stage_bytes(args, L, threads), stream);
  return hip_status(hipGetLastError(), "real_kernel launch");
}

BatchLaunch pick_batch(int tile, int tm, int variant) {
  switch (tile) {
#define TSIM_BATCH_CASE(ID, SL, L) \
  case ID:                         \
    return pick_batch_tile_##ID(tm, variant);
    TSIM_TILES(TSIM_BATCH_CASE)
#undef TSIM_BATCH_CASE
    default:
      return nullptr;
  }
}

// Normal equations from a stored Jacobian: one wavefront per voxel, each lane
// a stride-64 slice of the samples (so a wave reads 64 consecutive complex
// samples of a row at a time), FP64 products and sums, then a butterfly over
// the wave.
__global__ void __launch_bounds__(256) normal_from_jacobian(const float2* __restrict__ signal,
                                                            const float2* __restrict__ jacobian,
                                                            const float2* __restrict__ data, int64_t voxels,
                                                            int64_t outputs, int cols,
                                                            double* __restrict__ normal) {
  const int64_t v = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int lane = (int)(threadIdx.x & 63);
  if (v >= voxels) return;  // whole waves leave together
  double acc[NORMAL_STRIDE];
#pragma unroll
  for (int i = 0; i < NORMAL_STRIDE; ++i) acc[i] = 0.0;
  const float2* srow = signal + v * outputs;
  const float2* drow = data + v * outputs;
  const float2* jrow = jacobian + v * (int64_t)cols * outputs;
  for (int64_t n = lane; n < outputs; n += 64) {
    const float2 s = srow[n];
    const float2 y = drow[n];
    const double rx = (double)s.x - (double)y.x;
    const double ry = (double)s.y - (double)y.y;
    acc[0] += rx * rx + ry * ry;
    double jx[NORMAL_COLUMNS], jy[NORMAL_COLUMNS];
#pragma unroll
    for (int c = 0; c < NORMAL_COLUMNS; ++c) {
      const float2 j = c < cols ? jrow[(int64_t)c * outputs + n] : make_float2(0.0f, 0.0f);
      jx[c] = j.x;
      jy[c] = j.y;
      acc[1 + c] += jx[c] * rx + jy[c] * ry;  // Re(conj(J_c) r)
    }
    int h = 0;
#pragma unroll
    for (int c = 0; c < NORMAL_COLUMNS; ++c) {
#pragma unroll
      for (int q = c; q < NORMAL_COLUMNS; ++q) acc[1 + NORMAL_COLUMNS + h++] += jx[c] * jx[q] + jy[c] * jy[q];
    }
  }
#pragma unroll
  for (int i = 0; i < NORMAL_STRIDE; ++i) {
    double x = acc[i];
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) x += __shfl_xor(x, offset, 64);
    acc[i] = x;
  }
  if (lane == 0) {
#pragma unroll
    for (int i = 0; i < NORMAL_STRIDE; ++i) normal[v * NORMAL_STRIDE + i] = acc[i];
  }
}

// One thread per voxel: H_free = D H D, g_free = D g from the accumulated
// normal equations and the chain-rule scale D of the bound transform, then
// the damped step (H_free + lambda I) d = -g_free by Cholesky in FP64, and
// the reduction that step predicts, d . (lambda d - g_free).
__global__ void solve_damped(const double* __restrict__ normal, const double* __restrict__ scale,
                             const double* __restrict__ damping, const double* __restrict__ offset,
                             int64_t voxels, int cols,
                             double* __restrict__ step, double* __restrict__ predicted,
                             double* __restrict__ gradient, double* __restrict__ curvature) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= voxels) return;
  const double* n = normal + v * NORMAL_STRIDE;
  double D[NORMAL_COLUMNS], g[NORMAL_COLUMNS], H[NORMAL_COLUMNS][NORMAL_COLUMNS];
  for (int c = 0; c < NORMAL_COLUMNS; ++c) D[c] = (c < cols) ? (scale ? scale[v * cols + c] : 1.0) : 0.0;
  int h = 0;
  for (int c = 0; c < NORMAL_COLUMNS; ++c) {
    g[c] = D[c] * n[1 + c];
    for (int k = c; k < NORMAL_COLUMNS; ++k) {
      const double x = D[c] * n[1 + NORMAL_COLUMNS + h++] * D[k];
      H[c][k] = x;
      H[k][c] = x;
    }
  }
  double peak = 0.0;
  for (int c = 0; c < cols; ++c) peak = fmax(peak, H[c][c]);
  if (curvature) curvature[v] = peak;
  if (gradient)
    for (int c = 0; c < cols; ++c) gradient[v * cols + c] = g[c];
  if (!step) return;
  const double lam = damping ? damping[v] : 0.0;
  // Cholesky of H + lam I (cols <= 4)
  double Lm[NORMAL_COLUMNS][NORMAL_COLUMNS] = {};
  bool ok = true;
  for (int i = 0; i < cols; ++i) {
    for (int j = 0; j <= i; ++j) {
      double sum = H[i][j] + (i == j ? lam : 0.0);
      for (int k = 0; k < j; ++k) sum -= Lm[i][k] * Lm[j][k];
      if (i == j) {
        if (!(sum > 0.0)) ok = false;
        Lm[i][i] = sqrt(fmax(sum, 0.0));
      } else {
        Lm[i][j] = sum / Lm[j][j];
      }
    }
  }
  double y[NORMAL_COLUMNS], d[NORMAL_COLUMNS];
  for (int i = 0; i < cols; ++i) {
    double sum = -g[i] + (offset ? offset[v * cols + i] : 0.0);
    for (int k = 0; k < i; ++k) sum -= Lm[i][k] * y[k];
    y[i] = sum / Lm[i][i];
  }
  for (int i = cols - 1; i >= 0; --i) {
    double sum = y[i];
    for (int k = i + 1; k < cols; ++k) sum -= Lm[k][i] * d[k];
    d[i] = sum / Lm[i][i];
  }
  double pred = 0.0;
  for (int c = 0; c < cols; ++c) {
    const double dc = ok ? d[c] : __builtin_nan("");
    step[v * cols + c] = dc;
    pred += dc * (lam * dc - g[c]);
  }
  if (predicted) predicted[v] = pred;
}

// Change of variables of a bounded unknown (two-sided: scaled sigmoid;
// one-sided: softplus), as the host-side transforms define it.
struct LMBounds {
  int kind[NORMAL_COLUMNS];  // 0 free, 1 two-sided, 2 lower only, 3 upper only
  double low[NORMAL_COLUMNS];
  double high[NORMAL_COLUMNS];
};

__device__ inline double lm_sigmoid(double x) { return 1.0 / (1.0 + exp(-x)); }
__device__ inline double lm_softplus(double x) { return x > 20.0 ? x : log1p(exp(x)); }

__device__ inline double lm_natural(const LMBounds& b, int c, double x) {
  switch (b.kind[c]) {
    case 1: return b.low[c] + (b.high[c] - b.low[c]) * lm_sigmoid(x);
    case 2: return b.low[c] + lm_softplus(x);
    case 3: return b.high[c] - lm_softplus(-x);
    default: return x;
  }
}

__device__ inline double lm_widen(const LMBounds& b, int c, double x) {
  switch (b.kind[c]) {
    case 1: {
      const double s = lm_sigmoid(x);
      return (b.high[c] - b.low[c]) * s * (1.0 - s);
    }
    case 2: return lm_sigmoid(x);
    case 3: return lm_sigmoid(-x);
    default: return 1.0;
  }
}

// One Levenberg-Marquardt proposal per voxel: the chain-rule factors of the
// bounded unknowns at x, the damped step of solve_damped (same arithmetic),
// its predicted reduction, whether it failed, the candidate x + step (x where
// it failed) and the candidate's natural values as float32 planes
// [column][voxel], which are the tissue fields of the next normal equations.
__global__ void lm_propose(const double* __restrict__ normal, const double* __restrict__ x,
                           const double* __restrict__ damping, LMBounds bounds, int64_t voxels, int cols,
                           double* __restrict__ candidate, float* __restrict__ natural,
                           double* __restrict__ predicted, uint8_t* __restrict__ singular) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= voxels) return;
  const double* n = normal + v * NORMAL_STRIDE;
  double xs[NORMAL_COLUMNS], D[NORMAL_COLUMNS], g[NORMAL_COLUMNS], H[NORMAL_COLUMNS][NORMAL_COLUMNS];
  for (int c = 0; c < NORMAL_COLUMNS; ++c) {
    xs[c] = c < cols ? x[v * cols + c] : 0.0;
    D[c] = c < cols ? lm_widen(bounds, c, xs[c]) : 0.0;
  }
  int h = 0;
  for (int c = 0; c < NORMAL_COLUMNS; ++c) {
    g[c] = D[c] * n[1 + c];
    for (int k = c; k < NORMAL_COLUMNS; ++k) {
      const double e = D[c] * n[1 + NORMAL_COLUMNS + h++] * D[k];
      H[c][k] = e;
      H[k][c] = e;
    }
  }
  const double lam = damping[v];
  double Lm[NORMAL_COLUMNS][NORMAL_COLUMNS] = {};
  bool ok = true;
  for (int i = 0; i < cols; ++i) {
    for (int j = 0; j <= i; ++j) {
      double sum = H[i][j] + (i == j ? lam : 0.0);
      for (int k = 0; k < j; ++k) sum -= Lm[i][k] * Lm[j][k];
      if (i == j) {
        if (!(sum > 0.0)) ok = false;
        Lm[i][i] = sqrt(fmax(sum, 0.0));
      } else {
        Lm[i][j] = sum / Lm[j][j];
      }
    }
  }
  double y[NORMAL_COLUMNS], d[NORMAL_COLUMNS];
  for (int i = 0; i < cols; ++i) {
    double sum = -g[i];
    for (int k = 0; k < i; ++k) sum -= Lm[i][k] * y[k];
    y[i] = sum / Lm[i][i];
  }
  for (int i = cols - 1; i >= 0; --i) {
    double sum = y[i];
    for (int k = i + 1; k < cols; ++k) sum -= Lm[k][i] * d[k];
    d[i] = sum / Lm[i][i];
  }
  double pred = 0.0;
  bool failed = false;
  for (int c = 0; c < cols; ++c) {
    const double dc = ok ? d[c] : __builtin_nan("");
    d[c] = dc;
    pred += dc * (lam * dc - g[c]);
    failed = failed || !isfinite(dc);
  }
  predicted[v] = pred;
  singular[v] = failed ? 1 : 0;
  for (int c = 0; c < cols; ++c) {
    const double next = failed ? xs[c] : xs[c] + d[c];
    candidate[v * cols + c] = next;
    natural[(int64_t)c * voxels + v] = (float)lm_natural(bounds, c, next);
  }
}

// Accept or reject each voxel's candidate from the normal equations at it
// (Nielsen's damping update on the gain ratio), in place.
__global__ void lm_accept(const double* __restrict__ trial, const double* __restrict__ predicted,
                          const uint8_t* __restrict__ singular, const double* __restrict__ candidate,
                          int64_t voxels, int cols, double* __restrict__ x, double* __restrict__ normal,
                          double* __restrict__ cost, double* __restrict__ damping, double* __restrict__ rising) {
  const int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= voxels) return;
  const double tiny = 1.1754943508222875e-38;  // float32's smallest normal, the host-side floor
  const double trial_cost = trial[v * NORMAL_STRIDE];
  const double pred = predicted[v];
  const double gain = pred > 0.0 ? (cost[v] - trial_cost) / fmax(pred, tiny) : 0.0;
  const bool better = gain > 0.0 && singular[v] == 0;
  const double lam = damping[v];
  const double grow = rising[v];
  const double t = 2.0 * gain - 1.0;
  const double lowered = lam * fmax(1.0 - t * t * t, 1.0 / 3.0);
  damping[v] = better ? lowered : lam * grow;
  rising[v] = better ? 2.0 : grow * 2.0;
  if (better) {
    for (int c = 0; c < cols; ++c) x[v * cols + c] = candidate[v * cols + c];
    for (int k = 0; k < NORMAL_STRIDE; ++k) normal[v * NORMAL_STRIDE + k] = trial[v * NORMAL_STRIDE + k];
    cost[v] = trial_cost;
  }
}

}  // namespace tsim

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

using namespace tsim;

namespace {

bool tangent_mask(const int32_t* columns, int64_t column_count, int& tm, int cols[5]) {
  tm = 0;
  for (int i = 0; i < 5; ++i) cols[i] = -1;
  for (int64_t c = 0; c < column_count; ++c) {
    const int prop = columns[c];
    int slot;
    switch (prop) {
      case T_T1: tm |= TAN_T1; slot = 0; break;
      case T_T2: tm |= TAN_T2; slot = 1; break;
      case T_B1: tm |= TAN_B1; slot = 2; break;
      case T_INV_EFF: tm |= TAN_IE; slot = 3; break;
      case T_M0: slot = 4; break;
      default:
        set_error("column %lld: property %d carries no real-subspace tangent", (long long)c, prop);
        return false;
    }
    if (cols[slot] >= 0) {
      set_error("property %d appears twice among the Jacobian columns", prop);
      return false;
    }
    cols[slot] = (int)c;
  }
  // Masks carrying the inversion efficiency run the full kernel; the unused
  // directions are computed and not written.
  if (tm & TAN_IE) tm = 15;
  return true;
}

}  // namespace

namespace tsim {

Compiled::~Compiled() {
  if (ops) (void)hipFreeAsync(ops, stream);
  if (codes) (void)hipFreeAsync(codes, stream);
}

int scratch_alloc(void** ptr, size_t bytes, hipStream_t stream, const char* what) {
  static std::atomic<uint64_t> pools_ready{0};
  int device = -1;
  if (hipGetDevice(&device) == hipSuccess && device >= 0 && device < 64 &&
      !(pools_ready.load(std::memory_order_acquire) & (1ull << device))) {
    hipMemPool_t pool;
    if (hipDeviceGetDefaultMemPool(&pool, device) == hipSuccess) {
      uint64_t keep = 2ull << 30;  // bytes the pool keeps mapped when idle
      (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    pools_ready.fetch_or(1ull << device, std::memory_order_acq_rel);
  }
  return hip_status(hipMallocAsync(ptr, bytes, stream), what);
}

int compile_packed(void* const* packed, int64_t event_count, hipStream_t stream, Compiled& out) {
  out.stream = stream;
  if (event_count <= 0) return OK;
  for (int i : {E_DURATION, E_KIND, E_FLIP, E_PHASE, E_ACTION, E_OUTPUT_INDEX}) {
    if (packed[TISSUE_COUNT + i] == nullptr) {
      set_error("event buffer %d is null", i);
      return BAD_ARGUMENT;
    }
  }
  int st = scratch_alloc((void**)&out.codes, sizeof(EventCode) * (size_t)event_count, stream, "event codes");
  if (st != OK) return st;
  return compile_stream((const float*)packed[TISSUE_COUNT + E_DURATION],
                        (const int32_t*)packed[TISSUE_COUNT + E_KIND],
                        (const float*)packed[TISSUE_COUNT + E_FLIP],
                        (const float*)packed[TISSUE_COUNT + E_PHASE],
                        (const uint8_t*)packed[TISSUE_COUNT + E_ACTION],
                        (const int32_t*)packed[TISSUE_COUNT + E_OUTPUT_INDEX], event_count, out.codes,
                        stream);
}

}  // namespace tsim

namespace {

int fill_real_args(RealArgs& args, void* const* packed, int64_t atom_count, int64_t event_count,
                   int64_t state_count, int64_t output_count, int features, int64_t tissue_stride) {
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
  if (packed[T_T1] == nullptr || packed[T_T2] == nullptr) {
    set_error("t1 and t2 buffers are required");
    return BAD_ARGUMENT;
  }
  memset(&args, 0, sizeof(args));
  args.stage_k_log2 = -1;
  args.event_count = event_count;
  args.t1 = (const float*)packed[T_T1];
  args.t2 = (const float*)packed[T_T2];
  args.m0 = (const float*)packed[T_M0];
  args.b1 = (const float*)packed[T_B1];
  args.inv_eff = (const float*)packed[T_INV_EFF];
  args.diffusion = (const float*)packed[T_DIFFUSION];
  args.stride_t1 = 1;
  args.stride_t2 = 1;
  args.stride_opt = tissue_stride;
  args.atoms = atom_count;
  args.outputs = output_count;
  args.states = (int)state_count;
  args.density = (features & FEATURE_DENSITY) && args.m0 ? 1 : 0;
  args.inverting = (features & FEATURE_INVERTING) && args.inv_eff ? 1 : 0;
  args.factor_re = 1.0f;
  args.factor_im = 0.0f;
  args.col_t1 = args.col_t2 = args.col_b1 = args.col_ie = args.col_m0 = -1;
  if ((features & FEATURE_TRANSMIT) && args.b1 == nullptr) {
    set_error("transmit feature set but the b1 buffer is null");
    return BAD_ARGUMENT;
  }
  if ((features & FEATURE_DIFFUSING) && args.diffusion == nullptr) {
    set_error("diffusing feature set but the diffusion buffer is null");
    return BAD_ARGUMENT;
  }
  return OK;
}

}  // namespace

extern "C" {

const char* tsim_last_error(void) { return tsim::last_error(); }

int tsim_version(void) { return 3; }

int tsim_simulate(void* const* pointers, int64_t pointer_count, int64_t atom_count,
                  int64_t train_count, int64_t event_count, int64_t state_count,
                  int64_t output_count, int real_axis, int64_t shim_count, int pool_kind,
                  int features, int64_t tissue_stride, void* stream) {
  if (pointers == nullptr || pointer_count != PACKED_COUNT + 2) {
    set_error("wrong number of buffer pointers: %lld, expected %d", (long long)pointer_count,
              PACKED_COUNT + 2);
    return BAD_ARGUMENT;
  }
  if (train_count != 1) {
    set_error("one train per launch (batched trains run one launch each)");
    return UNSUPPORTED;
  }
  if (real_axis != 1 || pool_kind != 0 || shim_count != 1) {
    return tsim_complex_simulate(pointers, pointer_count, atom_count, event_count, state_count,
                                 output_count, pool_kind, features, tissue_stride, stream);
  }
  RealArgs args;
  int st = fill_real_args(args, pointers, atom_count, event_count, state_count, output_count,
                          features, tissue_stride);
  if (st != OK) return st;
  args.out_real = (float*)pointers[PACKED_COUNT];
  args.out_imag = (float*)pointers[PACKED_COUNT + 1];
  if (args.out_real == nullptr || args.out_imag == nullptr) {
    set_error("output planes are required");
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(pointers, event_count, s, compiled);
  if (st != OK) return st;
  st = compile_ops(event_count, output_count, compiled);
  if (st != OK) return st;
  args.ops = compiled.ops;
  return run_real(args, 0, features, -1, s);
}

}  // extern "C"

namespace {

// Arguments of a fused signal + Jacobian launch, checked; tm receives the
// tangent mask.
int jacobian_args(RealArgs& args, int& tm, void* const* packed, int64_t atom_count, int64_t event_count,
                  int64_t state_count, int64_t output_count, int features, int64_t tissue_stride,
                  const int32_t* columns, int64_t column_count, float factor_re, float factor_im,
                  const float* sample_sign, void* signal, void* jacobian) {
  if (packed == nullptr) {
    set_error("packed buffers are null");
    return BAD_ARGUMENT;
  }
  if (column_count < 0 || column_count > 5 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 5 Jacobian columns, got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (column_count > 0 && jacobian == nullptr && atom_count > 0 && output_count > 0) {
    set_error("Jacobian columns asked for but the Jacobian buffer is null");
    return BAD_ARGUMENT;
  }
  int st = fill_real_args(args, packed, atom_count, event_count, state_count, output_count,
                          features, tissue_stride);
  if (st != OK) return st;
  tm = 0;
  int cols[5];
  if (!tangent_mask(columns, column_count, tm, cols)) return BAD_ARGUMENT;
  if (cols[2] >= 0 && !(features & FEATURE_TRANSMIT)) {
    set_error("a B1 column needs the transmit feature and a b1 buffer");
    return BAD_ARGUMENT;
  }
  if (cols[3] >= 0 && !args.inverting) {
    set_error("an inversion-efficiency column needs the inverting feature");
    return BAD_ARGUMENT;
  }
  args.col_t1 = cols[0];
  args.col_t2 = cols[1];
  args.col_b1 = cols[2];
  args.col_ie = cols[3];
  args.col_m0 = cols[4];
  args.columns = (int)column_count;
  args.signal = (float2*)signal;
  args.jacobian = column_count > 0 ? (float2*)jacobian : nullptr;
  args.sample_sign = sample_sign;
  args.factor_re = factor_re;
  args.factor_im = factor_im;
  return OK;
}

// Workspace of tsim_simulate_jacobian_noalloc: the compiled event codes and
// macro-ops of one stream of at most kNoAllocEvents events, in a device
// array that exists for the life of the process (one copy per GPU). Its
// host-side name is not a device address; each device's address comes from
// hipGetSymbolAddress once, with that device current.
constexpr int64_t kNoAllocEvents = 16384;
constexpr size_t kNoAllocCodeBytes = sizeof(EventCode) * (size_t)kNoAllocEvents;
constexpr size_t kNoAllocBytes = kNoAllocCodeBytes + sizeof(MacroOp) * (size_t)(5 * kNoAllocEvents + 2);
static_assert(kNoAllocCodeBytes % 256 == 0, "macro-ops start 256-byte aligned");

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
unsigned char* g_scratch_address[kMaxDevices] = {};

// The current device's copy of `scratch`, looked up on the first call there.
int scratch_address(unsigned char** out) {
  int device = -1;
  int st = hip_status(hipGetDevice(&device), "hipGetDevice");
  if (st != OK) return st;
  if (device < 0 || device >= kMaxDevices) {
    set_error("device %d is beyond the %d this library tracks", device, kMaxDevices);
    return UNSUPPORTED;
  }
  std::lock_guard<std::mutex> hold(g_scratch_mutex);
  if (g_scratch_address[device] == nullptr) {
    void* address = nullptr;
    st = hip_status(hipGetSymbolAddress(&address, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (st != OK) return st;
    if (address == nullptr) {
      set_error("hipGetSymbolAddress(scratch) returned no address");
      return HIP_FAILURE;
    }
    g_scratch_address[device] = (unsigned char*)address;
  }
  *out = g_scratch_address[device];
  return OK;
}

}  // namespace

extern "C" {

int tsim_simulate_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                           int64_t state_count, int64_t output_count, int features,
                           int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                           float factor_re, float factor_im, const float* sample_sign,
                           void* signal, void* jacobian, int tile, void* stream) {
  RealArgs args;
  int tm = 0;
  int st = jacobian_args(args, tm, packed, atom_count, event_count, state_count, output_count, features,
                         tissue_stride, columns, column_count, factor_re, factor_im, sample_sign, signal, jacobian);
  if (st != OK) return st;
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  st = compile_ops(event_count, output_count, compiled);
  if (st != OK) return st;
  args.ops = compiled.ops;
  return run_real(args, tm, features, tile, s);
}

int64_t tsim_noalloc_max_events(void) { return kNoAllocEvents; }

int tsim_simulate_jacobian_noalloc(void* const* packed, int64_t atom_count, int64_t event_count,
                                   int64_t state_count, int64_t output_count, int features,
                                   int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                                   float factor_re, float factor_im, const float* sample_sign,
                                   void* signal, void* jacobian, int tile, void* stream) {
  RealArgs args;
  int tm = 0;
  int st = jacobian_args(args, tm, packed, atom_count, event_count, state_count, output_count, features,
                         tissue_stride, columns, column_count, factor_re, factor_im, sample_sign, signal, jacobian);
  if (st != OK) return st;
  if (event_count > kNoAllocEvents) {
    set_error("tsim_simulate_jacobian_noalloc takes at most %lld events, got %lld",
              (long long)kNoAllocEvents, (long long)event_count);
    return BAD_ARGUMENT;
  }
  for (int i : {E_DURATION, E_KIND, E_FLIP, E_PHASE, E_ACTION, E_OUTPUT_INDEX}) {
    if (event_count > 0 && packed[TISSUE_COUNT + i] == nullptr) {
      set_error("event buffer %d is null", i);
      return BAD_ARGUMENT;
    }
  }
  unsigned char* base = nullptr;
  st = scratch_address(&base);
  if (st != OK) return st;
  EventCode* codes = (EventCode*)base;
  MacroOp* ops = (MacroOp*)(base + kNoAllocCodeBytes);
  hipStream_t s = (hipStream_t)stream;
  if (event_count > 0) {
    st = compile_stream((const float*)packed[TISSUE_COUNT + E_DURATION],
                        (const int32_t*)packed[TISSUE_COUNT + E_KIND], (const float*)packed[TISSUE_COUNT + E_FLIP],
                        (const float*)packed[TISSUE_COUNT + E_PHASE],
                        (const uint8_t*)packed[TISSUE_COUNT + E_ACTION],
                        (const int32_t*)packed[TISSUE_COUNT + E_OUTPUT_INDEX], event_count, codes, s);
    if (st != OK) return st;
  }
  // pack_ops writes every op it emits and the end markers after them, so
  // nothing a previous call left in the workspace is read.
  st = pack_stream(codes, event_count, output_count, ops, s);
  if (st != OK) return st;
  args.ops = ops;
  return run_real(args, tm, features, tile, s);
}

int tsim_simulate_projected(void* const* packed, int64_t atom_count, int64_t event_count,
                            int64_t state_count, int64_t output_count, int features,
                            int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                            float factor_re, float factor_im, const float* sample_sign,
                            const void* basis, int64_t rank, void* signal, void* jacobian, int tile,
                            void* stream) {
  if (packed == nullptr) {
    set_error("packed buffers are null");
    return BAD_ARGUMENT;
  }
  const int64_t max_rank = (int64_t)PROJECT_PER_LANE * kTiles[kTileCount - 1].l;
  if (rank < 1 || rank > max_rank) {
    set_error("a projection keeps between 1 and %lld basis directions, got %lld", (long long)max_rank,
              (long long)rank);
    return UNSUPPORTED;
  }
  if (basis == nullptr && atom_count > 0 && output_count > 0) {
    set_error("the basis buffer is null");
    return BAD_ARGUMENT;
  }
  if (column_count < 0 || column_count > 4 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 4 Jacobian columns, got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (column_count > 0 && jacobian == nullptr && atom_count > 0 && output_count > 0) {
    set_error("Jacobian columns asked for but the Jacobian buffer is null");
    return BAD_ARGUMENT;
  }
  if (features & FEATURE_DIFFUSING) {
    set_error("a projected launch carries no diffusion");
    return UNSUPPORTED;
  }
  RealArgs args;
  int st = fill_real_args(args, packed, atom_count, event_count, state_count, output_count, features,
                          tissue_stride);
  if (st != OK) return st;
  int tm = 0;
  int cols[5];
  if (!tangent_mask(columns, column_count, tm, cols)) return BAD_ARGUMENT;
  if (cols[3] >= 0) {
    set_error("a projected launch carries no inversion-efficiency column");
    return UNSUPPORTED;
  }
  if (cols[2] >= 0 && !(features & FEATURE_TRANSMIT)) {
    set_error("a B1 column needs the transmit feature and a b1 buffer");
    return BAD_ARGUMENT;
  }
  const int min_lanes = (int)((rank + PROJECT_PER_LANE - 1) / PROJECT_PER_LANE);
  if (tile < 0) tile = default_tile(args.states, popcount4(tangent_class(tm)), min_lanes);
  if (tile < 0 || tile >= kTileCount || kTiles[tile].sl * kTiles[tile].l < args.states) {
    set_error("no state tile covers %d orders", args.states);
    return UNSUPPORTED;
  }
  if (kTiles[tile].l < min_lanes) {
    set_error("tile %d has %d lanes per atom, fewer than the %d that %lld directions need", tile,
              kTiles[tile].l, min_lanes, (long long)rank);
    return BAD_ARGUMENT;
  }
  args.col_t1 = cols[0];
  args.col_t2 = cols[1];
  args.col_b1 = cols[2];
  args.col_ie = -1;
  args.col_m0 = cols[4];
  args.columns = (int)column_count;
  args.signal = (float2*)signal;
  args.jacobian = column_count > 0 ? (float2*)jacobian : nullptr;
  args.sample_sign = sample_sign;
  args.factor_re = factor_re;
  args.factor_im = factor_im;
  args.basis = (const float2*)basis;
  args.rank = (int)rank;
  args.stage_k_log2 = -1;  // no sample leaves the registers
  args.stage_s_log2 = 0;
  if (args.atoms == 0) return OK;
  if (args.outputs == 0) {
    set_error("a projection needs at least one recorded sample");
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  st = compile_ops(event_count, output_count, compiled);
  if (st != OK) return st;
  args.ops = compiled.ops;
  const int L = kTiles[tile].l;
  const int threads = real_block_threads(atom_count * L);
  const int64_t blocks = (atom_count * L + threads - 1) / threads;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)atom_count);
    return BAD_ARGUMENT;
  }
  RealLaunch launch = pick_real(tile, tm, variant_of(features, false, tile, args.states) | V_PROJECT);
  if (!launch) {
    set_error("no projecting kernel for this variant");
    return UNSUPPORTED;
  }
  launch(args, dim3((unsigned)blocks), dim3(threads), 0, s);
  return hip_status(hipGetLastError(), "real_kernel launch");
}

int tsim_tile_count(void) { return kTileCount; }

int tsim_tile_shape(int tile, int* orders_per_lane, int* lanes_per_atom) {
  if (tile < 0 || tile >= kTileCount) {
    set_error("no tile %d", tile);
    return BAD_ARGUMENT;
  }
  *orders_per_lane = kTiles[tile].sl;
  *lanes_per_atom = kTiles[tile].l;
  return OK;
}

}  // extern "C"

extern "C" {

int tsim_gauss_newton_normal(void* const* packed, int64_t atom_count, int64_t event_count,
                             int64_t state_count, int64_t output_count, int features,
                             int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                             float factor_re, float factor_im, const float* sample_sign,
                             const void* data, double* normal, int tile, void* stream) {
  if (packed == nullptr) {
    set_error("packed buffers are null");
    return BAD_ARGUMENT;
  }
  if (column_count < 1 || column_count > NORMAL_COLUMNS || columns == nullptr) {
    set_error("between 1 and %d fitted columns, got %lld", NORMAL_COLUMNS, (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (atom_count > 0 && output_count > 0 && (data == nullptr || normal == nullptr)) {
    set_error("measured data and the normal-equation buffer are required");
    return BAD_ARGUMENT;
  }
  if (features & FEATURE_DIFFUSING) {
    set_error("the normal-equation kernel carries no diffusion term");
    return UNSUPPORTED;
  }
  RealArgs args;
  int st = fill_real_args(args, packed, atom_count, event_count, state_count, output_count, features,
                          tissue_stride);
  if (st != OK) return st;
  int tm = 0;
  int cols[5];
  if (!tangent_mask(columns, column_count, tm, cols)) return BAD_ARGUMENT;
  if (cols[3] >= 0 || tm > 7) {
    set_error("the normal-equation kernel fits t1, t2, b1 and m0");
    return UNSUPPORTED;
  }
  if (cols[2] >= 0 && !(features & FEATURE_TRANSMIT)) {
    set_error("a B1 column needs the transmit feature and a b1 buffer");
    return BAD_ARGUMENT;
  }
  args.col_t1 = cols[0];
  args.col_t2 = cols[1];
  args.col_b1 = cols[2];
  args.col_m0 = cols[4];
  args.columns = (int)column_count;
  args.sample_sign = sample_sign;
  args.factor_re = factor_re;
  args.factor_im = factor_im;
  args.data = (const float2*)data;
  args.normal = normal;
  if (atom_count == 0 || output_count == 0) return OK;
  if (tile < 0) tile = default_tile(args.states, popcount4(tangent_class(tm)), 1, true);
  if (tile < 0 || tile >= kTileCount || kTiles[tile].sl * kTiles[tile].l < args.states) {
    set_error("no state tile covers %d orders", args.states);
    return UNSUPPORTED;
  }
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  st = compile_ops(event_count, output_count, compiled);
  if (st != OK) return st;
  args.ops = compiled.ops;
  const int L = kTiles[tile].l;
  const int threads = real_block_threads(atom_count * L);
  const int64_t blocks = (atom_count * L + threads - 1) / threads;
  RealLaunch launch = pick_real(tile, tm, variant_of(features, true, tile, args.states));
  if (!launch) {
    set_error("no normal-equation kernel for this variant");
    return UNSUPPORTED;
  }
  set_staging(args, L);
  launch(args, dim3((unsigned)blocks), dim3(threads), stage_bytes(args, L, threads), s);
  return hip_status(hipGetLastError(), "normal-equation kernel launch");
}

int tsim_normal_from_jacobian(const void* signal, const void* jacobian, const void* data, int64_t voxels,
                              int64_t outputs, int columns, double* normal, void* stream) {
  if (columns < 1 || columns > NORMAL_COLUMNS || voxels < 0 || outputs < 0) {
    set_error("between 1 and %d columns and non-negative sizes", NORMAL_COLUMNS);
    return BAD_ARGUMENT;
  }
  if (voxels == 0) return OK;
  if (signal == nullptr || jacobian == nullptr || data == nullptr || normal == nullptr) {
    set_error("signal, jacobian, data and normal buffers are required");
    return BAD_ARGUMENT;
  }
  const int64_t blocks = (voxels + 3) / 4;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many voxels for one launch: %lld", (long long)voxels);
    return BAD_ARGUMENT;
  }
  hipLaunchKernelGGL(normal_from_jacobian, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)signal, (const float2*)jacobian, (const float2*)data, voxels, outputs, columns,
                     normal);
  return hip_status(hipGetLastError(), "normal_from_jacobian launch");
}

int tsim_solve_damped(const double* normal, const double* scale, const double* damping, const double* offset,
                      int64_t voxels, int columns, double* step, double* predicted, double* gradient,
                      double* curvature, void* stream) {
  if (columns < 1 || columns > NORMAL_COLUMNS || voxels < 0) {
    set_error("between 1 and %d columns and a non-negative voxel count", NORMAL_COLUMNS);
    return BAD_ARGUMENT;
  }
  if (voxels == 0) return OK;
  if (normal == nullptr) {
    set_error("normal-equation buffer is null");
    return BAD_ARGUMENT;
  }
  const int64_t blocks = (voxels + 255) / 256;
  hipLaunchKernelGGL(solve_damped, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, normal, scale,
                     damping, offset, voxels, columns, step, predicted, gradient, curvature);
  return hip_status(hipGetLastError(), "solve_damped launch");
}

int tsim_normal_stride(void) { return NORMAL_STRIDE; }

int tsim_lm_propose(const double* normal, const double* x, const double* damping, const int32_t* bound_kind,
                    const double* low, const double* high, int64_t voxels, int columns, double* candidate,
                    float* natural, double* predicted, uint8_t* singular, void* stream) {
  if (columns < 1 || columns > NORMAL_COLUMNS || voxels < 0) {
    set_error("between 1 and %d columns and a non-negative voxel count", NORMAL_COLUMNS);
    return BAD_ARGUMENT;
  }
  if (voxels == 0) return OK;
  if (!normal || !x || !damping || !bound_kind || !low || !high || !candidate || !natural || !predicted ||
      !singular) {
    set_error("every lm_propose buffer is required");
    return BAD_ARGUMENT;
  }
  LMBounds b = {};
  for (int c = 0; c < columns; ++c) {
    if (bound_kind[c] < 0 || bound_kind[c] > 3) {
      set_error("column %d: bound kind %d (0 free, 1 two-sided, 2 lower, 3 upper)", c, bound_kind[c]);
      return BAD_ARGUMENT;
    }
    b.kind[c] = bound_kind[c];
    b.low[c] = low[c];
    b.high[c] = high[c];
  }
  const int64_t blocks = (voxels + 255) / 256;
  hipLaunchKernelGGL(lm_propose, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, normal, x, damping, b,
                     voxels, columns, candidate, natural, predicted, singular);
  return hip_status(hipGetLastError(), "lm_propose launch");
}

int tsim_lm_accept(const double* trial, const double* predicted, const uint8_t* singular, const double* candidate,
                   int64_t voxels, int columns, double* x, double* normal, double* cost, double* damping,
                   double* rising, void* stream) {
  if (columns < 1 || columns > NORMAL_COLUMNS || voxels < 0) {
    set_error("between 1 and %d columns and a non-negative voxel count", NORMAL_COLUMNS);
    return BAD_ARGUMENT;
  }
  if (voxels == 0) return OK;
  if (!trial || !predicted || !singular || !candidate || !x || !normal || !cost || !damping || !rising) {
    set_error("every lm_accept buffer is required");
    return BAD_ARGUMENT;
  }
  const int64_t blocks = (voxels + 255) / 256;
  hipLaunchKernelGGL(lm_accept, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream, trial, predicted,
                     singular, candidate, voxels, columns, x, normal, cost, damping, rising);
  return hip_status(hipGetLastError(), "lm_accept launch");
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Uneven-batch launch
// ---------------------------------------------------------------------------

namespace {

// Pinned staging for the request and work tables. Buffers are allocated once
// and reused: a buffer returns to the free list when the copy that read it
// has completed (checked at the next batch call). Allocating and freeing
// pinned memory per launch would stall the host (hipHostFree waits for the
// device), so nothing is freed while the library is loaded.
struct Staged {
  char* host = nullptr;
  size_t bytes = 0;
  hipEvent_t copied = nullptr;
};
std::mutex g_staged_mutex;
std::vector<Staged> g_in_flight;
std::vector<Staged> g_landed;

int allocate_staging(size_t bytes, Staged& fresh) {
  size_t size = 65536;
  while (size < bytes) size <<= 1;
  int st = hip_status(hipHostMalloc((void**)&fresh.host, size, hipHostMallocDefault), "hipHostMalloc(staging)");
  if (st != OK) return st;
  fresh.bytes = size;
  st = hip_status(hipEventCreateWithFlags(&fresh.copied, hipEventDisableTiming), "hipEventCreate");
  if (st != OK) fresh.copied = nullptr;
  return st;
}

// Keep one idle buffer of at least `bytes` in reserve (called with the lock
// held): a capture of the same batch then finds it without allocating.
void keep_spare(size_t bytes) {
  for (const Staged& idle : g_landed)
    if (idle.bytes >= bytes) return;
  Staged spare;
  if (allocate_staging(bytes, spare) == OK) g_landed.push_back(spare);
}

int take_staging(size_t bytes, Staged& out) {
  std::lock_guard<std::mutex> lock(g_staged_mutex);
  for (size_t i = 0; i < g_in_flight.size();) {
    if (hipEventQuery(g_in_flight[i].copied) == hipSuccess) {
      g_landed.push_back(g_in_flight[i]);
      g_in_flight[i] = g_in_flight.back();
      g_in_flight.pop_back();
    } else {
      ++i;
    }
  }
  size_t best = g_landed.size();
  for (size_t i = 0; i < g_landed.size(); ++i)
    if (g_landed[i].bytes >= bytes && (best == g_landed.size() || g_landed[i].bytes < g_landed[best].bytes)) best = i;
  if (best < g_landed.size()) {
    out = g_landed[best];
    g_landed[best] = g_landed.back();
    g_landed.pop_back();
    keep_spare(bytes);
    return OK;
  }
  Staged fresh;
  const int st = allocate_staging(bytes, fresh);
  if (st != OK) {
    if (fresh.host != nullptr) g_landed.push_back(Staged{fresh.host, fresh.bytes, nullptr});
    return st;
  }
  out = fresh;
  keep_spare(bytes);
  return OK;
}

// A buffer a captured graph reads at every replay: kept until the library
// is unloaded. Captures take an idle buffer (no allocation happens while a
// stream captures), so a batch is launched once eagerly before it is captured.
std::vector<Staged> g_retained;

int take_staging_for_capture(size_t bytes, Staged& out) {
  std::lock_guard<std::mutex> lock(g_staged_mutex);
  size_t best = g_landed.size();
  for (size_t i = 0; i < g_landed.size(); ++i)
    if (g_landed[i].bytes >= bytes && (best == g_landed.size() || g_landed[i].bytes < g_landed[best].bytes)) best = i;
  if (best == g_landed.size()) {
    set_error("capturing a batch needs an idle staging buffer: launch the batch once before capturing it");
    return UNSUPPORTED;
  }
  out = g_landed[best];
  g_landed[best] = g_landed.back();
  g_landed.pop_back();
  g_retained.push_back(out);
  return OK;
}

// Hand a buffer back after its copy was enqueued (its event recorded), or,
// with `landed`, when nothing was enqueued from it.
void give_staging(const Staged& used, bool landed) {
  std::lock_guard<std::mutex> lock(g_staged_mutex);
  if (landed || used.copied == nullptr)
    g_landed.push_back(used);
  else
    g_in_flight.push_back(used);
}

}  // namespace

namespace tsim {

int upload_table(const void* host, size_t bytes, void* device, hipStream_t s) {
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  int st = hip_status(hipStreamIsCapturing(s, &capture), "hipStreamIsCapturing");
  if (st != OK) return st;
  Staged staging;
  bool have_staging = false;
  if (capture != hipStreamCaptureStatusNone) {
    // the graph replays the copy from this buffer: it is never reused
    st = take_staging_for_capture(bytes, staging);
  } else {
    st = take_staging(bytes, staging);
    have_staging = st == OK;
  }
  bool enqueued = false;
  if (st == OK) {
    memcpy(staging.host, host, bytes);
    st = hip_status(hipMemcpyAsync(device, staging.host, bytes, hipMemcpyHostToDevice, s), "hipMemcpyAsync(table)");
    enqueued = st == OK;
  }
  if (st == OK && have_staging) st = hip_status(hipEventRecord(staging.copied, s), "hipEventRecord");
  if (have_staging) {
    if (enqueued && st != OK) (void)hipStreamSynchronize(s);
    give_staging(staging, st != OK);
  }
  return st;
}

}  // namespace tsim

extern "C" {

typedef struct tsim_batch_request {
  void* packed[26];
  int64_t atom_count;
  int64_t event_count;
  int64_t state_count;
  int64_t output_count;
  int64_t tissue_stride;
  int32_t features;
  int32_t column_count;
  int32_t columns[5];
  float factor_re;
  float factor_im;
  const float* sample_sign;
  void* signal;
  void* jacobian;
  const void* data;
  double* normal;
} tsim_batch_request;

int tsim_simulate_batch(const tsim_batch_request* requests, int64_t count, int mode, int tile,
                        int64_t grid_blocks, void* stream) {
  if (count < 0 || (count > 0 && requests == nullptr)) {
    set_error("a batch needs its request table");
    return BAD_ARGUMENT;
  }
  if (mode != 0 && mode != 1) {
    set_error("mode 0 (signal and Jacobian) or 1 (normal equations), got %d", mode);
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<RealArgs> args;
  std::vector<int64_t> code_offset;
  args.reserve((size_t)count);
  int key_tm = -1, key_diff = -1, key_b1 = -1, max_states = 0;
  int64_t total_events = 0;
  for (int64_t r = 0; r < count; ++r) {
    const tsim_batch_request& q = requests[r];
    if (q.atom_count == 0 || q.output_count == 0) continue;
    RealArgs a;
    int st = fill_real_args(a, q.packed, q.atom_count, q.event_count, q.state_count, q.output_count, q.features,
                            q.tissue_stride);
    if (st != OK) return st;
    int tm = 0;
    int cols[5];
    if (q.column_count < 0 || q.column_count > 5 || !tangent_mask(q.columns, q.column_count, tm, cols)) {
      if (q.column_count < 0 || q.column_count > 5) set_error("request %lld: 0 to 5 columns", (long long)r);
      return BAD_ARGUMENT;
    }
    if (cols[2] >= 0 && !(q.features & FEATURE_TRANSMIT)) {
      set_error("request %lld: a B1 column needs the transmit feature", (long long)r);
      return BAD_ARGUMENT;
    }
    const int diff = (q.features & FEATURE_DIFFUSING) ? 1 : 0;
    const int b1 = (q.features & FEATURE_TRANSMIT) ? 1 : 0;
    if (mode == 1 && (tm > 7 || diff || q.column_count < 1 || q.column_count > NORMAL_COLUMNS || !q.data || !q.normal)) {
      set_error("request %lld: normal-equation mode fits 1-4 of t1, t2, m0, b1 with data and output", (long long)r);
      return BAD_ARGUMENT;
    }
    if (key_tm < 0) {
      key_tm = tm;
      key_diff = diff;
      key_b1 = b1;
    } else if (tangent_class(tm) != tangent_class(key_tm) || diff != key_diff || b1 != key_b1) {
      set_error("request %lld runs another kernel variant than the batch (group requests by variant)",
                (long long)r);
      return BAD_ARGUMENT;
    }
    a.col_t1 = cols[0];
    a.col_t2 = cols[1];
    a.col_b1 = cols[2];
    a.col_ie = cols[3];
    a.col_m0 = cols[4];
    a.columns = q.column_count;
    a.sample_sign = q.sample_sign;
    a.factor_re = q.factor_re;
    a.factor_im = q.factor_im;
    a.signal = (float2*)q.signal;
    a.jacobian = q.column_count > 0 ? (float2*)q.jacobian : nullptr;
    a.data = (const float2*)q.data;
    a.normal = q.normal;
    max_states = a.states > max_states ? a.states : max_states;
    code_offset.push_back(total_events);
    total_events += q.event_count;
    args.push_back(a);
  }
  if (args.empty()) return OK;
  if (tile < 0) tile = default_tile(max_states, popcount4(tangent_class(key_tm)), 1, mode == 1);
  if (tile < 0 || tile >= kTileCount || kTiles[tile].sl * kTiles[tile].l < max_states) {
    set_error("no state tile covers %d orders", max_states);
    return UNSUPPORTED;
  }
  const int L = kTiles[tile].l;
  const int64_t per_block = 256 / L;
  size_t shared = 0;
  for (RealArgs& a : args) {
    set_staging(a, L);
    const size_t bytes = stage_bytes(a, L);
    shared = bytes > shared ? bytes : shared;
  }
  std::vector<BatchWork> work;
  for (size_t r = 0; r < args.size(); ++r) {
    for (int64_t first = 0; first < args[r].atoms; first += per_block) {
      BatchWork w;
      w.request = (int32_t)r;
      w.pad = 0;
      w.first_atom = first;
      work.push_back(w);
    }
  }
  // one workspace: compiled events, their macro-ops, the request table, the
  // work table
  int64_t total_ops = 0;
  for (const RealArgs& a : args) total_ops += op_capacity(a.event_count);
  const size_t codes_bytes = sizeof(EventCode) * (size_t)total_events;
  const size_t ops_bytes = sizeof(MacroOp) * (size_t)total_ops;
  const size_t args_bytes = sizeof(RealArgs) * args.size();
  const size_t work_bytes = sizeof(BatchWork) * work.size();
  const size_t o_off = (codes_bytes + 255) & ~(size_t)255;
  const size_t a_off = (o_off + ops_bytes + 255) & ~(size_t)255;
  const size_t w_off = (a_off + args_bytes + 255) & ~(size_t)255;
  char* space = nullptr;
  int st = scratch_alloc((void**)&space, w_off + work_bytes, s, "batch workspace");
  if (st != OK) return st;
  EventCode* codes = (EventCode*)space;
  MacroOp* ops = (MacroOp*)(space + o_off);
  size_t k = 0;
  int64_t op_offset = 0;
  for (int64_t r = 0; r < count && st == OK; ++r) {
    const tsim_batch_request& q = requests[r];
    if (q.atom_count == 0 || q.output_count == 0) continue;
    EventCode* mine = codes + code_offset[k];
    st = compile_stream((const float*)q.packed[TISSUE_COUNT + E_DURATION],
                        (const int32_t*)q.packed[TISSUE_COUNT + E_KIND],
                        (const float*)q.packed[TISSUE_COUNT + E_FLIP],
                        (const float*)q.packed[TISSUE_COUNT + E_PHASE],
                        (const uint8_t*)q.packed[TISSUE_COUNT + E_ACTION],
                        (const int32_t*)q.packed[TISSUE_COUNT + E_OUTPUT_INDEX], q.event_count, mine, s);
    if (st == OK) st = pack_stream(mine, q.event_count, q.output_count, ops + op_offset, s);
    args[k].ops = ops + op_offset;
    op_offset += op_capacity(q.event_count);
    ++k;
  }
  hipStreamCaptureStatus capture = hipStreamCaptureStatusNone;
  if (st == OK) st = hip_status(hipStreamIsCapturing(s, &capture), "hipStreamIsCapturing");
  const bool capturing = capture != hipStreamCaptureStatusNone;
  Staged staging;
  bool have_staging = false;
  if (st == OK && capturing) {
    // the graph replays the table copy from this buffer: it is never reused
    st = take_staging_for_capture(w_off + work_bytes, staging);
  } else if (st == OK) {
    st = take_staging(w_off + work_bytes, staging);
    have_staging = st == OK;
  }
  bool enqueued = false;
  if (st == OK) {
    memcpy(staging.host + a_off, args.data(), args_bytes);
    memcpy(staging.host + w_off, work.data(), work_bytes);
    st = hip_status(hipMemcpyAsync(space + a_off, staging.host + a_off, w_off + work_bytes - a_off,
                                   hipMemcpyHostToDevice, s),
                    "hipMemcpyAsync(batch tables)");
    enqueued = st == OK;
  }
  if (st == OK && have_staging) st = hip_status(hipEventRecord(staging.copied, s), "hipEventRecord");
  if (have_staging) {
    // a copy that was enqueued but whose event could not be recorded is waited
    // for here before the buffer can be reused
    if (enqueued && st != OK) (void)hipStreamSynchronize(s);
    give_staging(staging, st != OK);
  }
  if (st == OK) {
    const int64_t blocks = (int64_t)work.size();
    int64_t grid = grid_blocks > 0 ? grid_blocks : blocks;
    if (grid > blocks) grid = blocks;
    int variant = (key_diff ? V_DIFF : 0) | (key_b1 ? V_B1MAP : 0) | (mode == 1 ? V_ACCUM : 0);
    for (const RealArgs& a : args) {
      if ((int64_t)kTiles[tile].sl * kTiles[tile].l != a.states) variant |= V_PADDED;
    }
    BatchLaunch launch = pick_batch(tile, key_tm, variant);
    if (!launch) {
      set_error("no batch kernel for this variant");
      st = UNSUPPORTED;
    } else {
      launch((const RealArgs*)(space + a_off), (const BatchWork*)(space + w_off), blocks, dim3((unsigned)grid),
             dim3(256), shared, s);
      st = hip_status(hipGetLastError(), "real_batch_kernel launch");
    }
  }
  (void)hipFreeAsync(space, s);
  return st;
}

int tsim_batch_request_size(void) { return (int)sizeof(tsim_batch_request); }

}  // extern "C"
