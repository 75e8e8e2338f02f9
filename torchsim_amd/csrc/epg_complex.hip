// Host side of the complex EPG kernel: argument checks, lane-count choice and
// the C ABI entry points (declared in include/torchsim_hip.h).
#include "epg_complex_impl.h"

#include <cstring>
#include <vector>

#include "../../include/torchsim_hip.h"

namespace tsim {
namespace {

ComplexLaunch pick_complex(int lanes, int tm, bool exch) {
  switch (lanes) {
#define TSIM_COMPLEX_CASE(LANES) \
  case LANES:                    \
    return pick_complex_##LANES(tm, exch);
    TSIM_COMPLEX_LANES(TSIM_COMPLEX_CASE)
#undef TSIM_COMPLEX_CASE
    default:
      return nullptr;
  }
}

// Fewest lanes covering the orders; 10, 12 and 20 fit 6, 5 and 3 atoms in a
// wave (60 of 64 lanes busy) where 16 and 32 would fit 4 and 2.
int lanes_for(int64_t states) {
  for (int l : {4, 8, 10, 12, 16, 20, 32, 64})
    if (states <= l) return l;
  return -1;
}

int fill_complex(ComplexArgs& a, void* const* packed, int64_t atoms, int64_t events, int64_t states,
                 int64_t outputs, int pool_kind, int features, int64_t stride) {
  if (atoms < 0 || events < 0 || states < 1 || outputs < 0) {
    set_error("invalid EPG buffer dimensions");
    return BAD_ARGUMENT;
  }
  if (states > 64) {
    set_error("at most 64 configuration orders are carried, %lld asked", (long long)states);
    return UNSUPPORTED;
  }
  if (pool_kind != 0 && pool_kind != 2) {
    set_error("pool kind %d is not built (single pool 0 or one exchanging pool 2)", pool_kind);
    return UNSUPPORTED;
  }
  if (stride != 0 && stride != 1) {
    set_error("tissue_stride must be 0 or 1");
    return BAD_ARGUMENT;
  }
  if (packed[T_T1] == nullptr || packed[T_T2] == nullptr) {
    set_error("t1 and t2 buffers are required");
    return BAD_ARGUMENT;
  }
  memset(&a, 0, sizeof(a));
  a.event_count = events;
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
  a.stride_opt = stride;
  a.atoms = atoms;
  a.outputs = outputs;
  a.states = (int)states;
  a.density = (features & FEATURE_DENSITY) && a.m0 ? 1 : 0;
  a.inverting = (features & FEATURE_INVERTING) && a.inv_eff ? 1 : 0;
  a.transmit = (features & FEATURE_TRANSMIT) && a.b1 ? 1 : 0;
  a.off_axis = (features & FEATURE_OFF_AXIS) ? 1 : 0;
  a.diffusing = (features & FEATURE_DIFFUSING) && a.diffusion ? 1 : 0;
  a.factor_re = 1.0f;
  a.col_t1 = a.col_t2 = a.col_b1 = a.col_m0 = -1;
  if (pool_kind == 2 && (!a.pool_fraction || !a.pool_exchange || !a.t1_pool || !a.t2_pool)) {
    set_error("an exchanging pool needs its fraction, exchange rate, T1 and T2 buffers");
    return BAD_ARGUMENT;
  }
  return OK;
}

// Output staging: K samples per flush with the padded stream count, one
// wave's stage within COMPLEX_STAGE_BYTES_PER_WAVE; K < 2 stores directly.
// Returns the dynamic LDS bytes a block of `threads` needs.
size_t set_staging(ComplexArgs& a, int lanes, int threads, int min_s_log2 = 0) {
  a.stage_k_log2 = -1;
  a.stage_s_log2 = 0;
  if (a.out_real != nullptr || (a.signal == nullptr && a.jacobian == nullptr)) return 0;
  int s_log2 = min_s_log2;
  while ((1 << s_log2) < 1 + a.columns) ++s_log2;
  const int64_t slot = (int64_t)(64 / lanes) * (1 << s_log2) * (int64_t)sizeof(float2);
  int k_log2 = 0;
  while (k_log2 < 4 && slot * (2 << k_log2) <= COMPLEX_STAGE_BYTES_PER_WAVE) ++k_log2;
  if (k_log2 < 1) return 0;
  a.stage_k_log2 = k_log2;
  a.stage_s_log2 = s_log2;
  return (size_t)(threads / 64) * (size_t)slot * ((size_t)1 << k_log2);
}

ComplexBatchLaunch pick_complex_batch(int lanes, int tm) {
  switch (lanes) {
#define TSIM_COMPLEX_BATCH_CASE(LANES) \
  case LANES:                          \
    return pick_complex_batch_##LANES(tm);
    TSIM_COMPLEX_LANES(TSIM_COMPLEX_BATCH_CASE)
#undef TSIM_COMPLEX_BATCH_CASE
    default:
      return nullptr;
  }
}

// The tangent mask and column slots of a request's columns (t1, t2, m0, b1).
int complex_columns(const int32_t* columns, int64_t count, int& tm, int (&cols)[4]) {
  tm = 0;
  for (int k = 0; k < 4; ++k) cols[k] = -1;
  for (int64_t c = 0; c < count; ++c) {
    int slot;
    switch (columns[c]) {
      case T_T1: tm |= CT1; slot = 0; break;
      case T_T2: tm |= CT2; slot = 1; break;
      case T_B1: tm |= CB1; slot = 2; break;
      case T_M0: slot = 3; break;
      default:
        set_error("column %lld: property %d carries no complex-kernel tangent", (long long)c, columns[c]);
        return BAD_ARGUMENT;
    }
    if (cols[slot] >= 0) {
      set_error("property %d appears twice among the Jacobian columns", columns[c]);
      return BAD_ARGUMENT;
    }
    cols[slot] = (int)c;
  }
  return OK;
}

int run_complex(ComplexArgs& a, int tm, bool exch, hipStream_t stream) {
  if (a.atoms == 0 || a.outputs == 0) return OK;
  const int lanes = lanes_for(a.states);
  ComplexLaunch launch = pick_complex(lanes, tm, exch);
  if (!launch) {
    set_error("no complex kernel for %d orders", a.states);
    return UNSUPPORTED;
  }
  // one-wave blocks when four-wave blocks would occupy fewer than all 256
  // compute units, so a small launch's waves spread over every unit (the
  // kernel indexes by global wave and wave-within-block only)
  const int64_t waves = (a.atoms + 64 / lanes - 1) / (64 / lanes);
  const int threads = waves < 4LL * 256 ? 64 : 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * (64 / lanes);
  const int64_t blocks = (a.atoms + atoms_per_block - 1) / atoms_per_block;
  if (blocks > 0x7FFFFFFF) {
    set_error("too many atoms for one launch: %lld", (long long)a.atoms);
    return BAD_ARGUMENT;
  }
  const size_t shared = set_staging(a, lanes, threads);
  launch(a, dim3((unsigned)blocks), dim3(threads), shared, stream);
  return hip_status(hipGetLastError(), "complex_kernel launch");
}

}  // namespace
}  // namespace tsim

using namespace tsim;

extern "C" {

int tsim_complex_simulate(void* const* pointers, int64_t pointer_count, int64_t atom_count,
                          int64_t event_count, int64_t state_count, int64_t output_count,
                          int pool_kind, int features, int64_t tissue_stride, void* stream) {
  if (pointers == nullptr || pointer_count != PACKED_COUNT + 2) {
    set_error("wrong number of buffer pointers: %lld, expected %d", (long long)pointer_count,
              PACKED_COUNT + 2);
    return BAD_ARGUMENT;
  }
  ComplexArgs a;
  int st = fill_complex(a, pointers, atom_count, event_count, state_count, output_count, pool_kind,
                        features, tissue_stride);
  if (st != OK) return st;
  a.out_real = (float*)pointers[PACKED_COUNT];
  a.out_imag = (float*)pointers[PACKED_COUNT + 1];
  if ((a.out_real == nullptr || a.out_imag == nullptr) && atom_count > 0 && output_count > 0) {
    set_error("output planes are required");
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(pointers, event_count, s, compiled);
  if (st != OK) return st;
  a.events = compiled.codes;
  return run_complex(a, 0, pool_kind == 2, s);
}

int tsim_complex_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                          int64_t state_count, int64_t output_count, int pool_kind, int features,
                          int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                          float factor_re, float factor_im, void* signal, void* jacobian,
                          void* stream) {
  if (packed == nullptr) {
    set_error("packed buffers are null");
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
  ComplexArgs a;
  int st = fill_complex(a, packed, atom_count, event_count, state_count, output_count, pool_kind,
                        features, tissue_stride);
  if (st != OK) return st;
  int tm = 0;
  int cols[4];
  st = complex_columns(columns, column_count, tm, cols);
  if (st != OK) return st;
  if (cols[2] >= 0 && !a.transmit) {
    set_error("a B1 column needs the transmit feature and a b1 buffer");
    return BAD_ARGUMENT;
  }
  a.col_t1 = cols[0];
  a.col_t2 = cols[1];
  a.col_b1 = cols[2];
  a.col_m0 = cols[3];
  a.columns = (int)column_count;
  a.signal = (float2*)signal;
  a.jacobian = column_count > 0 ? (float2*)jacobian : nullptr;
  a.factor_re = factor_re;
  a.factor_im = factor_im;
  hipStream_t s = (hipStream_t)stream;
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  a.events = compiled.codes;
  return run_complex(a, tm, pool_kind == 2, s);
}

}  // extern "C"

extern "C" int tsim_complex_batch(const tsim_batch_request* requests, int64_t count, int64_t grid_blocks,
                                  void* stream) {
  if (count < 0 || (count > 0 && requests == nullptr)) {
    set_error("a batch needs its request table");
    return BAD_ARGUMENT;
  }
  hipStream_t s = (hipStream_t)stream;
  std::vector<ComplexArgs> args;
  std::vector<int64_t> code_offset;
  std::vector<int64_t> source;
  int key_tm = -1, max_states = 0;
  int64_t total_events = 0;
  for (int64_t r = 0; r < count; ++r) {
    const tsim_batch_request& q = requests[r];
    if (q.atom_count == 0 || q.output_count == 0) continue;
    const int pool_kind = q.packed[T_POOL_B_FRACTION] != nullptr ? 2 : 0;
    ComplexArgs a;
    int st = fill_complex(a, q.packed, q.atom_count, q.event_count, q.state_count, q.output_count, pool_kind,
                          q.features, q.tissue_stride);
    if (st != OK) return st;
    if (q.column_count < 0 || q.column_count > 4) {
      set_error("request %lld: 0 to 4 columns", (long long)r);
      return BAD_ARGUMENT;
    }
    int tm = 0;
    int cols[4];
    st = complex_columns(q.columns, q.column_count, tm, cols);
    if (st != OK) return st;
    if (cols[2] >= 0 && !a.transmit) {
      set_error("request %lld: a B1 column needs the transmit feature and a b1 buffer", (long long)r);
      return BAD_ARGUMENT;
    }
    if (q.sample_sign != nullptr || q.data != nullptr || q.normal != nullptr) {
      set_error("request %lld: the complex batch writes signal and Jacobian (no sample signs, data or normal "
                "equations)", (long long)r);
      return BAD_ARGUMENT;
    }
    if (q.column_count > 0 && q.jacobian == nullptr) {
      set_error("request %lld: Jacobian columns asked for but the Jacobian buffer is null", (long long)r);
      return BAD_ARGUMENT;
    }
    if (key_tm < 0) {
      key_tm = tm;
    } else if (tm != key_tm) {
      set_error("request %lld carries another tangent set than the batch (group requests by tangent set)",
                (long long)r);
      return BAD_ARGUMENT;
    }
    a.col_t1 = cols[0];
    a.col_t2 = cols[1];
    a.col_b1 = cols[2];
    a.col_m0 = cols[3];
    a.columns = (int)q.column_count;
    a.signal = (float2*)q.signal;
    a.jacobian = q.column_count > 0 ? (float2*)q.jacobian : nullptr;
    a.factor_re = q.factor_re;
    a.factor_im = q.factor_im;
    a.exchanging = pool_kind == 2 ? 1 : 0;
    max_states = a.states > max_states ? a.states : max_states;
    code_offset.push_back(total_events);
    source.push_back(r);
    total_events += q.event_count;
    args.push_back(a);
  }
  if (args.empty()) return OK;
  const int lanes = lanes_for(max_states);
  ComplexBatchLaunch launch = pick_complex_batch(lanes, key_tm);
  if (!launch) {
    set_error("no complex batch kernel for %d orders", max_states);
    return UNSUPPORTED;
  }
  const int threads = 256;
  const int64_t atoms_per_block = (int64_t)(threads / 64) * (64 / lanes);
  // The waves of a block may work on different requests at once, each in
  // its own LDS stage region, whose place follows the streams per atom: every
  // request stages as many streams as the widest of the batch, so the regions
  // of all requests lie alike.
  int streams_log2 = 0;
  for (const ComplexArgs& a : args)
    while ((1 << streams_log2) < 1 + a.columns) ++streams_log2;
  size_t shared = 0;
  std::vector<ComplexWork> work;
  for (size_t k = 0; k < args.size(); ++k) {
    const size_t bytes = set_staging(args[k], lanes, threads, streams_log2);
    shared = bytes > shared ? bytes : shared;
    for (int64_t first = 0; first < args[k].atoms; first += atoms_per_block) {
      ComplexWork w;
      w.request = (int32_t)k;
      w.pad = 0;
      w.first_atom = first;
      work.push_back(w);
    }
  }
  // one workspace: the compiled events of every request, then the request and
  // work tables
  const size_t codes_bytes = sizeof(EventCode) * (size_t)total_events;
  const size_t a_off = (codes_bytes + 255) & ~(size_t)255;
  const size_t args_bytes = sizeof(ComplexArgs) * args.size();
  const size_t w_off = (a_off + args_bytes + 255) & ~(size_t)255;
  const size_t work_bytes = sizeof(ComplexWork) * work.size();
  char* space = nullptr;
  int st = scratch_alloc((void**)&space, w_off + work_bytes, s, "complex batch workspace");
  if (st != OK) return st;
  EventCode* codes = (EventCode*)space;
  for (size_t k = 0; k < args.size() && st == OK; ++k) {
    const tsim_batch_request& q = requests[source[k]];
    EventCode* mine = codes + code_offset[k];
    st = compile_stream((const float*)q.packed[TISSUE_COUNT + E_DURATION],
                        (const int32_t*)q.packed[TISSUE_COUNT + E_KIND],
                        (const float*)q.packed[TISSUE_COUNT + E_FLIP],
                        (const float*)q.packed[TISSUE_COUNT + E_PHASE],
                        (const uint8_t*)q.packed[TISSUE_COUNT + E_ACTION],
                        (const int32_t*)q.packed[TISSUE_COUNT + E_OUTPUT_INDEX], q.event_count, mine, s);
    args[k].events = mine;
  }
  if (st == OK) {
    std::vector<char> tables(w_off + work_bytes - a_off);
    memcpy(tables.data(), args.data(), args_bytes);
    memcpy(tables.data() + (w_off - a_off), work.data(), work_bytes);
    st = upload_table(tables.data(), tables.size(), space + a_off, s);
  }
  if (st == OK) {
    const int64_t blocks = (int64_t)work.size();
    int64_t grid = grid_blocks > 0 ? grid_blocks : blocks;
    if (grid > blocks) grid = blocks;
    if (grid > 0x7FFFFFFF) grid = 0x7FFFFFFF;
    launch((const ComplexArgs*)(space + a_off), (const ComplexWork*)(space + w_off), blocks, dim3((unsigned)grid),
           dim3(threads), shared, s);
    st = hip_status(hipGetLastError(), "complex_batch_kernel launch");
  }
  (void)hipFreeAsync(space, s);
  return st;
}
