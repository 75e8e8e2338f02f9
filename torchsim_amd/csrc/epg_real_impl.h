// The real-subspace kernel template, shared by the per-tile translation units.
// See epg_real.hip for what it computes and how an atom is laid out.
#pragma once

#include "epg_common.h"

#include <type_traits>
#include <utility>

namespace tsim {

// The real kernel's unit of work: the stream regrouped so that every op is
// [relax dt] [rotation | inversion | spoil] [record] [shift], each part
// optional, in that order. A pulse of a gradient-spoiled train (relax TR,
// pulse, sample, spoiler shift) is one op, whatever number of reference events
// it was written as. Ops are built on the device by compile_ops from the
// per-event codes; the last one carries OP_END.
constexpr uint32_t OP_RELAX = 1;
constexpr uint32_t OP_ROTATE = 2;
constexpr uint32_t OP_INVERT = 4;
constexpr uint32_t OP_RECORD = 8;
constexpr uint32_t OP_SPOIL = 16;
constexpr uint32_t OP_SHIFT = 32;
constexpr uint32_t OP_END = 64;

struct alignas(32) MacroOp {
  float dt;      // relaxation interval, seconds (OP_RELAX)
  float flip;    // radians (OP_ROTATE)
  float versine; // 1 - cos(flip)
  float sine;    // sin(flip)
  uint32_t ctrl;
  int32_t out;   // output index (OP_RECORD)
  int32_t event; // packed position of the rotation's event
  int32_t relax_event;  // packed position of the event whose interval is the relaxation
};

// Upper bound on the ops E events compile to (each event contributes at most
// five parts), plus the end marker and one more end marker that the kernel's
// one-op-ahead load may read.
inline int64_t op_capacity(int64_t events) { return 5 * events + 2; }

// Ops are read-only for a launch; the constant address space turns the
// wave-uniform index into scalar loads.
__device__ __forceinline__ MacroOp load_op(const MacroOp* ops, int64_t i) {
  using Word = const __attribute__((address_space(4))) uint32_t;
  Word* w = (Word*)(ops + i);
  MacroOp m;
  m.dt = __uint_as_float(w[0]);
  m.flip = __uint_as_float(w[1]);
  m.versine = __uint_as_float(w[2]);
  m.sine = __uint_as_float(w[3]);
  m.ctrl = w[4];
  m.out = (int32_t)w[5];
  m.event = (int32_t)w[6];
  m.relax_event = (int32_t)w[7];
  return m;
}

struct RealArgs {
  const MacroOp* ops;
  int64_t event_count;
  const float* t1;
  const float* t2;
  const float* m0;
  const float* b1;
  const float* inv_eff;
  const float* diffusion;
  int64_t stride_t1;   // 1 = one value per atom, 0 = one value for all
  int64_t stride_t2;
  int64_t stride_opt;  // the optional properties share one stride
  int64_t atoms;
  int64_t outputs;
  int states;
  int density;
  int inverting;
  // Complex64 results in the API layout, each optional.
  float2* signal;    // (atoms, outputs)
  float2* jacobian;  // (atoms, columns, outputs)
  // Reference-layout planes (raw recorded sample, no factor), optional.
  float* out_real;
  float* out_imag;
  const float* sample_sign;  // (outputs) or null
  float factor_re;
  float factor_im;
  int columns;
  int col_t1, col_t2, col_b1, col_ie, col_m0;
  // Normal-equation mode: measured samples and the per-atom accumulators
  // (cost, gradient[4], upper triangle of J^T J [10]) in FP64.
  const float2* data;  // (atoms, outputs) complex64
  double* normal;      // (atoms, NORMAL_STRIDE)
  // Staged output writes: log2 of the pulses buffered per flush and of the
  // padded stream count (signal + columns); stage_k_log2 < 0 writes each
  // sample directly.
  int stage_k_log2;
  int stage_s_log2;
  // Projection mode: the recorded samples are contracted on the fly with a
  // temporal basis (outputs, rank) complex64, and signal / jacobian receive
  // (atoms, rank) and (atoms, columns, rank) coefficients instead of samples.
  const float2* basis;
  int rank;
};

// Staged output writes. Storing one complex sample per atom per pulse puts
// every atom of a wave on its own cache line (the rows of the (atom, output)
// layout are a whole row apart), so lines leave L2 part-written and HBM sees
// several times the bytes. Instead each wave keeps its atoms' last K samples
// of every stream in LDS and writes them every K pulses, K consecutive
// samples of one row together. This needs the records to arrive in output
// order 0, 1, 2, ... (the op compiler checks it and marks the stream).
constexpr int STAGE_BYTES_PER_WAVE = 8192;
constexpr uint32_t OPS_SEQUENTIAL = 128;  // in MacroOp[0].ctrl: records come in output order

// Accumulators one atom hands back in normal-equation mode: the cost
// sum |model - data|^2, J^T r (4), and the upper triangle of J^T J (10),
// columns in the order the caller named them.
constexpr int NORMAL_COLUMNS = 4;
constexpr int NORMAL_STRIDE = 1 + NORMAL_COLUMNS + NORMAL_COLUMNS * (NORMAL_COLUMNS + 1) / 2;

// Tangent directions a kernel carries, as bits of its template mask.
constexpr int TAN_T1 = 1;
constexpr int TAN_T2 = 2;
constexpr int TAN_B1 = 4;
constexpr int TAN_IE = 8;

// Projection mode: each of an atom's L lanes keeps PROJECT_PER_LANE basis
// directions (r = lane-in-atom + L j), so a tile projects onto at most
// PROJECT_PER_LANE * L directions.
constexpr int PROJECT_PER_LANE = 4;

constexpr int popcount4(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1); }
// Slot of direction `bit` among those present in `mask`.
constexpr int slot_of(int mask, int bit) { return popcount4(mask & (bit - 1)); }

// The value of the first lane of this lane's atom (L adjacent lanes, L <= 4,
// so an atom never straddles a DPP quad).
template <int L>
__device__ __forceinline__ float atom_broadcast(float x) {
  static_assert(L == 1 || L == 2 || L == 4, "atoms of 1, 2 or 4 lanes");
  if constexpr (L == 1) {
    return x;
  } else if constexpr (L == 2) {
    // quad_perm [0, 0, 2, 2]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0xA0, 0xF, 0xF, false));
  } else {
    // quad_perm [0, 0, 0, 0]
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(x), 0x00, 0xF, 0xF, false));
  }
}

__device__ __forceinline__ float dpp_from_lower(float x) {
  // row_shr:1 -- lane i receives lane i-1 within its 16-lane row.
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x111, 0xF, 0xF, true));
}
__device__ __forceinline__ float dpp_from_upper(float x) {
  // row_shl:1 -- lane i receives lane i+1 within its 16-lane row.
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(x), 0x101, 0xF, 0xF, true));
}

// The gradient shift of one plane pair (reference shift_real,
// _epg_cpu.cpp:3010-3027): minus moves down an order and loses nothing at the
// top but gains a zero there; plus moves up, drops its top order, and takes
// -minus[0] at order zero.
template <int SL, int L>
__device__ __forceinline__ void shift_pair(float (&a)[SL], float (&b)[SL], bool first, bool last) {
  float b_in = 0.0f;
  float a_in = 0.0f;
  if constexpr (L > 1) {
    b_in = dpp_from_upper(b[0]);
    b_in = last ? 0.0f : b_in;
    a_in = dpp_from_lower(a[SL - 1]);
  }
#pragma unroll
  for (int r = 0; r + 1 < SL; ++r) b[r] = b[r + 1];
  b[SL - 1] = b_in;
#pragma unroll
  for (int r = SL - 1; r > 0; --r) a[r] = a[r - 1];
  if constexpr (L > 1) {
    a[0] = first ? -b[0] : a_in;
  } else {
    a[0] = -b[0];
  }
}

template <int SL>
__device__ __forceinline__ void clear_padding(float (&a)[SL], unsigned keep) {
#pragma unroll
  for (int r = 0; r < SL; ++r) a[r] = ((keep >> r) & 1u) ? a[r] : 0.0f;
}

// ---------------------------------------------------------------------------
// Rotating register file.
//
// Logical order j of this lane's F+ block (a) lives in physical register
// (j + R) mod SL, of F- (b) in (j - R) mod SL, of Z (c) in j, for every plane
// set (value and tangents). A gradient shift then moves no register: R drops
// by one, and only the one value per plane that crosses a block edge travels,
// by DPP, into the register it lands in. The event loop runs in segments
// that each end at a shift, with R a compile-time constant inside a segment:
// SL segment bodies in a row make one turn of the rotation. Tiles wider than
// ROTATE_MAX orders per lane keep R = 0 and move registers on a shift (their
// SL-fold code would not fit the instruction cache).
// ---------------------------------------------------------------------------

constexpr int ROTATE_MAX = 12;

template <int SL, int R>
struct Rot {
  static constexpr int a(int j) { return (j + R) % SL; }
  static constexpr int b(int j) { return ((j - R) % SL + SL) % SL; }
};

// The shift of one plane pair at rotation R (reference shift_real,
// _epg_cpu.cpp:3010-3027): F- moves down an order, F+ up, the top order drops
// and F+[0] = -F-[0]. Afterwards the rotation is R - 1 (mod SL).
template <int SL, int L, int R>
__device__ __forceinline__ void shift_rotating(float (&a)[SL], float (&b)[SL], bool first, bool last) {
  constexpr int RN = (R + SL - 1) % SL;  // rotation after the shift
  constexpr int KA = RN;                 // old logical a[SL-1], new logical a[0]
  constexpr int KB0 = (SL - R) % SL;     // old logical b[0], new logical b[SL-1]
  constexpr int KB1 = (SL - RN) % SL;    // new logical b[0]
  if constexpr (L > 1) {
    const float b_in = dpp_from_upper(b[KB0]);
    const float a_in = dpp_from_lower(a[KA]);
    b[KB0] = last ? 0.0f : b_in;
    a[KA] = first ? -b[KB1] : a_in;
  } else {
    b[KB0] = 0.0f;
    a[KA] = -b[KB1];
  }
}

// Zero the F+ register that holds logical order `pad` (the first order past
// the state count in this lane, or SL when the lane has none) at rotation R.
template <int SL, int R>
__device__ __forceinline__ void clear_order(float (&a)[SL], int pad) {
#pragma unroll
  for (int j = 0; j < SL; ++j) a[Rot<SL, R>::a(j)] = (pad == j) ? 0.0f : a[Rot<SL, R>::a(j)];
}

// e * m + u and e * p - u, computed in the register of m (of p). Written
// with fmaf the compiler picks the two-address v_fmac with u as its
// accumulator wherever u dies, and the result then has to be copied back
// into the plane's register before the next op.
__device__ __forceinline__ float scale_add_in_place(float m, float e, float u) {
  asm("v_fma_f32 %0, %0, %1, %2" : "+v"(m) : "v"(e), "v"(u));
  return m;
}
__device__ __forceinline__ float scale_sub_in_place(float p, float e, float u) {
  asm("v_fma_f32 %0, %0, %1, -%2" : "+v"(p) : "v"(e), "v"(u));
  return p;
}

// sin and cos of x for |x| up to a few thousand radians, branch-free (so the
// scheduler can interleave it with other work). Cody-Waite reduction by
// pi/2 split in three parts (the first two exact in float for |q| < 2^12),
// then the Taylor series on [-pi/4, pi/4] to x^11 and x^12, whose truncation
// error (below 2e-10) is far under float rounding, and the quadrant's swap
// and signs.
__device__ __forceinline__ void sincos_reduced(float x, float& sn, float& cs) {
  constexpr float kTwoOverPi = 0.636619772367581343f;
  constexpr float kP1 = 1.5703125f;                 // 12 significant bits
  constexpr float kP2 = 4.83751296997070312e-4f;    // next 12 bits
  constexpr float kP3 = 7.54978995489188216e-8f;    // remainder of pi/2
  const float q = rintf(x * kTwoOverPi);
  float r = fmaf(-q, kP1, x);
  r = fmaf(-q, kP2, r);
  r = fmaf(-q, kP3, r);
  const float r2 = r * r;
  float ps = fmaf(r2, 2.75573192239858907e-6f, -1.98412698412698413e-4f);
  ps = fmaf(r2, ps, 8.33333333333333333e-3f);
  ps = fmaf(r2, ps, -1.66666666666666667e-1f);
  const float sr = fmaf(r * r2, ps, r);
  float pc = fmaf(r2, -2.75573192239858907e-7f, 2.48015873015873016e-5f);
  pc = fmaf(r2, pc, -1.38888888888888889e-3f);
  pc = fmaf(r2, pc, 4.16666666666666667e-2f);
  pc = fmaf(r2, pc, -0.5f);
  const float cr = fmaf(r2, pc, 1.0f);
  const int iq = (int)q;
  const bool swap = (iq & 1) != 0;
  const float s0 = swap ? cr : sr;
  const float c0 = swap ? sr : cr;
  sn = (iq & 2) ? -s0 : s0;
  cs = ((iq + 1) & 2) ? -c0 : c0;
}

// One turn of the rotation: segments at R = 0, SL-1, ..., 1 (a shift takes R
// to R - 1), stopping at the first that finds the stream ended.
template <int SL, typename F, int... I>
__device__ __forceinline__ bool run_rotations(F& f, std::integer_sequence<int, I...>) {
  return (f(std::integral_constant<int, (SL - I) % SL>{}) && ...);
}

// One atom group's walk of its event stream. `group` is the atom this lane's
// group stands for (lanes past the last atom run clamped and write nothing).
// Kernel variant bits (the fourth template argument of the real kernels).
constexpr int V_DIFF = 1;    // diffusion damps each order by its own factor
constexpr int V_B1MAP = 2;   // a per-atom B1 scales every flip
constexpr int V_ACCUM = 4;   // accumulate normal equations instead of writing outputs
constexpr int V_PADDED = 8;  // the tile covers more orders than the state count
constexpr int V_PROJECT = 16;  // contract the samples with a temporal basis instead of writing them

// Normal equations carry at most (T1, T2, B1) and no diffusion; so does a
// projection, which never accumulates normal equations.
template <int TM, int V>
constexpr bool valid_variant() {
  if ((V & V_ACCUM) && ((V & V_DIFF) || TM == 15)) return false;
  if ((V & V_PROJECT) && ((V & V_ACCUM) || (V & V_DIFF) || TM == 15)) return false;
  return true;
}

template <int SL, int L, int TM, int V>
__device__ __forceinline__ void real_body(const RealArgs& p, int64_t group) {
  constexpr bool DIFF = (V & V_DIFF) != 0;
  constexpr bool B1MAP = (V & V_B1MAP) != 0;
  constexpr bool ACCUM = (V & V_ACCUM) != 0;
  constexpr bool PADDED = (V & V_PADDED) != 0;
  constexpr bool PROJECT = (V & V_PROJECT) != 0;
  constexpr int NT = popcount4(TM);
  constexpr int NTA = NT > 0 ? NT : 1;
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
  constexpr int SB1 = slot_of(TM, TAN_B1);
  constexpr int SIE = slot_of(TM, TAN_IE);
  constexpr bool ROTATE = SL <= ROTATE_MAX;

  const int sub = (int)(threadIdx.x % L);
  const bool first = sub == 0;
  const bool last = sub == L - 1;
  const bool live = group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;

  // The first logical order of this lane past the state count (SL if none):
  // after a shift F+ carries a value into it, which the truncation drops.
  const int base = sub * SL;
  const int pad = (p.states - base >= 0 && p.states - base < SL) ? (int)(p.states - base) : SL;

  const float t1 = p.t1[atom * p.stride_t1];
  const float t2 = p.t2[atom * p.stride_t2];
  const float r1 = 1000.0f / t1;
  const float r2 = 1000.0f / t2;
  const float m0 = p.density ? p.m0[atom * p.stride_opt] : 1.0f;
  const float b1 = B1MAP ? p.b1[atom * p.stride_opt] : 1.0f;
  const float eta = p.inverting ? p.inv_eff[atom * p.stride_opt] : 1.0f;
  const float drate = DIFF ? p.diffusion[atom * p.stride_opt] : 0.0f;
  // d r1 / d T1 = -1000 / T1^2, which the tangent of e1 = exp(-r1 dt) carries.
  const float k1 = 1000.0f / (t1 * t1);
  const float k2 = 1000.0f / (t2 * t2);

  float a[SL], b[SL], c[SL];
  float ta[NTA][SL], tb[NTA][SL], tc[NTA][SL];
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    a[r] = 0.0f;
    b[r] = 0.0f;
    c[r] = 0.0f;
#pragma unroll
    for (int k = 0; k < NTA; ++k) {
      ta[k][r] = 0.0f;
      tb[k][r] = 0.0f;
      tc[k][r] = 0.0f;
    }
  }
  c[0] = first ? 1.0f : 0.0f;

  double acc_cost = 0.0;
  // A fit names at most one column per carried tangent plus M0, so a kernel
  // with NT tangents accumulates 1 + NT columns (the output keeps the
  // NORMAL_COLUMNS layout, zero beyond).
  constexpr int NCOL = ACCUM ? (1 + NT < NORMAL_COLUMNS ? 1 + NT : NORMAL_COLUMNS) : 1;
  double acc_g[NCOL];
  double acc_h[NCOL * (NCOL + 1) / 2];
#pragma unroll
  for (int i = 0; i < NCOL; ++i) acc_g[i] = 0.0;
#pragma unroll
  for (int i = 0; i < NCOL * (NCOL + 1) / 2; ++i) acc_h[i] = 0.0;

  // Projection accumulators: stream s (0 the unscaled sample a0, 1 + k the
  // carried tangent k) against direction sub + L j, summed over the samples
  // as sum_n sign_n x_n basis[n][r]; and this op's basis values.
  constexpr int PJ = PROJECT ? PROJECT_PER_LANE : 1;
  constexpr int NPS = PROJECT ? 1 + NT : 1;
  float2 proj[NPS][PJ];
  float2 pbasis[PJ];
#pragma unroll
  for (int j = 0; j < PJ; ++j) {
    pbasis[j] = make_float2(0.0f, 0.0f);
#pragma unroll
    for (int q = 0; q < NPS; ++q) proj[q][j] = make_float2(0.0f, 0.0f);
  }

  uint32_t last_dt = 0xFFFFFFFFu;  // bit pattern of the cached interval (a NaN: never matches)
  float e1 = 1.0f, e2 = 1.0f, de1 = 0.0f, de2 = 0.0f;
  // diffusion damping per logical order of this lane's block
  float dl[DIFF ? SL : 1], dtr[DIFF ? SL : 1];
#pragma unroll
  for (int r = 0; r < (DIFF ? SL : 1); ++r) {
    dl[r] = 1.0f;
    dtr[r] = 1.0f;
  }

  const MacroOp* __restrict__ ops = p.ops;

  // Decay factors over dt, recomputed only when dt changes: e1 = exp(-dt/T1),
  // e2 = exp(-dt/T2), their T1/T2 derivatives, k2 dt = de2/e2, and the
  // recovery 1 - e1 on the first lane. The recovery comes from expm1: 1 - e1
  // formed from a rounded e1 loses about five digits when dt << T1, and that
  // error recurs identically on every pulse of a steady train.
  float k2dt = 0.0f, rec = 0.0f, e1m = 0.0f;
  auto decay = [&](float dt) __attribute__((always_inline)) {
    // dt is wave-uniform: compare bit patterns on the scalar unit
    if (__float_as_uint(dt) == last_dt) return;
    last_dt = __float_as_uint(dt);
    e1m = expm1f(-r1 * dt);
    e1 = 1.0f + e1m;
    e2 = expf(-r2 * dt);
    k2dt = k2 * dt;
    de1 = e1 * (k1 * dt);
    de2 = e2 * k2dt;
    rec = first ? -e1m : 0.0f;
    if constexpr (DIFF) {
      const float bf = drate * dt;
#pragma unroll
      for (int r = 0; r < SL; ++r) {
        const float l = (float)(base + r);
        dl[r] = expf(-(l * l) * bf);
        dtr[r] = expf(-(l * l + l + (1.0f / 3.0f)) * bf);
      }
    }
  };

  // Relaxation and recovery over dt (reference relax_real,
  // _epg_cpu.cpp:2930-2975), register map of rotation R.
  auto relax = [&](auto rc, float dt) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    decay(dt);
    // Every update below reads a register's old value before it overwrites
    // that register, so the allocator can keep each plane where it lives.
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int ia = Rot<SL, R>::a(j), ib = Rot<SL, R>::b(j);
      const float tra = DIFF ? e2 * dtr[j] : e2;
      const float lo = DIFF ? e1 * dl[j] : e1;
      const float dlo = DIFF ? de1 * dl[j] : de1;
      const float dtra = DIFF ? de2 * dtr[j] : de2;
      if constexpr ((TM & TAN_T1) != 0) {
        ta[ST1][ia] *= tra;
        tb[ST1][ib] *= tra;
        tc[ST1][j] = fmaf(c[j], dlo, tc[ST1][j] * lo);
      }
      if constexpr ((TM & TAN_T2) != 0) {
        ta[ST2][ia] = fmaf(a[ia], dtra, ta[ST2][ia] * tra);
        tb[ST2][ib] = fmaf(b[ib], dtra, tb[ST2][ib] * tra);
        tc[ST2][j] *= lo;
      }
      if constexpr ((TM & TAN_B1) != 0) {
        ta[SB1][ia] *= tra;
        tb[SB1][ib] *= tra;
        tc[SB1][j] *= lo;
      }
      if constexpr ((TM & TAN_IE) != 0) {
        ta[SIE][ia] *= tra;
        tb[SIE][ib] *= tra;
        tc[SIE][j] *= lo;
      }
      a[ia] *= tra;
      b[ib] *= tra;
      c[j] *= lo;
    }
    // Recovery toward equilibrium lands on order zero only.
    c[0] += rec;
    if constexpr ((TM & TAN_T1) != 0) tc[ST1][0] -= first ? de1 : 0.0f;
  };

  // 1 - cos and sin of the op's flip as this atom plays it; with a B1 map
  // both from the half angle, so the versine keeps its relative accuracy
  // when the flip is small.
  auto angle = [&](const MacroOp& ev, float& vs, float& sn) __attribute__((always_inline)) {
    if constexpr (B1MAP) {
      float sh, ch;
      sincos_reduced(0.5f * (ev.flip * b1), sh, ch);
      const float s2 = sh + sh;
      vs = s2 * sh;
      sn = s2 * ch;
    } else {
      vs = ev.versine;
      sn = ev.sine;
    }
  };

  // Everything an op does to the planes before its record: a relaxation over
  // dt (kernels without diffusion), then a rotation about x (reference
  // rf_real, _epg_cpu.cpp:2980-3005), an inversion Z -> -eta Z, or a spoil
  // F+ = F- = 0. Z relaxes first, in place, as an increment:
  // Z' = Z + (e1 - 1)(Z - Z_eq) with e1 - 1 from expm1 (Z_eq = 1 at order
  // 0 of the first lane, else 0). A rounded e1 is off by up to 6e-8, which
  // against 1 - e1 ~ TR/T1 ~ 1e-3 is a 1e-4 relative error in the recovery
  // per pulse, and a train with a long memory (a balanced one) sums it over
  // hundreds of pulses; the increment keeps e1 - 1 to its own rounding.
  // The transverse decay folds into the rotation, written with d = p - m as
  //   u = A d + B z,  p' = E p - u,  m' = E m + u,  z' = C d + z - W z,
  // where a rotation by a has A = e2 (1 - cos)/2, B = sin, C = e2 sin/2,
  // W = 1 - cos (the versine, from the half angle, for the same reason),
  // E = e2; an inversion A = B = C = 0, W = 1 + eta; a spoil E = A = B = C =
  // W = 0; a bare relaxation A = B = C = W = 0 (so p' = e2 p, exactly
  // rounded). The tangent sets take the same coefficients with D = 1 - W.
  // The T2 derivative of the transverse decay is applied
  // before the scale: a k2 dt and b k2 dt on the T2 tangents of F+ and F-
  // (E multiplies them by e2). Along B1 the rotation angle moves by the flip
  // per unit B1, which adds bA d + bB z to u and bC d + bD z to z' of the B1
  // set; along the inversion efficiency z' of that set gains -z. Each plane
  // register is read before it is written, so every plane is updated where
  // it lives. With small flips this form keeps p' = p - u close to p, where
  // a sum/difference form would round p and m together.
  auto apply = [&](auto rc, const MacroOp& m, uint32_t ctl, float vs, float sn) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    const bool relaxing = !DIFF && (ctl & OP_RELAX);
    if (relaxing) decay(m.dt);
    const float f2 = relaxing ? e2 : 1.0f;
    const float g1 = relaxing ? de1 : 0.0f;
    const float q2 = relaxing ? k2dt : 0.0f;
    const float h1 = relaxing ? e1m : 0.0f;
    const float dal = (ctl & OP_ROTATE) ? m.flip : 0.0f;
    const bool spoil = (ctl & OP_SPOIL) != 0;
    const bool invert = (ctl & OP_INVERT) != 0;
    const float cs = 1.0f - vs;
    const float E = spoil ? 0.0f : f2;
    const float A = f2 * (0.5f * vs);
    const float B = sn;
    const float C = f2 * (0.5f * sn);
    const float D = invert ? -eta : cs;
    const float W = invert ? 1.0f + eta : vs;  // 1 - D
    const float bA = f2 * (0.5f * sn * dal), bB = cs * dal, bC = f2 * (0.5f * cs * dal), bD = -sn * dal;

    // longitudinal relaxation and recovery, and the T1/T2 decay derivatives
#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int ia = Rot<SL, R>::a(j), ib = Rot<SL, R>::b(j);
      if constexpr ((TM & TAN_T2) != 0) {
        ta[ST2][ia] = fmaf(a[ia], q2, ta[ST2][ia]);
        tb[ST2][ib] = fmaf(b[ib], q2, tb[ST2][ib]);
      }
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        if ((TM & TAN_T1) != 0 && k == ST1) {
          tc[k][j] = fmaf(c[j], g1, fmaf(tc[k][j], h1, tc[k][j]));
        } else {
          tc[k][j] = fmaf(tc[k][j], h1, tc[k][j]);
        }
      }
      c[j] = fmaf(j == 0 && first ? c[j] - 1.0f : c[j], h1, c[j]);
    }
    if constexpr ((TM & TAN_T1) != 0) tc[ST1][0] -= first ? g1 : 0.0f;

#pragma unroll
    for (int j = 0; j < SL; ++j) {
      const int ia = Rot<SL, R>::a(j), ib = Rot<SL, R>::b(j);
      const float pv = a[ia], mv = b[ib], zv = c[j];
      float ub = 0.0f, zb = 0.0f;
      if constexpr ((TM & TAN_B1) != 0) {
        const float dv = pv - mv;
        ub = fmaf(bA, dv, bB * zv);
        zb = fmaf(bC, dv, bD * zv);
      }
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        const float dp = ta[k][ia], dm = tb[k][ib], dz = tc[k][j];
        const float d = dp - dm;
        float u = fmaf(A, d, B * dz);
        float z = fmaf(C, d, D * dz);
        if ((TM & TAN_B1) != 0 && k == SB1) {
          u += ub;
          z += zb;
        }
        if ((TM & TAN_IE) != 0 && k == SIE) z -= invert ? zv : 0.0f;
        // m' before p': u is still live at m', so neither update can take
        // u's register as its destination
        tb[k][ib] = scale_add_in_place(dm, E, u);
        ta[k][ia] = scale_sub_in_place(dp, E, u);
        tc[k][j] = z;
      }
      const float d = pv - mv;
      const float u = fmaf(A, d, B * zv);
      c[j] = fmaf(C, d, fmaf(-W, zv, zv));
      b[ib] = scale_add_in_place(mv, E, u);
      a[ia] = scale_sub_in_place(pv, E, u);
    }
  };

  // Output staging (see STAGE_BYTES_PER_WAVE). `nrec` counts the records so
  // far; every stage_k of them the wave writes its stage out.
  extern __shared__ float2 stage_memory[];
  // Staged writes (outputs) or, in normal-equation mode, staged reads of the
  // measured samples: the same per-atom-row traffic, the other direction.
  const bool staged = p.stage_k_log2 > 0 && (load_op(ops, 0).ctrl & OPS_SEQUENTIAL) != 0;
  const int stage_k = 1 << (p.stage_k_log2 > 0 ? p.stage_k_log2 : 0);
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  float2* const stage =
      stage_memory + (((threadIdx.x >> 6) * (64 / L)) << ((p.stage_s_log2 > 0 ? p.stage_s_log2 : 0) +
                                                         (p.stage_k_log2 > 0 ? p.stage_k_log2 : 0)));
  int64_t nrec = 0;

  // Write the first `count` slots of every row of this wave's stage: row r
  // holds stream (r mod S) of the wave's group (r / S), samples
  // nrec - count ... nrec - 1. Two samples per lane per store, K samples of
  // a row contiguous.
  auto flush = [&](int count) __attribute__((always_inline)) {
    const int64_t g0 = group - wave_group;
    const int64_t o0 = nrec - count;
    const int rows_log2 = p.stage_s_log2 + p.stage_k_log2;
    const int total = (64 / L) << rows_log2;
    for (int i = 2 * lane; i < total; i += 128) {
      const int kk = i & (stage_k - 1);
      if (kk >= count) continue;
      const int row = i >> p.stage_k_log2;
      const int64_t at = g0 + (row >> p.stage_s_log2);
      if (at >= p.atoms) continue;
      const int stream = row & ((1 << p.stage_s_log2) - 1);
      const int64_t o = o0 + kk;
      float2* dst;
      if (stream == 0) {
        if (!p.signal) continue;
        dst = p.signal + at * p.outputs + o;
      } else {
        if (stream > p.columns) continue;
        dst = p.jacobian + (at * p.columns + (stream - 1)) * p.outputs + o;
      }
      const float2 lo = stage[i];
      if (kk + 1 < count) {
        const float2 hi = stage[i + 1];
        dst[0] = lo;
        dst[1] = hi;
      } else {
        dst[0] = lo;
      }
    }
  };

  // Normal-equation mode: load samples nrec ... nrec + K - 1 of every atom
  // of the wave into its stage (one row per atom, K consecutive samples read
  // together).
  auto fill = [&]() __attribute__((always_inline)) {
    const int64_t g0 = group - wave_group;
    const int total = (64 / L) << p.stage_k_log2;
    for (int i = lane; i < total; i += 64) {
      const int kk = i & (stage_k - 1);
      const int64_t at = g0 + (i >> p.stage_k_log2);
      const int64_t o = nrec + kk;
      if (at < p.atoms && o < p.outputs) stage[i] = p.data[at * p.outputs + o];
    }
  };

  // The sample i m0 F+[0] and its tangents (reference record,
  // _epg_cpu.cpp:3090-3120); F+[0] is register R of the first lane.
  auto record = [&](auto rc, const MacroOp& ev) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    const float a0 = a[R];
    float t0[NTA];
#pragma unroll
    for (int k = 0; k < NTA; ++k) t0[k] = NT > 0 ? ta[k][R] : 0.0f;
    if constexpr (PROJECT) {
      // every lane of the atom takes the first lane's sample and tangents
      // and adds them along its own basis directions (loaded at op start)
      const float sign = p.sample_sign ? p.sample_sign[ev.out] : 1.0f;
      float x[NPS];
      x[0] = atom_broadcast<L>(a0) * sign;
#pragma unroll
      for (int k = 0; k < NPS - 1; ++k) x[1 + k] = atom_broadcast<L>(t0[k]) * sign;
#pragma unroll
      for (int q = 0; q < NPS; ++q) {
#pragma unroll
        for (int j = 0; j < PJ; ++j) {
          proj[q][j].x = fmaf(x[q], pbasis[j].x, proj[q][j].x);
          proj[q][j].y = fmaf(x[q], pbasis[j].y, proj[q][j].y);
        }
      }
      return;
    }
    if (!(first && live)) return;
    const int64_t o = ev.out;
    if constexpr (ACCUM) {
      const float sign = p.sample_sign ? p.sample_sign[o] : 1.0f;
      const float fr = p.factor_re * sign, fi = p.factor_im * sign;
      const float v = m0 * a0;
      const float2 y = staged ? stage[(wave_group << p.stage_k_log2) + (nrec & (stage_k - 1))]
                              : p.data[atom * p.outputs + o];
      const double rx = (double)(-fi * v) - (double)y.x;
      const double ry = (double)(fr * v) - (double)y.y;
      acc_cost += rx * rx + ry * ry;
      // J_c = (-fi d_c, fr d_c): J_c . r = d_c q, J_c . J_k = |factor|^2 d_c d_k
      const double q = -(double)fi * rx + (double)fr * ry;
      const double ff = (double)fr * fr + (double)fi * fi;
      double d[NCOL];
#pragma unroll
      for (int cc = 0; cc < NCOL; ++cc) {
        float x = 0.0f;
        if constexpr ((TM & TAN_T1) != 0) x = (p.col_t1 == cc) ? m0 * t0[ST1] : x;
        if constexpr ((TM & TAN_T2) != 0) x = (p.col_t2 == cc) ? m0 * t0[ST2] : x;
        if constexpr ((TM & TAN_B1) != 0) x = (p.col_b1 == cc) ? m0 * t0[SB1] : x;
        x = (p.col_m0 == cc) ? a0 : x;
        d[cc] = (double)x;
      }
      int h = 0;
#pragma unroll
      for (int cc = 0; cc < NCOL; ++cc) {
        acc_g[cc] += d[cc] * q;
#pragma unroll
        for (int kk = cc; kk < NCOL; ++kk) acc_h[h++] += ff * d[cc] * d[kk];
      }
    } else if (staged) {
      // this group's row of the wave's stage: stream s at slot s * K
      const float v = m0 * a0;
      const float sign = p.sample_sign ? p.sample_sign[o] : 1.0f;
      const float fr = p.factor_re * sign, fi = p.factor_im * sign;
      float2* row = stage + ((wave_group << p.stage_s_log2) << p.stage_k_log2) + (nrec & (stage_k - 1));
      row[0] = make_float2(-fi * v, fr * v);
      if constexpr ((TM & TAN_T1) != 0) {
        if (p.col_t1 >= 0) {
          const float d = m0 * t0[ST1];
          row[(1 + p.col_t1) << p.stage_k_log2] = make_float2(-fi * d, fr * d);
        }
      }
      if constexpr ((TM & TAN_T2) != 0) {
        if (p.col_t2 >= 0) {
          const float d = m0 * t0[ST2];
          row[(1 + p.col_t2) << p.stage_k_log2] = make_float2(-fi * d, fr * d);
        }
      }
      if constexpr ((TM & TAN_B1) != 0) {
        if (p.col_b1 >= 0) {
          const float d = m0 * t0[SB1];
          row[(1 + p.col_b1) << p.stage_k_log2] = make_float2(-fi * d, fr * d);
        }
      }
      if constexpr ((TM & TAN_IE) != 0) {
        if (p.col_ie >= 0) {
          const float d = m0 * t0[SIE];
          row[(1 + p.col_ie) << p.stage_k_log2] = make_float2(-fi * d, fr * d);
        }
      }
      if (p.col_m0 >= 0) row[(1 + p.col_m0) << p.stage_k_log2] = make_float2(-fi * a0, fr * a0);
    } else {
      const float v = m0 * a0;
      if (p.out_real) {
        p.out_real[atom * p.outputs + o] = 0.0f;
        p.out_imag[atom * p.outputs + o] = v;
      }
      const float sign = p.sample_sign ? p.sample_sign[o] : 1.0f;
      const float fr = p.factor_re * sign, fi = p.factor_im * sign;
      // factor * (0 + i v) = (-fi v, fr v)
      if (p.signal) p.signal[atom * p.outputs + o] = make_float2(-fi * v, fr * v);
      if (p.jacobian) {
        float2* row = p.jacobian + atom * (int64_t)p.columns * p.outputs + o;
        if constexpr ((TM & TAN_T1) != 0) {
          if (p.col_t1 >= 0) {
            const float d = m0 * t0[ST1];
            row[(int64_t)p.col_t1 * p.outputs] = make_float2(-fi * d, fr * d);
          }
        }
        if constexpr ((TM & TAN_T2) != 0) {
          if (p.col_t2 >= 0) {
            const float d = m0 * t0[ST2];
            row[(int64_t)p.col_t2 * p.outputs] = make_float2(-fi * d, fr * d);
          }
        }
        if constexpr ((TM & TAN_B1) != 0) {
          if (p.col_b1 >= 0) {
            const float d = m0 * t0[SB1];
            row[(int64_t)p.col_b1 * p.outputs] = make_float2(-fi * d, fr * d);
          }
        }
        if constexpr ((TM & TAN_IE) != 0) {
          if (p.col_ie >= 0) {
            const float d = m0 * t0[SIE];
            row[(int64_t)p.col_ie * p.outputs] = make_float2(-fi * d, fr * d);
          }
        }
        if (p.col_m0 >= 0) row[(int64_t)p.col_m0 * p.outputs] = make_float2(-fi * a0, fr * a0);
      }
    }
  };

  // The shift that ends a segment at rotation R.
  auto shift = [&](auto rc) __attribute__((always_inline)) {
    constexpr int R = decltype(rc)::value;
    if constexpr (ROTATE) {
      constexpr int RN = (R + SL - 1) % SL;
      shift_rotating<SL, L, R>(a, b, first, last);
#pragma unroll
      for (int k = 0; k < NT; ++k) shift_rotating<SL, L, R>(ta[k], tb[k], first, last);
      if constexpr (PADDED) {
        clear_order<SL, RN>(a, pad);
#pragma unroll
        for (int k = 0; k < NT; ++k) clear_order<SL, RN>(ta[k], pad);
      }
    } else {
      shift_pair<SL, L>(a, b, first, last);
#pragma unroll
      for (int k = 0; k < NT; ++k) shift_pair<SL, L>(ta[k], tb[k], first, last);
      if constexpr (PADDED) {
        clear_order<SL, 0>(a, pad);
#pragma unroll
        for (int k = 0; k < NT; ++k) clear_order<SL, 0>(ta[k], pad);
      }
    }
  };

  // Ops up to and including the next shift, at rotation R; false once the
  // end marker is reached. The loop has one exit (a flag tells the end
  // marker from a shift): with two, the compiler re-homes every plane
  // register at each segment boundary. It runs one op ahead: the next op's
  // scalar load and its cos/sin are issued before this op's plane work, so
  // their latency hides behind it.
  MacroOp next = load_op(ops, 0);
  int64_t e = 1;
  float next_vs, next_sn;
  angle(next, next_vs, next_sn);
  auto segment = [&](auto rc) __attribute__((always_inline)) -> bool {
    bool ended = false;
    for (;;) {
      const MacroOp m = next;
      const float vs = next_vs, sn = next_sn;
      const uint32_t ctl = m.ctrl;
      if (ctl & OP_END) {
        ended = true;
        break;
      }
      next = load_op(ops, e);
      ++e;
      if constexpr (PROJECT) {
        // this op's basis row, issued before the plane work hides its latency
        if ((ctl & OP_RECORD) && m.out >= 0 && m.out < p.outputs) {
          const float2* row = p.basis + (int64_t)m.out * p.rank;
#pragma unroll
          for (int j = 0; j < PJ; ++j) {
            const int r = sub + L * j;
            pbasis[j] = r < p.rank ? row[r] : make_float2(0.0f, 0.0f);
          }
        }
      }
      if constexpr (DIFF) {
        if (ctl & OP_RELAX) relax(rc, m.dt);
      }
      // every op runs the plane transform (the identity where it has nothing
      // to apply), and the next op's angle in the same straight-line block;
      // an op that does not rotate has flip 0, versine 0, sin 0
      apply(rc, m, ctl, vs, sn);
      angle(next, next_vs, next_sn);
      if (ctl & OP_RECORD) {
        if (m.out >= 0 && m.out < p.outputs) {
          if constexpr (ACCUM) {
            if (staged && (nrec & (stage_k - 1)) == 0) fill();
          }
          record(rc, m);
          if (!PROJECT && staged) {
            ++nrec;
            if constexpr (!ACCUM) {
              if ((nrec & (stage_k - 1)) == 0) flush(stage_k);
            }
          }
        }
      }
      if (ctl & OP_SHIFT) break;
    }
    if (ended) return false;
    shift(rc);
    return true;
  };

  if constexpr (ROTATE) {
    while (run_rotations<SL>(segment, std::make_integer_sequence<int, SL>{})) {
    }
  } else {
    while (segment(std::integral_constant<int, 0>{})) {
    }
  }
  if constexpr (!ACCUM && !PROJECT) {
    if (staged && (nrec & (stage_k - 1)) != 0) flush((int)(nrec & (stage_k - 1)));
  }

  if constexpr (PROJECT) {
    // coefficient = factor * i * (m0 or 1) * accumulator: the projection of
    // the samples factor * sign * (0 + i m0 a0) and of their tangents
    if (live) {
      const float gr = -p.factor_im, gi = p.factor_re;
      auto scaled = [&](float2 v, float k) __attribute__((always_inline)) {
        return make_float2(k * (gr * v.x - gi * v.y), k * (gr * v.y + gi * v.x));
      };
#pragma unroll
      for (int j = 0; j < PJ; ++j) {
        const int r = sub + L * j;
        if (r >= p.rank) continue;
        if (p.signal) p.signal[atom * p.rank + r] = scaled(proj[0][j], m0);
        if (p.jacobian) {
          float2* row = p.jacobian + atom * (int64_t)p.columns * p.rank + r;
          if constexpr ((TM & TAN_T1) != 0) {
            if (p.col_t1 >= 0) row[(int64_t)p.col_t1 * p.rank] = scaled(proj[1 + ST1][j], m0);
          }
          if constexpr ((TM & TAN_T2) != 0) {
            if (p.col_t2 >= 0) row[(int64_t)p.col_t2 * p.rank] = scaled(proj[1 + ST2][j], m0);
          }
          if constexpr ((TM & TAN_B1) != 0) {
            if (p.col_b1 >= 0) row[(int64_t)p.col_b1 * p.rank] = scaled(proj[1 + SB1][j], m0);
          }
          if (p.col_m0 >= 0) row[(int64_t)p.col_m0 * p.rank] = scaled(proj[0][j], 1.0f);
        }
      }
    }
  }

  if constexpr (ACCUM) {
    if (first && live) {
      double* out = p.normal + atom * NORMAL_STRIDE;
      out[0] = acc_cost;
#pragma unroll
      for (int cc = 0; cc < NORMAL_COLUMNS; ++cc) out[1 + cc] = cc < NCOL ? acc_g[cc] : 0.0;
      int hn = 0, h4 = 0;
#pragma unroll
      for (int cc = 0; cc < NORMAL_COLUMNS; ++cc) {
#pragma unroll
        for (int kk = cc; kk < NORMAL_COLUMNS; ++kk, ++h4) {
          const bool held = cc < NCOL && kk < NCOL;
          out[1 + NORMAL_COLUMNS + h4] = held ? acc_h[held ? hn : 0] : 0.0;
          hn += held ? 1 : 0;
        }
      }
    }
  }
}

template <int SL, int L, int TM, int V>
__global__ void __launch_bounds__(256) real_kernel(const RealArgs p) {
  const int64_t group = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
  real_body<SL, L, TM, V>(p, group);
}

// The uneven-batch launch: a table of requests, each with its own event
// stream, tissue, outputs and state count, and a table of work blocks
// (request, first atom). A persistent grid walks the work blocks; every block
// runs exactly the arithmetic a launch of its request alone would, so a
// result is bitwise the same alone and in the mix.
struct BatchWork {
  int32_t request;
  int32_t pad;
  int64_t first_atom;
};

template <int SL, int L, int TM, int V>
__global__ void __launch_bounds__(256) real_batch_kernel(const RealArgs* __restrict__ requests,
                                                        const BatchWork* __restrict__ work, int64_t blocks) {
  for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
    const BatchWork w = work[b];
    const RealArgs p = requests[w.request];
    const int64_t group = w.first_atom + (int64_t)threadIdx.x / L;
    real_body<SL, L, TM, V>(p, group);
  }
}

// Launchers take the dynamic LDS bytes of the output stage.
using RealLaunch = void (*)(const RealArgs&, dim3, dim3, size_t, hipStream_t);
using BatchLaunch = void (*)(const RealArgs*, const BatchWork*, int64_t, dim3, dim3, size_t, hipStream_t);

template <int SL, int L, int TM, int V>
void launch_real(const RealArgs& args, dim3 grid, dim3 block, size_t shared, hipStream_t stream) {
  hipLaunchKernelGGL((real_kernel<SL, L, TM, V>), grid, block, shared, stream, args);
}

template <int SL, int L, int TM, int V>
void launch_batch(const RealArgs* r, const BatchWork* w, int64_t blocks, dim3 grid, dim3 block, size_t shared,
                  hipStream_t s) {
  hipLaunchKernelGGL((real_batch_kernel<SL, L, TM, V>), grid, block, shared, s, r, w, blocks);
}

// Tangent sets come in four kernels: none, (T1, T2), (T1, T2, B1) and all
// four. A request for a subset runs the smallest superset; the extra
// directions are carried and never written (their column index is -1).
__host__ __device__ constexpr int tangent_class(int tm) {
  return tm == 0 ? 0 : (tm & ~3) == 0 ? 3 : (tm & ~7) == 0 ? 7 : 15;
}

// The launcher of variant v (a run-time value) among the 16 compiled ones;
// null for the combinations no kernel carries.
template <int SL, int L, int TM, int V = 0>
RealLaunch pick_real_variant(int v) {
  if constexpr (V < 32) {
    if (v == V) {
      if constexpr (valid_variant<TM, V>()) return &launch_real<SL, L, TM, V>;
      return nullptr;
    }
    return pick_real_variant<SL, L, TM, V + 1>(v);
  } else {
    return nullptr;
  }
}

template <int SL, int L, int TM, int V = 0>
BatchLaunch pick_batch_variant(int v) {
  if constexpr (V < 16) {
    if (v == V) {
      if constexpr (valid_variant<TM, V>()) return &launch_batch<SL, L, TM, V>;
      return nullptr;
    }
    return pick_batch_variant<SL, L, TM, V + 1>(v);
  } else {
    return nullptr;
  }
}

template <int SL, int L>
RealLaunch pick_real_kernel(int tm, int v) {
  switch (tangent_class(tm)) {
    case 0: return pick_real_variant<SL, L, 0>(v);
    case 3: return pick_real_variant<SL, L, 3>(v);
    case 7: return pick_real_variant<SL, L, 7>(v);
    default: return pick_real_variant<SL, L, 15>(v);
  }
}

template <int SL, int L>
BatchLaunch pick_batch_kernel(int tm, int v) {
  switch (tangent_class(tm)) {
    case 0: return pick_batch_variant<SL, L, 0>(v);
    case 3: return pick_batch_variant<SL, L, 3>(v);
    case 7: return pick_batch_variant<SL, L, 7>(v);
    default: return pick_batch_variant<SL, L, 15>(v);
  }
}

// One entry per state tile, defined by epg_real_tile.hip built once per tile.
#define TSIM_TILES(X) \
  X(0, 4, 1)          \
  X(1, 8, 1)          \
  X(2, 10, 1)         \
  X(3, 5, 2)          \
  X(4, 6, 2)          \
  X(5, 8, 2)          \
  X(6, 10, 2)         \
  X(7, 12, 2)         \
  X(8, 5, 4)          \
  X(9, 6, 4)          \
  X(10, 8, 4)         \
  X(11, 10, 4)        \
  X(12, 12, 4)        \
  X(13, 16, 4)

#define TSIM_DECLARE_TILE(ID, SL, L)                \
  RealLaunch pick_real_tile_##ID(int tm, int variant); \
  BatchLaunch pick_batch_tile_##ID(int tm, int variant);
TSIM_TILES(TSIM_DECLARE_TILE)
#undef TSIM_DECLARE_TILE

}  // namespace tsim
