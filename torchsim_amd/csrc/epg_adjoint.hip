// Checkpointed adjoint of the fused real-subspace EPG system, for gfx950.
//
// Gradients of a scalar loss with respect to every event's flip angle and
// duration and to each atom's T1, T2, M0, B1 and inversion efficiency, given
// the loss's cotangents on the recorded signal and on the Jacobian columns
// along T1/T2/B1/M0. It computes the reference's reverse passes: first order
// (src/torchsim/_epg_cpu.cpp:5985-6227 simulate_real_vjp_range) when no
// tangent column carries a cotangent, forward-over-reverse
// (:7028-7314 simulate_real_vjp_jvp_range) otherwise -- the path
// `crlb(J).backward()` takes in sequence design -- with the per-atom tissue
// reduction of :11366-11395.
//
// The kernel walks the same macro-op stream as the forward kernel (one op
// per MRF pulse: relaxation, rotation, record, shift), not the packed
// events. Each atom keeps its dual state (the value planes and one tangent
// set per carried column) in registers, S orders split into L blocks of SL
// over L adjacent lanes as in the forward kernel. Pass 1 runs the train
// forward and stores the dual state every `segment` ops (a checkpoint); pass
// 2 takes the segments last to first, recomputes one forward from its
// checkpoint storing each op's entry state to scratch ([op][float][lane],
// coalesced), and walks the segment's ops in reverse, loading the next
// state one op ahead. Per-op gradient contributions of the lanes go to LDS
// and are summed FLUSH ops at a time (one LDS transpose instead of two wave
// reductions per op); per-wave partials are reduced over waves in a fixed
// order by a second launch, so the result does not depend on scheduling.
#include "epg_real_impl.h"

#include <cstring>

namespace tsim {
namespace {

constexpr int FLUSH = 16;  // ops whose per-lane gradient partials LDS holds before a flush
constexpr int G_T1 = 0, G_T2 = 1, G_M0 = 2, G_B1 = 3, G_IE = 4, G_COUNT = 5;

struct AdjointArgs {
  const MacroOp* ops;
  const int64_t* op_count;  // device: ops in the stream (the kernel clamps it to op_bound)
  int64_t op_bound;         // ops the checkpoints and partials are sized for
  const float* t1;
  const float* t2;
  const float* m0;
  const float* b1;
  const float* inv_eff;
  int64_t stride_opt;
  int64_t atoms;
  int64_t outputs;
  int states;
  int density;
  int inverting;
  int transmit;
  const float* w_signal;  // (atoms, outputs) cotangent of the recorded m0*a0, or null
  const float* w_jac;     // (atoms, columns, outputs) cotangent of the columns, or null
  int columns;
  int col_t1, col_t2, col_b1, col_m0;
  int64_t segment;        // ops per checkpoint
  int64_t threads_total;
  float* checkpoints;     // (segments, floats, threads_total)
  float* scratch;         // (segment, floats, threads_total)
  float* partial;         // (waves, op_bound, 2): d/dflip, d/dduration
  float* grad[G_COUNT];   // per-atom tissue gradients, or null
};

// Transpose of shift_pair: the adjoint of F+ moves DOWN an order, of F- UP;
// the coupling F+[0] = -F-[0] (after the shift) sends -lambda_a'[0] to
// lambda_b[1] -- the first lane's second register, or with one order per
// lane the second lane's first.
template <int SL, int L>
__device__ __forceinline__ void shift_pair_t(float (&la)[SL], float (&lb)[SL], bool first, bool last, int sub) {
  const float a0 = la[0];
  float a_in = 0.0f;  // lambda_a'[0] of the next lane -> lambda_a[SL-1]
  float b_in = 0.0f;  // lambda_b'[SL-1] of the previous lane -> lambda_b[0]
  float a0_lower = 0.0f;
  if constexpr (L > 1) {
    a_in = dpp_from_upper(la[0]);
    a_in = last ? 0.0f : a_in;
    b_in = dpp_from_lower(lb[SL - 1]);
    b_in = first ? 0.0f : b_in;
    if constexpr (SL == 1) a0_lower = dpp_from_lower(a0);
  }
#pragma unroll
  for (int r = 0; r + 1 < SL; ++r) la[r] = la[r + 1];
  la[SL - 1] = a_in;
#pragma unroll
  for (int r = SL - 1; r > 0; --r) lb[r] = lb[r - 1];
  lb[0] = b_in;
  if constexpr (SL >= 2) {
    if (first) lb[1] -= a0;
  } else {
    if (sub == 1) lb[0] -= a0_lower;
  }
}

template <int SL, int TM>
struct Dual {
  static constexpr int NT = popcount4(TM);
  static constexpr int NTA = NT > 0 ? NT : 1;
  float a[SL], b[SL], c[SL];
  float ta[NTA][SL], tb[NTA][SL], tc[NTA][SL];
  static constexpr int floats = 3 * SL * (1 + NT);

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      a[r] = b[r] = c[r] = 0.0f;
#pragma unroll
      for (int k = 0; k < NTA; ++k) ta[k][r] = tb[k][r] = tc[k][r] = 0.0f;
    }
  }
  __device__ __forceinline__ void store(float* base, int64_t stride) const {
    int i = 0;
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      base[(i++) * stride] = a[r];
      base[(i++) * stride] = b[r];
      base[(i++) * stride] = c[r];
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        base[(i++) * stride] = ta[k][r];
        base[(i++) * stride] = tb[k][r];
        base[(i++) * stride] = tc[k][r];
      }
    }
  }
  __device__ __forceinline__ void load(const float* base, int64_t stride) {
    int i = 0;
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      a[r] = base[(i++) * stride];
      b[r] = base[(i++) * stride];
      c[r] = base[(i++) * stride];
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        ta[k][r] = base[(i++) * stride];
        tb[k][r] = base[(i++) * stride];
        tc[k][r] = base[(i++) * stride];
      }
    }
  }
};

// One atom's constants.
struct Atom {
  float r1, r2, k1, k2, t1, t2, m0, b1, eta;
  bool transmit;
};

// Decay over one interval, memoized on the interval's bits: e1 - 1 (expm1),
// e1, e2, and their T1 / T2 derivatives de1 = e1 k1 dt, de2 = e2 k2 dt.
struct Decay {
  uint32_t bits = 0xFFFFFFFFu;
  float e1m = 0.0f, e1 = 1.0f, e2 = 1.0f, de1 = 0.0f, de2 = 0.0f;
  __device__ __forceinline__ void set(float dt, const Atom& at) {
    if (__float_as_uint(dt) == bits) return;
    bits = __float_as_uint(dt);
    e1m = expm1f(-at.r1 * dt);
    e1 = 1.0f + e1m;
    e2 = expf(-at.r2 * dt);
    de1 = e1 * (at.k1 * dt);
    de2 = e2 * (at.k2 * dt);
  }
};

// 1 - cos and sin of an op's flip as the atom plays it (the forward kernel's
// convention: with a B1 map both from the half angle).
__device__ __forceinline__ void op_angle(const MacroOp& m, const Atom& at, float& vs, float& sn) {
  if (at.transmit) {
    float sh, ch;
    sincos_reduced(0.5f * (m.flip * at.b1), sh, ch);
    const float s2 = sh + sh;
    vs = s2 * sh;
    sn = s2 * ch;
  } else {
    vs = m.versine;
    sn = m.sine;
  }
}

// Relaxation and recovery over the op's interval (reference relax_real,
// _epg_cpu.cpp:2930-2975), with the tangents' source terms: along T1 the
// longitudinal decay moves by de1 (Z - Z_eq), along T2 the transverse by de2.
template <int SL, int TM>
__device__ __forceinline__ void relax_forward(Dual<SL, TM>& x, const Decay& f, bool first) {
  constexpr int NT = Dual<SL, TM>::NT;
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    const float cm = (r == 0 && first) ? x.c[r] - 1.0f : x.c[r];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      if ((TM & TAN_T2) != 0 && k == ST2) {
        x.ta[k][r] = fmaf(x.ta[k][r], f.e2, x.a[r] * f.de2);
        x.tb[k][r] = fmaf(x.tb[k][r], f.e2, x.b[r] * f.de2);
      } else {
        x.ta[k][r] *= f.e2;
        x.tb[k][r] *= f.e2;
      }
      if ((TM & TAN_T1) != 0 && k == ST1) {
        x.tc[k][r] = fmaf(x.tc[k][r], f.e1, cm * f.de1);
      } else {
        x.tc[k][r] *= f.e1;
      }
    }
    x.a[r] *= f.e2;
    x.b[r] *= f.e2;
    x.c[r] = fmaf(cm, f.e1m, x.c[r]);
  }
}

// Rotation about x by alpha (reference rf_real, _epg_cpu.cpp:2980-3005):
// u = (vs/2) d + sn z, p' = p - u, m' = m + u, z' = (sn/2) d + z - vs z with
// d = p - m. Along B1 the angle moves by the flip per unit B1, which adds
// flip (du/dalpha, dz'/dalpha) of the value planes to that tangent set.
template <int SL, int TM>
__device__ __forceinline__ void rotate_forward(Dual<SL, TM>& x, float vs, float sn, float flip) {
  constexpr int NT = Dual<SL, TM>::NT;
  constexpr int SB1 = slot_of(TM, TAN_B1);
  const float hv = 0.5f * vs, hs = 0.5f * sn, cs = 1.0f - vs;
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    const float d = x.a[r] - x.b[r];
    const float zv = x.c[r];
#pragma unroll
    for (int k = 0; k < NT; ++k) {
      const float dk = x.ta[k][r] - x.tb[k][r];
      float u = fmaf(hv, dk, sn * x.tc[k][r]);
      float z = fmaf(hs, dk, fmaf(-vs, x.tc[k][r], x.tc[k][r]));
      if ((TM & TAN_B1) != 0 && k == SB1) {
        u = fmaf(flip, fmaf(hs, d, cs * zv), u);
        z = fmaf(flip, fmaf(0.5f * cs, d, -sn * zv), z);
      }
      x.ta[k][r] -= u;
      x.tb[k][r] += u;
      x.tc[k][r] = z;
    }
    const float u = fmaf(hv, d, sn * zv);
    x.c[r] = fmaf(hs, d, fmaf(-vs, zv, zv));
    x.a[r] -= u;
    x.b[r] += u;
  }
}

template <int SL, int L, int TM>
__device__ __forceinline__ void shift_forward(Dual<SL, TM>& x, bool first, bool last, bool padded, unsigned keep) {
  shift_pair<SL, L>(x.a, x.b, first, last);
  if (padded) clear_padding<SL>(x.a, keep);
#pragma unroll
  for (int k = 0; k < Dual<SL, TM>::NT; ++k) {
    shift_pair<SL, L>(x.ta[k], x.tb[k], first, last);
    if (padded) clear_padding<SL>(x.ta[k], keep);
  }
}

template <int SL, int L, int TM>
__device__ __forceinline__ void shift_adjoint(Dual<SL, TM>& l, bool first, bool last, int sub, bool padded,
                                              unsigned keep) {
  shift_pair_t<SL, L>(l.a, l.b, first, last, sub);
  if (padded) clear_padding<SL>(l.b, keep);
#pragma unroll
  for (int k = 0; k < Dual<SL, TM>::NT; ++k) {
    shift_pair_t<SL, L>(l.ta[k], l.tb[k], first, last, sub);
    if (padded) clear_padding<SL>(l.tb[k], keep);
  }
}

// The op's plane transform before its record and shift: relaxation, then a
// rotation, an inversion or a spoil.
template <int SL, int TM>
__device__ __forceinline__ void transform_forward(Dual<SL, TM>& x, const MacroOp& m, uint32_t ctl, Decay& f,
                                                  const Atom& at, bool first) {
  constexpr int NT = Dual<SL, TM>::NT;
  if (ctl & OP_RELAX) {
    f.set(m.dt, at);
    relax_forward<SL, TM>(x, f, first);
  }
  if (ctl & OP_ROTATE) {
    float vs, sn;
    op_angle(m, at, vs, sn);
    rotate_forward<SL, TM>(x, vs, sn, m.flip);
  } else if (ctl & OP_INVERT) {
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      x.c[r] *= -at.eta;
#pragma unroll
      for (int k = 0; k < NT; ++k) x.tc[k][r] *= -at.eta;
    }
  } else if (ctl & OP_SPOIL) {
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      x.a[r] = x.b[r] = 0.0f;
#pragma unroll
      for (int k = 0; k < NT; ++k) x.ta[k][r] = x.tb[k][r] = 0.0f;
    }
  }
}

__global__ void clear_gradients(float* __restrict__ a, float* __restrict__ b, int64_t events) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= events) return;
  if (a) a[e] = 0.0f;
  if (b) b[e] = 0.0f;
}

template <int SL, int L, int TM>
__global__ void __launch_bounds__(256) adjoint_kernel(const AdjointArgs p) {
  using D = Dual<SL, TM>;
  constexpr int NT = D::NT;
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
  constexpr int SB1 = slot_of(TM, TAN_B1);
  extern __shared__ float s_part_all[];  // per wave: [2 * FLUSH rows][64 lanes]
  const int sub = (int)(threadIdx.x % L);
  const bool first = sub == 0;
  const bool last = sub == L - 1;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t group = tid / L;
  const bool live = group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const int64_t wave = tid / 64;
  const int lane = (int)(threadIdx.x & 63);
  float* const s_part = s_part_all + (threadIdx.x >> 6) * (2 * FLUSH * 64);
  const int base = sub * SL;
  unsigned keep = 0;
#pragma unroll
  for (int r = 0; r < SL; ++r) keep |= (base + r < p.states) ? (1u << r) : 0u;
  const bool padded = SL * L != p.states;

  Atom at;
  at.t1 = p.t1[atom];
  at.t2 = p.t2[atom];
  at.r1 = 1000.0f / at.t1;
  at.r2 = 1000.0f / at.t2;
  at.k1 = 1000.0f / (at.t1 * at.t1);
  at.k2 = 1000.0f / (at.t2 * at.t2);
  at.m0 = p.density ? p.m0[atom * p.stride_opt] : 1.0f;
  at.b1 = p.transmit ? p.b1[atom * p.stride_opt] : 1.0f;
  at.eta = p.inverting ? p.inv_eff[atom * p.stride_opt] : 1.0f;
  at.transmit = p.transmit != 0;
  const int64_t N = *p.op_count < p.op_bound ? *p.op_count : p.op_bound;
  const int64_t segments = (N + p.segment - 1) / p.segment;
  const int64_t T = p.threads_total;
  const float* ws_row = p.w_signal ? p.w_signal + atom * p.outputs : nullptr;
  const float* wj_row = p.w_jac ? p.w_jac + atom * (int64_t)p.columns * p.outputs : nullptr;

  // Pass 1: forward, one checkpoint per segment.
  D x;
  x.zero();
  x.c[0] = first ? 1.0f : 0.0f;
  Decay f;
  for (int64_t i = 0; i < N; ++i) {
    if (i % p.segment == 0) x.store(p.checkpoints + (i / p.segment) * D::floats * T + tid, T);
    const MacroOp m = load_op(p.ops, i);
    transform_forward<SL, TM>(x, m, m.ctrl, f, at, first);
    if (m.ctrl & OP_SHIFT) shift_forward<SL, L, TM>(x, first, last, padded, keep);
  }

  // Pass 2: reverse, one segment at a time.
  float g[G_COUNT] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  D lam;
  lam.zero();
  int rows = 0;          // ops held in s_part
  int64_t flush_top = 0; // op index of row 0
  auto flush = [&]() __attribute__((always_inline)) {
    __syncthreads();
    // lanes 0-31 and 32-63 each sum half of one row (2 rows per op)
    const int row = lane & 31;
    const int half = lane >> 5;
    float v = 0.0f;
    if (row < 2 * rows) {
      const float* src = s_part + row * 64 + half * 32;
#pragma unroll 8
      for (int j = 0; j < 32; ++j) v += src[j];
    }
    v += __shfl_xor(v, 32, 64);
    if (half == 0 && row < 2 * rows) {
      const int64_t op = flush_top - (row >> 1);
      p.partial[(wave * p.op_bound + op) * 2 + (row & 1)] = v;
    }
    __syncthreads();
    rows = 0;
  };

  for (int64_t seg = segments - 1; seg >= 0; --seg) {
    const int64_t i0 = seg * p.segment;
    const int64_t i1 = i0 + p.segment < N ? i0 + p.segment : N;
    float* const states = p.scratch + tid;
    x.load(p.checkpoints + seg * D::floats * T + tid, T);
    Decay fr;
    for (int64_t i = i0; i < i1; ++i) {
      x.store(states + (i - i0) * D::floats * T, T);
      const MacroOp m = load_op(p.ops, i);
      transform_forward<SL, TM>(x, m, m.ctrl, fr, at, first);
      if (m.ctrl & OP_SHIFT) shift_forward<SL, L, TM>(x, first, last, padded, keep);
    }
    D xn;
    xn.load(states + (i1 - 1 - i0) * D::floats * T, T);
    MacroOp mn = load_op(p.ops, i1 - 1);
    for (int64_t i = i1 - 1; i >= i0; --i) {
      const D xe = xn;  // the op's entry state
      const MacroOp m = mn;
      if (i > i0) {
        xn.load(states + (i - 1 - i0) * D::floats * T, T);
        mn = load_op(p.ops, i - 1);
      }
      const uint32_t ctl = m.ctrl;
      if (rows == 0) flush_top = i;
      float g_alpha = 0.0f;  // d loss / d alpha of this op (alpha = flip * b1)
      float g_coef = 0.0f;   // d loss / d flip through the B1 direction's own coefficient
      float g_dt = 0.0f;

      if (ctl & OP_SHIFT) shift_adjoint<SL, L, TM>(lam, first, last, sub, padded, keep);

      // the state the pulse acts on
      D xr = xe;
      if (ctl & OP_RELAX) {
        fr.set(m.dt, at);
        relax_forward<SL, TM>(xr, fr, first);
      }
      float vs = 0.0f, sn = 0.0f;
      if (ctl & OP_ROTATE) op_angle(m, at, vs, sn);
      const float hv = 0.5f * vs, hs = 0.5f * sn, cs = 1.0f - vs;

      // the record reads F+[0] (and its tangents) after the pulse
      if ((ctl & OP_RECORD) && m.out >= 0 && m.out < p.outputs && first && live) {
        const int64_t o = m.out;
        const float ws = ws_row ? ws_row[o] : 0.0f;
        const float wm0 = (wj_row && p.col_m0 >= 0) ? wj_row[(int64_t)p.col_m0 * p.outputs + o] : 0.0f;
        float wk[D::NTA];
#pragma unroll
        for (int k = 0; k < D::NTA; ++k) wk[k] = 0.0f;
        if (wj_row) {
          if constexpr ((TM & TAN_T1) != 0) wk[ST1] = p.col_t1 >= 0 ? wj_row[(int64_t)p.col_t1 * p.outputs + o] : 0.0f;
          if constexpr ((TM & TAN_T2) != 0) wk[ST2] = p.col_t2 >= 0 ? wj_row[(int64_t)p.col_t2 * p.outputs + o] : 0.0f;
          if constexpr ((TM & TAN_B1) != 0) wk[SB1] = p.col_b1 >= 0 ? wj_row[(int64_t)p.col_b1 * p.outputs + o] : 0.0f;
        }
        lam.a[0] += at.m0 * ws + wm0;
#pragma unroll
        for (int k = 0; k < NT; ++k) lam.ta[k][0] += at.m0 * wk[k];
        if (p.grad[G_M0]) {
          // F+[0] after the pulse, of the value planes and each tangent set
          float a0 = xr.a[0];
          float ta0[D::NTA];
#pragma unroll
          for (int k = 0; k < NT; ++k) ta0[k] = xr.ta[k][0];
          if (ctl & OP_ROTATE) {
            const float d = xr.a[0] - xr.b[0];
            a0 -= fmaf(hv, d, sn * xr.c[0]);
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              float u = fmaf(hv, xr.ta[k][0] - xr.tb[k][0], sn * xr.tc[k][0]);
              if ((TM & TAN_B1) != 0 && k == SB1) u = fmaf(m.flip, fmaf(hs, d, cs * xr.c[0]), u);
              ta0[k] -= u;
            }
          } else if (ctl & OP_SPOIL) {
            a0 = 0.0f;
#pragma unroll
            for (int k = 0; k < NT; ++k) ta0[k] = 0.0f;
          }
          float gm = a0 * ws;
#pragma unroll
          for (int k = 0; k < NT; ++k) gm = fmaf(ta0[k], wk[k], gm);
          g[G_M0] += gm;
        }
      }

      // the pulse
      if (ctl & OP_ROTATE) {
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          const float d = xr.a[r] - xr.b[r];
          const float zv = xr.c[r];
          const float w1 = fmaf(hs, d, cs * zv);          // du/dalpha
          const float w2 = fmaf(0.5f * cs, d, -sn * zv);  // dz'/dalpha
          const float lu = lam.b[r] - lam.a[r];
          const float lz = lam.c[r];
          g_alpha = fmaf(lu, w1, fmaf(lz, w2, g_alpha));
          float xa = 0.0f, xc = 0.0f;  // the B1 direction's term reaching the value planes' adjoint
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            const float dk = xr.ta[k][r] - xr.tb[k][r];
            const float zk = xr.tc[k][r];
            const float lut = lam.tb[k][r] - lam.ta[k][r];
            const float lzt = lam.tc[k][r];
            g_alpha = fmaf(lut, fmaf(hs, dk, cs * zk), fmaf(lzt, fmaf(0.5f * cs, dk, -sn * zk), g_alpha));
            if ((TM & TAN_B1) != 0 && k == SB1) {
              g_coef = fmaf(lut, w1, fmaf(lzt, w2, g_coef));
              g_alpha = fmaf(m.flip, fmaf(lut, w2, -lzt * w1), g_alpha);
              xa = m.flip * fmaf(hs, lut, 0.5f * cs * lzt);
              xc = m.flip * fmaf(cs, lut, -sn * lzt);
            }
            const float ld = fmaf(hv, lut, hs * lzt);
            lam.ta[k][r] += ld;
            lam.tb[k][r] -= ld;
            lam.tc[k][r] = fmaf(sn, lut, cs * lzt);
          }
          const float ld = fmaf(hv, lu, hs * lz);
          lam.a[r] += ld + xa;
          lam.b[r] -= ld + xa;
          lam.c[r] = fmaf(sn, lu, cs * lz) + xc;
        }
      } else if (ctl & OP_INVERT) {
        float gi = 0.0f;
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          gi = fmaf(lam.c[r], xr.c[r], gi);
          lam.c[r] *= -at.eta;
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            gi = fmaf(lam.tc[k][r], xr.tc[k][r], gi);
            lam.tc[k][r] *= -at.eta;
          }
        }
        g[G_IE] -= gi;
      } else if (ctl & OP_SPOIL) {
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          lam.a[r] = lam.b[r] = 0.0f;
#pragma unroll
          for (int k = 0; k < NT; ++k) lam.ta[k][r] = lam.tb[k][r] = 0.0f;
        }
      }

      // the relaxation before it: gradients through e1, e2 and their T1 / T2
      // derivatives, then lambda <- R^T lambda
      if (ctl & OP_RELAX) {
        const float dt = m.dt;
        const float e1 = fr.e1, e2 = fr.e2, de1 = fr.de1, de2 = fr.de2;
        float gE1 = 0.0f, gE2 = 0.0f, gS1 = 0.0f, gS2 = 0.0f;  // sources: along T1 (de1), along T2 (de2)
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          const float cm = (r == 0 && first) ? xe.c[r] - 1.0f : xe.c[r];
          gE2 = fmaf(lam.a[r], xe.a[r], fmaf(lam.b[r], xe.b[r], gE2));
          gE1 = fmaf(lam.c[r], cm, gE1);
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            gE2 = fmaf(lam.ta[k][r], xe.ta[k][r], fmaf(lam.tb[k][r], xe.tb[k][r], gE2));
            gE1 = fmaf(lam.tc[k][r], xe.tc[k][r], gE1);
            if ((TM & TAN_T1) != 0 && k == ST1) gS1 = fmaf(lam.tc[k][r], cm, gS1);
            if ((TM & TAN_T2) != 0 && k == ST2) gS2 = fmaf(lam.ta[k][r], xe.a[r], fmaf(lam.tb[k][r], xe.b[r], gS2));
          }
          float na = e2 * lam.a[r], nb = e2 * lam.b[r], nc = e1 * lam.c[r];
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            if ((TM & TAN_T2) != 0 && k == ST2) {
              na = fmaf(de2, lam.ta[k][r], na);
              nb = fmaf(de2, lam.tb[k][r], nb);
            }
            if ((TM & TAN_T1) != 0 && k == ST1) nc = fmaf(de1, lam.tc[k][r], nc);
            lam.ta[k][r] *= e2;
            lam.tb[k][r] *= e2;
            lam.tc[k][r] *= e1;
          }
          lam.a[r] = na;
          lam.b[r] = nb;
          lam.c[r] = nc;
        }
        // d e1/d dt = -r1 e1, d e2/d dt = -r2 e2; d de1/d dt = k1 e1 (1 - r1 dt),
        // d de1/d T1 = de1 (k1 dt - 2/T1), and the same for T2
        g_dt = -at.r1 * e1 * gE1 - at.r2 * e2 * gE2 + at.k1 * e1 * (1.0f - at.r1 * dt) * gS1 +
               at.k2 * e2 * (1.0f - at.r2 * dt) * gS2;
        g[G_T1] += de1 * gE1 + de1 * (at.k1 * dt - 2.0f / at.t1) * gS1;
        g[G_T2] += de2 * gE2 + de2 * (at.k2 * dt - 2.0f / at.t2) * gS2;
      }
      g[G_B1] = fmaf(m.flip, g_alpha, g[G_B1]);

      const float g_flip = live ? fmaf(at.b1, g_alpha, g_coef) : 0.0f;
      s_part[(2 * rows) * 64 + lane] = g_flip;
      s_part[(2 * rows + 1) * 64 + lane] = live ? g_dt : 0.0f;
      ++rows;
      if (rows == FLUSH || i == i0) flush();
    }
  }

  // tissue gradients: sum over the atom's lanes
#pragma unroll
  for (int q = 0; q < G_COUNT; ++q) {
    float v = g[q];
#pragma unroll
    for (int offset = L / 2; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
    if (first && live && p.grad[q]) p.grad[q][group] = v;
  }
}

// Per-op partials summed over waves in a fixed order, then scattered to the
// events the op's rotation and relaxation came from.
__global__ void reduce_partials(const float* __restrict__ partial, int64_t waves, int64_t ops,
                                const int64_t* __restrict__ op_count, const MacroOp* __restrict__ codes,
                                float* __restrict__ grad_flip, float* __restrict__ grad_duration) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= ops || i >= *op_count) return;
  float gf = 0.0f, gd = 0.0f;
  for (int64_t w = 0; w < waves; ++w) {
    gf += partial[(w * ops + i) * 2];
    gd += partial[(w * ops + i) * 2 + 1];
  }
  const MacroOp m = codes[i];
  if (grad_flip && (m.ctrl & OP_ROTATE) && m.event >= 0) grad_flip[m.event] = gf;
  if (grad_duration && (m.ctrl & OP_RELAX) && m.relax_event >= 0) grad_duration[m.relax_event] = gd;
}

using AdjointLaunch = void (*)(const AdjointArgs&, dim3, dim3, size_t, hipStream_t);

template <int SL, int L, int TM>
void launch_adjoint(const AdjointArgs& a, dim3 g, dim3 b, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((adjoint_kernel<SL, L, TM>), g, b, lds, s, a);
}

// Tangent sets compiled: none, T1 + T2, T1 + T2 + B1; a request runs the
// smallest that covers it (an unseeded set adds work, not error).
template <int SL, int L>
AdjointLaunch pick_adjoint_mask(int tm) {
  if (tm == 0) return &launch_adjoint<SL, L, 0>;
  if ((tm & TAN_B1) == 0) return &launch_adjoint<SL, L, TAN_T1 | TAN_T2>;
  return &launch_adjoint<SL, L, TAN_T1 | TAN_T2 | TAN_B1>;
}

struct AdjointTile {
  int sl, l;
};
// Lanes per atom divide a 16-lane DPP row.
constexpr AdjointTile kAdjointTiles[] = {{1, 4}, {1, 8}, {1, 16}, {2, 4}, {2, 8}, {2, 16},
                                         {3, 4}, {3, 16}, {4, 16}, {5, 4}};
constexpr int kAdjointTileCount = (int)(sizeof(kAdjointTiles) / sizeof(kAdjointTiles[0]));

AdjointLaunch pick_adjoint(int tile, int tm) {
  switch (tile) {
    case 0: return pick_adjoint_mask<1, 4>(tm);
    case 1: return pick_adjoint_mask<1, 8>(tm);
    case 2: return pick_adjoint_mask<1, 16>(tm);
    case 3: return pick_adjoint_mask<2, 4>(tm);
    case 4: return pick_adjoint_mask<2, 8>(tm);
    case 5: return pick_adjoint_mask<2, 16>(tm);
    case 6: return pick_adjoint_mask<3, 4>(tm);
    case 7: return pick_adjoint_mask<3, 16>(tm);
    case 8: return pick_adjoint_mask<4, 16>(tm);
    default: return pick_adjoint_mask<5, 4>(tm);
  }
}

// With enough atoms to fill the chip the tile with the least padding (then
// the most orders per lane) wins. With few atoms the adjoint is a latency
// chain per atom -- every op's recompute and reverse step walk the lane's SL
// orders in sequence -- so the tile with the fewest orders per lane among
// those padding at most 2x wins instead (more lanes per atom, a shorter
// chain).
int adjoint_tile(int states, int64_t atoms) {
  constexpr int64_t kFillLanes = 1024 * 64 * 2;  // two waves on every SIMD
  int best = -1;
  for (int t = 0; t < kAdjointTileCount; ++t) {
    const int cover = kAdjointTiles[t].sl * kAdjointTiles[t].l;
    if (cover < states) continue;
    if (best < 0) {
      best = t;
      continue;
    }
    const int best_cover = kAdjointTiles[best].sl * kAdjointTiles[best].l;
    if (cover < best_cover || (cover == best_cover && kAdjointTiles[t].sl > kAdjointTiles[best].sl)) best = t;
  }
  if (best < 0 || atoms * kAdjointTiles[best].l >= kFillLanes) return best;
  const int least = kAdjointTiles[best].sl * kAdjointTiles[best].l;
  for (int t = 0; t < kAdjointTileCount; ++t) {
    const int cover = kAdjointTiles[t].sl * kAdjointTiles[t].l;
    if (cover < states || cover > 2 * least) continue;
    const bool shorter = kAdjointTiles[t].sl < kAdjointTiles[best].sl;
    const bool tighter = kAdjointTiles[t].sl == kAdjointTiles[best].sl &&
                         cover < kAdjointTiles[best].sl * kAdjointTiles[best].l;
    if (shorter || tighter) best = t;
  }
  return best;
}

}  // namespace
}  // namespace tsim

using namespace tsim;

extern "C" int tsim_simulate_vjp_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                                          int64_t state_count, int64_t output_count, int features,
                                          int64_t tissue_stride, const int32_t* columns,
                                          int64_t column_count, const float* signal_cotangent,
                                          const float* jacobian_cotangent, int64_t segment_ops, int64_t op_bound,
                                          float* grad_flip, float* grad_duration, float* const* tissue_gradients,
                                          int64_t tissue_gradient_count, const uint8_t* duration_mask,
                                          void* stream) {
  if (packed == nullptr || packed[T_T1] == nullptr || packed[T_T2] == nullptr) {
    set_error("packed buffers with t1 and t2 are required");
    return BAD_ARGUMENT;
  }
  if (atom_count < 0 || event_count < 0 || state_count < 2 || output_count < 0) {
    set_error("invalid dimensions (the adjoint needs at least 2 orders)");
    return BAD_ARGUMENT;
  }
  if (segment_ops < 1) {
    set_error("segment_ops must be positive");
    return BAD_ARGUMENT;
  }
  if (column_count < 0 || column_count > 4 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 4 Jacobian columns (t1, t2, m0, b1), got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (features & (FEATURE_DIFFUSING | FEATURE_OFF_AXIS | FEATURE_MOVING)) {
    set_error("the adjoint carries no diffusion, off-resonance or flow term");
    return UNSUPPORTED;
  }
  if (tissue_gradient_count != 0 && (tissue_gradients == nullptr || tissue_gradient_count != TISSUE_COUNT)) {
    set_error("tissue_gradients holds %d pointers (tissue order) or tissue_gradient_count is 0", TISSUE_COUNT);
    return BAD_ARGUMENT;
  }
  const int64_t capacity = op_capacity(event_count);
  if (op_bound > capacity) op_bound = capacity;
  AdjointArgs a;
  memset(&a, 0, sizeof(a));
  if (tissue_gradient_count == TISSUE_COUNT) {
    for (int q = 0; q < TISSUE_COUNT; ++q) {
      if (tissue_gradients[q] == nullptr) continue;
      switch (q) {
        case T_T1: a.grad[G_T1] = tissue_gradients[q]; break;
        case T_T2: a.grad[G_T2] = tissue_gradients[q]; break;
        case T_M0: a.grad[G_M0] = tissue_gradients[q]; break;
        case T_B1: a.grad[G_B1] = tissue_gradients[q]; break;
        case T_INV_EFF: a.grad[G_IE] = tissue_gradients[q]; break;
        default:
          set_error("tissue property %d carries no gradient in the real-subspace adjoint", q);
          return UNSUPPORTED;
      }
    }
  }
  int tm = 0;
  a.col_t1 = a.col_t2 = a.col_b1 = a.col_m0 = -1;
  for (int64_t c = 0; c < column_count; ++c) {
    int* slot = nullptr;
    switch (columns[c]) {
      case T_T1: tm |= TAN_T1; slot = &a.col_t1; break;
      case T_T2: tm |= TAN_T2; slot = &a.col_t2; break;
      case T_B1: tm |= TAN_B1; slot = &a.col_b1; break;
      case T_M0: slot = &a.col_m0; break;
      default:
        set_error("column %lld: the adjoint carries tangents along t1, t2, b1 and m0 only", (long long)c);
        return BAD_ARGUMENT;
    }
    if (*slot >= 0) {
      set_error("column %lld repeats a property", (long long)c);
      return BAD_ARGUMENT;
    }
    *slot = (int)c;
  }
  if (jacobian_cotangent == nullptr) tm = 0;  // no column carries a cotangent: first order
  const int tile = adjoint_tile((int)state_count, atom_count);
  if (tile < 0) {
    set_error("no adjoint tile covers %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  hipStream_t s = (hipStream_t)stream;
  if (event_count > 0) {
    const int64_t rb = (event_count + 255) / 256;
    hipLaunchKernelGGL(clear_gradients, dim3((unsigned)rb), dim3(256), 0, s, grad_flip, grad_duration,
                       event_count);
    int st = hip_status(hipGetLastError(), "clear_gradients launch");
    if (st != OK) return st;
  }
  if (atom_count == 0 || event_count == 0 || op_bound == 0) {
    // no atom or op contributes: every gradient is zero
    for (int q = 0; q < G_COUNT; ++q) {
      if (a.grad[q] && atom_count > 0) {
        int st = hip_status(hipMemsetAsync(a.grad[q], 0, sizeof(float) * (size_t)atom_count, s), "tissue clear");
        if (st != OK) return st;
      }
    }
    return OK;
  }
  const int L = kAdjointTiles[tile].l;
  const int SL = kAdjointTiles[tile].sl;
  const int NT = popcount4(tm == 0 ? 0 : ((tm & TAN_B1) ? 7 : 3));
  const int64_t nfloat = 3LL * SL * (1 + NT);
  // Few atoms leave most compute units idle with 4-wave blocks: one-wave
  // blocks spread the same waves over four times as many units. Every index
  // derives from the global thread id and the launch total, not the block.
  const int threads = atom_count * L < 256LL * 256 ? 64 : 256;
  const int64_t blocks = (atom_count * L + threads - 1) / threads;
  const int64_t total = blocks * threads;
  const int64_t waves = total / 64;

  // The macro-op stream and its length on the device. Without a bound from
  // the caller the length is read back once (one synchronization); it sizes
  // the checkpoints and the per-op partials.
  Compiled compiled;
  int st = compile_packed(packed, event_count, s, compiled);
  if (st != OK) return st;
  char* head = nullptr;
  st = scratch_alloc((void**)&head, sizeof(MacroOp) * (size_t)capacity + 256, s, "adjoint ops");
  if (st != OK) return st;
  MacroOp* ops = (MacroOp*)head;
  int64_t* op_count_device = (int64_t*)(head + sizeof(MacroOp) * (size_t)capacity);
  st = pack_stream(compiled.codes, event_count, output_count, ops, s, op_count_device, duration_mask,
                   duration_mask ? 2 : 1);
  if (st == OK && op_bound < 0) {
    int64_t counted = 0;
    st = hip_status(hipMemcpyAsync(&counted, op_count_device, sizeof(int64_t), hipMemcpyDeviceToHost, s),
                    "op count read-back");
    if (st == OK) st = hip_status(hipStreamSynchronize(s), "op count read-back");
    op_bound = counted;
  }
  if (st != OK || op_bound == 0) {
    for (int q = 0; q < G_COUNT && st == OK; ++q) {
      if (a.grad[q]) st = hip_status(hipMemsetAsync(a.grad[q], 0, sizeof(float) * (size_t)atom_count, s),
                                     "tissue clear");
    }
    (void)hipFreeAsync(head, s);
    return st;
  }

  a.ops = ops;
  a.op_count = op_count_device;
  a.op_bound = op_bound;
  a.t1 = (const float*)packed[T_T1];
  a.t2 = (const float*)packed[T_T2];
  a.m0 = (const float*)packed[T_M0];
  a.b1 = (const float*)packed[T_B1];
  a.inv_eff = (const float*)packed[T_INV_EFF];
  a.stride_opt = tissue_stride;
  a.atoms = atom_count;
  a.outputs = output_count;
  a.states = (int)state_count;
  a.density = (features & FEATURE_DENSITY) && a.m0 ? 1 : 0;
  a.inverting = (features & FEATURE_INVERTING) && a.inv_eff ? 1 : 0;
  a.transmit = (features & FEATURE_TRANSMIT) && a.b1 ? 1 : 0;
  a.w_signal = signal_cotangent;
  a.w_jac = column_count > 0 ? jacobian_cotangent : nullptr;
  a.columns = (int)column_count;
  a.segment = segment_ops;
  a.threads_total = total;
  const int64_t segments = (op_bound + segment_ops - 1) / segment_ops;
  const int64_t seg_ops = segment_ops < op_bound ? segment_ops : op_bound;
  const size_t ck = sizeof(float) * (size_t)(segments * nfloat * total);
  const size_t sc = sizeof(float) * (size_t)(seg_ops * nfloat * total);
  const size_t pa = sizeof(float) * (size_t)(waves * op_bound * 2);
  char* work = nullptr;
  st = scratch_alloc((void**)&work, ck + sc + pa, s, "adjoint workspace");
  if (st != OK) {
    (void)hipFreeAsync(head, s);
    return st;
  }
  a.checkpoints = (float*)work;
  a.scratch = (float*)(work + ck);
  a.partial = (float*)(work + ck + sc);
  const size_t lds = sizeof(float) * (size_t)(threads / 64) * 2 * FLUSH * 64;
  pick_adjoint(tile, tm)(a, dim3((unsigned)blocks), dim3(threads), lds, s);
  st = hip_status(hipGetLastError(), "adjoint_kernel launch");
  if (st == OK) {
    const int64_t rb = (op_bound + 255) / 256;
    hipLaunchKernelGGL(reduce_partials, dim3((unsigned)rb), dim3(256), 0, s, a.partial, waves, op_bound,
                       (const int64_t*)op_count_device, (const MacroOp*)ops, grad_flip, grad_duration);
    st = hip_status(hipGetLastError(), "reduce_partials launch");
  }
  (void)hipFreeAsync(work, s);
  (void)hipFreeAsync(head, s);
  return st;
}
