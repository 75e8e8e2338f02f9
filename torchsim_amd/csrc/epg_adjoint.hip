// Checkpointed adjoint of the fused real-subspace EPG system, for gfx950.
//
// Gradients of a scalar loss with respect to every event's flip angle and
// duration, given the loss's cotangents on the recorded signal and on the
// Jacobian columns along T1/T2/M0. What it computes is the reference's
// forward-over-reverse pass (src/torchsim/_epg_cpu.cpp:7028-7314
// simulate_real_vjp_jvp_range; first order :5985-6227; Triton
// sequence/_epg_triton.py:13155-13844) -- the path `crlb(J).backward()` takes
// in sequence design. Where the reference records every event's state for
// the whole train, this kernel keeps one dual state per segment of
// `segment` events (64 pulses = 256 events of an MRF train), recomputes a
// segment forward into per-thread scratch, and walks it in reverse.
//
// State layout as in the forward kernel: S orders split into L blocks of SL
// registers over L adjacent lanes; the transposed shift moves the adjoint of
// F+ down and of F- up, with one DPP crossing per plane at block edges.
#include "epg_real_impl.h"

#include <cstring>

namespace tsim {
namespace {

struct AdjointArgs {
  const EventCode* events;
  int64_t event_count;
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
  int col_t1, col_t2, col_m0;
  int64_t segment;        // events per checkpoint
  int64_t segments;
  int64_t threads_total;
  float* checkpoints;     // (segments, nfloat, threads_total)
  float* scratch;         // (segment, nfloat, threads_total)
  float* partial;         // (waves, events, 2): d/dflip, d/dduration
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int offset = 32; offset > 0; offset >>= 1) v += __shfl_xor(v, offset, 64);
  return v;
}

// Transpose of shift_pair: the adjoint of F+ moves DOWN an order, of F- UP;
// the coupling F+[0] = -F-[0] (after the shift) sends -lambda_a'[0] to
// lambda_b[1]. Needs SL >= 2 so that order 1 sits in the atom's first lane.
template <int SL, int L>
__device__ __forceinline__ void shift_pair_t(float (&la)[SL], float (&lb)[SL], bool first, bool last) {
  const float a0 = la[0];
  float a_in = 0.0f;  // lambda_a'[0] of the next lane -> lambda_a[SL-1]
  float b_in = 0.0f;  // lambda_b'[SL-1] of the previous lane -> lambda_b[0]
  if constexpr (L > 1) {
    a_in = dpp_from_upper(la[0]);
    a_in = last ? 0.0f : a_in;
    b_in = dpp_from_lower(lb[SL - 1]);
    b_in = first ? 0.0f : b_in;
  }
#pragma unroll
  for (int r = 0; r + 1 < SL; ++r) la[r] = la[r + 1];
  la[SL - 1] = a_in;
#pragma unroll
  for (int r = SL - 1; r > 0; --r) lb[r] = lb[r - 1];
  lb[0] = b_in;
  if (first) lb[1] -= a0;
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

// Interval factors of one atom, memoized on the interval.
struct Relax {
  float dt = -1.0f;
  float e1 = 1.0f, e2 = 1.0f, de1 = 0.0f, de2 = 0.0f, om1 = 0.0f;
  __device__ __forceinline__ void set(float t, float r1, float r2, float k1, float k2) {
    if (t == dt) return;
    dt = t;
    // 1 - e1 from expm1: a rounded e1 would cost it five digits when t << T1
    const float e1m = expm1f(-r1 * t);
    e1 = 1.0f + e1m;
    om1 = -e1m;
    e2 = expf(-r2 * t);
    de1 = e1 * (k1 * t);
    de2 = e2 * (k2 * t);
  }
};

template <int SL, int L, int TM>
__device__ __forceinline__ void relax_forward(Dual<SL, TM>& x, const Relax& f, bool first) {
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    if constexpr ((TM & TAN_T1) != 0) {
      x.ta[ST1][r] *= f.e2;
      x.tb[ST1][r] *= f.e2;
      x.tc[ST1][r] = fmaf(x.tc[ST1][r], f.e1, x.c[r] * f.de1);
    }
    if constexpr ((TM & TAN_T2) != 0) {
      x.ta[ST2][r] = fmaf(x.ta[ST2][r], f.e2, x.a[r] * f.de2);
      x.tb[ST2][r] = fmaf(x.tb[ST2][r], f.e2, x.b[r] * f.de2);
      x.tc[ST2][r] *= f.e1;
    }
    x.a[r] *= f.e2;
    x.b[r] *= f.e2;
    x.c[r] *= f.e1;
  }
  x.c[0] += first ? f.om1 : 0.0f;
  if constexpr ((TM & TAN_T1) != 0) x.tc[ST1][0] -= first ? f.de1 : 0.0f;
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
__device__ __forceinline__ void shift_adjoint(Dual<SL, TM>& l, bool first, bool last, bool padded, unsigned keep) {
  shift_pair_t<SL, L>(l.a, l.b, first, last);
  if (padded) clear_padding<SL>(l.b, keep);
#pragma unroll
  for (int k = 0; k < Dual<SL, TM>::NT; ++k) {
    shift_pair_t<SL, L>(l.ta[k], l.tb[k], first, last);
    if (padded) clear_padding<SL>(l.tb[k], keep);
  }
}

struct Rotation {
  float ch, sh, sn, cs, hs;
  __device__ __forceinline__ void set(float versine, float sine) {
    cs = 1.0f - versine;
    sn = sine;
    sh = 0.5f * versine;
    ch = 1.0f - sh;
    hs = 0.5f * sn;
  }
};

template <int SL>
__device__ __forceinline__ void rotate3(float (&a)[SL], float (&b)[SL], float (&c)[SL], const Rotation& t) {
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    const float pv = a[r], mv = b[r], zv = c[r];
    a[r] = t.ch * pv + t.sh * mv - t.sn * zv;
    b[r] = t.sh * pv + t.ch * mv + t.sn * zv;
    c[r] = t.hs * (pv - mv) + t.cs * zv;
  }
}

// lambda <- T^T lambda, and the angle derivative lambda' . (dT/dalpha) x.
template <int SL>
__device__ __forceinline__ float rotate3_adjoint(float (&la)[SL], float (&lb)[SL], float (&lc)[SL],
                                                 const float (&a)[SL], const float (&b)[SL],
                                                 const float (&c)[SL], const Rotation& t) {
  float g = 0.0f;
#pragma unroll
  for (int r = 0; r < SL; ++r) {
    const float u = a[r] - b[r];
    const float da = -t.hs * u - t.cs * c[r];  // d a'/d alpha; d b'/d alpha = -da
    const float dc = 0.5f * t.cs * u - t.sn * c[r];
    g += (la[r] - lb[r]) * da + lc[r] * dc;
    const float xa = la[r], xb = lb[r], xc = lc[r];
    la[r] = t.ch * xa + t.sh * xb + t.hs * xc;
    lb[r] = t.sh * xa + t.ch * xb - t.hs * xc;
    lc[r] = -t.sn * xa + t.sn * xb + t.cs * xc;
  }
  return g;
}

template <int SL, int L, int TM>
__device__ __forceinline__ void event_forward(Dual<SL, TM>& x, const EventCode& ev, Relax& f, float r1, float r2,
                                              float k1, float k2, float b1, float eta, bool transmit, bool first,
                                              bool last, bool padded, unsigned keep) {
  const uint32_t act = ev.ctrl & 0xFFu;
  const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
  if (ev.dt != 0.0f) {
    f.set(ev.dt, r1, r2, k1, k2);
    relax_forward<SL, L, TM>(x, f, first);
  }
  if (act & PRE_SHIFT) shift_forward<SL, L, TM>(x, first, last, padded, keep);
  if (kind == KIND_RF) {
    if (act & INVERSION) {
#pragma unroll
      for (int r = 0; r < SL; ++r) {
        x.c[r] *= -eta;
#pragma unroll
        for (int k = 0; k < Dual<SL, TM>::NT; ++k) x.tc[k][r] *= -eta;
      }
    } else {
      Rotation t;
      if (transmit) {
        const float s = sinf(ev.flip * b1), h = sinf(0.5f * (ev.flip * b1));
        t.set(2.0f * (h * h), s);
      } else {
        t.set(ev.versine, ev.sine);
      }
      rotate3<SL>(x.a, x.b, x.c, t);
#pragma unroll
      for (int k = 0; k < Dual<SL, TM>::NT; ++k) rotate3<SL>(x.ta[k], x.tb[k], x.tc[k], t);
    }
  }
  if (act & POST_SHIFT) shift_forward<SL, L, TM>(x, first, last, padded, keep);
  if (act & SPOIL_AFTER) {
#pragma unroll
    for (int r = 0; r < SL; ++r) {
      x.a[r] = x.b[r] = 0.0f;
#pragma unroll
      for (int k = 0; k < Dual<SL, TM>::NT; ++k) x.ta[k][r] = x.tb[k][r] = 0.0f;
    }
  } else if (act & SHIFT_AFTER) {
    shift_forward<SL, L, TM>(x, first, last, padded, keep);
  }
}

template <int SL, int L, int TM>
__global__ void __launch_bounds__(256) adjoint_kernel(const AdjointArgs p) {
  using D = Dual<SL, TM>;
  constexpr int NT = D::NT;
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
  const int sub = (int)(threadIdx.x % L);
  const bool first = sub == 0;
  const bool last = sub == L - 1;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t group = tid / L;
  const bool live = group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const int64_t wave = tid / 64;
  const int lane = (int)(threadIdx.x & 63);
  const int base = sub * SL;
  unsigned keep = 0;
#pragma unroll
  for (int r = 0; r < SL; ++r) keep |= (base + r < p.states) ? (1u << r) : 0u;
  const bool padded = SL * L != p.states;

  const float t1 = p.t1[atom];
  const float t2 = p.t2[atom];
  const float r1 = 1000.0f / t1, r2 = 1000.0f / t2;
  const float k1 = 1000.0f / (t1 * t1), k2 = 1000.0f / (t2 * t2);
  const float m0 = p.density ? p.m0[atom * p.stride_opt] : 1.0f;
  const float b1 = p.transmit ? p.b1[atom * p.stride_opt] : 1.0f;
  const float eta = p.inverting ? p.inv_eff[atom * p.stride_opt] : 1.0f;
  const bool transmit = p.transmit != 0;
  const int64_t E = p.event_count;
  const int64_t T = p.threads_total;

  // Phase 1: forward, one checkpoint per segment.
  D x;
  x.zero();
  x.c[0] = first ? 1.0f : 0.0f;
  Relax f;
  for (int64_t e = 0; e < E; ++e) {
    if (e % p.segment == 0) x.store(p.checkpoints + (e / p.segment) * D::floats * T + tid, T);
    const EventCode ev = load_event(p.events, e);
    event_forward<SL, L, TM>(x, ev, f, r1, r2, k1, k2, b1, eta, transmit, first, last, padded, keep);
  }

  // Phase 2: reverse, one segment at a time.
  D lam;
  lam.zero();
  for (int64_t seg = p.segments - 1; seg >= 0; --seg) {
    const int64_t e0 = seg * p.segment;
    const int64_t e1 = e0 + p.segment < E ? e0 + p.segment : E;
    x.load(p.checkpoints + seg * D::floats * T + tid, T);
    Relax fr;
    for (int64_t e = e0; e < e1; ++e) {
      x.store(p.scratch + (e - e0) * D::floats * T + tid, T);
      const EventCode ev = load_event(p.events, e);
      event_forward<SL, L, TM>(x, ev, fr, r1, r2, k1, k2, b1, eta, transmit, first, last, padded, keep);
    }
    for (int64_t e = e1 - 1; e >= e0; --e) {
      const EventCode ev = load_event(p.events, e);
      const uint32_t act = ev.ctrl & 0xFFu;
      const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
      float g_flip = 0.0f, g_dt = 0.0f;
      // after-event actions, in reverse
      if (act & SPOIL_AFTER) {
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          lam.a[r] = lam.b[r] = 0.0f;
#pragma unroll
          for (int k = 0; k < NT; ++k) lam.ta[k][r] = lam.tb[k][r] = 0.0f;
        }
      } else if (act & SHIFT_AFTER) {
        shift_adjoint<SL, L, TM>(lam, first, last, padded, keep);
      }
      if (act & POST_SHIFT) shift_adjoint<SL, L, TM>(lam, first, last, padded, keep);
      // the sample reads a0 (and its tangents) of the state after the pulse
      if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs && first && live) {
        const int64_t o = ev.out;
        float ws = p.w_signal ? p.w_signal[atom * p.outputs + o] : 0.0f;
        const float* wj = p.w_jac ? p.w_jac + atom * (int64_t)p.columns * p.outputs + o : nullptr;
        float wm0 = 0.0f;
        if (wj && p.col_m0 >= 0) wm0 = wj[(int64_t)p.col_m0 * p.outputs];
        lam.a[0] += m0 * ws + wm0;
        if constexpr ((TM & TAN_T1) != 0) {
          if (wj && p.col_t1 >= 0) lam.ta[ST1][0] += m0 * wj[(int64_t)p.col_t1 * p.outputs];
        }
        if constexpr ((TM & TAN_T2) != 0) {
          if (wj && p.col_t2 >= 0) lam.ta[ST2][0] += m0 * wj[(int64_t)p.col_t2 * p.outputs];
        }
      }
      // state before the event, then the state the pulse acts on
      D xb;
      xb.load(p.scratch + (e - e0) * D::floats * T + tid, T);
      if (ev.dt != 0.0f) fr.set(ev.dt, r1, r2, k1, k2);
      D x2 = xb;
      if (ev.dt != 0.0f) relax_forward<SL, L, TM>(x2, fr, first);
      if (act & PRE_SHIFT) shift_forward<SL, L, TM>(x2, first, last, padded, keep);
      if (kind == KIND_RF) {
        if (act & INVERSION) {
#pragma unroll
          for (int r = 0; r < SL; ++r) {
            lam.c[r] *= -eta;
#pragma unroll
            for (int k = 0; k < NT; ++k) lam.tc[k][r] *= -eta;
          }
        } else {
          Rotation t;
          if (transmit) {
            const float s = sinf(ev.flip * b1), h = sinf(0.5f * (ev.flip * b1));
            t.set(2.0f * (h * h), s);
          } else {
            t.set(ev.versine, ev.sine);
          }
          float g = rotate3_adjoint<SL>(lam.a, lam.b, lam.c, x2.a, x2.b, x2.c, t);
#pragma unroll
          for (int k = 0; k < NT; ++k) g += rotate3_adjoint<SL>(lam.ta[k], lam.tb[k], lam.tc[k], x2.ta[k], x2.tb[k], x2.tc[k], t);
          g_flip = g * b1;
        }
      }
      if (act & PRE_SHIFT) shift_adjoint<SL, L, TM>(lam, first, last, padded, keep);
      // relaxation over the interval before the event: the duration's
      // derivative, then lambda <- R^T lambda
      {
        const float dt = ev.dt;
        float e1v = 1.0f, e2v = 1.0f, d1 = 0.0f, d2 = 0.0f;
        if (dt != 0.0f) {
          e1v = fr.e1;
          e2v = fr.e2;
          d1 = fr.de1;
          d2 = fr.de2;
        }
        const float dtr = -r2 * e2v, dlo = -r1 * e1v;
        const float dde1 = k1 * e1v * (1.0f - r1 * dt);
        const float dde2 = k2 * e2v * (1.0f - r2 * dt);
        float g = 0.0f;
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          g += dtr * (lam.a[r] * xb.a[r] + lam.b[r] * xb.b[r]) + dlo * lam.c[r] * xb.c[r];
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            g += dtr * (lam.ta[k][r] * xb.ta[k][r] + lam.tb[k][r] * xb.tb[k][r]) + dlo * lam.tc[k][r] * xb.tc[k][r];
            if ((TM & TAN_T1) != 0 && k == ST1) g += dde1 * lam.tc[k][r] * xb.c[r];
            if ((TM & TAN_T2) != 0 && k == ST2) g += dde2 * (lam.ta[k][r] * xb.a[r] + lam.tb[k][r] * xb.b[r]);
          }
        }
        if (first) {
          g += lam.c[0] * r1 * e1v;
          if constexpr ((TM & TAN_T1) != 0) g -= lam.tc[ST1][0] * dde1;
        }
        g_dt = g;
        if (dt != 0.0f) {
#pragma unroll
          for (int r = 0; r < SL; ++r) {
            float na = e2v * lam.a[r], nb = e2v * lam.b[r], nc = e1v * lam.c[r];
            if constexpr ((TM & TAN_T2) != 0) {
              na += d2 * lam.ta[ST2][r];
              nb += d2 * lam.tb[ST2][r];
            }
            if constexpr ((TM & TAN_T1) != 0) nc += d1 * lam.tc[ST1][r];
            lam.a[r] = na;
            lam.b[r] = nb;
            lam.c[r] = nc;
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              lam.ta[k][r] *= e2v;
              lam.tb[k][r] *= e2v;
              lam.tc[k][r] *= e1v;
            }
          }
        }
      }
      if (!live) {
        g_flip = 0.0f;
        g_dt = 0.0f;
      }
      g_flip = wave_sum(g_flip);
      g_dt = wave_sum(g_dt);
      if (lane == 0) {
        float* out = p.partial + (wave * E + e) * 2;
        out[0] = g_flip;
        out[1] = g_dt;
      }
    }
  }
}

__global__ void reduce_partials(const float* __restrict__ partial, int64_t waves, int64_t events,
                                float* __restrict__ grad_flip, float* __restrict__ grad_duration) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= events) return;
  float gf = 0.0f, gd = 0.0f;
  for (int64_t w = 0; w < waves; ++w) {
    gf += partial[(w * events + e) * 2];
    gd += partial[(w * events + e) * 2 + 1];
  }
  if (grad_flip) grad_flip[e] = gf;
  if (grad_duration) grad_duration[e] = gd;
}

using AdjointLaunch = void (*)(const AdjointArgs&, dim3, dim3, hipStream_t);

template <int SL, int L, int TM>
void launch_adjoint(const AdjointArgs& a, dim3 g, dim3 b, hipStream_t s) {
  hipLaunchKernelGGL((adjoint_kernel<SL, L, TM>), g, b, 0, s, a);
}

template <int SL, int L>
AdjointLaunch pick_adjoint_mask(int tm) {
  switch (tm) {
    case 0: return &launch_adjoint<SL, L, 0>;
    case 1: return &launch_adjoint<SL, L, 1>;
    case 2: return &launch_adjoint<SL, L, 2>;
    default: return &launch_adjoint<SL, L, 3>;
  }
}

struct AdjointTile {
  int sl, l;
};
constexpr AdjointTile kAdjointTiles[] = {{2, 2}, {4, 2}, {5, 2}, {8, 2}, {5, 4}, {8, 4}, {16, 4}, {3, 4}, {2, 8}};

AdjointLaunch pick_adjoint(int tile, int tm) {
  switch (tile) {
    case 0: return pick_adjoint_mask<2, 2>(tm);
    case 1: return pick_adjoint_mask<4, 2>(tm);
    case 2: return pick_adjoint_mask<5, 2>(tm);
    case 3: return pick_adjoint_mask<8, 2>(tm);
    case 4: return pick_adjoint_mask<5, 4>(tm);
    case 5: return pick_adjoint_mask<8, 4>(tm);
    case 6: return pick_adjoint_mask<16, 4>(tm);
    case 7: return pick_adjoint_mask<3, 4>(tm);
    default: return pick_adjoint_mask<2, 8>(tm);
  }
}

int best_cover_of(int t) { return kAdjointTiles[t].sl * kAdjointTiles[t].l; }

// With enough atoms to fill the chip the smallest covering tile wins (least
// padding). With few atoms the adjoint is a latency chain per atom -- every
// event's recompute and reverse step walk the lane's SL orders in sequence --
// so the tile with the fewest orders per lane among those padding at most 2x
// wins instead (more lanes per atom, same work, a shorter chain).
int adjoint_tile(int states, int64_t atoms) {
  constexpr int kTiles = (int)(sizeof(kAdjointTiles) / sizeof(kAdjointTiles[0]));
  constexpr int64_t kFillLanes = 1024 * 64 * 2;  // two waves on every SIMD
  int smallest = -1, smallest_cover = 1 << 30;
  for (int t = 0; t < kTiles; ++t) {
    const int cover = kAdjointTiles[t].sl * kAdjointTiles[t].l;
    if (cover >= states && cover < smallest_cover) {
      smallest = t;
      smallest_cover = cover;
    }
  }
  if (smallest < 0 || atoms * kAdjointTiles[smallest].l >= kFillLanes) return smallest;
  int best = smallest;
  for (int t = 0; t < kTiles; ++t) {
    const int cover = kAdjointTiles[t].sl * kAdjointTiles[t].l;
    if (cover < states || cover > 2 * smallest_cover) continue;
    const bool shorter = kAdjointTiles[t].sl < kAdjointTiles[best].sl;
    const bool tighter = kAdjointTiles[t].sl == kAdjointTiles[best].sl && cover < best_cover_of(best);
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
                                          const float* jacobian_cotangent, int64_t segment_events,
                                          float* grad_flip, float* grad_duration, void* stream) {
  if (packed == nullptr || packed[T_T1] == nullptr || packed[T_T2] == nullptr) {
    set_error("packed buffers with t1 and t2 are required");
    return BAD_ARGUMENT;
  }
  if (atom_count < 0 || event_count < 0 || state_count < 2 || output_count < 0) {
    set_error("invalid dimensions (the adjoint needs at least 2 orders)");
    return BAD_ARGUMENT;
  }
  if (segment_events < 1) {
    set_error("segment_events must be positive");
    return BAD_ARGUMENT;
  }
  if (column_count < 0 || column_count > 3 || (column_count > 0 && columns == nullptr)) {
    set_error("between 0 and 3 Jacobian columns (t1, t2, m0), got %lld", (long long)column_count);
    return BAD_ARGUMENT;
  }
  if (features & (FEATURE_DIFFUSING | FEATURE_OFF_AXIS | FEATURE_MOVING)) {
    set_error("the adjoint carries no diffusion, off-resonance or flow term");
    return UNSUPPORTED;
  }
  AdjointArgs a;
  memset(&a, 0, sizeof(a));
  int tm = 0;
  a.col_t1 = a.col_t2 = a.col_m0 = -1;
  for (int64_t c = 0; c < column_count; ++c) {
    switch (columns[c]) {
      case T_T1: tm |= TAN_T1; a.col_t1 = (int)c; break;
      case T_T2: tm |= TAN_T2; a.col_t2 = (int)c; break;
      case T_M0: a.col_m0 = (int)c; break;
      default:
        set_error("column %lld: the adjoint carries tangents along t1, t2 and m0 only", (long long)c);
        return BAD_ARGUMENT;
    }
  }
  const int tile = adjoint_tile((int)state_count, atom_count);
  if (tile < 0) {
    set_error("no adjoint tile covers %lld orders", (long long)state_count);
    return UNSUPPORTED;
  }
  hipStream_t s = (hipStream_t)stream;
  if (atom_count == 0 || event_count == 0) {
    // no atom contributes: the reduction over zero waves writes zeros
    if (event_count > 0) {
      const int64_t rb = (event_count + 255) / 256;
      hipLaunchKernelGGL(reduce_partials, dim3((unsigned)rb), dim3(256), 0, s, nullptr, (int64_t)0, event_count,
                         grad_flip, grad_duration);
    }
    return hip_status(hipGetLastError(), "adjoint clear");
  }
  const int L = kAdjointTiles[tile].l;
  const int SL = kAdjointTiles[tile].sl;
  const int NT = popcount4(tm);
  const int64_t nfloat = 3LL * SL * (1 + NT);
  // Few atoms leave most compute units idle with 4-wave blocks (1,024 atoms x
  // 8 lanes fill 32 of 256): one-wave blocks spread the same waves over four
  // times as many units, each with its own caches and memory path. Every index below
  // derives from the global thread id and the launch total, not the block.
  const int threads = atom_count * L < 256LL * 256 ? 64 : 256;
  const int64_t blocks = (atom_count * L + threads - 1) / threads;
  const int64_t total = blocks * threads;
  const int64_t waves = total / 64;
  a.event_count = event_count;
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
  a.segment = segment_events;
  a.segments = (event_count + segment_events - 1) / segment_events;
  a.threads_total = total;
  const size_t ck = sizeof(float) * (size_t)(a.segments * nfloat * total);
  const size_t sc = sizeof(float) * (size_t)(segment_events * nfloat * total);
  const size_t pa = sizeof(float) * (size_t)(waves * event_count * 2);
  char* work = nullptr;
  int st = scratch_alloc((void**)&work, ck + sc + pa, s, "adjoint workspace");
  if (st != OK) return st;
  a.checkpoints = (float*)work;
  a.scratch = (float*)(work + ck);
  a.partial = (float*)(work + ck + sc);
  Compiled compiled;
  st = compile_packed(packed, event_count, s, compiled);
  if (st == OK) {
    a.events = compiled.codes;
    pick_adjoint(tile, tm)(a, dim3((unsigned)blocks), dim3(threads), s);
    st = hip_status(hipGetLastError(), "adjoint_kernel launch");
  }
  if (st == OK) {
    const int64_t rb = (event_count + 255) / 256;
    hipLaunchKernelGGL(reduce_partials, dim3((unsigned)rb), dim3(256), 0, s, a.partial, waves, event_count,
                       grad_flip, grad_duration);
    st = hip_status(hipGetLastError(), "reduce_partials launch");
  }
  (void)hipFreeAsync(work, s);
  return st;
}
