// Forward-over-reverse of the complex EPG kernel, for gfx950: the gradient,
// with respect to every event's flip angle, phase and duration, of a loss
// that reads the Jacobian columns (T1, T2 or B1) of the complex signal. It
// computes the event part of the reference's simulate_vjp_jvp
// (src/torchsim/_epg_cpu.cpp:7759-9013) that `_NativeEpgJvp.backward`
// reaches (sequence/_accelerators.py:1787-1850) when a design loss takes the
// Jacobian of a phase-cycled, off-resonant, exchanging or diffusing protocol.
//
// One launch per tangent direction th. The forward carries the state x and
// its tangent t = dx/dth (t' = A t + A_th x + b_th); the reverse carries
// their adjoints: mu, driven by the cotangent W on the Jacobian column the
// tangent records, with mu <- A^H mu, and lambda, which the tangent's coupling
// feeds, lambda <- A^H lambda + A_th^H mu. An event parameter q gets
//   <lambda', A_q x> + <mu', A_q t + A_{th q} x>
// (plus the recovery terms at order 0). The signal's own cotangent is the
// first-order adjoint's (tsim_complex_vjp), run beside these launches; an M0
// column is the signal over M0, so its cotangent joins the signal's.
//
// Mixed derivatives. Along an interval dt every operator is exp(G dt) with G
// its generator, so d/d(dt) of an operator is G times it and the mixed
// derivative along th is G_th A + G A_th: single pool, F+- carry
// tp k2 (1 + dt c) for T2 (c the transverse rate of the order, complex with
// the off-resonance), Z carries lo k1 (1 + dt cz) for T1 and the recovery
// -k1 e1 (1 - r1 dt); exchanging pool, (cl + M) dX/dT2 + E X and
// (M_l - wl D) dY/dT1 + E_l Y with E the rate matrix's derivative, and the
// recovery's E Y u + M_l dY u + M_l Y du. A pulse's operator depends on B1
// through the angle flip * b1: its mixed derivatives are T_a + flip b1 T_aa
// along the flip and flip T_a,phi along the phase.
#pragma once

#include "epg_complex_adjoint_impl.h"

namespace tsim {

struct ComplexFovrArgs {
  ComplexAdjointArgs base;           // events, tissue, sizes, features, factor, workspace, partials
  int tangent;                       // CG_T1, CG_T2 or CG_B1
  const float2* jacobian_cotangent;  // (atoms, columns, outputs)
  int64_t column;                    // this tangent's column
  int64_t columns;
};

namespace cadj {

// The exchanging pool's equilibrium split u = G^{-1} s of the longitudinal
// generator and its derivative along T1 (constants of the atom).
struct PoolSplit {
  float u0 = 0.0f, u1 = 0.0f, du0 = 0.0f, du1 = 0.0f;
  __device__ __forceinline__ void set(const Atom& at) {
    const float g00 = -at.kab - at.r1, g01 = at.kba, g10 = at.kab, g11 = -at.kba - at.r1b;
    const float det = g00 * g11 - g01 * g10;
    const float s0 = (1.0f - at.frac) * at.r1, s1 = at.frac * at.r1b;
    u0 = (g11 * s0 - g01 * s1) / det;
    u1 = (-g10 * s0 + g00 * s1) / det;
    const float w0 = -(1.0f - at.frac) * at.k1 - at.k1 * u0;
    du0 = (g11 * w0) / det;
    du1 = (-g10 * w0) / det;
  }
};

// One event forward with the tangent along direction d: relaxation, shifts,
// pulse or inversion (a sample records nothing here).
template <int L, bool EXCH>
__device__ __forceinline__ void forward_event_t(State<EXCH>& x, State<EXCH>& t, const EventCode& ev,
                                                Interval<EXCH, true>& I, const Atom& at, const Direction& d,
                                                const ComplexAdjointArgs& p, int sub, int lane, bool beyond) {
  const uint32_t act = ev.ctrl & 0xFFu;
  const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
  if (ev.dt != 0.0f) {
    I.set(ev.dt, at);
    relax_jvp<EXCH, true>(x, t, I, at, d, ev.dt, 0.0f, sub == 0);
  }
  if (act & PRE_SHIFT) {
    shift_all<L, EXCH>(x, sub, lane, beyond);
    shift_all<L, EXCH>(t, sub, lane, beyond);
  }
  if (kind == KIND_RF) {
    if (act & INVERSION) {
#pragma unroll
      for (int q = 0; q < State<EXCH>::NP; ++q) {
        t.z[q] = scl(t.z[q], -at.eta);
        x.z[q] = scl(x.z[q], -at.eta);
      }
    } else {
      float sa, va;
      pulse_angle(ev, at, p.transmit, sa, va);
      Rotation Tm;
      Tm.set(sa, va, ev.phase + at.b1_phase);
      const float da = ev.flip * d.b1;
      if (da != 0.0f) {
        const float ca = Tm.ca, dch = -0.5f * sa, dca = -sa;
        const float2 e1p = Tm.e1p;
        const float2 d01 = scl(mul(e1p, e1p), 0.5f * sa);
        const float2 d02 = mul(mk(0.0f, -ca), e1p);
        const float2 d10 = cj(d01);
        const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
        const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
        const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
#pragma unroll
        for (int q = 0; q < State<EXCH>::NP; ++q) {
          const float2 fp = x.fp[q], fm = x.fm[q], z = x.z[q];
          const float2 ap = mac(mac(scl(fp, dch), d01, fm), d02, z);
          const float2 am = mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z);
          const float2 az = mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z);
          const float2 tp = t.fp[q], tm = t.fm[q], tz = t.z[q];
          t.fp[q] = add(mac(mac(scl(tp, Tm.chh), Tm.t01, tm), Tm.t02, tz), scl(ap, da));
          t.fm[q] = add(mac(mac(mul(Tm.t10, tp), mk(Tm.chh, 0.0f), tm), Tm.t12, tz), scl(am, da));
          t.z[q] = add(mac(mac(mul(Tm.t20, tp), Tm.t21, tm), mk(ca, 0.0f), tz), scl(az, da));
        }
      } else {
        rotate<EXCH>(t, Tm);
      }
      rotate<EXCH>(x, Tm);
    }
  }
  if (act & POST_SHIFT) {
    shift_all<L, EXCH>(x, sub, lane, beyond);
    shift_all<L, EXCH>(t, sub, lane, beyond);
  }
  if (act & SPOIL_AFTER) {
#pragma unroll
    for (int q = 0; q < State<EXCH>::NP; ++q) x.fp[q] = x.fm[q] = t.fp[q] = t.fm[q] = mk(0.0f, 0.0f);
  } else if (act & SHIFT_AFTER) {
    shift_all<L, EXCH>(x, sub, lane, beyond);
    shift_all<L, EXCH>(t, sub, lane, beyond);
  }
}

// The reverse of one relaxation over interval I (dt) for the pair (x, t):
// returns the gradient along the interval's duration,
//   <lam', G x'> + <mu', G t' + G_th x'> (as operators applied to the entry
// states, see the header), and (TIMED) applies lam <- A^H lam + A_th^H mu,
// mu <- A^H mu. At dt = 0 the operators are the identity and only the
// duration gradient remains.
template <bool EXCH, bool TIMED>
__device__ __forceinline__ float relax_reverse_fovr(const Interval<EXCH, true>& I, State<EXCH>& lam, State<EXCH>& mu,
                                                    const State<EXCH>& xe, const State<EXCH>& te, const Atom& at,
                                                    const PoolSplit& ps, bool first, float dt, int th) {
  float unused[CG_COUNT];
  float mixed = 0.0f;
  State<EXCH> extra;  // A_th^H mu'
  extra.zero();
  if constexpr (!EXCH) {
    if (th == CG_T2) {
      const float2 tp = scl(I.turn, I.e2), tm = scl(cj(I.turn), I.e2);
      const float2 P = mul(cmul_conj(mu.fp[0], tp), xe.fp[0]);
      const float2 M = mul(cmul_conj(mu.fm[0], tm), xe.fm[0]);
      const float cr = 1.0f - dt * (at.r2 + at.wt * at.drate);
      const float ci = -kTwoPi * at.b0 * dt;
      // Re((1 + dt c+) P) + Re((1 + dt c-) M), c- the conjugate turn's rate
      mixed = at.k2 * (cr * P.x - ci * P.y + cr * M.x + ci * M.y);
      if constexpr (TIMED) {
        extra.fp[0] = scl(cmul_conj(tp, mu.fp[0]), at.k2 * dt);
        extra.fm[0] = scl(cmul_conj(tm, mu.fm[0]), at.k2 * dt);
      }
    } else if (th == CG_T1) {
      const float lo = I.e1 * I.dlong;
      const float Q = lo * dotr(mu.z[0], xe.z[0]);
      mixed = at.k1 * (1.0f - dt * (at.r1 + at.wl * at.drate)) * Q;
      if (first) mixed -= mu.z[0].x * at.k1 * I.e1 * (1.0f - at.r1 * dt);
      if constexpr (TIMED) extra.z[0] = scl(mu.z[0], lo * at.k1 * dt);
    }
  } else {
    if (th == CG_T2) {
      const float2 cl = mk(-at.wt * at.drate, -kTwoPi * at.b0);
      const float2 m[4] = {mk(-at.kab - at.r2, 0.0f), mk(at.kba, 0.0f), mk(at.kab, 0.0f),
                           mk(-at.kba - at.r2b, -kTwoPi * at.shift_hz)};
      float2 acc = mk(0.0f, 0.0f);
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float2 V = add(cmul_conj(mu.fp[i], xe.fp[j]), cmul_conj(xe.fm[j], mu.fm[i]));
          float2 D = add(mul(cl, I.dX2[2 * i + j]),
                         add(mul(m[2 * i], I.dX2[j]), mul(m[2 * i + 1], I.dX2[2 + j])));
          if (i == 0) D = add(D, scl(I.X[j], at.k2));
          acc = mac(acc, V, D);
        }
      }
      mixed = acc.x;
      if constexpr (TIMED) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          extra.fp[j] = add(cmul_conj(I.dX2[j], mu.fp[0]), cmul_conj(I.dX2[2 + j], mu.fp[1]));
          extra.fm[j] = add(mul(I.dX2[j], mu.fm[0]), mul(I.dX2[2 + j], mu.fm[1]));
        }
      }
    } else if (th == CG_T1) {
      const float m[4] = {-at.kab - at.r1, at.kba, at.kab, -at.kba - at.r1b};
      const float wd = -at.wl * at.drate;
      float acc = 0.0f;
#pragma unroll
      for (int i = 0; i < 2; ++i) {
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const float U = dotr(mu.z[i], xe.z[j]);
          float D = wd * I.dY1[2 * i + j] + m[2 * i] * I.dY1[j] + m[2 * i + 1] * I.dY1[2 + j];
          if (i == 0) D += at.k1 * I.Y[j];
          acc = fmaf(U, D, acc);
        }
      }
      if (first) {
        // the recovery's dt derivative M_l Y u, along T1 (order 0: Y is the
        // bare exponential)
        const float yu0 = I.Y[0] * ps.u0 + I.Y[1] * ps.u1;
        const float dy0 = I.dY1[0] * ps.u0 + I.dY1[1] * ps.u1 + I.Y[0] * ps.du0 + I.Y[1] * ps.du1;
        const float dy1 = I.dY1[2] * ps.u0 + I.dY1[3] * ps.u1 + I.Y[2] * ps.du0 + I.Y[3] * ps.du1;
        const float r0 = at.k1 * yu0 + m[0] * dy0 + m[1] * dy1;
        const float r1 = m[2] * dy0 + m[3] * dy1;
        acc += mu.z[0].x * r0 + mu.z[1].x * r1;
      }
      mixed = acc;
      if constexpr (TIMED) {
#pragma unroll
        for (int j = 0; j < 2; ++j) extra.z[j] = add(scl(mu.z[0], I.dY1[j]), scl(mu.z[1], I.dY1[2 + j]));
      }
    }
  }
  // the tangent's recursion has no recovery of its own (its b_th is in the
  // mixed part), so mu pairs with the operators as at an order above 0
  const float g = relax_reverse<EXCH, TIMED>(I, lam, xe, at, first, dt, unused) +
                  relax_reverse<EXCH, TIMED>(I, mu, te, at, false, dt, unused) + mixed;
  if constexpr (TIMED) {
#pragma unroll
    for (int q = 0; q < State<EXCH>::NP; ++q) {
      lam.fp[q] = add(lam.fp[q], extra.fp[q]);
      lam.fm[q] = add(lam.fm[q], extra.fm[q]);
      lam.z[q] = add(lam.z[q], extra.z[q]);
    }
  }
  return g;
}

}  // namespace cadj

template <int L, bool EXCH>
__global__ void __launch_bounds__(256) complex_fovr_kernel(const ComplexFovrArgs q) {
  using namespace cadj;
  using S = State<EXCH>;
  constexpr int NP = S::NP;
  constexpr int APW = 64 / L;
  const ComplexAdjointArgs& p = q.base;
  extern __shared__ float f_part_all[];  // per wave: [3 * CFLUSH rows][64 lanes]
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave = tid >> 6;
  const int64_t group = wave * APW + wave_group;
  const bool live = wave_group < APW && group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const bool beyond = sub >= p.states;
  const bool first = sub == 0;
  const int64_t ia = atom * p.stride_opt;
  float* const s_part = f_part_all + (threadIdx.x >> 6) * (3 * CFLUSH * 64);
  const int th = q.tangent;

  Atom at;
  at.t1 = p.t1[atom];
  at.t2 = p.t2[atom];
  at.r1 = 1000.0f / at.t1;
  at.r2 = 1000.0f / at.t2;
  at.k1 = 1000.0f / (at.t1 * at.t1);
  at.k2 = 1000.0f / (at.t2 * at.t2);
  at.m0 = p.density ? p.m0[ia] : 1.0f;
  at.b1 = p.transmit ? p.b1[ia] : 1.0f;
  at.b1_phase = p.off_axis && p.b1_phase ? p.b1_phase[ia] : 0.0f;
  at.b0 = p.off_axis && p.b0 ? p.b0[ia] : 0.0f;
  at.eta = p.inverting ? p.inv_eff[ia] : 1.0f;
  at.drate = p.diffusing ? p.diffusion[ia] : 0.0f;
  at.frac = at.kab = at.kba = at.r1b = at.r2b = at.shift_hz = 0.0f;
  PoolSplit ps;
  if constexpr (EXCH) {
    at.frac = p.pool_fraction[ia];
    const float kex = p.pool_exchange[ia];
    at.kab = kex * at.frac;
    at.kba = kex * (1.0f - at.frac);
    at.r1b = 1000.0f / p.t1_pool[ia];
    at.r2b = 1000.0f / p.t2_pool[ia];
    at.shift_hz = p.pool_shift ? p.pool_shift[ia] : 0.0f;
    ps.set(at);
  }
  const float order = (float)sub;
  at.wt = order * order + order + (1.0f / 3.0f);
  at.wl = order * order;
  Direction d;
  d.t1 = th == CG_T1 ? 1.0f : 0.0f;
  d.t2 = th == CG_T2 ? 1.0f : 0.0f;
  d.b1 = th == CG_B1 ? 1.0f : 0.0f;
  d.m0 = d.b1_phase = d.b0 = d.eta = d.drate = 0.0f;

  const int64_t E = p.event_count;
  const int64_t K = p.segment;
  const int64_t segments = (E + K - 1) / K;
  const int64_t T = p.threads_total;
  constexpr int64_t F2 = 2 * S::floats;  // floats of (x, t)
  const float2 factor = mk(p.factor_re, p.factor_im);
  const float2* w_row = q.jacobian_cotangent + (atom * q.columns + q.column) * p.outputs;

  // Pass 1: forward, one checkpoint of (x, t) per segment.
  S x, t;
  x.zero();
  t.zero();
  if (first) {
    x.z[0] = mk(EXCH ? 1.0f - at.frac : 1.0f, 0.0f);
    if constexpr (EXCH) x.z[1] = mk(at.frac, 0.0f);
  }
  Interval<EXCH, true> If;
  for (int64_t e = 0; e < E; ++e) {
    if (e % K == 0) {
      float* ck = p.checkpoints + (e / K) * F2 * T + tid;
      x.store(ck, T);
      t.store(ck + S::floats * T, T);
    }
    const EventCode ev = load_event(p.events, e);
    forward_event_t<L, EXCH>(x, t, ev, If, at, d, p, sub, lane, beyond);
  }

  // Pass 2: reverse, one segment at a time.
  S lam, mu;
  lam.zero();
  mu.zero();
  Interval<EXCH, true> Ir;
  Interval<EXCH, true> I0;
  I0.set(0.0f, at);
  int rows = 0;
  int64_t flush_top = 0;
  auto flush = [&]() __attribute__((always_inline)) {
    __syncthreads();
    if (lane < 3 * rows) {
      const float* src = s_part + lane * 64;
      float v = 0.0f;
#pragma unroll 8
      for (int j = 0; j < 64; ++j) v += src[j];
      const int64_t e = flush_top - lane / 3;
      p.partial[(wave * E + e) * 3 + lane % 3] = v;
    }
    __syncthreads();
    rows = 0;
  };

  for (int64_t seg = segments - 1; seg >= 0; --seg) {
    const int64_t e0 = seg * K;
    const int64_t e1 = e0 + K < E ? e0 + K : E;
    float* const states = p.scratch + tid;
    const float* ck = p.checkpoints + seg * F2 * T + tid;
    x.load(ck, T);
    t.load(ck + S::floats * T, T);
    for (int64_t e = e0; e < e1; ++e) {
      float* row = states + (e - e0) * F2 * T;
      x.store(row, T);
      t.store(row + S::floats * T, T);
      const EventCode ev = load_event(p.events, e);
      forward_event_t<L, EXCH>(x, t, ev, If, at, d, p, sub, lane, beyond);
    }
    for (int64_t e = e1 - 1; e >= e0; --e) {
      S xe, te;
      const float* row = states + (e - e0) * F2 * T;
      xe.load(row, T);
      te.load(row + S::floats * T, T);
      const EventCode ev = load_event(p.events, e);
      const uint32_t act = ev.ctrl & 0xFFu;
      const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
      if (rows == 0) flush_top = e;
      float g_flip = 0.0f, g_phase = 0.0f, g_dt = 0.0f;
      const bool timed = ev.dt != 0.0f;
      if (timed) Ir.set(ev.dt, at);
      const bool record = kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs;
      const bool pulse = kind == KIND_RF && !(act & INVERSION);
      const bool invert = kind == KIND_RF && (act & INVERSION);

      // the states the pulse or sample acts on
      S xr = xe, tr = te;
      if (pulse || invert || record) {
        if (timed) relax_jvp<EXCH, true>(xr, tr, Ir, at, d, ev.dt, 0.0f, first);
        if (act & PRE_SHIFT) {
          shift_all<L, EXCH>(xr, sub, lane, beyond);
          shift_all<L, EXCH>(tr, sub, lane, beyond);
        }
      }

      if (act & SPOIL_AFTER) {
#pragma unroll
        for (int k = 0; k < NP; ++k) lam.fp[k] = lam.fm[k] = mu.fp[k] = mu.fm[k] = mk(0.0f, 0.0f);
      } else if (act & SHIFT_AFTER) {
        shift_all_t<L, EXCH>(lam, sub, lane, p.states);
        shift_all_t<L, EXCH>(mu, sub, lane, p.states);
      }
      if (act & POST_SHIFT) {
        shift_all_t<L, EXCH>(lam, sub, lane, p.states);
        shift_all_t<L, EXCH>(mu, sub, lane, p.states);
      }

      if (pulse) {
        float sa, va;
        pulse_angle(ev, at, p.transmit, sa, va);
        Rotation Tm;
        Tm.set(sa, va, ev.phase + at.b1_phase);
        const float ca = Tm.ca, dch = -0.5f * sa, dca = -sa;
        const float2 e1p = Tm.e1p;
        const float2 d01 = scl(mul(e1p, e1p), 0.5f * sa);
        const float2 d02 = mul(mk(0.0f, -ca), e1p);
        const float2 d10 = cj(d01);
        const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
        const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
        const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
        // second derivatives along the angle
        const float ddch = -0.5f * ca, ddca = -ca;
        const float2 dd01 = scl(mul(e1p, e1p), 0.5f * ca);
        const float2 dd02 = mul(mk(0.0f, sa), e1p);
        const float2 dd10 = cj(dd01);
        const float2 dd12 = mul(mk(0.0f, -sa), cj(e1p));
        const float2 dd20 = mul(mk(0.0f, 0.5f * sa), cj(e1p));
        const float2 dd21 = mul(mk(0.0f, -0.5f * sa), e1p);
        const bool coupled = th == CG_B1;
        const float fb = ev.flip * at.b1;
        float ga = 0.0f, gm = 0.0f;  // per unit angle: the pairs' first-order part; the mixed part along the flip
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const float2 fp = xr.fp[k], fm = xr.fm[k], z = xr.z[k];
          const float2 up = tr.fp[k], um = tr.fm[k], uz = tr.z[k];
          const float2 lp = lam.fp[k], lm = lam.fm[k], lz = lam.z[k];
          const float2 mp = mu.fp[k], mm = mu.fm[k], mz = mu.z[k];
          // T_a x and T_a t
          const float2 ap = mac(mac(scl(fp, dch), d01, fm), d02, z);
          const float2 am = mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z);
          const float2 az = mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z);
          const float2 bp = mac(mac(scl(up, dch), d01, um), d02, uz);
          const float2 bm = mac(mac(mul(d10, up), mk(dch, 0.0f), um), d12, uz);
          const float2 bz = mac(mac(mul(d20, up), d21, um), mk(dca, 0.0f), uz);
          ga += dotr(lp, ap) + dotr(lm, am) + dotr(lz, az) + dotr(mp, bp) + dotr(mm, bm) + dotr(mz, bz);
          // along the phase: i (2 t01 fm + t02 z), -i (2 t10 fp + t12 z), i (t21 fm - t20 fp)
          const float2 qp = add(scl(mul(Tm.t01, fm), 2.0f), mul(Tm.t02, z));
          const float2 qm = neg(add(scl(mul(Tm.t10, fp), 2.0f), mul(Tm.t12, z)));
          const float2 qz = sub_of(mul(Tm.t21, fm), mul(Tm.t20, fp));
          const float2 rp = add(scl(mul(Tm.t01, um), 2.0f), mul(Tm.t02, uz));
          const float2 rm = neg(add(scl(mul(Tm.t10, up), 2.0f), mul(Tm.t12, uz)));
          const float2 rz = sub_of(mul(Tm.t21, um), mul(Tm.t20, up));
          g_phase -= cmul_conj(lp, qp).y + cmul_conj(lm, qm).y + cmul_conj(lz, qz).y + cmul_conj(mp, rp).y +
                      cmul_conj(mm, rm).y + cmul_conj(mz, rz).y;
          if (coupled) {
            // along the flip: <mu', (T_a + flip b1 T_aa) x>
            const float2 cp = mac(mac(scl(fp, ddch), dd01, fm), dd02, z);
            const float2 cm = mac(mac(mul(dd10, fp), mk(ddch, 0.0f), fm), dd12, z);
            const float2 cz = mac(mac(mul(dd20, fp), dd21, fm), mk(ddca, 0.0f), z);
            gm += dotr(mp, add(ap, scl(cp, fb))) + dotr(mm, add(am, scl(cm, fb))) + dotr(mz, add(az, scl(cz, fb)));
            // along the phase: flip <mu', T_a,phi x>, T_a,phi x = i (2 d01 fm + d02 z),
            // -i (2 d10 fp + d12 z), i (d21 fm - d20 fp)
            const float2 sp = add(scl(mul(d01, fm), 2.0f), mul(d02, z));
            const float2 sm = neg(add(scl(mul(d10, fp), 2.0f), mul(d12, z)));
            const float2 sz = sub_of(mul(d21, fm), mul(d20, fp));
            g_phase -= ev.flip * (cmul_conj(mp, sp).y + cmul_conj(mm, sm).y + cmul_conj(mz, sz).y);
          }
          // lam <- T^H lam (+ flip T_a^H mu), mu <- T^H mu
          float2 nlp = add(add(scl(lp, Tm.chh), cmul_conj(Tm.t10, lm)), cmul_conj(Tm.t20, lz));
          float2 nlm = add(add(cmul_conj(Tm.t01, lp), scl(lm, Tm.chh)), cmul_conj(Tm.t21, lz));
          float2 nlz = add(add(cmul_conj(Tm.t02, lp), cmul_conj(Tm.t12, lm)), scl(lz, ca));
          if (coupled) {
            nlp = add(nlp, scl(add(add(scl(mp, dch), cmul_conj(d10, mm)), cmul_conj(d20, mz)), ev.flip));
            nlm = add(nlm, scl(add(add(cmul_conj(d01, mp), scl(mm, dch)), cmul_conj(d21, mz)), ev.flip));
            nlz = add(nlz, scl(add(add(cmul_conj(d02, mp), cmul_conj(d12, mm)), scl(mz, dca)), ev.flip));
          }
          lam.fp[k] = nlp;
          lam.fm[k] = nlm;
          lam.z[k] = nlz;
          mu.fp[k] = add(add(scl(mp, Tm.chh), cmul_conj(Tm.t10, mm)), cmul_conj(Tm.t20, mz));
          mu.fm[k] = add(add(cmul_conj(Tm.t01, mp), scl(mm, Tm.chh)), cmul_conj(Tm.t21, mz));
          mu.z[k] = add(add(cmul_conj(Tm.t02, mp), cmul_conj(Tm.t12, mm)), scl(mz, ca));
        }
        g_flip = at.b1 * ga + gm;
      } else if (invert) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          lam.z[k] = scl(lam.z[k], -at.eta);
          mu.z[k] = scl(mu.z[k], -at.eta);
        }
      } else if (record && first && live) {
        float sp, cp;
        sincosf(-ev.phase, &sp, &cp);
        const float2 w = w_row[ev.out];
        float2 total = tr.fp[0];
        if constexpr (EXCH) total = add(total, tr.fp[1]);
        const float2 c = scl(mul(factor, mk(cp, sp)), at.m0);
        const float2 lin = cmul_conj(c, w);
#pragma unroll
        for (int k = 0; k < NP; ++k) mu.fp[k] = add(mu.fp[k], lin);
        g_phase += cmul_conj(w, mul(c, total)).y;
      }

      if (act & PRE_SHIFT) {
        shift_all_t<L, EXCH>(lam, sub, lane, p.states);
        shift_all_t<L, EXCH>(mu, sub, lane, p.states);
      }

      if (timed) {
        g_dt = relax_reverse_fovr<EXCH, true>(Ir, lam, mu, xe, te, at, ps, first, ev.dt, th);
      } else {
        g_dt = relax_reverse_fovr<EXCH, false>(I0, lam, mu, xe, te, at, ps, first, 0.0f, th);
      }

      const bool counts = live && !beyond;
      s_part[(3 * rows + 0) * 64 + lane] = counts ? g_flip : 0.0f;
      s_part[(3 * rows + 1) * 64 + lane] = counts ? g_phase : 0.0f;
      s_part[(3 * rows + 2) * 64 + lane] = counts ? g_dt : 0.0f;
      ++rows;
      if (rows == CFLUSH || e == e0) flush();
    }
  }
}

using ComplexFovrLaunch = void (*)(const ComplexFovrArgs&, dim3, dim3, size_t, hipStream_t);

template <int L, bool EXCH>
void launch_complex_fovr(const ComplexFovrArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((complex_fovr_kernel<L, EXCH>), grid, block, lds, s, a);
}

#define TSIM_DECLARE_COMPLEX_FOVR(LANES) ComplexFovrLaunch pick_complex_fovr_##LANES(bool exch);
TSIM_COMPLEX_LANES(TSIM_DECLARE_COMPLEX_FOVR)
#undef TSIM_DECLARE_COMPLEX_FOVR

}  // namespace tsim
