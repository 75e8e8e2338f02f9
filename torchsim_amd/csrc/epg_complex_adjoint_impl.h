// Checkpointed first-order adjoint of the complex EPG kernel, for gfx950.
//
// Gradients of a scalar loss with respect to every event's flip angle, phase
// and duration and to each atom's T1, T2, M0, B1, transmit phase, B0,
// inversion efficiency and diffusion rate, given the loss's complex cotangent
// on the recorded samples. It computes the reference's first-order reverse
// pass of the complex state machine (src/torchsim/_epg_cpu.cpp:4839-5944
// simulate_vjp_range with real_axis -1, single pool or one chemically
// exchanging pool; the tissue reduction of :11366-11395) that
// `_NativeEpg.backward` reaches (sequence/_accelerators.py:1544-1586).
//
// Layout as the forward complex kernel: one configuration order per lane, an
// atom on L adjacent lanes. Pass 1 runs the events forward and stores the
// state every `segment` events (a checkpoint); pass 2 takes the segments last
// to first, recomputes one forward from its checkpoint storing each event's
// entry state to scratch ([event][float][thread], coalesced), and walks the
// segment backward with the adjoint state lambda (the transposed operators:
// T^H for a pulse, X^H / X^T / Y^T for the relaxation of F+ / F- / Z, the
// transposed shift with its conjugate coupling at order 0). Per-event
// gradient contributions of the lanes go to LDS and are summed CFLUSH events
// at a time; per-wave partials are reduced over waves in a fixed order by a
// second launch, so the result does not depend on scheduling.
//
// Conventions. The loss is real; the cotangent w of a complex output s is
// dL/dRe(s) + i dL/dIm(s), so dL = Re(conj(w) ds), and the adjoint of a state
// entry x is lambda with dL = Re(conj(lambda) dx).
#pragma once

#include "epg_complex_impl.h"

namespace tsim {

constexpr int CFLUSH = 16;  // events whose per-lane gradient partials LDS holds before a flush

// Tissue gradients the complex adjoint carries.
enum ComplexGradient : int { CG_T1 = 0, CG_T2, CG_M0, CG_B1, CG_B1_PHASE, CG_B0, CG_IE, CG_D, CG_COUNT };

struct ComplexAdjointArgs {
  const EventCode* events;
  int64_t event_count;
  const float* t1;
  const float* t2;
  const float* m0;
  const float* b1;
  const float* b1_phase;
  const float* b0;
  const float* inv_eff;
  const float* diffusion;
  const float* pool_fraction;
  const float* pool_exchange;
  const float* t1_pool;
  const float* t2_pool;
  const float* pool_shift;
  int64_t stride_opt;
  int64_t atoms;
  int64_t outputs;
  int states;
  int density;
  int inverting;
  int transmit;
  int off_axis;
  int diffusing;
  const float2* cotangent;  // (atoms, outputs) complex cotangent of the public output
  float factor_re;
  float factor_im;
  int64_t segment;        // events per checkpoint
  int64_t threads_total;  // threads of the launch (the scratch row stride)
  float* checkpoints;     // (segments, floats, threads_total)
  float* scratch;         // (segment, floats, threads_total)
  float* partial;         // (waves, event_count, 3): d/dflip, d/dphase, d/dduration
  float* grad[CG_COUNT];  // per-atom tissue gradients, or null
};

namespace cadj {

using namespace cx;

template <bool EXCH>
struct State {
  static constexpr int NP = EXCH ? 2 : 1;
  static constexpr int floats = 6 * NP;
  float2 fp[NP], fm[NP], z[NP];

  __device__ __forceinline__ void zero() {
#pragma unroll
    for (int q = 0; q < NP; ++q) fp[q] = fm[q] = z[q] = mk(0.0f, 0.0f);
  }
  __device__ __forceinline__ void store(float* base, int64_t stride) const {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      base[(6 * q + 0) * stride] = fp[q].x;
      base[(6 * q + 1) * stride] = fp[q].y;
      base[(6 * q + 2) * stride] = fm[q].x;
      base[(6 * q + 3) * stride] = fm[q].y;
      base[(6 * q + 4) * stride] = z[q].x;
      base[(6 * q + 5) * stride] = z[q].y;
    }
  }
  __device__ __forceinline__ void load(const float* base, int64_t stride) {
#pragma unroll
    for (int q = 0; q < NP; ++q) {
      fp[q] = mk(base[(6 * q + 0) * stride], base[(6 * q + 1) * stride]);
      fm[q] = mk(base[(6 * q + 2) * stride], base[(6 * q + 3) * stride]);
      z[q] = mk(base[(6 * q + 4) * stride], base[(6 * q + 5) * stride]);
    }
  }
};

// One atom's constants, seen from one lane (its order).
struct Atom {
  float t1, t2, r1, r2, k1, k2, m0, b1, b1_phase, b0, eta, drate;
  float frac, kab, kba, r1b, r2b, shift_hz;
  float wt, wl;  // transverse and longitudinal diffusion weights of this order: l^2 + l + 1/3, l^2
};

constexpr float kTwoPi = 6.28318530717958647692f;

// The operators of one interval on this lane, memoized on the interval's
// bits. Values: the off-resonance turn times transverse damping, the
// longitudinal damping, and the pool's relaxation (single: e1, 1 - e1, e2;
// exchanging: the 2x2 transverse X and longitudinal Y exponentials with the
// turn and damping folded in, and the recovery at order 0). With GRAD also
// what the reverse sweep needs: dX/dT2, dY/dT1 and drec/dT1 (exchanging), and
// the derivatives of X, Y and the recovery along the interval itself.
template <bool EXCH, bool GRAD>
struct Interval {
  uint32_t bits = 0xFFFFFFFFu;
  float2 turn = mk(1.0f, 0.0f);
  float dlong = 1.0f;
  float e1 = 1.0f, om1 = 0.0f, e2 = 1.0f;  // single pool
  float2 X[EXCH ? 4 : 1], dX2[EXCH && GRAD ? 4 : 1], GX[EXCH && GRAD ? 4 : 1];
  float Y[EXCH ? 4 : 1], dY1[EXCH && GRAD ? 4 : 1], GY[EXCH && GRAD ? 4 : 1];
  float rec[2] = {0.0f, 0.0f}, drec1[2] = {0.0f, 0.0f}, drecdt[2] = {0.0f, 0.0f};

  __device__ __forceinline__ void set(float dt, const Atom& at) {
    if (__float_as_uint(dt) == bits) return;
    bits = __float_as_uint(dt);
    const float bf = at.drate * dt;
    const float dtr = expf(-at.wt * bf);
    dlong = expf(-at.wl * bf);
    float s, c;
    sincosf(-kTwoPi * at.b0 * dt, &s, &c);
    turn = mk(c * dtr, s * dtr);
    if constexpr (!EXCH) {
      const float e1m = expm1f(-at.r1 * dt);
      e1 = 1.0f + e1m;
      om1 = -e1m;
      e2 = expf(-at.r2 * dt);
    } else {
      const float kab = at.kab, kba = at.kba;
      const float2 mt[4] = {mk((-kab - at.r2) * dt, 0.0f), mk(kba * dt, 0.0f), mk(kab * dt, 0.0f),
                            mk((-kba - at.r2b) * dt, -kTwoPi * at.shift_hz * dt)};
      const float2 et[4] = {mk(at.k2 * dt, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
      float2 xe[4], dxe[4];
      expm2<GRAD>(mt, et, xe, dxe);
      const float2 ml[4] = {mk((-kab - at.r1) * dt, 0.0f), mk(kba * dt, 0.0f), mk(kab * dt, 0.0f),
                            mk((-kba - at.r1b) * dt, 0.0f)};
      const float2 el[4] = {mk(at.k1 * dt, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
      float2 yl[4], dyl[4];
      expm2<GRAD>(ml, el, yl, dyl);
      float ye[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) ye[i] = yl[i].x;
      // recovery (Y - I) u, u = G^{-1} s: s = ((1-f) R1a, f R1b), G the longitudinal generator
      const float g00 = -kab - at.r1, g01 = kba, g10 = kab, g11 = -kba - at.r1b;
      const float det = g00 * g11 - g01 * g10;
      const float s0 = (1.0f - at.frac) * at.r1, s1 = at.frac * at.r1b;
      const float u0 = (g11 * s0 - g01 * s1) / det;
      const float u1 = (-g10 * s0 + g00 * s1) / det;
      rec[0] = (ye[0] - 1.0f) * u0 + ye[1] * u1;
      rec[1] = ye[2] * u0 + (ye[3] - 1.0f) * u1;
      if constexpr (GRAD) {
        // along T1: du = G^{-1} (ds - dG u), ds = (-(1-f) k1, 0), dG = diag(k1, 0)
        const float w0 = -(1.0f - at.frac) * at.k1 - at.k1 * u0;
        const float du0 = (g11 * w0) / det;
        const float du1 = (-g10 * w0) / det;
        drec1[0] = dyl[0].x * u0 + dyl[1].x * u1 + (ye[0] - 1.0f) * du0 + ye[1] * du1;
        drec1[1] = dyl[2].x * u0 + dyl[3].x * u1 + ye[2] * du0 + (ye[3] - 1.0f) * du1;
        // along the interval: d expm(M dt)/d dt = M expm(M dt) (M per unit time)
        const float2 mtr[4] = {mk(-kab - at.r2, 0.0f), mk(kba, 0.0f), mk(kab, 0.0f),
                               mk(-kba - at.r2b, -kTwoPi * at.shift_hz)};
        float2 mx[4];
        mx[0] = add(mul(mtr[0], xe[0]), mul(mtr[1], xe[2]));
        mx[1] = add(mul(mtr[0], xe[1]), mul(mtr[1], xe[3]));
        mx[2] = add(mul(mtr[2], xe[0]), mul(mtr[3], xe[2]));
        mx[3] = add(mul(mtr[2], xe[1]), mul(mtr[3], xe[3]));
        float my[4];
        my[0] = g00 * ye[0] + g01 * ye[2];
        my[1] = g00 * ye[1] + g01 * ye[3];
        my[2] = g10 * ye[0] + g11 * ye[2];
        my[3] = g10 * ye[1] + g11 * ye[3];
        drecdt[0] = my[0] * u0 + my[1] * u1;
        drecdt[1] = my[2] * u0 + my[3] * u1;
        // the turn and damping depend on dt too: d(turn)/d dt = c turn,
        // c = -2 pi i b0 - wt D; d(dlong)/d dt = -wl D dlong
        const float2 cl = mk(-at.wt * at.drate, -kTwoPi * at.b0);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          X[i] = mul(turn, xe[i]);
          dX2[i] = mul(turn, dxe[i]);
          GX[i] = add(mul(cl, X[i]), mul(turn, mx[i]));
          Y[i] = dlong * ye[i];
          dY1[i] = dlong * dyl[i].x;
          GY[i] = dlong * (my[i] - at.wl * at.drate * ye[i]);
        }
      } else {
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          X[i] = mul(turn, xe[i]);
          Y[i] = dlong * ye[i];
        }
      }
    }
  }
};

// Relaxation over the interval (the forward kernel's arithmetic).
template <bool EXCH, bool GRAD>
__device__ __forceinline__ void relax(State<EXCH>& x, const Interval<EXCH, GRAD>& I, bool first) {
  if constexpr (!EXCH) {
    const float2 tp = scl(I.turn, I.e2), tm = scl(cj(I.turn), I.e2);
    x.fp[0] = mul(x.fp[0], tp);
    x.fm[0] = mul(x.fm[0], tm);
    x.z[0] = scl(x.z[0], I.e1 * I.dlong);
    if (first) x.z[0].x += I.om1;
  } else {
    const float2 a0 = add(mul(I.X[0], x.fp[0]), mul(I.X[1], x.fp[1]));
    const float2 a1 = add(mul(I.X[2], x.fp[0]), mul(I.X[3], x.fp[1]));
    const float2 b0 = add(mul(cj(I.X[0]), x.fm[0]), mul(cj(I.X[1]), x.fm[1]));
    const float2 b1 = add(mul(cj(I.X[2]), x.fm[0]), mul(cj(I.X[3]), x.fm[1]));
    const float2 z0 = add(scl(x.z[0], I.Y[0]), scl(x.z[1], I.Y[1]));
    const float2 z1 = add(scl(x.z[0], I.Y[2]), scl(x.z[1], I.Y[3]));
    x.fp[0] = a0;
    x.fp[1] = a1;
    x.fm[0] = b0;
    x.fm[1] = b1;
    x.z[0] = z0;
    x.z[1] = z1;
    if (first) {
      x.z[0].x += I.rec[0];
      x.z[1].x += I.rec[1];
    }
  }
}

template <int L, bool EXCH>
__device__ __forceinline__ void shift_all(State<EXCH>& x, int sub, int lane, bool beyond) {
#pragma unroll
  for (int q = 0; q < State<EXCH>::NP; ++q) shift_complex<L>(x.fp[q], x.fm[q], sub, lane, beyond);
}

// Transpose of shift_complex: the adjoint of F+ moves DOWN an order (from the
// lane above, nothing above the last carried order), of F- UP; the coupling
// F+[0] = conj(F-[1]) sends conj(lambda_F+[0]) to lambda_F-[1]. Orders past
// the carried ones hold no state, so their adjoint is cleared.
template <int L>
__device__ __forceinline__ void shift_complex_t(float2& lp, float2& lm, int sub, int lane, int states) {
  const float2 p_up = mk(from_upper<L>(lp.x, lane), from_upper<L>(lp.y, lane));
  const float2 m_down = mk(from_lower<L>(lm.x, lane), from_lower<L>(lm.y, lane));
  const float2 p_down = mk(from_lower<L>(lp.x, lane), from_lower<L>(lp.y, lane));
  lp = sub + 1 < states ? p_up : mk(0.0f, 0.0f);
  lm = sub == 0 ? mk(0.0f, 0.0f) : m_down;
  if (sub == 1) lm = add(lm, cj(p_down));
  if (sub >= states) lm = mk(0.0f, 0.0f);
}

template <int L, bool EXCH>
__device__ __forceinline__ void shift_all_t(State<EXCH>& l, int sub, int lane, int states) {
#pragma unroll
  for (int q = 0; q < State<EXCH>::NP; ++q) shift_complex_t<L>(l.fp[q], l.fm[q], sub, lane, states);
}

// sin and 1 - cos of the pulse as the atom plays it (the forward kernel's
// convention: with a B1 map both from the half angle).
__device__ __forceinline__ void pulse_angle(const EventCode& ev, const Atom& at, int transmit, float& sa, float& va) {
  if (transmit) {
    const float alpha = ev.flip * at.b1;
    const float h = sinf(0.5f * alpha);
    sa = sinf(alpha);
    va = 2.0f * (h * h);
  } else {
    sa = ev.sine;
    va = ev.versine;
  }
}

// The RF T-matrix (reference rotate, _epg_cpu.cpp:2176-2208) of angle
// (sa, va) and phase phi: rows (F+, F-, Z) of
//   [[chh, t01, t02], [t10, chh, t12], [t20, t21, ca]].
struct Rotation {
  float chh, ca;
  float2 e1p;  // e^{i phi}
  float2 t01, t02, t10, t12, t20, t21;
  __device__ __forceinline__ void set(float sa, float va, float phi) {
    float sp, cp;
    sincosf(phi, &sp, &cp);
    e1p = mk(cp, sp);
    const float2 e2p = mul(e1p, e1p);
    const float shh = 0.5f * va;
    chh = 1.0f - shh;
    ca = 1.0f - va;
    t01 = scl(e2p, shh);
    t02 = mul(mk(0.0f, -sa), e1p);
    t10 = cj(t01);
    t12 = mul(mk(0.0f, sa), cj(e1p));
    t20 = mul(mk(0.0f, -0.5f * sa), cj(e1p));
    t21 = mul(mk(0.0f, 0.5f * sa), e1p);
  }
};

template <bool EXCH>
__device__ __forceinline__ void rotate(State<EXCH>& x, const Rotation& T) {
#pragma unroll
  for (int q = 0; q < State<EXCH>::NP; ++q) {
    const float2 fp = x.fp[q], fm = x.fm[q], z = x.z[q];
    x.fp[q] = mac(mac(scl(fp, T.chh), T.t01, fm), T.t02, z);
    x.fm[q] = mac(mac(mul(T.t10, fp), mk(T.chh, 0.0f), fm), T.t12, z);
    x.z[q] = mac(mac(mul(T.t20, fp), T.t21, fm), mk(T.ca, 0.0f), z);
  }
}

// Re(conj(a) b)
__device__ __forceinline__ float dotr(float2 a, float2 b) { return fmaf(a.x, b.x, a.y * b.y); }
// a - b (cx::sub under a name a lane's `sub` index does not hide)
__device__ __forceinline__ float2 sub_of(float2 a, float2 b) { return mk(a.x - b.x, a.y - b.y); }
// conj(a) b
__device__ __forceinline__ float2 cmul_conj(float2 a, float2 b) {
  return mk(fmaf(a.x, b.x, a.y * b.y), fmaf(a.x, b.y, -a.y * b.x));
}

// The reverse of one relaxation over interval I (dt): returns the gradient
// along the interval's duration; with TIMED, also adds the tissue gradients
// to g and applies lambda <- R^H lambda (at dt = 0 both are the identity, so
// an untimed event skips them).
//   single pool: with P = conj(l+) tp F+, M = conj(l-) tm F-,
//     Q = e1 dlong Re(conj(lz) Z), Q0 = Re(lz[0]) at order 0,
//     d/dT2 = k2 dt Re(P + M), d/dT1 = k1 dt (Q - e1 Q0),
//     d/dB0 = 2 pi dt Im(P - conj M), d/dD = -dt (wt Re(P + M) + wl Q),
//     d/ddt = -(r2 + wt D) Re(P + M) + 2 pi b0 Im(P - conj M) - (r1 + wl D) Q + r1 e1 Q0;
//   exchanging pool: with V_ij = conj(l+_i) F+_j + l-_i conj(F-_j) and
//     U_ij = Re(conj(lz_i) Z_j), a derivative dX, dY, drec of the operators
//     contributes Re sum V dX + sum U dY + Re(lz[0]) . drec.
template <bool EXCH, bool TIMED>
__device__ __forceinline__ float relax_reverse(const Interval<EXCH, true>& I, State<EXCH>& lam, const State<EXCH>& xe,
                                               const Atom& at, bool first, float dt, float (&g)[CG_COUNT]) {
  if constexpr (!EXCH) {
    const float2 tp = scl(I.turn, I.e2), tm = scl(cj(I.turn), I.e2);
    const float lo = I.e1 * I.dlong;
    const float2 P = mul(cmul_conj(lam.fp[0], tp), xe.fp[0]);
    const float2 M = mul(cmul_conj(lam.fm[0], tm), xe.fm[0]);
    const float Q = lo * dotr(lam.z[0], xe.z[0]);
    const float Q0 = first ? lam.z[0].x : 0.0f;
    const float re = P.x + M.x;
    const float im = P.y - M.y;
    if constexpr (TIMED) {
      g[CG_T2] += re * at.k2 * dt;
      g[CG_T1] += (Q - Q0 * I.e1) * at.k1 * dt;
      g[CG_B0] += kTwoPi * dt * im;
      g[CG_D] -= dt * (at.wt * re + at.wl * Q);
      lam.fp[0] = cmul_conj(tp, lam.fp[0]);
      lam.fm[0] = cmul_conj(tm, lam.fm[0]);
      lam.z[0] = scl(lam.z[0], lo);
    }
    return -(at.r2 + at.wt * at.drate) * re + kTwoPi * at.b0 * im - (at.r1 + at.wl * at.drate) * Q +
           Q0 * at.r1 * I.e1;
  } else {
    float2 V[4];
    float U[4];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        V[2 * i + j] = add(cmul_conj(lam.fp[i], xe.fp[j]), cmul_conj(xe.fm[j], lam.fm[i]));
        U[2 * i + j] = dotr(lam.z[i], xe.z[j]);
      }
    }
    float2 vg = mk(0.0f, 0.0f);
    float ug = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      vg = mac(vg, V[i], I.GX[i]);
      ug = fmaf(U[i], I.GY[i], ug);
    }
    const float rdg = first ? lam.z[0].x * I.drecdt[0] + lam.z[1].x * I.drecdt[1] : 0.0f;
    if constexpr (TIMED) {
      float2 vx = mk(0.0f, 0.0f), vd2 = mk(0.0f, 0.0f);
      float uy = 0.0f, ud1 = 0.0f;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        vx = mac(vx, V[i], I.X[i]);
        vd2 = mac(vd2, V[i], I.dX2[i]);
        uy = fmaf(U[i], I.Y[i], uy);
        ud1 = fmaf(U[i], I.dY1[i], ud1);
      }
      const float r1g = first ? lam.z[0].x * I.drec1[0] + lam.z[1].x * I.drec1[1] : 0.0f;
      g[CG_T2] += vd2.x;
      g[CG_T1] += ud1 + r1g;
      g[CG_B0] += kTwoPi * dt * vx.y;
      g[CG_D] -= dt * (at.wt * vx.x + at.wl * uy);
      // lambda+ <- X^H lambda+, lambda- <- X^T lambda-, lambda_z <- Y^T lambda_z
      const float2 p0 = add(cmul_conj(I.X[0], lam.fp[0]), cmul_conj(I.X[2], lam.fp[1]));
      const float2 p1 = add(cmul_conj(I.X[1], lam.fp[0]), cmul_conj(I.X[3], lam.fp[1]));
      const float2 m0v = add(mul(I.X[0], lam.fm[0]), mul(I.X[2], lam.fm[1]));
      const float2 m1v = add(mul(I.X[1], lam.fm[0]), mul(I.X[3], lam.fm[1]));
      const float2 z0 = add(scl(lam.z[0], I.Y[0]), scl(lam.z[1], I.Y[2]));
      const float2 z1 = add(scl(lam.z[0], I.Y[1]), scl(lam.z[1], I.Y[3]));
      lam.fp[0] = p0;
      lam.fp[1] = p1;
      lam.fm[0] = m0v;
      lam.fm[1] = m1v;
      lam.z[0] = z0;
      lam.z[1] = z1;
    }
    return vg.x + ug + rdg;
  }
}

// One event forward, value only: relaxation, shifts, pulse or inversion
// (a sample records nothing here).
template <int L, bool EXCH>
__device__ __forceinline__ void forward_event(State<EXCH>& x, const EventCode& ev, Interval<EXCH, false>& I,
                                              const Atom& at, const ComplexAdjointArgs& p, int sub, int lane,
                                              bool beyond) {
  const uint32_t act = ev.ctrl & 0xFFu;
  const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
  if (ev.dt != 0.0f) {
    I.set(ev.dt, at);
    relax<EXCH, false>(x, I, sub == 0);
  }
  if (act & PRE_SHIFT) shift_all<L, EXCH>(x, sub, lane, beyond);
  if (kind == KIND_RF) {
    if (act & INVERSION) {
#pragma unroll
      for (int q = 0; q < State<EXCH>::NP; ++q) x.z[q] = scl(x.z[q], -at.eta);
    } else {
      float sa, va;
      pulse_angle(ev, at, p.transmit, sa, va);
      Rotation T;
      T.set(sa, va, ev.phase + at.b1_phase);
      rotate<EXCH>(x, T);
    }
  }
  if (act & POST_SHIFT) shift_all<L, EXCH>(x, sub, lane, beyond);
  if (act & SPOIL_AFTER) {
#pragma unroll
    for (int q = 0; q < State<EXCH>::NP; ++q) x.fp[q] = x.fm[q] = mk(0.0f, 0.0f);
  } else if (act & SHIFT_AFTER) {
    shift_all<L, EXCH>(x, sub, lane, beyond);
  }
}

}  // namespace cadj

template <int L, bool EXCH>
__global__ void __launch_bounds__(256) complex_adjoint_kernel(const ComplexAdjointArgs p) {
  using namespace cadj;
  using S = State<EXCH>;
  constexpr int NP = S::NP;
  constexpr int APW = 64 / L;
  extern __shared__ float c_part_all[];  // per wave: [3 * CFLUSH rows][64 lanes]
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
  float* const s_part = c_part_all + (threadIdx.x >> 6) * (3 * CFLUSH * 64);

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
  if constexpr (EXCH) {
    at.frac = p.pool_fraction[ia];
    const float kex = p.pool_exchange[ia];
    at.kab = kex * at.frac;
    at.kba = kex * (1.0f - at.frac);
    at.r1b = 1000.0f / p.t1_pool[ia];
    at.r2b = 1000.0f / p.t2_pool[ia];
    at.shift_hz = p.pool_shift ? p.pool_shift[ia] : 0.0f;
  }
  const float order = (float)sub;
  at.wt = order * order + order + (1.0f / 3.0f);
  at.wl = order * order;

  const int64_t E = p.event_count;
  const int64_t K = p.segment;
  const int64_t segments = (E + K - 1) / K;
  const int64_t T = p.threads_total;
  const float2 factor = mk(p.factor_re, p.factor_im);
  const float2* w_row = p.cotangent + atom * p.outputs;

  // Pass 1: forward, one checkpoint per segment.
  S x;
  x.zero();
  if (first) {
    x.z[0] = mk(EXCH ? 1.0f - at.frac : 1.0f, 0.0f);
    if constexpr (EXCH) x.z[1] = mk(at.frac, 0.0f);
  }
  Interval<EXCH, false> If;
  for (int64_t e = 0; e < E; ++e) {
    if (e % K == 0) x.store(p.checkpoints + (e / K) * S::floats * T + tid, T);
    const EventCode ev = load_event(p.events, e);
    forward_event<L, EXCH>(x, ev, If, at, p, sub, lane, beyond);
  }

  // Pass 2: reverse, one segment at a time.
  float g[CG_COUNT];
#pragma unroll
  for (int q = 0; q < CG_COUNT; ++q) g[q] = 0.0f;
  S lam;
  lam.zero();
  Interval<EXCH, true> Ir;
  Interval<EXCH, true> I0;  // the operators of a zero interval (identity, with their dt derivatives)
  I0.set(0.0f, at);
  int rows = 0;           // events held in s_part
  int64_t flush_top = 0;  // event index of row 0
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
    x.load(p.checkpoints + seg * S::floats * T + tid, T);
    for (int64_t e = e0; e < e1; ++e) {
      x.store(states + (e - e0) * S::floats * T, T);
      const EventCode ev = load_event(p.events, e);
      forward_event<L, EXCH>(x, ev, If, at, p, sub, lane, beyond);
    }
    S xn;
    xn.load(states + (e1 - 1 - e0) * S::floats * T, T);
    for (int64_t e = e1 - 1; e >= e0; --e) {
      const S xe = xn;  // the event's entry state
      if (e > e0) xn.load(states + (e - 1 - e0) * S::floats * T, T);
      const EventCode ev = load_event(p.events, e);
      const uint32_t act = ev.ctrl & 0xFFu;
      const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
      if (rows == 0) flush_top = e;
      float g_alpha = 0.0f, g_phase = 0.0f, g_dt = 0.0f;
      const bool timed = ev.dt != 0.0f;
      if (timed) Ir.set(ev.dt, at);
      const bool record = kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs;
      const bool pulse = kind == KIND_RF && !(act & INVERSION);
      const bool invert = kind == KIND_RF && (act & INVERSION);

      // the state the pulse or sample acts on
      S xr = xe;
      if (pulse || invert || record) {
        if (timed) relax<EXCH, true>(xr, Ir, first);
        if (act & PRE_SHIFT) shift_all<L, EXCH>(xr, sub, lane, beyond);
      }

      // after the pulse: spoil or shift, then the post-shift
      if (act & SPOIL_AFTER) {
#pragma unroll
        for (int q = 0; q < NP; ++q) lam.fp[q] = lam.fm[q] = mk(0.0f, 0.0f);
      } else if (act & SHIFT_AFTER) {
        shift_all_t<L, EXCH>(lam, sub, lane, p.states);
      }
      if (act & POST_SHIFT) shift_all_t<L, EXCH>(lam, sub, lane, p.states);

      if (pulse) {
        float sa, va;
        pulse_angle(ev, at, p.transmit, sa, va);
        Rotation Tm;
        Tm.set(sa, va, ev.phase + at.b1_phase);
        const float ca = Tm.ca;
        // derivatives of the T-matrix along the angle
        const float dch = -0.5f * sa;
        const float2 e1p = Tm.e1p;
        const float2 d01 = scl(mul(e1p, e1p), 0.5f * sa);
        const float2 d02 = mul(mk(0.0f, -ca), e1p);
        const float2 d10 = cj(d01);
        const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
        const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
        const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
        const float dca = -sa;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          const float2 fp = xr.fp[q], fm = xr.fm[q], z = xr.z[q];
          const float2 lp = lam.fp[q], lm = lam.fm[q], lz = lam.z[q];
          // along the angle
          const float2 ap = mac(mac(scl(fp, dch), d01, fm), d02, z);
          const float2 am = mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z);
          const float2 az = mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z);
          g_alpha += dotr(lp, ap) + dotr(lm, am) + dotr(lz, az);
          // along the phase: d t01 = 2i t01, d t02 = i t02, d t10 = -2i t10,
          // d t12 = -i t12, d t20 = -i t20, d t21 = i t21
          // the phase derivatives of the three rows are i qp, i qm, i qz
          const float2 qp = add(scl(mul(Tm.t01, fm), 2.0f), mul(Tm.t02, z));
          const float2 qm = neg(add(scl(mul(Tm.t10, fp), 2.0f), mul(Tm.t12, z)));
          const float2 qz = cx::sub(mul(Tm.t21, fm), mul(Tm.t20, fp));
          // Re(conj(l) i v) = -Im(conj(l) v)
          g_phase -= cmul_conj(lp, qp).y + cmul_conj(lm, qm).y + cmul_conj(lz, qz).y;
          // lambda <- T^H lambda
          lam.fp[q] = add(add(scl(lp, Tm.chh), cmul_conj(Tm.t10, lm)), cmul_conj(Tm.t20, lz));
          lam.fm[q] = add(add(cmul_conj(Tm.t01, lp), scl(lm, Tm.chh)), cmul_conj(Tm.t21, lz));
          lam.z[q] = add(add(cmul_conj(Tm.t02, lp), cmul_conj(Tm.t12, lm)), scl(lz, ca));
        }
      } else if (invert) {
        float gi = 0.0f;
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          gi -= dotr(lam.z[q], xr.z[q]);
          lam.z[q] = scl(lam.z[q], -at.eta);
        }
        g[CG_IE] += gi;
      } else if (record && first && live) {
        float sp, cp;
        sincosf(-ev.phase, &sp, &cp);
        const float2 demod = mk(cp, sp);
        const float2 w = w_row[ev.out];
        float2 total = xr.fp[0];
        if constexpr (EXCH) total = add(total, xr.fp[1]);
        const float2 fd = mul(factor, demod);
        const float2 c = scl(fd, at.m0);
        const float2 lin = cmul_conj(c, w);  // conj(c) w
#pragma unroll
        for (int q = 0; q < NP; ++q) lam.fp[q] = add(lam.fp[q], lin);
        g[CG_M0] += dotr(w, mul(fd, total));
        // d sample / d phase = -i sample: Re(conj(w) (-i) s) = Im(conj(w) s)
        g_phase += cmul_conj(w, mul(c, total)).y;
      }

      if (act & PRE_SHIFT) shift_all_t<L, EXCH>(lam, sub, lane, p.states);

      // the relaxation: its duration's gradient, then (a nonzero interval)
      // the tissue gradients and lambda <- R^H lambda. An event with no
      // interval still has a duration gradient: the generator's pairing
      // <lambda, G x> at the event, from the operators at dt = 0.
      if (timed) {
        g_dt = relax_reverse<EXCH, true>(Ir, lam, xe, at, first, ev.dt, g);
      } else {
        g_dt = relax_reverse<EXCH, false>(I0, lam, xe, at, first, 0.0f, g);
      }
      g[CG_B1] = fmaf(ev.flip, g_alpha, g[CG_B1]);
      g[CG_B1_PHASE] += kind == KIND_RF ? g_phase : 0.0f;

      const bool counts = live && !beyond;
      s_part[(3 * rows + 0) * 64 + lane] = counts ? at.b1 * g_alpha : 0.0f;
      s_part[(3 * rows + 1) * 64 + lane] = counts ? g_phase : 0.0f;
      s_part[(3 * rows + 2) * 64 + lane] = counts ? g_dt : 0.0f;
      ++rows;
      if (rows == CFLUSH || e == e0) flush();
    }
  }

  // tissue gradients: each atom's lanes summed in order
  const int base = wave_group * L;
#pragma unroll
  for (int q = 0; q < CG_COUNT; ++q) {
    if (p.grad[q] == nullptr) continue;
    const float v = (live && !beyond) ? g[q] : 0.0f;
    float total = 0.0f;
    for (int j = 0; j < L; ++j) total += __shfl(v, (base + j) & 63, 64);
    if (first && live) p.grad[q][group] = total;
  }
}

// Forward mode along one arbitrary direction: a tangent for every carried
// tissue property (per atom, or one value for all under stride 0) and for
// every event's duration, flip and phase. It computes the reference's
// `simulate_jvp` (src/torchsim/_epg_cpu.cpp:2469-3005, entry :11017-11148)
// for the complex state machine, which is also exact on real-subspace
// inputs: the state and its tangent walk the events together, each operator
// applied to the tangent plus the operator's own derivative along the
// direction applied to the state.
struct ComplexJvpArgs {
  ComplexAdjointArgs base;        // events, tissue, sizes, features, factor
  const float* tissue_tangent[CG_COUNT];  // CG order; null = zero
  const float* duration_tangent;  // (events) or null
  const float* flip_tangent;      // (events) or null
  const float* phase_tangent;     // (events) or null
  float2* signal;                 // (atoms, outputs) or null
  float2* tangent;                // (atoms, outputs)
};

namespace cadj {

// One atom's tangent direction.
struct Direction {
  float t1, t2, m0, b1, b1_phase, b0, eta, drate;
};

// The tangent of one relaxation over interval I (dt) with duration tangent
// ddt: t <- R t + dR x, then (APPLY) x <- R x. At dt = 0, R is the identity
// and dR = ddt G, the interval's generator.
template <bool EXCH, bool APPLY>
__device__ __forceinline__ void relax_jvp(State<EXCH>& x, State<EXCH>& t, const Interval<EXCH, true>& I,
                                          const Atom& at, const Direction& d, float dt, float ddt, bool first) {
  if constexpr (!EXCH) {
    const float2 tp = scl(I.turn, I.e2), tm = scl(cj(I.turn), I.e2);
    const float lo = I.e1 * I.dlong;
    // d ln(tp) and d ln(tm): real parts equal, imaginary parts opposite
    const float lr = d.t2 * at.k2 * dt - ddt * (at.r2 + at.wt * at.drate) - d.drate * at.wt * dt;
    const float li = -kTwoPi * (d.b0 * dt + at.b0 * ddt);
    const float lz = d.t1 * at.k1 * dt - ddt * (at.r1 + at.wl * at.drate) - d.drate * at.wl * dt;
    const float dom1 = -I.e1 * (d.t1 * at.k1 * dt - ddt * at.r1);
    t.fp[0] = mul(tp, add(t.fp[0], mul(mk(lr, li), x.fp[0])));
    t.fm[0] = mul(tm, add(t.fm[0], mul(mk(lr, -li), x.fm[0])));
    t.z[0] = scl(add(t.z[0], scl(x.z[0], lz)), lo);
    if (first) t.z[0].x += dom1;
    if constexpr (APPLY) relax<EXCH, true>(x, I, first);
  } else {
    const float2 cx = mk(-at.wt * d.drate * dt, -kTwoPi * d.b0 * dt);
    const float cy = -at.wl * d.drate * dt;
    float2 dX[4];
    float dY[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      dX[i] = add(add(scl(I.dX2[i], d.t2), scl(I.GX[i], ddt)), mul(cx, I.X[i]));
      dY[i] = d.t1 * I.dY1[i] + ddt * I.GY[i] + cy * I.Y[i];
    }
    const float2 a0 = add(add(mul(I.X[0], t.fp[0]), mul(I.X[1], t.fp[1])), add(mul(dX[0], x.fp[0]), mul(dX[1], x.fp[1])));
    const float2 a1 = add(add(mul(I.X[2], t.fp[0]), mul(I.X[3], t.fp[1])), add(mul(dX[2], x.fp[0]), mul(dX[3], x.fp[1])));
    const float2 b0 = add(add(mul(cj(I.X[0]), t.fm[0]), mul(cj(I.X[1]), t.fm[1])),
                          add(mul(cj(dX[0]), x.fm[0]), mul(cj(dX[1]), x.fm[1])));
    const float2 b1 = add(add(mul(cj(I.X[2]), t.fm[0]), mul(cj(I.X[3]), t.fm[1])),
                          add(mul(cj(dX[2]), x.fm[0]), mul(cj(dX[3]), x.fm[1])));
    float2 z0 = add(add(scl(t.z[0], I.Y[0]), scl(t.z[1], I.Y[1])), add(scl(x.z[0], dY[0]), scl(x.z[1], dY[1])));
    float2 z1 = add(add(scl(t.z[0], I.Y[2]), scl(t.z[1], I.Y[3])), add(scl(x.z[0], dY[2]), scl(x.z[1], dY[3])));
    if (first) {
      z0.x += d.t1 * I.drec1[0] + ddt * I.drecdt[0];
      z1.x += d.t1 * I.drec1[1] + ddt * I.drecdt[1];
    }
    t.fp[0] = a0;
    t.fp[1] = a1;
    t.fm[0] = b0;
    t.fm[1] = b1;
    t.z[0] = z0;
    t.z[1] = z1;
    if constexpr (APPLY) relax<EXCH, true>(x, I, first);
  }
}

}  // namespace cadj

template <int L, bool EXCH>
__global__ void __launch_bounds__(256) complex_jvp_kernel(const ComplexJvpArgs q) {
  using namespace cadj;
  using S = State<EXCH>;
  constexpr int NP = S::NP;
  constexpr int APW = 64 / L;
  const ComplexAdjointArgs& p = q.base;
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t group = (tid >> 6) * APW + wave_group;
  const bool live = wave_group < APW && group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const bool beyond = sub >= p.states;
  const bool first = sub == 0;
  const int64_t ia = atom * p.stride_opt;

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
  if constexpr (EXCH) {
    at.frac = p.pool_fraction[ia];
    const float kex = p.pool_exchange[ia];
    at.kab = kex * at.frac;
    at.kba = kex * (1.0f - at.frac);
    at.r1b = 1000.0f / p.t1_pool[ia];
    at.r2b = 1000.0f / p.t2_pool[ia];
    at.shift_hz = p.pool_shift ? p.pool_shift[ia] : 0.0f;
  }
  const float order = (float)sub;
  at.wt = order * order + order + (1.0f / 3.0f);
  at.wl = order * order;
  // t1 and t2 tangents are per atom; the others follow the tissue stride
  auto tan = [&](int k, int64_t index) -> float {
    return q.tissue_tangent[k] ? q.tissue_tangent[k][index] : 0.0f;
  };
  Direction d;
  d.t1 = tan(CG_T1, atom);
  d.t2 = tan(CG_T2, atom);
  d.m0 = tan(CG_M0, ia);
  d.b1 = tan(CG_B1, ia);
  d.b1_phase = tan(CG_B1_PHASE, ia);
  d.b0 = tan(CG_B0, ia);
  d.eta = tan(CG_IE, ia);
  d.drate = tan(CG_D, ia);

  S x, t;
  x.zero();
  t.zero();
  if (first) {
    x.z[0] = mk(EXCH ? 1.0f - at.frac : 1.0f, 0.0f);
    if constexpr (EXCH) x.z[1] = mk(at.frac, 0.0f);
  }
  Interval<EXCH, true> I;
  Interval<EXCH, true> I0;
  I0.set(0.0f, at);
  const float2 factor = mk(p.factor_re, p.factor_im);
  const int64_t E = p.event_count;
  for (int64_t e = 0; e < E; ++e) {
    const EventCode ev = load_event(p.events, e);
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
    const float ddt = q.duration_tangent ? q.duration_tangent[e] : 0.0f;
    if (ev.dt != 0.0f) {
      I.set(ev.dt, at);
      relax_jvp<EXCH, true>(x, t, I, at, d, ev.dt, ddt, first);
    } else if (ddt != 0.0f) {
      relax_jvp<EXCH, false>(x, t, I0, at, d, 0.0f, ddt, first);
    }
    if (act & PRE_SHIFT) {
      shift_all<L, EXCH>(x, sub, lane, beyond);
      shift_all<L, EXCH>(t, sub, lane, beyond);
    }
    if (kind == KIND_RF) {
      if (act & INVERSION) {
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          t.z[k] = sub_of(scl(t.z[k], -at.eta), scl(x.z[k], d.eta));
          x.z[k] = scl(x.z[k], -at.eta);
        }
      } else {
        float sa, va;
        pulse_angle(ev, at, p.transmit, sa, va);
        Rotation Tm;
        Tm.set(sa, va, ev.phase + at.b1_phase);
        const float dflip = q.flip_tangent ? q.flip_tangent[e] : 0.0f;
        const float dphase = q.phase_tangent ? q.phase_tangent[e] : 0.0f;
        const float da = dflip * at.b1 + ev.flip * d.b1;
        const float dphi = dphase + d.b1_phase;
        const float ca = Tm.ca, dch = -0.5f * sa, dca = -sa;
        const float2 e1p = Tm.e1p;
        const float2 d01 = scl(mul(e1p, e1p), 0.5f * sa);
        const float2 d02 = mul(mk(0.0f, -ca), e1p);
        const float2 d10 = cj(d01);
        const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
        const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
        const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
#pragma unroll
        for (int k = 0; k < NP; ++k) {
          const float2 fp = x.fp[k], fm = x.fm[k], z = x.z[k];
          // along the angle
          const float2 ap = mac(mac(scl(fp, dch), d01, fm), d02, z);
          const float2 am = mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z);
          const float2 az = mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z);
          // along the phase: i (2 t01 fm + t02 z), i (-2 t10 fp - t12 z), i (t21 fm - t20 fp)
          const float2 qp = add(scl(mul(Tm.t01, fm), 2.0f), mul(Tm.t02, z));
          const float2 qm = neg(add(scl(mul(Tm.t10, fp), 2.0f), mul(Tm.t12, z)));
          const float2 qz = sub_of(mul(Tm.t21, fm), mul(Tm.t20, fp));
          const float2 tp = t.fp[k], tm = t.fm[k], tz = t.z[k];
          t.fp[k] = add(mac(mac(scl(tp, Tm.chh), Tm.t01, tm), Tm.t02, tz),
                        add(scl(ap, da), scl(mk(-qp.y, qp.x), dphi)));
          t.fm[k] = add(mac(mac(mul(Tm.t10, tp), mk(Tm.chh, 0.0f), tm), Tm.t12, tz),
                        add(scl(am, da), scl(mk(-qm.y, qm.x), dphi)));
          t.z[k] = add(mac(mac(mul(Tm.t20, tp), Tm.t21, tm), mk(ca, 0.0f), tz),
                       add(scl(az, da), scl(mk(-qz.y, qz.x), dphi)));
        }
        rotate<EXCH>(x, Tm);
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs && first && live) {
      float sp, cp;
      sincosf(-ev.phase, &sp, &cp);
      const float2 fd = mul(factor, mk(cp, sp));
      float2 total = x.fp[0], ttotal = t.fp[0];
      if constexpr (EXCH) {
        total = add(total, x.fp[1]);
        ttotal = add(ttotal, t.fp[1]);
      }
      const float2 sample = mul(fd, scl(total, at.m0));
      const float dphase = q.phase_tangent ? q.phase_tangent[e] : 0.0f;
      // d sample = fd (dm0 total + m0 dtotal) - i dphase sample
      const float2 ds = add(mul(fd, add(scl(total, d.m0), scl(ttotal, at.m0))), scl(mk(sample.y, -sample.x), dphase));
      const int64_t o = atom * p.outputs + ev.out;
      if (q.signal) q.signal[o] = sample;
      q.tangent[o] = ds;
    }
    if (act & POST_SHIFT) {
      shift_all<L, EXCH>(x, sub, lane, beyond);
      shift_all<L, EXCH>(t, sub, lane, beyond);
    }
    if (act & SPOIL_AFTER) {
#pragma unroll
      for (int k = 0; k < NP; ++k) x.fp[k] = x.fm[k] = t.fp[k] = t.fm[k] = mk(0.0f, 0.0f);
    } else if (act & SHIFT_AFTER) {
      shift_all<L, EXCH>(x, sub, lane, beyond);
      shift_all<L, EXCH>(t, sub, lane, beyond);
    }
  }
}

using ComplexJvpLaunch = void (*)(const ComplexJvpArgs&, dim3, dim3, hipStream_t);

template <int L, bool EXCH>
void launch_complex_jvp(const ComplexJvpArgs& a, dim3 grid, dim3 block, hipStream_t s) {
  hipLaunchKernelGGL((complex_jvp_kernel<L, EXCH>), grid, block, 0, s, a);
}

#define TSIM_DECLARE_COMPLEX_JVP(LANES) ComplexJvpLaunch pick_complex_jvp_##LANES(bool exch);
TSIM_COMPLEX_LANES(TSIM_DECLARE_COMPLEX_JVP)
#undef TSIM_DECLARE_COMPLEX_JVP

using ComplexAdjointLaunch = void (*)(const ComplexAdjointArgs&, dim3, dim3, size_t, hipStream_t);

template <int L, bool EXCH>
void launch_complex_adjoint(const ComplexAdjointArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t s) {
  hipLaunchKernelGGL((complex_adjoint_kernel<L, EXCH>), grid, block, lds, s, a);
}

#define TSIM_DECLARE_COMPLEX_ADJOINT(LANES) ComplexAdjointLaunch pick_complex_adjoint_##LANES(bool exch);
TSIM_COMPLEX_LANES(TSIM_DECLARE_COMPLEX_ADJOINT)
#undef TSIM_DECLARE_COMPLEX_ADJOINT

}  // namespace tsim
