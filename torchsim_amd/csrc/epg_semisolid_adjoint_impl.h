// First-order adjoint of the semisolid-pool / slice-profiled-pulse kernel,
// for gfx950: the gradient of a real loss with respect to every event's flip
// angle, phase and duration, given the loss's cotangent on the recorded
// samples (the reference's simulate_vjp_range<INSTANT | PROFILED, SEMISOLID>
// and simulate_vjp_single_pool<PROFILED>, src/torchsim/_epg_cpu.cpp:
// 4839-5944 and the dispatch of :11310-11340, for the event gradients
// `_NativeEpg.backward` hands on, sequence/_accelerators.py:1544-1586).
//
// Layout as the forward: one configuration order per lane, an atom on L
// adjacent lanes. Pass 1 runs the events forward and stores every event's
// entry state (free F+, F-, Z and bound Z: 8 floats) to scratch,
// [event][float][lane] so a wave's stores coalesce; pass 2 walks the events
// last to first, recomputes the event's relaxation and pre-shift from its
// entry state and applies the transposed operators to the adjoint state
// lambda: the shift's transpose, T^H for a pulse (and the bound pool's
// saturation factor, real), conj(tp) / tp / Y^T for the relaxation of
// F+ / F- / (Z, Z_bound). A launch covers as many atoms as its scratch
// holds; the caller walks the atoms launch by launch.
//
// Conventions. The loss is real; the cotangent w of a complex output s is
// dL/dRe(s) + i dL/dIm(s), so dL = Re(conj(w) ds), and the adjoint of a state
// entry x is lambda with dL = Re(conj(lambda) dx). Each event's three
// gradient terms are summed over the wave's lanes (a butterfly) and stored
// per wave; a second launch sums the waves in a fixed order.
#pragma once

#include "epg_semisolid_impl.h"

namespace tsim {

struct SemisolidVjpArgs {
  SemisolidArgs p;          // the forward's arguments (its outputs unused)
  const float2* cotangent;  // (atoms, outputs)
  float* trajectory;        // (events, 8, threads): each event's entry state, this launch's lanes
  float* partial;           // (waves, events, 3): flip, phase, duration terms per wave
  int64_t first_atom;       // this launch's atoms: [first_atom, first_atom + launch_atoms)
  int64_t launch_atoms;
  int64_t threads;          // lanes of this launch (the trajectory's stride)
};

namespace mtv {

using namespace cx;
using mt::Order;

// Re(conj(a) b)
__device__ __forceinline__ float dotr(float2 a, float2 b) { return fmaf(a.x, b.x, a.y * b.y); }

// One interval on this lane: the free water's transverse factor tp (F+;
// F- takes its conjugate), the (Z, Z_bound) operator Y with this order's
// damping, the recovery at order 0, and their derivatives along dt.
struct Relax {
  float2 tp, dtp;
  float Y[4], dY[4];
  float rec[2], drec[2];
};

struct Atom {
  float r1, r2, r1b, kab, kba, f, free, b0, drate, order;
};

__device__ __forceinline__ void form_relax(const Atom& at, float dt, Relax& R) {
  const float wt = at.order * at.order + at.order + (1.0f / 3.0f);
  const float wl = at.order * at.order;
  const float bf = at.drate * dt;
  const float dtr = expf(-wt * bf);
  const float dlong = expf(-wl * bf);
  float s, c;
  sincosf(-2.0f * 3.14159265358979323846f * at.b0 * dt, &s, &c);
  const float e2 = expf(-at.r2 * dt);
  R.tp = mk(c * dtr * e2, s * dtr * e2);
  R.dtp = mul(R.tp, mk(-(at.r2 + wt * at.drate), -2.0f * 3.14159265358979323846f * at.b0));
  const float g00 = -at.kab - at.r1, g01 = at.kba, g10 = at.kab, g11 = -at.kba - at.r1b;
  const float2 ml[4] = {mk(g00 * dt, 0.0f), mk(g01 * dt, 0.0f), mk(g10 * dt, 0.0f), mk(g11 * dt, 0.0f)};
  const float2 el[4] = {mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
  float2 yl[4], unused[4];
  expm2<false>(ml, el, yl, unused);
  const float y0 = yl[0].x, y1 = yl[1].x, y2 = yl[2].x, y3 = yl[3].x;
  // recovery toward the equilibrium split (1 - f, f); d/dt = -G Y eq
  R.rec[0] = at.free - (y0 * at.free + y1 * at.f);
  R.rec[1] = at.f - (y2 * at.free + y3 * at.f);
  const float q0 = y0 * at.free + y1 * at.f, q1 = y2 * at.free + y3 * at.f;
  R.drec[0] = -(g00 * q0 + g01 * q1);
  R.drec[1] = -(g10 * q0 + g11 * q1);
  R.Y[0] = y0 * dlong;
  R.Y[1] = y1 * dlong;
  R.Y[2] = y2 * dlong;
  R.Y[3] = y3 * dlong;
  // d(Y dlong)/dt = (G - wl D) Y dlong
  const float d = wl * at.drate;
  R.dY[0] = (g00 - d) * R.Y[0] + g01 * R.Y[2];
  R.dY[1] = (g00 - d) * R.Y[1] + g01 * R.Y[3];
  R.dY[2] = g10 * R.Y[0] + (g11 - d) * R.Y[2];
  R.dY[3] = g10 * R.Y[1] + (g11 - d) * R.Y[3];
}

__device__ __forceinline__ void relax(Order& x, const Relax& R, bool first) {
  x.fp = mul(x.fp, R.tp);
  x.fm = mul(x.fm, cj(R.tp));
  const float2 z = add(scl(x.z, R.Y[0]), scl(x.zb, R.Y[1]));
  const float2 zb = add(scl(x.z, R.Y[2]), scl(x.zb, R.Y[3]));
  x.z = z;
  x.zb = zb;
  if (first) {
    x.z.x += R.rec[0];
    x.zb.x += R.rec[1];
  }
}

// The reverse of relax from entry state x: returns the term along dt and
// turns lambda (the adjoint after the interval) into the one before it.
__device__ __forceinline__ float relax_reverse(const Relax& R, const Order& x, Order& l, bool first) {
  float g = dotr(l.fp, mul(R.dtp, x.fp)) + dotr(l.fm, mul(cj(R.dtp), x.fm)) +
            dotr(l.z, add(scl(x.z, R.dY[0]), scl(x.zb, R.dY[1]))) +
            dotr(l.zb, add(scl(x.z, R.dY[2]), scl(x.zb, R.dY[3])));
  if (first) g += l.z.x * R.drec[0] + l.zb.x * R.drec[1];
  l.fp = mul(cj(R.tp), l.fp);
  l.fm = mul(R.tp, l.fm);
  const float2 lz = add(scl(l.z, R.Y[0]), scl(l.zb, R.Y[2]));
  const float2 lb = add(scl(l.z, R.Y[1]), scl(l.zb, R.Y[3]));
  l.z = lz;
  l.zb = lb;
  return g;
}

// Transpose of shift_complex: the adjoint of F+ moves down an order (from
// the lane above, nothing past the last carried order), of F- up; F+[0] =
// conj(F-[1]) sends conj(lambda_F+[0]) to lambda_F-[1].
template <int L>
__device__ __forceinline__ void shift_reverse(Order& l, int sub, int lane, int states) {
  const float2 p_up = mk(from_upper<L>(l.fp.x, lane), from_upper<L>(l.fp.y, lane));
  const float2 m_down = mk(from_lower<L>(l.fm.x, lane), from_lower<L>(l.fm.y, lane));
  const float2 p_down = mk(from_lower<L>(l.fp.x, lane), from_lower<L>(l.fp.y, lane));
  l.fp = sub + 1 < states ? p_up : mk(0.0f, 0.0f);
  l.fm = sub == 0 ? mk(0.0f, 0.0f) : m_down;
  if (sub == 1) l.fm = add(l.fm, cj(p_down));
  if (sub >= states) l.fm = mk(0.0f, 0.0f);
}

// A pulse's T-matrix on (F+, F-, Z) and its derivative along the effective
// flip theta: rows [[t00, t01, t02], [t10, t11, t12], [t20, t21, t22]]
// (t22 real), for an instantaneous pulse from (sin, 1 - cos) of theta and the
// phase, for a profiled one from the Cayley-Klein pair its table holds at
// the atom's position (reference rotate / rotate_spinor).
struct Pulse {
  float2 t00, t01, t02, t10, t11, t12, t20, t21;
  float t22;
  float2 u00, u01, u02, u10, u11, u12, u20, u21;
  float u22;
};

template <bool PROFILED>
__device__ __forceinline__ void form_pulse(const SemisolidArgs& p, const EventCode& ev, int64_t atom, float theta,
                                           float phase, int transmit, Pulse& T) {
  if constexpr (PROFILED) {
    float2 a, bz, da, dbz;
    profile_pair(p, atom % p.locations, theta, a, bz, da, dbz);
    float sp, cp;
    sincosf(-phase, &sp, &cp);
    const float2 spin = mk(cp, sp);
    const float2 b = mul(bz, spin), db = mul(dbz, spin);
    const float2 ca = cj(a), cb = cj(b), cda = cj(da), cdb = cj(db);
    T.t00 = mul(ca, ca);
    T.t01 = neg(mul(cb, cb));
    T.t02 = scl(cj(mul(a, b)), -2.0f);
    T.t10 = neg(mul(b, b));
    T.t11 = mul(a, a);
    T.t12 = scl(mul(a, b), -2.0f);
    T.t20 = mul(ca, b);
    T.t21 = mul(a, cb);
    T.t22 = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
    const float2 ab = add(mul(da, b), mul(a, db));
    T.u00 = scl(mul(ca, cda), 2.0f);
    T.u01 = scl(mul(cb, cdb), -2.0f);
    T.u02 = scl(cj(ab), -2.0f);
    T.u10 = scl(mul(b, db), -2.0f);
    T.u11 = scl(mul(a, da), 2.0f);
    T.u12 = scl(ab, -2.0f);
    T.u20 = add(mul(cda, b), mul(ca, db));
    T.u21 = add(mul(da, cb), mul(a, cdb));
    T.u22 = 2.0f * (a.x * da.x + a.y * da.y) - 2.0f * (b.x * db.x + b.y * db.y);
  } else {
    float sa, va;
    if (transmit) {
      const float h = sinf(0.5f * theta);
      sa = sinf(theta);
      va = 2.0f * (h * h);
    } else {
      sa = ev.sine;
      va = ev.versine;
    }
    const float ca = 1.0f - va;
    float sp, cp;
    sincosf(phase, &sp, &cp);
    const float2 e1p = mk(cp, sp);
    const float2 e2p = mul(e1p, e1p);
    const float shh = 0.5f * va, chh = 1.0f - shh;
    T.t00 = T.t11 = mk(chh, 0.0f);
    T.t01 = scl(e2p, shh);
    T.t02 = mul(mk(0.0f, -sa), e1p);
    T.t10 = cj(T.t01);
    T.t12 = mul(mk(0.0f, sa), cj(e1p));
    T.t20 = mul(mk(0.0f, -0.5f * sa), cj(e1p));
    T.t21 = mul(mk(0.0f, 0.5f * sa), e1p);
    T.t22 = ca;
    T.u00 = T.u11 = mk(-0.5f * sa, 0.0f);
    T.u01 = scl(e2p, 0.5f * sa);
    T.u02 = mul(mk(0.0f, -ca), e1p);
    T.u10 = cj(T.u01);
    T.u12 = mul(mk(0.0f, ca), cj(e1p));
    T.u20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
    T.u21 = mul(mk(0.0f, 0.5f * ca), e1p);
    T.u22 = -sa;
  }
}

__device__ __forceinline__ void rotate(Order& x, const Pulse& T) {
  const float2 fp = x.fp, fm = x.fm, z = x.z;
  x.fp = add(add(mul(T.t00, fp), mul(T.t01, fm)), mul(T.t02, z));
  x.fm = add(add(mul(T.t10, fp), mul(T.t11, fm)), mul(T.t12, z));
  x.z = add(add(mul(T.t20, fp), mul(T.t21, fm)), scl(z, T.t22));
}

}  // namespace mtv

template <int L, bool PROFILED>
__global__ void __launch_bounds__(64) semisolid_vjp_kernel(const SemisolidVjpArgs q) {
  using namespace cx;
  using mt::Order;
  using mtv::dotr;
  const SemisolidArgs& p = q.p;
  constexpr int APW = 64 / L;
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t wave = tid >> 6;
  const int64_t local = wave * APW + wave_group;
  const bool live = wave_group < APW && local < q.launch_atoms;
  const int64_t atom = q.first_atom + (live ? local : q.launch_atoms - 1);
  const bool first = sub == 0;
  const bool beyond = sub >= p.states;
  const int64_t ia = atom * p.stride_opt;

  mtv::Atom at;
  const float t1 = p.t1[atom], t2 = p.t2[atom];
  at.r1 = 1000.0f / t1;
  at.r2 = 1000.0f / t2;
  const float m0 = p.density ? p.m0[ia] : 1.0f;
  const float b1 = p.transmit ? p.b1[ia] : 1.0f;
  const float b1_phase = p.off_axis && p.b1_phase ? p.b1_phase[ia] : 0.0f;
  at.b0 = p.off_axis && p.b0 ? p.b0[ia] : 0.0f;
  const float eta = p.inverting ? p.inv_eff[ia] : 1.0f;
  at.drate = p.diffusing ? p.diffusion[ia] : 0.0f;
  at.f = p.bound_fraction ? p.bound_fraction[ia] : 0.0f;
  const float kex = p.bound_exchange ? p.bound_exchange[ia] : 0.0f;
  at.r1b = p.t1_bound ? 1000.0f / p.t1_bound[ia] : at.r1;
  at.free = 1.0f - at.f;
  at.kab = kex * at.f;
  at.kba = kex * at.free;
  at.order = (float)sub;
  const float mask = live ? 1.0f : 0.0f;

  const int64_t E = p.event_count;
  const int64_t T = q.threads;
  float* const traj = q.trajectory + tid;
  const float2 factor = mk(p.factor_re, p.factor_im);

  // pass 1: forward, storing each event's entry state
  Order x;
  x.fp = x.fm = x.z = x.zb = mk(0.0f, 0.0f);
  if (first) {
    x.z = mk(at.free, 0.0f);
    x.zb = mk(at.f, 0.0f);
  }
  mtv::Relax R;
  float last_dt = -1.0f;
  for (int64_t e = 0; e < E; ++e) {
    const EventCode ev = load_event(p.events, e);
    float* slot = traj + e * 8 * T;
    slot[0] = x.fp.x;
    slot[T] = x.fp.y;
    slot[2 * T] = x.fm.x;
    slot[3 * T] = x.fm.y;
    slot[4 * T] = x.z.x;
    slot[5 * T] = x.z.y;
    slot[6 * T] = x.zb.x;
    slot[7 * T] = x.zb.y;
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
    if (ev.dt != 0.0f) {
      if (ev.dt != last_dt) {
        last_dt = ev.dt;
        mtv::form_relax(at, ev.dt, R);
      }
      mtv::relax(x, R, first);
    }
    if (act & PRE_SHIFT) shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
    if (kind == KIND_RF) {
      if (act & INVERSION) {
        x.z = scl(x.z, -eta);
      } else {
        const float theta = ev.flip * b1;
        const float deposit = p.bound_fraction ? p.saturation[e] : 0.0f;
        if (deposit != 0.0f) x.zb = scl(x.zb, expf(deposit * theta * theta * lineshape_at(p, p.rf_frequency[e] - at.b0)));
        mtv::Pulse P;
        mtv::form_pulse<PROFILED>(p, ev, atom, theta, ev.phase + b1_phase, p.transmit, P);
        mtv::rotate(x, P);
      }
    }
    if (act & POST_SHIFT) shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
    if (act & SPOIL_AFTER) {
      x.fp = x.fm = mk(0.0f, 0.0f);
    } else if (act & SHIFT_AFTER) {
      shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
    }
  }

  // pass 2: the events last to first. Every event has a duration gradient,
  // an untimed one too (the reference's convention): the relaxation at
  // dt = 0 is the identity and its derivative there the generator.
  mtv::Relax R0;
  mtv::form_relax(at, 0.0f, R0);
  Order l;
  l.fp = l.fm = l.z = l.zb = mk(0.0f, 0.0f);
  last_dt = -1.0f;
  // each event's entry state is read one event ahead, so that the read's
  // latency overlaps the work on the event before it
  auto entry_state = [&](int64_t e, Order& o) {
    const float* slot = traj + e * 8 * T;
    o.fp = mk(slot[0], slot[T]);
    o.fm = mk(slot[2 * T], slot[3 * T]);
    o.z = mk(slot[4 * T], slot[5 * T]);
    o.zb = mk(slot[6 * T], slot[7 * T]);
  };
  Order ahead;
  entry_state(E - 1, ahead);
  for (int64_t e = E - 1; e >= 0; --e) {
    const EventCode ev = load_event(p.events, e);
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
    const Order x0 = ahead;
    if (e > 0) entry_state(e - 1, ahead);
    // the state the event's action sees
    Order xa = x0;
    if (ev.dt != 0.0f) {
      if (ev.dt != last_dt) {
        last_dt = ev.dt;
        mtv::form_relax(at, ev.dt, R);
      }
      mtv::relax(xa, R, first);
    }
    if (act & PRE_SHIFT) shift_complex<L>(xa.fp, xa.fm, sub, lane, beyond);

    float g_flip = 0.0f, g_phase = 0.0f, g_dt = 0.0f;
    if (act & SPOIL_AFTER) {
      l.fp = l.fm = mk(0.0f, 0.0f);
    } else if (act & SHIFT_AFTER) {
      mtv::shift_reverse<L>(l, sub, lane, p.states);
    }
    if (act & POST_SHIFT) mtv::shift_reverse<L>(l, sub, lane, p.states);
    if (kind == KIND_RF) {
      if (act & INVERSION) {
        l.z = scl(l.z, -eta);
      } else {
        const float theta = ev.flip * b1;
        const float deposit = p.bound_fraction ? p.saturation[e] : 0.0f;
        if (deposit != 0.0f) {
          const float g = lineshape_at(p, p.rf_frequency[e] - at.b0);
          const float w = expf(deposit * theta * theta * g);
          g_flip += dotr(l.zb, xa.zb) * w * deposit * 2.0f * theta * b1 * g;
          l.zb = scl(l.zb, w);
        }
        mtv::Pulse P;
        mtv::form_pulse<PROFILED>(p, ev, atom, theta, ev.phase + b1_phase, p.transmit, P);
        const float2 fp = xa.fp, fm = xa.fm, z = xa.z;
        // along the effective flip (the flip's is b1 times it)
        const float2 ap = add(add(mul(P.u00, fp), mul(P.u01, fm)), mul(P.u02, z));
        const float2 am = add(add(mul(P.u10, fp), mul(P.u11, fm)), mul(P.u12, z));
        const float2 az = add(add(mul(P.u20, fp), mul(P.u21, fm)), scl(z, P.u22));
        g_flip += b1 * (dotr(l.fp, ap) + dotr(l.fm, am) + dotr(l.z, az));
        // along the phase: the entries carry e^{i k phi} (k = 2, 1, -2, -1,
        // -1, 1 off the diagonal) and gain the factor i k
        const float2 pp = add(mul(mk(0.0f, 2.0f), mul(P.t01, fm)), mul(mk(0.0f, 1.0f), mul(P.t02, z)));
        const float2 pm = add(mul(mk(0.0f, -2.0f), mul(P.t10, fp)), mul(mk(0.0f, -1.0f), mul(P.t12, z)));
        const float2 pz = add(mul(mk(0.0f, -1.0f), mul(P.t20, fp)), mul(mk(0.0f, 1.0f), mul(P.t21, fm)));
        g_phase += dotr(l.fp, pp) + dotr(l.fm, pm) + dotr(l.z, pz);
        // lambda <- T^H lambda
        const float2 lp = l.fp, lm = l.fm, lz = l.z;
        l.fp = add(add(mul(cj(P.t00), lp), mul(cj(P.t10), lm)), mul(cj(P.t20), lz));
        l.fm = add(add(mul(cj(P.t01), lp), mul(cj(P.t11), lm)), mul(cj(P.t21), lz));
        l.z = add(add(mul(cj(P.t02), lp), mul(cj(P.t12), lm)), scl(lz, P.t22));
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs && first && live) {
      float sp, cp;
      sincosf(-ev.phase, &sp, &cp);
      const float2 c = scl(mul(factor, mk(cp, sp)), m0);
      const float2 w = q.cotangent[atom * p.outputs + ev.out];
      const float2 s = mul(c, xa.fp);
      g_phase += dotr(w, mk(s.y, -s.x));  // ds/dphase = -i s
      l.fp = add(l.fp, mul(cj(c), w));
    }
    if (act & PRE_SHIFT) mtv::shift_reverse<L>(l, sub, lane, p.states);
    g_dt += mtv::relax_reverse(ev.dt != 0.0f ? R : R0, x0, l, first);

    g_flip *= mask;
    g_phase *= mask;
    g_dt *= mask;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
      g_flip += __shfl_xor(g_flip, off, 64);
      g_phase += __shfl_xor(g_phase, off, 64);
      g_dt += __shfl_xor(g_dt, off, 64);
    }
    if (lane == 0) {
      float* out = q.partial + (wave * E + e) * 3;
      out[0] = g_flip;
      out[1] = g_phase;
      out[2] = g_dt;
    }
  }
}

// Forward mode along the events: the directional derivative of the signal
// for per-event tangents of the flips, phases and durations (the event rows
// of the reference's simulate_jvp, src/torchsim/_epg_cpu.cpp:11017-11148,
// for pool kind 1 and RfMode::PROFILED). The tangent state t follows the
// value x through every operator A and takes A's derivative along the event
// direction applied to x: dA/d(dt) d_dt for a relaxation, (b1 dT/dtheta
// d_flip + dT/dphi d_phase) for a pulse (and the saturation factor's flip
// term), -i s d_phase for a record's demodulation.
struct SemisolidEventJvpArgs {
  SemisolidArgs p;
  const float* d_flip;      // (events,) or null
  const float* d_phase;     // (events,) or null
  const float* d_duration;  // (events,) or null
  float2* tangent;          // (atoms, outputs)
};

template <int L, bool PROFILED>
__global__ void __launch_bounds__(64) semisolid_event_jvp_kernel(const SemisolidEventJvpArgs q) {
  using namespace cx;
  using mt::Order;
  const SemisolidArgs& p = q.p;
  constexpr int APW = 64 / L;
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t group = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * APW + wave_group;
  const bool live = wave_group < APW && group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const bool first = sub == 0;
  const bool beyond = sub >= p.states;
  const int64_t ia = atom * p.stride_opt;

  mtv::Atom at;
  const float t1 = p.t1[atom], t2 = p.t2[atom];
  at.r1 = 1000.0f / t1;
  at.r2 = 1000.0f / t2;
  const float m0 = p.density ? p.m0[ia] : 1.0f;
  const float b1 = p.transmit ? p.b1[ia] : 1.0f;
  const float b1_phase = p.off_axis && p.b1_phase ? p.b1_phase[ia] : 0.0f;
  at.b0 = p.off_axis && p.b0 ? p.b0[ia] : 0.0f;
  const float eta = p.inverting ? p.inv_eff[ia] : 1.0f;
  at.drate = p.diffusing ? p.diffusion[ia] : 0.0f;
  at.f = p.bound_fraction ? p.bound_fraction[ia] : 0.0f;
  const float kex = p.bound_exchange ? p.bound_exchange[ia] : 0.0f;
  at.r1b = p.t1_bound ? 1000.0f / p.t1_bound[ia] : at.r1;
  at.free = 1.0f - at.f;
  at.kab = kex * at.f;
  at.kba = kex * at.free;
  at.order = (float)sub;
  const float2 factor = mk(p.factor_re, p.factor_im);

  Order x, t;
  x.fp = x.fm = x.z = x.zb = mk(0.0f, 0.0f);
  t = x;
  if (first) {
    x.z = mk(at.free, 0.0f);
    x.zb = mk(at.f, 0.0f);
  }
  mtv::Relax R, R0;
  mtv::form_relax(at, 0.0f, R0);
  float last_dt = -1.0f;
  for (int64_t e = 0; e < p.event_count; ++e) {
    const EventCode ev = load_event(p.events, e);
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);
    const float d_dt = q.d_duration ? q.d_duration[e] : 0.0f;
    // relaxation: t <- R t + dR/d(dt) x d_dt (an untimed event's derivative too)
    {
      const mtv::Relax* use = &R0;
      if (ev.dt != 0.0f) {
        if (ev.dt != last_dt) {
          last_dt = ev.dt;
          mtv::form_relax(at, ev.dt, R);
        }
        use = &R;
      }
      Order dx;
      dx.fp = mul(use->dtp, x.fp);
      dx.fm = mul(cj(use->dtp), x.fm);
      dx.z = add(scl(x.z, use->dY[0]), scl(x.zb, use->dY[1]));
      dx.zb = add(scl(x.z, use->dY[2]), scl(x.zb, use->dY[3]));
      if (first) {
        dx.z.x += use->drec[0];
        dx.zb.x += use->drec[1];
      }
      if (ev.dt != 0.0f) {
        mtv::relax(t, R, false);  // the recovery is constant: no tangent term of its own
        mtv::relax(x, R, first);
      }
      t.fp = add(t.fp, scl(dx.fp, d_dt));
      t.fm = add(t.fm, scl(dx.fm, d_dt));
      t.z = add(t.z, scl(dx.z, d_dt));
      t.zb = add(t.zb, scl(dx.zb, d_dt));
    }
    if (act & PRE_SHIFT) {
      shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
      shift_complex<L>(t.fp, t.fm, sub, lane, beyond);
    }
    if (kind == KIND_RF) {
      if (act & INVERSION) {
        x.z = scl(x.z, -eta);
        t.z = scl(t.z, -eta);
      } else {
        const float theta = ev.flip * b1;
        const float d_theta = (q.d_flip ? q.d_flip[e] : 0.0f) * b1;
        const float d_phi = q.d_phase ? q.d_phase[e] : 0.0f;
        const float deposit = p.bound_fraction ? p.saturation[e] : 0.0f;
        if (deposit != 0.0f) {
          const float g = lineshape_at(p, p.rf_frequency[e] - at.b0);
          const float w = expf(deposit * theta * theta * g);
          t.zb = add(scl(t.zb, w), scl(x.zb, w * deposit * 2.0f * theta * g * d_theta));
          x.zb = scl(x.zb, w);
        }
        mtv::Pulse P;
        mtv::form_pulse<PROFILED>(p, ev, atom, theta, ev.phase + b1_phase, p.transmit, P);
        const float2 fp = x.fp, fm = x.fm, z = x.z;
        const float2 ap = add(add(mul(P.u00, fp), mul(P.u01, fm)), mul(P.u02, z));
        const float2 am = add(add(mul(P.u10, fp), mul(P.u11, fm)), mul(P.u12, z));
        const float2 az = add(add(mul(P.u20, fp), mul(P.u21, fm)), scl(z, P.u22));
        const float2 pp = add(mul(mk(0.0f, 2.0f), mul(P.t01, fm)), mul(mk(0.0f, 1.0f), mul(P.t02, z)));
        const float2 pm = add(mul(mk(0.0f, -2.0f), mul(P.t10, fp)), mul(mk(0.0f, -1.0f), mul(P.t12, z)));
        const float2 pz = add(mul(mk(0.0f, -1.0f), mul(P.t20, fp)), mul(mk(0.0f, 1.0f), mul(P.t21, fm)));
        mtv::rotate(t, P);
        t.fp = add(t.fp, add(scl(ap, d_theta), scl(pp, d_phi)));
        t.fm = add(t.fm, add(scl(am, d_theta), scl(pm, d_phi)));
        t.z = add(t.z, add(scl(az, d_theta), scl(pz, d_phi)));
        mtv::rotate(x, P);
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs && first && live) {
      float sp, cp;
      sincosf(-ev.phase, &sp, &cp);
      const float2 c = scl(mul(factor, mk(cp, sp)), m0);
      const float d_phi = q.d_phase ? q.d_phase[e] : 0.0f;
      const float2 s = mul(c, x.fp);
      q.tangent[atom * p.outputs + ev.out] = add(mul(c, t.fp), scl(mk(s.y, -s.x), d_phi));
    }
    if (act & POST_SHIFT) {
      shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
      shift_complex<L>(t.fp, t.fm, sub, lane, beyond);
    }
    if (act & SPOIL_AFTER) {
      x.fp = x.fm = t.fp = t.fm = mk(0.0f, 0.0f);
    } else if (act & SHIFT_AFTER) {
      shift_complex<L>(x.fp, x.fm, sub, lane, beyond);
      shift_complex<L>(t.fp, t.fm, sub, lane, beyond);
    }
  }
}

// Per-event partials summed over the waves in a fixed order, added to the
// gradients with `accumulate` (a launch after the first).
__global__ void semisolid_sum_partials(const float* __restrict__ partial, int64_t waves, int64_t events,
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

}  // namespace tsim
