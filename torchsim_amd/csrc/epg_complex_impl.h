// Complex EPG kernel template for gfx950: single pool or one chemically
// exchanging pool, per-order diffusion damping, off-resonance, RF phase, with
// forward-mode tangents along T1, T2 and B1 carried in the same walk.
//
// What it computes is the reference's complex state machine
// (src/torchsim/_epg_cpu.cpp:4216-4566 forward, :2469-3005 JVP; two-pool steps
// :875-932 longitudinal and :1706-1764 transverse; rotate :2176-2208; shift
// :2164-2174; damping :677-715), restricted to instantaneous pulses.
//
// Layout: one configuration order per lane, an atom on L = 4..64 adjacent
// lanes (orders >= S stay zero). The gradient shift is a cross-lane move:
// DPP row shifts inside 16-lane rows, ds_bpermute beyond. Every per-atom
// operator that depends only on the interval (relaxation factors, the 2x2
// exchange exponentials and their T1/T2 derivatives, off-resonance turn,
// diffusion factor of the lane's own order) is recomputed only when the
// interval changes.
#pragma once

#include <type_traits>

#include "epg_common.h"

namespace tsim {

struct ComplexArgs {
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
  int64_t stride_opt;  // optional tissue buffers: 1 per atom, 0 one value
  int64_t atoms;
  int64_t outputs;
  int states;
  int density;
  int inverting;
  int transmit;
  int off_axis;
  int diffusing;
  float2* signal;    // (atoms, outputs) complex64, optional
  float2* jacobian;  // (atoms, columns, outputs) complex64, optional
  float* out_real;   // reference planes, optional
  float* out_imag;
  float factor_re;
  float factor_im;
  int columns;
  int col_t1, col_t2, col_b1, col_m0;
  // Staged output writes (as the real kernel's): log2 of the samples per
  // flush and of the padded stream count; stage_k_log2 < 1 writes directly.
  int stage_k_log2;
  int stage_s_log2;
  int exchanging;  // one exchanging pool (the batch kernel picks the body per request)
};

// Bytes of one wave's output stage.
constexpr int COMPLEX_STAGE_BYTES_PER_WAVE = 8192;

namespace cx {

__device__ __forceinline__ float2 mk(float r, float i) { return make_float2(r, i); }
__device__ __forceinline__ float2 add(float2 a, float2 b) { return mk(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 sub(float2 a, float2 b) { return mk(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 mul(float2 a, float2 b) {
  return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 scl(float2 a, float s) { return mk(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cj(float2 a) { return mk(a.x, -a.y); }
__device__ __forceinline__ float2 neg(float2 a) { return mk(-a.x, -a.y); }
// multiply-accumulate: acc + a * b
__device__ __forceinline__ float2 mac(float2 acc, float2 a, float2 b) {
  return mk(fmaf(a.x, b.x, fmaf(-a.y, b.y, acc.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, acc.y)));
}
__device__ __forceinline__ float2 cexp(float2 z) {
  const float m = expf(z.x);
  float s, c;
  sincosf(z.y, &s, &c);
  return mk(m * c, m * s);
}
__device__ __forceinline__ float2 csqrt(float2 z) {
  const float r = hypotf(z.x, z.y);
  if (r == 0.0f) return mk(0.0f, 0.0f);
  float re = sqrtf(0.5f * (r + fabsf(z.x)));
  float im = 0.5f * z.y / re;
  if (z.x < 0.0f) {
    const float t = re;
    re = fabsf(im);
    im = copysignf(t, z.y);
  }
  return mk(re, im);
}
__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {
  const float d = b.x * b.x + b.y * b.y;
  return mk((a.x * b.x + a.y * b.y) / d, (a.y * b.x - a.x * b.y) / d);
}

// One row of a 2x2 operator on a pool pair, m0 f0 + m1 f1 (CONJ: with the
// entries conjugated); with REAL the entries' imaginary parts are zero and
// not read. The real form's fused multiply-adds are written out, so every
// instantiation rounds alike whatever else it computes.
template <bool REAL, bool CONJ>
__device__ __forceinline__ float2 row(float2 m0, float2 f0, float2 m1, float2 f1) {
  if constexpr (REAL) {
    return mk(fmaf(m1.x, f1.x, m0.x * f0.x), fmaf(m1.x, f1.y, m0.x * f0.y));
  } else if constexpr (CONJ) {
    return add(mul(cj(m0), f0), mul(cj(m1), f1));
  } else {
    return add(mul(m0, f0), mul(m1, f1));
  }
}

// The RF T-matrix on one pool's (F+, F-, Z) for a pulse whose phase is
// zero: c = cos^2(a/2) and h = sin^2(a/2) couple F+ and F- as real numbers,
// s = sin(a) couples them to Z as imaginary ones, k = cos(a) keeps Z. With
// (c, h, s, k) replaced by their derivatives along a, the derivative.
// The fused multiply-adds are written out (as in `row`).
__device__ __forceinline__ void rotate_zero_phase(float2 fp, float2 fm, float2 z, float c, float h, float s,
                                                  float k, float2& np, float2& nm, float2& nz) {
  np = mk(fmaf(s, z.y, fmaf(h, fm.x, c * fp.x)), fmaf(-s, z.x, fmaf(h, fm.y, c * fp.y)));
  nm = mk(fmaf(-s, z.y, fmaf(c, fm.x, h * fp.x)), fmaf(s, z.x, fmaf(c, fm.y, h * fp.y)));
  const float hs = 0.5f * s;
  nz = mk(fmaf(k, z.x, hs * (fp.y - fm.y)), fmaf(k, z.y, hs * (fm.x - fp.x)));
}

// The even functions of delta a 2x2 exponential needs, from delta^2:
// ch = cosh(delta), sh = sinh(delta)/delta, c2 = (cosh(delta) - sh)/delta^2.
// Series below |delta^2| < 0.05 (relative error < 1e-8 there).
__device__ __forceinline__ void even_terms(float2 d2, float2& ch, float2& sh, float2& c2) {
  const float mag = hypotf(d2.x, d2.y);
  if (mag < 0.05f) {
    const float2 d4 = mul(d2, d2);
    const float2 d6 = mul(d4, d2);
    ch = add(add(mk(1.0f, 0.0f), scl(d2, 0.5f)), add(scl(d4, 1.0f / 24.0f), scl(d6, 1.0f / 720.0f)));
    sh = add(add(mk(1.0f, 0.0f), scl(d2, 1.0f / 6.0f)), add(scl(d4, 1.0f / 120.0f), scl(d6, 1.0f / 5040.0f)));
    c2 = add(add(mk(1.0f / 3.0f, 0.0f), scl(d2, 1.0f / 30.0f)), add(scl(d4, 1.0f / 840.0f), scl(d6, 1.0f / 45360.0f)));
    return;
  }
  const float2 d = csqrt(d2);
  const float2 ep = cexp(d);
  const float2 em = cexp(neg(d));
  ch = scl(add(ep, em), 0.5f);
  const float2 sinh_d = scl(sub(ep, em), 0.5f);
  sh = cdiv(sinh_d, d);
  c2 = cdiv(sub(ch, sh), d2);
}

// 2x2 complex matrix exponential of M (row-major m00 m01 m10 m11):
// e^{tau} (cosh(delta) I + sinh(delta)/delta (M - tau I)), tau = tr/2,
// delta^2 = ((m00 - m11)/2)^2 + m01 m10 (Cayley-Hamilton for 2x2).
// With a direction E, also the derivative dX of the exponential along E.
template <bool WITH_DERIVATIVE>
__device__ void expm2(const float2 (&m)[4], const float2 (&e)[4], float2 (&x)[4], float2 (&dx)[4]) {
  const float2 tau = scl(add(m[0], m[3]), 0.5f);
  const float2 half = scl(sub(m[0], m[3]), 0.5f);
  const float2 d2 = add(mul(half, half), mul(m[1], m[2]));
  float2 ch, sh, c2;
  even_terms(d2, ch, sh, c2);
  const float2 et = cexp(tau);
  const float2 n0 = sub(m[0], tau), n3 = sub(m[3], tau);
  x[0] = mul(et, add(ch, mul(sh, n0)));
  x[1] = mul(et, mul(sh, m[1]));
  x[2] = mul(et, mul(sh, m[2]));
  x[3] = mul(et, add(ch, mul(sh, n3)));
  if constexpr (WITH_DERIVATIVE) {
    const float2 dtau = scl(add(e[0], e[3]), 0.5f);
    const float2 dd2 = add(mul(half, sub(e[0], e[3])), add(mul(e[1], m[2]), mul(m[1], e[2])));
    const float2 dch = scl(mul(sh, dd2), 0.5f);
    const float2 dsh = scl(mul(c2, dd2), 0.5f);
    // d(M - tau I) = E - dtau I
    const float2 en0 = sub(e[0], dtau), en3 = sub(e[3], dtau);
    dx[0] = add(mul(dtau, x[0]), mul(et, add(dch, add(mul(dsh, n0), mul(sh, en0)))));
    dx[1] = add(mul(dtau, x[1]), mul(et, add(mul(dsh, m[1]), mul(sh, e[1]))));
    dx[2] = add(mul(dtau, x[2]), mul(et, add(mul(dsh, m[2]), mul(sh, e[2]))));
    dx[3] = add(mul(dtau, x[3]), mul(et, add(dch, add(mul(dsh, n3), mul(sh, en3)))));
  }
}

}  // namespace cx

// Cross-lane moves of one value inside an atom's group of L lanes: the
// value of the lane below (from_lower) or above (from_upper). DPP row moves
// (row_shr:1 / row_shl:1) where every atom sits inside one 16-lane row (L a
// divisor of 16), whole-wave DPP moves (wave_shr:1 / wave_shl:1, GFX9) where
// atoms straddle rows; both give 0 past the ends of the row or wave, lanes
// whose value the shift discards (an atom's order 0 takes conj(F-), its top
// order drops, and the wave's last lanes are an atom's top or idle).
template <int L>
constexpr bool dpp_lanes() { return L <= 16 && 16 % L == 0; }

template <int L>
__device__ __forceinline__ float from_lower(float v, int lane) {
  (void)lane;
  constexpr int ctrl = dpp_lanes<L>() ? 0x111 : 0x138;
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, true));
}
template <int L>
__device__ __forceinline__ float from_upper(float v, int lane) {
  (void)lane;
  constexpr int ctrl = dpp_lanes<L>() ? 0x101 : 0x130;
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, true));
}

// One pool's transverse pair at this lane's order, and the gradient shift
// (reference shift, _epg_cpu.cpp:2164-2174): F- moves down an order, F+ up;
// F+[0] = conj(F-[0]); the top order is dropped.
template <int L>
__device__ __forceinline__ void shift_complex(float2& fp, float2& fm, int sub, int lane, bool beyond) {
  float2 up = cx::mk(from_lower<L>(fp.x, lane), from_lower<L>(fp.y, lane));
  float2 down = cx::mk(from_upper<L>(fm.x, lane), from_upper<L>(fm.y, lane));
  if (sub == L - 1) down = cx::mk(0.0f, 0.0f);
  fm = down;
  fp = sub == 0 ? cx::cj(down) : up;
  if (beyond) fp = cx::mk(0.0f, 0.0f);
}

constexpr int CT1 = 1;
constexpr int CT2 = 2;
constexpr int CB1 = 4;

constexpr int cpop(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1); }
constexpr int cslot(int mask, int bit) { return cpop(mask & (bit - 1)); }

// The per-lane state: one pool, or two (EXCH), value and NT tangent sets.
template <bool EXCH>
struct Order {
  float2 fp[EXCH ? 2 : 1];
  float2 fm[EXCH ? 2 : 1];
  float2 z[EXCH ? 2 : 1];
};

// The atoms of one wave: atom `first + lane / L` on lanes [L (lane / L), ...).
template <int L, int TM, bool EXCH>
__device__ __forceinline__ void complex_atoms(const ComplexArgs& p, int64_t first) {
  constexpr int NT = cpop(TM);
  constexpr int NTA = NT > 0 ? NT : 1;
  constexpr int NP = EXCH ? 2 : 1;
  constexpr int S1 = cslot(TM, CT1);
  constexpr int S2 = cslot(TM, CT2);
  constexpr int SB = cslot(TM, CB1);
  using namespace cx;

  // A wave holds 64 / L atoms on adjacent lanes; with L not a divisor of 64
  // the last 64 mod L lanes are idle (they run, but record nothing).
  constexpr int APW = 64 / L;
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t group = first + wave_group;
  const bool live = wave_group < APW && group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const bool beyond = sub >= p.states;  // padding order: held at zero
  const int64_t ia = atom * p.stride_opt;

  const float t1 = p.t1[atom];
  const float t2 = p.t2[atom];
  const float r1 = 1000.0f / t1, r2 = 1000.0f / t2;
  const float k1 = 1000.0f / (t1 * t1), k2 = 1000.0f / (t2 * t2);
  const float m0 = p.density ? p.m0[ia] : 1.0f;
  const float b1 = p.transmit ? p.b1[ia] : 1.0f;
  const float b1_phase = p.off_axis && p.b1_phase ? p.b1_phase[ia] : 0.0f;
  const float b0 = p.off_axis && p.b0 ? p.b0[ia] : 0.0f;
  const float eta = p.inverting ? p.inv_eff[ia] : 1.0f;
  const float drate = p.diffusing ? p.diffusion[ia] : 0.0f;
  float frac = 0.0f, kex = 0.0f, r1b = 0.0f, r2b = 0.0f, shift_hz = 0.0f;
  if constexpr (EXCH) {
    frac = p.pool_fraction[ia];
    kex = p.pool_exchange[ia];
    r1b = 1000.0f / p.t1_pool[ia];
    r2b = 1000.0f / p.t2_pool[ia];
    shift_hz = p.pool_shift ? p.pool_shift[ia] : 0.0f;
  }
  const float free_frac = 1.0f - frac;
  const float kab = kex * frac, kba = kex * free_frac;
  const bool real_transverse = shift_hz == 0.0f && b0 == 0.0f;
  const float order = (float)sub;

  Order<EXCH> v;
  Order<EXCH> t[NTA];
#pragma unroll
  for (int q = 0; q < NP; ++q) {
    v.fp[q] = v.fm[q] = v.z[q] = mk(0.0f, 0.0f);
#pragma unroll
    for (int k = 0; k < NTA; ++k) t[k].fp[q] = t[k].fm[q] = t[k].z[q] = mk(0.0f, 0.0f);
  }
  if (sub == 0) {
    v.z[0] = mk(EXCH ? free_frac : 1.0f, 0.0f);
    if constexpr (EXCH) v.z[1] = mk(frac, 0.0f);
  }

  // Interval operators, recomputed when the interval changes.
  float last_dt = -1.0f;
  float2 turn = mk(1.0f, 0.0f);   // e^{-2 pi i b0 dt} x transverse damping of this order
  float dlong = 1.0f;             // longitudinal damping of this order
  // single pool
  float e1 = 1.0f, e2 = 1.0f, de1 = 0.0f, de2 = 0.0f, om1 = 0.0f;
  // exchanging pool: transverse X (complex 2x2) and dX/dT2, longitudinal
  // Y (2x2) and dY/dT1, recovery r and dr/dT1; once formed for an interval,
  // X and dX carry this order's turn and Y and dY its longitudinal damping
  float2 X[4], dX[4];
  float Y[4], dY[4], rec[2], drec[2];
  // A clock of float32 durations gives one nominal interval a few values a
  // rounding apart, so consecutive intervals differ often by ~1e-5 of their
  // length. The operators over dt = base + delta are those over the base
  // interval times those over delta (each generator commutes with itself),
  // the latter to second order in |G delta| <= 1e-3 (remainder < 2e-10);
  // a larger step forms them afresh and becomes the base. An exchanging
  // pool's base operators before this order's turn and damping are the
  // atom's: they sit in LDS, one slot per atom of the wave (and one for its
  // idle lanes), the order's factors over the base in registers; a single
  // pool's base factors are a few scalars in registers. `reach` bounds the
  // generators' norm for every order of the atom (exchange, relaxation,
  // shift, B0, diffusion at the top order), so the atom's lanes step alike.
  constexpr int SLOTS = 64 / L + 1;
  constexpr int BASE_WORDS = 24;  // X (8), dX/dT2 (8), Y (4), dY/dT1 (4)
  __shared__ float4 interval_base[EXCH ? 4 * SLOTS * (BASE_WORDS / 4) : 1];
  float4* const base_slot = interval_base + ((threadIdx.x >> 6) * SLOTS + wave_group) * (BASE_WORDS / 4);
  float2 turn_b = mk(1.0f, 0.0f);
  float dlong_b = 1.0f;
  float base_dt = 0.0f;
  bool have_base = false;
  // recovery (Y - I) G^{-1} s, s = ((1-f) R1a, f R1b): u = G^{-1} s and its
  // T1 derivative du = G^{-1} (ds - dG u), ds = (-(1-f) k1, 0), dG = diag(k1, 0)
  float u0 = 0.0f, u1 = 0.0f, du0 = 0.0f, du1 = 0.0f, reach = 0.0f;
  // single pool: the base interval's factors
  float e1b = 1.0f, om1b = 0.0f, e2b = 1.0f;
  if constexpr (EXCH) {
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      X[i] = dX[i] = mk(0.0f, 0.0f);
      Y[i] = dY[i] = 0.0f;
    }
    rec[0] = rec[1] = drec[0] = drec[1] = 0.0f;
    const float g00 = -kab - r1, g01 = kba, g10 = kab, g11 = -kba - r1b;
    const float det = g00 * g11 - g01 * g10;
    const float s0 = free_frac * r1, s1 = frac * r1b;
    u0 = (g11 * s0 - g01 * s1) / det;
    u1 = (-g10 * s0 + g00 * s1) / det;
    const float w0 = -free_frac * k1 - k1 * u0;
    du0 = (g11 * w0) / det;
    du1 = (-g10 * w0) / det;
    const float top = (float)(L - 1);
    reach = kex + r1 + r2 + r1b + r2b + 2.0f * 3.14159265358979323846f * (fabsf(shift_hz) + fabsf(b0)) +
            (top * top + top + (1.0f / 3.0f)) * drate;
  } else {
    const float top = (float)(L - 1);
    reach = r1 + r2 + 2.0f * 3.14159265358979323846f * fabsf(b0) + (top * top + top + (1.0f / 3.0f)) * drate;
  }

  // Output staging. Each wave keeps its atoms' last K samples of every
  // stream (signal, then the Jacobian columns) in LDS and writes them K at a
  // time, K consecutive samples of a row together, instead of one 8-byte
  // store per atom per sample on its own cache line. It holds while records
  // arrive in output order 0, 1, 2, ...; the first that does not flushes what
  // is staged and the rest are stored directly.
  extern __shared__ float2 complex_stage[];
  bool staging = p.stage_k_log2 > 0 && p.out_real == nullptr;
  const int k_log2 = p.stage_k_log2 > 0 ? p.stage_k_log2 : 0;
  const int s_log2 = p.stage_s_log2 > 0 ? p.stage_s_log2 : 0;
  const int stage_k = 1 << k_log2;
  float2* const stage = complex_stage + (((threadIdx.x >> 6) * APW) << (s_log2 + k_log2));
  int64_t nrec = 0;
  auto flush = [&](int count) __attribute__((always_inline)) {
    const int64_t g0 = group - wave_group;
    const int64_t o0 = nrec - count;
    const int total = APW << (s_log2 + k_log2);
    for (int i = 2 * lane; i < total; i += 128) {
      const int kk = i & (stage_k - 1);
      if (kk >= count) continue;
      const int row = i >> k_log2;
      const int64_t at = g0 + (row >> s_log2);
      if (at >= p.atoms) continue;
      const int stream = row & ((1 << s_log2) - 1);
      float2* dst;
      if (stream == 0) {
        if (!p.signal) continue;
        dst = p.signal + at * p.outputs + o0 + kk;
      } else {
        if (stream > p.columns || !p.jacobian) continue;
        dst = p.jacobian + (at * p.columns + (stream - 1)) * p.outputs + o0 + kk;
      }
      dst[0] = stage[i];
      if (kk + 1 < count) dst[1] = stage[i + 1];
    }
  };

  const EventCode* __restrict__ stream = p.events;
  const int64_t E = p.event_count;
  for (int64_t e = 0; e < E; ++e) {
    const EventCode ev = load_event(stream, e);
    const float dt = ev.dt;
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);

    if (dt != 0.0f) {
      if (dt != last_dt) {
        last_dt = dt;
        bool stepped = false;
        if constexpr (EXCH) {
          const float d = dt - base_dt;
          if (have_base && fabsf(d) * reach <= 1e-3f) {
            stepped = true;
            // this order's scalar factors over d: diffusion damping, B0 turn
            const float xo = -(order * order + order + (1.0f / 3.0f)) * drate * d;
            const float xl = -(order * order) * drate * d;
            const float yo = -2.0f * 3.14159265358979323846f * b0 * d;
            const float eo = 1.0f + xo + 0.5f * (xo * xo);
            const float2 tau = mul(turn_b, mk(eo * (1.0f - 0.5f * (yo * yo)), eo * yo));
            const float dl = dlong_b * (1.0f + xl + 0.5f * (xl * xl));
            float2 Xb[4], dXb[4];
            float Yb[4], dYb[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
              const float4 w = base_slot[i];
              Xb[2 * i] = mk(w.x, w.y);
              Xb[2 * i + 1] = mk(w.z, w.w);
            }
            {
              const float4 w = base_slot[4];
              Yb[0] = w.x;
              Yb[1] = w.y;
              Yb[2] = w.z;
              Yb[3] = w.w;
            }
            // transverse: E = I + A + A^2/2, A = G d; dE = dA + (dA A + A dA)/2, dA = diag(k2 d, 0)
            const float a00 = (-kab - r2) * d, a01 = kba * d, a10 = kab * d;
            const float2 a11 = mk((-kba - r2b) * d, -2.0f * 3.14159265358979323846f * shift_hz * d);
            const float2 half_tr = mk(1.0f + 0.5f * (a00 + a11.x), 0.5f * a11.y);
            const float2 e[4] = {mk(1.0f + a00 + 0.5f * (a00 * a00 + a01 * a10), 0.0f), scl(half_tr, a01),
                                 scl(half_tr, a10),
                                 add(mk(1.0f + 0.5f * (a01 * a10), 0.0f), add(a11, scl(mul(a11, a11), 0.5f)))};
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const float2 x0 = Xb[2 * r], x1 = Xb[2 * r + 1];
              X[2 * r] = mul(tau, add(mul(x0, e[0]), mul(x1, e[2])));
              X[2 * r + 1] = mul(tau, add(mul(x0, e[1]), mul(x1, e[3])));
            }
            if constexpr ((TM & CT2) != 0) {
#pragma unroll
              for (int i = 0; i < 2; ++i) {
                const float4 w = base_slot[2 + i];
                dXb[2 * i] = mk(w.x, w.y);
                dXb[2 * i + 1] = mk(w.z, w.w);
              }
              const float g = k2 * d;
              const float de00 = g * (1.0f + a00), de01 = 0.5f * g * a01, de10 = 0.5f * g * a10;
#pragma unroll
              for (int r = 0; r < 2; ++r) {
                const float2 x0 = Xb[2 * r], x1 = Xb[2 * r + 1];
                const float2 y0 = dXb[2 * r], y1 = dXb[2 * r + 1];
                const float2 c0 = add(add(mul(y0, e[0]), mul(y1, e[2])), add(scl(x0, de00), scl(x1, de10)));
                const float2 c1 = add(add(mul(y0, e[1]), mul(y1, e[3])), scl(x0, de01));
                dX[2 * r] = mul(tau, c0);
                dX[2 * r + 1] = mul(tau, c1);
              }
            }
            // longitudinal, real: the same with B = G d and dB = diag(k1 d, 0)
            const float b00 = (-kab - r1) * d, b01 = kba * d, b10 = kab * d, b11 = (-kba - r1b) * d;
            const float hl = 1.0f + 0.5f * (b00 + b11);
            const float l00 = 1.0f + b00 + 0.5f * (b00 * b00 + b01 * b10), l01 = b01 * hl, l10 = b10 * hl;
            const float l11 = 1.0f + b11 + 0.5f * (b11 * b11 + b01 * b10);
#pragma unroll
            for (int r = 0; r < 2; ++r) {
              const float y0 = Yb[2 * r], y1 = Yb[2 * r + 1];
              Y[2 * r] = y0 * l00 + y1 * l10;
              Y[2 * r + 1] = y0 * l01 + y1 * l11;
            }
            if constexpr ((TM & CT1) != 0) {
              const float4 w = base_slot[5];
              dYb[0] = w.x;
              dYb[1] = w.y;
              dYb[2] = w.z;
              dYb[3] = w.w;
              const float h = k1 * d;
              const float m00 = h * (1.0f + b00), m01 = 0.5f * h * b01, m10 = 0.5f * h * b10;
#pragma unroll
              for (int r = 0; r < 2; ++r) {
                const float y0 = Yb[2 * r], y1 = Yb[2 * r + 1];
                const float z0 = dYb[2 * r], z1 = dYb[2 * r + 1];
                dY[2 * r] = z0 * l00 + z1 * l10 + y0 * m00 + y1 * m10;
                dY[2 * r + 1] = z0 * l01 + z1 * l11 + y0 * m01;
              }
            }
            // recovery from the atom's Y, then this order's damping
            rec[0] = (Y[0] - 1.0f) * u0 + Y[1] * u1;
            rec[1] = Y[2] * u0 + (Y[3] - 1.0f) * u1;
            if constexpr ((TM & CT1) != 0) {
              drec[0] = dY[0] * u0 + dY[1] * u1 + (Y[0] - 1.0f) * du0 + Y[1] * du1;
              drec[1] = dY[2] * u0 + dY[3] * u1 + Y[2] * du0 + (Y[3] - 1.0f) * du1;
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              Y[i] *= dl;
              dY[i] *= dl;
            }
          }
        } else {
          const float d = dt - base_dt;
          if (have_base && fabsf(d) * reach <= 1e-3f) {
            stepped = true;
            // the base's factors times those over d, to second order
            const float x1 = -r1 * d, x2 = -r2 * d;
            const float g1 = x1 + 0.5f * (x1 * x1);  // e^{x1} - 1
            e1 = fmaf(e1b, g1, e1b);
            om1 = fmaf(-e1b, g1, om1b);                // 1 - e1 without cancellation
            e2 = e2b * (1.0f + x2 + 0.5f * (x2 * x2));
            de1 = e1 * (k1 * dt);
            de2 = e2 * (k2 * dt);
            const float xo = -(order * order + order + (1.0f / 3.0f)) * drate * d;
            const float xl = -(order * order) * drate * d;
            const float yo = -2.0f * 3.14159265358979323846f * b0 * d;
            const float eo = 1.0f + xo + 0.5f * (xo * xo);
            turn = mul(turn_b, mk(eo * (1.0f - 0.5f * (yo * yo)), eo * yo));
            dlong = dlong_b * (1.0f + xl + 0.5f * (xl * xl));
          }
        }
        if (!stepped) {
          const float bf = drate * dt;
          const float dtr = expf(-(order * order + order + (1.0f / 3.0f)) * bf);
          dlong = expf(-(order * order) * bf);
          float s, c;
          sincosf(-2.0f * 3.14159265358979323846f * b0 * dt, &s, &c);
          turn = mk(c * dtr, s * dtr);
          if constexpr (!EXCH) {
            // 1 - e1 from expm1: a rounded e1 would cost it five digits when dt << T1
            const float e1m = expm1f(-r1 * dt);
            e1 = 1.0f + e1m;
            om1 = -e1m;
            e2 = expf(-r2 * dt);
            de1 = e1 * (k1 * dt);
            de2 = e2 * (k2 * dt);
            e1b = e1;
            om1b = om1;
            e2b = e2;
            turn_b = turn;
            dlong_b = dlong;
            base_dt = dt;
            have_base = true;
          } else {
            // transverse generator K - diag(R2) - 2 pi i diag(0, df), times dt
            const float2 mt[4] = {mk((-kab - r2) * dt, 0.0f), mk(kba * dt, 0.0f), mk(kab * dt, 0.0f),
                                  mk((-kba - r2b) * dt, -2.0f * 3.14159265358979323846f * shift_hz * dt)};
            const float2 et[4] = {mk(k2 * dt, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
            expm2<(TM & CT2) != 0>(mt, et, X, dX);
            // longitudinal generator K - diag(R1)
            const float2 ml[4] = {mk((-kab - r1) * dt, 0.0f), mk(kba * dt, 0.0f), mk(kab * dt, 0.0f),
                                  mk((-kba - r1b) * dt, 0.0f)};
            const float2 el[4] = {mk(k1 * dt, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
            float2 yl[4], dyl[4];
            expm2<(TM & CT1) != 0>(ml, el, yl, dyl);
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              Y[i] = yl[i].x;
              dY[i] = dyl[i].x;
            }
            rec[0] = (Y[0] - 1.0f) * u0 + Y[1] * u1;
            rec[1] = Y[2] * u0 + (Y[3] - 1.0f) * u1;
            if constexpr ((TM & CT1) != 0) {
              drec[0] = dY[0] * u0 + dY[1] * u1 + (Y[0] - 1.0f) * du0 + Y[1] * du1;
              drec[1] = dY[2] * u0 + dY[3] * u1 + Y[2] * du0 + (Y[3] - 1.0f) * du1;
            }
            // the atom's operators become the base (every lane of the atom
            // writes the same words to its atom's slot and reads them back)
            base_slot[0] = make_float4(X[0].x, X[0].y, X[1].x, X[1].y);
            base_slot[1] = make_float4(X[2].x, X[2].y, X[3].x, X[3].y);
            if constexpr ((TM & CT2) != 0) {
              base_slot[2] = make_float4(dX[0].x, dX[0].y, dX[1].x, dX[1].y);
              base_slot[3] = make_float4(dX[2].x, dX[2].y, dX[3].x, dX[3].y);
            }
            base_slot[4] = make_float4(Y[0], Y[1], Y[2], Y[3]);
            if constexpr ((TM & CT1) != 0) base_slot[5] = make_float4(dY[0], dY[1], dY[2], dY[3]);
            turn_b = turn;
            dlong_b = dlong;
            base_dt = dt;
            have_base = true;
            // every event of this interval applies turn X (F+), its conjugate
            // (F-) and dlong Y (Z): fold the per-order factors in once here
#pragma unroll
            for (int i = 0; i < 4; ++i) {
              X[i] = mul(turn, X[i]);
              dX[i] = mul(turn, dX[i]);
              Y[i] *= dlong;
              dY[i] *= dlong;
            }
          }
        }
      }
      const float2 turn_c = cj(turn);
      if constexpr (!EXCH) {
        const float2 tp = scl(turn, e2), tm = scl(turn_c, e2);
        const float lo = e1 * dlong;
        // tangents first: they read the value before it relaxes
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          t[k].fp[0] = mul(t[k].fp[0], tp);
          t[k].fm[0] = mul(t[k].fm[0], tm);
          t[k].z[0] = scl(t[k].z[0], lo);
          if ((TM & CT2) != 0 && k == S2) {
            t[k].fp[0] = add(t[k].fp[0], mul(v.fp[0], scl(turn, de2)));
            t[k].fm[0] = add(t[k].fm[0], mul(v.fm[0], scl(turn_c, de2)));
          }
          if ((TM & CT1) != 0 && k == S1) {
            t[k].z[0] = add(t[k].z[0], scl(v.z[0], de1 * dlong));
            if (sub == 0) t[k].z[0].x -= de1;
          }
        }
        v.fp[0] = mul(v.fp[0], tp);
        v.fm[0] = mul(v.fm[0], tm);
        v.z[0] = scl(v.z[0], lo);
        if (sub == 0) v.z[0].x += om1;
      } else {
        // transverse: F+ <- (turn X) F+, F- <- conj(turn X) F-; longitudinal
        // Z <- (dlong Y) Z + recovery. Without B0 and pool shift the
        // transverse operators are real (the atom's own property, so an atom
        // takes the same path alone and beside others).
        auto relax = [&](auto real_tag) __attribute__((always_inline)) {
          constexpr bool R = decltype(real_tag)::value;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          float2 a0 = row<R, false>(X[0], t[k].fp[0], X[1], t[k].fp[1]);
          float2 a1 = row<R, false>(X[2], t[k].fp[0], X[3], t[k].fp[1]);
          float2 b0v = row<R, true>(X[0], t[k].fm[0], X[1], t[k].fm[1]);
          float2 b1v = row<R, true>(X[2], t[k].fm[0], X[3], t[k].fm[1]);
          float z0 = 0.0f, z1 = 0.0f;
          float2 zz0 = add(scl(t[k].z[0], Y[0]), scl(t[k].z[1], Y[1]));
          float2 zz1 = add(scl(t[k].z[0], Y[2]), scl(t[k].z[1], Y[3]));
          if ((TM & CT2) != 0 && k == S2) {
            a0 = add(a0, row<R, false>(dX[0], v.fp[0], dX[1], v.fp[1]));
            a1 = add(a1, row<R, false>(dX[2], v.fp[0], dX[3], v.fp[1]));
            b0v = add(b0v, row<R, true>(dX[0], v.fm[0], dX[1], v.fm[1]));
            b1v = add(b1v, row<R, true>(dX[2], v.fm[0], dX[3], v.fm[1]));
          }
          if ((TM & CT1) != 0 && k == S1) {
            zz0 = add(zz0, add(scl(v.z[0], dY[0]), scl(v.z[1], dY[1])));
            zz1 = add(zz1, add(scl(v.z[0], dY[2]), scl(v.z[1], dY[3])));
            if (sub == 0) {
              z0 = drec[0];
              z1 = drec[1];
            }
          }
          t[k].fp[0] = a0;
          t[k].fp[1] = a1;
          t[k].fm[0] = b0v;
          t[k].fm[1] = b1v;
          t[k].z[0] = zz0;
          t[k].z[1] = zz1;
          t[k].z[0].x += z0;
          t[k].z[1].x += z1;
        }
        const float2 a0 = row<R, false>(X[0], v.fp[0], X[1], v.fp[1]);
        const float2 a1 = row<R, false>(X[2], v.fp[0], X[3], v.fp[1]);
        const float2 b0v = row<R, true>(X[0], v.fm[0], X[1], v.fm[1]);
        const float2 b1v = row<R, true>(X[2], v.fm[0], X[3], v.fm[1]);
        const float2 zz0 = add(scl(v.z[0], Y[0]), scl(v.z[1], Y[1]));
        const float2 zz1 = add(scl(v.z[0], Y[2]), scl(v.z[1], Y[3]));
        v.fp[0] = a0;
        v.fp[1] = a1;
        v.fm[0] = b0v;
        v.fm[1] = b1v;
        v.z[0] = zz0;
        v.z[1] = zz1;
        if (sub == 0) {
          v.z[0].x += rec[0];
          v.z[1].x += rec[1];
        }
        };
        if (real_transverse) {
          relax(std::true_type{});
        } else {
          relax(std::false_type{});
        }
      }
    }

    if (act & PRE_SHIFT) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        shift_complex<L>(v.fp[q], v.fm[q], sub, lane, beyond);
#pragma unroll
        for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp[q], t[k].fm[q], sub, lane, beyond);
      }
    }

    if (kind == KIND_RF) {
      if (act & INVERSION) {
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          v.z[q] = scl(v.z[q], -eta);
#pragma unroll
          for (int k = 0; k < NT; ++k) t[k].z[q] = scl(t[k].z[q], -eta);
        }
      } else {
        const float alpha = ev.flip * b1;
        float sa, va;  // sin(alpha), 1 - cos(alpha)
        if (p.transmit) {
          const float h = sinf(0.5f * alpha);
          sa = sinf(alpha);
          va = 2.0f * (h * h);
        } else {
          sa = ev.sine;
          va = ev.versine;
        }
        const float ca = 1.0f - va;
        const float shh = 0.5f * va, chh = 1.0f - shh;
        const float dch = -0.5f * sa, dsh = 0.5f * sa;
        const float dca = -sa;
        const float dal = ev.flip;  // d alpha / d B1
        if (ev.phase + b1_phase == 0.0f) {
          // zero phase: real F+/F- couplings, imaginary Z couplings
#pragma unroll
          for (int q = 0; q < NP; ++q) {
            const float2 fp = v.fp[q], fm = v.fm[q], z = v.z[q];
#pragma unroll
            for (int k = 0; k < NT; ++k) {
              float2 np, nm, nz;
              rotate_zero_phase(t[k].fp[q], t[k].fm[q], t[k].z[q], chh, shh, sa, ca, np, nm, nz);
              if ((TM & CB1) != 0 && k == SB) {
                float2 ep, em, ez;
                rotate_zero_phase(fp, fm, z, dch, dsh, ca, dca, ep, em, ez);
                np = add(np, scl(ep, dal));
                nm = add(nm, scl(em, dal));
                nz = add(nz, scl(ez, dal));
              }
              t[k].fp[q] = np;
              t[k].fm[q] = nm;
              t[k].z[q] = nz;
            }
            rotate_zero_phase(fp, fm, z, chh, shh, sa, ca, v.fp[q], v.fm[q], v.z[q]);
          }
        } else {
        float sp, cp;
        sincosf(ev.phase + b1_phase, &sp, &cp);
        const float2 e1p = mk(cp, sp);            // e^{i phi}
        const float2 e2p = mul(e1p, e1p);         // e^{2 i phi}
        // T-matrix (reference rotate, _epg_cpu.cpp:2176-2208)
        const float2 t01 = scl(e2p, shh);
        const float2 t02 = mul(mk(0.0f, -sa), e1p);
        const float2 t10 = cj(t01);
        const float2 t12 = mul(mk(0.0f, sa), cj(e1p));
        const float2 t20 = mul(mk(0.0f, -0.5f * sa), cj(e1p));
        const float2 t21 = mul(mk(0.0f, 0.5f * sa), e1p);
        // derivatives along the angle, per unit angle
        const float2 d01 = scl(e2p, dsh);
        const float2 d02 = mul(mk(0.0f, -ca), e1p);
        const float2 d10 = cj(d01);
        const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
        const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
        const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
#pragma unroll
        for (int q = 0; q < NP; ++q) {
          const float2 fp = v.fp[q], fm = v.fm[q], z = v.z[q];
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            const float2 dp = t[k].fp[q], dm = t[k].fm[q], dz = t[k].z[q];
            float2 np = mac(mac(scl(dp, chh), t01, dm), t02, dz);
            float2 nm = mac(mac(mul(t10, dp), mk(chh, 0.0f), dm), t12, dz);
            float2 nz = mac(mac(mul(t20, dp), t21, dm), mk(ca, 0.0f), dz);
            if ((TM & CB1) != 0 && k == SB) {
              np = add(np, scl(mac(mac(scl(fp, dch), d01, fm), d02, z), dal));
              nm = add(nm, scl(mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z), dal));
              nz = add(nz, scl(mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z), dal));
            }
            t[k].fp[q] = np;
            t[k].fm[q] = nm;
            t[k].z[q] = nz;
          }
          v.fp[q] = mac(mac(scl(fp, chh), t01, fm), t02, z);
          v.fm[q] = mac(mac(mul(t10, fp), mk(chh, 0.0f), fm), t12, z);
          v.z[q] = mac(mac(mul(t20, fp), t21, fm), mk(ca, 0.0f), z);
        }
        }
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs) {
      if (staging && ev.out != nrec) {
        if ((nrec & (stage_k - 1)) != 0) flush((int)(nrec & (stage_k - 1)));
        staging = false;
      }
      if (staging) {
        if (sub == 0 && live) {
          float sp = 0.0f, cp = 1.0f;
          if (ev.phase != 0.0f) sincosf(-ev.phase, &sp, &cp);
          const float2 demod = mk(cp, sp);
          float2 total = v.fp[0];
          if constexpr (EXCH) total = add(total, v.fp[1]);
          const float2 sample = mul(scl(total, m0), demod);
          const float2 factor = mk(p.factor_re, p.factor_im);
          float2* row = stage + ((wave_group << s_log2) << k_log2) + (nrec & (stage_k - 1));
          row[0] = mul(factor, sample);
          const float2 fm0 = mul(factor, scl(demod, m0));
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            float2 d = t[k].fp[0];
            if constexpr (EXCH) d = add(d, t[k].fp[1]);
            int col = -1;
            if ((TM & CT1) != 0 && k == S1) col = p.col_t1;
            if ((TM & CT2) != 0 && k == S2) col = p.col_t2;
            if ((TM & CB1) != 0 && k == SB) col = p.col_b1;
            if (col >= 0) row[(1 + col) << k_log2] = mul(fm0, d);
          }
          if (p.col_m0 >= 0) row[(1 + p.col_m0) << k_log2] = mul(factor, mul(total, demod));
        }
        ++nrec;
        if ((nrec & (stage_k - 1)) == 0) flush(stage_k);
      } else if (sub == 0 && live) {
        float sp = 0.0f, cp = 1.0f;
        if (ev.phase != 0.0f) sincosf(-ev.phase, &sp, &cp);
        const float2 demod = mk(cp, sp);
        const int64_t o = ev.out;
        float2 total = v.fp[0];
        if constexpr (EXCH) total = add(total, v.fp[1]);
        const float2 sample = mul(scl(total, m0), demod);
        if (p.out_real) {
          p.out_real[atom * p.outputs + o] = sample.x;
          p.out_imag[atom * p.outputs + o] = sample.y;
        }
        const float2 factor = mk(p.factor_re, p.factor_im);
        if (p.signal) p.signal[atom * p.outputs + o] = mul(factor, sample);
        if (p.jacobian) {
          float2* row = p.jacobian + atom * (int64_t)p.columns * p.outputs + o;
          const float2 fm0 = mul(factor, scl(demod, m0));
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            float2 d = t[k].fp[0];
            if constexpr (EXCH) d = add(d, t[k].fp[1]);
            int col = -1;
            if ((TM & CT1) != 0 && k == S1) col = p.col_t1;
            if ((TM & CT2) != 0 && k == S2) col = p.col_t2;
            if ((TM & CB1) != 0 && k == SB) col = p.col_b1;
            if (col >= 0) row[(int64_t)col * p.outputs] = mul(fm0, d);
          }
          if (p.col_m0 >= 0) row[(int64_t)p.col_m0 * p.outputs] = mul(factor, mul(total, demod));
        }
      }
    }

    if (act & POST_SHIFT) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        shift_complex<L>(v.fp[q], v.fm[q], sub, lane, beyond);
#pragma unroll
        for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp[q], t[k].fm[q], sub, lane, beyond);
      }
    }
    if (act & SPOIL_AFTER) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        v.fp[q] = v.fm[q] = mk(0.0f, 0.0f);
#pragma unroll
        for (int k = 0; k < NT; ++k) t[k].fp[q] = t[k].fm[q] = mk(0.0f, 0.0f);
      }
    } else if (act & SHIFT_AFTER) {
#pragma unroll
      for (int q = 0; q < NP; ++q) {
        shift_complex<L>(v.fp[q], v.fm[q], sub, lane, beyond);
#pragma unroll
        for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp[q], t[k].fm[q], sub, lane, beyond);
      }
    }
  }
  if (staging && (nrec & (stage_k - 1)) != 0) flush((int)(nrec & (stage_k - 1)));
}

// Three waves per SIMD: the exchanging pool's two paths (complex operators,
// and real ones where B0 and the pool shift vanish) fit 168 VGPRs.
template <int L, int TM, bool EXCH>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
complex_kernel(const ComplexArgs p) {
  constexpr int APW = 64 / L;
  complex_atoms<L, TM, EXCH>(p, (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * APW);
}

// One work block of an uneven batch: a request and its first atom; a block of
// 256 threads takes 4 (64 / L) atoms from there.
struct ComplexWork {
  int32_t request;
  int32_t pad;
  int64_t first_atom;
};

// The uneven-batch launch of the complex kernel: a persistent grid walks the
// work blocks; each block runs its request's single- or exchanging-pool body
// (uniform per block), so requests with different pool counts, sequences,
// state counts (padded to the batch's lanes) and sizes share one launch. Every
// atom runs the arithmetic it would run alone.
template <int L, int TM>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3)))
complex_batch_kernel(const ComplexArgs* __restrict__ args, const ComplexWork* __restrict__ work, int64_t blocks) {
  constexpr int APW = 64 / L;
  for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
    const ComplexWork w = work[b];
    const ComplexArgs p = args[w.request];
    const int64_t first = w.first_atom + (int64_t)(threadIdx.x >> 6) * APW;
    if (p.exchanging) {
      complex_atoms<L, TM, true>(p, first);
    } else {
      complex_atoms<L, TM, false>(p, first);
    }
  }
}

using ComplexBatchLaunch = void (*)(const ComplexArgs*, const ComplexWork*, int64_t, dim3, dim3, size_t,
                                    hipStream_t);

template <int L, int TM>
void launch_complex_batch(const ComplexArgs* args, const ComplexWork* work, int64_t blocks, dim3 grid, dim3 block,
                          size_t shared, hipStream_t stream) {
  hipLaunchKernelGGL((complex_batch_kernel<L, TM>), grid, block, shared, stream, args, work, blocks);
}

template <int L>
ComplexBatchLaunch pick_complex_batch_mask(int tm) {
  switch (tm & 7) {
    case 0: return &launch_complex_batch<L, 0>;
    case 1: return &launch_complex_batch<L, 1>;
    case 2: return &launch_complex_batch<L, 2>;
    case 3: return &launch_complex_batch<L, 3>;
    case 4: return &launch_complex_batch<L, 4>;
    case 5: return &launch_complex_batch<L, 5>;
    case 6: return &launch_complex_batch<L, 6>;
    default: return &launch_complex_batch<L, 7>;
  }
}

// Launchers take the dynamic LDS bytes of the output stage.
using ComplexLaunch = void (*)(const ComplexArgs&, dim3, dim3, size_t, hipStream_t);

template <int L, int TM, bool EXCH>
void launch_complex(const ComplexArgs& args, dim3 grid, dim3 block, size_t shared, hipStream_t stream) {
  hipLaunchKernelGGL((complex_kernel<L, TM, EXCH>), grid, block, shared, stream, args);
}

template <int L, int TM>
ComplexLaunch pick_complex_exch(bool exch) {
  return exch ? &launch_complex<L, TM, true> : &launch_complex<L, TM, false>;
}

template <int L>
ComplexLaunch pick_complex_mask(int tm, bool exch) {
  switch (tm & 7) {
    case 0: return pick_complex_exch<L, 0>(exch);
    case 1: return pick_complex_exch<L, 1>(exch);
    case 2: return pick_complex_exch<L, 2>(exch);
    case 3: return pick_complex_exch<L, 3>(exch);
    case 4: return pick_complex_exch<L, 4>(exch);
    case 5: return pick_complex_exch<L, 5>(exch);
    case 6: return pick_complex_exch<L, 6>(exch);
    default: return pick_complex_exch<L, 7>(exch);
  }
}

// Lanes per atom the complex kernel is built for, one translation unit each.
#define TSIM_COMPLEX_LANES(X) X(4) X(8) X(10) X(12) X(16) X(20) X(32) X(64)

#define TSIM_DECLARE_COMPLEX(LANES)                            \
  ComplexLaunch pick_complex_##LANES(int tm, bool exch);       \
  ComplexBatchLaunch pick_complex_batch_##LANES(int tm);
TSIM_COMPLEX_LANES(TSIM_DECLARE_COMPLEX)
#undef TSIM_DECLARE_COMPLEX

}  // namespace tsim
