// The complex EPG state machine with a semisolid (magnetization-transfer)
// pool and/or slice-profiled pulses, for gfx950: signal and fused Jacobian
// columns (T1, T2, B1, and M0 by its closed form), complex64 in the public
// layout.
//
// It computes the reference's `simulate_range<INSTANT, SEMISOLID>` and its
// forward-mode tangents (src/torchsim/_epg_cpu.cpp:4216-4566 with pool kind
// 1; `two_pool_step`, :875-932; the saturation read, :4463-4480 and the
// lineshape's Hermite read, :1888-1910). The bound pool holds only
// longitudinal states: F+ and F- stay the free water's (one pool: e2, the
// off-resonance turn, the per-order diffusion damping), while each order's
// (Z_free, Z_bound) pair relaxes and exchanges through the 2x2 exponential
// of [[-k f - R1, k (1 - f)], [k f, -k (1 - f) - R1b]] (f the bound fraction,
// k the exchange rate) and recovers toward (1 - f, f) at order 0. A pulse
// rotates the free water and saturates the bound pool by
// exp(s theta^2 G(rf_frequency - b0)), s the pulse's deposited power per
// flip squared (an event buffer), theta = flip * b1 and G the absorption
// lineshape read from a table; an inversion leaves the bound pool alone.
// Without a bound pool the pair reduces to the free water's own Z.
//
// A profiled pulse (PROFILED) turns the states through the Cayley-Klein pair
// (a, b) a shaped pulse performs at the atom's slice position, read by cubic
// Hermite from a table over (position, effective flip theta = flip * b1) at
// zero RF phase, the phase applied as b exp(-i phi) (reference
// rotate_spinor / profile_pair / table_row, src/torchsim/_epg_cpu.cpp:
// 2041-2049, 2107-2160, 2273-2304, 4491-4520). Atoms are laid out voxel-major
// with one atom per tabulated position; the caller averages over positions.
//
// Layout as the complex kernel: one configuration order per lane, an atom on
// L adjacent lanes; orders past the state count are padding held at zero.
// Records are stored directly (a model outside the throughput path).
#pragma once

#include "epg_complex_impl.h"

namespace tsim {

struct SemisolidArgs {
  const EventCode* events;
  int64_t event_count;
  const float* saturation;    // per event, 1/s per rad^2 (packed event buffer 7)
  const float* rf_frequency;  // per event, Hz (packed event buffer 8)
  const float* t1;
  const float* t2;
  const float* m0;
  const float* b1;
  const float* b1_phase;
  const float* b0;
  const float* inv_eff;
  const float* diffusion;
  const float* bound_fraction;
  const float* bound_exchange;
  const float* t1_bound;
  const float* lineshape;  // (bins, 2): value, slope per knot
  int64_t lineshape_bins;
  float lineshape_step;
  const float* profile;  // (locations, bins, 8): a, b, da/dtheta, db/dtheta as (re, im)
  int64_t profile_bins;
  float profile_step;
  int64_t locations;
  int64_t stride_opt;
  int64_t atoms;
  int64_t outputs;
  int states;
  int density;
  int inverting;
  int transmit;
  int off_axis;
  int diffusing;
  float factor_re;
  float factor_im;
  int columns;
  int col_t1, col_t2, col_b1, col_m0;
  float2* signal;    // (atoms, outputs)
  float2* jacobian;  // (atoms, columns, outputs) or null
};

// The lineshape at |offset| by cubic Hermite between knots, clamped to the
// table (the reference's read, _epg_cpu.cpp:1888-1910).
__device__ __forceinline__ float lineshape_at(const SemisolidArgs& p, float offset_hz) {
  const float last = (float)(p.lineshape_bins - 1);
  const float scaled = fminf(fabsf(offset_hz) / p.lineshape_step, last);
  const float lower = fminf(floorf(scaled), last - 1.0f);
  const float u = scaled - lower;
  const float* near = p.lineshape + (int64_t)lower * 2;
  const float u2 = u * u, u3 = u2 * u;
  const float h = p.lineshape_step;
  return (2.0f * u3 - 3.0f * u2 + 1.0f) * near[0] + (u3 - 2.0f * u2 + u) * h * near[1] +
         (-2.0f * u3 + 3.0f * u2) * near[2] + (u3 - u2) * h * near[3];
}

namespace mt {

using namespace cx;

// One order's state: the free water's F+, F-, Z and the bound pool's Z.
struct Order {
  float2 fp, fm, z, zb;
};

}  // namespace mt

// The pair (a, b) a profiled pulse performs at position row `loc` and
// effective flip theta, and its derivative along theta, by cubic Hermite
// between the knots bracketing theta (clamped to the table).
__device__ __forceinline__ void profile_pair(const SemisolidArgs& p, int64_t loc, float theta, float2& a, float2& b,
                                             float2& da, float2& db) {
  const float h = p.profile_step;
  const float last = (float)(p.profile_bins - 1);
  const float scaled = fminf(fmaxf(theta / h, 0.0f), last);
  const float lower = fminf(floorf(scaled), last - 1.0f);
  const float u = scaled - lower;
  const float* near = p.profile + (loc * p.profile_bins + (int64_t)lower) * 8;
  const float* far = near + 8;
  const float u2 = u * u, u3 = u2 * u;
  const float h00 = 2.0f * u3 - 3.0f * u2 + 1.0f, h10 = (u3 - 2.0f * u2 + u) * h;
  const float h01 = -2.0f * u3 + 3.0f * u2, h11 = (u3 - u2) * h;
  const float g00 = (6.0f * u2 - 6.0f * u) / h, g10 = 3.0f * u2 - 4.0f * u + 1.0f;
  const float g01 = (6.0f * u - 6.0f * u2) / h, g11 = 3.0f * u2 - 2.0f * u;
  auto read = [&](int i, float w00, float w10, float w01, float w11) {
    return make_float2(w00 * near[i] + w10 * near[i + 4] + w01 * far[i] + w11 * far[i + 4],
                       w00 * near[i + 1] + w10 * near[i + 5] + w01 * far[i + 1] + w11 * far[i + 5]);
  };
  a = read(0, h00, h10, h01, h11);
  b = read(2, h00, h10, h01, h11);
  da = read(0, g00, g10, g01, g11);
  db = read(2, g00, g10, g01, g11);
}

template <int L, int TM, bool PROFILED>
__global__ void __launch_bounds__(256) semisolid_kernel(const SemisolidArgs p) {
  using namespace cx;
  using mt::Order;
  constexpr int NT = cpop(TM);
  constexpr int NTA = NT > 0 ? NT : 1;
  constexpr int S1 = cslot(TM, CT1);
  constexpr int S2 = cslot(TM, CT2);
  constexpr int SB = cslot(TM, CB1);
  constexpr int APW = 64 / L;
  const int lane = (int)(threadIdx.x & 63);
  const int wave_group = lane / L;
  const int sub = lane - wave_group * L;
  const int64_t group = (((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6) * APW + wave_group;
  const bool live = wave_group < APW && group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;
  const bool beyond = sub >= p.states;
  const bool first = sub == 0;
  const int64_t ia = atom * p.stride_opt;

  const float t1 = p.t1[atom], t2 = p.t2[atom];
  const float r1 = 1000.0f / t1, r2 = 1000.0f / t2;
  const float k1 = 1000.0f / (t1 * t1), k2 = 1000.0f / (t2 * t2);
  const float m0 = p.density ? p.m0[ia] : 1.0f;
  const float b1 = p.transmit ? p.b1[ia] : 1.0f;
  const float b1_phase = p.off_axis && p.b1_phase ? p.b1_phase[ia] : 0.0f;
  const float b0 = p.off_axis && p.b0 ? p.b0[ia] : 0.0f;
  const float eta = p.inverting ? p.inv_eff[ia] : 1.0f;
  const float drate = p.diffusing ? p.diffusion[ia] : 0.0f;
  // without a bound pool: f = 0, and the pair's second row holds nothing
  const float f = p.bound_fraction ? p.bound_fraction[ia] : 0.0f;
  const float kex = p.bound_exchange ? p.bound_exchange[ia] : 0.0f;
  const float r1b = p.t1_bound ? 1000.0f / p.t1_bound[ia] : r1;
  const float free = 1.0f - f;
  const float kab = kex * f, kba = kex * free;
  const float order = (float)sub;

  Order v;
  Order t[NTA];
  v.fp = v.fm = v.z = v.zb = mk(0.0f, 0.0f);
#pragma unroll
  for (int k = 0; k < NTA; ++k) t[k].fp = t[k].fm = t[k].z = t[k].zb = mk(0.0f, 0.0f);
  if (first) {
    v.z = mk(free, 0.0f);
    v.zb = mk(f, 0.0f);
  }

  // interval operators, recomputed when the interval changes
  float last_dt = -1.0f;
  float2 tp = mk(1.0f, 0.0f), tm = mk(1.0f, 0.0f), dtp = mk(0.0f, 0.0f), dtm = mk(0.0f, 0.0f);
  float Y[4] = {1.0f, 0.0f, 0.0f, 1.0f}, dY[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  float rec[2] = {0.0f, 0.0f}, drec[2] = {0.0f, 0.0f};

  const float2 factor = mk(p.factor_re, p.factor_im);
  const int64_t E = p.event_count;
  for (int64_t e = 0; e < E; ++e) {
    const EventCode ev = load_event(p.events, e);
    const float dt = ev.dt;
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);

    if (dt != 0.0f) {
      if (dt != last_dt) {
        last_dt = dt;
        const float bf = drate * dt;
        const float dtr = expf(-(order * order + order + (1.0f / 3.0f)) * bf);
        const float dlong = expf(-(order * order) * bf);
        float s, c;
        sincosf(-2.0f * 3.14159265358979323846f * b0 * dt, &s, &c);
        const float e2 = expf(-r2 * dt);
        tp = mk(c * dtr * e2, s * dtr * e2);
        tm = cj(tp);
        dtp = scl(tp, k2 * dt);
        dtm = cj(dtp);
        // (Z_free, Z_bound): relaxation and exchange, and its derivative along T1
        const float2 ml[4] = {mk((-kab - r1) * dt, 0.0f), mk(kba * dt, 0.0f), mk(kab * dt, 0.0f),
                              mk((-kba - r1b) * dt, 0.0f)};
        const float2 el[4] = {mk(k1 * dt, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f), mk(0.0f, 0.0f)};
        float2 yl[4], dyl[4];
        expm2<(TM & CT1) != 0>(ml, el, yl, dyl);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          Y[i] = yl[i].x;
          dY[i] = (TM & CT1) != 0 ? dyl[i].x : 0.0f;
        }
        // recovery toward the equilibrium split (1 - f, f)
        rec[0] = free - (Y[0] * free + Y[1] * f);
        rec[1] = f - (Y[2] * free + Y[3] * f);
        drec[0] = -(dY[0] * free + dY[1] * f);
        drec[1] = -(dY[2] * free + dY[3] * f);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          Y[i] *= dlong;
          dY[i] *= dlong;
        }
      }
      // tangents first: they read the value before it relaxes
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        float2 np = mul(t[k].fp, tp), nm = mul(t[k].fm, tm);
        float2 nz = add(scl(t[k].z, Y[0]), scl(t[k].zb, Y[1]));
        float2 nb = add(scl(t[k].z, Y[2]), scl(t[k].zb, Y[3]));
        if ((TM & CT2) != 0 && k == S2) {
          np = add(np, mul(v.fp, dtp));
          nm = add(nm, mul(v.fm, dtm));
        }
        if ((TM & CT1) != 0 && k == S1) {
          nz = add(nz, add(scl(v.z, dY[0]), scl(v.zb, dY[1])));
          nb = add(nb, add(scl(v.z, dY[2]), scl(v.zb, dY[3])));
          if (first) {
            nz.x += drec[0];
            nb.x += drec[1];
          }
        }
        t[k].fp = np;
        t[k].fm = nm;
        t[k].z = nz;
        t[k].zb = nb;
      }
      v.fp = mul(v.fp, tp);
      v.fm = mul(v.fm, tm);
      const float2 z = add(scl(v.z, Y[0]), scl(v.zb, Y[1]));
      const float2 zb = add(scl(v.z, Y[2]), scl(v.zb, Y[3]));
      v.z = z;
      v.zb = zb;
      if (first) {
        v.z.x += rec[0];
        v.zb.x += rec[1];
      }
    }

    if (act & PRE_SHIFT) {
      shift_complex<L>(v.fp, v.fm, sub, lane, beyond);
#pragma unroll
      for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp, t[k].fm, sub, lane, beyond);
    }

    if (kind == KIND_RF) {
      if (act & INVERSION) {
        // the free water inverts; the bound pool is saturated rather than
        // turned over, by nothing this event carries
        v.z = scl(v.z, -eta);
#pragma unroll
        for (int k = 0; k < NT; ++k) t[k].z = scl(t[k].z, -eta);
      } else {
        const float theta = ev.flip * b1;
        // the bound pool absorbs the pulse's power at its offset
        const float deposit = p.bound_fraction ? p.saturation[e] : 0.0f;
        if (deposit != 0.0f) {
          const float g = lineshape_at(p, p.rf_frequency[e] - b0);
          const float absorbed = expf(deposit * theta * theta * g);
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            float2 nb = scl(t[k].zb, absorbed);
            if ((TM & CB1) != 0 && k == SB) nb = add(nb, scl(v.zb, absorbed * deposit * 2.0f * theta * ev.flip * g));
            t[k].zb = nb;
          }
          v.zb = scl(v.zb, absorbed);
        }
        if constexpr (PROFILED) {
          float2 a, bz, da, dbz;
          profile_pair(p, atom % p.locations, theta, a, bz, da, dbz);
          float sp, cp;
          sincosf(-(ev.phase + b1_phase), &sp, &cp);
          const float2 spin = mk(cp, sp);
          const float2 b = mul(bz, spin), db = mul(dbz, spin);
          // rows (F+, F-, Z) of the spinor's T-matrix (reference rotate_spinor)
          const float2 ca = cj(a), cb = cj(b);
          const float2 t00 = mul(ca, ca), t01 = neg(mul(cb, cb)), t02 = scl(cj(mul(a, b)), -2.0f);
          const float2 t10 = neg(mul(b, b)), t11 = mul(a, a), t12 = scl(mul(a, b), -2.0f);
          const float2 t20 = mul(ca, b), t21 = mul(a, cb);
          const float t22 = (a.x * a.x + a.y * a.y) - (b.x * b.x + b.y * b.y);
          const float2 fp = v.fp, fm = v.fm, z = v.z;
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            const float2 dp = t[k].fp, dm = t[k].fm, dz = t[k].z;
            float2 np = add(add(mul(t00, dp), mul(t01, dm)), mul(t02, dz));
            float2 nm = add(add(mul(t10, dp), mul(t11, dm)), mul(t12, dz));
            float2 nz = add(add(mul(t20, dp), mul(t21, dm)), scl(dz, t22));
            if ((TM & CB1) != 0 && k == SB) {
              // d/dB1 = flip * d/dtheta of every entry, by the product rule
              const float2 cda = cj(da), cdb = cj(db);
              const float2 u00 = scl(mul(ca, cda), 2.0f), u01 = scl(mul(cb, cdb), -2.0f);
              const float2 ab = add(mul(da, b), mul(a, db));
              const float2 u02 = scl(cj(ab), -2.0f);
              const float2 u10 = scl(mul(b, db), -2.0f), u11 = scl(mul(a, da), 2.0f), u12 = scl(ab, -2.0f);
              const float2 u20 = add(mul(cda, b), mul(ca, db)), u21 = add(mul(da, cb), mul(a, cdb));
              const float u22 = 2.0f * (a.x * da.x + a.y * da.y) - 2.0f * (b.x * db.x + b.y * db.y);
              np = add(np, scl(add(add(mul(u00, fp), mul(u01, fm)), mul(u02, z)), ev.flip));
              nm = add(nm, scl(add(add(mul(u10, fp), mul(u11, fm)), mul(u12, z)), ev.flip));
              nz = add(nz, scl(add(add(mul(u20, fp), mul(u21, fm)), scl(z, u22)), ev.flip));
            }
            t[k].fp = np;
            t[k].fm = nm;
            t[k].z = nz;
          }
          v.fp = add(add(mul(t00, fp), mul(t01, fm)), mul(t02, z));
          v.fm = add(add(mul(t10, fp), mul(t11, fm)), mul(t12, z));
          v.z = add(add(mul(t20, fp), mul(t21, fm)), scl(z, t22));
        } else {
        float sa, va;
        if (p.transmit) {
          const float h = sinf(0.5f * theta);
          sa = sinf(theta);
          va = 2.0f * (h * h);
        } else {
          sa = ev.sine;
          va = ev.versine;
        }
        const float ca = 1.0f - va;
        float sp, cp;
        sincosf(ev.phase + b1_phase, &sp, &cp);
        const float2 e1p = mk(cp, sp);
        const float2 e2p = mul(e1p, e1p);
        const float shh = 0.5f * va, chh = 1.0f - shh;
        const float2 t01 = scl(e2p, shh);
        const float2 t02 = mul(mk(0.0f, -sa), e1p);
        const float2 t10 = cj(t01);
        const float2 t12 = mul(mk(0.0f, sa), cj(e1p));
        const float2 t20 = mul(mk(0.0f, -0.5f * sa), cj(e1p));
        const float2 t21 = mul(mk(0.0f, 0.5f * sa), e1p);
        const float2 fp = v.fp, fm = v.fm, z = v.z;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          const float2 dp = t[k].fp, dm = t[k].fm, dz = t[k].z;
          float2 np = mac(mac(scl(dp, chh), t01, dm), t02, dz);
          float2 nm = mac(mac(mul(t10, dp), mk(chh, 0.0f), dm), t12, dz);
          float2 nz = mac(mac(mul(t20, dp), t21, dm), mk(ca, 0.0f), dz);
          if ((TM & CB1) != 0 && k == SB) {
            // d/dB1 = flip * d/dtheta of the T-matrix, applied to the value
            const float dch = -0.5f * sa, dca = -sa;
            const float2 d01 = scl(e2p, 0.5f * sa);
            const float2 d02 = mul(mk(0.0f, -ca), e1p);
            const float2 d10 = cj(d01);
            const float2 d12 = mul(mk(0.0f, ca), cj(e1p));
            const float2 d20 = mul(mk(0.0f, -0.5f * ca), cj(e1p));
            const float2 d21 = mul(mk(0.0f, 0.5f * ca), e1p);
            np = add(np, scl(mac(mac(scl(fp, dch), d01, fm), d02, z), ev.flip));
            nm = add(nm, scl(mac(mac(mul(d10, fp), mk(dch, 0.0f), fm), d12, z), ev.flip));
            nz = add(nz, scl(mac(mac(mul(d20, fp), d21, fm), mk(dca, 0.0f), z), ev.flip));
          }
          t[k].fp = np;
          t[k].fm = nm;
          t[k].z = nz;
        }
        v.fp = mac(mac(scl(fp, chh), t01, fm), t02, z);
        v.fm = mac(mac(mul(t10, fp), mk(chh, 0.0f), fm), t12, z);
        v.z = mac(mac(mul(t20, fp), t21, fm), mk(ca, 0.0f), z);
        }
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs && first && live) {
      float sp, cp;
      sincosf(-ev.phase, &sp, &cp);
      const float2 fd = mul(factor, mk(cp, sp));
      const int64_t o = atom * p.outputs + ev.out;
      p.signal[o] = mul(fd, scl(v.fp, m0));
      if (p.jacobian) {
        float2* row = p.jacobian + atom * (int64_t)p.columns * p.outputs + ev.out;
        const float2 fm0 = scl(fd, m0);
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          int col = -1;
          if ((TM & CT1) != 0 && k == S1) col = p.col_t1;
          if ((TM & CT2) != 0 && k == S2) col = p.col_t2;
          if ((TM & CB1) != 0 && k == SB) col = p.col_b1;
          if (col >= 0) row[(int64_t)col * p.outputs] = mul(fm0, t[k].fp);
        }
        if (p.col_m0 >= 0) row[(int64_t)p.col_m0 * p.outputs] = mul(fd, v.fp);
      }
    }

    if (act & POST_SHIFT) {
      shift_complex<L>(v.fp, v.fm, sub, lane, beyond);
#pragma unroll
      for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp, t[k].fm, sub, lane, beyond);
    }
    if (act & SPOIL_AFTER) {
      v.fp = v.fm = mk(0.0f, 0.0f);
#pragma unroll
      for (int k = 0; k < NT; ++k) t[k].fp = t[k].fm = mk(0.0f, 0.0f);
    } else if (act & SHIFT_AFTER) {
      shift_complex<L>(v.fp, v.fm, sub, lane, beyond);
#pragma unroll
      for (int k = 0; k < NT; ++k) shift_complex<L>(t[k].fp, t[k].fm, sub, lane, beyond);
    }
  }
}

using SemisolidLaunch = void (*)(const SemisolidArgs&, dim3, dim3, hipStream_t);

template <int L, int TM, bool PROFILED>
void launch_semisolid(const SemisolidArgs& a, dim3 grid, dim3 block, hipStream_t s) {
  hipLaunchKernelGGL((semisolid_kernel<L, TM, PROFILED>), grid, block, 0, s, a);
}

template <int L, bool PROFILED>
SemisolidLaunch pick_semisolid_mode(int tm) {
  switch (tm & 7) {
    case 0: return &launch_semisolid<L, 0, PROFILED>;
    case 1: return &launch_semisolid<L, 1, PROFILED>;
    case 2: return &launch_semisolid<L, 2, PROFILED>;
    case 3: return &launch_semisolid<L, 3, PROFILED>;
    case 4: return &launch_semisolid<L, 4, PROFILED>;
    case 5: return &launch_semisolid<L, 5, PROFILED>;
    case 6: return &launch_semisolid<L, 6, PROFILED>;
    default: return &launch_semisolid<L, 7, PROFILED>;
  }
}

template <int L>
SemisolidLaunch pick_semisolid_mask(int tm, bool profiled) {
  return profiled ? pick_semisolid_mode<L, true>(tm) : pick_semisolid_mode<L, false>(tm);
}

#define TSIM_DECLARE_SEMISOLID(LANES) SemisolidLaunch pick_semisolid_##LANES(int tm, bool profiled);
TSIM_COMPLEX_LANES(TSIM_DECLARE_SEMISOLID)
#undef TSIM_DECLARE_SEMISOLID

}  // namespace tsim
