// The real-subspace kernel template, shared by the per-tile translation units.
// See epg_real.hip for what it computes and how an atom is laid out.
#pragma once

#include "epg_common.h"

namespace tsim {

struct RealArgs {
  const EventCode* events;
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
};

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

constexpr int popcount4(int m) { return (m & 1) + ((m >> 1) & 1) + ((m >> 2) & 1) + ((m >> 3) & 1); }
// Slot of direction `bit` among those present in `mask`.
constexpr int slot_of(int mask, int bit) { return popcount4(mask & (bit - 1)); }

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

// One atom group's walk of its event stream. `group` is the atom this lane's
// group stands for (clamped by the caller's `live` test inside).
template <int SL, int L, int TM, bool DIFF, bool B1MAP, bool ACCUM>
__device__ __forceinline__ void real_body(const RealArgs& p, int64_t group) {
  constexpr int NT = popcount4(TM);
  constexpr int NTA = NT > 0 ? NT : 1;
  constexpr int ST1 = slot_of(TM, TAN_T1);
  constexpr int ST2 = slot_of(TM, TAN_T2);
  constexpr int SB1 = slot_of(TM, TAN_B1);
  constexpr int SIE = slot_of(TM, TAN_IE);

  const int sub = (int)(threadIdx.x % L);
  const bool first = sub == 0;
  const bool last = sub == L - 1;
  const bool live = group < p.atoms;
  const int64_t atom = live ? group : p.atoms - 1;

  // Which of this lane's registers stand for an order below S.
  const int base = sub * SL;
  unsigned keep = 0;
#pragma unroll
  for (int r = 0; r < SL; ++r) keep |= (base + r < p.states) ? (1u << r) : 0u;
  const bool padded = SL * L != p.states;

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
  double acc_g[ACCUM ? NORMAL_COLUMNS : 1];
  double acc_h[ACCUM ? NORMAL_COLUMNS * (NORMAL_COLUMNS + 1) / 2 : 1];
#pragma unroll
  for (int i = 0; i < (ACCUM ? NORMAL_COLUMNS : 1); ++i) acc_g[i] = 0.0;
#pragma unroll
  for (int i = 0; i < (ACCUM ? NORMAL_COLUMNS * (NORMAL_COLUMNS + 1) / 2 : 1); ++i) acc_h[i] = 0.0;

  float last_dt = -1.0f;
  float e1 = 1.0f, e2 = 1.0f, de1 = 0.0f, de2 = 0.0f;
  float dl[DIFF ? SL : 1], dtr[DIFF ? SL : 1];
#pragma unroll
  for (int r = 0; r < (DIFF ? SL : 1); ++r) {
    dl[r] = 1.0f;
    dtr[r] = 1.0f;
  }

  const EventCode* __restrict__ stream = p.events;
  const int64_t E = p.event_count;
  for (int64_t e = 0; e < E; ++e) {
    const EventCode ev = stream[e];
    const float dt = ev.dt;
    const uint32_t act = ev.ctrl & 0xFFu;
    const int32_t kind = (int32_t)(ev.ctrl >> CTRL_KIND_SHIFT);

    if (dt != 0.0f) {
      if (dt != last_dt) {
        last_dt = dt;
        e1 = expf(-r1 * dt);
        e2 = expf(-r2 * dt);
        de1 = e1 * (k1 * dt);
        de2 = e2 * (k2 * dt);
        if constexpr (DIFF) {
          const float bf = drate * dt;
#pragma unroll
          for (int r = 0; r < SL; ++r) {
            const float l = (float)(base + r);
            dl[r] = expf(-(l * l) * bf);
            dtr[r] = expf(-(l * l + l + (1.0f / 3.0f)) * bf);
          }
        }
      }
#pragma unroll
      for (int r = 0; r < SL; ++r) {
        const float tr = DIFF ? e2 * dtr[r] : e2;
        const float lo = DIFF ? e1 * dl[r] : e1;
        if constexpr ((TM & TAN_T1) != 0) {
          const float dlo = DIFF ? de1 * dl[r] : de1;
          ta[ST1][r] *= tr;
          tb[ST1][r] *= tr;
          tc[ST1][r] = fmaf(tc[ST1][r], lo, c[r] * dlo);
        }
        if constexpr ((TM & TAN_T2) != 0) {
          const float dtv = DIFF ? de2 * dtr[r] : de2;
          ta[ST2][r] = fmaf(ta[ST2][r], tr, a[r] * dtv);
          tb[ST2][r] = fmaf(tb[ST2][r], tr, b[r] * dtv);
          tc[ST2][r] *= lo;
        }
        if constexpr ((TM & TAN_B1) != 0) {
          ta[SB1][r] *= tr;
          tb[SB1][r] *= tr;
          tc[SB1][r] *= lo;
        }
        if constexpr ((TM & TAN_IE) != 0) {
          ta[SIE][r] *= tr;
          tb[SIE][r] *= tr;
          tc[SIE][r] *= lo;
        }
        a[r] *= tr;
        b[r] *= tr;
        c[r] *= lo;
      }
      // Recovery toward equilibrium lands on order zero only.
      c[0] += first ? 1.0f - e1 : 0.0f;
      if constexpr ((TM & TAN_T1) != 0) tc[ST1][0] -= first ? de1 : 0.0f;
    }

    if (act & PRE_SHIFT) {
      shift_pair<SL, L>(a, b, first, last);
      if (padded) clear_padding<SL>(a, keep);
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        shift_pair<SL, L>(ta[k], tb[k], first, last);
        if (padded) clear_padding<SL>(ta[k], keep);
      }
    }

    if (kind == KIND_RF) {
      if (act & INVERSION) {
        const float eff = -eta;
#pragma unroll
        for (int r = 0; r < SL; ++r) {
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            if ((TM & TAN_IE) != 0 && k == SIE) {
              tc[k][r] = fmaf(tc[k][r], eff, -c[r]);
            } else {
              tc[k][r] *= eff;
            }
          }
          c[r] *= eff;
        }
      } else {
        float cs, sn;
        float alpha = ev.flip;
        if constexpr (B1MAP) {
          alpha = ev.flip * b1;
          sincosf(alpha, &sn, &cs);
        } else {
          cs = ev.cosine;
          sn = ev.sine;
        }
        const float ch = 0.5f * (1.0f + cs);
        const float sh = 0.5f * (1.0f - cs);
        const float hs = 0.5f * sn;
        // Along B1 the angle moves by the flip itself per unit B1.
        const float dal = ev.flip;
        const float dch = -0.5f * sn * dal;
        const float dhs = 0.5f * cs * dal;
        const float dsn = cs * dal;
        const float dcs = -sn * dal;
#pragma unroll
        for (int r = 0; r < SL; ++r) {
          const float pv = a[r], mv = b[r], zv = c[r];
#pragma unroll
          for (int k = 0; k < NT; ++k) {
            const float dp = ta[k][r], dm = tb[k][r], dz = tc[k][r];
            float na = ch * dp + sh * dm - sn * dz;
            float nb = sh * dp + ch * dm + sn * dz;
            float nc = hs * (dp - dm) + cs * dz;
            if ((TM & TAN_B1) != 0 && k == SB1) {
              // product rule on the angle: d(ch) = -d(sh) = dch
              const float u = pv - mv;
              const float xa = dch * u - dsn * zv;
              na += xa;
              nb -= xa;
              nc += dhs * u + dcs * zv;
            }
            ta[k][r] = na;
            tb[k][r] = nb;
            tc[k][r] = nc;
          }
          a[r] = ch * pv + sh * mv - sn * zv;
          b[r] = sh * pv + ch * mv + sn * zv;
          c[r] = hs * (pv - mv) + cs * zv;
        }
      }
    } else if (kind == KIND_ADC && (act & RECORD) && ev.out >= 0 && ev.out < p.outputs) {
      if constexpr (ACCUM) {
        if (first && live) {
          const int64_t o = ev.out;
          const float sign = p.sample_sign ? p.sample_sign[o] : 1.0f;
          const float fr = p.factor_re * sign, fi = p.factor_im * sign;
          const float v = m0 * a[0];
          const float2 y = p.data[atom * p.outputs + o];
          const double rx = (double)(-fi * v) - (double)y.x;
          const double ry = (double)(fr * v) - (double)y.y;
          acc_cost += rx * rx + ry * ry;
          // J_c = (-fi d_c, fr d_c): J_c . r = d_c q, J_c . J_k = |factor|^2 d_c d_k
          const double q = -(double)fi * rx + (double)fr * ry;
          const double ff = (double)fr * fr + (double)fi * fi;
          double d[NORMAL_COLUMNS];
#pragma unroll
          for (int c = 0; c < NORMAL_COLUMNS; ++c) {
            float x = 0.0f;
            if constexpr ((TM & TAN_T1) != 0) x = (p.col_t1 == c) ? m0 * ta[ST1][0] : x;
            if constexpr ((TM & TAN_T2) != 0) x = (p.col_t2 == c) ? m0 * ta[ST2][0] : x;
            if constexpr ((TM & TAN_B1) != 0) x = (p.col_b1 == c) ? m0 * ta[SB1][0] : x;
            x = (p.col_m0 == c) ? a[0] : x;
            d[c] = (double)x;
          }
          int h = 0;
#pragma unroll
          for (int c = 0; c < NORMAL_COLUMNS; ++c) {
            acc_g[c] += d[c] * q;
#pragma unroll
            for (int k = c; k < NORMAL_COLUMNS; ++k) acc_h[h++] += ff * d[c] * d[k];
          }
        }
      } else if (first && live) {
        const int64_t o = ev.out;
        const float v = m0 * a[0];
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
              const float d = m0 * ta[ST1][0];
              row[(int64_t)p.col_t1 * p.outputs] = make_float2(-fi * d, fr * d);
            }
          }
          if constexpr ((TM & TAN_T2) != 0) {
            if (p.col_t2 >= 0) {
              const float d = m0 * ta[ST2][0];
              row[(int64_t)p.col_t2 * p.outputs] = make_float2(-fi * d, fr * d);
            }
          }
          if constexpr ((TM & TAN_B1) != 0) {
            if (p.col_b1 >= 0) {
              const float d = m0 * ta[SB1][0];
              row[(int64_t)p.col_b1 * p.outputs] = make_float2(-fi * d, fr * d);
            }
          }
          if constexpr ((TM & TAN_IE) != 0) {
            if (p.col_ie >= 0) {
              const float d = m0 * ta[SIE][0];
              row[(int64_t)p.col_ie * p.outputs] = make_float2(-fi * d, fr * d);
            }
          }
          if (p.col_m0 >= 0) {
            const float d = a[0];
            row[(int64_t)p.col_m0 * p.outputs] = make_float2(-fi * d, fr * d);
          }
        }
      }
    }

    if (act & POST_SHIFT) {
      shift_pair<SL, L>(a, b, first, last);
      if (padded) clear_padding<SL>(a, keep);
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        shift_pair<SL, L>(ta[k], tb[k], first, last);
        if (padded) clear_padding<SL>(ta[k], keep);
      }
    }
    if (act & SPOIL_AFTER) {
#pragma unroll
      for (int r = 0; r < SL; ++r) {
        a[r] = 0.0f;
        b[r] = 0.0f;
#pragma unroll
        for (int k = 0; k < NT; ++k) {
          ta[k][r] = 0.0f;
          tb[k][r] = 0.0f;
        }
      }
    } else if (act & SHIFT_AFTER) {
      shift_pair<SL, L>(a, b, first, last);
      if (padded) clear_padding<SL>(a, keep);
#pragma unroll
      for (int k = 0; k < NT; ++k) {
        shift_pair<SL, L>(ta[k], tb[k], first, last);
        if (padded) clear_padding<SL>(ta[k], keep);
      }
    }
  }
  if constexpr (ACCUM) {
    if (first && live) {
      double* out = p.normal + atom * NORMAL_STRIDE;
      out[0] = acc_cost;
#pragma unroll
      for (int c = 0; c < NORMAL_COLUMNS; ++c) out[1 + c] = acc_g[c];
#pragma unroll
      for (int h = 0; h < NORMAL_COLUMNS * (NORMAL_COLUMNS + 1) / 2; ++h) out[1 + NORMAL_COLUMNS + h] = acc_h[h];
    }
  }
}


template <int SL, int L, int TM, bool DIFF, bool B1MAP, bool ACCUM = false>
__global__ void __launch_bounds__(256) real_kernel(const RealArgs p) {
  const int64_t group = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / L;
  real_body<SL, L, TM, DIFF, B1MAP, ACCUM>(p, group);
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

template <int SL, int L, int TM, bool DIFF, bool B1MAP, bool ACCUM = false>
__global__ void __launch_bounds__(256) real_batch_kernel(const RealArgs* __restrict__ requests,
                                                        const BatchWork* __restrict__ work, int64_t blocks) {
  for (int64_t b = blockIdx.x; b < blocks; b += gridDim.x) {
    const BatchWork w = work[b];
    const RealArgs p = requests[w.request];
    const int64_t group = w.first_atom + (int64_t)threadIdx.x / L;
    real_body<SL, L, TM, DIFF, B1MAP, ACCUM>(p, group);
  }
}

using BatchLaunch = void (*)(const RealArgs*, const BatchWork*, int64_t, dim3, dim3, hipStream_t);

template <int SL, int L, int TM, bool DIFF, bool B1MAP, bool ACCUM>
void launch_batch(const RealArgs* r, const BatchWork* w, int64_t blocks, dim3 grid, dim3 block, hipStream_t s) {
  hipLaunchKernelGGL((real_batch_kernel<SL, L, TM, DIFF, B1MAP, ACCUM>), grid, block, 0, s, r, w, blocks);
}

template <int SL, int L, int TM, bool ACCUM>
BatchLaunch pick_batch_flags(bool diff, bool b1map) {
  if constexpr (ACCUM) {
    return b1map ? &launch_batch<SL, L, TM, false, true, true> : &launch_batch<SL, L, TM, false, false, true>;
  } else {
    if (diff) return b1map ? &launch_batch<SL, L, TM, true, true, false> : &launch_batch<SL, L, TM, true, false, false>;
    return b1map ? &launch_batch<SL, L, TM, false, true, false> : &launch_batch<SL, L, TM, false, false, false>;
  }
}

template <int SL, int L>
BatchLaunch pick_batch_mask(int tm, bool diff, bool b1map, bool accum) {
  switch (tm) {
#define TSIM_BATCH_TM(M)                                                     \
  case M:                                                                   \
    return accum ? pick_batch_flags<SL, L, M, true>(diff, b1map)            \
                 : pick_batch_flags<SL, L, M, false>(diff, b1map);
    TSIM_BATCH_TM(0)
    TSIM_BATCH_TM(1)
    TSIM_BATCH_TM(2)
    TSIM_BATCH_TM(3)
    TSIM_BATCH_TM(4)
    TSIM_BATCH_TM(5)
    TSIM_BATCH_TM(6)
    TSIM_BATCH_TM(7)
#undef TSIM_BATCH_TM
    default:
      return accum ? nullptr : pick_batch_flags<SL, L, 15, false>(diff, b1map);
  }
}

using RealLaunch = void (*)(const RealArgs&, dim3, dim3, hipStream_t);

template <int SL, int L, int TM, bool DIFF, bool B1MAP>
void launch_real(const RealArgs& args, dim3 grid, dim3 block, hipStream_t stream) {
  hipLaunchKernelGGL((real_kernel<SL, L, TM, DIFF, B1MAP>), grid, block, 0, stream, args);
}

template <int SL, int L, int TM, bool B1MAP>
void launch_normal(const RealArgs& args, dim3 grid, dim3 block, hipStream_t stream) {
  hipLaunchKernelGGL((real_kernel<SL, L, TM, false, B1MAP, true>), grid, block, 0, stream, args);
}

template <int SL, int L, int TM>
RealLaunch pick_normal_b1(bool b1map) {
  return b1map ? &launch_normal<SL, L, TM, true> : &launch_normal<SL, L, TM, false>;
}

template <int SL, int L>
RealLaunch pick_normal_mask(int tm, bool b1map) {
  switch (tm) {
    case 0: return pick_normal_b1<SL, L, 0>(b1map);
    case 1: return pick_normal_b1<SL, L, 1>(b1map);
    case 2: return pick_normal_b1<SL, L, 2>(b1map);
    case 3: return pick_normal_b1<SL, L, 3>(b1map);
    case 4: return pick_normal_b1<SL, L, 4>(b1map);
    case 5: return pick_normal_b1<SL, L, 5>(b1map);
    case 6: return pick_normal_b1<SL, L, 6>(b1map);
    default: return pick_normal_b1<SL, L, 7>(b1map);
  }
}

template <int SL, int L, int TM, bool DIFF>
RealLaunch pick_b1(bool b1map) {
  return b1map ? &launch_real<SL, L, TM, DIFF, true> : &launch_real<SL, L, TM, DIFF, false>;
}

template <int SL, int L, int TM>
RealLaunch pick_diff(bool diff, bool b1map) {
  return diff ? pick_b1<SL, L, TM, true>(b1map) : pick_b1<SL, L, TM, false>(b1map);
}

template <int SL, int L>
RealLaunch pick_mask(int tm, bool diff, bool b1map) {
  switch (tm) {
    case 0: return pick_diff<SL, L, 0>(diff, b1map);
    case 1: return pick_diff<SL, L, 1>(diff, b1map);
    case 2: return pick_diff<SL, L, 2>(diff, b1map);
    case 3: return pick_diff<SL, L, 3>(diff, b1map);
    case 4: return pick_diff<SL, L, 4>(diff, b1map);
    case 5: return pick_diff<SL, L, 5>(diff, b1map);
    case 6: return pick_diff<SL, L, 6>(diff, b1map);
    case 7: return pick_diff<SL, L, 7>(diff, b1map);
    default: return pick_diff<SL, L, 15>(diff, b1map);
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

#define TSIM_DECLARE_TILE(ID, SL, L)                                   \
  RealLaunch pick_real_tile_##ID(int tm, bool diff, bool b1map); \
  RealLaunch pick_normal_tile_##ID(int tm, bool b1map);          \
  BatchLaunch pick_batch_tile_##ID(int tm, bool diff, bool b1map, bool accum);
TSIM_TILES(TSIM_DECLARE_TILE)
#undef TSIM_DECLARE_TILE

}  // namespace tsim
