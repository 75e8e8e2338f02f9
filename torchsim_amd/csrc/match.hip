// Dictionary matching on the matrix cores.
//
// For every voxel y and every atom d of a dictionary the match score is the
// normalized complex inner product |<y, d>| / (|y| |d|) (reference
// estimators/_dictionary.py:308-387 and estimators/_grouped.py:219-245, which
// form it as a torch matmul per chunk and keep torch.topk of each chunk). Here
// no score matrix ever reaches memory:
//
// * atoms and voxels are normalized and laid out once as two fp16 planes,
//   hi = fp16(256 x) and lo = fp16(256 x - hi), in the lane order of the
//   32x32x16 fp16 MFMA, so every operand load of a wave is one coalesced
//   1 KiB row;
// * match_candidates accumulates hi*hi + hi*lo + lo*hi on the MFMA (products
//   of fp16 are exact in the FP32 accumulator; the dropped lo*lo term and the
//   split leave ~2^-22 of the unit score), forms re^2 + im^2 per (voxel, atom)
//   and keeps, per voxel, the best KC atoms in registers across its whole
//   share of the dictionary; lanes and splits are merged afterwards;
// * match_refine recomputes the inner product of each surviving candidate
//   exactly (FP64 over the caller's complex64 / float32 data), ranks them by
//   the exact score and writes indices, scores, complex least-squares scales
//   and densities (reference _dictionary.py:389-431).
//
// Keeping KC >= top_k + 1 candidates means an atom that the approximate pass
// ranks just below the k-th is still compared exactly.

#include <hip/hip_runtime.h>

#include <climits>
#include <cmath>
#include <cstdint>

#include "epg_common.h"

namespace tsim {
namespace {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

constexpr int KSTEP = 16;              // k per MFMA (v_mfma_f32_32x32x16_f16)
constexpr int TILE = 32;               // MFMA rows (16 voxels x 2 projections) and columns (atoms)
constexpr int VOXELS_PER_TILE = 16;
constexpr int ATOM_TILES = 4;          // per step: 128 atoms
// The 4 waves of a workgroup share its voxel rows and take consecutive
// 128-atom tiles of each 512-atom step: the voxel operand is fetched once per
// workgroup step (and reused from cache by the other three waves) instead of
// once per wave.
constexpr int WAVES = 4;
constexpr int ATOMS_PER_STEP = WAVES * ATOM_TILES * TILE;
// Row tiles per wave: 2 (32 voxels) when two candidates are kept per voxel,
// 1 (16 voxels) with four, so lists and accumulators fit without spilling.
template <int KC>
constexpr int row_tiles() { return KC <= 2 ? 2 : 1; }
template <int KC>
constexpr int voxels_per_block() { return row_tiles<KC>() * VOXELS_PER_TILE; }
constexpr float SPLIT_SCALE = 256.0f;
constexpr int MAX_CANDIDATES_PER_LANE = 8;  // refine: splits * KC <= 512
constexpr float NORM_EPS = 1.1920928955078125e-07f;  // float32 epsilon, the reference's norm floor

// Modes of the atom operand. 0: real atoms; 1: complex atoms whose imaginary
// parts are all zero; 2: complex atoms whose real parts are all zero; in
// these three the operand is one plane of N columns and a voxel's two rows
// are its real and imaginary parts. 3: complex atoms, 2N columns [Re d, Im d];
// a voxel's rows are [Re y, Im y] and [Im y, -Re y].
enum MatchMode : int { REAL = 0, REAL_PLANE = 1, IMAG_PLANE = 2, COMPLEX = 3 };

__device__ inline float atom_value(const void* dictionary, int complex_input, int mode, int64_t atom,
                                   int64_t contrasts, int64_t k) {
  const float* d = static_cast<const float*>(dictionary);
  if (!complex_input) return d[atom * contrasts + k];
  if (mode == REAL_PLANE) return d[(atom * contrasts + k) * 2];
  if (mode == IMAG_PLANE) return d[(atom * contrasts + k) * 2 + 1];
  if (k < contrasts) return d[(atom * contrasts + k) * 2];
  return d[(atom * contrasts + (k - contrasts)) * 2 + 1];
}

// The value of row `part` (0 or 1) of voxel v at column k of the query operand.
__device__ inline float query_value(const void* signals, int complex_input, int mode, int64_t v, int part,
                                    int64_t contrasts, int64_t k) {
  const float* y = static_cast<const float*>(signals);
  if (mode != COMPLEX) {
    if (!complex_input) return part == 0 ? y[v * contrasts + k] : 0.0f;
    return y[(v * contrasts + k) * 2 + part];
  }
  const int64_t n = k < contrasts ? k : k - contrasts;
  const float re = complex_input ? y[(v * contrasts + n) * 2] : y[v * contrasts + n];
  const float im = complex_input ? y[(v * contrasts + n) * 2 + 1] : 0.0f;
  if (part == 0) return k < contrasts ? re : im;
  return k < contrasts ? im : -re;
}

// Which plane entries of a complex dictionary are nonzero: bit 0 real, bit 1 imaginary.
__global__ void match_planes(const float2* __restrict__ d, int64_t count, int* __restrict__ flags) {
  int local = 0;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (int64_t)gridDim.x * blockDim.x) {
    const float2 v = d[i];
    local |= (v.x != 0.0f ? 1 : 0) | (v.y != 0.0f ? 2 : 0);
  }
  const bool any_re = __any(local & 1), any_im = __any(local & 2);
  if ((threadIdx.x & 63) == 0 && (any_re || any_im)) atomicOr(flags, (any_re ? 1 : 0) | (any_im ? 2 : 0));
}

// |row| in FP64, one wave per row (a row of complex or real values).
__global__ void row_norms(const float* __restrict__ x, int complex_input, int64_t rows, int64_t contrasts,
                          double* __restrict__ norm) {
  const int64_t row = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= rows) return;
  const int64_t width = complex_input ? 2 * contrasts : contrasts;
  const float* r = x + row * width;
  double sum = 0.0;
  for (int64_t k = lane; k < width; k += 64) {
    const double v = r[k];
    sum = fma(v, v, sum);
  }
  for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
  if (lane == 0) norm[row] = sqrt(sum);
}

__device__ inline void split_half(float x, _Float16& hi, _Float16& lo) {
  hi = (_Float16)x;
  lo = (_Float16)(x - (float)hi);
}

// Atom operand: element ((tile * ks + s) * 64 + lane) holds atom tile*32 + lane%32,
// columns 16 s + 8 (lane/32) ... + 7.
__global__ void layout_atoms(const void* __restrict__ dictionary, int complex_input, int mode, int64_t atoms,
                             int64_t contrasts, int64_t columns, const double* __restrict__ norm, int64_t ks,
                             int64_t total, half8* __restrict__ hi, half8* __restrict__ lo) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int lane = (int)(e & 63);
  const int64_t ts = e >> 6;
  const int64_t s = ts % ks, tile = ts / ks;
  const int64_t atom = tile * TILE + (lane & 31);
  const int64_t k0 = s * KSTEP + 8 * (lane >> 5);
  half8 h = {}, l = {};
  if (atom < atoms) {
    const float scale = (float)((double)SPLIT_SCALE / fmax(norm[atom], (double)NORM_EPS));
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = k0 + i;
      if (k < columns) {
        _Float16 a, b;
        split_half(atom_value(dictionary, complex_input, mode, atom, contrasts, k) * scale, a, b);
        h[i] = a;
        l[i] = b;
      }
    }
  }
  hi[e] = h;
  lo[e] = l;
}

// Query operand: 32-row tile t holds voxels 16 t ... 16 t + 15; row r of it is
// voxel 16 t + (r % 8) + 8 (r / 16), projection (r / 8) % 2, so that the two
// projections of a voxel land in the same lane of the MFMA result (registers
// j and j + 4).
__global__ void layout_queries(const void* __restrict__ signals, int complex_input, int mode, int64_t voxels,
                               int64_t contrasts, int64_t columns, const double* __restrict__ norm, int64_t ks,
                               int64_t total, half8* __restrict__ hi, half8* __restrict__ lo) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int lane = (int)(e & 63);
  const int64_t ts = e >> 6;
  const int64_t s = ts % ks, tile = ts / ks;
  const int r = lane & 31;
  const int64_t v = tile * VOXELS_PER_TILE + (r % 8) + 8 * (r / 16);
  const int part = (r / 8) % 2;
  const int64_t k0 = s * KSTEP + 8 * (lane >> 5);
  half8 h = {}, l = {};
  if (v < voxels) {
    const float scale = (float)((double)SPLIT_SCALE / fmax(norm[v], (double)NORM_EPS));
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int64_t k = k0 + i;
      if (k < columns) {
        _Float16 a, b;
        split_half(query_value(signals, complex_input, mode, v, part, contrasts, k) * scale, a, b);
        h[i] = a;
        l[i] = b;
      }
    }
  }
  hi[e] = h;
  lo[e] = l;
}

// Candidate lists are sorted best first; better = higher score, then lower index.
__device__ inline bool better(float s, int i, float t, int j) { return s > t || (s == t && i < j); }

template <int KC>
__device__ inline void insert(float (&s)[KC], int (&ix)[KC], float v, int a) {
  // candidates of one lane arrive in increasing atom order, so an equal score
  // never displaces an earlier (lower) index: strict comparison suffices
#pragma unroll
  for (int i = KC - 1; i > 0; --i) {
    const bool above = v > s[i - 1];
    const bool here = v > s[i];
    s[i] = above ? s[i - 1] : (here ? v : s[i]);
    ix[i] = above ? ix[i - 1] : (here ? a : ix[i]);
  }
  const bool top = v > s[0];
  s[0] = top ? v : s[0];
  ix[0] = top ? a : ix[0];
}

// The best KC of two sorted lists (bitonic merge; every operand index static).
template <int KC>
__device__ inline void merge(float (&s)[KC], int (&ix)[KC], const float (&t)[KC], const int (&jx)[KC]) {
#pragma unroll
  for (int i = 0; i < KC; ++i) {
    const bool mine = better(s[i], ix[i], t[KC - 1 - i], jx[KC - 1 - i]);
    s[i] = mine ? s[i] : t[KC - 1 - i];
    ix[i] = mine ? ix[i] : jx[KC - 1 - i];
  }
#pragma unroll
  for (int stride = KC / 2; stride > 0; stride >>= 1) {
#pragma unroll
    for (int i = 0; i < KC; ++i) {
      if ((i & stride) == 0) {
        const bool keep = better(s[i], ix[i], s[i + stride], ix[i + stride]);
        const float a = s[i], b = s[i + stride];
        const int p = ix[i], q = ix[i + stride];
        s[i] = keep ? a : b;
        ix[i] = keep ? p : q;
        s[i + stride] = keep ? b : a;
        ix[i + stride] = keep ? q : p;
      }
    }
  }
}

// One workgroup: 4 waves x ROW_TILES x 16 voxels, walking atom steps
// [begin, end) of its split; each step is 128 atoms (4 MFMA column tiles)
// over the full depth.
template <int KC>
__global__ __launch_bounds__(256) void match_candidates(const half8* __restrict__ q_hi,
                                                        const half8* __restrict__ q_lo,
                                                        const half8* __restrict__ d_hi,
                                                        const half8* __restrict__ d_lo, int64_t ks,
                                                        int64_t atoms, int64_t atom_steps, int64_t steps_per_split,
                                                        int64_t voxels_pad, float* __restrict__ part_score,
                                                        int* __restrict__ part_index) {
  constexpr int ROW_TILES = row_tiles<KC>();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int h = lane >> 5, col = lane & 31;
  const int64_t row_tile0 = (int64_t)blockIdx.x * ROW_TILES;
  const int64_t split = blockIdx.y;
  const int64_t begin = split * steps_per_split;
  const int64_t end = begin + steps_per_split < atom_steps ? begin + steps_per_split : atom_steps;

  float best[ROW_TILES][8][KC];
  int which[ROW_TILES][8][KC];
#pragma unroll
  for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
    for (int q = 0; q < 8; ++q)
#pragma unroll
      for (int i = 0; i < KC; ++i) {
        best[r][q][i] = -1.0f;
        which[r][q][i] = INT_MAX;
      }

  const half8* qh[ROW_TILES];
  const half8* ql[ROW_TILES];
#pragma unroll
  for (int r = 0; r < ROW_TILES; ++r) {
    qh[r] = q_hi + (row_tile0 + r) * ks * 64 + lane;
    ql[r] = q_lo + (row_tile0 + r) * ks * 64 + lane;
  }

  for (int64_t step = begin; step < end; ++step) {
    const int64_t tile0 = (step * WAVES + wave) * ATOM_TILES;
    const half8* dh = d_hi + tile0 * ks * 64 + lane;
    const half8* dl = d_lo + tile0 * ks * 64 + lane;
    const int64_t tile_stride = ks * 64;
    floatx16 acc[ROW_TILES][ATOM_TILES];
#pragma unroll
    for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
      for (int c = 0; c < ATOM_TILES; ++c) acc[r][c] = floatx16{};

    // two operand buffers: the loads of k-step s + 2 are issued as soon as
    // step s has been multiplied, so a load has two steps of MFMA work
    // (~1,500 cycles) to land in; the depth is padded to an even step count
    struct Operands {
      half8 ah[ROW_TILES], al[ROW_TILES], bh[ATOM_TILES], bl[ATOM_TILES];
    };
    // loads issued in the order the MFMAs first read them, so each wait
    // leaves the later loads of the buffer in flight
    auto load = [&](Operands& o, int64_t step_k) {
      const int64_t k = (step_k < ks ? step_k : ks - 1) * 64;  // clamped: no branch at the tail
      o.ah[0] = qh[0][k];
#pragma unroll
      for (int c = 0; c < ATOM_TILES; ++c) o.bh[c] = dh[c * tile_stride + k];
#pragma unroll
      for (int r = 1; r < ROW_TILES; ++r) o.ah[r] = qh[r][k];
#pragma unroll
      for (int c = 0; c < ATOM_TILES; ++c) o.bl[c] = dl[c * tile_stride + k];
#pragma unroll
      for (int r = 0; r < ROW_TILES; ++r) o.al[r] = ql[r][k];
    };
    auto multiply = [&](const Operands& o) {
#pragma unroll
      for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
        for (int c = 0; c < ATOM_TILES; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.ah[r], o.bh[c], acc[r][c], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
        for (int c = 0; c < ATOM_TILES; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.ah[r], o.bl[c], acc[r][c], 0, 0, 0);
#pragma unroll
      for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
        for (int c = 0; c < ATOM_TILES; ++c)
          acc[r][c] = __builtin_amdgcn_mfma_f32_32x32x16_f16(o.al[r], o.bh[c], acc[r][c], 0, 0, 0);
    };
    Operands even, odd;
    load(even, 0);
    __builtin_amdgcn_sched_barrier(0);
    load(odd, 1);
    __builtin_amdgcn_sched_barrier(0);
    // the scheduling barriers keep the compiler from sinking each load next
    // to its use, which would put the wait for it inside the same step
    for (int64_t s = 0; s < ks; s += 2) {
      multiply(even);
      __builtin_amdgcn_sched_barrier(0);
      load(even, s + 2);
      __builtin_amdgcn_sched_barrier(0);
      multiply(odd);
      __builtin_amdgcn_sched_barrier(0);
      load(odd, s + 3);
      __builtin_amdgcn_sched_barrier(0);
    }
    // epilogue: lane holds column `col` of each tile; registers j and j + 4
    // (j = 0..3, 8..11) are the two projections of voxel 4 h + j of the row tile
#pragma unroll
    for (int c = 0; c < ATOM_TILES; ++c) {
      const int64_t atom = (tile0 + c) * TILE + col;
      if (atom < atoms) {
#pragma unroll
        for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
          for (int q = 0; q < 8; ++q) {
            const int j = q < 4 ? q : q + 4;
            const float re = acc[r][c][j], im = acc[r][c][j + 4];
            insert<KC>(best[r][q], which[r][q], fmaf(re, re, im * im), (int)atom);
          }
      }
    }
  }

  // merge the 32 lanes that hold the same voxels
#pragma unroll
  for (int off = 1; off < 32; off <<= 1) {
#pragma unroll
    for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        float t[KC];
        int jx[KC];
#pragma unroll
        for (int i = 0; i < KC; ++i) {
          t[i] = __shfl_xor(best[r][q][i], off);
          jx[i] = __shfl_xor(which[r][q][i], off);
        }
        merge<KC>(best[r][q], which[r][q], t, jx);
      }
  }
  // lane `col` < 8 ROW_TILES writes the list of (row tile col / 8, slot col % 8)
#pragma unroll
  for (int r = 0; r < ROW_TILES; ++r)
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      if (col == r * 8 + q) {
        const int j = q < 4 ? q : q + 4;
        const int64_t voxel = (row_tile0 + r) * VOXELS_PER_TILE + 4 * h + j;
        // each wave's lists are a split of their own for the refine pass
        const int64_t at = ((split * WAVES + wave) * voxels_pad + voxel) * KC;
#pragma unroll
        for (int i = 0; i < KC; ++i) {
          part_score[at + i] = best[r][q][i];
          part_index[at + i] = which[r][q][i];
        }
      }
    }
}

__device__ inline double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
  return v;
}

// One wave per voxel: best KC of the splits' candidates by the approximate
// score, each recomputed exactly, the best top_k written.
template <int KC>
__global__ __launch_bounds__(256) void match_refine(
    const float* __restrict__ part_score, const int* __restrict__ part_index, int64_t splits, int64_t voxels_pad,
    int64_t voxels, const void* __restrict__ signals, int signals_complex, const double* __restrict__ voxel_norm,
    const void* __restrict__ dictionary, int dictionary_complex, const double* __restrict__ atom_norm,
    int64_t contrasts, int top_k, int64_t* __restrict__ out_index, float* __restrict__ out_score,
    float2* __restrict__ out_scale, float* __restrict__ out_density) {
  const int64_t v = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (v >= voxels) return;
  const int64_t count = splits * KC;
  float s[MAX_CANDIDATES_PER_LANE];
  int ix[MAX_CANDIDATES_PER_LANE];
#pragma unroll
  for (int m = 0; m < MAX_CANDIDATES_PER_LANE; ++m) {
    const int64_t e = lane + 64 * (int64_t)m;
    if (e < count) {
      const int64_t sp = e / KC, i = e % KC;
      s[m] = part_score[(sp * voxels_pad + v) * KC + i];
      ix[m] = part_index[(sp * voxels_pad + v) * KC + i];
    } else {
      s[m] = -1.0f;
      ix[m] = INT_MAX;
    }
  }
  const float* y = static_cast<const float*>(signals);
  const float* d = static_cast<const float*>(dictionary);
  const double ynorm = fmax(voxel_norm[v], (double)NORM_EPS);

  double exact[KC];
  int chosen[KC];
  double dot_re[KC], dot_im[KC];
#pragma unroll
  for (int c = 0; c < KC; ++c) {
    // wave-wide best remaining candidate
    float bs = s[0];
    int bi = ix[0];
#pragma unroll
    for (int m = 1; m < MAX_CANDIDATES_PER_LANE; ++m) {
      const bool b = better(s[m], ix[m], bs, bi);
      bs = b ? s[m] : bs;
      bi = b ? ix[m] : bi;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const float ts = __shfl_xor(bs, off);
      const int ti = __shfl_xor(bi, off);
      const bool b = better(ts, ti, bs, bi);
      bs = b ? ts : bs;
      bi = b ? ti : bi;
    }
#pragma unroll
    for (int m = 0; m < MAX_CANDIDATES_PER_LANE; ++m)
      if (ix[m] == bi) {
        s[m] = -1.0f;
        ix[m] = INT_MAX;
      }
    chosen[c] = bi;
    double re = 0.0, im = 0.0;
    if (bi != INT_MAX) {
      const int64_t a = bi;
      for (int64_t n = lane; n < contrasts; n += 64) {
        const double yr = signals_complex ? y[(v * contrasts + n) * 2] : y[v * contrasts + n];
        const double yi = signals_complex ? y[(v * contrasts + n) * 2 + 1] : 0.0;
        const double dr = dictionary_complex ? d[(a * contrasts + n) * 2] : d[a * contrasts + n];
        const double di = dictionary_complex ? d[(a * contrasts + n) * 2 + 1] : 0.0;
        // y conj(d)
        re = fma(yr, dr, fma(yi, di, re));
        im = fma(yi, dr, fma(-yr, di, im));
      }
      re = wave_sum(re);
      im = wave_sum(im);
      exact[c] = sqrt(re * re + im * im) / (ynorm * fmax(atom_norm[a], (double)NORM_EPS));
    } else {
      exact[c] = -1.0;
    }
    dot_re[c] = re;
    dot_im[c] = im;
  }
  if (lane != 0) return;
  // the top_k of the refined candidates: higher exact score, then lower index
  for (int t = 0; t < top_k; ++t) {
    int pick = -1;
#pragma unroll
    for (int c = 0; c < KC; ++c) {
      if (chosen[c] == INT_MAX || exact[c] < -0.5) continue;
      if (pick < 0) {
        pick = c;
        continue;
      }
      double pe = 0.0;
      int pi = 0;
#pragma unroll
      for (int e = 0; e < KC; ++e)
        if (e == pick) {
          pe = exact[e];
          pi = chosen[e];
        }
      if (exact[c] > pe || (exact[c] == pe && chosen[c] < pi)) pick = c;
    }
    const int64_t at = v * top_k + t;
    if (pick < 0) {
      out_index[at] = -1;
      out_score[at] = 0.0f;
      out_scale[at] = make_float2(0.0f, 0.0f);
      out_density[at] = 0.0f;
      continue;
    }
    double pe = 0.0, pr = 0.0, pm = 0.0;
    int pi = 0;
#pragma unroll
    for (int e = 0; e < KC; ++e)
      if (e == pick) {
        pe = exact[e];
        pi = chosen[e];
        pr = dot_re[e];
        pm = dot_im[e];
        exact[e] = -1.0;
      }
    const double anorm = fmax(atom_norm[pi], (double)NORM_EPS);
    const double power = anorm * anorm;
    out_index[at] = pi;
    out_score[at] = (float)pe;
    out_scale[at] = make_float2((float)(pr / power), (float)(pm / power));
    out_density[at] = (float)(ynorm * pe / anorm);
  }
}

template <int KC>
int run_match(const half8* d_hi, const half8* d_lo, int mode, int64_t atoms, int64_t atoms_pad, int64_t k_pad,
              const void* dictionary, int dictionary_complex, const double* atom_norm, const void* signals,
              int signals_complex, int64_t voxels, int64_t contrasts, int top_k, int64_t splits_hint,
              int64_t* out_index, float* out_score, float2* out_scale, float* out_density, hipStream_t stream) {
  const int64_t ks = k_pad / KSTEP;
  const int64_t columns = mode == COMPLEX ? 2 * contrasts : contrasts;
  const int64_t atom_steps = atoms_pad / ATOMS_PER_STEP;
  // voxels per pass: bounds the query operand (2 rows x k_pad x 4 B per voxel)
  constexpr int VOXELS_PER_BLOCK = voxels_per_block<KC>();
  const int64_t chunk = 131072;
  const int64_t chunk_pad = ((voxels < chunk ? voxels : chunk) + VOXELS_PER_BLOCK - 1) / VOXELS_PER_BLOCK *
                            VOXELS_PER_BLOCK;
  const int64_t blocks_x = chunk_pad / VOXELS_PER_BLOCK;
  int64_t splits = splits_hint > 0 ? splits_hint : (2048 + blocks_x - 1) / blocks_x;
  const int64_t max_splits = (int64_t)MAX_CANDIDATES_PER_LANE * 64 / (KC * WAVES);
  if (splits > max_splits) splits = max_splits;
  if (splits > atom_steps) splits = atom_steps;
  if (splits < 1) splits = 1;
  const int64_t steps_per_split = (atom_steps + splits - 1) / splits;
  splits = (atom_steps + steps_per_split - 1) / steps_per_split;

  const size_t q_elems = (size_t)(chunk_pad / VOXELS_PER_TILE) * ks * 64;  // half8 per plane
  half8* q_hi = nullptr;
  half8* q_lo = nullptr;
  double* y_norm = nullptr;
  float* p_score = nullptr;
  int* p_index = nullptr;
  const size_t part = (size_t)splits * WAVES * chunk_pad * KC;
  int status = scratch_alloc((void**)&q_hi, q_elems * sizeof(half8), stream, "match scratch");
  if (status == OK) status = scratch_alloc((void**)&q_lo, q_elems * sizeof(half8), stream, "match scratch");
  if (status == OK) status = scratch_alloc((void**)&y_norm, (size_t)chunk_pad * sizeof(double), stream, "match scratch");
  if (status == OK) status = scratch_alloc((void**)&p_score, part * sizeof(float), stream, "match scratch");
  if (status == OK) status = scratch_alloc((void**)&p_index, part * sizeof(int), stream, "match scratch");
  const size_t width = signals_complex ? 2 : 1;
  for (int64_t first = 0; status == OK && first < voxels; first += chunk) {
    const int64_t count = voxels - first < chunk ? voxels - first : chunk;
    const int64_t count_pad = (count + VOXELS_PER_BLOCK - 1) / VOXELS_PER_BLOCK * VOXELS_PER_BLOCK;
    const float* y = static_cast<const float*>(signals) + (size_t)first * contrasts * width;
    hipLaunchKernelGGL(row_norms, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, stream, y, signals_complex,
                       count, contrasts, y_norm);
    const int64_t total = (count_pad / VOXELS_PER_TILE) * ks * 64;
    hipLaunchKernelGGL(layout_queries, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, stream,
                       (const void*)y, signals_complex, mode, count, contrasts, columns, y_norm, ks, total, q_hi,
                       q_lo);
    hipLaunchKernelGGL(match_candidates<KC>, dim3((unsigned)(count_pad / VOXELS_PER_BLOCK), (unsigned)splits),
                       dim3(256), 0, stream, q_hi, q_lo, d_hi, d_lo, ks, atoms, atom_steps, steps_per_split,
                       count_pad, p_score, p_index);
    hipLaunchKernelGGL(match_refine<KC>, dim3((unsigned)((count + 3) / 4)), dim3(256), 0, stream, p_score,
                       p_index, splits * WAVES, count_pad, count, (const void*)y, signals_complex, y_norm, dictionary,
                       dictionary_complex, atom_norm, contrasts, top_k, out_index + first * top_k,
                       out_score + first * top_k, out_scale + first * top_k, out_density + first * top_k);
    status = hip_status(hipGetLastError(), "match launch");
  }
  hipFreeAsync(q_hi, stream);
  hipFreeAsync(q_lo, stream);
  hipFreeAsync(y_norm, stream);
  hipFreeAsync(p_score, stream);
  hipFreeAsync(p_index, stream);
  return status;
}

}  // namespace
}  // namespace tsim

using namespace tsim;

extern "C" {

int tsim_match_operand_size(int64_t atoms, int64_t contrasts, int mode, int64_t* atoms_pad, int64_t* k_pad) {
  if (atoms < 1 || contrasts < 1 || mode < REAL || mode > COMPLEX || atoms_pad == nullptr || k_pad == nullptr) {
    set_error("positive atom and contrast counts, a mode in 0..3 and output pointers are required");
    return BAD_ARGUMENT;
  }
  const int64_t columns = mode == COMPLEX ? 2 * contrasts : contrasts;
  *atoms_pad = (atoms + ATOMS_PER_STEP - 1) / ATOMS_PER_STEP * ATOMS_PER_STEP;
  // an even number of k-steps (the candidate kernel double-buffers them)
  *k_pad = (columns + 2 * KSTEP - 1) / (2 * KSTEP) * (2 * KSTEP);
  if (*atoms_pad > INT_MAX) {
    set_error("at most %d atoms", INT_MAX - ATOMS_PER_STEP);
    return BAD_ARGUMENT;
  }
  return OK;
}

int tsim_match_planes(const void* dictionary, int64_t atoms, int64_t contrasts, int* flags, void* stream) {
  if (dictionary == nullptr || flags == nullptr || atoms < 0 || contrasts < 0) {
    set_error("a complex dictionary, a flag word and non-negative sizes are required");
    return BAD_ARGUMENT;
  }
  const int64_t count = atoms * contrasts;
  if (count == 0) return OK;
  int64_t blocks = (count + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(match_planes, dim3((unsigned)blocks), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)dictionary, count, flags);
  return hip_status(hipGetLastError(), "match_planes launch");
}

int tsim_match_layout_atoms(const void* dictionary, int dictionary_complex, int64_t atoms, int64_t contrasts,
                            int mode, double* atom_norm, void* hi, void* lo, void* stream) {
  int64_t atoms_pad = 0, k_pad = 0;
  int status = tsim_match_operand_size(atoms, contrasts, mode, &atoms_pad, &k_pad);
  if (status != OK) return status;
  if (dictionary == nullptr || atom_norm == nullptr || hi == nullptr || lo == nullptr) {
    set_error("dictionary, norm and operand buffers are required");
    return BAD_ARGUMENT;
  }
  if ((mode != REAL) != (dictionary_complex != 0)) {
    set_error("mode 0 takes a real dictionary, modes 1-3 a complex one");
    return BAD_ARGUMENT;
  }
  const hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(row_norms, dim3((unsigned)((atoms + 3) / 4)), dim3(256), 0, s, (const float*)dictionary,
                     dictionary_complex, atoms, contrasts, atom_norm);
  const int64_t ks = k_pad / KSTEP;
  const int64_t total = (atoms_pad / TILE) * ks * 64;
  const int64_t columns = mode == COMPLEX ? 2 * contrasts : contrasts;
  hipLaunchKernelGGL(layout_atoms, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, s, dictionary,
                     dictionary_complex, mode, atoms, contrasts, columns, (const double*)atom_norm, ks, total,
                     (half8*)hi, (half8*)lo);
  return hip_status(hipGetLastError(), "match layout launch");
}

int tsim_match(const void* atoms_hi, const void* atoms_lo, int mode, int64_t atoms, const void* dictionary,
               int dictionary_complex, const double* atom_norm, const void* signals, int signals_complex,
               int64_t voxels, int64_t contrasts, int top_k, int64_t splits, int64_t* indices, float* scores,
               void* scales, float* densities, void* stream) {
  int64_t atoms_pad = 0, k_pad = 0;
  int status = tsim_match_operand_size(atoms, contrasts, mode, &atoms_pad, &k_pad);
  if (status != OK) return status;
  if (voxels < 0) {
    set_error("a non-negative voxel count is required");
    return BAD_ARGUMENT;
  }
  if (top_k < 1 || top_k > 3 || top_k > atoms) {
    set_error("top_k must be between 1 and min(3, atoms), got %d", top_k);
    return top_k > 3 ? UNSUPPORTED : BAD_ARGUMENT;
  }
  if (voxels == 0) return OK;
  if (atoms_hi == nullptr || atoms_lo == nullptr || dictionary == nullptr || atom_norm == nullptr ||
      signals == nullptr || indices == nullptr || scores == nullptr || scales == nullptr || densities == nullptr) {
    set_error("operand, dictionary, norm, signal and output buffers are required");
    return BAD_ARGUMENT;
  }
  const hipStream_t s = (hipStream_t)stream;
  const half8* hi = static_cast<const half8*>(atoms_hi);
  const half8* lo = static_cast<const half8*>(atoms_lo);
  float2* sc = static_cast<float2*>(scales);
  if (top_k + 1 <= 2)
    return run_match<2>(hi, lo, mode, atoms, atoms_pad, k_pad, dictionary, dictionary_complex, atom_norm, signals,
                        signals_complex, voxels, contrasts, top_k, splits, indices, scores, sc, densities, s);
  return run_match<4>(hi, lo, mode, atoms, atoms_pad, k_pad, dictionary, dictionary_complex, atom_norm, signals,
                      signals_complex, voxels, contrasts, top_k, splits, indices, scores, sc, densities, s);
}

}  // extern "C"
