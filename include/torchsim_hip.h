/*
 * torchsim_hip.h -- C ABI of the gfx950 EPG library (libtorchsim_hip.so).
 *
 * Plain pointers and sizes only. Every pointer to array data is a DEVICE
 * pointer unless stated otherwise; `stream` is a hipStream_t passed as void*.
 * Every entry point returns 0 on success or a nonzero status:
 *   1 bad argument, 2 unsupported, 3 HIP failure;
 * the message of the last failure on the calling thread is returned by
 * tsim_last_error(). No entry point throws across the ABI, and none allocates
 * caller-visible memory: the caller owns every input and output buffer.
 * Launches are asynchronous on `stream`; entry points are safe to call from
 * several host threads on distinct streams.
 *
 * Packed buffer order (26 pointers) is the reference's ABI:
 *   tissue (17, FP32, one value per atom, or one value in total where the
 *   caller says so):  t1 t2 m0 b1 b1_phase b0 inversion_efficiency diffusion
 *   velocity bound_fraction bound_exchange t1_bound pool_b_fraction
 *   pool_b_exchange t1_pool_b t2_pool_b pool_b_shift
 *   events (9, one entry per event): duration(f32, s) kind(i32: 0 wait,
 *   1 rf, 2 adc) flip(f32, rad) phase(f32, rad) action(u8) output_index(i32,
 *   -1 when not recorded) shim_index(i32) saturation(f32) rf_frequency(f32)
 * (reference src/torchsim/_epg_cpu.cpp:613-631, sequence/_parameters.py:105-195).
 * Feature bits: 1 off_axis, 2 moving, 4 diffusing, 8 transmit, 16 density,
 * 32 inverting (reference _epg_cpu.cpp:89-96).
 */
#ifndef TORCHSIM_HIP_H
#define TORCHSIM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Message of the last failure on the calling thread ("" if none). */
const char* tsim_last_error(void);

/* ABI revision of the loaded library. */
int tsim_version(void);

/*
 * Forward EPG of one train for every atom, into the reference's two output
 * planes. Replaces the reference's CPython entry `_epg_cpu.simulate`
 * (src/torchsim/_epg_cpu.cpp:10740-10926) and the Triton launcher
 * `_epg_triton.simulate` (sequence/_epg_triton.py:17120).
 *
 * pointers: 26 packed buffers (order above) then output_real, output_imag,
 *           each (atom_count, output_count) FP32; pointer_count must be 28.
 * real_axis: 1 when every pulse and sample share one phase and the tissue has
 *           no off-resonance (the reference's verdict, sequence/
 *           _accelerators.py:1953-2017); then the real-subspace kernel runs.
 *           Otherwise the complex kernel runs.
 * pool_kind: 0 single pool, 2 one chemically exchanging pool.
 * tissue_stride: 1 if every optional tissue buffer holds one value per atom,
 *           0 if each holds one value for all atoms (t1, t2 are per atom).
 */
int tsim_simulate(void* const* pointers, int64_t pointer_count, int64_t atom_count,
                  int64_t train_count, int64_t event_count, int64_t state_count,
                  int64_t output_count, int real_axis, int64_t shim_count, int pool_kind,
                  int features, int64_t tissue_stride, void* stream);

/*
 * Complex forward EPG (single pool or one exchanging pool, diffusion,
 * off-resonance, RF phase). Same pointer layout as tsim_simulate.
 * Replaces `_epg_cpu.simulate` with pool_kind 0/2 and real_axis -1
 * (src/torchsim/_epg_cpu.cpp:4216-4566, sequence/_epg_triton.py:14676-15370).
 */
int tsim_complex_simulate(void* const* pointers, int64_t pointer_count, int64_t atom_count,
                          int64_t event_count, int64_t state_count, int64_t output_count,
                          int pool_kind, int features, int64_t tissue_stride, void* stream);

/*
 * Complex signal and Jacobian in ONE pass, complex64 in the public layout.
 * Replaces `_epg_cpu.simulate` + `_epg_cpu.simulate_jvp` per property for
 * the complex kernels (src/torchsim/_epg_cpu.cpp:4216-4566, 2469-3005).
 * columns: HOST array of column_count (<= 4) indices among {0 t1, 1 t2,
 * 2 m0, 3 b1}; pool_kind 0 single pool, 2 one exchanging pool.
 */
int tsim_complex_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                          int64_t state_count, int64_t output_count, int pool_kind, int features,
                          int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                          float factor_re, float factor_im, void* signal, void* jacobian,
                          void* stream);

/*
 * Signal and Jacobian of a tissue with a semisolid (magnetization-transfer)
 * pool, complex64 in the public layout. Replaces `_epg_cpu.simulate` +
 * `_epg_cpu.simulate_jvp` with pool kind 1 and real_axis -1
 * (src/torchsim/_epg_cpu.cpp:4216-4566, two_pool_step :875-932, the
 * saturation :4463-4480, the lineshape read :1888-1910; the lineshape table
 * is the reference's Tail::ABSORPTION slot).
 * The packed buffers bound_fraction, bound_exchange and t1_bound must be set
 * (pool_b_fraction must not: three pools are not built), and the event
 * buffers saturation (1/s per rad^2 of flip) and rf_frequency (Hz).
 * columns: HOST array of column_count (<= 4) indices among {0 t1, 1 t2,
 *          2 m0, 3 b1}.
 * lineshape: DEVICE (lineshape_bins, 2) FP32, value then slope (per Hz) of
 *          the absorption lineshape at knots lineshape_step Hz apart from
 *          resonance.
 */
int tsim_semisolid_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                            int64_t state_count, int64_t output_count, int features, int64_t tissue_stride,
                            const int32_t* columns, int64_t column_count, float factor_re, float factor_im,
                            const float* lineshape, int64_t lineshape_bins, float lineshape_step, void* signal,
                            void* jacobian, void* stream);

/*
 * First-order adjoint through the semisolid pool (instantaneous pulses):
 * the gradient of a real loss with respect to every event's flip (rad),
 * phase (rad) and duration (s), given the loss's cotangent on the signal
 * tsim_semisolid_jacobian writes (dL/dRe + i dL/dIm per sample). Replaces the
 * event gradients of `_epg_cpu.simulate_vjp` with pool kind 1 and real_axis
 * -1 (src/torchsim/_epg_cpu.cpp:4839-5944 via simulate_vjp_range<INSTANT,
 * SEMISOLID>, :11316-11340); the tissue gradients come from that entry's
 * Jacobian columns. Buffers and lineshape as for tsim_semisolid_jacobian.
 * cotangent: DEVICE (atom_count, output_count) complex64.
 * grad_flip, grad_phase, grad_duration: DEVICE (event_count,) FP32,
 *          overwritten. Entry states are kept for at most 4 GiB of lanes at a
 *          time (stream-ordered scratch); more atoms take several launches.
 */
int tsim_semisolid_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                       int64_t output_count, int features, int64_t tissue_stride, float factor_re, float factor_im,
                       const float* lineshape, int64_t lineshape_bins, float lineshape_step, const void* cotangent,
                       float* grad_flip, float* grad_phase, float* grad_duration, void* stream);

/*
 * First-order adjoint with slice-profiled pulses (and a semisolid pool where
 * its packed buffers are set): as tsim_semisolid_vjp, the pulses turning the
 * states as for tsim_profiled_jacobian (the reference's simulate_vjp in
 * RfMode::PROFILED, src/torchsim/_epg_cpu.cpp:11310-11340). Atoms and the
 * cotangent are voxel-major with `locations` atoms per voxel, each
 * position's cotangent the voxel's over `locations` (the caller's average).
 */
int tsim_profiled_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                      int64_t output_count, int features, int64_t tissue_stride, float factor_re, float factor_im,
                      const float* profile, int64_t locations, int64_t profile_bins, float profile_step,
                      const float* lineshape, int64_t lineshape_bins, float lineshape_step, const void* cotangent,
                      float* grad_flip, float* grad_phase, float* grad_duration, void* stream);

/*
 * Forward mode along the events through the semisolid pool and/or
 * slice-profiled pulses: the signal's directional derivative for per-event
 * tangents of the flips (rad), phases (rad) and durations (s) -- the event
 * rows of `_epg_cpu.simulate_jvp` with pool kind 1 / RfMode::PROFILED
 * (src/torchsim/_epg_cpu.cpp:11017-11148); the tissue rows are the columns of
 * tsim_semisolid_jacobian / tsim_profiled_jacobian. Buffers as for those
 * (profile NULL for instantaneous pulses; a semisolid pool where its packed
 * buffers are set, then the lineshape required).
 * d_flip, d_phase, d_duration: DEVICE (event_count,) FP32, each may be NULL.
 * tangent: DEVICE (atom_count, output_count) complex64; the recorded samples
 *          are written (the caller clears the rest).
 */
int tsim_semisolid_event_jvp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                             int64_t output_count, int features, int64_t tissue_stride, float factor_re,
                             float factor_im, const float* profile, int64_t locations, int64_t profile_bins,
                             float profile_step, const float* lineshape, int64_t lineshape_bins, float lineshape_step,
                             const float* d_flip, const float* d_phase, const float* d_duration, void* tangent,
                             void* stream);

/*
 * Signal and Jacobian with slice-profiled pulses: every pulse but an
 * inversion turns the states through the Cayley-Klein pair (a, b) its shape
 * performs at the atom's slice position and effective flip flip * b1, read
 * by cubic Hermite from `profile`, the RF phase applied as b exp(-i phi).
 * Replaces `_epg_cpu.simulate` / `simulate_jvp` in RfMode::PROFILED with the
 * table in the Tail::TABLE slot, one shape (src/torchsim/_epg_cpu.cpp:
 * 2041-2049, 2107-2160, 2273-2304, 4491-4520).
 * Atoms are voxel-major with `locations` atoms per voxel (atom_count a
 * multiple of it; atom i sits at position i % locations); the caller
 * averages each voxel's signal and columns over its positions.
 * profile: DEVICE (locations, profile_bins, 8) FP32: a, b, da/dtheta,
 *          db/dtheta as (real, imag) at flips k * profile_step.
 * A semisolid pool runs too where its packed buffers are set (then the
 * lineshape is required as for tsim_semisolid_jacobian; else it may be NULL).
 */
int tsim_profiled_jacobian(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                           int64_t output_count, int features, int64_t tissue_stride, const int32_t* columns,
                           int64_t column_count, float factor_re, float factor_im, const float* profile,
                           int64_t locations, int64_t profile_bins, float profile_step, const float* lineshape,
                           int64_t lineshape_bins, float lineshape_step, void* signal, void* jacobian,
                           void* stream);

/*
 * Real-subspace signal and Jacobian in ONE pass, written as complex64 in the
 * public layout. Replaces the P pairs of `_epg_cpu.simulate` +
 * `_epg_cpu.simulate_jvp` (src/torchsim/_epg_cpu.cpp:10740, 11017) that
 * `Simulator.jacobian` issues through `torch.func.jvp`, one per property
 * (reference model/_signal.py:230-238).
 *
 * packed:   the 26 packed buffers.
 * columns:  HOST array of column_count (<= 5) tissue indices among
 *           {0 t1, 1 t2, 2 m0, 3 b1, 6 inversion_efficiency}; column p of the
 *           Jacobian is the derivative along property columns[p].
 * factor:   complex factor (factor_re + i factor_im) every recorded sample
 *           i*m0*a0 is multiplied by (MRF: i, reference simulators/mrf.py:84).
 * sample_sign: DEVICE (output_count) signs from bringing pulses onto one axis
 *           (reference sequence/_accelerators.py:549-571), or NULL.
 * signal:   complex64 (atom_count, output_count), or NULL.
 * jacobian: complex64 (atom_count, column_count, output_count), or NULL when
 *           column_count is 0.
 * tile:     state tile index (tsim_tile_shape), or -1 for the library's choice.
 */
int tsim_simulate_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                           int64_t state_count, int64_t output_count, int features,
                           int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                           float factor_re, float factor_im, const float* sample_sign,
                           void* signal, void* jacobian, int tile, void* stream);

/*
 * tsim_simulate_jacobian for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc). No counterpart in the reference: it exists for those callers.
 * Same arguments, checks and results (bitwise) as tsim_simulate_jacobian,
 * but the compiled event stream lives in a fixed device workspace that the
 * library holds on every GPU it runs on until the process ends (about
 * 3.1 MB), so the call allocates and frees nothing. Limit: event_count <=
 * tsim_noalloc_max_events() (16384), else status 1. Every call on one GPU
 * shares that GPU's workspace: calls that overlap on one GPU (two streams,
 * or two host threads) race and must be serialized by the caller.
 */
int tsim_simulate_jacobian_noalloc(void* const* packed, int64_t atom_count, int64_t event_count,
                                   int64_t state_count, int64_t output_count, int features,
                                   int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                                   float factor_re, float factor_im, const float* sample_sign,
                                   void* signal, void* jacobian, int tile, void* stream);

/* Largest event_count tsim_simulate_jacobian_noalloc accepts. */
int64_t tsim_noalloc_max_events(void);

/*
 * tsim_simulate_jacobian with the samples projected onto a temporal basis
 * inside the kernel: what the reference computes as
 * `Subspace.project(signal)` and `Subspace.project(derivative)` after a full
 * simulate / jacobian (recon/_operator.py:446-473, estimators/_nlls.py:268-273,
 * _subspace.py:85-95: coefficients = samples @ basis). The samples never
 * reach HBM; each atom's L lanes accumulate 4 basis directions per lane in
 * FP32 over the recorded samples.
 *
 * basis:    DEVICE complex64 (output_count, rank), row-major.
 * rank:     1 .. 16 directions (a tile of L lanes per atom carries up to 4 L;
 *           the library's tile choice takes that into account).
 * columns:  HOST array of column_count (<= 4) indices among {0 t1, 1 t2, 2 m0,
 *           3 b1}; no diffusion feature.
 * signal:   complex64 (atom_count, rank), or NULL.
 * jacobian: complex64 (atom_count, column_count, rank), or NULL when
 *           column_count is 0.
 * Other arguments as tsim_simulate_jacobian; output_count must be >= 1.
 */
int tsim_simulate_projected(void* const* packed, int64_t atom_count, int64_t event_count,
                            int64_t state_count, int64_t output_count, int features,
                            int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                            float factor_re, float factor_im, const float* sample_sign,
                            const void* basis, int64_t rank, void* signal, void* jacobian, int tile,
                            void* stream);

/*
 * Adjoint of the fused real-subspace system: gradients of a scalar loss with
 * respect to every event's flip angle and duration and to each atom's
 * tissue, given the loss's real cotangents on the recorded m0*a0 and on the
 * Jacobian columns. Replaces the reference's `_epg_cpu.simulate_vjp`
 * (first order, when jacobian_cotangent is NULL) and
 * `_epg_cpu.simulate_vjp_jvp` (forward-over-reverse along the columns)
 * (src/torchsim/_epg_cpu.cpp:11150-11755; kernels :5985-6227, :7028-7314;
 * tissue reduction :11366-11395) that `_NativeEpg.backward` /
 * `_NativeEpgJvp.backward` reach (sequence/_accelerators.py:1544-1586,
 * 1787-1850). The kernel walks the macro-op stream (one op per MRF pulse);
 * the dual state is checkpointed every segment_ops ops and recomputed per
 * segment. op_bound: the number of macro-ops the stream compiles to, or an
 * upper bound on it (the workspace is sized by it; the kernels read the true
 * count on the device and never pass the bound); -1 reads the count back
 * from the device (one stream synchronization).
 *
 * columns:  HOST array of column_count (<= 4) indices among {0 t1, 1 t2,
 *           2 m0, 3 b1}.
 * signal_cotangent:   DEVICE (atom_count, output_count) FP32, or NULL.
 * jacobian_cotangent: DEVICE (atom_count, column_count, output_count), or NULL.
 * grad_flip, grad_duration: DEVICE (event_count) FP32, summed over atoms in a
 *           fixed order (deterministic); events with no rotation get a zero
 *           flip gradient. An event's duration gradient is the derivative
 *           along the interval before it, also where that interval is zero
 *           (the reference's convention, _epg_cpu.cpp:6189-6214).
 * duration_mask: DEVICE (event_count) bytes, or NULL. NULL computes every
 *           event's duration gradient; otherwise an event whose interval is
 *           zero gets its duration gradient only where its byte is nonzero
 *           (0 elsewhere) -- a zero interval the caller will not
 *           differentiate costs the kernel an op of its own, so a caller
 *           passes the ones its loss reaches. Timed events always get theirs.
 * tissue_gradients: HOST array of tissue_gradient_count (17, tissue order)
 *           DEVICE pointers, each NULL or (atom_count) FP32 written with the
 *           gradient per atom; t1, t2, m0, b1 and inversion_efficiency are
 *           carried (another non-null entry is status 2); count 0 for none.
 * No diffusion, off-resonance or flow feature (status 2): those run on the
 * complex kernel, whose first-order adjoint is tsim_complex_vjp.
 */
int tsim_simulate_vjp_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                               int64_t state_count, int64_t output_count, int features,
                               int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                               const float* signal_cotangent, const float* jacobian_cotangent,
                               int64_t segment_ops, int64_t op_bound, float* grad_flip, float* grad_duration,
                               float* const* tissue_gradients, int64_t tissue_gradient_count,
                               const uint8_t* duration_mask, void* stream);

/*
 * First-order adjoint of the complex kernel (single pool or one exchanging
 * pool, diffusion, off-resonance, RF and transmit phase): gradients of a
 * scalar loss with respect to every event's flip, phase and duration and to
 * each atom's tissue, given the loss's complex cotangent on the signal.
 * Replaces the reference's `_epg_cpu.simulate_vjp` with real_axis -1
 * (src/torchsim/_epg_cpu.cpp:4839-5944, entry :11150-11416, tissue reduction
 * :11366-11395) that `_NativeEpg.backward` reaches for phase-cycled,
 * off-resonant, exchanging or diffusing protocols
 * (sequence/_accelerators.py:1544-1586). The state is checkpointed every
 * segment_events events and recomputed per segment.
 *
 * signal_cotangent: DEVICE complex64 (atom_count, output_count): dL/dRe(s) +
 *           i dL/dIm(s) of the output s = factor * m0 * F+[0] e^{-i phase}
 *           (both pools' F+[0] summed), the public signal tsim_complex_jacobian
 *           writes with the same factor.
 * segment_events: events per checkpoint, 0 for about sqrt(event_count).
 * grad_flip, grad_phase, grad_duration: DEVICE (event_count) FP32 or NULL,
 *           summed over atoms in a fixed order (deterministic); an event
 *           whose flip, phase or interval touches nothing gets 0.
 * tissue_gradients: HOST array of tissue_gradient_count (17, tissue order)
 *           DEVICE pointers, each NULL or (atom_count) FP32 written with the
 *           gradient per atom; t1, t2, m0, b1, b1_phase, b0,
 *           inversion_efficiency and diffusion (the scaled rate the kernel
 *           reads) are carried, another non-null entry is status 2; count 0
 *           for none. Other arguments as tsim_complex_jacobian.
 * No flow feature (status 2).
 */
int tsim_complex_vjp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                     int64_t output_count, int pool_kind, int features, int64_t tissue_stride, float factor_re,
                     float factor_im, const void* signal_cotangent, int64_t segment_events, float* grad_flip,
                     float* grad_phase, float* grad_duration, float* const* tissue_gradients,
                     int64_t tissue_gradient_count, void* stream);

/*
 * Forward-over-reverse of the complex kernel: the gradient, with respect to
 * every event's flip, phase and duration, of a loss that reads the signal
 * and its Jacobian columns, given their cotangents. Replaces the event part
 * of _epg_cpu.simulate_vjp_jvp with real_axis -1 (_epg_cpu.cpp:7759-9013,
 * entry :11572-11755) as `_NativeEpgJvp.backward` reaches it
 * (sequence/_accelerators.py:1787-1850) for phase-cycled, off-resonant,
 * exchanging (pool kind 2) and diffusing protocols.
 * columns: HOST array of column_count (<= 4) indices among T_T1, T_T2,
 *          T_M0, T_B1 (B1 needs the transmit feature).
 * signal_cotangent: DEVICE (atom_count, output_count) complex64, or NULL.
 * jacobian_cotangent: DEVICE (atom_count, column_count, output_count)
 *          complex64 (the Jacobian as tsim_complex_jacobian writes it).
 * grad_flip, grad_phase, grad_duration: DEVICE (event_count) FP32, summed
 *          over atoms in a fixed order; every event gets its duration
 *          gradient, zero intervals included.
 * One first-order pass for the signal (an M0 column's cotangent joins it:
 * the column is the signal over M0) and one forward-over-reverse pass per
 * T1, T2 or B1 column. Tissue gradients are not carried here (status 2 is
 * not returned for them: the call has no tissue outputs).
 */
int tsim_complex_vjp_jacobian(void* const* packed, int64_t atom_count, int64_t event_count,
                              int64_t state_count, int64_t output_count, int pool_kind, int features,
                              int64_t tissue_stride, float factor_re, float factor_im,
                              const void* signal_cotangent, const int32_t* columns, int64_t column_count,
                              const void* jacobian_cotangent, int64_t segment_events, float* grad_flip,
                              float* grad_phase, float* grad_duration, void* stream);

/*
 * Forward mode along one arbitrary direction: the signal and its directional
 * derivative for a tangent on any carried tissue property and on every
 * event's duration, flip and phase. Replaces the reference's
 * `_epg_cpu.simulate_jvp` (src/torchsim/_epg_cpu.cpp:2469-3005 complex,
 * 3136-3285 real, entry :11017-11148) that `_NativeEpg.jvp` reaches
 * (sequence/_accelerators.py:1511-1542); runs the complex state machine,
 * which is exact on real-subspace inputs too.
 *
 * tangents: HOST array of tangent_count (20, or 0 for none) DEVICE pointers in
 *           the reference's JVP order: the 17 tissue tangents (tissue order;
 *           t1 and t2 one value per atom, the others as tissue_stride says),
 *           then duration, flip, phase (one value per event); each NULL for
 *           zero. t1, t2, m0, b1, b1_phase, b0, inversion_efficiency and
 *           diffusion are carried; another non-null tissue tangent is
 *           status 2.
 * signal:   complex64 (atom_count, output_count) = factor * sample, or NULL.
 * tangent:  complex64 (atom_count, output_count), its derivative.
 * Other arguments as tsim_complex_jacobian.
 */
int tsim_complex_jvp(void* const* packed, int64_t atom_count, int64_t event_count, int64_t state_count,
                     int64_t output_count, int pool_kind, int features, int64_t tissue_stride, float factor_re,
                     float factor_im, const float* const* tangents, int64_t tangent_count, void* signal,
                     void* tangent, void* stream);

/*
 * Gauss-Newton normal equations, fused: for each voxel, run the real-subspace
 * signal and its Jacobian along `columns` and accumulate, at every sample, in
 * FP64: the cost sum |model - data|^2, the gradient J^T r and the upper
 * triangle of J^T J -- the Jacobian never reaches memory. Replaces the
 * reference's per-iteration `operator.A(x)` + `operator.jacobian(x)` (one
 * forward + one JVP kernel per property) and the lstsq on the stacked blocks
 * (recon/_gauss_newton.py:247-307, 633-651; recon/_operator.py:342-361).
 *
 * columns: HOST array of 1..4 indices among {0 t1, 1 t2, 2 m0, 3 b1}.
 * data:    DEVICE complex64 (atom_count, output_count), the measured samples.
 * normal:  DEVICE FP64 (atom_count, tsim_normal_stride()): [cost, g0..g3,
 *          H00 H01 H02 H03 H11 H12 H13 H22 H23 H33] (unused columns zero).
 * The model sample is factor * i * m0 * a0 (times sample_sign), as in
 * tsim_simulate_jacobian.
 */
int tsim_gauss_newton_normal(void* const* packed, int64_t atom_count, int64_t event_count,
                             int64_t state_count, int64_t output_count, int features,
                             int64_t tissue_stride, const int32_t* columns, int64_t column_count,
                             float factor_re, float factor_im, const float* sample_sign,
                             const void* data, double* normal, int tile, void* stream);

/*
 * The same normal equations from a Jacobian that passed through memory (any
 * sequence the fused kernel does not carry: complex states, exchange,
 * diffusion): signal, data DEVICE complex64 (voxels, outputs), jacobian
 * DEVICE complex64 (voxels, columns, outputs), columns 1..4; normal as above.
 * One wavefront per voxel reads its rows once, accumulating in FP64.
 * Replaces the stacked least-squares set-up of the reference's solve
 * (recon/_gauss_newton.py:247-307).
 */
int tsim_normal_from_jacobian(const void* signal, const void* jacobian, const void* data, int64_t voxels,
                              int64_t outputs, int columns, double* normal, void* stream);

/*
 * The damped Gauss-Newton step of every voxel in FP64, one thread per voxel:
 * H = D H D, g = D g with D the per-voxel chain-rule scale of the bound
 * transform (DEVICE FP64 (voxels, columns), NULL for none); then
 * (H + damping I) step = -g + offset by Cholesky (NaN where not positive
 * definite; offset DEVICE FP64 (voxels, columns) or NULL -- a damping that
 * pulls toward a reference passes damping * (reference - x)),
 * predicted = step . (damping step - g), gradient = g, curvature = max diag H.
 * Any output may be NULL. Replaces `direct` (recon/_gauss_newton.py:247-307).
 */
int tsim_solve_damped(const double* normal, const double* scale, const double* damping, const double* offset,
                      int64_t voxels, int columns, double* step, double* predicted, double* gradient,
                      double* curvature, void* stream);

/* Doubles per voxel in the normal-equation buffer. */
int tsim_normal_stride(void);

/*
 * One request of an uneven batch: its 26 packed DEVICE buffers and sizes, the
 * Jacobian columns it wants (HOST indices as tsim_simulate_jacobian), the
 * complex factor and optional sample signs, and either the complex64 outputs
 * (mode 0) or the measured data and normal-equation output (mode 1).
 */
typedef struct tsim_batch_request {
  void* packed[26];
  int64_t atom_count;
  int64_t event_count;
  int64_t state_count;
  int64_t output_count;
  int64_t tissue_stride;
  int32_t features;
  int32_t column_count;
  int32_t columns[5];
  float factor_re;
  float factor_im;
  const float* sample_sign;
  void* signal;
  void* jacobian;
  const void* data;
  double* normal;
} tsim_batch_request;

/*
 * The uneven-batch launcher: requests with their own event streams, state
 * counts (padded to the batch's tile), atom counts and outputs, run by ONE
 * persistent launch of `grid_blocks` workgroups (<= 0: one per work block)
 * walking a table of (request, first atom) work blocks. All requests of a
 * batch must share one kernel variant: the same tangent set, diffusion and
 * transmit features (the caller groups them). Every atom runs exactly the
 * arithmetic its request alone would, so results are bitwise the same alone
 * and in the mix. No reference counterpart: the reference runs one launch per
 * call (sequence/_accelerators.py:3262-3449).
 * mode: 0 signal + Jacobian (as tsim_simulate_jacobian), 1 normal equations
 * (as tsim_gauss_newton_normal). tile: -1 for the library's choice.
 * Under HIP graph capture the graph keeps the device address of every buffer
 * a request names (the 26 packed buffers, sample_sign, signal, jacobian,
 * data, normal): each must stay allocated, and must not be handed back to an
 * allocator that could give its memory to another tensor, for as long as the
 * graph may be replayed; replays read and write them in place.
 */
int tsim_simulate_batch(const tsim_batch_request* requests, int64_t count, int mode, int tile,
                        int64_t grid_blocks, void* stream);

/* sizeof(tsim_batch_request) as compiled into the library. */
int tsim_batch_request_size(void);

/*
 * The uneven-batch launcher of the complex kernel: requests with their own
 * event streams, state counts (padded to the lanes of the batch's largest),
 * atom counts, outputs and POOL COUNTS -- a request whose packed buffer
 * T_POOL_B_FRACTION is set runs the exchanging two-pool body, the others the
 * single-pool one -- in ONE persistent launch of `grid_blocks` workgroups
 * (<= 0: one per work block of 4 (64 / lanes) atoms). Each request writes
 * its signal (times factor) and Jacobian as tsim_complex_jacobian does,
 * columns among T_T1, T_T2, T_M0, T_B1; all requests of a batch share one
 * tangent set (the caller groups them). sample_sign, data and normal must be
 * NULL. Every atom runs the arithmetic it would run alone, so results are
 * bitwise those of tsim_complex_jacobian. No reference counterpart: the
 * reference runs one launch per call and pool model
 * (sequence/_accelerators.py:3262-3449, model/_physics.py:214-262).
 */
int tsim_complex_batch(const tsim_batch_request* requests, int64_t count, int64_t grid_blocks,
                       void* stream);

/*
 * One fixed-iteration Levenberg-Marquardt step of every voxel, in two
 * launches around the normal equations at the candidate (no reference
 * counterpart as entry points: they fuse the per-voxel bookkeeping of
 * recon/_gauss_newton.py:365-415 and 474-631 -- bounds chain rule, damped
 * solve, gain ratio and Nielsen's damping update -- that the serving loop
 * runs per request and iteration).
 * tsim_lm_propose: from x (voxels x columns, free variables), its normal
 * equations (voxels x tsim_normal_stride(), FP64) and damping, writes the
 * candidate (free), its natural values as float32 planes [column][voxel],
 * the predicted reduction and a failed-solve flag (1 byte per voxel).
 * bound_kind per column: 0 free, 1 two-sided (low + (high - low) sigmoid),
 * 2 lower (low + softplus), 3 upper (high - softplus(-x)).
 * tsim_lm_accept: from the normal equations at the candidate, accepts or
 * rejects each voxel's step and updates x, normal, cost, damping and its
 * growth factor in place.
 */
int tsim_lm_propose(const double* normal, const double* x, const double* damping, const int32_t* bound_kind,
                    const double* low, const double* high, int64_t voxels, int columns, double* candidate,
                    float* natural, double* predicted, uint8_t* singular, void* stream);
int tsim_lm_accept(const double* trial, const double* predicted, const uint8_t* singular, const double* candidate,
                   int64_t voxels, int columns, double* x, double* normal, double* cost, double* damping,
                   double* rising, void* stream);

/*
 * Dictionary matching (replaces DictionaryMatcher._match_here with its
 * correlate / topk chunks, _density and _scaled: reference
 * estimators/_dictionary.py:308-431, estimators/_grouped.py:219-245).
 * A dictionary is laid out once as two fp16 operand planes of
 * atoms_pad x k_pad halves each (tsim_match_operand_size), with its FP64 row
 * norms; tsim_match then scores every voxel against every atom on the MFMA
 * (normalized |<y, d>|, never stored), keeps the best top_k + 1 per voxel,
 * recomputes those exactly in FP64 and writes, per voxel and rank, the atom
 * index (int64), score (float32), complex least-squares scale (complex64)
 * and density (float32). Ties rank the lower atom index first.
 * mode: 0 real dictionary (float32); complex64 dictionaries 1 (imaginary
 * parts all zero), 2 (real parts all zero) or 3 (general), chosen from the
 * flags tsim_match_planes ORs into a device int (bit 0: some real part is
 * nonzero, bit 1: some imaginary part is). Signals are float32 or complex64
 * rows of `contrasts` values. top_k 1..3; splits <= 0 picks the split of the
 * dictionary across workgroups.
 */
int tsim_match_operand_size(int64_t atoms, int64_t contrasts, int mode, int64_t* atoms_pad, int64_t* k_pad);
int tsim_match_planes(const void* dictionary, int64_t atoms, int64_t contrasts, int* flags, void* stream);
int tsim_match_layout_atoms(const void* dictionary, int dictionary_complex, int64_t atoms, int64_t contrasts,
                            int mode, double* atom_norm, void* hi, void* lo, void* stream);
int tsim_match(const void* atoms_hi, const void* atoms_lo, int mode, int64_t atoms, const void* dictionary,
               int dictionary_complex, const double* atom_norm, const void* signals, int signals_complex,
               int64_t voxels, int64_t contrasts, int top_k, int64_t splits, int64_t* indices, float* scores,
               void* scales, float* densities, void* stream);

/* Number of state tiles compiled into the library, and the shape of one:
 * orders per lane and lanes per atom. No reference counterpart (tuning). */
int tsim_tile_count(void);
int tsim_tile_shape(int tile, int* orders_per_lane, int* lanes_per_atom);

#ifdef __cplusplus
}
#endif

#endif /* TORCHSIM_HIP_H */
