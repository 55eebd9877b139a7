/*
 * fdsp.h -- C ABI of the MI355X (gfx950) streaming-channelizer library.
 *
 * Every entry point takes plain pointers and sizes and returns an int status
 * (FDSP_OK or a negative FDSP_E* code; fdsp_last_error() describes the last
 * failure on the calling thread). Nothing here ever calls exit().
 *
 * Two calling forms per operation:
 *   *_apply        host pointers, synchronous: the drop-in form that the
 *                  reference's std::vector based classes map onto.
 *   *_apply_device device pointers, asynchronous on `stream` (a hipStream_t;
 *                  NULL means the handle's own stream). Output counts are
 *                  exact and known on return, before the kernels finish.
 *
 * Complex samples are interleaved float pairs {re, im} (std::complex<float>).
 * Handles are not thread-safe: one owner per handle.
 */
#ifndef FDSP_H
#define FDSP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define FDSP_OK        0
#define FDSP_EINVAL   -1  /* bad argument */
#define FDSP_ENOTINIT -2  /* handle not initialised */
#define FDSP_EBUSY    -3  /* resource in use */
#define FDSP_ENOMEM   -4  /* device or host allocation failed / over budget */
#define FDSP_EHIP     -5  /* HIP runtime error */
#define FDSP_ERANGE   -6  /* size above a limit, or output capacity too small */
#define FDSP_EAGAIN   -7  /* serving: not admitted yet, retry later */

typedef struct fdsp_cf32 { float re, im; } fdsp_cf32;
/* complex int16 sample (std::complex<int16_t>): read exactly as float; results
 * written as int16 are truncated toward zero, as the reference's float ->
 * int16_t conversions are, and saturated outside the int16 range (where the
 * reference's conversion is undefined) */
typedef struct fdsp_ci16 { int16_t re, im; } fdsp_ci16;
typedef void *fdsp_stream_t; /* hipStream_t */
typedef void *fdsp_event_t;  /* hipEvent_t */

const char *fdsp_last_error(void);
int fdsp_version(void);

/* ---- device plumbing (plain HIP calls behind a C ABI; used by hosts that
 *      have no HIP headers, e.g. ctypes) ---- */
int fdsp_device_count(int *count);
int fdsp_set_device(int device);
int fdsp_malloc(void **ptr, size_t bytes);
int fdsp_free(void *ptr);
int fdsp_memcpy_h2d(void *dst, const void *src, size_t bytes, fdsp_stream_t stream);
int fdsp_memcpy_d2h(void *dst, const void *src, size_t bytes, fdsp_stream_t stream);
int fdsp_memcpy_d2d(void *dst, const void *src, size_t bytes, fdsp_stream_t stream);
int fdsp_memset(void *dst, int value, size_t bytes, fdsp_stream_t stream);
int fdsp_host_alloc_pinned(void **ptr, size_t bytes);
int fdsp_host_free_pinned(void *ptr);
int fdsp_stream_create(fdsp_stream_t *stream);
int fdsp_stream_destroy(fdsp_stream_t stream);
int fdsp_stream_sync(fdsp_stream_t stream);
int fdsp_device_sync(void);
int fdsp_event_create(fdsp_event_t *ev);
int fdsp_event_destroy(fdsp_event_t ev);
int fdsp_event_record(fdsp_event_t ev, fdsp_stream_t stream);
int fdsp_event_sync(fdsp_event_t ev);
int fdsp_event_elapsed_ms(fdsp_event_t start, fdsp_event_t stop, float *ms);

/* ---- rate helpers ----
 * fdsp_gcd / fdsp_lcm replace calculate_gcd / calculate_lcm
 *   (reference src/utilities/falcon_dsp_utils.cc:123-137).
 * fdsp_rat_approx replaces rat_approx (falcon_dsp_utils.cc:231-294).
 * fdsp_freq_shift_params replaces falcon_dsp_freq_shift::get_freq_shift_params
 *   (reference src/transform/falcon_dsp_freq_shift.cc:194-235); f_hz == 0 is
 *   defined here as the identity shift (rollover 1, omega 0). */
uint64_t fdsp_gcd(uint64_t a, uint64_t b);
uint64_t fdsp_lcm(uint64_t a, uint64_t b);
void fdsp_rat_approx(double f, int64_t max_denom, int64_t *num, int64_t *denom);
int fdsp_freq_shift_params(uint32_t fs, int32_t f_hz, uint32_t *rollover, double *omega);

/* ---- NCO frequency shift ----
 * Replaces falcon_dsp_freq_shift::{initialize, apply(cf32), reset_state}
 * (reference hdr/transform/falcon_dsp_freq_shift.h:98-121, src .cc:107-192) and
 * falcon_dsp_freq_shift_cuda::{initialize(fs, vector<int32>),
 * apply(in, vector<vector<cf32>>)} (hdr/transform/falcon_dsp_freq_shift_cuda.h:187-212,
 * src/transform/falcon_dsp_freq_shift_cuda.cu:355-488): one input, n_shifts
 * outputs, y_c[n] = x[n] * exp(j*w_c*((s_c + n) mod R_c)), s_c carried. */
typedef struct fdsp_freq_shift_s *fdsp_freq_shift_t;
int fdsp_freq_shift_create(uint32_t fs, const int32_t *shifts_hz, uint32_t n_shifts,
                           fdsp_freq_shift_t *out);
int fdsp_freq_shift_destroy(fdsp_freq_shift_t h);
int fdsp_freq_shift_reset(fdsp_freq_shift_t h);
/* samples handled modulo the rollover index for shift `chan` */
int fdsp_freq_shift_get_state(fdsp_freq_shift_t h, uint32_t chan, uint64_t *samples_handled);
/* outs[c] receives n samples for each shift c */
int fdsp_freq_shift_apply(fdsp_freq_shift_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *const *outs);
/* output of shift c starts at d_out + c * out_stride (in samples) */
int fdsp_freq_shift_apply_device(fdsp_freq_shift_t h, const fdsp_cf32 *d_in, size_t n,
                                 fdsp_cf32 *d_out, size_t out_stride, fdsp_stream_t stream);
/* n_chunks consecutive chunks of chunk_len samples from d_in in one launch
 * (no reference counterpart: the reference makes one apply call per chunk,
 * falcon_dsp_freq_shift_cuda.cu:355-488). The same outputs as n_chunks
 * apply_device calls: shift c of chunk k at d_out + c*out_stride +
 * k*chunk_len, the index carried across the chunks. */
int fdsp_freq_shift_apply_chunks_device(fdsp_freq_shift_t h, const fdsp_cf32 *d_in, size_t chunk_len,
                                        size_t n_chunks, fdsp_cf32 *d_out, size_t out_stride,
                                        fdsp_stream_t stream);
/* Restores a stream position: every shift continues as if samples_handled
 * samples had been processed since reset (its index samples_handled mod R),
 * as the reference's m_samples_handled carries it
 * (hdr/transform/falcon_dsp_freq_shift.h:117; the reference has no setter).
 * The position of a one-shift handle round-trips through get_state. */
int fdsp_freq_shift_set_state(fdsp_freq_shift_t h, uint64_t samples_handled);
/* ci16 in, ci16 out. Replaces falcon_dsp_freq_shift::apply(vector<complex<int16_t>>&, ...)
 * (reference src/transform/falcon_dsp_freq_shift.cc:133-162) and
 * falcon_dsp_freq_shift_cuda::apply(ci16) (falcon_dsp_freq_shift_cuda.cu:308-329):
 * the shift of the float-converted input, each part truncated back to int16.
 * Shares the stream position with the cf32 calls. */
int fdsp_freq_shift_apply_ci16(fdsp_freq_shift_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *const *outs);
int fdsp_freq_shift_apply_ci16_device(fdsp_freq_shift_t h, const fdsp_ci16 *d_in, size_t n,
                                      fdsp_ci16 *d_out, size_t out_stride, fdsp_stream_t stream);

/* ---- polyphase rational resampler ----
 * Replaces falcon_dsp_polyphase_resampler::{initialize, apply(cf32),
 * needed_out_count, coeffs_per_phase, reset_state}
 * (reference hdr/transform/falcon_dsp_polyphase_resampler.h:168-195,
 * src .cc:114-346) and falcon_dsp_polyphase_resampler_cuda::apply
 * (src/transform/falcon_dsp_polyphase_resampler_cuda.cu:459-619). Only the
 * real part of each tap is used, as in the reference (.cc:288,298). */
typedef struct fdsp_resampler_s *fdsp_resampler_t;
int fdsp_resampler_create(uint32_t up, uint32_t down, const fdsp_cf32 *taps, uint32_t n_taps,
                          fdsp_resampler_t *out);
int fdsp_resampler_destroy(fdsp_resampler_t h);
int fdsp_resampler_reset(fdsp_resampler_t h);
int fdsp_resampler_needed_out_count(fdsp_resampler_t h, uint32_t in_count, uint32_t *out_count);
int fdsp_resampler_coeffs_per_phase(fdsp_resampler_t h, uint32_t *cpp);
/* coeff_phase and xOffset exactly as the reference carries them, plus the
 * totals of inputs consumed and outputs produced since the last reset */
int fdsp_resampler_get_state(fdsp_resampler_t h, uint32_t *coeff_phase, uint32_t *x_offset,
                             uint64_t *in_total, uint64_t *out_total);
int fdsp_resampler_apply(fdsp_resampler_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out,
                         size_t cap, size_t *n_out);
int fdsp_resampler_apply_device(fdsp_resampler_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                size_t cap, size_t *n_out, fdsp_stream_t stream);
/* n_chunks consecutive chunks of chunk_len samples from d_in in one launch
 * (no reference counterpart: the reference makes one apply call per chunk,
 * falcon_dsp_polyphase_resampler_cuda.cu:459-619). The same outputs, back to
 * back in d_out, as n_chunks apply_device calls; counts[k] receives chunk k's
 * output count. */
int fdsp_resampler_apply_chunks_device(fdsp_resampler_t h, const fdsp_cf32 *d_in, size_t chunk_len,
                                       size_t n_chunks, fdsp_cf32 *d_out, size_t cap, size_t *counts,
                                       fdsp_stream_t stream);
/* Restores the reference's carried state (polyphase_resampler_params_s
 * coeff_phase, xOffset and state, hdr/transform/falcon_dsp_polyphase_resampler.h:157-163;
 * the reference has no setter): history holds the last coeffs_per_phase - 1
 * inputs, oldest first. FDSP_EINVAL when no stream position yields the two
 * fields. A state read with get_state plus the last inputs round-trips. */
int fdsp_resampler_set_state(fdsp_resampler_t h, uint32_t coeff_phase, uint32_t x_offset,
                             const fdsp_cf32 *history, size_t n_history);

/* ---- FIR filter (complex taps) ----
 * Replaces falcon_dsp_fir_filter(coeffs) / apply(cf32) / reset_state
 * (reference hdr/transform/falcon_dsp_fir_filter.h:119-140, src .cc:136-284)
 * and falcon_dsp_fir_filter_cuda (src/transform/falcon_dsp_fir_filter_cuda.cu:256-373):
 * y[n] = sum_k h[k] x[n-k], history carried. Empty input is FDSP_EINVAL. */
typedef struct fdsp_fir_s *fdsp_fir_t;
int fdsp_fir_create(const fdsp_cf32 *coeffs, uint32_t n_coeffs, fdsp_fir_t *out);
int fdsp_fir_destroy(fdsp_fir_t h);
int fdsp_fir_reset(fdsp_fir_t h);
int fdsp_fir_apply(fdsp_fir_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out);
int fdsp_fir_apply_device(fdsp_fir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                          fdsp_stream_t stream);
/* ci16 in, ci16 out. Replaces falcon_dsp_fir_filter::apply(vector<complex<int16_t>>&, ...)
 * (reference src/transform/falcon_dsp_fir_filter.cc:151-175) and
 * falcon_dsp_fir_filter_cuda::apply(ci16) (falcon_dsp_fir_filter_cuda.cu:233-254):
 * the filter over the float-converted input, outputs truncated to int16. The
 * history stays float and is shared with the cf32 calls. */
int fdsp_fir_apply_ci16(fdsp_fir_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out);
int fdsp_fir_apply_ci16_device(fdsp_fir_t h, const fdsp_ci16 *d_in, size_t n, fdsp_ci16 *d_out,
                               fdsp_stream_t stream);

/* ---- FIR bank: many independent channels, real taps each ----
 * The multi-channel form of falcon_dsp_fir_filter::apply (reference
 * src/transform/falcon_dsp_fir_filter.cc:177-243) with real taps: channel c
 * filters d_in + c*in_stride with taps[c*n_taps ...], history carried per
 * channel. One launch covers all channels. */
typedef struct fdsp_fir_bank_s *fdsp_fir_bank_t;
int fdsp_fir_bank_create(uint32_t n_chan, uint32_t n_taps, const float *taps, fdsp_fir_bank_t *out);
int fdsp_fir_bank_destroy(fdsp_fir_bank_t h);
int fdsp_fir_bank_reset(fdsp_fir_bank_t h);
int fdsp_fir_bank_apply_device(fdsp_fir_bank_t h, const fdsp_cf32 *d_in, size_t n, size_t in_stride,
                               fdsp_cf32 *d_out, size_t out_stride, fdsp_stream_t stream);

/* ---- polar (FM) discriminator ----
 * Replaces falcon_dsp_polar_discriminator::{apply(cf32), reset_state}
 * (reference hdr/transform/falcon_dsp_polar_discriminator.h:97-114, src .cc:108-171):
 * out = atan2(Im z, Re z), z = x[n]*conj(x[n-1]); the first call after a reset
 * emits n-1 values, later calls n. */
typedef struct fdsp_discriminator_s *fdsp_discriminator_t;
int fdsp_discriminator_create(fdsp_discriminator_t *out);
int fdsp_discriminator_destroy(fdsp_discriminator_t h);
int fdsp_discriminator_reset(fdsp_discriminator_t h);
int fdsp_discriminator_apply(fdsp_discriminator_t h, const fdsp_cf32 *in, size_t n, float *out,
                             size_t *n_out);
int fdsp_discriminator_apply_device(fdsp_discriminator_t h, const fdsp_cf32 *d_in, size_t n, float *d_out,
                                    size_t *n_out, fdsp_stream_t stream);
/* ci16 in. Replaces falcon_dsp_polar_discriminator::apply(const vector<complex<int16_t>>&,
 * vector<float>&) (reference src/transform/falcon_dsp_polar_discriminator.cc:117-132):
 * the float-converted input through the same discriminator; the carried
 * sample is shared with the cf32 calls. */
int fdsp_discriminator_apply_ci16(fdsp_discriminator_t h, const fdsp_ci16 *in, size_t n, float *out,
                                  size_t *n_out);
int fdsp_discriminator_apply_ci16_device(fdsp_discriminator_t h, const fdsp_ci16 *d_in, size_t n, float *d_out,
                                         size_t *n_out, fdsp_stream_t stream);

/* ---- IIR filter ----
 * Replaces falcon_dsp_iir_filter::{initialize, apply (complex<float>, float,
 * complex<int16_t>, int16_t), reset_state} (reference
 * hdr/transform/falcon_dsp_iir_filter.h:97-130, src/transform/
 * falcon_dsp_iir_filter.cc:104-396). y[n] = sum_k b[k] x[n-k] -
 * sum_{j>=1} a[j] y[n-j]; when a[0]'s real part is not 1 every b and a[1:] is
 * first divided by a[0], as the reference does (its complex<float> division).
 * Inputs and outputs are carried between calls; one handle's calls may mix
 * the four sample types. int16 outputs are truncated toward zero and
 * saturated (out of range is undefined in the reference). Limits: 1 <= nb <=
 * 256, 1 <= na <= 17, a[0] != 0 (the reference would divide by zero).
 *
 * The recurrence runs in parallel segments, each started `warm` samples
 * early from a zero state. When such segments meet the exact trajectory bit
 * for bit (a trial at create decides; first-order filters such as the FM
 * de-emphasis do), each is verified against its predecessor's exact end
 * state and recomputed if it differs: outputs are the reference's bit for
 * bit (up to the sign of zero). Higher-order filters whose feedback amplifies
 * rounding differences keep a few-ulp gap for ever; for them the modes are:
 *   FDSP_IIR_AUTO (default): float outputs from the unverified segments
 *     (within float rounding of the reference, as the library's other float
 *     kernels are), int16 outputs from one sequential wave (bit for bit);
 *     a filter that does not decay within 2048 samples runs sequentially;
 *   FDSP_IIR_EXACT: always bit for bit (a single stream's exact recurrence is
 *     inherently serial and runs as one wave);
 *   FDSP_IIR_FAST: always the unverified segments. */
#define FDSP_IIR_FAST 0
#define FDSP_IIR_EXACT 1
#define FDSP_IIR_AUTO 2
#define FDSP_IIR_PLAN_FAST 0
#define FDSP_IIR_PLAN_SEGMENTS 1
#define FDSP_IIR_PLAN_SERIAL 2
typedef struct fdsp_iir_s *fdsp_iir_t;
int fdsp_iir_create(const fdsp_cf32 *b, uint32_t nb, const fdsp_cf32 *a, uint32_t na, fdsp_iir_t *out);
int fdsp_iir_destroy(fdsp_iir_t h);
int fdsp_iir_reset(fdsp_iir_t h);
/* FDSP_IIR_AUTO (default), FDSP_IIR_EXACT or FDSP_IIR_FAST for the following calls */
int fdsp_iir_set_mode(fdsp_iir_t h, int mode);
/* warm: samples each segment's recurrence runs before its outputs; order:
 * na - 1; plan: the FDSP_IIR_PLAN_* a call with complex float outputs runs */
int fdsp_iir_info(fdsp_iir_t h, uint32_t *warm, uint32_t *order, uint32_t *plan);
int fdsp_iir_apply(fdsp_iir_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out);
int fdsp_iir_apply_f32(fdsp_iir_t h, const float *in, size_t n, float *out);
int fdsp_iir_apply_ci16(fdsp_iir_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out);
int fdsp_iir_apply_i16(fdsp_iir_t h, const int16_t *in, size_t n, int16_t *out);
int fdsp_iir_apply_device(fdsp_iir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out, fdsp_stream_t stream);
int fdsp_iir_apply_f32_device(fdsp_iir_t h, const float *d_in, size_t n, float *d_out, fdsp_stream_t stream);
int fdsp_iir_apply_ci16_device(fdsp_iir_t h, const fdsp_ci16 *d_in, size_t n, fdsp_ci16 *d_out,
                               fdsp_stream_t stream);
int fdsp_iir_apply_i16_device(fdsp_iir_t h, const int16_t *d_in, size_t n, int16_t *d_out, fdsp_stream_t stream);
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops). The same
 * filtering as fdsp_iir_apply_device, with the kernels' per-segment state
 * records and block flags in a 16 MiB array that the library holds on each
 * device until the process ends, in place of buffers grown per handle. It
 * calls no allocator. A chunk needs 2 * roundup256(ceil(n/seg) * P * 8) +
 * 8 * ceil(n / (64 seg)) bytes of it (P: the order rounded up to a power of
 * two; seg: 8 for n <= 2^21, 16 for n <= 2^22, 32 for n <= 2^23, else 64):
 * chunks of up to 2^22 samples fit at orders 1-2, 2^20 at orders 3-4, 2^19 at
 * 5-8 and 2^18 at 9-16; a larger chunk fails with FDSP_ERANGE and changes
 * nothing. Every such call on one device
 * shares the array: calls that overlap on a device (other streams, other
 * threads) race and must be serialised by the caller. */
int fdsp_iir_apply_device_noalloc(fdsp_iir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                  fdsp_stream_t stream);

/* ---- IIR bank: many streams, each exact ----
 * n_streams independent falcon_dsp_iir_filter instances sharing one (b, a)
 * (reference src/transform/falcon_dsp_iir_filter.cc:104-396 per stream):
 * each stream's outputs equal the reference's sequential loop on that
 * stream, bit for bit up to the sign of zero, for every filter (the bank
 * runs one stream per GPU lane, so no filter needs the serial plan). Rows:
 * stream s's n samples at d_in + s * in_stride, its outputs at d_out + s *
 * out_stride (strides in samples, >= n); the host forms take n_streams
 * contiguous rows of n. Every stream advances by n samples per call; its
 * carried inputs and outputs stay on the device. int16 outputs are truncated
 * toward zero and saturated, as fdsp_iir_apply_ci16. Limits: 1 <= nb <= 17,
 * 1 <= na <= 17, a[0] != 0. Calls on one bank must be ordered (one stream). */
typedef struct fdsp_iir_bank_s *fdsp_iir_bank_t;
int fdsp_iir_bank_create(const fdsp_cf32 *b, uint32_t nb, const fdsp_cf32 *a, uint32_t na, uint32_t n_streams,
                         fdsp_iir_bank_t *out);
int fdsp_iir_bank_destroy(fdsp_iir_bank_t h);
int fdsp_iir_bank_reset(fdsp_iir_bank_t h);
int fdsp_iir_bank_apply(fdsp_iir_bank_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out);
int fdsp_iir_bank_apply_ci16(fdsp_iir_bank_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out);
int fdsp_iir_bank_apply_device(fdsp_iir_bank_t h, const fdsp_cf32 *d_in, size_t n, size_t in_stride,
                               fdsp_cf32 *d_out, size_t out_stride, fdsp_stream_t stream);
int fdsp_iir_bank_apply_ci16_device(fdsp_iir_bank_t h, const fdsp_ci16 *d_in, size_t n, size_t in_stride,
                                    fdsp_ci16 *d_out, size_t out_stride, fdsp_stream_t stream);

/* ---- first-order IIR (real samples) ----
 * Replaces falcon_dsp_iir_filter::{initialize(b = [b0], a = [1, a1]),
 * apply(vector<float>&, vector<float>&), reset_state} (reference
 * src/transform/falcon_dsp_iir_filter.cc:104-203,235-396) for the first-order
 * filters the FM de-emphasis uses: y[n] = b0 x[n] - a1 y[n-1]. A shorthand
 * for fdsp_iir_create with b = [b0], a = [1, a1]; bit-identical outputs. */
typedef struct fdsp_iir1_s *fdsp_iir1_t;
int fdsp_iir1_create(float b0, float a1, fdsp_iir1_t *out);
int fdsp_iir1_destroy(fdsp_iir1_t h);
int fdsp_iir1_reset(fdsp_iir1_t h);
int fdsp_iir1_apply(fdsp_iir1_t h, const float *in, size_t n, float *out);
int fdsp_iir1_apply_device(fdsp_iir1_t h, const float *d_in, size_t n, float *d_out, fdsp_stream_t stream);

/* ---- predefined resampling filters ----
 * Replaces get_resample_fir_params (reference src/transform/
 * falcon_dsp_fir_filter.cc:111-130) and its table (src/transform/
 * falcon_dsp_predefined_fir_filter.cc, keys as in hdr/transform/
 * falcon_dsp_predefined_fir_filter.h:40-110). The filters are designed at build
 * time by the reference's documented procedure (firls with a Kaiser window,
 * or remez) for 1,140,000 -> 228,000 / 200,000, 1,000,000 -> 200,000,
 * 228,000 -> 45,600 and 200,000 -> 45,000 / 44,000 sps. taps: 32, 64, 128,
 * 256 or FDSP_FILTER_TAPS_OPTIMAL; source: FDSP_FILTER_REMEZ or
 * FDSP_FILTER_FIRLS. FDSP_EINVAL when no such filter exists (the reference
 * returns false); FDSP_ERANGE, with *n_coeffs set, when cap is too small. */
#define FDSP_FILTER_REMEZ 0
#define FDSP_FILTER_FIRLS 1
#define FDSP_FILTER_TAPS_OPTIMAL -1
int fdsp_resample_fir_params(uint32_t in_rate, uint32_t out_rate, int32_t taps, uint32_t source, uint32_t *up,
                             uint32_t *down, fdsp_cf32 *coeffs, size_t cap, size_t *n_coeffs);

/* ---- FM mono demodulator ----
 * Replaces falcon_dsp_fm_demodulator::{initialize, demod_mono, reset_state}
 * (reference src/signals/falcon_dsp_fm_demod.cc:98-271): shift by the signal's
 * offset, resample to the FM band rate, polar discriminator, single-pole
 * de-emphasis (time constant tau, coefficients from fdsp_fm_deemphasis),
 * resample to the audio rate, then scale each chunk to a peak of 10000 and
 * convert to int16 (truncation; all zero for an all-zero chunk). The
 * reference reads both resamplers' (up, down, taps) from its predefined
 * REMEZ-64 table for 1,140,000 -> 228,000 -> 45,600 sps; here the caller
 * passes them. audio (optional) receives the samples before scaling. */
int fdsp_fm_deemphasis(double tau, float *b0, float *a1);
typedef struct fdsp_fm_demod_s *fdsp_fm_demod_t;
int fdsp_fm_demod_create(uint32_t fs, int32_t offset_hz, double tau, uint32_t iso_up, uint32_t iso_down,
                         const fdsp_cf32 *iso_taps, uint32_t n_iso, uint32_t mono_up, uint32_t mono_down,
                         const fdsp_cf32 *mono_taps, uint32_t n_mono, fdsp_fm_demod_t *out);
int fdsp_fm_demod_destroy(fdsp_fm_demod_t h);
int fdsp_fm_demod_reset(fdsp_fm_demod_t h);
/* most outputs the next n-sample chunk can emit */
int fdsp_fm_demod_out_count(fdsp_fm_demod_t h, size_t n, size_t *count);
int fdsp_fm_demod_apply(fdsp_fm_demod_t h, const fdsp_cf32 *in, size_t n, int16_t *out, float *audio, size_t cap,
                        size_t *n_out);
int fdsp_fm_demod_apply_device(fdsp_fm_demod_t h, const fdsp_cf32 *d_in, size_t n, int16_t *d_out, float *d_audio,
                               size_t cap, size_t *n_out, fdsp_stream_t stream);

/* ---- multi-rate channelizer ----
 * Replaces falcon_dsp_multi_rate_channelizer{,_cuda}::{initialize, apply, reset_state}
 * (reference hdr/transform/falcon_dsp_multi_rate_channelizer.h:70-134,
 * src .cc:81-173; hdr/transform/falcon_dsp_multi_rate_channelizer_cuda.h:102-167,
 * src/transform/falcon_dsp_multi_rate_channelizer_cuda.cu:87-577). The channel
 * descriptor mirrors multi_rate_channelizer_channel_s (.h:70-77); the shift is
 * narrowed to int32 as the reference does.
 *
 * FDSP_CHAN_PLAIN keeps the reference's structure on the GPU: one multi-shift
 * NCO launch writing a full-rate copy per channel, then one resampler launch
 * per channel. FDSP_CHAN_BATCHED runs every channel in one fused launch that
 * reads the input once (shift folded into modulated taps, input staged in LDS),
 * in FP32 arithmetic throughout, as the reference's multiply-accumulates
 * (src/transform/falcon_dsp_polyphase_resampler.cc:288,298). Both produce the
 * reference's samples and counts.
 * FDSP_CHAN_BATCHED_BF16X2 is the batched form with reduced-precision matrix
 * products, opt-in only: channel groups with 32-tap windows and an even down
 * below 32 multiply bf16 hi/lo splits of the input and taps (about 16
 * significant bits; the hi*hi, hi*lo and lo*hi products, FP32 accumulation).
 * Its outputs stay within 1e-4 x input RMS of the reference but are not FP32
 * results. */
#define FDSP_CHAN_PLAIN          0
#define FDSP_CHAN_BATCHED        1
#define FDSP_CHAN_BATCHED_BF16X2 2

#define FDSP_TAIL_NONE          0
#define FDSP_TAIL_DISCRIMINATOR 1 /* channel output is float angles (polar discriminator) */

typedef struct fdsp_channel_desc {
    uint32_t output_sample_rate_in_sps;
    int64_t freq_shift_in_hz;
    uint32_t up_rate;
    uint32_t down_rate;
    const fdsp_cf32 *taps;
    uint32_t n_taps;
    uint32_t tail; /* FDSP_TAIL_* (batched mode only) */
} fdsp_channel_desc;

typedef struct fdsp_channelizer_s *fdsp_channelizer_t;
int fdsp_channelizer_create(uint32_t fs, const fdsp_channel_desc *chans, uint32_t n_chan, int mode,
                            fdsp_channelizer_t *out);
int fdsp_channelizer_destroy(fdsp_channelizer_t h);
int fdsp_channelizer_reset(fdsp_channelizer_t h);
int fdsp_channelizer_num_channels(fdsp_channelizer_t h, uint32_t *n_chan);
/* exact count channel `chan` emits for the next in_count inputs */
int fdsp_channelizer_needed_out_count(fdsp_channelizer_t h, uint32_t chan, uint32_t in_count,
                                      uint32_t *out_count);
/* per-channel carried state: NCO samples handled (mod R), resampler
 * coeff_phase / xOffset, totals since reset */
int fdsp_channelizer_get_state(fdsp_channelizer_t h, uint32_t chan, uint64_t *nco_samples_handled,
                               uint32_t *coeff_phase, uint32_t *x_offset, uint64_t *in_total,
                               uint64_t *out_total);
/* outs[c] holds caps[c] samples (float for discriminator channels); counts[c]
 * receives the number written */
int fdsp_channelizer_apply(fdsp_channelizer_t h, const fdsp_cf32 *in, size_t n, void *const *outs,
                           const size_t *caps, size_t *counts);
int fdsp_channelizer_apply_device(fdsp_channelizer_t h, const fdsp_cf32 *d_in, size_t n,
                                  void *const *d_outs, const size_t *caps, size_t *counts,
                                  fdsp_stream_t stream);

/* ---- batched channelizer as a replayable HIP graph ----
 * Captures one chunk of chunk_len samples from d_in into d_outs (fixed
 * buffers) as a graph whose last node advances the carried state on the
 * device, so consecutive launches process consecutive chunks with no host
 * work. Counts of the last launch are read back with ..._graph_counts. */
int fdsp_channelizer_graph_create(fdsp_channelizer_t h, size_t chunk_len, const fdsp_cf32 *d_in,
                                  void *const *d_outs, const size_t *caps);
int fdsp_channelizer_graph_launch(fdsp_channelizer_t h, fdsp_stream_t stream);
int fdsp_channelizer_graph_counts(fdsp_channelizer_t h, size_t *counts);
int fdsp_channelizer_graph_destroy(fdsp_channelizer_t h);

/* Work split of the batched kernel (picked by the autotuner). The first two
 * forms launch as many blocks as the GPU holds at once, each starting in a
 * channel group picked in proportion to the group's estimated cost.
 *  - FDSP_TILING_STRIDED (1, the default): block s of a group runs its tiles
 *    s, s + b, ... (b = the group's block count).
 *  - FDSP_TILING_CLAIM (0): blocks claim output tiles one at a time from
 *    per-channelizer device counters and move to other groups when theirs
 *    is done. Launches of one channelizer must not overlap in time (they
 *    share the counters); the last block of a launch resets them.
 *  - a power of two in [256, 2^20]: one block runs consecutive output tiles
 *    of one channel group until it has covered about segment_len inputs.
 * block_threads is 128, 256 or 512. Outputs agree across settings to float32
 * rounding of the NCO phasor (a few ulp). */
#define FDSP_TILING_CLAIM 0u
#define FDSP_TILING_STRIDED 1u
int fdsp_channelizer_set_tiling(fdsp_channelizer_t h, uint32_t segment_len, uint32_t block_threads);
int fdsp_channelizer_get_tiling(fdsp_channelizer_t h, uint32_t *segment_len, uint32_t *block_threads);
/* Which kernel a batched channelizer launches (no reference counterpart:
 * the reference has one CUDA structure):
 *  - FDSP_KERNEL_SEGMENT_F32: the FDSP_CHAN_BATCHED default whenever every
 *    channel's window (taps per phase) is at most 64. One workgroup per input
 *    segment stages it once and computes every channel's outputs in it
 *    (groups of 2..8 channels with up 1, equal down and <= 32 taps on FP32
 *    matrix cores, the others on FP32 vector ALUs). Its tiling takes
 *    segment_len a multiple of 256 from 1024 to 4096 (256 threads; default
 *    2560); the autotuner times 1024, 2048, 2560, 3072, 3584 and 4096.
 *  - FDSP_KERNEL_GROUPS_F32: the channel-group kernel in FP32 (windows
 *    above 64 taps), with the work splits described above.
 *  - FDSP_KERNEL_GROUPS_BF16X2: FDSP_CHAN_BATCHED_BF16X2.
 *  - FDSP_KERNEL_PLAIN: FDSP_CHAN_PLAIN.
 * groups / mfma_groups (optional): the plan's work groups and how many of
 * them run on matrix cores. */
#define FDSP_KERNEL_SEGMENT_F32   0u
#define FDSP_KERNEL_GROUPS_F32    1u
#define FDSP_KERNEL_GROUPS_BF16X2 2u
#define FDSP_KERNEL_PLAIN         3u
int fdsp_channelizer_kernel_info(fdsp_channelizer_t h, uint32_t *kernel, uint32_t *groups, uint32_t *mfma_groups);
/* Switch an FDSP_CHAN_BATCHED handle between FDSP_KERNEL_SEGMENT_F32 (when
 * the channel set fits it) and FDSP_KERNEL_GROUPS_F32; both are FP32 and
 * carry the same stream state. No graph may exist. */
int fdsp_channelizer_set_kernel(fdsp_channelizer_t h, uint32_t kernel);
/* The batched kernel's work split for a chunk of chunk_len samples: blocks
 * launched, blocks the GPU holds at once, channel groups, LDS bytes per block
 * (diagnostics for the autotuner; any output may be null). */
int fdsp_channelizer_plan_info(fdsp_channelizer_t h, size_t chunk_len, uint32_t *blocks, uint32_t *capacity,
                               uint32_t *groups, uint32_t *lds_bytes);
/* Autotuner: first times each channel group alone (on all and on half of
 * the resident blocks) to weigh the groups' shares of the blocks by measured
 * work, then times the batched kernel on the device chunk d_in (chunk_len
 * samples, the caller's real input) under each candidate work split
 * (FDSP_TILING_STRIDED / FDSP_TILING_CLAIM at 128, 256 and 512 threads with
 * estimated group weights, at 256 and 512 with measured ones; segments of
 * 4096 ... 65536 inputs at 256 threads and of 16384 at 512), reps launches
 * each after a warm-up, and keeps the fastest; then it times that split with
 * fusion on (bit 1 of the trial flags) and keeps it if faster. The launches
 * write scratch outputs and leave the stream position and history as they
 * were. trials
 * (optional, cap entries, up to 17) receives every candidate's average kernel
 * time (us < 0: the split does not fit this channel set).
 * The pick is stored in a tuning cache file, keyed by the channel set (fs,
 * mode, every channel's description and taps), chunk_len, the input's 16-byte
 * alignment, the kernel and the device (architecture and CU count); a later
 * call with the same key, in any process, applies the stored pick without
 * running trials and reports it as the one trial, with flag bit 3 set. The
 * file is FDSP_TUNE_CACHE when that is set ("off": no cache), else
 * $XDG_CACHE_HOME/fdsp/tuning.txt or $HOME/.cache/fdsp/tuning.txt; one text
 * line per pick, the last line of a key wins (delete it to re-tune). */
typedef struct fdsp_tiling_trial {
    uint32_t segment_len, block_threads;
    float us;
    uint32_t flags; /* bit 0: groups weighed by measured work; bit 1: bf16-plane groups fused;
                       bit 2: segment kernel (segment_len is its segment); bit 3: from the tuning cache */
} fdsp_tiling_trial;
int fdsp_channelizer_autotune(fdsp_channelizer_t h, const fdsp_cf32 *d_in, size_t chunk_len, uint32_t reps,
                              fdsp_tiling_trial *trials, size_t cap, size_t *n_trials);
/* Phase grouping: with fusion on, the bf16-plane channel groups whose rows
 * tile 1920-sample spans evenly share one staged span per tile (the input is
 * read and converted once for all of them); it applies to 16-byte aligned
 * input. Off by default; the autotuner turns it on when it measures faster.
 * fusable_groups: groups in the fused record (0: nothing to fuse). */
int fdsp_channelizer_set_fusion(fdsp_channelizer_t h, int on);
int fdsp_channelizer_get_fusion(fdsp_channelizer_t h, int *on, uint32_t *fusable_groups);

/* ---------------------------------------------------------------------------
 * Serving: channelizer streams admitted as they arrive, their chunks batched
 * continuously into shared launches. No counterpart in the reference, whose
 * callers run one falcon_dsp_multi_rate_channelizer object per stream
 * (hdr/transform/falcon_dsp_multi_rate_channelizer.h:106-110). A stream is
 * one channel (fdsp_channel_desc, up to and including a discriminator tail)
 * of an input stream at cfg.fs; each chunk's results equal that channel of a
 * channelizer fed the same chunks. Its filter history and NCO state stay in
 * device memory between chunks.
 *
 * Requests are asynchronous: fdsp_serve_submit queues a chunk and returns a
 * ticket; a worker thread packs waiting chunks (one per stream per launch,
 * streams taken in turn, at most cfg.max_batch input and output samples)
 * into one launch while the previous launch runs. `in` and `out` must stay
 * valid and untouched until fdsp_serve_poll reports the ticket done. Chunks
 * of one stream are processed in submission order. All calls are
 * thread-safe. */
typedef struct fdsp_serve_s *fdsp_serve_t;
typedef struct fdsp_serve_config {
    uint32_t fs;          /* input sample rate of every stream (Hz) */
    uint32_t max_streams; /* live streams at most */
    uint32_t max_chunk;   /* samples per request at most */
    uint32_t max_batch;   /* input (and output) samples per launch at most; >= max_chunk */
    uint64_t hbm_budget;  /* device bytes for batch buffers plus the stream state pool (0: no limit) */
    uint64_t state_bytes; /* the stream state pool (0: what hbm_budget leaves after the batch buffers, at
                             most 1 GiB; 64 KiB per stream when hbm_budget is 0) */
} fdsp_serve_config;
typedef struct fdsp_serve_result {
    size_t n_out;       /* samples written: fdsp_cf32, or float angles for a discriminator tail */
    int status;         /* FDSP_OK, or the error the request ended with */
    uint64_t submit_ns; /* steady-clock time of submission */
    uint64_t done_ns;   /* steady-clock time the result was in `out` */
    uint64_t batch;     /* launch sequence number */
    uint64_t pack_ns;   /* steady-clock time the worker took it into a batch */
    uint64_t issue_ns;  /* ... the batch's copy-in was issued */
    uint64_t gpu_ns;    /* ... the worker saw the batch's kernels and copy-out done */
} fdsp_serve_result;
typedef struct fdsp_serve_stats {
    uint64_t hbm_used, hbm_budget; /* device bytes in use: batch buffers + live streams' state */
    uint32_t live_streams;
    size_t queued;                 /* requests waiting or running */
    uint64_t batches, chunks, in_samples, out_samples;
    uint64_t state_pool;           /* bytes of the stream state pool */
} fdsp_serve_stats;

/* Allocates the batch buffers (two staging slots) and the stream state pool
 * up front; FDSP_ENOMEM when they exceed the budget. Copies in run on one
 * HIP stream and kernels plus copies out on another, ordered by events. */
int fdsp_serve_create(const fdsp_serve_config *cfg, fdsp_serve_t *out);
/* Finishes every queued request, then frees everything. */
int fdsp_serve_destroy(fdsp_serve_t s);
/* Admit a stream: its state takes a slice of the pool (no device
 * allocation); FDSP_EAGAIN when max_streams are live or no slice of its size
 * is free (retry after a release). */
int fdsp_serve_admit(fdsp_serve_t s, const fdsp_channel_desc *chan, uint32_t *stream);
/* Output capacity a chunk of n samples may need (samples). */
int fdsp_serve_max_out(fdsp_serve_t s, uint32_t stream, size_t n, size_t *count);
/* Queue one chunk (1 <= n <= max_chunk, cap >= fdsp_serve_max_out);
 * FDSP_EAGAIN when too many requests are outstanding. */
int fdsp_serve_submit(fdsp_serve_t s, uint32_t stream, const fdsp_cf32 *in, size_t n, void *out, size_t cap,
                      uint64_t *ticket);
/* FDSP_EAGAIN while the request runs (after waiting up to wait_ms; < 0
 * waits for it). A finished ticket is consumed and returns its status. */
int fdsp_serve_poll(fdsp_serve_t s, uint64_t ticket, int wait_ms, fdsp_serve_result *res);
/* End a stream and free its state; FDSP_EBUSY while it has requests. */
int fdsp_serve_release(fdsp_serve_t s, uint32_t stream);
int fdsp_serve_info(fdsp_serve_t s, fdsp_serve_stats *stats);

#ifdef __cplusplus
}
#endif
#endif /* FDSP_H */
