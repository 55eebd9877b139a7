/*
 * fdsp_oracle.h -- CPU restatement of libfalcon_dsp's streaming-channelizer
 * algorithms. TEST INFRASTRUCTURE ONLY: the tests, the smoke check and the
 * CPU-baseline leg of bench.py call it as the checker; nothing in the
 * product (libfalcon_dsp_amd) links or calls it.
 *
 * Complex samples are interleaved float pairs {re, im} (std::complex<float>
 * layout). Each stateful object mirrors one reference class; the citations
 * name the reference file:line whose semantics it restates.
 *
 * Parity pinning (what checks each function):
 *   orc_gcd/lcm/rat_approx    - reference KATs (utils_unit_tests.cc:116-218)
 *   orc_freq_shift_*          - closed form x*exp(j*w*(n mod R)) in FP64, R
 *                               brute force, compiled reference (oracle/_ref)
 *   orc_resampler_*           - scipy.signal.upfirdn, reference KATs
 *                               (polyphase_resample_*_unit_tests:165-301),
 *                               chunk invariance, compiled reference
 *   orc_fir_*                 - scipy.signal.lfilter, reference KATs
 *                               (fir_filter_cpp_unit_tests.cc:147-306)
 *   orc_discrim_*             - np.angle(x[1:]*conj(x[:-1])), reference KAT
 *                               (polar_discriminator_cpp_unit_tests.cc:78-138)
 *   orc_channelizer_*         - composition of the above, compiled reference
 */
#ifndef FDSP_ORACLE_H
#define FDSP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- utilities: src/utilities/falcon_dsp_utils.cc ---- */
uint64_t orc_gcd(uint64_t a, uint64_t b);
uint64_t orc_lcm(uint64_t a, uint64_t b);
void orc_rat_approx(double f, int64_t md, int64_t *num, int64_t *denom);

/* ---- NCO frequency shift: src/transform/falcon_dsp_freq_shift.cc ---- */
int orc_freq_shift_params(uint32_t fs, int32_t f_hz, uint32_t *rollover, double *omega);

typedef struct orc_freq_shift {
    int initialized;
    double samples_handled;     /* a double in the reference (freq_shift.h:118) */
    uint32_t rollover;
    double omega;
} orc_freq_shift;

int orc_freq_shift_init(orc_freq_shift *s, uint32_t fs, int32_t f_hz);
void orc_freq_shift_reset(orc_freq_shift *s);
/* returns number of outputs (== n), 0 when not initialised or n == 0 */
size_t orc_freq_shift_apply(orc_freq_shift *s, const float *in, size_t n, float *out);

/* ---- polyphase resampler: src/transform/falcon_dsp_polyphase_resampler.cc ---- */
typedef struct orc_resampler {
    int initialized;
    uint32_t up, down;
    uint32_t padded_count, cpp;
    float *bank;           /* up*cpp complex (interleaved), transposed + flipped */
    float *state;          /* cpp-1 complex */
    uint32_t coeff_phase;
    uint32_t x_offset;
} orc_resampler;

int orc_resampler_init(orc_resampler *r, uint32_t up, uint32_t down,
                       const float *taps, size_t n_taps);
void orc_resampler_free(orc_resampler *r);
void orc_resampler_reset(orc_resampler *r);
uint32_t orc_resampler_needed_out_count(const orc_resampler *r, uint32_t in_count);
/* out must hold needed_out_count(n) samples; returns the count written */
size_t orc_resampler_apply(orc_resampler *r, const float *in, size_t n, float *out);

/* ---- FIR filter: src/transform/falcon_dsp_fir_filter.cc ---- */
typedef struct orc_fir {
    float *coeffs;       /* M complex */
    size_t m;
    float *state;        /* up to M complex, oldest first */
    size_t state_len;
} orc_fir;

int orc_fir_init(orc_fir *f, const float *coeffs, size_t m);
void orc_fir_free(orc_fir *f);
void orc_fir_reset(orc_fir *f);
/* out holds n samples; returns n (0 for empty input or no taps) */
size_t orc_fir_apply(orc_fir *f, const float *in, size_t n, float *out);

/* ---- polar discriminator: src/transform/falcon_dsp_polar_discriminator.cc ---- */
typedef struct orc_discrim {
    int has_state;
    float last[2];
} orc_discrim;

void orc_discrim_reset(orc_discrim *d);
/* out holds n floats; returns n-1 on the first call after reset, n otherwise */
size_t orc_discrim_apply(orc_discrim *d, const float *in, size_t n, float *out);

/* IIR filter (src/transform/falcon_dsp_iir_filter.cc), complex coefficients */
typedef struct orc_iir {
    float *b, *a;   /* nb and na complex coefficients, normalised by a[0] */
    size_t nb, na;
    float *xs;      /* input state, oldest first, up to nb samples */
    size_t xs_len;
    float *ys;      /* previous outputs, oldest first, up to na-1 */
    size_t ys_len;
} orc_iir;
int orc_iir_init(orc_iir *f, const float *b, size_t nb, const float *a, size_t na);
void orc_iir_free(orc_iir *f);
void orc_iir_reset(orc_iir *f);
size_t orc_iir_apply(orc_iir *f, const float *in, size_t n, float *out);
size_t orc_iir_apply_f32(orc_iir *f, const float *in, size_t n, float *out);
size_t orc_iir_apply_ci16(orc_iir *f, const int16_t *in, size_t n, int16_t *out);
size_t orc_iir_apply_i16(orc_iir *f, const int16_t *in, size_t n, int16_t *out);

/* FM mono demodulator (src/signals/falcon_dsp_fm_demod.cc), resampler taps
 * given by the caller */
typedef struct orc_fm {
    orc_freq_shift shift;
    orc_resampler iso, mono;
    orc_discrim disc;
    orc_iir deemph;
} orc_fm;
void orc_fm_deemphasis(double tau, float *b0, float *a1);
int orc_fm_init(orc_fm *f, uint32_t fs, int32_t offset_hz, double tau, uint32_t iso_up, uint32_t iso_down,
                const float *iso_taps, size_t n_iso, uint32_t mono_up, uint32_t mono_down, const float *mono_taps,
                size_t n_mono);
void orc_fm_free(orc_fm *f);
void orc_fm_reset(orc_fm *f);
size_t orc_fm_out_cap(const orc_fm *f, size_t n);
size_t orc_fm_demod(orc_fm *f, const float *in, size_t n, int16_t *out, float *audio);
size_t orc_sizeof_iir(void);
size_t orc_sizeof_fm(void);

/* complex int16 overloads (interleaved re, im) */
int16_t orc_i16(float v);
size_t orc_freq_shift_apply_ci16(orc_freq_shift *s, const int16_t *in, size_t n, int16_t *out);
size_t orc_fir_apply_ci16(orc_fir *f, const int16_t *in, size_t n, int16_t *out);
size_t orc_discrim_apply_ci16(orc_discrim *d, const int16_t *in, size_t n, float *out);

/* ---- multi-rate channelizer: src/transform/falcon_dsp_multi_rate_channelizer.cc ---- */
typedef struct orc_channelizer {
    uint32_t n_chan;
    orc_freq_shift *shifters;
    orc_resampler *resamplers;
    float *scratch;      /* full-rate shifted samples of one channel */
    size_t scratch_len;
} orc_channelizer;

int orc_channelizer_init(orc_channelizer *c, uint32_t fs, uint32_t n_chan,
                         const int64_t *shift_hz, const uint32_t *up, const uint32_t *down,
                         const float *const *taps, const uint32_t *n_taps);
void orc_channelizer_free(orc_channelizer *c);
void orc_channelizer_reset(orc_channelizer *c);
uint32_t orc_channelizer_needed_out_count(const orc_channelizer *c, uint32_t chan, uint32_t in_count);
/* outs[c] must hold needed_out_count(c, n); counts[c] receives the count */
int orc_channelizer_apply(orc_channelizer *c, const float *in, size_t n,
                          float *const *outs, size_t *counts);

/* carried state of channel `chan`: NCO samples handled, resampler
 * coeff_phase and xOffset */
int orc_channelizer_state(const orc_channelizer *c, uint32_t chan, double *samples_handled,
                          uint32_t *coeff_phase, uint32_t *x_offset);

/* sizes of the opaque structs, so ctypes callers can allocate them */
size_t orc_sizeof_freq_shift(void);
size_t orc_sizeof_resampler(void);
size_t orc_sizeof_fir(void);
size_t orc_sizeof_discrim(void);
size_t orc_sizeof_channelizer(void);

#ifdef __cplusplus
}
#endif
#endif
