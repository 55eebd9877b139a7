/*
 * fdsp_oracle.c -- CPU restatement of libfalcon_dsp's streaming-channelizer
 * algorithms (NCO frequency shift, polyphase resampler, FIR filter, polar
 * discriminator, multi-rate channelizer, up/down rate helpers).
 *
 * TEST INFRASTRUCTURE ONLY. This file is the checker that the HIP product is
 * compared against: only tests, the smoke check and bench.py's cpu_baseline
 * leg load it. The product library never links or calls it.
 *
 * Each function follows the reference source cited beside it (paths relative
 * to the reference repository root). Arithmetic mirrors the reference:
 * FP64 angles with libm cos/sin cast to float, std::complex<float> products
 * written out as (ac - bd, ad + bc), FP32 accumulation in the reference's
 * tap order. Build with -ffp-contract=off so no FMA contraction changes the
 * rounding relative to the plain IEEE statement.
 */
#include "fdsp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* complex<float> product, the finite-value branch of the C/C++ complex
 * multiply used by std::complex<float>::operator* */
static inline void cmul(float ar, float ai, float br, float bi, float *cr, float *ci)
{
    *cr = ar * br - ai * bi;
    *ci = ar * bi + ai * br;
}

/* (*xr + j *xi) / (c + j d) in place, as the reference's toolchain (g++ 11)
 * divides std::complex<float> values: Smith's method (R. L. Smith,
 * "Algorithm 116: Complex division", CACM 5(8), 1962), scaling by the larger
 * of |c| and |d|, evaluated in double and rounded to float once. (Checked
 * against g++'s std::complex<float> operator/= on 4e6 random operands.) The
 * non-finite recovery branch is not restated: the callers never divide by
 * zero. */
static void orc_cdiv(float *xr, float *xi, float c_, float d_)
{
    const double a = *xr, b = *xi, c = c_, d = d_;
    if (fabs(c) < fabs(d)) {
        const double r = c / d, den = c * r + d;
        *xr = (float)((a * r + b) / den);
        *xi = (float)((b * r - a) / den);
    } else {
        const double r = d / c, den = d * r + c;
        *xr = (float)((b * r + a) / den);
        *xi = (float)((b - a * r) / den);
    }
}

/* ------------------------------------------------------------------------ */
/* utilities -- src/utilities/falcon_dsp_utils.cc                            */
/* ------------------------------------------------------------------------ */

/* falcon_dsp_utils.cc:123-126 (Euclid, recursive in the reference) */
uint64_t orc_gcd(uint64_t a, uint64_t b)
{
    while (b != 0) {
        uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}

/* falcon_dsp_utils.cc:133-137 */
uint64_t orc_lcm(uint64_t a, uint64_t b)
{
    uint64_t g = orc_gcd(a, b);
    return (a / g) * b;
}

/* falcon_dsp_utils.cc:231-294: continued-fraction expansion of f with the
 * denominator bounded by md. The input is first scaled by powers of two until
 * it is an integer (d/n with n a power of two); convergents h/k are built
 * until the next one would exceed md, in which case a semiconvergent is
 * taken when it is at least half the next partial quotient. */
void orc_rat_approx(double f, int64_t md, int64_t *num, int64_t *denom)
{
    int64_t a, h[3] = {0, 1, 0}, k[3] = {1, 0, 0};
    int64_t x, d, n = 1;
    int i, neg = 0;

    if (md <= 1) {
        *denom = 1;
        *num = (int64_t)f;
        return;
    }
    if (f < 0) {
        neg = 1;
        f = -f;
    }
    while (f != floor(f)) {
        n <<= 1;
        f *= 2;
    }
    d = (int64_t)f;

    for (i = 0; i < 64; i++) {
        a = n ? d / n : 0;
        if (i && !a)
            break;
        x = d;
        d = n;
        n = x % n;
        x = a;
        if (k[1] * a + k[0] >= md) {
            x = (md - k[0]) / k[1];
            if (x * 2 >= a || k[1] >= md)
                i = 65;
            else
                break;
        }
        h[2] = x * h[1] + h[0]; h[0] = h[1]; h[1] = h[2];
        k[2] = x * k[1] + k[0]; k[0] = k[1]; k[1] = k[2];
    }
    *denom = k[1];
    *num = neg ? -h[1] : h[1];
}

/* ------------------------------------------------------------------------ */
/* NCO frequency shift -- src/transform/falcon_dsp_freq_shift.cc              */
/* ------------------------------------------------------------------------ */

/* falcon_dsp_freq_shift.cc:194-235. The rollover index is fs/|f| scaled by
 * ten until it reaches 1e5, then rounded; the angular step is 2*pi*f/fs.
 * f == 0 divides by zero in the reference (undefined); here it is defined as
 * the identity shift: rollover 1, omega 0. Returns 0 on success. */
int orc_freq_shift_params(uint32_t fs, int32_t f_hz, uint32_t *rollover, double *omega)
{
    if (fs == 0)
        return -1;
    if (f_hz == 0) {
        *rollover = 1;
        *omega = 0.0;
        return 0;
    }
    double r = (double)fs / (double)abs(f_hz);
    while (r < 1e5)
        r *= 10;
    *rollover = (uint32_t)round(r);
    *omega = ((double)f_hz / (double)fs) * 2.0 * M_PI;
    return 0;
}

/* falcon_dsp_freq_shift.cc:107-122 */
int orc_freq_shift_init(orc_freq_shift *s, uint32_t fs, int32_t f_hz)
{
    if (orc_freq_shift_params(fs, f_hz, &s->rollover, &s->omega) != 0)
        return -1;
    s->samples_handled = 0.0;
    s->initialized = 1;
    return 0;
}

/* falcon_dsp_freq_shift.cc:124-131 */
void orc_freq_shift_reset(orc_freq_shift *s)
{
    s->samples_handled = 0.0;
}

/* falcon_dsp_freq_shift.cc:164-192: y[n] = x[n] * cf32(cos(w*s), sin(w*s)),
 * s advanced by one per sample and reset to zero when it reaches R. */
size_t orc_freq_shift_apply(orc_freq_shift *s, const float *in, size_t n, float *out)
{
    if (!s->initialized)
        return 0;
    for (size_t i = 0; i < n; ++i) {
        double angle = s->omega * s->samples_handled;
        float c = (float)cos(angle);
        float sn = (float)sin(angle);
        cmul(in[2 * i], in[2 * i + 1], c, sn, &out[2 * i], &out[2 * i + 1]);
        s->samples_handled += 1.0;
        if (s->samples_handled >= (double)s->rollover)
            s->samples_handled = 0.0;
    }
    return n;
}

/* ------------------------------------------------------------------------ */
/* polyphase resampler -- src/transform/falcon_dsp_polyphase_resampler.cc    */
/* ------------------------------------------------------------------------ */

/* polyphase_resampler.cc:114-162: pad the tap count up to a multiple of up,
 * cpp = padded/up taps per phase, bank[p*cpp + (cpp-1-j)] = h[j*up + p]
 * (transposed and flipped), state = cpp-1 zeros. */
int orc_resampler_init(orc_resampler *r, uint32_t up, uint32_t down,
                       const float *taps, size_t n_taps)
{
    memset(r, 0, sizeof(*r));
    if (up == 0 || down == 0 || n_taps == 0)
        return -1;
    r->up = up;
    r->down = down;
    uint32_t padded = (uint32_t)n_taps;
    while (padded % up)
        padded++;
    r->padded_count = padded;
    r->cpp = padded / up;
    r->bank = (float *)calloc(2 * (size_t)padded, sizeof(float));
    r->state = (float *)calloc(2 * (size_t)(r->cpp > 1 ? r->cpp - 1 : 1), sizeof(float));
    if (!r->bank || !r->state)
        return -1;
    for (uint32_t p = 0; p < up; ++p) {
        for (uint32_t j = 0; j < r->cpp; ++j) {
            size_t src = (size_t)j * up + p;
            if (src < n_taps) {
                size_t dst = (size_t)(r->cpp - 1 - j) + (size_t)p * r->cpp;
                r->bank[2 * dst] = taps[2 * src];
                r->bank[2 * dst + 1] = taps[2 * src + 1];
            }
        }
    }
    r->coeff_phase = 0;
    r->x_offset = 0;
    r->initialized = 1;
    return 0;
}

void orc_resampler_free(orc_resampler *r)
{
    free(r->bank);
    free(r->state);
    r->bank = NULL;
    r->state = NULL;
    r->initialized = 0;
}

/* polyphase_resampler.cc:164-172 */
void orc_resampler_reset(orc_resampler *r)
{
    if (r->cpp > 1)
        memset(r->state, 0, 2 * sizeof(float) * (size_t)(r->cpp - 1));
    r->coeff_phase = 0;
    r->x_offset = 0;
}

/* polyphase_resampler.cc:216-228 */
uint32_t orc_resampler_needed_out_count(const orc_resampler *r, uint32_t in_count)
{
    if (!r->initialized)
        return 0;
    uint64_t np = (uint64_t)in_count * (uint64_t)r->up;
    uint32_t need = (uint32_t)(np / r->down);
    if ((uint32_t)(r->coeff_phase + r->up * r->x_offset) < (np % r->down))
        need++;
    return need;
}

/* polyphase_resampler.cc:321-346: the state becomes the last cpp-1 samples
 * of [state | in]. */
static void resampler_manage_state(orc_resampler *r, const float *in, size_t n)
{
    size_t slen = r->cpp - 1;
    if (slen == 0)
        return;
    if (n < slen) {
        size_t retain = slen - n;
        memmove(r->state, r->state + 2 * (slen - retain), 2 * sizeof(float) * retain);
        memcpy(r->state + 2 * retain, in, 2 * sizeof(float) * n);
    } else {
        memcpy(r->state, in + 2 * (n - slen), 2 * sizeof(float) * slen);
    }
}

/* polyphase_resampler.cc:256-319. For each output while x_idx < n:
 * acc = sum_j Re(bank[phase][j]) * X[x_idx - cpp + 1 + j], X = [state | in],
 * only the real part of each tap is used (:288, :298); then
 * phase += down, x_idx += phase / up, phase %= up; finally
 * xOffset = x_idx - n and the state is updated. */
size_t orc_resampler_apply(orc_resampler *r, const float *in, size_t n, float *out)
{
    if (!r->initialized)
        return 0;
    const int64_t cpp = r->cpp;
    size_t produced = 0;
    int64_t x_idx = r->x_offset;
    while ((uint64_t)x_idx < n) {
        float acc_r = 0.0f, acc_i = 0.0f;
        const float *coeff = r->bank + 2 * (size_t)r->coeff_phase * (size_t)cpp;
        int64_t back = x_idx - cpp + 1;
        int64_t offset = -back;
        if (offset > 0) {
            const float *st = r->state + 2 * ((cpp - 1) - offset);
            for (int64_t t = 0; t < offset; ++t) {
                acc_r += st[2 * t] * coeff[0];
                acc_i += st[2 * t + 1] * coeff[0];
                coeff += 2;
            }
            back += offset;
        }
        while (back <= x_idx) {
            acc_r += in[2 * back] * coeff[0];
            acc_i += in[2 * back + 1] * coeff[0];
            back++;
            coeff += 2;
        }
        out[2 * produced] = acc_r;
        out[2 * produced + 1] = acc_i;
        produced++;
        r->coeff_phase += r->down;
        x_idx += r->coeff_phase / r->up;
        r->coeff_phase %= r->up;
    }
    r->x_offset = (uint32_t)(x_idx - (int64_t)n);
    resampler_manage_state(r, in, n);
    return produced;
}

/* ------------------------------------------------------------------------ */
/* FIR filter -- src/transform/falcon_dsp_fir_filter.cc                       */
/* ------------------------------------------------------------------------ */

/* fir_filter.cc:136-140 */
int orc_fir_init(orc_fir *f, const float *coeffs, size_t m)
{
    memset(f, 0, sizeof(*f));
    f->m = m;
    f->coeffs = (float *)malloc(2 * sizeof(float) * (m ? m : 1));
    f->state = (float *)malloc(2 * sizeof(float) * (m ? m : 1));
    if (!f->coeffs || !f->state)
        return -1;
    if (m)
        memcpy(f->coeffs, coeffs, 2 * sizeof(float) * m);
    f->state_len = 0;
    return 0;
}

void orc_fir_free(orc_fir *f)
{
    free(f->coeffs);
    free(f->state);
    f->coeffs = NULL;
    f->state = NULL;
}

/* fir_filter.cc:142-149 */
void orc_fir_reset(orc_fir *f)
{
    f->state_len = 0;
}

/* fir_filter.cc:245-284: keep the newest M inputs, oldest first */
static void fir_update_state(orc_fir *f, const float *in, size_t n)
{
    size_t m = f->m;
    if (n >= m) {
        memcpy(f->state, in + 2 * (n - m), 2 * sizeof(float) * m);
        f->state_len = m;
    } else {
        size_t total = f->state_len + n;
        if (total > m) {
            size_t drop = total - m;
            memmove(f->state, f->state + 2 * drop, 2 * sizeof(float) * (f->state_len - drop));
            f->state_len -= drop;
        }
        memcpy(f->state + 2 * f->state_len, in, 2 * sizeof(float) * n);
        f->state_len += n;
    }
}

/* fir_filter.cc:177-243: y[n] = sum_k h[k] * x[n-k] (complex x complex),
 * accumulated for k = 0..M-1; samples before the input come from the state
 * (newest first), missing history contributes nothing. */
size_t orc_fir_apply(orc_fir *f, const float *in, size_t n, float *out)
{
    if (f->m == 0 || n == 0)
        return 0;
    const size_t m = f->m;
    for (size_t i = 0; i < n; ++i) {
        float ar = 0.0f, ai = 0.0f, pr, pi;
        size_t k = 0;
        if (i < m) {
            for (int64_t d = (int64_t)i; d >= 0; --d, ++k) {
                cmul(in[2 * d], in[2 * d + 1], f->coeffs[2 * k], f->coeffs[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
            for (int64_t s = (int64_t)f->state_len - 1; k < m && s >= 0; --s, ++k) {
                cmul(f->state[2 * s], f->state[2 * s + 1], f->coeffs[2 * k], f->coeffs[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
        } else {
            for (; k < m; ++k) {
                size_t d = i - k;
                cmul(in[2 * d], in[2 * d + 1], f->coeffs[2 * k], f->coeffs[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
        }
        out[2 * i] = ar;
        out[2 * i + 1] = ai;
    }
    fir_update_state(f, in, n);
    return n;
}

/* ------------------------------------------------------------------------ */
/* polar discriminator -- src/transform/falcon_dsp_polar_discriminator.cc    */
/* ------------------------------------------------------------------------ */

/* polar_discriminator.cc:108-115 */
void orc_discrim_reset(orc_discrim *d)
{
    d->has_state = 0;
}

/* polar_discriminator.cc:134-171: out = atan2(Im z, Re z), z = x[n]*conj(x[n-1]);
 * the first sample pairs with the saved last sample when one exists. The
 * reference's unqualified atan2 on float parts is the double overload, its
 * result rounded to float when stored. */
size_t orc_discrim_apply(orc_discrim *d, const float *in, size_t n, float *out)
{
    if (n == 0)
        return 0;
    size_t k = 0;
    float zr, zi;
    if (d->has_state) {
        cmul(in[0], in[1], d->last[0], -d->last[1], &zr, &zi);
        out[k++] = (float)atan2((double)zi, (double)zr);
    }
    for (size_t i = 1; i < n; ++i) {
        cmul(in[2 * i], in[2 * i + 1], in[2 * i - 2], -in[2 * i - 1], &zr, &zi);
        out[k++] = (float)atan2((double)zi, (double)zr);
    }
    d->last[0] = in[2 * (n - 1)];
    d->last[1] = in[2 * (n - 1) + 1];
    d->has_state = 1;
    return k;
}

/* ------------------------------------------------------------------------ */
/* multi-rate channelizer -- src/transform/falcon_dsp_multi_rate_channelizer.cc */
/* ------------------------------------------------------------------------ */

/* multi_rate_channelizer.cc:81-93,109-144: one shifter + resampler per
 * channel; the shift in Hz is narrowed to int32 as the reference's
 * initialize(uint32, int32) call does. */
int orc_channelizer_init(orc_channelizer *c, uint32_t fs, uint32_t n_chan,
                         const int64_t *shift_hz, const uint32_t *up, const uint32_t *down,
                         const float *const *taps, const uint32_t *n_taps)
{
    memset(c, 0, sizeof(*c));
    if (fs == 0 || n_chan == 0)
        return -1;
    c->n_chan = n_chan;
    c->shifters = (orc_freq_shift *)calloc(n_chan, sizeof(orc_freq_shift));
    c->resamplers = (orc_resampler *)calloc(n_chan, sizeof(orc_resampler));
    if (!c->shifters || !c->resamplers)
        return -1;
    for (uint32_t i = 0; i < n_chan; ++i) {
        if (orc_freq_shift_init(&c->shifters[i], fs, (int32_t)shift_hz[i]) != 0)
            return -1;
        if (orc_resampler_init(&c->resamplers[i], up[i], down[i], taps[i], n_taps[i]) != 0)
            return -1;
    }
    return 0;
}

void orc_channelizer_free(orc_channelizer *c)
{
    for (uint32_t i = 0; c->resamplers && i < c->n_chan; ++i)
        orc_resampler_free(&c->resamplers[i]);
    free(c->shifters);
    free(c->resamplers);
    free(c->scratch);
    memset(c, 0, sizeof(*c));
}

/* multi_rate_channelizer.cc:166-173 */
void orc_channelizer_reset(orc_channelizer *c)
{
    for (uint32_t i = 0; i < c->n_chan; ++i) {
        orc_freq_shift_reset(&c->shifters[i]);
        orc_resampler_reset(&c->resamplers[i]);
    }
}

uint32_t orc_channelizer_needed_out_count(const orc_channelizer *c, uint32_t chan, uint32_t in_count)
{
    if (chan >= c->n_chan)
        return 0;
    return orc_resampler_needed_out_count(&c->resamplers[chan], in_count);
}

/* multi_rate_channelizer.cc:146-164: per channel, shift the whole input then
 * resample the shifted copy. */
int orc_channelizer_apply(orc_channelizer *c, const float *in, size_t n,
                          float *const *outs, size_t *counts)
{
    if (n > c->scratch_len) {
        free(c->scratch);
        c->scratch = (float *)malloc(2 * sizeof(float) * n);
        if (!c->scratch)
            return -1;
        c->scratch_len = n;
    }
    for (uint32_t i = 0; i < c->n_chan; ++i) {
        orc_freq_shift_apply(&c->shifters[i], in, n, c->scratch);
        counts[i] = orc_resampler_apply(&c->resamplers[i], c->scratch, n, outs[i]);
    }
    return 0;
}

int orc_channelizer_state(const orc_channelizer *c, uint32_t chan, double *samples_handled,
                          uint32_t *coeff_phase, uint32_t *x_offset)
{
    if (chan >= c->n_chan)
        return -1;
    *samples_handled = c->shifters[chan].samples_handled;
    *coeff_phase = c->resamplers[chan].coeff_phase;
    *x_offset = c->resamplers[chan].x_offset;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* complex int16 overloads                                                   */
/* ------------------------------------------------------------------------ */

/* The reference builds its std::complex<int16_t> outputs from float parts, an
 * implicit float -> int16_t conversion: truncation toward zero (e.g.
 * fir_filter.cc:168-171). Outside the int16 range that conversion is
 * undefined; this restatement saturates there (NaN reads as -32768), the
 * choice the GPU build makes too. */
int16_t orc_i16(float v)
{
    if (v != v || v <= -32768.0f)
        return -32768;
    if (v >= 32767.0f)
        return 32767;
    return (int16_t)(int32_t)v;
}

/* freq_shift.cc:133-162: each sample converted to float and shifted exactly
 * as in the float overload (one angle, one counter step per sample), the
 * product converted back part by part */
size_t orc_freq_shift_apply_ci16(orc_freq_shift *s, const int16_t *in, size_t n, int16_t *out)
{
    if (!s->initialized)
        return 0;
    for (size_t i = 0; i < n; ++i) {
        const float x[2] = {(float)in[2 * i], (float)in[2 * i + 1]};
        float y[2];
        orc_freq_shift_apply(s, x, 1, y);
        out[2 * i] = orc_i16(y[0]);
        out[2 * i + 1] = orc_i16(y[1]);
    }
    return n;
}

/* fir_filter.cc:151-175: the input converted to float, filtered by the float
 * overload (whose history it joins), the outputs converted back */
size_t orc_fir_apply_ci16(orc_fir *f, const int16_t *in, size_t n, int16_t *out)
{
    if (f->m == 0 || n == 0)
        return 0;
    float *x = (float *)malloc(2 * n * sizeof(float));
    float *y = (float *)malloc(2 * n * sizeof(float));
    size_t k = 0;
    if (x && y) {
        for (size_t i = 0; i < 2 * n; ++i)
            x[i] = (float)in[i];
        k = orc_fir_apply(f, x, n, y);
        for (size_t i = 0; i < 2 * k; ++i)
            out[i] = orc_i16(y[i]);
    }
    free(x);
    free(y);
    return k;
}

/* polar_discriminator.cc:117-132: the input converted to float, then the float
 * overload */
size_t orc_discrim_apply_ci16(orc_discrim *d, const int16_t *in, size_t n, float *out)
{
    if (n == 0)
        return 0;
    float *x = (float *)malloc(2 * n * sizeof(float));
    size_t k = 0;
    if (x) {
        for (size_t i = 0; i < 2 * n; ++i)
            x[i] = (float)in[i];
        k = orc_discrim_apply(d, x, n, out);
    }
    free(x);
    return k;
}

/* ------------------------------------------------------------------------ */
/* IIR filter -- src/transform/falcon_dsp_iir_filter.cc                      */
/* ------------------------------------------------------------------------ */

/* iir_filter.cc:119-155: coefficients copied; when Re a[0] is not 1, b and
 * a[1..] are divided by a[0] (complex division, textbook form) */
int orc_iir_init(orc_iir *f, const float *b, size_t nb, const float *a, size_t na)
{
    memset(f, 0, sizeof(*f));
    f->b = (float *)malloc(2 * (nb ? nb : 1) * sizeof(float));
    f->a = (float *)malloc(2 * (na ? na : 1) * sizeof(float));
    f->xs = (float *)malloc(2 * (nb ? nb : 1) * sizeof(float));
    f->ys = (float *)malloc(2 * (na ? na : 1) * sizeof(float));
    if (!f->b || !f->a || !f->xs || !f->ys)
        return -1;
    memcpy(f->b, b, 2 * nb * sizeof(float));
    memcpy(f->a, a, 2 * na * sizeof(float));
    f->nb = nb;
    f->na = na;
    /* iir_filter.cc:137-150: when a[0]'s real part is not 1, every b and
     * a[1:] are divided by a[0] with std::complex<float> division */
    if (na >= 1 && a[0] != 1.0f) {
        for (size_t i = 0; i < nb; ++i)
            orc_cdiv(&f->b[2 * i], &f->b[2 * i + 1], a[0], a[1]);
        for (size_t i = 1; i < na; ++i)
            orc_cdiv(&f->a[2 * i], &f->a[2 * i + 1], a[0], a[1]);
    }
    return 0;
}

void orc_iir_free(orc_iir *f)
{
    free(f->b);
    free(f->a);
    free(f->xs);
    free(f->ys);
    memset(f, 0, sizeof(*f));
}

/* iir_filter.cc:108-116 */
void orc_iir_reset(orc_iir *f)
{
    f->xs_len = 0;
    f->ys_len = 0;
}

/* iir_filter.cc:235-396: y = sum_k b[k] x[n-k] (the newest inputs from the
 * chunk, older ones from the input state, newest first) minus
 * sum_j a[j] y[n-j] over the kept outputs (newest first), each product
 * formed, then added or subtracted, in that order. The state keeps the last
 * nb inputs and the last na-1 outputs. */
size_t orc_iir_apply(orc_iir *f, const float *in, size_t n, float *out)
{
    if (f->nb == 0 || f->na == 0 || n == 0)
        return 0;
    float pr, pi;
    for (size_t i = 0; i < n; ++i) {
        float ar = 0.0f, ai = 0.0f;
        size_t k = 0;
        if (i < f->nb) {
            for (int64_t d = (int64_t)i; d >= 0; --d, ++k) {
                cmul(in[2 * d], in[2 * d + 1], f->b[2 * k], f->b[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
            for (int64_t s = (int64_t)f->xs_len - 1; k < f->nb && s >= 0; --s, ++k) {
                cmul(f->xs[2 * s], f->xs[2 * s + 1], f->b[2 * k], f->b[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
        } else {
            for (; k < f->nb; ++k) {
                cmul(in[2 * (i - k)], in[2 * (i - k) + 1], f->b[2 * k], f->b[2 * k + 1], &pr, &pi);
                ar += pr;
                ai += pi;
            }
        }
        size_t j = 1;
        for (int64_t s = (int64_t)f->ys_len - 1; j < f->na && s >= 0; --s, ++j) {
            cmul(f->ys[2 * s], f->ys[2 * s + 1], f->a[2 * j], f->a[2 * j + 1], &pr, &pi);
            ar -= pr;
            ai -= pi;
        }
        out[2 * i] = ar;
        out[2 * i + 1] = ai;
        if (f->na > 1) {
            if (f->ys_len == f->na - 1) {
                memmove(f->ys, f->ys + 2, 2 * (f->ys_len - 1) * sizeof(float));
                f->ys_len--;
            }
            f->ys[2 * f->ys_len] = ar;
            f->ys[2 * f->ys_len + 1] = ai;
            f->ys_len++;
        }
    }
    if (n >= f->nb) {
        memcpy(f->xs, in + 2 * (n - f->nb), 2 * f->nb * sizeof(float));
        f->xs_len = f->nb;
    } else {
        /* append, keep the newest nb */
        for (size_t i = 0; i < n; ++i) {
            if (f->xs_len == f->nb) {
                memmove(f->xs, f->xs + 2, 2 * (f->xs_len - 1) * sizeof(float));
                f->xs_len--;
            }
            f->xs[2 * f->xs_len] = in[2 * i];
            f->xs[2 * f->xs_len + 1] = in[2 * i + 1];
            f->xs_len++;
        }
    }
    return n;
}

/* iir_filter.cc:183-203: real input as (x, 0), real parts out */
size_t orc_iir_apply_f32(orc_iir *f, const float *in, size_t n, float *out)
{
    float *x = (float *)calloc(2 * (n ? n : 1), sizeof(float));
    float *y = (float *)malloc(2 * (n ? n : 1) * sizeof(float));
    size_t k = 0;
    if (x && y) {
        for (size_t i = 0; i < n; ++i)
            x[2 * i] = in[i];
        k = orc_iir_apply(f, x, n, y);
        for (size_t i = 0; i < k; ++i)
            out[i] = y[2 * i];
    }
    free(x);
    free(y);
    return k;
}

/* iir_filter.cc:157-233: int16 samples converted to complex<float> (real ones
 * with a zero imaginary part), filtered by the complex overload (whose state
 * they share), the outputs converted back part by part (the real part only
 * for real samples) */
size_t orc_iir_apply_ci16(orc_iir *f, const int16_t *in, size_t n, int16_t *out)
{
    float *x = (float *)malloc(2 * (n ? n : 1) * sizeof(float));
    float *y = (float *)malloc(2 * (n ? n : 1) * sizeof(float));
    size_t k = 0;
    if (x && y) {
        for (size_t i = 0; i < 2 * n; ++i)
            x[i] = (float)in[i];
        k = orc_iir_apply(f, x, n, y);
        for (size_t i = 0; i < 2 * k; ++i)
            out[i] = orc_i16(y[i]);
    }
    free(x);
    free(y);
    return k;
}

size_t orc_iir_apply_i16(orc_iir *f, const int16_t *in, size_t n, int16_t *out)
{
    float *x = (float *)calloc(2 * (n ? n : 1), sizeof(float));
    float *y = (float *)malloc(2 * (n ? n : 1) * sizeof(float));
    size_t k = 0;
    if (x && y) {
        for (size_t i = 0; i < n; ++i)
            x[2 * i] = (float)in[i];
        k = orc_iir_apply(f, x, n, y);
        for (size_t i = 0; i < k; ++i)
            out[i] = orc_i16(y[2 * i]);
    }
    free(x);
    free(y);
    return k;
}

/* ------------------------------------------------------------------------ */
/* FM mono demodulator -- src/signals/falcon_dsp_fm_demod.cc                 */
/* ------------------------------------------------------------------------ */

/* fm_demod.cc:146-165: single-pole de-emphasis at the 200 kHz design rate,
 * x = exp(-1 / (200000 tau)), b = [1 - x], a = [1, -x] (the reference keeps
 * the two values for 75 and 50 us as float literals of these) */
void orc_fm_deemphasis(double tau, float *b0, float *a1)
{
    const double x = exp(-1.0 / (200000.0 * tau));
    *b0 = (float)(1.0 - x);
    *a1 = (float)(-x);
}

/* fm_demod.cc:98-198, with the two resamplers' taps passed in (the reference
 * reads them from its predefined table: REMEZ, 64 taps) */
int orc_fm_init(orc_fm *f, uint32_t fs, int32_t offset_hz, double tau, uint32_t iso_up, uint32_t iso_down,
                const float *iso_taps, size_t n_iso, uint32_t mono_up, uint32_t mono_down, const float *mono_taps,
                size_t n_mono)
{
    memset(f, 0, sizeof(*f));
    if (orc_freq_shift_init(&f->shift, fs, offset_hz) != 0)
        return -1;
    if (orc_resampler_init(&f->iso, iso_up, iso_down, iso_taps, n_iso) != 0)
        return -1;
    if (orc_resampler_init(&f->mono, mono_up, mono_down, mono_taps, n_mono) != 0)
        return -1;
    orc_discrim_reset(&f->disc);
    float b0, a1;
    orc_fm_deemphasis(tau, &b0, &a1);
    const float b[2] = {b0, 0.0f}, a[4] = {1.0f, 0.0f, a1, 0.0f};
    return orc_iir_init(&f->deemph, b, 1, a, 2);
}

void orc_fm_free(orc_fm *f)
{
    orc_resampler_free(&f->iso);
    orc_resampler_free(&f->mono);
    orc_iir_free(&f->deemph);
}

/* fm_demod.cc:198-208 */
void orc_fm_reset(orc_fm *f)
{
    orc_freq_shift_reset(&f->shift);
    orc_resampler_reset(&f->iso);
    orc_discrim_reset(&f->disc);
    orc_iir_reset(&f->deemph);
    orc_resampler_reset(&f->mono);
}

/* at most this many outputs for an n-sample chunk */
size_t orc_fm_out_cap(const orc_fm *f, size_t n)
{
    const size_t a = n * f->iso.up / f->iso.down + 2;
    return a * f->mono.up / f->mono.down + 2;
}

/* fm_demod.cc:210-271: shift, isolate (resample), discriminate, de-emphasise,
 * decimate to audio (real samples through the complex resampler), then scale
 * the chunk so its largest |sample| is 10000: s = (float)(10000.0 / max), each
 * output the float product converted to int16 (truncation; saturated outside
 * int16, and all zero when the chunk is all zero, where the reference's
 * conversion is undefined). audio, when given, receives the samples before
 * scaling. */
size_t orc_fm_demod(orc_fm *f, const float *in, size_t n, int16_t *out, float *audio)
{
    if (n == 0)
        return 0;
    float *shifted = (float *)malloc(2 * n * sizeof(float));
    const size_t n1cap = orc_resampler_needed_out_count(&f->iso, (uint32_t)n) + 1;
    float *iso = (float *)malloc(2 * n1cap * sizeof(float));
    float *disc = (float *)malloc(n1cap * sizeof(float));
    float *emph = (float *)malloc(2 * n1cap * sizeof(float));
    float *emph_c = (float *)calloc(2 * n1cap, sizeof(float));
    const size_t n3cap = orc_resampler_needed_out_count(&f->mono, (uint32_t)n1cap) + 1;
    float *mono = (float *)malloc(2 * n3cap * sizeof(float));
    size_t n_out = 0;
    if (shifted && iso && disc && emph && emph_c && mono) {
        orc_freq_shift_apply(&f->shift, in, n, shifted);
        const size_t n1 = orc_resampler_apply(&f->iso, shifted, n, iso);
        const size_t n2 = n1 ? orc_discrim_apply(&f->disc, iso, n1, disc) : 0;
        const size_t n3 = n2 ? orc_iir_apply_f32(&f->deemph, disc, n2, emph) : 0;
        for (size_t i = 0; i < n3; ++i)
            emph_c[2 * i] = emph[i];
        const size_t n4 = n3 ? orc_resampler_apply(&f->mono, emph_c, n3, mono) : 0;
        float amax = 0.0f;
        for (size_t i = 0; i < n4; ++i)
            amax = fabsf(mono[2 * i]) > amax ? fabsf(mono[2 * i]) : amax;
        const float scale = amax > 0.0f ? (float)(10000.0 / (double)amax) : 0.0f;
        for (size_t i = 0; i < n4; ++i) {
            out[i] = orc_i16(mono[2 * i] * scale);
            if (audio)
                audio[i] = mono[2 * i];
        }
        n_out = n4;
    }
    free(shifted);
    free(iso);
    free(disc);
    free(emph);
    free(emph_c);
    free(mono);
    return n_out;
}

size_t orc_sizeof_iir(void) { return sizeof(orc_iir); }
size_t orc_sizeof_fm(void) { return sizeof(orc_fm); }

size_t orc_sizeof_freq_shift(void) { return sizeof(orc_freq_shift); }
size_t orc_sizeof_resampler(void) { return sizeof(orc_resampler); }
size_t orc_sizeof_fir(void) { return sizeof(orc_fir); }
size_t orc_sizeof_discrim(void) { return sizeof(orc_discrim); }
size_t orc_sizeof_channelizer(void) { return sizeof(orc_channelizer); }
