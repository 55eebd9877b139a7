/*
 * ref_shim.cc -- extern "C" wrappers around the reference's own CPU classes,
 * compiled from the reference sources where they lie (see oracle/Makefile).
 * TEST INFRASTRUCTURE ONLY: used to validate the oracle restatement and as
 * the CPU baseline in bench.py. Never linked into the product.
 *
 * Wrapped classes (reference headers):
 *   falcon_dsp::falcon_dsp_freq_shift            hdr/transform/falcon_dsp_freq_shift.h:98-121
 *   falcon_dsp::falcon_dsp_polyphase_resampler   hdr/transform/falcon_dsp_polyphase_resampler.h:168-195
 *   falcon_dsp::falcon_dsp_fir_filter            hdr/transform/falcon_dsp_fir_filter.h:119-140
 *   falcon_dsp::falcon_dsp_polar_discriminator   hdr/transform/falcon_dsp_polar_discriminator.h:97-114
 *   falcon_dsp::falcon_dsp_multi_rate_channelizer hdr/transform/falcon_dsp_multi_rate_channelizer.h:97-134
 *   calculate_gcd / rat_approx                   hdr/utilities/falcon_dsp_utils.h
 */
#include <algorithm>
#include <complex>
#include <string>
#include <cstdint>
#include <cstring>
#include <vector>

#include "signals/falcon_dsp_fm_demod.h"
#include "transform/falcon_dsp_fir_filter.h"
#include "transform/falcon_dsp_iir_filter.h"
#include "transform/falcon_dsp_predefined_fir_filter.h"
#include "transform/falcon_dsp_freq_shift.h"
#include "transform/falcon_dsp_multi_rate_channelizer.h"
#include "transform/falcon_dsp_polar_discriminator.h"
#include "transform/falcon_dsp_polyphase_resampler.h"
#include "utilities/falcon_dsp_utils.h"

using cf32 = std::complex<float>;

static std::vector<cf32> to_vec(const float *p, size_t n)
{
    std::vector<cf32> v(n);
    if (n)
        std::memcpy(v.data(), p, n * sizeof(cf32));
    return v;
}

extern "C" {

uint64_t ref_gcd(uint64_t a, uint64_t b) { return falcon_dsp::calculate_gcd(a, b); }
uint64_t ref_lcm(uint64_t a, uint64_t b) { return falcon_dsp::calculate_lcm(a, b); }
void ref_rat_approx(double f, int64_t md, int64_t *num, int64_t *den)
{
    falcon_dsp::rat_approx(f, md, *num, *den);
}

void ref_freq_shift_params(uint32_t fs, int32_t f, uint32_t *r, double *w)
{
    auto p = falcon_dsp::falcon_dsp_freq_shift::get_freq_shift_params(fs, f);
    *r = p.first;
    *w = p.second;
}

/* ---- freq shift ---- */
void *ref_fs_create(uint32_t fs, int32_t f)
{
    auto *o = new falcon_dsp::falcon_dsp_freq_shift();
    if (!o->initialize(fs, f)) {
        delete o;
        return nullptr;
    }
    return o;
}
size_t ref_fs_apply(void *h, const float *in, size_t n, float *out)
{
    auto v = to_vec(in, n);
    std::vector<cf32> o;
    static_cast<falcon_dsp::falcon_dsp_freq_shift *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(cf32));
    return o.size();
}
void ref_fs_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_freq_shift *>(h)->reset_state(); }
void ref_fs_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_freq_shift *>(h); }

/* ---- polyphase resampler ---- */
void *ref_rs_create(uint32_t up, uint32_t down, const float *taps, size_t n_taps)
{
    auto *o = new falcon_dsp::falcon_dsp_polyphase_resampler();
    if (!o->initialize(up, down, to_vec(taps, n_taps))) {
        delete o;
        return nullptr;
    }
    return o;
}
uint32_t ref_rs_needed_out_count(void *h, uint32_t n)
{
    return static_cast<falcon_dsp::falcon_dsp_polyphase_resampler *>(h)->needed_out_count(n);
}
uint32_t ref_rs_coeffs_per_phase(void *h)
{
    return static_cast<falcon_dsp::falcon_dsp_polyphase_resampler *>(h)->coeffs_per_phase();
}
/* returns the output count, or -1 when cap is too small */
int64_t ref_rs_apply(void *h, const float *in, size_t n, float *out, size_t cap)
{
    auto v = to_vec(in, n);
    std::vector<cf32> o;
    static_cast<falcon_dsp::falcon_dsp_polyphase_resampler *>(h)->apply(v, o);
    if (o.size() > cap)
        return -1;
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(cf32));
    return (int64_t)o.size();
}
void ref_rs_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_polyphase_resampler *>(h)->reset_state(); }
void ref_rs_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_polyphase_resampler *>(h); }

/* ---- FIR ---- */
void *ref_fir_create(const float *coeffs, size_t m)
{
    auto c = to_vec(coeffs, m);
    return new falcon_dsp::falcon_dsp_fir_filter(c);
}
size_t ref_fir_apply(void *h, const float *in, size_t n, float *out)
{
    auto v = to_vec(in, n);
    std::vector<cf32> o;
    static_cast<falcon_dsp::falcon_dsp_fir_filter *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(cf32));
    return o.size();
}
void ref_fir_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_fir_filter *>(h)->reset_state(); }
void ref_fir_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_fir_filter *>(h); }

/* ---- polar discriminator ---- */
void *ref_disc_create(void) { return new falcon_dsp::falcon_dsp_polar_discriminator(); }
size_t ref_disc_apply(void *h, const float *in, size_t n, float *out)
{
    auto v = to_vec(in, n);
    std::vector<float> o;
    static_cast<falcon_dsp::falcon_dsp_polar_discriminator *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    return o.size();
}
void ref_disc_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_polar_discriminator *>(h)->reset_state(); }
void ref_disc_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_polar_discriminator *>(h); }

/* ---- complex int16 overloads of the classes above ---- */
typedef std::complex<int16_t> ci16;
static std::vector<ci16> to_vec16(const int16_t *p, size_t n)
{
    std::vector<ci16> v(n);
    for (size_t i = 0; i < n; ++i)
        v[i] = ci16(p[2 * i], p[2 * i + 1]);
    return v;
}
size_t ref_fs_apply_ci16(void *h, const int16_t *in, size_t n, int16_t *out)
{
    auto v = to_vec16(in, n);
    std::vector<ci16> o;
    static_cast<falcon_dsp::falcon_dsp_freq_shift *>(h)->apply(v, o);
    for (size_t i = 0; i < o.size(); ++i) {
        out[2 * i] = o[i].real();
        out[2 * i + 1] = o[i].imag();
    }
    return o.size();
}
size_t ref_fir_apply_ci16(void *h, const int16_t *in, size_t n, int16_t *out)
{
    auto v = to_vec16(in, n);
    std::vector<ci16> o;
    static_cast<falcon_dsp::falcon_dsp_fir_filter *>(h)->apply(v, o);
    for (size_t i = 0; i < o.size(); ++i) {
        out[2 * i] = o[i].real();
        out[2 * i + 1] = o[i].imag();
    }
    return o.size();
}
size_t ref_disc_apply_ci16(void *h, const int16_t *in, size_t n, float *out)
{
    auto v = to_vec16(in, n);
    std::vector<float> o;
    static_cast<falcon_dsp::falcon_dsp_polar_discriminator *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    return o.size();
}

/* ---- sample files (falcon_dsp_utils.cc:301-811) ----
 * kind: 0 ci16, 1 cf32, 2 int16, 3 float; ascii selects file_type_e::ASCII.
 * data holds n items of interleaved values (2 per complex item). */
static falcon_dsp::file_type_e ftype(int ascii)
{
    return ascii ? falcon_dsp::file_type_e::ASCII : falcon_dsp::file_type_e::BINARY;
}
int ref_write_file(int kind, int ascii, const char *path, const void *data, size_t n)
{
    const std::string name(path);
    switch (kind) {
    case 0: {
        auto v = to_vec16(static_cast<const int16_t *>(data), n);
        return falcon_dsp::write_complex_data_to_file(name, ftype(ascii), v) ? 0 : -1;
    }
    case 1: {
        auto v = to_vec(static_cast<const float *>(data), n);
        return falcon_dsp::write_complex_data_to_file(name, ftype(ascii), v) ? 0 : -1;
    }
    case 2: {
        std::vector<int16_t> v(static_cast<const int16_t *>(data), static_cast<const int16_t *>(data) + n);
        return falcon_dsp::write_data_to_file(name, ftype(ascii), v) ? 0 : -1;
    }
    default: {
        std::vector<float> v(static_cast<const float *>(data), static_cast<const float *>(data) + n);
        return falcon_dsp::write_data_to_file(name, ftype(ascii), v) ? 0 : -1;
    }
    }
}
/* items read (at most cap are copied out), or -1 when the reader returns false */
int64_t ref_read_file(int kind, int ascii, const char *path, void *out, size_t cap)
{
    const std::string name(path);
    switch (kind) {
    case 0: {
        std::vector<ci16> v;
        if (!falcon_dsp::read_complex_data_from_file(name, ftype(ascii), v))
            return -1;
        for (size_t i = 0; i < v.size() && i < cap; ++i) {
            static_cast<int16_t *>(out)[2 * i] = v[i].real();
            static_cast<int16_t *>(out)[2 * i + 1] = v[i].imag();
        }
        return (int64_t)v.size();
    }
    case 1: {
        std::vector<cf32> v;
        if (!falcon_dsp::read_complex_data_from_file(name, ftype(ascii), v))
            return -1;
        std::memcpy(out, v.data(), std::min(v.size(), cap) * sizeof(cf32));
        return (int64_t)v.size();
    }
    case 2: {
        std::vector<int16_t> v;
        if (!falcon_dsp::read_data_from_file(name, ftype(ascii), v))
            return -1;
        std::memcpy(out, v.data(), std::min(v.size(), cap) * sizeof(int16_t));
        return (int64_t)v.size();
    }
    default: {
        std::vector<float> v;
        if (!falcon_dsp::read_data_from_file(name, ftype(ascii), v))
            return -1;
        std::memcpy(out, v.data(), std::min(v.size(), cap) * sizeof(float));
        return (int64_t)v.size();
    }
    }
}

/* ---- IIR filter (complex coefficients), float overload ---- */
void *ref_iir_create(const float *b, size_t nb, const float *a, size_t na)
{
    auto bv = to_vec(b, nb), av = to_vec(a, na);
    auto *o = new falcon_dsp::falcon_dsp_iir_filter();
    if (!o->initialize(bv, av)) {
        delete o;
        return nullptr;
    }
    return o;
}
size_t ref_iir_apply_f32(void *h, const float *in, size_t n, float *out)
{
    std::vector<float> v(in, in + n), o;
    static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(float));
    return o.size();
}
size_t ref_iir_apply(void *h, const float *in, size_t n, float *out)
{
    auto v = to_vec(in, n);
    std::vector<cf32> o;
    static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(cf32));
    return o.size();
}
size_t ref_iir_apply_ci16(void *h, const int16_t *in, size_t n, int16_t *out)
{
    auto v = to_vec16(in, n);
    std::vector<ci16> o;
    static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h)->apply(v, o);
    for (size_t i = 0; i < o.size(); ++i) {
        out[2 * i] = o[i].real();
        out[2 * i + 1] = o[i].imag();
    }
    return o.size();
}
size_t ref_iir_apply_i16(void *h, const int16_t *in, size_t n, int16_t *out)
{
    std::vector<int16_t> v(in, in + n), o;
    static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h)->apply(v, o);
    if (!o.empty())
        std::memcpy(out, o.data(), o.size() * sizeof(int16_t));
    return o.size();
}
void ref_iir_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h)->reset_state(); }
void ref_iir_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_iir_filter *>(h); }

/* ---- predefined resample filters (64-tap REMEZ) and the FM demodulator ---- */
/* taps copied out when cap allows; returns the tap count, or -1 */
int64_t ref_resample_fir_params(uint32_t in_rate, uint32_t out_rate, uint32_t *up, uint32_t *down, float *taps,
                                size_t cap)
{
    std::vector<cf32> c;
    if (!falcon_dsp::get_resample_fir_params(in_rate, out_rate, filter_taps_e::FILTER_TAPS_64,
                                             filter_source_type_e::REMEZ, *up, *down, c))
        return -1;
    if (c.size() <= cap)
        std::memcpy(taps, c.data(), c.size() * sizeof(cf32));
    return (int64_t)c.size();
}
/* any table entry: taps is the filter_taps_e value, source 0 REMEZ / 1 FIRLS */
int64_t ref_resample_fir_params_key(uint32_t in_rate, uint32_t out_rate, int32_t taps, uint32_t source, uint32_t *up,
                                    uint32_t *down, float *coeffs, size_t cap)
{
    std::vector<cf32> c;
    if (!falcon_dsp::get_resample_fir_params(in_rate, out_rate, static_cast<filter_taps_e>(taps),
                                             static_cast<filter_source_type_e>(source), *up, *down, c))
        return -1;
    if (c.size() <= cap)
        std::memcpy(coeffs, c.data(), c.size() * sizeof(cf32));
    return (int64_t)c.size();
}
void *ref_fm_create(uint32_t fs, int32_t offset_hz, int europe)
{
    auto *o = new falcon_dsp::falcon_dsp_fm_demodulator();
    const float tc = europe ? falcon_dsp::falcon_dsp_fm_demodulator::EUROPE_ASIA_FM_DEEMPHASIS_TIME_CONSTANT
                            : falcon_dsp::falcon_dsp_fm_demodulator::AMERICAS_FM_DEEMPHASIS_TIME_CONSTANT;
    if (!o->initialize(fs, offset_hz, tc)) {
        delete o;
        return nullptr;
    }
    return o;
}
/* outputs (at most cap copied), or -1 when demod_mono returns false */
int64_t ref_fm_demod(void *h, const float *in, size_t n, int16_t *out, size_t cap)
{
    auto v = to_vec(in, n);
    std::vector<int16_t> o;
    if (!static_cast<falcon_dsp::falcon_dsp_fm_demodulator *>(h)->demod_mono(v, o))
        return -1;
    std::memcpy(out, o.data(), std::min(o.size(), cap) * sizeof(int16_t));
    return (int64_t)o.size();
}
void ref_fm_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_fm_demodulator *>(h)->reset_state(); }
void ref_fm_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_fm_demodulator *>(h); }

/* ---- multi-rate channelizer ---- */
void *ref_chan_create(uint32_t fs, uint32_t n_chan, const int64_t *shift_hz, const uint32_t *out_rate,
                      const uint32_t *up, const uint32_t *down, const float *const *taps,
                      const uint32_t *n_taps)
{
    std::vector<multi_rate_channelizer_channel_s> chans;
    for (uint32_t i = 0; i < n_chan; ++i) {
        multi_rate_channelizer_channel_s c;
        c.output_sample_rate_in_sps = out_rate[i];
        c.freq_shift_in_hz = shift_hz[i];
        c.up_rate = up[i];
        c.down_rate = down[i];
        c.resample_filter_coeffs = to_vec(taps[i], n_taps[i]);
        chans.push_back(c);
    }
    auto *o = new falcon_dsp::falcon_dsp_multi_rate_channelizer();
    if (!o->initialize(fs, chans)) {
        delete o;
        return nullptr;
    }
    return o;
}
/* counts[i] receives each channel's output count; -1 when a cap is too small */
int ref_chan_apply(void *h, const float *in, size_t n, float *const *outs, const size_t *caps,
                   size_t *counts)
{
    auto v = to_vec(in, n);
    std::vector<std::vector<cf32>> o;
    static_cast<falcon_dsp::falcon_dsp_multi_rate_channelizer *>(h)->apply(v, o);
    for (size_t i = 0; i < o.size(); ++i) {
        if (o[i].size() > caps[i])
            return -1;
        if (!o[i].empty())
            std::memcpy(outs[i], o[i].data(), o[i].size() * sizeof(cf32));
        counts[i] = o[i].size();
    }
    return 0;
}
void ref_chan_reset(void *h) { static_cast<falcon_dsp::falcon_dsp_multi_rate_channelizer *>(h)->reset_state(); }
void ref_chan_destroy(void *h) { delete static_cast<falcon_dsp::falcon_dsp_multi_rate_channelizer *>(h); }

} /* extern "C" */
