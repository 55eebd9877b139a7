// falcon_dsp_dropin.cc -- the reference's class API (hdr/transform,
// hdr/utilities) as thin host code over the C ABI in include/fdsp.h. Every
// call is synchronous, as the reference's are: inputs are copied to the GPU,
// the kernels run, outputs come back before the call returns. Errors surface
// as the reference's bool / count returns; the C ABI's fdsp_last_error()
// holds the reason.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>

#include "fdsp.h"
#include "transform/falcon_dsp_fir_filter_cuda.h"
#include "transform/falcon_dsp_freq_shift_cuda.h"
#include "transform/falcon_dsp_iir_filter.h"
#include "signals/falcon_dsp_fm_demod.h"
#include "signals/falcon_dsp_fm_demod_cuda.h"
#include "transform/falcon_dsp_multi_rate_channelizer_cuda.h"
#include "transform/falcon_dsp_polar_discriminator.h"
#include "transform/falcon_dsp_polyphase_resampler_cuda.h"
#include "transform/falcon_dsp_resample_cuda.h"
#include "utilities/falcon_dsp_utils.h"

namespace
{
// std::complex<float> is layout-compatible with float[2], as is fdsp_cf32
inline const fdsp_cf32 *cf(const std::complex<float> *p) { return reinterpret_cast<const fdsp_cf32 *>(p); }
inline fdsp_cf32 *cf(std::complex<float> *p) { return reinterpret_cast<fdsp_cf32 *>(p); }
// and std::complex<int16_t> with int16_t[2] and fdsp_ci16
inline const fdsp_ci16 *ci(const std::complex<int16_t> *p) { return reinterpret_cast<const fdsp_ci16 *>(p); }
inline fdsp_ci16 *ci(std::complex<int16_t> *p) { return reinterpret_cast<fdsp_ci16 *>(p); }

// little-endian encoding of the sample files, independent of the host's order
void put_le(std::ofstream &f, uint32_t bits, int bytes)
{
    char b[4];
    for (int i = 0; i < bytes; ++i)
        b[i] = (char)((bits >> (8 * i)) & 0xFFu);
    f.write(b, bytes);
}
uint32_t get_le(const unsigned char *b, int bytes)
{
    uint32_t v = 0;
    for (int i = 0; i < bytes; ++i)
        v |= (uint32_t)b[i] << (8 * i);
    return v;
}
uint32_t f32_bits(float v)
{
    uint32_t u;
    std::memcpy(&u, &v, 4);
    return u;
}
float bits_f32(uint32_t u)
{
    float v;
    std::memcpy(&v, &u, 4);
    return v;
}

// Write n items of `parts` values each: binary through enc(value) -> bits,
// ASCII as the values separated by spaces, one item per line
template <typename T, typename Enc>
bool write_items(const std::string &name, falcon_dsp::file_type_e type, const T *vals, size_t n, int parts, int bytes,
                 Enc enc)
{
    if (type == falcon_dsp::file_type_e::BINARY) {
        std::ofstream f(name, std::ios::out | std::ios::binary);
        if (!f.is_open())
            return false;
        for (size_t i = 0; i < n * parts && f.good(); ++i)
            put_le(f, enc(vals[i]), bytes);
        return f.good();
    }
    if (type == falcon_dsp::file_type_e::ASCII) {
        std::ofstream f(name, std::ios::out);
        if (!f.is_open())
            return false;
        for (size_t i = 0; i < n && f.good(); ++i) {
            for (int k = 0; k < parts; ++k)
                f << (k ? " " : "") << vals[i * parts + k];
            f << std::endl;
        }
        return f.good();
    }
    return false;
}

// Binary file of whole items of `bytes_per_item`, or false
bool read_binary(const std::string &name, size_t bytes_per_item, std::vector<unsigned char> &raw)
{
    std::ifstream f(name, std::ios::in | std::ios::binary);
    if (!f.is_open())
        return false;
    f.seekg(0, std::ios::end);
    const std::streamoff size = f.tellg();
    f.seekg(0, std::ios::beg);
    raw.resize((size_t)(size < 0 ? 0 : size) / bytes_per_item * bytes_per_item);
    if (!raw.empty())
        f.read(reinterpret_cast<char *>(raw.data()), (std::streamsize)raw.size());
    return f.good() || raw.empty();
}
}  // namespace

namespace falcon_dsp
{
// ---------------------------------------------------------------------------
// utilities (reference src/utilities/falcon_dsp_utils.cc:86-137,231-294)
// ---------------------------------------------------------------------------
uint64_t calculate_gcd(uint64_t a, uint64_t b) { return fdsp_gcd(a, b); }
uint64_t calculate_lcm(uint64_t a, uint64_t b) { return fdsp_lcm(a, b); }
void rat_approx(double f, int64_t md, int64_t &num, int64_t &denom) { fdsp_rat_approx(f, md, &num, &denom); }

uint32_t calculate_filter_delay_from_up_down_rates(uint32_t num_coeffs, uint32_t up_rate, uint32_t down_rate)
{
    const uint32_t divisor = up_rate > down_rate ? down_rate : up_rate;
    return divisor ? num_coeffs / divisor / 2 : 0;
}

uint32_t calculate_filter_delay_from_sample_rates(uint32_t num_coeffs, uint32_t in_sample_rate_in_sps,
                                                  uint32_t out_sample_rate_in_sps)
{
    int64_t up = 0, down = 0;
    fdsp_rat_approx((double)out_sample_rate_in_sps / (double)in_sample_rate_in_sps, 1024, &up, &down);
    return calculate_filter_delay_from_up_down_rates(num_coeffs, (uint32_t)up, (uint32_t)down);
}

// ---------------------------------------------------------------------------
// frequency shift (reference src/transform/falcon_dsp_freq_shift.cc:107-235,
// falcon_dsp_freq_shift_cuda.cu:275-488)
// ---------------------------------------------------------------------------
std::pair<uint32_t, double> falcon_dsp_freq_shift::get_freq_shift_params(uint32_t input_sample_rate_in_sps,
                                                                         int32_t freq_shift_in_hz)
{
    uint32_t r = 0;
    double w = 0.0;
    fdsp_freq_shift_params(input_sample_rate_in_sps, freq_shift_in_hz, &r, &w);
    return {r, w};
}

falcon_dsp_freq_shift::falcon_dsp_freq_shift(void) : m_handle(nullptr), m_num_shifts(0) {}

falcon_dsp_freq_shift::~falcon_dsp_freq_shift(void) { fdsp_freq_shift_destroy(m_handle); }

bool falcon_dsp_freq_shift::_create(uint32_t fs, const int32_t *shifts, uint32_t n)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle || n == 0)
        return false;
    if (fdsp_freq_shift_create(fs, shifts, n, &m_handle) != FDSP_OK) {
        m_handle = nullptr;
        return false;
    }
    m_num_shifts = n;
    return true;
}

bool falcon_dsp_freq_shift::initialize(uint32_t input_sample_rate_in_sps, int32_t freq_shift_in_hz)
{
    return _create(input_sample_rate_in_sps, &freq_shift_in_hz, 1);
}

void falcon_dsp_freq_shift::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_freq_shift_reset(m_handle);
}

bool falcon_dsp_freq_shift::apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || m_num_shifts != 1 || in.empty())
        return false;
    out.resize(in.size());
    fdsp_cf32 *o = cf(out.data());
    if (fdsp_freq_shift_apply(m_handle, cf(in.data()), in.size(), &o) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool falcon_dsp_freq_shift_cuda::initialize(uint32_t input_sample_rate_in_sps, int32_t freq_shift_in_hz)
{
    return _create(input_sample_rate_in_sps, &freq_shift_in_hz, 1);
}

bool falcon_dsp_freq_shift_cuda::initialize(uint32_t input_sample_rate_in_sps, std::vector<int32_t> freq_shift_in_hz)
{
    return _create(input_sample_rate_in_sps, freq_shift_in_hz.data(), (uint32_t)freq_shift_in_hz.size());
}

bool falcon_dsp_freq_shift_cuda::apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out)
{
    return falcon_dsp_freq_shift::apply(in, out);
}

bool falcon_dsp_freq_shift_cuda::apply(std::vector<std::complex<float>>& in,
                                       std::vector<std::vector<std::complex<float>>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.assign(m_num_shifts, std::vector<std::complex<float>>(in.size()));
    std::vector<fdsp_cf32 *> outs(m_num_shifts);
    for (uint32_t c = 0; c < m_num_shifts; ++c)
        outs[c] = cf(out[c].data());
    if (fdsp_freq_shift_apply(m_handle, cf(in.data()), in.size(), outs.data()) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool freq_shift(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in, int32_t freq_shift_in_hz,
                std::vector<std::complex<float>>& out)
{
    falcon_dsp_freq_shift s;
    return s.initialize(in_sample_rate_in_sps, freq_shift_in_hz) && s.apply(in, out);
}

// reference falcon_dsp_freq_shift.cc:133-162: out cleared, false when not
// initialised or when nothing was produced
bool falcon_dsp_freq_shift::apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || m_num_shifts != 1 || in.empty())
        return false;
    out.resize(in.size());
    fdsp_ci16 *o = ci(out.data());
    if (fdsp_freq_shift_apply_ci16(m_handle, ci(in.data()), in.size(), &o) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

// reference falcon_dsp_freq_shift_cuda.cu:308-329: appended to out, which is
// not cleared first
bool falcon_dsp_freq_shift_cuda::apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    std::vector<std::complex<int16_t>> got;
    const bool ok = falcon_dsp_freq_shift::apply(in, got);
    out.insert(out.end(), got.begin(), got.end());
    return ok;
}

bool freq_shift(uint32_t in_sample_rate_in_sps, std::vector<std::complex<int16_t>>& in, int32_t freq_shift_in_hz,
                std::vector<std::complex<int16_t>>& out)
{
    falcon_dsp_freq_shift s;
    return s.initialize(in_sample_rate_in_sps, freq_shift_in_hz) && s.apply(in, out);
}

bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<int16_t>>& in, int32_t freq_shift_in_hz,
                     std::vector<std::complex<int16_t>>& out)
{
    falcon_dsp_freq_shift_cuda s;
    return s.initialize(in_sample_rate_in_sps, freq_shift_in_hz) && s.apply(in, out);
}

bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in, int32_t freq_shift_in_hz,
                     std::vector<std::complex<float>>& out)
{
    return freq_shift(in_sample_rate_in_sps, in, freq_shift_in_hz, out);
}

bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                     std::vector<int32_t>& freq_shift_in_hz, std::vector<std::vector<std::complex<float>>>& out)
{
    falcon_dsp_freq_shift_cuda s;
    return s.initialize(in_sample_rate_in_sps, freq_shift_in_hz) && s.apply(in, out);
}

// ---------------------------------------------------------------------------
// polyphase resampler (reference src/transform/falcon_dsp_polyphase_resampler.cc:114-346)
// ---------------------------------------------------------------------------
falcon_dsp_polyphase_resampler::falcon_dsp_polyphase_resampler(void) : m_handle(nullptr) {}

falcon_dsp_polyphase_resampler::~falcon_dsp_polyphase_resampler(void) { fdsp_resampler_destroy(m_handle); }

bool falcon_dsp_polyphase_resampler::initialize(uint32_t up_rate, uint32_t down_rate,
                                                const std::vector<std::complex<float>>& filter_coeffs)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle || filter_coeffs.empty())
        return false;
    if (fdsp_resampler_create(up_rate, down_rate, cf(filter_coeffs.data()), (uint32_t)filter_coeffs.size(),
                              &m_handle) != FDSP_OK) {
        m_handle = nullptr;
        return false;
    }
    return true;
}

void falcon_dsp_polyphase_resampler::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_resampler_reset(m_handle);
}

uint32_t falcon_dsp_polyphase_resampler::needed_out_count(uint32_t in_count)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    uint32_t n = 0;
    if (m_handle)
        fdsp_resampler_needed_out_count(m_handle, in_count, &n);
    return n;
}

uint32_t falcon_dsp_polyphase_resampler::coeffs_per_phase(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    uint32_t cpp = 0;
    if (m_handle)
        fdsp_resampler_coeffs_per_phase(m_handle, &cpp);
    return cpp;
}

int32_t falcon_dsp_polyphase_resampler::apply(std::vector<std::complex<float>>& in,
                                              std::vector<std::complex<float>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return 0;
    uint32_t need = 0;
    fdsp_resampler_needed_out_count(m_handle, (uint32_t)in.size(), &need);
    out.resize(need);
    size_t got = 0;
    if (fdsp_resampler_apply(m_handle, cf(in.data()), in.size(), cf(out.data()), out.size(), &got) != FDSP_OK) {
        out.clear();
        return 0;
    }
    out.resize(got);
    return (int32_t)got;
}

int32_t falcon_dsp_polyphase_resampler::apply(std::vector<float>& in, std::vector<float>& out)
{
    // real input: the imaginary parts stay zero, the real parts come back
    std::vector<std::complex<float>> cin(in.begin(), in.end()), cout;
    const int32_t n = apply(cin, cout);
    out.clear();
    out.reserve(cout.size());
    for (const auto &v : cout)
        out.push_back(v.real());
    return n;
}

// ---------------------------------------------------------------------------
// FIR filter (reference src/transform/falcon_dsp_fir_filter.cc:136-284)
// ---------------------------------------------------------------------------
falcon_dsp_fir_filter::falcon_dsp_fir_filter(std::vector<std::complex<float>>& coeffs) : m_handle(nullptr)
{
    if (!coeffs.empty() && fdsp_fir_create(cf(coeffs.data()), (uint32_t)coeffs.size(), &m_handle) != FDSP_OK)
        m_handle = nullptr;
}

falcon_dsp_fir_filter::~falcon_dsp_fir_filter(void) { fdsp_fir_destroy(m_handle); }

void falcon_dsp_fir_filter::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_fir_reset(m_handle);
}

bool falcon_dsp_fir_filter::apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_fir_apply(m_handle, cf(in.data()), in.size(), cf(out.data())) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                std::vector<std::complex<float>>& out)
{
    falcon_dsp_fir_filter f(coeffs);
    return f.apply(in, out);
}

bool fir_filter_cuda(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                     std::vector<std::complex<float>>& out)
{
    return fir_filter(coeffs, in, out);
}

// reference falcon_dsp_fir_filter.cc:151-175: out cleared; the float
// overload's result
bool falcon_dsp_fir_filter::apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_fir_apply_ci16(m_handle, ci(in.data()), in.size(), ci(out.data())) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

// reference falcon_dsp_fir_filter_cuda.cu:233-254: appended to out
bool falcon_dsp_fir_filter_cuda::apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    std::vector<std::complex<int16_t>> got;
    const bool ok = falcon_dsp_fir_filter::apply(in, got);
    out.insert(out.end(), got.begin(), got.end());
    return ok;
}

bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<int16_t>>& in,
                std::vector<std::complex<int16_t>>& out)
{
    falcon_dsp_fir_filter f(coeffs);
    return f.apply(in, out);
}

bool fir_filter_cuda(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<int16_t>>& in,
                     std::vector<std::complex<int16_t>>& out)
{
    falcon_dsp_fir_filter_cuda f(coeffs);
    return f.apply(in, out);
}

// ---------------------------------------------------------------------------
// predefined resampling filters (reference src/transform/falcon_dsp_fir_filter.cc:111-130)
// ---------------------------------------------------------------------------
bool get_resample_fir_params(uint32_t input_sample_rate, uint32_t output_sample_rate, filter_taps_e taps,
                             filter_source_type_e source, uint32_t& up_rate, uint32_t& down_rate,
                             std::vector<std::complex<float>> &coeffs)
{
    coeffs.clear();
    uint32_t up = 0, down = 0;
    size_t n = 0;
    std::vector<std::complex<float>> c(512);
    if (fdsp_resample_fir_params(input_sample_rate, output_sample_rate, static_cast<int32_t>(taps),
                                 static_cast<uint32_t>(source), &up, &down, cf(c.data()), c.size(), &n) != FDSP_OK)
        return false;
    c.resize(n);
    up_rate = up;
    down_rate = down;
    coeffs = c;
    return true;
}

// ---------------------------------------------------------------------------
// FM mono demodulator (reference src/signals/falcon_dsp_fm_demod.cc:98-271)
// ---------------------------------------------------------------------------
falcon_dsp_fm_demodulator::falcon_dsp_fm_demodulator(void) : m_initialized(false), m_handle(nullptr) {}

falcon_dsp_fm_demodulator::~falcon_dsp_fm_demodulator(void) { fdsp_fm_demod_destroy(m_handle); }

bool falcon_dsp_fm_demodulator::initialize(uint32_t input_sample_rate_in_sps, int32_t signal_offset_from_dc_in_hz,
                                           float deemphasis_time_constant_in_usecs)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_initialized)
        return false;
    // the two supported time constants, as the exact values the de-emphasis
    // coefficients are defined from
    double tau;
    if (deemphasis_time_constant_in_usecs == AMERICAS_FM_DEEMPHASIS_TIME_CONSTANT)
        tau = 75e-6;
    else if (deemphasis_time_constant_in_usecs == EUROPE_ASIA_FM_DEEMPHASIS_TIME_CONSTANT)
        tau = 50e-6;
    else
        return false;
    uint32_t iso_up, iso_down, mono_up, mono_down;
    std::vector<std::complex<float>> iso, mono;
    if (!get_resample_fir_params(input_sample_rate_in_sps, FM_RADIO_SAMPLE_RATE_IN_SPS, filter_taps_e::FILTER_TAPS_64,
                                 filter_source_type_e::REMEZ, iso_up, iso_down, iso) ||
        !get_resample_fir_params(FM_RADIO_SAMPLE_RATE_IN_SPS, FM_RADIO_AUDIO_SAMPLE_RATE_IN_SPS,
                                 filter_taps_e::FILTER_TAPS_64, filter_source_type_e::REMEZ, mono_up, mono_down, mono))
        return false;
    if (fdsp_fm_demod_create(input_sample_rate_in_sps, signal_offset_from_dc_in_hz, tau, iso_up, iso_down,
                             cf(iso.data()), (uint32_t)iso.size(), mono_up, mono_down, cf(mono.data()),
                             (uint32_t)mono.size(), &m_handle) != FDSP_OK)
        return false;
    m_initialized = true;
    return true;
}

void falcon_dsp_fm_demodulator::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_fm_demod_reset(m_handle);
}

bool falcon_dsp_fm_demodulator::demod_mono(std::vector<std::complex<float>>& in, std::vector<int16_t>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    size_t most = 0, got = 0;
    if (fdsp_fm_demod_out_count(m_handle, in.size(), &most) != FDSP_OK)
        return false;
    out.resize(std::max<size_t>(most, 1));
    if (fdsp_fm_demod_apply(m_handle, cf(in.data()), in.size(), out.data(), nullptr, out.size(), &got) != FDSP_OK) {
        out.clear();
        return false;
    }
    out.resize(got);
    return got > 0;
}

// GPU-named FM demodulator (reference src/signals/falcon_dsp_fm_demod_cuda.cu:91-265):
// the same pipeline as the base class, which already runs on the GPU
falcon_dsp_fm_demodulator_cuda::falcon_dsp_fm_demodulator_cuda(void) : falcon_dsp_fm_demodulator() {}

bool falcon_dsp_fm_demodulator_cuda::initialize(uint32_t input_sample_rate_in_sps,
                                                int32_t signal_offset_from_dc_in_hz,
                                                float deemphasis_time_constant_in_usecs)
{
    return falcon_dsp_fm_demodulator::initialize(input_sample_rate_in_sps, signal_offset_from_dc_in_hz,
                                                 deemphasis_time_constant_in_usecs);
}

void falcon_dsp_fm_demodulator_cuda::reset_state(void) { falcon_dsp_fm_demodulator::reset_state(); }

bool falcon_dsp_fm_demodulator_cuda::demod_mono(std::vector<std::complex<float>>& in, std::vector<int16_t>& out)
{
    return falcon_dsp_fm_demodulator::demod_mono(in, out);
}

// ---------------------------------------------------------------------------
// IIR filter (reference src/transform/falcon_dsp_iir_filter.cc:104-396)
// ---------------------------------------------------------------------------
falcon_dsp_iir_filter::falcon_dsp_iir_filter(void) : m_initialized(false), m_handle(nullptr) {}

falcon_dsp_iir_filter::~falcon_dsp_iir_filter(void) { fdsp_iir_destroy(m_handle); }

bool falcon_dsp_iir_filter::initialize(std::vector<std::complex<float>> &b_coeffs,
                                       std::vector<std::complex<float>> &a_coeffs)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_initialized)
        return false;
    if (!b_coeffs.empty() && !a_coeffs.empty() &&
        fdsp_iir_create(cf(b_coeffs.data()), (uint32_t)b_coeffs.size(), cf(a_coeffs.data()),
                        (uint32_t)a_coeffs.size(), &m_handle) != FDSP_OK)
        return false;
    m_initialized = true;
    return true;
}

void falcon_dsp_iir_filter::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_iir_reset(m_handle);
}

bool falcon_dsp_iir_filter::apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_iir_apply(m_handle, cf(in.data()), in.size(), cf(out.data())) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool falcon_dsp_iir_filter::apply(std::vector<float>& in, std::vector<float>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_iir_apply_f32(m_handle, in.data(), in.size(), out.data()) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool falcon_dsp_iir_filter::apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_iir_apply_ci16(m_handle, ci(in.data()), in.size(), ci(out.data())) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

bool falcon_dsp_iir_filter::apply(std::vector<int16_t>& in, std::vector<int16_t>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    if (fdsp_iir_apply_i16(m_handle, in.data(), in.size(), out.data()) != FDSP_OK) {
        out.clear();
        return false;
    }
    return true;
}

// reference falcon_dsp_iir_filter.cc:85-100: a fresh filter over in
bool iir_filter(std::vector<std::complex<float>> &b_coeffs, std::vector<std::complex<float>> &a_coeffs,
                std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out)
{
    falcon_dsp_iir_filter f;
    return f.initialize(b_coeffs, a_coeffs) && f.apply(in, out);
}

bool iir_filter(std::vector<std::complex<float>> &b_coeffs, std::vector<std::complex<float>> &a_coeffs,
                std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out)
{
    falcon_dsp_iir_filter f;
    return f.initialize(b_coeffs, a_coeffs) && f.apply(in, out);
}

// ---------------------------------------------------------------------------
// polar discriminator (reference src/transform/falcon_dsp_polar_discriminator.cc:103-171)
// ---------------------------------------------------------------------------
falcon_dsp_polar_discriminator::falcon_dsp_polar_discriminator(void) : m_handle(nullptr)
{
    if (fdsp_discriminator_create(&m_handle) != FDSP_OK)
        m_handle = nullptr;
}

falcon_dsp_polar_discriminator::~falcon_dsp_polar_discriminator(void) { fdsp_discriminator_destroy(m_handle); }

void falcon_dsp_polar_discriminator::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_discriminator_reset(m_handle);
}

bool falcon_dsp_polar_discriminator::apply(const std::vector<std::complex<float>>& in, std::vector<float>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    size_t got = 0;
    if (fdsp_discriminator_apply(m_handle, cf(in.data()), in.size(), out.data(), &got) != FDSP_OK) {
        out.clear();
        return false;
    }
    out.resize(got);
    return true;
}

bool polar_discriminator(const std::vector<std::complex<float>>& in, std::vector<float>& out)
{
    falcon_dsp_polar_discriminator d;
    return d.apply(in, out);
}

// reference falcon_dsp_polar_discriminator.cc:117-132
bool falcon_dsp_polar_discriminator::apply(const std::vector<std::complex<int16_t>>& in, std::vector<float>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle || in.empty())
        return false;
    out.resize(in.size());
    size_t got = 0;
    if (fdsp_discriminator_apply_ci16(m_handle, ci(in.data()), in.size(), out.data(), &got) != FDSP_OK) {
        out.clear();
        return false;
    }
    out.resize(got);
    return true;
}

bool polar_discriminator(const std::vector<std::complex<int16_t>>& in, std::vector<float>& out)
{
    falcon_dsp_polar_discriminator d;
    return d.apply(in, out);
}

// ---------------------------------------------------------------------------
// sample files (reference src/utilities/falcon_dsp_utils.cc:301-811)
// ---------------------------------------------------------------------------
namespace
{
uint32_t enc16(int16_t v) { return (uint16_t)v; }
uint32_t encf(float v) { return f32_bits(v); }
}  // namespace

bool write_data_to_file(std::string file_name, file_type_e type, std::vector<int16_t>& data)
{
    return write_items(file_name, type, data.data(), data.size(), 1, 2, enc16);
}

bool write_data_to_file(std::string file_name, file_type_e type, std::vector<float>& data)
{
    return write_items(file_name, type, data.data(), data.size(), 1, 4, encf);
}

bool write_complex_data_to_file(std::string file_name, file_type_e type, std::vector<std::complex<int16_t>>& data)
{
    return write_items(file_name, type, reinterpret_cast<const int16_t *>(data.data()), data.size(), 2, 2, enc16);
}

bool write_complex_data_to_file(std::string file_name, file_type_e type, std::vector<std::complex<float>>& data)
{
    return write_items(file_name, type, reinterpret_cast<const float *>(data.data()), data.size(), 2, 4, encf);
}

bool read_data_from_file(std::string file_name, file_type_e type, std::vector<int16_t>& data)
{
    data.clear();
    if (type == file_type_e::BINARY) {
        std::vector<unsigned char> raw;
        if (!read_binary(file_name, 2, raw))
            return false;
        for (size_t i = 0; i < raw.size(); i += 2)
            data.push_back((int16_t)(uint16_t)get_le(&raw[i], 2));
        return true;
    }
    if (type == file_type_e::ASCII) {
        std::ifstream f(file_name, std::ios::in);
        if (!f.is_open())
            return false;
        int16_t v;
        while (f >> v)
            data.push_back(v);
        return true;
    }
    return false;
}

bool read_data_from_file(std::string file_name, file_type_e type, std::vector<float>& data)
{
    data.clear();
    if (type != file_type_e::ASCII)
        return false;  // binary floats: not supported, as in the reference (.cc:498-502)
    std::ifstream f(file_name, std::ios::in);
    if (!f.is_open())
        return false;
    float v;
    while (f >> v)
        data.push_back(v);
    return true;
}

bool read_complex_data_from_file(std::string file_name, file_type_e type, std::vector<std::complex<int16_t>>& data)
{
    data.clear();
    if (type == file_type_e::BINARY) {
        std::vector<unsigned char> raw;
        if (!read_binary(file_name, 4, raw))
            return false;
        for (size_t i = 0; i < raw.size(); i += 4)
            data.emplace_back((int16_t)(uint16_t)get_le(&raw[i], 2), (int16_t)(uint16_t)get_le(&raw[i + 2], 2));
        return true;
    }
    if (type == file_type_e::ASCII) {
        std::ifstream f(file_name, std::ios::in);
        if (!f.is_open())
            return false;
        int16_t re, im;
        while (f >> re >> im)
            data.emplace_back(re, im);
        return true;
    }
    return false;
}

bool read_complex_data_from_file(std::string file_name, file_type_e type, std::vector<std::complex<float>>& data)
{
    data.clear();
    if (type == file_type_e::BINARY) {
        std::vector<unsigned char> raw;
        if (!read_binary(file_name, 8, raw))
            return false;
        for (size_t i = 0; i < raw.size(); i += 8)
            data.emplace_back(bits_f32(get_le(&raw[i], 4)), bits_f32(get_le(&raw[i + 4], 4)));
        return true;
    }
    if (type == file_type_e::ASCII) {
        std::ifstream f(file_name, std::ios::in);
        if (!f.is_open())
            return false;
        float re, im;
        while (f >> re >> im)
            data.emplace_back(re, im);
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// multi-rate channelizer (reference src/transform/falcon_dsp_multi_rate_channelizer.cc:81-173,
// falcon_dsp_multi_rate_channelizer_cuda.cu:375-577)
// ---------------------------------------------------------------------------
falcon_dsp_multi_rate_channelizer::falcon_dsp_multi_rate_channelizer(void) : m_handle(nullptr), m_num_channels(0) {}

falcon_dsp_multi_rate_channelizer::~falcon_dsp_multi_rate_channelizer(void) { fdsp_channelizer_destroy(m_handle); }

bool falcon_dsp_multi_rate_channelizer::_initialize(uint32_t input_sample_rate,
                                                    const std::vector<multi_rate_channelizer_channel_s>& channels,
                                                    int mode)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle || channels.empty())
        return false;
    std::vector<fdsp_channel_desc> d(channels.size());
    for (size_t c = 0; c < channels.size(); ++c) {
        const auto &ch = channels[c];
        d[c] = fdsp_channel_desc{ch.output_sample_rate_in_sps, ch.freq_shift_in_hz, ch.up_rate, ch.down_rate,
                                 cf(ch.resample_filter_coeffs.data()), (uint32_t)ch.resample_filter_coeffs.size(),
                                 FDSP_TAIL_NONE};
    }
    if (fdsp_channelizer_create(input_sample_rate, d.data(), (uint32_t)d.size(), mode, &m_handle) != FDSP_OK) {
        m_handle = nullptr;
        return false;
    }
    m_num_channels = (uint32_t)channels.size();
    return true;
}

bool falcon_dsp_multi_rate_channelizer::initialize(uint32_t input_sample_rate,
                                                   std::vector<multi_rate_channelizer_channel_s> channels)
{
    return _initialize(input_sample_rate, channels, FDSP_CHAN_BATCHED);
}

bool falcon_dsp_multi_rate_channelizer_cuda::initialize(uint32_t input_sample_rate,
                                                        std::vector<multi_rate_channelizer_channel_s> channels)
{
    return _initialize(input_sample_rate, channels, FDSP_CHAN_PLAIN);
}

bool falcon_dsp_multi_rate_channelizer::apply(std::vector<std::complex<float>>& in,
                                              std::vector<std::vector<std::complex<float>>>& out)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    out.clear();
    if (!m_handle)
        return false;
    out.resize(m_num_channels);
    if (in.empty())
        return m_num_channels > 0;
    std::vector<void *> ptrs(m_num_channels);
    std::vector<size_t> caps(m_num_channels), counts(m_num_channels);
    for (uint32_t c = 0; c < m_num_channels; ++c) {
        uint32_t need = 0;
        fdsp_channelizer_needed_out_count(m_handle, c, (uint32_t)in.size(), &need);
        out[c].resize(need);
        ptrs[c] = out[c].data();
        caps[c] = need;
    }
    if (fdsp_channelizer_apply(m_handle, cf(in.data()), in.size(), ptrs.data(), caps.data(), counts.data()) !=
        FDSP_OK) {
        out.clear();
        return false;
    }
    for (uint32_t c = 0; c < m_num_channels; ++c)
        out[c].resize(counts[c]);
    return true;
}

void falcon_dsp_multi_rate_channelizer::reset_state(void)
{
    std::lock_guard<std::mutex> lock(m_mutex);
    if (m_handle)
        fdsp_channelizer_reset(m_handle);
}

// ---------------------------------------------------------------------------
// one-shot resampling (reference src/transform/falcon_dsp_resample.cc:126-148)
// ---------------------------------------------------------------------------
bool resample(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
              std::vector<std::complex<float>>& filter_coeffs, uint32_t out_sample_rate_in_sps,
              std::vector<std::complex<float>>& out)
{
    int64_t up = 0, down = 0;
    fdsp_rat_approx((double)out_sample_rate_in_sps / (double)in_sample_rate_in_sps, 1024, &up, &down);
    falcon_dsp_polyphase_resampler r;
    return r.initialize((uint32_t)up, (uint32_t)down, filter_coeffs) && r.apply(in, out) > 0;
}

bool upfirdn(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
             std::vector<std::complex<float>>& filter_coeffs, uint32_t out_sample_rate_in_sps,
             std::vector<std::complex<float>>& out)
{
    return resample(in_sample_rate_in_sps, in, filter_coeffs, out_sample_rate_in_sps, out);
}

bool upfirdn(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in, std::vector<float>& filter_coeffs,
             uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out)
{
    std::vector<std::complex<float>> c(filter_coeffs.begin(), filter_coeffs.end());
    return resample(in_sample_rate_in_sps, in, c, out_sample_rate_in_sps, out);
}

bool resample_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                   std::vector<std::complex<float>>& filter_coeffs, uint32_t out_sample_rate_in_sps,
                   std::vector<std::complex<float>>& out)
{
    return resample(in_sample_rate_in_sps, in, filter_coeffs, out_sample_rate_in_sps, out);
}

bool upfirdn_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                  std::vector<float>& filter_coeffs, uint32_t out_sample_rate_in_sps,
                  std::vector<std::complex<float>>& out)
{
    return upfirdn(in_sample_rate_in_sps, in, filter_coeffs, out_sample_rate_in_sps, out);
}
}  // namespace falcon_dsp
