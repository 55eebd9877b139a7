// falcon_dsp_freq_shift.h -- NCO frequency shift, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_freq_shift.h:85-121;
//  the work runs on the GPU through the C ABI in include/fdsp.h.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <utility>
#include <vector>

struct fdsp_freq_shift_s;

namespace falcon_dsp
{
    // y[n] = x[n] * e^{j w s}, s = samples handled modulo the rollover index
    bool freq_shift(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                    int32_t freq_shift_in_hz, std::vector<std::complex<float>>& out);
    bool freq_shift(uint32_t in_sample_rate_in_sps, std::vector<std::complex<int16_t>>& in,
                    int32_t freq_shift_in_hz, std::vector<std::complex<int16_t>>& out);

    class falcon_dsp_freq_shift
    {
    public:
        // (R, w): R = fs/|f| scaled by ten until >= 1e5, rounded; w = 2 pi f / fs.
        // f = 0 gives (1, 0), the identity shift.
        static std::pair<uint32_t, double> get_freq_shift_params(uint32_t input_sample_rate_in_sps,
                                                                 int32_t freq_shift_in_hz);

        falcon_dsp_freq_shift(void);
        virtual ~falcon_dsp_freq_shift(void);
        falcon_dsp_freq_shift(const falcon_dsp_freq_shift&) = delete;

        // false on a second call or bad parameters
        virtual bool initialize(uint32_t input_sample_rate_in_sps, int32_t freq_shift_in_hz);
        void reset_state(void);
        // false when not initialised or the input is empty
        virtual bool apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out);
        // complex int16: shifted as floats, each output part truncated toward
        // zero back to int16 (saturated outside its range); out is cleared first
        virtual bool apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out);

    protected:
        bool _create(uint32_t fs, const int32_t *shifts, uint32_t n);

        std::mutex         m_mutex;
        fdsp_freq_shift_s *m_handle;
        uint32_t           m_num_shifts;
    };
}
