// falcon_dsp_fir_filter.h -- streaming complex FIR, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_fir_filter.h:87-140;
// get_resample_fir_params as .h:95-108.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

#include "transform/falcon_dsp_predefined_fir_filter.h"

struct fdsp_fir_s;

namespace falcon_dsp
{
    bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                    std::vector<std::complex<float>>& out);
    bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<int16_t>>& in,
                    std::vector<std::complex<int16_t>>& out);

    // the predefined anti-alias filter for in -> out sps with the given tap
    // count and design; false (coeffs empty) when there is none
    bool get_resample_fir_params(uint32_t input_sample_rate, uint32_t output_sample_rate,
                                 filter_taps_e taps, filter_source_type_e source,
                                 uint32_t& up_rate, uint32_t& down_rate,
                                 std::vector<std::complex<float>> &coeffs);

    class falcon_dsp_fir_filter
    {
    public:
        explicit falcon_dsp_fir_filter(std::vector<std::complex<float>>& coeffs);
        virtual ~falcon_dsp_fir_filter(void);
        falcon_dsp_fir_filter(void) = delete;
        falcon_dsp_fir_filter(const falcon_dsp_fir_filter&) = delete;

        void reset_state(void);
        // y[n] = sum_k h[k] x[n-k] with the history carried between calls;
        // false for an empty input or taps
        virtual bool apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out);
        // complex int16: filtered as floats (sharing the history above), each
        // output part truncated toward zero back to int16 (saturated outside
        // its range); out is cleared first
        virtual bool apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out);

    protected:
        std::mutex  m_mutex;
        fdsp_fir_s *m_handle;
    };
}
