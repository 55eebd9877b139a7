// falcon_dsp_fir_filter.h -- streaming complex FIR, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_fir_filter.h:87-140.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_fir_s;

namespace falcon_dsp
{
    bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                    std::vector<std::complex<float>>& out);
    bool fir_filter(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<int16_t>>& in,
                    std::vector<std::complex<int16_t>>& out);

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
