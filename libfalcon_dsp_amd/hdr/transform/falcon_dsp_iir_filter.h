// falcon_dsp_iir_filter.h -- streaming IIR filter, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_iir_filter.h:75-130.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_iir_s;

namespace falcon_dsp
{
    bool iir_filter(std::vector<std::complex<float>> &b_coeffs, std::vector<std::complex<float>> &a_coeffs,
                    std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out);
    bool iir_filter(std::vector<std::complex<float>> &b_coeffs, std::vector<std::complex<float>> &a_coeffs,
                    std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out);

    class falcon_dsp_iir_filter
    {
    public:
        falcon_dsp_iir_filter(void);
        virtual ~falcon_dsp_iir_filter(void);
        falcon_dsp_iir_filter(const falcon_dsp_iir_filter&) = delete;

        // false when already initialised. Empty b or a are accepted and make
        // every apply return false, as in the reference. Beyond the reference:
        // false for a[0] == 0, more than 256 b or 17 a coefficients.
        virtual bool initialize(std::vector<std::complex<float>> &b_coeffs, std::vector<std::complex<float>> &a_coeffs);
        void reset_state(void);

        // y[n] = sum_k b[k] x[n-k] - sum_{j>=1} a[j] y[n-j], inputs and outputs
        // carried between calls (all four overloads share them); out is
        // cleared first; false when not initialised or the input is empty.
        // int16 outputs are truncated toward zero (saturated outside int16).
        virtual bool apply(std::vector<int16_t>& in, std::vector<int16_t>& out);
        virtual bool apply(std::vector<float>& in, std::vector<float>& out);
        virtual bool apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out);
        virtual bool apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out);

    protected:
        bool        m_initialized;
        std::mutex  m_mutex;
        fdsp_iir_s *m_handle;
    };
}
