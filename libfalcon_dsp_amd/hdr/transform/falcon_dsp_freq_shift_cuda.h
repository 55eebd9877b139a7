// falcon_dsp_freq_shift_cuda.h -- single and multi-shift NCO, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_freq_shift_cuda.h:142-212.)
#pragma once

#include "transform/falcon_dsp_freq_shift.h"

namespace falcon_dsp
{
    bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                         int32_t freq_shift_in_hz, std::vector<std::complex<float>>& out);
    // (appends to out, as the reference's GPU class does)
    bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<int16_t>>& in,
                         int32_t freq_shift_in_hz, std::vector<std::complex<int16_t>>& out);
    // one output vector per shift, all from one pass over the input
    bool freq_shift_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                         std::vector<int32_t>& freq_shift_in_hz, std::vector<std::vector<std::complex<float>>>& out);

    class falcon_dsp_freq_shift_cuda : public falcon_dsp_freq_shift
    {
    public:
        falcon_dsp_freq_shift_cuda(void) = default;
        ~falcon_dsp_freq_shift_cuda(void) = default;
        falcon_dsp_freq_shift_cuda(const falcon_dsp_freq_shift_cuda&) = delete;

        bool initialize(uint32_t input_sample_rate_in_sps, int32_t freq_shift_in_hz) override;
        bool initialize(uint32_t input_sample_rate_in_sps, std::vector<int32_t> freq_shift_in_hz);

        bool apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out) override;
        bool apply(std::vector<std::complex<float>>& in, std::vector<std::vector<std::complex<float>>>& out);
        // complex int16, single shift; the results are appended to out without
        // clearing it, as in the reference (falcon_dsp_freq_shift_cuda.cu:308-329)
        bool apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out) override;
    };
}
