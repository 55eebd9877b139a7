// falcon_dsp_fir_filter_cuda.h -- drop-in API of the GPU FIR.
// (Same class and signatures as reference hdr/transform/falcon_dsp_fir_filter_cuda.h:80-140.)
#pragma once

#include "transform/falcon_dsp_fir_filter.h"

namespace falcon_dsp
{
    bool fir_filter_cuda(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                         std::vector<std::complex<float>>& out);

    class falcon_dsp_fir_filter_cuda : public falcon_dsp_fir_filter
    {
    public:
        explicit falcon_dsp_fir_filter_cuda(std::vector<std::complex<float>>& coeffs)
            : falcon_dsp_fir_filter(coeffs) {}
        ~falcon_dsp_fir_filter_cuda(void) = default;
        falcon_dsp_fir_filter_cuda(void) = delete;
        falcon_dsp_fir_filter_cuda(const falcon_dsp_fir_filter_cuda&) = delete;
    };
}
