// falcon_dsp_fir_filter_cuda.h -- drop-in API of the GPU FIR.
// (Same class and signatures as reference hdr/transform/falcon_dsp_fir_filter_cuda.h:80-140.)
#pragma once

#include "transform/falcon_dsp_fir_filter.h"

namespace falcon_dsp
{
    bool fir_filter_cuda(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<float>>& in,
                         std::vector<std::complex<float>>& out);
    // (appends to out, as the reference's GPU class does)
    bool fir_filter_cuda(std::vector<std::complex<float>>& coeffs, std::vector<std::complex<int16_t>>& in,
                         std::vector<std::complex<int16_t>>& out);

    class falcon_dsp_fir_filter_cuda : public falcon_dsp_fir_filter
    {
    public:
        explicit falcon_dsp_fir_filter_cuda(std::vector<std::complex<float>>& coeffs)
            : falcon_dsp_fir_filter(coeffs) {}
        ~falcon_dsp_fir_filter_cuda(void) = default;
        falcon_dsp_fir_filter_cuda(void) = delete;
        falcon_dsp_fir_filter_cuda(const falcon_dsp_fir_filter_cuda&) = delete;

        using falcon_dsp_fir_filter::apply;
        // complex int16; the results are appended to out without clearing it,
        // as in the reference (falcon_dsp_fir_filter_cuda.cu:233-254)
        bool apply(std::vector<std::complex<int16_t>>& in, std::vector<std::complex<int16_t>>& out) override;
    };
}
