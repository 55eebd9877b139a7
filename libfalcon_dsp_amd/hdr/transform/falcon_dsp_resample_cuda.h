// falcon_dsp_resample_cuda.h -- one-shot rational resampling, drop-in API.
// (Same functions as reference hdr/transform/falcon_dsp_resample_cuda.h:85-95.)
#pragma once

#include "transform/falcon_dsp_resample.h"

namespace falcon_dsp
{
    bool resample_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                       std::vector<std::complex<float>>& filter_coeffs,
                       uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out);
    bool upfirdn_cuda(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                      std::vector<float>& filter_coeffs,
                      uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out);
}
