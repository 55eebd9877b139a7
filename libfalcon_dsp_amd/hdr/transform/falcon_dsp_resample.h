// falcon_dsp_resample.h -- one-shot rational resampling, drop-in API.
// (Same functions as reference hdr/transform/falcon_dsp_resample.h:85-95.)
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

namespace falcon_dsp
{
    // up/down = rat_approx(out_rate / in_rate, 1024), as the reference does
    // (src/transform/falcon_dsp_resample.cc:126-137); true when outputs were made
    bool resample(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                  std::vector<std::complex<float>>& filter_coeffs,
                  uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out);
    bool upfirdn(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                 std::vector<std::complex<float>>& filter_coeffs,
                 uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out);
    bool upfirdn(uint32_t in_sample_rate_in_sps, std::vector<std::complex<float>>& in,
                 std::vector<float>& filter_coeffs,
                 uint32_t out_sample_rate_in_sps, std::vector<std::complex<float>>& out);
}
