// falcon_dsp_utils.h -- rate helpers of the drop-in API.
// (Same names and signatures as reference hdr/utilities/falcon_dsp_utils.h:94-146;
//  implemented over the C ABI in include/fdsp.h.)
#pragma once

#include <complex>
#include <cstdint>
#include <string>
#include <vector>

namespace falcon_dsp
{
    // greatest common divisor and least common multiple (reference
    // src/utilities/falcon_dsp_utils.cc:123-137)
    uint64_t calculate_gcd(uint64_t a, uint64_t b);
    uint64_t calculate_lcm(uint64_t a, uint64_t b);

    // best rational approximation num/denom of f with denom <= md, by
    // continued fractions (falcon_dsp_utils.cc:231-294)
    void rat_approx(double f, int64_t md, int64_t &num, int64_t &denom);

    // filter delay in samples for a resampling ratio (falcon_dsp_utils.cc:86-116)
    uint32_t calculate_filter_delay_from_sample_rates(uint32_t num_coeffs, uint32_t in_sample_rate_in_sps,
                                                      uint32_t out_sample_rate_in_sps);
    uint32_t calculate_filter_delay_from_up_down_rates(uint32_t num_coeffs, uint32_t up_rate, uint32_t down_rate);

    // sample files (reference falcon_dsp_utils.h:75-79,153-177, .cc:301-811):
    // BINARY is little-endian int16 / float32 values, complex ones as (re, im)
    // pairs; ASCII is one value (or one "re im" pair) per line in the default
    // iostream format. Binary files of float samples cannot be read back as
    // floats (false), as in the reference. Unlike the reference's reader,
    // int16 binary values whose low byte has its top bit set read correctly.
    enum class file_type_e : uint32_t
    {
        BINARY = 0,
        ASCII
    };
    bool write_data_to_file(std::string file_name, file_type_e type, std::vector<int16_t>& data);
    bool write_data_to_file(std::string file_name, file_type_e type, std::vector<float>& data);
    bool read_data_from_file(std::string file_name, file_type_e type, std::vector<int16_t>& data);
    bool read_data_from_file(std::string file_name, file_type_e type, std::vector<float>& data);
    bool write_complex_data_to_file(std::string file_name, file_type_e type,
                                    std::vector<std::complex<int16_t>>& data);
    bool write_complex_data_to_file(std::string file_name, file_type_e type, std::vector<std::complex<float>>& data);
    bool read_complex_data_from_file(std::string file_name, file_type_e type,
                                     std::vector<std::complex<int16_t>>& data);
    bool read_complex_data_from_file(std::string file_name, file_type_e type, std::vector<std::complex<float>>& data);
}
