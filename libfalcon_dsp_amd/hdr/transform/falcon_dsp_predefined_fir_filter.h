// falcon_dsp_predefined_fir_filter.h -- keys of the predefined resampling
// filters, drop-in API. (Same enums and key struct as reference
// hdr/transform/falcon_dsp_predefined_fir_filter.h:40-110; the filters
// themselves come from get_resample_fir_params, declared in
// falcon_dsp_fir_filter.h. The reference's table object is not exported.)
#pragma once

#include <complex>
#include <cstdint>
#include <vector>

enum class filter_source_type_e : uint32_t
{
    REMEZ = 0,
    FIRLS
};

enum class filter_taps_e : int32_t
{
    FILTER_TAPS_OPTIMAL = -1,
    FILTER_TAPS_16 = 16,
    FILTER_TAPS_32 = 32,
    FILTER_TAPS_64 = 64,
    FILTER_TAPS_128 = 128,
    FILTER_TAPS_256 = 256
};

struct predefined_resample_filter_key_s
{
    predefined_resample_filter_key_s(uint32_t in_sps, uint32_t out_sps, filter_taps_e taps, filter_source_type_e src)
      : input_sample_rate(in_sps), output_sample_rate(out_sps), num_taps(taps), source(src)
    { }

    // ordered by input rate, output rate, tap count, then source, each descending
    bool operator<(const predefined_resample_filter_key_s& o) const
    {
        if (o.input_sample_rate != input_sample_rate)
            return o.input_sample_rate < input_sample_rate;
        if (o.output_sample_rate != output_sample_rate)
            return o.output_sample_rate < output_sample_rate;
        if (o.num_taps != num_taps)
            return static_cast<int32_t>(o.num_taps) < static_cast<int32_t>(num_taps);
        return static_cast<uint32_t>(o.source) < static_cast<uint32_t>(source);
    }

    uint32_t             input_sample_rate;
    uint32_t             output_sample_rate;
    filter_taps_e        num_taps;
    filter_source_type_e source;
};

struct predefined_resample_filter_params_s
{
    uint32_t up_rate;
    uint32_t down_rate;
    std::vector<std::complex<float>> coeffs;
};
