// falcon_dsp_multi_rate_channelizer.h -- multi-rate channelizer, drop-in API.
// (Same struct, class and signatures as reference
//  hdr/transform/falcon_dsp_multi_rate_channelizer.h:70-134.)
//
// This class runs the batched form: every channel's shift and resample in one
// fused GPU launch per call.
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_channelizer_s;

struct multi_rate_channelizer_channel_s
{
    uint32_t                             output_sample_rate_in_sps;
    int64_t                              freq_shift_in_hz;
    uint32_t                             up_rate;
    uint32_t                             down_rate;
    std::vector<std::complex<float>>     resample_filter_coeffs;
};

namespace falcon_dsp
{
    class falcon_dsp_multi_rate_channelizer
    {
    public:
        falcon_dsp_multi_rate_channelizer(void);
        virtual ~falcon_dsp_multi_rate_channelizer(void);
        falcon_dsp_multi_rate_channelizer(const falcon_dsp_multi_rate_channelizer&) = delete;

        // false on a second call, no channels, or channels the GPU form rejects
        bool initialize(uint32_t input_sample_rate, std::vector<multi_rate_channelizer_channel_s> channels);
        // out gets one vector per channel (resample(shift(in)) of this chunk)
        bool apply(std::vector<std::complex<float>>& in, std::vector<std::vector<std::complex<float>>>& out);
        void reset_state(void);

    protected:
        bool _initialize(uint32_t input_sample_rate, const std::vector<multi_rate_channelizer_channel_s>& channels,
                         int mode);

        std::mutex          m_mutex;
        fdsp_channelizer_s *m_handle;
        uint32_t            m_num_channels;
    };
}
