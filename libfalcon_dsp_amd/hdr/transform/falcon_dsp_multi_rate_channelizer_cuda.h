// falcon_dsp_multi_rate_channelizer_cuda.h -- drop-in API of the GPU channelizer.
// (Same class and signatures as reference hdr/transform/falcon_dsp_multi_rate_channelizer_cuda.h:102-167.)
//
// This class keeps the reference's per-channel structure: one multi-shift NCO
// launch writing a full-rate copy per channel, then one resampler launch per
// channel. falcon_dsp_multi_rate_channelizer is the fused batched form.
#pragma once

#include "transform/falcon_dsp_multi_rate_channelizer.h"

namespace falcon_dsp
{
    class falcon_dsp_multi_rate_channelizer_cuda : public falcon_dsp_multi_rate_channelizer
    {
    public:
        falcon_dsp_multi_rate_channelizer_cuda(void) = default;
        ~falcon_dsp_multi_rate_channelizer_cuda(void) = default;
        falcon_dsp_multi_rate_channelizer_cuda(const falcon_dsp_multi_rate_channelizer_cuda&) = delete;

        bool initialize(uint32_t input_sample_rate, std::vector<multi_rate_channelizer_channel_s> channels);
    };
}
