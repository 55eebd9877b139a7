// falcon_dsp_polyphase_resampler_cuda.h -- drop-in API of the GPU resampler.
// (Same class as reference hdr/transform/falcon_dsp_polyphase_resampler_cuda.h:160-200.)
#pragma once

#include "transform/falcon_dsp_polyphase_resampler.h"

namespace falcon_dsp
{
    class falcon_dsp_polyphase_resampler_cuda : public falcon_dsp_polyphase_resampler
    {
    public:
        falcon_dsp_polyphase_resampler_cuda(void) = default;
        ~falcon_dsp_polyphase_resampler_cuda(void) = default;
        falcon_dsp_polyphase_resampler_cuda(const falcon_dsp_polyphase_resampler_cuda&) = delete;
    };
}
