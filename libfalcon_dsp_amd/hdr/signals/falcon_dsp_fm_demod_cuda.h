// falcon_dsp_fm_demod_cuda.h -- drop-in API of the GPU FM mono demodulator.
// (Same class and signatures as reference hdr/signals/falcon_dsp_fm_demod_cuda.h:92-106.)
//
// The reference's CUDA class runs shift and isolation resample as one
// channelizer (src/signals/falcon_dsp_fm_demod_cuda.cu:113-189), then the
// discriminator, de-emphasis, audio decimator and rescale of the base class
// (.cu:210-265). Here both classes run the same gfx950 pipeline
// (fdsp_fm_demod_*: the fused channelizer with a discriminator tail, the IIR
// de-emphasis, the resampler and the rescale kernels), so their outputs are
// identical for the same chunks.
#pragma once

#include "signals/falcon_dsp_fm_demod.h"

namespace falcon_dsp
{
    class falcon_dsp_fm_demodulator_cuda : public falcon_dsp_fm_demodulator
    {
    public:
        falcon_dsp_fm_demodulator_cuda(void);
        virtual ~falcon_dsp_fm_demodulator_cuda(void) = default;
        falcon_dsp_fm_demodulator_cuda(const falcon_dsp_fm_demodulator_cuda&) = delete;

        // false when already initialised, for a time constant other than the
        // two of the base class, or when no 64-tap REMEZ filter exists for
        // input_sample_rate_in_sps -> 228,000 sps
        bool initialize(uint32_t input_sample_rate_in_sps,
                        int32_t signal_offset_from_dc_in_hz,
                        float deemphasis_time_constant_in_usecs = AMERICAS_FM_DEEMPHASIS_TIME_CONSTANT) override;
        void reset_state(void) override;
        // out cleared first; false when not initialised, for empty input, or
        // when no audio sample came out
        virtual bool demod_mono(std::vector<std::complex<float>>& in, std::vector<int16_t>& out) override;
    };
}
