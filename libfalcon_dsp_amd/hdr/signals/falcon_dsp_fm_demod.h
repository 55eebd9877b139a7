// falcon_dsp_fm_demod.h -- FM mono demodulator, drop-in API.
// (Same class and signatures as reference hdr/signals/falcon_dsp_fm_demod.h:50-90.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_fm_demod_s;

namespace falcon_dsp
{
    class falcon_dsp_fm_demodulator
    {
    public:
        static constexpr float AMERICAS_FM_DEEMPHASIS_TIME_CONSTANT = 75e-6;
        static constexpr float EUROPE_ASIA_FM_DEEMPHASIS_TIME_CONSTANT = 50e-6;
        static const uint32_t FM_RADIO_SAMPLE_RATE_IN_SPS = 228000;
        static const uint32_t FM_RADIO_AUDIO_SAMPLE_RATE_IN_SPS = 45600;

        falcon_dsp_fm_demodulator(void);
        virtual ~falcon_dsp_fm_demodulator(void);
        falcon_dsp_fm_demodulator(const falcon_dsp_fm_demodulator&) = delete;

        // false when already initialised, for a time constant other than the
        // two above, or when no 64-tap REMEZ filter exists for
        // input_sample_rate_in_sps -> 228,000 sps
        virtual bool initialize(uint32_t input_sample_rate_in_sps,
                                int32_t signal_offset_from_dc_in_hz,
                                float deemphasis_time_constant_in_usecs = AMERICAS_FM_DEEMPHASIS_TIME_CONSTANT);
        virtual void reset_state(void);
        // shift, isolate, discriminate, de-emphasise, decimate to 45.6 ksps and
        // scale the chunk to a peak of 10000; out cleared first; false when
        // not initialised or no audio sample came out
        virtual bool demod_mono(std::vector<std::complex<float>>& in, std::vector<int16_t>& out);

    protected:
        std::mutex       m_mutex;
        bool             m_initialized;
        fdsp_fm_demod_s *m_handle;
    };
}
