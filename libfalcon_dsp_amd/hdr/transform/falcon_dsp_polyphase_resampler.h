// falcon_dsp_polyphase_resampler.h -- streaming rational resampler, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_polyphase_resampler.h:170-200.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_resampler_s;

namespace falcon_dsp
{
    class falcon_dsp_polyphase_resampler
    {
    public:
        falcon_dsp_polyphase_resampler(void);
        virtual ~falcon_dsp_polyphase_resampler(void);
        falcon_dsp_polyphase_resampler(const falcon_dsp_polyphase_resampler&) = delete;

        // up/down > 0 and at least one tap; only the taps' real parts filter
        virtual bool initialize(uint32_t up_rate, uint32_t down_rate,
                                const std::vector<std::complex<float>>& filter_coeffs);
        void reset_state(void);

        // returns the number of outputs written to out (0 when not initialised)
        virtual int32_t apply(std::vector<float>& in, std::vector<float>& out);
        virtual int32_t apply(std::vector<std::complex<float>>& in, std::vector<std::complex<float>>& out);

        // outputs the next apply() of in_count samples will emit
        uint32_t needed_out_count(uint32_t in_count);
        uint32_t coeffs_per_phase(void);

    protected:
        std::mutex        m_mutex;
        fdsp_resampler_s *m_handle;
    };
}
