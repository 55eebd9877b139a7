// falcon_dsp_polar_discriminator.h -- FM discriminator, drop-in API.
// (Same class and signatures as reference hdr/transform/falcon_dsp_polar_discriminator.h:81-114.)
#pragma once

#include <complex>
#include <cstdint>
#include <mutex>
#include <vector>

struct fdsp_discriminator_s;

namespace falcon_dsp
{
    bool polar_discriminator(const std::vector<std::complex<float>>& in, std::vector<float>& out);
    bool polar_discriminator(const std::vector<std::complex<int16_t>>& in, std::vector<float>& out);

    class falcon_dsp_polar_discriminator
    {
    public:
        falcon_dsp_polar_discriminator(void);
        virtual ~falcon_dsp_polar_discriminator(void);
        falcon_dsp_polar_discriminator(const falcon_dsp_polar_discriminator&) = delete;

        void reset_state(void);
        // out[n] = arg(x[n] conj(x[n-1])): N-1 outputs on the first call, N
        // afterwards (the last sample is carried); false for an empty input
        virtual bool apply(const std::vector<std::complex<float>>& in, std::vector<float>& out);
        // complex int16 input, read as floats; shares the carried sample
        virtual bool apply(const std::vector<std::complex<int16_t>>& in, std::vector<float>& out);

    protected:
        std::mutex            m_mutex;
        fdsp_discriminator_s *m_handle;
    };
}
