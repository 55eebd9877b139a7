// fdsp_internal.h -- shared declarations of the gfx950 kernels' launchers and
// device helpers. Host code in fdsp_capi.hip calls the launch_* functions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdsp {

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// acc += a * b (complex), four FP32 FMAs
__device__ __forceinline__ void cmac(float2 &acc, float2 a, float2 b)
{
    acc.x = fmaf(a.x, b.x, acc.x);
    acc.x = fmaf(-a.y, b.y, acc.x);
    acc.y = fmaf(a.x, b.y, acc.y);
    acc.y = fmaf(a.y, b.x, acc.y);
}

// e^{j 2 pi nu k}: the phase is formed and reduced to [-1/2, 1/2] revolutions
// in FP64 (k < 2^32 and |nu * k| < 2^21, so the reduced phase carries
// < 1e-9 revolutions of error), then FP32 sincospi evaluates it.
__device__ __forceinline__ float2 nco_phasor(int64_t k, double nu)
{
    double t = (double)k * nu;
    t -= rint(t);
    float s, c;
    sincospif(2.0f * (float)t, &s, &c);
    return make_float2(c, s);
}

// a mod m for a < 2^52, m >= 1, using an FP64 quotient estimate plus one
// correction step either way (exact for these ranges).
__device__ __forceinline__ uint64_t mod_u64(uint64_t a, uint32_t m, double inv_m)
{
    int64_t q = (int64_t)((double)a * inv_m);
    int64_t r = (int64_t)a - q * (int64_t)m;
    if (r < 0)
        r += m;
    else if (r >= (int64_t)m)
        r -= m;
    return (uint64_t)r;
}

// ---------------------------------------------------------------------------
// parameter blocks
// ---------------------------------------------------------------------------

// one NCO output of the multi-shift kernel
struct NcoChan {
    double nu;        // f/fs, signed, revolutions per sample
    double inv_r;     // 1/R
    uint32_t r;       // rollover index R (UINT32_MAX for the identity shift)
    uint32_t pad;
    float2 step[4];   // e^{j 2 pi nu i}, i = 0..3
};

// static description of one channel of the fused channelizer
struct ChanStatic {
    double nu;        // f/fs (0 for the identity shift)
    double inv_r;     // 1/R
    uint32_t r;       // rollover index
    uint32_t up, down, cpp;
    uint32_t taps_off;  // offset of this channel's modulated taps (up*cpp complex)
    uint32_t tail;
    float2 wrap;      // e^{j 2 pi nu R}: phase jump across a rollover inside a window
    void *out;        // chunk output base (set per launch)
};

// device-resident stream position of a batched channelizer (graph mode)
struct ChanPos {
    uint64_t p;       // inputs consumed since reset
    uint32_t parity;  // which history buffer holds the current history
    uint32_t pad;
};

// ---------------------------------------------------------------------------
// launchers (all asynchronous on `s`)
// ---------------------------------------------------------------------------
hipError_t launch_nco_multi(const float2 *in, uint32_t n, uint64_t start, const NcoChan *chans,
                            uint32_t n_chan, float2 *out, uint64_t out_stride, hipStream_t s);

// polyphase resampler over one chunk (global input position p); `taps` are
// the real parts of the transposed bank, up*cpp floats. hist_in holds cpp-1
// samples (may be null when cpp == 1); hist_out receives the new history.
hipError_t launch_resampler(const float2 *in, uint32_t n, uint64_t p, uint32_t up, uint32_t down,
                            uint32_t cpp, const float *taps, const float2 *hist_in, float2 *hist_out,
                            float2 *out, uint32_t count, hipStream_t s);

// FIR with complex taps over one chunk; hist_in/hist_out hold m-1 samples
hipError_t launch_fir(const float2 *in, uint32_t n, const float2 *taps, uint32_t m, const float2 *hist_in,
                      float2 *hist_out, float2 *out, hipStream_t s);

// FIR bank with real taps, n_chan channels of n samples each
hipError_t launch_fir_bank(const float2 *in, uint32_t n, uint64_t in_stride, const float *taps, uint32_t m,
                           uint32_t n_chan, const float2 *hist_in, float2 *hist_out, float2 *out,
                           uint64_t out_stride, hipStream_t s);

// polar discriminator; prev (device, 1 sample) is read when has_prev, and
// prev_out receives the last input sample
hipError_t launch_discriminator(const float2 *in, uint32_t n, const float2 *prev, int has_prev,
                                float2 *prev_out, float *out, hipStream_t s);

// fused multi-channel shift + resample (batched channelizer)
struct FusedArgs {
    const float2 *in;
    uint32_t n;
    uint32_t hist_len;          // H: history samples kept (max cpp - 1, rounded up)
    const float2 *hist[2];      // ping-pong history buffers
    float2 *hist_w[2];
    const ChanStatic *chans;
    uint32_t n_chan;
    const float2 *taps;         // all channels' modulated taps
    uint32_t taps_total;        // complex entries
    uint32_t seg_len;           // input samples per block
    uint32_t block;             // threads per block
    ChanPos *pos;               // device position (read by the kernel)
};
hipError_t launch_chan_fused(const FusedArgs &a, hipStream_t s);
// advance the device position by n inputs and flip the history parity
hipError_t launch_chan_advance(ChanPos *pos, uint32_t n, hipStream_t s);
// LDS bytes one block of the fused kernel needs
size_t chan_fused_lds_bytes(uint32_t n_chan, uint32_t taps_total, uint32_t seg_len, uint32_t hist_len);

}  // namespace fdsp
