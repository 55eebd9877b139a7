// fdsp_fm.hip -- the FM mono demodulation chain after the discriminator, on
// gfx950: a first-order IIR (the de-emphasis), the audio decimator and the
// rescale to int16 (reference src/signals/falcon_dsp_fm_demod.cc:98-271,
// src/transform/falcon_dsp_iir_filter.cc:104-396).
//
// The chain: the batched channelizer with one channel and a discriminator
// tail does shift -> isolation resampler -> discriminator (fm_demod.cc:
// 225-237); the general IIR filter (fdsp_iir.hip) de-emphasises the angles; a
// resampler handle decimates to the audio rate; peak_rescale_kernel (or, for
// long chunks, absmax_kernel and rescale_kernel) scales the chunk to a peak of
// 10000 and converts to int16.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/fdsp.h"
#include "fdsp_host.h"
#include "fdsp_iir.h"
#include "fdsp_internal.h"

using namespace fdsp;

namespace {

// ---------------------------------------------------------------------------
// rescale to int16 (fm_demod.cc:251-266)
// ---------------------------------------------------------------------------
__global__ void zero_u32_kernel(uint32_t *p) { *p = 0u; }

// the largest |re| of n complex samples, as the bits of a non-negative float
// (which order as unsigned integers) in *amax
__global__ __launch_bounds__(256) void absmax_kernel(const float2 *__restrict__ v, uint32_t n, uint32_t *amax)
{
    float m = 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        m = fmaxf(m, fabsf(v[i].x));
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o));
    __shared__ float wm[4];
    if ((threadIdx.x & 63) == 0)
        wm[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fmaxf(fmaxf(wm[0], wm[1]), fmaxf(wm[2], wm[3]));
        atomicMax(amax, __float_as_uint(m));
    }
}

// out[i] = int16(re v[i] * s), s = (float)(10000.0 / max) as the reference
// forms it (a double quotient stored in a float); truncation toward zero,
// saturated, and all zero when the chunk's peak is zero
__device__ __forceinline__ float peak_scale(float m) { return m > 0.f ? (float)(10000.0 / (double)m) : 0.f; }

__global__ __launch_bounds__(256) void rescale_kernel(const float2 *__restrict__ v, uint32_t n,
                                                      const uint32_t *__restrict__ amax, int16_t *__restrict__ out,
                                                      float *__restrict__ audio)
{
    const float s = peak_scale(__uint_as_float(*amax));
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float a = v[i].x;
        out[i] = i16_trunc(a * s);
        if (audio)
            audio[i] = a;
    }
}

// both steps in one block, for chunks of at most kPeakOneBlock samples (the
// second pass re-reads the chunk from L2)
constexpr uint32_t kPeakThreads = 1024, kPeakOneBlock = 1u << 16;

// kPer > 0: at most kPer samples a thread, kept in registers between the
// two passes (one read of the chunk); kPer = 0: any n up to kPeakOneBlock
template <int kPer>
__global__ __launch_bounds__(kPeakThreads) void peak_rescale_kernel(const float2 *__restrict__ v, uint32_t n,
                                                                   int16_t *__restrict__ out,
                                                                   float *__restrict__ audio)
{
    __shared__ float wm[kPeakThreads / 64];
    float m = 0.f;
    float a[kPer > 0 ? kPer : 1];
    if constexpr (kPer > 0) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const uint32_t i = threadIdx.x + e * kPeakThreads;
            a[e] = i < n ? v[i].x : 0.f;
            m = fmaxf(m, fabsf(a[e]));
        }
    } else {
#pragma unroll 8
        for (uint32_t i = threadIdx.x; i < n; i += kPeakThreads)
            m = fmaxf(m, fabsf(v[i].x));
    }
    for (int o = 32; o > 0; o >>= 1)
        m = fmaxf(m, __shfl_xor(m, o));
    if ((threadIdx.x & 63) == 0)
        wm[threadIdx.x >> 6] = m;
    __syncthreads();
    m = wm[0];
#pragma unroll
    for (uint32_t w = 1; w < kPeakThreads / 64; ++w)
        m = fmaxf(m, wm[w]);
    const float s = peak_scale(m);
    if constexpr (kPer > 0) {
#pragma unroll
        for (int e = 0; e < kPer; ++e) {
            const uint32_t i = threadIdx.x + e * kPeakThreads;
            if (i < n) {
                out[i] = i16_trunc(a[e] * s);
                if (audio)
                    audio[i] = a[e];
            }
        }
    } else {
#pragma unroll 8
        for (uint32_t i = threadIdx.x; i < n; i += kPeakThreads) {
            const float x = v[i].x;
            out[i] = i16_trunc(x * s);
            if (audio)
                audio[i] = x;
        }
    }
}

}  // namespace

// ===========================================================================
// FM mono demodulator
// ===========================================================================
struct fdsp_fm_demod_s {
    fdsp_channelizer_t chan = nullptr;  // shift + isolation resampler + discriminator
    fdsp_resampler_t mono = nullptr;    // audio decimator
    fdsp_iir_t emph = nullptr;          // de-emphasis, b = [b0], a = [1, a1]
    DevBuf d_disc, d_emph, d_mono, d_max, d_in, d_out, d_audio;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_fm_deemphasis(double tau, float *b0, float *a1)
{
    if (!b0 || !a1 || !(tau > 0.0))
        return set_err(FDSP_EINVAL, "de-emphasis needs tau > 0");
    // fm_demod.cc:146-165: x = exp(-1 / (200000 tau)), b = [1 - x], a = [1, -x]
    const double x = std::exp(-1.0 / (200000.0 * tau));
    *b0 = (float)(1.0 - x);
    *a1 = (float)(-x);
    return FDSP_OK;
}

extern "C" int fdsp_fm_demod_create(uint32_t fs, int32_t offset_hz, double tau, uint32_t iso_up, uint32_t iso_down,
                                    const fdsp_cf32 *iso_taps, uint32_t n_iso, uint32_t mono_up, uint32_t mono_down,
                                    const fdsp_cf32 *mono_taps, uint32_t n_mono, fdsp_fm_demod_t *out)
{
    if (!out || !iso_taps || !mono_taps || n_iso == 0 || n_mono == 0 || fs == 0)
        return set_err(FDSP_EINVAL, "fm_demod_create: need fs and both resamplers' taps");
    float b0, a1;
    if (int rc = fdsp_fm_deemphasis(tau, &b0, &a1))
        return rc;
    auto *h = new (std::nothrow) fdsp_fm_demod_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    fdsp_channel_desc d{};
    d.output_sample_rate_in_sps = (uint32_t)((uint64_t)fs * iso_up / iso_down);
    d.freq_shift_in_hz = offset_hz;
    d.up_rate = iso_up;
    d.down_rate = iso_down;
    d.taps = iso_taps;
    d.n_taps = n_iso;
    d.tail = FDSP_TAIL_DISCRIMINATOR;
    int rc = fdsp_channelizer_create(fs, &d, 1, FDSP_CHAN_BATCHED, &h->chan);
    if (!rc)
        rc = fdsp_resampler_create(mono_up, mono_down, mono_taps, n_mono, &h->mono);
    if (!rc) {
        const fdsp_cf32 b[1] = {{b0, 0.f}}, a[2] = {{1.f, 0.f}, {a1, 0.f}};
        rc = fdsp_iir_create(b, 1, a, 2, &h->emph);
    }
    if (!rc)
        rc = h->d_max.ensure(sizeof(uint32_t));
    if (!rc && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        rc = set_err(FDSP_EHIP, "fm_demod_create: HIP setup failed");
    if (rc) {
        fdsp_fm_demod_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_fm_demod_destroy(fdsp_fm_demod_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    fdsp_channelizer_destroy(h->chan);
    fdsp_resampler_destroy(h->mono);
    fdsp_iir_destroy(h->emph);
    for (DevBuf *b : {&h->d_disc, &h->d_emph, &h->d_mono, &h->d_max, &h->d_in, &h->d_out, &h->d_audio})
        b->release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_fm_demod_reset(fdsp_fm_demod_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    if (int rc = fdsp_channelizer_reset(h->chan))
        return rc;
    if (int rc = fdsp_resampler_reset(h->mono))
        return rc;
    return fdsp_iir_reset(h->emph);
}

extern "C" int fdsp_fm_demod_out_count(fdsp_fm_demod_t h, size_t n, size_t *count)
{
    if (!h || !count)
        return set_err(FDSP_EINVAL, "null argument");
    // the isolation stage's count for n inputs, then the decimator's for as
    // many (the discriminator passes one fewer on the first chunk)
    uint32_t n1 = 0, n2 = 0;
    if (int rc = fdsp_channelizer_needed_out_count(h->chan, 0, (uint32_t)n, &n1))
        return rc;
    if (int rc = fdsp_resampler_needed_out_count(h->mono, n1, &n2))
        return rc;
    *count = n2;  // at most; the exact count comes back from apply
    return FDSP_OK;
}

extern "C" int fdsp_fm_demod_apply_device(fdsp_fm_demod_t h, const fdsp_cf32 *d_in, size_t n, int16_t *d_out,
                                          float *d_audio, size_t cap, size_t *n_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!n_out)
        return set_err(FDSP_EINVAL, "null n_out");
    *n_out = 0;
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out)
        return set_err(FDSP_EINVAL, "null buffers");
    hipStream_t s = pick(h->stream, stream);
    // capacity first, so a failure leaves every stage's state as it was
    uint32_t n1max = 0, n3max = 0;
    if (int rc = fdsp_channelizer_needed_out_count(h->chan, 0, (uint32_t)n, &n1max))
        return rc;
    if (int rc = fdsp_resampler_needed_out_count(h->mono, n1max, &n3max))
        return rc;
    if (n3max > cap)
        return set_err(FDSP_ERANGE, "fm_demod: capacity %zu below the %u samples this chunk can emit", cap, n3max);
    if (int rc = h->d_disc.ensure(std::max<size_t>(n1max, 1) * sizeof(float)))
        return rc;
    if (int rc = h->d_emph.ensure(std::max<size_t>(n1max, 1) * sizeof(float2)))
        return rc;
    if (int rc = h->d_mono.ensure(std::max<size_t>(n3max, 1) * sizeof(float2)))
        return rc;
    // shift, isolate, discriminate (fm_demod.cc:225-237)
    void *outs[1] = {h->d_disc.p};
    size_t caps[1] = {n1max}, n1 = 0;
    if (int rc = fdsp_channelizer_apply_device(h->chan, d_in, n, outs, caps, &n1, s))
        return rc;
    if (n1 == 0)
        return FDSP_OK;  // nothing reached the audio stages
    // de-emphasis (fm_demod.cc:239-242)
    if (int rc = iir_apply_typed(h->emph, kSampleF32, h->d_disc.p, n1, kSampleCf32, h->d_emph.p, s))
        return rc;
    // to the audio rate (fm_demod.cc:244-247): real samples through the complex resampler
    size_t n3 = 0;
    if (int rc = fdsp_resampler_apply_device(h->mono, h->d_emph.as<fdsp_cf32>(), n1, h->d_mono.as<fdsp_cf32>(),
                                             n3max, &n3, s))
        return rc;
    if (n3 == 0)
        return FDSP_OK;
    // rescale to a peak of 10000 (fm_demod.cc:251-266)
    if (n3 <= 16 * kPeakThreads) {
        hipLaunchKernelGGL(peak_rescale_kernel<16>, dim3(1), dim3(kPeakThreads), 0, s, h->d_mono.as<float2>(),
                           (uint32_t)n3, d_out, d_audio);
    } else if (n3 <= kPeakOneBlock) {
        hipLaunchKernelGGL(peak_rescale_kernel<0>, dim3(1), dim3(kPeakThreads), 0, s, h->d_mono.as<float2>(),
                           (uint32_t)n3, d_out, d_audio);
    } else {
        hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, h->d_max.as<uint32_t>());
        const uint32_t blocks = std::min<uint32_t>(1024, (uint32_t)((n3 + 255) / 256));
        hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(256), 0, s, h->d_mono.as<float2>(), (uint32_t)n3,
                           h->d_max.as<uint32_t>());
        hipLaunchKernelGGL(rescale_kernel, dim3(blocks), dim3(256), 0, s, h->d_mono.as<float2>(), (uint32_t)n3,
                           h->d_max.as<uint32_t>(), d_out, d_audio);
    }
    HIP_TRY(hipGetLastError());
    *n_out = n3;
    return FDSP_OK;
}

extern "C" int fdsp_fm_demod_apply(fdsp_fm_demod_t h, const fdsp_cf32 *in, size_t n, int16_t *out, float *audio,
                                   size_t cap, size_t *n_out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!n_out)
        return set_err(FDSP_EINVAL, "null n_out");
    *n_out = 0;
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!in || !out)
        return set_err(FDSP_EINVAL, "null buffers");
    size_t most = 0;
    if (int rc = fdsp_fm_demod_out_count(h, n, &most))
        return rc;
    if (most > cap)
        return set_err(FDSP_ERANGE, "fm_demod: capacity %zu below the %zu samples this chunk can emit", cap, most);
    if (int rc = h->d_in.ensure(n * sizeof(float2)))
        return rc;
    if (int rc = h->d_out.ensure(std::max<size_t>(most, 1) * sizeof(int16_t)))
        return rc;
    if (audio)
        if (int rc = h->d_audio.ensure(std::max<size_t>(most, 1) * sizeof(float)))
            return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    size_t got = 0;
    if (int rc = fdsp_fm_demod_apply_device(h, h->d_in.as<fdsp_cf32>(), n, h->d_out.as<int16_t>(),
                                            audio ? h->d_audio.as<float>() : nullptr, most, &got, h->stream))
        return rc;
    if (got) {
        HIP_TRY(hipMemcpyAsync(out, h->d_out.p, got * sizeof(int16_t), hipMemcpyDeviceToHost, h->stream));
        if (audio)
            HIP_TRY(hipMemcpyAsync(audio, h->d_audio.p, got * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    *n_out = got;
    return FDSP_OK;
}
