// fdsp_fm.hip -- the FM mono demodulation chain after the discriminator, on
// gfx950: a first-order IIR (the de-emphasis), the audio decimator and the
// rescale to int16 (reference src/signals/falcon_dsp_fm_demod.cc:98-271,
// src/transform/falcon_dsp_iir_filter.cc:104-396).
//
// The chain: the batched channelizer with one channel and a discriminator
// tail does shift -> isolation resampler -> discriminator in one launch
// (fm_demod.cc:225-237); iir1_kernel de-emphasises the angles; a resampler handle
// decimates to the audio rate; absmax_kernel and rescale_kernel scale the
// chunk to a peak of 10000 and convert to int16.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/fdsp.h"
#include "fdsp_host.h"
#include "fdsp_internal.h"

using namespace fdsp;

namespace {

// ---------------------------------------------------------------------------
// first-order IIR: y[n] = b0 x[n] - a1 y[n-1]
// ---------------------------------------------------------------------------
// The reference filters sequentially (iir_filter.cc:235-396 with b = [b0],
// a = [1, a1]): each output is fl(fl(b0 x) - fl(a1 y)), its imaginary parts
// zero. Here each thread owns kSeg outputs and starts the same recurrence
// kWarm samples earlier from y = 0 (or at the chunk start from the carried
// output). Two float trajectories of a recurrence with |a1| < 1 draw
// together by |a1| per step and, once a step's products round alike, are the
// same from then on; kWarm = 1024 steps leave |a1|^1024 < 1e-27 of the start
// difference for |a1| <= 0.94, so the segment's outputs are the sequential ones.
constexpr uint32_t kSeg = 128, kWarm = 1024, kIirThreads = 64;
constexpr uint32_t kSpan = kIirThreads * kSeg + kWarm;  // inputs a block stages in LDS (36 KB)

// A block stages the inputs of its 64 segments and their warm-up in LDS with
// coalesced loads, then each thread runs its recurrence from LDS.
template <typename TO>
__global__ __launch_bounds__(kIirThreads) void iir1_kernel(const float *__restrict__ x, uint32_t n, float b0,
                                                           float a1, const float *__restrict__ y_in,
                                                           float *__restrict__ y_out, TO *__restrict__ out)
{
    __shared__ __align__(16) float xs[kSpan];
    const uint32_t first = blockIdx.x * kIirThreads * kSeg;  // the block's first output
    const int64_t lo = (int64_t)first - kWarm;              // xs[t] holds input lo + t
    for (uint32_t t = threadIdx.x; t < kSpan; t += kIirThreads) {
        const int64_t i = lo + t;
        xs[t] = i >= 0 && i < (int64_t)n ? x[i] : 0.f;
    }
    __syncthreads();
    const uint32_t s0 = first + threadIdx.x * kSeg;
    if (s0 >= n)
        return;
    const uint32_t s1 = min(n, s0 + kSeg);
    uint32_t j = s0 > kWarm ? s0 - kWarm : 0u;
    float y = j == 0 ? *y_in : 0.f;
    // j - lo is a multiple of 4 and s0 - j of 32: whole groups of 32 inputs
    // come from LDS as eight 16-byte reads before their 32 serial steps
    const float *xw = xs - lo;  // xw[j] = x[j] for lo <= j < lo + kSpan
    auto step = [&](float xv) {
#pragma clang fp contract(off)
        const float p = b0 * xv;
        const float q = a1 * y;
        y = p - q;
    };
    auto put = [&](uint32_t i) {
        if constexpr (std::is_same<TO, float2>::value)
            out[i] = make_float2(y, 0.f);
        else
            out[i] = y;
    };
    for (; j < s0; j += 32) {
        float4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = *reinterpret_cast<const float4 *>(xw + j + 4 * q);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            step(v[q].x);
            step(v[q].y);
            step(v[q].z);
            step(v[q].w);
        }
    }
    for (; j + 32 <= s1; j += 32) {
        float4 v[8];
#pragma unroll
        for (int q = 0; q < 8; ++q)
            v[q] = *reinterpret_cast<const float4 *>(xw + j + 4 * q);
#pragma unroll
        for (int q = 0; q < 8; ++q) {
            step(v[q].x);
            put(j + 4 * q);
            step(v[q].y);
            put(j + 4 * q + 1);
            step(v[q].z);
            put(j + 4 * q + 2);
            step(v[q].w);
            put(j + 4 * q + 3);
        }
    }
    for (; j < s1; ++j) {
        step(xw[j]);
        put(j);
    }
    if (s1 == n)
        *y_out = y;
}

template <typename TO>
hipError_t launch_iir1(const float *x, uint32_t n, float b0, float a1, const float *y_in, float *y_out, TO *out,
                       hipStream_t s)
{
    const uint32_t blocks = (n + kIirThreads * kSeg - 1) / (kIirThreads * kSeg);
    hipLaunchKernelGGL((iir1_kernel<TO>), dim3(blocks), dim3(kIirThreads), 0, s, x, n, b0, a1, y_in, y_out, out);
    return hipGetLastError();
}

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
__global__ __launch_bounds__(256) void rescale_kernel(const float2 *__restrict__ v, uint32_t n,
                                                      const uint32_t *__restrict__ amax, int16_t *__restrict__ out,
                                                      float *__restrict__ audio)
{
    const float m = __uint_as_float(*amax);
    const float s = m > 0.f ? (float)(10000.0 / (double)m) : 0.f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float a = v[i].x;
        out[i] = i16_trunc(a * s);
        if (audio)
            audio[i] = a;
    }
}

int check_iir1(float a1)
{
    if (!(std::fabs(a1) <= 0.94f))
        return set_err(FDSP_EINVAL, "first-order IIR: |a1| must be at most 0.94 (the segmented recurrence's "
                                    "warm-up is sized for it)");
    return FDSP_OK;
}

}  // namespace

// ===========================================================================
// first-order IIR handle
// ===========================================================================
struct fdsp_iir1_s {
    float b0 = 0.f, a1 = 0.f;
    DevBuf d_y[2], d_in, d_out;  // carried output (two slots), host-path staging
    uint32_t parity = 0;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_iir1_create(float b0, float a1, fdsp_iir1_t *out)
{
    if (!out)
        return set_err(FDSP_EINVAL, "null output");
    if (int rc = check_iir1(a1))
        return rc;
    auto *h = new (std::nothrow) fdsp_iir1_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->b0 = b0;
    h->a1 = a1;
    int rc = h->d_y[0].ensure(sizeof(float));
    if (!rc)
        rc = h->d_y[1].ensure(sizeof(float));
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemset(h->d_y[0].p, 0, sizeof(float)) != hipSuccess))
        rc = set_err(FDSP_EHIP, "iir1_create: HIP setup failed");
    if (rc) {
        fdsp_iir1_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_iir1_destroy(fdsp_iir1_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->d_y[0].release();
    h->d_y[1].release();
    h->d_in.release();
    h->d_out.release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_iir1_reset(fdsp_iir1_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemset(h->d_y[h->parity].p, 0, sizeof(float)));
    return FDSP_OK;
}

extern "C" int fdsp_iir1_apply_device(fdsp_iir1_t h, const float *d_in, size_t n, float *d_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out)
        return set_err(FDSP_EINVAL, "null buffers");
    if (int rc = check_chunk(n, 1))
        return rc;
    HIP_TRY(launch_iir1(d_in, (uint32_t)n, h->b0, h->a1, h->d_y[h->parity].as<float>(),
                        h->d_y[h->parity ^ 1].as<float>(), d_out, pick(h->stream, stream)));
    h->parity ^= 1;
    return FDSP_OK;
}

extern "C" int fdsp_iir1_apply(fdsp_iir1_t h, const float *in, size_t n, float *out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!in || !out)
        return set_err(FDSP_EINVAL, "null buffers");
    if (int rc = h->d_in.ensure(n * sizeof(float)))
        return rc;
    if (int rc = h->d_out.ensure(n * sizeof(float)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(float), hipMemcpyHostToDevice, h->stream));
    if (int rc = fdsp_iir1_apply_device(h, h->d_in.as<float>(), n, h->d_out.as<float>(), h->stream))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, h->d_out.p, n * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

// ===========================================================================
// FM mono demodulator
// ===========================================================================
struct fdsp_fm_demod_s {
    fdsp_channelizer_t chan = nullptr;  // shift + isolation resampler + discriminator
    fdsp_resampler_t mono = nullptr;    // audio decimator
    float b0 = 0.f, a1 = 0.f;           // de-emphasis
    DevBuf d_disc, d_emph, d_mono, d_max, d_y[2], d_in, d_out, d_audio;
    uint32_t parity = 0;
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
    if (int rc = check_iir1(a1))
        return rc;
    auto *h = new (std::nothrow) fdsp_fm_demod_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->b0 = b0;
    h->a1 = a1;
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
    if (!rc)
        rc = h->d_max.ensure(sizeof(uint32_t));
    if (!rc)
        rc = h->d_y[0].ensure(sizeof(float));
    if (!rc)
        rc = h->d_y[1].ensure(sizeof(float));
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemset(h->d_y[0].p, 0, sizeof(float)) != hipSuccess))
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
    for (DevBuf *b : {&h->d_disc, &h->d_emph, &h->d_mono, &h->d_max, &h->d_y[0], &h->d_y[1], &h->d_in, &h->d_out,
                      &h->d_audio})
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
    HIP_TRY(hipMemset(h->d_y[h->parity].p, 0, sizeof(float)));
    return FDSP_OK;
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
    HIP_TRY(launch_iir1(h->d_disc.as<float>(), (uint32_t)n1, h->b0, h->a1, h->d_y[h->parity].as<float>(),
                        h->d_y[h->parity ^ 1].as<float>(), h->d_emph.as<float2>(), s));
    h->parity ^= 1;
    // to the audio rate (fm_demod.cc:244-247): real samples through the complex resampler
    size_t n3 = 0;
    if (int rc = fdsp_resampler_apply_device(h->mono, h->d_emph.as<fdsp_cf32>(), n1, h->d_mono.as<fdsp_cf32>(),
                                             n3max, &n3, s))
        return rc;
    if (n3 == 0)
        return FDSP_OK;
    // rescale to a peak of 10000 (fm_demod.cc:251-266)
    hipLaunchKernelGGL(zero_u32_kernel, dim3(1), dim3(1), 0, s, h->d_max.as<uint32_t>());
    const uint32_t blocks = std::min<uint32_t>(1024, (uint32_t)((n3 + 255) / 256));
    hipLaunchKernelGGL(absmax_kernel, dim3(blocks), dim3(256), 0, s, h->d_mono.as<float2>(), (uint32_t)n3,
                       h->d_max.as<uint32_t>());
    hipLaunchKernelGGL(rescale_kernel, dim3(blocks), dim3(256), 0, s, h->d_mono.as<float2>(), (uint32_t)n3,
                       h->d_max.as<uint32_t>(), d_out, d_audio);
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
