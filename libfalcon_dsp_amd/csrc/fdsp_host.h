// fdsp_host.h -- host-side helpers shared by the C-ABI translation units:
// error reporting, device buffers, the closed-form stream counts and the
// per-channel tables (NCO parameters, polyphase bank, modulated taps).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/fdsp.h"
#include "fdsp_internal.h"

namespace fdsp {

// the calling thread's last error message (fdsp_last_error)
inline thread_local std::string g_err;

inline int set_err(int code, const char *fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIP_TRY(expr)                                                                       \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            return set_err(FDSP_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                           __LINE__);                                                       \
    } while (0)

constexpr double kTwoPi = 6.283185307179586476925286766559;

// device buffer grown on demand
struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want)
    {
        if (want <= bytes)
            return FDSP_OK;
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (want == 0)
            return FDSP_OK;
        if (hipMalloc(&p, want) != hipSuccess)
            return set_err(FDSP_ENOMEM, "hipMalloc(%zu) failed", want);
        bytes = want;
        return FDSP_OK;
    }
    void release()
    {
        if (p)
            (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T *as() const { return static_cast<T *>(p); }
};

inline uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// outputs a chunk of n inputs starting at global input p emits
inline uint32_t chunk_count(uint64_t p, uint64_t n, uint32_t up, uint32_t down)
{
    return (uint32_t)(ceil_div((p + n) * up, down) - ceil_div(p * up, down));
}

// reference coeff_phase / xOffset after p inputs
inline void resampler_fields(uint64_t p, uint32_t up, uint32_t down, uint32_t *phase, uint32_t *xoff,
                             uint64_t *n_next)
{
    const uint64_t nn = ceil_div(p * up, down);
    const uint64_t num = nn * down;
    *phase = (uint32_t)(num % up);
    *xoff = (uint32_t)(num / up - p);
    *n_next = nn;
}

inline float2 phasor64(double revolutions)
{
    const double t = revolutions - std::floor(revolutions);
    return make_float2((float)std::cos(kTwoPi * t), (float)std::sin(kTwoPi * t));
}

inline int check_chunk(size_t n, uint32_t up)
{
    // chunk-local index arithmetic runs in 32 bits
    if (n >= (1ull << 31) || (uint64_t)n * up >= (1ull << 32))
        return set_err(FDSP_ERANGE, "chunk of %zu samples is above the per-call limit (n < 2^31, n*up < 2^32)", n);
    return FDSP_OK;
}

inline hipStream_t pick(hipStream_t own, fdsp_stream_t s) { return s ? static_cast<hipStream_t>(s) : own; }

// per-shift NCO parameters shared by the freq-shift, channelizer and serving
// paths (R and the shift as fdsp_freq_shift_params gives them)
inline NcoChan make_nco(uint32_t fs, int32_t f)
{
    NcoChan c{};
    c.tab = kNoTable;
    uint32_t r = 1;
    double w = 0;
    fdsp_freq_shift_params(fs, f, &r, &w);
    if (f == 0) {
        c.nu = 0.0;
        c.r = 0xFFFFFFFFu;  // identity shift: the index never matters
    } else {
        c.nu = (double)f / (double)fs;
        c.r = r;
    }
    c.inv_r = 1.0 / (double)c.r;
    for (int i = 0; i < kNcoSamples; ++i)
        c.step[i] = phasor64(c.nu * kNcoOffsets[i]);
    c.omega = w;
    return c;
}

// transposed, flipped real bank: bank[ph*cpp + (cpp-1-j)] = Re h[j*up + ph]
// (reference src/transform/falcon_dsp_polyphase_resampler.cc:131-157)
inline std::vector<float> real_bank(const fdsp_cf32 *taps, uint32_t n_taps, uint32_t up, uint32_t *cpp_out)
{
    uint32_t padded = n_taps;
    while (padded % up)
        ++padded;
    const uint32_t cpp = padded / up;
    std::vector<float> bank((size_t)up * cpp, 0.f);
    for (uint32_t ph = 0; ph < up; ++ph)
        for (uint32_t j = 0; j < cpp; ++j)
            if ((size_t)j * up + ph < n_taps)
                bank[(size_t)ph * cpp + (cpp - 1 - j)] = taps[(size_t)j * up + ph].re;
    *cpp_out = cpp;
    return bank;
}

// The shift folded into the bank: g[ph][t] = bank[ph][t] * e^{-j 2 pi nu d},
// d = cpp-1-t the distance of window sample t from the window's anchor.
// Appended to *taps (up*cpp entries).
inline void append_modulated_taps(const std::vector<float> &bank, uint32_t up, uint32_t cpp, double nu,
                                  std::vector<float2> *taps)
{
    for (uint32_t ph = 0; ph < up; ++ph)
        for (uint32_t t = 0; t < cpp; ++t) {
            const double d = (double)(cpp - 1 - t);
            const double g = bank[(size_t)ph * cpp + t];
            const double rev = -nu * d;
            const double tt = rev - std::floor(rev);
            taps->push_back(make_float2((float)(g * std::cos(kTwoPi * tt)), (float)(g * std::sin(kTwoPi * tt))));
        }
}

}  // namespace fdsp
