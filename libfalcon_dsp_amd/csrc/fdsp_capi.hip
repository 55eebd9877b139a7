// fdsp_capi.hip -- the C ABI (include/fdsp.h) over the gfx950 kernels.
//
// Handles own their device buffers and carry the reference's stream state in
// closed form: the only per-stream integers are the inputs consumed since the
// last reset (p) and the ping-pong parity of the history buffers; coeff_phase,
// xOffset, output counts and NCO indices are all derived from p exactly
// (reference src/transform/falcon_dsp_polyphase_resampler.cc:216-228,256-319,
// src/transform/falcon_dsp_freq_shift.cc:177-189).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <sys/stat.h>
#include <numeric>
#include <string>
#include <vector>

#include <type_traits>

#include "../../include/fdsp.h"
#include "fdsp_host.h"
#include "fdsp_internal.h"

using namespace fdsp;

namespace {
// 64-bit FNV-1a over byte ranges
struct Fnv64 {
    uint64_t h = 0xcbf29ce484222325ull;
    void add(const void *p, size_t n)
    {
        const unsigned char *b = static_cast<const unsigned char *>(p);
        for (size_t i = 0; i < n; ++i)
            h = (h ^ b[i]) * 0x100000001b3ull;
    }
};
}  // namespace

// ===========================================================================
// misc + plumbing
// ===========================================================================
extern "C" const char *fdsp_last_error(void) { return g_err.c_str(); }
extern "C" int fdsp_version(void) { return 100; }

extern "C" int fdsp_device_count(int *count)
{
    if (!count)
        return set_err(FDSP_EINVAL, "null count");
    HIP_TRY(hipGetDeviceCount(count));
    return FDSP_OK;
}
extern "C" int fdsp_set_device(int device)
{
    HIP_TRY(hipSetDevice(device));
    return FDSP_OK;
}
extern "C" int fdsp_malloc(void **ptr, size_t bytes)
{
    if (!ptr)
        return set_err(FDSP_EINVAL, "null ptr");
    if (hipMalloc(ptr, bytes ? bytes : 1) != hipSuccess)
        return set_err(FDSP_ENOMEM, "hipMalloc(%zu) failed", bytes);
    return FDSP_OK;
}
extern "C" int fdsp_free(void *ptr)
{
    if (ptr)
        HIP_TRY(hipFree(ptr));
    return FDSP_OK;
}
extern "C" int fdsp_memcpy_h2d(void *dst, const void *src, size_t bytes, fdsp_stream_t s)
{
    if (s)
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, (hipStream_t)s));
    else
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
    return FDSP_OK;
}
extern "C" int fdsp_memcpy_d2h(void *dst, const void *src, size_t bytes, fdsp_stream_t s)
{
    if (s)
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, (hipStream_t)s));
    else
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
    return FDSP_OK;
}
extern "C" int fdsp_memcpy_d2d(void *dst, const void *src, size_t bytes, fdsp_stream_t s)
{
    if (s)
        HIP_TRY(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, (hipStream_t)s));
    else
        HIP_TRY(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToDevice));
    return FDSP_OK;
}
extern "C" int fdsp_memset(void *dst, int value, size_t bytes, fdsp_stream_t s)
{
    if (s)
        HIP_TRY(hipMemsetAsync(dst, value, bytes, (hipStream_t)s));
    else
        HIP_TRY(hipMemset(dst, value, bytes));
    return FDSP_OK;
}
extern "C" int fdsp_host_alloc_pinned(void **ptr, size_t bytes)
{
    if (!ptr)
        return set_err(FDSP_EINVAL, "null ptr");
    if (hipHostMalloc(ptr, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess)
        return set_err(FDSP_ENOMEM, "hipHostMalloc(%zu) failed", bytes);
    return FDSP_OK;
}
extern "C" int fdsp_host_free_pinned(void *ptr)
{
    if (ptr)
        HIP_TRY(hipHostFree(ptr));
    return FDSP_OK;
}
extern "C" int fdsp_stream_create(fdsp_stream_t *stream)
{
    if (!stream)
        return set_err(FDSP_EINVAL, "null stream");
    hipStream_t s;
    HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    *stream = s;
    return FDSP_OK;
}
extern "C" int fdsp_stream_destroy(fdsp_stream_t stream)
{
    if (stream)
        HIP_TRY(hipStreamDestroy((hipStream_t)stream));
    return FDSP_OK;
}
extern "C" int fdsp_stream_sync(fdsp_stream_t stream)
{
    HIP_TRY(hipStreamSynchronize((hipStream_t)stream));
    return FDSP_OK;
}
extern "C" int fdsp_device_sync(void)
{
    HIP_TRY(hipDeviceSynchronize());
    return FDSP_OK;
}
extern "C" int fdsp_event_create(fdsp_event_t *ev)
{
    if (!ev)
        return set_err(FDSP_EINVAL, "null event");
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    *ev = e;
    return FDSP_OK;
}
extern "C" int fdsp_event_destroy(fdsp_event_t ev)
{
    if (ev)
        HIP_TRY(hipEventDestroy((hipEvent_t)ev));
    return FDSP_OK;
}
extern "C" int fdsp_event_record(fdsp_event_t ev, fdsp_stream_t stream)
{
    HIP_TRY(hipEventRecord((hipEvent_t)ev, (hipStream_t)stream));
    return FDSP_OK;
}
extern "C" int fdsp_event_sync(fdsp_event_t ev)
{
    HIP_TRY(hipEventSynchronize((hipEvent_t)ev));
    return FDSP_OK;
}
extern "C" int fdsp_event_elapsed_ms(fdsp_event_t a, fdsp_event_t b, float *ms)
{
    if (!ms)
        return set_err(FDSP_EINVAL, "null ms");
    HIP_TRY(hipEventElapsedTime(ms, (hipEvent_t)a, (hipEvent_t)b));
    return FDSP_OK;
}

// ===========================================================================
// rate helpers (reference src/utilities/falcon_dsp_utils.cc:123-137,231-294;
// src/transform/falcon_dsp_freq_shift.cc:194-235)
// ===========================================================================
extern "C" uint64_t fdsp_gcd(uint64_t a, uint64_t b)
{
    while (b) {
        const uint64_t t = a % b;
        a = b;
        b = t;
    }
    return a;
}
extern "C" uint64_t fdsp_lcm(uint64_t a, uint64_t b)
{
    const uint64_t g = fdsp_gcd(a, b);
    return g ? (a / g) * b : 0;
}

// Best rational approximation num/denom of f with denom < max_denom, by the
// continued-fraction expansion of f's binary expansion; when the next
// convergent would reach the bound, the largest admissible semiconvergent is
// taken if it is at least half the next partial quotient.
extern "C" void fdsp_rat_approx(double f, int64_t max_denom, int64_t *num, int64_t *denom)
{
    if (!num || !denom)
        return;
    if (max_denom <= 1) {
        *denom = 1;
        *num = (int64_t)f;
        return;
    }
    const bool neg = f < 0;
    if (neg)
        f = -f;
    int64_t scale = 1;  // f = d / scale exactly, scale a power of two
    while (f != std::floor(f)) {
        scale <<= 1;
        f *= 2;
    }
    int64_t d = (int64_t)f, n = scale;
    int64_t hm2 = 0, hm1 = 1;  // numerators of the last two convergents
    int64_t km2 = 1, km1 = 0;  // denominators
    for (int i = 0; i < 64; ++i) {
        const int64_t a = n ? d / n : 0;
        if (i && !a)
            break;
        const int64_t rem = n ? d % n : 0;
        d = n;
        n = rem;
        int64_t x = a;
        bool last = false;
        if (km1 * a + km2 >= max_denom) {
            x = (max_denom - km2) / km1;
            if (x * 2 >= a || km1 >= max_denom)
                last = true;
            else
                break;
        }
        const int64_t h = x * hm1 + hm2, k = x * km1 + km2;
        hm2 = hm1;
        hm1 = h;
        km2 = km1;
        km1 = k;
        if (last)
            break;
    }
    *denom = km1;
    *num = neg ? -hm1 : hm1;
}

extern "C" int fdsp_freq_shift_params(uint32_t fs, int32_t f_hz, uint32_t *rollover, double *omega)
{
    if (!rollover || !omega || fs == 0)
        return set_err(FDSP_EINVAL, "fs must be nonzero and outputs non-null");
    if (f_hz == 0) {
        *rollover = 1;
        *omega = 0.0;
        return FDSP_OK;
    }
    // the period fs/|f| in samples, scaled by ten until it is at least 1e5
    double r = (double)fs / std::fabs((double)f_hz);
    while (r < 1e5)
        r *= 10;
    *rollover = (uint32_t)std::llround(r);
    *omega = ((double)f_hz / (double)fs) * 2.0 * M_PI;
    return FDSP_OK;
}


// ===========================================================================
// NCO frequency shift
// ===========================================================================
struct fdsp_freq_shift_s {
    uint32_t fs = 0;
    std::vector<NcoChan> chans;
    DevBuf d_chans, d_in, d_out;
    uint64_t start = 0;  // samples since reset (the NCO index is start mod R)
    hipStream_t stream = nullptr;
    DevBuf d_tab;            // int16 output: the shifts' phasor tables, built at the first such call
    bool tab_built = false;
};

extern "C" int fdsp_freq_shift_create(uint32_t fs, const int32_t *shifts_hz, uint32_t n_shifts,
                                      fdsp_freq_shift_t *out)
{
    if (!out || !shifts_hz || n_shifts == 0 || fs == 0)
        return set_err(FDSP_EINVAL, "freq_shift_create: need fs > 0 and at least one shift");
    auto *h = new (std::nothrow) fdsp_freq_shift_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->fs = fs;
    for (uint32_t i = 0; i < n_shifts; ++i)
        h->chans.push_back(make_nco(fs, shifts_hz[i]));
    int rc = h->d_chans.ensure(sizeof(NcoChan) * n_shifts);
    if (rc) {
        delete h;
        return rc;
    }
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
        hipMemcpy(h->d_chans.p, h->chans.data(), sizeof(NcoChan) * n_shifts, hipMemcpyHostToDevice) !=
            hipSuccess) {
        h->d_chans.release();
        delete h;
        return set_err(FDSP_EHIP, "freq_shift_create: HIP setup failed");
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_freq_shift_destroy(fdsp_freq_shift_t h)
{
    if (!h)
        return FDSP_OK;
    (void)hipStreamSynchronize(h->stream);
    h->d_chans.release();
    h->d_tab.release();
    h->d_in.release();
    h->d_out.release();
    (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_freq_shift_reset(fdsp_freq_shift_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    h->start = 0;
    return FDSP_OK;
}

extern "C" int fdsp_freq_shift_set_state(fdsp_freq_shift_t h, uint64_t samples_handled)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (samples_handled >= (1ull << 52))
        return set_err(FDSP_ERANGE, "set_state: position above 2^52");
    // launches queued on caller streams read the position as an argument,
    // so only the host copy changes
    h->start = samples_handled;
    return FDSP_OK;
}

extern "C" int fdsp_freq_shift_get_state(fdsp_freq_shift_t h, uint32_t chan, uint64_t *samples_handled)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (chan >= h->chans.size() || !samples_handled)
        return set_err(FDSP_EINVAL, "bad channel");
    *samples_handled = h->chans[chan].nu == 0.0 ? 0 : h->start % h->chans[chan].r;
    return FDSP_OK;
}

namespace {

// Phasor tables for int16 output: for each shift with R <= 2^22, the
// reference's ((float)cos(omega s), (float)sin(omega s)), s = 0 .. R-1, formed
// on the host exactly as falcon_dsp_freq_shift.cc:146-148 forms them (FP64
// angle and libm, then float). Up to 32 MiB per shift; larger R keep the
// device's FP64 sincos.
constexpr uint32_t kTableMaxR = 1u << 22;
int build_phasor_tables(fdsp_freq_shift_t h)
{
    std::vector<float2> tab;
    std::vector<NcoChan> chans = h->chans;
    for (auto &c : chans) {
        if (c.nu == 0.0 || c.r > kTableMaxR || tab.size() + c.r >= kNoTable)
            continue;
        c.tab = (uint32_t)tab.size();
        for (uint32_t s = 0; s < c.r; ++s) {
            const double ang = c.omega * (double)s;
            tab.push_back(make_float2((float)std::cos(ang), (float)std::sin(ang)));
        }
    }
    if (!tab.empty()) {
        if (int rc = h->d_tab.ensure(tab.size() * sizeof(float2)))
            return rc;
        // queued launches read d_chans; their fields other than tab do not change
        HIP_TRY(hipStreamSynchronize(h->stream));
        HIP_TRY(hipMemcpy(h->d_tab.p, tab.data(), tab.size() * sizeof(float2), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(h->d_chans.p, chans.data(), sizeof(NcoChan) * chans.size(), hipMemcpyHostToDevice));
        h->chans = chans;
    }
    h->tab_built = true;
    return FDSP_OK;
}

// T: float2 (cf32) or short2 (ci16) samples, in and out alike
template <typename T>
int freq_shift_device(fdsp_freq_shift_t h, const T *d_in, size_t n, T *d_out, size_t out_stride,
                      fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out || (h->chans.size() > 1 && out_stride < n))
        return set_err(FDSP_EINVAL, "bad buffers");
    if (int rc = check_chunk(n, 1))
        return rc;
    constexpr bool kInt = std::is_same<T, short2>::value;
    if (kInt && !h->tab_built)
        if (int rc = build_phasor_tables(h))
            return rc;
    HIP_TRY(launch_nco_multi(d_in, (uint32_t)n, h->start, h->d_chans.as<NcoChan>(), (uint32_t)h->chans.size(), d_out,
                             out_stride, pick(h->stream, stream), kInt ? h->d_tab.as<float2>() : nullptr));
    h->start += n;
    return FDSP_OK;
}

template <typename T>
int freq_shift_host(fdsp_freq_shift_t h, const T *in, size_t n, T *const *outs)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!in || !outs)
        return set_err(FDSP_EINVAL, "null buffers");
    const size_t C = h->chans.size();
    if (int rc = h->d_in.ensure(n * sizeof(T)))
        return rc;
    if (int rc = h->d_out.ensure(n * C * sizeof(T)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    if (int rc = freq_shift_device(h, h->d_in.as<T>(), n, h->d_out.as<T>(), n, h->stream))
        return rc;
    for (size_t c = 0; c < C; ++c)
        HIP_TRY(hipMemcpyAsync(outs[c], h->d_out.as<T>() + c * n, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_freq_shift_apply_device(fdsp_freq_shift_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                            size_t out_stride, fdsp_stream_t stream)
{
    return freq_shift_device(h, reinterpret_cast<const float2 *>(d_in), n, reinterpret_cast<float2 *>(d_out),
                             out_stride, stream);
}

extern "C" int fdsp_freq_shift_apply_chunks_device(fdsp_freq_shift_t h, const fdsp_cf32 *d_in, size_t chunk_len,
                                                   size_t n_chunks, fdsp_cf32 *d_out, size_t out_stride,
                                                   fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n_chunks == 0)
        return FDSP_OK;
    if (chunk_len == 0)
        return set_err(FDSP_EINVAL, "apply_chunks: chunk_len > 0 needed");
    // the shift of consecutive chunks continues the index across them, so the
    // chunks' outputs (shift s, chunk k at d_out + s*out_stride + k*chunk_len)
    // are those of one call over their concatenation
    return fdsp_freq_shift_apply_device(h, d_in, chunk_len * n_chunks, d_out, out_stride, stream);
}

extern "C" int fdsp_freq_shift_apply(fdsp_freq_shift_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *const *outs)
{
    return freq_shift_host(h, reinterpret_cast<const float2 *>(in), n, reinterpret_cast<float2 *const *>(outs));
}

extern "C" int fdsp_freq_shift_apply_ci16_device(fdsp_freq_shift_t h, const fdsp_ci16 *d_in, size_t n,
                                                 fdsp_ci16 *d_out, size_t out_stride, fdsp_stream_t stream)
{
    return freq_shift_device(h, reinterpret_cast<const short2 *>(d_in), n, reinterpret_cast<short2 *>(d_out),
                             out_stride, stream);
}

extern "C" int fdsp_freq_shift_apply_ci16(fdsp_freq_shift_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *const *outs)
{
    return freq_shift_host(h, reinterpret_cast<const short2 *>(in), n, reinterpret_cast<short2 *const *>(outs));
}

// ===========================================================================
// polyphase resampler
// ===========================================================================
struct fdsp_resampler_s {
    uint32_t up = 1, down = 1, cpp = 1;
    DevBuf d_taps, d_hist[2], d_in, d_out;
    uint32_t parity = 0;
    uint64_t p = 0;  // inputs consumed since reset
    hipStream_t stream = nullptr;
};


extern "C" int fdsp_resampler_create(uint32_t up, uint32_t down, const fdsp_cf32 *taps, uint32_t n_taps,
                                     fdsp_resampler_t *out)
{
    if (!out || !taps || up == 0 || down == 0 || n_taps == 0)
        return set_err(FDSP_EINVAL, "resampler_create: need up, down > 0 and taps");
    auto *h = new (std::nothrow) fdsp_resampler_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->up = up;
    h->down = down;
    auto bank = real_bank(taps, n_taps, up, &h->cpp);
    const size_t hbytes = (size_t)(h->cpp > 1 ? h->cpp - 1 : 1) * sizeof(float2);
    int rc = h->d_taps.ensure(bank.size() * sizeof(float));
    if (!rc)
        rc = h->d_hist[0].ensure(hbytes);
    if (!rc)
        rc = h->d_hist[1].ensure(hbytes);
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemcpy(h->d_taps.p, bank.data(), bank.size() * sizeof(float), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                hipMemset(h->d_hist[0].p, 0, hbytes) != hipSuccess ||
                hipMemset(h->d_hist[1].p, 0, hbytes) != hipSuccess))
        rc = set_err(FDSP_EHIP, "resampler_create: HIP setup failed");
    if (rc) {
        fdsp_resampler_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_destroy(fdsp_resampler_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->d_taps.release();
    h->d_hist[0].release();
    h->d_hist[1].release();
    h->d_in.release();
    h->d_out.release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_reset(fdsp_resampler_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    h->p = 0;
    h->parity = 0;
    HIP_TRY(hipMemsetAsync(h->d_hist[0].p, 0, h->d_hist[0].bytes, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

extern "C" int fdsp_resampler_needed_out_count(fdsp_resampler_t h, uint32_t in_count, uint32_t *out_count)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!out_count)
        return set_err(FDSP_EINVAL, "null output");
    *out_count = chunk_count(h->p, in_count, h->up, h->down);
    return FDSP_OK;
}

extern "C" int fdsp_resampler_coeffs_per_phase(fdsp_resampler_t h, uint32_t *cpp)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!cpp)
        return set_err(FDSP_EINVAL, "null output");
    *cpp = h->cpp;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_get_state(fdsp_resampler_t h, uint32_t *coeff_phase, uint32_t *x_offset,
                                        uint64_t *in_total, uint64_t *out_total)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    uint32_t ph, xo;
    uint64_t nn;
    resampler_fields(h->p, h->up, h->down, &ph, &xo, &nn);
    if (coeff_phase)
        *coeff_phase = ph;
    if (x_offset)
        *x_offset = xo;
    if (in_total)
        *in_total = h->p;
    if (out_total)
        *out_total = nn;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_set_state(fdsp_resampler_t h, uint32_t coeff_phase, uint32_t x_offset,
                                        const fdsp_cf32 *history, size_t n_history)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    const size_t hlen = h->cpp > 1 ? h->cpp - 1 : 0;
    if (coeff_phase >= h->up)
        return set_err(FDSP_EINVAL, "set_state: coeff_phase %u is not below up = %u", coeff_phase, h->up);
    if (n_history != hlen || (hlen && !history))
        return set_err(FDSP_EINVAL, "set_state: the history holds coeffs_per_phase - 1 = %zu samples", hlen);
    // The carried fields follow from the inputs consumed p: the next output m
    // = ceil(p up/down) reads anchor (m down)/up, so m down = up (p + xOffset)
    // + coeff_phase. The smallest m >= 1 giving an integer p >= 0 that
    // reproduces (coeff_phase, xOffset) is taken (any such p yields the same
    // outputs; they repeat with period up*down in m).
    uint64_t p = 0;
    bool found = coeff_phase == 0 && x_offset == 0;  // the state after reset
    for (uint64_t m = 1; !found && m <= (uint64_t)h->up * h->down + h->up; ++m) {
        const uint64_t num = m * h->down;
        if (num % h->up != coeff_phase || num / h->up < x_offset)
            continue;
        p = num / h->up - x_offset;
        uint32_t ph, xo;
        uint64_t nn;
        resampler_fields(p, h->up, h->down, &ph, &xo, &nn);
        found = ph == coeff_phase && xo == x_offset;
    }
    if (!found)
        return set_err(FDSP_EINVAL, "set_state: no stream position gives coeff_phase %u and xOffset %u (up %u, down %u)",
                       coeff_phase, x_offset, h->up, h->down);
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipDeviceSynchronize());  // launches on caller streams may still read the history
    if (hlen)
        HIP_TRY(hipMemcpy(h->d_hist[h->parity].p, history, hlen * sizeof(float2), hipMemcpyHostToDevice));
    h->p = p;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_apply_device(fdsp_resampler_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                           size_t cap, size_t *n_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!n_out)
        return set_err(FDSP_EINVAL, "null n_out");
    *n_out = 0;
    if (n == 0)
        return FDSP_OK;  // the reference emits nothing and keeps its state
    if (!d_in)
        return set_err(FDSP_EINVAL, "null input");
    if (int rc = check_chunk(n, h->up))
        return rc;
    const uint32_t count = chunk_count(h->p, n, h->up, h->down);
    if (count > cap || (count && !d_out))
        return set_err(FDSP_ERANGE, "output capacity %zu below the %u samples this chunk emits", cap, count);
    HIP_TRY(launch_resampler(reinterpret_cast<const float2 *>(d_in), (uint32_t)n, h->p, h->up, h->down, h->cpp,
                             h->d_taps.as<float>(), h->d_hist[h->parity].as<float2>(),
                             h->d_hist[h->parity ^ 1].as<float2>(), reinterpret_cast<float2 *>(d_out), count,
                             pick(h->stream, stream)));
    h->parity ^= 1;
    h->p += n;
    *n_out = count;
    return FDSP_OK;
}

extern "C" int fdsp_resampler_apply_chunks_device(fdsp_resampler_t h, const fdsp_cf32 *d_in, size_t chunk_len,
                                                  size_t n_chunks, fdsp_cf32 *d_out, size_t cap, size_t *counts,
                                                  fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!counts || (n_chunks && (!d_in || chunk_len == 0)))
        return set_err(FDSP_EINVAL, "apply_chunks: need counts, an input and chunk_len > 0");
    // the resampler is chunk-invariant: n_chunks consecutive chunks emit, back
    // to back, exactly the outputs of their concatenation, so one launch over
    // the concatenation is the same as n_chunks calls
    const uint64_t total = (uint64_t)chunk_len * n_chunks;
    uint64_t p = h->p;
    size_t sum = 0;
    for (size_t k = 0; k < n_chunks; ++k, p += chunk_len) {
        counts[k] = chunk_count(p, chunk_len, h->up, h->down);
        sum += counts[k];
    }
    if (sum > cap || (sum && !d_out))
        return set_err(FDSP_ERANGE, "output capacity %zu below the %zu samples these chunks emit", cap, sum);
    size_t got = 0;
    if (int rc = fdsp_resampler_apply_device(h, d_in, total, d_out, cap, &got, stream))
        return rc;
    return got == sum ? FDSP_OK : set_err(FDSP_EHIP, "apply_chunks: %zu outputs, expected %zu", got, sum);
}

extern "C" int fdsp_resampler_apply(fdsp_resampler_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out, size_t cap,
                                    size_t *n_out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!n_out)
        return set_err(FDSP_EINVAL, "null n_out");
    *n_out = 0;
    if (n == 0)
        return FDSP_OK;
    if (!in)
        return set_err(FDSP_EINVAL, "null input");
    if (int rc = check_chunk(n, h->up))
        return rc;
    const uint32_t count = chunk_count(h->p, n, h->up, h->down);
    if (count > cap || (count && !out))
        return set_err(FDSP_ERANGE, "output capacity %zu below the %u samples this chunk emits", cap, count);
    if (int rc = h->d_in.ensure(n * sizeof(float2)))
        return rc;
    if (int rc = h->d_out.ensure((size_t)(count ? count : 1) * sizeof(float2)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    size_t got = 0;
    if (int rc = fdsp_resampler_apply_device(h, reinterpret_cast<const fdsp_cf32 *>(h->d_in.p), n,
                                             h->d_out.as<fdsp_cf32>(), count, &got, h->stream))
        return rc;
    if (got)
        HIP_TRY(hipMemcpyAsync(out, h->d_out.p, got * sizeof(float2), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *n_out = got;
    return FDSP_OK;
}

// ===========================================================================
// FIR (complex taps)
// ===========================================================================
struct fdsp_fir_s {
    uint32_t m = 0;
    DevBuf d_taps, d_hist[2], d_in, d_out;
    uint32_t parity = 0;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_fir_create(const fdsp_cf32 *coeffs, uint32_t n_coeffs, fdsp_fir_t *out)
{
    if (!out || !coeffs || n_coeffs == 0)
        return set_err(FDSP_EINVAL, "fir_create: need at least one coefficient");
    auto *h = new (std::nothrow) fdsp_fir_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->m = n_coeffs;
    const size_t hbytes = (size_t)(n_coeffs > 1 ? n_coeffs - 1 : 1) * sizeof(float2);
    int rc = h->d_taps.ensure(n_coeffs * sizeof(float2));
    if (!rc)
        rc = h->d_hist[0].ensure(hbytes);
    if (!rc)
        rc = h->d_hist[1].ensure(hbytes);
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemcpy(h->d_taps.p, coeffs, n_coeffs * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(h->d_hist[0].p, 0, hbytes) != hipSuccess ||
                hipMemset(h->d_hist[1].p, 0, hbytes) != hipSuccess))
        rc = set_err(FDSP_EHIP, "fir_create: HIP setup failed");
    if (rc) {
        fdsp_fir_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_fir_destroy(fdsp_fir_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->d_taps.release();
    h->d_hist[0].release();
    h->d_hist[1].release();
    h->d_in.release();
    h->d_out.release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_fir_reset(fdsp_fir_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    h->parity = 0;
    HIP_TRY(hipMemsetAsync(h->d_hist[0].p, 0, h->d_hist[0].bytes, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

namespace {

// T: float2 (cf32) or short2 (ci16) samples, in and out alike; the history
// is float either way
template <typename T>
int fir_device(fdsp_fir_t h, const T *d_in, size_t n, T *d_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out)
        return set_err(FDSP_EINVAL, "null buffers");
    if (int rc = check_chunk(n, 1))
        return rc;
    hipError_t e = launch_fir(d_in, (uint32_t)n, h->d_taps.as<float2>(), h->m, h->d_hist[h->parity].as<float2>(),
                              h->d_hist[h->parity ^ 1].as<float2>(), d_out, pick(h->stream, stream));
    if (e == hipErrorInvalidValue)
        return set_err(FDSP_ERANGE, "fir: %u taps exceed the LDS tile", h->m);
    HIP_TRY(e);
    h->parity ^= 1;
    return FDSP_OK;
}

template <typename T>
int fir_host(fdsp_fir_t h, const T *in, size_t n, T *out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!in || !out)
        return set_err(FDSP_EINVAL, "null buffers");
    if (int rc = h->d_in.ensure(n * sizeof(T)))
        return rc;
    if (int rc = h->d_out.ensure(n * sizeof(T)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    if (int rc = fir_device(h, h->d_in.as<T>(), n, h->d_out.as<T>(), h->stream))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, h->d_out.p, n * sizeof(T), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_fir_apply_device(fdsp_fir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                     fdsp_stream_t stream)
{
    return fir_device(h, reinterpret_cast<const float2 *>(d_in), n, reinterpret_cast<float2 *>(d_out), stream);
}

extern "C" int fdsp_fir_apply(fdsp_fir_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out)
{
    return fir_host(h, reinterpret_cast<const float2 *>(in), n, reinterpret_cast<float2 *>(out));
}

extern "C" int fdsp_fir_apply_ci16_device(fdsp_fir_t h, const fdsp_ci16 *d_in, size_t n, fdsp_ci16 *d_out,
                                          fdsp_stream_t stream)
{
    return fir_device(h, reinterpret_cast<const short2 *>(d_in), n, reinterpret_cast<short2 *>(d_out), stream);
}

extern "C" int fdsp_fir_apply_ci16(fdsp_fir_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out)
{
    return fir_host(h, reinterpret_cast<const short2 *>(in), n, reinterpret_cast<short2 *>(out));
}

// ===========================================================================
// FIR bank (real taps)
// ===========================================================================
struct fdsp_fir_bank_s {
    uint32_t n_chan = 0, m = 0;
    DevBuf d_taps, d_hist[2];
    uint32_t parity = 0;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_fir_bank_create(uint32_t n_chan, uint32_t n_taps, const float *taps, fdsp_fir_bank_t *out)
{
    if (!out || !taps || n_chan == 0 || n_taps == 0 || n_chan > 65535)
        return set_err(FDSP_EINVAL, "fir_bank_create: need 1..65535 channels and taps");
    auto *h = new (std::nothrow) fdsp_fir_bank_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->n_chan = n_chan;
    h->m = n_taps;
    const size_t hbytes = (size_t)n_chan * (n_taps > 1 ? n_taps - 1 : 1) * sizeof(float2);
    int rc = h->d_taps.ensure((size_t)n_chan * n_taps * sizeof(float));
    if (!rc)
        rc = h->d_hist[0].ensure(hbytes);
    if (!rc)
        rc = h->d_hist[1].ensure(hbytes);
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemcpy(h->d_taps.p, taps, (size_t)n_chan * n_taps * sizeof(float), hipMemcpyHostToDevice) !=
                    hipSuccess ||
                hipMemset(h->d_hist[0].p, 0, hbytes) != hipSuccess ||
                hipMemset(h->d_hist[1].p, 0, hbytes) != hipSuccess))
        rc = set_err(FDSP_EHIP, "fir_bank_create: HIP setup failed");
    if (rc) {
        fdsp_fir_bank_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_fir_bank_destroy(fdsp_fir_bank_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->d_taps.release();
    h->d_hist[0].release();
    h->d_hist[1].release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_fir_bank_reset(fdsp_fir_bank_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    h->parity = 0;
    HIP_TRY(hipMemsetAsync(h->d_hist[0].p, 0, h->d_hist[0].bytes, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

extern "C" int fdsp_fir_bank_apply_device(fdsp_fir_bank_t h, const fdsp_cf32 *d_in, size_t n, size_t in_stride,
                                          fdsp_cf32 *d_out, size_t out_stride, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out || (h->n_chan > 1 && (in_stride < n || out_stride < n)))
        return set_err(FDSP_EINVAL, "bad buffers or strides");
    if (int rc = check_chunk(n, 1))
        return rc;
    hipError_t e = launch_fir_bank(reinterpret_cast<const float2 *>(d_in), (uint32_t)n, in_stride,
                                   h->d_taps.as<float>(), h->m, h->n_chan, h->d_hist[h->parity].as<float2>(),
                                   h->d_hist[h->parity ^ 1].as<float2>(), reinterpret_cast<float2 *>(d_out),
                                   out_stride, pick(h->stream, stream));
    if (e == hipErrorInvalidValue)
        return set_err(FDSP_ERANGE, "fir_bank: %u taps exceed the LDS tile", h->m);
    HIP_TRY(e);
    h->parity ^= 1;
    return FDSP_OK;
}

// ===========================================================================
// polar discriminator
// ===========================================================================
struct fdsp_discriminator_s {
    DevBuf d_prev[2], d_in, d_out;
    uint32_t parity = 0;
    bool has_prev = false;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_discriminator_create(fdsp_discriminator_t *out)
{
    if (!out)
        return set_err(FDSP_EINVAL, "null output");
    auto *h = new (std::nothrow) fdsp_discriminator_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    int rc = h->d_prev[0].ensure(sizeof(float2));
    if (!rc)
        rc = h->d_prev[1].ensure(sizeof(float2));
    if (!rc && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        rc = set_err(FDSP_EHIP, "stream creation failed");
    if (rc) {
        fdsp_discriminator_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_discriminator_destroy(fdsp_discriminator_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    h->d_prev[0].release();
    h->d_prev[1].release();
    h->d_in.release();
    h->d_out.release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_discriminator_reset(fdsp_discriminator_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    h->has_prev = false;
    return FDSP_OK;
}

namespace {

// T: float2 (cf32) or short2 (ci16) input samples
template <typename T>
int discriminator_device(fdsp_discriminator_t h, const T *d_in, size_t n, float *d_out, size_t *n_out,
                         fdsp_stream_t stream)
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
    if (int rc = check_chunk(n, 1))
        return rc;
    HIP_TRY(launch_discriminator(d_in, (uint32_t)n, h->d_prev[h->parity].as<float2>(), h->has_prev ? 1 : 0,
                                 h->d_prev[h->parity ^ 1].as<float2>(), d_out, pick(h->stream, stream)));
    *n_out = h->has_prev ? n : n - 1;
    h->has_prev = true;
    h->parity ^= 1;
    return FDSP_OK;
}

template <typename T>
int discriminator_host(fdsp_discriminator_t h, const T *in, size_t n, float *out, size_t *n_out)
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
    if (int rc = h->d_in.ensure(n * sizeof(T)))
        return rc;
    if (int rc = h->d_out.ensure(n * sizeof(float)))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(T), hipMemcpyHostToDevice, h->stream));
    size_t got = 0;
    if (int rc = discriminator_device(h, h->d_in.as<T>(), n, h->d_out.as<float>(), &got, h->stream))
        return rc;
    if (got)
        HIP_TRY(hipMemcpyAsync(out, h->d_out.p, got * sizeof(float), hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    *n_out = got;
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_discriminator_apply_device(fdsp_discriminator_t h, const fdsp_cf32 *d_in, size_t n, float *d_out,
                                               size_t *n_out, fdsp_stream_t stream)
{
    return discriminator_device(h, reinterpret_cast<const float2 *>(d_in), n, d_out, n_out, stream);
}

extern "C" int fdsp_discriminator_apply(fdsp_discriminator_t h, const fdsp_cf32 *in, size_t n, float *out,
                                        size_t *n_out)
{
    return discriminator_host(h, reinterpret_cast<const float2 *>(in), n, out, n_out);
}

extern "C" int fdsp_discriminator_apply_ci16_device(fdsp_discriminator_t h, const fdsp_ci16 *d_in, size_t n,
                                                    float *d_out, size_t *n_out, fdsp_stream_t stream)
{
    return discriminator_device(h, reinterpret_cast<const short2 *>(d_in), n, d_out, n_out, stream);
}

extern "C" int fdsp_discriminator_apply_ci16(fdsp_discriminator_t h, const fdsp_ci16 *in, size_t n, float *out,
                                             size_t *n_out)
{
    return discriminator_host(h, reinterpret_cast<const short2 *>(in), n, out, n_out);
}

// ===========================================================================
// multi-rate channelizer
// ===========================================================================
namespace {
struct ChanHost {
    int32_t shift;
    uint32_t out_rate, up, down, cpp, tail;
    NcoChan nco;
    std::vector<float> bank;    // real transposed bank (plain mode)
    DevBuf d_bank;              // plain mode
    DevBuf d_hist[2];           // plain mode: shifted-sample history
    DevBuf d_scratch;           // complex outputs before the discriminator tail
    DevBuf d_prev[2];           // discriminator previous sample
    uint32_t prev_parity = 0;
    bool has_prev = false;
};
}  // namespace

struct fdsp_channelizer_s {
    uint32_t fs = 0;
    int mode = FDSP_CHAN_BATCHED;
    std::vector<ChanHost> ch;
    uint64_t p = 0;             // inputs consumed since reset
    uint32_t parity = 0;        // plain-mode history parity (all channels together)
    hipStream_t stream = nullptr;
    // plain mode
    DevBuf d_nco, d_full;       // NCO params, full-rate shifted copies (C x n)
    // batched mode
    std::vector<ChanStatic> stat;
    DevBuf d_stat, d_stat_graph, d_taps, d_hist[2], d_pos;
    ChanStatic *pinned_stat = nullptr;
    hipEvent_t stat_ev = nullptr;   // last table upload
    hipEvent_t use_ev = nullptr;    // last fused launch that read the table
    bool stat_valid = false;
    bool pos_dirty = false;         // device position behind the host's (eager calls since the last graph launch)
    uint32_t taps_total = 0, hist_len = 0;
    uint32_t seg_len = FDSP_TILING_STRIDED, block = 256;
    uint64_t tune_hash = 0;  // FNV-1a of fs, the mode and every channel's description (tuning-cache key)
    // FDSP_CHAN_BATCHED_BF16X2: the group kernel's MFMA groups with 8-aligned
    // windows run on bf16 hi/lo planes (about 16 significant bits)
    bool bf16_planes = false;
    // segment-major FP32 kernel (fdsp_chan_seg.hip): the batched form's kernel
    // whenever every channel fits it (windows of at most 64 taps)
    bool seg_ok = false;   // the segment plan fits this channel set
    bool use_seg = false;  // launches use it (fdsp_channelizer_set_kernel)
    uint32_t seg_len_in = 2560, seg_halo = 0;
    std::vector<SegGroup> seg_groups;
    std::vector<float> seg_btab;
    DevBuf d_seg_groups, d_seg_btab, d_ptaps;
    // block-time per tile of each group, measured by the autotuner (empty:
    // the tile_cost estimate weighs the groups)
    std::vector<double> gcost;
    int only_group = -1;       // autotuner: only this group, on only_blocks blocks
    uint32_t only_blocks = 0;
    uint32_t capacity = 1024;  // fused-kernel blocks resident at once (occupancy x CUs)
    // work plan: channel groups (kGroupWords u32 each) and their MFMA B tables
    std::vector<uint32_t> groups;
    std::vector<float> btab;
    DevBuf d_groups, d_btab;
    DevBuf d_claims;  // per-group tile claim counters + a finished-block counter (zero between launches)
    DevBuf d_sink;    // stores of rows and lanes with no output
    uint32_t n_groups = 0, lds_bytes = 0;
    uint32_t lds_plan[2] = {0, 0};  // launch LDS without / with the fused record scheduled
    std::vector<uint32_t> tps;
    // block tables by chunk length (eager calls), and the graph's own
    std::vector<std::pair<uint32_t, DevBuf>> tables;  // key: chunk length * 2 + input 16-byte aligned
    // groups each launch schedules: [1] when the input is 16-byte aligned (a
    // fused record in place of its members), [0] otherwise (the members)
    std::vector<uint32_t> sched[2];
    bool fuse = true;      // FDSP_FUSE=0: no fused records are built
    bool fuse_on = false;  // launches schedule the fused record (fdsp_channelizer_set_fusion, autotuner)
    DevBuf d_table_graph;
    // host-pointer path
    DevBuf d_in;
    std::vector<DevBuf> d_outs;
    // graph mode
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    size_t graph_chunk = 0;
    std::vector<size_t> graph_counts;
};

namespace {

int chan_upload_outs(fdsp_channelizer_s *h, void *const *outs)
{
    // batched mode: output pointers live in the device channel table
    bool same = h->stat_valid;
    for (size_t c = 0; c < h->ch.size(); ++c) {
        // the group kernel writes a tail channel's complex samples to scratch
        // for the discriminator launch after it; the segment kernel fuses the
        // discriminator and writes the angles itself
        void *o = h->ch[c].tail && !h->use_seg ? h->ch[c].d_scratch.p : outs[c];
        if (h->stat[c].out != o) {
            h->stat[c].out = o;
            same = false;
        }
    }
    if (same)
        return FDSP_OK;
    // the pinned staging copy may still be in flight from the previous upload,
    // and a launch on another stream may still be reading the device table
    HIP_TRY(hipEventSynchronize(h->stat_ev));
    HIP_TRY(hipStreamWaitEvent(h->stream, h->use_ev, 0));
    std::memcpy(h->pinned_stat, h->stat.data(), sizeof(ChanStatic) * h->stat.size());
    HIP_TRY(hipMemcpyAsync(h->d_stat.p, h->pinned_stat, sizeof(ChanStatic) * h->stat.size(),
                           hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipEventRecord(h->stat_ev, h->stream));
    h->stat_valid = true;
    return FDSP_OK;
}

constexpr size_t kLdsCap = 64 * 1024;

// window length of a channel on the MFMA path (its taps padded in front with
// zeros to 32 or 64), or 0 when the channel does not qualify
uint32_t mfma_window(const ChanHost &k)
{
    if (k.up != 1)
        return 0;
    return k.cpp <= 32 ? 32 : k.cpp <= 64 ? 64 : 0;
}

// LDS bytes of one tile of `rows` outputs of a group, per kind
size_t group_lds(uint32_t kind, uint32_t up, uint32_t down, uint32_t w, uint32_t rows, uint32_t block)
{
    switch (kind) {
    case kGroupMfma:
        // the group's taps, then the whole staging capacity: dense spans
        // transposed by down (mc columns, rounded up to odd), sparse windows
        // at a pitch of w + 1
        if (down < w) {
            // (bf16 planes: 4 planes of cap samples of 2 bytes, never more)
            const size_t cap = (size_t)mfma_stage(block) * block;
            return (kMfmaMaxChan * w + down * (((cap + down - 1) / down) | 1u)) * sizeof(float2);
        }
        return (kMfmaMaxChan * w + (size_t)mfma_stage(block) * (block / w) * (w + 1)) * sizeof(float2);
    case kGroupVDense: {
        const size_t mc = (((size_t)(rows - 1) * down + w + down - 1) / down) | 1u;
        return (down * mc + 64 * kVDenseOpl) * sizeof(float2);
    }
    case kGroupVSparse:
        return (size_t)kVSparsePass * (block + 1) * sizeof(float2);
    default:
        return (((size_t)(rows - 1) * down + up - 1) / up + w + 1) * sizeof(float2);
    }
}

// Rows per tile of each group for `block` threads, written into word 5 of
// the group records; the launch's LDS size (the largest need) goes to lds[0]
// for the member groups alone and to lds[1] with the fused record as well.
// MFMA tiles hold what one staging pass of mfma_stage(block) samples per
// thread brings in; the one-channel tiles are sized to the same LDS budget where
// they can be (all blocks of the launch get the same LDS allocation).
int size_groups(std::vector<uint32_t> &groups, uint32_t block, uint32_t lds[2])
{
    const size_t G = groups.size() / kGroupWords;
    size_t budget = 32 * 1024;
    for (int pass = 0; pass < 2; ++pass) {
        for (size_t g = 0; g < G; ++g) {
            uint32_t *gw = groups.data() + g * kGroupWords;
            const uint32_t kind = gw[0], up = gw[1], down = gw[2], w = gw[3];
            if ((kind == kGroupMfma) != (pass == 0))
                continue;
            const size_t cap = pass == 0 ? kLdsCap : budget;
            uint32_t rows, step, lo;
            switch (kind) {
            case kGroupMfma: {
                const uint32_t stage = mfma_stage(block) * block;
                // bf16 planes: the span starts up to 7 samples before the first
                // window and class windows read 40 slots; rows come in whole
                // sub-tiles of every class
                const bool bf = mfma_bf16_planes(w, down);
                const uint32_t reach = bf ? w + 16 : w;
                rows = down < w ? ((stage - reach) / down + 1) : stage / w;
                step = lo = bf ? 16 * mfma_bf16_classes(down) : 16;
                rows = std::max(lo, rows / step * step);
                break;
            }
            case kGroupVDense:
                step = lo = 64;
                rows = block * kVDenseOpl;
                break;
            case kGroupVSparse:
                step = lo = rows = block;
                break;
            default:
                step = lo = 1;
                rows = 2 * block;
                break;
            }
            while (rows > lo && group_lds(kind, up, down, w, rows, block) > cap)
                rows = kind == kGroupGeneric ? rows / 2 : rows - step;
            const size_t need = group_lds(kind, up, down, w, rows, block);
            if (need > kLdsCap || (uint64_t)rows * down + up >= (1ull << 31))
                return set_err(FDSP_ERANGE, "batched channelizer: a tile of channel %u (up %u, down %u, %u taps per "
                                            "phase) does not fit one block; use FDSP_CHAN_PLAIN",
                               gw[8], up, down, w);
            gw[5] = rows;
            budget = std::max(budget, need);
        }
    }
    lds[0] = (uint32_t)budget;
    for (size_t g = 0; g < G; ++g)
        if (groups[g * kGroupWords] == kGroupFused) {
            // B tables (6 KB per member class), planes, taps, NCO steps,
            // channels, and (in the 512 spare bytes) 32-byte member descriptors
            const uint32_t *fw = groups.data() + g * kGroupWords;
            size_t nb = 0;
            for (uint32_t i = 0; i < fw[4]; ++i)
                nb += mfma_bf16_classes(groups[fw[8 + i] * kGroupWords + 2]);
            const size_t need = nb * kBfWordsPerBase * 4 + (size_t)mfma_stage(block) * block * 8 +
                                fw[4] * kMfmaMaxChan * (32 * sizeof(float2) + 2 * sizeof(float2) + sizeof(ChanStatic)) +
                                512;
            budget = std::max(budget, need);
        }
    lds[1] = (uint32_t)budget;
    return FDSP_OK;
}

// blocks of the fused kernel the GPU holds at once for the current plan
int update_capacity(fdsp_channelizer_s *h)
{
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    bool wide = false;
    for (uint32_t g = 0; g < h->n_groups; ++g)
        wide |= h->groups[g * kGroupWords] == kGroupMfma && h->groups[g * kGroupWords + 3] == 64;
    const int per_cu = chan_groups_occupancy(h->block, wide, h->lds_bytes);
    h->capacity = (uint32_t)std::max(1, per_cu) * (uint32_t)std::max(1, cus);
    return FDSP_OK;
}

// Group the channels. Channels with up 1, the same down and the same padded
// window (32 or 64) form MFMA groups of up to kMfmaMaxChan when there are at
// least kMfmaMinChan of them; every other channel is a one-channel VALU group.
// `taps` holds every channel's modulated taps (ChanStatic::taps_off).
// round-to-nearest-even bf16 of a finite float, and back
uint16_t bf16_rne(float f)
{
    uint32_t u;
    std::memcpy(&u, &f, sizeof(u));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
float bf16_to_float(uint16_t b)
{
    const uint32_t u = (uint32_t)b << 16;
    float f;
    std::memcpy(&f, &u, sizeof(f));
    return f;
}

// Work plan of the segment-major FP32 kernel. Channels with up 1, at most 32
// taps and the same down form MFMA groups of 2..8 (a larger set of one down
// splits into near-equal groups); every other channel is a one-channel VALU
// group. Each channel's modulated taps are padded in front to its window
// (32 or 64) per filter phase. Leaves use_seg false when a window exceeds 64
// taps or the plan exceeds the kernel's limits (the group kernel then runs).
void build_seg_plan(fdsp_channelizer_s *h, const std::vector<float2> &taps)
{
    const uint32_t C = (uint32_t)h->ch.size();
    h->seg_ok = h->use_seg = false;
    h->seg_groups.clear();
    h->seg_btab.clear();
    if (C > kSegMaxChan)
        return;
    uint32_t max_win = 32, reach = 31;
    for (auto &k : h->ch) {
        if (k.cpp > 64)
            return;
        const uint32_t win = k.cpp <= 32 ? 32u : 64u;
        max_win = std::max(max_win, win);
        // a discriminator tail recomputes the output before a segment's first:
        // its window reaches ceil(down/up) inputs further back
        reach = std::max(reach, win - 1 + (k.tail ? (uint32_t)ceil_div(k.down, k.up) : 0u));
    }
    if (reach > kSegMaxHalo)
        return;
    // the kernel's NCO index arithmetic is 32-bit signed: periods up to 2^30
    for (uint32_t c = 0; c < C; ++c)
        if (h->stat[c].r > kNoShiftPeriod)
            return;
    // padded modulated taps
    std::vector<float2> ptaps;
    for (uint32_t c = 0; c < C; ++c) {
        ChanHost &k = h->ch[c];
        ChanStatic &s = h->stat[c];
        s.win = k.cpp <= 32 ? 32u : 64u;
        s.poff = (uint32_t)ptaps.size();
        s.prev[0] = k.d_prev[0].as<float2>();
        s.prev[1] = k.d_prev[1].as<float2>();
        s.dk = (uint32_t)((16ull * kSegWaves * k.down) % s.r);
        s.d1 = k.down % s.r;
        s.e_dk = phasor64(s.nu * (double)s.dk);
        s.e_d1 = phasor64(s.nu * (double)s.d1);
        for (uint32_t ph = 0; ph < k.up; ++ph) {
            for (uint32_t t = 0; t < s.win - k.cpp; ++t)
                ptaps.push_back(make_float2(0.f, 0.f));
            for (uint32_t t = 0; t < k.cpp; ++t)
                ptaps.push_back(taps[s.taps_off + ph * k.cpp + t]);
        }
    }
    // MFMA groups by down
    std::vector<bool> done(C, false);
    std::vector<SegGroup> mfma, valu;
    for (uint32_t c = 0; c < C; ++c) {
        const ChanHost &k = h->ch[c];
        if (done[c] || k.up != 1 || k.cpp > 32 || k.tail)
            continue;
        std::vector<uint32_t> same;
        for (uint32_t d = c; d < C; ++d) {
            const ChanHost &o = h->ch[d];
            if (!done[d] && o.up == 1 && o.cpp <= 32 && !o.tail && o.down == k.down)
                same.push_back(d);
        }
        if (same.size() < 2)
            continue;
        const size_t n_g = (same.size() + kMfmaMaxChan - 1) / kMfmaMaxChan;
        for (size_t g = 0, at = 0; g < n_g; ++g) {
            const size_t take = (same.size() - at) / (n_g - g);
            SegGroup gr{};
            gr.kind = kSegMfma;
            gr.m = (uint32_t)take;
            gr.up = 1;
            gr.down = k.down;
            gr.win = 32;
            gr.btab = (uint32_t)h->seg_btab.size();
            for (size_t i = 0; i < take; ++i) {
                gr.chan[i] = same[at + i];
                done[same[at + i]] = true;
            }
            // B operand of K step s in lane (col, g): window position
            // t = g + 4(s/2), part s%2 (0 real, 1 imaginary); column 2c gives
            // channel c's real output (Gr xr - Gi xi), 2c+1 its imaginary
            // output (Gi xr + Gr xi)
            for (uint32_t st = 0; st < 16; ++st)
                for (uint32_t lane = 0; lane < 64; ++lane) {
                    const uint32_t col = lane % 16, t = lane / 16 + 4 * (st / 2), part = st % 2;
                    float v = 0.f;
                    if (col / 2 < gr.m) {
                        const ChanStatic &cs = h->stat[gr.chan[col / 2]];
                        const float2 G = ptaps[cs.poff + t];
                        if (col % 2 == 0)
                            v = part == 0 ? G.x : -G.y;
                        else
                            v = part == 0 ? G.y : G.x;
                    }
                    h->seg_btab.push_back(v);
                }
            mfma.push_back(gr);
            at += take;
        }
    }
    for (uint32_t c = 0; c < C; ++c) {
        if (done[c])
            continue;
        SegGroup gr{};
        gr.kind = kSegValu;
        gr.m = 1;
        gr.up = h->ch[c].up;
        gr.down = h->ch[c].down;
        gr.win = h->stat[c].win;
        gr.tail = h->ch[c].tail;
        gr.chan[0] = c;
        valu.push_back(gr);
    }
    // the waves take work items in plan order: the VALU channels first (their
    // one-output-per-lane batches are the longest latency chains, so they
    // start first), then the MFMA groups, largest output share first
    std::stable_sort(mfma.begin(), mfma.end(), [](const SegGroup &a, const SegGroup &b) { return a.down < b.down; });
    h->seg_groups = valu;
    h->seg_groups.insert(h->seg_groups.end(), mfma.begin(), mfma.end());
    if (h->seg_groups.size() > kSegMaxGroups)
        return;
    // staged history: every window's reach back, rounded to 16 samples
    h->seg_halo = (reach + 15) & ~15u;
    if (h->d_seg_groups.ensure(h->seg_groups.size() * sizeof(SegGroup)) ||
        h->d_seg_btab.ensure(std::max<size_t>(h->seg_btab.size(), 1) * sizeof(float)) ||
        h->d_ptaps.ensure(ptaps.size() * sizeof(float2)))
        return;
    if (hipMemcpy(h->d_seg_groups.p, h->seg_groups.data(), h->seg_groups.size() * sizeof(SegGroup),
                  hipMemcpyHostToDevice) != hipSuccess ||
        (!h->seg_btab.empty() && hipMemcpy(h->d_seg_btab.p, h->seg_btab.data(), h->seg_btab.size() * sizeof(float),
                                           hipMemcpyHostToDevice) != hipSuccess) ||
        hipMemcpy(h->d_ptaps.p, ptaps.data(), ptaps.size() * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess)
        return;
    h->seg_ok = h->use_seg = true;
}

// launch arguments of the segment kernel for a chunk (device_pos: graph mode,
// position and parity from the device record)
SegArgs seg_args(const fdsp_channelizer_s *h, const fdsp_cf32 *d_in, uint32_t n, bool device_pos,
                 const ChanStatic *chans)
{
    SegArgs a{};
    a.in = reinterpret_cast<const float2 *>(d_in);
    a.n = n;
    a.seg = h->seg_len_in;
    a.halo = h->seg_halo;
    a.hist_len = h->hist_len;
    a.hist[0] = h->d_hist[0].as<float2>();
    a.hist[1] = h->d_hist[1].as<float2>();
    a.hist_w[0] = h->d_hist[0].as<float2>();
    a.hist_w[1] = h->d_hist[1].as<float2>();
    a.chans = chans;
    a.ptaps = h->d_ptaps.as<float2>();
    a.btab = h->d_seg_btab.as<float>();
    a.groups = h->d_seg_groups.as<SegGroup>();
    a.n_groups = (uint32_t)h->seg_groups.size();
    a.n_chan = (uint32_t)h->ch.size();
    a.pos = device_pos ? h->d_pos.as<ChanPos>() : nullptr;
    a.p = h->p;
    a.parity = h->parity;
    a.out16 = true;
    for (const ChanStatic &st : h->stat)
        a.out16 = a.out16 && reinterpret_cast<uintptr_t>(st.out) % 16 == 0;
    a.tails = false;
    for (const SegGroup &g : h->seg_groups)
        a.tails = a.tails || g.tail != 0;
    return a;
}

void build_groups(fdsp_channelizer_s *h, const std::vector<float2> &taps)
{
    const uint32_t C = (uint32_t)h->ch.size();
    std::vector<bool> done(C, false);
    h->groups.clear();
    h->btab.clear();
    auto record = [&](uint32_t kind, const ChanHost &k, uint32_t w, const std::vector<uint32_t> &members) {
        uint32_t gw[kGroupWords] = {kind, k.up, k.down, w, (uint32_t)members.size(), 0, 0, 0};
        for (size_t i = 0; i < members.size(); ++i)
            gw[8 + i] = members[i];
        if (kind == kGroupMfma) {
            // B operand of step s, lane l: B[k = 4(s mod S) + l/16][j = l mod 16],
            // S = w/4 steps over the real parts then S over the imaginary parts.
            // Column 2c gives channel c's real output, 2c+1 its imaginary output.
            gw[6] = (uint32_t)h->btab.size();
            const uint32_t S = w / 4;
            for (uint32_t s = 0; s < 2 * S; ++s)
                for (uint32_t l = 0; l < 64; ++l) {
                    const uint32_t t = 4 * (s % S) + l / 16, j = l % 16, c = j / 2;
                    float v = 0.f;
                    if (c < members.size()) {
                        const ChanStatic &st = h->stat[members[c]];
                        const uint32_t pad = w - st.cpp;
                        if (t >= pad) {
                            const float2 g = taps[st.taps_off + t - pad];
                            if (s < S)
                                v = (j & 1) ? g.y : g.x;
                            else
                                v = (j & 1) ? g.x : -g.y;
                        }
                    }
                    h->btab.push_back(v);
                }
            if (mfma_bf16_planes(w, k.down) && h->bf16_planes) {
                // bf16 form: for each alignment base b (the windows start b
                // samples past an 8-aligned slot), K step q and lane l, the 8
                // values B[kk = 32q + 8(l/16) + e][j = l mod 16], e = 0..7, as
                // hi and lo bf16 parts. K slot blocks 0-4 are the real parts of
                // window slots 0..39, blocks 5-9 the imaginary parts, 10-11 zero;
                // slot s holds window tap s - b.
                gw[7] = (uint32_t)h->btab.size() + 1;
                for (uint32_t base = 0; base < 8; ++base)
                    for (uint32_t q = 0; q < 3; ++q)
                        for (uint32_t part = 0; part < 2; ++part)
                            for (uint32_t l = 0; l < 64; ++l)
                                for (uint32_t wd = 0; wd < 4; ++wd) {
                                    uint32_t packed = 0;
                                    for (uint32_t hlf = 0; hlf < 2; ++hlf) {
                                        const uint32_t kb = 4 * q + l / 16, e = 2 * wd + hlf, j = l % 16, c = j / 2;
                                        float v = 0.f;
                                        if (kb < 10 && c < members.size()) {
                                            const bool re = kb < 5;
                                            const int t = (int)(8 * (re ? kb : kb - 5) + e) - (int)base;
                                            const ChanStatic &st = h->stat[members[c]];
                                            const int pad = (int)(w - st.cpp);
                                            if (t >= pad && t < (int)w) {
                                                const float2 g = taps[st.taps_off + t - pad];
                                                v = re ? ((j & 1) ? g.y : g.x) : ((j & 1) ? g.x : -g.y);
                                            }
                                        }
                                        const uint16_t hi = bf16_rne(v);
                                        const uint16_t bits = part == 0 ? hi : bf16_rne(v - bf16_to_float(hi));
                                        packed |= (uint32_t)bits << (16 * hlf);
                                    }
                                    float word;
                                    std::memcpy(&word, &packed, sizeof(word));
                                    h->btab.push_back(word);
                                }
            }
        }
        h->groups.insert(h->groups.end(), gw, gw + kGroupWords);
    };
    for (uint32_t c = 0; c < C; ++c) {
        if (done[c])
            continue;
        const ChanHost &k = h->ch[c];
        const uint32_t w = mfma_window(k);
        std::vector<uint32_t> same;
        if (w)
            for (uint32_t d = c; d < C; ++d)
                if (!done[d] && mfma_window(h->ch[d]) == w && h->ch[d].down == k.down)
                    same.push_back(d);
        const bool bf = h->bf16_planes && w && mfma_bf16_planes(w, k.down);
        if (same.size() >= (bf ? 1u : kMfmaMinChan) && k.down <= kMfmaMaxDown) {
            // near-equal split into ceil(n / 8) groups
            const size_t ng = (same.size() + kMfmaMaxChan - 1) / kMfmaMaxChan;
            for (size_t i = 0, at = 0; i < ng; ++i) {
                const size_t take = (same.size() - at) / (ng - i);
                std::vector<uint32_t> members(same.begin() + at, same.begin() + at + take);
                record(kGroupMfma, k, w, members);
                at += take;
            }
            for (uint32_t d : same)
                done[d] = true;
            continue;
        }
        done[c] = true;
        const uint32_t kind = k.up > 1 ? kGroupGeneric : k.down < k.cpp ? kGroupVDense : kGroupVSparse;
        record(kind, k, k.cpp, {c});
    }
    h->n_groups = (uint32_t)(h->groups.size() / kGroupWords);
}

// Estimated cost of one tile of a group, in GPU-wide cycles: the larger of
// its matrix-core or vector-ALU time spread over all SIMDs and its HBM bytes
// at about 6 TB/s (2500 B per 2.4 GHz cycle).
double tile_cost(const uint32_t *gw, uint32_t block)
{
    const uint32_t kind = gw[0], up = gw[1], down = gw[2], w = gw[3], m = gw[4], rows = gw[5];
    const double simds = 1024.0, bytes_per_cycle = 2500.0;
    double compute, in_samples;
    if (kind == kGroupMfma) {
        // bf16 planes: 9 bf16 steps of 16 cycles per 16 rows; else 2w/4 FP32 steps of 32
        const double steps = gw[7] ? 9.0 * 16.0 : (2.0 * (w / 4)) * 32.0;
        compute = std::ceil(rows / 16.0) * steps + rows / 16.0 * 200.0;
        in_samples = down < w ? (double)rows * down + w : (double)rows * w;
    } else {
        // four FMAs per complex tap per output, 64 lanes per 4-cycle wave instruction, 2x for the rest
        compute = (double)rows * w * up * 4.0 / 64.0 * 4.0 * 2.0 / (up > 1 ? up : 1);
        in_samples = kind == kGroupVSparse ? (double)rows * w : (double)rows * down / up + w;
    }
    const double bytes = 8.0 * in_samples + 8.0 * rows * m;
    (void)block;
    return std::max(compute / simds, bytes / bytes_per_cycle);
}

// A fused record runs on 16-byte aligned input only, and needs a wave per
// residue class of each member
bool fused_ok(const fdsp_channelizer_s *h, bool in16)
{
    if (!h->fuse_on || !in16 || h->sched[1] == h->sched[0])
        return false;
    for (uint32_t g : h->sched[1]) {
        const uint32_t *gw = h->groups.data() + g * kGroupWords;
        if (gw[0] == kGroupFused)
            for (uint32_t i = 0; i < gw[4]; ++i)
                if (mfma_bf16_classes(h->groups[gw[8 + i] * kGroupWords + 2]) > h->block / 64)
                    return false;
    }
    return true;
}

// Schedule the fused record (on) or its members (off): the launch's LDS
// size and the resident-block capacity follow, and the cached block tables
// are dropped. The caller makes sure no queued launch still reads them.
int use_fusion(fdsp_channelizer_s *h, bool on)
{
    h->fuse_on = on;
    h->lds_bytes = h->lds_plan[on ? 1 : 0];
    for (auto &e : h->tables)
        e.second.release();
    h->tables.clear();
    return update_capacity(h);
}

// estimated cost of one tile of group g (a fused tile: its members' work on L inputs)
double group_tile_cost(const fdsp_channelizer_s *h, uint32_t g)
{
    const uint32_t *gw = h->groups.data() + g * kGroupWords;
    if (gw[0] != kGroupFused)
        return tile_cost(gw, h->block);
    double c = 0;
    for (uint32_t i = 0; i < gw[4]; ++i) {
        const uint32_t *mw = h->groups.data() + gw[8 + i] * kGroupWords;
        c += tile_cost(mw, h->block) * gw[2] / ((double)mw[5] * mw[2]);
    }
    return c * 0.8;  // one staging pass for all members
}

// Fuse the bf16-plane groups whose rows per tile of L inputs are whole
// sub-tiles of every residue class (L a multiple of 16 * classes * down,
// at most 1920 so a tile's span fits the staging capacity with its windows):
// the largest such set up to kFusedMax members, largest groups first.
void fuse_groups(fdsp_channelizer_s *h)
{
    std::vector<uint32_t> all(h->n_groups);
    for (uint32_t g = 0; g < h->n_groups; ++g)
        all[g] = g;
    h->sched[0] = h->sched[1] = all;
    if (!h->fuse || h->n_groups + 1 > kGroupsInline)
        return;
    std::vector<uint32_t> cand;
    for (uint32_t g = 0; g < h->n_groups; ++g) {
        const uint32_t *gw = h->groups.data() + g * kGroupWords;
        if (gw[0] == kGroupMfma && gw[7] && gw[3] == 32)
            cand.push_back(g);
    }
    // largest first: most outputs per input
    std::stable_sort(cand.begin(), cand.end(), [&](uint32_t x, uint32_t y) {
        const uint32_t *a = h->groups.data() + x * kGroupWords, *b = h->groups.data() + y * kGroupWords;
        return (double)a[4] / a[2] > (double)b[4] / b[2];
    });
    constexpr uint64_t kMaxL = 1920;
    std::vector<uint32_t> members;
    uint64_t M = 1;
    for (uint32_t g : cand) {
        if (members.size() == kFusedMax)
            break;
        const uint32_t down = h->groups[g * kGroupWords + 2];
        const uint64_t need = 16ull * mfma_bf16_classes(down) * down;
        const uint64_t l = std::lcm(M, need);
        if (l <= kMaxL) {
            M = l;
            members.push_back(g);
        }
    }
    if (members.size() < 2)
        return;
    const uint32_t L = (uint32_t)(kMaxL / M * M);
    uint32_t fw[kGroupWords] = {kGroupFused, 1, L, 32, (uint32_t)members.size(), L, 0, 0};
    for (size_t i = 0; i < members.size(); ++i)
        fw[8 + i] = members[i];
    const uint32_t fg = h->n_groups++;
    h->groups.insert(h->groups.end(), fw, fw + kGroupWords);
    h->sched[1].clear();
    for (uint32_t g = 0; g < fg; ++g)
        if (std::find(members.begin(), members.end(), g) == members.end())
            h->sched[1].push_back(g);
    h->sched[1].push_back(fg);
}

// Work split of a chunk of n inputs (a chunk emits at most floor(n*up/down)+1
// outputs per channel): each block runs a stripe of tps consecutive tiles of
// one group.
//  - FDSP_TILING_STRIDED (default) and FDSP_TILING_CLAIM: the grid is the
//    number of blocks the GPU holds at once; each group gets blocks in
//    proportion to its estimated cost. Strided: the group's tiles are split
//    among its blocks (block s runs s, s + blocks, ...). Claimed: the table
//    only names each block's first group; blocks take tiles one at a time
//    from per-group counters and move on to the other groups when theirs
//    runs out, so they all finish together whatever the cost estimate said.
//  - seg_len > 0: stripes cover about seg_len inputs, dealt to the kXcds
//    queues by the input segment they start in (all groups' stripes over the
//    same input share one XCD's L2), the queues interleaved to follow the
//    round-robin dispatch of workgroups to XCDs.
// table: (group, stripe) per block, kNoGroup where a queue is short.
void block_table(fdsp_channelizer_s *h, uint32_t n, bool in16, std::vector<uint32_t> *table)
{
    h->tps.assign(h->n_groups, 1);
    const std::vector<uint32_t> &sched = h->sched[fused_ok(h, in16) ? 1 : 0];
    struct Item {
        uint64_t start;
        uint32_t g, s;
    };
    std::vector<Item> items;
    std::vector<uint64_t> tiles(h->n_groups), span(h->n_groups);
    double total = 0;
    for (uint32_t g : sched) {
        const uint32_t *gw = h->groups.data() + g * kGroupWords;
        if (gw[0] == kGroupFused) {
            // global tiles of L inputs: the chunk's first and last may be partial
            tiles[g] = n / gw[2] + 2;
            span[g] = gw[2];
        } else {
            const uint64_t most = (uint64_t)n * gw[1] / gw[2] + 1;
            tiles[g] = (most + gw[5] - 1) / gw[5];
            span[g] = std::max<uint64_t>(1, (uint64_t)gw[5] * gw[2] / gw[1]);  // inputs per tile
        }
        total += tiles[g] * (h->gcost.empty() ? group_tile_cost(h, g) : h->gcost[g]);
    }
    for (uint32_t g : sched) {
        uint64_t tps;
        if (h->seg_len <= FDSP_TILING_STRIDED) {
            // strided split: the group's blocks sweep its tiles together, tps = block count
            const double share = tiles[g] * (h->gcost.empty() ? group_tile_cost(h, g) : h->gcost[g]) / total;
            uint64_t blocks = std::min<uint64_t>(tiles[g],
                                                 std::max<uint64_t>(1, (uint64_t)std::llround(share * h->capacity)));
            if (h->only_group >= 0)
                blocks = (int)g == h->only_group ? std::min<uint64_t>(tiles[g], h->only_blocks) : 0;
            h->tps[g] = (uint32_t)blocks;
            for (uint64_t st = 0; st < blocks; ++st)
                items.push_back({st * span[g], g, (uint32_t)st});
            continue;
        } else {
            tps = std::min<uint64_t>(std::max<uint64_t>(1, h->seg_len / span[g]), tiles[g]);
        }
        h->tps[g] = (uint32_t)tps;
        for (uint64_t st = 0; st * tps < tiles[g]; ++st)
            items.push_back({st * tps * span[g], g, (uint32_t)st});
    }
    if (h->seg_len <= FDSP_TILING_STRIDED) {
        // all blocks are resident together: interleave the groups so each
        // XCD gets a share of every group
        std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.start < y.start; });
        table->clear();
        for (const Item &it : items) {
            table->push_back(it.g);
            table->push_back(it.s);
        }
        return;
    }
    std::stable_sort(items.begin(), items.end(), [](const Item &x, const Item &y) { return x.start < y.start; });
    std::vector<std::vector<uint32_t>> q(kXcds);
    for (const Item &it : items) {
        auto &qq = q[(it.start / h->seg_len) % kXcds];
        qq.push_back(it.g);
        qq.push_back(it.s);
    }
    size_t most = 0;
    for (auto &qq : q)
        most = std::max(most, qq.size() / 2);
    table->assign(2 * most * kXcds, 0);
    for (size_t j = 0; j < most; ++j)
        for (uint32_t x = 0; x < kXcds; ++x) {
            uint32_t *e = table->data() + 2 * (j * kXcds + x);
            if (j < q[x].size() / 2) {
                e[0] = q[x][2 * j];
                e[1] = q[x][2 * j + 1];
            } else {
                e[0] = kNoGroup;
            }
        }
}

// device copy of the block table for chunk length n (cached per length)
int table_for(fdsp_channelizer_s *h, uint32_t n, bool in16, bool for_graph, const uint32_t **dev, uint32_t *blocks)
{
    std::vector<uint32_t> t;
    block_table(h, n, in16, &t);
    const uint32_t key = n * 2 + (in16 ? 1 : 0);
    *blocks = (uint32_t)(t.size() / 2);
    DevBuf *buf = nullptr;
    if (for_graph) {
        buf = &h->d_table_graph;
    } else {
        for (auto &e : h->tables)
            if (e.first == key) {
                *dev = e.second.as<uint32_t>();
                return FDSP_OK;
            }
        if (h->tables.size() >= 8) {
            // launches in flight may still read the cached tables
            HIP_TRY(hipDeviceSynchronize());
            for (auto &e : h->tables)
                e.second.release();
            h->tables.clear();
        }
        h->tables.emplace_back(key, DevBuf{});
        buf = &h->tables.back().second;
    }
    if (int rc = buf->ensure(t.size() * sizeof(uint32_t)))
        return rc;
    HIP_TRY(hipMemcpy(buf->p, t.data(), t.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    *dev = buf->as<uint32_t>();
    return FDSP_OK;
}

int groups_args(fdsp_channelizer_s *h, const fdsp_cf32 *d_in, uint32_t n, bool device_pos, GroupsArgs *out)
{
    GroupsArgs f{};
    f.in = reinterpret_cast<const float2 *>(d_in);
    f.n = n;
    f.hist_len = h->hist_len;
    f.hist[0] = h->d_hist[0].as<float2>();
    f.hist[1] = h->d_hist[1].as<float2>();
    f.hist_w[0] = h->d_hist[0].as<float2>();
    f.hist_w[1] = h->d_hist[1].as<float2>();
    f.chans = h->d_stat.as<ChanStatic>();
    f.taps = h->d_taps.as<float2>();
    f.btab = h->d_btab.as<float>();
    f.groups = h->d_groups.as<uint32_t>();
    f.n_groups = h->n_groups;
    f.block = h->block;
    f.lds_bytes = h->lds_bytes;
    f.wide = false;
    for (uint32_t g = 0; g < h->n_groups; ++g)
        f.wide |= h->groups[g * kGroupWords] == kGroupMfma && h->groups[g * kGroupWords + 3] == 64;
    f.pos = device_pos ? h->d_pos.as<ChanPos>() : nullptr;
    f.p = h->p;
    f.parity = h->parity;
    const bool in16 = reinterpret_cast<uintptr_t>(d_in) % 16 == 0;
    if (int rc = table_for(h, n, in16, device_pos, &f.table, &f.blocks))
        return rc;
    f.sched = h->sched[fused_ok(h, in16) ? 1 : 0].data();
    f.n_sched = (uint32_t)h->sched[fused_ok(h, in16) ? 1 : 0].size();
    f.tps = h->tps.data();
    f.strided = h->seg_len <= FDSP_TILING_STRIDED;
    f.claims = h->seg_len == FDSP_TILING_CLAIM ? h->d_claims.as<uint32_t>() : nullptr;
    f.sink = h->d_sink.as<float2>();
    f.out16 = true;
    for (const ChanStatic &st : h->stat)
        f.out16 = f.out16 && reinterpret_cast<uintptr_t>(st.out) % 16 == 0;
    f.in16 = reinterpret_cast<uintptr_t>(d_in) % 16 == 0;
    *out = f;
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_channelizer_destroy(fdsp_channelizer_t h);

extern "C" int fdsp_channelizer_create(uint32_t fs, const fdsp_channel_desc *chans, uint32_t n_chan, int mode,
                                       fdsp_channelizer_t *out)
{
    if (!out || !chans || n_chan == 0 || fs == 0)
        return set_err(FDSP_EINVAL, "channelizer_create: need fs > 0 and at least one channel");
    if (mode != FDSP_CHAN_PLAIN && mode != FDSP_CHAN_BATCHED && mode != FDSP_CHAN_BATCHED_BF16X2)
        return set_err(FDSP_EINVAL, "unknown channelizer mode %d", mode);
    for (uint32_t c = 0; c < n_chan; ++c)
        if (chans[c].up_rate == 0 || chans[c].down_rate == 0 || !chans[c].taps || chans[c].n_taps == 0 ||
            chans[c].tail > FDSP_TAIL_DISCRIMINATOR)
            return set_err(FDSP_EINVAL, "channel %u: need up, down > 0, taps and a known tail", c);
    auto *h = new (std::nothrow) fdsp_channelizer_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->fs = fs;
    h->mode = mode == FDSP_CHAN_PLAIN ? FDSP_CHAN_PLAIN : FDSP_CHAN_BATCHED;
    h->bf16_planes = mode == FDSP_CHAN_BATCHED_BF16X2;
    {
        Fnv64 f;
        f.add(&fs, sizeof fs);
        f.add(&mode, sizeof mode);
        for (uint32_t c = 0; c < n_chan; ++c) {
            const fdsp_channel_desc &d = chans[c];
            const uint32_t w[6] = {d.output_sample_rate_in_sps, (uint32_t)d.freq_shift_in_hz, d.up_rate,
                                   d.down_rate, d.n_taps, d.tail};
            f.add(w, sizeof w);
            f.add(d.taps, (size_t)d.n_taps * sizeof(fdsp_cf32));
        }
        h->tune_hash = f.h;
    }
    h->ch.resize(n_chan);
    int rc = FDSP_OK;
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess)
        rc = set_err(FDSP_EHIP, "stream creation failed");
    uint32_t max_cpp = 1;
    for (uint32_t c = 0; c < n_chan && !rc; ++c) {
        ChanHost &k = h->ch[c];
        k.shift = (int32_t)chans[c].freq_shift_in_hz;  // narrowed as the reference does
        k.out_rate = chans[c].output_sample_rate_in_sps;
        k.up = chans[c].up_rate;
        k.down = chans[c].down_rate;
        k.tail = chans[c].tail;
        k.nco = make_nco(fs, k.shift);
        k.bank = real_bank(chans[c].taps, chans[c].n_taps, k.up, &k.cpp);
        max_cpp = std::max(max_cpp, k.cpp);
        if (k.tail) {
            rc = k.d_prev[0].ensure(sizeof(float2));
            if (!rc)
                rc = k.d_prev[1].ensure(sizeof(float2));
        }
        if (!rc && h->mode == FDSP_CHAN_PLAIN) {
            const size_t hb = (size_t)(k.cpp > 1 ? k.cpp - 1 : 1) * sizeof(float2);
            rc = k.d_bank.ensure(k.bank.size() * sizeof(float));
            if (!rc)
                rc = k.d_hist[0].ensure(hb);
            if (!rc)
                rc = k.d_hist[1].ensure(hb);
            if (!rc && (hipMemcpy(k.d_bank.p, k.bank.data(), k.bank.size() * sizeof(float),
                                  hipMemcpyHostToDevice) != hipSuccess ||
                        hipMemset(k.d_hist[0].p, 0, hb) != hipSuccess ||
                        hipMemset(k.d_hist[1].p, 0, hb) != hipSuccess))
                rc = set_err(FDSP_EHIP, "channel %u: HIP setup failed", c);
        }
    }
    if (!rc && h->mode == FDSP_CHAN_PLAIN) {
        std::vector<NcoChan> nco;
        for (auto &k : h->ch)
            nco.push_back(k.nco);
        rc = h->d_nco.ensure(sizeof(NcoChan) * n_chan);
        if (!rc && hipMemcpy(h->d_nco.p, nco.data(), sizeof(NcoChan) * n_chan, hipMemcpyHostToDevice) != hipSuccess)
            rc = set_err(FDSP_EHIP, "NCO table upload failed");
    }
    if (!rc && h->mode == FDSP_CHAN_BATCHED) {
        // modulated taps of every channel, back to back
        std::vector<float2> taps;
        h->stat.resize(n_chan);
        for (uint32_t c = 0; c < n_chan; ++c) {
            ChanHost &k = h->ch[c];
            ChanStatic &s = h->stat[c];
            s.nu = k.nco.nu;
            s.r = k.nco.r;
            s.inv_r = k.nco.inv_r;
            if (s.nu == 0.0) {
                // no shift: the phase is 0 at every index, so any period
                // gives the same outputs; a long one keeps the kernels'
                // rollover paths (a window straddling an index restart) idle
                s.r = kNoShiftPeriod;
                s.inv_r = 1.0 / (double)kNoShiftPeriod;
            }
            s.inv_up = 1.0 / (double)k.up;
            s.inv_down = 1.0 / (double)k.down;
            s.up = k.up;
            s.down = k.down;
            s.cpp = k.cpp;
            s.taps_off = (uint32_t)taps.size();
            s.tail = k.tail;
            s.wrap = phasor64(s.nu * (double)s.r);
            s.out = nullptr;
            append_modulated_taps(k.bank, k.up, k.cpp, s.nu, &taps);
        }
        h->taps_total = (uint32_t)taps.size();
        // history: every window's reach back (MFMA windows are padded to 32/64),
        // even so staged spans stay 16-byte aligned
        uint32_t max_w = max_cpp;
        for (auto &k : h->ch)
            max_w = std::max(max_w, mfma_window(k));
        if (const char *e = std::getenv("FDSP_FUSE"))
            h->fuse = e[0] != '0';
        build_groups(h, taps);
        fuse_groups(h);
        // bf16-plane tiles start up to 7 samples before their first window
        for (uint32_t g = 0; g < h->n_groups; ++g)
            if (h->groups[g * kGroupWords + 7])
                max_w = std::max(max_w, h->groups[g * kGroupWords + 3] + 8);
        h->hist_len = (max_w + 1) & ~1u;
        if (!h->bf16_planes)
            build_seg_plan(h, taps);
        h->hist_len = std::max(h->hist_len, h->seg_halo);
        if (h->n_groups > kGroupsInline)
            rc = set_err(FDSP_ERANGE, "batched channelizer: %u channel groups exceed %u; use FDSP_CHAN_PLAIN",
                         h->n_groups, kGroupsInline);
        if (!rc)
            rc = size_groups(h->groups, h->block, h->lds_plan);
        if (!rc)
            rc = use_fusion(h, false);
        if (!rc)
            rc = h->d_groups.ensure(h->groups.size() * sizeof(uint32_t));
        if (!rc)
            rc = h->d_btab.ensure(std::max<size_t>(h->btab.size(), 1) * sizeof(float));
        if (!rc && (hipMemcpy(h->d_groups.p, h->groups.data(), h->groups.size() * sizeof(uint32_t),
                              hipMemcpyHostToDevice) != hipSuccess ||
                    (!h->btab.empty() && hipMemcpy(h->d_btab.p, h->btab.data(), h->btab.size() * sizeof(float),
                                                   hipMemcpyHostToDevice) != hipSuccess)))
            rc = set_err(FDSP_EHIP, "work plan upload failed");
        const size_t hb = (size_t)(h->hist_len ? h->hist_len : 1) * sizeof(float2);
        if (!rc)
            rc = h->d_taps.ensure(taps.size() * sizeof(float2));
        if (!rc)
            rc = h->d_stat.ensure(sizeof(ChanStatic) * n_chan);
        if (!rc)
            rc = h->d_hist[0].ensure(hb);
        if (!rc)
            rc = h->d_hist[1].ensure(hb);
        if (!rc)
            rc = h->d_pos.ensure(sizeof(ChanPos));
        if (!rc)
            rc = h->d_claims.ensure(kClaimWords * sizeof(uint32_t));
        if (!rc)
            rc = h->d_sink.ensure(kSinkWords * sizeof(float2));
        if (!rc && (hipMemcpy(h->d_taps.p, taps.data(), taps.size() * sizeof(float2), hipMemcpyHostToDevice) !=
                        hipSuccess ||
                    hipMemset(h->d_claims.p, 0, kClaimWords * sizeof(uint32_t)) != hipSuccess ||
                    hipMemset(h->d_hist[0].p, 0, hb) != hipSuccess || hipMemset(h->d_hist[1].p, 0, hb) != hipSuccess ||
                    hipMemset(h->d_pos.p, 0, sizeof(ChanPos)) != hipSuccess ||
                    hipHostMalloc((void **)&h->pinned_stat, sizeof(ChanStatic) * n_chan, hipHostMallocDefault) !=
                        hipSuccess ||
                    hipEventCreateWithFlags(&h->stat_ev, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&h->use_ev, hipEventDisableTiming) != hipSuccess ||
                    hipEventRecord(h->stat_ev, h->stream) != hipSuccess ||
                    hipEventRecord(h->use_ev, h->stream) != hipSuccess))
            rc = set_err(FDSP_EHIP, "batched channelizer: HIP setup failed");
    }
    if (rc) {
        fdsp_channelizer_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_graph_destroy(fdsp_channelizer_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->graph_exec)
        (void)hipGraphExecDestroy(h->graph_exec);
    if (h->graph)
        (void)hipGraphDestroy(h->graph);
    h->graph_exec = nullptr;
    h->graph = nullptr;
    h->graph_chunk = 0;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_destroy(fdsp_channelizer_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    fdsp_channelizer_graph_destroy(h);
    for (auto &k : h->ch) {
        k.d_bank.release();
        k.d_hist[0].release();
        k.d_hist[1].release();
        k.d_scratch.release();
        k.d_prev[0].release();
        k.d_prev[1].release();
    }
    h->d_nco.release();
    h->d_full.release();
    h->d_stat.release();
    h->d_stat_graph.release();
    h->d_taps.release();
    h->d_hist[0].release();
    h->d_hist[1].release();
    h->d_pos.release();
    h->d_claims.release();
    h->d_sink.release();
    h->d_groups.release();
    h->d_btab.release();
    h->d_table_graph.release();
    for (auto &e : h->tables)
        e.second.release();
    h->d_in.release();
    for (auto &b : h->d_outs)
        b.release();
    if (h->pinned_stat)
        (void)hipHostFree(h->pinned_stat);
    if (h->stat_ev)
        (void)hipEventDestroy(h->stat_ev);
    if (h->use_ev)
        (void)hipEventDestroy(h->use_ev);
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_reset(fdsp_channelizer_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    // launches on caller streams may still read the history being cleared
    HIP_TRY(hipDeviceSynchronize());
    h->p = 0;
    h->parity = 0;
    h->pos_dirty = false;
    for (auto &k : h->ch) {
        k.has_prev = false;
        k.prev_parity = 0;
        if (h->mode == FDSP_CHAN_PLAIN)
            HIP_TRY(hipMemsetAsync(k.d_hist[0].p, 0, k.d_hist[0].bytes, h->stream));
    }
    if (h->mode == FDSP_CHAN_BATCHED) {
        HIP_TRY(hipMemsetAsync(h->d_hist[0].p, 0, h->d_hist[0].bytes, h->stream));
        HIP_TRY(hipMemsetAsync(h->d_hist[1].p, 0, h->d_hist[1].bytes, h->stream));
        HIP_TRY(hipMemsetAsync(h->d_pos.p, 0, sizeof(ChanPos), h->stream));
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_num_channels(fdsp_channelizer_t h, uint32_t *n_chan)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!n_chan)
        return set_err(FDSP_EINVAL, "null output");
    *n_chan = (uint32_t)h->ch.size();
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_needed_out_count(fdsp_channelizer_t h, uint32_t chan, uint32_t in_count,
                                                 uint32_t *out_count)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (chan >= h->ch.size() || !out_count)
        return set_err(FDSP_EINVAL, "bad channel");
    const ChanHost &k = h->ch[chan];
    uint32_t cnt = chunk_count(h->p, in_count, k.up, k.down);
    // the stream's first output (emitted with its first input) has no
    // predecessor for the discriminator
    if (k.tail && cnt && h->p == 0)
        --cnt;
    *out_count = cnt;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_get_state(fdsp_channelizer_t h, uint32_t chan, uint64_t *nco_samples_handled,
                                          uint32_t *coeff_phase, uint32_t *x_offset, uint64_t *in_total,
                                          uint64_t *out_total)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (chan >= h->ch.size())
        return set_err(FDSP_EINVAL, "bad channel");
    const ChanHost &k = h->ch[chan];
    uint32_t ph, xo;
    uint64_t nn;
    resampler_fields(h->p, k.up, k.down, &ph, &xo, &nn);
    if (nco_samples_handled)
        *nco_samples_handled = k.nco.nu == 0.0 ? 0 : h->p % k.nco.r;
    if (coeff_phase)
        *coeff_phase = ph;
    if (x_offset)
        *x_offset = xo;
    if (in_total)
        *in_total = h->p;
    if (out_total)
        *out_total = nn;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_apply_device(fdsp_channelizer_t h, const fdsp_cf32 *d_in, size_t n,
                                             void *const *d_outs, const size_t *caps, size_t *counts,
                                             fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!d_outs || !caps || !counts)
        return set_err(FDSP_EINVAL, "null output arrays");
    const size_t C = h->ch.size();
    for (size_t c = 0; c < C; ++c)
        counts[c] = 0;
    if (n == 0)
        return FDSP_OK;
    if (!d_in)
        return set_err(FDSP_EINVAL, "null input");
    uint32_t max_up = 1;
    for (auto &k : h->ch)
        max_up = std::max(max_up, k.up);
    if (int rc = check_chunk(n, max_up))
        return rc;
    hipStream_t s = pick(h->stream, stream);
    // exact per-channel counts, checked against the capacities up front
    std::vector<uint32_t> raw(C);
    for (size_t c = 0; c < C; ++c) {
        const ChanHost &k = h->ch[c];
        raw[c] = chunk_count(h->p, n, k.up, k.down);
        size_t emit = raw[c];
        // the stream's first output has no predecessor for the discriminator
        if (k.tail && emit && h->p == 0)
            --emit;
        if (emit > caps[c] || (emit && !d_outs[c]))
            return set_err(FDSP_ERANGE, "channel %zu: capacity %zu below the %zu samples this chunk emits", c,
                           caps[c], emit);
        counts[c] = emit;
    }
    for (size_t c = 0; c < C; ++c)
        if (h->ch[c].tail && !(h->mode == FDSP_CHAN_BATCHED && h->use_seg))
            if (int rc = h->ch[c].d_scratch.ensure((size_t)(raw[c] ? raw[c] : 1) * sizeof(float2)))
                return rc;

    if (h->mode == FDSP_CHAN_PLAIN) {
        // reference structure: one multi-shift NCO pass writing a full-rate copy
        // per channel, then one resampler launch per channel
        if (int rc = h->d_full.ensure(n * C * sizeof(float2)))
            return rc;
        HIP_TRY(launch_nco_multi(reinterpret_cast<const float2 *>(d_in), (uint32_t)n, h->p, h->d_nco.as<NcoChan>(),
                                 (uint32_t)C, h->d_full.as<float2>(), n, s));
        for (size_t c = 0; c < C; ++c) {
            ChanHost &k = h->ch[c];
            float2 *dst = k.tail ? k.d_scratch.as<float2>() : static_cast<float2 *>(d_outs[c]);
            HIP_TRY(launch_resampler(h->d_full.as<float2>() + c * n, (uint32_t)n, h->p, k.up, k.down, k.cpp,
                                     k.d_bank.as<float>(), k.d_hist[h->parity].as<float2>(),
                                     k.d_hist[h->parity ^ 1].as<float2>(), dst, raw[c], s));
        }
        h->parity ^= 1;
    } else {
        if (s != h->stream) {
            // the channel table upload runs on the handle's stream; order it before `s`
            if (int rc = chan_upload_outs(h, d_outs))
                return rc;
            HIP_TRY(hipStreamWaitEvent(s, h->stat_ev, 0));
        } else if (int rc = chan_upload_outs(h, d_outs)) {
            return rc;
        }
        // eager form: position and parity travel as kernel arguments, one launch
        if (h->use_seg) {
            HIP_TRY(launch_chan_seg(seg_args(h, d_in, (uint32_t)n, false, h->d_stat.as<ChanStatic>()), s));
        } else {
            GroupsArgs f;
            if (int rc = groups_args(h, d_in, (uint32_t)n, false, &f))
                return rc;
            HIP_TRY(launch_chan_groups(f, s));
        }
        HIP_TRY(hipEventRecord(h->use_ev, s));
        h->parity ^= 1;
        h->pos_dirty = true;
    }
    // discriminator tails (fused into the segment kernel)
    for (size_t c = 0; c < C; ++c) {
        ChanHost &k = h->ch[c];
        if (!k.tail || raw[c] == 0 || (h->mode == FDSP_CHAN_BATCHED && h->use_seg))
            continue;
        HIP_TRY(launch_discriminator(k.d_scratch.as<float2>(), raw[c], k.d_prev[k.prev_parity].as<float2>(),
                                     k.has_prev ? 1 : 0, k.d_prev[k.prev_parity ^ 1].as<float2>(),
                                     static_cast<float *>(d_outs[c]), s));
        k.has_prev = true;
        k.prev_parity ^= 1;
    }
    h->p += n;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_apply(fdsp_channelizer_t h, const fdsp_cf32 *in, size_t n, void *const *outs,
                                      const size_t *caps, size_t *counts)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!outs || !caps || !counts)
        return set_err(FDSP_EINVAL, "null output arrays");
    const size_t C = h->ch.size();
    if (n == 0) {
        for (size_t c = 0; c < C; ++c)
            counts[c] = 0;
        return FDSP_OK;
    }
    if (!in)
        return set_err(FDSP_EINVAL, "null input");
    if (int rc = h->d_in.ensure(n * sizeof(float2)))
        return rc;
    h->d_outs.resize(C);
    std::vector<void *> dptr(C);
    std::vector<size_t> dcap(C);
    for (size_t c = 0; c < C; ++c) {
        const ChanHost &k = h->ch[c];
        const size_t need = chunk_count(h->p, n, k.up, k.down) + 1;
        if (int rc = h->d_outs[c].ensure(need * sizeof(float2)))
            return rc;
        dptr[c] = h->d_outs[c].p;
        dcap[c] = need;
    }
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, n * sizeof(float2), hipMemcpyHostToDevice, h->stream));
    std::vector<size_t> got(C);
    // validate the caller's capacities first so a failure leaves the state untouched
    for (size_t c = 0; c < C; ++c) {
        uint32_t cnt = 0;
        fdsp_channelizer_needed_out_count(h, (uint32_t)c, (uint32_t)n, &cnt);
        if (cnt > caps[c] || (cnt && !outs[c]))
            return set_err(FDSP_ERANGE, "channel %zu: capacity %zu below the %u samples this chunk emits", c,
                           caps[c], cnt);
    }
    if (int rc = fdsp_channelizer_apply_device(h, reinterpret_cast<const fdsp_cf32 *>(h->d_in.p), n, dptr.data(),
                                               dcap.data(), got.data(), h->stream))
        return rc;
    for (size_t c = 0; c < C; ++c) {
        const size_t bytes = got[c] * (h->ch[c].tail ? sizeof(float) : sizeof(float2));
        if (bytes)
            HIP_TRY(hipMemcpyAsync(outs[c], dptr[c], bytes, hipMemcpyDeviceToHost, h->stream));
        counts[c] = got[c];
    }
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

// ---------------------------------------------------------------------------
// graph mode
// ---------------------------------------------------------------------------
extern "C" int fdsp_channelizer_graph_create(fdsp_channelizer_t h, size_t chunk_len, const fdsp_cf32 *d_in,
                                             void *const *d_outs, const size_t *caps)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED)
        return set_err(FDSP_EINVAL, "graph mode needs FDSP_CHAN_BATCHED");
    if (!d_in || !d_outs || !caps || chunk_len == 0)
        return set_err(FDSP_EINVAL, "null buffers or empty chunk");
    const size_t C = h->ch.size();
    uint32_t max_up = 1;
    for (size_t c = 0; c < C; ++c) {
        const ChanHost &k = h->ch[c];
        if (k.tail && !h->use_seg)
            return set_err(FDSP_EINVAL, "graph mode takes discriminator-tail channels on the segment kernel only");
        max_up = std::max(max_up, k.up);
        // any chunk of chunk_len inputs emits at most ceil(chunk_len*up/down) outputs
        const uint64_t most = ceil_div((uint64_t)chunk_len * k.up, k.down);
        if (caps[c] < most || !d_outs[c])
            return set_err(FDSP_ERANGE, "channel %zu: graph capacity %zu below %llu", c, caps[c],
                           (unsigned long long)most);
    }
    if (int rc = check_chunk(chunk_len, max_up))
        return rc;
    fdsp_channelizer_graph_destroy(h);
    // the graph keeps its own channel table (its output buffers are fixed)
    std::vector<ChanStatic> gstat = h->stat;
    for (size_t c = 0; c < C; ++c)
        gstat[c].out = d_outs[c];
    if (int rc = h->d_stat_graph.ensure(sizeof(ChanStatic) * C))
        return rc;
    HIP_TRY(hipStreamSynchronize(h->stream));
    HIP_TRY(hipMemcpy(h->d_stat_graph.p, gstat.data(), sizeof(ChanStatic) * C, hipMemcpyHostToDevice));
    GroupsArgs f{};
    SegArgs sa{};
    bool out16 = true;
    for (const ChanStatic &st : gstat)
        out16 = out16 && reinterpret_cast<uintptr_t>(st.out) % 16 == 0;
    if (h->use_seg) {
        sa = seg_args(h, d_in, (uint32_t)chunk_len, true, h->d_stat_graph.as<ChanStatic>());
        sa.out16 = out16;
    } else {
        if (int rc = groups_args(h, d_in, (uint32_t)chunk_len, true, &f))
            return rc;
        f.chans = h->d_stat_graph.as<ChanStatic>();
        f.out16 = out16;
    }
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e1 = h->use_seg ? launch_chan_seg(sa, h->stream) : launch_chan_groups(f, h->stream);
    hipError_t e2 = launch_chan_advance(h->d_pos.as<ChanPos>(), (uint32_t)chunk_len, h->stream);
    hipGraph_t g = nullptr;
    hipError_t e3 = hipStreamEndCapture(h->stream, &g);
    HIP_TRY(e1);
    HIP_TRY(e2);
    HIP_TRY(e3);
    h->graph = g;
    HIP_TRY(hipGraphInstantiate(&h->graph_exec, g, nullptr, nullptr, 0));
    h->graph_chunk = chunk_len;
    h->graph_counts.assign(C, 0);
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_graph_launch(fdsp_channelizer_t h, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!h->graph_exec)
        return set_err(FDSP_ENOTINIT, "no graph: call fdsp_channelizer_graph_create first");
    hipStream_t s = pick(h->stream, stream);
    if (h->pos_dirty) {
        // eager calls advanced the host position only: bring the device copy along
        const ChanPos pos{h->p, h->parity, 0};
        HIP_TRY(hipDeviceSynchronize());
        HIP_TRY(hipMemcpy(h->d_pos.p, &pos, sizeof(pos), hipMemcpyHostToDevice));
        h->pos_dirty = false;
    }
    for (size_t c = 0; c < h->ch.size(); ++c) {
        size_t cnt = chunk_count(h->p, h->graph_chunk, h->ch[c].up, h->ch[c].down);
        if (h->ch[c].tail && cnt && h->p == 0)
            --cnt;  // the stream's first output has no predecessor
        h->graph_counts[c] = cnt;
    }
    HIP_TRY(hipGraphLaunch(h->graph_exec, s));
    h->p += h->graph_chunk;
    h->parity ^= 1;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_graph_counts(fdsp_channelizer_t h, size_t *counts)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (!counts)
        return set_err(FDSP_EINVAL, "null counts");
    for (size_t c = 0; c < h->graph_counts.size(); ++c)
        counts[c] = h->graph_counts[c];
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_set_tiling(fdsp_channelizer_t h, uint32_t segment_len, uint32_t block_threads)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED)
        return set_err(FDSP_EINVAL, "tiling applies to the batched mode");
    if (h->use_seg) {
        // segment kernel: the segment length is the only parameter
        if (!chan_seg_len_ok(segment_len))
            return set_err(FDSP_EINVAL, "the segment kernel takes segment_len a multiple of 256 from 1024 to 4096");
        if (h->graph_exec)
            return set_err(FDSP_EBUSY, "destroy the graph before retiling");
        h->seg_len_in = segment_len;
        return FDSP_OK;
    }
    if (segment_len > FDSP_TILING_STRIDED &&
        (segment_len < 256 || segment_len > (1u << 20) || (segment_len & (segment_len - 1))))
        return set_err(FDSP_EINVAL, "segment_len must be FDSP_TILING_CLAIM, FDSP_TILING_STRIDED or a power of two in "
                                    "[256, 2^20]");
    if (block_threads != 128 && block_threads != 256 && block_threads != 512)
        return set_err(FDSP_EINVAL, "block_threads must be 128, 256 or 512");
    if (h->graph_exec)
        return set_err(FDSP_EBUSY, "destroy the graph before retiling");
    std::vector<uint32_t> groups = h->groups;
    uint32_t lds[2] = {0, 0};
    if (int rc = size_groups(groups, block_threads, lds))
        return rc;
    // launches already queued read the device plan: let them finish first
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemcpy(h->d_groups.p, groups.data(), groups.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    h->groups.swap(groups);
    h->lds_plan[0] = lds[0];
    h->lds_plan[1] = lds[1];
    h->seg_len = segment_len;
    h->block = block_threads;
    return use_fusion(h, h->fuse_on);
}

extern "C" int fdsp_channelizer_plan_info(fdsp_channelizer_t h, size_t chunk_len, uint32_t *blocks,
                                          uint32_t *capacity, uint32_t *groups, uint32_t *lds_bytes)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED)
        return set_err(FDSP_EINVAL, "plan info applies to the batched mode");
    if (h->use_seg) {
        const SegGeom geo = chan_seg_geometry((uint32_t)chunk_len, h->seg_len_in, h->seg_halo, (uint32_t)h->ch.size());
        if (blocks)
            *blocks = chunk_len ? geo.blocks : 0u;
        if (capacity)
            *capacity = h->capacity;
        if (groups)
            *groups = (uint32_t)h->seg_groups.size();
        if (lds_bytes)
            *lds_bytes = (uint32_t)geo.lds;
        return FDSP_OK;
    }
    std::vector<uint32_t> t;
    block_table(h, (uint32_t)chunk_len, true, &t);
    if (blocks)
        *blocks = (uint32_t)(t.size() / 2);
    if (capacity)
        *capacity = h->capacity;
    if (groups)
        *groups = h->n_groups;
    if (lds_bytes)
        *lds_bytes = h->lds_bytes;
    return FDSP_OK;
}

// ---------------------------------------------------------------------------
// tuning cache: one line per tuned setting,
//   v1 <key> = <segment_len> <block_threads> <fused> <us> [<group weight> ...]
// the last line of a key wins. The file: FDSP_TUNE_CACHE when set ("" or
// "off": no cache), else $XDG_CACHE_HOME/fdsp/tuning.txt or
// $HOME/.cache/fdsp/tuning.txt. A cache that cannot be read or written is
// skipped silently (the trials then run every time).
// ---------------------------------------------------------------------------
struct TuneRecord {
    uint32_t seg = 0, block = 256, fused = 0;
    float us = 0.f;
    std::vector<double> weights;
};

static std::string tune_cache_path(bool create_dirs)
{
    if (const char *e = std::getenv("FDSP_TUNE_CACHE"))
        return std::strcmp(e, "off") == 0 ? std::string() : std::string(e);
    std::string dir;
    if (const char *x = std::getenv("XDG_CACHE_HOME"); x && *x)
        dir = x;
    else if (const char *home = std::getenv("HOME"); home && *home)
        dir = std::string(home) + "/.cache";
    else
        return std::string();
    if (create_dirs) {
        (void)mkdir(dir.c_str(), 0755);
        (void)mkdir((dir + "/fdsp").c_str(), 0755);
    }
    return dir + "/fdsp/tuning.txt";
}

static int tune_key(const fdsp_channelizer_s *h, size_t chunk_len, bool in16, std::string *key)
{
    int dev = 0, cus = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, dev));
    char arch[64];
    std::snprintf(arch, sizeof arch, "%s", prop.gcnArchName);
    for (char *c = arch; *c; ++c)
        if (*c == ' ' || *c == '=' || *c == ':')
            *c = '_';
    char buf[256];
    std::snprintf(buf, sizeof buf, "%016llx %s cu%d n%zu a%d %s", (unsigned long long)h->tune_hash, arch, cus,
                  chunk_len, in16 ? 1 : 0, h->use_seg ? "seg" : h->bf16_planes ? "groups-bf16x2" : "groups");
    *key = buf;
    return FDSP_OK;
}

static bool tune_cache_get(const std::string &key, TuneRecord *out)
{
    const std::string path = tune_cache_path(false);
    if (path.empty())
        return false;
    FILE *f = std::fopen(path.c_str(), "r");
    if (!f)
        return false;
    const std::string head = "v1 " + key + " = ";
    char line[4096];
    bool found = false;
    while (std::fgets(line, sizeof line, f)) {
        if (std::strncmp(line, head.c_str(), head.size()) != 0)
            continue;
        TuneRecord r;
        const char *p = line + head.size();
        char *end = nullptr;
        const unsigned long seg = std::strtoul(p, &end, 10);
        if (end == p)
            continue;
        p = end;
        const unsigned long block = std::strtoul(p, &end, 10);
        if (end == p)
            continue;
        p = end;
        const unsigned long fused = std::strtoul(p, &end, 10);
        if (end == p)
            continue;
        p = end;
        r.us = std::strtof(p, &end);
        if (end == p)
            continue;
        p = end;
        for (;;) {
            const double w = std::strtod(p, &end);
            if (end == p)
                break;
            r.weights.push_back(w);
            p = end;
        }
        r.seg = (uint32_t)seg;
        r.block = (uint32_t)block;
        r.fused = (uint32_t)fused;
        *out = r;
        found = true;  // later lines override earlier ones
    }
    std::fclose(f);
    return found;
}

static void tune_cache_put(const std::string &key, const TuneRecord &r)
{
    const std::string path = tune_cache_path(true);
    if (path.empty())
        return;
    FILE *f = std::fopen(path.c_str(), "a");
    if (!f)
        return;
    std::string line = "v1 " + key + " = " + std::to_string(r.seg) + " " + std::to_string(r.block) + " " +
                       std::to_string(r.fused);
    char num[64];
    std::snprintf(num, sizeof num, " %.3f", (double)r.us);
    line += num;
    for (double w : r.weights) {
        std::snprintf(num, sizeof num, " %.17g", w);
        line += num;
    }
    line += "\n";
    // one write of the whole line: appends from concurrent processes do not interleave
    (void)std::fwrite(line.data(), 1, line.size(), f);
    std::fclose(f);
}

// The autotuner's trials for chunk_len on d_in; the fastest setting is
// applied, and its time returned in *best_us.
static int autotune_run(fdsp_channelizer_t h, const fdsp_cf32 *d_in, size_t chunk_len, uint32_t reps,
                        fdsp_tiling_trial *trials, size_t cap, size_t *n_trials, float *best_us)
{
    *best_us = -1.f;
    // the splits are timed unfused; fusion is a trial of its own at the end
    HIP_TRY(hipDeviceSynchronize());
    if (int rc = use_fusion(h, false))
        return rc;
    const uint32_t n = (uint32_t)chunk_len;
    // scratch outputs: the stream position (p, parity) is passed to every
    // launch unchanged and never advanced, so the tuning launches leave the
    // channelizer's state as it was
    const size_t C = h->ch.size();
    std::vector<DevBuf> outs(C);
    std::vector<void *> ptrs(C);
    for (size_t c = 0; c < C; ++c) {
        ChanHost &k = h->ch[c];
        const size_t most = (size_t)chunk_len * k.up / k.down + 1;
        DevBuf &dst = k.tail ? k.d_scratch : outs[c];
        if (int rc = dst.ensure(most * sizeof(float2)))
            return rc;
        ptrs[c] = dst.p;
    }
    if (int rc = chan_upload_outs(h, ptrs.data()))
        return rc;
    if (h->use_seg) {
        // segment kernel: time each segment length, keep the fastest
        hipEvent_t t0 = nullptr, t1 = nullptr;
        HIP_TRY(hipEventCreate(&t0));
        if (hipEventCreate(&t1) != hipSuccess) {
            (void)hipEventDestroy(t0);
            return set_err(FDSP_EHIP, "event creation failed");
        }
        static const uint32_t kSegCand[] = {2560, 3072, 2048, 3584, 4096, 1024};
        const uint32_t seg0 = h->seg_len_in;
        float best = -1.f;
        uint32_t best_seg = seg0;
        size_t k = 0;
        int rc = FDSP_OK;
        for (uint32_t L : kSegCand) {
            h->seg_len_in = L;
            const SegArgs sa = seg_args(h, d_in, n, false, h->d_stat.as<ChanStatic>());
            float ms = 0.f;
            bool ok = launch_chan_seg(sa, h->stream) == hipSuccess && hipEventRecord(t0, h->stream) == hipSuccess;
            for (uint32_t r = 0; ok && r < reps; ++r)
                ok = launch_chan_seg(sa, h->stream) == hipSuccess;
            ok = ok && hipEventRecord(t1, h->stream) == hipSuccess && hipEventSynchronize(t1) == hipSuccess &&
                 hipEventElapsedTime(&ms, t0, t1) == hipSuccess;
            if (!ok) {
                rc = set_err(FDSP_EHIP, "autotune launch failed");
                break;
            }
            const float us = 1e3f * ms / (float)reps;
            if (best < 0.f || us < best) {
                best = us;
                best_seg = L;
            }
            if (trials && k < cap)
                trials[k] = fdsp_tiling_trial{L, 256u, us, 4u};
            ++k;
        }
        (void)hipStreamSynchronize(h->stream);
        (void)hipEventDestroy(t0);
        (void)hipEventDestroy(t1);
        h->stat_valid = false;
        h->seg_len_in = rc ? seg0 : best_seg;
        *best_us = best;
        if (n_trials)
            *n_trials = k;
        return rc;
    }
    // candidates: {split, threads, group weights measured (1) or estimated (0)}
    static const uint32_t kCand[][3] = {
        {FDSP_TILING_STRIDED, 256, 0}, {FDSP_TILING_STRIDED, 512, 0}, {FDSP_TILING_STRIDED, 128, 0},
        {FDSP_TILING_CLAIM, 256, 0},   {FDSP_TILING_CLAIM, 512, 0},   {FDSP_TILING_CLAIM, 128, 0},
        {4096, 256, 0},                {8192, 256, 0},                {16384, 256, 0},
        {32768, 256, 0},               {65536, 256, 0},               {16384, 512, 0},
        {FDSP_TILING_STRIDED, 256, 1}, {FDSP_TILING_STRIDED, 512, 1}, {FDSP_TILING_CLAIM, 256, 1},
        {FDSP_TILING_CLAIM, 512, 1},
    };
    const uint32_t seg0 = h->seg_len, block0 = h->block;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    HIP_TRY(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return set_err(FDSP_EHIP, "event creation failed");
    }
    int rc = FDSP_OK;
    // average time of reps back-to-back launches of the current plan
    auto time_plan = [&](float *us) -> int {
        for (auto &e : h->tables)
            e.second.release();
        h->tables.clear();
        GroupsArgs f;
        if (int r = groups_args(h, d_in, n, false, &f))
            return r;
        if (launch_chan_groups(f, h->stream) != hipSuccess || hipEventRecord(e0, h->stream) != hipSuccess)
            return set_err(FDSP_EHIP, "autotune launch failed");
        for (uint32_t r = 0; r < reps; ++r)
            if (launch_chan_groups(f, h->stream) != hipSuccess)
                return set_err(FDSP_EHIP, "autotune launch failed");
        float ms = 0.f;
        if (hipEventRecord(e1, h->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
            return set_err(FDSP_EHIP, "autotune timing failed");
        *us = 1e3f * ms / (float)reps;
        return FDSP_OK;
    };
    // Group weights: each group alone (strided split, 256 threads) on every
    // block the GPU holds and on half of them. The difference of the two times
    // is the group's work per block at full occupancy, free of the launch's
    // fixed costs; the splits below give each group blocks in proportion to it.
    h->gcost.clear();
    std::vector<double> measured;
    if (h->n_groups > 1 && (rc = fdsp_channelizer_set_tiling(h, FDSP_TILING_STRIDED, 256)) == FDSP_OK) {
        std::vector<double> work(h->n_groups, 0.0);
        double most = 0;
        // the groups a launch on this input schedules
        const std::vector<uint32_t> sched =
            h->sched[fused_ok(h, reinterpret_cast<uintptr_t>(d_in) % 16 == 0) ? 1 : 0];
        for (uint32_t g : sched) {
            if (rc)
                break;
            h->only_group = (int)g;
            float full = 0.f, half = 0.f;
            h->only_blocks = h->capacity;
            rc = time_plan(&full);
            h->only_blocks = std::max(1u, h->capacity / 2);
            if (!rc)
                rc = time_plan(&half);
            work[g] = std::max(0.0, (double)half - (double)full);
            most = std::max(most, work[g]);
        }
        h->only_group = -1;
        if (!rc && most > 0) {
            measured.assign(h->n_groups, 0.0);
            for (uint32_t g : sched) {
                const uint32_t *gw = h->groups.data() + g * kGroupWords;
                const uint64_t tiles = gw[0] == kGroupFused ? n / gw[2] + 2
                                                            : ((uint64_t)n * gw[1] / gw[2] + 1 + gw[5] - 1) / gw[5];
                measured[g] = std::max(work[g], 0.02 * most) / (double)std::max<uint64_t>(1, tiles);
            }
        }
    }
    float best = -1.f;
    uint32_t best_seg = seg0, best_block = block0;
    bool best_measured = false;
    size_t k = 0;
    for (const auto &cand : kCand) {
        float us = -1.f;
        if (cand[2] && measured.empty())
            continue;  // one group, or no weights: nothing to compare
        if (cand[2])
            h->gcost = measured;
        else
            h->gcost.clear();
        if (!rc && fdsp_channelizer_set_tiling(h, cand[0], cand[1]) == FDSP_OK) {
            // one untimed launch (table upload, caches), then reps back to back
            if ((rc = time_plan(&us)) != FDSP_OK)
                break;
            if (best < 0.f || us < best) {
                best = us;
                best_seg = cand[0];
                best_block = cand[1];
                best_measured = cand[2] != 0;
            }
        }
        if (trials && k < cap)
            trials[k] = fdsp_tiling_trial{cand[0], cand[1], us, cand[2]};
        ++k;
    }
    if (best_measured && !rc)
        h->gcost = measured;
    else
        h->gcost.clear();
    // phase grouping: the best split again with the bf16-plane groups fused
    // into shared tiles, kept if faster
    bool best_fused = false;
    if (!rc && h->sched[1] != h->sched[0] && reinterpret_cast<uintptr_t>(d_in) % 16 == 0 &&
        fdsp_channelizer_set_tiling(h, best_seg, best_block) == FDSP_OK) {
        float us = -1.f;
        if ((rc = use_fusion(h, true)) == FDSP_OK && fused_ok(h, true)) {
            rc = time_plan(&us);
            if (!rc && us < best) {
                best = us;
                best_fused = true;
            }
        }
        (void)hipStreamSynchronize(h->stream);
        if (int r = use_fusion(h, false))
            rc = rc ? rc : r;
        if (trials && k < cap)
            trials[k] = fdsp_tiling_trial{best_seg, best_block, us, (best_measured ? 1u : 0u) | 2u};
        ++k;
    }
    (void)hipStreamSynchronize(h->stream);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // the device table points at the scratch outputs: the next call re-uploads its own
    h->stat_valid = false;
    h->fuse_on = !rc && best_fused;
    const int rc2 = fdsp_channelizer_set_tiling(h, rc ? seg0 : best_seg, rc ? block0 : best_block);
    *best_us = best;
    if (n_trials)
        *n_trials = k;
    return rc ? rc : rc2;
}

extern "C" int fdsp_channelizer_autotune(fdsp_channelizer_t h, const fdsp_cf32 *d_in, size_t chunk_len,
                                         uint32_t reps, fdsp_tiling_trial *trials, size_t cap, size_t *n_trials)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED)
        return set_err(FDSP_EINVAL, "autotuning applies to the batched mode");
    if (!d_in || chunk_len == 0 || reps == 0)
        return set_err(FDSP_EINVAL, "autotune: need a device chunk and reps > 0");
    if (h->graph_exec)
        return set_err(FDSP_EBUSY, "destroy the graph before tuning");
    uint32_t max_up = 1;
    for (auto &k : h->ch)
        max_up = std::max(max_up, k.up);
    if (int rc = check_chunk(chunk_len, max_up))
        return rc;
    // a setting tuned before for this channel set, chunk length, input
    // alignment, kernel and device is taken from the tuning cache
    std::string key;
    if (int rc = tune_key(h, chunk_len, reinterpret_cast<uintptr_t>(d_in) % 16 == 0, &key))
        return rc;
    TuneRecord rec;
    if (tune_cache_get(key, &rec)) {
        HIP_TRY(hipDeviceSynchronize());
        if (h->use_seg) {
            h->seg_len_in = rec.seg;
            h->stat_valid = false;
        } else {
            h->gcost = rec.weights.size() == h->n_groups ? rec.weights : std::vector<double>();
            h->fuse_on = rec.fused != 0;
            if (int rc = fdsp_channelizer_set_tiling(h, rec.seg, rec.block))
                return rc;
        }
        if (trials && cap)
            trials[0] = fdsp_tiling_trial{rec.seg, rec.block, rec.us,
                                          (rec.weights.empty() ? 0u : 1u) | (rec.fused ? 2u : 0u) |
                                              (h->use_seg ? 4u : 0u) | 8u};
        if (n_trials)
            *n_trials = 1;
        return FDSP_OK;
    }
    float best = -1.f;
    const int rc = autotune_run(h, d_in, chunk_len, reps, trials, cap, n_trials, &best);
    if (!rc && best >= 0.f) {
        rec.seg = h->use_seg ? h->seg_len_in : h->seg_len;
        rec.block = h->use_seg ? 256u : h->block;
        rec.fused = h->fuse_on ? 1u : 0u;
        rec.us = best;
        rec.weights = h->use_seg ? std::vector<double>() : h->gcost;
        tune_cache_put(key, rec);
    }
    return rc;
}

extern "C" int fdsp_channelizer_set_fusion(fdsp_channelizer_t h, int on)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED)
        return set_err(FDSP_EINVAL, "fusion applies to the batched mode");
    if (h->graph_exec)
        return set_err(FDSP_EBUSY, "destroy the graph before changing the plan");
    // launches already queued read the current tables
    HIP_TRY(hipDeviceSynchronize());
    return use_fusion(h, on != 0);
}

extern "C" int fdsp_channelizer_get_fusion(fdsp_channelizer_t h, int *on, uint32_t *fusable_groups)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (on)
        *on = h->fuse_on ? 1 : 0;
    if (fusable_groups) {
        *fusable_groups = 0;
        for (uint32_t g : h->sched[1])
            if (h->groups[g * kGroupWords] == kGroupFused)
                *fusable_groups = h->groups[g * kGroupWords + 4];
    }
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_get_tiling(fdsp_channelizer_t h, uint32_t *segment_len, uint32_t *block_threads)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (segment_len)
        *segment_len = h->use_seg ? h->seg_len_in : h->seg_len;
    if (block_threads)
        *block_threads = h->use_seg ? 256u : h->block;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_set_kernel(fdsp_channelizer_t h, uint32_t kernel)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (h->mode != FDSP_CHAN_BATCHED || h->bf16_planes)
        return set_err(FDSP_EINVAL, "set_kernel applies to the FP32 batched mode");
    if (kernel != FDSP_KERNEL_SEGMENT_F32 && kernel != FDSP_KERNEL_GROUPS_F32)
        return set_err(FDSP_EINVAL, "kernel must be FDSP_KERNEL_SEGMENT_F32 or FDSP_KERNEL_GROUPS_F32");
    if (kernel == FDSP_KERNEL_SEGMENT_F32 && !h->seg_ok)
        return set_err(FDSP_EINVAL, "the segment kernel does not fit this channel set (a window above 64 taps)");
    if (h->graph_exec)
        return set_err(FDSP_EBUSY, "destroy the graph before switching kernels");
    for (auto &k : h->ch)
        if (k.tail && h->p != 0)
            return set_err(FDSP_EBUSY, "discriminator tails: switch kernels before the first chunk or after a reset");
    // launches already queued on caller streams use the current tables
    HIP_TRY(hipDeviceSynchronize());
    h->use_seg = kernel == FDSP_KERNEL_SEGMENT_F32;
    return FDSP_OK;
}

extern "C" int fdsp_channelizer_kernel_info(fdsp_channelizer_t h, uint32_t *kernel, uint32_t *groups,
                                            uint32_t *mfma_groups)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    uint32_t kind = FDSP_KERNEL_PLAIN, ng = 0, nm = 0;
    if (h->mode == FDSP_CHAN_BATCHED) {
        if (h->use_seg) {
            kind = FDSP_KERNEL_SEGMENT_F32;
            ng = (uint32_t)h->seg_groups.size();
            for (const SegGroup &g : h->seg_groups)
                nm += g.kind == kSegMfma ? 1u : 0u;
        } else {
            kind = h->bf16_planes ? FDSP_KERNEL_GROUPS_BF16X2 : FDSP_KERNEL_GROUPS_F32;
            ng = h->n_groups;
            for (uint32_t g = 0; g < h->n_groups; ++g)
                nm += h->groups[g * kGroupWords] == kGroupMfma || h->groups[g * kGroupWords] == kGroupFused ? 1u : 0u;
        }
    }
    if (kernel)
        *kernel = kind;
    if (groups)
        *groups = ng;
    if (mfma_groups)
        *mfma_groups = nm;
    return FDSP_OK;
}
