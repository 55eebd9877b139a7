// fdsp_kernels.hip -- gfx950 kernels of the streaming channelizer.
//
// Every streaming kernel works from the closed form of the reference's
// carried state (reference src/transform/falcon_dsp_polyphase_resampler.cc:256-319):
// output n of a stream reads input A(n) = floor(n*down/up) (the "x_idx") with
// filter phase (n*down) mod up, and is emitted in the chunk whose input range
// holds A(n). A chunk starting at global input p therefore emits outputs
// n0 = ceil(p*up/down) ... ceil((p+N)*up/down)-1, and the reference's
// coeff_phase / xOffset are functions of (p, n0). The NCO index of input i is
// i mod R (reference src/transform/falcon_dsp_freq_shift.cc:177-189).
#include <algorithm>
#include <type_traits>

#include "fdsp_internal.h"

namespace fdsp {

__device__ __forceinline__ uint64_t ceil_div_u64(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

// ---------------------------------------------------------------------------
// NCO, one input -> n_chan shifted outputs
// (reference kernels __freq_shift / __freq_shift_multi_chan,
//  src/transform/falcon_dsp_freq_shift_cuda.cu:124-216)
// Each thread owns 16 samples (pairs 128 apart): one exact phasor per thread and
// channel, the others by a product with e^{j 2 pi nu o_i} (two roundings).
// Complex int16 output instead takes the reference's own arithmetic
// (falcon_dsp_freq_shift.cc:146-151): per sample, angle = omega * s in FP64
// with s the index modulo R, cos and sin in FP64 rounded to float (from the
// host-built table where there is one, else sincos here), and the float
// complex product, so the truncated outputs equal the reference's.
// ---------------------------------------------------------------------------
constexpr int kNcoL = kNcoSamples;
// A thread's samples: pairs 128 apart (lane l: its wave's samples 2l, 2l+1,
// 128+2l, 129+2l, ...), so each 16-byte load or store instruction of a wave
// covers 1 KB of consecutive samples; a wave covers 64 kNcoL samples
constexpr const uint32_t (&kNcoOff)[kNcoL] = kNcoOffsets;

// a thread's samples as whole aligned accesses, one per pair
__device__ __forceinline__ void load_pairs(const float2 *p, float2 (&x)[kNcoL])
{
#pragma unroll
    for (int i = 0; i < kNcoL; i += 2) {
        const float4 a = *reinterpret_cast<const float4 *>(p + kNcoOff[i]);
        x[i] = make_float2(a.x, a.y);
        x[i + 1] = make_float2(a.z, a.w);
    }
}
__device__ __forceinline__ void load_pairs(const short2 *p, float2 (&x)[kNcoL])
{
#pragma unroll
    for (int i = 0; i < kNcoL; i += 2) {
        const uint2 a = *reinterpret_cast<const uint2 *>(p + kNcoOff[i]);
        x[i] = make_float2((float)(short)(a.x & 0xffffu), (float)(short)(a.x >> 16));
        x[i + 1] = make_float2((float)(short)(a.y & 0xffffu), (float)(short)(a.y >> 16));
    }
}
__device__ __forceinline__ void store_pairs(float2 *p, const float2 (&y)[kNcoL])
{
#pragma unroll
    for (int i = 0; i < kNcoL; i += 2)
        *reinterpret_cast<float4 *>(p + kNcoOff[i]) = make_float4(y[i].x, y[i].y, y[i + 1].x, y[i + 1].y);
}
__device__ __forceinline__ void store_pairs(short2 *p, const float2 (&y)[kNcoL])
{
    const auto word = [](float2 v) {
        return (uint32_t)(uint16_t)i16_trunc(v.x) | ((uint32_t)(uint16_t)i16_trunc(v.y) << 16);
    };
#pragma unroll
    for (int i = 0; i < kNcoL; i += 2)
        *reinterpret_cast<uint2 *>(p + kNcoOff[i]) = make_uint2(word(y[i]), word(y[i + 1]));
}

template <bool kVec, typename TI, typename TO>
__global__ __launch_bounds__(256) void nco_multi_kernel(const TI *__restrict__ in, uint32_t n, uint64_t start,
                                                        const NcoChan *__restrict__ chans, uint32_t n_chan,
                                                        TO *__restrict__ out, uint64_t out_stride,
                                                        const float2 *__restrict__ tab, uint32_t cpb)
{
    // blockIdx.y picks the shifts c0 .. c1-1 (cpb per block): small chunks
    // spread their shifts over more blocks, re-reading the input from L2
    const uint32_t c0 = blockIdx.y * cpb, c1 = min(n_chan, c0 + cpb);
    // the wave's 256 samples, and this lane's first pair in them
    const uint32_t i0 = (blockIdx.x * blockDim.x + (threadIdx.x & ~63u)) * kNcoL + 2 * (threadIdx.x & 63u);
    if (i0 >= n)
        return;
    const bool full = i0 + kNcoOff[kNcoL - 1] < n;
    float2 x[kNcoL];
    if (kVec && full) {
        load_pairs(in + i0, x);
    } else {
#pragma unroll
        for (int i = 0; i < kNcoL; ++i)
            x[i] = (i0 + kNcoOff[i] < n) ? sample_f32(in[i0 + kNcoOff[i]]) : make_float2(0.f, 0.f);
    }
    for (uint32_t c = c0; c < c1; ++c) {
        const double nu = chans[c].nu;
        const uint32_t r = chans[c].r;
        const uint64_t k0 = mod_u64(start + i0, r, chans[c].inv_r);
        float2 y[kNcoL];
        if constexpr (std::is_same<TO, short2>::value) {
            const double w = chans[c].omega;
            const uint32_t t0 = chans[c].tab;
#pragma unroll
            for (int i = 0; i < kNcoL; ++i) {
                const uint64_t k = k0 + kNcoOff[i] >= r ? k0 + kNcoOff[i] - r : k0 + kNcoOff[i];
                float2 ph;
                if (tab && t0 != kNoTable) {
                    ph = tab[t0 + k];
                } else {
                    double sn, cs;
                    sincos(w * (double)k, &sn, &cs);
                    ph = make_float2((float)cs, (float)sn);
                }
                y[i] = cmul_ref(x[i], ph);
            }
        } else if (k0 + kNcoOff[kNcoL - 1] < r) {
            const float2 base = nco_phasor((int64_t)k0, nu);
#pragma unroll
            for (int i = 0; i < kNcoL; ++i)
                y[i] = cmul(x[i], cmul(base, chans[c].step[i]));
        } else {
            // the rollover falls inside this thread's samples: index restarts at 0
            const float2 base = nco_phasor((int64_t)k0, nu);
            const float2 basew = nco_phasor((int64_t)k0 - (int64_t)r, nu);
#pragma unroll
            for (int i = 0; i < kNcoL; ++i)
                y[i] = cmul(x[i], cmul((k0 + kNcoOff[i] < r) ? base : basew, chans[c].step[i]));
        }
        TO *o = out + (uint64_t)c * out_stride;
        if (kVec && full) {
            store_pairs(o + i0, y);
        } else {
#pragma unroll
            for (int i = 0; i < kNcoL; ++i)
                if (i0 + kNcoOff[i] < n)
                    sample_put(o + i0 + kNcoOff[i], y[i]);
        }
    }
}

template <typename TI, typename TO>
hipError_t launch_nco_multi(const TI *in, uint32_t n, uint64_t start, const NcoChan *chans, uint32_t n_chan, TO *out,
                            uint64_t out_stride, hipStream_t s, const float2 *tab)
{
    if (n == 0)
        return hipSuccess;
    const uint32_t threads = (n + kNcoL - 1) / kNcoL;
    const uint32_t blocks = (threads + 255) / 256;
    // shifts per block: split so the grid reaches about `target` blocks.
    // Measured on 2^18-sample chunks of 16 shifts: complex float outputs run
    // best at 512 (8 shifts per thread), complex int16 outputs, with their
    // per-sample reference phasors, at 1024 (4 per thread).
    const uint32_t target = std::is_same<TO, short2>::value ? 1024u : 512u;
    const uint32_t groups = min(n_chan, max(1u, target / blocks));
    const uint32_t cpb = (n_chan + groups - 1) / groups;
    const dim3 grid(blocks, (n_chan + cpb - 1) / cpb);
    // 16-byte accesses of sample pairs: every channel's output must start 16-byte aligned
    const bool vec = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) % 16 == 0) &&
                     (out_stride * sizeof(TO)) % 16 == 0;
    if (vec)
        hipLaunchKernelGGL((nco_multi_kernel<true, TI, TO>), grid, dim3(256), 0, s, in, n, start, chans, n_chan, out,
                           out_stride, tab, cpb);
    else
        hipLaunchKernelGGL((nco_multi_kernel<false, TI, TO>), grid, dim3(256), 0, s, in, n, start, chans, n_chan,
                           out, out_stride, tab, cpb);
    return hipGetLastError();
}
template hipError_t launch_nco_multi<float2, float2>(const float2 *, uint32_t, uint64_t, const NcoChan *, uint32_t,
                                                     float2 *, uint64_t, hipStream_t, const float2 *);
template hipError_t launch_nco_multi<short2, short2>(const short2 *, uint32_t, uint64_t, const NcoChan *, uint32_t,
                                                     short2 *, uint64_t, hipStream_t, const float2 *);

// ---------------------------------------------------------------------------
// Polyphase resampler, one stream
// (reference falcon_dsp_polyphase_resampler::apply, .cc:256-319, and the
//  __polyphase_resampler_single_out kernel, _cuda.cu:165-235; here the
//  per-output (x_idx, phase) pairs are computed on the device in closed form)
// A block owns a tile of consecutive outputs. Dense tiles stage their input
// span in LDS and read the real taps as scalars; sparse tiles (long
// decimation) read each window straight from global memory. kCpp > 0 is a
// build for exactly kCpp taps per phase: the phase's taps are loaded into
// scalar registers once and the window reads issued in groups (a scalar load
// and an LDS read share one wait counter, so the loop over a runtime tap
// count waits for both every four taps).
// ---------------------------------------------------------------------------
typedef const __attribute__((address_space(4))) float *ctaps_ptr;
typedef float rs_fv2 __attribute__((ext_vector_type(2)));

// one output: the window xw[0 .. cpp-1] (LDS) against the phase's taps g
// (wavefront-uniform), accumulated in the reference's tap order
template <int kCpp>
__device__ __forceinline__ float2 resampler_dot(const float2 *xw, ctaps_ptr g, uint32_t cpp)
{
    float2 acc = make_float2(0.f, 0.f);
    if constexpr (kCpp > 0) {
        // all taps into scalar registers, then the window eight reads at a time
        static_assert(kCpp % 8 == 0, "unrolled builds take multiples of 8 taps per phase");
        float gr[kCpp];
#pragma unroll
        for (int t = 0; t < kCpp; ++t)
            gr[t] = g[t];
#pragma unroll
        for (int t0 = 0; t0 < kCpp; t0 += 8) {
            float2 xr[8];
#pragma unroll
            for (int t = 0; t < 8; ++t)
                xr[t] = xw[t0 + t];
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                acc.x = fmaf(xr[t].x, gr[t0 + t], acc.x);
                acc.y = fmaf(xr[t].y, gr[t0 + t], acc.y);
            }
        }
    } else {
        // four taps per step (scalar operands)
        uint32_t t = 0;
        for (; t + 4 <= cpp; t += 4) {
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const float2 x = xw[t + e];
                acc.x = fmaf(x.x, g[t + e], acc.x);
                acc.y = fmaf(x.y, g[t + e], acc.y);
            }
        }
        for (; t < cpp; ++t) {
            const float2 x = xw[t];
            acc.x = fmaf(x.x, g[t], acc.x);
            acc.y = fmaf(x.y, g[t], acc.y);
        }
    }
    return acc;
}

// outputs are written once and not read back by the launch: streaming stores
__device__ __forceinline__ void store_stream(float2 *dst, float2 v)
{
    __builtin_nontemporal_store(rs_fv2{v.x, v.y}, (__attribute__((address_space(1))) rs_fv2 *)dst);
}

template <int kBlock, bool kStaged, int kCpp = 0>
__global__ __launch_bounds__(kBlock) void resampler_kernel(const float2 *__restrict__ in, uint32_t n, uint64_t n0,
                                                           uint64_t p, uint32_t up, uint32_t down, uint32_t cpp,
                                                           const float *__restrict__ taps,
                                                           const float2 *__restrict__ hist_in,
                                                           float2 *__restrict__ hist_out, float2 *__restrict__ out,
                                                           uint32_t count, uint32_t tile)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t hlen = cpp - 1;
    // new history = last cpp-1 samples of [history | chunk]
    if (blockIdx.x == 0) {
        for (uint32_t t = threadIdx.x; t < hlen; t += kBlock) {
            const int64_t i = (int64_t)n - hlen + t;
            hist_out[t] = i >= 0 ? in[i] : hist_in[hlen + i];
        }
    }
    const uint32_t j0 = blockIdx.x * tile;
    if (j0 >= count)
        return;
    const uint32_t nt = min(tile, count - j0);
    const uint64_t num0 = (n0 + j0) * (uint64_t)down;
    const uint64_t q0 = num0 / up;
    const uint32_t r0 = (uint32_t)(num0 - q0 * up);
    const int64_t a_first = (int64_t)(q0 - p);  // chunk-local anchor of the tile's first output

    if (kStaged) {
        const uint32_t last_num = r0 + (nt - 1) * down;
        const int64_t a_last = a_first + last_num / up;
        const int64_t lo = a_first - hlen;
        const uint32_t span = (uint32_t)(a_last - lo + 1);
        float2 *xs = reinterpret_cast<float2 *>(smem);
        // four loads in flight a thread: unconditional, with the index
        // clamped into the span and the source picked by pointer, so the
        // one wait is at the first LDS write
        for (uint32_t t0 = 0; t0 < span; t0 += 4 * kBlock) {
            float2 v[4];
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const int64_t i = lo + min(t0 + threadIdx.x + e * kBlock, span - 1);
                const float2 *src = i >= 0 ? in + i : hist_in + (hlen + i);
                v[e] = *src;
            }
#pragma unroll
            for (uint32_t e = 0; e < 4; ++e) {
                const uint32_t t = t0 + threadIdx.x + e * kBlock;
                if (t < span)
                    xs[t] = v[e];
            }
        }
        // Outputs q = c (mod up) share the filter phase (r0 + c down) mod up:
        // a wave takes 64 outputs of one class (q = c + up (64 i + lane)), so
        // its taps are wavefront-uniform scalar loads and only the window
        // samples come from LDS. Work item k: class k mod up, slice k / up.
        __syncthreads();
        constexpr uint32_t kWaves = kBlock / 64;
        const uint32_t wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), lane = threadIdx.x & 63;
        const uint32_t per_class = (nt + up - 1) / up, slices = (per_class + 63) / 64;
        for (uint32_t k = wave; k < up * slices; k += kWaves) {
            const uint32_t c = k % up, i = k / up;
            // r0 + q down = (r0 + c down) + up down (64 i + lane): the class's
            // first anchor (uniform) plus down per output of the class
            const uint64_t nc = (uint64_t)r0 + (uint64_t)c * down;
            const uint32_t ph = (uint32_t)(nc % up), ac = (uint32_t)(nc / up);
            const uint32_t oc = 64 * i + lane;  // the output's index within its class
            const uint32_t q = c + up * oc;
            if (q >= nt)
                continue;
            const float2 *xw = xs + ac + down * oc;  // window start, relative to lo
            store_stream(out + j0 + q, resampler_dot<kCpp>(xw, (ctaps_ptr)taps + ph * cpp, cpp));
        }
    } else {
        for (uint32_t q = threadIdx.x; q < nt; q += kBlock) {
            const uint32_t num = r0 + q * down;
            const uint32_t adv = (up == 1) ? num : num / up;
            const uint32_t ph = num - adv * up;
            const int64_t a = a_first + adv;
            const float *g = taps + ph * cpp;
            float2 acc = make_float2(0.f, 0.f);
            const int64_t lo = a - hlen;
            if (lo >= 0) {
                const float2 *xw = in + lo;
                for (uint32_t t = 0; t < cpp; ++t) {
                    const float2 x = xw[t];
                    acc.x = fmaf(x.x, g[t], acc.x);
                    acc.y = fmaf(x.y, g[t], acc.y);
                }
            } else {
                for (uint32_t t = 0; t < cpp; ++t) {
                    const int64_t i = lo + t;
                    const float2 x = i >= 0 ? in[i] : hist_in[hlen + i];
                    acc.x = fmaf(x.x, g[t], acc.x);
                    acc.y = fmaf(x.y, g[t], acc.y);
                }
            }
            out[j0 + q] = acc;
        }
    }
}

// LDS budget for a staged resampler tile's input span
constexpr uint32_t kResamplerLdsBytes = 48 * 1024;

hipError_t launch_resampler(const float2 *in, uint32_t n, uint64_t p, uint32_t up, uint32_t down, uint32_t cpp,
                            const float *taps, const float2 *hist_in, float2 *hist_out, float2 *out, uint32_t count,
                            hipStream_t s)
{
    constexpr int kBlock = 256;
    const uint64_t n0 = (p * up + down - 1) / down;
    // staged tile: outputs whose input span (tile*down/up + cpp) fits the LDS budget
    uint64_t tile = 0;
    if ((uint64_t)(cpp + 2) * sizeof(float2) < kResamplerLdsBytes) {
        const uint64_t span_cap = kResamplerLdsBytes / sizeof(float2) - cpp - 1;
        tile = span_cap * up / down;
        if (tile > 4096)
            tile = 4096;
        // smaller tiles, so the grid reaches about 2048 blocks (8 per CU): a
        // chunk of 52k outputs otherwise ran on 13 CUs, and 838k outputs in
        // 512 blocks of 1639 took 13 us a launch, in 2048 blocks of 410 about 9
        const uint64_t cap = tile;
        tile = std::min<uint64_t>(cap, std::max<uint64_t>(kBlock, (count + 2047) / 2048));
        // whole waves per phase class: the next multiple of 64 up outputs
        // (838k outputs at up 4: 512 a tile, 7.3 us, against 7.8 at 256)
        const uint64_t cls = 64ull * up;
        if (cap >= cls)
            tile = std::min((tile + cls - 1) / cls * cls, cap / cls * cls);
    }
    const bool staged = tile >= kBlock;
    if (!staged) {
        tile = 1024;
        // keep r0 + q*down inside 32 bits
        while (tile > 1 && (uint64_t)tile * down + up >= (1ull << 31))
            tile >>= 1;
    }
    const uint32_t blocks = count ? (uint32_t)((count + tile - 1) / tile) : 1;
    if (staged) {
        const uint32_t span_max = (uint32_t)(((tile - 1) * down + up) / up + cpp + 1);
        const size_t lds = (size_t)span_max * sizeof(float2);
        // unrolled builds for common filter lengths per phase
#define FDSP_RS_LAUNCH(C)                                                                                          \
    hipLaunchKernelGGL((resampler_kernel<kBlock, true, C>), dim3(blocks), dim3(kBlock), lds, s, in, n, n0, p, up, \
                       down, cpp, taps, hist_in, hist_out, out, count, (uint32_t)tile)
        switch (cpp) {
        case 16: FDSP_RS_LAUNCH(16); break;
        case 24: FDSP_RS_LAUNCH(24); break;
        case 32: FDSP_RS_LAUNCH(32); break;
        case 64: FDSP_RS_LAUNCH(64); break;
        default: FDSP_RS_LAUNCH(0); break;
        }
#undef FDSP_RS_LAUNCH
    } else {
        hipLaunchKernelGGL((resampler_kernel<kBlock, false>), dim3(blocks), dim3(kBlock), 0, s, in, n, n0, p, up,
                           down, cpp, taps, hist_in, hist_out, out, count, (uint32_t)tile);
    }
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// FIR with complex taps, one stream
// (reference falcon_dsp_fir_filter::apply, .cc:177-243; CUDA __fir_filter,
//  _cuda.cu:102-187). y[n] = sum_k h[k] x[n-k], accumulated k = 0..M-1.
// Input window and taps staged in LDS.
// ---------------------------------------------------------------------------
template <int kBlock, bool kTapsLds, typename TI, typename TO>
__global__ __launch_bounds__(kBlock) void fir_kernel(const TI *__restrict__ in, uint32_t n,
                                                     const float2 *__restrict__ taps, uint32_t m,
                                                     const float2 *__restrict__ hist_in,
                                                     float2 *__restrict__ hist_out, TO *__restrict__ out,
                                                     uint32_t tile)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint32_t hlen = m - 1;
    if (blockIdx.x == 0) {
        for (uint32_t t = threadIdx.x; t < hlen; t += kBlock) {
            const int64_t i = (int64_t)n - hlen + t;
            hist_out[t] = i >= 0 ? sample_f32(in[i]) : hist_in[hlen + i];
        }
    }
    const uint32_t o0 = blockIdx.x * tile;
    const uint32_t nt = min(tile, n - o0);
    float2 *xs = reinterpret_cast<float2 *>(smem);
    float2 *ts = xs + tile + hlen;
    const int64_t lo = (int64_t)o0 - hlen;
    for (uint32_t t = threadIdx.x; t < nt + hlen; t += kBlock) {
        const int64_t i = lo + t;
        xs[t] = i >= 0 ? sample_f32(in[i]) : hist_in[hlen + i];
    }
    if (kTapsLds)
        for (uint32_t t = threadIdx.x; t < m; t += kBlock)
            ts[t] = taps[t];
    __syncthreads();
    const float2 *g = kTapsLds ? ts : taps;
    for (uint32_t q = threadIdx.x; q < nt; q += kBlock) {
        const float2 *xw = xs + q + hlen;  // xw[-k] = x[o0 + q - k]
        float2 acc = make_float2(0.f, 0.f);
        if constexpr (std::is_same<TO, short2>::value) {
            // integer output: the reference's rounding, product then sum, no
            // fused multiply-add (falcon_dsp_fir_filter.cc:208-230)
            for (uint32_t k = 0; k < m; ++k) {
                const float2 p = cmul_ref(xw[-(int32_t)k], g[k]);
                {
#pragma clang fp contract(off)
                    acc.x = acc.x + p.x;
                    acc.y = acc.y + p.y;
                }
            }
        } else {
            for (uint32_t k = 0; k < m; ++k) {
                const float2 x = xw[-(int32_t)k];
                const float2 h = g[k];
                // complex<float> product added to the accumulator, reference order
                acc.x += x.x * h.x - x.y * h.y;
                acc.y += x.x * h.y + x.y * h.x;
            }
        }
        sample_put(out + o0 + q, acc);
    }
}

template <typename TI, typename TO>
hipError_t launch_fir(const TI *in, uint32_t n, const float2 *taps, uint32_t m, const float2 *hist_in,
                      float2 *hist_out, TO *out, hipStream_t s)
{
    constexpr int kBlock = 256;
    const uint32_t tile = 2048;
    const bool taps_lds = m <= 2048;
    const size_t lds = (size_t)(tile + m - 1) * sizeof(float2) + (taps_lds ? m * sizeof(float2) : 0);
    if (lds > 64 * 1024)
        return hipErrorInvalidValue;
    const uint32_t blocks = (n + tile - 1) / tile;
    if (taps_lds)
        hipLaunchKernelGGL((fir_kernel<kBlock, true, TI, TO>), dim3(blocks), dim3(kBlock), lds, s, in, n, taps, m,
                           hist_in, hist_out, out, tile);
    else
        hipLaunchKernelGGL((fir_kernel<kBlock, false, TI, TO>), dim3(blocks), dim3(kBlock), lds, s, in, n, taps, m,
                           hist_in, hist_out, out, tile);
    return hipGetLastError();
}
template hipError_t launch_fir<float2, float2>(const float2 *, uint32_t, const float2 *, uint32_t, const float2 *,
                                               float2 *, float2 *, hipStream_t);
template hipError_t launch_fir<short2, short2>(const short2 *, uint32_t, const float2 *, uint32_t, const float2 *,
                                               float2 *, short2 *, hipStream_t);

// ---------------------------------------------------------------------------
// FIR bank, real taps, many channels (compute-bound FP32 VALU kernel).
// blockIdx.y is the channel. Each thread keeps kR consecutive outputs in
// registers and slides a kR+kR-1 sample register window down the input, so
// every tap costs one LDS read and 2*kR FMAs; taps are channel-uniform and
// come through the scalar cache. The LDS window is padded by one sample per
// kR so the stride-kR reads of a wavefront fall in distinct banks.
// ---------------------------------------------------------------------------
constexpr int kBankR = 8;
__host__ __device__ __forceinline__ uint32_t bank_pad(uint32_t i) { return i + i / kBankR; }

template <int kBlock>
__global__ __launch_bounds__(kBlock) void fir_bank_kernel(const float2 *__restrict__ in, uint32_t n,
                                                          uint64_t in_stride, const float *__restrict__ taps,
                                                          uint32_t m, const float2 *__restrict__ hist_in,
                                                          float2 *__restrict__ hist_out,
                                                          float2 *__restrict__ out, uint64_t out_stride)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *xs = reinterpret_cast<float2 *>(smem);
    const uint32_t c = blockIdx.y;
    const uint32_t hlen = m - 1;
    const float2 *x = in + c * in_stride;
    const float2 *hi = hist_in + (uint64_t)c * hlen;
    const float *h = taps + (uint64_t)c * m;
    float2 *y = out + c * out_stride;
    if (blockIdx.x == 0) {
        float2 *ho = hist_out + (uint64_t)c * hlen;
        for (uint32_t t = threadIdx.x; t < hlen; t += kBlock) {
            const int64_t i = (int64_t)n - hlen + t;
            ho[t] = i >= 0 ? x[i] : hi[hlen + i];
        }
    }
    constexpr uint32_t tile = kBlock * kBankR;
    const uint32_t o0 = blockIdx.x * tile;
    // window: x[o0 - mpad .. o0 + tile), mpad = m rounded up to kR (extra zero-tap slack)
    const uint32_t mpad = (m + kBankR - 1) / kBankR * kBankR;
    const int64_t lo = (int64_t)o0 - mpad;
    const uint32_t span = tile + mpad;
    for (uint32_t t = threadIdx.x; t < span; t += kBlock) {
        const int64_t i = lo + t;
        float2 v = make_float2(0.f, 0.f);
        if (i >= 0) {
            if (i < n)
                v = x[i];
        } else if (i >= -(int64_t)hlen) {
            v = hi[hlen + i];
        }
        xs[bank_pad(t)] = v;
    }
    __syncthreads();
    // outputs o0 + tid*kR + r; sample x[o0 + tid*kR + r - k] lives at window
    // index mpad + tid*kR + r - k
    const uint32_t wb = mpad + threadIdx.x * kBankR;
    float2 acc[kBankR];
#pragma unroll
    for (int r = 0; r < kBankR; ++r)
        acc[r] = make_float2(0.f, 0.f);
    // w[i] = x at window index wb - k0 - (kR-1) + i, i in [0, 2kR-1). Taps
    // k0..k0+kR-1 need all of it; the next block of taps reuses the lowest
    // kR-1 entries (moved up by kR) and loads kR new ones.
    float2 w[2 * kBankR - 1];
#pragma unroll
    for (int i = 1; i < kBankR; ++i)
        w[kBankR - 1 + i] = xs[bank_pad(wb + i)];
    for (uint32_t k0 = 0; k0 < mpad; k0 += kBankR) {
#pragma unroll
        for (int i = 0; i < kBankR; ++i)
            w[i] = xs[bank_pad(wb - k0 - (kBankR - 1) + i)];
        float hk[kBankR];
#pragma unroll
        for (int kk = 0; kk < kBankR; ++kk)
            hk[kk] = (k0 + kk < m) ? h[k0 + kk] : 0.f;
#pragma unroll
        for (int kk = 0; kk < kBankR; ++kk) {
#pragma unroll
            for (int r = 0; r < kBankR; ++r) {
                const float2 v = w[kBankR - 1 + r - kk];
                acc[r].x = fmaf(hk[kk], v.x, acc[r].x);
                acc[r].y = fmaf(hk[kk], v.y, acc[r].y);
            }
        }
        // slide: the lowest kR-1 samples become the upper part for the next taps
#pragma unroll
        for (int i = 0; i < kBankR - 1; ++i)
            w[kBankR + i] = w[i];
    }
    const uint32_t ob = o0 + threadIdx.x * kBankR;
#pragma unroll
    for (int r = 0; r < kBankR; ++r)
        if (ob + r < n)
            y[ob + r] = acc[r];
}

// FIR bank on FP32 matrix cores. For 16 consecutive outputs of a channel,
// y[16b + r] = sum_t H[r][t] W[t][b] with H[r][t] = h[t - 15 + r] (a banded
// Toeplitz matrix of the taps, zero outside 0 <= t - 15 + r < m) and
// W[t][b] = x[16b + 15 - t]: a GEMM per channel with M = 16 (r), N = output
// blocks b, K = m + 15 rounded up to 4 (kp). v_mfma_f32_16x16x4f32 takes
// A = H (lane l: row l%16, k l/16) and B = W (lane l: column l%16, k l/16)
// and leaves lane l outputs 4(l/16) .. 4(l/16)+3 of block l%16. The real and
// imaginary parts of the samples are two GEMMs sharing A. A block stages
// its input window and the zero-padded taps in LDS; each wave accumulates
// kFbG column groups of 16 blocks.
typedef float fb_floatx4 __attribute__((ext_vector_type(4)));
constexpr uint32_t kFbG = 4, kFbWaves = 4;
constexpr uint32_t kFbTile = kFbWaves * kFbG * 256;  // outputs per block
// one padding sample per 16, so the 16 columns of a B read (16 samples apart)
// fall in distinct banks
__host__ __device__ __forceinline__ uint32_t fb_pad(uint32_t q) { return q + (q >> 4); }

__global__ __launch_bounds__(kFbWaves * 64) void fir_bank_mfma_kernel(
    const float2 *__restrict__ in, uint32_t n, uint64_t in_stride, const float *__restrict__ taps, uint32_t m,
    uint32_t kp, const float2 *__restrict__ hist_in, float2 *__restrict__ hist_out, float2 *__restrict__ out,
    uint64_t out_stride)
{
    extern __shared__ __align__(16) unsigned char smem[];
    float2 *xs = reinterpret_cast<float2 *>(smem);
    const uint32_t span = kFbTile + kp - 16;
    float *hs = reinterpret_cast<float *>(xs + fb_pad(span) + 1);  // kp + 16 taps
    const uint32_t c = blockIdx.y, tid = threadIdx.x;
    const uint32_t hlen = m - 1;
    const float2 *x = in + c * in_stride;
    const float2 *hi = hist_in + (uint64_t)c * hlen;
    const float *h = taps + (uint64_t)c * m;
    float2 *y = out + c * out_stride;
    if (blockIdx.x == 0) {
        float2 *ho = hist_out + (uint64_t)c * hlen;
        for (uint32_t t = tid; t < hlen; t += kFbWaves * 64) {
            const int64_t i = (int64_t)n - hlen + t;
            ho[t] = i >= 0 ? x[i] : hi[hlen + i];
        }
    }
    const uint32_t o0 = blockIdx.x * kFbTile;
    // window: x[lo .. lo + span), lo = o0 + 16 - kp; W[t][b] for the block's
    // blocks bb = 0 .. kFbTile/16 - 1 sits at window index 16 bb + kp - 1 - t
    const int64_t lo = (int64_t)o0 + 16 - (int64_t)kp;
    for (uint32_t u = tid; u < kp + 16; u += kFbWaves * 64)
        hs[u] = u >= 15 && u - 15 < m ? h[u - 15] : 0.f;
    if (lo >= 0 && lo + span <= n) {
        // the whole window inside the chunk: eight independent loads per thread in flight
        constexpr uint32_t kU = 8, kT = kFbWaves * 64;
        for (uint32_t t0 = tid; t0 < span; t0 += kT * kU) {
            float2 v[kU];
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u)
                v[u] = t0 + u * kT < span ? x[lo + t0 + u * kT] : make_float2(0.f, 0.f);
#pragma unroll
            for (uint32_t u = 0; u < kU; ++u)
                if (t0 + u * kT < span)
                    xs[fb_pad(t0 + u * kT)] = v[u];
        }
    } else {
        for (uint32_t t = tid; t < span; t += kFbWaves * 64) {
            const int64_t i = lo + t;
            float2 v = make_float2(0.f, 0.f);
            if (i >= 0) {
                if (i < n)
                    v = x[i];
            } else if (i >= -(int64_t)hlen) {
                v = hi[hlen + i];
            }
            xs[fb_pad(t)] = v;
        }
    }
    __syncthreads();
    const uint32_t wave = tid >> 6, lane = tid & 63, kq = lane >> 4, j = lane & 15;
    fb_floatx4 dr[kFbG], di[kFbG];
#pragma unroll
    for (uint32_t g = 0; g < kFbG; ++g) {
        dr[g] = fb_floatx4{0.f, 0.f, 0.f, 0.f};
        di[g] = fb_floatx4{0.f, 0.f, 0.f, 0.f};
    }
    // window index of W[kq][bb] at t0 = 0, for the lane's block in group 0
    const uint32_t q0 = 16 * (wave * kFbG * 16 + j) + kp - 1 - kq;
    // K-steps in rounds of four, t0 = 16u + 4s. W[t0 + kq][block 16g + j]
    // sits at fb_pad(q0 + 256g - t0) = fb_pad(q0) + 272g - e_s - 17u with
    // e_s = 4s + (4s > q0 mod 16): one offset per s, the groups at constant
    // distances, 17 slots back per round; H[j][t0 + kq] at t0 + kq + j.
    uint32_t bo[4];
#pragma unroll
    for (uint32_t st = 0; st < 4; ++st)
        bo[st] = fb_pad(q0) - (4 * st + (4 * st > (q0 & 15) ? 1u : 0u));
    uint32_t ao = kq + j;
    const uint32_t rounds = kp / 16;
    for (uint32_t u = 0; u < rounds; ++u) {
#pragma unroll
        for (uint32_t st = 0; st < 4; ++st) {
            const float a = hs[ao + 4 * st];
#pragma unroll
            for (uint32_t g = 0; g < kFbG; ++g) {
                const float2 b = xs[bo[st] + 272 * g];
                dr[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.x, dr[g], 0, 0, 0);
                di[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.y, di[g], 0, 0, 0);
            }
        }
        ao += 16;
#pragma unroll
        for (uint32_t st = 0; st < 4; ++st)
            bo[st] -= 17;
    }
    for (uint32_t t0 = rounds * 16; t0 < kp; t0 += 4) {  // kp not a multiple of 16
        const float a = hs[t0 + kq + j];
#pragma unroll
        for (uint32_t g = 0; g < kFbG; ++g) {
            const float2 b = xs[fb_pad(q0 + 256 * g - t0)];
            dr[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.x, dr[g], 0, 0, 0);
            di[g] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b.y, di[g], 0, 0, 0);
        }
    }
    const bool vec = (reinterpret_cast<uintptr_t>(y) & 15) == 0;
#pragma unroll
    for (uint32_t g = 0; g < kFbG; ++g) {
        const uint32_t o = o0 + 16 * ((wave * kFbG + g) * 16 + j) + 4 * kq;
        if (vec && o + 4 <= n) {
            float4 *p = reinterpret_cast<float4 *>(y + o);
            p[0] = make_float4(dr[g][0], di[g][0], dr[g][1], di[g][1]);
            p[1] = make_float4(dr[g][2], di[g][2], dr[g][3], di[g][3]);
        } else {
#pragma unroll
            for (uint32_t v = 0; v < 4; ++v)
                if (o + v < n)
                    y[o + v] = make_float2(dr[g][v], di[g][v]);
        }
    }
}

hipError_t launch_fir_bank(const float2 *in, uint32_t n, uint64_t in_stride, const float *taps, uint32_t m,
                           uint32_t n_chan, const float2 *hist_in, float2 *hist_out, float2 *out,
                           uint64_t out_stride, hipStream_t s)
{
    {
        // the matrix-core form whenever its window and taps fit the LDS
        const uint32_t kp = (m + 15 + 3) / 4 * 4;
        const size_t lds = (size_t)(fb_pad(kFbTile + kp - 16) + 1) * sizeof(float2) + (kp + 16) * sizeof(float);
        if (lds <= 64 * 1024) {
            dim3 grid((n + kFbTile - 1) / kFbTile, n_chan);
            hipLaunchKernelGGL(fir_bank_mfma_kernel, grid, dim3(kFbWaves * 64), lds, s, in, n, in_stride, taps, m,
                               kp, hist_in, hist_out, out, out_stride);
            return hipGetLastError();
        }
    }
    constexpr int kBlock = 256;
    constexpr uint32_t tile = kBlock * kBankR;
    const uint32_t mpad = (m + kBankR - 1) / kBankR * kBankR;
    const size_t lds = (size_t)(bank_pad(tile + mpad) + 1) * sizeof(float2);
    if (lds > 64 * 1024)
        return hipErrorInvalidValue;
    dim3 grid((n + tile - 1) / tile, n_chan);
    hipLaunchKernelGGL((fir_bank_kernel<kBlock>), grid, dim3(kBlock), lds, s, in, n, in_stride, taps, m, hist_in,
                       hist_out, out, out_stride);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------
// Polar discriminator
// (reference falcon_dsp_polar_discriminator::apply, .cc:134-171)
// ---------------------------------------------------------------------------
template <typename TI>
__global__ __launch_bounds__(256) void discriminator_kernel(const TI *__restrict__ in, uint32_t n,
                                                            const float2 *__restrict__ prev, int has_prev,
                                                            float2 *__restrict__ prev_out, float *__restrict__ out)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    if (i == n - 1)
        *prev_out = sample_f32(in[i]);
    if (i == 0 && !has_prev)
        return;
    const float2 a = sample_f32(in[i]);
    const float2 b = (i == 0) ? *prev : sample_f32(in[i - 1]);
    // the reference's float complex product (unfused), whatever the input type
    const float2 z = cmul_ref(a, make_float2(b.x, -b.y));
    out[has_prev ? i : i - 1] = atan2f(z.y, z.x);
}

template <typename TI>
hipError_t launch_discriminator(const TI *in, uint32_t n, const float2 *prev, int has_prev, float2 *prev_out,
                                float *out, hipStream_t s)
{
    if (n == 0)
        return hipSuccess;
    hipLaunchKernelGGL(discriminator_kernel<TI>, dim3((n + 255) / 256), dim3(256), 0, s, in, n, prev, has_prev,
                       prev_out, out);
    return hipGetLastError();
}
template hipError_t launch_discriminator<float2>(const float2 *, uint32_t, const float2 *, int, float2 *, float *,
                                                 hipStream_t);
template hipError_t launch_discriminator<short2>(const short2 *, uint32_t, const float2 *, int, float2 *, float *,
                                                 hipStream_t);

}  // namespace fdsp
