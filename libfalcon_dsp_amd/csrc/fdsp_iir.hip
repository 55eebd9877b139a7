// General IIR filter y = (sum_k b[k] x[n-k]) - sum_j a[j] y[n-j], bit-identical
// to the reference's sequential loop (src/transform/falcon_dsp_iir_filter.cc:
// 119-155 for the coefficients, 235-396 for the recurrence).
//
// The reference forms each output as ((0 + b0 x[n]) + b1 x[n-1] + ...) and
// then subtracts a1 y[n-1], a2 y[n-2], ... in that order, every product a
// std::complex<float> product. The b-part v[n] depends on inputs only, so a
// block computes it for all its positions in parallel; the feedback is a
// recurrence on v. Each lane owns `seg` consecutive outputs and starts the same
// recurrence `warm` samples earlier from a zero state (or at the chunk start
// from the carried state). For a stable filter two float trajectories draw
// together and, once one step's products round alike, are identical from then
// on; `warm` is sized on the host from the filter's own decay (the shortest
// length between the 2^-30 and 2^-44 decays at which trial segments meet,
// plus a margin).
//
// Nothing rests on that argument: every lane records its state when its
// segment starts and when it ends. A segment whose start state equals its
// predecessor's end state bit for bit (up to the sign of zero) computed the
// sequential outputs exactly, by induction from the chunk's first segment,
// which starts from the carried state. Lanes compare within their block; a
// one-wave kernel launched next compares across blocks and, if any segment
// failed, recomputes the failed segments in order from their predecessors'
// exact end states. Filters whose segments do not merge run as one
// sequential wave instead (iir_serial_kernel), or, when the caller accepts
// float rounding, as unverified segments.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <new>
#include <type_traits>
#include <vector>

#include "../../include/fdsp.h"
#include "fdsp_host.h"
#include "fdsp_iir.h"
#include "fdsp_internal.h"

using namespace fdsp;

namespace {

constexpr uint32_t kLanes = 64;                  // one wave per block (lockstep in-place LDS use)
constexpr uint32_t kMaxSeg = 64;                 // outputs per lane: 8 to 64, by chunk length
constexpr uint32_t kMinWarm = 32, kMaxWarm = 2048;
// added to the shortest warm-up that merges on the host's trial starts: real
// signals miss more often than the trial noise (FM angles: 0.5 us of repairs a
// chunk without it)
constexpr uint32_t kWarmMargin = 32;
constexpr uint32_t kMaxB = 256, kMaxFb = 16;

template <bool C> using Val = typename std::conditional<C, float2, float>::type;

template <bool C> __device__ __forceinline__ Val<C> vzero()
{
    if constexpr (C)
        return make_float2(0.f, 0.f);
    else
        return 0.f;
}
template <bool C> __device__ __forceinline__ Val<C> from2(float2 v)
{
    if constexpr (C)
        return v;
    else
        return v.x;
}
__device__ __forceinline__ float2 to2(float2 v) { return v; }
__device__ __forceinline__ float2 to2(float v) { return make_float2(v, 0.f); }

// LDS slot of span position q: 16 bytes of padding after every 64 values, so
// lanes reading at a stride of `seg` values start on different banks
template <bool C> __device__ __forceinline__ uint32_t slot(uint32_t q) { return q + (C ? 2u : 4u) * (q >> 6); }
template <bool C> constexpr uint32_t slots(uint32_t span) { return span + (C ? 2u : 4u) * ((span + 63) >> 6); }

// equal bit patterns, or equal values (+0 and -0 lead to the same outputs)
__device__ __forceinline__ bool same(float a, float b) { return a == b || __float_as_uint(a) == __float_as_uint(b); }
__device__ __forceinline__ bool same(float2 a, float2 b) { return same(a.x, b.x) && same(a.y, b.y); }

struct IirLaunch {
    const void *x;            // n inputs of type TIN
    void *out;                // n outputs of type tout
    const float2 *b;          // nb taps
    const float2 *xh;         // nb - 1 carried inputs, oldest first
    const float2 *yh;         // kMaxFb carried outputs, newest first
    float2 *xh_next, *yh_next;  // the same after this chunk (the other copy: no read of xh, yh races)
    float2 *rs, *re;          // per segment: P start-state values, P end-state values
    unsigned long long *bad;  // per block: bit t set when segment t's start differs from t-1's end
    uint32_t *failed;         // blocks with a failed segment (the finishing wave clears it)
    uint32_t n, nb, warm, tout;
    uint32_t verify;          // 0: segments are trusted (outputs within float rounding)
    uint32_t split;           // serial form, complex samples, real a: re and im run on lanes 0 and 1
    uint32_t seg;             // outputs per lane (8, 16, 32 or 64); a block owns 64 segments
};
template <int P> struct Fb {
    float2 a[P];  // a[1..P] of the normalised denominator (zero past na - 1)
};

template <uint32_t TIN, bool C> __device__ __forceinline__ Val<C> load_x(const IirLaunch &g, int32_t i)
{
    float2 v;
    if (i < 0) {
        v = g.xh[(int32_t)g.nb - 1 + i];
    } else {
        if constexpr (TIN == kSampleCf32)
            v = static_cast<const float2 *>(g.x)[i];
        else if constexpr (TIN == kSampleF32)
            v = make_float2(static_cast<const float *>(g.x)[i], 0.f);
        else if constexpr (TIN == kSampleCi16) {
            const short2 s = static_cast<const short2 *>(g.x)[i];
            v = make_float2((float)s.x, (float)s.y);
        } else
            v = make_float2((float)static_cast<const short *>(g.x)[i], 0.f);
    }
    return from2<C>(v);
}

// acc + x b, the product as std::complex<float> rounds it. With a real
// coefficient (RA) the products with its zero imaginary part only add signed
// zeros, so each part is one real product (the same values for finite
// samples); a real-coefficient filter on real samples (C false) likewise
// keeps zero imaginary parts throughout.
template <bool C, bool RA = false> __device__ __forceinline__ Val<C> mac(Val<C> acc, Val<C> x, float2 b)
{
#pragma clang fp contract(off)
    if constexpr (C && RA) {
        return make_float2(acc.x + x.x * b.x, acc.y + x.y * b.x);
    } else if constexpr (C) {
        const float2 p = cmul_ref(x, b);
        return make_float2(acc.x + p.x, acc.y + p.y);
    } else {
        return acc + x * b.x;
    }
}

// one output: v - a1 y[n-1] - a2 y[n-2] - ..., subtracted in that order;
// st[0] is the newest output
template <int P, bool C, bool RA = false>
__device__ __forceinline__ Val<C> fb_step(Val<C> v, Val<C> (&st)[P], const Fb<P> &fb)
{
#pragma clang fp contract(off)
    Val<C> y = v;
#pragma unroll
    for (int k = 0; k < P; ++k) {
        if constexpr (C && RA) {
            y.x = y.x - st[k].x * fb.a[k].x;
            y.y = y.y - st[k].y * fb.a[k].x;
        } else if constexpr (C) {
            const float2 p = cmul_ref(st[k], fb.a[k]);
            y.x = y.x - p.x;
            y.y = y.y - p.y;
        } else {
            y = y - st[k] * fb.a[k].x;
        }
    }
#pragma unroll
    for (int k = P - 1; k > 0; --k)
        st[k] = st[k - 1];
    st[0] = y;
    return y;
}

template <bool C> __device__ __forceinline__ void put_out(const IirLaunch &g, uint32_t i, Val<C> y)
{
    const float2 v = to2(y);
    switch (g.tout) {
    case kSampleCf32:
        static_cast<float2 *>(g.out)[i] = v;
        break;
    case kSampleF32:
        static_cast<float *>(g.out)[i] = v.x;
        break;
    case kSampleCi16:
        static_cast<short2 *>(g.out)[i] = make_short2(i16_trunc(v.x), i16_trunc(v.y));
        break;
    default:
        static_cast<short *>(g.out)[i] = i16_trunc(v.x);
    }
}

// the b-part of position i >= 0
template <uint32_t TIN, bool C, bool RA = false> __device__ __forceinline__ Val<C> bpart(const IirLaunch &g, int32_t i)
{
    Val<C> acc = vzero<C>();
    for (uint32_t k = 0; k < g.nb; ++k)
        acc = mac<C, RA>(acc, load_x<TIN, C>(g, i - (int32_t)k), g.b[k]);
    return acc;
}

// The b-part of positions lo .. lo + span - 1 into buf[slot(0 .. span - 1)].
// The inputs lo - (nb - 1) .. lo + span - 1 are first staged at buf[slot(t)],
// t = position - lo + nb - 1, with coalesced loads, 16 per lane in flight.
// Then rows of 64 positions, in ascending order, replace them: the b-part of
// q goes to slot q, whose input (position lo + q - nb + 1) no later row
// reads, and a row's reads all precede its writes (one wave, in order).
template <uint32_t TIN, bool C, bool RA = false>
__device__ void stage_bpart(const IirLaunch &g, Val<C> *buf, int32_t lo, uint32_t span)
{
    const uint32_t lane = threadIdx.x, h = g.nb - 1, tot = span + h;
    const int32_t base = lo - (int32_t)h;
    constexpr uint32_t kU = 16;
    for (uint32_t t0 = 0; t0 < tot; t0 += kLanes * kU) {
        Val<C> xv[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t t = t0 + u * kLanes + lane;
            const int32_t p = base + (int32_t)t;
            xv[u] = t < tot && p >= -(int32_t)h && p < (int32_t)g.n ? load_x<TIN, C>(g, p) : vzero<C>();
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t t = t0 + u * kLanes + lane;
            if (t < tot)
                buf[slot<C>(t)] = xv[u];
        }
    }
    __syncthreads();
    // eight rows at a time: their reads (all at or above the group's first
    // row) before their writes
    constexpr uint32_t kR = 8;
    for (uint32_t q0 = 0; q0 < span; q0 += kLanes * kR) {
        Val<C> acc[kR];
#pragma unroll
        for (uint32_t r = 0; r < kR; ++r)
            acc[r] = vzero<C>();
        for (uint32_t k = 0; k < g.nb; ++k) {
            const float2 bk = g.b[k];
#pragma unroll
            for (uint32_t r = 0; r < kR; ++r) {
                const uint32_t q = q0 + r * kLanes + lane;
                if (q < span)
                    acc[r] = mac<C, RA>(acc[r], buf[slot<C>(q + h - k)], bk);
            }
        }
#pragma unroll
        for (uint32_t r = 0; r < kR; ++r) {
            const uint32_t q = q0 + r * kLanes + lane;
            if (q < span)
                buf[slot<C>(q)] = acc[r];
        }
    }
    __syncthreads();
}

// 8 values: two (real) or four (complex) 16-byte LDS accesses; a batch
// starts at a multiple of 8 and so never crosses a padding gap
constexpr int kBatch = 8;

template <bool C> __device__ __forceinline__ void lds_read(const Val<C> *p, Val<C> (&v)[kBatch])
{
    const float4 *q = reinterpret_cast<const float4 *>(p);
#pragma unroll
    for (int i = 0; i < (C ? 4 : 2); ++i) {
        const float4 w = q[i];
        if constexpr (C) {
            v[2 * i] = make_float2(w.x, w.y);
            v[2 * i + 1] = make_float2(w.z, w.w);
        } else {
            v[4 * i] = w.x;
            v[4 * i + 1] = w.y;
            v[4 * i + 2] = w.z;
            v[4 * i + 3] = w.w;
        }
    }
}
template <bool C> __device__ __forceinline__ void lds_write(Val<C> *p, const Val<C> (&v)[kBatch])
{
    float4 *q = reinterpret_cast<float4 *>(p);
#pragma unroll
    for (int i = 0; i < (C ? 4 : 2); ++i) {
        if constexpr (C)
            q[i] = make_float4(v[2 * i].x, v[2 * i].y, v[2 * i + 1].x, v[2 * i + 1].y);
        else
            q[i] = make_float4(v[4 * i], v[4 * i + 1], v[4 * i + 2], v[4 * i + 3]);
    }
}

__device__ __forceinline__ unsigned long long shfl64(unsigned long long v, int lane)
{
    const uint32_t lo = __shfl((int)(uint32_t)v, lane), hi = __shfl((int)(uint32_t)(v >> 32), lane);
    return ((unsigned long long)hi << 32) | lo;
}

// carried inputs: the last nb - 1 of (old history, this chunk), into the
// next copy (one wave; nb - 1 <= 255)
template <uint32_t TIN> __device__ void carry_inputs(const IirLaunch &g)
{
    const uint32_t lane = threadIdx.x, h = g.nb - 1;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const uint32_t k = lane + r * kLanes;
        if (k < h)
            g.xh_next[k] = load_x<TIN, true>(g, (int32_t)g.n - (int32_t)h + (int32_t)k);
    }
}

// One wave, launched after iir_kernel on the same stream (so every block's
// outputs, records and failure bits are visible): check the block
// boundaries, repair failed segments in order, then carry the stream state to
// the next call.
template <int P, bool C, uint32_t TIN, bool RA>
__global__ __launch_bounds__(kLanes) void iir_finish_kernel(IirLaunch g, Fb<P> fb)
{
    using V = Val<C>;
    __shared__ V fixed_end[P];  // end state of the most recently repaired segment
    __shared__ V stash[kMaxSeg];
    __shared__ int32_t fixed_seg;
    const uint32_t lane = threadIdx.x, S = g.seg, nblk = (g.n + kLanes * S - 1) / (kLanes * S);
    const uint32_t nseg = (g.n + S - 1) / S;
    if (lane == 0)
        fixed_seg = -1;
    auto boundary_bad = [&](uint32_t b) {  // block b's first segment against its predecessor
        const uint32_t s = b * kLanes;
        bool d = false;
#pragma unroll
        for (int k = 0; k < P; ++k)
            d |= !same(from2<C>(g.rs[(size_t)s * P + k]), from2<C>(g.re[(size_t)(s - 1) * P + k]));
        return d;
    };
    constexpr uint32_t kU = P >= 8 ? 1u : 8u / P;  // boundary checks in flight per lane
    bool any = false;
    for (uint32_t b0 = 1; g.verify && b0 < nblk; b0 += kLanes * kU) {
        bool d[kU];
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u) {
            const uint32_t b = b0 + u * kLanes + lane;
            d[u] = b < nblk && boundary_bad(b);
        }
#pragma unroll
        for (uint32_t u = 0; u < kU; ++u)
            any |= d[u];
    }
    any = g.verify && (__ballot(any) != 0ull || atomicAdd(g.failed, 0u) != 0u);
    __syncthreads();

    if (any) {
        // Repair, in segment order. Segments below `next` are final. A set bit
        // whose predecessor was never repaired is a real failure; after a
        // repair the successor is checked against the new end state directly.
        auto repair = [&](uint32_t f) {
            const uint32_t p0 = f * S, cnt = min(S, g.n - p0);
            stash[lane] = lane < cnt ? bpart<TIN, C, RA>(g, (int32_t)(p0 + lane)) : vzero<C>();
            __syncthreads();
            if (lane == 0) {
                V st[P];
#pragma unroll
                for (int k = 0; k < P; ++k)
                    st[k] = fixed_seg == (int32_t)f - 1 ? fixed_end[k] : from2<C>(g.re[(size_t)(f - 1) * P + k]);
                for (uint32_t i = 0; i < cnt; ++i)
                    stash[i] = fb_step<P, C, RA>(stash[i], st, fb);
#pragma unroll
                for (int k = 0; k < P; ++k) {
                    fixed_end[k] = st[k];
                    g.re[(size_t)f * P + k] = to2(st[k]);
                }
                fixed_seg = (int32_t)f;
            }
            __syncthreads();
            if (lane < cnt)
                put_out<C>(g, p0 + lane, stash[lane]);
            __syncthreads();
        };
        uint32_t next = 1;
        for (uint32_t w0 = 0; w0 < nblk; w0 += kLanes) {
            const uint32_t b = w0 + lane;
            unsigned long long w = 0ull;
            if (b < nblk) {
                w = g.bad[b];
                if (b > 0 && boundary_bad(b))
                    w |= 1ull;
            }
            unsigned long long nz = __ballot(w != 0ull);
            while (nz) {
                const int l = __ffsll((long long)nz) - 1;
                nz &= nz - 1;
                unsigned long long bits = shfl64(w, l);
                while (bits) {
                    const uint32_t f0 = (w0 + (uint32_t)l) * kLanes + (uint32_t)(__ffsll((long long)bits) - 1);
                    bits &= bits - 1;
                    if (f0 < next)
                        continue;
                    uint32_t f = f0;
                    for (;;) {
                        repair(f);
                        if (++f >= nseg)
                            break;
                        bool ok = true;
#pragma unroll
                        for (int k = 0; k < P; ++k)
                            ok &= same(from2<C>(g.rs[(size_t)f * P + k]), fixed_end[k]);
                        if (ok)
                            break;
                    }
                    next = f + 1;
                }
            }
        }
    }

    // carried outputs: the last segment's end state
    const uint32_t last = nseg - 1;
    if (lane < (uint32_t)P)
        g.yh_next[lane] = to2(fixed_seg == (int32_t)last ? fixed_end[lane] : from2<C>(g.re[(size_t)last * P + lane]));
    carry_inputs<TIN>(g);
    if (lane == 0)
        atomicExch(g.failed, 0u);
}

template <int P, bool C, uint32_t TIN, bool RA>
__global__ __launch_bounds__(kLanes) void iir_kernel(IirLaunch g, Fb<P> fb)
{
    using V = Val<C>;
    constexpr int kB = kBatch;
    extern __shared__ __align__(16) unsigned char lds_raw[];
    V *vs = reinterpret_cast<V *>(lds_raw);
    __shared__ V ends[P][kLanes];
    const uint32_t lane = threadIdx.x;
    const uint32_t S = g.seg;
    const uint32_t first = blockIdx.x * kLanes * S;
    const uint32_t cnt = min(kLanes * S, g.n - first);
    const uint32_t span = cnt + g.warm;
    const int32_t lo = (int32_t)first - (int32_t)g.warm;  // vs slot(q) holds position lo + q

    // 1. b-part of every position the block's recurrences read
    stage_bpart<TIN, C, RA>(g, vs, lo, span);

    // 2. each lane's recurrence. The segment's outputs overwrite its own
    // b-part slots. Other lanes read those slots only in their warm-up loop,
    // which the wave finishes before any lane enters the segment loop.
    const uint32_t s0 = first + lane * S;
    const bool active = s0 < g.n;
    V st[P], st0[P];
    if (active) {
        const uint32_t s1 = min(g.n, s0 + S);
        uint32_t j = s0 > g.warm ? s0 - g.warm : 0u;
#pragma unroll
        for (int k = 0; k < P; ++k)
            st[k] = j == 0 ? from2<C>(g.yh[k]) : vzero<C>();
        for (; j < s0; j += kB) {  // s0 - j and j - lo are multiples of kB
            V v[kB];
            lds_read<C>(vs + slot<C>((uint32_t)((int32_t)j - lo)), v);
#pragma unroll
            for (int u = 0; u < kB; ++u)
                (void)fb_step<P, C, RA>(v[u], st, fb);
        }
#pragma unroll
        for (int k = 0; k < P; ++k)
            st0[k] = st[k];
        for (; j + kB <= s1; j += kB) {
            V v[kB];
            V *p = vs + slot<C>((uint32_t)((int32_t)j - lo));
            lds_read<C>(p, v);
#pragma unroll
            for (int u = 0; u < kB; ++u)
                v[u] = fb_step<P, C, RA>(v[u], st, fb);
            lds_write<C>(p, v);
        }
        for (; j < s1; ++j) {
            V *p = vs + slot<C>((uint32_t)((int32_t)j - lo));
            *p = fb_step<P, C, RA>(*p, st, fb);
        }
        // records: every segment's when verifying, else the last one's (the carried state)
        const size_t seg = s0 / S;
        const bool keep = g.verify || s1 == g.n;
#pragma unroll
        for (int k = 0; k < P; ++k) {
            if (keep) {
                g.rs[seg * P + k] = to2(st0[k]);
                g.re[seg * P + k] = to2(st[k]);
            }
            ends[k][lane] = st[k];
        }
    }
    __syncthreads();

    // 3. outputs, coalesced
    for (uint32_t q = lane; q < cnt; q += kLanes)
        put_out<C>(g, first + q, vs[slot<C>(q + g.warm)]);

    // 4. segments whose start state differs from the previous lane's end state
    bool bad = false;
    if (g.verify && active && lane > 0) {
#pragma unroll
        for (int k = 0; k < P; ++k)
            bad |= !same(st0[k], ends[k][lane - 1]);
    }
    const unsigned long long m = __ballot(bad);
    if (lane == 0) {
        g.bad[blockIdx.x] = m;
        if (m)
            atomicAdd(g.failed, 1u);
    }
    // verifying: iir_finish_kernel, next on the stream, checks across blocks,
    // repairs and carries the state. Unverified: the block holding the
    // chunk's last segment carries it (its lane's end state is final).
    if (!g.verify && blockIdx.x == gridDim.x - 1) {
        if (active && s0 + S >= g.n) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                g.yh_next[k] = to2(st[k]);
        }
        carry_inputs<TIN>(g);
    }
}

// The sequential form, for filters whose segments do not merge: one wave
// walks the chunk in tiles. All lanes form a tile's b-part in LDS, lane 0 runs
// the recurrence over it in place, all lanes write the outputs.
constexpr uint32_t kTile = 4096;

template <int P, bool C, uint32_t TIN>
__global__ __launch_bounds__(kLanes) void iir_serial_kernel(IirLaunch g, Fb<P> fb)
{
    using V = Val<C>;
    constexpr int kB = kBatch;
    __shared__ __align__(16) V tile[slots<C>(kTile + kMaxB - 1)];
    const uint32_t lane = threadIdx.x;
    V st[P];
    float sr[P];  // the split form's real state: lane 0 the real parts, lane 1 the imaginary
#pragma unroll
    for (int k = 0; k < P; ++k) {
        st[k] = from2<C>(g.yh[k]);
        sr[k] = lane == 1 ? g.yh[k].y : g.yh[k].x;
    }
    for (uint32_t t0 = 0; t0 < g.n; t0 += kTile) {
        const uint32_t cnt = min(kTile, g.n - t0);
        stage_bpart<TIN, C>(g, tile, (int32_t)t0, cnt);
        __syncthreads();
        if constexpr (C) {
            if (g.split) {
                // real feedback on complex samples: the real and imaginary
                // parts are two independent real recurrences (the products
                // with the zero imaginary parts of a only add signed zeros)
                if (lane < 2) {
                    uint32_t q = 0;
                    for (; q + 16 <= cnt; q += 16) {  // 16 values never cross a padding gap
                        float *part = reinterpret_cast<float *>(tile + slot<C>(q)) + lane;
                        float v[16];
#pragma unroll
                        for (int u = 0; u < 16; ++u)
                            v[u] = part[2 * u];
#pragma unroll
                        for (int u = 0; u < 16; ++u)
                            part[2 * u] = fb_step<P, false>(v[u], sr, fb);
                    }
                    for (; q < cnt; ++q) {
                        float *part = reinterpret_cast<float *>(tile + slot<C>(q)) + lane;
                        *part = fb_step<P, false>(*part, sr, fb);
                    }
                }
                __syncthreads();
                for (uint32_t q = lane; q < cnt; q += kLanes)
                    put_out<C>(g, t0 + q, tile[slot<C>(q)]);
                __syncthreads();
                continue;
            }
        }
        if (lane == 0) {
            uint32_t q = 0;
            for (; q + kB <= cnt; q += kB) {
                V v[kB];
                lds_read<C>(tile + slot<C>(q), v);
#pragma unroll
                for (int u = 0; u < kB; ++u)
                    v[u] = fb_step<P, C>(v[u], st, fb);
                lds_write<C>(tile + slot<C>(q), v);
            }
            for (; q < cnt; ++q)
                tile[slot<C>(q)] = fb_step<P, C>(tile[slot<C>(q)], st, fb);
        }
        __syncthreads();
        for (uint32_t q = lane; q < cnt; q += kLanes)
            put_out<C>(g, t0 + q, tile[slot<C>(q)]);
        __syncthreads();
    }
    if (C && g.split) {
        if (lane < 2) {
#pragma unroll
            for (int k = 0; k < P; ++k)
                reinterpret_cast<float *>(g.yh_next + k)[lane] = sr[k];
        }
    } else if (lane == 0) {
#pragma unroll
        for (int k = 0; k < P; ++k)
            g.yh_next[k] = to2(st[k]);
    }
    carry_inputs<TIN>(g);
}

template <int P, bool C, uint32_t TIN, bool RA>
hipError_t launch_p(const IirLaunch &g, const float2 *a, bool serial, hipStream_t s)
{
    Fb<P> fb;
    for (int k = 0; k < P; ++k)
        fb.a[k] = a[k];
    if (serial) {
        hipLaunchKernelGGL((iir_serial_kernel<P, C, TIN>), dim3(1), dim3(kLanes), 0, s, g, fb);
        return hipGetLastError();
    }
    const uint32_t blocks = (g.n + kLanes * g.seg - 1) / (kLanes * g.seg);
    const size_t lds = slots<C>(kLanes * g.seg + g.warm + g.nb - 1) * sizeof(Val<C>);
    hipLaunchKernelGGL((iir_kernel<P, C, TIN, RA>), dim3(blocks), dim3(kLanes), lds, s, g, fb);
    if (g.verify)
        hipLaunchKernelGGL((iir_finish_kernel<P, C, TIN, RA>), dim3(1), dim3(kLanes), 0, s, g, fb);
    return hipGetLastError();
}

template <bool C, uint32_t TIN, bool RA>
hipError_t launch_pa(const IirLaunch &g, uint32_t p, const float2 *a, bool serial, hipStream_t s)
{
    switch (p) {
    case 1:
        return launch_p<1, C, TIN, RA>(g, a, serial, s);
    case 2:
        return launch_p<2, C, TIN, RA>(g, a, serial, s);
    case 4:
        return launch_p<4, C, TIN, RA>(g, a, serial, s);
    case 8:
        return launch_p<8, C, TIN, RA>(g, a, serial, s);
    default:
        return launch_p<16, C, TIN, RA>(g, a, serial, s);
    }
}

// complex samples with real coefficients take the two-real-recurrence form
template <bool C, uint32_t TIN>
hipError_t launch_c(const IirLaunch &g, uint32_t p, const float2 *a, bool serial, hipStream_t s)
{
    if constexpr (C)
        if (g.split)
            return launch_pa<C, TIN, true>(g, p, a, serial, s);
    return launch_pa<C, TIN, false>(g, p, a, serial, s);
}

// the reference's std::complex<float> division (g++: Smith's method in
// double, rounded to float once)
std::complex<float> cdiv(std::complex<float> x, std::complex<float> y)
{
    const double a = x.real(), b = x.imag(), c = y.real(), d = y.imag();
    if (std::fabs(c) < std::fabs(d)) {
        const double r = c / d, den = c * r + d;
        return {(float)((a * r + b) / den), (float)((b * r - a) / den)};
    }
    const double r = d / c, den = d * r + c;
    return {(float)((b * r + a) / den), (float)((b - a * r) / den)};
}

// samples until the homogeneous recurrence shrinks a unit state below
// `floor` (two starting states, so a mode missing from one is seen by the
// other): 2^-44 for segments that must meet the exact trajectory, 2^-26 for
// unverified ones (a start error that far below float rounding)
uint32_t warm_for(const std::vector<float2> &a, uint32_t p, bool *decays, double floor = 0x1p-44)
{
    *decays = true;
    bool zero = true;
    for (uint32_t k = 0; k < p; ++k)
        zero &= a[k].x == 0.f && a[k].y == 0.f;
    if (zero)
        return kMinWarm;
    uint32_t worst = 0;
    for (int trial = 0; trial < 2; ++trial) {
        std::vector<std::complex<double>> y(p);
        for (uint32_t k = 0; k < p; ++k)
            y[k] = trial ? std::complex<double>(k & 1 ? -0.5 : 1.0, k & 2 ? 0.75 : -0.25) : 1.0;
        uint32_t k = 1;
        for (; k <= kMaxWarm; ++k) {
            std::complex<double> v = 0.0;
            for (uint32_t j = 0; j < p; ++j)
                v -= std::complex<double>(a[j].x, a[j].y) * y[j];
            for (uint32_t j = p - 1; j > 0; --j)
                y[j] = y[j - 1];
            y[0] = v;
            double mx = 0.0;
            for (uint32_t j = 0; j < p; ++j)
                mx = std::max(mx, std::abs(y[j]));
            if (mx < floor)
                break;
        }
        worst = std::max(worst, k);
    }
    *decays = worst <= kMaxWarm;
    return std::min(kMaxWarm, std::max(kMinWarm, (worst + 31) / 32 * 32));
}

// Whether zero-started segments meet the true trajectory bit for bit within
// `warm` samples, tried on `trials` starts (at most `max_miss` may miss) in a
// deterministic noise input with the GPU's rounding (complex products and
// sums in float, no fused multiply-add).
// Filters whose rounding differences are amplified by their feedback keep a
// few-ulp gap for ever; those run sequentially.
bool segments_merge(const std::vector<float2> &b, const std::vector<float2> &a, uint32_t p, uint32_t warm,
                    uint32_t trials = 16, uint32_t max_miss = 1)
{
#pragma clang fp contract(off)
    using cx = std::complex<float>;
    auto mul = [](cx x, cx y) {
        return cx(x.real() * y.real() - x.imag() * y.imag(), x.real() * y.imag() + x.imag() * y.real());
    };
    auto step = [&](cx v, std::vector<cx> &st) {
        cx y = v;
        for (uint32_t k = 0; k < p; ++k) {
            const cx q = mul(st[k], cx(a[k].x, a[k].y));
            y = cx(y.real() - q.real(), y.imag() - q.imag());
        }
        st.insert(st.begin(), y);
        st.pop_back();
    };
    const uint32_t first = 256, gap = 96, n = first + trials * gap + warm;
    std::vector<cx> v(n);
    uint64_t r = 0x9e3779b97f4a7c15ull;
    std::vector<cx> x(n);
    for (uint32_t i = 0; i < n; ++i) {
        auto u = [&]() {
            r = r * 6364136223846793005ull + 1442695040888963407ull;
            return (float)((double)(r >> 11) * 0x1p-53 * 2.0 - 1.0);
        };
        const float re = u(), im = u();
        x[i] = cx(re, im);
        cx acc(0.f, 0.f);
        for (uint32_t k = 0; k < b.size() && k <= i; ++k) {
            const cx q = mul(x[i - k], cx(b[k].x, b[k].y));
            acc = cx(acc.real() + q.real(), acc.imag() + q.imag());
        }
        v[i] = acc;
    }
    std::vector<std::vector<cx>> truth(n);
    std::vector<cx> st(p, cx(0.f, 0.f));
    for (uint32_t i = 0; i < n; ++i) {
        step(v[i], st);
        truth[i] = st;
    }
    auto same = [](cx u, cx w) {
        return (u.real() == w.real() || std::memcmp(&u, &w, 4) == 0) &&
               (u.imag() == w.imag() || std::memcmp(reinterpret_cast<const float *>(&u) + 1,
                                                    reinterpret_cast<const float *>(&w) + 1, 4) == 0);
    };
    uint32_t met = 0;
    for (uint32_t t = 0; t < trials; ++t) {
        const uint32_t j0 = first + t * gap;
        std::vector<cx> z(p, cx(0.f, 0.f));
        for (uint32_t i = j0; i < j0 + warm; ++i)
            step(v[i], z);
        bool eq = true;
        for (uint32_t k = 0; k < p; ++k)
            eq &= same(z[k], truth[j0 + warm - 1][k]);
        met += eq;
    }
    return met + max_miss >= trials;  // the device verifies every segment regardless
}

// The shortest warm-up, in steps of 32 between the 2^-30 decay and `warm`
// (the 2^-44 decay), whose zero-started segments meet the trajectory on all
// of 512 trial starts (bisection; meeting is monotone in the warm-up). The
// 2^-44 rule is conservative: a start error only has to fall below half an
// ulp of the state, and the de-emphasis (a pole at 0.936) meets 100% of
// trials at 384 samples against 480 for the rule. Segments that still miss
// are caught by the device's verification and recomputed.
uint32_t shortest_merging_warm(const std::vector<float2> &b, const std::vector<float2> &a, uint32_t p, uint32_t warm)
{
    bool unused;
    uint32_t lo = warm_for(a, p, &unused, 0x1p-30) / 32, hi = warm / 32;  // hi is taken to meet
    if (lo >= hi)
        return warm;
    if (segments_merge(b, a, p, lo * 32, 512, 0))
        return lo * 32;
    while (hi - lo > 1) {  // lo misses, hi meets
        const uint32_t mid = (lo + hi) / 2;
        if (segments_merge(b, a, p, mid * 32, 512, 0))
            hi = mid;
        else
            lo = mid;
    }
    return hi * 32;
}

}  // namespace

// ===========================================================================
// handle
// ===========================================================================
struct fdsp_iir_s {
    std::mutex mu;
    uint32_t nb = 0, nfb = 0, p = 1, warm = kMinWarm;
    bool real_coeffs = true;
    int mode = FDSP_IIR_AUTO;  // FDSP_IIR_AUTO, FDSP_IIR_EXACT or FDSP_IIR_FAST
    bool merges = true;        // whether verified segments are expected to hold
    bool decays = true;        // whether `warm` reaches the 2^-44 decay (else it is capped)
    uint32_t warm_fast = kMinWarm;  // the unverified segments' warm-up (2^-26 decay)
    std::vector<float2> b, a;  // normalised taps; a[1..p] (zero past nfb)
    DevBuf d_b, d_xh[2], d_yh[2], d_rs, d_re, d_bad, d_failed, d_in, d_out;
    uint32_t parity = 0;  // which copy of the carried state is current
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_iir_create(const fdsp_cf32 *b, uint32_t nb, const fdsp_cf32 *a, uint32_t na, fdsp_iir_t *out)
{
    if (!out || !b || !a)
        return set_err(FDSP_EINVAL, "iir_create: null argument");
    if (nb == 0 || na == 0)
        return set_err(FDSP_EINVAL, "iir_create: need at least one b and one a coefficient");
    if (nb > kMaxB || na - 1 > kMaxFb)
        return set_err(FDSP_EINVAL, "iir_create: at most %u b and %u + 1 a coefficients", kMaxB, kMaxFb);
    if (a[0].re == 0.f && a[0].im == 0.f)
        return set_err(FDSP_EINVAL, "iir_create: a[0] must be nonzero");
    auto *h = new (std::nothrow) fdsp_iir_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->nb = nb;
    h->nfb = na - 1;
    h->p = 1;
    while (h->p < h->nfb)
        h->p *= 2;
    // iir_filter.cc:137-150: divide through by a[0] unless its real part is 1
    const std::complex<float> a0(a[0].re, a[0].im);
    const bool scale = a0.real() != 1.0f;
    for (uint32_t k = 0; k < nb; ++k) {
        std::complex<float> v(b[k].re, b[k].im);
        if (scale)
            v = cdiv(v, a0);
        h->b.push_back(make_float2(v.real(), v.imag()));
    }
    h->a.assign(h->p, make_float2(0.f, 0.f));
    for (uint32_t k = 1; k < na; ++k) {
        std::complex<float> v(a[k].re, a[k].im);
        if (scale)
            v = cdiv(v, a0);
        h->a[k - 1] = make_float2(v.real(), v.imag());
    }
    for (const float2 &v : h->b)
        h->real_coeffs &= v.y == 0.f;
    for (const float2 &v : h->a)
        h->real_coeffs &= v.y == 0.f;
    h->warm = warm_for(h->a, h->p, &h->decays);
    bool unused;
    h->warm_fast = warm_for(h->a, h->p, &unused, 0x1p-26);
    h->merges = segments_merge(h->b, h->a, h->p, h->warm);
    if (h->merges && h->decays)
        h->warm = std::min(h->warm, shortest_merging_warm(h->b, h->a, h->p, h->warm) + kWarmMargin);
    int rc = h->d_b.ensure(nb * sizeof(float2));
    if (!rc)
        rc = h->d_xh[0].ensure(std::max<uint32_t>(nb - 1, 1) * sizeof(float2));
    if (!rc)
        rc = h->d_xh[1].ensure(std::max<uint32_t>(nb - 1, 1) * sizeof(float2));
    if (!rc)
        rc = h->d_yh[0].ensure(kMaxFb * sizeof(float2));
    if (!rc)
        rc = h->d_yh[1].ensure(kMaxFb * sizeof(float2));
    if (!rc)
        rc = h->d_failed.ensure(sizeof(uint32_t));
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemcpy(h->d_b.p, h->b.data(), nb * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(h->d_xh[0].p, 0, h->d_xh[0].bytes) != hipSuccess ||
                hipMemset(h->d_yh[0].p, 0, h->d_yh[0].bytes) != hipSuccess ||
                hipMemset(h->d_failed.p, 0, h->d_failed.bytes) != hipSuccess))
        rc = set_err(FDSP_EHIP, "iir_create: HIP setup failed");
    if (rc) {
        fdsp_iir_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_iir_destroy(fdsp_iir_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    for (DevBuf *d : {&h->d_b, &h->d_xh[0], &h->d_xh[1], &h->d_yh[0], &h->d_yh[1], &h->d_rs, &h->d_re, &h->d_bad,
                      &h->d_failed, &h->d_in, &h->d_out})
        d->release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_iir_reset(fdsp_iir_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    // launches on caller streams may still write the copy being cleared
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(h->d_xh[h->parity].p, 0, h->d_xh[h->parity].bytes));
    HIP_TRY(hipMemset(h->d_yh[h->parity].p, 0, h->d_yh[h->parity].bytes));
    return FDSP_OK;
}

// The plan a call with outputs of layout tout runs. Exact: verified segments
// when they merge, else one sequential wave. Fast: segments unverified.
// Auto: verified segments when they merge; otherwise float outputs take the
// unverified segments (float rounding, like the library's other float
// kernels) if the warm-up reaches its decay target, and int16 outputs (or a
// filter that does not decay) the sequential wave.
static uint32_t plan_of(const fdsp_iir_s *h, uint32_t tout)
{
    if (h->mode == FDSP_IIR_FAST)
        return FDSP_IIR_PLAN_FAST;
    if (h->merges)
        return FDSP_IIR_PLAN_SEGMENTS;
    const bool float_out = tout == kSampleCf32 || tout == kSampleF32;
    if (h->mode == FDSP_IIR_AUTO && float_out && h->decays)
        return FDSP_IIR_PLAN_FAST;
    return FDSP_IIR_PLAN_SERIAL;
}

extern "C" int fdsp_iir_info(fdsp_iir_t h, uint32_t *warm, uint32_t *order, uint32_t *plan)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    if (warm)
        *warm = h->warm;
    if (order)
        *order = h->nfb;
    if (plan)
        *plan = plan_of(h, kSampleCf32);
    return FDSP_OK;
}

extern "C" int fdsp_iir_set_mode(fdsp_iir_t h, int mode)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (mode != FDSP_IIR_FAST && mode != FDSP_IIR_EXACT && mode != FDSP_IIR_AUTO)
        return set_err(FDSP_EINVAL, "iir_set_mode: FDSP_IIR_AUTO, FDSP_IIR_EXACT or FDSP_IIR_FAST");
    std::lock_guard<std::mutex> lk(h->mu);
    h->mode = mode;
    return FDSP_OK;
}

namespace fdsp {

int iir_apply_typed(fdsp_iir_t h, uint32_t tin, const void *d_in, size_t n, uint32_t tout, void *d_out,
                    hipStream_t s, void *work, size_t work_bytes)
{
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!d_in || !d_out)
        return set_err(FDSP_EINVAL, "null buffers");
    if (int rc = check_chunk(n, 1))
        return rc;
    // outputs per lane: short segments give short chunks enough waves, long
    // ones keep the warm-up overhead down on long chunks (8 below 2^21
    // samples: measured fastest on 2^20-sample chunks of a 4th-order filter)
    uint32_t seg = 8;
    while (seg < kMaxSeg && (size_t)seg * 2 * 131072 < n)
        seg *= 2;
    const size_t nseg = (n + seg - 1) / seg, nblk = (n + kLanes * seg - 1) / (kLanes * seg);
    // every record is written by this call's iir_kernel before anything reads
    // it, so the buffers need no clearing
    const size_t rec_bytes = (nseg * h->p * sizeof(float2) + 255) & ~(size_t)255;
    float2 *rs, *re;
    unsigned long long *bad;
    if (work) {
        const size_t need = 2 * rec_bytes + nblk * sizeof(unsigned long long);
        if (need > work_bytes)
            return set_err(FDSP_ERANGE,
                           "iir: a %zu-sample chunk of an order-%u filter needs %zu bytes of work memory, %zu given",
                           n, h->nfb, need, work_bytes);
        rs = static_cast<float2 *>(work);
        re = reinterpret_cast<float2 *>(static_cast<unsigned char *>(work) + rec_bytes);
        bad = reinterpret_cast<unsigned long long *>(static_cast<unsigned char *>(work) + 2 * rec_bytes);
    } else {
        if (int rc = h->d_rs.ensure(nseg * h->p * sizeof(float2)))
            return rc;
        if (int rc = h->d_re.ensure(nseg * h->p * sizeof(float2)))
            return rc;
        if (int rc = h->d_bad.ensure(nblk * sizeof(unsigned long long)))
            return rc;
        rs = h->d_rs.as<float2>();
        re = h->d_re.as<float2>();
        bad = h->d_bad.as<unsigned long long>();
    }
    IirLaunch g{};
    g.x = d_in;
    g.out = d_out;
    g.b = h->d_b.as<float2>();
    g.xh = h->d_xh[h->parity].as<float2>();
    g.yh = h->d_yh[h->parity].as<float2>();
    g.xh_next = h->d_xh[h->parity ^ 1].as<float2>();
    g.yh_next = h->d_yh[h->parity ^ 1].as<float2>();
    g.rs = rs;
    g.re = re;
    g.bad = bad;
    g.failed = h->d_failed.as<uint32_t>();
    g.n = (uint32_t)n;
    g.nb = h->nb;
    g.warm = h->warm;
    g.seg = seg;
    g.tout = tout;
    const uint32_t plan = plan_of(h, tout);
    g.verify = plan == FDSP_IIR_PLAN_FAST ? 0u : 1u;
    if (plan == FDSP_IIR_PLAN_FAST)
        g.warm = h->warm_fast;
    const bool serial = plan == FDSP_IIR_PLAN_SERIAL;
    g.split = h->real_coeffs ? 1u : 0u;
    const float2 *a = h->a.data();
    const bool real_in = tin == kSampleF32 || tin == kSampleI16;
    hipError_t e;
    if (h->real_coeffs && real_in)
        e = tin == kSampleF32 ? launch_c<false, kSampleF32>(g, h->p, a, serial, s)
                              : launch_c<false, kSampleI16>(g, h->p, a, serial, s);
    else if (tin == kSampleCf32)
        e = launch_c<true, kSampleCf32>(g, h->p, a, serial, s);
    else if (tin == kSampleCi16)
        e = launch_c<true, kSampleCi16>(g, h->p, a, serial, s);
    else if (tin == kSampleF32)
        e = launch_c<true, kSampleF32>(g, h->p, a, serial, s);
    else
        e = launch_c<true, kSampleI16>(g, h->p, a, serial, s);
    HIP_TRY(e);
    h->parity ^= 1;  // the launch wrote the next carried state into the other copy
    return FDSP_OK;
}

}  // namespace fdsp

namespace {

size_t sample_bytes(uint32_t t)
{
    return t == kSampleCf32 ? 8 : t == kSampleCi16 || t == kSampleF32 ? 4 : 2;
}

int apply_device(fdsp_iir_t h, uint32_t t, const void *d_in, size_t n, void *d_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    return iir_apply_typed(h, t, d_in, n, t, d_out, pick(h->stream, stream));
}

int apply_host(fdsp_iir_t h, uint32_t t, const void *in, size_t n, void *out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return set_err(FDSP_EINVAL, "empty input");
    if (!in || !out)
        return set_err(FDSP_EINVAL, "null buffers");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t bytes = n * sample_bytes(t);
    if (int rc = h->d_in.ensure(bytes))
        return rc;
    if (int rc = h->d_out.ensure(bytes))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, bytes, hipMemcpyHostToDevice, h->stream));
    if (int rc = iir_apply_typed(h, t, h->d_in.p, n, t, h->d_out.p, h->stream))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, h->d_out.p, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_iir_apply(fdsp_iir_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out)
{
    return apply_host(h, kSampleCf32, in, n, out);
}
extern "C" int fdsp_iir_apply_f32(fdsp_iir_t h, const float *in, size_t n, float *out)
{
    return apply_host(h, kSampleF32, in, n, out);
}
extern "C" int fdsp_iir_apply_ci16(fdsp_iir_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out)
{
    return apply_host(h, kSampleCi16, in, n, out);
}
extern "C" int fdsp_iir_apply_i16(fdsp_iir_t h, const int16_t *in, size_t n, int16_t *out)
{
    return apply_host(h, kSampleI16, in, n, out);
}
extern "C" int fdsp_iir_apply_device(fdsp_iir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                     fdsp_stream_t stream)
{
    return apply_device(h, kSampleCf32, d_in, n, d_out, stream);
}
// ---------------------------------------------------------------------------
// the form that allocates nothing
// ---------------------------------------------------------------------------
namespace {

// Work memory of fdsp_iir_apply_device_noalloc: one copy of this array on
// every device that runs the library, held until the process ends. All such
// calls on one device share it, so calls that overlap on a device race.
constexpr size_t kScratchBytes = (size_t)16 << 20;
constexpr int kScratchDevices = 64;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_scratch_mu;
void *g_scratch_ptr[kScratchDevices];  // the array's address on each device, looked up once

// the current device's copy of `scratch` (the host-side name of a __device__
// array is no device address)
int scratch_on_current_device(void **out)
{
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kScratchDevices)
        return set_err(FDSP_EINVAL, "iir noalloc: device %d beyond the %d supported", dev, kScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p)
            return set_err(FDSP_EHIP, "iir noalloc: no device address for the work array");
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_iir_apply_device_noalloc(fdsp_iir_t h, const fdsp_cf32 *d_in, size_t n, fdsp_cf32 *d_out,
                                             fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    void *work = nullptr;
    if (int rc = scratch_on_current_device(&work))
        return rc;
    std::lock_guard<std::mutex> lk(h->mu);
    return iir_apply_typed(h, kSampleCf32, d_in, n, kSampleCf32, d_out, pick(h->stream, stream), work,
                           kScratchBytes);
}

extern "C" int fdsp_iir_apply_f32_device(fdsp_iir_t h, const float *d_in, size_t n, float *d_out,
                                         fdsp_stream_t stream)
{
    return apply_device(h, kSampleF32, d_in, n, d_out, stream);
}
extern "C" int fdsp_iir_apply_ci16_device(fdsp_iir_t h, const fdsp_ci16 *d_in, size_t n, fdsp_ci16 *d_out,
                                          fdsp_stream_t stream)
{
    return apply_device(h, kSampleCi16, d_in, n, d_out, stream);
}
extern "C" int fdsp_iir_apply_i16_device(fdsp_iir_t h, const int16_t *d_in, size_t n, int16_t *d_out,
                                         fdsp_stream_t stream)
{
    return apply_device(h, kSampleI16, d_in, n, d_out, stream);
}

// ===========================================================================
// first-order IIR (the FM de-emphasis form): a general handle with
// b = [b0], a = [1, a1]
// ===========================================================================
struct fdsp_iir1_s {
    fdsp_iir_t iir = nullptr;
};

extern "C" int fdsp_iir1_create(float b0, float a1, fdsp_iir1_t *out)
{
    if (!out)
        return set_err(FDSP_EINVAL, "null output");
    auto *h = new (std::nothrow) fdsp_iir1_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    const fdsp_cf32 b[1] = {{b0, 0.f}}, a[2] = {{1.f, 0.f}, {a1, 0.f}};
    if (int rc = fdsp_iir_create(b, 1, a, 2, &h->iir)) {
        delete h;
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_iir1_destroy(fdsp_iir1_t h)
{
    if (!h)
        return FDSP_OK;
    fdsp_iir_destroy(h->iir);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_iir1_reset(fdsp_iir1_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    return fdsp_iir_reset(h->iir);
}

extern "C" int fdsp_iir1_apply_device(fdsp_iir1_t h, const float *d_in, size_t n, float *d_out, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    return fdsp_iir_apply_f32_device(h->iir, d_in, n, d_out, stream);
}

extern "C" int fdsp_iir1_apply(fdsp_iir1_t h, const float *in, size_t n, float *out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    return fdsp_iir_apply_f32(h->iir, in, n, out);
}

// ===========================================================================
// IIR bank: many independent streams through one filter. One stream's exact
// recurrence is sequential, so each stream runs on one lane. A block of four
// waves owns 64 streams and walks their samples in tiles of kBankT as a
// pipeline: in phase q, wave 0 runs the recurrence of tile q-2 (lane =
// stream, in place over the tile's b-parts), while waves 1-3 write tile q's
// inputs (loaded in phase q-1) and its carried inputs into LDS, issue the
// loads of tile q+1, form the b-parts of tile q-1 and store the outputs of
// tile q-3, all with accesses coalesced along each stream's row. A barrier
// ends each phase. The arithmetic is that of the single-stream kernels (mac,
// fb_step): outputs equal the reference's loop (falcon_dsp_iir_filter.cc:
// 235-396) run on each stream, up to the sign of zero.
// ===========================================================================
namespace {

constexpr uint32_t kBankT = 16;          // samples per tile
constexpr uint32_t kBankMaxB = 17;       // b taps: up to 16 carried inputs (LDS below 64 KiB)
constexpr uint32_t kBankH = kBankMaxB - 1;  // carried-input slots per stream in the state array
constexpr uint32_t kBankVRow = kBankT + 1;  // LDS row of b-parts / outputs (odd)
constexpr uint32_t kBankProd = 192;      // producer threads (waves 1-3)
constexpr uint32_t kBankPer = (64 * kBankT + kBankProd - 1) / kBankProd;  // tile entries per producer

// LDS row of inputs: h carried inputs, then the tile (odd length)
__host__ __device__ constexpr uint32_t bank_xrow(uint32_t h) { return (h + kBankT) | 1u; }
__host__ __device__ constexpr size_t bank_lds_bytes(uint32_t h)
{
    return (size_t)(2 * 64 * bank_xrow(h) + 3 * 64 * kBankVRow) * sizeof(float2);
}

struct BankLaunch {
    const void *in;
    void *out;
    size_t in_stride, out_stride;  // samples from one stream's row to the next
    uint32_t n, n_streams, nb;
    const float2 *b;  // nb taps
    float2 *xh;       // per stream kBankH carried inputs, oldest first (the last nb - 1 used)
    float2 *yh;       // per stream kMaxFb carried outputs, newest first
};

template <uint32_t T> __device__ __forceinline__ float2 bank_load(const void *p, size_t i)
{
    if constexpr (T == kSampleCf32) {
        return static_cast<const float2 *>(p)[i];
    } else {
        const short2 s = static_cast<const short2 *>(p)[i];
        return make_float2((float)s.x, (float)s.y);
    }
}
template <uint32_t T> __device__ __forceinline__ void bank_store(void *p, size_t i, float2 v)
{
    if constexpr (T == kSampleCf32)
        static_cast<float2 *>(p)[i] = v;
    else
        static_cast<short2 *>(p)[i] = make_short2(i16_trunc(v.x), i16_trunc(v.y));
}

template <int P, bool RA, uint32_t TS>
__global__ __launch_bounds__(256) void iir_bank_kernel(BankLaunch g, Fb<P> fb)
{
    extern __shared__ float2 bank_lds[];
    const uint32_t h = g.nb - 1, xrow = bank_xrow(h);
    float2 *xt = bank_lds;                     // [2][64][xrow]: tile q in xt + (q % 2) * 64 * xrow
    float2 *vt = bank_lds + 2 * 64 * xrow;     // [3][64][kBankVRow]: tile q in vt + (q % 3) * 64 * kBankVRow
    const uint32_t tid = threadIdx.x, s0 = blockIdx.x * 64;
    const uint32_t rows = min(64u, g.n_streams - s0);
    const uint32_t nt = (g.n + kBankT - 1) / kBankT;
    const auto tile_cnt = [&](uint32_t q) { return min(kBankT, g.n - q * kBankT); };
    __shared__ float2 bsh[kBankMaxB];  // the b taps, read by all lanes at once
    if (tid < g.nb)
        bsh[tid] = g.b[tid];
    float2 st[P];  // wave 0: the stream's carried outputs, newest first
    // waves 1-3: the inputs of the next two tiles in flight (tile q in
    // pf[q % 2]), so each load has two phases to land
    float2 pf0[kBankPer], pf1[kBankPer];
    if (tid < 64) {
#pragma unroll
        for (int k = 0; k < P; ++k)
            st[k] = tid < rows ? g.yh[(size_t)(s0 + tid) * kMaxFb + k] : make_float2(0.f, 0.f);
    }
    const uint32_t p = tid - 64;  // producer index (waves 1-3)
    const auto load_tile = [&](uint32_t q, float2 (&pf)[kBankPer]) {
        const uint32_t t0 = q * kBankT, cnt = tile_cnt(q);
#pragma unroll
        for (uint32_t i = 0; i < kBankPer; ++i) {
            const uint32_t e = i * kBankProd + p, r = e / kBankT, c = e % kBankT;
            pf[i] = r < rows && c < cnt ? bank_load<TS>(g.in, (size_t)(s0 + r) * g.in_stride + t0 + c)
                                        : make_float2(0.f, 0.f);
        }
    };
    if (tid >= 64 && nt > 0)
        load_tile(0, pf0);
    if (tid >= 64 && nt > 1)
        load_tile(1, pf1);
    const auto phase = [&](uint32_t q, float2 (&pf)[kBankPer]) {
        if (tid < 64) {
            // the recurrence of tile q - 2 along the lane's stream
            if (q >= 2 && q - 2 < nt && tid < rows) {
                const uint32_t cnt = tile_cnt(q - 2);
                float2 *vr = vt + ((q - 2) % 3) * 64 * kBankVRow + tid * kBankVRow;
#pragma unroll 4
                for (uint32_t t = 0; t < cnt; ++t)
                    vr[t] = fb_step<P, true, RA>(vr[t], st, fb);
            }
        } else {
            if (q < nt) {
                // tile q's inputs after its carried inputs: the last h of
                // (carried inputs, tile q - 1), or the state array for tile 0
                float2 *xq = xt + (q % 2) * 64 * xrow;
                const uint32_t cnt = tile_cnt(q);
#pragma unroll
                for (uint32_t i = 0; i < kBankPer; ++i) {
                    const uint32_t e = i * kBankProd + p, r = e / kBankT, c = e % kBankT;
                    if (r < 64 && c < cnt)
                        xq[r * xrow + h + c] = pf[i];
                }
                for (uint32_t e = p; e < 64 * h; e += kBankProd) {
                    const uint32_t r = e / h, k = e - r * h;
                    if (r < rows)
                        xq[r * xrow + k] = q == 0 ? g.xh[(size_t)(s0 + r) * kBankH + kBankH - h + k]
                                                  : xt[((q - 1) % 2) * 64 * xrow + r * xrow + tile_cnt(q - 1) + k];
                }
                if (q + 2 < nt)
                    load_tile(q + 2, pf);
            }
            if (q >= 1 && q - 1 < nt) {
                // b-parts of tile q - 1: ((0 + b0 x[n]) + b1 x[n-1]) + ...,
                // the lane's entries side by side, one tap at a time
                const float2 *xq = xt + ((q - 1) % 2) * 64 * xrow;
                float2 *vq = vt + ((q - 1) % 3) * 64 * kBankVRow;
                const uint32_t cnt = tile_cnt(q - 1);
                float2 acc[kBankPer];
                uint32_t base[kBankPer];
#pragma unroll
                for (uint32_t i = 0; i < kBankPer; ++i) {
                    const uint32_t e = i * kBankProd + p, r = min(e / kBankT, 63u), c = e % kBankT;
                    acc[i] = make_float2(0.f, 0.f);
                    base[i] = r * xrow + h + c;
                }
                for (uint32_t k = 0; k < g.nb; ++k) {
                    const float2 bk = bsh[k];
#pragma unroll
                    for (uint32_t i = 0; i < kBankPer; ++i)
                        acc[i] = mac<true, RA>(acc[i], xq[base[i] - k], bk);
                }
#pragma unroll
                for (uint32_t i = 0; i < kBankPer; ++i) {
                    const uint32_t e = i * kBankProd + p, r = e / kBankT, c = e % kBankT;
                    if (r < rows && c < cnt)
                        vq[r * kBankVRow + c] = acc[i];
                }
            }
            if (q >= 3 && q - 3 < nt) {
                // outputs of tile q - 3
                const float2 *vq = vt + ((q - 3) % 3) * 64 * kBankVRow;
                const uint32_t t0 = (q - 3) * kBankT, cnt = tile_cnt(q - 3);
#pragma unroll
                for (uint32_t i = 0; i < kBankPer; ++i) {
                    const uint32_t e = i * kBankProd + p, r = e / kBankT, c = e % kBankT;
                    if (r < rows && c < cnt)
                        bank_store<TS>(g.out, (size_t)(s0 + r) * g.out_stride + t0 + c, vq[r * kBankVRow + c]);
                }
            }
        }
        __syncthreads();
    };
    for (uint32_t q = 0; q < nt + 3; q += 2) {
        phase(q, pf0);
        if (q + 1 < nt + 3)
            phase(q + 1, pf1);
    }
    // carried state: the last h inputs of (carried inputs, last tile), and the outputs
    if (tid < 64 && tid < rows) {
        const float2 *xl = xt + ((nt - 1) % 2) * 64 * xrow + tid * xrow + tile_cnt(nt - 1);
        for (uint32_t k = 0; k < h; ++k)
            g.xh[(size_t)(s0 + tid) * kBankH + kBankH - h + k] = xl[k];
#pragma unroll
        for (int k = 0; k < P; ++k)
            g.yh[(size_t)(s0 + tid) * kMaxFb + k] = st[k];
    }
}

template <int P, uint32_t TS> hipError_t launch_bank_p(const BankLaunch &g, const float2 *a, bool ra, hipStream_t s)
{
    Fb<P> fb;
    for (int k = 0; k < P; ++k)
        fb.a[k] = a[k];
    const dim3 grid((g.n_streams + 63) / 64);
    const size_t lds = bank_lds_bytes(g.nb - 1);
    if (ra)
        hipLaunchKernelGGL((iir_bank_kernel<P, true, TS>), grid, dim3(256), lds, s, g, fb);
    else
        hipLaunchKernelGGL((iir_bank_kernel<P, false, TS>), grid, dim3(256), lds, s, g, fb);
    return hipGetLastError();
}

template <uint32_t TS> hipError_t launch_bank(const BankLaunch &g, uint32_t p, const float2 *a, bool ra, hipStream_t s)
{
    switch (p) {
    case 1:
        return launch_bank_p<1, TS>(g, a, ra, s);
    case 2:
        return launch_bank_p<2, TS>(g, a, ra, s);
    case 4:
        return launch_bank_p<4, TS>(g, a, ra, s);
    case 8:
        return launch_bank_p<8, TS>(g, a, ra, s);
    default:
        return launch_bank_p<16, TS>(g, a, ra, s);
    }
}

}  // namespace

struct fdsp_iir_bank_s {
    std::mutex mu;
    uint32_t nb = 0, nfb = 0, p = 1, n_streams = 0;
    bool real_coeffs = true;
    std::vector<float2> b, a;  // normalised taps; a[1..p] (zero past nfb)
    DevBuf d_b, d_xh, d_yh, d_in, d_out;
    hipStream_t stream = nullptr;
};

extern "C" int fdsp_iir_bank_create(const fdsp_cf32 *b, uint32_t nb, const fdsp_cf32 *a, uint32_t na,
                                    uint32_t n_streams, fdsp_iir_bank_t *out)
{
    if (!out || !b || !a)
        return set_err(FDSP_EINVAL, "iir_bank_create: null argument");
    if (nb == 0 || na == 0 || n_streams == 0)
        return set_err(FDSP_EINVAL, "iir_bank_create: need b and a coefficients and at least one stream");
    if (nb > kBankMaxB || na - 1 > kMaxFb)
        return set_err(FDSP_EINVAL, "iir_bank_create: at most %u b and %u + 1 a coefficients", kBankMaxB, kMaxFb);
    if (a[0].re == 0.f && a[0].im == 0.f)
        return set_err(FDSP_EINVAL, "iir_bank_create: a[0] must be nonzero");
    auto *h = new (std::nothrow) fdsp_iir_bank_s;
    if (!h)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    h->nb = nb;
    h->nfb = na - 1;
    h->n_streams = n_streams;
    while (h->p < h->nfb)
        h->p *= 2;
    // iir_filter.cc:137-150: divide through by a[0] unless its real part is 1
    const std::complex<float> a0(a[0].re, a[0].im);
    const bool scale = a0.real() != 1.0f;
    for (uint32_t k = 0; k < nb; ++k) {
        std::complex<float> v(b[k].re, b[k].im);
        if (scale)
            v = cdiv(v, a0);
        h->b.push_back(make_float2(v.real(), v.imag()));
    }
    h->a.assign(h->p, make_float2(0.f, 0.f));
    for (uint32_t k = 1; k < na; ++k) {
        std::complex<float> v(a[k].re, a[k].im);
        if (scale)
            v = cdiv(v, a0);
        h->a[k - 1] = make_float2(v.real(), v.imag());
    }
    for (const float2 &v : h->b)
        h->real_coeffs &= v.y == 0.f;
    for (const float2 &v : h->a)
        h->real_coeffs &= v.y == 0.f;
    int rc = h->d_b.ensure(nb * sizeof(float2));
    if (!rc)
        rc = h->d_xh.ensure((size_t)n_streams * kBankH * sizeof(float2));
    if (!rc)
        rc = h->d_yh.ensure((size_t)n_streams * kMaxFb * sizeof(float2));
    if (!rc && (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess ||
                hipMemcpy(h->d_b.p, h->b.data(), nb * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess ||
                hipMemset(h->d_xh.p, 0, h->d_xh.bytes) != hipSuccess ||
                hipMemset(h->d_yh.p, 0, h->d_yh.bytes) != hipSuccess))
        rc = set_err(FDSP_EHIP, "iir_bank_create: HIP setup failed");
    if (rc) {
        fdsp_iir_bank_destroy(h);
        return rc;
    }
    *out = h;
    return FDSP_OK;
}

extern "C" int fdsp_iir_bank_destroy(fdsp_iir_bank_t h)
{
    if (!h)
        return FDSP_OK;
    if (h->stream)
        (void)hipStreamSynchronize(h->stream);
    for (DevBuf *d : {&h->d_b, &h->d_xh, &h->d_yh, &h->d_in, &h->d_out})
        d->release();
    if (h->stream)
        (void)hipStreamDestroy(h->stream);
    delete h;
    return FDSP_OK;
}

extern "C" int fdsp_iir_bank_reset(fdsp_iir_bank_t h)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    // launches on caller streams may still write the state being cleared
    HIP_TRY(hipDeviceSynchronize());
    HIP_TRY(hipMemset(h->d_xh.p, 0, h->d_xh.bytes));
    HIP_TRY(hipMemset(h->d_yh.p, 0, h->d_yh.bytes));
    return FDSP_OK;
}

namespace {

int bank_apply_device(fdsp_iir_bank_t h, uint32_t t, const void *d_in, size_t n, size_t in_stride, void *d_out,
                      size_t out_stride, hipStream_t s)
{
    if (n == 0)
        return FDSP_OK;  // nothing to filter; the state stays
    if (!d_in || !d_out)
        return set_err(FDSP_EINVAL, "iir_bank: null buffers");
    if (in_stride < n || out_stride < n)
        return set_err(FDSP_EINVAL, "iir_bank: row strides (%zu in, %zu out) below the %zu samples", in_stride,
                       out_stride, n);
    if (int rc = check_chunk(n, 1))
        return rc;
    BankLaunch g{};
    g.in = d_in;
    g.out = d_out;
    g.in_stride = in_stride;
    g.out_stride = out_stride;
    g.n = (uint32_t)n;
    g.n_streams = h->n_streams;
    g.nb = h->nb;
    g.b = h->d_b.as<float2>();
    g.xh = h->d_xh.as<float2>();
    g.yh = h->d_yh.as<float2>();
    const hipError_t e = t == kSampleCf32 ? launch_bank<kSampleCf32>(g, h->p, h->a.data(), h->real_coeffs, s)
                                          : launch_bank<kSampleCi16>(g, h->p, h->a.data(), h->real_coeffs, s);
    HIP_TRY(e);
    return FDSP_OK;
}

int bank_apply_host(fdsp_iir_bank_t h, uint32_t t, const void *in, size_t n, void *out)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    if (n == 0)
        return FDSP_OK;
    if (!in || !out)
        return set_err(FDSP_EINVAL, "iir_bank: null buffers");
    std::lock_guard<std::mutex> lk(h->mu);
    const size_t bytes = (size_t)h->n_streams * n * (t == kSampleCf32 ? 8 : 4);
    if (int rc = h->d_in.ensure(bytes))
        return rc;
    if (int rc = h->d_out.ensure(bytes))
        return rc;
    HIP_TRY(hipMemcpyAsync(h->d_in.p, in, bytes, hipMemcpyHostToDevice, h->stream));
    if (int rc = bank_apply_device(h, t, h->d_in.p, n, n, h->d_out.p, n, h->stream))
        return rc;
    HIP_TRY(hipMemcpyAsync(out, h->d_out.p, bytes, hipMemcpyDeviceToHost, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
    return FDSP_OK;
}

}  // namespace

extern "C" int fdsp_iir_bank_apply(fdsp_iir_bank_t h, const fdsp_cf32 *in, size_t n, fdsp_cf32 *out)
{
    return bank_apply_host(h, kSampleCf32, in, n, out);
}
extern "C" int fdsp_iir_bank_apply_ci16(fdsp_iir_bank_t h, const fdsp_ci16 *in, size_t n, fdsp_ci16 *out)
{
    return bank_apply_host(h, kSampleCi16, in, n, out);
}
extern "C" int fdsp_iir_bank_apply_device(fdsp_iir_bank_t h, const fdsp_cf32 *d_in, size_t n, size_t in_stride,
                                          fdsp_cf32 *d_out, size_t out_stride, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    return bank_apply_device(h, kSampleCf32, d_in, n, in_stride, d_out, out_stride, pick(h->stream, stream));
}
extern "C" int fdsp_iir_bank_apply_ci16_device(fdsp_iir_bank_t h, const fdsp_ci16 *d_in, size_t n, size_t in_stride,
                                               fdsp_ci16 *d_out, size_t out_stride, fdsp_stream_t stream)
{
    if (!h)
        return set_err(FDSP_ENOTINIT, "null handle");
    std::lock_guard<std::mutex> lk(h->mu);
    return bank_apply_device(h, kSampleCi16, d_in, n, in_stride, d_out, out_stride, pick(h->stream, stream));
}
