// fdsp_chan_groups.hip -- fused multi-rate channelizer: every channel's NCO
// shift and polyphase resample in one launch.
// (Replaces falcon_dsp_multi_rate_channelizer{,_cuda}::apply, reference
//  src/transform/falcon_dsp_multi_rate_channelizer.cc:146-164 and
//  src/transform/falcon_dsp_multi_rate_channelizer_cuda.cu:375-514.)
//
// Shift-then-filter equals filter-with-modulated-taps then one phasor per
// output: with k(a) the NCO index of the output's anchor sample a and
// d = a - i the distance of window sample i,
//   y = e^{j2pi nu k(a)} * sum_t [h_t e^{-j2pi nu d_t}] x[i_t]
// as long as no rollover (the index restarting at 0 after R-1) falls inside
// the window. When one does (k(a) < cpp-1; rare, R >= 1e5), the samples before
// it carry an extra factor e^{j2pi nu R}; such outputs are recomputed on a
// slow path.
//
// Output n of a channel reads input A(n) = floor(n*down/up) with filter phase
// (n*down) mod up; a chunk starting at global input P emits the outputs whose
// anchors fall inside it (reference polyphase_resampler.cc:256-319). Channels
// with the same up == 1 and down share their anchors.
//
// Work is (group, row tile): a block computes `rows` consecutive outputs of
// one channel group. Group kinds:
//   MFMA     2..8 channels with up 1 and equal down and padded tap count
//            (32 or 64). Their windows form one real matrix X (rows x 2cpp:
//            real then imaginary parts) and the modulated taps one matrix B
//            (2cpp x 16: columns 2c, 2c+1 are channel c's real and imaginary
//            output), so Y = X B runs on v_mfma_f32_16x16x4f32.
//   VDENSE   one channel, up 1, down < cpp: the tile's input span is staged
//            transposed by `down` ((i mod down)*mc + i/down) so the lanes,
//            which own consecutive outputs `down` samples apart, read
//            consecutive LDS words; taps are wavefront-uniform SGPRs.
//   VSPARSE  one channel, up 1, down >= cpp: the windows are gathered with
//            coalesced loads into a tap-major LDS tile, a few taps at a time.
//   GENERIC  one channel, up > 1: natural-order span, flattened outputs with
//            per-output filter phase.
// Every staging step issues its loads as one batch (a handful per thread)
// before writing LDS, so memory latency overlaps across the batch.
#include "fdsp_internal.h"

namespace fdsp {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

struct GroupsDev {
    const float2 *in;
    uint32_t n;
    uint32_t hist_len;
    const float2 *hist0, *hist1;
    float2 *histw0, *histw1;
    const ChanStatic *chans;
    const float2 *taps;
    const float *btab;
    const ChanPos *pos;
    uint64_t p;
    uint32_t parity;
    const uint32_t *groups;
    uint32_t n_groups;
    // first block of each group (n_groups + 1 entries) when n_groups <= kGroupsInline
    uint32_t prefix[kGroupsInline + 1];
};

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }

// floor(a / d) for a < 2^52 from an FP64 quotient estimate and one exact
// correction step (inv_d = 1/d rounded to double)
__device__ __forceinline__ uint64_t udiv(uint64_t a, uint32_t d, double inv_d)
{
    int64_t q = (int64_t)((double)a * inv_d);
    const int64_t r = (int64_t)a - q * (int64_t)d;
    if (r < 0)
        --q;
    else if (r >= (int64_t)d)
        ++q;
    return (uint64_t)q;
}
__device__ __forceinline__ uint64_t cdiv(uint64_t a, uint32_t d, double inv_d) { return udiv(a + d - 1, d, inv_d); }

// Taps read at wavefront-uniform addresses go through the constant address
// space so they become scalar (SMEM) loads into SGPRs.
typedef const __attribute__((address_space(4))) float *taps_ptr;
__device__ __forceinline__ taps_ptr const_taps(const float2 *p) { return (taps_ptr)p; }
__device__ __forceinline__ float2 tap(taps_ptr g, uint32_t t) { return make_float2(g[2 * t], g[2 * t + 1]); }

__device__ __forceinline__ void cadd(float2 &a, float2 b)
{
    a.x += b.x;
    a.y += b.y;
}

// Input sample i of the chunk (i < 0: carried history). The source is chosen
// by selecting the address, not by branching, so the load stays a single
// unconditional instruction and a batch of them can all be in flight.
__device__ __forceinline__ float2 sample(const GroupsDev &a, const float2 *__restrict__ hist, int64_t i)
{
    const float2 *p = i >= 0 ? a.in + i : hist + (a.hist_len + i);
    return *p;
}

// Slow path: the window of anchor a (chunk-local) holds a rollover; samples
// t < split precede it and carry the extra e^{j2pi nu R}.
template <class G>
__device__ float2 window_rollover(const GroupsDev &a, const float2 *__restrict__ hist, G g, int64_t anchor,
                                  uint32_t cpp, uint32_t k, float2 wrap)
{
    const uint32_t split = cpp - 1 - k;
    const int64_t lo = anchor - (int64_t)(cpp - 1);
    float2 acc = make_float2(0.f, 0.f), acc2 = make_float2(0.f, 0.f);
    for (uint32_t t = 0; t < split; ++t)
        cmac(acc2, g(t), sample(a, hist, lo + t));
    for (uint32_t t = split; t < cpp; ++t)
        cmac(acc, g(t), sample(a, hist, lo + t));
    cadd(acc, cmul(wrap, acc2));
    return acc;
}

// Final step of one output: NCO index of the anchor, rollover slow path when
// the window straddles a restart, phasor, store.
__device__ __forceinline__ void finish(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                       const ChanStatic &ch, int64_t anchor, uint32_t j, float2 acc)
{
    const uint64_t k = mod_u64(P + (uint64_t)anchor, ch.r, ch.inv_r);
    if (k < ch.cpp - 1) {
        const float2 *__restrict__ g = a.taps + ch.taps_off;
        acc = window_rollover(a, hist, [&](uint32_t t) { return g[t]; }, anchor, ch.cpp, (uint32_t)k, ch.wrap);
    }
    reinterpret_cast<float2 *>(ch.out)[j] = cmul(nco_phasor((int64_t)k, ch.nu), acc);
}

// Stage the span x[lo .. lo+len) into LDS transposed by d: sample lo+i goes
// to (i mod d)*mc + i/d (mc odd, so the writes of consecutive i miss no
// bank). kPer loads per thread per batch.
template <int kBlock>
__device__ __forceinline__ void stage_transposed(const GroupsDev &a, const float2 *__restrict__ hist, int64_t lo,
                                                 uint32_t len, uint32_t d, uint32_t mc, float2 *xs)
{
    constexpr uint32_t kPer = 8;
    const uint32_t bq = kBlock / d, br = kBlock - bq * d;
    for (uint32_t i0 = 0; i0 < len; i0 += kPer * kBlock) {
        float2 v[kPer];
        const uint32_t last = len - 1;
        if (lo >= 0) {
            const float2 *__restrict__ base = a.in + lo;
#pragma unroll
            for (uint32_t s = 0; s < kPer; ++s)
                v[s] = base[min(i0 + s * kBlock + threadIdx.x, last)];
        } else {
#pragma unroll
            for (uint32_t s = 0; s < kPer; ++s)
                v[s] = sample(a, hist, lo + (int64_t)min(i0 + s * kBlock + threadIdx.x, last));
        }
        const uint32_t i = i0 + threadIdx.x;
        uint32_t q = i / d, rr = i - q * d;
#pragma unroll
        for (uint32_t s = 0; s < kPer; ++s) {
            if (i + s * kBlock < len)
                xs[rr * mc + q] = v[s];
            q += bq;
            rr += br;
            if (rr >= d) {
                rr -= d;
                ++q;
            }
        }
    }
}

// Gather the windows of nt outputs (anchor a0 + r*down, r < nt) into the
// tap-major tile tile[t*pitch + r] for taps t0 .. t0+nt_taps (window sample
// t of output r is x[a0 + r*down - (cpp-1) + t]). Element e = r*nt_taps + t,
// t fastest, so each window is read as one contiguous run.
template <int kBlock, uint32_t kPer>
__device__ __forceinline__ void gather_windows(const GroupsDev &a, const float2 *__restrict__ hist, int64_t a0,
                                               uint32_t down, uint32_t cpp, uint32_t nt, uint32_t t0,
                                               uint32_t nt_taps, uint32_t pitch, float2 *tile)
{
    const uint32_t total = nt * nt_taps;
    const int64_t lo = a0 - (int64_t)(cpp - 1) + t0;
    const uint32_t bq = kBlock / nt_taps, br = kBlock - bq * nt_taps;
    for (uint32_t e0 = 0; e0 < total; e0 += kPer * kBlock) {
        const uint32_t e = e0 + threadIdx.x;
        uint32_t r = e / nt_taps, t = e - r * nt_taps;
        float2 v[kPer];
        {
            uint32_t rr = r, tt = t;
#pragma unroll
            for (uint32_t s = 0; s < kPer; ++s) {
                const uint32_t rc = min(rr, nt - 1);
                const int64_t idx = lo + (int64_t)rc * down + tt;
                v[s] = (lo >= 0) ? a.in[idx] : sample(a, hist, idx);
                rr += bq;
                tt += br;
                if (tt >= nt_taps) {
                    tt -= nt_taps;
                    ++rr;
                }
            }
        }
#pragma unroll
        for (uint32_t s = 0; s < kPer; ++s) {
            if (e + s * kBlock < total)
                tile[t * pitch + r] = v[s];
            r += bq;
            t += br;
            if (t >= nt_taps) {
                t -= nt_taps;
                ++r;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// MFMA groups
// ---------------------------------------------------------------------------
template <int kBlock, uint32_t kS>
__device__ __forceinline__ void mfma_group(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                           const uint32_t *gw, uint32_t j0, uint32_t nt, int64_t a0,
                                           unsigned char *smem)
{
    constexpr uint32_t kCpp = 4 * kS;  // padded taps per window (zero taps in front)
    const uint32_t down = gw[2], m = gw[4], rows = gw[5];
    ChanStatic *sc = reinterpret_cast<ChanStatic *>(smem);  // the group's channels
    float2 *xs = reinterpret_cast<float2 *>(smem + kMfmaMaxChan * sizeof(ChanStatic));
    if (threadIdx.x < m)
        sc[threadIdx.x] = a.chans[gw[8 + threadIdx.x]];

    const bool dense = down < kCpp;
    uint32_t mc = 0, pitch = 0;
    if (dense) {
        const uint32_t len = (nt - 1) * down + kCpp;
        mc = ((len + down - 1) / down) | 1u;
        stage_transposed<kBlock>(a, hist, a0 - (int64_t)(kCpp - 1), len, down, mc, xs);
    } else {
        pitch = rows + 1;
        gather_windows<kBlock, 8>(a, hist, a0, down, kCpp, nt, 0, kCpp, pitch, xs);
    }

    // B operands: step s holds B[k = 4s + lane/16][j = lane%16]; steps 0..kS-1
    // cover the real parts of the window, kS..2kS-1 the imaginary parts
    const uint32_t lane = threadIdx.x & 63;
    float bv[2 * kS];
    {
        const float *__restrict__ bt = a.btab + gw[6];
#pragma unroll
        for (uint32_t s = 0; s < 2 * kS; ++s)
            bv[s] = bt[s * 64 + lane];
    }
    // LDS offset of window tap t = 4s + lane/16 of row 0
    uint32_t off[kS];
    {
        const uint32_t t0 = lane >> 4;
        if (dense) {
            uint32_t q = t0 / down, rr = t0 - q * down;
            const uint32_t sq = 4 / down, sr = 4 - sq * down;
#pragma unroll
            for (uint32_t s = 0; s < kS; ++s) {
                off[s] = rr * mc + q;
                q += sq;
                rr += sr;
                if (rr >= down) {
                    rr -= down;
                    ++q;
                }
            }
        } else {
#pragma unroll
            for (uint32_t s = 0; s < kS; ++s)
                off[s] = (4 * s + t0) * pitch;
        }
    }
    __syncthreads();

    constexpr uint32_t kWaves = kBlock / 64;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t subtiles = (nt + 15) >> 4;
    const uint32_t col = lane & 15, cprime = col >> 1;
    const bool even = (col & 1) == 0;
    for (uint32_t st = wave; st < subtiles; st += kWaves) {
        const uint32_t rbase = st * 16 + (lane & 15);
        floatx4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t s = 0; s < kS; ++s) {
            const float2 x = xs[off[s] + rbase];
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.x, bv[s], acc, 0, 0, 0);
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(x.y, bv[kS + s], acc, 0, 0, 0);
        }
        // D[i][j] sits in lane j + 16*(i/4), element i%4. Lanes 2c and 2c+1
        // hold channel c's real and imaginary parts of the same rows: swap
        // through DPP (quad_perm 1,0,3,2); the even lane finishes rows 0, 1
        // of its quad of rows, the odd lane rows 2, 3.
        float oth[4];
#pragma unroll
        for (uint32_t v = 0; v < 4; ++v)
            oth[v] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[v]), 0xB1, 0xF, 0xF, false));
        if (cprime < m) {
            const ChanStatic &ch = sc[cprime];
#pragma unroll
            for (uint32_t vv = 0; vv < 2; ++vv) {
                const uint32_t v = even ? vv : 2 + vv;
                const float mine = even ? (vv ? acc[1] : acc[0]) : (vv ? acc[3] : acc[2]);
                const float other = even ? (vv ? oth[1] : oth[0]) : (vv ? oth[3] : oth[2]);
                const uint32_t row = st * 16 + 4 * (lane >> 4) + v;
                if (row < nt) {
                    const float2 y = even ? make_float2(mine, other) : make_float2(other, mine);
                    finish(a, hist, P, ch, a0 + (int64_t)row * down, j0 + row, y);
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// one channel, dense (VALU, transposed span)
// ---------------------------------------------------------------------------
constexpr uint32_t kOpl = kVDenseOpl;

template <int kBlock>
__device__ __forceinline__ void vdense_group(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                             const uint32_t *gw, uint32_t j0, uint32_t nt, int64_t a0,
                                             unsigned char *smem)
{
    const uint32_t down = gw[2], cpp = gw[3];
    const uint32_t c = gw[8];
    float2 *xs = reinterpret_cast<float2 *>(smem);
    const uint32_t len = (nt - 1) * down + cpp;
    const uint32_t mc = ((len + down - 1) / down) | 1u;
    stage_transposed<kBlock>(a, hist, a0 - (int64_t)(cpp - 1), len, down, mc, xs);
    __syncthreads();

    const ChanStatic ch = a.chans[c];
    const taps_ptr g = const_taps(a.taps + uniform(ch.taps_off));
    constexpr uint32_t kWaves = kBlock / 64;
    constexpr uint32_t kTask = 64 * kOpl;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63;
    for (uint32_t tt = wave; tt * kTask < nt; tt += kWaves) {
        const uint32_t rl = tt * kTask + lane;  // rows rl + 64*o, o < kOpl
        if (rl >= nt)
            continue;
        uint32_t row_t = 0, colq = 0;  // tap t's sample: xs[row_t*mc + colq + r]
        float2 acc[kOpl];
#pragma unroll
        for (uint32_t o = 0; o < kOpl; ++o)
            acc[o] = make_float2(0.f, 0.f);
#pragma unroll 4
        for (uint32_t t = 0; t < cpp; ++t) {
            const float2 h = tap(g, t);
            const uint32_t xi = row_t * mc + colq + rl;
#pragma unroll
            for (uint32_t o = 0; o < kOpl; ++o)
                cmac(acc[o], h, xs[xi + 64 * o]);
            if (++row_t == down) {
                row_t = 0;
                ++colq;
            }
        }
#pragma unroll
        for (uint32_t o = 0; o < kOpl; ++o) {
            const uint32_t row = rl + 64 * o;
            if (row < nt)
                finish(a, hist, P, ch, a0 + (int64_t)row * down, j0 + row, acc[o]);
        }
    }
}

// ---------------------------------------------------------------------------
// one channel, sparse (VALU, gathered windows)
// ---------------------------------------------------------------------------
template <int kBlock>
__device__ __forceinline__ void vsparse_group(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                              const uint32_t *gw, uint32_t j0, uint32_t nt, int64_t a0,
                                              unsigned char *smem)
{
    // one output per thread (nt <= kBlock); windows gathered kVSparsePass taps at a time
    constexpr uint32_t kPitch = kBlock + 1;
    const uint32_t down = gw[2], cpp = gw[3];
    float2 *tile = reinterpret_cast<float2 *>(smem);
    const ChanStatic ch = a.chans[gw[8]];
    const taps_ptr g = const_taps(a.taps + uniform(ch.taps_off));
    float2 acc = make_float2(0.f, 0.f);
    for (uint32_t t0 = 0; t0 < cpp; t0 += kVSparsePass) {
        const uint32_t np = min(kVSparsePass, cpp - t0);
        if (t0)
            __syncthreads();
        gather_windows<kBlock, kVSparsePass>(a, hist, a0, down, cpp, nt, t0, np, kPitch, tile);
        __syncthreads();
        for (uint32_t t = 0; t < np; ++t)
            cmac(acc, tap(g, t0 + t), tile[t * kPitch + threadIdx.x]);
    }
    if (threadIdx.x < nt)
        finish(a, hist, P, ch, a0 + (int64_t)threadIdx.x * down, j0 + threadIdx.x, acc);
}

// ---------------------------------------------------------------------------
// one channel, up > 1 (natural span, per-output phase)
// ---------------------------------------------------------------------------
template <int kBlock>
__device__ __forceinline__ void generic_group(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                              const uint32_t *gw, uint32_t j0, uint32_t nt, uint64_t n0,
                                              unsigned char *smem)
{
    const uint32_t up = gw[1], down = gw[2], cpp = gw[3];
    const uint32_t c = gw[8];
    const ChanStatic ch = a.chans[c];
    // output j0 + q: numerator (n0 + j0 + q)*down = (qb + P)*up + rb + q*down
    const uint64_t num0 = (n0 + j0) * down;
    const uint64_t qb = udiv(num0, up, ch.inv_up);
    const uint32_t rb = (uint32_t)(num0 - qb * up);
    const int64_t a_first = (int64_t)(qb - P);  // chunk-local anchor of the tile's first output
    const uint32_t span = (uint32_t)((rb + (uint64_t)(nt - 1) * down) / up) + cpp;
    float2 *xs = reinterpret_cast<float2 *>(smem);
    const int64_t lo = a_first - (int64_t)(cpp - 1);
    for (uint32_t i0 = 0; i0 < span; i0 += 8 * kBlock) {
        float2 v[8];
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s)
            v[s] = sample(a, hist, lo + (int64_t)min(i0 + s * kBlock + threadIdx.x, span - 1));
#pragma unroll
        for (uint32_t s = 0; s < 8; ++s)
            if (i0 + s * kBlock + threadIdx.x < span)
                xs[i0 + s * kBlock + threadIdx.x] = v[s];
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < nt; q += kBlock) {
        const uint32_t num = rb + q * down;
        const uint32_t adv = num / up;
        const uint32_t ph = num - adv * up;
        const float2 *xw = xs + adv;
        const float2 *__restrict__ gg = a.taps + ch.taps_off + ph * cpp;
        float2 acc = make_float2(0.f, 0.f);
        for (uint32_t t = 0; t < cpp; ++t)
            cmac(acc, gg[t], xw[t]);
        const int64_t anchor = a_first + adv;
        const uint64_t k = mod_u64(P + (uint64_t)anchor, ch.r, ch.inv_r);
        if (k < cpp - 1)
            acc = window_rollover(a, hist, [&](uint32_t t) { return gg[t]; }, anchor, cpp, (uint32_t)k, ch.wrap);
        reinterpret_cast<float2 *>(ch.out)[j0 + q] = cmul(nco_phasor((int64_t)k, ch.nu), acc);
    }
}

template <int kBlock>
__global__ __launch_bounds__(kBlock) void chan_groups_kernel(GroupsDev a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    const uint64_t P = a.pos ? a.pos->p : a.p;
    const uint32_t par = a.pos ? a.pos->parity : a.parity;
    const float2 *hist = par ? a.hist1 : a.hist0;
    const uint32_t H = a.hist_len;
    if (blockIdx.x == 0) {
        // next chunk's history: the last H samples of [history | chunk]
        float2 *hw = par ? a.histw0 : a.histw1;
        for (uint32_t t = threadIdx.x; t < H; t += kBlock) {
            const int64_t i = (int64_t)a.n - H + t;
            hw[t] = i >= 0 ? a.in[i] : hist[H + i];
        }
    }
    // which group, which row tile
    uint32_t lo = 0, hi = a.n_groups;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (a.prefix[mid] <= blockIdx.x)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t g = lo, tb = blockIdx.x - a.prefix[lo];
    const uint32_t *gw = a.groups + g * kGroupWords;
    const uint32_t kind = gw[0], up = gw[1], down = gw[2], rows = gw[5];
    const ChanStatic &c0 = a.chans[gw[8]];
    const uint64_t n0 = cdiv(P * up, down, c0.inv_down);
    const uint32_t count = (uint32_t)(cdiv((P + a.n) * up, down, c0.inv_down) - n0);
    const uint32_t j0 = tb * rows;
    if (j0 >= count)
        return;  // uniform across the block
    const uint32_t nt = min(rows, count - j0);
    const int64_t a0 = (int64_t)((n0 + j0) * down - P);  // up == 1: anchor of row 0
    switch (kind) {
    case kGroupMfma:
        if (gw[3] == 32)
            mfma_group<kBlock, 8>(a, hist, P, gw, j0, nt, a0, smem);
        else
            mfma_group<kBlock, 16>(a, hist, P, gw, j0, nt, a0, smem);
        break;
    case kGroupVDense:
        vdense_group<kBlock>(a, hist, P, gw, j0, nt, a0, smem);
        break;
    case kGroupVSparse:
        vsparse_group<kBlock>(a, hist, P, gw, j0, nt, a0, smem);
        break;
    default:
        generic_group<kBlock>(a, hist, P, gw, j0, nt, n0, smem);
        break;
    }
}

__global__ void chan_advance_kernel(ChanPos *pos, uint32_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        pos->p += n;
        pos->parity ^= 1u;
    }
}

}  // namespace

hipError_t launch_chan_groups(const GroupsArgs &f, hipStream_t s)
{
    if (f.n == 0)
        return hipSuccess;
    if (f.lds_bytes > 64 * 1024 || f.n_groups > kGroupsInline)
        return hipErrorInvalidValue;
    GroupsDev d{};
    d.in = f.in;
    d.n = f.n;
    d.hist_len = f.hist_len;
    d.hist0 = f.hist[0];
    d.hist1 = f.hist[1];
    d.histw0 = f.hist_w[0];
    d.histw1 = f.hist_w[1];
    d.chans = f.chans;
    d.taps = f.taps;
    d.btab = f.btab;
    d.pos = f.pos;
    d.p = f.p;
    d.parity = f.parity;
    d.groups = f.groups;
    d.n_groups = f.n_groups;
    for (uint32_t k = 0; k <= f.n_groups; ++k)
        d.prefix[k] = f.prefix[k];
    const uint32_t blocks = f.prefix[f.n_groups];
    switch (f.block) {
    case 128:
        hipLaunchKernelGGL(chan_groups_kernel<128>, dim3(blocks), dim3(128), f.lds_bytes, s, d);
        break;
    case 512:
        hipLaunchKernelGGL(chan_groups_kernel<512>, dim3(blocks), dim3(512), f.lds_bytes, s, d);
        break;
    default:
        hipLaunchKernelGGL(chan_groups_kernel<256>, dim3(blocks), dim3(256), f.lds_bytes, s, d);
        break;
    }
    return hipGetLastError();
}

hipError_t launch_chan_advance(ChanPos *pos, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(chan_advance_kernel, dim3(1), dim3(64), 0, s, pos, n);
    return hipGetLastError();
}

}  // namespace fdsp
