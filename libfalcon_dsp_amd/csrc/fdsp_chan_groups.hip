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
#include <type_traits>

#include "fdsp_internal.h"

namespace fdsp {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
// Output pointers come from the channel table in memory, so the compiler
// cannot infer their address space; flat stores would also count against
// the LDS wait counter and stall every following LDS read until they land.
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef __attribute__((address_space(1))) fvec2 *gvec2_ptr;
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) fvec4 *gvec4_ptr;
__device__ __forceinline__ void store_out(void *base, uint32_t j, float2 v)
{
    ((gvec2_ptr)base)[j] = fvec2{v.x, v.y};
}

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
    const uint32_t *table;  // (group, stripe) of each block
    uint32_t strided;       // 0: tps consecutive tiles per block; 1: every tps-th tile
    uint32_t *claims;       // non-null: blocks claim tiles from per-group counters
    float2 *sink;           // kSinkWords samples that absorb the stores of rows and lanes with no output
    uint32_t out16;         // every channel's output is 16-byte aligned
    uint32_t in16;          // the chunk's input is 16-byte aligned
    uint32_t tps[kGroupsInline];  // tiles per stripe, or blocks, of each group
    uint16_t sched[kGroupsInline];  // the groups this launch schedules
    uint32_t n_sched;
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
    store_out(ch.out, j, cmul(nco_phasor((int64_t)k, ch.nu), acc));
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
// tile sequences
// ---------------------------------------------------------------------------
// A block's tiles of one group come either from a fixed sequence (tb, tb+step,
// ... below end) or from the group's claim counters: shard s hands out tiles
// s, s + kClaimShards, ... in turn, and a block draws from one shard until it
// is spent, then from the next. Each claim is issued one tile ahead (its
// result is read at the next claim), so its latency hides behind a tile of
// work. claim() is called by all threads; only thread 0 touches the
// counters. It leaves the next tile (or ~0 when there is none) in *slot,
// which needs a barrier before it is read.
struct TileSeq {
    uint32_t *ctr;  // the group's counters, or null for the fixed sequence
    uint32_t next, step, end;
    uint32_t tiles;
    uint32_t first;   // shard drawn from first
    uint32_t *spent;  // LDS word (thread 0): shards found spent; the shard drawn from is first + *spent
    uint32_t pend;    // thread 0: the claim in flight
    __device__ __forceinline__ void start(uint32_t first_shard, uint32_t *spent_word)
    {
        first = first_shard;
        spent = spent_word;
        if (ctr && threadIdx.x == 0) {
            *spent = 0;
            pend = atomicAdd(ctr + (first % kClaimShards) * kClaimPitch, 1u);
        }
    }
    __device__ __forceinline__ void claim(uint32_t *slot)
    {
        if (ctr) {
            if (threadIdx.x == 0) {
                uint32_t n_spent = *spent;
                uint32_t shard = (first + n_spent) % kClaimShards;
                uint32_t t = shard + kClaimShards * pend;
                // at most kClaimShards - 1 more shards to try
                while (t >= tiles && n_spent + 1 < kClaimShards) {
                    ++n_spent;
                    shard = (first + n_spent) % kClaimShards;
                    t = shard + kClaimShards * atomicAdd(ctr + shard * kClaimPitch, 1u);
                }
                if (t < tiles) {
                    *slot = t;
                    pend = atomicAdd(ctr + shard * kClaimPitch, 1u);
                } else {
                    *slot = 0xffffffffu;
                    n_spent = kClaimShards;
                }
                *spent = n_spent;
            }
        } else {
            if (threadIdx.x == 0)
                *slot = next < end ? next : 0xffffffffu;
            next += step;
        }
    }
};

// ---------------------------------------------------------------------------
// MFMA groups
// ---------------------------------------------------------------------------
// A tile's input passes through registers (kStage samples per thread) on
// its way to LDS, so the loads of tile i+1 are in flight while tile i runs on
// the matrix cores. Dense tiles stage the span transposed by down; sparse
// tiles gather the windows tap-major (thread tid always carries tap
// tid mod w, rows tid/w + s*(kBlock/w)).
// Raw buffer loads: 32-bit offsets from a base held in SGPRs, and an offset
// at or past `bytes` reads as zero instead of touching memory.
constexpr int kBufferFlags = 0x00020000;  // gfx9 descriptor word 3: 32-bit data format
__device__ __forceinline__ __amdgpu_buffer_rsrc_t buffer_of(const float2 *p, int64_t count)
{
    const int64_t bytes = count * (int64_t)sizeof(float2);
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<float2 *>(p), 0, (int)min(bytes, (int64_t)0x7ffffff8),
                                             kBufferFlags);
}
__device__ __forceinline__ float2 buffer_sample(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff)
{
    const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0);
    return make_float2(__int_as_float(v[0]), __int_as_float(v[1]));
}

// Loads of one MFMA tile (anchor a0 of row 0). Element s of a thread is span
// sample s*kBlock + tid (dense) or window tap tid mod w of row
// s*(kBlock/w) + tid/w (sparse). Elements past the tile read what lies there
// (zero past the chunk) and are never used. Only element 0 can precede the
// chunk (kBlock > w - 1, and sparse rows are down >= w apart), so a tile that
// reaches into the history takes element 0 from there.
template <int kBlock, uint32_t kW, uint32_t kStage = mfma_stage(kBlock)>
__device__ __forceinline__ void mfma_load(const GroupsDev &a, const float2 *__restrict__ hist, bool dense,
                                          int64_t a0, uint32_t down, float2 (&v)[kStage])
{
    constexpr uint32_t kRowsPer = kBlock / kW;
    const int64_t lo = a0 - (int64_t)(kW - 1);
    // base may lie before the chunk: offsets below -lo are never read through it
    const __amdgpu_buffer_rsrc_t r = buffer_of(a.in + lo, (int64_t)a.n - lo);
    uint32_t e0 = dense ? threadIdx.x : threadIdx.x % kW + (threadIdx.x / kW) * down;
    // Opaque per call: keeps the per-element offsets from being hoisted out
    // of the tile loop into registers that would stay live across it.
    asm volatile("" : "+v"(e0));
    const uint32_t stride = dense ? kBlock * 8u : kRowsPer * down * 8u;
    if (lo >= 0) {
        v[0] = buffer_sample(r, e0 * 8u, 0);
    } else {
        // from the history, or from the chunk through a descriptor bounded
        // by the chunk (short chunks end before the tile's span does)
        const int64_t i = lo + (int64_t)e0;
        v[0] = i < 0 ? hist[a.hist_len + i] : buffer_sample(buffer_of(a.in, a.n), (uint32_t)i * 8u, 0);
    }
    // the whole offset goes in the VGPR operand, which the range check covers
#pragma unroll
    for (uint32_t s = 1; s < kStage; ++s)
        v[s] = buffer_sample(r, e0 * 8u + s * stride, 0);
}

// LDS layout of a staged MFMA tile. Dense: span sample i at
// (i mod down)*mc + i/down with mc = ceil(capacity/down) rounded up to odd,
// so the 16 rows of a sub-tile (same tap, consecutive rows) read 16
// consecutive slots and the residue rows start on different banks. Sparse:
// window r at r*(w+1), tap t at +t (rows w+1 apart fall on different banks).
// The whole staging capacity is written.
__host__ __device__ constexpr uint32_t mfma_dense_mc(uint32_t capacity, uint32_t down)
{
    return ((capacity + down - 1) / down) | 1u;
}

template <int kBlock, uint32_t kW, uint32_t kStage = mfma_stage(kBlock)>
__device__ __forceinline__ void mfma_commit(bool dense, uint32_t down, uint32_t mc, const float2 (&v)[kStage],
                                            float2 *xs)
{
    constexpr uint32_t kRowsPer = kBlock / kW;
    if (dense) {
        uint32_t i0 = threadIdx.x;
        asm volatile("" : "+v"(i0));  // per tile: keeps the slots out of long-lived registers
        const uint32_t bq = kBlock / down, br = kBlock - bq * down;
        uint32_t q = i0 / down, rr = i0 - q * down;
#pragma unroll
        for (uint32_t s = 0; s < kStage; ++s) {
            xs[rr * mc + q] = v[s];
            q += bq;
            rr += br;
            if (rr >= down) {
                rr -= down;
                ++q;
            }
        }
    } else {
        float2 *p = xs + (threadIdx.x / kW) * (kW + 1) + threadIdx.x % kW;
#pragma unroll
        for (uint32_t s = 0; s < kStage; ++s)
            p[s * kRowsPer * (kW + 1)] = v[s];
    }
}

// bf16 planes of a staged tile: the span's real and imaginary parts, each as
// a hi and a lo bf16 part (x = hi + lo to about 2^-16), in natural order from
// an 8-aligned start s0, planes [re hi | im hi | re lo | im lo] of kPlane
// samples each. A thread stages pairs of samples (16-byte loads): pair
// p = e*kBlock + tid holds samples s0 + 2p, s0 + 2p + 1.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void bf16_split(float a, float b, uint32_t &hi, uint32_t &lo)
{
    const bf16x2 h = __builtin_convertvector(fvec2{a, b}, bf16x2);
    const bf16x2 l = __builtin_convertvector(fvec2{a - (float)h[0], b - (float)h[1]}, bf16x2);
    hi = __builtin_bit_cast(uint32_t, h);
    lo = __builtin_bit_cast(uint32_t, l);
}
__device__ __forceinline__ float bf16_word_half(uint32_t w, uint32_t i)
{
    return __uint_as_float(i ? (w & 0xffff0000u) : (w << 16));
}

template <int kBlock, uint32_t kPairs = mfma_stage(kBlock) / 2>
__device__ __forceinline__ void bf_load(const GroupsDev &a, const float2 *__restrict__ hist, int64_t s0,
                                        u32x4 (&v)[kPairs])
{
    const __amdgpu_buffer_rsrc_t r = buffer_of(a.in + s0, (int64_t)a.n - s0);
    uint32_t p0 = threadIdx.x;
    asm volatile("" : "+v"(p0));
    if (s0 >= 0) {
        v[0] = __builtin_bit_cast(u32x4, __builtin_amdgcn_raw_buffer_load_b128(r, (int)(p0 * 16u), 0, 0));
    } else {
        // s0 and the history length are even, so a pair never straddles the
        // chunk start; only pair 0 can precede it (s0 >= -38 > -2*kBlock)
        const int64_t i = s0 + 2 * (int64_t)p0;
        if (i < 0)
            v[0] = *reinterpret_cast<const u32x4 *>(hist + (a.hist_len + i));
        else  // bounded by the chunk, which may end before the span does
            v[0] = __builtin_bit_cast(
                u32x4, __builtin_amdgcn_raw_buffer_load_b128(buffer_of(a.in, a.n), (int)((uint32_t)i * 8u), 0, 0));
    }
#pragma unroll
    for (uint32_t e = 1; e < kPairs; ++e)
        v[e] = __builtin_bit_cast(u32x4,
                                  __builtin_amdgcn_raw_buffer_load_b128(r, (int)(p0 * 16u + e * kBlock * 16u), 0, 0));
}

template <int kBlock, uint32_t kPairs = mfma_stage(kBlock) / 2>
__device__ __forceinline__ void bf_commit(const u32x4 (&v)[kPairs], uint32_t *planes)
{
    constexpr uint32_t kWords = kPairs * kBlock;  // words per plane
    uint32_t p0 = threadIdx.x;
    asm volatile("" : "+v"(p0));
#pragma unroll
    for (uint32_t e = 0; e < kPairs; ++e) {
        uint32_t rh, rl, ih, il;
        bf16_split(__uint_as_float(v[e][0]), __uint_as_float(v[e][2]), rh, rl);
        bf16_split(__uint_as_float(v[e][1]), __uint_as_float(v[e][3]), ih, il);
        const uint32_t p = p0 + e * kBlock;
        planes[p] = rh;
        planes[kWords + p] = ih;
        planes[2 * kWords + p] = rl;
        planes[3 * kWords + p] = il;
    }
}

// sample i of a bf16-plane tile
template <uint32_t kWords>
__device__ __forceinline__ float2 bf_sample(const uint32_t *planes, uint32_t i)
{
    const uint32_t w = i >> 1, h = i & 1;
    return make_float2(bf16_word_half(planes[w], h) + bf16_word_half(planes[2 * kWords + w], h),
                       bf16_word_half(planes[kWords + w], h) + bf16_word_half(planes[3 * kWords + w], h));
}

// The tiles of an MFMA group that seq hands this block, the first one already
// in *slot (tile i: rows i*rows .. of the chunk's `count` outputs, first
// output n0 of the stream). The next tile is claimed while the current one
// is staged, so its loads overlap the current tile's MFMA work.
template <int kBlock, uint32_t kS, bool kDense, bool kBf = false>
__device__ __forceinline__ void mfma_stripe(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                            const uint32_t *gw, uint64_t n0, uint32_t count, TileSeq seq,
                                            uint32_t *slot, unsigned char *smem)
{
    static_assert(!kBf || (kS == 8 && kDense), "bf16 planes: 32-tap dense windows");
    constexpr uint32_t kW = 4 * kS;  // padded taps per window (zero taps in front)
    constexpr uint32_t kPairs = mfma_stage(kBlock) / 2;
    constexpr uint32_t kWords = kPairs * kBlock;  // bf16 plane length in words
    constexpr uint32_t kWaves = kBlock / 64;
    const uint32_t down = gw[2], m = gw[4], rows = gw[5];
    // LDS: the group's modulated taps [channel][w] (zero-padded in front, as
    // the windows are), then the staged tile
    float2 *taps_l = reinterpret_cast<float2 *>(smem);
    float2 *xs = taps_l + kMfmaMaxChan * kW;
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t wave = uniform(threadIdx.x >> 6);
    const uint32_t mc = kDense ? mfma_dense_mc(mfma_stage(kBlock) * kBlock, down) : 0;
    // Rows of a tile fall in nc residue classes (bf16 planes: nc = 8/gcd(down,
    // 8), whose rows' windows share their offset from an 8-aligned slot;
    // otherwise nc = 1): class c holds rows c, c+nc, c+2nc, ... Wave w works
    // on class w mod nc, on every wpc-th sub-tile of 16 of its rows.
    const uint32_t nc = kBf ? 8u >> min(3u, (uint32_t)__builtin_ctz(down)) : 1u;
    const uint32_t cls = wave % nc, wpc = kWaves / nc, wi = wave / nc;
    const uint32_t dn = down * nc;  // inputs between consecutive rows of a class
    for (uint32_t e = threadIdx.x; e < m * kW; e += kBlock) {
        const ChanStatic &c = a.chans[gw[8 + e / kW]];
        const uint32_t t = e % kW, pad = kW - c.cpp;
        taps_l[e] = t < pad ? make_float2(0.f, 0.f) : a.taps[c.taps_off + t - pad];
    }

    // B operands: step s holds B[k = 4s + lane/16][j = lane%16]; steps 0..kS-1
    // cover the real parts of the window, kS..2kS-1 the imaginary parts
    float bv[kBf ? 1 : 2 * kS];
    bf16x8 bh[3], bl[3];  // bf16 planes: B of K step q, hi and lo parts
    // bf16 planes: the windows start `base` samples past an 8-aligned slot,
    // the same for every tile (tile anchors differ by multiples of down)
    const uint32_t base = (uint32_t)((int64_t)((n0 + uniform(*slot) * rows) * down - P) - (int64_t)(kW - 1)) & 7u;
    // this wave's class: windows start cstart + dn*m + cbase samples into the
    // staged span (cstart a multiple of 8, cbase < 8)
    const uint32_t cstart = (base + down * cls) & ~7u, cbase = (base + down * cls) & 7u;
    if constexpr (kBf) {
        const u32x4 *bt = reinterpret_cast<const u32x4 *>(a.btab + (gw[7] - 1) + cbase * kBfWordsPerBase);
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            bh[q] = __builtin_bit_cast(bf16x8, bt[(q * 2 + 0) * 64 + lane]);
            bl[q] = __builtin_bit_cast(bf16x8, bt[(q * 2 + 1) * 64 + lane]);
        }
    } else {
        const float *__restrict__ bt = a.btab + gw[6];
#pragma unroll
        for (uint32_t s = 0; s < 2 * kS; ++s)
            bv[s] = bt[s * 64 + lane];
    }
    uint32_t *planes = reinterpret_cast<uint32_t *>(xs);
    // bf16 A operand of K step q: class row lane%16, K block kb = 4q +
    // lane/16: the real (kb < 5) or imaginary (kb 5..9) plane from slot
    // 8*(kb mod 5) of the row's aligned window; the zero-B blocks 10, 11 read
    // the real plane's slot 0 (finite)
    uint32_t aoff[3];
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t kb = 4 * q + (lane >> 4);
        const uint32_t plane = kb >= 5 && kb < 10 ? 1u : 0u;
        const uint32_t slot = kb < 5 ? 8 * kb : kb < 10 ? 8 * (kb - 5) : 0u;
        aoff[q] = plane * kWords * 4 + (cstart + dn * (lane & 15) + slot) * 2;  // bytes, hi part
    }
    // A operand of step s: row r = st*16 + lane%16, window tap 4s + lane/16,
    // read at xa[s] + 16*st
    const float2 *xa[kS];
    {
        const uint32_t t0 = lane >> 4, r0 = lane & 15;
        if constexpr (kDense) {
            uint32_t q = t0 / down, rr = t0 - q * down;
            const uint32_t sq = 4 / down, sr = 4 - sq * down;
#pragma unroll
            for (uint32_t s = 0; s < kS; ++s) {
                xa[s] = xs + rr * mc + q + r0;
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
                xa[s] = xs + r0 * (kW + 1) + t0 + 4 * s;
        }
    }

    // D[i][j] sits in lane j + 16*(i/4), element i%4: lanes 2c and 2c+1 hold
    // channel c's real and imaginary parts of rows 4*(lane/16) .. +3. The even
    // lane finishes the first two of those rows (e = 0, 1), the odd lane the
    // last two.
    const uint32_t col = lane & 15;
    const bool odd = col & 1, live = (col >> 1) < m;
    const uint32_t cl = min(col >> 1, m - 1);
    const ChanStatic *ch = a.chans + gw[8 + cl];
    const uint64_t R = ch->r;
    const uint32_t cpp1 = ch->cpp - 1;
    const double nu = ch->nu, inv_r = ch->inv_r;
    const float2 wrap = ch->wrap;
    gvec2_ptr out = (gvec2_ptr)ch->out;
    // rows and lanes with no output store into this block's part of the sink
    // (16 bytes per thread)
    gvec2_ptr sink = (gvec2_ptr)(a.sink + 2 * ((blockIdx.x % (kSinkWords / (2 * kBlock))) * kBlock + threadIdx.x));
    const bool out16 = a.out16 != 0;
    const float2 *my_taps = taps_l + cl * kW;
    const uint32_t ro = 4 * (lane >> 4) + (odd ? 2 : 0);
    // NCO index k of the lane's first row (class row ro) in its current
    // sub-tile, in [0, R), with the phasor e^{j2pi nu k}; the lane's second
    // row is dn inputs later and the wave's next sub-tile du inputs later.
    // While cpp-1 <= k < R - du - dn neither row's window holds a rollover and
    // neither index nor the next sub-tile's wraps: the fast path steps the
    // phasor by recurrence. Otherwise the wave takes the exact path.
    const uint32_t du = 16 * wpc * dn;
    const float2 s_row = nco_phasor((int64_t)dn, nu);
    const float2 s_sub = nco_phasor((int64_t)du, nu);
    const uint32_t k_hi = R > (uint64_t)du + dn ? (uint32_t)(R - du - dn) : 0u;

    // A operands and MFMA chains of one sub-tile: FP32, two independent
    // chains (the real-part and the imaginary-part steps), summed by the
    // consumer; bf16 planes, one chain in c0
    auto mma = [&](uint32_t st, floatx4 &c0, floatx4 &c1) {
        if constexpr (kBf) {
            const unsigned char *pb = reinterpret_cast<const unsigned char *>(planes) + st * 16 * dn * 2;
            bf16x8 ah[3], al[3];
#pragma unroll
            for (uint32_t q = 0; q < 3; ++q) {
                ah[q] = *reinterpret_cast<const bf16x8 *>(pb + aoff[q]);
                al[q] = *reinterpret_cast<const bf16x8 *>(pb + aoff[q] + 2 * kWords * 4);
            }
            // one chain (the lo terms first, then hi x hi): the epilogue
            // needs no sum of two chains
            c0 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (uint32_t q = 0; q < 3; ++q) {
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bl[q], c0, 0, 0, 0);
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[q], bh[q], c0, 0, 0, 0);
                c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bh[q], c0, 0, 0, 0);
            }
            c1 = c0;  // unused: the bf16 epilogue reads c0 alone
            return;
        }
        float2 x[kS];
#pragma unroll
        for (uint32_t s = 0; s < kS; ++s)
            x[s] = xa[s][st * 16 * (kDense ? 1 : kW + 1)];
        c0 = floatx4{0.f, 0.f, 0.f, 0.f};
        c1 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t s = 0; s < kS; ++s) {
            c0 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[s].x, bv[s], c0, 0, 0, 0);
            c1 = __builtin_amdgcn_mfma_f32_16x16x4f32(x[s].y, bv[kS + s], c1, 0, 0, 0);
        }
    };

    // tiles of this chunk: ceil(count / rows); the first one was claimed by
    // the caller (in *slot)
    const uint32_t tiles = (count + rows - 1) / rows;
    uint32_t tb = uniform(*slot);
    float2 v[kBf ? 1 : mfma_stage(kBlock)];
    u32x4 vb[kBf ? kPairs : 1];
    // bf16 planes: the staged span of a tile starts at its first window's
    // first sample less `base`
    auto load = [&](uint32_t t) {
        const int64_t a0t = (int64_t)((n0 + t * rows) * down - P);
        if constexpr (kBf)
            bf_load<kBlock>(a, hist, a0t - (int64_t)(kW - 1) - (int64_t)base, vb);
        else
            mfma_load<kBlock, kW>(a, hist, kDense, a0t, down, v);
    };
    load(tb);
    for (;;) {
        const uint32_t j0 = tb * rows;
        const uint32_t nt = min(rows, count - j0);
        const int64_t a0 = (int64_t)((n0 + j0) * down - P);  // chunk-local anchor of row 0
        __syncthreads();  // the previous tile's reads of LDS (and of *slot) are done
        if constexpr (kBf)
            bf_commit<kBlock>(vb, planes);
        else
            mfma_commit<kBlock, kW>(kDense, down, mc, v, xs);
        seq.claim(slot);
        __syncthreads();
        const uint32_t tn = uniform(*slot);
        if (tn < tiles)
            load(tn);
        // rows of this wave's class in the tile, and their sub-tiles of 16;
        // the class's waves take sub-tiles wt, wt + wpc, ... with wt rotating
        // from tile to tile, so a tile's odd sub-tile does not always fall to
        // the same wave (and SIMD)
        const uint32_t crows = nt > cls ? (nt - cls + nc - 1) / nc : 0u;
        const uint32_t subtiles = (crows + 15) >> 4;
        const uint32_t wt = (wi + tb) & (wpc - 1);
        if (wt >= subtiles)
            goto next_tile;  // uniform
        {
            uint32_t k = (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)(cls + nc * (wt * 16 + ro)) * down, (uint32_t)R,
                                           inv_r);
            float2 ph = nco_phasor((int64_t)k, nu);
            gvec2_ptr out_t = out + j0;
            // Finish sub-tile st from its two chains: combine real and imaginary
            // parts across the lane pair, apply the NCO phasor, store. With
            // kPipe, the next sub-tile's LDS reads and MFMAs were issued just
            // before in the same block of code and interleave with this work.
            auto finish = [&](auto pipe, uint32_t st, const floatx4 &c0, const floatx4 &c1) {
                floatx4 acc;
                if constexpr (kBf)
                    acc = c0;
                else
                    acc = c0 + c1;
                // partner lane's values (quad_perm 1,0,3,2 swaps lanes 2c and 2c+1)
                float oth[4];
#pragma unroll
                for (uint32_t e = 0; e < 4; ++e)
                    oth[e] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[e]), 0xB1, 0xF, 0xF, false));
                const float2 y0 = odd ? make_float2(oth[2], acc[2]) : make_float2(acc[0], oth[0]);
                const float2 y1 = odd ? make_float2(oth[3], acc[3]) : make_float2(acc[1], oth[1]);
                const uint32_t r0 = cls + nc * (st * 16 + ro);  // tile row of the lane's first output
                float2 o0 = cmul(ph, y0);
                float2 o1 = cmul(cmul(ph, s_row), y1);
                // fast path: a whole sub-tile of rows (the lane's two rows are
                // then one 16-byte store), no rollover near the lane's rows
                const bool fast = cls + nc * (st * 16 + 15) < nt && k >= cpp1 && k < k_hi;
                if constexpr (decltype(pipe)::value) {
                    __builtin_amdgcn_sched_group_barrier(0x100, kS, 0);
#pragma unroll
                    for (uint32_t i = 0; i < 2 * kS; ++i) {
                        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                        __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
                    }
                }
                if (__builtin_amdgcn_ballot_w64(!fast) != 0) {
                    // exact path (uniform branch, rare): FP64 NCO index and phasor
                    // of both rows, the rollover correction from the staged tile,
                    // and a fresh start for the next sub-tile
                    if (!fast) {
#pragma unroll
                        for (uint32_t e = 0; e < 2; ++e) {
                            const uint32_t row = r0 + e * nc;
                            const uint32_t kk =
                                (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)row * down, (uint32_t)R, inv_r);
                            float2 y = e ? y1 : y0;
                            if (kk < cpp1) {
                                // window taps before the rollover (t < w-1-kk) carry
                                // e^{j2pi nu R}: add (e^{j2pi nu R} - 1) times their sum
                                const uint32_t split = kW - 1 - kk;
                                const uint32_t rw = min(row, rows - 1);
                                float2 part = make_float2(0.f, 0.f);
                                uint32_t q = 0, rr = 0;
                                for (uint32_t t = 0; t < split; ++t) {
                                    const float2 xv = kBf      ? bf_sample<kWords>(planes, rw * down + base + t)
                                                      : kDense ? xs[rr * mc + q + rw]
                                                               : xs[rw * (kW + 1) + t];
                                    cmac(part, my_taps[t], xv);
                                    if (++rr == down) {
                                        rr = 0;
                                        ++q;
                                    }
                                }
                                cmac(y, make_float2(wrap.x - 1.f, wrap.y), part);
                            }
                            const float2 o = cmul(nco_phasor((int64_t)kk, nu), y);
                            if (e)
                                o1 = o;
                            else
                                o0 = o;
                        }
                        k = (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)(r0 + 16 * wpc * nc) * down, (uint32_t)R,
                                              inv_r);
                        ph = nco_phasor((int64_t)k, nu);
                    } else {
                        k += du;
                        ph = cmul(ph, s_sub);
                    }
                    (live && r0 < nt ? out_t + r0 : sink)[0] = fvec2{o0.x, o0.y};
                    (live && r0 + nc < nt ? out_t + r0 + nc : sink + 1)[0] = fvec2{o1.x, o1.y};
                } else {
                    k += du;
                    ph = cmul(ph, s_sub);
                    if (nc == 1 && out16) {
                        // rows r0, r0+1 of one channel: the 8 lanes of a
                        // channel write its 16 rows, 128 contiguous bytes
                        ((gvec4_ptr)(live ? out_t + r0 : sink))[0] = fvec4{o0.x, o0.y, o1.x, o1.y};
                    } else {
                        (live ? out_t + r0 : sink)[0] = fvec2{o0.x, o0.y};
                        (live ? out_t + r0 + nc : sink + 1)[0] = fvec2{o1.x, o1.y};
                    }
                }
            };
            floatx4 pa, pb;  // the pending sub-tile's chains
            mma(wt, pa, pb);
            for (uint32_t st = wt;;) {
                const uint32_t sn = st + wpc;
                if (sn >= subtiles) {
                    finish(std::false_type{}, st, pa, pb);
                    break;
                }
                floatx4 na, nb;
                mma(sn, na, nb);
                finish(std::true_type{}, st, pa, pb);
                pa = na;
                pb = nb;
                st = sn;
            }
        }
    next_tile:
        if (tn >= tiles)
            break;  // uniform
        tb = tn;
    }
}

// ---------------------------------------------------------------------------
// fused bf16-plane tiles: every bf16-plane group on one staged span
// ---------------------------------------------------------------------------
// A fused tile is L consecutive inputs of the stream (global tile T covers
// stream inputs T*L .. T*L+L-1). Its span is staged once as bf16 planes
// (from s0 = T*L - P - 40, even, not before the 40-sample history) and every
// member group then computes its outputs whose anchors fall in the tile.

// The rows jt .. jt+nt-1 of bf16-plane group gw, whose windows lie in the
// staged span (s0: chunk-local index of plane sample 0); this wave's class.
// A fused record's member group, as its stripe uses it on every tile
struct FusedMember {
    uint64_t n0;        // the chunk's first output of the group
    uint32_t count;     // the chunk's outputs of the group
    uint32_t per_tile;  // outputs per fused tile (L / down)
    uint32_t down, m;   // decimation, channels
    uint32_t lnc;       // log2 of the residue classes
    uint32_t btab;      // offset of the group's B tables in btab
};

template <int kBlock>
__device__ __forceinline__ void bf_rows(const GroupsDev &a, uint64_t P, const FusedMember &fm, uint32_t jt,
                                        uint32_t nt, uint32_t rot, int64_t s0, const uint32_t *planes, const float2 *taps_m,
                                        const float2 *steps_m, const ChanStatic *chans_m, const u32x4 *b_m,
                                        uint32_t b_base, gvec2_ptr sink)
{
    constexpr uint32_t kW = 32, kS = 8, kWaves = kBlock / 64;
    constexpr uint32_t kWords = mfma_stage(kBlock) / 2 * kBlock;
    const uint32_t down = fm.down, m = fm.m, lnc = fm.lnc;
    const uint64_t n0 = fm.n0;
    const uint32_t lane = threadIdx.x & 63, wave = uniform(threadIdx.x >> 6);
    const uint32_t nc = 1u << lnc;
    const uint32_t cls = wave & (nc - 1), wpc = kWaves >> lnc, wi = wave >> lnc, dn = down << lnc;
    const uint32_t crows = nt > cls ? (nt - cls + nc - 1) >> lnc : 0u;
    const uint32_t subtiles = (crows + 15) >> 4;
    const uint32_t wt = (wi + rot) & (wpc - 1);  // rotating sub-tile start, as in mfma_stripe
    if (wt >= subtiles)
        return;  // uniform
    const int64_t a0 = (int64_t)((n0 + jt) * down - P);  // chunk-local anchor of row jt
    const uint32_t w0 = (uint32_t)(a0 - (int64_t)(kW - 1) - s0);  // its window's start in the span
    const uint32_t cstart = (w0 + down * cls) & ~7u, cbase = (w0 + down * cls) & 7u;
    bf16x8 bh[3], bl[3];
    // the class's B operands: from the copy in LDS when the tile has the
    // stripe's usual window offset (every tile but a chunk's first), else
    // from the group's tables in memory (uniform branch)
    if (cbase == ((b_base + down * cls) & 7u)) {
        const u32x4 *bt = b_m + cls * (kBfWordsPerBase / 4);
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            bh[q] = __builtin_bit_cast(bf16x8, bt[(q * 2 + 0) * 64 + lane]);
            bl[q] = __builtin_bit_cast(bf16x8, bt[(q * 2 + 1) * 64 + lane]);
        }
    } else {
        // (buffer loads: a pointer that could be either would make flat loads of both)
        const __amdgpu_buffer_rsrc_t bt = __builtin_amdgcn_make_buffer_rsrc(
            const_cast<float *>(a.btab + fm.btab + cbase * kBfWordsPerBase), 0, (int)(kBfWordsPerBase * 4), kBufferFlags);
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            bh[q] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(bt, (int)(((q * 2 + 0) * 64 + lane) * 16), 0, 0));
            bl[q] = __builtin_bit_cast(bf16x8, __builtin_amdgcn_raw_buffer_load_b128(bt, (int)(((q * 2 + 1) * 64 + lane) * 16), 0, 0));
        }
    }
    uint32_t aoff[3];
#pragma unroll
    for (uint32_t q = 0; q < 3; ++q) {
        const uint32_t kb = 4 * q + (lane >> 4);
        const uint32_t plane = kb >= 5 && kb < 10 ? 1u : 0u;
        const uint32_t slot = kb < 5 ? 8 * kb : kb < 10 ? 8 * (kb - 5) : 0u;
        aoff[q] = plane * kWords * 4 + (cstart + dn * (lane & 15) + slot) * 2;
    }
    const uint32_t col = lane & 15;
    const bool odd = col & 1, live = (col >> 1) < m;
    const uint32_t cl = min(col >> 1, m - 1);
    const ChanStatic *ch = chans_m + cl;  // LDS copy
    const uint64_t R = ch->r;
    const uint32_t cpp1 = ch->cpp - 1;
    const double nu = ch->nu, inv_r = ch->inv_r;
    const float2 wrap = ch->wrap;
    gvec2_ptr out_t = (gvec2_ptr)ch->out + jt;
    const bool pair16 = nc == 1 && a.out16 && (jt & 1) == 0;
    const float2 *my_taps = taps_m + cl * kW;
    const float2 s_row = steps_m[2 * cl], s_sub = steps_m[2 * cl + 1];
    const uint32_t ro = 4 * (lane >> 4) + (odd ? 2 : 0);
    const uint32_t du = 16 * wpc * dn;
    const uint32_t k_hi = R > (uint64_t)du + dn ? (uint32_t)(R - du - dn) : 0u;

    auto mma = [&](uint32_t st, floatx4 &c0, floatx4 &c1) {
        const unsigned char *pb = reinterpret_cast<const unsigned char *>(planes) + st * 16 * dn * 2;
        bf16x8 ah[3], al[3];
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            ah[q] = *reinterpret_cast<const bf16x8 *>(pb + aoff[q]);
            al[q] = *reinterpret_cast<const bf16x8 *>(pb + aoff[q] + 2 * kWords * 4);
        }
        // one chain, as in mfma_stripe
        c0 = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (uint32_t q = 0; q < 3; ++q) {
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bl[q], c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[q], bh[q], c0, 0, 0, 0);
            c0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[q], bh[q], c0, 0, 0, 0);
        }
        c1 = c0;
    };

    uint32_t k = (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)(cls + nc * (wt * 16 + ro)) * down, (uint32_t)R, inv_r);
    float2 ph = nco_phasor((int64_t)k, nu);
    // as in mfma_stripe: combine across the lane pair, phasor, store; the
    // next sub-tile's MFMAs interleave with this work when kPipe
    auto finish = [&](auto pipe, uint32_t st, const floatx4 &c0, const floatx4 &c1) {
        const floatx4 acc = c0;
        (void)c1;
        float oth[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e)
            oth[e] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[e]), 0xB1, 0xF, 0xF, false));
        const float2 y0 = odd ? make_float2(oth[2], acc[2]) : make_float2(acc[0], oth[0]);
        const float2 y1 = odd ? make_float2(oth[3], acc[3]) : make_float2(acc[1], oth[1]);
        const uint32_t r0 = cls + nc * (st * 16 + ro);
        float2 o0 = cmul(ph, y0);
        float2 o1 = cmul(cmul(ph, s_row), y1);
        const bool fast = cls + nc * (st * 16 + 15) < nt && k >= cpp1 && k < k_hi;
        if constexpr (decltype(pipe)::value) {
            __builtin_amdgcn_sched_group_barrier(0x100, kS, 0);
#pragma unroll
            for (uint32_t i = 0; i < 2 * kS; ++i) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
                __builtin_amdgcn_sched_group_barrier(0x002, 2, 0);
            }
        }
        if (__builtin_amdgcn_ballot_w64(!fast) != 0) {
            if (!fast) {
#pragma unroll
                for (uint32_t e = 0; e < 2; ++e) {
                    const uint32_t row = r0 + e * nc;
                    const uint32_t kk = (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)row * down, (uint32_t)R, inv_r);
                    float2 y = e ? y1 : y0;
                    if (kk < cpp1) {
                        const uint32_t split = kW - 1 - kk;
                        const uint32_t rw = min(row, nt - 1);  // rows past nt store nothing
                        float2 part = make_float2(0.f, 0.f);
                        for (uint32_t t = 0; t < split; ++t)
                            cmac(part, my_taps[t], bf_sample<kWords>(planes, w0 + rw * down + t));
                        cmac(y, make_float2(wrap.x - 1.f, wrap.y), part);
                    }
                    const float2 o = cmul(nco_phasor((int64_t)kk, nu), y);
                    if (e)
                        o1 = o;
                    else
                        o0 = o;
                }
                k = (uint32_t)mod_u64(P + (uint64_t)a0 + (uint64_t)(r0 + 16 * wpc * nc) * down, (uint32_t)R, inv_r);
                ph = nco_phasor((int64_t)k, nu);
            } else {
                k += du;
                ph = cmul(ph, s_sub);
            }
            (live && r0 < nt ? out_t + r0 : sink)[0] = fvec2{o0.x, o0.y};
            (live && r0 + nc < nt ? out_t + r0 + nc : sink + 1)[0] = fvec2{o1.x, o1.y};
        } else {
            k += du;
            ph = cmul(ph, s_sub);
            if (pair16) {
                ((gvec4_ptr)(live ? out_t + r0 : sink))[0] = fvec4{o0.x, o0.y, o1.x, o1.y};
            } else {
                (live ? out_t + r0 : sink)[0] = fvec2{o0.x, o0.y};
                (live ? out_t + r0 + nc : sink + 1)[0] = fvec2{o1.x, o1.y};
            }
        }
    };
    floatx4 pa, pb;
    mma(wt, pa, pb);
    for (uint32_t st = wt;;) {
        const uint32_t sn = st + wpc;
        if (sn >= subtiles) {
            finish(std::false_type{}, st, pa, pb);
            break;
        }
        floatx4 na, nb;
        mma(sn, na, nb);
        finish(std::true_type{}, st, pa, pb);
        pa = na;
        pb = nb;
        st = sn;
    }
}

// the stream's global fused tiles that reach into this chunk: T0 ..
__device__ __forceinline__ uint32_t fused_tiles(uint64_t P, uint32_t n, uint32_t L)
{
    return (uint32_t)((P + n - 1) / L - P / L + 1);
}

template <int kBlock>
__device__ __forceinline__ void fused_stripe(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P,
                                             const uint32_t *fw, TileSeq seq, uint32_t *slot, unsigned char *smem)
{
    constexpr uint32_t kW = 32, kWaves = kBlock / 64;
    constexpr uint32_t kPairs = mfma_stage(kBlock) / 2;
    const uint32_t L = fw[2], nm = fw[4];
    // LDS: members' B operands (one 6 KB table per residue class, for the
    // stripe's usual window offset), the planes, taps, NCO steps, channels
    u32x4 *btl = reinterpret_cast<u32x4 *>(smem);
    uint32_t nb = 0;  // B tables
    for (uint32_t i = 0; i < nm; ++i)
        nb += mfma_bf16_classes(a.groups[uniform(fw[8 + i]) * kGroupWords + 2]);
    uint32_t *planes = reinterpret_cast<uint32_t *>(btl + nb * (kBfWordsPerBase / 4));
    float2 *taps_l = reinterpret_cast<float2 *>(planes + mfma_stage(kBlock) * kBlock * 2);
    float2 *steps = taps_l + nm * kMfmaMaxChan * kW;
    ChanStatic *chl = reinterpret_cast<ChanStatic *>(steps + nm * kMfmaMaxChan * 2);
    FusedMember *mem = reinterpret_cast<FusedMember *>(chl + nm * kMfmaMaxChan);
    gvec2_ptr sink = (gvec2_ptr)(a.sink + 2 * ((blockIdx.x % (kSinkWords / (2 * kBlock))) * kBlock + threadIdx.x));
    // a full tile's first row has its window at w_std = 9 or 10 samples into
    // the span (the span starts 40 samples before the tile, rounded down to
    // an even index)
    const uint32_t w_std = 9u + (uint32_t)(P & 1u);
    // members' taps (zero-padded in front), channels, B tables, and NCO steps:
    // one row of the member's class (down*nc inputs) and one sub-tile of its
    // wave (16*wpc rows)
    for (uint32_t i = 0, bi = 0; i < nm; ++i) {
        const uint32_t *gw = a.groups + uniform(fw[8 + i]) * kGroupWords;
        const uint32_t m = gw[4], down = gw[2];
        for (uint32_t e = threadIdx.x; e < m * (sizeof(ChanStatic) / 4); e += kBlock)
            reinterpret_cast<uint32_t *>(chl + i * kMfmaMaxChan)[e] =
                reinterpret_cast<const uint32_t *>(a.chans + gw[8 + e / (sizeof(ChanStatic) / 4)])[e % (sizeof(ChanStatic) / 4)];
        const uint32_t nc = mfma_bf16_classes(down);
        for (uint32_t c = 0; c < nc; ++c, ++bi) {
            const u32x4 *src =
                reinterpret_cast<const u32x4 *>(a.btab + (gw[7] - 1) + ((w_std + down * c) & 7u) * kBfWordsPerBase);
            for (uint32_t e = threadIdx.x; e < kBfWordsPerBase / 4; e += kBlock)
                btl[bi * (kBfWordsPerBase / 4) + e] = src[e];
        }
        for (uint32_t e = threadIdx.x; e < m * kW; e += kBlock) {
            const ChanStatic &c = a.chans[gw[8 + e / kW]];
            const uint32_t t = e % kW, pad = kW - c.cpp;
            taps_l[i * kMfmaMaxChan * kW + e] = t < pad ? make_float2(0.f, 0.f) : a.taps[c.taps_off + t - pad];
        }
        if (threadIdx.x == 0) {
            // L is a multiple of down: the tile's outputs are per_tile whole rows
            const double inv_down = a.chans[gw[8]].inv_down;
            const uint64_t n0 = cdiv(P, down, inv_down);
            mem[i] = FusedMember{n0, (uint32_t)(cdiv(P + a.n, down, inv_down) - n0), L / down, down, m,
                                 3u - min(3u, (uint32_t)__builtin_ctz(down)), gw[7] - 1};
        }
        if (threadIdx.x < m) {
            const double nu = a.chans[gw[8 + threadIdx.x]].nu;
            const uint32_t nc = mfma_bf16_classes(down);
            steps[(i * kMfmaMaxChan + threadIdx.x) * 2] = nco_phasor((int64_t)(down * nc), nu);
            steps[(i * kMfmaMaxChan + threadIdx.x) * 2 + 1] = nco_phasor((int64_t)(16 * (kWaves / nc) * down * nc), nu);
        }
    }
    const uint64_t T0 = P / L;
    const uint32_t tiles = fused_tiles(P, a.n, L);
    // span start of tile t (relative to T0): 40 samples before the tile, even,
    // and not before the history
    auto span0 = [&](uint32_t t) -> int64_t {
        int64_t s0 = (int64_t)((T0 + t) * L) - (int64_t)P - 40;
        s0 = s0 < -40 ? -40 : s0;
        return s0 & ~(int64_t)1;
    };
    uint32_t tb = uniform(*slot);
    u32x4 vb[kPairs];
    bf_load<kBlock>(a, hist, span0(tb), vb);
    for (;;) {
        const int64_t s0 = span0(tb);
        __syncthreads();  // the previous tile's reads of LDS (and of *slot) are done; taps and steps written
        bf_commit<kBlock>(vb, planes);
        seq.claim(slot);
        __syncthreads();
        const uint32_t tn = uniform(*slot);
        if (tn < tiles)
            bf_load<kBlock>(a, hist, span0(tn), vb);
        const uint64_t Tg = T0 + tb;
        for (uint32_t i = 0, bi = 0; i < nm; ++i) {
            FusedMember fm;
            {
                const FusedMember &f = mem[i];
                fm.n0 = ((uint64_t)uniform((uint32_t)(f.n0 >> 32)) << 32) | uniform((uint32_t)f.n0);
                fm.count = uniform(f.count);
                fm.per_tile = uniform(f.per_tile);
                fm.down = uniform(f.down);
                fm.m = uniform(f.m);
                fm.lnc = uniform(f.lnc);
                fm.btab = uniform(f.btab);
            }
            const uint32_t bi_i = bi;
            bi += 1u << fm.lnc;
            // outputs whose stream anchors (n0 + j)*down fall in the tile
            const uint64_t jl = Tg * fm.per_tile, jh = jl + fm.per_tile;
            const uint32_t jt = jl > fm.n0 ? (uint32_t)min<uint64_t>(jl - fm.n0, fm.count) : 0u;
            const uint32_t je = jh > fm.n0 ? (uint32_t)min<uint64_t>(jh - fm.n0, fm.count) : 0u;
            if (je > jt)
                bf_rows<kBlock>(a, P, fm, jt, je - jt, tb, s0, planes, taps_l + i * kMfmaMaxChan * kW,
                                steps + i * kMfmaMaxChan * 2, chl + i * kMfmaMaxChan, btl + bi_i * (kBfWordsPerBase / 4),
                                w_std, sink);
        }
        if (tn >= tiles)
            break;  // uniform
        tb = tn;
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
        store_out(ch.out, j0 + q, cmul(nco_phasor((int64_t)k, ch.nu), acc));
    }
}

// All of this block's tiles of group g from seq.
template <int kBlock, bool kWide>
__device__ __forceinline__ void run_group(const GroupsDev &a, const float2 *__restrict__ hist, uint64_t P, uint32_t g, TileSeq seq,
                          uint32_t *slot, unsigned char *smem)
{
    const uint32_t *gw = a.groups + g * kGroupWords;
    const uint32_t kind = gw[0], up = gw[1], down = gw[2], rows = gw[5];
    if (kind == kGroupFused) {
        seq.tiles = fused_tiles(P, a.n, down);
        seq.start(blockIdx.x, slot + 1);
        seq.claim(slot);
        __syncthreads();
        if (uniform(*slot) < seq.tiles)
            fused_stripe<kBlock>(a, hist, P, gw, seq, slot, smem);
        return;
    }
    const double inv_down = a.chans[gw[8]].inv_down;
    const uint64_t n0 = cdiv(P * up, down, inv_down);
    const uint32_t count = (uint32_t)(cdiv((P + a.n) * up, down, inv_down) - n0);
    const uint32_t tiles = (count + rows - 1) / rows;
    seq.tiles = tiles;
    seq.start(blockIdx.x, slot + 1);
    seq.claim(slot);
    __syncthreads();
    uint32_t tb = uniform(*slot);
    if (tb >= tiles)
        return;  // uniform: nothing (left) for this block
    if (kind == kGroupMfma) {
        if constexpr (kWide) {
            if (gw[3] != 32) {
                if (down < 64)
                    mfma_stripe<kBlock, 16, true>(a, hist, P, gw, n0, count, seq, slot, smem);
                else
                    mfma_stripe<kBlock, 16, false>(a, hist, P, gw, n0, count, seq, slot, smem);
                return;
            }
        }
        if (down < 32) {
            if (gw[7] && a.in16 && (8u >> min(3u, (uint32_t)__builtin_ctz(down))) <= kBlock / 64)
                mfma_stripe<kBlock, 8, true, true>(a, hist, P, gw, n0, count, seq, slot, smem);
            else
                mfma_stripe<kBlock, 8, true>(a, hist, P, gw, n0, count, seq, slot, smem);
        } else
            mfma_stripe<kBlock, 8, false>(a, hist, P, gw, n0, count, seq, slot, smem);
        return;
    }
    for (;;) {
        const uint32_t j0 = tb * rows;
        const uint32_t nt = min(rows, count - j0);
        const int64_t a0 = (int64_t)((n0 + j0) * down - P);  // up == 1: anchor of row 0
        switch (kind) {
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
        __syncthreads();  // this tile's reads of LDS (and of *slot) are done
        seq.claim(slot);
        __syncthreads();
        tb = uniform(*slot);
        if (tb >= tiles)
            return;
    }
}

// kWide: the plan holds 64-tap MFMA groups (their B operands and window
// reads take 32 more registers, so plans without them get a leaner build
// that fits 128 registers: four waves per SIMD)
template <int kBlock, bool kWide>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kWide ? 1 : 4))) void chan_groups_kernel(
    GroupsDev a)
{
    extern __shared__ __align__(16) unsigned char smem[];
    // stream position and history parity: from the arguments, or (graph
    // replays) from device memory; wavefront-uniform either way
    uint64_t P = a.p;
    uint32_t par = a.parity;
    if (a.pos) {
        P = a.pos->p;
        par = a.pos->parity;
    }
    P = ((uint64_t)uniform((uint32_t)(P >> 32)) << 32) | uniform((uint32_t)P);
    par = uniform(par);
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
    __shared__ uint32_t s_tile[2];  // next tile, claim shards spent
    const uint32_t g0 = uniform(a.table[2 * blockIdx.x]);
    if (!a.claims) {
        if (g0 == kNoGroup)
            return;  // padding of an XCD's queue
        // fixed split: the stripe s*tps .. s*tps + tps - 1, or (strided) every
        // tps-th tile from s on
        const uint32_t s_idx = uniform(a.table[2 * blockIdx.x + 1]);
        TileSeq seq{nullptr, a.strided ? s_idx : s_idx * a.tps[g0], a.strided ? a.tps[g0] : 1u,
                    a.strided ? 0xffffffffu : s_idx * a.tps[g0] + a.tps[g0]};
        run_group<kBlock, kWide>(a, hist, P, g0, seq, s_tile, smem);
        return;
    }
    // claimed tiles: start with this block's group, then help the others;
    // every block reaches the finished-block count below
    // this block's group's place in the schedule (a padding entry starts at 0)
    uint32_t i0 = 0;
    for (uint32_t i = 0; i < a.n_sched; ++i)
        if (a.sched[i] == g0)
            i0 = i;
    for (uint32_t k = 0, i = i0; k < a.n_sched; ++k, i = i + 1 == a.n_sched ? 0 : i + 1) {
        const uint32_t g = a.sched[i];
        __syncthreads();  // the previous group's reads of LDS are done
        run_group<kBlock, kWide>(a, hist, P, g, TileSeq{a.claims + g * kClaimShards * kClaimPitch}, s_tile,
                                 smem);
    }
    // the last block out resets the counters for the next launch
    uint32_t *const finished = a.claims + (size_t)kGroupsInline * kClaimShards * kClaimPitch;
    __syncthreads();
    if (threadIdx.x == 0) {
        __threadfence();
        s_tile[0] = atomicAdd(finished, 1u) == gridDim.x - 1 ? 1u : 0u;
    }
    __syncthreads();
    if (uniform(s_tile[0])) {
        for (uint32_t i = threadIdx.x; i < a.n_groups * kClaimShards; i += kBlock)
            atomicExch(a.claims + i * kClaimPitch, 0u);
        if (threadIdx.x == 0)
            atomicExch(finished, 0u);
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
    d.n_sched = f.n_sched;
    for (uint32_t i = 0; i < f.n_sched && i < kGroupsInline; ++i)
        d.sched[i] = (uint16_t)f.sched[i];
    d.table = f.table;
    d.strided = f.strided ? 1u : 0u;
    d.claims = f.claims;
    d.sink = f.sink;
    d.out16 = f.out16 ? 1u : 0u;
    d.in16 = f.in16 ? 1u : 0u;
    for (uint32_t k = 0; k < f.n_groups; ++k)
        d.tps[k] = f.tps[k];
    const uint32_t blocks = f.blocks;
    auto go = [&](auto kernel, uint32_t threads) {
        hipLaunchKernelGGL(kernel, dim3(blocks), dim3(threads), f.lds_bytes, s, d);
    };
    switch (f.block * 2 + (f.wide ? 1 : 0)) {
    case 256:
        go(chan_groups_kernel<128, false>, 128);
        break;
    case 257:
        go(chan_groups_kernel<128, true>, 128);
        break;
    case 1024:
        go(chan_groups_kernel<512, false>, 512);
        break;
    case 1025:
        go(chan_groups_kernel<512, true>, 512);
        break;
    case 513:
        go(chan_groups_kernel<256, true>, 256);
        break;
    default:
        go(chan_groups_kernel<256, false>, 256);
        break;
    }
    return hipGetLastError();
}

int chan_groups_occupancy(uint32_t block, bool wide, size_t lds_bytes)
{
    int n = 0;
    hipError_t e = hipSuccess;
    switch (block * 2 + (wide ? 1 : 0)) {
    case 256:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<128, false>, 128, lds_bytes);
        break;
    case 257:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<128, true>, 128, lds_bytes);
        break;
    case 1024:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<512, false>, 512, lds_bytes);
        break;
    case 1025:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<512, true>, 512, lds_bytes);
        break;
    case 513:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<256, true>, 256, lds_bytes);
        break;
    default:
        e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, chan_groups_kernel<256, false>, 256, lds_bytes);
        break;
    }
    return e == hipSuccess ? n : 0;
}

hipError_t launch_chan_advance(ChanPos *pos, uint32_t n, hipStream_t s)
{
    hipLaunchKernelGGL(chan_advance_kernel, dim3(1), dim3(64), 0, s, pos, n);
    return hipGetLastError();
}

}  // namespace fdsp
