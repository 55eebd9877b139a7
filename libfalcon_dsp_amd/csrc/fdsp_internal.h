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

// Complex int16 samples (fdsp_ci16 on the host, short2 here) are read
// exactly as floats. Results go back to int16 as the reference's
// float -> int16_t conversions do (std::complex<int16_t> built from float
// parts, e.g. src/transform/falcon_dsp_fir_filter.cc:168-171): truncation
// toward zero. Outside the int16 range, where that conversion is undefined,
// they saturate.
__device__ __forceinline__ float2 sample_f32(float2 v) { return v; }
__device__ __forceinline__ float2 sample_f32(short2 v) { return make_float2((float)v.x, (float)v.y); }
__device__ __forceinline__ short i16_trunc(float v) { return (short)(int)fminf(fmaxf(v, -32768.f), 32767.f); }
__device__ __forceinline__ void sample_put(float2 *p, float2 v) { *p = v; }
__device__ __forceinline__ void sample_put(short2 *p, float2 v) { *p = make_short2(i16_trunc(v.x), i16_trunc(v.y)); }

// a * b rounded as std::complex<float> multiplication rounds it, four
// products and two sums with no fused multiply-add (the integer-output paths
// compute as the reference does, so their truncated results agree bit for bit)
__device__ __forceinline__ float2 cmul_ref(float2 a, float2 b)
{
#pragma clang fp contract(off)
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// e^{j 2 pi nu k}: the phase is formed and reduced to [-1/2, 1/2] revolutions
// in FP64 (k < 2^32 and |nu * k| < 2^21, so the reduced phase carries
// < 1e-9 revolutions of error), then the hardware v_sin_f32 / v_cos_f32,
// which take their argument in revolutions, evaluate it (max abs error
// 1.3e-7 over [-1/2, 1/2], measured against FP64 on gfx950).
__device__ __forceinline__ float2 nco_phasor(int64_t k, double nu)
{
    double t = (double)k * nu;
    t -= rint(t);
    const float tf = (float)t;
    return make_float2(__builtin_amdgcn_cosf(tf), __builtin_amdgcn_sinf(tf));
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
// the offsets of an NCO thread's samples from its first: pairs 128 apart, so
// a wave's 16-byte accesses each cover 1 KB of consecutive samples
constexpr int kNcoSamples = 16;
constexpr uint32_t kNcoOffsets[kNcoSamples] = {0, 1, 128, 129, 256, 257, 384, 385,
                                               512, 513, 640, 641, 768, 769, 896, 897};

struct NcoChan {
    double nu;        // f/fs, signed, revolutions per sample
    double inv_r;     // 1/R
    uint32_t r;       // rollover index R (UINT32_MAX for the identity shift)
    uint32_t tab;     // first entry of this shift's phasor table (kNoTable: none)
    float2 step[kNcoSamples];  // e^{j 2 pi nu o_i}, o = kNcoOffsets: a thread's samples relative to its first
    double omega;     // 2 pi f / fs rad/sample, formed as the reference forms it
};

// static description of one channel of the fused channelizer
struct ChanStatic {
    double nu;        // f/fs (0 for the identity shift)
    double inv_r;     // 1/R
    double inv_up, inv_down;
    uint32_t r;       // rollover index
    uint32_t up, down, cpp;
    uint32_t taps_off;  // offset of this channel's modulated taps (up*cpp complex)
    uint32_t tail;
    float2 wrap;      // e^{j 2 pi nu R}: phase jump across a rollover inside a window
    void *out;        // chunk output base (set per launch)
    uint32_t win;     // segment kernel: window (cpp padded in front to 32 or 64)
    uint32_t poff;    // segment kernel: offset of the padded modulated taps (up*win complex)
    float2 *prev[2];  // segment kernel, discriminator tail: the stream's last complex output,
                      // read from [parity], written to [parity ^ 1]
    uint32_t dk;      // segment kernel, MFMA groups: (16 kSegWaves down) mod R, a wave's NCO index step
    uint32_t d1;      // down mod R
    float2 e_dk;      // e^{j 2 pi nu dk}
    float2 e_d1;      // e^{j 2 pi nu d1}
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
// TI, TO: float2 (cf32) or short2 (ci16) samples, in the pairs
// (float2, float2) and (short2, short2). tab (int16 output only, may be
// null): per-shift tables of the reference's phasors, entry NcoChan::tab + s
// holding ((float)cos(omega s), (float)sin(omega s)) for s < R.
constexpr uint32_t kNoTable = 0xFFFFFFFFu;
template <typename TI, typename TO>
hipError_t launch_nco_multi(const TI *in, uint32_t n, uint64_t start, const NcoChan *chans, uint32_t n_chan,
                            TO *out, uint64_t out_stride, hipStream_t s, const float2 *tab = nullptr);

// polyphase resampler over one chunk (global input position p); `taps` are
// the real parts of the transposed bank, up*cpp floats. hist_in holds cpp-1
// samples (may be null when cpp == 1); hist_out receives the new history.
hipError_t launch_resampler(const float2 *in, uint32_t n, uint64_t p, uint32_t up, uint32_t down,
                            uint32_t cpp, const float *taps, const float2 *hist_in, float2 *hist_out,
                            float2 *out, uint32_t count, hipStream_t s);

// FIR with complex taps over one chunk; hist_in/hist_out hold m-1 samples
// (float whatever the sample type); TI, TO as for launch_nco_multi
template <typename TI, typename TO>
hipError_t launch_fir(const TI *in, uint32_t n, const float2 *taps, uint32_t m, const float2 *hist_in,
                      float2 *hist_out, TO *out, hipStream_t s);

// FIR bank with real taps, n_chan channels of n samples each
hipError_t launch_fir_bank(const float2 *in, uint32_t n, uint64_t in_stride, const float *taps, uint32_t m,
                           uint32_t n_chan, const float2 *hist_in, float2 *hist_out, float2 *out,
                           uint64_t out_stride, hipStream_t s);

// polar discriminator; prev (device, 1 sample) is read when has_prev, and
// prev_out receives the last input sample
// (float2 or short2 input; the carried sample is float)
template <typename TI>
hipError_t launch_discriminator(const TI *in, uint32_t n, const float2 *prev, int has_prev, float2 *prev_out,
                                float *out, hipStream_t s);

// Work plan of the fused channelizer, built on the host at create time: the
// channels fall in groups, each computed by its own row tiles (one block per
// tile of `rows` consecutive outputs). Group record, kGroupWords u32 each:
//   [0] kind  [1] up  [2] down  [3] taps per window (padded to 32/64 for MFMA)
//   [4] channels m  [5] rows per tile  [6] offset of the MFMA B table (floats)
//   [7] unused  [8 .. 8+m) channel indices
enum : uint32_t {
    kGroupMfma = 0,    // 3..8 channels, up 1, equal down and padded taps: FP32 MFMA
    kGroupVDense = 1,  // one channel, up 1, down < cpp: VALU over a transposed span
    kGroupVSparse = 2, // one channel, up 1, down >= cpp: VALU over gathered windows
    kGroupGeneric = 3, // one channel, up > 1: VALU with per-output filter phase
    kGroupFused = 4,   // bf16-plane MFMA groups sharing staged spans: [2] L inputs per
                       // tile, [4] member count, [8 ..] member group indices
};
// member groups of a fused record at most
constexpr uint32_t kFusedMax = 4;
constexpr uint32_t kGroupWords = 16;
constexpr uint32_t kMfmaMaxChan = 8;
constexpr uint32_t kMfmaMinChan = 3;
// largest down of an MFMA group (keeps a tile's buffer offsets below 2^30)
constexpr uint32_t kMfmaMaxDown = 1u << 20;
// groups whose tiles-per-block count travels in the kernel arguments
constexpr uint32_t kGroupsInline = 256;
// MI355X: 256 CUs in 8 XCDs, each with its own L2; consecutive workgroup ids
// are dealt round-robin to the XCDs
constexpr uint32_t kXcds = 8;
// tile claim counters of the batched kernel: kClaimShards per group (shard s
// hands out tiles s, s + kClaimShards, ...), each on its own 128-byte line,
// then one finished-block counter
constexpr uint32_t kClaimShards = 8;
constexpr uint32_t kClaimPitch = 32;  // u32
constexpr size_t kClaimWords = ((size_t)kGroupsInline * kClaimShards + 1) * kClaimPitch;
// samples of the batched kernel's sink: MFMA lanes and rows with no output
// store there instead of branching around the store
constexpr uint32_t kSinkWords = 1u << 16;
// block-table entry of a block with no work (pads the per-XCD queues)
constexpr uint32_t kNoGroup = 0xffffffffu;
// taps per LDS pass of the one-channel sparse path
constexpr uint32_t kVSparsePass = 8;
// outputs per lane of the one-channel dense path
constexpr uint32_t kVDenseOpl = 2;
// MFMA groups of 32-tap windows with an even down below 32 can run on bf16
// matrix cores: the staged span is kept as bf16 hi/lo planes from an
// 8-aligned start, rows are taken in nc = 8/gcd(down, 8) <= 4 classes whose
// windows share one offset from an 8-aligned slot, and each product is
// hi*hi + hi*lo + lo*hi (error about 2^-16 of each term). Such channels form
// groups from one channel up. Their group record holds the offset of the
// bf16 B tables + 1 in word 7.
__host__ __device__ constexpr bool mfma_bf16_planes(uint32_t w, uint32_t down)
{
    return w == 32 && down < 32 && down % 2 == 0;
}
__host__ __device__ constexpr uint32_t mfma_bf16_classes(uint32_t down)
{
    return down % 8 == 0 ? 1u : down % 4 == 0 ? 2u : 4u;
}
// MFMA B words per base (8 bf16 per lane, 3 K steps of 32, hi and lo)
constexpr uint32_t kBfWordsPerBase = 3 * 2 * 64 * 4;
// input samples per thread staged for one MFMA tile (4096 per block at most)
constexpr uint32_t mfma_stage(uint32_t block) { return block >= 256 ? 8 : 16; }

struct GroupsArgs {
    const float2 *in;
    uint32_t n;
    uint32_t hist_len;          // H: history samples kept (>= every window - 1, even)
    const float2 *hist[2];      // ping-pong history buffers
    float2 *hist_w[2];
    const ChanStatic *chans;
    const float2 *taps;         // all channels' modulated taps
    const float *btab;          // MFMA B tables in lane order
    const uint32_t *groups;     // n_groups records of kGroupWords
    uint32_t n_groups;
    uint32_t block;             // threads per block: 128, 256 or 512
    uint32_t lds_bytes;         // dynamic LDS per block
    bool wide;                  // the plan holds 64-tap MFMA groups
    const ChanPos *pos;         // device position (graph mode), or null:
    uint64_t p;                 //   then the position and parity come from here
    uint32_t parity;
    const uint32_t *table;      // device: (group, stripe) of each block, 2 x blocks u32
    uint32_t blocks;
    const uint32_t *tps;        // host: tiles per block (one stripe) of each group, or
    bool strided;               //   (strided) the group's block count: block s runs tiles s, s+tps, ...
    float2 *sink;               // device: kSinkWords samples, written and never read
    bool out16;                 // every channel's output is 16-byte aligned (paired-row stores)
    bool in16;                  // the chunk's input is 16-byte aligned (bf16-plane staging)
    const uint32_t *sched;      // host: the groups this launch schedules (claimed tiles visit only these)
    uint32_t n_sched;
    uint32_t *claims;           // device: kClaimWords zeroed words, or null: blocks then claim
                                //   tiles one by one, starting in their table group
};
hipError_t launch_chan_groups(const GroupsArgs &a, hipStream_t s);

// Work plan of the segment-major FP32 channelizer (fdsp_chan_seg.hip): one
// workgroup per segment of `seg` inputs, staged once with `halo` samples of
// history before it; records in the order the waves take their work.
enum : uint32_t {
    kSegMfma = 0,  // 2..8 channels, up 1, equal down, <= 32 taps: FP32 MFMA sub-tiles of 16 outputs
    kSegValu = 1,  // one channel, window <= 64: one output per lane
};
constexpr uint32_t kSegMaxChan = 256;    // channels of one plan at most
constexpr uint32_t kSegWaves = 4;        // waves per workgroup (256 threads)
constexpr uint32_t kSegMaxGroups = 64;   // work groups of one plan at most
constexpr uint32_t kSegMaxHalo = 256;    // staged history samples at most
constexpr uint32_t kNoShiftPeriod = 1u << 30;  // NCO period used for a zero shift (any period is exact there)
struct SegGroup {
    uint32_t kind;     // kSegMfma or kSegValu
    uint32_t m;        // channels
    uint32_t up, down;
    uint32_t win;      // window (padded taps per phase): 32 or 64
    uint32_t btab;     // MFMA: offset of the group's B operands (16 x 64 floats)
    uint32_t tail;     // VALU: the channel ends in the polar discriminator
    uint32_t pad;
    uint32_t chan[8];  // channel indices
};
struct SegArgs {
    const float2 *in;
    uint32_t n;
    uint32_t seg, halo, hist_len;
    const float2 *hist[2];
    float2 *hist_w[2];
    const ChanStatic *chans;
    const float2 *ptaps;
    const float *btab;
    const SegGroup *groups;
    uint32_t n_groups, n_chan;
    const ChanPos *pos;
    uint64_t p;
    uint32_t parity;
    bool out16;
    bool tails;  // some group ends in the polar discriminator
};
hipError_t launch_chan_seg(const SegArgs &a, hipStream_t s);
size_t chan_seg_lds_bytes(uint32_t seg, uint32_t halo, uint32_t n_chan);
bool chan_seg_len_ok(uint32_t seg);  // a segment length the segment kernel takes
// what a launch of the segment kernel over n inputs runs: the segment length
// used, the workgroups of full length and the shorter length after them, the
// grid and the LDS per workgroup
struct SegGeom {
    uint32_t seg, n_big, small, blocks;
    size_t lds;
};
SegGeom chan_seg_geometry(uint32_t n, uint32_t seg, uint32_t halo, uint32_t n_chan);
// resident blocks of the fused kernel per CU for a block size, build and LDS size
int chan_groups_occupancy(uint32_t block, bool wide, size_t lds_bytes);
// advance the device position by n inputs and flip the history parity
hipError_t launch_chan_advance(ChanPos *pos, uint32_t n, hipStream_t s);

}  // namespace fdsp
