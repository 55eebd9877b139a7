// fdsp_chan_seg.hip -- segment-major fused multi-rate channelizer in FP32.
// (Replaces falcon_dsp_multi_rate_channelizer{,_cuda}::apply, reference
//  src/transform/falcon_dsp_multi_rate_channelizer.cc:146-164 and
//  src/transform/falcon_dsp_multi_rate_channelizer_cuda.cu:375-514, whose
//  per-channel work is falcon_dsp_freq_shift.cc:164-192 followed by
//  falcon_dsp_polyphase_resampler.cc:256-319.)
//
// One workgroup owns one segment of L consecutive inputs of the chunk. It
// stages the segment, plus the H samples before it, in LDS once, then
// computes every channel's outputs whose anchors fall inside the segment.
// The input is read from HBM once for all channels, every workgroup carries
// the same mix of work, and no arithmetic is narrower than FP32.
//
// Shift-then-filter equals filter-with-modulated-taps then one phasor per
// output (k(a): the NCO index of anchor a, d_t = cpp-1-t):
//   y = e^{j2pi nu k(a)} * sum_t [Re(h_t) e^{-j2pi nu d_t}] x[a - d_t]
// as long as no rollover (the index restarting at 0 after R-1) falls inside
// the window. When one does (k(a) < cpp-1), the samples before it carry an
// extra e^{j2pi nu R}, added as (e^{j2pi nu R} - 1) times their partial sum.
//
// Work groups (SegGroup records, built on the host):
//   MFMA  2..8 channels with up 1, equal down and at most 32 taps. Their
//         32-sample windows (taps zero-padded in front) are the A operand of
//         v_mfma_f32_16x16x4f32: 16 outputs x 64 real K values. Lane i + 16g
//         holds window samples g, g+4, ..., g+28 of row i, the real part of
//         sample g+4j at K step 2j and its imaginary part at step 2j+1. The
//         modulated taps are the B operand in the same K order: channel c's
//         real output in column 2c, its imaginary output in 2c+1. A sub-tile
//         is 16 outputs of every channel of the group.
//   VALU  one channel (any up), window of at most 64 taps: one output per
//         lane, taps zero-padded in front to kW = 32 or 64. With up 1 the taps
//         are wavefront-uniform scalar loads; with up > 1 each lane reads the
//         bank row of its filter phase (n*down) mod up.
//
// LDS layout of the staged span: transposed by 4, span sample q at slot
// (q mod 4)*km + q/4, km odd and sized for the segment and the plan's halo.
// Samples 4 apart are adjacent, so a lane's eight MFMA samples (or a VALU
// window's samples of one residue) are eight consecutive slots, read from one
// base address with immediate offsets; the 16 rows of a down-8 sub-tile read
// slots 2 apart, on distinct banks.
#include <algorithm>

#include "fdsp_internal.h"

namespace fdsp {

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef float fvec2 __attribute__((ext_vector_type(2)));
typedef float fvec4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) fvec2 *gvec2_ptr;
typedef __attribute__((address_space(1))) fvec4 *gvec4_ptr;
// records read at wavefront-uniform addresses: scalar (SMEM) loads
template <class T> using cptr = const __attribute__((address_space(4))) T *;
template <class T> __device__ __forceinline__ cptr<T> as_const(const T *p) { return (cptr<T>)p; }

constexpr uint32_t kBlock = 256;
constexpr uint32_t kWaves = kBlock / 64;
constexpr int kBufferFlags = 0x00020000;  // gfx9 buffer descriptor word 3: 32-bit data format

// segment lengths: multiples of kSegSplit inputs from 1024 to kSegMax
constexpr uint32_t kSegMax = 4096, kSegSplit = 256;
// rows of the transposed span (odd) for a segment length and halo
__host__ __device__ constexpr uint32_t seg_rows(uint32_t seg, uint32_t halo) { return ((seg + halo) / 4 + 1) | 1u; }

__device__ __forceinline__ uint32_t uniform(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint32_t slot(uint32_t q, uint32_t km) { return (q & 3u) * km + (q >> 2); }

// floor(a / d) for a < 2^52 from an FP64 quotient estimate and one exact
// correction step (inv_d = 1/d rounded to double)
__device__ __forceinline__ uint64_t udiv64(uint64_t a, uint32_t d, double inv_d)
{
    int64_t q = (int64_t)((double)a * inv_d);
    const int64_t r = (int64_t)a - q * (int64_t)d;
    if (r < 0)
        --q;
    else if (r >= (int64_t)d)
        ++q;
    return (uint64_t)q;
}

// Stream bookkeeping of one channel for this chunk (global input position
// P), the same for every block but j_lo / j_hi: the chunk's outputs start at
// stream output n0 = ceil(P*up/down); its output j has chunk-local anchor
// floor((j*down + rho)/up) with rho = n0*down - up*P in [0, down); outputs
// j_lo .. j_hi-1 have their anchors in this block's segment, the first of
// them at chunk-local input a_lo, whose NCO index is k_lo = (P + a_lo) mod R;
// the chunk emits j_end outputs in all.
struct ChanRun {
    uint32_t rho;
    uint32_t j_lo, j_hi;
    uint32_t j_end;
    uint32_t a_lo, k_lo;
    uint32_t pad[2];
};

// the ChanStatic fields chan_run reads
struct RunParams {
    double inv_down, inv_r, inv_up;
    uint32_t up, down, r;
};

__device__ __forceinline__ RunParams run_params(const ChanStatic &c)
{
    return RunParams{c.inv_down, c.inv_r, c.inv_up, c.up, c.down, c.r};
}

// a mod m for a below about 2m (one subtraction), else the FP64 form
__device__ __forceinline__ uint32_t mod_near(uint64_t a, uint32_t m, double inv_m)
{
    if (a < m)
        return (uint32_t)a;
    if (a - m < m)
        return (uint32_t)(a - m);
    return (uint32_t)mod_u64(a, m, inv_m);
}

__device__ __forceinline__ ChanRun chan_run(const RunParams &c, uint64_t P, uint32_t lo, uint32_t hi, uint32_t n)
{
    const uint64_t up = c.up, down = c.down;
    const uint64_t uP = up * P;
    const uint64_t n0 = udiv64(uP + down - 1, c.down, c.inv_down);
    ChanRun r;
    r.rho = (uint32_t)(n0 * down - uP);
    const auto first_at = [&](uint32_t x) -> uint32_t {
        const uint64_t ux = up * x;
        return ux > r.rho ? (uint32_t)udiv64(ux - r.rho + down - 1, c.down, c.inv_down) : 0u;
    };
    r.j_lo = first_at(lo);
    r.j_hi = first_at(hi);
    r.j_end = first_at(n);
    const uint64_t num = (uint64_t)r.j_lo * down + r.rho;
    r.a_lo = (uint32_t)(c.up == 1 ? num : udiv64(num, c.up, c.inv_up));
    r.k_lo = (uint32_t)mod_u64(P + r.a_lo, c.r, c.inv_r);
    return r;
}

// The lane's channel, for the epilogue of its outputs
struct LaneChan {
    double nu, inv_r;
    uint32_t r, win, poff;
    float2 wrap_m1;  // e^{j2pi nu R} - 1
    gvec2_ptr out;
};

__device__ __forceinline__ LaneChan lane_chan(const ChanStatic &c)
{
    LaneChan l;
    l.nu = c.nu;
    l.inv_r = c.inv_r;
    l.r = c.r;
    l.win = c.win;
    l.poff = c.poff;
    l.wrap_m1 = make_float2(c.wrap.x - 1.f, c.wrap.y);
    l.out = (gvec2_ptr)c.out;
    return l;
}

struct SegDev {
    const float2 *in;
    uint32_t n;
    uint32_t halo;      // H: staged samples before the segment (>= every window - 1, multiple of 16)
    uint32_t hist_len;  // samples in each history buffer (>= H)
    const float2 *hist0, *hist1;
    float2 *histw0, *histw1;
    const ChanStatic *chans;
    const float2 *ptaps;  // per channel: up phases x win padded modulated taps, from ChanStatic::poff
    const float *btab;    // MFMA B operands, 16 steps x 64 lanes per group
    const SegGroup *groups;
    uint32_t n_groups;
    uint32_t n_chan;
    const ChanPos *pos;  // device position (graph mode), or null: p and parity below
    uint64_t p;
    uint32_t parity;
    uint32_t out16;      // every channel's output is 16-byte aligned
    uint32_t seg;        // segment length (a multiple of kSegSplit, at most kSegMax)
    uint32_t km;         // rows of the transposed span, seg_rows(seg, halo)
    uint32_t n_big;      // workgroups 0 .. n_big-1 take segments of length seg,
    uint32_t small;      // later ones (the chunk's end) segments of this length
};

// the chunk's position, from the device record in graph mode
struct SegPos {
    uint64_t P;
    uint32_t parity;
};

// y (filtered, not yet phased) -> the output of anchor NCO index k: rollover
// correction when the window straddles a restart (window positions t <
// win-1-k precede it; taps of phase ph; q: the window's first span sample),
// then the phasor.
__device__ __forceinline__ float2 finish_out(const SegDev &a, const LaneChan &c, uint32_t ph, const float2 *xs,
                                             uint32_t q, uint32_t k, float2 y)
{
    if (k + 1 < c.win) {
        const float2 *g = a.ptaps + c.poff + ph * c.win;
        float2 part = make_float2(0.f, 0.f);
        for (uint32_t t = 0; t + 1 + k < c.win; ++t)
            cmac(part, g[t], xs[slot(q + t, a.km)]);
        cmac(y, c.wrap_m1, part);
    }
    return cmul(nco_phasor((int64_t)k, c.nu), y);
}

// ---------------------------------------------------------------------------
// MFMA group: this wave's sub-tiles st, st + kWaves, ... < n_st (sub-tile s:
// segment rows 16 s .. 16 s + 15 of every channel of the group)
// ---------------------------------------------------------------------------
// One sub-tile's A operand (eight consecutive slots from p) through one MFMA
// chain: the K steps of the real and imaginary parts alternate (dependent
// MFMAs issue back to back).
__device__ __forceinline__ floatx4 mfma_tile(const float2 *p, const float (&bv)[16])
{
    float2 x[8];
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j)
        x[j] = p[j];
    floatx4 c = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (uint32_t j = 0; j < 8; ++j) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].x, bv[2 * j], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(x[j].y, bv[2 * j + 1], c, 0, 0, 0);
    }
    return c;
}

__device__ __forceinline__ void mfma_group(const SegDev &a, cptr<SegGroup> gr, const ChanRun &run,
                                           const ChanRun *runs, const float2 *xs, uint32_t span_lo, uint32_t st,
                                           uint32_t n_st, uint32_t lane)
{
    const uint32_t down = gr->down, m = gr->m;
    const uint32_t n_rows = run.j_hi - run.j_lo;
    // opaque per call: keeps the group's loads inside the persistent segment
    // loop instead of hoisted (and every group's operands held) across it
    asm volatile("" : "+v"(lane));
    float bv[16];
    {
        // one lane pointer, the steps at immediate offsets
        const float *bt = a.btab + gr->btab + lane;
#pragma unroll
        for (uint32_t s = 0; s < 16; ++s)
            bv[s] = bt[s * 64];
    }
    // D[i][col] sits in lane col + 16*(i/4), element i%4: lanes 2c and 2c+1
    // hold channel c's real and imaginary parts of rows 4*(lane/16) .. +3.
    // The even lane finishes the first two of those rows, the odd lane the
    // last two (quad_perm 1,0,3,2 swaps the pair's values). Lanes past the
    // group's columns run with its last channel and store nothing.
    const uint32_t col = lane & 15, g = lane >> 4;
    const bool odd = col & 1, live = (col >> 1) < m;
    const uint32_t ci = gr->chan[min(col >> 1, m - 1)];
    // the constants of the quiet steps stay in registers; the rare paths
    // (index wraps, rollover corrections) read the others from the record
    const ChanStatic *csp = a.chans + ci;
    const uint32_t R = csp->r, dk = csp->dk, d1 = csp->d1;
    const float2 e_dk = csp->e_dk, e_d1 = csp->e_d1;
    const gvec2_ptr out = (gvec2_ptr)csp->out;
    const auto rec = [&]() {
        const ChanStatic *p = csp;
        asm volatile("" : "+v"(p));  // read at the use, not held across the loop
        return p;
    };
    const uint32_t ro = 4 * g + (odd ? 2u : 0u);  // sub-tile row of the lane's first output
    // NCO index k of the lane's first row, advanced by dk = (16 kWaves down)
    // mod R per step, and its phasor ph = e^{j2pi nu k}, advanced by e^{j2pi
    // nu dk}; the second row is d1 = down mod R later. Where an index wraps
    // past R the phase jumps (the reference's rollover): that lane's phasor
    // is evaluated afresh in FP64. The lane tracks kw = k - (win - 1) (signed;
    // MFMA groups have 32-tap windows): its rows need no rollover correction
    // and its second row does not wrap while 0 <= kw < quiet = R - d1 - 31.
    constexpr int32_t w1 = 31;
    int32_t kw =
        (int32_t)mod_near((uint64_t)runs[ci].k_lo + (uint64_t)(16 * st + ro) * down, R, csp->inv_r) - w1;
    const int32_t r_w = (int32_t)R - w1;  // kw at which k wraps
    const uint32_t quiet = (uint32_t)max((int32_t)(R - d1) - w1, 0);
    float2 ph = nco_phasor((int64_t)(kw + w1), csp->nu);
    // A operand: row lane%16 reads window samples g + 4j, eight consecutive
    // slots; rows of a whole sub-tile are unclamped, so the next one's base is
    // (16 kWaves down)/4 slots on. A partial sub-tile clamps its rows to the
    // segment's last.
    const uint32_t i = lane & 15;
    const auto a_base = [&](uint32_t s) -> uint32_t {
        const uint32_t row = min(16 * s + i, n_rows - 1);
        return slot((run.j_lo + row) * down + run.rho - 31 - span_lo + g, a.km);
    };
    const bool pair16 = a.out16 && (run.j_lo & 1) == 0;
    // the lane's two outputs of sub-tile s (rows ro, ro + 1), not yet phased
    const auto rows_of = [&](const floatx4 &acc, float2 &y0, float2 &y1) {
        float oth[4];
#pragma unroll
        for (uint32_t e = 0; e < 4; ++e)
            oth[e] = __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(acc[e]), 0xB1, 0xF, 0xF, false));
        y0 = odd ? make_float2(oth[2], acc[2]) : make_float2(acc[0], oth[0]);
        y1 = odd ? make_float2(oth[3], acc[3]) : make_float2(acc[1], oth[1]);
    };
    // outputs are streamed (non-temporal): nothing in this launch reads them
    const auto store = [&](uint32_t rr, float2 o0, float2 o1, bool two) {
        const uint32_t j = run.j_lo + rr;
        if (two) {
            if (pair16)
                __builtin_nontemporal_store(fvec4{o0.x, o0.y, o1.x, o1.y}, (gvec4_ptr)(out + j));
            else {
                __builtin_nontemporal_store(fvec2{o0.x, o0.y}, out + j);
                __builtin_nontemporal_store(fvec2{o1.x, o1.y}, out + j + 1);
            }
        } else {
            __builtin_nontemporal_store(fvec2{o0.x, o0.y}, out + j);
        }
    };
    // epilogue of sub-tile s. Quiet (every row of the sub-tile exists and no
    // lane is near a wrap): phase and store. Otherwise the general form:
    // partial rows, a second row that wraps, rollover corrections.
    const auto finish = [&](uint32_t s, const floatx4 &acc) {
        float2 y0, y1;
        rows_of(acc, y0, y1);
        const uint32_t rr = 16 * s + ro;  // segment row of y0
        const bool quiet_here = (uint32_t)kw < quiet;
        if (16 * (s + 1) <= n_rows && __builtin_amdgcn_ballot_w64(!quiet_here) == 0) {
            const float2 ph1 = cmul(ph, e_d1);
            if (live)
                store(rr, cmul(ph, y0), cmul(ph1, y1), true);
            return;
        }
        const uint32_t k = (uint32_t)(kw + w1);
        uint32_t k1 = k + d1;
        float2 ph1 = cmul(ph, e_d1);
        if (k1 >= R) {  // rare: the second row's index wraps
            k1 -= R;
            ph1 = nco_phasor((int64_t)k1, rec()->nu);
        }
        const bool roll0 = rr < n_rows && k < (uint32_t)w1, roll1 = rr + 1 < n_rows && k1 < (uint32_t)w1;
        if (roll0 || roll1) {
            // rare: a window of a row that exists straddles a rollover; its
            // samples before it carry e^{j2pi nu R}
            const ChanStatic *c = rec();
            const uint32_t q0 = (run.j_lo + rr) * down + run.rho - 31 - span_lo;
            const float2 *gt = a.ptaps + c->poff;
            const float2 wrap_m1 = make_float2(c->wrap.x - 1.f, c->wrap.y);
            float2 p0 = make_float2(0.f, 0.f), p1 = make_float2(0.f, 0.f);
            for (uint32_t t = 0; roll0 && t + k < (uint32_t)w1; ++t)
                cmac(p0, gt[t], xs[slot(q0 + t, a.km)]);
            for (uint32_t t = 0; roll1 && t + k1 < (uint32_t)w1; ++t)
                cmac(p1, gt[t], xs[slot(q0 + down + t, a.km)]);
            cmac(y0, wrap_m1, p0);
            cmac(y1, wrap_m1, p1);
        }
        if (live && rr < n_rows)
            store(rr, cmul(ph, y0), cmul(ph1, y1), rr + 1 < n_rows);
    };
    // software pipeline: the next sub-tile's LDS reads and MFMAs are issued
    // before this one's epilogue, so no wave waits on its own products
    uint32_t ab = a_base(st);
    floatx4 c = mfma_tile(xs + ab, bv);
    for (uint32_t sn = st + kWaves; sn < n_st; sn += kWaves) {
        ab = 16 * (sn + 1) <= n_rows ? ab + 4 * kWaves * down : a_base(sn);
        const floatx4 n = mfma_tile(xs + ab, bv);
        finish(st, c);
        c = n;
        st = sn;
        kw += (int32_t)dk;
        ph = cmul(ph, e_dk);
        if (kw >= r_w) {  // rare: the index wrapped
            kw -= (int32_t)R;
            ph = nco_phasor((int64_t)(kw + w1), rec()->nu);
        }
    }
    finish(st, c);
}

// ---------------------------------------------------------------------------
// VALU channel: batch b of one channel's outputs in the segment
// ---------------------------------------------------------------------------
// Without a tail, lane l computes segment row 64 b + l. With a discriminator
// tail (reference src/transform/falcon_dsp_polar_discriminator.cc:134-171),
// batch b covers rows 63 b - 1 .. 63 b + 62: lane 0 recomputes the row before
// the batch (the segment's row -1 sits in the halo, or, at the chunk's first
// output, is the carried last sample of the previous chunk), so every other
// lane takes y[j-1] from its neighbour and emits angle(y[j] conj(y[j-1])).
// The stream's first output has no predecessor and emits nothing; the
// chunk's last output is carried to the next chunk.
// The window's samples of residue u (window positions u, u+4, ...) are
// consecutive slots from base u.
template <uint32_t kW>
__device__ __forceinline__ void valu_group(const SegDev &a, uint32_t ci, bool tail, const ChanRun &run,
                                           const SegPos &pos, const float2 *xs, uint32_t span_lo, uint32_t b,
                                           uint32_t n_b, uint32_t lane)
{
    // opaque per call: keeps the channel's loads inside the segment loop
    asm volatile("" : "+s"(ci));
    const cptr<ChanStatic> cs = as_const(a.chans) + ci;
    const uint32_t up = cs->up, down = cs->down;
    const int32_t n_rows = (int32_t)(run.j_hi - run.j_lo);
    const LaneChan lc = lane_chan(a.chans[ci]);
    // NCO index of the segment's first output's anchor (output 0's anchor is
    // floor(rho/up)); later anchors of the segment are fewer than R inputs on
    // (the host keeps channels with ceil(down/up) + segment >= R off this path)
    const uint32_t a_lo = run.a_lo, k_lo = run.k_lo;
    for (; b < n_b; b += kWaves) {
        const int32_t rr = tail ? (int32_t)(b * 63 + lane) - 1 : (int32_t)(b * 64 + lane);
        // the row computed: rows past the segment's last take the last; row -1
        // of the chunk's first segment (no window: the carried sample replaces
        // it) takes row 0
        const int32_t rc = min(rr, n_rows - 1);
        const uint32_t j = (uint32_t)((int32_t)run.j_lo + (rc < 0 && run.j_lo == 0 ? 0 : rc));
        uint32_t anchor, ph;
        if (up == 1) {
            anchor = j * down + run.rho;
            ph = 0;
        } else {
            const uint64_t num = (uint64_t)j * down + run.rho;
            anchor = (uint32_t)udiv64(num, up, cs->inv_up);
            ph = (uint32_t)(num - (uint64_t)anchor * up);
        }
        const uint32_t q = anchor + 1 - kW - span_lo;
        const float2 *base[4];
#pragma unroll
        for (uint32_t u = 0; u < 4; ++u)
            base[u] = xs + slot(q + u, a.km);
        // four independent chains: even and odd taps, real and imaginary;
        // eight taps (two per residue) per step
        float2 acc0 = make_float2(0.f, 0.f), acc1 = make_float2(0.f, 0.f);
        if (up == 1) {
            const cptr<float> g = (cptr<float>)(a.ptaps + cs->poff);
#pragma unroll 1
            for (uint32_t m = 0; m < kW / 4; m += 2) {
#pragma unroll
                for (uint32_t e = 0; e < 8; e += 2) {
                    const uint32_t t = 4 * m + e;  // taps t, t+1 of residues e%4, e%4 + 1
                    cmac(acc0, make_float2(g[2 * t], g[2 * t + 1]), base[e & 3][m + (e >> 2)]);
                    cmac(acc1, make_float2(g[2 * t + 2], g[2 * t + 3]), base[(e + 1) & 3][m + (e >> 2)]);
                }
            }
        } else {
            // per-lane taps (the lane's filter phase): four per step, so the
            // step's operands stay within the registers of the up 1 path
            const float2 *g = a.ptaps + cs->poff + ph * kW;
#pragma unroll 1
            for (uint32_t m = 0; m < kW / 4; ++m) {
#pragma unroll
                for (uint32_t e = 0; e < 4; e += 2) {
                    const uint32_t t = 4 * m + e;
                    cmac(acc0, g[t], base[e][m]);
                    cmac(acc1, g[t + 1], base[e + 1][m]);
                }
            }
        }
        const float2 acc = make_float2(acc0.x + acc1.x, acc0.y + acc1.y);
        // NCO index of the anchor (the overlap row before the segment is at
        // most ceil(down/up) inputs earlier)
        uint32_t k = k_lo + (anchor - a_lo);
        if ((int32_t)(anchor - a_lo) < 0)
            k += lc.r;
        if (k >= lc.r)
            k -= lc.r;
        float2 y = finish_out(a, lc, ph, xs, q, k, acc);
        if (!tail) {
            if (rr < n_rows)
                __builtin_nontemporal_store(fvec2{y.x, y.y}, lc.out + j);
            continue;
        }
        // discriminator tail (see above)
        const bool first_row = rr < 0 && run.j_lo == 0;  // the previous chunk's last output
        if (first_row && pos.P > 0)
            y = cs->prev[pos.parity][0];
        const float2 yp = make_float2(__shfl_up(y.x, 1), __shfl_up(y.y, 1));
        // lane 0 only supplies its row (the previous batch's last) to lane 1
        if (lane == 0 || rr >= n_rows)
            continue;
        // the chunk's last output is the next chunk's predecessor
        if (j + 1 == run.j_end)
            cs->prev[pos.parity ^ 1u][0] = y;
        // output j of the chunk has a predecessor unless it is the stream's first
        const uint32_t drop = pos.P == 0 ? 1u : 0u;
        if (j < drop)
            continue;
        const float2 z = cmul(y, make_float2(yp.x, -yp.y));
        __builtin_nontemporal_store(atan2f(z.y, z.x), (__attribute__((address_space(1))) float *)lc.out + (j - drop));
    }
}

// ---------------------------------------------------------------------------
// the kernel: one segment per workgroup
// ---------------------------------------------------------------------------
// kTails: the plan has discriminator tails. Their code (the angle, the
// neighbour exchange, the carried sample) sets the kernel's register peak, so
// plans without them run an instance that leaves it out (fewer VGPRs, more
// workgroups per CU).
template <bool kTails>
__global__ __launch_bounds__(kBlock, 5) void chan_seg_kernel(SegDev a)
{
    const uint32_t km = a.km;
    // dynamic LDS: the transposed span (4 km samples), then one ChanRun per channel
    extern __shared__ float2 xs[];
    ChanRun *runs = reinterpret_cast<ChanRun *>(xs + 4 * km);
    __shared__ uint32_t item0[kSegMaxGroups + 1];

    SegPos pos{a.p, a.parity};
    if (a.pos) {
        pos.P = a.pos->p;
        pos.parity = a.pos->parity;
    }
    const uint64_t P = pos.P;
    const uint32_t parity = pos.parity;
    const float2 *hist = parity ? a.hist1 : a.hist0;
    // chunk-local first input and nominal length of the segment
    const uint32_t seg_lo =
        blockIdx.x < a.n_big ? blockIdx.x * a.seg : a.n_big * a.seg + (blockIdx.x - a.n_big) * a.small;
    const uint32_t seg_len = blockIdx.x < a.n_big ? a.seg : a.small;
    const uint32_t seg_hi = min(a.n, seg_lo + seg_len);
    const uint32_t H = a.halo;
    // chunk-local index of span sample 0 (wraps below zero in the first
    // segment; window starts q = anchor - (w-1) - span_lo stay >= 0 in
    // modular arithmetic since H >= w-1)
    const uint32_t span_lo = seg_lo - H;
    const uint32_t tid = threadIdx.x, lane = tid & 63, wave = uniform(tid >> 6);

    if (wave == 0) {
        // 1. bookkeeping, wave 0 (while waves 1.. stage the span): one lane
        // per channel, then the groups' work item counts (from the channels'
        // runs this wave has just written) scanned across the wave. Lanes
        // past the last channel / group repeat it.
        const RunParams rp = run_params(a.chans[min(lane, a.n_chan - 1)]);
        uint32_t g_chan, g_kind, g_tail;
        {
            const SegGroup &gr = a.groups[min(lane, a.n_groups - 1)];
            g_chan = gr.chan[0];
            g_kind = gr.kind;
            g_tail = gr.tail;
        }
        if (lane < a.n_chan)
            runs[lane] = chan_run(rp, P, seg_lo, seg_hi, a.n);
        for (uint32_t c = lane + 64; c < a.n_chan; c += 64)
            runs[c] = chan_run(run_params(a.chans[c]), P, seg_lo, seg_hi, a.n);
        // the lanes' LDS writes above precede the reads below (one wave)
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        uint32_t cnt = 0;
        if (lane < a.n_groups) {
            const ChanRun &r = runs[g_chan];
            const uint32_t rows = r.j_hi - r.j_lo;
            cnt = g_kind == kSegMfma ? (rows + 15) / 16 : g_tail ? (rows + 62) / 63 : (rows + 63) / 64;
        }
#pragma unroll
        for (uint32_t d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(cnt, d);
            if (lane >= d)
                cnt += o;
        }
        if (lane < a.n_groups)
            item0[lane + 1] = cnt;
        if (lane == 0)
            item0[0] = 0;
    } else {
        // 2. stage the span (chunk-local inputs seg_lo - H .. seg_hi - 1, the
        // part before the chunk from the history), waves 1..: two samples per
        // thread and load, into LDS transposed by 4
        constexpr uint32_t kStage = kBlock - 64;
        const uint32_t st = tid - 64;
        const int64_t base = (int64_t)seg_lo - (int64_t)H;
        const uint32_t pairs = (seg_hi - seg_lo + H + 1) / 2;
        const bool in16 = (reinterpret_cast<uintptr_t>(a.in) & 15) == 0;
        constexpr uint32_t kRounds = ((kSegMax + kSegMaxHalo) / 2 + kStage - 1) / kStage;
        float4 v[kRounds];
        if (base >= 0 && seg_lo + seg_len <= a.n && in16) {
            // interior segment: every pair lies in the chunk, 16-byte aligned
            // (base and the input are even / 16-byte aligned). Buffer loads:
            // 32-bit offsets on the segment's base, zero past the span.
            const __amdgpu_buffer_rsrc_t rs = __builtin_amdgcn_make_buffer_rsrc(
                const_cast<float2 *>(a.in + base), 0, (int)(pairs * 16u), kBufferFlags);
#pragma unroll
            for (uint32_t r = 0; r < kRounds; ++r) {
                const auto w = __builtin_amdgcn_raw_buffer_load_b128(rs, (int)((st + r * kStage) * 16u), 0, 0);
                v[r] = make_float4(__uint_as_float(w[0]), __uint_as_float(w[1]), __uint_as_float(w[2]),
                                   __uint_as_float(w[3]));
            }
        } else {
#pragma unroll
            for (uint32_t r = 0; r < kRounds; ++r) {
                const uint32_t pr = st + r * kStage;
                const int64_t i = base + 2 * (int64_t)pr;
                float4 w = make_float4(0.f, 0.f, 0.f, 0.f);
                if (pr < pairs) {
                    if (i >= 0) {
                        if (in16 && i + 1 < (int64_t)a.n) {
                            w = *reinterpret_cast<const float4 *>(a.in + i);
                        } else {
                            if (i < (int64_t)a.n) {
                                const float2 u = a.in[i];
                                w.x = u.x;
                                w.y = u.y;
                            }
                            if (i + 1 < (int64_t)a.n) {
                                const float2 u = a.in[i + 1];
                                w.z = u.x;
                                w.w = u.y;
                            }
                        }
                    } else {
                        // i even and i + 1 <= -1: both samples from the history
                        const float2 u0 = hist[a.hist_len + i], u1 = hist[a.hist_len + i + 1];
                        w = make_float4(u0.x, u0.y, u1.x, u1.y);
                    }
                }
                v[r] = w;
            }
        }
        // pair pr = st + r kStage holds samples q = 2 pr and q + 1, at slots
        // (q mod 4) km + q/4 = 2 (st mod 2) km + st/2 + r kStage/2 and that
        // plus km: two bases per thread, constant offsets per round
        const uint32_t s0 = 2u * (st & 1u) * km + (st >> 1), s1 = s0 + km;
        static_assert(kStage % 2 == 0, "pairs of one thread keep their residue");
#pragma unroll
        for (uint32_t r = 0; r < kRounds; ++r) {
            if (st + r * kStage < pairs) {
                xs[s0 + r * (kStage / 2)] = make_float2(v[r].x, v[r].y);
                xs[s1 + r * (kStage / 2)] = make_float2(v[r].z, v[r].w);
            }
        }
    }
    __syncthreads();

    // 3. work items in group order, item t to wave (t + block) mod kWaves: the
    // rotation spreads the segment's leftover items over the SIMDs (wave w of
    // every block runs on the same SIMD)
    const cptr<SegGroup> groups = as_const(a.groups);
    const uint32_t rot = blockIdx.x % kWaves;
    for (uint32_t g = 0; g < a.n_groups; ++g) {
        const uint32_t i0 = item0[g], i1 = item0[g + 1];
        uint32_t i = i0 + (wave + 2 * kWaves - (i0 + rot) % kWaves) % kWaves;
        if (i >= i1)
            continue;
        const cptr<SegGroup> gr = groups + g;
        const ChanRun run = runs[gr->chan[0]];
        if (gr->kind == kSegMfma) {
            mfma_group(a, gr, run, runs, xs, span_lo, i - i0, i1 - i0, lane);
        } else {
            const uint32_t ci = gr->chan[0];
            const bool tail = kTails && gr->tail != 0;
            if (gr->win == 32)
                valu_group<32>(a, ci, tail, run, pos, xs, span_lo, i - i0, i1 - i0, lane);
            else
                valu_group<64>(a, ci, tail, run, pos, xs, span_lo, i - i0, i1 - i0, lane);
        }
    }

    // 4. the block holding the chunk's end writes the next history (the
    // other buffer: every block of this launch reads `hist`), and carries the
    // discriminator's last sample of every tail channel with no output in
    // this chunk
    if (seg_hi == a.n) {
        for (uint32_t g = tid; g < a.n_groups; g += kBlock) {
            const SegGroup &gr = a.groups[g];
            if (gr.tail && runs[gr.chan[0]].j_end == 0) {
                const ChanStatic &c = a.chans[gr.chan[0]];
                c.prev[parity ^ 1u][0] = c.prev[parity][0];
            }
        }
        float2 *hw = parity ? a.histw0 : a.histw1;
        for (uint32_t t = tid; t < a.hist_len; t += kBlock) {
            const int64_t i = (int64_t)a.n - (int64_t)a.hist_len + t;
            hw[t] = i >= 0 ? a.in[i] : hist[a.hist_len + i];
        }
    }
}

// The chunk's last 1/kTailDiv or so runs in segments kTailSplit times
// shorter: the workgroups dispatched last are short, so the launch drains
// sooner (every workgroup of the full length has to end before it does).
// 16 and 2 measured best on the 32-channel workload (1/8 and 1/32, and
// quarter-length segments, were slower).
constexpr uint32_t kTailDiv = 16, kTailSplit = 2;

static_assert(kSegSplit % (kTailSplit * 16) == 0, "short segments stay multiples of 16 inputs (16-byte aligned spans)");

}  // namespace

bool chan_seg_len_ok(uint32_t seg) { return seg >= 1024 && seg <= kSegMax && seg % kSegSplit == 0; }

SegGeom chan_seg_geometry(uint32_t n, uint32_t seg, uint32_t halo, uint32_t n_chan)
{
    SegGeom g{};
    // a chunk too short to give the GPU about kFillBlocks workgroups at the
    // set length runs shorter segments (never below 1024): a 2^18-sample
    // chunk of the 32-channel workload took 14.0 us at 1024 against 16.6 at 2560
    constexpr uint32_t kFillBlocks = 1024;
    g.seg = seg;
    if (n / seg < kFillBlocks)
        g.seg = std::max(1024u, std::min(seg, n / kFillBlocks / kSegSplit * kSegSplit));
    const uint32_t whole = n / g.seg, tail = (n / kTailDiv + g.seg - 1) / g.seg;
    g.n_big = whole > tail ? whole - tail : 0;
    g.small = g.seg / kTailSplit;
    g.blocks = g.n_big + (n - g.n_big * g.seg + g.small - 1) / g.small;
    g.lds = chan_seg_lds_bytes(g.seg, halo, n_chan);
    return g;
}

size_t chan_seg_lds_bytes(uint32_t seg, uint32_t halo, uint32_t n_chan)
{
    return (size_t)4 * seg_rows(seg, halo) * sizeof(float2) + (size_t)n_chan * sizeof(ChanRun);
}

hipError_t launch_chan_seg(const SegArgs &s, hipStream_t stream)
{
    if (s.n == 0)
        return hipSuccess;
    SegDev a{};
    a.in = s.in;
    a.n = s.n;
    a.halo = s.halo;
    a.hist_len = s.hist_len;
    a.hist0 = s.hist[0];
    a.hist1 = s.hist[1];
    a.histw0 = s.hist_w[0];
    a.histw1 = s.hist_w[1];
    a.chans = s.chans;
    a.ptaps = s.ptaps;
    a.btab = s.btab;
    a.groups = s.groups;
    a.n_groups = s.n_groups;
    a.n_chan = s.n_chan;
    a.pos = s.pos;
    a.p = s.p;
    a.parity = s.parity;
    a.out16 = s.out16 ? 1u : 0u;
    if (!chan_seg_len_ok(s.seg) || s.halo > kSegMaxHalo)
        return hipErrorInvalidValue;
    const SegGeom geo = chan_seg_geometry(s.n, s.seg, s.halo, s.n_chan);
    a.seg = geo.seg;
    a.km = seg_rows(geo.seg, s.halo);
    a.n_big = geo.n_big;
    a.small = geo.small;
    const uint32_t blocks = geo.blocks;
    const size_t lds = geo.lds;
    if (s.tails)
        hipLaunchKernelGGL(chan_seg_kernel<true>, dim3(blocks), dim3(kBlock), lds, stream, a);
    else
        hipLaunchKernelGGL(chan_seg_kernel<false>, dim3(blocks), dim3(kBlock), lds, stream, a);
    return hipGetLastError();
}

}  // namespace fdsp
