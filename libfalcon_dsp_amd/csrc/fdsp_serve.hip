// fdsp_serve.hip -- serving loop for channelizer streams: admission by a
// device-memory budget, continuous batching of chunks from many streams into
// one launch, device-resident stream state.
//
// A stream is one channel of the multi-rate channelizer (reference
// falcon_dsp_multi_rate_channelizer.cc:146-164: shift, then polyphase
// resample), optionally followed by the polar discriminator
// (falcon_dsp_polar_discriminator.cc:134-171). Its state between chunks is
// the reference's (NCO index, coeff_phase, xOffset, filter history), kept in
// closed form: inputs consumed p plus the last hist_len inputs in device
// memory, so a chunk's work depends only on (p, history, chunk).
//
// Runtime: clients submit chunks from any thread; the scheduler
// (fdsp_serve_sched.h, free of HIP) queues them and one worker thread forms
// batches, packs each into pinned staging and issues it: the copy-in on one
// HIP stream, then (ordered by an event) the kernels and the copy-out on a
// second, so batch k+1's copy-in overlaps batch k's kernels. Two staging
// slots alternate: while batch k runs, batch k+1 is packed and issued, then
// batch k's results are delivered. Stream state lives in slices of one
// device pool allocated at create, so admission allocates nothing on the
// device.
//
// Kernels: one block per (chunk, 256-output tile); thread j of the tile
// computes output j as y = e^{j2pi nu k(A)} sum_t g_t x[A - cpp + 1 + t]
// with the shift folded into modulated taps g (the channelizer's
// formulation; the window of an output that straddles an NCO rollover is
// recomputed with the rollover's phase jump). Serving is bound by host
// copies and PCIe, not by this kernel, so windows are read straight from
// the batch buffer through the caches rather than staged in LDS.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>
#include <vector>

#include "fdsp_host.h"
#include "fdsp_internal.h"
#include "fdsp_serve_sched.h"

using namespace fdsp;

namespace {

constexpr uint32_t kServeBlock = 256;  // outputs per tile
constexpr size_t kAlign = 256;
constexpr size_t kMaxQueued = 1u << 16;  // requests queued or running at once

size_t align_up(size_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// one chunk of a batch (device)
struct ServeItem {
    const float2 *in;      // the chunk's inputs in the batch buffer
    const float2 *hist;    // the stream's last hist_len inputs before the chunk
    float2 *hist_next;     // receives the last hist_len inputs after it
    const float2 *taps;    // modulated taps [up][cpp]
    float2 *out;           // complex outputs (discriminator input for tails)
    float *angles;         // tails: discriminator outputs; else null
    const float2 *prev;    // tails: last complex output of the previous chunk
    float2 *prev_next;     // tails: receives this chunk's last complex output
    double nu, inv_r, inv_up;
    uint64_t p;            // stream inputs consumed before the chunk
    uint64_t n0;           // stream index of the chunk's first output: ceil(p*up/down)
    float2 wrap;           // e^{j2pi nu R}
    uint32_t n, count, r, up, down, cpp, hist_len, has_prev;
};

__device__ __forceinline__ float2 serve_sample(const ServeItem &a, int64_t i)
{
    const float2 *p = i >= 0 ? a.in + i : a.hist + (a.hist_len + i);
    return *p;
}

// table: (item, tile) per block
__global__ __launch_bounds__(kServeBlock) void serve_chan_kernel(const ServeItem *__restrict__ items,
                                                                  const uint32_t *__restrict__ table)
{
    const uint32_t it = __builtin_amdgcn_readfirstlane(table[2 * blockIdx.x]);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(table[2 * blockIdx.x + 1]);
    const ServeItem &a = items[it];
    if (tile == 0) {
        // next chunk's history: the last hist_len samples of [history | chunk]
        for (uint32_t t = threadIdx.x; t < a.hist_len; t += kServeBlock) {
            const int64_t i = (int64_t)a.n - a.hist_len + t;
            a.hist_next[t] = serve_sample(a, i);
        }
    }
    const uint32_t j = tile * kServeBlock + threadIdx.x;
    if (j >= a.count)
        return;
    // anchor A (stream index of the window's newest sample) and filter phase
    const uint64_t num = (a.n0 + j) * (uint64_t)a.down;
    uint64_t A = num;
    uint32_t ph = 0;
    if (a.up != 1) {
        int64_t q = (int64_t)((double)num * a.inv_up);
        const int64_t r = (int64_t)num - q * (int64_t)a.up;
        if (r < 0)
            --q;
        else if (r >= (int64_t)a.up)
            ++q;
        A = (uint64_t)q;
        ph = (uint32_t)(num - A * a.up);
    }
    const int64_t lo = (int64_t)(A - a.p) - (int64_t)(a.cpp - 1);  // chunk-local, >= -hist_len
    const float2 *__restrict__ g = a.taps + (size_t)ph * a.cpp;
    const uint64_t k = mod_u64(A, a.r, a.inv_r);
    float2 acc = make_float2(0.f, 0.f);
    if (k + 1 >= a.cpp) {
        for (uint32_t t = 0; t < a.cpp; ++t)
            cmac(acc, g[t], serve_sample(a, lo + t));
    } else {
        // an NCO rollover inside the window: the samples before it carry an
        // extra e^{j2pi nu R}
        const uint32_t split = a.cpp - 1 - (uint32_t)k;
        float2 acc2 = make_float2(0.f, 0.f);
        for (uint32_t t = 0; t < split; ++t)
            cmac(acc2, g[t], serve_sample(a, lo + t));
        for (uint32_t t = split; t < a.cpp; ++t)
            cmac(acc, g[t], serve_sample(a, lo + t));
        const float2 w = cmul(a.wrap, acc2);
        acc.x += w.x;
        acc.y += w.y;
    }
    a.out[j] = cmul(nco_phasor((int64_t)k, a.nu), acc);
}

// polar discriminator tail: angle of y[i] conj(y[i-1]); the stream's first
// output has no predecessor and emits nothing (reference
// falcon_dsp_polar_discriminator.cc:134-171). table: (item, tile) of the
// tail items with outputs.
__global__ __launch_bounds__(kServeBlock) void serve_discriminator_kernel(const ServeItem *__restrict__ items,
                                                                           const uint32_t *__restrict__ table)
{
    const uint32_t it = __builtin_amdgcn_readfirstlane(table[2 * blockIdx.x]);
    const uint32_t tile = __builtin_amdgcn_readfirstlane(table[2 * blockIdx.x + 1]);
    const ServeItem &a = items[it];
    const uint32_t i = tile * kServeBlock + threadIdx.x;
    if (i >= a.count)
        return;
    if (i == a.count - 1)
        *a.prev_next = a.out[i];
    if (i == 0 && !a.has_prev)
        return;
    const float2 y = a.out[i];
    const float2 b = i == 0 ? *a.prev : a.out[i - 1];
    const float2 z = cmul(y, make_float2(b.x, -b.y));
    a.angles[a.has_prev ? i : i - 1] = atan2f(z.y, z.x);
}

struct ServeStream {
    bool live = false;
    uint32_t up = 1, down = 1, cpp = 1, hist_len = 1, tail = 0;
    double nu = 0, inv_r = 1;
    uint32_t r = 1;
    float2 wrap{1.f, 0.f};
    char *dev = nullptr;  // the stream's pool slice: taps | hist[2] | prev[2]
    size_t dev_off = 0, dev_bytes = 0;
    size_t taps_off = 0, hist_off[2] = {0, 0}, prev_off[2] = {0, 0};
    template <class T> T *at(size_t off) const { return reinterpret_cast<T *>(dev + off); }
};

struct Slot {
    char *h_buf = nullptr;  // pinned: items | table | tail table | inputs
    char *h_out = nullptr;  // pinned: packed outputs
    DevBuf d_buf, d_out, d_scratch;
    hipEvent_t ev_in = nullptr;    // the copy-in has landed
    hipEvent_t ev_done = nullptr;  // kernels and copy-out are done
    std::vector<SchedEntry> entries;
    std::vector<size_t> out_off;  // bytes into the output region, per entry
    uint64_t batch = 0;
};

}  // namespace

struct fdsp_serve_s {
    fdsp_serve_config cfg{};
    int device = 0;
    hipStream_t s_in = nullptr;     // copies in
    hipStream_t s_run = nullptr;    // kernels, then copies out
    hipStream_t s_admin = nullptr;  // stream state uploads at admission
    size_t table_cap = 0;           // tiles per batch at most
    size_t buf_bytes = 0, out_bytes = 0;
    uint64_t fixed_bytes = 0;  // batch buffers
    Slot slot[2];
    std::unique_ptr<ServeScheduler> sched;
    // stream state pool: one allocation, slices handed out first-fit
    std::mutex pool_mu;
    DevBuf d_pool;
    std::map<size_t, size_t> free_blocks;  // offset -> bytes, coalesced
    uint64_t state_used = 0;
    std::vector<ServeStream> streams;  // max_streams entries, never resized
    std::vector<uint32_t> free_ids;
    uint32_t live = 0;
    std::thread worker;
};

namespace {

// the device bytes a stream holds
void stream_layout(ServeStream &st)
{
    size_t off = 0;
    st.taps_off = off;
    off = align_up(off + (size_t)st.up * st.cpp * sizeof(float2));
    for (int i = 0; i < 2; ++i) {
        st.hist_off[i] = off;
        off = align_up(off + (size_t)st.hist_len * sizeof(float2));
    }
    for (int i = 0; i < 2; ++i) {
        st.prev_off[i] = off;
        off = align_up(off + sizeof(float2));
    }
    st.dev_bytes = off;
}

// first fit in the pool; false when no free block holds `bytes`
bool pool_take(fdsp_serve_s *s, size_t bytes, size_t *off)
{
    for (auto it = s->free_blocks.begin(); it != s->free_blocks.end(); ++it) {
        if (it->second < bytes)
            continue;
        *off = it->first;
        const size_t rest = it->second - bytes;
        const size_t at = it->first + bytes;
        s->free_blocks.erase(it);
        if (rest)
            s->free_blocks.emplace(at, rest);
        return true;
    }
    return false;
}

void pool_give(fdsp_serve_s *s, size_t off, size_t bytes)
{
    auto it = s->free_blocks.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != s->free_blocks.end() && it->first + it->second == next->first) {
        it->second += next->second;
        s->free_blocks.erase(next);
    }
    if (it != s->free_blocks.begin()) {
        auto prev = std::prev(it);
        if (prev->first + prev->second == it->first) {
            prev->second += it->second;
            s->free_blocks.erase(it);
        }
    }
}

// Take the next batch into slot `si` and issue its work. *launched is false
// when no stream has a chunk ready.
int form_and_launch(fdsp_serve_s *s, int si, bool *launched)
{
    Slot &sl = s->slot[si];
    *launched = false;
    s->sched->form_batch(&sl.entries, &sl.batch);
    if (sl.entries.empty())
        return FDSP_OK;
    std::vector<ServeItem> items;
    std::vector<uint32_t> table, dtable;
    size_t total_in = 0, total_scratch = 0, total_bytes_out = 0;
    sl.out_off.clear();
    for (const SchedEntry &e : sl.entries) {
        // the stream's slice stays put while it has a chunk in flight
        // (release refuses busy streams)
        const ServeStream &st = s->streams[e.stream];
        sl.out_off.push_back(total_bytes_out);
        total_bytes_out += (e.emit * e.elem + 7) / 8 * 8;
        ServeItem it{};
        it.hist = st.at<float2>(st.hist_off[e.parity]);
        it.hist_next = st.at<float2>(st.hist_off[e.parity ^ 1]);
        it.taps = st.at<float2>(st.taps_off);
        it.nu = st.nu;
        it.inv_r = st.inv_r;
        it.inv_up = 1.0 / (double)st.up;
        it.p = e.p;
        it.n0 = ceil_div(e.p * st.up, st.down);
        it.wrap = st.wrap;
        it.n = (uint32_t)e.n;
        it.count = e.count;
        it.r = st.r;
        it.up = st.up;
        it.down = st.down;
        it.cpp = st.cpp;
        it.hist_len = st.hist_len;
        it.has_prev = e.has_prev ? 1u : 0u;
        if (e.tail) {
            it.prev = st.at<float2>(st.prev_off[e.prev_parity]);
            it.prev_next = st.at<float2>(st.prev_off[e.prev_parity ^ 1]);
            // complex outputs go to scratch at this offset (patched below)
            it.out = reinterpret_cast<float2 *>(total_scratch);
            total_scratch += e.count;
        }
        const uint32_t idx = (uint32_t)items.size();
        items.push_back(it);
        const uint32_t tiles = std::max<uint32_t>(1, (e.count + kServeBlock - 1) / kServeBlock);
        for (uint32_t t = 0; t < tiles; ++t) {
            table.push_back(idx);
            table.push_back(t);
        }
        if (e.tail)
            for (uint32_t t = 0; t * kServeBlock < e.count; ++t) {
                dtable.push_back(idx);
                dtable.push_back(t);
            }
        total_in += e.n;
    }
    // pack: items | table | tail table | inputs, each 256-byte aligned
    const size_t items_off = 0;
    const size_t table_off = align_up(items_off + items.size() * sizeof(ServeItem));
    const size_t dtable_off = align_up(table_off + table.size() * sizeof(uint32_t));
    const size_t in_off = align_up(dtable_off + dtable.size() * sizeof(uint32_t));
    const size_t used = in_off + total_in * sizeof(float2);
    char *dbase = static_cast<char *>(sl.d_buf.p);
    char *obase = static_cast<char *>(sl.d_out.p);
    float2 *scratch = sl.d_scratch.as<float2>();
    size_t in_pos = 0;
    for (size_t i = 0; i < items.size(); ++i) {
        const SchedEntry &e = sl.entries[i];
        ServeItem &it = items[i];
        std::memcpy(sl.h_buf + in_off + in_pos * sizeof(float2), e.in, e.n * sizeof(float2));
        it.in = reinterpret_cast<const float2 *>(dbase + in_off) + in_pos;
        in_pos += e.n;
        if (it.prev) {
            it.out = scratch + reinterpret_cast<size_t>(it.out);
            it.angles = reinterpret_cast<float *>(obase + sl.out_off[i]);
        } else {
            it.out = reinterpret_cast<float2 *>(obase + sl.out_off[i]);
        }
    }
    std::memcpy(sl.h_buf + items_off, items.data(), items.size() * sizeof(ServeItem));
    std::memcpy(sl.h_buf + table_off, table.data(), table.size() * sizeof(uint32_t));
    if (!dtable.empty())
        std::memcpy(sl.h_buf + dtable_off, dtable.data(), dtable.size() * sizeof(uint32_t));
    s->sched->mark_issued(sl.entries);
    HIP_TRY(hipMemcpyAsync(dbase, sl.h_buf, used, hipMemcpyHostToDevice, s->s_in));
    HIP_TRY(hipEventRecord(sl.ev_in, s->s_in));
    HIP_TRY(hipStreamWaitEvent(s->s_run, sl.ev_in, 0));
    hipLaunchKernelGGL(serve_chan_kernel, dim3((uint32_t)(table.size() / 2)), dim3(kServeBlock), 0, s->s_run,
                       reinterpret_cast<const ServeItem *>(dbase + items_off),
                       reinterpret_cast<const uint32_t *>(dbase + table_off));
    HIP_TRY(hipGetLastError());
    if (!dtable.empty()) {
        hipLaunchKernelGGL(serve_discriminator_kernel, dim3((uint32_t)(dtable.size() / 2)), dim3(kServeBlock), 0,
                           s->s_run, reinterpret_cast<const ServeItem *>(dbase + items_off),
                           reinterpret_cast<const uint32_t *>(dbase + dtable_off));
        HIP_TRY(hipGetLastError());
    }
    if (total_bytes_out)
        HIP_TRY(hipMemcpyAsync(sl.h_out, obase, total_bytes_out, hipMemcpyDeviceToHost, s->s_run));
    HIP_TRY(hipEventRecord(sl.ev_done, s->s_run));
    *launched = true;
    return FDSP_OK;
}

// wait for slot si's batch and deliver its results
int complete(fdsp_serve_s *s, int si)
{
    Slot &sl = s->slot[si];
    HIP_TRY(hipEventSynchronize(sl.ev_done));
    const uint64_t gpu_ns = sched_now_ns();
    for (size_t i = 0; i < sl.entries.size(); ++i) {
        const SchedEntry &e = sl.entries[i];
        if (e.emit)
            std::memcpy(e.out, sl.h_out + sl.out_off[i], e.emit * e.elem);
    }
    s->sched->complete(sl.entries, gpu_ns);
    sl.entries.clear();
    return FDSP_OK;
}

void worker_main(fdsp_serve_s *s)
{
    if (hipSetDevice(s->device) != hipSuccess) {
        s->sched->fail_all(FDSP_EHIP, "serve worker: hipSetDevice failed");
        return;
    }
    int cur = 0;
    bool running[2] = {false, false};
    for (;;) {
        bool launched = false;
        int rc = form_and_launch(s, cur, &launched);
        if (rc == FDSP_OK && launched)
            running[cur] = true;
        const int other = cur ^ 1;
        if (rc == FDSP_OK && running[other]) {
            rc = complete(s, other);
            running[other] = false;
        }
        if (rc != FDSP_OK) {
            s->sched->fail_all(rc, g_err);
            // drain what the GPU still holds, then wait for the end
            (void)hipStreamSynchronize(s->s_in);
            (void)hipStreamSynchronize(s->s_run);
            while (s->sched->wait_work()) {
            }
            return;
        }
        if (launched) {
            cur = other;  // the slot just completed takes the next batch
            continue;
        }
        if (running[cur]) {
            // nothing new could be formed: deliver the batch in flight
            if (complete(s, cur) != FDSP_OK) {
                s->sched->fail_all(FDSP_EHIP, g_err);
                return;
            }
            running[cur] = false;
            continue;
        }
        if (!s->sched->wait_work())
            return;
    }
}

void release_slot(Slot &sl)
{
    if (sl.h_buf)
        (void)hipHostFree(sl.h_buf);
    if (sl.h_out)
        (void)hipHostFree(sl.h_out);
    sl.h_buf = sl.h_out = nullptr;
    sl.d_buf.release();
    sl.d_out.release();
    sl.d_scratch.release();
    for (hipEvent_t *e : {&sl.ev_in, &sl.ev_done})
        if (*e) {
            (void)hipEventDestroy(*e);
            *e = nullptr;
        }
}

}  // namespace

extern "C" int fdsp_serve_destroy(fdsp_serve_t s);

extern "C" int fdsp_serve_create(const fdsp_serve_config *cfg, fdsp_serve_t *out)
{
    if (!cfg || !out)
        return set_err(FDSP_EINVAL, "serve_create: null argument");
    *out = nullptr;
    if (cfg->fs == 0 || cfg->max_streams == 0 || cfg->max_chunk == 0 || cfg->max_batch < cfg->max_chunk)
        return set_err(FDSP_EINVAL, "serve_create: need fs, max_streams, max_chunk > 0 and max_batch >= max_chunk");
    if (cfg->max_batch >= (1u << 31))
        return set_err(FDSP_ERANGE, "serve_create: max_batch must stay below 2^31 samples");
    auto *s = new (std::nothrow) fdsp_serve_s;
    if (!s)
        return set_err(FDSP_ENOMEM, "host allocation failed");
    s->cfg = *cfg;
    int rc = FDSP_OK;
    if (hipGetDevice(&s->device) != hipSuccess ||
        hipStreamCreateWithFlags(&s->s_in, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->s_run, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&s->s_admin, hipStreamNonBlocking) != hipSuccess)
        rc = set_err(FDSP_EHIP, "serve_create: HIP stream setup failed");
    // per slot: items, tables and inputs; outputs (<= 8 B per output sample);
    // complex scratch of the discriminator tails
    s->table_cap = cfg->max_batch / kServeBlock + cfg->max_streams + 1;
    s->buf_bytes = align_up((size_t)cfg->max_streams * sizeof(ServeItem)) +
                   2 * align_up(s->table_cap * 2 * sizeof(uint32_t)) + (size_t)cfg->max_batch * sizeof(float2);
    s->out_bytes = (size_t)cfg->max_batch * sizeof(float2) + (size_t)cfg->max_streams * 8;
    s->fixed_bytes = 2 * (s->buf_bytes + s->out_bytes + (size_t)cfg->max_batch * sizeof(float2));
    if (!rc && cfg->hbm_budget && s->fixed_bytes > cfg->hbm_budget)
        rc = set_err(FDSP_ENOMEM, "serve_create: batch buffers need %llu bytes, above the %llu-byte budget",
                     (unsigned long long)s->fixed_bytes, (unsigned long long)cfg->hbm_budget);
    // the stream state pool: state_bytes, else what the budget leaves (at most 1 GiB), else 64 KiB a stream
    uint64_t pool = cfg->state_bytes;
    if (!pool)
        pool = cfg->hbm_budget ? std::min<uint64_t>(cfg->hbm_budget - std::min(cfg->hbm_budget, s->fixed_bytes),
                                                    (uint64_t)1 << 30)
                               : (uint64_t)cfg->max_streams << 16;
    pool = pool / kAlign * kAlign;
    if (!rc && cfg->hbm_budget && s->fixed_bytes + pool > cfg->hbm_budget)
        rc = set_err(FDSP_ENOMEM, "serve_create: batch buffers and a %llu-byte state pool exceed the %llu-byte budget",
                     (unsigned long long)pool, (unsigned long long)cfg->hbm_budget);
    if (!rc && pool == 0)
        rc = set_err(FDSP_ENOMEM, "serve_create: no device memory left for stream state");
    if (!rc)
        rc = s->d_pool.ensure(pool);
    if (!rc)
        s->free_blocks.emplace(0, (size_t)pool);
    for (int i = 0; i < 2 && !rc; ++i) {
        Slot &sl = s->slot[i];
        rc = sl.d_buf.ensure(s->buf_bytes);
        if (!rc)
            rc = sl.d_out.ensure(s->out_bytes);
        if (!rc)
            rc = sl.d_scratch.ensure((size_t)cfg->max_batch * sizeof(float2));
        if (!rc && (hipHostMalloc((void **)&sl.h_buf, s->buf_bytes, hipHostMallocDefault) != hipSuccess ||
                    hipHostMalloc((void **)&sl.h_out, s->out_bytes, hipHostMallocDefault) != hipSuccess ||
                    hipEventCreateWithFlags(&sl.ev_in, hipEventDisableTiming) != hipSuccess ||
                    hipEventCreateWithFlags(&sl.ev_done, hipEventDisableTiming) != hipSuccess))
            rc = set_err(FDSP_ENOMEM, "serve_create: pinned staging allocation failed");
    }
    // Warm-up, so the first batches do not pay one-time costs (measured at
    // several ms each): touch every page of the pinned staging (first-touch
    // faults), copy it both ways once (the device's mappings of it), and
    // launch both kernels once on an empty item (their code objects load).
    for (int i = 0; i < 2 && !rc; ++i) {
        Slot &sl = s->slot[i];
        std::memset(sl.h_buf, 0, s->buf_bytes);
        std::memset(sl.h_out, 0, s->out_bytes);
        const uint32_t table[2] = {0, 0};
        std::memcpy(sl.h_buf + align_up(sizeof(ServeItem)), table, sizeof table);
        const ServeItem *items = sl.d_buf.as<ServeItem>();  // item 0: all zero, count 0
        const uint32_t *dtab = reinterpret_cast<const uint32_t *>(sl.d_buf.as<char>() + align_up(sizeof(ServeItem)));
        if (hipMemcpyAsync(sl.d_buf.p, sl.h_buf, s->buf_bytes, hipMemcpyHostToDevice, s->s_in) != hipSuccess ||
            hipStreamSynchronize(s->s_in) != hipSuccess)
            rc = set_err(FDSP_EHIP, "serve_create: staging warm-up failed");
        if (!rc) {
            hipLaunchKernelGGL(serve_chan_kernel, dim3(1), dim3(kServeBlock), 0, s->s_run, items, dtab);
            hipLaunchKernelGGL(serve_discriminator_kernel, dim3(1), dim3(kServeBlock), 0, s->s_run, items, dtab);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(sl.h_out, sl.d_out.p, s->out_bytes, hipMemcpyDeviceToHost, s->s_run) != hipSuccess ||
                hipStreamSynchronize(s->s_run) != hipSuccess)
                rc = set_err(FDSP_EHIP, "serve_create: kernel warm-up failed");
        }
    }
    if (!rc) {
        try {
            s->streams.resize(cfg->max_streams);
            for (uint32_t i = cfg->max_streams; i > 0; --i)
                s->free_ids.push_back(i - 1);
            s->sched.reset(new ServeScheduler(cfg->max_batch, kMaxQueued));
            s->worker = std::thread(worker_main, s);
        } catch (...) {
            rc = set_err(FDSP_ENOMEM, "serve_create: host setup failed");
        }
    }
    if (rc) {
        fdsp_serve_destroy(s);
        return rc;
    }
    *out = s;
    return FDSP_OK;
}

extern "C" int fdsp_serve_destroy(fdsp_serve_t s)
{
    if (!s)
        return FDSP_OK;
    if (s->worker.joinable()) {
        s->sched->stop();
        s->worker.join();
    }
    for (hipStream_t st : {s->s_in, s->s_run, s->s_admin})
        if (st)
            (void)hipStreamSynchronize(st);
    for (auto &sl : s->slot)
        release_slot(sl);
    s->d_pool.release();
    for (hipStream_t st : {s->s_in, s->s_run, s->s_admin})
        if (st)
            (void)hipStreamDestroy(st);
    delete s;
    return FDSP_OK;
}

extern "C" int fdsp_serve_admit(fdsp_serve_t s, const fdsp_channel_desc *chan, uint32_t *stream)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!chan || !stream || !chan->taps || chan->n_taps == 0 || chan->up_rate == 0 || chan->down_rate == 0 ||
        chan->tail > FDSP_TAIL_DISCRIMINATOR)
        return set_err(FDSP_EINVAL, "serve_admit: need up, down > 0, taps and a known tail");
    ServeStream st;
    st.up = chan->up_rate;
    st.down = chan->down_rate;
    st.tail = chan->tail;
    const NcoChan nco = make_nco(s->cfg.fs, (int32_t)chan->freq_shift_in_hz);  // narrowed as the reference does
    st.nu = nco.nu;
    st.r = nco.r;
    st.inv_r = nco.inv_r;
    st.wrap = phasor64(st.nu * (double)st.r);
    const std::vector<float> bank = real_bank(chan->taps, chan->n_taps, st.up, &st.cpp);
    st.hist_len = std::max<uint32_t>(1, st.cpp - 1);
    if ((uint64_t)s->cfg.max_chunk * st.up / st.down + 1 > s->cfg.max_batch)
        return set_err(FDSP_EINVAL, "serve_admit: a %u-sample chunk emits more than max_batch (%u) outputs",
                       s->cfg.max_chunk, s->cfg.max_batch);
    if (int rc = check_chunk(s->cfg.max_chunk, st.up))
        return rc;
    stream_layout(st);
    std::string why;
    if (int f = s->sched->fatal(&why))
        return set_err(f, "server stopped: %s", why.c_str());
    std::vector<float2> taps;
    append_modulated_taps(bank, st.up, st.cpp, st.nu, &taps);
    // a slice of the pool, zeroed, with the modulated taps: the upload runs
    // on the admission stream and finishes before the stream is live
    std::lock_guard<std::mutex> lk(s->pool_mu);
    if (s->live >= s->cfg.max_streams || s->free_ids.empty())
        return set_err(FDSP_EAGAIN, "serve_admit: %u live streams, the limit", s->live);
    size_t off = 0;
    if (!pool_take(s, st.dev_bytes, &off))
        return set_err(FDSP_EAGAIN, "serve_admit: no %zu-byte slice free in the %zu-byte state pool (%llu in use)",
                       st.dev_bytes, s->d_pool.bytes, (unsigned long long)s->state_used);
    st.dev = static_cast<char *>(s->d_pool.p) + off;
    st.dev_off = off;
    if (hipMemsetAsync(st.dev, 0, st.dev_bytes, s->s_admin) != hipSuccess ||
        hipMemcpyAsync(st.at<float2>(st.taps_off), taps.data(), taps.size() * sizeof(float2),
                       hipMemcpyHostToDevice, s->s_admin) != hipSuccess ||
        hipStreamSynchronize(s->s_admin) != hipSuccess) {
        pool_give(s, off, st.dev_bytes);
        return set_err(FDSP_EHIP, "serve_admit: state upload failed");
    }
    const uint32_t id = s->free_ids.back();
    s->free_ids.pop_back();
    st.live = true;
    s->streams[id] = st;
    s->state_used += st.dev_bytes;
    ++s->live;
    s->sched->add_stream(id, st.up, st.down, st.tail != 0);
    *stream = id;
    return FDSP_OK;
}

extern "C" int fdsp_serve_max_out(fdsp_serve_t s, uint32_t stream, size_t n, size_t *count)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!count)
        return set_err(FDSP_EINVAL, "null count");
    std::lock_guard<std::mutex> lk(s->pool_mu);
    if (stream >= s->streams.size() || !s->streams[stream].live)
        return set_err(FDSP_EINVAL, "stream %u is not live", stream);
    const ServeStream &st = s->streams[stream];
    *count = (size_t)((uint64_t)n * st.up / st.down + 1);
    return FDSP_OK;
}

extern "C" int fdsp_serve_submit(fdsp_serve_t s, uint32_t stream, const fdsp_cf32 *in, size_t n, void *out,
                                 size_t cap, uint64_t *ticket)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!in || !out || !ticket || n == 0)
        return set_err(FDSP_EINVAL, "serve_submit: need input, output and ticket pointers and n > 0");
    if (n > s->cfg.max_chunk)
        return set_err(FDSP_ERANGE, "serve_submit: chunk of %zu samples above max_chunk (%u)", n, s->cfg.max_chunk);
    std::string why;
    if (int rc = s->sched->submit(stream, in, n, out, cap, ticket, &why))
        return set_err(rc, "serve_submit: %s", why.c_str());
    return FDSP_OK;
}

extern "C" int fdsp_serve_poll(fdsp_serve_t s, uint64_t ticket, int wait_ms, fdsp_serve_result *res)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    SchedReq rq;
    std::string why;
    const int rc = s->sched->poll(ticket, wait_ms, &rq, &why);
    if (rc == FDSP_EAGAIN)
        return FDSP_EAGAIN;
    if (rc == FDSP_EINVAL && why.rfind("unknown", 0) == 0)
        return set_err(rc, "%s", why.c_str());
    if (res) {
        res->n_out = rq.n_out;
        res->status = rq.status;
        res->submit_ns = rq.submit_ns;
        res->done_ns = rq.done_ns;
        res->batch = rq.batch;
        res->pack_ns = rq.pack_ns;
        res->issue_ns = rq.issue_ns;
        res->gpu_ns = rq.gpu_ns;
    }
    if (rc != FDSP_OK)
        return set_err(rc, "%s", why.c_str());
    return FDSP_OK;
}

extern "C" int fdsp_serve_release(fdsp_serve_t s, uint32_t stream)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    std::lock_guard<std::mutex> lk(s->pool_mu);
    if (stream >= s->streams.size() || !s->streams[stream].live)
        return set_err(FDSP_EINVAL, "stream %u is not live", stream);
    if (int rc = s->sched->remove_stream(stream))
        return set_err(rc, "stream %u has chunks queued or running", stream);
    ServeStream &st = s->streams[stream];
    pool_give(s, st.dev_off, st.dev_bytes);
    s->state_used -= st.dev_bytes;
    --s->live;
    st = ServeStream{};
    s->free_ids.push_back(stream);
    return FDSP_OK;
}

extern "C" int fdsp_serve_info(fdsp_serve_t s, fdsp_serve_stats *st)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!st)
        return set_err(FDSP_EINVAL, "null stats");
    {
        std::lock_guard<std::mutex> lk(s->pool_mu);
        st->hbm_used = s->fixed_bytes + s->state_used;
        st->hbm_budget = s->cfg.hbm_budget;
        st->live_streams = s->live;
        st->state_pool = s->d_pool.bytes;
    }
    s->sched->stats(&st->queued, &st->batches, &st->chunks, &st->in_samples, &st->out_samples);
    return s->sched->fatal(nullptr);
}
