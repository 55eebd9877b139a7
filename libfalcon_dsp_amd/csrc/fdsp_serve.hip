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
// Runtime: clients submit chunks from any thread; one worker thread forms
// batches (at most one chunk per stream per batch, streams taken round-robin,
// bounded by max_batch input and output samples), packs them into pinned
// staging, and issues copy-in, kernels and copy-out on the server's HIP
// stream. Two staging slots alternate: while batch k runs on the GPU, batch
// k+1 is packed and queued behind it, then batch k's results are delivered.
// Requests that arrive while a batch runs join the next one.
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
#include <deque>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

#include "fdsp_host.h"
#include "fdsp_internal.h"

using namespace fdsp;

namespace {

constexpr uint32_t kServeBlock = 256;  // outputs per tile
constexpr size_t kAlign = 256;
constexpr uint32_t kMaxQueued = 1u << 16;  // requests queued or running at once

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

using Clock = std::chrono::steady_clock;
uint64_t now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(Clock::now().time_since_epoch()).count();
}

struct ServeStream {
    bool live = false;
    uint32_t up = 1, down = 1, cpp = 1, hist_len = 1, tail = 0;
    double nu = 0, inv_r = 1;
    uint32_t r = 1;
    float2 wrap{1.f, 0.f};
    void *dev = nullptr;  // taps | hist[2] | prev[2]
    size_t dev_bytes = 0;
    size_t taps_off = 0, hist_off[2] = {0, 0}, prev_off[2] = {0, 0};
    // host view of the stream position (advanced when a chunk joins a batch)
    uint64_t p = 0;
    uint32_t parity = 0, prev_parity = 0;
    bool has_prev = false;
    std::deque<uint64_t> queue;  // tickets waiting, oldest first
    bool busy = false;           // a chunk of this stream is in a launched batch
    template <class T> T *at(size_t off) const { return reinterpret_cast<T *>(static_cast<char *>(dev) + off); }
};

struct ServeReq {
    uint32_t stream;
    const fdsp_cf32 *in;
    size_t n;
    void *out;
    size_t cap;
    uint64_t submit_ns = 0, done_ns = 0, batch = 0;
    size_t n_out = 0;
    int status = FDSP_OK;
    bool done = false;
};

// what the worker needs of one chunk of a packed batch
struct BatchEntry {
    uint64_t ticket;
    uint32_t stream;
    const fdsp_cf32 *in;
    size_t n;
    void *out;
    uint32_t count;    // complex outputs computed
    size_t emit;       // samples delivered (count, or count - 1 on a tail's first chunk)
    size_t elem;       // bytes per delivered sample
    size_t out_off;    // bytes into the output region
};

struct Slot {
    char *h_buf = nullptr;  // pinned: items | table | tail table | inputs
    char *h_out = nullptr;  // pinned: packed outputs
    DevBuf d_buf, d_out, d_scratch;
    hipEvent_t ev = nullptr;
    std::vector<BatchEntry> entries;
    uint64_t batch = 0;
};

}  // namespace

struct fdsp_serve_s {
    fdsp_serve_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    size_t table_cap = 0;   // tiles per batch at most
    size_t buf_bytes = 0, out_bytes = 0;
    uint64_t fixed_bytes = 0, used_bytes = 0;
    Slot slot[2];

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<ServeStream> streams;
    std::vector<uint32_t> free_ids;
    uint32_t live = 0;
    std::unordered_map<uint64_t, ServeReq> reqs;
    uint64_t next_ticket = 1;
    size_t queued = 0;
    uint32_t rr = 0;
    uint64_t batches = 0, chunks = 0, in_samples = 0, out_samples = 0;
    int fatal = FDSP_OK;
    std::string fatal_msg;
    bool stop = false;
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

bool eligible(const fdsp_serve_s *s)
{
    for (const ServeStream &st : s->streams)
        if (st.live && !st.busy && !st.queue.empty())
            return true;
    return false;
}

void fail_all(fdsp_serve_s *s, int code, const std::string &msg)
{
    // a HIP failure: every waiting and running request ends with it
    std::lock_guard<std::mutex> lk(s->mu);
    s->fatal = code;
    s->fatal_msg = msg;
    for (auto &kv : s->reqs)
        if (!kv.second.done) {
            kv.second.done = true;
            kv.second.status = code;
            kv.second.done_ns = now_ns();
        }
    for (ServeStream &st : s->streams) {
        st.queue.clear();
        st.busy = false;
    }
    s->queued = 0;
    s->cv_done.notify_all();
}

// Take the next batch into slot `si` and queue its work on the GPU. Returns
// false when no stream has a chunk ready.
int form_and_launch(fdsp_serve_s *s, int si, bool *launched)
{
    Slot &sl = s->slot[si];
    *launched = false;
    sl.entries.clear();
    std::vector<ServeItem> items;
    std::vector<uint32_t> table, dtable;
    size_t total_in = 0, total_out = 0, total_scratch = 0, total_bytes_out = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal)
            return FDSP_OK;
        const uint32_t S = (uint32_t)s->streams.size();
        for (uint32_t k = 0; k < S; ++k) {
            const uint32_t id = (s->rr + k) % S;
            ServeStream &st = s->streams[id];
            if (!st.live || st.busy || st.queue.empty())
                continue;
            const uint64_t ticket = st.queue.front();
            ServeReq &rq = s->reqs.at(ticket);
            const uint32_t count = chunk_count(st.p, rq.n, st.up, st.down);
            if (total_in + rq.n > s->cfg.max_batch || total_out + count > s->cfg.max_batch)
                continue;  // does not fit this batch; it leads the next one (admission keeps one chunk within)
            st.queue.pop_front();
            st.busy = true;
            BatchEntry e{};
            e.ticket = ticket;
            e.stream = id;
            e.in = rq.in;
            e.n = rq.n;
            e.out = rq.out;
            e.count = count;
            e.elem = st.tail ? sizeof(float) : sizeof(float2);
            e.emit = (st.tail && count && !st.has_prev) ? count - 1 : count;
            e.out_off = total_bytes_out;
            total_bytes_out += (e.emit * e.elem + 7) / 8 * 8;
            ServeItem it{};
            it.in = nullptr;  // set after packing (device address of the slot)
            it.hist = st.at<float2>(st.hist_off[st.parity]);
            it.hist_next = st.at<float2>(st.hist_off[st.parity ^ 1]);
            it.taps = st.at<float2>(st.taps_off);
            it.nu = st.nu;
            it.inv_r = st.inv_r;
            it.inv_up = 1.0 / (double)st.up;
            it.p = st.p;
            it.n0 = ceil_div(st.p * st.up, st.down);
            it.wrap = st.wrap;
            it.n = (uint32_t)rq.n;
            it.count = count;
            it.r = st.r;
            it.up = st.up;
            it.down = st.down;
            it.cpp = st.cpp;
            it.hist_len = st.hist_len;
            it.has_prev = st.has_prev ? 1u : 0u;
            if (st.tail) {
                it.prev = st.at<float2>(st.prev_off[st.prev_parity]);
                it.prev_next = st.at<float2>(st.prev_off[st.prev_parity ^ 1]);
                // complex outputs go to scratch; `out` and `angles` are
                // patched with device addresses below
                it.out = reinterpret_cast<float2 *>(total_scratch);
                total_scratch += count;
            }
            const uint32_t idx = (uint32_t)items.size();
            items.push_back(it);
            const uint32_t tiles = std::max<uint32_t>(1, (count + kServeBlock - 1) / kServeBlock);
            for (uint32_t t = 0; t < tiles; ++t) {
                table.push_back(idx);
                table.push_back(t);
            }
            if (st.tail)
                for (uint32_t t = 0; t * kServeBlock < count; ++t) {
                    dtable.push_back(idx);
                    dtable.push_back(t);
                }
            // advance the host view of the stream
            st.p += rq.n;
            st.parity ^= 1;
            if (st.tail && count) {
                st.has_prev = true;
                st.prev_parity ^= 1;
            }
            total_in += rq.n;
            total_out += count;
            sl.entries.push_back(e);
        }
        if (!sl.entries.empty()) {
            s->rr = (s->rr + 1) % std::max<uint32_t>(1, S);
            sl.batch = ++s->batches;
        }
    }
    if (sl.entries.empty())
        return FDSP_OK;
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
        const BatchEntry &e = sl.entries[i];
        ServeItem &it = items[i];
        std::memcpy(sl.h_buf + in_off + in_pos * sizeof(float2), e.in, e.n * sizeof(float2));
        it.in = reinterpret_cast<const float2 *>(dbase + in_off) + in_pos;
        in_pos += e.n;
        if (it.prev) {
            it.out = scratch + reinterpret_cast<size_t>(it.out);
            it.angles = reinterpret_cast<float *>(obase + e.out_off);
        } else {
            it.out = reinterpret_cast<float2 *>(obase + e.out_off);
        }
    }
    std::memcpy(sl.h_buf + items_off, items.data(), items.size() * sizeof(ServeItem));
    std::memcpy(sl.h_buf + table_off, table.data(), table.size() * sizeof(uint32_t));
    if (!dtable.empty())
        std::memcpy(sl.h_buf + dtable_off, dtable.data(), dtable.size() * sizeof(uint32_t));
    HIP_TRY(hipMemcpyAsync(dbase, sl.h_buf, used, hipMemcpyHostToDevice, s->stream));
    hipLaunchKernelGGL(serve_chan_kernel, dim3((uint32_t)(table.size() / 2)), dim3(kServeBlock), 0, s->stream,
                       reinterpret_cast<const ServeItem *>(dbase + items_off),
                       reinterpret_cast<const uint32_t *>(dbase + table_off));
    HIP_TRY(hipGetLastError());
    if (!dtable.empty()) {
        hipLaunchKernelGGL(serve_discriminator_kernel, dim3((uint32_t)(dtable.size() / 2)), dim3(kServeBlock), 0,
                           s->stream, reinterpret_cast<const ServeItem *>(dbase + items_off),
                           reinterpret_cast<const uint32_t *>(dbase + dtable_off));
        HIP_TRY(hipGetLastError());
    }
    if (total_bytes_out)
        HIP_TRY(hipMemcpyAsync(sl.h_out, obase, total_bytes_out, hipMemcpyDeviceToHost, s->stream));
    HIP_TRY(hipEventRecord(sl.ev, s->stream));
    *launched = true;
    return FDSP_OK;
}

// wait for slot si's batch and deliver its results
int complete(fdsp_serve_s *s, int si)
{
    Slot &sl = s->slot[si];
    HIP_TRY(hipEventSynchronize(sl.ev));
    for (const BatchEntry &e : sl.entries)
        if (e.emit)
            std::memcpy(e.out, sl.h_out + e.out_off, e.emit * e.elem);
    const uint64_t t = now_ns();
    std::lock_guard<std::mutex> lk(s->mu);
    for (const BatchEntry &e : sl.entries) {
        ServeReq &rq = s->reqs.at(e.ticket);
        rq.done = true;
        rq.status = FDSP_OK;
        rq.n_out = e.emit;
        rq.done_ns = t;
        rq.batch = sl.batch;
        s->streams[e.stream].busy = false;
        --s->queued;
        ++s->chunks;
        s->in_samples += e.n;
        s->out_samples += e.emit;
    }
    sl.entries.clear();
    s->cv_done.notify_all();
    return FDSP_OK;
}

void worker_main(fdsp_serve_s *s)
{
    if (hipSetDevice(s->device) != hipSuccess) {
        fail_all(s, FDSP_EHIP, "serve worker: hipSetDevice failed");
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
            fail_all(s, rc, g_err);
            // drain what the GPU still holds before going idle
            (void)hipStreamSynchronize(s->stream);
            running[0] = running[1] = false;
            std::unique_lock<std::mutex> lk(s->mu);
            s->cv_work.wait(lk, [&] { return s->stop; });
            return;
        }
        if (launched) {
            cur = other;  // the slot just completed takes the next batch
            continue;
        }
        if (running[cur]) {
            // nothing new could be formed: deliver the batch in flight
            if (complete(s, cur) != FDSP_OK) {
                fail_all(s, FDSP_EHIP, g_err);
                return;
            }
            running[cur] = false;
            continue;
        }
        std::unique_lock<std::mutex> lk(s->mu);
        s->cv_work.wait(lk, [&] { return s->stop || s->fatal || eligible(s); });
        if (s->fatal || (s->stop && !eligible(s)))
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
    if (sl.ev)
        (void)hipEventDestroy(sl.ev);
    sl.ev = nullptr;
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
        hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess)
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
    for (int i = 0; i < 2 && !rc; ++i) {
        Slot &sl = s->slot[i];
        rc = sl.d_buf.ensure(s->buf_bytes);
        if (!rc)
            rc = sl.d_out.ensure(s->out_bytes);
        if (!rc)
            rc = sl.d_scratch.ensure((size_t)cfg->max_batch * sizeof(float2));
        if (!rc && (hipHostMalloc((void **)&sl.h_buf, s->buf_bytes, hipHostMallocDefault) != hipSuccess ||
                    hipHostMalloc((void **)&sl.h_out, s->out_bytes, hipHostMallocDefault) != hipSuccess ||
                    hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming) != hipSuccess))
            rc = set_err(FDSP_ENOMEM, "serve_create: pinned staging allocation failed");
    }
    s->used_bytes = s->fixed_bytes;
    if (!rc) {
        try {
            s->worker = std::thread(worker_main, s);
        } catch (...) {
            rc = set_err(FDSP_ENOMEM, "serve_create: worker thread could not start");
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
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->stop = true;
        }
        s->cv_work.notify_all();
        s->worker.join();
    }
    if (s->stream)
        (void)hipStreamSynchronize(s->stream);
    for (auto &sl : s->slot)
        release_slot(sl);
    for (auto &st : s->streams)
        if (st.dev)
            (void)hipFree(st.dev);
    if (s->stream)
        (void)hipStreamDestroy(s->stream);
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
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->fatal)
            return set_err(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
        if (s->live >= s->cfg.max_streams)
            return set_err(FDSP_EAGAIN, "serve_admit: %u live streams, the limit", s->live);
        if (s->cfg.hbm_budget && s->used_bytes + st.dev_bytes > s->cfg.hbm_budget)
            return set_err(FDSP_EAGAIN, "serve_admit: %zu more bytes exceed the budget (%llu of %llu in use)",
                           st.dev_bytes, (unsigned long long)s->used_bytes,
                           (unsigned long long)s->cfg.hbm_budget);
        s->used_bytes += st.dev_bytes;
        ++s->live;
    }
    std::vector<float2> taps;
    append_modulated_taps(bank, st.up, st.cpp, st.nu, &taps);
    int rc = FDSP_OK;
    if (hipMalloc(&st.dev, st.dev_bytes) != hipSuccess) {
        st.dev = nullptr;
        rc = set_err(FDSP_ENOMEM, "serve_admit: hipMalloc(%zu) failed", st.dev_bytes);
    } else if (hipMemset(st.dev, 0, st.dev_bytes) != hipSuccess ||
               hipMemcpy(st.at<float2>(st.taps_off), taps.data(), taps.size() * sizeof(float2),
                         hipMemcpyHostToDevice) != hipSuccess) {
        rc = set_err(FDSP_EHIP, "serve_admit: state upload failed");
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (rc) {
        if (st.dev)
            (void)hipFree(st.dev);
        s->used_bytes -= st.dev_bytes;
        --s->live;
        return rc;
    }
    st.live = true;
    uint32_t id;
    if (!s->free_ids.empty()) {
        id = s->free_ids.back();
        s->free_ids.pop_back();
        s->streams[id] = std::move(st);
    } else {
        id = (uint32_t)s->streams.size();
        s->streams.push_back(std::move(st));
    }
    *stream = id;
    return FDSP_OK;
}

extern "C" int fdsp_serve_max_out(fdsp_serve_t s, uint32_t stream, size_t n, size_t *count)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!count)
        return set_err(FDSP_EINVAL, "null count");
    std::lock_guard<std::mutex> lk(s->mu);
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
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal)
        return set_err(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
    if (stream >= s->streams.size() || !s->streams[stream].live)
        return set_err(FDSP_EINVAL, "stream %u is not live", stream);
    ServeStream &st = s->streams[stream];
    const size_t bound = (size_t)((uint64_t)n * st.up / st.down + 1);
    if (cap < bound)
        return set_err(FDSP_ERANGE, "serve_submit: output capacity %zu below the %zu samples a chunk may emit", cap,
                       bound);
    if (s->queued >= kMaxQueued)
        return set_err(FDSP_EAGAIN, "serve_submit: %zu requests in flight, the limit", s->queued);
    ServeReq rq;
    rq.stream = stream;
    rq.in = in;
    rq.n = n;
    rq.out = out;
    rq.cap = cap;
    rq.submit_ns = now_ns();
    const uint64_t t = s->next_ticket++;
    s->reqs.emplace(t, rq);
    st.queue.push_back(t);
    ++s->queued;
    *ticket = t;
    s->cv_work.notify_one();
    return FDSP_OK;
}

extern "C" int fdsp_serve_poll(fdsp_serve_t s, uint64_t ticket, int wait_ms, fdsp_serve_result *res)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end())
        return set_err(FDSP_EINVAL, "unknown ticket %llu", (unsigned long long)ticket);
    if (!it->second.done && wait_ms != 0) {
        auto ready = [&] { return s->reqs.at(ticket).done; };
        if (wait_ms < 0)
            s->cv_done.wait(lk, ready);
        else
            s->cv_done.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
        it = s->reqs.find(ticket);
    }
    if (!it->second.done)
        return FDSP_EAGAIN;
    const ServeReq rq = it->second;
    s->reqs.erase(it);
    if (res) {
        res->n_out = rq.n_out;
        res->status = rq.status;
        res->submit_ns = rq.submit_ns;
        res->done_ns = rq.done_ns;
        res->batch = rq.batch;
    }
    if (rq.status != FDSP_OK)
        return set_err(rq.status, "request failed: %s", s->fatal_msg.c_str());
    return FDSP_OK;
}

extern "C" int fdsp_serve_release(fdsp_serve_t s, uint32_t stream)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    void *dev = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (stream >= s->streams.size() || !s->streams[stream].live)
            return set_err(FDSP_EINVAL, "stream %u is not live", stream);
        ServeStream &st = s->streams[stream];
        if (st.busy || !st.queue.empty())
            return set_err(FDSP_EBUSY, "stream %u has chunks queued or running", stream);
        dev = st.dev;
        s->used_bytes -= st.dev_bytes;
        --s->live;
        st = ServeStream{};
        s->free_ids.push_back(stream);
    }
    if (dev)
        HIP_TRY(hipFree(dev));
    return FDSP_OK;
}

extern "C" int fdsp_serve_info(fdsp_serve_t s, fdsp_serve_stats *st)
{
    if (!s)
        return set_err(FDSP_ENOTINIT, "null server");
    if (!st)
        return set_err(FDSP_EINVAL, "null stats");
    std::lock_guard<std::mutex> lk(s->mu);
    st->hbm_used = s->used_bytes;
    st->hbm_budget = s->cfg.hbm_budget;
    st->live_streams = s->live;
    st->queued = s->queued;
    st->batches = s->batches;
    st->chunks = s->chunks;
    st->in_samples = s->in_samples;
    st->out_samples = s->out_samples;
    return s->fatal;
}
