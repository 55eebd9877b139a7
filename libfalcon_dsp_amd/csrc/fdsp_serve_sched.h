// fdsp_serve_sched.h -- the serving loop's host-side scheduler: request
// queues, stream positions, batch formation and completion, independent of
// HIP (so it builds and runs on its own, e.g. under ThreadSanitizer).
//
// Clients submit chunks from any thread. One worker forms batches: at most
// one chunk per stream per batch, streams taken round-robin from a rotating
// start, bounded by max_batch input and output samples. Forming a batch
// advances each taken stream's position (inputs consumed p, the history
// copy's parity, the discriminator's carried sample) as the reference's
// per-chunk state (falcon_dsp_multi_rate_channelizer.cc:146-164,
// falcon_dsp_polar_discriminator.cc:134-171) would advance, and marks the
// stream busy until the batch completes, so a stream's chunks run one at a
// time and in submission order.
#pragma once

#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <cstdio>
#include <deque>
#include <mutex>
#include <string>
#include <unordered_map>
#include <vector>

#include "../../include/fdsp.h"

namespace fdsp {

inline uint64_t sched_now_ns()
{
    return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

// outputs of a chunk of n inputs starting at stream input p: ceil((p+n) up /
// down) - ceil(p up / down)
inline uint32_t sched_chunk_count(uint64_t p, uint64_t n, uint32_t up, uint32_t down)
{
    const auto cdiv = [](uint64_t a, uint64_t b) { return (a + b - 1) / b; };
    return (uint32_t)(cdiv((p + n) * up, down) - cdiv(p * up, down));
}

struct SchedReq {
    uint32_t stream = 0;
    const void *in = nullptr;
    size_t n = 0;
    void *out = nullptr;
    size_t cap = 0;
    // host steady-clock times: submitted, taken into a batch, its copy-in
    // issued, its completion seen, its outputs delivered
    uint64_t submit_ns = 0, pack_ns = 0, issue_ns = 0, gpu_ns = 0, done_ns = 0;
    uint64_t batch = 0;
    size_t n_out = 0;
    int status = FDSP_OK;
    bool done = false;
};

// one chunk of a formed batch, with its stream's state before the chunk
struct SchedEntry {
    uint64_t ticket = 0;
    uint32_t stream = 0;
    const void *in = nullptr;
    size_t n = 0;
    void *out = nullptr;
    uint32_t count = 0;  // complex outputs computed
    size_t emit = 0;     // samples delivered (count, or count - 1 on a tail's first chunk)
    size_t elem = 0;     // bytes per delivered sample
    uint64_t p = 0;      // stream inputs consumed before the chunk
    uint32_t parity = 0, prev_parity = 0;
    bool has_prev = false, tail = false;
};

class ServeScheduler {
public:
    explicit ServeScheduler(uint32_t max_batch, size_t max_queued = 1u << 16)
        : max_batch_(max_batch), max_queued_(max_queued)
    {
    }

    // a stream slot (id from the caller's allocation of ids)
    void add_stream(uint32_t id, uint32_t up, uint32_t down, bool tail)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (id >= streams_.size())
            streams_.resize(id + 1);
        Stream &st = streams_[id];
        st = Stream{};
        st.live = true;
        st.up = up;
        st.down = down;
        st.tail = tail;
    }

    // FDSP_EBUSY while the stream has requests, FDSP_EINVAL if not live
    int remove_stream(uint32_t id)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (id >= streams_.size() || !streams_[id].live)
            return FDSP_EINVAL;
        if (streams_[id].busy || !streams_[id].queue.empty())
            return FDSP_EBUSY;
        streams_[id] = Stream{};
        return FDSP_OK;
    }

    bool live(uint32_t id)
    {
        std::lock_guard<std::mutex> lk(mu_);
        return id < streams_.size() && streams_[id].live;
    }

    // queue one chunk; on failure *why holds the reason
    int submit(uint32_t id, const void *in, size_t n, void *out, size_t cap, uint64_t *ticket, std::string *why)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (fatal_) {
            *why = "server stopped: " + fatal_msg_;
            return fatal_;
        }
        if (id >= streams_.size() || !streams_[id].live) {
            *why = "stream " + std::to_string(id) + " is not live";
            return FDSP_EINVAL;
        }
        Stream &st = streams_[id];
        const size_t bound = (size_t)((uint64_t)n * st.up / st.down + 1);
        if (cap < bound) {
            *why = "output capacity " + std::to_string(cap) + " below the " + std::to_string(bound) +
                   " samples a chunk may emit";
            return FDSP_ERANGE;
        }
        if (queued_ >= max_queued_) {
            *why = std::to_string(queued_) + " requests in flight, the limit";
            return FDSP_EAGAIN;
        }
        SchedReq rq;
        rq.stream = id;
        rq.in = in;
        rq.n = n;
        rq.out = out;
        rq.cap = cap;
        rq.submit_ns = sched_now_ns();
        const uint64_t t = next_ticket_++;
        reqs_.emplace(t, rq);
        st.queue.push_back(t);
        ++queued_;
        *ticket = t;
        cv_work_.notify_one();
        return FDSP_OK;
    }

    // worker: wait until a chunk can be batched, or the scheduler stops;
    // false when the worker should exit
    bool wait_work()
    {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&] { return stop_ || fatal_ || eligible(); });
        return !fatal_ && !(stop_ && !eligible());
    }

    bool has_work()
    {
        std::lock_guard<std::mutex> lk(mu_);
        return !fatal_ && eligible();
    }

    // Form the next batch into *out (cleared first); *batch receives its
    // sequence number. Empty when no stream has a chunk ready.
    void form_batch(std::vector<SchedEntry> *out, uint64_t *batch)
    {
        out->clear();
        std::lock_guard<std::mutex> lk(mu_);
        if (fatal_)
            return;
        const uint32_t S = (uint32_t)streams_.size();
        size_t total_in = 0, total_out = 0;
        const uint64_t t_now = sched_now_ns();
        for (uint32_t k = 0; k < S; ++k) {
            const uint32_t id = (rr_ + k) % S;
            Stream &st = streams_[id];
            if (!st.live || st.busy || st.queue.empty())
                continue;
            const uint64_t ticket = st.queue.front();
            SchedReq &rq = reqs_.at(ticket);
            const uint32_t count = sched_chunk_count(st.p, rq.n, st.up, st.down);
            if (total_in + rq.n > max_batch_ || total_out + count > max_batch_)
                continue;  // does not fit this batch; it leads a later one
            st.queue.pop_front();
            st.busy = true;
            SchedEntry e;
            e.ticket = ticket;
            e.stream = id;
            e.in = rq.in;
            e.n = rq.n;
            e.out = rq.out;
            e.count = count;
            e.tail = st.tail;
            e.elem = st.tail ? sizeof(float) : 2 * sizeof(float);
            e.emit = (st.tail && count && !st.has_prev) ? count - 1 : count;
            e.p = st.p;
            e.parity = st.parity;
            e.prev_parity = st.prev_parity;
            e.has_prev = st.has_prev;
            rq.pack_ns = t_now;
            st.p += rq.n;
            st.parity ^= 1;
            if (st.tail && count) {
                st.has_prev = true;
                st.prev_parity ^= 1;
            }
            total_in += rq.n;
            total_out += count;
            out->push_back(e);
        }
        if (!out->empty()) {
            rr_ = (rr_ + 1) % (S ? S : 1);
            *batch = ++batches_;
            for (const SchedEntry &e : *out)
                reqs_.at(e.ticket).batch = *batch;
        }
    }

    void mark_issued(const std::vector<SchedEntry> &entries)
    {
        const uint64_t t = sched_now_ns();
        std::lock_guard<std::mutex> lk(mu_);
        for (const SchedEntry &e : entries)
            reqs_.at(e.ticket).issue_ns = t;
    }

    // the batch's outputs are in the callers' buffers: its requests are done
    void complete(const std::vector<SchedEntry> &entries, uint64_t gpu_ns)
    {
        const uint64_t t = sched_now_ns();
        std::lock_guard<std::mutex> lk(mu_);
        for (const SchedEntry &e : entries) {
            SchedReq &rq = reqs_.at(e.ticket);
            rq.done = true;
            rq.status = FDSP_OK;
            rq.n_out = e.emit;
            rq.gpu_ns = gpu_ns;
            rq.done_ns = t;
            streams_[e.stream].busy = false;
            --queued_;
            ++chunks_;
            in_samples_ += e.n;
            out_samples_ += e.emit;
        }
        cv_done_.notify_all();
        cv_work_.notify_one();
    }

    // a failure: every waiting and running request ends with it
    void fail_all(int code, const std::string &msg)
    {
        std::lock_guard<std::mutex> lk(mu_);
        fatal_ = code;
        fatal_msg_ = msg;
        const uint64_t t = sched_now_ns();
        for (auto &kv : reqs_)
            if (!kv.second.done) {
                kv.second.done = true;
                kv.second.status = code;
                kv.second.done_ns = t;
            }
        for (Stream &st : streams_) {
            st.queue.clear();
            st.busy = false;
        }
        queued_ = 0;
        cv_done_.notify_all();
        cv_work_.notify_all();
    }

    void stop()
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
        cv_work_.notify_all();
    }

    // FDSP_EAGAIN while running (after waiting up to wait_ms; < 0 waits);
    // a finished ticket is consumed and *rq receives it
    int poll(uint64_t ticket, int wait_ms, SchedReq *rq, std::string *why)
    {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = reqs_.find(ticket);
        if (it == reqs_.end()) {
            *why = "unknown ticket " + std::to_string(ticket);
            return FDSP_EINVAL;
        }
        if (!it->second.done && wait_ms != 0) {
            auto ready = [&] { return reqs_.at(ticket).done; };
            if (wait_ms < 0)
                cv_done_.wait(lk, ready);
            else
                cv_done_.wait_for(lk, std::chrono::milliseconds(wait_ms), ready);
            it = reqs_.find(ticket);
        }
        if (!it->second.done)
            return FDSP_EAGAIN;
        *rq = it->second;
        reqs_.erase(it);
        if (rq->status != FDSP_OK)
            *why = "request failed: " + fatal_msg_;
        return rq->status;
    }

    int fatal(std::string *msg)
    {
        std::lock_guard<std::mutex> lk(mu_);
        if (msg)
            *msg = fatal_msg_;
        return fatal_;
    }

    void stats(size_t *queued, uint64_t *batches, uint64_t *chunks, uint64_t *in_samples, uint64_t *out_samples)
    {
        std::lock_guard<std::mutex> lk(mu_);
        *queued = queued_;
        *batches = batches_;
        *chunks = chunks_;
        *in_samples = in_samples_;
        *out_samples = out_samples_;
    }

private:
    struct Stream {
        bool live = false, tail = false;
        uint32_t up = 1, down = 1;
        uint64_t p = 0;
        uint32_t parity = 0, prev_parity = 0;
        bool has_prev = false;
        std::deque<uint64_t> queue;  // tickets waiting, oldest first
        bool busy = false;           // a chunk of this stream is in a formed batch
    };

    bool eligible() const
    {
        for (const Stream &st : streams_)
            if (st.live && !st.busy && !st.queue.empty())
                return true;
        return false;
    }

    const uint32_t max_batch_;
    const size_t max_queued_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::vector<Stream> streams_;
    std::unordered_map<uint64_t, SchedReq> reqs_;
    uint64_t next_ticket_ = 1;
    size_t queued_ = 0;
    uint32_t rr_ = 0;
    uint64_t batches_ = 0, chunks_ = 0, in_samples_ = 0, out_samples_ = 0;
    int fatal_ = FDSP_OK;
    std::string fatal_msg_;
    bool stop_ = false;
};

}  // namespace fdsp
