// Host-only run of the serving scheduler (libfalcon_dsp_amd/csrc/
// fdsp_serve_sched.h), built with -fsanitize=thread by tests/test_sanitizers_cpu.py.
//
// Client threads submit chunks to their own streams and collect them, some
// by polling without waiting, some by blocking; a worker thread forms
// batches and keeps up to two in flight, as the server's worker does, with a
// stand-in for the GPU that writes each chunk's stream position and batch
// number into its output. Checks: every request completes exactly once with
// the output count the chunk emits; a stream's chunks run in submission
// order, each starting where the previous ended; no batch holds two chunks of
// one stream or a stream whose chunk is still in flight; batches respect
// max_batch; stop ends the worker; fail_all ends every waiting request with
// its error.
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <deque>
#include <random>
#include <set>
#include <thread>
#include <vector>

#include "fdsp_serve_sched.h"

using namespace fdsp;

static std::atomic<int> g_fail{0};
#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            std::fprintf(stderr, "CHECK failed: %s (line %d)\n", #c, __LINE__); \
            g_fail = 1;                                                   \
        }                                                                 \
    } while (0)

struct Out {
    uint64_t p;      // the stream position the chunk started at
    uint64_t batch;  // the batch it ran in
};

constexpr uint32_t kStreams = 24, kClients = 6, kChunks = 60, kMaxChunk = 5000, kMaxBatch = 20000;

static void worker(ServeScheduler *s, std::atomic<uint64_t> *batches)
{
    std::deque<std::pair<std::vector<SchedEntry>, uint64_t>> flight;
    std::set<uint32_t> busy;
    for (;;) {
        std::vector<SchedEntry> e;
        uint64_t b = 0;
        s->form_batch(&e, &b);
        const bool formed = !e.empty();
        if (formed) {
            size_t in = 0, out = 0;
            std::set<uint32_t> seen;
            for (const SchedEntry &x : e) {
                CHECK(seen.insert(x.stream).second);
                CHECK(!busy.count(x.stream));
                in += x.n;
                out += x.count;
            }
            CHECK(in <= kMaxBatch && out <= kMaxBatch);
            for (const SchedEntry &x : e)
                busy.insert(x.stream);
            s->mark_issued(e);
            for (const SchedEntry &x : e)
                *static_cast<Out *>(x.out) = Out{x.p, b};
            flight.emplace_back(std::move(e), b);
            ++*batches;
        }
        if (!flight.empty() && (!formed || flight.size() == 2)) {
            for (const SchedEntry &x : flight.front().first)
                busy.erase(x.stream);
            s->complete(flight.front().first, sched_now_ns());
            flight.pop_front();
            continue;
        }
        if (!formed && flight.empty() && !s->wait_work())
            return;
    }
}

static void client(ServeScheduler *s, uint32_t c, std::vector<uint32_t> downs)
{
    std::mt19937 rng(1234 + c);
    std::vector<uint32_t> mine;
    for (uint32_t id = c; id < kStreams; id += kClients)
        mine.push_back(id);
    struct Sent {
        uint32_t stream;
        uint64_t ticket;
        size_t n;
        std::vector<Out> out;
    };
    std::vector<Sent> sent;
    sent.reserve(mine.size() * kChunks);
    for (uint32_t k = 0; k < kChunks; ++k)
        for (uint32_t id : mine) {
            Sent x{id, 0, 1 + rng() % kMaxChunk, std::vector<Out>(1)};
            std::string why;
            const int rc = s->submit(id, nullptr, x.n, x.out.data(), x.n + 1, &x.ticket, &why);
            CHECK(rc == FDSP_OK);
            sent.push_back(std::move(x));
        }
    std::vector<uint64_t> pos(kStreams, 0);
    std::vector<bool> has_prev(kStreams, false);  // tails: a previous output exists
    for (Sent &x : sent) {
        SchedReq rq;
        std::string why;
        int rc;
        if (c % 2) {
            while ((rc = s->poll(x.ticket, 0, &rq, &why)) == FDSP_EAGAIN)
                std::this_thread::yield();
        } else {
            rc = s->poll(x.ticket, -1, &rq, &why);
        }
        CHECK(rc == FDSP_OK && rq.done);
        CHECK(x.out[0].p == pos[x.stream]);
        const uint32_t count = sched_chunk_count(pos[x.stream], x.n, 1, downs[x.stream]);
        const bool tail = x.stream % 4 == 0;
        // a tail's stream-first output has no predecessor and is not delivered
        CHECK(rq.n_out == (tail && count && !has_prev[x.stream] ? count - 1 : count));
        if (tail && count)
            has_prev[x.stream] = true;
        CHECK(rq.submit_ns <= rq.pack_ns && rq.pack_ns <= rq.issue_ns && rq.issue_ns <= rq.gpu_ns &&
              rq.gpu_ns <= rq.done_ns);
        CHECK(rq.batch == x.out[0].batch);
        pos[x.stream] += x.n;
        // a consumed ticket is gone
        CHECK(s->poll(x.ticket, 0, &rq, &why) == FDSP_EINVAL);
    }
}

int main()
{
    // normal operation
    {
        ServeScheduler s(kMaxBatch);
        std::vector<uint32_t> downs(kStreams);
        for (uint32_t id = 0; id < kStreams; ++id) {
            downs[id] = 1 + id % 7;
            s.add_stream(id, 1, downs[id], id % 4 == 0);
        }
        std::atomic<uint64_t> batches{0};
        std::thread w(worker, &s, &batches);
        std::vector<std::thread> cs;
        for (uint32_t c = 0; c < kClients; ++c)
            cs.emplace_back(client, &s, c, downs);
        for (auto &t : cs)
            t.join();
        size_t queued;
        uint64_t b, chunks, in, out;
        s.stats(&queued, &b, &chunks, &in, &out);
        CHECK(queued == 0 && chunks == (uint64_t)kStreams * kChunks && b == batches.load());
        for (uint32_t id = 0; id < kStreams; ++id)
            CHECK(s.remove_stream(id) == FDSP_OK);
        s.stop();
        w.join();
        std::printf("scheduler: %llu chunks in %llu batches\n", (unsigned long long)chunks, (unsigned long long)b);
    }
    // failure: requests waiting when fail_all runs end with its code
    {
        ServeScheduler s(kMaxBatch);
        s.add_stream(0, 1, 2, false);
        std::vector<Out> outs(50);
        std::vector<uint64_t> tickets(50);
        std::string why;
        for (int i = 0; i < 50; ++i)
            CHECK(s.submit(0, nullptr, 100, &outs[i], 101, &tickets[i], &why) == FDSP_OK);
        std::vector<std::thread> waiters;
        std::atomic<int> failed{0};
        for (int i = 0; i < 50; ++i)
            waiters.emplace_back([&, i] {
                SchedReq rq;
                std::string w;
                if (s.poll(tickets[i], -1, &rq, &w) == FDSP_EHIP)
                    ++failed;
            });
        std::thread killer([&] { s.fail_all(FDSP_EHIP, "stand-in failure"); });
        killer.join();
        for (auto &t : waiters)
            t.join();
        CHECK(failed == 50);
        uint64_t t;
        CHECK(s.submit(0, nullptr, 100, &outs[0], 101, &t, &why) == FDSP_EHIP);
        CHECK(!s.wait_work());
    }
    if (g_fail) {
        std::fprintf(stderr, "scheduler checks failed\n");
        return 1;
    }
    std::printf("scheduler checks passed\n");
    return 0;
}
