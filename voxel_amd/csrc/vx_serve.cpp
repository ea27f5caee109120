// vx_serve.cpp -- continuous-batching query server over resident columns.
//
// No counterpart in the reference (its queries are one Engine::Run each,
// exec/engine.hpp:109-141).  The server owns `inflight` batch engines
// (BatchRunner, vx_batch.hip: own stream, persistent kernel, graph launch).
//
// Live serving: voxel_gpu_server_start runs the request loop on a thread of
// its own.  Clients on any thread submit one scan descriptor at a time and
// get a ticket; the loop takes the submitted queries in arrival order into
// the next free engine while their scratch footprint -- one 64-byte partial
// per 8192-row unit plus the result and the 128-byte descriptor -- fits the
// memory budget left by the batches already on the device, launches, and
// keeps forming the next batch while earlier ones run, so requests that
// arrive during a batch join the following one instead of waiting for the
// queue to drain.  Completed batches fill their tickets, which poll / wait
// read.  A query's bits do not depend on its batch (fixed reduction units),
// so a ticket's result equals the same query run alone.
//
// voxel_gpu_server_replay drives the same loop from a recorded arrival
// schedule (mode 0), and for comparison replays it with one synchronous
// voxel_gpu_scan per query (mode 1) or with one single-query launch per
// query queued without waiting and read back through per-query events
// (mode 2).
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_batch.hpp"
#include "vx_runtime.hpp"

using namespace vx;

namespace {

using Clock = std::chrono::steady_clock;

struct Ticket {
    voxel_gpu_scan_desc q{};
    double t_submit = 0.0, t_done = 0.0;   // seconds from the loop's start
    voxel_gpu_agg res{};
    int state = 0;                          // 0 pending, 1 done, 2 failed
    std::string err;
};

struct Server {
    std::vector<std::unique_ptr<BatchRunner>> slots;
    size_t max_batch = 0;
    uint64_t arena = 0;

    // live loop
    std::thread loop;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Ticket> tickets;             // index = ticket id (stable references)
    std::deque<int64_t> submitted;          // not yet taken by the loop
    std::atomic<uint64_t> nsubmit{0};       // submissions so far (read without the lock)
    bool running = false, stopping = false;
    Clock::time_point t0;
    uint64_t batches = 0, max_seen = 0, waits = 0;

    // mode 2 comparator: a stream and scan scratch of its own
    hipStream_t cmp_stream = nullptr;
    Scratch cmp_scr{};

    double now() const { return std::chrono::duration<double>(Clock::now() - t0).count(); }
    ~Server();
};

uint64_t footprint(const voxel_gpu_scan_desc& q) {
    return (units_of(q.n) + 1) * sizeof(ScanOut) + sizeof(BatchQuery);
}

uint64_t query_bytes(const voxel_gpu_scan_desc& q) {
    const uint64_t w = (q.dtype == 0 || q.dtype == 2) ? 8 : 4;
    const bool two = q.b != nullptr && q.op_b != 6;
    return q.n * w * (two ? 2 : 1);
}

void finish_stats(const std::vector<double>& lat, double makespan, uint64_t bytes, uint64_t batches,
                  uint64_t max_seen, uint64_t waits, voxel_gpu_serve_stats* st) {
    if (!st) return;
    std::memset(st, 0, sizeof(*st));
    const size_t n = lat.size();
    st->queries = n;
    st->batches = batches;
    st->makespan_s = makespan;
    st->qps = makespan > 0 ? (double)n / makespan : 0.0;
    st->gbps = makespan > 0 ? (double)bytes / makespan / 1e9 : 0.0;
    st->bytes = bytes;
    st->mean_batch = batches ? (double)n / (double)batches : 0.0;
    st->max_batch_seen = max_seen;
    st->admission_waits = waits;
    if (n) {
        std::vector<double> s = lat;
        std::sort(s.begin(), s.end());
        auto pct = [&](double p) { size_t i = (size_t)(p * (double)(n - 1) + 0.5); return s[std::min(i, n - 1)]; };
        st->p50_s = pct(0.50);
        st->p99_s = pct(0.99);
        st->max_s = s.back();
        double m = 0;
        for (double v : s) m += v;
        st->mean_s = m / (double)n;
    }
}

// The request loop (one thread per running server).
void request_loop(Server* s) {
    const size_t K = s->slots.size();
    std::vector<std::vector<int64_t>> inq(K);
    std::vector<uint64_t> used_by(K, 0);
    std::vector<voxel_gpu_scan_desc> bq;
    std::vector<voxel_gpu_agg> res(s->max_batch);
    std::deque<std::pair<int64_t, voxel_gpu_scan_desc>> pending;
    uint64_t used = 0, seen = 0;
    bool waited_head = false;
    for (;;) {
        bool inflight = false;
        for (size_t k = 0; k < K; ++k) inflight = inflight || !inq[k].empty();
        if (!inflight && pending.empty()) {
            // idle: spin (yielding) for a while -- a sleeping thread takes
            // tens of microseconds to wake -- then sleep on the condition
            const Clock::time_point until = Clock::now() + std::chrono::milliseconds(2);
            while (s->nsubmit.load(std::memory_order_acquire) == seen && Clock::now() < until) {
                std::this_thread::yield();
                if (s->nsubmit.load(std::memory_order_acquire) == seen) {
                    std::lock_guard<std::mutex> lk(s->mu);
                    if (s->stopping) break;
                }
            }
        }
        {
            std::unique_lock<std::mutex> lk(s->mu);
            if (!inflight && pending.empty() && s->submitted.empty()) {
                if (s->stopping) break;
                s->cv_submit.wait_for(lk, std::chrono::microseconds(500),
                                      [&] { return s->stopping || !s->submitted.empty(); });
            }
            seen = s->nsubmit.load(std::memory_order_acquire);
            while (!s->submitted.empty()) {
                const int64_t id = s->submitted.front();
                s->submitted.pop_front();
                pending.emplace_back(id, s->tickets[(size_t)id].q);
            }
        }
        bool progress = false;
        // completions
        for (size_t k = 0; k < K; ++k) {
            if (inq[k].empty() || !s->slots[k]->done()) continue;
            const bool ok = s->slots[k]->wait(res.data());
            const std::string err = ok ? std::string() : std::string(last_error());
            const double tc = s->now();
            {
                std::lock_guard<std::mutex> lk(s->mu);
                for (size_t j = 0; j < inq[k].size(); ++j) {
                    Ticket& t = s->tickets[(size_t)inq[k][j]];
                    t.res = res[j];
                    t.t_done = tc;
                    t.state = ok ? 1 : 2;
                    if (!ok) t.err = err;
                }
            }
            s->cv_done.notify_all();
            used -= used_by[k];
            used_by[k] = 0;
            inq[k].clear();
            progress = true;
        }
        // admissions: the next free engine takes arrived queries in order
        for (size_t k = 0; k < K && !pending.empty(); ++k) {
            if (!inq[k].empty()) continue;
            BatchRunner& r = *s->slots[k];
            bq.clear();
            uint64_t add = 0;
            while (!pending.empty() && bq.size() < s->max_batch) {
                const uint64_t fp = footprint(pending.front().second);
                const bool alone = bq.empty() && used == 0;      // an oversized query runs by itself
                if (used + add + fp > s->arena && !alone) break;
                bq.push_back(pending.front().second);
                inq[k].push_back(pending.front().first);
                pending.pop_front();
                add += fp;
            }
            if (bq.empty()) {
                if (!waited_head) { ++s->waits; waited_head = true; }
                break;
            }
            waited_head = false;
            if (!r.launch(bq.data(), bq.size())) {             // descriptors were checked at submit
                const std::string err = last_error();
                const double tc = s->now();
                {
                    std::lock_guard<std::mutex> lk(s->mu);
                    for (int64_t id : inq[k]) {
                        Ticket& t = s->tickets[(size_t)id];
                        t.state = 2;
                        t.err = err;
                        t.t_done = tc;
                    }
                }
                s->cv_done.notify_all();
                inq[k].clear();
                continue;
            }
            used += add;
            used_by[k] = add;
            ++s->batches;
            s->max_seen = std::max<uint64_t>(s->max_seen, bq.size());
            progress = true;
        }
        if (!progress) std::this_thread::yield();              // batches on the GPU: poll again soon
    }
}

bool start_loop(Server* s) {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->running) { set_error("server: already running"); return false; }
    s->tickets.clear();
    s->submitted.clear();
    s->nsubmit.store(0);
    s->stopping = false;
    s->batches = s->max_seen = s->waits = 0;
    s->t0 = Clock::now();
    s->running = true;
    s->loop = std::thread(request_loop, s);
    return true;
}

bool stop_loop(Server* s, voxel_gpu_serve_stats* stats) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->running) { set_error("server: not running"); return false; }
        s->stopping = true;
    }
    s->cv_submit.notify_all();
    s->loop.join();
    std::lock_guard<std::mutex> lk(s->mu);
    s->running = false;
    std::vector<double> lat;
    uint64_t bytes = 0;
    double first = 0.0, last = 0.0;
    for (size_t i = 0; i < s->tickets.size(); ++i) {
        const Ticket& t = s->tickets[i];
        if (t.state != 1) continue;
        lat.push_back(t.t_done - t.t_submit);
        bytes += query_bytes(t.q);
        if (lat.size() == 1 || t.t_submit < first) first = t.t_submit;
        last = std::max(last, t.t_done);
    }
    finish_stats(lat, lat.empty() ? 0.0 : last - first, bytes, s->batches, s->max_seen, s->waits, stats);
    return true;
}

Server::~Server() {
    if (running) stop_loop(this, nullptr);
    if (cmp_stream) (void)hipStreamDestroy(cmp_stream);
    if (cmp_scr.partials) (void)hipFree(cmp_scr.partials);
}

// Submits n queries that arrive together (at = their arrival times, or NULL
// for now); ids[i] receives their tickets.  Descriptors were checked.
bool submit_n(Server* s, const voxel_gpu_scan_desc* q, const double* at, size_t n, int64_t* ids) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (!s->running || s->stopping) { set_error("server: not running (call voxel_gpu_server_start)"); return false; }
        const double t = s->now();
        for (size_t i = 0; i < n; ++i) {
            ids[i] = (int64_t)s->tickets.size();
            s->tickets.emplace_back();
            Ticket& tk = s->tickets.back();
            tk.q = q[i];
            tk.t_submit = at ? at[i] : t;
            s->submitted.push_back(ids[i]);
        }
        s->nsubmit.fetch_add(n, std::memory_order_release);
    }
    s->cv_submit.notify_one();
    return true;
}

int64_t submit(Server* s, const voxel_gpu_scan_desc* q) {
    ScanArgs g;
    if (!make_scan_args(q, g)) return -1;                      // refused before it reaches the loop
    int64_t id = -1;
    return submit_n(s, q, nullptr, 1, &id) ? id : -1;
}

// Mode 2: one single-query launch per query (k_scan + k_scan_final into the
// query's own result slot, its read-back and an event), queued on the
// comparator stream as the query arrives; completions are read in order
// from the events.
int replay_async(Server* s, const voxel_gpu_scan_desc* q, const double* arrival_s, size_t n, voxel_gpu_agg* out,
                 std::vector<double>& lat, double* makespan) {
    if (!s->cmp_stream) {
        if (!hip_ok(hipStreamCreateWithFlags(&s->cmp_stream, hipStreamNonBlocking), "hipStreamCreate")) return -1;
        s->cmp_scr.max_blocks = (size_t)device_cu_count() * 16;
        if (!hip_ok(hipMalloc(&s->cmp_scr.partials, s->cmp_scr.max_blocks * 64), "hipMalloc(partials)")) return -1;
        s->cmp_scr.ticket = nullptr;
    }
    std::vector<ScanArgs> args(n);
    for (size_t i = 0; i < n; ++i) {
        if (!make_scan_args(&q[i], args[i])) return -1;
        args[i].layout = 0;                                    // split layouts: no ticket
    }
    ScanOut* dres = nullptr;
    ScanOut* hres = nullptr;
    std::vector<hipEvent_t> ev(n, nullptr);
    int rc = -1;
    if (!hip_ok(hipMalloc((void**)&dres, n * sizeof(ScanOut)), "hipMalloc") ||
        !hip_ok(hipHostMalloc((void**)&hres, n * sizeof(ScanOut), hipHostMallocDefault), "hipHostMalloc")) {
        if (dres) (void)hipFree(dres);
        return -1;
    }
    for (size_t i = 0; i < n; ++i)
        if (!hip_ok(hipEventCreateWithFlags(&ev[i], hipEventDisableTiming), "hipEventCreate")) goto out;
    {
        const Clock::time_point t0 = Clock::now();
        auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
        size_t next = 0, done = 0;
        while (done < n) {
            const double t = now();
            while (next < n && arrival_s[next] <= t) {
                const ScanArgs& g = args[next];
                if (g.end > g.begin) {
                    if (!hip_ok(launch_scan(g, s->cmp_scr, &dres[next], s->cmp_stream), "scan launch")) goto out;
                } else if (!hip_ok(hipMemsetAsync(&dres[next], 0, sizeof(ScanOut), s->cmp_stream), "clear")) {
                    goto out;
                }
                if (!hip_ok(hipMemcpyAsync(&hres[next], &dres[next], sizeof(ScanOut), hipMemcpyDeviceToHost,
                                           s->cmp_stream), "read back") ||
                    !hip_ok(hipEventRecord(ev[next], s->cmp_stream), "hipEventRecord"))
                    goto out;
                ++next;
            }
            bool any = false;
            while (done < next && hipEventQuery(ev[done]) == hipSuccess) {
                lat[done] = now() - arrival_s[done];
                fill_agg(q[done].dtype, hres[done], &out[done]);
                ++done;
                any = true;
            }
            if (!any && next < n && done == next) {
                const double wait = arrival_s[next] - now();
                if (wait > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
            } else if (!any) {
                std::this_thread::yield();
            }
        }
        *makespan = now() - arrival_s[0];
        rc = 0;
    }
out:
    if (rc != 0) (void)hipStreamSynchronize(s->cmp_stream);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    (void)hipFree(dres);
    (void)hipHostFree(hres);
    return rc;
}

}  // namespace

extern "C" {

void* voxel_gpu_server_create(size_t max_batch, uint64_t memory_budget_bytes, int inflight) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    if (inflight < 1 || inflight > 8) { set_error("server: inflight must be 1..8"); return nullptr; }
    auto* s = new Server();
    s->max_batch = max_batch ? max_batch : 1024;
    s->arena = memory_budget_bytes ? memory_budget_bytes : (64ull << 20);
    for (int i = 0; i < inflight; ++i) {
        s->slots.emplace_back(new BatchRunner(s->max_batch));
        if (!s->slots.back()->ok()) { delete s; return nullptr; }
    }
    return s;
}

void voxel_gpu_server_destroy(void* h) { delete static_cast<Server*>(h); }

// Same tile/unroll/grid configuration on every engine (a query's bits do not
// depend on it).
int voxel_gpu_server_config(void* h, int unroll, uint32_t tile_rows, int blocks_per_cu) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->running) { set_error("server: stop it before changing the configuration"); return -1; }
    }
    for (auto& r : s->slots) {
        BatchConfig c = r->config();
        if (unroll) c.unroll = unroll;
        if (tile_rows) c.tile_rows = tile_rows;
        if (blocks_per_cu) c.blocks_per_cu = blocks_per_cu;
        if (!r->set_config(c)) return -1;
    }
    return 0;
}

int voxel_gpu_server_start(void* h) {
    set_error("%s", "");
    return start_loop(static_cast<Server*>(h)) ? 0 : -1;
}

int64_t voxel_gpu_server_submit(void* h, const voxel_gpu_scan_desc* q) {
    set_error("%s", "");
    return submit(static_cast<Server*>(h), q);
}

int voxel_gpu_server_poll(void* h, int64_t ticket, voxel_gpu_agg* out) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    std::lock_guard<std::mutex> lk(s->mu);
    if (ticket < 0 || (size_t)ticket >= s->tickets.size()) { set_error("server: unknown ticket %lld", (long long)ticket); return -1; }
    const Ticket& t = s->tickets[(size_t)ticket];
    if (t.state == 0) return 0;
    if (t.state == 2) { set_error("server: query failed: %s", t.err.c_str()); return -1; }
    if (out) *out = t.res;
    return 1;
}

int voxel_gpu_server_wait(void* h, int64_t ticket, voxel_gpu_agg* out, double* latency_s) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket < 0 || (size_t)ticket >= s->tickets.size()) { set_error("server: unknown ticket %lld", (long long)ticket); return -1; }
    s->cv_done.wait(lk, [&] { return s->tickets[(size_t)ticket].state != 0; });
    const Ticket& t = s->tickets[(size_t)ticket];
    if (t.state == 2) { set_error("server: query failed: %s", t.err.c_str()); return -1; }
    if (out) *out = t.res;
    if (latency_s) *latency_s = t.t_done - t.t_submit;
    return 0;
}

int voxel_gpu_server_stop(void* h, voxel_gpu_serve_stats* stats) {
    set_error("%s", "");
    return stop_loop(static_cast<Server*>(h), stats) ? 0 : -1;
}

int voxel_gpu_server_replay(void* h, const voxel_gpu_scan_desc* q, const double* arrival_s, size_t n, int mode,
                            voxel_gpu_agg* out, double* latency_s, voxel_gpu_serve_stats* stats) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    if (mode < 0 || mode > 2) { set_error("server: mode must be 0 (batched), 1 (synchronous) or 2 (asynchronous)"); return -1; }
    if (n == 0) { finish_stats({}, 0, 0, 0, 0, 0, stats); return 0; }
    for (size_t i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) { set_error("server: arrivals must be non-decreasing"); return -1; }
    for (size_t i = 0; i < n; ++i) {
        ScanArgs g;
        if (!make_scan_args(&q[i], g)) return -1;
    }
    uint64_t bytes = 0;
    for (size_t i = 0; i < n; ++i) bytes += query_bytes(q[i]);
    std::vector<double> lat(n, 0.0);

    if (mode == 2) {
        double makespan = 0.0;
        if (replay_async(s, q, arrival_s, n, out, lat, &makespan) != 0) return -1;
        finish_stats(lat, makespan, bytes, n, 1, 0, stats);
        if (latency_s) std::copy(lat.begin(), lat.end(), latency_s);
        return 0;
    }
    if (mode == 1) {                          // one synchronous single-query launch per query
        const Clock::time_point t0 = Clock::now();
        auto now = [&]() { return std::chrono::duration<double>(Clock::now() - t0).count(); };
        for (size_t i = 0; i < n; ++i) {
            while (now() < arrival_s[i])
                if (arrival_s[i] - now() > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
            if (voxel_gpu_scan(&q[i], nullptr, &out[i]) != 0) return -1;
            lat[i] = now() - arrival_s[i];
        }
        finish_stats(lat, now() - arrival_s[0], bytes, n, 1, 0, stats);
        if (latency_s) std::copy(lat.begin(), lat.end(), latency_s);
        return 0;
    }

    // mode 0: the live request loop, fed from this thread at the arrivals;
    // queries whose arrival time has come enter together (latency is
    // measured from the scheduled arrival)
    if (!start_loop(s)) return -1;
    std::vector<int64_t> ids(n, -1);
    int rc = 0;
    for (size_t i = 0; i < n && rc == 0;) {
        while (s->now() < arrival_s[i])
            if (arrival_s[i] - s->now() > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
        const double t = s->now();
        size_t j = i + 1;
        while (j < n && arrival_s[j] <= t) ++j;
        if (!submit_n(s, &q[i], &arrival_s[i], j - i, &ids[i])) rc = -1;
        i = j;
    }
    for (size_t i = 0; i < n && rc == 0; ++i) {
        if (voxel_gpu_server_wait(s, ids[i], &out[i], &lat[i]) != 0) rc = -1;
    }
    voxel_gpu_serve_stats st{};
    const std::string err = rc ? std::string(last_error()) : std::string();
    if (!stop_loop(s, &st)) return -1;
    if (rc) { set_error("%s", err.c_str()); return -1; }
    st.bytes = bytes;
    st.gbps = st.makespan_s > 0 ? (double)bytes / st.makespan_s / 1e9 : 0.0;
    if (stats) *stats = st;
    if (latency_s) std::copy(lat.begin(), lat.end(), latency_s);
    return 0;
}

// Autotunes engine 0 on the given queries (voxel_gpu_batch_autotune) and
// applies the chosen configuration to every engine.
int voxel_gpu_server_autotune(void* h, const voxel_gpu_scan_desc* q, size_t nq, int reps, double* best_ms,
                              int* configs_tried) {
    Server* s = static_cast<Server*>(h);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->running) { set_error("server: stop it before autotuning"); return -1; }
    }
    if (voxel_gpu_batch_autotune(s->slots[0].get(), q, nq, reps, best_ms, configs_tried) != 0) return -1;
    const BatchConfig& c = s->slots[0]->config();
    return voxel_gpu_server_config(h, c.unroll, c.tile_rows, c.blocks_per_cu);
}

}  // extern "C"
