// vx_serve.cpp -- continuous-batching query server over resident columns.
//
// No counterpart in the reference (its queries are one Engine::Run each,
// exec/engine.hpp:109-141).  The server owns `inflight` batch engines
// (BatchRunner, vx_batch.hip: own stream, persistent kernel, graph launch).
// A request loop admits arrived queries in arrival order into the next free
// engine while their scratch footprint -- one 64-byte partial per tile plus
// the result and the 128-byte descriptor -- fits the memory budget left by
// the batches already on the device, launches, and keeps forming the next
// batch while earlier ones run, so requests that arrive during a batch join
// the following one instead of waiting for the queue to drain.
//
// voxel_gpu_server_replay drives the loop from a recorded arrival schedule
// (seconds from the start) and reports per-query latency (arrival ->
// result on the host) and throughput.  Mode 1 replays the same schedule with
// one voxel_gpu_scan launch per query (single-query kernel, synchronous) as
// the unbatched reference point.
#include <hip/hip_runtime_api.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <memory>
#include <thread>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_batch.hpp"
#include "vx_runtime.hpp"

using namespace vx;

namespace {

struct Server {
    std::vector<std::unique_ptr<BatchRunner>> slots;
    size_t max_batch = 0;
    uint64_t arena = 0;
};

uint64_t footprint(const BatchRunner& r, const voxel_gpu_scan_desc& q) {
    return (r.tiles_of(q.n) + 1) * sizeof(ScanOut) + sizeof(BatchQuery);
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

// Same tile/unroll/grid configuration on every engine (results of a query
// depend on tile_rows only; keep it fixed while serving).
int voxel_gpu_server_config(void* h, int unroll, uint32_t tile_rows, int blocks_per_cu) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    for (auto& r : s->slots) {
        BatchConfig c = r->config();
        if (unroll) c.unroll = unroll;
        if (tile_rows) c.tile_rows = tile_rows;
        if (blocks_per_cu) c.blocks_per_cu = blocks_per_cu;
        if (!r->set_config(c)) return -1;
    }
    return 0;
}

int voxel_gpu_server_replay(void* h, const voxel_gpu_scan_desc* q, const double* arrival_s, size_t n, int mode,
                            voxel_gpu_agg* out, double* latency_s, voxel_gpu_serve_stats* stats) {
    set_error("%s", "");
    Server* s = static_cast<Server*>(h);
    if (n == 0) { finish_stats({}, 0, 0, 0, 0, 0, stats); return 0; }
    for (size_t i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) { set_error("server: arrivals must be non-decreasing"); return -1; }
    uint64_t bytes = 0;
    for (size_t i = 0; i < n; ++i) bytes += query_bytes(q[i]);
    std::vector<double> lat(n, 0.0);
    using clk = std::chrono::steady_clock;
    const auto t0 = clk::now();
    auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
    auto wait_until = [&](double t) {
        while (now() < t) {
            if (t - now() > 200e-6) std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
    };

    if (mode == 1) {                          // one synchronous single-query launch per query
        for (size_t i = 0; i < n; ++i) {
            wait_until(arrival_s[i]);
            if (voxel_gpu_scan(&q[i], nullptr, &out[i]) != 0) return -1;
            lat[i] = now() - arrival_s[i];
        }
        const double ms = now() - arrival_s[0];
        finish_stats(lat, ms, bytes, n, 1, 0, stats);
        if (latency_s) std::copy(lat.begin(), lat.end(), latency_s);
        return 0;
    }
    if (mode != 0) { set_error("server: mode must be 0 (batched) or 1 (one launch per query)"); return -1; }

    const size_t K = s->slots.size();
    std::vector<std::vector<size_t>> inq(K);
    std::vector<uint64_t> used_by(K, 0);
    std::vector<voxel_gpu_scan_desc> bq;
    std::vector<voxel_gpu_agg> res(s->max_batch);
    std::deque<size_t> pending;
    size_t next = 0, done = 0;
    uint64_t used = 0, batches = 0, max_seen = 0, waits = 0;
    bool waited_head = false;
    while (done < n) {
        bool progress = false;
        const double t = now();
        while (next < n && arrival_s[next] <= t) pending.push_back(next++);
        for (size_t k = 0; k < K; ++k) {
            if (inq[k].empty() || !s->slots[k]->done()) continue;
            if (!s->slots[k]->wait(res.data())) return -1;
            const double tc = now();
            for (size_t j = 0; j < inq[k].size(); ++j) {
                const size_t qi = inq[k][j];
                out[qi] = res[j];
                lat[qi] = tc - arrival_s[qi];
            }
            done += inq[k].size();
            used -= used_by[k];
            used_by[k] = 0;
            inq[k].clear();
            progress = true;
        }
        for (size_t k = 0; k < K && !pending.empty(); ++k) {
            if (!inq[k].empty()) continue;
            BatchRunner& r = *s->slots[k];
            bq.clear();
            uint64_t add = 0;
            while (!pending.empty() && bq.size() < s->max_batch) {
                const uint64_t fp = footprint(r, q[pending.front()]);
                const bool alone = bq.empty() && used == 0;      // an oversized query runs by itself
                if (used + add + fp > s->arena && !alone) break;
                bq.push_back(q[pending.front()]);
                inq[k].push_back(pending.front());
                pending.pop_front();
                add += fp;
            }
            if (bq.empty()) {
                if (!waited_head) { ++waits; waited_head = true; }
                break;
            }
            waited_head = false;
            if (!r.launch(bq.data(), bq.size())) return -1;
            used += add;
            used_by[k] = add;
            ++batches;
            max_seen = std::max<uint64_t>(max_seen, bq.size());
            progress = true;
        }
        if (!progress && pending.empty() && next < n) {
            bool busy = false;
            for (size_t k = 0; k < K; ++k) busy = busy || !inq[k].empty();
            if (!busy) wait_until(arrival_s[next]);
        }
    }
    const double makespan = now() - arrival_s[0];
    finish_stats(lat, makespan, bytes, batches, max_seen, waits, stats);
    if (latency_s) std::copy(lat.begin(), lat.end(), latency_s);
    return 0;
}

// Autotunes engine 0 on the given queries (voxel_gpu_batch_autotune) and
// applies the chosen configuration to every engine.
int voxel_gpu_server_autotune(void* h, const voxel_gpu_scan_desc* q, size_t nq, int reps, double* best_ms,
                              int* configs_tried) {
    Server* s = static_cast<Server*>(h);
    if (voxel_gpu_batch_autotune(s->slots[0].get(), q, nq, reps, best_ms, configs_tried) != 0) return -1;
    const BatchConfig& c = s->slots[0]->config();
    return voxel_gpu_server_config(h, c.unroll, c.tile_rows, c.blocks_per_cu);
}

}  // extern "C"
