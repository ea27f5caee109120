// vx_batch.hpp -- batched filter/aggregate queries: many scan descriptors
// (voxel_gpu_scan_desc) executed by one persistent kernel launch, driven
// through a HIP graph (descriptor upload, kernel, result download).  Host
// side of vx_batch.hip, used by the C ABI there and by the serving loop
// (vx_serve.cpp).
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"

namespace vx {

// Queries of one batch that read the same rows (same column(s), length and
// element type) share a scan: one job loads a tile once and feeds up to
// kBatchShare queries' predicates and accumulators.
constexpr int kBatchShare = 4;

// A query's rows are reduced in units of kUnitRows rows counted from its
// first row, one wave per unit: one compensated partial per unit, merged in
// unit order.  Chunks (the rows a wave takes at a time, tile_rows) are whole
// units, so a query's bits depend on its rows only -- not on the chunk size,
// the batch it runs in or the configuration.
constexpr uint32_t kUnitRows = 8192;

inline uint64_t units_of(uint64_t n) {
    const uint64_t u = (n + kUnitRows - 1) / kUnitRows;
    return u ? u : 1;
}

// Device descriptor of one query (128 bytes).
struct BatchQuery {
    const void* a;
    const void* b;
    uint64_t n;
    uint64_t unit0;          // first unit partial of the query
    PredDev pa, pb;
    int32_t dtype, mode;     // mode 0: pa on a; 2: pa on a and pb on b
    uint32_t nunits, orig;   // unit partials; index of the query in the caller's list
    uint64_t pad2[2];
};
static_assert(sizeof(BatchQuery) == 128, "BatchQuery layout");

// One shared scan (64 bytes): its rows, tiles and member queries.
struct BatchJob {
    const void* a;
    const void* b;
    uint64_t n;
    uint64_t tile0;          // first global tile of the job
    int32_t dtype, mode;
    uint32_t ntiles, nmem;
    uint32_t mem[kBatchShare];   // members: indices into the query records
};
static_assert(sizeof(BatchJob) == 64, "BatchJob layout");

// Header in front of the descriptors (64 bytes); the job records follow the
// nq query records.
struct BatchHeader {
    uint32_t nq, total_tiles, tile_rows, njobs;
    ScanOut* parts;          // one partial per query and tile
    ScanOut* results;        // one result per query
    const BatchJob* jobs;
    uint64_t pad2[3];
};
static_assert(sizeof(BatchHeader) == 64, "BatchHeader layout");

struct BatchConfig {
    int unroll = 4;              // 16-byte loads in flight per lane and column: 2, 4, 8
    uint32_t tile_rows = 65536;  // largest rows per chunk (multiple of kUnitRows)
    int blocks_per_cu = 2;       // resident workgroups per CU
};

// One batch engine: its own stream, pinned descriptor/result buffers, tile
// partials and graph.  Not thread-safe; one launch in flight at a time.
class BatchRunner {
  public:
    explicit BatchRunner(size_t max_queries);
    ~BatchRunner();
    bool ok() const { return ok_; }
    size_t max_queries() const { return max_q_; }
    const BatchConfig& config() const { return cfg_; }
    bool set_config(const BatchConfig& c);
    // Rows of a chunk (the rows one wave takes at a time) in a launch of
    // `units` unit partials on `grid` workgroups: the configured tile_rows,
    // made smaller (down to one unit) while the batch would give fewer than
    // two chunks per wave.
    uint32_t tile_rows_for(uint64_t units, uint64_t grid) const;
    uint32_t last_tile_rows() const { return last_tile_rows_; }
    // Validates and enqueues nq queries; false (error set) on bad input.
    bool launch(const voxel_gpu_scan_desc* q, size_t nq);
    bool done();                                   // non-blocking
    bool wait(voxel_gpu_agg* out);                 // blocks, fills nq results
    float last_ms();                               // kernel + copies of the last batch
    hipStream_t stream() const { return st_; }

  private:
    bool ensure_parts(uint64_t units);
    bool build_graph();
    size_t max_q_;
    bool ok_ = false;
    BatchConfig cfg_;
    hipStream_t st_ = nullptr;
    char* hq_ = nullptr;           // pinned: header + descriptors
    char* dq_ = nullptr;           // device copy
    ScanOut* hres_ = nullptr;      // pinned results
    ScanOut* dres_ = nullptr;
    ScanOut* parts_ = nullptr;
    uint64_t parts_cap_ = 0;
    uint32_t* ctr_ = nullptr;      // [0] next tile, [1] exited workgroups, [2..] per-query tile tickets
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
    hipGraphNode_t n_up_ = nullptr, n_kern_ = nullptr, n_down_ = nullptr;
    hipEvent_t ev0_ = nullptr, ev1_ = nullptr;
    std::vector<int> dtypes_;
    std::vector<uint32_t> order_;
    std::vector<BatchJob> jobs_;
    size_t nq_ = 0;
    uint32_t last_tile_rows_ = 0;
    bool inflight_ = false;
    bool graph_valid_ = false;
};

}  // namespace vx
