// vx_batch.hip -- many filter/aggregate queries in one persistent kernel.
//
// The reference runs one program per Engine::Run call (exec/engine.hpp:
// 109-141, one column, one predicate); a batch here is a list of scan
// descriptors (the voxel_gpu_scan descriptor: column(s), dtype, predicate(s))
// over columns of any length.  Queries that read the same rows (same
// column(s), length and element type) are grouped into jobs of up to
// kBatchShare queries: a job loads each row once and feeds every member's
// predicate and accumulator (a shared scan).  Every job is cut into chunks
// of tile_rows rows (whole units of kUnitRows); the chunks of all jobs form
// one global list.  Resident waves take chunks round-robin (wave g: chunks
// g, g + waves, ...) and reduce each unit of a chunk with the scan's
// compensated COUNT/SUM/MIN/MAX accumulator (vx_acc.hpp), one per member,
// then a wave reduction, storing one partial per member and unit; a
// per-query pass after the kernel boundary (k_batch_finish) combines each
// query's partials in unit order.  A query's accumulator sees the same rows
// of a unit in the same lane order whatever the chunk size, the unroll, the
// grid or the other members of its job, so its result depends only on its
// own rows: it is the same bits alone, in any mix and under any
// configuration, and each launch may pick its own chunk size.
//
// Host side: BatchRunner keeps pinned descriptor/result buffers and a graph
// of four nodes -- descriptor upload, tile kernel, finish kernel, result
// download -- whose copy sizes are updated per launch (no memset node); a
// batch of at most kSmallQ queries instead runs as two kernel launches with
// their descriptors as arguments (k_batch_small, k_batch_small_finish).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_acc.hpp"
#include "vx_batch.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

// Rows [s, e) of a job, one unit (local unit index lu = s / kUnitRows), run
// by one wave: each of the job's NM members' predicate(s) (TWO: a predicate
// on column b too) and accumulator on every row loaded, lane l taking the
// 16-byte vectors l, l + 64, ... of the unit in order; then a wave
// reduction (shuffles, no barrier) and lane 0 stores member m's partial in
// its slot unit0 + lu.  NM and TWO are template parameters, so the row loop
// carries no runtime branch.
template <typename T, int U, bool NT, int NM, bool TWO>
__device__ __forceinline__ void wave_unit(const BatchJob& job, const BatchQuery* __restrict__ qs, uint64_t s,
                                          uint64_t e, uint64_t lu, ScanOut* __restrict__ parts) {
    using V = typename TyInfo<T>::V;
    constexpr int E = TyInfo<T>::E;
    const uint32_t lane = threadIdx.x & 63;
    Pred<T, false> pa[NM], pb[NM];
    Acc<T> acc[NM];
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        acc[m].init();
        pa[m].load(qs[job.mem[m]].pa);
        if constexpr (TWO) pb[m].load(qs[job.mem[m]].pb);
    }
    const T* __restrict__ a = reinterpret_cast<const T*>(job.a);
    const T* __restrict__ b = reinterpret_cast<const T*>(TWO ? job.b : job.a);
    auto one = [&](T x, T y) {
#pragma unroll
        for (int m = 0; m < NM; ++m) {
            bool p = pa[m](x);
            if constexpr (TWO) p = p & pb[m](y);
            acc[m].add(p, x);
        }
    };
    const bool aligned = ((reinterpret_cast<uintptr_t>(a + s) | reinterpret_cast<uintptr_t>(b + s)) & 15u) == 0;
    const uint64_t nvec = aligned ? (e - s) / E : 0;
    const V* __restrict__ av = reinterpret_cast<const V*>(a + s);
    const V* __restrict__ bv = reinterpret_cast<const V*>(b + s);
    for (uint64_t base = 0; base < nvec; base += 64ull * U) {
        V xa[U], xb[U];
        if (base + 64ull * U <= nvec) {
#pragma unroll
            for (int k = 0; k < U; ++k) xa[k] = ld16<V, NT>(&av[base + k * 64 + lane]);
            if constexpr (TWO) {
#pragma unroll
                for (int k = 0; k < U; ++k) xb[k] = ld16<V, NT>(&bv[base + k * 64 + lane]);
            }
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const T* pxa = reinterpret_cast<const T*>(&xa[k]);
                const T* pxb = reinterpret_cast<const T*>(&xb[k]);
#pragma unroll
                for (int r = 0; r < E; ++r) one(pxa[r], TWO ? pxb[r] : T(0));
            }
        } else {
#pragma unroll
            for (int k = 0; k < U; ++k) {
                const uint64_t vi = base + k * 64 + lane;
                if (vi < nvec) {
                    xa[k] = av[vi];
                    if constexpr (TWO) xb[k] = bv[vi];
                    const T* pxa = reinterpret_cast<const T*>(&xa[k]);
                    const T* pxb = reinterpret_cast<const T*>(&xb[k]);
#pragma unroll
                    for (int r = 0; r < E; ++r) one(pxa[r], TWO ? pxb[r] : T(0));
                }
            }
        }
    }
    for (uint64_t i = s + nvec * E + lane; i < e; i += 64) one(a[i], TWO ? b[i] : T(0));
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        acc[m].fold();
        for (int mask = 32; mask > 0; mask >>= 1) acc[m].shfl_merge(mask);
        if (lane == 0) acc[m].store(&parts[qs[job.mem[m]].unit0 + lu]);
    }
}

template <typename T, int U, bool NT, int NM, bool TWO>
__device__ __forceinline__ void wave_units(const BatchJob& job, const BatchQuery* __restrict__ qs, uint64_t s,
                                           uint64_t e, ScanOut* __restrict__ parts) {
    uint64_t us = s;
    do {
        const uint64_t ue = us + kUnitRows < e ? us + kUnitRows : e;
        wave_unit<T, U, NT, NM, TWO>(job, qs, us, ue, us / kUnitRows, parts);
        us = ue;
    } while (us < e);
}

// One chunk of a job (rows [s, e), whole units; an empty query has one
// empty unit) run by one wave, with the row loop specialised on the job's
// member count and mode.
template <typename T, int U, bool NT>
__device__ __forceinline__ void wave_chunk(const BatchJob& job, const BatchQuery* __restrict__ qs, uint64_t s,
                                           uint64_t e, ScanOut* __restrict__ parts) {
    const bool two = job.mode == 2;
    // one member on one column: twice the loads in flight (its accumulator
    // leaves the registers for them)
    constexpr int U1 = U * 2 < 16 ? U * 2 : 16;
    switch (job.nmem * 2 + (two ? 1 : 0)) {
        case 2: wave_units<T, U1, NT, 1, false>(job, qs, s, e, parts); break;
        case 3: wave_units<T, U, NT, 1, true>(job, qs, s, e, parts); break;
        case 4: wave_units<T, U, NT, 2, false>(job, qs, s, e, parts); break;
        case 5: wave_units<T, U, NT, 2, true>(job, qs, s, e, parts); break;
        case 6: wave_units<T, U, NT, 3, false>(job, qs, s, e, parts); break;
        case 7: wave_units<T, U, NT, 3, true>(job, qs, s, e, parts); break;
        case 8: wave_units<T, U, NT, 4, false>(job, qs, s, e, parts); break;
        default: wave_units<T, U, NT, 4, true>(job, qs, s, e, parts); break;
    }
}

// Chunks of all jobs (tile_rows rows each, whole units; the list runs largest
// job first): wave g of the grid takes chunks g, g + waves, ..., where g
// counts the first wave of every workgroup before any second wave (so a
// short list spreads over all CUs) -- no shared counter, whose atomics
// serialise on one address, and no barrier: each wave reduces its own
// units.  Job of a chunk: the last job with tile0 <= chunk;
// a wave's chunks increase, so it steps forward from the previous one and
// jumps by binary search when far.
template <int U, bool NT>
__device__ __forceinline__ void run_chunks(const BatchJob* __restrict__ jobs, uint32_t nj, uint32_t total,
                                           uint32_t tile_rows, const BatchQuery* __restrict__ qs,
                                           ScanOut* __restrict__ parts) {
    constexpr uint32_t kWaves = kBlock / 64;
    const uint32_t waves = gridDim.x * kWaves;
    uint32_t lo = 0;
    for (uint32_t c = (threadIdx.x >> 6) * gridDim.x + blockIdx.x; c < total; c += waves) {
        if (lo + 1 < nj && jobs[lo + 1].tile0 <= c) {
            if (lo + 2 < nj && jobs[lo + 2].tile0 <= c) {
                uint32_t a = lo + 2, z = nj;
                while (z - a > 1) {
                    const uint32_t mid = (a + z) >> 1;
                    if (jobs[mid].tile0 <= c) a = mid; else z = mid;
                }
                lo = a;
            } else {
                lo = lo + 1;
            }
        }
        const BatchJob& job = jobs[lo];
        const uint64_t s = (uint64_t)(c - job.tile0) * tile_rows;
        const uint64_t e = s + tile_rows < job.n ? s + tile_rows : job.n;
        switch (job.dtype) {
            case DT_F64: wave_chunk<double, U, NT>(job, qs, s, e, parts); break;
            case DT_F32: wave_chunk<float, U, NT>(job, qs, s, e, parts); break;
            case DT_I64: wave_chunk<int64_t, U, NT>(job, qs, s, e, parts); break;
            default: wave_chunk<int32_t, U, NT>(job, qs, s, e, parts); break;
        }
    }
}

// The chunk pass of a batch.  No wave waits on or reads another's partial
// here (that would need an agent-scope release and acquire per chunk: an L2
// write-back and invalidate on this multi-XCD part); k_batch_finish combines
// them after the kernel boundary.
template <int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_batch(const BatchHeader* __restrict__ hdr, const BatchQuery* __restrict__ qs,
                                                  uint32_t* __restrict__ ctr) {
    (void)ctr;
    run_chunks<U, NT>(hdr->jobs, hdr->njobs, hdr->total_tiles, hdr->tile_rows, qs, hdr->parts);
}

template <typename T>
__device__ __forceinline__ void finish_query(const BatchQuery& q, const ScanOut* __restrict__ parts,
                                             ScanOut* __restrict__ out) {
    __shared__ Acc<T> lds[kBlock / 64];
    Acc<T> r; r.init();
    for (uint32_t i = threadIdx.x; i < q.nunits; i += kBlock) {
        Acc<T> o; o.load(&parts[q.unit0 + i]);
        r.merge(o);
    }
    block_reduce(r, lds);
    if (threadIdx.x == 0) r.store(out);
    __syncthreads();
}

// One workgroup per query (grid-stride): the query's unit partials in a
// fixed order -> its result, in the caller's order.
__global__ __launch_bounds__(kBlock) void k_batch_finish(const BatchHeader* __restrict__ hdr,
                                                         const BatchQuery* __restrict__ qs, uint32_t* __restrict__ ctr) {
    (void)ctr;
    const uint32_t nq = hdr->nq;
    for (uint32_t qi = blockIdx.x; qi < nq; qi += gridDim.x) {
        const BatchQuery q = qs[qi];
        ScanOut* out = &hdr->results[q.orig];
        switch (q.dtype) {
            case DT_F64: finish_query<double>(q, hdr->parts, out); break;
            case DT_F32: finish_query<float>(q, hdr->parts, out); break;
            case DT_I64: finish_query<int64_t>(q, hdr->parts, out); break;
            default: finish_query<int32_t>(q, hdr->parts, out); break;
        }
    }
}

// Batches of at most kSmallQ queries: no graph, no descriptor upload.  The
// header, query and job records travel as the kernel's first argument (the
// kernarg segment, read through its segment pointer and copied to LDS by
// each workgroup); waves take chunks like k_batch at the deepest unroll, and
// a second kernel merges each query's unit partials with the same code as
// k_batch_finish -- the same bits -- and writes the results straight into
// pinned host memory.  (A last-workgroup ticket in one kernel needs an
// agent-scope release / acquire in every workgroup, an L2 write-back on this
// multi-XCD part, which cost more than the second launch.)
constexpr uint32_t kSmallQ = 8;

struct SmallArgs {
    BatchHeader h;
    BatchQuery q[kSmallQ];
    BatchJob j[kSmallQ];
};
static_assert(sizeof(SmallArgs) == 64 + kSmallQ * (128 + 64), "SmallArgs layout");
static_assert(sizeof(SmallArgs) <= 2048, "SmallArgs must stay a small kernel argument");

template <int U, bool NT>
__global__ __launch_bounds__(kBlock) void k_batch_small(const SmallArgs args, ScanOut* __restrict__ parts) {
    __shared__ SmallArgs sa;
    {
        // args is the first explicit argument: offset 0 of the kernarg segment
        const uint4* src = (const uint4*)__builtin_amdgcn_kernarg_segment_ptr();
        uint4* dst = reinterpret_cast<uint4*>(&sa);
        const uint32_t nq = src[0].x, nj = src[0].w;            // BatchHeader.nq, .njobs
        const uint32_t wq = 4 + nq * 8;
        for (uint32_t i = threadIdx.x; i < 4 + (uint32_t)kSmallQ * 8 + nj * 4; i += kBlock) {
            if (i >= wq && i < 4 + (uint32_t)kSmallQ * 8) continue;     // unused query records
            dst[i] = src[i];
        }
    }
    __syncthreads();
    run_chunks<U, NT>(sa.j, sa.h.njobs, sa.h.total_tiles, sa.h.tile_rows, sa.q, parts);
}

// The small batch's queries (at most kSmallQ), as a kernel argument.
struct SmallQueries {
    BatchQuery q[kSmallQ];
};

// After k_batch_small, at the kernel boundary (no release / acquire per
// workgroup): workgroup b merges query b's unit partials in unit order and
// writes its result to the host-visible slot of its caller index.
__global__ __launch_bounds__(kBlock) void k_batch_small_finish(const SmallQueries sq,
                                                               const ScanOut* __restrict__ parts,
                                                               ScanOut* __restrict__ hres) {
    const BatchQuery& q = sq.q[blockIdx.x];
    switch (q.dtype) {
        case DT_F64: finish_query<double>(q, parts, &hres[q.orig]); break;
        case DT_F32: finish_query<float>(q, parts, &hres[q.orig]); break;
        case DT_I64: finish_query<int64_t>(q, parts, &hres[q.orig]); break;
        default: finish_query<int32_t>(q, parts, &hres[q.orig]); break;
    }
}

void* batch_kernel(int unroll) {
    switch (unroll) {
        case 2: return reinterpret_cast<void*>(&k_batch<2, true>);
        case 8: return reinterpret_cast<void*>(&k_batch<8, true>);
        default: return reinterpret_cast<void*>(&k_batch<4, true>);
    }
}

}  // namespace

BatchRunner::BatchRunner(size_t max_queries) : max_q_(max_queries ? max_queries : 1) {
    if (!runtime_ok()) return;
    const size_t qbytes = sizeof(BatchHeader) + max_q_ * (sizeof(BatchQuery) + sizeof(BatchJob));
    ok_ = hip_ok(hipStreamCreateWithFlags(&st_, hipStreamNonBlocking), "hipStreamCreate") &&
          hip_ok(hipHostMalloc((void**)&hq_, qbytes, hipHostMallocDefault), "hipHostMalloc") &&
          hip_ok(hipMalloc((void**)&dq_, qbytes), "hipMalloc") &&
          hip_ok(hipHostMalloc((void**)&hres_, max_q_ * sizeof(ScanOut), hipHostMallocDefault), "hipHostMalloc") &&
          hip_ok(hipMalloc((void**)&dres_, max_q_ * sizeof(ScanOut)), "hipMalloc") &&
          hip_ok(hipMalloc((void**)&ctr_, (max_q_ + 2) * 4), "hipMalloc") &&
          hip_ok(hipMemset(ctr_, 0, (max_q_ + 2) * 4), "hipMemset") &&
          hip_ok(hipEventCreate(&ev0_), "hipEventCreate") && hip_ok(hipEventCreate(&ev1_), "hipEventCreate") &&
          ensure_parts(4096);
    dtypes_.assign(max_q_, 0);
}

BatchRunner::~BatchRunner() {
    if (inflight_) (void)hipStreamSynchronize(st_);
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    if (ev0_) (void)hipEventDestroy(ev0_);
    if (ev1_) (void)hipEventDestroy(ev1_);
    if (hq_) (void)hipHostFree(hq_);
    if (hres_) (void)hipHostFree(hres_);
    if (dq_) (void)hipFree(dq_);
    if (dres_) (void)hipFree(dres_);
    if (parts_) (void)hipFree(parts_);
    if (ctr_) (void)hipFree(ctr_);
    if (st_) (void)hipStreamDestroy(st_);
}

bool BatchRunner::set_config(const BatchConfig& c) {
    if (c.unroll != 2 && c.unroll != 4 && c.unroll != 8) { set_error("batch: unroll must be 2, 4 or 8"); return false; }
    if (c.tile_rows < kUnitRows || c.tile_rows % kUnitRows != 0) {
        set_error("batch: tile_rows must be a multiple of %u", kUnitRows);
        return false;
    }
    if (c.blocks_per_cu < 1 || c.blocks_per_cu > 8) { set_error("batch: blocks_per_cu must be 1..8"); return false; }
    if (inflight_) { set_error("batch: a launch is in flight"); return false; }
    if (c.unroll != cfg_.unroll || c.blocks_per_cu != cfg_.blocks_per_cu) graph_valid_ = false;
    cfg_ = c;
    return true;
}

uint32_t BatchRunner::tile_rows_for(uint64_t units, uint64_t grid) const {
    const uint64_t waves = grid * (kBlock / 64);
    const uint64_t max_units = cfg_.tile_rows / kUnitRows;
    uint64_t tu = units / (2 * waves);
    if (tu < 1) tu = 1;
    if (tu > max_units) tu = max_units;
    return (uint32_t)(tu * kUnitRows);
}

bool BatchRunner::ensure_parts(uint64_t units) {
    if (units <= parts_cap_) return true;
    uint64_t cap = std::max<uint64_t>(units, parts_cap_ * 2);
    if (parts_) (void)hipFree(parts_);
    parts_ = nullptr;
    parts_cap_ = 0;
    if (!hip_ok(hipMalloc((void**)&parts_, cap * sizeof(ScanOut)), "hipMalloc(batch partials)")) return false;
    parts_cap_ = cap;
    return true;
}

bool BatchRunner::build_graph() {
    if (exec_) (void)hipGraphExecDestroy(exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    exec_ = nullptr;
    graph_ = nullptr;
    if (!hip_ok(hipGraphCreate(&graph_, 0), "hipGraphCreate")) return false;
    const size_t qbytes = sizeof(BatchHeader) + max_q_ * (sizeof(BatchQuery) + sizeof(BatchJob));
    if (!hip_ok(hipGraphAddMemcpyNode1D(&n_up_, graph_, nullptr, 0, dq_, hq_, qbytes, hipMemcpyHostToDevice),
                "graph upload node"))
        return false;
    const BatchHeader* hdr = reinterpret_cast<const BatchHeader*>(dq_);
    const BatchQuery* qs = reinterpret_cast<const BatchQuery*>(dq_ + sizeof(BatchHeader));
    uint32_t* ctr = ctr_;
    void* args[] = {(void*)&hdr, (void*)&qs, (void*)&ctr};
    hipKernelNodeParams kp = {};
    kp.func = batch_kernel(cfg_.unroll);
    kp.gridDim = dim3((unsigned)(device_cu_count() * cfg_.blocks_per_cu));
    kp.blockDim = dim3(kBlock);
    kp.sharedMemBytes = 0;
    kp.kernelParams = args;
    kp.extra = nullptr;
    if (!hip_ok(hipGraphAddKernelNode(&n_kern_, graph_, &n_up_, 1, &kp), "graph kernel node")) return false;
    hipKernelNodeParams fp = kp;
    fp.func = reinterpret_cast<void*>(&k_batch_finish);
    fp.gridDim = dim3((unsigned)std::min<size_t>(max_q_, (size_t)device_cu_count() * 8));
    hipGraphNode_t n_fin;
    if (!hip_ok(hipGraphAddKernelNode(&n_fin, graph_, &n_kern_, 1, &fp), "graph finish node")) return false;
    if (!hip_ok(hipGraphAddMemcpyNode1D(&n_down_, graph_, &n_fin, 1, hres_, dres_, max_q_ * sizeof(ScanOut),
                                        hipMemcpyDeviceToHost), "graph download node"))
        return false;
    if (!hip_ok(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0), "hipGraphInstantiate")) return false;
    graph_valid_ = true;
    return true;
}

bool BatchRunner::launch(const voxel_gpu_scan_desc* q, size_t nq) {
    if (!ok_) { set_error("batch: not initialised"); return false; }
    if (inflight_) { set_error("batch: a launch is in flight"); return false; }
    if (nq == 0 || nq > max_q_) { set_error("batch: %zu queries (1..%zu allowed)", nq, max_q_); return false; }
    BatchHeader* h = reinterpret_cast<BatchHeader*>(hq_);
    BatchQuery* qs = reinterpret_cast<BatchQuery*>(hq_ + sizeof(BatchHeader));
    // Jobs: queries reading the same rows (column(s), length, element type)
    // share a scan, up to kBatchShare per job, in the caller's order.  Tiles
    // are listed largest job first, so the last tiles of every workgroup are
    // the small ones and no long tile runs alone at the end; results go back
    // to the caller's order (descriptor field orig).
    std::vector<ScanArgs> args(nq);
    for (size_t i = 0; i < nq; ++i)
        if (!make_scan_args(&q[i], args[i])) { set_error("batch query %zu: %s", i, last_error()); return false; }
    jobs_.clear();
    std::vector<std::vector<uint32_t>> members;
    {
        struct Key {
            const void* a; const void* b; uint64_t n; int dtype, mode;
            bool operator==(const Key& o) const {
                return a == o.a && b == o.b && n == o.n && dtype == o.dtype && mode == o.mode;
            }
        };
        struct KeyHash {
            size_t operator()(const Key& k) const {
                return std::hash<const void*>()(k.a) ^ (std::hash<const void*>()(k.b) * 31) ^
                       (std::hash<uint64_t>()(k.n) * 131) ^ (size_t)(k.dtype * 7 + k.mode);
            }
        };
        std::unordered_map<Key, size_t, KeyHash> open;      // key -> job still taking members
        for (size_t i = 0; i < nq; ++i) {
            const ScanArgs& g = args[i];
            // mode 1 (both predicates on column a) runs as a two-column job
            // whose second column is a itself
            const int jm = g.mode == 0 ? 0 : 2;
            const void* jb = g.mode == 2 ? g.b : g.a;
            Key k{g.a, jm ? jb : nullptr, q[i].n, g.dtype, jm};
            auto it = open.find(k);
            if (it == open.end() || members[it->second].size() == (size_t)kBatchShare) {
                BatchJob j;
                std::memset(&j, 0, sizeof(j));
                j.a = g.a;
                j.b = jb;
                j.n = q[i].n;
                j.dtype = g.dtype;
                j.mode = jm;
                jobs_.push_back(j);
                members.emplace_back();
                open[k] = jobs_.size() - 1;
                it = open.find(k);
            }
            members[it->second].push_back((uint32_t)i);
        }
    }
    const size_t nj = jobs_.size();
    uint64_t all_units = 0;
    for (const BatchJob& j : jobs_) all_units += units_of(j.n);
    // a small batch runs as one kernel on at most one workgroup per CU (its
    // chunks spread one wave per CU first); a large one on the configured grid
    const bool small = nq <= kSmallQ;
    const uint64_t grid_max = (uint64_t)device_cu_count() * (small ? 1 : cfg_.blocks_per_cu);
    const uint32_t tile_rows = tile_rows_for(all_units, grid_max);
    for (BatchJob& j : jobs_) {
        const uint64_t t = (j.n + tile_rows - 1) / tile_rows;
        j.ntiles = (uint32_t)(t ? t : 1);
    }
    order_.resize(nj);
    for (size_t i = 0; i < nj; ++i) order_[i] = (uint32_t)i;
    std::stable_sort(order_.begin(), order_.end(), [&](uint32_t x, uint32_t y) { return jobs_[x].n > jobs_[y].n; });
    BatchJob* jd = reinterpret_cast<BatchJob*>(hq_ + sizeof(BatchHeader) + nq * sizeof(BatchQuery));
    uint64_t tiles = 0, parts = 0;
    uint32_t qn = 0;                                    // query records in job order
    for (size_t jj = 0; jj < nj; ++jj) {
        BatchJob j = jobs_[order_[jj]];
        j.tile0 = tiles;
        tiles += j.ntiles;
        const std::vector<uint32_t>& mem = members[order_[jj]];
        j.nmem = (uint32_t)mem.size();
        for (size_t m = 0; m < mem.size(); ++m) {
            const uint32_t i = mem[m];
            const ScanArgs& g = args[i];
            BatchQuery& d = qs[qn];
            std::memset(&d, 0, sizeof(d));
            d.a = g.a;
            d.b = g.mode == 2 ? g.b : g.a;               // mode 1: column a again
            d.n = q[i].n;
            d.pa = g.pa;
            d.pb = g.mode != 0 ? g.pb : g.pa;
            d.dtype = g.dtype;
            d.mode = g.mode == 0 ? 0 : 2;
            d.unit0 = parts;
            d.nunits = (uint32_t)units_of(d.n);
            d.orig = i;
            parts += d.nunits;
            dtypes_[i] = g.dtype;
            j.mem[m] = qn++;
        }
        jd[jj] = j;
    }
    if (tiles >= (1ull << 32) || parts >= (1ull << 32)) { set_error("batch: too many tiles"); return false; }
    if (!ensure_parts(parts)) return false;
    std::memset(h, 0, sizeof(*h));
    h->nq = (uint32_t)nq;
    h->total_tiles = (uint32_t)tiles;
    h->tile_rows = tile_rows;
    last_tile_rows_ = tile_rows;
    h->njobs = (uint32_t)nj;
    h->parts = parts_;
    h->results = dres_;
    h->jobs = reinterpret_cast<const BatchJob*>(dq_ + sizeof(BatchHeader) + nq * sizeof(BatchQuery));
    if (small) {                                        // one kernel, descriptors as its argument
        const unsigned grid = (unsigned)std::min<uint64_t>(tiles, grid_max);
        SmallArgs sa;
        std::memset(&sa, 0, sizeof(sa));
        std::memcpy(&sa.h, h, sizeof(BatchHeader));
        std::memcpy(sa.q, qs, nq * sizeof(BatchQuery));
        std::memcpy(sa.j, jd, nj * sizeof(BatchJob));
        if (!hip_ok(hipEventRecord(ev0_, st_), "hipEventRecord")) return false;
        // a small batch has few units to spread over the waves, so each wave
        // keeps the deepest unroll in flight whatever the configuration (the
        // unroll does not change a lane's order of rows, so the bits are the
        // large path's)
        hipLaunchKernelGGL((k_batch_small<8, true>), dim3(grid), dim3(kBlock), 0, st_, sa, parts_);
        SmallQueries sq;
        std::memset(&sq, 0, sizeof(sq));
        std::memcpy(sq.q, qs, nq * sizeof(BatchQuery));
        hipLaunchKernelGGL(k_batch_small_finish, dim3((unsigned)nq), dim3(kBlock), 0, st_, sq, (const ScanOut*)parts_,
                           hres_);
        if (!hip_ok(hipGetLastError(), "small batch launch") || !hip_ok(hipEventRecord(ev1_, st_), "hipEventRecord"))
            return false;
        nq_ = nq;
        inflight_ = true;
        return true;
    }
    if (!graph_valid_ && !build_graph()) return false;
    const size_t up = sizeof(BatchHeader) + nq * sizeof(BatchQuery) + nj * sizeof(BatchJob);
    if (!hip_ok(hipGraphExecMemcpyNodeSetParams1D(exec_, n_up_, dq_, hq_, up, hipMemcpyHostToDevice), "graph upload size") ||
        !hip_ok(hipGraphExecMemcpyNodeSetParams1D(exec_, n_down_, hres_, dres_, nq * sizeof(ScanOut), hipMemcpyDeviceToHost),
                "graph download size"))
        return false;
    if (!hip_ok(hipEventRecord(ev0_, st_), "hipEventRecord") || !hip_ok(hipGraphLaunch(exec_, st_), "hipGraphLaunch") ||
        !hip_ok(hipEventRecord(ev1_, st_), "hipEventRecord"))
        return false;
    nq_ = nq;
    inflight_ = true;
    return true;
}

bool BatchRunner::done() {
    if (!inflight_) return true;
    return hipEventQuery(ev1_) == hipSuccess;
}

bool BatchRunner::wait(voxel_gpu_agg* out) {
    if (!inflight_) { set_error("batch: nothing launched"); return false; }
    inflight_ = false;
    if (!hip_ok(hipEventSynchronize(ev1_), "batch wait")) return false;
    for (size_t i = 0; i < nq_; ++i) fill_agg(dtypes_[i], hres_[i], &out[i]);
    return true;
}

float BatchRunner::last_ms() {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev0_, ev1_) != hipSuccess) return -1.0f;
    return ms;
}

}  // namespace vx

using namespace vx;

extern "C" {

void* voxel_gpu_batch_create(size_t max_queries) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    auto* r = new BatchRunner(max_queries);
    if (!r->ok()) { delete r; return nullptr; }
    return r;
}

void voxel_gpu_batch_destroy(void* h) { delete static_cast<BatchRunner*>(h); }

int voxel_gpu_batch_config(void* h, int unroll, uint32_t tile_rows, int blocks_per_cu) {
    set_error("%s", "");
    BatchRunner* r = static_cast<BatchRunner*>(h);
    BatchConfig c = r->config();
    if (unroll) c.unroll = unroll;
    if (tile_rows) c.tile_rows = tile_rows;
    if (blocks_per_cu) c.blocks_per_cu = blocks_per_cu;
    return r->set_config(c) ? 0 : -1;
}

int voxel_gpu_batch_get_config(void* h, int* unroll, uint32_t* tile_rows, int* blocks_per_cu) {
    const BatchConfig& c = static_cast<BatchRunner*>(h)->config();
    if (unroll) *unroll = c.unroll;
    if (tile_rows) *tile_rows = c.tile_rows;
    if (blocks_per_cu) *blocks_per_cu = c.blocks_per_cu;
    return 0;
}

int voxel_gpu_batch_launch(void* h, const voxel_gpu_scan_desc* q, size_t nq) {
    set_error("%s", "");
    return static_cast<BatchRunner*>(h)->launch(q, nq) ? 0 : -1;
}

int voxel_gpu_batch_done(void* h) { return static_cast<BatchRunner*>(h)->done() ? 1 : 0; }

int voxel_gpu_batch_wait(void* h, voxel_gpu_agg* out, float* ms) {
    set_error("%s", "");
    BatchRunner* r = static_cast<BatchRunner*>(h);
    if (!r->wait(out)) return -1;
    if (ms) *ms = r->last_ms();
    return 0;
}

int voxel_gpu_batch_run(void* h, const voxel_gpu_scan_desc* q, size_t nq, voxel_gpu_agg* out, float* ms) {
    if (voxel_gpu_batch_launch(h, q, nq) != 0) return -1;
    return voxel_gpu_batch_wait(h, out, ms);
}

// Autotuner: times every (unroll, tile_rows, blocks_per_cu) of the search
// space on the given queries (median of reps launches each) and keeps the
// fastest.  Per-query results of every configuration are checked against
// the first one bit for bit (the reduction units do not change with the
// configuration).
int voxel_gpu_batch_autotune(void* h, const voxel_gpu_scan_desc* q, size_t nq, int reps, double* best_ms,
                             int* configs_tried) {
    set_error("%s", "");
    BatchRunner* r = static_cast<BatchRunner*>(h);
    if (reps < 1) reps = 3;
    const int unrolls[] = {2, 4, 8};
    const uint32_t tiles[] = {16384, 32768, 65536, 131072, 262144};
    const int bpcs[] = {1, 2, 4};
    std::vector<voxel_gpu_agg> ref(nq), cur(nq);
    BatchConfig best = r->config();
    double best_t = 1e300;
    int tried = 0;
    bool have_ref = false;
    for (int u : unrolls)
        for (uint32_t t : tiles)
            for (int b : bpcs) {
                BatchConfig c{u, t, b};
                if (!r->set_config(c)) return -1;
                std::vector<float> ms;
                for (int k = 0; k < reps + 1; ++k) {
                    if (!r->launch(q, nq) || !r->wait(cur.data())) return -1;
                    if (k > 0) ms.push_back(r->last_ms());
                }
                ++tried;
                if (!have_ref) { ref = cur; have_ref = true; }
                for (size_t i = 0; i < nq; ++i) {
                    if (std::memcmp(&cur[i], &ref[i], sizeof(voxel_gpu_agg)) != 0) {
                        set_error("autotune: query %zu differs under unroll %d tile %u blocks/CU %d", i, u, t, b);
                        return -1;
                    }
                }
                std::sort(ms.begin(), ms.end());
                const double med = ms[ms.size() / 2];
                if (med < best_t) { best_t = med; best = c; }
            }
    if (!r->set_config(best)) return -1;
    if (best_ms) *best_ms = best_t;
    if (configs_tried) *configs_tried = tried;
    return 0;
}

}  // extern "C"
