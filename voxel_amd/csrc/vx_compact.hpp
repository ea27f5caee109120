// vx_compact.hpp -- order-preserving stream compaction on gfx950.
//
// One launch, one pass: workgroup w takes tile id t from an atomic ticket,
// evaluates the row predicate of its 4096 rows, ranks the passing rows in row
// order (wave ballots + mbcnt-style popcounts, then a 64-entry LDS scan),
// publishes its passing count and looks back over the status words of
// earlier tiles (decoupled look-back, flag | inclusive count in 64 bits) for
// its output offset, then writes the passing rows' outputs in row order.
// Sources are functors:
//   __device__ bool pass(uint64_t row) const;
//   __device__ void emit(uint64_t row, uint64_t out_index) const;
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {

constexpr int kCompactThreads = 256;
constexpr int kCompactRounds = 16;
constexpr int kCompactTile = kCompactThreads * kCompactRounds;
constexpr uint64_t kCFlagAgg = 1ull << 62;
constexpr uint64_t kCFlagPre = 2ull << 62;
constexpr uint64_t kCValMask = (1ull << 62) - 1;

__device__ __forceinline__ uint64_t c_lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <class Src>
__global__ __launch_bounds__(kCompactThreads) void k_compact(Src src, uint64_t n, uint64_t* __restrict__ status,
                                                             uint32_t* __restrict__ ticket,
                                                             unsigned long long* __restrict__ total_out) {
    __shared__ uint32_t tile_s;
    __shared__ uint32_t wtot[kCompactRounds][kCompactThreads / 64];
    __shared__ uint32_t wpre[kCompactRounds][kCompactThreads / 64];
    __shared__ unsigned long long base_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) tile_s = atomicAdd(ticket, 1u);
    __syncthreads();
    const uint64_t tile = tile_s;
    const uint64_t base = tile * kCompactTile;
    bool p[kCompactRounds];
    uint32_t below[kCompactRounds];
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        uint64_t row = base + (uint64_t)r * kCompactThreads + threadIdx.x;
        p[r] = row < n && src.pass(row);
        uint64_t m = __ballot(p[r]);
        below[r] = (uint32_t)__popcll(m & c_lanemask_lt());
        if (lane == 0) wtot[r][w] = (uint32_t)__popcll(m);
    }
    __syncthreads();
    // exclusive scan over (round, wave) in row order; one thread, 64 entries
    if (threadIdx.x == 0) {
        uint32_t run = 0;
        for (int r = 0; r < kCompactRounds; ++r)
            for (int ww = 0; ww < kCompactThreads / 64; ++ww) { wpre[r][ww] = run; run += wtot[r][ww]; }
        uint64_t tot = run;
        uint64_t excl = 0;
        if (tile == 0) {
            __hip_atomic_store(&status[0], kCFlagPre | tot, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        } else {
            __hip_atomic_store(&status[tile], kCFlagAgg | tot, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
            int64_t j = (int64_t)tile - 1;
            while (true) {
                uint64_t s = __hip_atomic_load(&status[j], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
                uint64_t flag = s & ~kCValMask;
                if (flag == 0) { __builtin_amdgcn_s_sleep(1); continue; }
                excl += s & kCValMask;
                if (flag == kCFlagPre) break;
                --j;
            }
            __hip_atomic_store(&status[tile], kCFlagPre | (excl + tot), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
        base_s = excl;
        if (tile == gridDim.x - 1) *total_out = excl + tot;
    }
    __syncthreads();
    const uint64_t ob = base_s;
#pragma unroll
    for (int r = 0; r < kCompactRounds; ++r) {
        if (p[r]) {
            uint64_t row = base + (uint64_t)r * kCompactThreads + threadIdx.x;
            src.emit(row, ob + wpre[r][w] + below[r]);
        }
    }
}

inline uint64_t compact_tiles(uint64_t n) { return (n + kCompactTile - 1) / kCompactTile; }

struct StreamCtx;

// Compaction work areas (status words + ticket + total) per stream context.
struct CompactWork {
    uint64_t* status = nullptr;
    uint64_t cap = 0;
    uint32_t* ticket = nullptr;
    unsigned long long* total = nullptr;
    unsigned long long* total_host = nullptr;
};
CompactWork& compact_work(StreamCtx* c);   // vx_filter.hip

// Runs one compaction of rows [0, n) and returns the number written.
template <class Src>
bool run_compact(const Src& src, uint64_t n, StreamCtx* c, uint64_t* count) {
    CompactWork& w = compact_work(c);
    uint64_t tiles = compact_tiles(n);
    if (!w.ticket) {
        if (!hip_ok(hipMalloc((void**)&w.ticket, 64), "hipMalloc")) return false;
        if (!hip_ok(hipMalloc((void**)&w.total, 64), "hipMalloc")) return false;
        if (!hip_ok(hipHostMalloc((void**)&w.total_host, 64, hipHostMallocDefault), "hipHostMalloc")) return false;
    }
    if (tiles > w.cap) {
        if (w.status) (void)hipFree(w.status);
        w.status = nullptr; w.cap = 0;
        if (!hip_ok(hipMalloc((void**)&w.status, tiles * 8), "hipMalloc(status)")) return false;
        w.cap = tiles;
    }
    if (n == 0) { *count = 0; return true; }
    if (!hip_ok(launch_zero32((uint32_t*)w.status, tiles * 2, c->stream), "zero")) return false;
    if (!hip_ok(launch_zero32(w.ticket, 1, c->stream), "zero")) return false;
    hipLaunchKernelGGL((k_compact<Src>), dim3((unsigned)tiles), dim3(kCompactThreads), 0, c->stream, src, n,
                       w.status, w.ticket, w.total);
    if (!hip_ok(hipGetLastError(), "compact launch")) return false;
    if (!hip_ok(hipMemcpyAsync(w.total_host, w.total, 8, hipMemcpyDeviceToHost, c->stream), "read count")) return false;
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return false;
    *count = *w.total_host;
    return true;
}

}  // namespace vx
