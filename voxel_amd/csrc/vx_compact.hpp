// vx_compact.hpp -- order-preserving stream compaction on gfx950.
//
// Reduce-then-scan, no inter-workgroup waits: rows are cut into at most
// 2048 ranges of whole 4096-row tiles; k_compact_count counts each range's
// passing rows, k_compact_scan turns the counts into each range's output
// offset (and the total), and k_compact_emit (one workgroup per range, tiles
// in order) ranks the passing rows of each tile in row order (wave ballots +
// popcounts, then a 64-entry LDS scan) and writes their outputs.  The
// predicate is evaluated twice (count, emit).  Sources are functors:
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
constexpr uint64_t kCompactMaxRanges = 2048;

__device__ __forceinline__ uint64_t c_lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

inline uint64_t compact_tiles_per_range(uint64_t n) {
    uint64_t tiles = (n + kCompactTile - 1) / kCompactTile;
    uint64_t r = (tiles + kCompactMaxRanges - 1) / kCompactMaxRanges;
    return r ? r : 1;
}

inline uint64_t compact_ranges(uint64_t n) {
    uint64_t tiles = (n + kCompactTile - 1) / kCompactTile;
    uint64_t per = compact_tiles_per_range(n);
    uint64_t r = (tiles + per - 1) / per;
    return r ? r : 1;
}

template <class Src>
__global__ __launch_bounds__(kCompactThreads) void k_compact_count(Src src, uint64_t n, uint64_t range_len,
                                                                   uint32_t* __restrict__ cnt) {
    __shared__ uint32_t ws[kCompactThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint64_t s = (uint64_t)blockIdx.x * range_len;
    const uint64_t e = s + range_len < n ? s + range_len : n;
    uint32_t c = 0;
    for (uint64_t i0 = s; i0 < e; i0 += kCompactThreads) {       // block-uniform trip count
        const uint64_t i = i0 + threadIdx.x;
        c += (uint32_t)__popcll(__ballot(i < e && src.pass(i)));
    }
    if (lane == 0) ws[w] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (int i = 0; i < kCompactThreads / 64; ++i) t += ws[i];
        cnt[blockIdx.x] = t;
    }
}

// one workgroup: off[r] = passing rows of ranges before r; *total = all
static __global__ __launch_bounds__(256) void k_compact_scan(const uint32_t* __restrict__ cnt, uint32_t nr,
                                                             uint64_t* __restrict__ off,
                                                             unsigned long long* __restrict__ total) {
    __shared__ uint64_t sc[256];
    const uint32_t t = threadIdx.x;
    const uint32_t per = (nr + 255) / 256;
    const uint32_t lo = t * per < nr ? t * per : nr, hi = lo + per < nr ? lo + per : nr;
    uint64_t sum = 0;
    for (uint32_t r = lo; r < hi; ++r) sum += cnt[r];
    sc[t] = sum;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {
        uint64_t x = t >= (uint32_t)o ? sc[t - o] : 0;
        __syncthreads();
        sc[t] += x;
        __syncthreads();
    }
    uint64_t run = sc[t] - sum;
    for (uint32_t r = lo; r < hi; ++r) {
        off[r] = run;
        run += cnt[r];
    }
    if (t == 255) *total = sc[255];
}

template <class Src>
__global__ __launch_bounds__(kCompactThreads) void k_compact_emit(Src src, uint64_t n, uint64_t tiles_per_range,
                                                                  const uint64_t* __restrict__ off) {
    __shared__ uint32_t wtot[kCompactRounds][kCompactThreads / 64];
    __shared__ uint32_t wpre[kCompactRounds][kCompactThreads / 64];
    __shared__ unsigned long long run_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) run_s = off[blockIdx.x];
    for (uint64_t t = 0; t < tiles_per_range; ++t) {
        const uint64_t base = ((uint64_t)blockIdx.x * tiles_per_range + t) * kCompactTile;
        if (base >= n) break;                                        // block-uniform
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
        }
        __syncthreads();
        const uint64_t ob = run_s;
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r) {
            if (p[r]) {
                uint64_t row = base + (uint64_t)r * kCompactThreads + threadIdx.x;
                src.emit(row, ob + wpre[r][w] + below[r]);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            run_s = ob + wpre[kCompactRounds - 1][kCompactThreads / 64 - 1] +
                    wtot[kCompactRounds - 1][kCompactThreads / 64 - 1];
        __syncthreads();
    }
}

struct StreamCtx;

// Compaction work areas (range counts + offsets + total) per stream context.
struct CompactWork {
    uint32_t* cnt = nullptr;
    uint64_t* off = nullptr;
    uint64_t cap = 0;                       // ranges
    unsigned long long* total = nullptr;
    unsigned long long* total_host = nullptr;
};
CompactWork& compact_work(StreamCtx* c);   // vx_filter.hip

// Runs one compaction of rows [0, n) and returns the number written.
template <class Src>
bool run_compact(const Src& src, uint64_t n, StreamCtx* c, uint64_t* count) {
    CompactWork& w = compact_work(c);
    if (!w.total) {
        if (!hip_ok(hipMalloc((void**)&w.total, 64), "hipMalloc")) return false;
        if (!hip_ok(hipHostMalloc((void**)&w.total_host, 64, hipHostMallocDefault), "hipHostMalloc")) return false;
    }
    if (n == 0) { *count = 0; return true; }
    const uint64_t nr = compact_ranges(n), per = compact_tiles_per_range(n);
    if (nr > w.cap) {
        if (w.cnt) (void)hipFree(w.cnt);
        if (w.off) (void)hipFree(w.off);
        w.cnt = nullptr; w.off = nullptr; w.cap = 0;
        if (!hip_ok(hipMalloc((void**)&w.cnt, kCompactMaxRanges * 4), "hipMalloc") ||
            !hip_ok(hipMalloc((void**)&w.off, kCompactMaxRanges * 8), "hipMalloc"))
            return false;
        w.cap = kCompactMaxRanges;
    }
    hipLaunchKernelGGL((k_compact_count<Src>), dim3((unsigned)nr), dim3(kCompactThreads), 0, c->stream, src, n,
                       per * kCompactTile, w.cnt);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(256), 0, c->stream, w.cnt, (uint32_t)nr, w.off, w.total);
    hipLaunchKernelGGL((k_compact_emit<Src>), dim3((unsigned)nr), dim3(kCompactThreads), 0, c->stream, src, n, per,
                       (const uint64_t*)w.off);
    if (!hip_ok(hipGetLastError(), "compact launch")) return false;
    if (!hip_ok(hipMemcpyAsync(w.total_host, w.total, 8, hipMemcpyDeviceToHost, c->stream), "read count")) return false;
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return false;
    *count = *w.total_host;
    return true;
}

}  // namespace vx
