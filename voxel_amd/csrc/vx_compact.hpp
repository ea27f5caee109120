// vx_compact.hpp -- order-preserving stream compaction on gfx950.
//
// Reduce-then-scan, no inter-workgroup waits: rows are cut into at most
// 16384 ranges of whole 4096-row tiles (one tile per range up to 2^26 rows); k_compact_count counts each range's
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

#include <type_traits>

#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {

constexpr int kCompactThreads = 256;
constexpr int kCompactRounds = 16;
constexpr int kCompactTile = kCompactThreads * kCompactRounds;
constexpr uint64_t kCompactMaxRanges = 1u << 14;     // ranges of 4096-row tiles: 4 tiles each at 2^28 rows

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

// cnt[r] = passing rows of range r; workgroups stride over the ranges.
template <class Src>
__global__ __launch_bounds__(kCompactThreads) void k_compact_count(Src src, uint64_t n, uint64_t range_len,
                                                                   uint64_t nr, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t ws[kCompactThreads / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint64_t rg = blockIdx.x; rg < nr; rg += gridDim.x) {
        const uint64_t s = rg * range_len;
        const uint64_t e = s + range_len < n ? s + range_len : n;
        uint32_t c = 0;
        constexpr int U = 8;                                       // predicates in flight per lane
        for (uint64_t i0 = s; i0 < e; i0 += (uint64_t)kCompactThreads * U) {   // block-uniform trip count
            bool p[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {                          // unconditional (clamped) loads: one
                const uint64_t i = i0 + (uint64_t)u * kCompactThreads + threadIdx.x;   // basic block
                const bool ok = src.pass(i < e ? i : e - 1);
                p[u] = ok & (i < e);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) c += (uint32_t)__popcll(__ballot(p[u]));
        }
        if (lane == 0) ws[w] = c;
        __syncthreads();
        if (threadIdx.x == 0) {
            uint32_t t = 0;
            for (int i = 0; i < kCompactThreads / 64; ++i) t += ws[i];
            cnt[rg] = t;
        }
        __syncthreads();
    }
}

// one workgroup: off[r] = passing rows of ranges before r; *total = all.
// Chunks of 4096 counts (16 consecutive per thread, coalesced loads), wave
// shuffle scans and a running carry: four barriers per chunk.
static __global__ __launch_bounds__(256) void k_compact_scan(const uint32_t* __restrict__ cnt, uint32_t nr,
                                                             uint64_t* __restrict__ off,
                                                             unsigned long long* __restrict__ total) {
    constexpr int kItems = 16, kChunk = 256 * kItems;
    __shared__ uint64_t tile[kChunk + kChunk / 16];
    __shared__ uint64_t wsum[4];
    auto at = [](int i) { return i + (i >> 4); };         // a thread's 16 entries avoid one bank
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    uint64_t carry = 0;
    for (uint64_t t0 = 0; t0 < nr; t0 += kChunk) {
        for (int k = 0; k < kItems; ++k) {
            const int i = k * 256 + t;
            tile[at(i)] = t0 + i < nr ? cnt[t0 + i] : 0;
        }
        __syncthreads();
        uint64_t sum = 0;
        for (int k = 0; k < kItems; ++k) sum += tile[at(t * kItems + k)];
        uint64_t incl = sum;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint64_t u = __shfl_up(incl, o, 64);
            if (lane >= o) incl += u;
        }
        if (lane == 63) wsum[wv] = incl;
        __syncthreads();
        uint64_t before = carry;
        for (int k = 0; k < wv; ++k) before += wsum[k];
        const uint64_t chunk_total = wsum[0] + wsum[1] + wsum[2] + wsum[3];
        uint64_t run = before + incl - sum;
        for (int k = 0; k < kItems; ++k) {
            const int i = at(t * kItems + k);
            const uint64_t v = tile[i];
            tile[i] = run;
            run += v;
        }
        __syncthreads();
        for (int k = 0; k < kItems; ++k) {
            const int i = k * 256 + t;
            if (t0 + i < nr) off[t0 + i] = tile[at(i)];
        }
        carry += chunk_total;
        __syncthreads();
    }
    if (t == 0) *total = carry;
}

// A source may split emit into fetch(row) -> V and put(row, out_index, V):
// the kernel then loads every row of a tile with its predicate (the same
// address, so one load when pass reads it too) before the tile's rank scan,
// instead of each store waiting for its own load.
template <class S, class = void>
struct CompactFetchType { using V = char; };
template <class S>
struct CompactFetchType<S, std::void_t<typename S::V>> { using V = typename S::V; };
template <class S, class = void>
struct CompactFetch : std::false_type {};
template <class S>
struct CompactFetch<S, std::void_t<typename S::V>> : std::true_type {};

template <class Src>
__global__ __launch_bounds__(kCompactThreads) void k_compact_emit(Src src, uint64_t n, uint64_t tiles_per_range,
                                                                  uint64_t nr, const uint64_t* __restrict__ off) {
    __shared__ uint32_t wtot[kCompactRounds][kCompactThreads / 64];
    __shared__ uint32_t wpre[kCompactRounds][kCompactThreads / 64];
    __shared__ unsigned long long run_s;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint64_t rg = blockIdx.x; rg < nr; rg += gridDim.x) {    // workgroups stride over the ranges
    __syncthreads();
    if (threadIdx.x == 0) run_s = off[rg];
    for (uint64_t t = 0; t < tiles_per_range; ++t) {
        const uint64_t base = (rg * tiles_per_range + t) * kCompactTile;
        if (base >= n) break;                                        // block-uniform
        bool p[kCompactRounds];
        uint32_t below[kCompactRounds];
        using FV = typename std::conditional<CompactFetch<Src>::value, typename CompactFetchType<Src>::V, char>::type;
        FV fv[kCompactRounds];
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r) {                 // every load of the tile issued first:
            const uint64_t row = base + (uint64_t)r * kCompactThreads + threadIdx.x;   // clamped, branch-free
            const uint64_t rr = row < n ? row : n - 1;
            const bool ok = src.pass(rr);
            if constexpr (CompactFetch<Src>::value) fv[r] = src.fetch(rr);
            p[r] = ok & (row < n);
        }
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r) {
            uint64_t m = __ballot(p[r]);
            below[r] = (uint32_t)__popcll(m & c_lanemask_lt());
            if (lane == 0) wtot[r][w] = (uint32_t)__popcll(m);
        }
        __syncthreads();
        // exclusive scan over the 64 (round, wave) counts in row order, by
        // the first wave: lane e holds entry e, six shuffle steps
        static_assert(kCompactRounds * (kCompactThreads / 64) == 64, "one entry per lane");
        if (threadIdx.x < 64) {
            const int e = threadIdx.x;
            const uint32_t v = wtot[e / (kCompactThreads / 64)][e % (kCompactThreads / 64)];
            uint32_t incl = v;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t u = __shfl_up(incl, o, 64);
                if (e >= o) incl += u;
            }
            wpre[e / (kCompactThreads / 64)][e % (kCompactThreads / 64)] = incl - v;
        }
        __syncthreads();
        const uint64_t ob = run_s;
#pragma unroll
        for (int r = 0; r < kCompactRounds; ++r) {
            if (p[r]) {
                uint64_t row = base + (uint64_t)r * kCompactThreads + threadIdx.x;
                if constexpr (CompactFetch<Src>::value) src.put(row, ob + wpre[r][w] + below[r], fv[r]);
                else src.emit(row, ob + wpre[r][w] + below[r]);
            }
        }
        __syncthreads();
        if (threadIdx.x == 0)
            run_s = ob + wpre[kCompactRounds - 1][kCompactThreads / 64 - 1] +
                    wtot[kCompactRounds - 1][kCompactThreads / 64 - 1];
        __syncthreads();
    }
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
        const uint64_t cap = nr < 2048 ? 2048 : nr;
        if (!hip_ok(hipMalloc((void**)&w.cnt, cap * 4), "hipMalloc") ||
            !hip_ok(hipMalloc((void**)&w.off, cap * 8), "hipMalloc"))
            return false;
        w.cap = cap;
    }
    const unsigned grid = (unsigned)nr;                              // one workgroup per range
    hipLaunchKernelGGL((k_compact_count<Src>), dim3(grid), dim3(kCompactThreads), 0, c->stream, src, n,
                       per * kCompactTile, nr, w.cnt);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(256), 0, c->stream, w.cnt, (uint32_t)nr, w.off, w.total);
    hipLaunchKernelGGL((k_compact_emit<Src>), dim3(grid), dim3(kCompactThreads), 0, c->stream, src, n, per, nr,
                       (const uint64_t*)w.off);
    if (!hip_ok(hipGetLastError(), "compact launch")) return false;
    if (!hip_ok(hipMemcpyAsync(w.total_host, w.total, 8, hipMemcpyDeviceToHost, c->stream), "read count")) return false;
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return false;
    *count = *w.total_host;
    return true;
}

}  // namespace vx
