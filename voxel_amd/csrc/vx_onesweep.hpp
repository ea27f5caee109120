// vx_onesweep.hpp -- the onesweep digit-pass kernels of vx_radix.hip as
// templates over the digit function, so that other kernel files can
// instantiate them (the learned partitioner's model digit, vx_learned.hip).
// See vx_radix.hip for the algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_kernels.hpp"
#include "vx_radix.hpp"

namespace vx {

constexpr int kThreads = 256;           // = number of digit values
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 16;             // keys per lane per tile
constexpr int kTile = kThreads * kRounds;   // 4096 keys per tile
constexpr uint32_t kFlagAgg = 1u << 30;
constexpr uint32_t kFlagPre = 2u << 30;
constexpr uint32_t kValMask = (1u << 30) - 1;

template <class Digit>
__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t* __restrict__ keys, uint64_t n, Digit dig,
                                                   int passes, unsigned long long* __restrict__ ghist) {
    __shared__ uint32_t h[kWaves][8][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 8 * 256; i += kThreads) (&h[0][0][0])[i] = 0;
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        uint64_t k = keys[i];
        for (int p = 0; p < passes; ++p) atomicAdd(&h[w][p][dig.digit(k, p)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += kThreads) {
        uint32_t s = 0;
        for (int ww = 0; ww < kWaves; ++ww) s += (&h[ww][0][0])[i];
        if (s) atomicAdd(&ghist[i], (unsigned long long)s);
    }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of this wave holding the same 8-bit digit (and a valid key).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return valid ? peers : 0ull;
}

template <class Digit>
__global__ __launch_bounds__(kThreads) void k_onesweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                       uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                       uint64_t n, Digit dig, int pass,
                                                       const unsigned long long* __restrict__ offsets,
                                                       uint32_t* __restrict__ status, uint32_t* __restrict__ ticket) {
    __shared__ uint64_t buf[kTile];                 // keys, then payloads (as u32)
    __shared__ uint32_t wcnt[kWaves][256];          // per-wave digit counters -> wave prefixes
    __shared__ uint32_t lstart[256];                // tile-local start of each digit
    __shared__ unsigned long long gbase[256];       // global start of each digit for this tile
    __shared__ uint32_t tile_s;
    __shared__ uint32_t scan_s[256];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (threadIdx.x == 0) tile_s = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&wcnt[0][0])[i] = 0;
    __syncthreads();
    const uint32_t tile = tile_s;
    const uint64_t base = (uint64_t)tile * kTile + (uint64_t)w * (64 * kRounds);

    uint64_t k[kRounds];
    uint32_t rank[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint64_t idx = base + r * 64 + lane;
        k[r] = idx < n ? kin[idx] : 0;
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint64_t idx = base + r * 64 + lane;
        bool valid = idx < n;
        uint32_t d = dig.digit(k[r], pass);
        uint64_t peers = match_digit(d, valid);
        uint32_t before = valid ? wcnt[w][d] : 0;
        uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        // the lowest lane of each peer group adds the group's size
        if (valid && below == 0) wcnt[w][d] = before + (uint32_t)__popcll(peers);
        rank[r] = before + below;
    }
    __syncthreads();

    // per digit: wave prefixes, tile count, tile-local digit start
    const int d = threadIdx.x;
    uint32_t tot = 0;
#pragma unroll
    for (int ww = 0; ww < kWaves; ++ww) {
        uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = tot;
        tot += c;
    }
    scan_s[d] = tot;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {
        uint32_t t = d >= off ? scan_s[d - off] : 0;
        __syncthreads();
        scan_s[d] += t;
        __syncthreads();
    }
    lstart[d] = scan_s[d] - tot;

    // decoupled look-back for digit d
    uint32_t* st = status + (uint64_t)tile * 256;
    if (tile == 0) {
        __hip_atomic_store(&st[d], kFlagPre | tot, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        gbase[d] = offsets[pass * 256 + d];
    } else {
        __hip_atomic_store(&st[d], kFlagAgg | tot, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        uint32_t excl = 0;
        int64_t j = (int64_t)tile - 1;
        while (true) {
            uint32_t s = __hip_atomic_load(&status[(uint64_t)j * 256 + d], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
            uint32_t flag = s & ~kValMask;
            if (flag == 0) { __builtin_amdgcn_s_sleep(1); continue; }
            excl += s & kValMask;
            if (flag == kFlagPre) break;
            --j;
        }
        __hip_atomic_store(&st[d], kFlagPre | (excl + tot), __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        gbase[d] = offsets[pass * 256 + d] + excl;
    }
    __syncthreads();

    // keys into LDS in digit order (stable: wave prefix + rank)
    uint32_t lpos[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint64_t idx = base + r * 64 + lane;
        if (idx < n) {
            uint32_t dd = dig.digit(k[r], pass);
            lpos[r] = lstart[dd] + wcnt[w][dd] + rank[r];
            if (lpos[r] < kTile) buf[lpos[r]] = k[r];
        }
    }
    __syncthreads();
    const uint64_t tbase = (uint64_t)tile * kTile;
    const uint32_t tn = tbase + kTile <= n ? kTile : (uint32_t)(n > tbase ? n - tbase : 0);
    uint64_t gpos[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint32_t i = r * kThreads + threadIdx.x;
        if (i < tn) {
            uint64_t key = buf[i];
            uint32_t dd = dig.digit(key, pass);
            gpos[r] = gbase[dd] + (i - lstart[dd]);
            if (gpos[r] < n) kout[gpos[r]] = key;     // always true when the histogram matches
        }
    }
    __syncthreads();
    // payloads through the same LDS slots
    uint32_t* vb = reinterpret_cast<uint32_t*>(buf);
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint64_t idx = base + r * 64 + lane;
        if (idx < n && lpos[r] < kTile) vb[lpos[r]] = vin[idx];
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        uint32_t i = r * kThreads + threadIdx.x;
        if (i < tn && gpos[r] < n) vout[gpos[r]] = vb[i];
    }
}

hipError_t launch_scan_hist(const unsigned long long* ghist, int passes, uint64_t n, unsigned long long* offsets,
                            uint32_t* skip, hipStream_t st);

template <class Digit>
hipError_t radix_histogram(const uint64_t* keys, uint64_t n, const Digit& dig, int passes, RadixWork& w,
                           hipStream_t st) {
    hipError_t e = launch_zero32((uint32_t*)w.hist, (uint64_t)passes * 256 * 2, st);
    if (e != hipSuccess) return e;
    if ((e = launch_zero32(w.skip, 8, st)) != hipSuccess) return e;
    unsigned blocks = (unsigned)(device_cu_count() * 4);
    uint64_t need = (n + kThreads * 16 - 1) / (kThreads * 16);
    if (need < blocks) blocks = (unsigned)(need ? need : 1);
    hipLaunchKernelGGL((k_hist<Digit>), dim3(blocks), dim3(kThreads), 0, st, keys, n, dig, passes, w.hist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_scan_hist(w.hist, passes, n, w.offsets, w.skip, st);
}

template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    hipError_t e = launch_zero32(w.status, tiles * 256, st);
    if (e != hipSuccess) return e;
    if ((e = launch_zero32(w.ticket, 1, st)) != hipSuccess) return e;
    hipLaunchKernelGGL((k_onesweep<Digit>), dim3((unsigned)tiles), dim3(kThreads), 0, st, kin, vin, kout, vout, n,
                       dig, pass, (const unsigned long long*)w.offsets, w.status, w.ticket);
    return hipGetLastError();
}

}  // namespace vx
