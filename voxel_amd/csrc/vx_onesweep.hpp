// vx_onesweep.hpp -- the digit-pass kernels of vx_radix.hip as templates
// over the digit function, so that other kernel files can instantiate them
// (the learned partitioner's model digit, vx_learned.hip).  See vx_radix.hip
// for the algorithm.
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

template <class Digit>
__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t* __restrict__ keys, uint64_t n, Digit dig,
                                                   int passes, unsigned long long* __restrict__ ghist) {
    __shared__ uint32_t h[kWaves][8][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 8 * 256; i += kThreads) (&h[0][0][0])[i] = 0;
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const uint64_t b = dig.bits(keys[i]);
        for (int p = 0; p < passes; ++p) atomicAdd(&h[w][p][dig.digit_of_bits(b, p)], 1u);
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

// Rows of the key array are cut into ranges of whole 4096-key tiles, at
// most kMaxRanges of them.  Per digit pass: k_upsweep counts each range's
// digits, k_range_scan turns the counts into each range's first output
// position per digit, and k_downsweep (one workgroup per range, tiles in
// order) scatters stably.  No workgroup waits on another.
constexpr uint64_t kMaxRanges = 2048;

__host__ __device__ inline uint64_t radix_tiles_per_range(uint64_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t r = (tiles + kMaxRanges - 1) / kMaxRanges;
    return r ? r : 1;
}

__host__ __device__ inline uint64_t radix_ranges(uint64_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t per = radix_tiles_per_range(n);
    uint64_t r = (tiles + per - 1) / per;
    return r ? r : 1;
}

// cnt[d * nr + range] = keys of the range with digit d
template <class Digit>
__global__ __launch_bounds__(kThreads) void k_upsweep(const uint64_t* __restrict__ keys, uint64_t n, Digit dig, int pass,
                                                      uint64_t range_len, uint32_t nr, uint32_t* __restrict__ cnt) {
    __shared__ uint32_t h[kWaves][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * range_len;
    const uint64_t e = s + range_len < n ? s + range_len : n;
    for (uint64_t i0 = s; i0 < e; i0 += kThreads) {         // block-uniform trip count
        const uint64_t i = i0 + threadIdx.x;
        const bool valid = i < e;
        const uint32_t d = valid ? dig.digit(keys[i], pass) : 0u;
        const uint64_t peers = match_digit(d, valid);
        if (valid && (peers & lanemask_lt()) == 0) h[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    const int d = threadIdx.x;
    cnt[(uint64_t)d * nr + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// One workgroup per digit: off[d * nr + r] = offsets[pass][d] + keys with
// digit d in ranges before r.  (Internal linkage: every file that
// instantiates the passes gets its own copy.)
static __global__ __launch_bounds__(kThreads) void k_range_scan(const uint32_t* __restrict__ cnt, uint32_t nr,
                                                         const unsigned long long* __restrict__ offsets, int pass,
                                                         uint32_t* __restrict__ off) {
    __shared__ uint32_t sc[kThreads];
    const uint32_t d = blockIdx.x, t = threadIdx.x;
    const uint32_t per = (nr + kThreads - 1) / kThreads;
    const uint32_t lo = t * per < nr ? t * per : nr, hi = lo + per < nr ? lo + per : nr;
    const uint32_t* c = cnt + (uint64_t)d * nr;
    uint32_t sum = 0;
    for (uint32_t r = lo; r < hi; ++r) sum += c[r];
    sc[t] = sum;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        uint32_t x = t >= (uint32_t)o ? sc[t - o] : 0;
        __syncthreads();
        sc[t] += x;
        __syncthreads();
    }
    uint32_t run = (uint32_t)offsets[pass * 256 + d] + sc[t] - sum;
    uint32_t* o = off + (uint64_t)d * nr;
    for (uint32_t r = lo; r < hi; ++r) {
        o[r] = run;
        run += c[r];
    }
}

template <class Digit>
__global__ __launch_bounds__(kThreads) void k_downsweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                        uint64_t n, Digit dig, int pass, uint64_t tiles_per_range,
                                                        uint32_t nr, const uint32_t* __restrict__ off) {
    __shared__ uint64_t buf[kTile];                 // keys, then payloads (as u32)
    __shared__ uint32_t wcnt[kWaves][256];          // per-wave digit counters -> wave prefixes
    __shared__ uint32_t lstart[256];                // tile-local start of each digit
    __shared__ uint32_t run[256];                   // output position of the digit's next key
    __shared__ uint32_t scan_s[256];
    __shared__ uint8_t dig8[kTile];                 // digit of each LDS slot

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, d = threadIdx.x;
    run[d] = off[(uint64_t)d * nr + blockIdx.x];
    for (uint64_t t = 0; t < tiles_per_range; ++t) {
        const uint64_t tbase = ((uint64_t)blockIdx.x * tiles_per_range + t) * kTile;
        if (tbase >= n) break;                                          // block-uniform
        for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&wcnt[0][0])[i] = 0;
        __syncthreads();
        const uint64_t base = tbase + (uint64_t)w * (64 * kRounds);
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
            uint32_t dd = dig.digit(k[r], pass);
            uint64_t peers = match_digit(dd, valid);
            uint32_t before = valid ? wcnt[w][dd] : 0;
            uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
            // the lowest lane of each peer group adds the group's size
            if (valid && below == 0) wcnt[w][dd] = before + (uint32_t)__popcll(peers);
            rank[r] = (before + below) | (dd << 16);        // rank < 4096, digit above it
        }
        __syncthreads();
        // per digit: wave prefixes, tile count, tile-local digit start
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
            uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = tot;
            tot += c;
        }
        scan_s[d] = tot;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            uint32_t x = d >= o ? scan_s[d - o] : 0;
            __syncthreads();
            scan_s[d] += x;
            __syncthreads();
        }
        lstart[d] = scan_s[d] - tot;
        __syncthreads();
        // keys into LDS in digit order (stable: wave prefix + rank)
        uint32_t lpos[kRounds];
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            uint64_t idx = base + r * 64 + lane;
            lpos[r] = kTile;
            if (idx < n) {
                const uint32_t dd = rank[r] >> 16;
                lpos[r] = lstart[dd] + wcnt[w][dd] + (rank[r] & 0xFFFFu);
                if (lpos[r] < kTile) {
                    buf[lpos[r]] = k[r];
                    dig8[lpos[r]] = (uint8_t)dd;
                }
            }
        }
        __syncthreads();
        const uint32_t tn = tbase + kTile <= n ? kTile : (uint32_t)(n - tbase);
        uint64_t gpos[kRounds];
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            uint32_t i = r * kThreads + threadIdx.x;
            gpos[r] = n;
            if (i < tn) {
                uint64_t key = buf[i];
                uint32_t dd = dig8[i];
                gpos[r] = (uint64_t)run[dd] + (i - lstart[dd]);
                if (gpos[r] < n) kout[gpos[r]] = key;     // always true when the counts match
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
        run[d] += tot;
        __syncthreads();
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

// One stable digit pass; needs radix_histogram's offsets for this pass and
// w.status of at least radix_work_words(n) words.
template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st) {
    if (n == 0) return hipSuccess;
    const uint64_t per = radix_tiles_per_range(n), nr = radix_ranges(n);
    uint32_t* cnt = w.status;
    uint32_t* off = w.status + nr * 256;
    hipLaunchKernelGGL((k_upsweep<Digit>), dim3((unsigned)nr), dim3(kThreads), 0, st, kin, n, dig, pass,
                       per * kTile, (uint32_t)nr, cnt);
    hipLaunchKernelGGL(k_range_scan, dim3(256), dim3(kThreads), 0, st, cnt, (uint32_t)nr,
                       (const unsigned long long*)w.offsets, pass, off);
    hipLaunchKernelGGL((k_downsweep<Digit>), dim3((unsigned)nr), dim3(kThreads), 0, st, kin, vin, kout, vout, n,
                       dig, pass, per, (uint32_t)nr, off);
    return hipGetLastError();
}

}  // namespace vx
