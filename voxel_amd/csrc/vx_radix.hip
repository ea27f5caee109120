// vx_radix.hip -- stable LSD radix sort of (u64 key, u32 payload) pairs on
// gfx950, replacing RadixSort<u64>::Sort(data, indices, count)
// (ops/sort.hpp:83-160) and, through an order-preserving key transform,
// MergeSort<f64>/SortOperator argsorts (ops/sort.hpp:196-213, 506-541).
//
// Structure:
//   1. k_hist: one read of the keys builds all eight 8-bit digit histograms
//      (per-wave LDS counters, then global 64-bit counters);
//   2. k_scan_hist: exclusive digit offsets per pass; a pass whose digit is
//      the same for every key is skipped (a stable pass over one digit is
//      the identity permutation);
//   3. per executed pass (reduce-then-scan, no inter-workgroup waits): the
//      keys are cut into at most 2048 ranges of whole 4096-key tiles;
//      k_upsweep counts each range's digits, k_range_scan gives every range
//      its first output position per digit, and k_downsweep (one workgroup
//      per range, tiles in order) ranks each tile's keys stably (wave
//      ballots of the digit bits + per-wave LDS counters) and writes keys and
//      payloads through LDS so that each digit run is written by consecutive
//      lanes.
// A decoupled look-back over 4096-key tiles (one status word per tile and
// digit) was measured at about 0.1 TB/s here: the status loads are coherent
// across the eight XCDs and the look-back walks them one at a time; the
// upsweep costs one extra read of the keys (8 of 32 bytes per key and pass)
// and leaves every workgroup independent.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_kernels.hpp"
#include "vx_onesweep.hpp"
#include "vx_radix.hpp"

namespace vx {
namespace {

__global__ __launch_bounds__(256) void k_zero32(uint32_t* __restrict__ p, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0;
}

// offsets[p][d] = sum of hist[p][d'] for d' < d; skip[p] = one digit holds all keys
__global__ __launch_bounds__(256) void k_scan_hist(const unsigned long long* __restrict__ ghist, int passes,
                                                   uint64_t n, unsigned long long* __restrict__ offsets,
                                                   uint32_t* __restrict__ skip) {
    __shared__ unsigned long long s[256];
    for (int p = 0; p < passes; ++p) {
        unsigned long long v = ghist[p * 256 + threadIdx.x];
        s[threadIdx.x] = v;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            unsigned long long t = threadIdx.x >= off ? s[threadIdx.x - off] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        offsets[p * 256 + threadIdx.x] = s[threadIdx.x] - v;
        if (v == n) skip[p] = 1;
        __syncthreads();
    }
}

// f64 -> u64 with the same order as the f64 compare (-0.0 folded onto +0.0);
// descending flips every bit.  Writes the row index payload.
__global__ __launch_bounds__(256) void k_f64_keys(const double* __restrict__ x, uint64_t n, int descending,
                                                  uint64_t* __restrict__ keys, uint32_t* __restrict__ idx) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t b = (uint64_t)__double_as_longlong(x[i]);
        if (b == 0x8000000000000000ull) b = 0;
        b = (b >> 63) ? ~b : (b | 0x8000000000000000ull);
        keys[i] = descending ? ~b : b;
        idx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(256) void k_gather_f64(const double* __restrict__ x, const uint32_t* __restrict__ idx,
                                                   uint64_t k, double* __restrict__ out) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < k; i += stride) out[i] = x[idx[i]];
}

__global__ __launch_bounds__(256) void k_iota(uint32_t* __restrict__ idx, uint64_t n) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) idx[i] = (uint32_t)i;
}

unsigned grid_for(uint64_t n, unsigned cap) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b < cap ? (b ? b : 1) : cap);
}

}  // namespace

hipError_t launch_zero32(uint32_t* p, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero32, dim3(grid_for(n, 8192)), dim3(256), 0, st, p, n);
    return hipGetLastError();
}

hipError_t launch_f64_keys(const double* x, uint64_t n, int descending, uint64_t* keys, uint32_t* idx, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f64_keys, dim3(grid_for(n, 16384)), dim3(256), 0, st, x, n, descending, keys, idx);
    return hipGetLastError();
}

hipError_t launch_gather_f64(const double* x, const uint32_t* idx, uint64_t k, double* out, hipStream_t st) {
    if (k == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gather_f64, dim3(grid_for(k, 16384)), dim3(256), 0, st, x, idx, k, out);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t* idx, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 16384)), dim3(256), 0, st, idx, n);
    return hipGetLastError();
}

uint64_t radix_work_words(uint64_t n) { return radix_ranges(n) * 256 * 2 + 256; }

hipError_t launch_scan_hist(const unsigned long long* ghist, int passes, uint64_t n, unsigned long long* offsets,
                            uint32_t* skip, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_hist, dim3(1), dim3(256), 0, st, ghist, passes, n, offsets, skip);
    return hipGetLastError();
}

template hipError_t radix_histogram<LsdDigit>(const uint64_t*, uint64_t, const LsdDigit&, int, RadixWork&, hipStream_t);
template hipError_t radix_pass<LsdDigit>(const uint64_t*, const uint32_t*, uint64_t*, uint32_t*, uint64_t,
                                         const LsdDigit&, int, RadixWork&, hipStream_t, const uint32_t*, uint32_t*);

}  // namespace vx
