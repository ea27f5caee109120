// vx_radix.hip -- stable LSD radix sort of (u64 key, u32 payload) pairs on
// gfx950, replacing RadixSort<u64>::Sort(data, indices, count)
// (ops/sort.hpp:83-160) and, through an order-preserving key transform,
// MergeSort<f64>/SortOperator argsorts (ops/sort.hpp:196-213, 506-541).
//
// Structure:
//   1. k_key_bits: one read of the keys gives the OR and the AND of all
//      keys; a pass whose byte is the same in every key is skipped (a
//      stable pass over one digit is the identity permutation);
//   2. (the digit offsets of each pass come from its own tile counts);
//   3. per executed pass (reduce-then-scan, no inter-workgroup waits): the
//      keys are cut into 8192-key tiles; k_tile_counts counts each tile's
//      digits, three small kernels turn the counts into every tile's first
//      output position per digit, and k_scatter_tile (one workgroup per
//      tile, tiles handed to the XCDs in contiguous blocks) ranks the tile's
//      keys stably (wave ballots of the digit bits + per-wave LDS counters)
//      and writes keys and payloads through LDS so that each digit run is
//      written by consecutive lanes.
// A decoupled look-back over tiles (one status word per tile and digit) was
// measured at about 0.1 TB/s here: the status loads are coherent across the
// eight XCDs and the look-back walks them one at a time; the count pass
// costs one extra read of the keys (8 of 32 bytes per key and pass) and
// leaves every workgroup independent.
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

// part[2 b] = OR, part[2 b + 1] = AND of the keys workgroup b reads.
__global__ __launch_bounds__(256) void k_key_bits(const uint64_t* __restrict__ keys, uint64_t n,
                                                  unsigned long long* __restrict__ part) {
    __shared__ unsigned long long so[4], sa[4];
    uint64_t o = 0, a = ~0ull;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    for (; i + 3 * stride < n; i += 4 * stride) {
        const uint64_t k0 = keys[i], k1 = keys[i + stride], k2 = keys[i + 2 * stride], k3 = keys[i + 3 * stride];
        o |= (k0 | k1) | (k2 | k3);
        a &= (k0 & k1) & (k2 & k3);
    }
    for (; i < n; i += stride) {
        o |= keys[i];
        a &= keys[i];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        o |= (uint64_t)__shfl_xor((unsigned long long)o, off, 64);
        a &= (uint64_t)__shfl_xor((unsigned long long)a, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { so[w] = o; sa[w] = a; }
    __syncthreads();
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = so[0] | so[1] | so[2] | so[3];
        part[2 * blockIdx.x + 1] = sa[0] & sa[1] & sa[2] & sa[3];
    }
}

// skip[p] = byte p of (OR ^ AND) is zero: every key has the same byte p
__global__ __launch_bounds__(256) void k_key_bits_final(const unsigned long long* __restrict__ part, unsigned nb,
                                                        uint64_t n, uint32_t* __restrict__ skip) {
    __shared__ unsigned long long so[4], sa[4];
    uint64_t o = 0, a = ~0ull;
    for (unsigned b = threadIdx.x; b < nb; b += 256) {
        o |= part[2 * b];
        a &= part[2 * b + 1];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        o |= (uint64_t)__shfl_xor((unsigned long long)o, off, 64);
        a &= (uint64_t)__shfl_xor((unsigned long long)a, off, 64);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { so[w] = o; sa[w] = a; }
    __syncthreads();
    if (threadIdx.x < 8) {
        const uint64_t d = (so[0] | so[1] | so[2] | so[3]) ^ (sa[0] & sa[1] & sa[2] & sa[3]);
        skip[threadIdx.x] = (n == 0 || ((d >> (8 * threadIdx.x)) & 0xFFu) == 0) ? 1u : 0u;
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

// (int64_t) of a double as x86-64's cvttsd2si converts it: toward zero,
// INT64_MIN for NaN and values outside int64's range.
__global__ __launch_bounds__(256) void k_f64_to_i64(const double* __restrict__ x, uint64_t n,
                                                    int64_t* __restrict__ out) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const double v = x[i];
        out[i] = (v != v || v >= 9223372036854775808.0 || v < -9223372036854775808.0) ? INT64_MIN : (int64_t)v;
    }
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

hipError_t launch_f64_to_i64(const double* x, uint64_t n, int64_t* out, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_f64_to_i64, dim3(grid_for(n, 16384)), dim3(256), 0, st, x, n, out);
    return hipGetLastError();
}

hipError_t launch_iota(uint32_t* idx, uint64_t n, hipStream_t st) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_iota, dim3(grid_for(n, 16384)), dim3(256), 0, st, idx, n);
    return hipGetLastError();
}

uint64_t radix_work_words(uint64_t n) { return (radix_tiles(n) + radix_chunks(n)) * 256 + 256; }

hipError_t launch_constant_bytes(const uint64_t* keys, uint64_t n, RadixWork& w, hipStream_t st) {
    unsigned nb = (unsigned)(n / (256 * 16));               // >= 16 keys per thread
    const unsigned cap = (unsigned)(device_cu_count() * 4) < 1024u ? (unsigned)(device_cu_count() * 4) : 1024u;
    nb = nb < 1 ? 1 : (nb > cap ? cap : nb);
    hipLaunchKernelGGL(k_key_bits, dim3(nb), dim3(256), 0, st, keys, n, w.hist);
    hipLaunchKernelGGL(k_key_bits_final, dim3(1), dim3(256), 0, st, (const unsigned long long*)w.hist, nb, n, w.skip);
    return hipGetLastError();
}

template hipError_t radix_pass<LsdDigit>(const uint64_t*, const uint32_t*, uint64_t*, uint32_t*, uint64_t,
                                         const LsdDigit&, int, RadixWork&, hipStream_t, const uint32_t*, uint32_t*,
                                         bool, uint32_t*);

}  // namespace vx
