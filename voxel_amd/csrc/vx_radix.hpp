// vx_radix.hpp -- onesweep radix sort building blocks (vx_radix.hip) and the
// digit functions they are instantiated with.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace vx {

// 8-bit LSD digit p of a u64 key.
struct LsdDigit {
    __host__ __device__ __forceinline__ uint32_t digit(uint64_t k, int p) const {
        return (uint32_t)(k >> (8 * p)) & 0xFFu;
    }
};

// Digit p of a monotone "fine bucket" id fb(k) computed from the key with a
// piecewise-linear table (learned CDF partitioner, vx_learned.hip).  Keys are
// clamped to [kmin, kmax] and mapped to a grid coordinate G(k): the offset
// k - kmin itself (linear grid), or the bit pattern of (float)(k - kmin)
// (logarithmic grid: 2^23 steps per octave).  Cell c = G >> cshift covers
// fine buckets [b0[c], b0[c+1]]; inside it fb = b0[c] + floor(slope[c] *
// (G - c << cshift)), clamped to b0[c+1].  Every step is non-decreasing in
// k, so fb is non-decreasing in k: a stable sort by fb followed by a stable
// sort inside each fb run sorts by key.  Partition pass p uses bits
// [8 + 8p, 16 + 8p) of fb (the coarse bucket is fb >> 8; the low 8 bits
// split a bucket once more inside LDS).
struct TableDigit {
    const uint32_t* b0;         // ncells + 1 entries, non-decreasing, b0[ncells] = total fine buckets
    const float* slope;         // ncells entries
    uint64_t kmin, kmax;
    uint32_t ncells;
    int cshift;
    int logmode;
    uint32_t nfine;             // total fine buckets
    __device__ __forceinline__ uint32_t fine(uint64_t k) const {
        uint64_t kc = k < kmin ? kmin : (k > kmax ? kmax : k);
        uint64_t d = kc - kmin;
        uint64_t g = logmode ? (uint64_t)__float_as_uint((float)d) : d;
        uint64_t c = g >> cshift;
        if (c >= ncells) c = ncells - 1;
        uint64_t off = g - (c << cshift);
        uint32_t lo = b0[c], hi = b0[c + 1];
        float f = floorf(slope[c] * (float)off);
        uint32_t add = f <= 0.0f ? 0u : (f >= (float)(hi - lo) ? hi - lo : (uint32_t)f);
        uint32_t fb = lo + add;
        return fb < nfine ? fb : nfine - 1;
    }
    __device__ __forceinline__ uint32_t digit(uint64_t k, int p) const {
        return (fine(k) >> (8 + 8 * p)) & 0xFFu;
    }
};

// Device work areas of one sort.
struct RadixWork {
    unsigned long long* hist = nullptr;     // 8 x 256 digit counts
    unsigned long long* offsets = nullptr;  // 8 x 256 exclusive digit offsets
    uint32_t* skip = nullptr;               // 8 flags: pass moves nothing
    uint32_t* status = nullptr;             // tiles x 256 look-back words
    uint32_t* ticket = nullptr;             // tile id counter
    uint64_t status_cap = 0;                // words
};

uint64_t radix_tiles(uint64_t n);
hipError_t launch_zero32(uint32_t* p, uint64_t n, hipStream_t st);
hipError_t launch_f64_keys(const double* x, uint64_t n, int descending, uint64_t* keys, uint32_t* idx, hipStream_t st);
hipError_t launch_iota(uint32_t* idx, uint64_t n, hipStream_t st);
hipError_t launch_gather_f64(const double* x, const uint32_t* idx, uint64_t k, double* out, hipStream_t st);

template <class Digit>
hipError_t radix_histogram(const uint64_t* keys, uint64_t n, const Digit& dig, int passes, RadixWork& w,
                           hipStream_t st);
template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st);

}  // namespace vx
