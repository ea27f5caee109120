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

// Digit p of a monotone bucket id computed from the key by a splitter/slope
// table (learned CDF partitioner, vx_learned.hip): bucket(k) is
// non-decreasing in k, so a stable sort by bucket followed by a stable sort
// inside each bucket sorts by key.
struct TableDigit {
    const uint64_t* cell_lo;    // key at the start of each cell (ascending), ncells entries
    const uint32_t* cell_b0;    // bucket at the start of each cell (non-decreasing)
    const float* cell_slope;    // buckets per key unit inside the cell
    uint32_t ncells;
    uint32_t nbuckets;
    int shift_bits;             // bits of bucket id per digit pass (8)
    uint64_t kmin;              // smallest key (cell grid origin)
    int cell_shift;             // cell index = float-bits grid (see vx_learned.hip)
    __device__ uint32_t bucket(uint64_t k) const;
    __device__ __forceinline__ uint32_t digit(uint64_t k, int p) const {
        return (bucket(k) >> (shift_bits * p)) & 0xFFu;
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
