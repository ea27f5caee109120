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
    // all passes' digits from one evaluation: bits(k) then digit p of it
    __host__ __device__ __forceinline__ uint64_t bits(uint64_t k) const { return k; }
    __host__ __device__ __forceinline__ uint32_t digit_of_bits(uint64_t b, int p) const {
        return (uint32_t)(b >> (8 * p)) & 0xFFu;
    }
};

// Digit p of a monotone "fine bucket" id fb(k) computed from the key with a
// piecewise-linear table (learned CDF partitioner, vx_learned.hip).  Keys are
// clamped to [kmin, kmax] and mapped to a grid coordinate G(d), d = k - kmin:
// d itself (linear grid), or the bit pattern of (float)d + 1 minus that of
// 1.0f (logarithmic grid: 2^23 steps per octave, G(0) = 0).  Cell c = G >>
// cshift covers fine buckets [lo, lo + width] (cell[c] = {lo, width}, lo of
// cell c + 1 = lo + width); inside it, with the offset o = G mod 2^cshift
// cut to its top w = min(cshift, 24) bits (o >> (cshift - w), exact as a
// float), fb = lo + floor(width * (o >> (cshift - w)) * 2^-w), clamped to
// lo + width.  Every step is non-decreasing in k, so fb is non-decreasing
// in k: a stable sort by fb followed by a stable sort inside each fb run
// sorts by key.  Partition pass p uses bits [8p, 8p + 8) of fb.
struct TableDigit {
    const uint2* cell;          // ncells entries {first fine bucket, fine buckets in the cell}
    uint64_t kmin, kmax;
    uint32_t ncells;
    int cshift;
    int logmode;
    uint32_t nfine;             // total buckets
    __host__ __device__ static __forceinline__ uint64_t grid(uint64_t d, int logmode) {
        if (!logmode) return d;
        float f = (float)d + 1.0f;
        return (uint64_t)(__builtin_bit_cast(uint32_t, f) - 0x3F800000u);
    }
    // fb(k) with the cell table read from tab (cell itself, or a copy of it
    // in LDS)
    __device__ __forceinline__ uint32_t fine_in(const uint2* tab, uint64_t k) const {
        const uint64_t kc = k < kmin ? kmin : (k > kmax ? kmax : k);
        const uint64_t g = grid(kc - kmin, logmode);
        const uint64_t c64 = g >> cshift;
        const bool past = c64 >= ncells;                     // not reached by keys in [kmin, kmax]
        const uint2 e = tab[past ? ncells - 1 : (uint32_t)c64];
        const int rs = cshift > 24 ? cshift - 24 : 0, w = cshift - rs;     // uniform: scalar work
        const uint32_t o = (uint32_t)(g >> rs) & ((1u << w) - 1u);
        const float f = floorf((float)e.y * ((float)o * __builtin_bit_cast(float, (uint32_t)(127 - w) << 23)));
        const uint32_t add = past ? e.y : min((uint32_t)f, e.y);
        const uint32_t fb = e.x + add;
        return fb < nfine ? fb : nfine - 1;
    }
    __device__ __forceinline__ uint32_t fine(uint64_t k) const { return fine_in(cell, k); }
    __device__ __forceinline__ uint32_t digit(uint64_t k, int p) const {
        return (fine(k) >> (8 * p)) & 0xFFu;
    }
    __device__ __forceinline__ uint64_t bits(uint64_t k) const { return fine(k); }
    __device__ __forceinline__ uint32_t digit_of_bits(uint64_t b, int p) const {
        return (uint32_t)(b >> (8 * p)) & 0xFFu;
    }
};

// Device work areas of one sort.
struct RadixWork {
    unsigned long long* hist = nullptr;     // 2048 words: (or, and) of the keys per workgroup
    uint32_t* skip = nullptr;               // 8 flags: pass moves nothing
    uint32_t* status = nullptr;             // per-range digit counts and offsets (radix_work_words)
    uint64_t status_cap = 0;                // words
};

uint64_t radix_work_words(uint64_t n);     // words of RadixWork::status a pass over n keys needs
hipError_t launch_zero32(uint32_t* p, uint64_t n, hipStream_t st);
hipError_t launch_f64_keys(const double* x, uint64_t n, int descending, uint64_t* keys, uint32_t* idx, hipStream_t st);
hipError_t launch_iota(uint32_t* idx, uint64_t n, hipStream_t st);
hipError_t launch_gather_f64(const double* x, const uint32_t* idx, uint64_t k, double* out, hipStream_t st);
hipError_t launch_f64_to_i64(const double* x, uint64_t n, int64_t* out, hipStream_t st);

hipError_t launch_constant_bytes(const uint64_t* keys, uint64_t n, RadixWork& w, hipStream_t st);
template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st, const uint32_t* fin,
                      uint32_t* fout, bool row_index = false, uint32_t* fpre = nullptr);

}  // namespace vx
