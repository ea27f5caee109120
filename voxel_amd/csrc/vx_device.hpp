// vx_device.hpp -- device helpers shared by the gfx950 kernel files:
// element traits, bit-pattern casts, the VM predicate in range / truth-table
// form (built on the host by make_pred, vx_runtime.cpp), and TwoSum.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "vx_kernels.hpp"

namespace vx {

template <typename T> struct TyInfo;
template <> struct TyInfo<double>  { using V = double2;   static constexpr int E = 2; using S = double;   static constexpr bool F = true; };
template <> struct TyInfo<float>   { using V = float4;    static constexpr int E = 4; using S = float;    static constexpr bool F = true; };
template <> struct TyInfo<int64_t> { using V = longlong2; static constexpr int E = 2; using S = uint64_t; static constexpr bool F = false; };
template <> struct TyInfo<int32_t> { using V = int4;      static constexpr int E = 4; using S = uint64_t; static constexpr bool F = false; };

template <typename T> __device__ __forceinline__ T from_bits(uint64_t b);
template <> __device__ __forceinline__ double from_bits<double>(uint64_t b) { return __longlong_as_double((long long)b); }
template <> __device__ __forceinline__ float from_bits<float>(uint64_t b) { return __uint_as_float((uint32_t)b); }
template <> __device__ __forceinline__ int64_t from_bits<int64_t>(uint64_t b) { return (int64_t)b; }
template <> __device__ __forceinline__ int32_t from_bits<int32_t>(uint64_t b) { return (int32_t)(uint32_t)b; }
template <> __device__ __forceinline__ uint64_t from_bits<uint64_t>(uint64_t b) { return b; }

// Sum of v over a 256-thread workgroup, returned to every thread (one
// atomic per workgroup instead of per wave keeps a persistent grid's
// closing atomics on one address from queueing at the L2).
__device__ __forceinline__ uint64_t block_sum_256(uint64_t v) {
    for (int m = 32; m > 0; m >>= 1) v += __shfl_xor(v, m, 64);
    __shared__ uint64_t part[4];
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
    __syncthreads();
    return part[0] + part[1] + part[2] + part[3];
}

__device__ __forceinline__ uint64_t to_bits(double v) { return (uint64_t)__double_as_longlong(v); }
__device__ __forceinline__ uint64_t to_bits(float v) { return (uint64_t)__float_as_uint(v); }
__device__ __forceinline__ uint64_t to_bits(int64_t v) { return (uint64_t)v; }
__device__ __forceinline__ uint64_t to_bits(int32_t v) { return (uint64_t)(uint32_t)v; }
__device__ __forceinline__ uint64_t to_bits(uint64_t v) { return v; }

template <typename T> struct Lim;
template <> struct Lim<double>  { __device__ static double lo() { return -__builtin_huge_val(); } __device__ static double hi() { return __builtin_huge_val(); } };
template <> struct Lim<float>   { __device__ static float lo() { return -__builtin_huge_valf(); } __device__ static float hi() { return __builtin_huge_valf(); } };
template <> struct Lim<int64_t> { __device__ static int64_t lo() { return INT64_MIN; } __device__ static int64_t hi() { return INT64_MAX; } };
template <> struct Lim<int32_t> { __device__ static int32_t lo() { return INT32_MIN; } __device__ static int32_t hi() { return INT32_MAX; } };

// Predicate, range or truth-table form (PredDev in vx_kernels.hpp).
template <typename T, bool TABLE>
struct Pred {
    T lo, hi, t;
    bool inv;
    uint32_t table;
    __device__ __forceinline__ void load(const PredDev& p) {
        lo = from_bits<T>(p.lo); hi = from_bits<T>(p.hi); t = from_bits<T>(p.t);
        inv = p.inv != 0; table = p.table;
    }
    __device__ __forceinline__ bool operator()(T x) const {
        if constexpr (TABLE) {
            uint32_t st = (x < t ? 1u : 0u) | (x == t ? 2u : 0u) | (x > t ? 4u : 0u);
            return (table >> st) & 1u;
        } else {
            return ((x >= lo) & (x <= hi)) != inv;
        }
    }
};

// Compensated summation step (Knuth TwoSum): s + c tracks the exact sum.
template <typename S>
__device__ __forceinline__ void two_sum(S& s, S& c, S v) {
    S t = s + v;
    S bp = t - s;
    S e = (s - (t - bp)) + (v - bp);
    s = t;
    c += e;
}

}  // namespace vx
