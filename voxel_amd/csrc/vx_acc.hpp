// vx_acc.hpp -- device accumulators shared by the scan kernels
// (vx_scan.hip) and the batched query kernel (vx_batch.hip): the
// compensated COUNT/SUM/MIN/MAX accumulator, its deterministic workgroup
// reduction, the last-workgroup ticket and the 16-byte streaming load.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_device.hpp"
#include "vx_kernels.hpp"

namespace vx {
namespace {

constexpr int kBlock = 256;

template <typename T>
struct Acc {
    using S = typename TyInfo<T>::S;
    S s, c;
    uint64_t n;        // rows counted by merges
    uint32_t k;        // rows counted by add (folded into n before any merge)
    T mn, mx;
    __device__ __forceinline__ void init() { s = S(0); c = S(0); n = 0; k = 0; mn = Lim<T>::hi(); mx = Lim<T>::lo(); }
    __device__ __forceinline__ void fold() { n += k; k = 0; }
    __device__ __forceinline__ void add(bool p, T x) {
        k += p ? 1u : 0u;
        if constexpr (TyInfo<T>::F) {
            two_sum<S>(s, c, p ? x : T(0));
            // NaN fails both compares, so NaN rows do not reach MIN/MAX
            mn = (p && x < mn) ? x : mn;
            mx = (p && x > mx) ? x : mx;
        } else {
            s += p ? (uint64_t)(int64_t)x : 0ull;
            mn = (p && x < mn) ? x : mn;
            mx = (p && x > mx) ? x : mx;
        }
    }
    __device__ __forceinline__ void merge(const Acc& o) {
        n += o.n + o.k;
        if constexpr (TyInfo<T>::F) {
            S cc = c + o.c;
            two_sum<S>(s, cc, o.s);
            c = cc;
            mn = o.mn < mn ? o.mn : mn;
            mx = o.mx > mx ? o.mx : mx;
        } else {
            s += o.s;
            mn = o.mn < mn ? o.mn : mn;
            mx = o.mx > mx ? o.mx : mx;
        }
    }
    __device__ __forceinline__ void shfl_merge(int mask) {
        Acc o;
        o.k = 0;
        o.s = __shfl_xor(s, mask, 64);
        o.c = __shfl_xor(c, mask, 64);
        o.n = __shfl_xor(n, mask, 64);
        o.mn = __shfl_xor(mn, mask, 64);
        o.mx = __shfl_xor(mx, mask, 64);
        merge(o);
    }
    __device__ __forceinline__ void store(ScanOut* p) const {
        p->s = to_bits(s); p->c = to_bits(c); p->cnt = n + k; p->mn = to_bits(mn); p->mx = to_bits(mx);
    }
    __device__ __forceinline__ void load(const ScanOut* p) {
        s = from_bits<S>(p->s); c = from_bits<S>(p->c); n = p->cnt; k = 0;
        mn = from_bits<T>(p->mn); mx = from_bits<T>(p->mx);
    }
};
// Workgroup reduction of an Acc into thread 0 (deterministic order).
template <typename T>
__device__ __forceinline__ void block_reduce(Acc<T>& a, Acc<T>* lds) {
    a.fold();
    for (int m = 32; m > 0; m >>= 1) a.shfl_merge(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < kBlock / 64; ++w) a.merge(lds[w]);
    }
}

// Publishes this workgroup's partial and returns true in the workgroup that
// finishes last; that workgroup then sees every other partial.
__device__ __forceinline__ bool publish_and_check_last(uint32_t* ticket, bool* s_last) {
    if (threadIdx.x == 0) {
        __threadfence();
        uint32_t prev = __hip_atomic_fetch_add(ticket, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        *s_last = (prev == gridDim.x - 1);
    }
    __syncthreads();
    bool last = *s_last;
    if (last) __threadfence();
    return last;
}

// 16-byte vector load; NT marks it non-temporal (streamed, not kept in cache).
template <typename V, bool NT>
__device__ __forceinline__ V ld16(const V* p) {
    if constexpr (NT && sizeof(V) == 16) {
        typedef uint32_t u4 __attribute__((ext_vector_type(4)));
        u4 r = __builtin_nontemporal_load(reinterpret_cast<const u4*>(p));
        V v;
        __builtin_memcpy(&v, &r, 16);
        return v;
    } else {
        return *p;
    }
}

// One workgroup: per-workgroup partials of a scan, merged in workgroup order
// (the same order as the last-workgroup combine of vx_scan.hip).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_final(const ScanOut* __restrict__ parts, uint32_t nblocks,
                                                       ScanOut* __restrict__ out) {
    __shared__ Acc<T> lds[kBlock / 64];
    Acc<T> r; r.init();
    for (uint32_t i = threadIdx.x; i < nblocks; i += kBlock) {
        Acc<T> o; o.load(&parts[i]);
        r.merge(o);
    }
    block_reduce(r, lds);
    if (threadIdx.x == 0) r.store(out);
}

}  // namespace
}  // namespace vx
