// vx_encoding.hip -- columnar encodings on the GPU (data/encoding.hpp:104-527):
// DecodeBatch of Dictionary, RLE, Delta, BitPacked and FrameOfReference
// columns, and filter + COUNT/SUM/MIN/MAX scans that read the encoded bytes
// and decode in registers (the decoded column is never written):
//   * Dictionary: 4-byte indices (16-byte loads, four rows each), values
//     looked up in the dictionary (LDS when it fits, else L2);
//   * FrameOfReference: 4-byte codes, value = Base + (f64)code * Scale;
//   * BitPacked: BitWidth bits per row read from the packed words;
//   * RLE: one thread per run, the run's rows counted and summed at once
//     (an f64 run sum v * len is kept exact as a product plus its FMA error);
//   * Delta: the running sum needs every earlier row, so rows are decoded by
//     a reduce-then-scan (wrapping integer sums: any order gives the same
//     bits) and the plain scan runs on the result.
// Aggregates use the scan's accumulator (vx_acc.hpp): compensated float sums,
// wrapping integer sums, per-workgroup partials merged in workgroup order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <utility>

#include "../../include/voxel_c.h"
#include "vx_acc.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

constexpr int kEncThreads = 256;
constexpr int kScanItems = 16;                            // rows per thread in the prefix sums
constexpr int kScanChunk = kEncThreads * kScanItems;      // rows per workgroup
constexpr uint32_t kDictLds = 4096;                       // dictionary entries staged in LDS
constexpr int kEncUnroll = 4;                             // 16-byte loads in flight per lane

unsigned grid_for(uint64_t n, unsigned per, unsigned cap) {
    uint64_t b = (n + per - 1) / per;
    if (b < 1) b = 1;
    return (unsigned)(b < cap ? b : cap);
}

// ---- decoders ---------------------------------------------------------------

// A code at or past the dictionary's size is never read through: its row
// decodes as 0 and *bad is set, so the call fails instead of reading past
// the dictionary.
template <typename U>
__global__ __launch_bounds__(kEncThreads) void k_dict_decode(const U* __restrict__ dict, uint32_t nd,
                                                             const uint32_t* __restrict__ idx, uint64_t off,
                                                             uint64_t cnt, U* __restrict__ out, uint32_t* __restrict__ bad) {
    const uint64_t stride = (uint64_t)gridDim.x * kEncThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kEncThreads + threadIdx.x; i < cnt; i += stride) {
        const uint32_t k = idx[off + i];
        if (k < nd) {
            out[i] = dict[k];
        } else {
            out[i] = U(0);
            *bad = 1u;
        }
    }
}

// RLE: run r covers rows [start[r], start[r] + len[r]); rows past the last
// run decode as 0 (RLEEncoding::DecodeBatch writes T{} there).
template <typename U>
__global__ __launch_bounds__(kEncThreads) void k_rle_decode(const U* __restrict__ vals, const uint64_t* __restrict__ start,
                                                            uint64_t nruns, uint64_t total, uint64_t off, uint64_t cnt,
                                                            U* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kEncThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kEncThreads + threadIdx.x; i < cnt; i += stride) {
        const uint64_t row = off + i;
        if (row >= total) { out[i] = U(0); continue; }
        uint64_t lo = 0, hi = nruns;                      // last run whose start <= row
        while (hi - lo > 1) {
            const uint64_t mid = (lo + hi) >> 1;
            if (start[mid] <= row) lo = mid; else hi = mid;
        }
        out[i] = vals[lo];
    }
}

__device__ __forceinline__ uint32_t unpack(const uint64_t* __restrict__ w, uint64_t nw, uint32_t bw, uint64_t row) {
    const uint64_t pos = row * bw;                        // BitPackedEncoding::Decode (encoding.hpp:397-409)
    const uint32_t sh = (uint32_t)(pos & 63);
    uint64_t r = w[pos >> 6] >> sh;
    if (sh + bw > 64 && (pos >> 6) + 1 < nw) r |= w[(pos >> 6) + 1] << (64 - sh);
    return (uint32_t)(r & ((1ull << bw) - 1));
}

__global__ __launch_bounds__(kEncThreads) void k_bitpack_decode(const uint64_t* __restrict__ w, uint64_t nw, uint32_t bw,
                                                                uint64_t off, uint64_t cnt, uint32_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kEncThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kEncThreads + threadIdx.x; i < cnt; i += stride) out[i] = unpack(w, nw, bw, off + i);
}

__global__ __launch_bounds__(kEncThreads) void k_for_decode(const int32_t* __restrict__ enc, double base, double scale,
                                                            uint64_t off, uint64_t cnt, double* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * kEncThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kEncThreads + threadIdx.x; i < cnt; i += stride)
        out[i] = base + (double)enc[off + i] * scale;     // FrameOfReferenceEncoding::Decode, two roundings
}

// ---- wrapping prefix sums (RLE run starts, Delta running values) -------------

// part[b] = sum of x[b*C, min((b+1)*C, n)) in U (wrapping).
template <typename In, typename U>
__global__ __launch_bounds__(kEncThreads) void k_chunk_sum(const In* __restrict__ x, uint64_t n, U* __restrict__ part) {
    __shared__ U red[kEncThreads];
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    U s = 0;
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kEncThreads + threadIdx.x;
        if (i < n) s += (U)x[i];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int m = kEncThreads / 2; m > 0; m >>= 1) {
        if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// One workgroup: part[] -> exclusive prefix sums, in place, in tiles of
// kScanChunk entries staged through LDS (coalesced reads and writes).
template <typename U>
__global__ __launch_bounds__(kEncThreads) void k_parts_scan(U* __restrict__ part, uint64_t np) {
    __shared__ U tile[kScanChunk + kScanChunk / 16];
    __shared__ U sc[kEncThreads];
    __shared__ U carry_s;
    auto at = [](int i) { return i + (i >> 4); };         // pad: a thread's 16 items avoid one bank
    U carry = 0;
    for (uint64_t t0 = 0; t0 < np; t0 += kScanChunk) {
        for (int k = 0; k < kScanItems; ++k) {
            const int t = k * kEncThreads + threadIdx.x;
            tile[at(t)] = t0 + t < np ? part[t0 + t] : U(0);
        }
        __syncthreads();
        U s = 0;
        for (int k = 0; k < kScanItems; ++k) s += tile[at(threadIdx.x * kScanItems + k)];
        sc[threadIdx.x] = s;
        __syncthreads();
        for (int o = 1; o < kEncThreads; o <<= 1) {
            const U v = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : U(0);
            __syncthreads();
            sc[threadIdx.x] += v;
            __syncthreads();
        }
        U run = carry + (threadIdx.x ? sc[threadIdx.x - 1] : U(0));
        for (int k = 0; k < kScanItems; ++k) {
            const int t = at(threadIdx.x * kScanItems + k);
            const U v = tile[t];
            tile[t] = run;
            run += v;
        }
        if (threadIdx.x == kEncThreads - 1) carry_s = run;
        __syncthreads();
        for (int k = 0; k < kScanItems; ++k) {
            const int t = k * kEncThreads + threadIdx.x;
            if (t0 + t < np) part[t0 + t] = tile[at(t)];
        }
        carry = carry_s;
        __syncthreads();
    }
}

// Inclusive prefix of chunk b's rows (carry = exclusive prefix of the chunk)
// handed to emit(row, inclusive, item): each thread owns kScanItems
// consecutive rows, staged through LDS so the global reads are coalesced.
template <typename In, typename U, class Emit>
__global__ __launch_bounds__(kEncThreads) void k_chunk_scan(const In* __restrict__ x, uint64_t n,
                                                            const U* __restrict__ carry, Emit emit) {
    __shared__ U tile[kScanChunk + kScanChunk / 16];
    __shared__ U sc[kEncThreads];
    auto at = [](int i) { return i + (i >> 4); };         // pad: a thread's 16 items avoid one bank
    const uint64_t base = (uint64_t)blockIdx.x * kScanChunk;
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = base + (uint64_t)k * kEncThreads + threadIdx.x;
        tile[at(k * kEncThreads + threadIdx.x)] = i < n ? (U)x[i] : U(0);
    }
    __syncthreads();
    U s = 0;
    for (int k = 0; k < kScanItems; ++k) s += tile[at(threadIdx.x * kScanItems + k)];
    sc[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < kEncThreads; o <<= 1) {
        const U v = (int)threadIdx.x >= o ? sc[threadIdx.x - o] : U(0);
        __syncthreads();
        sc[threadIdx.x] += v;
        __syncthreads();
    }
    U run = carry[blockIdx.x] + (threadIdx.x ? sc[threadIdx.x - 1] : U(0));
    for (int k = 0; k < kScanItems; ++k) {
        const int t = at(threadIdx.x * kScanItems + k);
        run += tile[t];
        tile[t] = run;
    }
    __syncthreads();
    for (int k = 0; k < kScanItems; ++k) {
        const int t = k * kEncThreads + threadIdx.x;
        const uint64_t i = base + (uint64_t)t;
        if (i < n) emit(i, tile[at(t)]);
    }
}

struct RunStarts {                                        // start[r] = rows before run r
    const uint32_t* len;
    uint64_t* start;
    uint32_t* zero;                                       // set when a run is empty
    __device__ __forceinline__ void operator()(uint64_t i, uint64_t incl) const {
        const uint32_t l = len[i];
        start[i] = incl - l;
        if (l == 0) *zero = 1;
    }
};

template <typename T, typename U>
struct DeltaOut {                                         // out[row - off] = base + running sum
    T* out;
    uint64_t off, cnt;
    U base;
    __device__ __forceinline__ void operator()(uint64_t i, U incl) const {
        if (i >= off && i - off < cnt) out[i - off] = (T)(U)(base + incl);
    }
};

// ---- Delta: filter + aggregate without the decoded column -------------------
// Workgroup b owns rows [b * span, min(n, (b + 1) * span)), span a multiple of
// kScanChunk.  k_delta_sum sums each workgroup's deltas (wrapping in U),
// k_parts_scan turns the sums into each workgroup's starting value, and
// k_delta_scan walks the workgroup's rows in kScanChunk tiles from that value:
// each thread owns 16 consecutive rows of a tile (staged through LDS so the
// loads stay coalesced), a wave shuffle scan and the four wave totals give
// each thread the value before its rows, and the rows enter the filter +
// aggregate in registers.  The deltas are read twice and nothing is written
// but one partial per workgroup.

// VEC: 16-byte loads (x 16-byte aligned).
template <typename T, typename U, bool VEC>
__global__ __launch_bounds__(kEncThreads) void k_delta_sum(const T* __restrict__ x, uint64_t n, uint64_t span,
                                                           U* __restrict__ part) {
    __shared__ U red[kEncThreads / 64];
    const uint64_t b0 = (uint64_t)blockIdx.x * span;
    const uint64_t b1 = b0 + span < n ? b0 + span : n;
    U s = 0;
    uint64_t i0 = b0;
    if constexpr (VEC) {
        using V = typename TyInfo<T>::V;
        constexpr int E = TyInfo<T>::E;
        constexpr int kU = 8;                             // vector loads in flight per lane
        const V* q = reinterpret_cast<const V*>(x);
        const uint64_t v0 = b0 / E, v1 = b1 / E;
        uint64_t v = v0 + threadIdx.x;
        for (; v + (kU - 1) * kEncThreads < v1; v += kU * kEncThreads) {
            V r[kU];
#pragma unroll
            for (int u = 0; u < kU; ++u) r[u] = ld16<V, true>(&q[v + u * kEncThreads]);
#pragma unroll
            for (int u = 0; u < kU; ++u) {
                const T* e = reinterpret_cast<const T*>(&r[u]);
#pragma unroll
                for (int k = 0; k < E; ++k) s += (U)e[k];
            }
        }
        for (; v < v1; v += kEncThreads) {
            const V r = ld16<V, true>(&q[v]);
            const T* e = reinterpret_cast<const T*>(&r);
#pragma unroll
            for (int k = 0; k < E; ++k) s += (U)e[k];
        }
        i0 = v1 * E;
    }
    for (uint64_t i = i0 + threadIdx.x; i < b1; i += kEncThreads) s += (U)x[i];
    for (int m = 32; m > 0; m >>= 1) s += __shfl_xor(s, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

template <typename T, typename U>
__global__ __launch_bounds__(kEncThreads) void k_delta_scan(const T* __restrict__ x, uint64_t n, uint64_t span,
                                                            const U* __restrict__ start, U base, PredDev pd,
                                                            ScanOut* __restrict__ parts) {
    __shared__ U tile[kScanChunk + kScanChunk / 16];
    __shared__ U wtot[kEncThreads / 64];
    __shared__ Acc<T> red[kEncThreads / 64];
    auto at = [](int i) { return i + (i >> 4); };         // pad: a thread's 16 items avoid one bank
    Pred<T, false> p; p.load(pd);
    Acc<T> acc; acc.init();
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    const uint64_t b0 = (uint64_t)blockIdx.x * span;
    const uint64_t b1 = b0 + span < n ? b0 + span : n;
    U run = base + start[blockIdx.x];                     // value before the tile's first row
    // the next tile's deltas are loaded into registers while this one is
    // scanned, so the loads of a workgroup overlap its LDS work
    U pf[kScanItems];
#pragma unroll
    for (int k = 0; k < kScanItems; ++k) {
        const uint64_t i = b0 + (uint64_t)k * kEncThreads + threadIdx.x;
        pf[k] = i < b1 ? (U)x[i] : U(0);
    }
    for (uint64_t t0 = b0; t0 < b1; t0 += kScanChunk) {
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) tile[at(k * kEncThreads + threadIdx.x)] = pf[k];
        __syncthreads();
        const uint64_t t1 = t0 + kScanChunk;
        if (t1 < b1) {
#pragma unroll
            for (int k = 0; k < kScanItems; ++k) {
                const uint64_t i = t1 + (uint64_t)k * kEncThreads + threadIdx.x;
                pf[k] = i < b1 ? (U)x[i] : U(0);
            }
        }
        U v[kScanItems];
        U s = 0;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = tile[at(threadIdx.x * kScanItems + k)];
            s += v[k];
        }
        U incl = s;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const U y = __shfl_up(incl, o, 64);
            if (lane >= o) incl += y;
        }
        if (lane == 63) wtot[wid] = incl;
        __syncthreads();
        U r = run + incl - s;
        for (int w = 0; w < wid; ++w) r += wtot[w];
        run += wtot[0] + wtot[1] + wtot[2] + wtot[3];
        const uint64_t row0 = t0 + (uint64_t)threadIdx.x * kScanItems;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            r += v[k];
            const T val = (T)r;
            acc.add(row0 + k < b1 && p(val), val);
        }
        __syncthreads();                                  // tile and wtot are rewritten by the next tile
    }
    block_reduce(acc, red);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

// ---- fused scans -------------------------------------------------------------

// Codes at or past nd are read as code 0 (in bounds: nd > kDictLds puts at
// least one entry in global memory, and a smaller dictionary is read from
// LDS) and set *bad, which fails the call.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_dict(const T* __restrict__ dict, uint32_t nd,
                                                      const uint32_t* __restrict__ idx, uint64_t n, PredDev pd,
                                                      ScanOut* __restrict__ parts, uint32_t* __restrict__ bad) {
    __shared__ T ld[kDictLds];
    __shared__ Acc<T> red[kBlock / 64];
    const bool in_lds = nd <= kDictLds;
    if (in_lds)
        for (uint32_t i = threadIdx.x; i < nd; i += kBlock) ld[i] = dict[i];
    __syncthreads();
    const T* dv = in_lds ? ld : dict;
    Pred<T, false> p; p.load(pd);
    Acc<T> acc; acc.init();
    // rows before the first 16-byte boundary, and the last n % 4, one each
    uint64_t head = ((16u - ((uintptr_t)idx & 15u)) & 15u) / 4;
    if (head > n) head = n;
    const uint64_t nq = (n - head) / 4;                   // 16-byte loads of four indices
    const uint64_t tail0 = head + nq * 4;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    bool oob = false;
    auto code = [&](uint32_t k) -> uint32_t {
        oob |= k >= nd;
        return k < nd ? k : 0u;
    };
    if (g < head + (n - tail0)) {
        const uint64_t r = g < head ? g : tail0 + (g - head);
        const T x = dv[code(idx[r])];
        acc.add(p(x), x);
    }
    const uint4* q = reinterpret_cast<const uint4*>(idx + head);
    const uint64_t tile = (uint64_t)kBlock * kEncUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < nq; base += (uint64_t)gridDim.x * tile) {
        uint4 k[kEncUnroll];
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u) {            // the tile's loads are in flight together
            const uint64_t v = base + u * kBlock + threadIdx.x;
            if (v < nq) k[u] = ld16<uint4, true>(&q[v]);
        }
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u) {
            if (base + u * kBlock + threadIdx.x < nq) {
                const T x0 = dv[code(k[u].x)], x1 = dv[code(k[u].y)], x2 = dv[code(k[u].z)], x3 = dv[code(k[u].w)];
                acc.add(p(x0), x0); acc.add(p(x1), x1); acc.add(p(x2), x2); acc.add(p(x3), x3);
            }
        }
    }
    if (oob) *bad = 1u;
    block_reduce(acc, red);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

// Dictionary scan over a dictionary of at most kDictLds entries: the
// predicate and the aggregates depend only on a row's code, so each
// workgroup counts its rows per code in an LDS histogram (one LDS atomic per
// row, no per-row float work) and adds the nonzero counts into a global
// per-code count; k_dict_final then filters the codes and aggregates count x
// value per code.  Codes at or past nd set *bad (the call fails).
__global__ __launch_bounds__(kBlock) void k_dict_hist(uint32_t nd, const uint32_t* __restrict__ idx, uint64_t n,
                                                      unsigned long long* __restrict__ counts,
                                                      uint32_t* __restrict__ bad) {
    __shared__ uint32_t h[kDictLds];
    for (uint32_t i = threadIdx.x; i < nd; i += kBlock) h[i] = 0u;
    __syncthreads();
    bool oob = false;
    auto add = [&](uint32_t k) {
        if (k < nd) atomicAdd(&h[k], 1u);
        else oob = true;
    };
    uint64_t head = ((16u - ((uintptr_t)idx & 15u)) & 15u) / 4;
    if (head > n) head = n;
    const uint64_t nq = (n - head) / 4;                   // 16-byte loads of four indices
    const uint64_t tail0 = head + nq * 4;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g < head + (n - tail0)) add(idx[g < head ? g : tail0 + (g - head)]);
    const uint4* q = reinterpret_cast<const uint4*>(idx + head);
    const uint64_t tile = (uint64_t)kBlock * kEncUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < nq; base += (uint64_t)gridDim.x * tile) {
        uint4 k[kEncUnroll];
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u) {            // the tile's loads are in flight together
            const uint64_t v = base + u * kBlock + threadIdx.x;
            if (v < nq) k[u] = ld16<uint4, true>(&q[v]);
        }
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u)
            if (base + u * kBlock + threadIdx.x < nq) { add(k[u].x); add(k[u].y); add(k[u].z); add(k[u].w); }
    }
    if (oob) *bad = 1u;
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < nd; i += kBlock)
        if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

// One workgroup: COUNT / SUM / MIN / MAX over the codes that pass, each code
// entering with its row count c: float sums as the exact product c * x (its
// rounded value plus the FMA error) in a compensated double pair, merged in
// thread order; integer sums wrapping; the result as a scan partial of T.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dict_final(const T* __restrict__ dict, uint32_t nd,
                                                       const unsigned long long* __restrict__ counts, PredDev pd,
                                                       ScanOut* __restrict__ out) {
    constexpr bool F = TyInfo<T>::F;
    __shared__ double ss[kBlock], sc[kBlock];
    __shared__ unsigned long long sn[kBlock], si[kBlock];
    __shared__ T smn[kBlock], smx[kBlock];
    Pred<T, false> p; p.load(pd);
    double s = 0.0, c = 0.0;
    unsigned long long cnt = 0, isum = 0;
    T mn = Lim<T>::hi(), mx = Lim<T>::lo();
    for (uint32_t k = threadIdx.x; k < nd; k += kBlock) {
        const unsigned long long ck = counts[k];
        const T x = dict[k];
        if (ck == 0 || !p(x)) continue;
        cnt += ck;
        if constexpr (F) {
            const double cd = (double)ck, xd = (double)x;
            const double ph = cd * xd;
            const double pl = __builtin_fma(cd, xd, -ph);
            two_sum<double>(s, c, ph);
            c += __builtin_isfinite(ph) ? pl : 0.0;
        } else {
            isum += (unsigned long long)ck * (unsigned long long)(long long)x;
        }
        mn = x < mn ? x : mn;
        mx = x > mx ? x : mx;
    }
    ss[threadIdx.x] = s; sc[threadIdx.x] = c; sn[threadIdx.x] = cnt; si[threadIdx.x] = isum;
    smn[threadIdx.x] = mn; smx[threadIdx.x] = mx;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int t = 1; t < kBlock; ++t) {                // thread order
            double cc = c + sc[t];
            two_sum<double>(s, cc, ss[t]);
            c = cc;
            cnt += sn[t];
            isum += si[t];
            mn = smn[t] < mn ? smn[t] : mn;
            mx = smx[t] > mx ? smx[t] : mx;
        }
        ScanOut r{};
        r.cnt = cnt;
        if constexpr (F) {
            const T hi = (T)(s + c);                      // the pair in T: rounded total and its remainder
            const T lo = (T)((s - (double)hi) + c);
            r.s = to_bits(hi);
            r.c = to_bits(lo);
        } else {
            r.s = isum;
            r.c = 0;
        }
        r.mn = to_bits(mn);
        r.mx = to_bits(mx);
        *out = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_scan_for(const int32_t* __restrict__ enc, double base, double scale,
                                                     uint64_t n, PredDev pd, ScanOut* __restrict__ parts) {
    __shared__ Acc<double> red[kBlock / 64];
    Pred<double, false> p; p.load(pd);
    Acc<double> acc; acc.init();
    auto one = [&](int32_t c) {
        const double x = base + (double)c * scale;
        acc.add(p(x), x);
    };
    uint64_t head = ((16u - ((uintptr_t)enc & 15u)) & 15u) / 4;
    if (head > n) head = n;
    const uint64_t nq = (n - head) / 4;
    const uint64_t tail0 = head + nq * 4;
    const uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (g < head + (n - tail0)) one(enc[g < head ? g : tail0 + (g - head)]);
    const int4* q = reinterpret_cast<const int4*>(enc + head);
    const uint64_t tile = (uint64_t)kBlock * kEncUnroll;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < nq; base += (uint64_t)gridDim.x * tile) {
        int4 k[kEncUnroll];
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u) {
            const uint64_t v = base + u * kBlock + threadIdx.x;
            if (v < nq) k[u] = ld16<int4, true>(&q[v]);
        }
#pragma unroll
        for (int u = 0; u < kEncUnroll; ++u)
            if (base + u * kBlock + threadIdx.x < nq) { one(k[u].x); one(k[u].y); one(k[u].z); one(k[u].w); }
    }
    block_reduce(acc, red);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

// BitPacked scan.  Row r's BW bits start at bit r * BW of the packed words
// (little-endian u64 words read as a stream of u32 words, which is the same
// bit stream), so 32 consecutive rows fill exactly BW u32 words.  Each thread
// takes such groups of 32 rows: it loads the group's BW words (16- or 8-byte
// vector loads when BW allows), and with BW a template constant every row's
// word and shift are constants, so a row costs one bit-field extract or one
// funnel shift.  The predicate is the 64-bit range predicate restricted to
// u32 values, a wrapped range (x - a) mod 2^32 <= d (en = 0: no row passes).
// Sums of a group of 32 rows stay in u32 while 32 * (2^BW - 1) fits; MIN/MAX
// run on the u32 values (their order is the i64 order).  COUNT, the wrapping
// i64 SUM, MIN and MAX do not depend on the order of the rows, so the result
// is the same as a row-by-row scan.
struct U32Range {
    uint32_t a, d, en;
};

template <int BW, int VW>
__global__ __launch_bounds__(kBlock) void k_scan_packed(const uint32_t* __restrict__ w, uint64_t groups,
                                                        const uint64_t* __restrict__ w64, uint64_t nw, uint64_t n,
                                                        U32Range pr, ScanOut* __restrict__ parts) {
    using GS = typename std::conditional<(BW <= 27), uint32_t, uint64_t>::type;
    constexpr uint32_t kMask = BW == 32 ? 0xFFFFFFFFu : ((1u << BW) - 1u);
    __shared__ Acc<int64_t> red[kBlock / 64];
    uint64_t sum = 0;
    uint32_t cnt = 0, mn = 0xFFFFFFFFu, mx = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t g = (uint64_t)blockIdx.x * kBlock + threadIdx.x; pr.en && g < groups; g += stride) {
        uint32_t v[BW];
        const uint32_t* src = w + g * BW;
        if constexpr (VW == 4) {
#pragma unroll
            for (int i = 0; i < BW / 4; ++i) {
                const uint4 q = ld16<uint4, true>(reinterpret_cast<const uint4*>(src) + i);
                v[4 * i] = q.x; v[4 * i + 1] = q.y; v[4 * i + 2] = q.z; v[4 * i + 3] = q.w;
            }
        } else if constexpr (VW == 2) {
#pragma unroll
            for (int i = 0; i < BW / 2; ++i) {
                const uint2 q = reinterpret_cast<const uint2*>(src)[i];
                v[2 * i] = q.x; v[2 * i + 1] = q.y;
            }
        } else {
#pragma unroll
            for (int i = 0; i < BW; ++i) v[i] = src[i];
        }
        GS gs = 0;
#pragma unroll
        for (int j = 0; j < 32; ++j) {
            const int o = j * BW, q = o >> 5, sh = o & 31;
            uint32_t x;
            if (sh + BW <= 32) x = (v[q] >> sh) & kMask;
            else x = __builtin_amdgcn_alignbit(v[q + 1], v[q], (uint32_t)sh) & kMask;
            const bool p = x - pr.a <= pr.d;
            const uint32_t sel = p ? x : 0u;
            gs += sel;
            cnt += p ? 1u : 0u;
            mx = sel > mx ? sel : mx;
            const uint32_t lo = p ? x : 0xFFFFFFFFu;
            mn = lo < mn ? lo : mn;
        }
        sum += gs;
    }
    // rows past the last whole group (fewer than 32), row by row
    if (pr.en && blockIdx.x == 0 && threadIdx.x == 0) {
        for (uint64_t r = groups * 32; r < n; ++r) {
            const uint32_t x = unpack(w64, nw, BW, r);
            if (x - pr.a <= pr.d) {
                sum += x; ++cnt;
                mx = x > mx ? x : mx;
                mn = x < mn ? x : mn;
            }
        }
    }
    Acc<int64_t> acc;
    acc.init();
    acc.s = sum;
    acc.k = cnt;
    if (mn <= mx) { acc.mn = (int64_t)mn; acc.mx = (int64_t)mx; }   // mn > mx: no row of this thread passed
    block_reduce(acc, red);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

// One thread per run: the run's len rows enter COUNT/SUM/MIN/MAX at once.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_scan_rle(const T* __restrict__ vals, const uint32_t* __restrict__ lens,
                                                     uint64_t nruns, PredDev pd, ScanOut* __restrict__ parts) {
    __shared__ Acc<T> red[kBlock / 64];
    Pred<T, false> p; p.load(pd);
    Acc<T> acc; acc.init();
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t r = (uint64_t)blockIdx.x * kBlock + threadIdx.x; r < nruns; r += stride) {
        const T x = vals[r];
        const uint32_t len = lens[r];
        const bool ok = p(x) && len != 0;
        acc.fold();
        acc.n += ok ? len : 0;
        if constexpr (TyInfo<T>::F) {
            // x * len as two exact products (len = hi + lo, each exact in the
            // float type), each as its rounded value plus the FMA error
            using S = typename TyInfo<T>::S;
            const uint32_t lo = len & 0xFFFu;
            const S L[2] = {(S)(len - lo), (S)lo};
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const S ph = (S)x * L[h];
                const S pl = __builtin_fma((S)x, L[h], -ph);
                two_sum<S>(acc.s, acc.c, ok ? ph : S(0));
                acc.c += ok && __builtin_isfinite(ph) ? pl : S(0);
            }
        } else {
            acc.s += ok ? (uint64_t)(int64_t)x * (uint64_t)len : 0ull;
        }
        acc.mn = (ok && x < acc.mn) ? x : acc.mn;
        acc.mx = (ok && x > acc.mx) ? x : acc.mx;
    }
    block_reduce(acc, red);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

// ---- host side ----------------------------------------------------------------

struct EncBufs {
    void* tmp = nullptr;           // decoded Delta column / prefix parts
    size_t tmp_bytes = 0;
    uint64_t* starts = nullptr;    // RLE run starts
    size_t starts_n = 0;
    void* parts = nullptr;         // chunk sums
    size_t parts_bytes = 0;
    uint32_t* flag_dev = nullptr;
    uint32_t* flag_host = nullptr;
    bool grow(void** p, size_t* have, size_t need) {
        if (need <= *have && *p) return true;
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        *have = 0;
        if (!hip_ok(hipMalloc(p, need ? need : 8), "hipMalloc")) return false;
        *have = need;
        return true;
    }
    bool flag() {
        if (flag_dev) return true;
        return hip_ok(hipMalloc((void**)&flag_dev, 4), "hipMalloc") &&
               hip_ok(hipHostMalloc((void**)&flag_host, 4, hipHostMallocDefault), "hipHostMalloc");
    }
};
EncBufs& enc_bufs() {                                     // kept to process end (no hipFree at exit)
    static EncBufs* b = new EncBufs();
    return *b;
}

__global__ void k_clear_u32(uint32_t* p) { *p = 0; }

bool wait(StreamCtx* c) { return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); }

// RLE run starts into b.starts; false (error set) for an empty run.
bool rle_starts(const uint32_t* lens, uint64_t nr, uint64_t* total, StreamCtx* c) {
    EncBufs& b = enc_bufs();
    size_t have = b.starts_n * 8;
    if (!b.grow((void**)&b.starts, &have, nr * 8)) return false;
    b.starts_n = have / 8;
    const uint64_t np = (nr + kScanChunk - 1) / kScanChunk;
    if (!b.grow(&b.parts, &b.parts_bytes, np * 8) || !b.flag()) return false;
    uint64_t* parts = static_cast<uint64_t*>(b.parts);
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(1), 0, c->stream, b.flag_dev);
    hipLaunchKernelGGL((k_chunk_sum<uint32_t, uint64_t>), dim3((unsigned)np), dim3(kEncThreads), 0, c->stream, lens, nr, parts);
    hipLaunchKernelGGL((k_parts_scan<uint64_t>), dim3(1), dim3(kEncThreads), 0, c->stream, parts, np);
    hipLaunchKernelGGL((k_chunk_scan<uint32_t, uint64_t, RunStarts>), dim3((unsigned)np), dim3(kEncThreads), 0, c->stream,
                       lens, nr, (const uint64_t*)parts, RunStarts{lens, b.starts, b.flag_dev});
    if (!hip_ok(hipGetLastError(), "RLE run starts launch")) return false;
    uint64_t last[2] = {0, 0};
    if (!hip_ok(hipMemcpyAsync(&last[0], b.starts + (nr - 1), 8, hipMemcpyDeviceToHost, c->stream), "read back") ||
        !hip_ok(hipMemcpyAsync(b.flag_host, b.flag_dev, 4, hipMemcpyDeviceToHost, c->stream), "read back"))
        return false;
    uint32_t lastlen = 0;
    if (!hip_ok(hipMemcpyAsync(&lastlen, lens + (nr - 1), 4, hipMemcpyDeviceToHost, c->stream), "read back") || !wait(c))
        return false;
    if (*b.flag_host) { set_error("RLE column has an empty run (RLEEncoding::Build never writes one)"); return false; }
    *total = last[0] + lastlen;
    return true;
}

// Delta rows [off, off + cnt) decoded into out (device).
template <typename T, typename U>
bool delta_decode(const T* deltas, uint64_t rows, U base, uint64_t off, uint64_t cnt, T* out, StreamCtx* c) {
    if (cnt == 0) return true;
    const uint64_t end = off + cnt;                       // rows needed: [0, end)
    EncBufs& b = enc_bufs();
    const uint64_t np = (end + kScanChunk - 1) / kScanChunk;
    if (!b.grow(&b.parts, &b.parts_bytes, np * sizeof(U))) return false;
    U* parts = static_cast<U*>(b.parts);
    hipLaunchKernelGGL((k_chunk_sum<T, U>), dim3((unsigned)np), dim3(kEncThreads), 0, c->stream, deltas, end, parts);
    hipLaunchKernelGGL((k_parts_scan<U>), dim3(1), dim3(kEncThreads), 0, c->stream, parts, np);
    hipLaunchKernelGGL((k_chunk_scan<T, U, DeltaOut<T, U>>), dim3((unsigned)np), dim3(kEncThreads), 0, c->stream, deltas,
                       end, (const U*)parts, DeltaOut<T, U>{out, off, cnt, base});
    (void)rows;
    return hip_ok(hipGetLastError(), "delta decode launch");
}

bool check_col(const voxel_gpu_encoded_col* e) {
    if (!e) { set_error("null encoded column"); return false; }
    switch (e->kind) {
    case VOXEL_ENC_DICTIONARY:
    case VOXEL_ENC_RLE:
        if (e->dtype != VOXEL_GPU_F64 && e->dtype != VOXEL_GPU_I64 && e->dtype != VOXEL_GPU_F32 &&
            e->dtype != VOXEL_GPU_I32) { set_error("encoded column: dtype %d not supported", e->dtype); return false; }
        if (e->nvalues && (!e->values || !e->index)) { set_error("encoded column: null array"); return false; }
        if (e->kind == VOXEL_ENC_DICTIONARY && e->rows && (!e->index || !e->values)) {
            set_error("encoded column: null array"); return false;
        }
        return true;
    case VOXEL_ENC_DELTA:
        if (e->dtype != VOXEL_GPU_I64 && e->dtype != VOXEL_GPU_I32) { set_error("Delta columns are i64 or i32"); return false; }
        if (e->rows && !e->values) { set_error("encoded column: null array"); return false; }
        return true;
    case VOXEL_ENC_BITPACKED:
        if (e->bit_width < 1 || e->bit_width > 32) { set_error("BitPacked width %u outside 1..32", e->bit_width); return false; }
        if (e->rows && (!e->values || e->nvalues < (e->rows * e->bit_width + 63) / 64)) {
            set_error("BitPacked column: %llu words hold fewer than %llu rows", (unsigned long long)e->nvalues,
                      (unsigned long long)e->rows);
            return false;
        }
        return true;
    case VOXEL_ENC_FOR:
        if (e->rows && !e->values) { set_error("encoded column: null array"); return false; }
        return true;
    default:
        set_error("encoding kind %d not supported", e->kind);
        return false;
    }
}

template <typename U>
void launch_dict_decode(const voxel_gpu_encoded_col* e, uint64_t off, uint64_t cnt, void* out, uint32_t* bad,
                        StreamCtx* c) {
    hipLaunchKernelGGL((k_dict_decode<U>), dim3(grid_for(cnt, kEncThreads, 65536)), dim3(kEncThreads), 0, c->stream,
                       static_cast<const U*>(e->values), (uint32_t)e->nvalues, e->index, off, cnt,
                       static_cast<U*>(out), bad);
}

// Clears the code-range flag before a dictionary launch.
bool dict_flag(StreamCtx* c) {
    EncBufs& b = enc_bufs();
    if (!b.flag()) return false;
    hipLaunchKernelGGL(k_clear_u32, dim3(1), dim3(1), 0, c->stream, b.flag_dev);
    return hip_ok(hipGetLastError(), "clear");
}
// After the stream has finished: false (error set) when a code was out of range.
bool dict_codes_ok(const voxel_gpu_encoded_col* e) {
    if (*enc_bufs().flag_host) {
        set_error("Dictionary column: an index is not below the dictionary's %llu values",
                  (unsigned long long)e->nvalues);
        return false;
    }
    return true;
}
template <typename U>
void launch_rle_decode(const voxel_gpu_encoded_col* e, uint64_t total, uint64_t off, uint64_t cnt, void* out,
                       StreamCtx* c) {
    hipLaunchKernelGGL((k_rle_decode<U>), dim3(grid_for(cnt, kEncThreads, 65536)), dim3(kEncThreads), 0, c->stream,
                       static_cast<const U*>(e->values), (const uint64_t*)enc_bufs().starts, e->nvalues, total, off,
                       cnt, static_cast<U*>(out));
}

int64_t decode(const voxel_gpu_encoded_col* e, uint64_t off, uint64_t cnt, void* out, StreamCtx* c) {
    const size_t es = (e->kind == VOXEL_ENC_BITPACKED) ? 4 : e->kind == VOXEL_ENC_FOR ? 8 : dtype_size(e->dtype);
    if (e->kind == VOXEL_ENC_RLE) {                       // writes cnt rows, 0 past the last run
        if (cnt == 0) return 0;
        if (!out) { set_error("decode: null output"); return -1; }
        if (e->nvalues == 0) {
            if (!hip_ok(hipMemsetAsync(out, 0, cnt * es, c->stream), "clear") || !wait(c)) return -1;
            return (int64_t)cnt;
        }
        uint64_t total = 0;
        if (!rle_starts(e->index, e->nvalues, &total, c)) return -1;
        if (es == 8) launch_rle_decode<uint64_t>(e, total, off, cnt, out, c);
        else launch_rle_decode<uint32_t>(e, total, off, cnt, out, c);
        if (!hip_ok(hipGetLastError(), "RLE decode launch") || !wait(c)) return -1;
        return (int64_t)cnt;
    }
    const uint64_t n = off < e->rows ? (cnt < e->rows - off ? cnt : e->rows - off) : 0;
    if (n == 0) return 0;
    if (!out) { set_error("decode: null output"); return -1; }
    switch (e->kind) {
    case VOXEL_ENC_DICTIONARY: {
        if (!dict_flag(c)) return -1;
        uint32_t* bad = enc_bufs().flag_dev;
        if (es == 8) launch_dict_decode<uint64_t>(e, off, n, out, bad, c);
        else launch_dict_decode<uint32_t>(e, off, n, out, bad, c);
        if (!hip_ok(hipGetLastError(), "decode launch") ||
            !hip_ok(hipMemcpyAsync(enc_bufs().flag_host, bad, 4, hipMemcpyDeviceToHost, c->stream), "read back") ||
            !wait(c))
            return -1;
        return dict_codes_ok(e) ? (int64_t)n : -1;
    }
    case VOXEL_ENC_DELTA:
        if (e->dtype == VOXEL_GPU_I64) {
            if (!delta_decode<int64_t, uint64_t>(static_cast<const int64_t*>(e->values), e->rows,
                                                 (uint64_t)e->delta_base, off, n, static_cast<int64_t*>(out), c))
                return -1;
        } else if (!delta_decode<int32_t, uint32_t>(static_cast<const int32_t*>(e->values), e->rows,
                                                    (uint32_t)(int32_t)e->delta_base, off, n,
                                                    static_cast<int32_t*>(out), c)) {
            return -1;
        }
        break;
    case VOXEL_ENC_BITPACKED:
        hipLaunchKernelGGL(k_bitpack_decode, dim3(grid_for(n, kEncThreads, 65536)), dim3(kEncThreads), 0, c->stream,
                           static_cast<const uint64_t*>(e->values), e->nvalues, e->bit_width, off, n,
                           static_cast<uint32_t*>(out));
        break;
    default:
        hipLaunchKernelGGL(k_for_decode, dim3(grid_for(n, kEncThreads, 65536)), dim3(kEncThreads), 0, c->stream,
                           static_cast<const int32_t*>(e->values), e->base, e->scale, off, n, static_cast<double*>(out));
        break;
    }
    if (!hip_ok(hipGetLastError(), "decode launch") || !wait(c)) return -1;
    return (int64_t)n;
}

// Dictionaries of at most kDictLds entries: per-code counts, then the
// aggregate from the dictionary (result straight into *res).
template <typename T>
bool launch_dict_hist(const voxel_gpu_encoded_col* e, PredDev pd, unsigned blocks, ScanOut* res, StreamCtx* c) {
    EncBufs& b = enc_bufs();
    if (!b.grow(&b.parts, &b.parts_bytes, (size_t)e->nvalues * 8)) return false;
    unsigned long long* counts = static_cast<unsigned long long*>(b.parts);
    if (e->nvalues &&
        !hip_ok(launch_zero32(reinterpret_cast<uint32_t*>(counts), e->nvalues * 2, c->stream), "clear")) return false;
    hipLaunchKernelGGL(k_dict_hist, dim3(blocks), dim3(kBlock), 0, c->stream, (uint32_t)e->nvalues, e->index, e->rows,
                       counts, enc_bufs().flag_dev);
    hipLaunchKernelGGL((k_dict_final<T>), dim3(1), dim3(kBlock), 0, c->stream, static_cast<const T*>(e->values),
                       (uint32_t)e->nvalues, (const unsigned long long*)counts, pd, res);
    return hip_ok(hipGetLastError(), "dictionary scan launch");
}

template <typename T>
void launch_dict_scan(const voxel_gpu_encoded_col* e, PredDev pd, ScanOut* parts, unsigned blocks, StreamCtx* c) {
    hipLaunchKernelGGL((k_scan_dict<T>), dim3(blocks), dim3(kBlock), 0, c->stream, static_cast<const T*>(e->values),
                       (uint32_t)e->nvalues, e->index, e->rows, pd, parts, enc_bufs().flag_dev);
}
template <typename T>
void launch_rle_scan(const voxel_gpu_encoded_col* e, PredDev pd, ScanOut* parts, unsigned blocks, StreamCtx* c) {
    hipLaunchKernelGGL((k_scan_rle<T>), dim3(blocks), dim3(kBlock), 0, c->stream, static_cast<const T*>(e->values),
                       e->index, e->nvalues, pd, parts);
}
// The i64 range predicate (lo, hi, inv) restricted to u32 values as a wrapped
// range: (x - a) mod 2^32 <= d.
U32Range u32_range(const PredDev& pd) {
    const int64_t lo = (int64_t)pd.lo, hi = (int64_t)pd.hi;
    const int64_t L = lo > 0 ? lo : 0, H = hi < 0xFFFFFFFFll ? hi : 0xFFFFFFFFll;
    const bool empty = L > H, full = L == 0 && H == 0xFFFFFFFFll;
    if (!pd.inv) {
        if (empty) return {0u, 0u, 0u};
        return {(uint32_t)L, (uint32_t)(H - L), 1u};
    }
    if (empty) return {0u, 0xFFFFFFFFu, 1u};             // the complement of nothing
    if (full) return {0u, 0u, 0u};
    return {(uint32_t)(H + 1), (uint32_t)(L - H - 2), 1u}; // [H + 1, L - 1] around the u32 circle
}

template <int BW>
void launch_packed_bw(const voxel_gpu_encoded_col* e, U32Range r, ScanOut* parts, unsigned blocks, StreamCtx* c) {
    const uint64_t* w64 = static_cast<const uint64_t*>(e->values);
    const uint32_t* w = reinterpret_cast<const uint32_t*>(w64);
    const uint64_t groups = e->rows / 32;
    if constexpr (BW % 4 == 0) {
        if ((((uintptr_t)w) & 15u) == 0) {
            hipLaunchKernelGGL((k_scan_packed<BW, 4>), dim3(blocks), dim3(kBlock), 0, c->stream, w, groups, w64,
                               e->nvalues, e->rows, r, parts);
            return;
        }
    }
    hipLaunchKernelGGL((k_scan_packed<BW, BW % 2 == 0 ? 2 : 1>), dim3(blocks), dim3(kBlock), 0, c->stream, w, groups,
                       w64, e->nvalues, e->rows, r, parts);
}
using PackedLaunch = void (*)(const voxel_gpu_encoded_col*, U32Range, ScanOut*, unsigned, StreamCtx*);
template <int... I>
constexpr std::array<PackedLaunch, sizeof...(I)> packed_table(std::integer_sequence<int, I...>) {
    return {&launch_packed_bw<I + 1>...};
}
constexpr std::array<PackedLaunch, 32> kPackedLaunch = packed_table(std::make_integer_sequence<int, 32>{});

// Delta filter + aggregate: per-workgroup partials into parts, their count
// into *blocks.
template <typename T, typename U>
bool delta_scan(const voxel_gpu_encoded_col* e, U base, PredDev pd, ScanOut* parts, unsigned* blocks, StreamCtx* c) {
    const T* x = static_cast<const T*>(e->values);
    const uint64_t n = e->rows;
    const uint64_t nch = n ? (n + kScanChunk - 1) / kScanChunk : 1;
    const uint64_t per = (nch + c->scr.max_blocks - 1) / c->scr.max_blocks;
    const uint64_t nb = (nch + per - 1) / per;
    const uint64_t span = per * (uint64_t)kScanChunk;
    EncBufs& b = enc_bufs();
    if (!b.grow(&b.parts, &b.parts_bytes, nb * sizeof(U))) return false;
    U* starts = static_cast<U*>(b.parts);
    if ((((uintptr_t)x) & 15u) == 0)
        hipLaunchKernelGGL((k_delta_sum<T, U, true>), dim3((unsigned)nb), dim3(kEncThreads), 0, c->stream, x, n, span,
                           starts);
    else
        hipLaunchKernelGGL((k_delta_sum<T, U, false>), dim3((unsigned)nb), dim3(kEncThreads), 0, c->stream, x, n, span,
                           starts);
    hipLaunchKernelGGL((k_parts_scan<U>), dim3(1), dim3(kEncThreads), 0, c->stream, starts, nb);
    hipLaunchKernelGGL((k_delta_scan<T, U>), dim3((unsigned)nb), dim3(kEncThreads), 0, c->stream, x, n, span,
                       (const U*)starts, base, pd, parts);
    *blocks = (unsigned)nb;
    return hip_ok(hipGetLastError(), "delta scan launch");
}

template <typename T>
void launch_final(ScanOut* parts, unsigned blocks, ScanOut* out, StreamCtx* c) {
    hipLaunchKernelGGL((k_scan_final<T>), dim3(1), dim3(kBlock), 0, c->stream, (const ScanOut*)parts, blocks, out);
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

int64_t voxel_gpu_decode(const voxel_gpu_encoded_col* col, uint64_t offset, uint64_t count, void* out_dev, void* stream) {
    set_error("%s", "");
    if (!check_col(col)) return -1;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    return decode(col, offset, count, out_dev, c);
}

int voxel_gpu_scan_encoded(const voxel_gpu_encoded_col* col, int op, uint64_t t, void* stream, voxel_gpu_agg* out) {
    set_error("%s", "");
    if (!check_col(col)) return -1;
    if (!out) { set_error("scan_encoded: null result"); return -1; }
    if (op < 0 || op > 6) { set_error("scan_encoded: predicate code out of range"); return -1; }
    if (col->kind == VOXEL_ENC_DICTIONARY && col->nvalues > 0xFFFFFFFFull) { set_error("dictionary too large"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    // the aggregate's type: the value type (BitPacked u32 values as i64)
    const int adt = col->kind == VOXEL_ENC_BITPACKED ? DT_I64 : col->kind == VOXEL_ENC_FOR ? DT_F64 : col->dtype;
    uint64_t tb = t;
    if (col->kind == VOXEL_ENC_BITPACKED) tb = (uint64_t)(int64_t)(t & 0xFFFFFFFFull);   // u32 threshold
    const PredDev pd = make_pred(adt, op, tb);
    ScanOut* parts = static_cast<ScanOut*>(c->scr.partials);
    const unsigned cap = (unsigned)std::min<uint64_t>((uint64_t)device_cu_count() * 4, c->scr.max_blocks);
    {
        unsigned blocks = 1;
        switch (col->kind) {
        case VOXEL_ENC_DELTA: {
            bool ok = col->dtype == VOXEL_GPU_I64
                          ? delta_scan<int64_t, uint64_t>(col, (uint64_t)col->delta_base, pd, parts, &blocks, c)
                          : delta_scan<int32_t, uint32_t>(col, (uint32_t)(int32_t)col->delta_base, pd, parts, &blocks, c);
            if (!ok) return -1;
            break;
        }
        case VOXEL_ENC_DICTIONARY:
            if (!dict_flag(c)) return -1;
            blocks = grid_for(col->rows / 4 + 1, kBlock * kEncUnroll, cap);
            if (col->nvalues <= kDictLds) {               // per-code counts; the result is written directly
                bool ok;
                if (col->dtype == VOXEL_GPU_F64) ok = launch_dict_hist<double>(col, pd, blocks, &c->dev->scan, c);
                else if (col->dtype == VOXEL_GPU_F32) ok = launch_dict_hist<float>(col, pd, blocks, &c->dev->scan, c);
                else if (col->dtype == VOXEL_GPU_I64) ok = launch_dict_hist<int64_t>(col, pd, blocks, &c->dev->scan, c);
                else ok = launch_dict_hist<int32_t>(col, pd, blocks, &c->dev->scan, c);
                if (!ok) return -1;
                blocks = 0;                               // no partials to merge
                break;
            }
            if (col->dtype == VOXEL_GPU_F64) launch_dict_scan<double>(col, pd, parts, blocks, c);
            else if (col->dtype == VOXEL_GPU_F32) launch_dict_scan<float>(col, pd, parts, blocks, c);
            else if (col->dtype == VOXEL_GPU_I64) launch_dict_scan<int64_t>(col, pd, parts, blocks, c);
            else launch_dict_scan<int32_t>(col, pd, parts, blocks, c);
            break;
        case VOXEL_ENC_RLE:
            blocks = grid_for(col->nvalues, kBlock, cap);
            if (col->dtype == VOXEL_GPU_F64) launch_rle_scan<double>(col, pd, parts, blocks, c);
            else if (col->dtype == VOXEL_GPU_F32) launch_rle_scan<float>(col, pd, parts, blocks, c);
            else if (col->dtype == VOXEL_GPU_I64) launch_rle_scan<int64_t>(col, pd, parts, blocks, c);
            else launch_rle_scan<int32_t>(col, pd, parts, blocks, c);
            break;
        case VOXEL_ENC_BITPACKED:
            // eight workgroups per CU keep enough 32-row groups in flight
            blocks = grid_for(col->rows / 32, kBlock,
                              (unsigned)std::min<uint64_t>((uint64_t)device_cu_count() * 8, c->scr.max_blocks));
            kPackedLaunch[col->bit_width - 1](col, u32_range(pd), parts, blocks, c);
            break;
        default:
            blocks = grid_for(col->rows / 4 + 1, kBlock * kEncUnroll, cap);
            hipLaunchKernelGGL(k_scan_for, dim3(blocks), dim3(kBlock), 0, c->stream,
                               static_cast<const int32_t*>(col->values), col->base, col->scale, col->rows, pd, parts);
            break;
        }
        if (blocks == 0) {
        } else if (adt == DT_F64) launch_final<double>(parts, blocks, &c->dev->scan, c);
        else if (adt == DT_F32) launch_final<float>(parts, blocks, &c->dev->scan, c);
        else if (adt == DT_I64) launch_final<int64_t>(parts, blocks, &c->dev->scan, c);
        else launch_final<int32_t>(parts, blocks, &c->dev->scan, c);
        if (!hip_ok(hipGetLastError(), "encoded scan launch")) return -1;
    }
    if (!hip_ok(hipMemcpyAsync(&c->host->scan, &c->dev->scan, sizeof(ScanOut), hipMemcpyDeviceToHost, c->stream),
                "read back result"))
        return -1;
    if (col->kind == VOXEL_ENC_DICTIONARY &&
        !hip_ok(hipMemcpyAsync(enc_bufs().flag_host, enc_bufs().flag_dev, 4, hipMemcpyDeviceToHost, c->stream),
                "read back"))
        return -1;
    if (!wait(c)) return -1;
    if (col->kind == VOXEL_ENC_DICTIONARY && !dict_codes_ok(col)) return -1;
    fill_agg(adt, c->host->scan, out);
    return 0;
}

}  // extern "C"
