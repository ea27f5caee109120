// vx_filter.hip -- bitmap filters, bitmap algebra and order-preserving
// compaction on gfx950, replacing
//   VectorFilter<T>::Apply{EQ,NE,LT,LE,GT,GE,Between}  ops/filter.hpp:42-288
//   ColumnFilter::Init/AndWith/OrWith/AndNotWith        ops/filter.hpp:528-570
//   SelectionVector::BuildFromBitmap                    ops/filter.hpp:378-399
//   simd::scalar::scalar_filter_gt_*                    simd/scalar.hpp:155,214,260,306
// A bitmap word is bit (i & 63) of word i / 64 = pred(x_i), tail bits 0,
// exactly the reference layout.  Each lane loads 16 bytes (E rows); the E
// wave ballots of one load cover E*64 rows = E words, assembled by spreading
// ballot bits (lane j's element e is row 64k + E*j' + e) on the scalar unit.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cstring>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_acc.hpp"
#include "vx_compact.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {
namespace {

// bits of x placed at every 2nd / 4th position
__device__ __forceinline__ uint64_t spread2(uint64_t x) {
    x &= 0xFFFFFFFFull;
    x = (x | (x << 16)) & 0x0000FFFF0000FFFFull;
    x = (x | (x << 8)) & 0x00FF00FF00FF00FFull;
    x = (x | (x << 4)) & 0x0F0F0F0F0F0F0F0Full;
    x = (x | (x << 2)) & 0x3333333333333333ull;
    x = (x | (x << 1)) & 0x5555555555555555ull;
    return x;
}
__device__ __forceinline__ uint64_t spread4(uint64_t x) {
    x &= 0xFFFFull;
    x = (x | (x << 24)) & 0x000000FF000000FFull;
    x = (x | (x << 12)) & 0x000F000F000F000Full;
    x = (x | (x << 6)) & 0x0303030303030303ull;
    x = (x | (x << 3)) & 0x1111111111111111ull;
    return x;
}

// Four 16-byte loads per column in flight per lane, eight workgroups of four
// waves per CU (measured on 2^28 f64 rows: 0.39 ms; eight loads with two
// workgroups per CU 0.41 ms; one load per step with sixteen 0.44 ms).
template <bool TWO> constexpr int kBitmapUnroll = 4;

template <typename T, bool TWO>
__global__ __launch_bounds__(256) void k_bitmap(const T* __restrict__ a, const T* __restrict__ b, uint64_t n,
                                                PredDev pda, PredDev pdb, uint64_t* __restrict__ words,
                                                unsigned long long* __restrict__ count) {
    using V = typename TyInfo<T>::V;
    constexpr int E = TyInfo<T>::E;
    constexpr int U = kBitmapUnroll<TWO>;          // chunks per wave step, their loads issued together
    Pred<T, false> pa; pa.load(pda);
    Pred<T, false> pb; pb.load(pdb);
    const int lane = threadIdx.x & 63;
    const uint64_t wave = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const uint64_t nwaves = ((uint64_t)gridDim.x * blockDim.x) >> 6;
    const uint64_t rows_per = 64ull * E;
    const uint64_t nw = (n + 63) / 64;
    uint32_t cnt = 0;
    for (uint64_t g = wave; g * rows_per * U < n; g += nwaves) {
        const uint64_t base = g * rows_per * U;
        T xa[U][E], xb[U][E];
        if (base + rows_per * U <= n) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t r0 = base + u * rows_per + (uint64_t)lane * E;
                V va = ld16<V, true>(reinterpret_cast<const V*>(a + r0));
                const T* pva = reinterpret_cast<const T*>(&va);
#pragma unroll
                for (int e = 0; e < E; ++e) xa[u][e] = pva[e];
                if constexpr (TWO) {
                    V vb = ld16<V, true>(reinterpret_cast<const V*>(b + r0));
                    const T* pvb = reinterpret_cast<const T*>(&vb);
#pragma unroll
                    for (int e = 0; e < E; ++e) xb[u][e] = pvb[e];
                }
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t r0 = base + u * rows_per + (uint64_t)lane * E;
#pragma unroll
                for (int e = 0; e < E; ++e) {
                    xa[u][e] = r0 + e < n ? a[r0 + e] : T(0);
                    if constexpr (TWO) xb[u][e] = r0 + e < n ? b[r0 + e] : T(0);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t blk = g * U + u;
            const uint64_t r0 = blk * rows_per + (uint64_t)lane * E;
            uint64_t bal[E];
#pragma unroll
            for (int e = 0; e < E; ++e) {
                bool p = (r0 + e < n) && pa(xa[u][e]);
                if constexpr (TWO) p = p && pb(xb[u][e]);
                cnt += p ? 1u : 0u;
                bal[e] = __ballot(p);
            }
            // word k covers rows blk*rows_per + 64k .. +63: lanes (64/E)k .. (64/E)(k+1)-1
            if (lane < E) {
                const int k = lane;
                uint64_t wd = 0;
                if constexpr (E == 2) {
#pragma unroll
                    for (int e = 0; e < 2; ++e) wd |= spread2(bal[e] >> (32 * k)) << e;
                } else {
#pragma unroll
                    for (int e = 0; e < 4; ++e) wd |= spread4(bal[e] >> (16 * k)) << e;
                }
                const uint64_t wi = blk * E + k;
                if (wi < nw) words[wi] = wd;
            }
        }
    }
    const uint64_t total = block_sum_256(cnt);
    if (threadIdx.x == 0 && total) atomicAdd(count, (unsigned long long)total);
}

// ColumnFilter combine: 0 AND, 1 OR, 2 AND-NOT; tail bits past `rows` cleared.
__global__ __launch_bounds__(256) void k_combine(uint64_t* __restrict__ act, const uint64_t* __restrict__ other,
                                                 uint64_t nw, uint64_t rows, int op,
                                                 unsigned long long* __restrict__ count) {
    uint32_t cnt = 0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nw; i += stride) {
        uint64_t v = op == 0 ? (act[i] & other[i]) : op == 1 ? (act[i] | other[i]) : (act[i] & ~other[i]);
        if (i == nw - 1 && (rows & 63)) v &= (1ull << (rows & 63)) - 1;
        act[i] = v;
        cnt += (uint32_t)__popcll(v);
    }
    const uint64_t total = block_sum_256(cnt);
    if (threadIdx.x == 0 && total) atomicAdd(count, (unsigned long long)total);
}

struct BitmapSrc {
    const uint64_t* words;
    uint32_t* out;
    __device__ __forceinline__ bool pass(uint64_t row) const { return (words[row >> 6] >> (row & 63)) & 1ull; }
    __device__ __forceinline__ void emit(uint64_t row, uint64_t k) const { out[k] = (uint32_t)row; }
};

template <typename T>
struct PredSrc {
    using V = T;
    const T* x;
    PredDev pd;
    T* vals;
    uint32_t* idx;
    __device__ __forceinline__ bool pass(uint64_t row) const {
        Pred<T, false> p;
        p.load(pd);
        return p(x[row]);
    }
    __device__ __forceinline__ T fetch(uint64_t row) const { return x[row]; }
    __device__ __forceinline__ void put(uint64_t row, uint64_t k, T v) const {
        if (vals) vals[k] = v;
        if (idx) idx[k] = (uint32_t)row;
    }
};

unsigned grid_cap(uint64_t work, unsigned cap) {
    uint64_t b = work ? work : 1;
    return (unsigned)(b < cap ? b : cap);
}

}  // namespace

CompactWork& compact_work(StreamCtx* c) {
    static std::vector<std::pair<StreamCtx*, CompactWork*>> all;
    for (auto& p : all) if (p.first == c) return *p.second;
    all.emplace_back(c, new CompactWork());
    return *all.back().second;
}

}  // namespace vx

using namespace vx;

namespace {
// VM comparison, or 7 = Between (t <= x && x <= t2, VectorFilter::ApplyBetween).
PredDev pred_for(int dtype, int op, uint64_t t, uint64_t t2) {
    if (op != 7) return make_pred(dtype, op, t);
    bool isf = dtype == DT_F64 || dtype == DT_F32;
    if (isf) {
        bool nan_t, nan_t2;
        if (dtype == DT_F64) {
            double a, b; std::memcpy(&a, &t, 8); std::memcpy(&b, &t2, 8);
            nan_t = a != a; nan_t2 = b != b;
        } else {
            uint32_t ua = (uint32_t)t, ub = (uint32_t)t2; float a, b;
            std::memcpy(&a, &ua, 4); std::memcpy(&b, &ub, 4);
            nan_t = a != a; nan_t2 = b != b;
        }
        if (nan_t || nan_t2) return make_never_pred(dtype);
    }
    PredDev p = make_pred(dtype, CMP_EQ, t);
    p.lo = t;
    p.hi = t2;
    p.inv = 0;
    return p;
}

template <typename T>
int bitmap_t(const void* a, const void* b, size_t n, PredDev pa, PredDev pb, uint64_t* words,
             uint64_t* pass_count, StreamCtx* c) {
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(&c->dev->last);
    if (!hip_ok(launch_zero32((uint32_t*)cnt, 2, c->stream), "zero")) return -1;
    if (n) {
        uint64_t rows_per = 64ull * TyInfo<T>::E * (b ? kBitmapUnroll<true> : kBitmapUnroll<false>);
        unsigned blocks = grid_cap(((n + rows_per - 1) / rows_per + 3) / 4, (unsigned)device_cu_count() * 8);
        bool aligned = (((uintptr_t)a) & 15u) == 0 && (!b || (((uintptr_t)b) & 15u) == 0);
        if (!aligned) { set_error("bitmap filter: columns must be 16-byte aligned"); return -1; }
        if (b)
            hipLaunchKernelGGL((k_bitmap<T, true>), dim3(blocks), dim3(256), 0, c->stream, (const T*)a, (const T*)b, n,
                               pa, pb, words, cnt);
        else
            hipLaunchKernelGGL((k_bitmap<T, false>), dim3(blocks), dim3(256), 0, c->stream, (const T*)a, (const T*)a,
                               n, pa, pb, words, cnt);
        if (!hip_ok(hipGetLastError(), "bitmap launch")) return -1;
    }
    if (!hip_ok(hipMemcpyAsync(&c->host->last, cnt, 8, hipMemcpyDeviceToHost, c->stream), "read count")) return -1;
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return -1;
    if (pass_count) *pass_count = c->host->last;
    return 0;
}

int bitmap_any(int dtype, const void* a, const void* b, size_t n, PredDev pa, PredDev pb, uint64_t* words,
               uint64_t* pc, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    switch (dtype) {
    case DT_F64: return bitmap_t<double>(a, b, n, pa, pb, words, pc, c);
    case DT_F32: return bitmap_t<float>(a, b, n, pa, pb, words, pc, c);
    case DT_I64: return bitmap_t<int64_t>(a, b, n, pa, pb, words, pc, c);
    case DT_I32: return bitmap_t<int32_t>(a, b, n, pa, pb, words, pc, c);
    default: set_error("bitmap filter: dtype %d not supported", dtype); return -1;
    }
}
}  // namespace

extern "C" {

int voxel_gpu_filter_bitmap(int dtype, const void* col, size_t n, int op, uint64_t t, uint64_t t2,
                            uint64_t* words_dev, uint64_t* pass_count, void* stream) {
    if (op < 0 || op > 7) { set_error("bitmap filter: op %d", op); return -1; }
    PredDev pa = pred_for(dtype, op, t, t2);
    return bitmap_any(dtype, col, nullptr, n, pa, pa, words_dev, pass_count, stream);
}

int voxel_gpu_filter_bitmap2(int dtype, const void* a, const void* b, size_t n, int op_a, uint64_t t_a,
                             int op_b, uint64_t t_b, uint64_t* words_dev, uint64_t* pass_count, void* stream) {
    if (op_a < 0 || op_a > 6 || op_b < 0 || op_b > 6) { set_error("bitmap filter: op out of range"); return -1; }
    return bitmap_any(dtype, a, b, n, make_pred(dtype, op_a, t_a), make_pred(dtype, op_b, t_b), words_dev,
                      pass_count, stream);
}

int voxel_gpu_bitmap_combine(uint64_t* active_dev, const uint64_t* other_dev, size_t rows, int combine,
                             uint64_t* active_count, void* stream) {
    set_error("%s", "");
    if (combine < 0 || combine > 2) { set_error("bitmap combine: op %d", combine); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    unsigned long long* cnt = reinterpret_cast<unsigned long long*>(&c->dev->last);
    if (!hip_ok(launch_zero32((uint32_t*)cnt, 2, c->stream), "zero")) return -1;
    uint64_t nw = (rows + 63) / 64;
    if (nw) {
        hipLaunchKernelGGL(k_combine, dim3(grid_cap((nw + 255) / 256, (unsigned)device_cu_count() * 4)), dim3(256), 0,
                           c->stream, active_dev,
                           other_dev, nw, (uint64_t)rows, combine, cnt);
        if (!hip_ok(hipGetLastError(), "combine launch")) return -1;
    }
    if (!hip_ok(hipMemcpyAsync(&c->host->last, cnt, 8, hipMemcpyDeviceToHost, c->stream), "read count")) return -1;
    if (!hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize")) return -1;
    if (active_count) *active_count = c->host->last;
    return 0;
}

int voxel_gpu_select_from_bitmap(const uint64_t* words_dev, size_t rows, uint32_t* idx_out_dev, uint64_t* count,
                                 void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    BitmapSrc src{words_dev, idx_out_dev};
    uint64_t k = 0;
    if (!run_compact(src, rows, c, &k)) return -1;
    if (count) *count = k;
    return 0;
}

int voxel_gpu_compact(int dtype, const void* col, size_t n, int op, uint64_t t, void* values_out_dev,
                      uint32_t* idx_out_dev, uint64_t* count, void* stream) {
    set_error("%s", "");
    if (op < 0 || op > 7) { set_error("compact: op %d", op); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    PredDev pd = pred_for(dtype, op, t, t);
    uint64_t k = 0;
    bool ok = false;
    auto go = [&](auto tag) -> bool {
        using T = decltype(tag);
        PredSrc<T> s{(const T*)col, pd, (T*)values_out_dev, idx_out_dev};
        return run_compact(s, n, c, &k);
    };
    switch (dtype) {
    case DT_F64: ok = go(double{}); break;
    case DT_F32: ok = go(float{}); break;
    case DT_I64: ok = go(int64_t{}); break;
    case DT_I32: ok = go(int32_t{}); break;
    default: set_error("compact: dtype %d not supported", dtype); return -1;
    }
    if (!ok) return -1;
    if (count) *count = k;
    return 0;
}

}  // extern "C"
