// vx_scan.hip -- fused filter + aggregate scans for gfx950 (MI355X).
//
// Replaces the reference's scan loops:
//   Engine<T>::RunFastPath            exec/engine.hpp:109-141
//   JIT fused filter+sum loop         codegen/x86_64.hpp:1632-1716
//   Tier-0 VLOAD/VFILTER/VSUM|VCOUNT  exec/engine.hpp:1947-1969,2870-2924,2980-3111
//   VectorFilter::Apply* + ColumnFilter::AndWith (compound predicates)
//                                     ops/filter.hpp:42-288,540-547
// with one streaming kernel family:
//   * 16-byte (dwordx4) coalesced loads, U vectors per column in flight per
//     lane, grid-stride over resident workgroups (<= 8 per CU);
//   * predicates as a closed range test (two compares -> SGPR lane masks);
//   * floats: per-lane TwoSum-compensated (s, c) pairs kept in T;
//     ints: wrapping 64-bit sums (exact);
//   * wave64 butterfly reduction and one partial per workgroup; by default a
//     second one-workgroup kernel (k_scan_final) merges the partials in
//     workgroup order after the kernel boundary, so the order of the final
//     combine depends only on the grid size.  Layouts 1 and 4 instead merge
//     in the last workgroup to finish (a self-resetting ticket), which costs
//     an agent-scope release/acquire per workgroup on this multi-XCD part.
#include "vx_acc.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>
#include <type_traits>

#include "../../include/voxel_c.h"

#pragma clang fp contract(off)

namespace vx {
namespace {

template <typename T, int MODE, bool TA, bool TB, bool VEC, int U, bool NT, bool CONTIG, bool FIN = true>
__global__ __launch_bounds__(kBlock)
void k_scan(const T* __restrict__ a, const T* __restrict__ b, uint64_t begin, uint64_t end,
            PredDev pda, PredDev pdb, ScanOut* __restrict__ parts, uint32_t* __restrict__ ticket,
            ScanOut* __restrict__ out) {
    using V = typename std::conditional<VEC, typename TyInfo<T>::V, T>::type;
    constexpr int E = VEC ? TyInfo<T>::E : 1;
    Pred<T, TA> pa; pa.load(pda);
    Pred<T, TB> pb; pb.load(pdb);
    Acc<T> acc; acc.init();

    auto one = [&](T x, T y) {
        bool p = pa(x);
        if constexpr (MODE == 1) p = p & pb(x);
        if constexpr (MODE == 2) p = p & pb(y);
        acc.add(p, x);
    };

    // Scalar head (up to the first 16-byte boundary of a) and tail.
    const uint64_t n = end - begin;
    uint64_t head = 0;
    if constexpr (VEC) {
        uintptr_t addr = (uintptr_t)(a + begin);
        head = ((16u - (addr & 15u)) & 15u) / sizeof(T);
        if (head > n) head = n;
    }
    const uint64_t vb = begin + head;
    const uint64_t nvec = (n - head) / E;
    const uint64_t ts = vb + nvec * E;
    const uint64_t ntail = end - ts;
    const uint64_t gtid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gtid < head + ntail) {
        uint64_t i = gtid < head ? begin + gtid : ts + (gtid - head);
        one(a[i], MODE == 2 ? b[i] : T(0));
    }

    const V* __restrict__ av = reinterpret_cast<const V*>(a + vb);
    const V* __restrict__ bv = reinterpret_cast<const V*>(MODE == 2 ? b + vb : a + vb);
    // tiles of kBlock*U vectors: grid-stride over tiles, or (CONTIG) one
    // contiguous run of tiles per workgroup
    const uint64_t tile = (uint64_t)kBlock * U;
    uint64_t first, stop, step;
    if constexpr (CONTIG) {
        const uint64_t ntiles = (nvec + tile - 1) / tile;
        const uint64_t per = (ntiles + gridDim.x - 1) / gridDim.x;
        first = (uint64_t)blockIdx.x * per * tile;
        stop = first + per * tile;
        if (stop > nvec) stop = nvec;
        step = tile;
    } else {
        first = (uint64_t)blockIdx.x * tile;
        stop = nvec;
        step = (uint64_t)gridDim.x * tile;
    }
    for (uint64_t base = first; base < stop; base += step) {
        V xa[U], xb[U];
        if (base + tile <= nvec) {
#pragma unroll
            for (int u = 0; u < U; ++u) xa[u] = ld16<V, NT>(&av[base + u * kBlock + threadIdx.x]);
            if constexpr (MODE == 2) {
#pragma unroll
                for (int u = 0; u < U; ++u) xb[u] = ld16<V, NT>(&bv[base + u * kBlock + threadIdx.x]);
            }
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const T* pxa = reinterpret_cast<const T*>(&xa[u]);
                const T* pxb = reinterpret_cast<const T*>(&xb[u]);
#pragma unroll
                for (int e = 0; e < E; ++e) one(pxa[e], MODE == 2 ? pxb[e] : T(0));
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                uint64_t vi = base + u * kBlock + threadIdx.x;
                if (vi < nvec) {
                    xa[u] = av[vi];
                    if constexpr (MODE == 2) xb[u] = bv[vi];
                    const T* pxa = reinterpret_cast<const T*>(&xa[u]);
                    const T* pxb = reinterpret_cast<const T*>(&xb[u]);
#pragma unroll
                    for (int e = 0; e < E; ++e) one(pxa[e], MODE == 2 ? pxb[e] : T(0));
                }
            }
        }
    }

    __shared__ Acc<T> lds[kBlock / 64];
    __shared__ bool s_last;
    block_reduce(acc, lds);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
    if constexpr (!FIN) return;                 // k_scan_final combines after the kernel boundary
    if (!publish_and_check_last(ticket, &s_last)) return;

    Acc<T> r; r.init();
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += kBlock) {
        Acc<T> o; o.load(&parts[i]);
        r.merge(o);
    }
    __syncthreads();
    block_reduce(r, lds);
    if (threadIdx.x == 0) {
        r.store(out);
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---------------------------------------------------------------------------
// Scans over nullable columns.  A row takes part only when its validity bit
// is 1 (NullBitmap layout, data/nulls.hpp:16-47: bit r & 63 of word r >> 6,
// 1 = valid; NULL = every row valid), for column a and, when given, column
// b; the predicates and aggregates then run as in k_scan.  This is the
// reference's filter bitmap ANDed with the validity words (ColumnFilter over
// the same word layout, ops/filter.hpp:540-547) followed by the aggregate, in
// one pass.  The validity words add 1/64 of the column's bytes per row of an
// 8-byte type; each lane reads the word(s) of its rows, which its wave
// neighbours share from the same cache line.
// ---------------------------------------------------------------------------
__device__ uint64_t g_all_valid = ~0ull;       // read only
// 16-byte loads in flight per lane and column: one column needs twice the
// loads of two to keep as many bytes in flight
template <int MODE> constexpr int kValidUnroll = MODE == 2 ? 4 : 8;

template <typename T, int MODE, bool VEC>
__global__ __launch_bounds__(kBlock)
void k_scan_valid(const T* __restrict__ a, const T* __restrict__ b, const uint64_t* __restrict__ va,
                  const uint64_t* __restrict__ vbm, uint64_t n, PredDev pda, PredDev pdb,
                  ScanOut* __restrict__ parts) {
    using V = typename std::conditional<VEC, typename TyInfo<T>::V, T>::type;
    constexpr int E = VEC ? TyInfo<T>::E : 1;
    constexpr int U = kValidUnroll<MODE>;
    Pred<T, false> pa; pa.load(pda);
    Pred<T, false> pb; pb.load(pdb);
    Acc<T> acc; acc.init();
    // An absent bitmap reads one all-valid word (index masked to 0), so the
    // loop has no branch on it.
    const uint64_t* wa = va ? va : &g_all_valid;
    const uint64_t* wb = vbm ? vbm : &g_all_valid;
    const uint64_t ka = va ? ~0ull : 0ull, kb = vbm ? ~0ull : 0ull;
    auto one = [&](T x, T y, bool valid) {
        bool p = pa(x) & valid;
        if constexpr (MODE == 1) p = p & pb(x);
        if constexpr (MODE == 2) p = p & pb(y);
        acc.add(p, x);
    };
    auto bit = [](uint64_t w, uint64_t r) -> bool { return (w >> (r & 63)) & 1ull; };
    uint64_t head = 0;
    if constexpr (VEC) {
        head = ((16u - ((uintptr_t)a & 15u)) & 15u) / sizeof(T);
        if (head > n) head = n;
    }
    const uint64_t nvec = (n - head) / E;
    const uint64_t ts = head + nvec * E;
    const uint64_t gtid = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gtid < head + (n - ts)) {
        const uint64_t i = gtid < head ? gtid : ts + (gtid - head);
        one(a[i], MODE == 2 ? b[i] : T(0), bit(wa[(i >> 6) & ka], i) & (MODE != 2 || bit(wb[(i >> 6) & kb], i)));
    }
    const V* __restrict__ av = reinterpret_cast<const V*>(a + head);
    const V* __restrict__ bv = reinterpret_cast<const V*>(MODE == 2 ? b + head : a + head);
    const uint64_t tile = (uint64_t)kBlock * U;
    // The E rows of a lane's vector lie in at most two words (two when a
    // misaligned column start puts them across a word boundary); both are
    // loaded with the data, and the wave's lanes share their cache lines.
    auto step = [&](uint64_t vi, const V& xa, const V& xb, uint64_t a0, uint64_t a1, uint64_t b0, uint64_t b1) {
        const T* pxa = reinterpret_cast<const T*>(&xa);
        const T* pxb = reinterpret_cast<const T*>(&xb);
        const uint64_t r0 = head + vi * E;
#pragma unroll
        for (int e = 0; e < E; ++e) {
            const uint64_t r = r0 + e;
            const bool lo = (r >> 6) == (r0 >> 6);
            one(pxa[e], MODE == 2 ? pxb[e] : T(0), bit(lo ? a0 : a1, r) & bit(lo ? b0 : b1, r));
        }
    };
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < nvec; base += (uint64_t)gridDim.x * tile) {
        V xa[U], xb[U];
        uint64_t a0[U], a1[U], b0[U], b1[U];
        if (base + tile <= nvec) {
#pragma unroll
            for (int u = 0; u < U; ++u) {            // every load of the tile is issued before any is used
                const uint64_t vi = base + u * kBlock + threadIdx.x;
                const uint64_t r0 = head + vi * E;
                xa[u] = ld16<V, true>(&av[vi]);
                if constexpr (MODE == 2) xb[u] = ld16<V, true>(&bv[vi]);
                a0[u] = wa[(r0 >> 6) & ka];
                a1[u] = wa[((r0 + E - 1) >> 6) & ka];
                if constexpr (MODE == 2) {              // column b's bitmap: only with a predicate on b
                    b0[u] = wb[(r0 >> 6) & kb];
                    b1[u] = wb[((r0 + E - 1) >> 6) & kb];
                } else {
                    b0[u] = b1[u] = ~0ull;
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u) step(base + u * kBlock + threadIdx.x, xa[u], xb[u], a0[u], a1[u], b0[u], b1[u]);
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint64_t vi = base + u * kBlock + threadIdx.x;
                if (vi < nvec) {
                    const uint64_t r0 = head + vi * E;
                    xa[u] = av[vi];
                    if constexpr (MODE == 2) xb[u] = bv[vi];
                    step(vi, xa[u], xb[u], wa[(r0 >> 6) & ka], wa[((r0 + E - 1) >> 6) & ka],
                         MODE == 2 ? wb[(r0 >> 6) & kb] : ~0ull, MODE == 2 ? wb[((r0 + E - 1) >> 6) & kb] : ~0ull);
                }
            }
        }
    }
    __shared__ Acc<T> lds[kBlock / 64];
    block_reduce(acc, lds);
    if (threadIdx.x == 0) acc.store(&parts[blockIdx.x]);
}

template <typename T, int MODE>
hipError_t launch_valid_t(const ScanArgs& g, const uint64_t* va, const uint64_t* vb, const Scratch& scr,
                          ScanOut* out, hipStream_t st) {
    const T* a = static_cast<const T*>(g.a) + g.begin;
    const T* b = MODE == 2 ? static_cast<const T*>(g.b) + g.begin : nullptr;
    const uint64_t n = g.end - g.begin;
    const bool vec = !(MODE == 2 && (((uintptr_t)a ^ (uintptr_t)b) & 15u));
    const uint64_t per_block = (uint64_t)kBlock * kValidUnroll<MODE> * (16 / sizeof(T));
    uint64_t want = (n + per_block - 1) / per_block;
    uint64_t cap = (uint64_t)device_cu_count() * (MODE == 2 ? 2 : 4);
    if (cap > scr.max_blocks) cap = scr.max_blocks;
    uint64_t blocks = want < cap ? want : cap;
    if (blocks < 1) blocks = 1;
    ScanOut* parts = static_cast<ScanOut*>(scr.partials);
    if (vec)
        hipLaunchKernelGGL((k_scan_valid<T, MODE, true>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a, b, va, vb, n,
                           g.pa, g.pb, parts);
    else
        hipLaunchKernelGGL((k_scan_valid<T, MODE, false>), dim3((unsigned)blocks), dim3(kBlock), 0, st, a, b, va, vb,
                           n, g.pa, g.pb, parts);
    hipLaunchKernelGGL((k_scan_final<T>), dim3(1), dim3(kBlock), 0, st, (const ScanOut*)parts, (uint32_t)blocks, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_valid_type(const ScanArgs& g, const uint64_t* va, const uint64_t* vb, const Scratch& scr,
                             ScanOut* out, hipStream_t st) {
    if (g.mode == 0) return launch_valid_t<T, 0>(g, va, vb, scr, out, st);
    if (g.mode == 1) return launch_valid_t<T, 1>(g, va, vb, scr, out, st);
    return launch_valid_t<T, 2>(g, va, vb, scr, out, st);
}

// ---------------------------------------------------------------------------
// Tier-0 VRED_MIN / VRED_MAX groups (engine.hpp:3071-3095) with the
// CMP acc,part / JLE|JGE +2 / MOVR fold of the voxel_query min()/max()
// programs (bindings/voxel_query.py:111-135).  Each group is 4 lanes of an
// 8-byte type; lane = pred(x) ? x : 0 (VFILTER, engine.hpp:2870-2924), rows
// past Count load as 0 (VLOAD, engine.hpp:1959-1962).
// ---------------------------------------------------------------------------
template <typename T, bool TAB>
__device__ __forceinline__ uint64_t group_value(const T* __restrict__ a, uint64_t count, uint64_t row,
                                                const Pred<T, TAB>& p, bool use_pred, int red) {
    T v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        uint64_t idx = row + i;
        T x = idx < count ? a[idx] : T(0);
        v[i] = (!use_pred || p(x)) ? x : T(0);
    }
    if (red == 0) {                       // VSUM: T sum = 0; sum += lane (engine.hpp:2980-2987)
        if constexpr (TyInfo<T>::F) {
            T sm = T(0);
#pragma unroll
            for (int i = 0; i < 4; ++i) sm = sm + v[i];
            return to_bits(sm);
        } else {
            uint64_t sm = 0;
#pragma unroll
            for (int i = 0; i < 4; ++i) sm += (uint64_t)(int64_t)v[i];
            return sm;
        }
    }
    if (red == 1) {                       // VCOUNT: lanes != 0 and not NaN (engine.hpp:3097-3111)
        uint64_t c = 0;
#pragma unroll
        for (int i = 0; i < 4; ++i) c += (v[i] == v[i] && v[i] != T(0)) ? 1u : 0u;
        return c;
    }
    T m = v[0];
    if (red == 2) {
#pragma unroll
        for (int i = 1; i < 4; ++i) {
            bool take = v[i] < m;
            if constexpr (TyInfo<T>::F) take = take || (v[i] != v[i]);
            m = take ? v[i] : m;
        }
    } else {
#pragma unroll
        for (int i = 1; i < 4; ++i) {
            bool take = v[i] > m;
            if constexpr (TyInfo<T>::F) take = take || (m != m);
            m = take ? v[i] : m;
        }
    }
    return to_bits(m);
}

struct GAcc {
    int64_t smin, smax;
    uint64_t umin, umax;
    __device__ __forceinline__ void init() { smin = INT64_MAX; smax = INT64_MIN; umin = UINT64_MAX; umax = 0; }
    __device__ __forceinline__ void add(uint64_t v) {
        int64_t s = (int64_t)v;
        smin = s < smin ? s : smin; smax = s > smax ? s : smax;
        umin = v < umin ? v : umin; umax = v > umax ? v : umax;
    }
    __device__ __forceinline__ void merge(const GAcc& o) {
        smin = o.smin < smin ? o.smin : smin; smax = o.smax > smax ? o.smax : smax;
        umin = o.umin < umin ? o.umin : umin; umax = o.umax > umax ? o.umax : umax;
    }
    __device__ __forceinline__ void shfl_merge(int m) {
        GAcc o;
        o.smin = __shfl_xor(smin, m, 64); o.smax = __shfl_xor(smax, m, 64);
        o.umin = __shfl_xor(umin, m, 64); o.umax = __shfl_xor(umax, m, 64);
        merge(o);
    }
};

__device__ __forceinline__ void gblock_reduce(GAcc& a, GAcc* lds) {
    for (int m = 32; m > 0; m >>= 1) a.shfl_merge(m);
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    if (lane == 0) lds[wid] = a;
    __syncthreads();
    if (threadIdx.x == 0)
        for (int w = 1; w < kBlock / 64; ++w) a.merge(lds[w]);
}

template <typename T, bool TAB>
__global__ __launch_bounds__(kBlock)
void k_group_minmax(const T* __restrict__ a, uint64_t count, uint64_t off, uint64_t groups,
                    PredDev pd, int use_pred, int red, uint64_t acc0,
                    GroupOut* __restrict__ parts, uint32_t* __restrict__ ticket, GroupOut* __restrict__ out) {
    Pred<T, TAB> p; p.load(pd);
    GAcc g; g.init();
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < groups; i += stride) {
        uint64_t v = group_value<T, TAB>(a, count, off + 4 * i, p, use_pred != 0, red);
        g.add(v);
        if (i == groups - 1) out->last = v;
    }
    __shared__ GAcc lds[kBlock / 64];
    __shared__ bool s_last;
    gblock_reduce(g, lds);
    if (threadIdx.x == 0) {
        GroupOut* q = &parts[blockIdx.x];
        q->smin = g.smin; q->smax = g.smax; q->umin = g.umin; q->umax = g.umax;
    }
    if (!publish_and_check_last(ticket, &s_last)) return;
    GAcc r; r.init();
    for (uint32_t i = threadIdx.x; i < gridDim.x; i += kBlock) {
        GAcc o; o.smin = parts[i].smin; o.smax = parts[i].smax; o.umin = parts[i].umin; o.umax = parts[i].umax;
        r.merge(o);
    }
    __syncthreads();
    gblock_reduce(r, lds);
    if (threadIdx.x == 0) {
        r.add(acc0);
        out->smin = r.smin; out->smax = r.smax; out->umin = r.umin; out->umax = r.umax;
        out->groups = groups;
        __hip_atomic_store(ticket, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// Exact sequential fold in one wave: lanes compute 64 group values at a time,
// then every lane applies them in row order (the loop of the reference).
template <typename T, bool TAB>
__global__ __launch_bounds__(64)
void k_group_fold(const T* __restrict__ a, uint64_t count, uint64_t off, uint64_t groups,
                  PredDev pd, int red, int fold, uint64_t acc0, uint64_t* __restrict__ result) {
    Pred<T, TAB> p; p.load(pd);
    uint64_t acc = acc0, last = 0;
    for (uint64_t base = 0; base < groups; base += 64) {
        uint64_t gi = base + threadIdx.x;
        uint64_t v = gi < groups ? group_value<T, TAB>(a, count, off + 4 * gi, p, true, red) : 0;
        uint64_t lim = groups - base < 64 ? groups - base : 64;
        for (uint64_t k = 0; k < lim; ++k) {
            uint64_t r = __shfl(v, (int)k, 64);
            int64_t diff = (int64_t)(acc - r);          // CMP acc, part
            bool take = (fold == 2) ? (diff > 0) : (diff < 0);
            acc = take ? r : acc;
            last = r;
        }
    }
    if (threadIdx.x == 0) { result[0] = acc; result[1] = last; }
}

template <typename T, bool TAB>
__global__ __launch_bounds__(64)
void k_group_last(const T* __restrict__ a, uint64_t count, uint64_t row, PredDev pd, int red,
                  uint64_t* __restrict__ result) {
    Pred<T, TAB> p; p.load(pd);
    if (threadIdx.x == 0) result[0] = group_value<T, TAB>(a, count, row, p, true, red);
}

template <typename T, int MODE, bool TA, bool TB, int U, bool NT, bool CONTIG>
hipError_t launch_u(const ScanArgs& g, const Scratch& scr, ScanOut* out, hipStream_t st) {
    const T* a = static_cast<const T*>(g.a);
    const T* b = static_cast<const T*>(g.b);
    uint64_t n = g.end - g.begin;
    bool vec = true;
    if (MODE == 2 && (((uintptr_t)(a + g.begin) ^ (uintptr_t)(b + g.begin)) & 15u)) vec = false;
    uint64_t per_block = (uint64_t)kBlock * U * (16 / sizeof(T));
    uint64_t want = (n + per_block - 1) / per_block;
    uint64_t cap = g.blocks > 0 ? (uint64_t)g.blocks : (uint64_t)device_cu_count() * 2;
    if (cap > scr.max_blocks) cap = scr.max_blocks;
    uint64_t blocks = want < cap ? want : cap;
    if (blocks < 1) blocks = 1;
    ScanOut* parts = static_cast<ScanOut*>(scr.partials);
    // the last-workgroup combine needs an agent-scope release/acquire per
    // workgroup (an L2 write-back + invalidate on this multi-XCD part); a
    // second one-workgroup kernel after the kernel boundary avoids it
    const bool split = g.layout == 0 || g.layout == 2 || g.layout == 3;
    if (split) {
        if (vec)
            hipLaunchKernelGGL((k_scan<T, MODE, TA, TB, true, U, NT, CONTIG, false>), dim3((unsigned)blocks), dim3(kBlock),
                               0, st, a, b, g.begin, g.end, g.pa, g.pb, parts, scr.ticket, out);
        else
            hipLaunchKernelGGL((k_scan<T, MODE, TA, TB, false, U, NT, CONTIG, false>), dim3((unsigned)blocks), dim3(kBlock),
                               0, st, a, b, g.begin, g.end, g.pa, g.pb, parts, scr.ticket, out);
        hipLaunchKernelGGL((k_scan_final<T>), dim3(1), dim3(kBlock), 0, st, (const ScanOut*)parts, (uint32_t)blocks, out);
        return hipGetLastError();
    }
    if (vec)
        hipLaunchKernelGGL((k_scan<T, MODE, TA, TB, true, U, NT, CONTIG>), dim3((unsigned)blocks), dim3(kBlock), 0, st,
                           a, b, g.begin, g.end, g.pa, g.pb, parts, scr.ticket, out);
    else
        hipLaunchKernelGGL((k_scan<T, MODE, TA, TB, false, U, NT, CONTIG>), dim3((unsigned)blocks), dim3(kBlock), 0, st,
                           a, b, g.begin, g.end, g.pa, g.pb, parts, scr.ticket, out);
    return hipGetLastError();
}

// Unroll (16-byte loads in flight per lane and column) and non-temporal
// loads are tunable; unroll 0 selects the default for the mode.
template <typename T, int MODE, bool TA, bool TB>
hipError_t launch_t(const ScanArgs& g, const Scratch& scr, ScanOut* out, hipStream_t st) {
    // defaults from the MI355X sweep (2^28-row compound scans): 2 workgroups
    // per CU, non-temporal loads, 4 (8-byte types) / 8 (4-byte types) loads
    int u = g.unroll > 0 ? g.unroll : (sizeof(T) == 8 ? 4 : 8);
    bool nt = g.nontemporal >= 0;             // default: non-temporal
    bool contig = g.layout == 1 || g.layout == 3;
    if (TA || TB) { u = 4; nt = false; contig = false; }    // table predicates: compat paths, one variant
#define VX_SCAN_U(UU)                                                                                   \
    if (u == UU) {                                                                                      \
        if (contig) return nt ? launch_u<T, MODE, TA, TB, UU, true, true>(g, scr, out, st)              \
                              : launch_u<T, MODE, TA, TB, UU, false, true>(g, scr, out, st);            \
        return nt ? launch_u<T, MODE, TA, TB, UU, true, false>(g, scr, out, st)                         \
                  : launch_u<T, MODE, TA, TB, UU, false, false>(g, scr, out, st);                       \
    }
    VX_SCAN_U(2)
    VX_SCAN_U(4)
    VX_SCAN_U(8)
#undef VX_SCAN_U
    return hipErrorInvalidValue;
}

template <typename T>
hipError_t launch_type(const ScanArgs& g, const Scratch& scr, ScanOut* out, hipStream_t st) {
    if (g.mode == 0) return g.table_a ? launch_t<T, 0, true, false>(g, scr, out, st)
                                      : launch_t<T, 0, false, false>(g, scr, out, st);
    if (g.mode == 1) return g.table_b ? launch_t<T, 1, false, true>(g, scr, out, st)
                                      : launch_t<T, 1, false, false>(g, scr, out, st);
    return launch_t<T, 2, false, false>(g, scr, out, st);
}

// ---------------------------------------------------------------------------
// Allocation-free scans (voxel_gpu_scan_noalloc).  The partials live in a
// file-scope static device buffer: kNoAllocMaxParts 64-byte partials, then
// the 64-byte ScanOut the combine writes.  Every partial k_scan_final reads
// was written by the same launch's k_scan and the result slot is written
// before it is read, so nothing has to start at zero.  Each GPU has its own
// copy at its own address, looked up once per device with
// hipGetSymbolAddress; host code uses only that address.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[(kNoAllocMaxParts + 1) * 64];
static_assert(sizeof(voxel_gpu_agg) == 80, "voxel_gpu_agg layout");

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

hipError_t scratch_of_current_device(void** out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return hipSuccess;
}

// ScanOut -> voxel_gpu_agg, the same conversion as the host's fill_agg
// (finish of the compensated pair with a zero initial value, NaN-free
// min <= max test for floats).  empty: no rows, every field zero.
template <typename T>
__global__ __launch_bounds__(64)
void k_agg_out(const ScanOut* __restrict__ r, int empty, voxel_gpu_agg* __restrict__ out) {
    if (threadIdx.x != 0) return;
    voxel_gpu_agg g;
    g.count = 0; g.sum = 0.0; g.sum_hi = 0.0; g.sum_lo = 0.0; g.isum = 0;
    g.fmin = 0.0; g.fmax = 0.0; g.imin = 0; g.imax = 0; g.has_minmax = 0; g.pad = 0;
    if (!empty) {
        const ScanOut o = *r;
        g.count = o.cnt;
        if constexpr (std::is_same<T, double>::value) {
            const double s = __longlong_as_double((long long)o.s), c = __longlong_as_double((long long)o.c);
            g.sum_hi = s; g.sum_lo = c;
            if (!__builtin_isfinite(s)) {
                g.sum = 0.0 + s;
            } else {
                const double t = 0.0 + s, bp = t - 0.0, e = (0.0 - (t - bp)) + (s - bp);
                g.sum = t + (e + c);
            }
            g.fmin = __longlong_as_double((long long)o.mn); g.fmax = __longlong_as_double((long long)o.mx);
            g.has_minmax = o.cnt > 0 && g.fmin <= g.fmax;
        } else if constexpr (std::is_same<T, float>::value) {
            const float s = __uint_as_float((uint32_t)o.s), c = __uint_as_float((uint32_t)o.c);
            g.sum_hi = s; g.sum_lo = c;
            if (!__builtin_isfinite(s)) {
                g.sum = 0.0f + s;
            } else {
                const float t = 0.0f + s, bp = t - 0.0f, e = (0.0f - (t - bp)) + (s - bp);
                g.sum = t + (e + c);
            }
            g.fmin = __uint_as_float((uint32_t)o.mn); g.fmax = __uint_as_float((uint32_t)o.mx);
            g.has_minmax = o.cnt > 0 && g.fmin <= g.fmax;
        } else if constexpr (std::is_same<T, int64_t>::value) {
            g.isum = (int64_t)o.s; g.imin = (int64_t)o.mn; g.imax = (int64_t)o.mx;
            g.has_minmax = o.cnt > 0;
        } else {
            g.isum = (int64_t)o.s; g.imin = (int32_t)(uint32_t)o.mn; g.imax = (int32_t)(uint32_t)o.mx;
            g.has_minmax = o.cnt > 0;
        }
    }
    *out = g;
}

}  // namespace

hipError_t launch_scan_noalloc(const ScanArgs& g, void* out_dev, hipStream_t st) {
    if (g.end < g.begin || !out_dev) return hipErrorInvalidValue;
    if (!(g.layout == 0 || g.layout == 2 || g.layout == 3)) return hipErrorInvalidValue;
    if (g.blocks < 0 || (size_t)g.blocks > kNoAllocMaxParts) return hipErrorInvalidValue;
    void* base = nullptr;
    hipError_t e = scratch_of_current_device(&base);
    if (e != hipSuccess) return e;
    Scratch scr{};
    scr.partials = base;
    scr.ticket = nullptr;                                   // split layouts never touch it
    scr.max_blocks = kNoAllocMaxParts;
    ScanOut* res = reinterpret_cast<ScanOut*>(static_cast<char*>(base) + kNoAllocMaxParts * 64);
    const int empty = g.end == g.begin;
    if (!empty) {
        e = launch_scan(g, scr, res, st);
        if (e != hipSuccess) return e;
    }
    voxel_gpu_agg* out = static_cast<voxel_gpu_agg*>(out_dev);
    switch (g.dtype) {
    case DT_F64: hipLaunchKernelGGL((k_agg_out<double>), dim3(1), dim3(64), 0, st, res, empty, out); break;
    case DT_F32: hipLaunchKernelGGL((k_agg_out<float>), dim3(1), dim3(64), 0, st, res, empty, out); break;
    case DT_I64: hipLaunchKernelGGL((k_agg_out<int64_t>), dim3(1), dim3(64), 0, st, res, empty, out); break;
    case DT_I32: hipLaunchKernelGGL((k_agg_out<int32_t>), dim3(1), dim3(64), 0, st, res, empty, out); break;
    default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

int device_cu_count() {
    static int cus = 0;
    if (cus == 0) {
        int dev = 0;
        if (hipGetDevice(&dev) != hipSuccess) return 256;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return 256;
        cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    }
    return cus;
}

hipError_t launch_scan_valid(const ScanArgs& g, const uint64_t* va, const uint64_t* vb, const Scratch& scr,
                             ScanOut* out, hipStream_t st) {
    if (g.end < g.begin || g.table_a || g.table_b) return hipErrorInvalidValue;
    if (g.begin & 63) return hipErrorInvalidValue;          // validity words start at row 0 of the range
    const uint64_t* wa = va ? va + (g.begin >> 6) : nullptr;
    const uint64_t* wb = vb ? vb + (g.begin >> 6) : nullptr;
    switch (g.dtype) {
    case DT_F64: return launch_valid_type<double>(g, wa, wb, scr, out, st);
    case DT_F32: return launch_valid_type<float>(g, wa, wb, scr, out, st);
    case DT_I64: return launch_valid_type<int64_t>(g, wa, wb, scr, out, st);
    case DT_I32: return launch_valid_type<int32_t>(g, wa, wb, scr, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_scan(const ScanArgs& g, const Scratch& scr, ScanOut* out, hipStream_t st) {
    if (g.end < g.begin) return hipErrorInvalidValue;
    if (g.mode == 2 && g.table_b) return hipErrorInvalidValue;
    switch (g.dtype) {
    case DT_F64: return launch_type<double>(g, scr, out, st);
    case DT_F32: return launch_type<float>(g, scr, out, st);
    case DT_I64: return launch_type<int64_t>(g, scr, out, st);
    case DT_I32: return launch_type<int32_t>(g, scr, out, st);
    default: return hipErrorInvalidValue;
    }
}

hipError_t launch_group_minmax(const GroupArgs& g, const Scratch& scr, GroupOut* out, hipStream_t st) {
    if (g.groups == 0) return hipErrorInvalidValue;
    uint64_t want = (g.groups + kBlock - 1) / kBlock;
    uint64_t cap = g.blocks > 0 ? (uint64_t)g.blocks : (uint64_t)device_cu_count() * 8;
    if (cap > scr.max_blocks) cap = scr.max_blocks;
    unsigned blocks = (unsigned)(want < cap ? want : cap);
    GroupOut* parts = static_cast<GroupOut*>(scr.partials);
    int use_pred = 1;
    if (g.dtype == DT_F64) {
        if (g.use_table)
            hipLaunchKernelGGL((k_group_minmax<double, true>), dim3(blocks), dim3(kBlock), 0, st,
                               (const double*)g.a, g.count, g.off, g.groups, g.p, use_pred, g.red, g.acc0, parts, scr.ticket, out);
        else
            hipLaunchKernelGGL((k_group_minmax<double, false>), dim3(blocks), dim3(kBlock), 0, st,
                               (const double*)g.a, g.count, g.off, g.groups, g.p, use_pred, g.red, g.acc0, parts, scr.ticket, out);
    } else if (g.dtype == DT_I64) {
        hipLaunchKernelGGL((k_group_minmax<int64_t, false>), dim3(blocks), dim3(kBlock), 0, st,
                           (const int64_t*)g.a, g.count, g.off, g.groups, g.p, use_pred, g.red, g.acc0, parts, scr.ticket, out);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_group_fold(const GroupArgs& g, uint64_t* result, hipStream_t st) {
    if (g.groups == 0) return hipErrorInvalidValue;
    if (g.dtype == DT_F64) {
        if (g.use_table)
            hipLaunchKernelGGL((k_group_fold<double, true>), dim3(1), dim3(64), 0, st,
                               (const double*)g.a, g.count, g.off, g.groups, g.p, g.red, g.fold, g.acc0, result);
        else
            hipLaunchKernelGGL((k_group_fold<double, false>), dim3(1), dim3(64), 0, st,
                               (const double*)g.a, g.count, g.off, g.groups, g.p, g.red, g.fold, g.acc0, result);
    } else if (g.dtype == DT_I64) {
        hipLaunchKernelGGL((k_group_fold<int64_t, false>), dim3(1), dim3(64), 0, st,
                           (const int64_t*)g.a, g.count, g.off, g.groups, g.p, g.red, g.fold, g.acc0, result);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_group_last(const GroupArgs& g, uint64_t* result, hipStream_t st) {
    if (g.dtype == DT_F64) {
        if (g.use_table)
            hipLaunchKernelGGL((k_group_last<double, true>), dim3(1), dim3(64), 0, st,
                               (const double*)g.a, g.count, g.off, g.p, g.red, result);
        else
            hipLaunchKernelGGL((k_group_last<double, false>), dim3(1), dim3(64), 0, st,
                               (const double*)g.a, g.count, g.off, g.p, g.red, result);
    } else if (g.dtype == DT_I64) {
        hipLaunchKernelGGL((k_group_last<int64_t, false>), dim3(1), dim3(64), 0, st,
                           (const int64_t*)g.a, g.count, g.off, g.p, g.red, result);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

}  // namespace vx
