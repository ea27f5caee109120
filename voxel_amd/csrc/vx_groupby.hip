// vx_groupby.hip -- GROUP BY key: COUNT, SUM, MIN, MAX (and AVG = SUM/COUNT)
// of an f64 value column on gfx950, replacing
// HashAggregator<TKey,f64>::Accumulate/Finalize (ops/aggregate.hpp:168-261)
// and the voxel_agg_* C ABI (bindings/voxel_c.cpp:167-182).
//
// Dense path (key range of at most kMaxDense consecutive values, e.g. the
// 265 taxi zones): every wave owns a private copy of the bins in LDS.  A
// wave takes 64 rows at a time; lanes holding the same bin are ranked by
// lane (wave ballots over the bin bits) and update the bin one rank after
// the other, so no two lanes of a wave write one bin at once and every bin
// sees its rows in row order.  Bins hold a TwoSum-compensated (s, c) pair.
// The 4 wave copies merge in wave order; workgroups write partials that one
// workgroup per bin merges in workgroup order.  The result does not depend
// on scheduling: it is bit-identical from run to run.
//
// General path (any key domain): keys become order-preserving u64 and are
// sorted with their row ids by the radix sort (vx_radix.hip); group starts
// are compacted (vx_compact.hpp); one wave per group reduces its rows
// (values gathered by row id) with compensated sums.  Groups come out in
// ascending key order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/voxel_c.h"
#include "vx_compact.hpp"
#include "vx_device.hpp"
#include "vx_groupby.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

constexpr int kGThreads = 256;
constexpr int kGWaves = kGThreads / 64;
constexpr int kGUnroll = 4;

__device__ __forceinline__ void bin_init(GBin& b) {
    b.s = 0.0; b.c = 0.0; b.cnt = 0;
    b.mn = __builtin_huge_val(); b.mx = -__builtin_huge_val();
}

__device__ __forceinline__ void bin_add(GBin& b, double v) {
    two_sum<double>(b.s, b.c, v);
    b.cnt += 1;
    b.mn = fmin(b.mn, v);
    b.mx = fmax(b.mx, v);
}

__device__ __forceinline__ void bin_merge(GBin& a, const GBin& o) {
    double cc = a.c + o.c;
    two_sum<double>(a.s, cc, o.s);
    a.c = cc;
    a.cnt += o.cnt;
    a.mn = fmin(a.mn, o.mn);
    a.mx = fmax(a.mx, o.mx);
}

__device__ __forceinline__ uint64_t g_lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

template <typename K>
__global__ __launch_bounds__(kGThreads) void k_gb_dense(const K* __restrict__ keys, const double* __restrict__ vals,
                                                        uint64_t n, int64_t kmin, uint32_t nbins, int nbits,
                                                        GBin* __restrict__ parts) {
    extern __shared__ GBin bins[];               // kGWaves * nbins
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (uint32_t i = threadIdx.x; i < kGWaves * nbins; i += kGThreads) bin_init(bins[i]);
    __syncthreads();
    GBin* mine = bins + (size_t)w * nbins;
    const uint64_t waves = (uint64_t)gridDim.x * kGWaves;
    const uint64_t wid = (uint64_t)blockIdx.x * kGWaves + w;
    for (uint64_t base = wid * 64 * kGUnroll; base < n; base += waves * 64 * kGUnroll) {
        uint32_t b[kGUnroll];
        double v[kGUnroll];
        bool ok[kGUnroll];
#pragma unroll
        for (int u = 0; u < kGUnroll; ++u) {
            uint64_t row = base + (uint64_t)u * 64 + lane;
            ok[u] = row < n;
            b[u] = ok[u] ? (uint32_t)((int64_t)keys[row] - kmin) : 0u;
            v[u] = ok[u] ? vals[row] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < kGUnroll; ++u) {
            uint64_t peers = __ballot(ok[u]);
            for (int bit = 0; bit < nbits; ++bit) {
                uint64_t m = __ballot((b[u] >> bit) & 1u);
                peers &= ((b[u] >> bit) & 1u) ? m : ~m;
            }
            int rank = ok[u] ? __popcll(peers & g_lanemask_lt()) : 64;
            for (int r = 0;; ++r) {
                if (rank == r) bin_add(mine[b[u]], v[u]);
                if (!__any(rank > r && rank < 64)) break;
            }
        }
    }
    __syncthreads();
    GBin* out = parts + (size_t)blockIdx.x * nbins;
    for (uint32_t i = threadIdx.x; i < nbins; i += kGThreads) {
        GBin acc = bins[i];
        for (int ww = 1; ww < kGWaves; ++ww) bin_merge(acc, bins[(size_t)ww * nbins + i]);
        out[i] = acc;
    }
}

// one workgroup per bin: partials of all workgroups merged in a fixed order
__global__ __launch_bounds__(256) void k_gb_final(const GBin* __restrict__ parts, uint32_t nblocks, uint32_t nbins,
                                                  GBin* __restrict__ out) {
    __shared__ GBin red[256];
    const uint32_t bin = blockIdx.x;
    GBin acc; bin_init(acc);
    for (uint32_t j = threadIdx.x; j < nblocks; j += 256) bin_merge(acc, parts[(size_t)j * nbins + bin]);
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) bin_merge(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[bin] = red[0];
}

template <typename K>
__global__ __launch_bounds__(256) void k_key_minmax(const K* __restrict__ keys, uint64_t n,
                                                    long long* __restrict__ mm) {
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        int64_t k = (int64_t)keys[i];
        lo = k < lo ? k : lo;
        hi = k > hi ? k : hi;
    }
    for (int m = 32; m > 0; m >>= 1) {
        int64_t a = __shfl_xor(lo, m, 64), b = __shfl_xor(hi, m, 64);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & 63) == 0) { atomicMin(&mm[0], (long long)lo); atomicMax(&mm[1], (long long)hi); }
}

__global__ void k_mm_init(long long* mm) { mm[0] = INT64_MAX; mm[1] = INT64_MIN; }

// non-empty dense bins -> compacted group records (ascending key)
struct DenseEmit {
    const GBin* bins;
    int64_t kmin;
    int64_t* okey;
    GBin* orec;
    __device__ __forceinline__ bool pass(uint64_t i) const { return bins[i].cnt > 0; }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t k) const {
        okey[k] = kmin + (int64_t)i;
        orec[k] = bins[i];
    }
};

// keys to order-preserving u64 (signed keys get the sign bit flipped), row ids
template <typename K>
__global__ __launch_bounds__(256) void k_gb_keys(const K* __restrict__ keys, uint64_t n, int is_signed,
                                                 uint64_t* __restrict__ ok, uint32_t* __restrict__ rows) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t u;
        if constexpr (sizeof(K) == 4) u = is_signed ? (uint64_t)((uint32_t)keys[i] ^ 0x80000000u) : (uint64_t)(uint32_t)keys[i];
        else u = is_signed ? ((uint64_t)keys[i] ^ 0x8000000000000000ull) : (uint64_t)keys[i];
        ok[i] = u;
        rows[i] = (uint32_t)i;
    }
}

struct StartSrc {
    const uint64_t* k;
    uint64_t* starts;
    __device__ __forceinline__ bool pass(uint64_t i) const { return i == 0 || k[i] != k[i - 1]; }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t j) const { starts[j] = i; }
};

// one wave per group over [starts[g], starts[g+1]); rec_in: pre-aggregated
// records (merge) or null (raw values gathered through rows)
__global__ __launch_bounds__(256) void k_gb_segments(const uint64_t* __restrict__ sk, const uint32_t* __restrict__ rows,
                                                     const uint64_t* __restrict__ starts, uint64_t ngroups, uint64_t n,
                                                     const double* __restrict__ vals, const GBin* __restrict__ rec_in,
                                                     int is_signed, int key_bytes, int64_t* __restrict__ okey,
                                                     GBin* __restrict__ orec) {
    const int lane = threadIdx.x & 63;
    const uint64_t waves = (uint64_t)gridDim.x * (blockDim.x / 64);
    for (uint64_t g = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) / 64; g < ngroups; g += waves) {
        uint64_t s = starts[g], e = g + 1 < ngroups ? starts[g + 1] : n;
        GBin acc; bin_init(acc);
        for (uint64_t j = s + lane; j < e; j += 64) {
            if (rec_in) bin_merge(acc, rec_in[rows[j]]);
            else bin_add(acc, vals[rows[j]]);
        }
        for (int m = 32; m > 0; m >>= 1) {
            GBin o;
            o.s = __shfl_xor(acc.s, m, 64); o.c = __shfl_xor(acc.c, m, 64);
            o.cnt = __shfl_xor(acc.cnt, m, 64);
            o.mn = __shfl_xor(acc.mn, m, 64); o.mx = __shfl_xor(acc.mx, m, 64);
            bin_merge(acc, o);
        }
        if (lane == 0) {
            uint64_t u = sk[s];
            int64_t key;
            if (key_bytes == 4) key = is_signed ? (int64_t)(int32_t)((uint32_t)u ^ 0x80000000u) : (int64_t)(uint32_t)u;
            else key = is_signed ? (int64_t)(u ^ 0x8000000000000000ull) : (int64_t)u;
            okey[g] = key;
            orec[g] = acc;
        }
    }
}

// raw rows -> records (for merging a batch into an existing aggregate)
__global__ __launch_bounds__(256) void k_gb_concat(const int64_t* __restrict__ ka, const GBin* __restrict__ ra,
                                                   uint64_t na, const int64_t* __restrict__ kb,
                                                   const GBin* __restrict__ rb, uint64_t nb,
                                                   uint64_t* __restrict__ ok, uint32_t* __restrict__ rows,
                                                   GBin* __restrict__ rec) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < na + nb; i += stride) {
        int64_t k = i < na ? ka[i] : kb[i - na];
        ok[i] = (uint64_t)k ^ 0x8000000000000000ull;
        rows[i] = (uint32_t)i;
        rec[i] = i < na ? ra[i] : rb[i - na];
    }
}

__global__ __launch_bounds__(256) void k_gb_finish(const int64_t* __restrict__ keys, const GBin* __restrict__ rec,
                                                   uint64_t n, int64_t* __restrict__ ok, uint64_t* __restrict__ oc,
                                                   double* __restrict__ os, double* __restrict__ oa,
                                                   double* __restrict__ omn, double* __restrict__ omx) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        GBin b = rec[i];
        double sum = b.s + b.c;
        if (ok) ok[i] = keys[i];
        if (oc) oc[i] = b.cnt;
        if (os) os[i] = sum;
        if (oa) oa[i] = b.cnt ? sum / (double)b.cnt : 0.0;     // AggregateState::Avg (aggregate.hpp:136-139)
        if (omn) omn[i] = b.mn;
        if (omx) omx[i] = b.mx;
    }
}

unsigned gridn(uint64_t n, unsigned cap) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b ? (b < cap ? b : cap) : 1);
}

}  // namespace

struct GroupbyWork {
    GBin* parts = nullptr;
    uint64_t parts_cap = 0;      // GBin records
    GBin* bins = nullptr;        // kMaxDense
    long long* mm = nullptr;
    long long* mm_host = nullptr;
    // general path
    uint64_t* k1 = nullptr; uint64_t* k2 = nullptr; uint32_t* r1 = nullptr; uint32_t* r2 = nullptr;
    uint64_t* starts = nullptr; GBin* rec = nullptr;
    uint64_t cap = 0;
    // results (ascending key), two sets so that a merge can read one
    int64_t* okey[2] = {nullptr, nullptr};
    GBin* orec[2] = {nullptr, nullptr};
    uint64_t ocap[2] = {0, 0};
    int cur = 0;
};

static GroupbyWork& gwork(StreamCtx* c) {
    static std::vector<std::pair<StreamCtx*, GroupbyWork*>> all;
    for (auto& p : all) if (p.first == c) return *p.second;
    all.emplace_back(c, new GroupbyWork());
    return *all.back().second;
}

static bool ensure_out(GroupbyWork& w, int which, uint64_t n) {
    if (n < (uint64_t)kMaxDense) n = kMaxDense;
    if (n <= w.ocap[which]) return true;
    if (w.okey[which]) (void)hipFree(w.okey[which]);
    if (w.orec[which]) (void)hipFree(w.orec[which]);
    w.okey[which] = nullptr; w.orec[which] = nullptr; w.ocap[which] = 0;
    if (!hip_ok(hipMalloc((void**)&w.okey[which], n * 8), "hipMalloc(group keys)")) return false;
    if (!hip_ok(hipMalloc((void**)&w.orec[which], n * sizeof(GBin)), "hipMalloc(group records)")) return false;
    w.ocap[which] = n;
    return true;
}

static bool ensure_general(GroupbyWork& w, uint64_t n) {
    if (n <= w.cap) return true;
    for (void* p : {(void*)w.k1, (void*)w.k2, (void*)w.r1, (void*)w.r2, (void*)w.starts, (void*)w.rec})
        if (p) (void)hipFree(p);
    w.k1 = w.k2 = nullptr; w.r1 = w.r2 = nullptr; w.starts = nullptr; w.rec = nullptr; w.cap = 0;
    bool ok = hip_ok(hipMalloc((void**)&w.k1, n * 8), "hipMalloc") && hip_ok(hipMalloc((void**)&w.k2, n * 8), "hipMalloc") &&
              hip_ok(hipMalloc((void**)&w.r1, n * 4), "hipMalloc") && hip_ok(hipMalloc((void**)&w.r2, n * 4), "hipMalloc") &&
              hip_ok(hipMalloc((void**)&w.starts, n * 8), "hipMalloc") &&
              hip_ok(hipMalloc((void**)&w.rec, n * sizeof(GBin)), "hipMalloc");
    if (ok) w.cap = n;
    return ok;
}

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

// Sort-based grouping of n items with u64 order keys in w.k1 / row ids in
// w.r1; values either raw (vals) or records (rec_in).  Returns groups.
static int64_t group_sorted(GroupbyWork& w, uint64_t n, const double* vals, const GBin* rec_in, int is_signed,
                            int key_bytes, int64_t* okey, GBin* orec, StreamCtx* c) {
    if (sort_pairs_u64(w.k1, w.r1, w.k2, w.r2, n, c) < 0) return -1;
    StartSrc src{w.k2, w.starts};
    uint64_t ng = 0;
    if (!run_compact(src, n, c, &ng)) return -1;
    hipLaunchKernelGGL(k_gb_segments, dim3(gridn(ng * 64, 8192)), dim3(256), 0, c->stream, w.k2, w.r2, w.starts, ng, n,
                       vals, rec_in, is_signed, key_bytes, okey, orec);
    if (!hip_ok(hipGetLastError(), "segment launch")) return -1;
    return (int64_t)ng;
}

// Group-by over device columns; group records in ascending key order.
int64_t groupby_f64(int key_dtype, const void* keys, const double* vals, uint64_t n, int64_t** okey_out,
                    GBin** orec_out, StreamCtx* c, int force_general) {
    GroupbyWork& w = gwork(c);
    hipStream_t st = c->stream;
    const int which = w.cur ^= 1;
    if (!ensure_out(w, which, 0)) return -1;
    int64_t* okey = w.okey[which];
    GBin* orec = w.orec[which];
    *okey_out = okey;
    *orec_out = orec;
    if (n == 0) return 0;
    const bool k32 = key_dtype == DT_I32 || key_dtype == DT_U32;
    const int is_signed = (key_dtype == DT_I32 || key_dtype == DT_I64) ? 1 : 0;
    if (!w.mm) {
        if (!hip_ok(hipMalloc((void**)&w.mm, 64), "hipMalloc")) return -1;
        if (!hip_ok(hipHostMalloc((void**)&w.mm_host, 64, hipHostMallocDefault), "hipHostMalloc")) return -1;
        if (!hip_ok(hipMalloc((void**)&w.bins, kMaxDense * sizeof(GBin)), "hipMalloc")) return -1;
    }
    hipLaunchKernelGGL(k_mm_init, dim3(1), dim3(1), 0, st, w.mm);
    unsigned mblocks = (unsigned)device_cu_count() * 4;
    if (key_dtype == DT_I32) hipLaunchKernelGGL(k_key_minmax<int32_t>, dim3(mblocks), dim3(256), 0, st, (const int32_t*)keys, n, w.mm);
    else if (key_dtype == DT_U32) hipLaunchKernelGGL(k_key_minmax<uint32_t>, dim3(mblocks), dim3(256), 0, st, (const uint32_t*)keys, n, w.mm);
    else if (key_dtype == DT_I64) hipLaunchKernelGGL(k_key_minmax<int64_t>, dim3(mblocks), dim3(256), 0, st, (const int64_t*)keys, n, w.mm);
    else { set_error("group-by: key dtype %d not supported (i32, u32, i64)", key_dtype); return -1; }
    if (!hip_ok(hipMemcpyAsync(w.mm_host, w.mm, 16, hipMemcpyDeviceToHost, st), "read key range") ||
        !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return -1;
    const int64_t kmin = w.mm_host[0], kmax = w.mm_host[1];
    const uint64_t span = (uint64_t)(kmax - kmin) + 1;
    if (!force_general && k32 && span <= (uint64_t)kMaxDense) {
        uint32_t nbins = (uint32_t)span;
        int nbits = 0;
        while ((1u << nbits) < nbins) ++nbits;
        unsigned blocks = (unsigned)device_cu_count() * 4;
        uint64_t need = (n + 64 * kGUnroll * kGWaves - 1) / (64 * kGUnroll * kGWaves);
        if (need < blocks) blocks = (unsigned)(need ? need : 1);
        uint64_t recs = (uint64_t)blocks * nbins;
        if (recs > w.parts_cap) {
            if (w.parts) (void)hipFree(w.parts);
            w.parts = nullptr; w.parts_cap = 0;
            if (!hip_ok(hipMalloc((void**)&w.parts, recs * sizeof(GBin)), "hipMalloc(partials)")) return -1;
            w.parts_cap = recs;
        }
        size_t lds = (size_t)kGWaves * nbins * sizeof(GBin);
        if (key_dtype == DT_I32)
            hipLaunchKernelGGL(k_gb_dense<int32_t>, dim3(blocks), dim3(kGThreads), lds, st, (const int32_t*)keys, vals, n,
                               kmin, nbins, nbits, w.parts);
        else
            hipLaunchKernelGGL(k_gb_dense<uint32_t>, dim3(blocks), dim3(kGThreads), lds, st, (const uint32_t*)keys, vals,
                               n, kmin, nbins, nbits, w.parts);
        if (!hip_ok(hipGetLastError(), "group-by launch")) return -1;
        hipLaunchKernelGGL(k_gb_final, dim3(nbins), dim3(256), 0, st, w.parts, blocks, nbins, w.bins);
        if (!hip_ok(hipGetLastError(), "group-by final launch")) return -1;
        DenseEmit src{w.bins, kmin, okey, orec};
        uint64_t ng = 0;
        if (!run_compact(src, nbins, c, &ng)) return -1;
        return (int64_t)ng;
    }
    // general path: sort (key, row) and reduce segments
    if (n >= (1ull << 30)) { set_error("group-by: general path limited to 2^30 rows"); return -1; }
    if (!ensure_general(w, n) || !ensure_out(w, which, n)) return -1;
    okey = *okey_out = w.okey[which];
    orec = *orec_out = w.orec[which];
    if (key_dtype == DT_I32) hipLaunchKernelGGL(k_gb_keys<int32_t>, dim3(gridn(n, 16384)), dim3(256), 0, st, (const int32_t*)keys, n, 1, w.k1, w.r1);
    else if (key_dtype == DT_U32) hipLaunchKernelGGL(k_gb_keys<uint32_t>, dim3(gridn(n, 16384)), dim3(256), 0, st, (const uint32_t*)keys, n, 0, w.k1, w.r1);
    else hipLaunchKernelGGL(k_gb_keys<int64_t>, dim3(gridn(n, 16384)), dim3(256), 0, st, (const int64_t*)keys, n, 1, w.k1, w.r1);
    if (!hip_ok(hipGetLastError(), "key launch")) return -1;
    return group_sorted(w, n, vals, nullptr, is_signed, k32 ? 4 : 8, okey, orec, c);
}

// Merges two ascending group-record sets (keys as i64) into the other result
// set (the inputs may be one of the internal sets).
int64_t groupby_merge(const int64_t* ka, const GBin* ra, uint64_t na, const int64_t* kb, const GBin* rb, uint64_t nb,
                      int64_t** okey_out, GBin** orec_out, StreamCtx* c) {
    GroupbyWork& w = gwork(c);
    uint64_t n = na + nb;
    int which = -1;
    for (int i = 0; i < 2; ++i)
        if (w.okey[i] != ka && w.okey[i] != kb) which = i;
    if (which < 0) { set_error("group-by merge: no free result set"); return -1; }
    w.cur = which;
    if (!ensure_out(w, which, n) || !ensure_general(w, n > 0 ? n : 1)) return -1;
    *okey_out = w.okey[which];
    *orec_out = w.orec[which];
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gb_concat, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, ka, ra, na, kb, rb, nb, w.k1, w.r1,
                       w.rec);
    if (!hip_ok(hipGetLastError(), "concat launch")) return -1;
    return group_sorted(w, n, nullptr, w.rec, 1, 8, w.okey[which], w.orec[which], c);
}

int groupby_finish(const int64_t* keys, const GBin* rec, uint64_t n, int64_t* ok, uint64_t* oc, double* os,
                   double* oa, double* omn, double* omx, void* ctx) {
    StreamCtx* c = static_cast<StreamCtx*>(ctx);
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_gb_finish, dim3(gridn(n, 4096)), dim3(256), 0, c->stream, keys, rec, n, ok, oc, os, oa, omn, omx);
    return hip_ok(hipGetLastError(), "finish launch") ? 0 : -1;
}

}  // namespace vx
