// vx_groupby.hip -- GROUP BY key: COUNT, SUM, MIN, MAX (and AVG = SUM/COUNT)
// of an f64 value column on gfx950, replacing
// HashAggregator<TKey,f64>::Accumulate/Finalize (ops/aggregate.hpp:168-261)
// and the voxel_agg_* C ABI (bindings/voxel_c.cpp:167-182).
//
// Dense path (key range of at most kMaxDense consecutive values, e.g. the
// 265 taxi zones): every wave owns a private copy of the bins in LDS, laid
// out as separate arrays (sum, compensation, min, max, u32 count).  Rows are
// added with LDS atomics: ds_add_rtn_f64 returns the bin's sum before the
// row's add, so the rounding error of that add is recovered exactly (TwoSum
// of the returned value and the row) and added to the compensation word;
// min, max and count are ds_min_f64 / ds_max_f64 / ds_add_u32.  Only one
// wave writes a copy and the LDS unit applies a wave instruction's lanes in
// a fixed order, so the result does not depend on scheduling and is
// bit-identical from run to run.  The wave copies merge in wave order;
// workgroups write partials that one workgroup per bin merges in workgroup
// order.
//
// General path (any key domain): keys become order-preserving u64 and are
// sorted with their row ids by the radix sort (vx_radix.hip); group starts
// are compacted (vx_compact.hpp); one wave per group reduces its rows
// (values gathered by row id) with compensated sums.  Groups come out in
// ascending key order.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <limits>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_acc.hpp"
#include "vx_compact.hpp"
#include "vx_device.hpp"
#include "vx_groupby.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

constexpr int kGThreads = 512;
constexpr int kGWaves = kGThreads / 64;
constexpr int kGChunk = 256;          // rows per wave step: 4 per lane
constexpr int kGUnroll = 2;           // wave steps in flight
constexpr size_t kGBinBytes = 4 * sizeof(double) + sizeof(uint32_t);

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

// Per-wave bin arrays in LDS: sum, compensation, (min, max) pairs and u32
// counts, each indexed by wave copy * bins + bin.  A row costs one
// ds_add_rtn_f64, one ds_read_b128 of its (min, max) pair and one ds_add_u32;
// the compensation add is skipped when the rounding error is zero, and the
// min / max atomics are issued only when the row beats the pair just read
// (only this wave writes the copy, so the pair read is current).
struct WaveBins {
    double* s;
    double* c;
    double2* mm;
    uint32_t* cnt;
    __device__ __forceinline__ void add(uint32_t i, double v) const {
        double old = __hip_atomic_fetch_add(&s[i], v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        double t = old + v;
        double bp = t - old;
        double e = (old - (t - bp)) + (v - bp);
        if (e != 0.0) __hip_atomic_fetch_add(&c[i], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        const double2 p = mm[i];
        if (v < p.x) __hip_atomic_fetch_min(&mm[i].x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (v > p.y) __hip_atomic_fetch_max(&mm[i].y, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        __hip_atomic_fetch_add(&cnt[i], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    __device__ __forceinline__ GBin get(uint32_t i) const {
        GBin b;
        const double2 p = mm[i];
        b.s = s[i]; b.c = c[i]; b.mn = p.x; b.mx = p.y; b.cnt = cnt[i];
        return b;
    }
};

// Rows [0, n): wave steps of kGChunk rows go to waves round-robin (wave
// wid takes steps wid, wid + waves, ...).  VEC: keys and values 16-byte
// aligned, so a lane reads its 4 consecutive rows with non-temporal 16-byte
// loads.  CHECK: the key window [kmin, kmin + nbins) came from a sample;
// rows outside it go to a spare bin (index nbins) that is never read and
// are counted in range[2], and the exact key range is folded into range[0]
// (min) and range[1] (max).  Each wave copy holds nbins + 1 bins.
template <typename K, bool VEC, bool CHECK>
__global__ __launch_bounds__(kGThreads) void k_gb_dense(const K* __restrict__ keys, const double* __restrict__ vals,
                                                        uint64_t n, int64_t kmin, uint32_t nbins,
                                                        GBin* __restrict__ parts, long long* __restrict__ range) {
    extern __shared__ double gsm[];
    const uint32_t nb1 = nbins + 1;
    const uint32_t tot = kGWaves * nb1;
    WaveBins all{gsm, gsm + tot, reinterpret_cast<double2*>(gsm + 2 * tot), reinterpret_cast<uint32_t*>(gsm + 4 * tot)};
    for (uint32_t i = threadIdx.x; i < tot; i += kGThreads) {
        all.s[i] = 0.0; all.c[i] = 0.0; all.cnt[i] = 0;
        all.mm[i] = make_double2(__builtin_huge_val(), -__builtin_huge_val());
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t off = (uint32_t)w * nb1;
    WaveBins mine{all.s + off, all.c + off, all.mm + off, all.cnt + off};
    K klo = std::numeric_limits<K>::max(), khi = std::numeric_limits<K>::lowest();
    uint32_t miss = 0;
    auto row = [&](K k, double v) {
        uint32_t b = (uint32_t)((int64_t)k - kmin);
        if constexpr (CHECK) {
            klo = k < klo ? k : klo;
            khi = k > khi ? k : khi;
            miss += b >= nbins ? 1u : 0u;
            b = b < nbins ? b : nbins;
        }
        mine.add(b, v);
    };
    const uint64_t waves = (uint64_t)gridDim.x * kGWaves;
    const uint64_t wid = (uint64_t)blockIdx.x * kGWaves + w;
    const uint64_t steps = n / kGChunk;
    uint64_t st = wid;
    for (; st + (kGUnroll - 1) * waves < steps; st += kGUnroll * waves) {
        K kk[kGUnroll][4];
        double v[kGUnroll][4];
#pragma unroll
        for (int u = 0; u < kGUnroll; ++u) {
            const uint64_t base = (st + (uint64_t)u * waves) * kGChunk;
            if constexpr (VEC) {
                uint4 k4 = ld16<uint4, true>(reinterpret_cast<const uint4*>(keys + base + 4 * lane));
                double2 v0 = ld16<double2, true>(reinterpret_cast<const double2*>(vals + base + 4 * lane));
                double2 v1 = ld16<double2, true>(reinterpret_cast<const double2*>(vals + base + 4 * lane + 2));
                kk[u][0] = (K)k4.x; kk[u][1] = (K)k4.y; kk[u][2] = (K)k4.z; kk[u][3] = (K)k4.w;
                v[u][0] = v0.x; v[u][1] = v0.y; v[u][2] = v1.x; v[u][3] = v1.y;
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    kk[u][j] = keys[base + 64 * j + lane];
                    v[u][j] = vals[base + 64 * j + lane];
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kGUnroll; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) row(kk[u][j], v[u][j]);
    }
    for (; st < steps; st += waves) {
        const uint64_t base = st * kGChunk;
#pragma unroll
        for (int j = 0; j < 4; ++j) row(keys[base + 64 * j + lane], vals[base + 64 * j + lane]);
    }
    if (wid == waves - 1)                                   // rows past the last whole step
        for (uint64_t r = steps * kGChunk + lane; r < n; r += 64) row(keys[r], vals[r]);
    if constexpr (CHECK) {
        int64_t l = (int64_t)klo, h = (int64_t)khi;
        for (int m = 32; m > 0; m >>= 1) {
            const int64_t x = __shfl_xor(l, m, 64), y = __shfl_xor(h, m, 64);
            l = x < l ? x : l;
            h = y > h ? y : h;
        }
        uint32_t msum = miss;
        if (__ballot(miss != 0))
            for (int m = 32; m > 0; m >>= 1) msum += __shfl_xor(msum, m, 64);
        if (lane == 0) {
            atomicMin(&range[0], (long long)l);
            atomicMax(&range[1], (long long)h);
            if (msum) atomicAdd((unsigned long long*)&range[2], (unsigned long long)msum);
        }
    }
    __syncthreads();
    GBin* out = parts + (size_t)blockIdx.x * nbins;
    for (uint32_t i = threadIdx.x; i < nbins; i += kGThreads) {
        GBin acc = all.get(i);
        for (int ww = 1; ww < kGWaves; ++ww) bin_merge(acc, all.get((uint32_t)ww * nb1 + i));
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

// range[0..1] = key range of a sample: rows i * n / S for i < S (every row
// when n <= S), one row per thread; range[0..4] start as (max, min, max,
// min, 0) from k_range_init, and the checked dense kernel folds the exact
// range and the rows outside the window into range[2..4].
constexpr uint32_t kSample = 65536;
__global__ void k_range_init(long long* range) {
    range[0] = INT64_MAX; range[1] = INT64_MIN; range[2] = INT64_MAX; range[3] = INT64_MIN; range[4] = 0;
}
template <typename K>
__global__ __launch_bounds__(256) void k_key_sample(const K* __restrict__ keys, uint64_t n,
                                                    long long* __restrict__ range) {
    const uint64_t S = n < kSample ? n : kSample;
    const uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    if (i < S) {
        const uint64_t r = S == n ? i : (uint64_t)(((unsigned __int128)i * n) / S);
        lo = hi = (int64_t)keys[r];
    }
    for (int m = 32; m > 0; m >>= 1) {
        const int64_t x = __shfl_xor(lo, m, 64), y = __shfl_xor(hi, m, 64);
        lo = x < lo ? x : lo;
        hi = y > hi ? y : hi;
    }
    if ((threadIdx.x & 63) == 0 && lo <= hi) {
        atomicMin(&range[0], (long long)lo);
        atomicMax(&range[1], (long long)hi);
    }
}

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

// One dense pass with key window [kmin, kmin + nbins).  check: the window
// is a guess; rows outside it are counted (w.mm_host[4] after return) and
// the exact key range comes back in w.mm_host[2..3].
static int64_t dense_pass(GroupbyWork& w, int key_dtype, const void* keys, const double* vals, uint64_t n,
                          int64_t kmin, uint32_t nbins, bool check, int64_t* okey, GBin* orec, StreamCtx* c) {
    hipStream_t st = c->stream;
    const size_t lds = (size_t)kGWaves * (nbins + 1) * kGBinBytes;
    const bool vec = ((uintptr_t)keys & 15) == 0 && ((uintptr_t)vals & 15) == 0;
    const bool i32 = key_dtype == DT_I32;
    const void* fns[2][2][2] = {
        {{(const void*)k_gb_dense<uint32_t, false, false>, (const void*)k_gb_dense<uint32_t, false, true>},
         {(const void*)k_gb_dense<uint32_t, true, false>, (const void*)k_gb_dense<uint32_t, true, true>}},
        {{(const void*)k_gb_dense<int32_t, false, false>, (const void*)k_gb_dense<int32_t, false, true>},
         {(const void*)k_gb_dense<int32_t, true, false>, (const void*)k_gb_dense<int32_t, true, true>}}};
    const void* fn = fns[i32][vec][check];
    int occ = 0;
    if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kGThreads, lds), "occupancy")) return -1;
    if (occ < 1) occ = 1;
    unsigned blocks = (unsigned)device_cu_count() * (unsigned)occ;
    const uint64_t need = (n + (uint64_t)kGChunk * kGUnroll * kGWaves - 1) / ((uint64_t)kGChunk * kGUnroll * kGWaves);
    if (need < blocks) blocks = (unsigned)(need ? need : 1);
    // a wave's bin counts are u32: it sees at most ceil(steps / waves) steps plus the tail
    const uint64_t per_wave =
        (n / kGChunk + (uint64_t)blocks * kGWaves - 1) / ((uint64_t)blocks * kGWaves) * kGChunk + kGChunk;
    if (per_wave >= (1ull << 32)) {
        set_error("group-by: %llu rows exceed the dense path's per-wave count", (unsigned long long)n);
        return -1;
    }
    const uint64_t recs = (uint64_t)blocks * nbins;
    if (recs > w.parts_cap) {
        if (w.parts) (void)hipFree(w.parts);
        w.parts = nullptr; w.parts_cap = 0;
        if (!hip_ok(hipMalloc((void**)&w.parts, recs * sizeof(GBin)), "hipMalloc(partials)")) return -1;
        w.parts_cap = recs;
    }
    long long* range = w.mm + 2;                  // exact min, max, rows outside the window
    void* args[] = {(void*)&keys, (void*)&vals, (void*)&n, (void*)&kmin, (void*)&nbins, (void*)&w.parts, (void*)&range};
    if (!hip_ok(hipLaunchKernel(fn, dim3(blocks), dim3(kGThreads), args, lds, st), "group-by launch")) return -1;
    hipLaunchKernelGGL(k_gb_final, dim3(nbins), dim3(256), 0, st, w.parts, blocks, nbins, w.bins);
    if (!hip_ok(hipGetLastError(), "group-by final launch")) return -1;
    if (check && !hip_ok(hipMemcpyAsync(w.mm_host + 2, w.mm + 2, 24, hipMemcpyDeviceToHost, st), "read key range"))
        return -1;
    DenseEmit src{w.bins, kmin, okey, orec};
    uint64_t ng = 0;
    if (!run_compact(src, nbins, c, &ng)) return -1;      // synchronises the stream
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
    if (key_dtype != DT_I32 && key_dtype != DT_U32 && key_dtype != DT_I64) {
        set_error("group-by: key dtype %d not supported (i32, u32, i64)", key_dtype);
        return -1;
    }
    if (!force_general && k32) {
        // key window from a sample of kSample rows (all rows when n <= kSample)
        const unsigned sblocks = (unsigned)(((n < kSample ? n : kSample) + 255) / 256);
        hipLaunchKernelGGL(k_range_init, dim3(1), dim3(1), 0, st, w.mm);
        if (key_dtype == DT_I32) hipLaunchKernelGGL(k_key_sample<int32_t>, dim3(sblocks), dim3(256), 0, st, (const int32_t*)keys, n, w.mm);
        else hipLaunchKernelGGL(k_key_sample<uint32_t>, dim3(sblocks), dim3(256), 0, st, (const uint32_t*)keys, n, w.mm);
        if (!hip_ok(hipGetLastError(), "sample launch") ||
            !hip_ok(hipMemcpyAsync(w.mm_host, w.mm, 16, hipMemcpyDeviceToHost, st), "read key range") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
            return -1;
        int64_t kmin = w.mm_host[0];
        uint64_t span = (uint64_t)(w.mm_host[1] - kmin) + 1;
        bool exact = n <= kSample;
        while (span <= (uint64_t)kMaxDense) {
            int64_t ng = dense_pass(w, key_dtype, keys, vals, n, kmin, (uint32_t)span, !exact, okey, orec, c);
            if (ng < 0) return -1;
            if (exact || w.mm_host[4] == 0) return ng;
            kmin = w.mm_host[2];                     // a key outside the sampled window: exact range known now
            span = (uint64_t)(w.mm_host[3] - kmin) + 1;
            exact = true;
        }
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
