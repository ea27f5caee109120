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
// Residue path (32-bit keys, voxel_gpu_groupby_f64): the same per-wave LDS
// bins indexed by key mod kMaxDense, so no key window is needed before the
// pass; each workgroup records its exact key range and the result stands
// when the whole range spans at most kMaxDense values (distinct residues).
// Bin merge and output kernels write the caller's arrays and the group
// count to pinned memory: one host synchronisation per call.
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
constexpr int kResidSteps = 2;       // k_gb_resid: wave steps loaded before the LDS updates

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
// min / max atomics are issued only when the row beats the pair read.
// add_n issues the sum, count and pair reads of N rows back to back and
// only then the dependent compensation and min / max atomics, so a wave
// keeps N LDS round trips in flight instead of one per row.  That is exact
// whatever the order: TwoSum of the value an atomic add returned recovers
// that add's rounding error, and min / max only move one way, so a pair
// read before another row's update is a conservative test.
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
    // rows whose bin is `skip` are not added (no LDS access)
    template <int N>
    __device__ __forceinline__ void add_n(const uint32_t (&b)[N], const double (&v)[N], uint32_t skip) const {
        double old[N];
        double2 p[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip)
                old[j] = __hip_atomic_fetch_add(&s[b[j]], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip) __hip_atomic_fetch_add(&cnt[b[j]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip) p[j] = mm[b[j]];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (b[j] == skip) continue;
            const double t = old[j] + v[j];
            const double bp = t - old[j];
            const double e = (old[j] - (t - bp)) + (v[j] - bp);
            if (e != 0.0) __hip_atomic_fetch_add(&c[b[j]], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (v[j] < p[j].x) __hip_atomic_fetch_min(&mm[b[j]].x, v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (v[j] > p[j].y) __hip_atomic_fetch_max(&mm[b[j]].y, v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        }
    }
    __device__ __forceinline__ GBin get(uint32_t i) const {
        GBin b;
        const double2 p = mm[i];
        b.s = s[i]; b.c = c[i]; b.mn = p.x; b.mx = p.y; b.cnt = cnt[i];
        return b;
    }
};

// Key range of a sample of the key column: rows i * n / S for i < S (every
// row when n <= S); workgroup b of kSampBlocks writes the (min, max) of its
// rows to samp[2b], samp[2b + 1] (INT64_MAX, INT64_MIN when it saw none).
constexpr uint32_t kSample = 65536;            // 2^16
static_assert(kSample == (1u << 16), "the sample row index divides by 2^16 with a shift");
constexpr int kSampBlocks = 32, kSampThreads = 512;
constexpr int kSampPer = (int)(kSample / (kSampBlocks * kSampThreads));

__device__ __forceinline__ void wave_minmax(int64_t& lo, int64_t& hi) {
    for (int m = 32; m > 0; m >>= 1) {
        const int64_t x = __shfl_xor(lo, m, 64), y = __shfl_xor(hi, m, 64);
        lo = x < lo ? x : lo;
        hi = y > hi ? y : hi;
    }
}

template <typename K>
__global__ __launch_bounds__(kSampThreads) void k_key_sample(const K* __restrict__ keys, uint64_t n,
                                                             long long* __restrict__ samp) {
    __shared__ int64_t red[2][kSampThreads / 64];
    const uint64_t S = n < kSample ? n : kSample;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
#pragma unroll
    for (int j = 0; j < kSampPer; ++j) {
        const uint64_t i = (uint64_t)j * (kSampBlocks * kSampThreads) + (uint64_t)blockIdx.x * kSampThreads + threadIdx.x;
        if (i < S) {
            // S == n, or S == kSample = 2^16: row floor(i * n / 2^16) from the 128-bit product
            const uint64_t r = S == n ? i : ((__umul64hi(i, n) << 48) | ((i * n) >> 16));
            const int64_t k = (int64_t)keys[r];
            lo = k < lo ? k : lo;
            hi = k > hi ? k : hi;
        }
    }
    wave_minmax(lo, hi);
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { red[0][w] = lo; red[1][w] = hi; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < kSampThreads / 64; ++ww) {
            lo = red[0][ww] < lo ? red[0][ww] : lo;
            hi = red[1][ww] > hi ? red[1][ww] : hi;
        }
        samp[2 * blockIdx.x] = lo;
        samp[2 * blockIdx.x + 1] = hi;
    }
}

// Exact key range and rows outside the window over the checked dense
// workgroups' (min, max, outside) records, by one whole wave; lane 0
// writes out[0..2].
__device__ __forceinline__ void fold_ranges(const long long* __restrict__ rngp, uint32_t nblocks,
                                            long long* __restrict__ out) {
    const int lane = threadIdx.x & 63;
    int64_t lo = INT64_MAX, hi = INT64_MIN;
    unsigned long long miss = 0;
    for (uint32_t j = lane; j < nblocks; j += 64) {
        lo = rngp[3 * j] < lo ? rngp[3 * j] : lo;
        hi = rngp[3 * j + 1] > hi ? rngp[3 * j + 1] : hi;
        miss += (unsigned long long)rngp[3 * j + 2];
    }
    wave_minmax(lo, hi);
    for (int m = 32; m > 0; m >>= 1) miss += __shfl_xor(miss, m, 64);
    if (lane == 0) { out[0] = lo; out[1] = hi; out[2] = (long long)miss; }
}

__device__ __forceinline__ bool window_ok(int64_t lo, int64_t hi) {
    return hi >= lo && (uint64_t)(hi - lo) < (uint64_t)kMaxDense;
}

// Rows [0, n): wave steps of kGChunk rows go to waves round-robin (wave
// wid takes steps wid, wid + waves, ...).  VEC: keys and values 16-byte
// aligned, so a lane reads its 4 consecutive rows with non-temporal 16-byte
// loads.  CHECK: the key window [kmin, kmin + nbins) came from a sample;
// rows outside it go to a spare bin (index nbins) that is never read, and
// each workgroup writes (min key, max key, rows outside) to rngp[3b..3b+2]
// (plain stores: no contended atomics).  Each wave copy holds nbins + 1
// bins.
template <typename K, bool VEC, bool CHECK>
__global__ __launch_bounds__(kGThreads) void k_gb_dense(const K* __restrict__ keys, const double* __restrict__ vals,
                                                        uint64_t n, int64_t kmin, uint32_t nbins,
                                                        GBin* __restrict__ parts, long long* __restrict__ rngp) {
    extern __shared__ double gsm[];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t nb1 = nbins + 1;
    const uint32_t tot = kGWaves * nb1;
    WaveBins all{gsm, gsm + tot, reinterpret_cast<double2*>(gsm + 2 * tot), reinterpret_cast<uint32_t*>(gsm + 4 * tot)};
    for (uint32_t i = threadIdx.x; i < tot; i += kGThreads) {
        all.s[i] = 0.0; all.c[i] = 0.0; all.cnt[i] = 0;
        all.mm[i] = make_double2(__builtin_huge_val(), -__builtin_huge_val());
    }
    __syncthreads();
    const uint32_t off = (uint32_t)w * nb1;
    WaveBins mine{all.s + off, all.c + off, all.mm + off, all.cnt + off};
    K klo = std::numeric_limits<K>::max(), khi = std::numeric_limits<K>::lowest();
    uint32_t miss = 0;
    auto bin_of = [&](K k) {
        uint32_t b = (uint32_t)((int64_t)k - kmin);
        if constexpr (CHECK) {
            klo = k < klo ? k : klo;
            khi = k > khi ? k : khi;
            miss += b >= nbins ? 1u : 0u;
            b = b < nbins ? b : nbins;
        }
        return b;
    };
    auto row = [&](K k, double v) { mine.add(bin_of(k), v); };
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
        uint32_t bb[kGUnroll * 4];
        double vv[kGUnroll * 4];
#pragma unroll
        for (int u = 0; u < kGUnroll; ++u)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                bb[u * 4 + j] = bin_of(kk[u][j]);
                vv[u * 4 + j] = v[u][j];
            }
        mine.add_n(bb, vv, nbins);                   // rows outside a checked window are dropped
    }
    for (; st < steps; st += waves) {
        const uint64_t base = st * kGChunk;
#pragma unroll
        for (int j = 0; j < 4; ++j) row(keys[base + 64 * j + lane], vals[base + 64 * j + lane]);
    }
    if (wid == waves - 1)                                   // rows past the last whole step
        for (uint64_t r = steps * kGChunk + lane; r < n; r += 64) row(keys[r], vals[r]);
    __shared__ int64_t s_rng[3][kGWaves];
    if constexpr (CHECK) {
        int64_t l = (int64_t)klo, h = (int64_t)khi;
        wave_minmax(l, h);
        uint32_t msum = miss;
        for (int m = 32; m > 0; m >>= 1) msum += __shfl_xor(msum, m, 64);
        if (lane == 0) { s_rng[0][w] = l; s_rng[1][w] = h; s_rng[2][w] = msum; }
    }
    __syncthreads();
    if constexpr (CHECK) {
        if (threadIdx.x == 0) {
            int64_t l = s_rng[0][0], h = s_rng[1][0], m = s_rng[2][0];
            for (int ww = 1; ww < kGWaves; ++ww) {
                l = s_rng[0][ww] < l ? s_rng[0][ww] : l;
                h = s_rng[1][ww] > h ? s_rng[1][ww] : h;
                m += s_rng[2][ww];
            }
            rngp[3 * blockIdx.x] = l;
            rngp[3 * blockIdx.x + 1] = h;
            rngp[3 * blockIdx.x + 2] = m;
        }
    }
    GBin* out = parts + (size_t)blockIdx.x * nbins;
    for (uint32_t i = threadIdx.x; i < nbins; i += kGThreads) {
        GBin acc = all.get(i);
        for (int ww = 1; ww < kGWaves; ++ww) bin_merge(acc, all.get((uint32_t)ww * nb1 + i));
        out[i] = acc;
    }
}

// f32 guards of an exact f64 minimum / maximum: guard_up(v) >= v and
// guard_dn(v) <= v, never an f32 subnormal (LDS float atomics may flush
// those to zero, which would break the bound).
__device__ __forceinline__ float guard_up(double v) {
    const float f = __double2float_ru(v);
    return (f > 0.0f && f < __FLT_MIN__) ? __FLT_MIN__ : f;
}
__device__ __forceinline__ float guard_dn(double v) {
    const float f = __double2float_rd(v);
    return (f < 0.0f && f > -__FLT_MIN__) ? -__FLT_MIN__ : f;
}

// Bins of the residue path.  Sum and compensation are per wave copy (f64
// adds round differently in another order, so only one wave writes a
// copy); count, exact (min, max) and the f32 guard pair are shared by the
// workgroup: integer adds and min / max give the same bits in any order.
// A row costs ds_add_rtn_f64 (sum), ds_add_f64 (compensation, skipped when
// the rounding error is zero), ds_add_u32 (count) and one ds_read_b64 of
// its guard pair; the exact min / max atomics (and the guard's) are issued
// only when the row is below the min guard or above the max guard.  The
// guards bound the exact values (gd.x >= mm.x, gd.y <= mm.y): a row's
// guard update is issued together with its exact update, so a row at or
// above a guard it reads is at or above the final minimum.
struct ResidBins {
    double* s;           // this wave's copy
    double* c;
    double2* mm;         // shared by the workgroup
    float2* gd;
    uint32_t* cnt;
    template <int N>
    __device__ __forceinline__ void add_n(const uint32_t (&b)[N], const double (&v)[N], uint32_t skip) const {
        double old[N];
        float2 g[N];
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip)
                old[j] = __hip_atomic_fetch_add(&s[b[j]], v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip) __hip_atomic_fetch_add(&cnt[b[j]], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
#pragma unroll
        for (int j = 0; j < N; ++j)
            if (b[j] != skip) g[j] = gd[b[j]];
#pragma unroll
        for (int j = 0; j < N; ++j) {
            if (b[j] == skip) continue;
            const double t = old[j] + v[j];
            const double bp = t - old[j];
            const double e = (old[j] - (t - bp)) + (v[j] - bp);
            if (e != 0.0) __hip_atomic_fetch_add(&c[b[j]], e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (v[j] < (double)g[j].x) {
                __hip_atomic_fetch_min(&mm[b[j]].x, v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_min(&gd[b[j]].x, guard_up(v[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
            if (v[j] > (double)g[j].y) {
                __hip_atomic_fetch_max(&mm[b[j]].y, v[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                __hip_atomic_fetch_max(&gd[b[j]].y, guard_dn(v[j]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            }
        }
    }
};
// LDS bytes of the residue path's bins: per wave copy (s, c), per workgroup
// (mm, gd, cnt), kMaxDense + 1 bins each (the last one the spare bin).
constexpr size_t kResidLds = ((size_t)kGWaves * 16 + 16 + 8 + 4) * (kMaxDense + 1);

// Dense group-by without a key window: a row goes to bin key mod kMaxDense
// (the low bits of the key's two's-complement value), so any key range of
// at most kMaxDense consecutive values lands in distinct bins and the range
// itself need not be known up front.  Each workgroup writes its exact
// (min key, max key) to rngp[2b..2b+1]; the merge kernel accepts the result
// only when max - min < kMaxDense.  Rows [0, n) in wave steps of kGChunk
// rows, round-robin over the waves, U steps (4U rows per lane) loaded
// before the LDS updates; the rows after the last whole step are read row
// by row with bounds checks by the wave whose turn that step is.  Steps
// past the end and rows past n go to the spare bin kMaxDense, which add_n
// skips.  VEC as in k_gb_dense.
template <typename K, bool VEC, int U>
__global__ __launch_bounds__(kGThreads) void k_gb_resid(const K* __restrict__ keys, const double* __restrict__ vals,
                                                        uint64_t n, GBin* __restrict__ parts,
                                                        long long* __restrict__ rngp) {
    static_assert(U % 2 == 0, "rows go to add_n eight at a time");
    extern __shared__ double gsm[];
    constexpr uint32_t nb1 = kMaxDense + 1, tot = kGWaves * nb1, spare = kMaxDense;
    constexpr uint32_t rmask = kMaxDense - 1;
    // the wave index through readfirstlane: step addresses become scalar, so
    // the loads take a scalar base and a 32-bit lane offset
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    double* s_all = gsm;
    double* c_all = gsm + tot;
    double2* mm = reinterpret_cast<double2*>(gsm + 2 * tot);
    float2* gd = reinterpret_cast<float2*>(mm + nb1);
    uint32_t* cnt = reinterpret_cast<uint32_t*>(gd + nb1);
    for (uint32_t i = threadIdx.x; i < tot; i += kGThreads) { s_all[i] = 0.0; c_all[i] = 0.0; }
    for (uint32_t i = threadIdx.x; i < nb1; i += kGThreads) {
        mm[i] = make_double2(__builtin_huge_val(), -__builtin_huge_val());
        gd[i] = make_float2(__builtin_huge_valf(), -__builtin_huge_valf());
        cnt[i] = 0;
    }
    __syncthreads();
    const uint32_t off = (uint32_t)w * nb1;
    ResidBins mine{s_all + off, c_all + off, mm, gd, cnt};
    K klo = std::numeric_limits<K>::max(), khi = std::numeric_limits<K>::lowest();
    const uint64_t waves = (uint64_t)gridDim.x * kGWaves;
    const uint64_t wid = (uint64_t)blockIdx.x * kGWaves + w;
    const uint64_t full = n / kGChunk;
    // whole steps; a step past `full` (the last iteration's) goes to the spare bin
    for (uint64_t st = wid; st < full; st += (uint64_t)U * waves) {
        K kk[U][4];
        double v[U][4];
        bool live[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t sx = st + (uint64_t)u * waves;
            const uint64_t base = sx * kGChunk;
            live[u] = sx < full;
            if (live[u]) {
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
            } else {
#pragma unroll
                for (int j = 0; j < 4; ++j) { kk[u][j] = 0; v[u][j] = 0.0; }
            }
        }
#pragma unroll
        for (int h = 0; h < U; h += 2) {
            uint32_t bb[8];
            double vv[8];
#pragma unroll
            for (int u = 0; u < 2; ++u)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    const K k = kk[h + u][j];
                    // after every load of the batch is issued: a use of a key next
                    // to its load would wait for it before the next step's loads
                    klo = live[h + u] && k < klo ? k : klo;
                    khi = live[h + u] && k > khi ? k : khi;
                    bb[u * 4 + j] = live[h + u] ? ((uint32_t)k & rmask) : spare;
                    vv[u * 4 + j] = v[h + u][j];
                }
            mine.add_n(bb, vv, spare);
        }
    }
    // rows past the last whole step: the wave whose turn step `full` is
    if (full * kGChunk < n && wid == full % waves) {
        uint32_t bb[4];
        double vv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const uint64_t r = full * kGChunk + 64 * j + lane;
            bb[j] = spare;
            vv[j] = 0.0;
            if (r < n) {
                const K k = keys[r];
                klo = k < klo ? k : klo;
                khi = k > khi ? k : khi;
                bb[j] = (uint32_t)k & rmask;
                vv[j] = vals[r];
            }
        }
        mine.add_n(bb, vv, spare);
    }
    __shared__ int64_t s_rng[2][kGWaves];
    int64_t l = (int64_t)klo, hh = (int64_t)khi;
    wave_minmax(l, hh);
    if (lane == 0) { s_rng[0][w] = l; s_rng[1][w] = hh; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int ww = 1; ww < kGWaves; ++ww) {
            l = s_rng[0][ww] < l ? s_rng[0][ww] : l;
            hh = s_rng[1][ww] > hh ? s_rng[1][ww] : hh;
        }
        rngp[2 * blockIdx.x] = l;
        rngp[2 * blockIdx.x + 1] = hh;
    }
    GBin* out = parts + (size_t)blockIdx.x * kMaxDense;
    for (uint32_t i = threadIdx.x; i < (uint32_t)kMaxDense; i += kGThreads) {
        GBin acc;
        acc.s = s_all[i]; acc.c = c_all[i];
        for (int ww = 1; ww < kGWaves; ++ww) {
            double cc = acc.c + c_all[(uint32_t)ww * nb1 + i];
            two_sum<double>(acc.s, cc, s_all[(uint32_t)ww * nb1 + i]);
            acc.c = cc;
        }
        const double2 p = mm[i];
        acc.mn = p.x; acc.mx = p.y; acc.cnt = cnt[i];
        out[i] = acc;
    }
}

// one workgroup per bin: partials of all workgroups merged in a fixed order;
// with rngp, the first wave of workgroup 0 also folds the checked dense
// workgroups' range records into mm[2..4]
__global__ __launch_bounds__(256) void k_gb_final(const GBin* __restrict__ parts, uint32_t nblocks, uint32_t nbins,
                                                  GBin* __restrict__ out, const long long* __restrict__ rngp,
                                                  long long* __restrict__ mm) {
    __shared__ GBin red[256];
    const uint32_t bin = blockIdx.x;
    if (rngp && bin == 0 && threadIdx.x < 64) fold_ranges(rngp, nblocks, mm + 2);
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

__device__ __forceinline__ void bin_shfl_merge(GBin& a, int m) {
    GBin o;
    o.s = __shfl_xor(a.s, m, 64);
    o.c = __shfl_xor(a.c, m, 64);
    o.mn = __shfl_xor(a.mn, m, 64);
    o.mx = __shfl_xor(a.mx, m, 64);
    o.cnt = __shfl_xor(a.cnt, m, 64);
    bin_merge(a, o);
}

// Bin merge of k_gb_resid's partials: one wave per bin; lane l merges the
// partials of workgroups l, l + 64, ... in order, then a fixed butterfly;
// lane 0 stores the bin.
constexpr int kMergeThreads = 256;
__global__ __launch_bounds__(kMergeThreads) void k_gb_merge_resid(const GBin* __restrict__ parts, uint32_t nblocks,
                                                                   GBin* __restrict__ bins) {
    const int lane = threadIdx.x & 63;
    const uint32_t bin = blockIdx.x * (kMergeThreads / 64) + (threadIdx.x >> 6);
    if (bin >= (uint32_t)kMaxDense) return;
    GBin acc; bin_init(acc);
    for (uint32_t j = lane; j < nblocks; j += 64) bin_merge(acc, parts[(size_t)j * kMaxDense + bin]);
    for (int m = 1; m < 64; m <<= 1) bin_shfl_merge(acc, m);
    if (lane == 0) bins[bin] = acc;
}

// Output of the residue path, one workgroup of kMaxDense threads: folds the
// workgroups' key ranges and, when max - min < kMaxDense, writes the
// non-empty bins in key order (key lo + t is bin (lo + t) mod kMaxDense)
// straight into the caller's arrays (the k_gb_finish formulas; rows past
// max_groups are not written).  Result in pinned host memory: hres[0] =
// group count, or -1 when the key range is wider than kMaxDense;
// hres[2..3] = the key range.
__global__ __launch_bounds__(kMaxDense) void k_gb_emit_resid(
    const GBin* __restrict__ bins, uint32_t nblocks, const long long* __restrict__ rngp, uint64_t max_groups,
    int64_t* __restrict__ ok, uint64_t* __restrict__ oc, double* __restrict__ os, double* __restrict__ oa,
    double* __restrict__ omn, double* __restrict__ omx, volatile long long* __restrict__ hres) {
    __shared__ uint32_t wc[kMaxDense / 64];
    __shared__ int64_t s_win[2];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (w == 0) {
        int64_t lo = INT64_MAX, hi = INT64_MIN;
        for (uint32_t j = lane; j < nblocks; j += 64) {
            lo = rngp[2 * j] < lo ? rngp[2 * j] : lo;
            hi = rngp[2 * j + 1] > hi ? rngp[2 * j + 1] : hi;
        }
        wave_minmax(lo, hi);
        if (lane == 0) { s_win[0] = lo; s_win[1] = hi; }
    }
    __syncthreads();
    const int64_t lo = s_win[0], hi = s_win[1];
    if (!window_ok(lo, hi)) {
        if (threadIdx.x == 0) { hres[0] = -1; hres[2] = lo; hres[3] = hi; }
        return;
    }
    const uint32_t nbins = (uint32_t)(hi - lo) + 1, t = threadIdx.x;
    GBin b;
    const bool live = t < nbins && (b = bins[(uint32_t)(lo + (int64_t)t) & (uint32_t)(kMaxDense - 1)], b.cnt > 0);
    const unsigned long long msk = __ballot(live);
    if (lane == 0) wc[w] = (uint32_t)__popcll(msk);
    __syncthreads();
    uint32_t pos = (uint32_t)__popcll(msk & (lane ? (~0ull >> (64 - lane)) : 0ull));
    for (int ww = 0; ww < w; ++ww) pos += wc[ww];
    if (live && pos < max_groups) {
        const double sum = b.s + b.c;
        if (ok) ok[pos] = lo + (int64_t)t;
        if (oc) oc[pos] = b.cnt;
        if (os) os[pos] = sum;
        if (oa) oa[pos] = b.cnt ? sum / (double)b.cnt : 0.0;     // AggregateState::Avg (aggregate.hpp:136-139)
        if (omn) omn[pos] = b.mn;
        if (omx) omx[pos] = b.mx;
    }
    if (t == 0) {
        uint32_t c = 0;
        for (int ww = 0; ww < kMaxDense / 64; ++ww) c += wc[ww];
        hres[0] = c; hres[2] = lo; hres[3] = hi;
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
    long long* mm = nullptr;     // device: [2..4] folded ranges of a checked dense pass
    long long* mm_host = nullptr;  // pinned: [0..4] key window and ranges, [8..11] residue-path result, [16..] sample records
    long long* samp = nullptr;   // kSampBlocks (min, max) sample records
    long long* rngp = nullptr;   // (min, max, outside) per checked dense workgroup
    uint32_t rng_cap = 0;
    // general path
    uint64_t* k1 = nullptr; uint64_t* k2 = nullptr; uint32_t* r1 = nullptr; uint32_t* r2 = nullptr;
    uint64_t* starts = nullptr; GBin* rec = nullptr;
    uint64_t cap = 0;
    // results (ascending key), two sets so that a merge can read one
    int64_t* okey[2] = {nullptr, nullptr};
    GBin* orec[2] = {nullptr, nullptr};
    uint64_t ocap[2] = {0, 0};
    int cur = 0;
    int dev_occ = 0;             // workgroups per CU of the device-window dense kernel
};

static GroupbyWork& gwork(StreamCtx* c) {
    static std::vector<std::pair<StreamCtx*, GroupbyWork*>> all;
    for (auto& p : all) if (p.first == c) return *p.second;
    all.emplace_back(c, new GroupbyWork());
    return *all.back().second;
}

static bool ensure_small(GroupbyWork& w, hipStream_t st) {
    if (w.mm) return true;
    if (!hip_ok(hipMalloc((void**)&w.mm, 64), "hipMalloc") || !hip_ok(hipMemsetAsync(w.mm, 0, 64, st), "hipMemsetAsync") ||
        !hip_ok(hipHostMalloc((void**)&w.mm_host, 1024, hipHostMallocDefault), "hipHostMalloc") ||
        !hip_ok(hipMalloc((void**)&w.bins, kMaxDense * sizeof(GBin)), "hipMalloc") ||
        !hip_ok(hipMalloc((void**)&w.samp, 2 * kSampBlocks * sizeof(long long)), "hipMalloc"))
        return false;
    return true;
}

static bool ensure_rng(GroupbyWork& w, uint32_t blocks) {
    if (blocks <= w.rng_cap) return true;
    if (w.rngp) (void)hipFree(w.rngp);
    w.rngp = nullptr; w.rng_cap = 0;
    if (!hip_ok(hipMalloc((void**)&w.rngp, (size_t)blocks * 3 * sizeof(long long)), "hipMalloc(ranges)")) return false;
    w.rng_cap = blocks;
    return true;
}

// Launches the key sample into w.samp.
static bool launch_sample(GroupbyWork& w, int key_dtype, const void* keys, uint64_t n, hipStream_t st) {
    if (key_dtype == DT_I32)
        hipLaunchKernelGGL(k_key_sample<int32_t>, dim3(kSampBlocks), dim3(kSampThreads), 0, st, (const int32_t*)keys, n, w.samp);
    else
        hipLaunchKernelGGL(k_key_sample<uint32_t>, dim3(kSampBlocks), dim3(kSampThreads), 0, st, (const uint32_t*)keys, n, w.samp);
    return hip_ok(hipGetLastError(), "sample launch");
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
    if (check && !ensure_rng(w, blocks)) return -1;
    long long* rngp = check ? w.rngp : nullptr;   // per-workgroup exact min, max, rows outside the window
    const long long* no_samp = nullptr;
    long long* no_mm = nullptr;
    void* args[] = {(void*)&keys, (void*)&vals, (void*)&n, (void*)&kmin, (void*)&nbins, (void*)&w.parts,
                    (void*)&rngp, (void*)&no_samp, (void*)&no_mm};
    if (!hip_ok(hipLaunchKernel(fn, dim3(blocks), dim3(kGThreads), args, lds, st), "group-by launch")) return -1;
    hipLaunchKernelGGL(k_gb_final, dim3(nbins), dim3(256), 0, st, w.parts, blocks, nbins, w.bins,
                       (const long long*)rngp, w.mm);
    if (!hip_ok(hipGetLastError(), "group-by final launch")) return -1;
    if (check && !hip_ok(hipMemcpyAsync(w.mm_host + 2, w.mm + 2, 24, hipMemcpyDeviceToHost, st), "read key range"))
        return -1;
    DenseEmit src{w.bins, kmin, okey, orec};
    uint64_t ng = 0;
    if (!run_compact(src, nbins, c, &ng)) return -1;      // synchronises the stream
    return (int64_t)ng;
}

// The dense path of a 32-bit key column decided on the device: residue
// bins (k_gb_resid), the bin merge (k_gb_merge_resid) and the output
// (k_gb_emit_resid) into the caller's arrays, one host synchronisation.  Returns the group count; -2
// when the key range is wider than kMaxDense (the caller runs groupby_f64).
int64_t groupby_dense_device(int key_dtype, const void* keys, const double* vals, uint64_t n, uint64_t max_groups,
                             int64_t* ok, uint64_t* oc, double* os, double* oa, double* omn, double* omx,
                             StreamCtx* c) {
    GroupbyWork& w = gwork(c);
    hipStream_t st = c->stream;
    if (!ensure_small(w, st)) return -1;
    const bool i32 = key_dtype == DT_I32;
    const bool vec = ((uintptr_t)keys & 15) == 0 && ((uintptr_t)vals & 15) == 0;
    constexpr int U = kResidSteps;
    const void* fns[2][2] = {{(const void*)k_gb_resid<uint32_t, false, U>, (const void*)k_gb_resid<uint32_t, true, U>},
                             {(const void*)k_gb_resid<int32_t, false, U>, (const void*)k_gb_resid<int32_t, true, U>}};
    const void* fn = fns[i32][vec];
    const size_t lds = kResidLds;
    if (w.dev_occ == 0) {
        int occ = 0;
        if (!hip_ok(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, kGThreads, lds), "occupancy")) return -1;
        w.dev_occ = occ < 1 ? 1 : occ;
    }
    // every resident workgroup for large inputs; one per CU (half the
    // partials to merge) until each wave would see 16 steps
    const uint64_t cus = (uint64_t)device_cu_count();
    const uint64_t steps16 = (n + (uint64_t)kGChunk * kGWaves * 16 - 1) / ((uint64_t)kGChunk * kGWaves * 16);
    uint64_t nb = steps16 > cus ? steps16 : cus;
    if (nb > cus * (uint64_t)w.dev_occ) nb = cus * (uint64_t)w.dev_occ;
    const uint64_t per_block = (uint64_t)kGChunk * U * kGWaves;       // rows one workgroup takes per iteration
    const uint64_t need = (n + per_block - 1) / per_block;
    if (need < nb) nb = need ? need : 1;
    const unsigned blocks = (unsigned)nb;
    // a workgroup's bin counts are u32: it sees at most ceil(steps / waves) steps per wave
    const uint64_t steps = (n + kGChunk - 1) / kGChunk;
    if ((steps + (uint64_t)blocks * kGWaves - 1) / ((uint64_t)blocks * kGWaves) * kGChunk * kGWaves >= (1ull << 32)) {
        set_error("group-by: %llu rows exceed the dense path's per-wave count", (unsigned long long)n);
        return -1;
    }
    const uint64_t recs = (uint64_t)blocks * kMaxDense;
    if (recs > w.parts_cap) {
        if (w.parts) (void)hipFree(w.parts);
        w.parts = nullptr; w.parts_cap = 0;
        if (!hip_ok(hipMalloc((void**)&w.parts, recs * sizeof(GBin)), "hipMalloc(partials)")) return -1;
        w.parts_cap = recs;
    }
    if (!ensure_rng(w, blocks)) return -1;
    void* args[] = {(void*)&keys, (void*)&vals, (void*)&n, (void*)&w.parts, (void*)&w.rngp};
    volatile long long* hres = w.mm_host + 8;
    if (!hip_ok(hipLaunchKernel(fn, dim3(blocks), dim3(kGThreads), args, lds, st), "group-by launch")) return -1;
    hipLaunchKernelGGL(k_gb_merge_resid, dim3(kMaxDense / (kMergeThreads / 64)), dim3(kMergeThreads), 0, st,
                       (const GBin*)w.parts, blocks, w.bins);
    hipLaunchKernelGGL(k_gb_emit_resid, dim3(1), dim3(kMaxDense), 0, st, (const GBin*)w.bins, blocks,
                       (const long long*)w.rngp, max_groups, ok, oc, os, oa, omn, omx, hres);
    if (!hip_ok(hipGetLastError(), "group-by launch") || !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
        return -1;
    if (hres[0] < 0) return -2;                            // key range wider than kMaxDense
    return hres[0];
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
    if (!ensure_small(w, st)) return -1;
    if (key_dtype != DT_I32 && key_dtype != DT_U32 && key_dtype != DT_I64) {
        set_error("group-by: key dtype %d not supported (i32, u32, i64)", key_dtype);
        return -1;
    }
    if (!force_general && k32) {
        // key window from a sample of kSample rows (all rows when n <= kSample)
        long long* hs = w.mm_host + 16;
        if (!launch_sample(w, key_dtype, keys, n, st) ||
            !hip_ok(hipMemcpyAsync(hs, w.samp, 2 * kSampBlocks * sizeof(long long), hipMemcpyDeviceToHost, st),
                    "read key sample") ||
            !hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize"))
            return -1;
        for (int b = 1; b < kSampBlocks; ++b) {
            hs[0] = hs[2 * b] < hs[0] ? hs[2 * b] : hs[0];
            hs[1] = hs[2 * b + 1] > hs[1] ? hs[2 * b + 1] : hs[1];
        }
        w.mm_host[0] = hs[0];
        w.mm_host[1] = hs[1];
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
