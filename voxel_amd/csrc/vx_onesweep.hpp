// vx_onesweep.hpp -- the digit-pass kernels of vx_radix.hip as templates
// over the digit function, so that other kernel files can instantiate them
// (the learned partitioner's model digit, vx_learned.hip).  See vx_radix.hip
// for the algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_kernels.hpp"
#include "vx_radix.hpp"

namespace vx {

constexpr int kThreads = 256;           // = number of digit values
constexpr int kWaves = kThreads / 64;
constexpr int kRounds = 16;             // keys per lane per tile
constexpr int kTile = kThreads * kRounds;   // 4096 keys per tile

template <class Digit>
__global__ __launch_bounds__(kThreads) void k_hist(const uint64_t* __restrict__ keys, uint64_t n, Digit dig,
                                                   int passes, unsigned long long* __restrict__ ghist) {
    __shared__ uint32_t h[kWaves][8][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 8 * 256; i += kThreads) (&h[0][0][0])[i] = 0;
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * kThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kThreads + threadIdx.x; i < n; i += stride) {
        const uint64_t b = dig.bits(keys[i]);
        for (int p = 0; p < passes; ++p) atomicAdd(&h[w][p][dig.digit_of_bits(b, p)], 1u);
    }
    __syncthreads();
    for (int i = threadIdx.x; i < passes * 256; i += kThreads) {
        uint32_t s = 0;
        for (int ww = 0; ww < kWaves; ++ww) s += (&h[ww][0][0])[i];
        if (s) atomicAdd(&ghist[i], (unsigned long long)s);
    }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
    const int lane = threadIdx.x & 63;
    return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Lanes of this wave holding the same 8-bit digit (and a valid key).
__device__ __forceinline__ uint64_t match_digit(uint32_t d, bool valid) {
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        uint64_t m = __ballot((d >> b) & 1u);
        peers &= ((d >> b) & 1u) ? m : ~m;
    }
    return valid ? peers : 0ull;
}

// Rows of the key array are cut into ranges of whole 4096-key tiles, at
// most kMaxRanges of them.  Per digit pass: k_upsweep counts each range's
// digits, k_range_scan turns the counts into each range's first output
// position per digit, and k_downsweep (one workgroup per range, tiles in
// order) scatters stably.  No workgroup waits on another.
constexpr uint64_t kMaxRanges = 2048;     // and words of RadixWork::status: 2 x ranges x 256 + 256

__host__ __device__ inline uint64_t radix_tiles_per_range(uint64_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t r = (tiles + kMaxRanges - 1) / kMaxRanges;
    return r ? r : 1;
}

__host__ __device__ inline uint64_t radix_ranges(uint64_t n) {
    uint64_t tiles = (n + kTile - 1) / kTile;
    uint64_t per = radix_tiles_per_range(n);
    uint64_t r = (tiles + per - 1) / per;
    return r ? r : 1;
}

// Digit source of a pass: the key through the digit function, or (CIN) a
// carried 32-bit digit word whose byte `pass` is the digit.
template <class Digit, bool CIN>
__device__ __forceinline__ uint32_t pass_digit(const Digit& dig, uint64_t key, uint32_t word, int pass) {
    if constexpr (CIN) return (word >> (8 * pass)) & 0xFFu;
    else return dig.digit(key, pass);
}

// cnt[d * nr + range] = keys of the range with digit d
template <class Digit, bool CIN>
__global__ __launch_bounds__(kThreads) void k_upsweep(const uint64_t* __restrict__ keys, const uint32_t* __restrict__ fin,
                                                      uint64_t n, Digit dig, int pass, uint64_t range_len, uint32_t nr,
                                                      uint32_t* __restrict__ cnt) {
    __shared__ uint32_t h[kWaves][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * range_len;
    const uint64_t e = s + range_len < n ? s + range_len : n;
    constexpr int U = 8;                                    // keys per lane in flight
    for (uint64_t i0 = s; i0 < e; i0 += kThreads * U) {      // block-uniform trip count
        uint64_t kk[U];
        uint32_t dd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kThreads + threadIdx.x;
            if constexpr (CIN) dd[u] = i < e ? fin[i] : 0u;
            else kk[u] = i < e ? keys[i] : 0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) dd[u] = pass_digit<Digit, CIN>(dig, CIN ? 0 : kk[u], dd[u], pass);
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const bool valid = i0 + (uint64_t)u * kThreads + threadIdx.x < e;
            const uint64_t peers = match_digit(dd[u], valid);
            if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&h[w][dd[u]], (uint32_t)__popcll(peers));
        }
    }
    __syncthreads();
    const int d = threadIdx.x;
    cnt[(uint64_t)d * nr + blockIdx.x] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// One workgroup per digit: off[d * nr + r] = keys with digit d in ranges
// before r; tot[d] = keys with digit d.  (Internal linkage: every file that
// instantiates the passes gets its own copy.)
static __global__ __launch_bounds__(kThreads) void k_range_scan(const uint32_t* __restrict__ cnt, uint32_t nr,
                                                                uint32_t* __restrict__ off, uint32_t* __restrict__ tot) {
    __shared__ uint32_t sc[kThreads];
    const uint32_t d = blockIdx.x, t = threadIdx.x;
    const uint32_t per = (nr + kThreads - 1) / kThreads;
    const uint32_t lo = t * per < nr ? t * per : nr, hi = lo + per < nr ? lo + per : nr;
    const uint32_t* c = cnt + (uint64_t)d * nr;
    uint32_t sum = 0;
    for (uint32_t r = lo; r < hi; ++r) sum += c[r];
    sc[t] = sum;
    __syncthreads();
    for (int o = 1; o < kThreads; o <<= 1) {
        uint32_t x = t >= (uint32_t)o ? sc[t - o] : 0;
        __syncthreads();
        sc[t] += x;
        __syncthreads();
    }
    uint32_t run = sc[t] - sum;
    uint32_t* o = off + (uint64_t)d * nr;
    for (uint32_t r = lo; r < hi; ++r) {
        o[r] = run;
        run += c[r];
    }
    if (t == kThreads - 1) tot[d] = sc[t];
}

// Stable scatter of one range's tiles.  Keys, payloads (and with CIN the
// carried digit words) of a tile are loaded together, ranked by digit, placed
// in LDS in digit order and written out so that each digit run is written
// by consecutive lanes.  COUT writes the digit word of every row (carried
// from fin, or dig.bits(key)) next to it.
template <class Digit, bool CIN, bool COUT, bool PAY = true>
__global__ __launch_bounds__(kThreads) void k_downsweep(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                        const uint32_t* __restrict__ fin, uint64_t* __restrict__ kout,
                                                        uint32_t* __restrict__ vout, uint32_t* __restrict__ fout,
                                                        uint64_t n, Digit dig, int pass, uint64_t tiles_per_range,
                                                        uint32_t nr, const uint32_t* __restrict__ off,
                                                        const uint32_t* __restrict__ tot) {
    __shared__ uint64_t kb[kTile];
    __shared__ uint32_t vb[PAY ? kTile : 1];
    __shared__ uint32_t fb[(CIN || COUT) ? kTile : 1];   // digit words of the LDS slots
    __shared__ uint32_t wcnt[kWaves][256];          // per-wave digit counters -> wave prefixes
    __shared__ uint32_t lstart[256];                // tile-local start of each digit
    __shared__ uint32_t run[256];                   // output position of the digit's next key
    __shared__ uint32_t scan_s[256];

    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, d = threadIdx.x;
    {   // digit base: exclusive scan of the digit totals
        const uint32_t t0 = tot[d];
        scan_s[d] = t0;
        __syncthreads();
        for (int o = 1; o < 256; o <<= 1) {
            uint32_t x = d >= o ? scan_s[d - o] : 0;
            __syncthreads();
            scan_s[d] += x;
            __syncthreads();
        }
        run[d] = scan_s[d] - t0 + off[(uint64_t)d * nr + blockIdx.x];
    }
    // rows of a tile: k/v/f registers; the next tile's rows are loaded into
    // them as soon as the current tile sits in LDS, so the loads overlap the
    // scatter of the current tile
    uint64_t k[kRounds];
    uint32_t v[kRounds], f[kRounds];
    auto load_tile = [&](uint64_t tb) {
        const uint64_t b0 = tb + (uint64_t)w * (64 * kRounds);
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint64_t idx = b0 + r * 64 + lane;
            const bool ok = idx < n;
            k[r] = ok ? kin[idx] : 0;
            if constexpr (PAY) v[r] = ok ? vin[idx] : 0;
            if constexpr (CIN) f[r] = ok ? fin[idx] : 0u;
        }
    };
    load_tile((uint64_t)blockIdx.x * tiles_per_range * kTile);
    for (uint64_t t = 0; t < tiles_per_range; ++t) {
        const uint64_t tbase = ((uint64_t)blockIdx.x * tiles_per_range + t) * kTile;
        if (tbase >= n) break;                                          // block-uniform
        for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&wcnt[0][0])[i] = 0;
        const uint64_t base = tbase + (uint64_t)w * (64 * kRounds);
        __syncthreads();                                                // wcnt cleared
        uint32_t rank[kRounds];
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if constexpr (COUT && !CIN) f[r] = (uint32_t)dig.bits(k[r]);
            if constexpr (COUT || CIN) rank[r] = (f[r] >> (8 * pass)) & 0xFFu;
            else rank[r] = dig.digit(k[r], pass);                       // all digits first: loads in flight together
        }
        // rank inside the wave: the lowest lane of each group of equal digits
        // adds the group's size to the wave's counter (LDS atomics of one wave
        // are performed in issue order, so round r + 1 sees round r's adds)
        // and hands the counter's old value to its group
        uint32_t leader[kRounds];
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint64_t idx = base + r * 64 + lane;
            const bool valid = idx < n;
            const uint32_t dd = rank[r];
            const uint64_t peers = match_digit(dd, valid);
            const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
            uint32_t old = 0;
            if (valid && below == 0) old = atomicAdd(&wcnt[w][dd], (uint32_t)__popcll(peers));
            leader[r] = old;
            rank[r] = below | (dd << 16) | (valid ? ((uint32_t)(__ffsll((unsigned long long)peers) - 1) << 24) : 0u);
        }
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t before = (uint32_t)__shfl((int)leader[r], (int)(rank[r] >> 24), 64);
            rank[r] = ((rank[r] & 0xFFu) + before) | (rank[r] & 0x00FF0000u);   // rank < 4096, digit in bits 16..23
        }
        __syncthreads();
        // per digit: wave prefixes, tile count, tile-local digit start
        uint32_t tcount = 0;
#pragma unroll
        for (int ww = 0; ww < kWaves; ++ww) {
            uint32_t c = wcnt[ww][d];
            wcnt[ww][d] = tcount;
            tcount += c;
        }
        {   // exclusive scan of the 256 tile counts: wave scans, then the wave totals
            uint32_t x = tcount;
#pragma unroll
            for (int o = 1; o < 64; o <<= 1) {
                const uint32_t y = __shfl_up(x, o, 64);
                if (lane >= o) x += y;
            }
            if (lane == 63) scan_s[w] = x;
            __syncthreads();
            uint32_t add = 0;
#pragma unroll
            for (int ww = 0; ww < kWaves; ++ww) add += ww < w ? scan_s[ww] : 0u;
            lstart[d] = x - tcount + add;
        }
        __syncthreads();
        // rows into LDS in digit order (stable: wave prefix + rank)
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint64_t idx = base + r * 64 + lane;
            if (idx < n) {
                const uint32_t dd = rank[r] >> 16;
                const uint32_t lp = lstart[dd] + wcnt[w][dd] + (rank[r] & 0xFFFFu);
                if (lp < kTile) {
                    kb[lp] = k[r];
                    if constexpr (PAY) vb[lp] = v[r];
                    if constexpr (CIN || COUT) fb[lp] = f[r];
                }
            }
        }
        if (t + 1 < tiles_per_range && tbase + kTile < n) load_tile(tbase + kTile);
        __syncthreads();
        const uint32_t tn = tbase + kTile <= n ? kTile : (uint32_t)(n - tbase);
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t i = r * kThreads + threadIdx.x;
            if (i < tn) {
                uint32_t dd;
                if constexpr (CIN || COUT) dd = (fb[i] >> (8 * pass)) & 0xFFu;
                else dd = dig.digit(kb[i], pass);
                const uint64_t g = (uint64_t)run[dd] + (i - lstart[dd]);
                if (g < n) {                                  // always true when the counts match
                    kout[g] = kb[i];
                    if constexpr (PAY) vout[g] = vb[i];
                    if constexpr (COUT) fout[g] = fb[i];
                }
            }
        }
        __syncthreads();
        run[d] += tcount;
    }
}

hipError_t launch_scan_hist(const unsigned long long* ghist, int passes, uint64_t n, unsigned long long* offsets,
                            uint32_t* skip, hipStream_t st);

template <class Digit>
hipError_t radix_histogram(const uint64_t* keys, uint64_t n, const Digit& dig, int passes, RadixWork& w,
                           hipStream_t st) {
    hipError_t e = launch_zero32((uint32_t*)w.hist, (uint64_t)passes * 256 * 2, st);
    if (e != hipSuccess) return e;
    if ((e = launch_zero32(w.skip, 8, st)) != hipSuccess) return e;
    unsigned blocks = (unsigned)(device_cu_count() * 4);
    uint64_t need = (n + kThreads * 16 - 1) / (kThreads * 16);
    if (need < blocks) blocks = (unsigned)(need ? need : 1);
    hipLaunchKernelGGL((k_hist<Digit>), dim3(blocks), dim3(kThreads), 0, st, keys, n, dig, passes, w.hist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    return launch_scan_hist(w.hist, passes, n, w.offsets, w.skip, st);
}

// One stable digit pass; needs w.status of at least radix_work_words(n)
// words.  fin: carried digit words of kin's rows (byte `pass` is the digit)
// instead of dig; fout: write the rows' digit words (carried, or
// dig.bits(key)) next to kout.  vin and vout null: keys only (no carried
// digit words then).
template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st, const uint32_t* fin,
                      uint32_t* fout) {
    if (n == 0) return hipSuccess;
    if ((!vin || !vout) && (vin || vout || fin || fout)) return hipErrorInvalidValue;
    const uint64_t per = radix_tiles_per_range(n), nr = radix_ranges(n);
    uint32_t* cnt = w.status;
    uint32_t* off = w.status + nr * 256;
    uint32_t* tot = w.status + nr * 512;
    if (fin) hipLaunchKernelGGL((k_upsweep<Digit, true>), dim3((unsigned)nr), dim3(kThreads), 0, st, kin, fin, n, dig,
                                pass, per * kTile, (uint32_t)nr, cnt);
    else hipLaunchKernelGGL((k_upsweep<Digit, false>), dim3((unsigned)nr), dim3(kThreads), 0, st, kin, fin, n, dig,
                            pass, per * kTile, (uint32_t)nr, cnt);
    hipLaunchKernelGGL(k_range_scan, dim3(256), dim3(kThreads), 0, st, cnt, (uint32_t)nr, off, tot);
    const dim3 g((unsigned)nr), b(kThreads);
    if (!vin || !vout) {
        hipLaunchKernelGGL((k_downsweep<Digit, false, false, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n,
                           dig, pass, per, (uint32_t)nr, (const uint32_t*)off, (const uint32_t*)tot);
    } else if (fin && fout)
        hipLaunchKernelGGL((k_downsweep<Digit, true, true>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig, pass,
                           per, (uint32_t)nr, (const uint32_t*)off, (const uint32_t*)tot);
    else if (fin)
        hipLaunchKernelGGL((k_downsweep<Digit, true, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, per, (uint32_t)nr, (const uint32_t*)off, (const uint32_t*)tot);
    else if (fout)
        hipLaunchKernelGGL((k_downsweep<Digit, false, true>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, per, (uint32_t)nr, (const uint32_t*)off, (const uint32_t*)tot);
    else
        hipLaunchKernelGGL((k_downsweep<Digit, false, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, per, (uint32_t)nr, (const uint32_t*)off, (const uint32_t*)tot);
    return hipGetLastError();
}

}  // namespace vx
