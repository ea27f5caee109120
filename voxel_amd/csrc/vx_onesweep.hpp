// vx_onesweep.hpp -- the digit-pass kernels of vx_radix.hip as templates
// over the digit function, so that other kernel files can instantiate them
// (the learned partitioner's model digit, vx_learned.hip).  See vx_radix.hip
// for the algorithm.
#pragma once
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>

#include "vx_kernels.hpp"
#include "vx_radix.hpp"

namespace vx {

constexpr int kThreads = 256;           // = number of digit values
constexpr int kWaves = kThreads / 64;

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

// Tiles of kTile rows are the unit of every digit pass.  Per pass:
// k_tile_counts writes each tile's 256 digit counts (tile-major, 1 KB per
// tile); k_chunk_sums, k_chunk_scan and k_tile_offsets turn them in place
// into each tile's first output position per digit (digits in order, then
// tiles in order inside a digit: the stable order); k_scatter_tile (one
// workgroup per tile) ranks the tile's rows by digit and writes them through
// LDS so that each digit run is written by consecutive lanes.  No workgroup
// waits on another and no tile depends on another tile's progress, so tiles
// are handed to the eight XCDs in contiguous blocks (workgroups are
// dispatched to the XCDs round-robin): the runs that neighbouring tiles
// write next to each other are then merged in one XCD's L2 before they
// reach HBM.
constexpr int kDsThreads = 512;                 // scatter workgroup
constexpr int kDsWaves = kDsThreads / 64;
constexpr int kRounds = 16;                     // rows per lane per tile
constexpr int kTile = kDsThreads * kRounds;     // 8192 rows per tile
constexpr uint64_t kChunkTiles = 256;           // tiles per chunk of the offset scan
constexpr uint64_t kXcds = 8;

__host__ __device__ inline uint64_t radix_tiles(uint64_t n) { return (n + kTile - 1) / kTile; }
__host__ __device__ inline uint64_t radix_chunks(uint64_t n) { return (radix_tiles(n) + kChunkTiles - 1) / kChunkTiles; }

// Tile of workgroup b: XCD x = b mod 8 takes tiles [x * per, (x + 1) * per)
// in order; the last ntiles mod 8 tiles map to themselves.  A bijection of
// [0, ntiles).
__device__ __forceinline__ uint64_t xcd_tile(uint64_t b, uint64_t ntiles) {
    const uint64_t per = ntiles / kXcds;
    return b < per * kXcds ? (b % kXcds) * per + b / kXcds : b;
}

// Digit source of a pass: the key through the digit function, or (CIN) a
// carried 32-bit digit word whose byte `pass` is the digit.
template <class Digit, bool CIN>
__device__ __forceinline__ uint32_t pass_digit(const Digit& dig, uint64_t key, uint32_t word, int pass) {
    if constexpr (CIN) return (word >> (8 * pass)) & 0xFFu;
    else return dig.digit(key, pass);
}

// Adds the wave's rows (digit d where valid) to the wave's counters hw.
// ROUNDS times, the rows sharing the digit of the first row not yet counted
// are added with one LDS atomic; the rows left over add one each.  Sorted
// or nearly sorted keys put a wave's rows on one to three digits, where
// per-row atomics on one counter would serialise.  Called by all lanes of
// the wave together.  (Three rounds cut the carried-id counts of 95 %-sorted
// keys from 1.07 to 0.80 ms at 2^29; the table-driven pass-0 counts, which
// are bound by their own instructions, run fastest with one.)
template <int ROUNDS>
__device__ __forceinline__ void count_digits(uint32_t* hw, uint32_t d, bool valid) {
    const int lane = (int)(threadIdx.x & 63);
    uint64_t rem = __ballot(valid);
#pragma unroll
    for (int k = 0; k < ROUNDS; ++k) {
        if (!rem) break;                                             // wave-uniform
        const int l = __ffsll((unsigned long long)rem) - 1;
        const uint32_t dk = (uint32_t)__builtin_amdgcn_readlane((int)d, l);
        const uint64_t m = __ballot(valid && d == dk) & rem;
        if (lane == l) atomicAdd(&hw[dk], (uint32_t)__popcll(m));
        rem &= ~m;
    }
    if ((rem >> lane) & 1) atomicAdd(&hw[d], 1u);
}

// cnt[t * 256 + d] = rows of tile t with digit d.  FOUT (keys through the
// digit function only): also fout[i] = dig.bits(key i), the row's digit
// word, so the scatter can read it instead of evaluating the function again.
template <class Digit, bool CIN, bool FOUT = false>
__global__ __launch_bounds__(kThreads) void k_tile_counts(const uint64_t* __restrict__ keys,
                                                          const uint32_t* __restrict__ fin, uint64_t n, Digit dig,
                                                          int pass, uint32_t* __restrict__ cnt,
                                                          uint32_t* __restrict__ fout = nullptr) {
    static_assert(!(CIN && FOUT), "a carried digit word is not written again");
    __shared__ uint32_t h[kWaves][256];
    const int w = threadIdx.x >> 6;
    for (int i = threadIdx.x; i < kWaves * 256; i += kThreads) (&h[0][0])[i] = 0;
    __syncthreads();
    const uint64_t s = (uint64_t)blockIdx.x * kTile;
    const uint64_t e = s + kTile < n ? s + kTile : n;
    constexpr int U = 8;                                    // rows per lane in flight
    // rows i0 + u * kThreads + tid; ALL: every one of them is below e (no
    // per-row bounds test)
    auto step = [&](uint64_t i0, auto all_c) {
        constexpr bool ALL = decltype(all_c)::value;
        auto in = [&](int u) { return ALL || i0 + (uint64_t)u * kThreads + threadIdx.x < e; };
        uint64_t kk[U];
        uint32_t dd[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = i0 + (uint64_t)u * kThreads + threadIdx.x;
            if constexpr (CIN) dd[u] = in(u) ? fin[i] : 0u;
            else kk[u] = in(u) ? keys[i] : 0;
        }
        if constexpr (FOUT) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const uint32_t f = (uint32_t)dig.bits(kk[u]);
                if (in(u)) fout[i0 + (uint64_t)u * kThreads + threadIdx.x] = f;
                dd[u] = (f >> (8 * pass)) & 0xFFu;
            }
        } else {
#pragma unroll
            for (int u = 0; u < U; ++u) dd[u] = pass_digit<Digit, CIN>(dig, CIN ? 0 : kk[u], dd[u], pass);
        }
        // counting needs no ranks: one LDS atomic per row, or per wave when
        // its rows share the digit
#pragma unroll
        for (int u = 0; u < U; ++u) count_digits<3>(h[w], dd[u], in(u));
    };
    uint64_t i0 = s;
    for (; i0 + (uint64_t)kThreads * U <= e; i0 += kThreads * U) step(i0, std::true_type{});   // block-uniform
    if (i0 < e) step(i0, std::false_type{});
    __syncthreads();
    const int d = threadIdx.x;
    cnt[(uint64_t)blockIdx.x * 256 + d] = h[0][d] + h[1][d] + h[2][d] + h[3][d];
}

// k_tile_counts<TableDigit, false, true> with the cell table in LDS, for
// tables of at most kLdsCells cells: one persistent workgroup of
// kTcThreads threads per CU loads the table once and takes tiles blockIdx.x,
// blockIdx.x + gridDim.x, ... (eight rows per thread per tile), issuing the
// next tile's key loads before it evaluates the current one.  Uniform keys
// read the table at random cells: from LDS that is one ds_read per row
// instead of one L2 line request per row.  Same outputs as k_tile_counts.
constexpr uint32_t kLdsCells = 16384;
constexpr int kTcThreads = 1024;
static_assert(kTile == 8 * kTcThreads, "eight rows per thread per tile");
static __global__ __launch_bounds__(kTcThreads) void k_table_counts(const uint64_t* __restrict__ keys, uint64_t n,
                                                             TableDigit dig, int pass, uint32_t* __restrict__ cnt,
                                                             uint32_t* __restrict__ fout, uint64_t ntiles) {
    constexpr int U = 8, NW = kTcThreads / 64;
    __shared__ uint2 tab[kLdsCells];
    __shared__ uint32_t h[NW][256];
    const int w = threadIdx.x >> 6;
    const uint32_t nc = dig.ncells < kLdsCells ? dig.ncells : kLdsCells;     // the host checks ncells
    for (uint32_t i = threadIdx.x; i < nc; i += kTcThreads) tab[i] = dig.cell[i];
    for (int i = threadIdx.x; i < NW * 256; i += kTcThreads) (&h[0][0])[i] = 0;
    auto load = [&](uint64_t t, uint64_t (&kk)[U]) {
        const uint64_t s = t * kTile;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = s + (uint64_t)u * kTcThreads + threadIdx.x;
            kk[u] = i < n ? keys[i] : 0;
        }
    };
    uint64_t kk[U];
    if (blockIdx.x < ntiles) load(blockIdx.x, kk);
    __syncthreads();                                                // table and counters ready
    for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint64_t nx[U];
        if (t + gridDim.x < ntiles) load(t + gridDim.x, nx);          // block-uniform
        const uint64_t s = t * kTile;
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = s + (uint64_t)u * kTcThreads + threadIdx.x;
            const uint32_t f = dig.fine_in(tab, kk[u]);
            if (i < n) fout[i] = f;
            count_digits<1>(h[w], (f >> (8 * pass)) & 0xFFu, i < n);
        }
        __syncthreads();
        if (threadIdx.x < 256) {
            const int d = threadIdx.x;
            uint32_t c = 0;
#pragma unroll
            for (int ww = 0; ww < NW; ++ww) { c += h[ww][d]; h[ww][d] = 0; }
            cnt[t * 256 + d] = c;
        }
        __syncthreads();                                            // counters cleared
#pragma unroll
        for (int u = 0; u < U; ++u) kk[u] = nx[u];
    }
}

// csum[j * 256 + d] = rows with digit d in the tiles of chunk j.  (Internal
// linkage for the scan kernels: every file that instantiates the passes gets
// its own copy.)
static __global__ __launch_bounds__(256) void k_chunk_sums(const uint32_t* __restrict__ cnt, uint64_t ntiles,
                                                           uint32_t* __restrict__ csum) {
    const uint32_t d = threadIdx.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * kChunkTiles;
    const uint64_t t1 = t0 + kChunkTiles < ntiles ? t0 + kChunkTiles : ntiles;
    uint32_t s = 0;
    uint64_t t = t0;
    for (; t + 8 <= t1; t += 8) {
        uint32_t c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = cnt[(t + u) * 256 + d];
#pragma unroll
        for (int u = 0; u < 8; ++u) s += c[u];
    }
    for (; t < t1; ++t) s += cnt[t * 256 + d];
    csum[(uint64_t)blockIdx.x * 256 + d] = s;
}

// One workgroup: csum[j * 256 + d] becomes the rows with digit d in chunks
// before j; dbase[d] = rows with a digit below d.
static __global__ __launch_bounds__(256) void k_chunk_scan(uint32_t* __restrict__ csum, uint64_t nch,
                                                           uint32_t* __restrict__ dbase) {
    __shared__ uint32_t ws[4];
    const uint32_t d = threadIdx.x, lane = d & 63, w = d >> 6;
    uint32_t run = 0;
    uint64_t j = 0;
    for (; j + 8 <= nch; j += 8) {
        uint32_t c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = csum[(j + u) * 256 + d];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            csum[(j + u) * 256 + d] = run;
            run += c[u];
        }
    }
    for (; j < nch; ++j) {
        const uint32_t c = csum[j * 256 + d];
        csum[j * 256 + d] = run;
        run += c;
    }
    uint32_t x = run;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t y = __shfl_up(x, o, 64);
        if (lane >= (uint32_t)o) x += y;
    }
    if (lane == 63) ws[w] = x;
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (uint32_t ww = 0; ww < 4; ++ww) add += ww < w ? ws[ww] : 0u;
    dbase[d] = x - run + add;
}

// cnt[t * 256 + d] becomes tile t's first output position for digit d.
static __global__ __launch_bounds__(256) void k_tile_offsets(uint32_t* __restrict__ cnt, uint64_t ntiles,
                                                             const uint32_t* __restrict__ csum,
                                                             const uint32_t* __restrict__ dbase) {
    const uint32_t d = threadIdx.x;
    const uint64_t t0 = (uint64_t)blockIdx.x * kChunkTiles;
    const uint64_t t1 = t0 + kChunkTiles < ntiles ? t0 + kChunkTiles : ntiles;
    uint32_t run = csum[(uint64_t)blockIdx.x * 256 + d] + dbase[d];
    uint64_t t = t0;
    for (; t + 8 <= t1; t += 8) {
        uint32_t c[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) c[u] = cnt[(t + u) * 256 + d];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            cnt[(t + u) * 256 + d] = run;
            run += c[u];
        }
    }
    for (; t < t1; ++t) {
        const uint32_t c = cnt[t * 256 + d];
        cnt[t * 256 + d] = run;
        run += c;
    }
}

// LDS of one scatter workgroup (75.8 KB: two workgroups per CU).
struct ScatterLds {
    uint64_t kb[kTile];                 // keys (load order, then digit order); u32 view for payloads / digit words
    union {
        uint32_t wcnt[kDsWaves][256];   // per-wave digit counters -> wave prefixes
        uint8_t sdig[kTile];            // then: digit of each LDS slot
    };
    uint32_t lstart[256];               // tile-local start of each digit
    uint32_t base[256];                 // output row of the tile's first row of each digit, minus its slot
    uint32_t scan_s[4];
};

// Stable scatter of one tile (FULL: all kTile rows present, no per-row
// checks).  Rows are loaded and their digits taken (from the key, or the
// carried digit word) and the rows are ranked by digit (per-wave LDS words
// of the lanes holding each digit, per-wave LDS counters), which gives every
// row its LDS slot in digit order and every slot its digit and output row.  Then each array in turn (digit words,
// keys, payloads) is placed in LDS in digit order and written out from
// there, so that consecutive lanes write consecutive rows of a digit run.
// One 64 KB LDS buffer serves the three arrays.  COUT writes the digit word
// of every row (carried from fin, or dig.bits(key)) next to it.  Output rows
// are clamped to n - 1, so that inconsistent counts could only misplace
// rows, never write outside the arrays.
template <class Digit, bool CIN, bool COUT, bool PAY, bool IDX, bool FULL>
__device__ __forceinline__ void scatter_tile(ScatterLds& L, const uint64_t* __restrict__ kin,
                                             const uint32_t* __restrict__ vin, const uint32_t* __restrict__ fin,
                                             uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                             uint32_t* __restrict__ fout, uint64_t n, const Digit& dig, int pass,
                                             uint64_t tile, const uint32_t* __restrict__ off) {
    uint32_t* const ub = reinterpret_cast<uint32_t*>(L.kb);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const uint64_t tbase = tile * kTile;
    const uint32_t tn = FULL ? (uint32_t)kTile : (uint32_t)(n - tbase);
    const uint32_t last = (uint32_t)(n - 1);                    // n <= 2^32
    // wave w reads rows [b0, b0 + 64 kRounds) of the tile: round r, lane l ->
    // row b0 + 64 r + l, present while 64 r + l < rem
    const uint64_t b0 = tbase + (uint64_t)w * (64 * kRounds);
    const int rem = FULL ? 64 * kRounds
                         : (b0 < n ? (n - b0 < (uint64_t)(64 * kRounds) ? (int)(n - b0) : 64 * kRounds) : 0);
    auto present = [&](int r) { return FULL || r * 64 + lane < rem; };
    uint64_t k[kRounds];
    uint32_t v[kRounds], f[kRounds];
    {
        const uint64_t* const kw = kin + b0 + lane;
        const uint32_t* const fw = fin + b0 + lane;
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            if constexpr (FULL) {
                k[r] = kw[r * 64];
                if constexpr (CIN) f[r] = fw[r * 64];
            } else {
                k[r] = present(r) ? kw[r * 64] : 0;
                if constexpr (CIN) f[r] = present(r) ? fw[r * 64] : 0u;
            }
        }
    }
    for (int i = tid; i < kDsWaves * 256; i += kDsThreads) (&L.wcnt[0][0])[i] = 0;
    if (tid < 256) L.base[tid] = off[tile * 256 + tid];
    uint32_t rank[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        if constexpr (COUT && !CIN) f[r] = (uint32_t)dig.bits(k[r]);
        if constexpr (COUT || CIN) rank[r] = (f[r] >> (8 * pass)) & 0xFFu;
        else rank[r] = dig.digit(k[r], pass);
    }
    // Lanes of a wave with equal digits come from one of two matchers.
    // With carried digit words (16 more live registers) the rows' keys stay
    // in registers and each lane ORs its lane bit into the wave's 64-bit LDS
    // word of its digit, reads the word back and clears it (LDS operations
    // of one wave are performed in issue order); the words sit in the key
    // buffer, unused until the rows are placed.  Without them the keys wait
    // in the key buffer in load order and the digit bits are matched with
    // eight wave ballots: at the same occupancy this leaves the LDS to the
    // other workgroup's scatter (measured about 4 % faster for plain key +
    // payload passes, while the ballots spill registers once digit words
    // are carried).
    constexpr bool kLdsMatch = CIN || COUT;
    const int nat = w * (64 * kRounds) + lane;           // load-order slot of row (w, 0, lane)
    uint64_t* const peer_w = L.kb + w * 256;
    if constexpr (kLdsMatch) {
        for (int i = tid; i < kDsWaves * 256; i += kDsThreads) L.kb[i] = 0;
    } else {
#pragma unroll
        for (int r = 0; r < kRounds; ++r) L.kb[nat + r * 64] = k[r];
    }
    __syncthreads();                                                // wcnt (and peer words) cleared
    // rank inside the wave: the lowest lane of each group of equal digits
    // adds the group's size to the wave's counter (LDS atomics of one wave
    // are performed in issue order, so round r + 1 sees round r's adds) and
    // hands the counter's old value to its group
    uint32_t leader[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const bool valid = present(r);
        const uint32_t dd = rank[r];
        uint64_t peers = 0;
        if constexpr (kLdsMatch) {
            // lanes with lane 0's digit find their peers by one ballot; the
            // others through the digit's LDS word (no lane of the first kind
            // touches it)
            const uint32_t d0 = __builtin_amdgcn_readfirstlane(dd);
            const uint64_t same = __ballot(valid && dd == d0);
            if (valid && dd == d0) {
                peers = same;
            } else if (valid) {
                atomicOr((unsigned long long*)&peer_w[dd], 1ull << lane);
                peers = peer_w[dd];
                peer_w[dd] = 0;
            }
        } else {
            peers = match_digit(dd, valid);
        }
        const uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
        uint32_t old = 0;
        if (valid && below == 0) old = atomicAdd(&L.wcnt[w][dd], (uint32_t)__popcll(peers));
        const int ldr = valid ? __ffsll((unsigned long long)peers) - 1 : 0;
        if constexpr (kLdsMatch) {
            // one round at a time: fewer live registers
            const uint32_t before = (uint32_t)__shfl((int)old, ldr, 64);
            rank[r] = (below + before) | (dd << 16);    // rank < 1024, digit in bits 16..23
        } else {
            // the counters' old values are gathered after all rounds' atomics
            leader[r] = old;
            rank[r] = below | (dd << 16) | ((uint32_t)ldr << 24);
        }
    }
    if constexpr (!kLdsMatch) {
#pragma unroll
        for (int r = 0; r < kRounds; ++r) {
            const uint32_t before = (uint32_t)__shfl((int)leader[r], (int)(rank[r] >> 24), 64);
            rank[r] = ((rank[r] & 0xFFu) + before) | (rank[r] & 0x00FF0000u);
        }
    }
    __syncthreads();
    // per digit (threads 0..255): wave prefixes, tile count, tile-local start
    uint32_t tcount = 0, x = 0;
    if (tid < 256) {
#pragma unroll
        for (int ww = 0; ww < kDsWaves; ++ww) {
            const uint32_t c = L.wcnt[ww][tid];
            L.wcnt[ww][tid] = tcount;
            tcount += c;
        }
        x = tcount;
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t y = __shfl_up(x, o, 64);
            if (lane >= o) x += y;
        }
        if (lane == 63) L.scan_s[w] = x;
    }
    __syncthreads();
    if (tid < 256) {
        uint32_t add = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) add += ww < w ? L.scan_s[ww] : 0u;
        const uint32_t ls = x - tcount + add;
        L.lstart[tid] = ls;
        L.base[tid] -= ls;                          // slot i of digit d goes to row base[d] + i (mod 2^32)
    }
    __syncthreads();
    // LDS slot of each row (stable: digit start + wave prefix + rank in the
    // wave); rows past the end get slot kTile and are not placed
    uint32_t lp[kRounds];
#pragma unroll
    for (int r = 0; r < kRounds; ++r) {
        const uint32_t dd = rank[r] >> 16;
        lp[r] = present(r) ? L.lstart[dd] + L.wcnt[w][dd] + (rank[r] & 0xFFFFu) : (uint32_t)kTile;
    }
    auto place = [&](int r) { return FULL || lp[r] < (uint32_t)kTile; };
    auto slot_ok = [&](int r) { return FULL || (uint32_t)(r * kDsThreads + tid) < tn; };
    __syncthreads();                                                // wcnt read: its space becomes sdig; peer words dead
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (place(r)) L.sdig[lp[r]] = (uint8_t)(rank[r] >> 16);
    if constexpr (!kLdsMatch) {
#pragma unroll
        for (int r = 0; r < kRounds; ++r) k[r] = L.kb[nat + r * 64];
    }
    auto load_payload = [&]() {
        if constexpr (PAY && IDX) {
#pragma unroll
            for (int r = 0; r < kRounds; ++r) v[r] = (uint32_t)(b0 + lane + r * 64);   // the row index
        } else if constexpr (PAY) {
            const uint32_t* const vw = vin + b0 + lane;
#pragma unroll
            for (int r = 0; r < kRounds; ++r) {
                if constexpr (FULL) v[r] = vw[r * 64];
                else v[r] = present(r) ? vw[r * 64] : 0u;
            }
        }
    };
    // payloads are loaded only once the digit words are out (fewer live
    // registers); they arrive while the keys go out
    if constexpr (!COUT) load_payload();
    __syncthreads();                                                // sdig written, load-order keys read
    // output row of slot i = r * kDsThreads + tid (looked up in each phase:
    // fewer live registers than keeping it)
    auto g = [&](int r) {
        const uint32_t i = r * kDsThreads + tid;
        return min(L.base[L.sdig[i]] + i, last);
    };
    if constexpr (COUT) {
#pragma unroll
        for (int r = 0; r < kRounds; ++r)
            if (place(r)) ub[lp[r]] = f[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r)
            if (slot_ok(r)) fout[g(r)] = ub[r * kDsThreads + tid];
        load_payload();
        __syncthreads();                                            // digit words read
    }
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (place(r)) L.kb[lp[r]] = k[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < kRounds; ++r)
        if (slot_ok(r)) kout[g(r)] = L.kb[r * kDsThreads + tid];
    if constexpr (PAY) {
        __syncthreads();                                            // keys read
#pragma unroll
        for (int r = 0; r < kRounds; ++r)
            if (place(r)) ub[lp[r]] = v[r];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < kRounds; ++r)
            if (slot_ok(r)) vout[g(r)] = ub[r * kDsThreads + tid];
    }
}

template <class Digit, bool CIN, bool COUT, bool PAY = true, bool IDX = false>
__global__ __launch_bounds__(kDsThreads, 4) void k_scatter_tile(const uint64_t* __restrict__ kin,
                                                                const uint32_t* __restrict__ vin,
                                                                const uint32_t* __restrict__ fin,
                                                                uint64_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout,
                                                                uint32_t* __restrict__ fout, uint64_t n, Digit dig,
                                                                int pass, uint64_t ntiles,
                                                                const uint32_t* __restrict__ off) {
    __shared__ ScatterLds L;
    const uint64_t tile = xcd_tile(blockIdx.x, ntiles);
    if ((tile + 1) * kTile <= n)
        scatter_tile<Digit, CIN, COUT, PAY, IDX, true>(L, kin, vin, fin, kout, vout, fout, n, dig, pass, tile, off);
    else
        scatter_tile<Digit, CIN, COUT, PAY, IDX, false>(L, kin, vin, fin, kout, vout, fout, n, dig, pass, tile, off);
}

// skip[p] = 1 when every key has the same byte p (a stable pass over that
// digit is the identity permutation); w.hist holds the per-workgroup
// partials.  All eight flags are set for n = 0.
hipError_t launch_constant_bytes(const uint64_t* keys, uint64_t n, RadixWork& w, hipStream_t st);

// One stable digit pass; needs w.status of at least radix_work_words(n)
// words.  fin: carried digit words of kin's rows (byte `pass` is the digit)
// instead of dig; fout: write the rows' digit words (carried, or
// dig.bits(key)) next to kout.  vin and vout null: keys only (no carried
// digit words then).  row_index: the payload of row i is i (vin unused,
// vout required, no fin).  fpre (with fout, without fin): n words the
// counting kernel fills with dig.bits(key) in kin's row order, which the
// scatter then reads as carried words (the digit function runs once per
// row instead of twice).
template <class Digit>
hipError_t radix_pass(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                      const Digit& dig, int pass, RadixWork& w, hipStream_t st, const uint32_t* fin,
                      uint32_t* fout, bool row_index, uint32_t* fpre) {
    if (n == 0) return hipSuccess;
    if (row_index ? (!vout || fin) : ((!vin || !vout) && (vin || vout || fin || fout))) return hipErrorInvalidValue;
    if (fpre && (fin || !fout)) return hipErrorInvalidValue;
    const uint64_t nt = radix_tiles(n), nch = radix_chunks(n);
    if (nt > 0xFFFFFFFFull || w.status_cap < nt * 256 + nch * 256 + 256) return hipErrorInvalidValue;
    uint32_t* cnt = w.status;
    uint32_t* csum = w.status + nt * 256;
    uint32_t* dbase = csum + nch * 256;
    if (fin) hipLaunchKernelGGL((k_tile_counts<Digit, true>), dim3((unsigned)nt), dim3(kThreads), 0, st, kin, fin, n,
                                dig, pass, cnt, nullptr);
    else if (fpre) {
        bool lds = false;
        if constexpr (std::is_same<Digit, TableDigit>::value) {
            if (dig.ncells <= kLdsCells) {
                const uint64_t cus = (uint64_t)device_cu_count();
                hipLaunchKernelGGL(k_table_counts, dim3((unsigned)(nt < cus ? nt : cus)), dim3(kTcThreads), 0, st, kin,
                                   n, dig, pass, cnt, fpre, nt);
                lds = true;
            }
        }
        if (!lds)
            hipLaunchKernelGGL((k_tile_counts<Digit, false, true>), dim3((unsigned)nt), dim3(kThreads), 0, st, kin, fin,
                               n, dig, pass, cnt, fpre);
    }
    else hipLaunchKernelGGL((k_tile_counts<Digit, false>), dim3((unsigned)nt), dim3(kThreads), 0, st, kin, fin, n, dig,
                            pass, cnt, nullptr);
    if (fpre) fin = fpre;
    hipLaunchKernelGGL(k_chunk_sums, dim3((unsigned)nch), dim3(256), 0, st, (const uint32_t*)cnt, nt, csum);
    hipLaunchKernelGGL(k_chunk_scan, dim3(1), dim3(256), 0, st, csum, nch, dbase);
    hipLaunchKernelGGL(k_tile_offsets, dim3((unsigned)nch), dim3(256), 0, st, cnt, nt, (const uint32_t*)csum,
                       (const uint32_t*)dbase);
    const dim3 g((unsigned)nt), b(kDsThreads);
    const uint32_t* off = cnt;
    if (row_index && fin)
        hipLaunchKernelGGL((k_scatter_tile<Digit, true, true, true, true>), g, b, 0, st, kin, vin, fin, kout, vout,
                           fout, n, dig, pass, nt, off);
    else if (row_index && fout)
        hipLaunchKernelGGL((k_scatter_tile<Digit, false, true, true, true>), g, b, 0, st, kin, vin, fin, kout, vout,
                           fout, n, dig, pass, nt, off);
    else if (row_index)
        hipLaunchKernelGGL((k_scatter_tile<Digit, false, false, true, true>), g, b, 0, st, kin, vin, fin, kout, vout,
                           fout, n, dig, pass, nt, off);
    else if (!vin || !vout)
        hipLaunchKernelGGL((k_scatter_tile<Digit, false, false, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout,
                           n, dig, pass, nt, off);
    else if (fin && fout)
        hipLaunchKernelGGL((k_scatter_tile<Digit, true, true>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, nt, off);
    else if (fin)
        hipLaunchKernelGGL((k_scatter_tile<Digit, true, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, nt, off);
    else if (fout)
        hipLaunchKernelGGL((k_scatter_tile<Digit, false, true>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, nt, off);
    else
        hipLaunchKernelGGL((k_scatter_tile<Digit, false, false>), g, b, 0, st, kin, vin, fin, kout, vout, fout, n, dig,
                           pass, nt, off);
    return hipGetLastError();
}

}  // namespace vx
