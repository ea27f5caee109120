// vx_join.hip -- i64 hash join build/probe on gfx950, replacing
// HashJoin<TKey>::Build/Probe (ops/join.hpp:36-80, 150-165) and the
// join + SUM(payload) harness of the configs.
//
// Table: open addressing with linear probing over 16-byte slots
// {key, payload}, slots = nextpow2(max(2 * build_rows, 64)) (the reference's
// Init sizing, join.hpp:37), hash = HashValue<i64>(key) (the splitmix64
// finaliser, ops/hash.hpp:328-333).  Empty slots hold INT64_MIN; a build key
// equal to INT64_MIN is kept in a side record.  The reference's chains are
// LIFO, so its Probe returns, for duplicate build keys, the highest build row
// id: the build keeps max(row) per key with atomicMax, then fills each
// slot's payload from that row.
//
// Partitioned layout: the slots are 2^pbits sub-tables of 2^sbits slots; a
// key lives in sub-table part(key) at (hash & (2^sbits - 1)) and probing
// wraps inside its sub-table.  part() is 0 (one table), the top pbits of the
// hash, or a learned range partition: the CDF model of the learned sort
// (vx_learned.hip) fitted to the build keys maps a key to one of 2^pbits
// buckets, non-decreasing in the key.  A large probe batch is first
// partitioned by part() with one stable digit pass (vx_onesweep.hpp, keys
// only), then probed in partition order, so that the probes of one sub-table
// (2 MB) run together.  A sub-table that overflows (a skewed range
// partition) rebuilds the table as one table.
//
// Measured on MI355X (2^24 build keys, 2^28 Zipf(1.0) probes, 90 % hits):
// one table 4.2 ms per probe batch, 83 B of HBM traffic per probe (one
// 128-byte line per L2 miss, 0.59 misses per probe); hash sub-tables 2.9 ms
// for the probe partition pass + 3.1 ms for the probe in partition order.
// Mode 0 therefore builds one table; the partitioned layouts are kept for
// callers that partition anyway (the learned range partition gives
// key-ordered sub-tables, e.g. as splitters for a multi-GPU join).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_compact.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"
#include "vx_onesweep.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {

struct __attribute__((aligned(16))) JoinSlot {
    int64_t key;
    int64_t payload;
};

constexpr int64_t kEmpty = INT64_MIN;

__host__ __device__ __forceinline__ uint64_t join_hash(int64_t key) {
    uint64_t x = (uint64_t)key;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

bool learned_partitioner(const uint64_t* kin, uint64_t n, uint64_t flip, int model, uint32_t nfine, uint2* cell_out,
                         TableDigit* dig, voxel_gpu_learned_stats* stats, StreamCtx* c);

enum JoinMode { kOneTable = 1, kHashParts = 2, kLearnedParts = 3, kEqualWidthParts = 4 };

// Where a key lives (see the file comment).
struct JoinLayout {
    int mode = kOneTable;
    int pbits = 0, sbits = 0;
    TableDigit t{};                      // range modes: bucket of the key's order bits
    __device__ __forceinline__ uint32_t part(int64_t k, uint64_t h) const {
        if (mode == kHashParts) return (uint32_t)(h >> (64 - pbits));
        if (mode == kLearnedParts || mode == kEqualWidthParts) return t.fine((uint64_t)k ^ 0x8000000000000000ull);
        return 0;
    }
    __device__ __forceinline__ uint64_t smask() const { return (1ull << sbits) - 1; }
    __device__ __forceinline__ uint64_t base(int64_t k, uint64_t h) const { return (uint64_t)part(k, h) << sbits; }
};

// Radix digit of the probe partition pass: the key's sub-table.
struct JoinDigit {
    JoinLayout L;
    __device__ __forceinline__ uint32_t digit(uint64_t k, int) const { return L.part((int64_t)k, join_hash((int64_t)k)); }
    __device__ __forceinline__ uint64_t bits(uint64_t k) const { return digit(k, 0); }
    __device__ __forceinline__ uint32_t digit_of_bits(uint64_t b, int) const { return (uint32_t)b & 0xFFu; }
};

namespace {

__global__ __launch_bounds__(256) void k_join_clear(JoinSlot* __restrict__ slots, uint32_t* __restrict__ rows,
                                                    uint64_t n, unsigned long long* __restrict__ side) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        slots[i].key = kEmpty;
        slots[i].payload = 0;
        rows[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 8) side[threadIdx.x] = 0;
}

// side[0] = 1 + highest row holding key INT64_MIN (0: none), side[1] = its
// payload, side[4] = keys that found their sub-table full
__global__ __launch_bounds__(256) void k_join_insert(const int64_t* __restrict__ keys, uint64_t nb,
                                                     JoinSlot* __restrict__ slots, uint32_t* __restrict__ rows,
                                                     JoinLayout L, unsigned long long* __restrict__ side) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    const uint64_t smask = L.smask();
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
        int64_t k = keys[i];
        if (k == kEmpty) { atomicMax(&side[0], (unsigned long long)(i + 1)); continue; }
        const uint64_t h = join_hash(k), base = L.base(k, h);
        uint64_t j = h & smask;
        bool placed = false;
        for (uint64_t step = 0; step <= smask; ++step) {
            unsigned long long prev = atomicCAS((unsigned long long*)&slots[base | j].key,
                                                (unsigned long long)kEmpty, (unsigned long long)k);
            if ((int64_t)prev == kEmpty || (int64_t)prev == k) {
                atomicMax(&rows[base | j], (uint32_t)i);
                placed = true;
                break;
            }
            j = (j + 1) & smask;
        }
        if (!placed) atomicAdd(&side[4], 1ull);
    }
}

__global__ __launch_bounds__(256) void k_join_fill(JoinSlot* __restrict__ slots, const uint32_t* __restrict__ rows,
                                                   uint64_t n, const int64_t* __restrict__ payload,
                                                   unsigned long long* __restrict__ side) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if (slots[i].key != kEmpty) slots[i].payload = payload ? payload[rows[i]] : (int64_t)rows[i];
    if (blockIdx.x == 0 && threadIdx.x == 0 && side[0]) side[1] = payload ? (unsigned long long)payload[side[0] - 1] : side[0] - 1;
}

// Lookup of one probe key: returns true and the slot index on a match.
// At most a sub-table's worth of slots is visited.
__device__ __forceinline__ bool join_find(const JoinSlot* __restrict__ slots, const JoinLayout& L, int64_t k,
                                          uint64_t& at) {
    const uint64_t h = join_hash(k), base = L.base(k, h), smask = L.smask();
    uint64_t j = h & smask;
    for (uint64_t step = 0; step <= smask; ++step) {
        JoinSlot s = slots[base | j];
        if (s.key == k) { at = base | j; return true; }
        if (s.key == kEmpty) return false;
        j = (j + 1) & smask;
    }
    return false;
}

// One 16-byte load of a slot (key and payload from one request).
__device__ __forceinline__ JoinSlot load_slot(const JoinSlot* __restrict__ slots, uint64_t at) {
    const uint4 q = *reinterpret_cast<const uint4*>(slots + at);
    JoinSlot r;
    r.key = (int64_t)(((uint64_t)q.y << 32) | q.x);
    r.payload = (int64_t)(((uint64_t)q.w << 32) | q.z);
    return r;
}

// Probes in input order, P per lane per iteration: the keys are loaded
// non-temporally (the probe stream is read once and should not evict table
// lines), then the first slot of each probe, before any is resolved; the
// rare continuation of a probe sequence is walked one slot at a time.  A
// lane visits at most a sub-table's worth of slots per probe.  P = 1: on
// the Zipf(1.0) benchmark probes, two or four probes in flight per lane ran
// 6-9 % slower (more table lines in flight evict the hot ones from L2).
__global__ __launch_bounds__(256) void k_join_probe_sum(const int64_t* __restrict__ probe, uint64_t np,
                                                        const JoinSlot* __restrict__ slots, JoinLayout L,
                                                        const unsigned long long* __restrict__ side,
                                                        unsigned long long* __restrict__ out) {
    constexpr int P = 1;
    uint64_t cnt = 0, sum = 0;
    const unsigned long long side_row = side[0], side_pay = side[1];
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x, smask = L.smask();
    for (uint64_t i0 = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i0 < np; i0 += P * stride) {
        int64_t k[P];
        bool v[P];
        uint64_t base[P], j[P];
        JoinSlot x[P];
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const uint64_t i = i0 + (uint64_t)p * stride;
            v[p] = i < np;
            k[p] = v[p] ? __builtin_nontemporal_load(probe + i) : kEmpty;
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            const uint64_t h = join_hash(k[p]);
            base[p] = L.base(k[p], h);
            j[p] = h & smask;
            x[p] = JoinSlot{kEmpty, 0};
            if (v[p] && k[p] != kEmpty) x[p] = load_slot(slots, base[p] | j[p]);
        }
#pragma unroll
        for (int p = 0; p < P; ++p) {
            if (!v[p]) continue;
            if (k[p] == kEmpty) {
                if (side_row) { ++cnt; sum += side_pay; }
                continue;
            }
            // the first slot's payload is used whether or not it matches, so the
            // compiler keeps one 16-byte load instead of a key load and a reload
            const bool m = x[p].key == k[p];
            cnt += m ? 1 : 0;
            sum += m ? (uint64_t)x[p].payload : 0;
            if (m || x[p].key == kEmpty) continue;
            uint64_t jj = j[p];
            for (uint64_t step = 1; step <= smask; ++step) {
                jj = (jj + 1) & smask;
                const JoinSlot sl = load_slot(slots, base[p] | jj);
                if (sl.key == k[p]) { ++cnt; sum += (uint64_t)sl.payload; break; }
                if (sl.key == kEmpty) break;
            }
        }
    }
    cnt = block_sum_256(cnt);
    __syncthreads();                                // the partial slots are reused by the second sum
    sum = block_sum_256(sum);
    if (threadIdx.x == 0) {
        if (cnt) atomicAdd(&out[0], (unsigned long long)cnt);
        if (sum) atomicAdd(&out[1], (unsigned long long)sum);
    }
}

// match[i] = build row of the first match of probe i, or 0xFFFFFFFF
__global__ __launch_bounds__(256) void k_join_match(const int64_t* __restrict__ probe, uint64_t np,
                                                    const JoinSlot* __restrict__ slots, const uint32_t* __restrict__ rows,
                                                    JoinLayout L, const unsigned long long* __restrict__ side,
                                                    uint32_t* __restrict__ match) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
        int64_t k = probe[i];
        uint32_t r = 0xFFFFFFFFu;
        if (k == kEmpty) {
            if (side[0]) r = (uint32_t)(side[0] - 1);
        } else {
            uint64_t at;
            if (join_find(slots, L, k, at)) r = rows[at];
        }
        match[i] = r;
    }
}

// keys per sub-table: blockIdx.y = sub-table, blockIdx.x = a slice of it
__global__ __launch_bounds__(256) void k_join_occupancy(const JoinSlot* __restrict__ slots, uint64_t per,
                                                        unsigned long long* __restrict__ occ) {
    __shared__ unsigned long long red[4];
    const uint64_t base = (uint64_t)blockIdx.y * per;
    unsigned long long c = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < per; i += (uint64_t)gridDim.x * 256)
        c += slots[base + i].key != kEmpty ? 1 : 0;
    for (int m = 32; m > 0; m >>= 1) c += __shfl_xor(c, m, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) atomicAdd(&occ[blockIdx.y], red[0] + red[1] + red[2] + red[3]);
}

struct PairSrc {
    const uint32_t* match;
    uint32_t* left;
    uint32_t* right;
    __device__ __forceinline__ bool pass(uint64_t i) const { return match[i] != 0xFFFFFFFFu; }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t k) const {
        left[k] = (uint32_t)i;
        right[k] = match[i];
    }
};

unsigned grid_for(uint64_t n, unsigned cap) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b ? (b < cap ? b : cap) : 1);
}

}  // namespace

struct JoinTable {
    JoinSlot* slots = nullptr;
    uint32_t* rows = nullptr;
    unsigned long long* side = nullptr;     // [0] row+1 of INT64_MIN key, [1] payload, [2..3] probe results,
    unsigned long long* side_host = nullptr;    // [4] keys that found their sub-table full
    uint32_t* match = nullptr;
    uint64_t match_cap = 0;
    uint64_t nslots = 0;
    uint64_t nbuild = 0;
    JoinLayout L;
    uint2* cells = nullptr;                  // range modes: the partition table
    unsigned long long* occ = nullptr;       // keys per sub-table
    int64_t* pk = nullptr;                   // probe keys in partition order
    uint64_t pk_cap = 0;
    RadixWork w;
    int partition_probes = 1;
    ~JoinTable() {
        for (void* p : {(void*)slots, (void*)rows, (void*)side, (void*)match, (void*)cells, (void*)occ, (void*)pk,
                        (void*)w.status})
            if (p) (void)hipFree(p);
        if (side_host) (void)hipHostFree(side_host);
    }
};

constexpr int kSubTableLog2 = 17;                 // 2^17 slots = 2 MB per sub-table
constexpr uint64_t kPartMinProbes = 1ull << 20;   // smaller probe batches are probed in input order

// Builds t's slots for layout t->L; returns false on a HIP error.  The
// number of keys that found their sub-table full is in t->side_host[4].
static bool join_fill_table(JoinTable* t, const int64_t* keys, const int64_t* payload, uint64_t nb, hipStream_t st) {
    const uint64_t slots = t->nslots;
    hipLaunchKernelGGL(k_join_clear, dim3(grid_for(slots, 8192)), dim3(256), 0, st, t->slots, t->rows, slots, t->side);
    if (nb) hipLaunchKernelGGL(k_join_insert, dim3(grid_for(nb, 8192)), dim3(256), 0, st, keys, nb, t->slots, t->rows,
                               t->L, t->side);
    hipLaunchKernelGGL(k_join_fill, dim3(grid_for(slots, 8192)), dim3(256), 0, st, t->slots, t->rows, slots, payload,
                       t->side);
    return hip_ok(hipGetLastError(), "join build launch") &&
           hip_ok(hipMemcpyAsync(t->side_host, t->side, 64, hipMemcpyDeviceToHost, st), "read join build") &&
           hip_ok(hipStreamSynchronize(st), "join build");
}

}  // namespace vx

using namespace vx;

extern "C" {

void* voxel_gpu_join_build_ex(const int64_t* build_keys_dev, const int64_t* payload_dev, size_t nb, int mode,
                              voxel_gpu_join_stats* stats, void* stream) {
    set_error("%s", "");
    if (nb >= (1ull << 32)) { set_error("join build: %zu rows exceed u32 row ids", nb); return nullptr; }
    if (mode < 0 || mode > 4) { set_error("join build: mode %d (0 auto, 1 one table, 2 hash, 3 learned, 4 equal width)", mode); return nullptr; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return nullptr;
    voxel_gpu_join_stats S;
    std::memset(&S, 0, sizeof(S));
    auto* t = new JoinTable();
    uint64_t want = nb * 2 > 64 ? nb * 2 : 64;
    uint64_t slots = 1;
    int lg = 0;
    while (slots < want) { slots <<= 1; ++lg; }
    t->nslots = slots;
    t->nbuild = nb;
    bool ok = hip_ok(hipMalloc((void**)&t->slots, slots * sizeof(JoinSlot)), "hipMalloc(join slots)") &&
              hip_ok(hipMalloc((void**)&t->rows, slots * 4), "hipMalloc(join rows)") &&
              hip_ok(hipMalloc((void**)&t->side, 64), "hipMalloc") &&
              hip_ok(hipMalloc((void**)&t->occ, 256 * 8), "hipMalloc") &&
              hip_ok(hipHostMalloc((void**)&t->side_host, 64, hipHostMallocDefault), "hipHostMalloc");
    if (!ok) { delete t; return nullptr; }
    hipStream_t st = c->stream;
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipEventRecord(e0, st);
    if (mode == 0) mode = kOneTable;         // measured fastest on MI355X (see the probe comment above)
    if ((mode == kLearnedParts || mode == kEqualWidthParts) && nb < 4096) mode = kHashParts;   // too few keys to fit
    int pbits = lg - kSubTableLog2;
    pbits = pbits < 1 ? 1 : (pbits > 8 ? 8 : pbits);
    if (lg - pbits < 5) mode = kOneTable;
    auto one_table = [&]() { t->L = JoinLayout{}; t->L.mode = kOneTable; t->L.pbits = 0; t->L.sbits = lg; };
    if (mode == kOneTable) {
        one_table();
    } else {
        t->L.mode = mode;
        t->L.pbits = pbits;
        t->L.sbits = lg - pbits;
        if (mode != kHashParts) {
            if (!hip_ok(hipMalloc((void**)&t->cells, (65536 + 2) * sizeof(uint2)), "hipMalloc(partition table)") ||
                !learned_partitioner((const uint64_t*)build_keys_dev, nb, 0x8000000000000000ull,
                                     mode == kLearnedParts ? 0 : 1, 1u << pbits, t->cells, &t->L.t, &S.model, c)) {
                delete t;
                return nullptr;
            }
        }
    }
    if (!join_fill_table(t, build_keys_dev, payload_dev, nb, st)) { delete t; return nullptr; }
    if (t->side_host[4] != 0) {
        // a sub-table overflowed: one table (it holds every key: slots >= 2 * rows)
        S.fell_back = t->L.mode;
        one_table();
        if (!join_fill_table(t, build_keys_dev, payload_dev, nb, st)) { delete t; return nullptr; }
    }
    S.mode = t->L.mode;
    S.partitions = 1u << t->L.pbits;
    S.slots_per_partition = 1ull << t->L.sbits;
    const unsigned slices = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>(S.slots_per_partition / 4096,
                                                                                    4096 / S.partitions));
    if (!hip_ok(launch_zero32((uint32_t*)t->occ, 2 * S.partitions, st), "zero")) { delete t; return nullptr; }
    hipLaunchKernelGGL(k_join_occupancy, dim3(slices, S.partitions), dim3(256), 0, st, t->slots, S.slots_per_partition,
                       t->occ);
    (void)hipEventRecord(e1, st);
    std::vector<unsigned long long> occ(S.partitions);
    if (!hip_ok(hipGetLastError(), "occupancy") || !hip_ok(hipStreamSynchronize(st), "sync") ||
        !hip_ok(hipMemcpy(occ.data(), t->occ, S.partitions * 8, hipMemcpyDeviceToHost), "read occupancy")) {
        delete t;
        return nullptr;
    }
    S.min_partition_keys = ~0ull;
    for (auto o : occ) {
        S.max_partition_keys = o > S.max_partition_keys ? o : S.max_partition_keys;
        S.min_partition_keys = o < S.min_partition_keys ? o : S.min_partition_keys;
    }
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    S.build_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (stats) *stats = S;
    return t;
}

void* voxel_gpu_join_build(const int64_t* build_keys_dev, const int64_t* payload_dev, size_t nb, void* stream) {
    return voxel_gpu_join_build_ex(build_keys_dev, payload_dev, nb, 0, nullptr, stream);
}

int voxel_gpu_join_set_probe_partitioning(void* j, int on) {
    set_error("%s", "");
    auto* t = static_cast<JoinTable*>(j);
    if (!t) { set_error("null join"); return -1; }
    t->partition_probes = on ? 1 : 0;
    return 0;
}

void voxel_gpu_join_destroy(void* j) { delete static_cast<JoinTable*>(j); }

int voxel_gpu_join_probe_sum_launch(void* j, const int64_t* probe_dev, size_t np, void* stream) {
    set_error("%s", "");
    auto* t = static_cast<JoinTable*>(j);
    if (!t) { set_error("null join"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    hipStream_t st = c->stream;
    if (!hip_ok(launch_zero32((uint32_t*)(t->side + 2), 4, st), "zero")) return -1;
    if (np == 0) return 0;
    const int64_t* keys = probe_dev;
    if (t->L.mode != kOneTable && t->partition_probes && np >= kPartMinProbes) {
        // probes in sub-table order: one stable digit pass, keys only
        if (np >= (1ull << 32)) { set_error("join probe: %zu rows exceed one partition pass", np); return -1; }
        if (np > t->pk_cap) {
            if (t->pk) (void)hipFree(t->pk);
            t->pk = nullptr; t->pk_cap = 0;
            if (!hip_ok(hipMalloc((void**)&t->pk, np * 8), "hipMalloc(partitioned probes)")) return -1;
            t->pk_cap = np;
        }
        const uint64_t words = radix_work_words(np);
        if (words > t->w.status_cap) {
            if (t->w.status) (void)hipFree(t->w.status);
            t->w.status = nullptr; t->w.status_cap = 0;
            if (!hip_ok(hipMalloc((void**)&t->w.status, words * 4), "hipMalloc(partition counts)")) return -1;
            t->w.status_cap = words;
        }
        if (!hip_ok(radix_pass((const uint64_t*)probe_dev, nullptr, (uint64_t*)t->pk, nullptr, np, JoinDigit{t->L}, 0,
                               t->w, st, nullptr, nullptr),
                    "probe partition pass"))
            return -1;
        keys = t->pk;
    }
    unsigned blocks = (unsigned)device_cu_count() * 8;
    uint64_t need = (np + 255) / 256;
    if (need < blocks) blocks = (unsigned)(need ? need : 1);
    hipLaunchKernelGGL(k_join_probe_sum, dim3(blocks), dim3(256), 0, st, keys, (uint64_t)np, t->slots, t->L, t->side,
                       t->side + 2);
    return hip_ok(hipGetLastError(), "join probe launch") ? 0 : -1;
}

int voxel_gpu_join_probe_sum_result(void* j, uint64_t* match_count, int64_t* payload_sum, void* stream) {
    auto* t = static_cast<JoinTable*>(j);
    if (!t) { set_error("null join"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (!hip_ok(hipMemcpyAsync(t->side_host, t->side, 32, hipMemcpyDeviceToHost, c->stream), "read join result") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return -1;
    if (match_count) *match_count = t->side_host[2];
    if (payload_sum) *payload_sum = (int64_t)t->side_host[3];
    return 0;
}

int voxel_gpu_join_probe_sum(void* j, const int64_t* probe_dev, size_t np, uint64_t* match_count,
                             int64_t* payload_sum, void* stream) {
    if (voxel_gpu_join_probe_sum_launch(j, probe_dev, np, stream) != 0) return -1;
    return voxel_gpu_join_probe_sum_result(j, match_count, payload_sum, stream);
}

int voxel_gpu_join_probe_pairs(void* j, const int64_t* probe_dev, size_t np, uint32_t* left_dev, uint32_t* right_dev,
                               uint64_t* match_count, void* stream) {
    set_error("%s", "");
    auto* t = static_cast<JoinTable*>(j);
    if (!t) { set_error("null join"); return -1; }
    if (np >= (1ull << 32)) { set_error("join probe: %zu rows exceed u32 row ids", np); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (np > t->match_cap) {
        if (t->match) (void)hipFree(t->match);
        t->match = nullptr; t->match_cap = 0;
        if (!hip_ok(hipMalloc((void**)&t->match, np * 4), "hipMalloc(match)")) return -1;
        t->match_cap = np;
    }
    if (np) {
        hipLaunchKernelGGL(k_join_match, dim3(grid_for(np, 16384)), dim3(256), 0, c->stream, probe_dev, (uint64_t)np,
                           t->slots, t->rows, t->L, t->side, t->match);
        if (!hip_ok(hipGetLastError(), "join match launch")) return -1;
    }
    PairSrc src{t->match, left_dev, right_dev};
    uint64_t k = 0;
    if (!run_compact(src, np, c, &k)) return -1;
    if (match_count) *match_count = k;
    return 0;
}

}  // extern "C"
