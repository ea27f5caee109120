// vx_join.hip -- i64 hash join build/probe on gfx950, replacing
// HashJoin<TKey>::Build/Probe (ops/join.hpp:36-80, 150-165) and the
// join + SUM(payload) harness of the configs.
//
// Table: open addressing with linear probing over 16-byte slots
// {key, payload}, slots = nextpow2(max(2 * build_rows, 64)) (the reference's
// Init sizing, join.hpp:37), slot = HashValue<i64>(key) & (slots - 1) (the
// splitmix64 finaliser, ops/hash.hpp:328-333).  Empty slots hold INT64_MIN;
// a build key equal to INT64_MIN is kept in a side record.  The reference's
// chains are LIFO, so its Probe returns, for duplicate build keys, the
// highest build row id: the build keeps max(row) per key with atomicMax, then
// fills each slot's payload from that row.  One probe = one coalesced 8-byte
// key read + one 16-byte slot read per probe step.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/voxel_c.h"
#include "vx_compact.hpp"
#include "vx_device.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {

struct __attribute__((aligned(16))) JoinSlot {
    int64_t key;
    int64_t payload;
};

constexpr int64_t kEmpty = INT64_MIN;

__device__ __forceinline__ uint64_t join_hash(int64_t key) {
    uint64_t x = (uint64_t)key;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
    return x ^ (x >> 31);
}

namespace {

__global__ __launch_bounds__(256) void k_join_clear(JoinSlot* __restrict__ slots, uint32_t* __restrict__ rows,
                                                    uint64_t n, unsigned long long* __restrict__ side) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        slots[i].key = kEmpty;
        slots[i].payload = 0;
        rows[i] = 0;
    }
    if (blockIdx.x == 0 && threadIdx.x < 4) side[threadIdx.x] = 0;
}

// side[0] = 1 + highest row holding key INT64_MIN (0: none), side[1] = its payload
__global__ __launch_bounds__(256) void k_join_insert(const int64_t* __restrict__ keys, uint64_t nb,
                                                     JoinSlot* __restrict__ slots, uint32_t* __restrict__ rows,
                                                     uint64_t mask, unsigned long long* __restrict__ side) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += stride) {
        int64_t k = keys[i];
        if (k == kEmpty) { atomicMax(&side[0], (unsigned long long)(i + 1)); continue; }
        uint64_t h = join_hash(k) & mask;
        while (true) {
            unsigned long long prev = atomicCAS((unsigned long long*)&slots[h].key, (unsigned long long)kEmpty,
                                                (unsigned long long)k);
            if ((int64_t)prev == kEmpty || (int64_t)prev == k) {
                atomicMax(&rows[h], (uint32_t)i);
                break;
            }
            h = (h + 1) & mask;
        }
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
__device__ __forceinline__ bool join_find(const JoinSlot* __restrict__ slots, uint64_t mask, int64_t k, uint64_t& at) {
    uint64_t h = join_hash(k) & mask;
    while (true) {
        JoinSlot s = slots[h];
        if (s.key == k) { at = h; return true; }
        if (s.key == kEmpty) return false;
        h = (h + 1) & mask;
    }
}

__global__ __launch_bounds__(256) void k_join_probe_sum(const int64_t* __restrict__ probe, uint64_t np,
                                                        const JoinSlot* __restrict__ slots, uint64_t mask,
                                                        const unsigned long long* __restrict__ side,
                                                        unsigned long long* __restrict__ out) {
    uint64_t cnt = 0, sum = 0;
    const unsigned long long side_row = side[0], side_pay = side[1];
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
        int64_t k = probe[i];
        if (k == kEmpty) {
            if (side_row) { ++cnt; sum += side_pay; }
            continue;
        }
        uint64_t h = join_hash(k) & mask;
        while (true) {
            JoinSlot s = slots[h];
            if (s.key == k) { ++cnt; sum += (uint64_t)s.payload; break; }
            if (s.key == kEmpty) break;
            h = (h + 1) & mask;
        }
    }
    for (int m = 32; m > 0; m >>= 1) {
        cnt += __shfl_xor(cnt, m, 64);
        sum += __shfl_xor(sum, m, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        if (cnt) atomicAdd(&out[0], (unsigned long long)cnt);
        if (sum) atomicAdd(&out[1], (unsigned long long)sum);
    }
}

// match[i] = build row of the first match of probe i, or 0xFFFFFFFF
__global__ __launch_bounds__(256) void k_join_match(const int64_t* __restrict__ probe, uint64_t np,
                                                    const JoinSlot* __restrict__ slots, const uint32_t* __restrict__ rows,
                                                    uint64_t mask, const unsigned long long* __restrict__ side,
                                                    uint32_t* __restrict__ match) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < np; i += stride) {
        int64_t k = probe[i];
        uint32_t r = 0xFFFFFFFFu;
        if (k == kEmpty) {
            if (side[0]) r = (uint32_t)(side[0] - 1);
        } else {
            uint64_t at;
            if (join_find(slots, mask, k, at)) r = rows[at];
        }
        match[i] = r;
    }
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
    unsigned long long* side = nullptr;     // [0] row+1 of INT64_MIN key, [1] payload, [2..3] probe results
    unsigned long long* side_host = nullptr;
    uint32_t* match = nullptr;
    uint64_t match_cap = 0;
    uint64_t nslots = 0;
    uint64_t nbuild = 0;
    ~JoinTable() {
        if (slots) (void)hipFree(slots);
        if (rows) (void)hipFree(rows);
        if (side) (void)hipFree(side);
        if (side_host) (void)hipHostFree(side_host);
        if (match) (void)hipFree(match);
    }
};

}  // namespace vx

using namespace vx;

extern "C" {

void* voxel_gpu_join_build(const int64_t* build_keys_dev, const int64_t* payload_dev, size_t nb, void* stream) {
    set_error("%s", "");
    if (nb >= (1ull << 32)) { set_error("join build: %zu rows exceed u32 row ids", nb); return nullptr; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return nullptr;
    auto* t = new JoinTable();
    uint64_t want = nb * 2 > 64 ? nb * 2 : 64;
    uint64_t slots = 1;
    while (slots < want) slots <<= 1;
    t->nslots = slots;
    t->nbuild = nb;
    bool ok = hip_ok(hipMalloc((void**)&t->slots, slots * sizeof(JoinSlot)), "hipMalloc(join slots)") &&
              hip_ok(hipMalloc((void**)&t->rows, slots * 4), "hipMalloc(join rows)") &&
              hip_ok(hipMalloc((void**)&t->side, 64), "hipMalloc") &&
              hip_ok(hipHostMalloc((void**)&t->side_host, 64, hipHostMallocDefault), "hipHostMalloc");
    if (!ok) { delete t; return nullptr; }
    hipStream_t st = c->stream;
    hipLaunchKernelGGL(k_join_clear, dim3(grid_for(slots, 8192)), dim3(256), 0, st, t->slots, t->rows, slots, t->side);
    if (nb) hipLaunchKernelGGL(k_join_insert, dim3(grid_for(nb, 8192)), dim3(256), 0, st, build_keys_dev, (uint64_t)nb,
                               t->slots, t->rows, slots - 1, t->side);
    hipLaunchKernelGGL(k_join_fill, dim3(grid_for(slots, 8192)), dim3(256), 0, st, t->slots, t->rows, slots,
                       payload_dev, t->side);
    if (!hip_ok(hipGetLastError(), "join build launch") || !hip_ok(hipStreamSynchronize(st), "join build")) {
        delete t;
        return nullptr;
    }
    return t;
}

void voxel_gpu_join_destroy(void* j) { delete static_cast<JoinTable*>(j); }

int voxel_gpu_join_probe_sum_launch(void* j, const int64_t* probe_dev, size_t np, void* stream) {
    set_error("%s", "");
    auto* t = static_cast<JoinTable*>(j);
    if (!t) { set_error("null join"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (!hip_ok(launch_zero32((uint32_t*)(t->side + 2), 4, c->stream), "zero")) return -1;
    if (np) {
        unsigned blocks = (unsigned)device_cu_count() * 16;
        uint64_t need = (np + 255) / 256;
        if (need < blocks) blocks = (unsigned)need;
        hipLaunchKernelGGL(k_join_probe_sum, dim3(blocks), dim3(256), 0, c->stream, probe_dev, (uint64_t)np, t->slots,
                           t->nslots - 1, t->side, t->side + 2);
    }
    return hip_ok(hipGetLastError(), "join probe launch") ? 0 : -1;
}

int voxel_gpu_join_probe_sum_result(void* j, uint64_t* match_count, int64_t* payload_sum, void* stream) {
    auto* t = static_cast<JoinTable*>(j);
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
                           t->slots, t->rows, t->nslots - 1, t->side, t->match);
        if (!hip_ok(hipGetLastError(), "join match launch")) return -1;
    }
    PairSrc src{t->match, left_dev, right_dev};
    uint64_t k = 0;
    if (!run_compact(src, np, c, &k)) return -1;
    if (match_count) *match_count = k;
    return 0;
}

}  // extern "C"
