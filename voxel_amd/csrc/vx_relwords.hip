// vx_relwords.hip -- the hash / sort / join words of straight-line programs
// on gfx950 (the reference's Engine<T>::HandleHashInit/Probe, HandleSort,
// HandleSortTopK, HandleJoinHash/Merge/Nested/AntiSemi, exec/engine.hpp:
// 704-842, and HashElement / HashU64, engine.hpp:1051-1071).
//
//   * SORT_ASC / SORT_DESC: values -> order keys (one pass, NaN flagged),
//     keys-only LSD radix sort (the onesweep passes of vx_onesweep.hpp),
//     keys -> values written in place.
//   * SORT_TOPK / SORT_BOTTOMK: libstdc++'s partial_sort (GCC 11
//     bits/stl_heap.h / bits/stl_algo.h: __make_heap, __heap_select,
//     __sort_heap over __adjust_heap / __push_heap), restated in one
//     workgroup: the heap of k <= 4095 rows lives in LDS, one lane runs the
//     heap steps, and the wave tests 512 rows per step against the heap top
//     so that only rows that replace the top cost sequential work.  Row
//     moves equal the reference's bit for bit (NaN and -0.0 included).
//   * JOIN_*: the right rows' order keys (-0.0 as +0.0) sorted, then one
//     thread per left row counts its equal run by two binary searches;
//     JOIN_MERGE also sorts the left keys and matches the i-th row of a run
//     while the right run lasts (the merge walk's pairing).
//   * HASH_*: (HashElement, row) pairs radix-sorted by hash (stable), the
//     first row of each hash kept by an order-preserving compaction; a probe
//     is one binary search.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "vx_compact.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_relwords.hpp"
#include "vx_runtime.hpp"

namespace vx {

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

namespace {

constexpr int kRelThreads = 256;

__host__ __device__ __forceinline__ bool nan_bits(uint64_t b) {
    return ((b >> 52) & 0x7FFu) == 0x7FFu && (b & 0xFFFFFFFFFFFFFull) != 0;
}

// Unsigned key in the value order of T: f64 by the total order of the bits
// (-0.0 just below +0.0; canon maps -0.0 to +0.0 first), i64 signed.
__host__ __device__ __forceinline__ uint64_t value_key(uint64_t b, bool f64, bool canon) {
    if (!f64) return b ^ 0x8000000000000000ull;
    if (canon && b == 0x8000000000000000ull) b = 0;
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__host__ __device__ __forceinline__ uint64_t value_unkey(uint64_t k, bool f64) {
    if (!f64) return k ^ 0x8000000000000000ull;
    return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
}

__host__ __device__ __forceinline__ uint64_t hash_u64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    x ^= x >> 33;
    return x;
}

unsigned grid_of(uint64_t n, unsigned cap = 8192) {
    const uint64_t b = (n + kRelThreads - 1) / kRelThreads;
    return (unsigned)(b == 0 ? 1 : (b < cap ? b : cap));
}

__global__ void k_rel_clear(unsigned long long* p) {
    if (threadIdx.x < 4) p[threadIdx.x] = 0;
}

// keys[i] = order key of x[i] (desc: complemented, so an ascending sort of
// the keys is the descending order of the values); flag |= NaN seen.
__global__ __launch_bounds__(kRelThreads) void k_order_keys(const uint64_t* __restrict__ x, uint64_t n, int f64,
                                                            int canon, int desc, uint64_t* __restrict__ keys,
                                                            unsigned long long* __restrict__ flag) {
    bool nan = false;
    const uint64_t stride = (uint64_t)gridDim.x * kRelThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kRelThreads + threadIdx.x; i < n; i += stride) {
        const uint64_t b = x[i];
        nan |= f64 && nan_bits(b);
        const uint64_t k = value_key(b, f64 != 0, canon != 0);
        keys[i] = desc ? ~k : k;
    }
    if (__ballot(nan) != 0 && (threadIdx.x & 63) == 0) *flag = 1ull;
}

__global__ __launch_bounds__(kRelThreads) void k_store_values(const uint64_t* __restrict__ keys, uint64_t n, int f64,
                                                              int desc, uint64_t* __restrict__ x) {
    const uint64_t stride = (uint64_t)gridDim.x * kRelThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kRelThreads + threadIdx.x; i < n; i += stride)
        x[i] = value_unkey(desc ? ~keys[i] : keys[i], f64 != 0);
}

// ---- partial sort ---------------------------------------------------------
template <typename T, bool DESC>
__device__ __forceinline__ bool before(T a, T b) {          // std::greater (DESC) / std::less
    return DESC ? a > b : a < b;
}

// __push_heap: move the hole up while its parent comes before v.
template <typename T, bool DESC>
__device__ void heap_push(T* f, int hole, int top, T v) {
    int parent = (hole - 1) / 2;
    while (hole > top && before<T, DESC>(f[parent], v)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = v;
}

// __adjust_heap: move the hole down to a leaf along the child that does not
// come before its sibling, then push v from there.
template <typename T, bool DESC>
__device__ void heap_adjust(T* f, int hole, int len, T v) {
    const int top = hole;
    int child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (before<T, DESC>(f[child], f[child - 1])) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    heap_push<T, DESC>(f, hole, top, v);
}

__device__ __forceinline__ uint64_t shfl64(uint64_t v, int src) {
    const int lo = __shfl((int)(uint32_t)v, src), hi = __shfl((int)(uint32_t)(v >> 32), src);
    return (uint64_t)(uint32_t)lo | ((uint64_t)(uint32_t)hi << 32);
}

template <typename T> __device__ __forceinline__ T from_bits(uint64_t b) { return __builtin_bit_cast(T, b); }
template <typename T> __device__ __forceinline__ uint64_t to_bits(T v) { return __builtin_bit_cast(uint64_t, v); }

// One workgroup of one wave; k <= kTopkMax rows of heap in LDS.
template <typename T, bool DESC>
__global__ __launch_bounds__(64) void k_partial_sort(uint64_t* __restrict__ x, uint64_t n, int k) {
    __shared__ T heap[kTopkMax];
    const int lane = threadIdx.x;
    for (int i = lane; i < k; i += 64) heap[i] = from_bits<T>(x[i]);
    __syncthreads();
    if (lane == 0 && k >= 2) {                              // __make_heap(first, middle)
        int parent = (k - 2) / 2;
        for (;;) {
            heap_adjust<T, DESC>(heap, parent, k, heap[parent]);
            if (parent == 0) break;
            parent--;
        }
    }
    __syncthreads();
    constexpr int U = 8;                                    // rows per lane and step, loaded together
    for (uint64_t base = (uint64_t)k; base < n; base += 64 * U) {
        uint64_t v[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const uint64_t i = base + (uint64_t)u * 64 + lane;
            ok[u] = i < n;
            v[u] = ok[u] ? x[i] : 0ull;
        }
        // rows in index order: u-major, then lane; a row whose value comes
        // before the heap top replaces it (__heap_select's __pop_heap)
#pragma unroll
        for (int u = 0; u < U; ++u) {
            unsigned long long m = __ballot(ok[u] && before<T, DESC>(from_bits<T>(v[u]), heap[0]));
            while (m) {
                const int j = __ffsll((long long)m) - 1;
                const uint64_t val = shfl64(v[u], j);
                if (lane == 0) {
                    x[base + (uint64_t)u * 64 + j] = to_bits<T>(heap[0]);
                    heap_adjust<T, DESC>(heap, 0, k, from_bits<T>(val));
                }
                __syncthreads();
                m = __ballot(ok[u] && lane > j && before<T, DESC>(from_bits<T>(v[u]), heap[0]));
            }
        }
    }
    __syncthreads();
    if (lane == 0) {                                        // __sort_heap(first, middle)
        int len = k;
        while (len > 1) {
            --len;
            const T val = heap[len];
            heap[len] = heap[0];
            heap_adjust<T, DESC>(heap, 0, len, val);
        }
    }
    __syncthreads();
    for (int i = lane; i < k; i += 64) x[i] = to_bits<T>(heap[i]);
}

// ---- joins ----------------------------------------------------------------
__device__ __forceinline__ uint64_t lower_bound(const uint64_t* __restrict__ s, uint64_t n, uint64_t key) {
    uint64_t a = 0, b = n;
    while (a < b) {
        const uint64_t m = a + ((b - a) >> 1);
        if (s[m] < key) a = m + 1; else b = m;
    }
    return a;
}
__device__ __forceinline__ uint64_t upper_bound(const uint64_t* __restrict__ s, uint64_t n, uint64_t key,
                                                uint64_t from) {
    uint64_t a = from, b = n;
    while (a < b) {
        const uint64_t m = a + ((b - a) >> 1);
        if (s[m] <= key) a = m + 1; else b = m;
    }
    return a;
}

// KIND 0 pairs, 1 merge (l = sorted left keys), 2 semi, 3 anti; the sum of
// the rows' counts is added to *out.
template <int KIND>
__global__ __launch_bounds__(kRelThreads) void k_join_count(const uint64_t* __restrict__ l, uint64_t nl,
                                                            const uint64_t* __restrict__ rs, uint64_t nr, int f64,
                                                            unsigned long long* __restrict__ out) {
    unsigned long long c = 0;
    const uint64_t stride = (uint64_t)gridDim.x * kRelThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kRelThreads + threadIdx.x; i < nl; i += stride) {
        if constexpr (KIND == 1) {
            const uint64_t key = l[i];
            const uint64_t rank = i - lower_bound(l, nl, key);
            const uint64_t lo = lower_bound(rs, nr, key), hi = upper_bound(rs, nr, key, lo);
            c += rank < hi - lo;
        } else {
            const uint64_t b = l[i];
            if (f64 && nan_bits(b)) {                       // NaN equals nothing
                c += KIND == 3;
                continue;
            }
            const uint64_t key = value_key(b, f64 != 0, true);
            const uint64_t lo = lower_bound(rs, nr, key);
            const uint64_t hi = upper_bound(rs, nr, key, lo);
            if constexpr (KIND == 0) c += hi - lo;
            else if constexpr (KIND == 2) c += hi > lo;
            else c += hi == lo;
        }
    }
    for (int o = 32; o > 0; o >>= 1) c += (unsigned long long)shfl64(c, (threadIdx.x & 63) ^ o);
    __shared__ unsigned long long ws[kRelThreads / 64];
    if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < kRelThreads / 64; ++w) t += ws[w];
        if (t) atomicAdd(out, t);
    }
}

// ---- hash tables ----------------------------------------------------------
__global__ __launch_bounds__(kRelThreads) void k_hash_elems(const uint64_t* __restrict__ x, uint64_t n, int f64,
                                                            uint64_t* __restrict__ h, uint32_t* __restrict__ row) {
    const uint64_t stride = (uint64_t)gridDim.x * kRelThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kRelThreads + threadIdx.x; i < n; i += stride) {
        const uint64_t b = x[i];
        h[i] = (f64 && nan_bits(b)) ? 0ull : hash_u64(b);
        row[i] = (uint32_t)i;
    }
}

// the first (lowest) row of each hash in the stably sorted pairs
struct FirstOfHash {
    const uint64_t* h;
    const uint32_t* row;
    uint64_t* th;
    uint64_t* trow;
    __device__ __forceinline__ bool pass(uint64_t i) const { return i == 0 || h[i] != h[i - 1]; }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t j) const { th[j] = h[i]; trow[j] = row[i]; }
};

__global__ void k_hash_probe(const uint64_t* __restrict__ th, const uint64_t* __restrict__ trow, uint64_t n,
                             uint64_t h, unsigned long long* __restrict__ out) {
    if (threadIdx.x != 0) return;
    const uint64_t i = lower_bound(th, n, h);
    *out = (i < n && th[i] == h) ? (unsigned long long)trow[i] : ~0ull;
}

bool grow64(uint64_t** p, uint64_t n) {
    if (*p) (void)hipFree(*p);
    *p = nullptr;
    return hip_ok(hipMalloc((void**)p, n * 8), "hipMalloc");
}

bool read_slots(RelBufs& b, StreamCtx* c) {
    return hip_ok(hipMemcpyAsync(b.host, b.dev, 32, hipMemcpyDeviceToHost, c->stream), "read back") &&
           hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // namespace

RelBufs::~RelBufs() {
    for (void* p : {(void*)a1, (void*)a2, (void*)b1, (void*)b2, (void*)v1, (void*)v2, (void*)dev})
        if (p) (void)hipFree(p);
    if (host) (void)hipHostFree(host);
}

bool RelBufs::ensure(uint64_t n) {
    if (!dev) {
        if (!hip_ok(hipMalloc((void**)&dev, 64), "hipMalloc") ||
            !hip_ok(hipHostMalloc((void**)&host, 64, hipHostMallocDefault), "hipHostMalloc"))
            return false;
    }
    if (n <= cap) return true;
    cap = 0;
    if (!grow64(&a1, n) || !grow64(&a2, n) || !grow64(&b1, n) || !grow64(&b2, n)) return false;
    for (uint32_t** p : {&v1, &v2}) {
        if (*p) (void)hipFree(*p);
        *p = nullptr;
        if (!hip_ok(hipMalloc((void**)p, n * 4), "hipMalloc")) return false;
    }
    cap = n;
    return true;
}

HashTab::~HashTab() {
    if (h) (void)hipFree(h);
    if (row) (void)hipFree(row);
}

bool HashTab::ensure(uint64_t n) {
    if (n <= cap && h) return true;
    cap = 0;
    if (!grow64(&h, n ? n : 1) || !grow64(&row, n ? n : 1)) return false;
    cap = n ? n : 1;
    return true;
}

uint64_t relw_hash_element(uint64_t bits, int dtype) {
    if (dtype == DT_F64 && nan_bits(bits)) return 0;
    return hash_u64(bits);
}

bool relw_sort(int dtype, void* d, uint64_t n, bool desc, RelBufs& b, StreamCtx* c) {
    if (n < 2) return true;
    if (n >= (1ull << 30)) { set_error("SORT over %llu rows: the GPU sort takes fewer than 2^30", (unsigned long long)n); return false; }
    if (!b.ensure(n)) return false;
    const bool f64 = dtype == DT_F64;
    uint64_t* x = static_cast<uint64_t*>(d);
    hipLaunchKernelGGL(k_rel_clear, dim3(1), dim3(64), 0, c->stream, b.dev);
    hipLaunchKernelGGL(k_order_keys, dim3(grid_of(n)), dim3(kRelThreads), 0, c->stream, x, n, (int)f64, 0,
                       (int)desc, b.a1, b.dev + 1);
    if (!hip_ok(hipGetLastError(), "sort key launch") || !read_slots(b, c)) return false;
    if (b.host[1]) {
        set_error("SORT_ASC / SORT_DESC over NaN values is not lowered (std::sort's order with NaN is unspecified)");
        return false;
    }
    if (sort_pairs_u64(b.a1, nullptr, b.a2, nullptr, n, c) < 0) return false;
    hipLaunchKernelGGL(k_store_values, dim3(grid_of(n)), dim3(kRelThreads), 0, c->stream, (const uint64_t*)b.a2, n,
                       (int)f64, (int)desc, x);
    return hip_ok(hipGetLastError(), "sort store launch");
}

bool relw_partial_sort(int dtype, void* d, uint64_t n, uint32_t k, bool desc, StreamCtx* c) {
    if (k < 1 || k > kTopkMax || k > n) { set_error("partial sort: k = %u outside 1..min(%llu, %u)", k,
                                                    (unsigned long long)n, kTopkMax); return false; }
    uint64_t* x = static_cast<uint64_t*>(d);
    if (dtype == DT_F64) {
        if (desc) hipLaunchKernelGGL((k_partial_sort<double, true>), dim3(1), dim3(64), 0, c->stream, x, n, (int)k);
        else hipLaunchKernelGGL((k_partial_sort<double, false>), dim3(1), dim3(64), 0, c->stream, x, n, (int)k);
    } else {
        if (desc) hipLaunchKernelGGL((k_partial_sort<int64_t, true>), dim3(1), dim3(64), 0, c->stream, x, n, (int)k);
        else hipLaunchKernelGGL((k_partial_sort<int64_t, false>), dim3(1), dim3(64), 0, c->stream, x, n, (int)k);
    }
    return hip_ok(hipGetLastError(), "partial sort launch");
}

bool relw_join_count(int kind, int dtype, const void* l, uint64_t nl, const void* r, uint64_t nr, uint64_t* out,
                     RelBufs& b, StreamCtx* c) {
    *out = 0;
    const uint64_t big = nl > nr ? nl : nr;
    if (big >= (1ull << 30)) { set_error("JOIN over %llu rows: the GPU sort takes fewer than 2^30", (unsigned long long)big); return false; }
    if (kind == 3 && nr == 0) { *out = nl; return true; }
    if (nl == 0 || nr == 0) return true;
    if (!b.ensure(big)) return false;
    const int f64 = dtype == DT_F64;
    hipLaunchKernelGGL(k_rel_clear, dim3(1), dim3(64), 0, c->stream, b.dev);
    hipLaunchKernelGGL(k_order_keys, dim3(grid_of(nr)), dim3(kRelThreads), 0, c->stream,
                       static_cast<const uint64_t*>(r), nr, f64, 1, 0, b.a1, b.dev + 1);
    if (!hip_ok(hipGetLastError(), "join key launch")) return false;
    if (sort_pairs_u64(b.a1, nullptr, b.a2, nullptr, nr, c) < 0) return false;
    const unsigned g = grid_of(nl);
    if (kind == 1) {
        hipLaunchKernelGGL(k_order_keys, dim3(grid_of(nl)), dim3(kRelThreads), 0, c->stream,
                           static_cast<const uint64_t*>(l), nl, f64, 1, 0, b.b1, b.dev + 1);
        if (!hip_ok(hipGetLastError(), "join key launch") || !read_slots(b, c)) return false;
        if (b.host[1]) {
            set_error("JOIN_MERGE over NaN values is not lowered (std::sort's order with NaN is unspecified)");
            return false;
        }
        if (sort_pairs_u64(b.b1, nullptr, b.b2, nullptr, nl, c) < 0) return false;
        hipLaunchKernelGGL(k_join_count<1>, dim3(g), dim3(kRelThreads), 0, c->stream, (const uint64_t*)b.b2, nl,
                           (const uint64_t*)b.a2, nr, f64, b.dev);
    } else if (kind == 0) {
        hipLaunchKernelGGL(k_join_count<0>, dim3(g), dim3(kRelThreads), 0, c->stream, static_cast<const uint64_t*>(l),
                           nl, (const uint64_t*)b.a2, nr, f64, b.dev);
    } else if (kind == 2) {
        hipLaunchKernelGGL(k_join_count<2>, dim3(g), dim3(kRelThreads), 0, c->stream, static_cast<const uint64_t*>(l),
                           nl, (const uint64_t*)b.a2, nr, f64, b.dev);
    } else {
        hipLaunchKernelGGL(k_join_count<3>, dim3(g), dim3(kRelThreads), 0, c->stream, static_cast<const uint64_t*>(l),
                           nl, (const uint64_t*)b.a2, nr, f64, b.dev);
    }
    if (!hip_ok(hipGetLastError(), "join count launch") || !read_slots(b, c)) return false;
    *out = b.host[0];
    return true;
}

bool relw_hash_build(int dtype, const void* d, uint64_t n, HashTab& t, uint64_t* distinct, RelBufs& b,
                     StreamCtx* c) {
    *distinct = 0;
    if (n >= (1ull << 30)) { set_error("HASH_INIT over %llu rows: the GPU sort takes fewer than 2^30", (unsigned long long)n); return false; }
    if (!t.ensure(n)) return false;
    if (n == 0) return true;
    if (!b.ensure(n)) return false;
    hipLaunchKernelGGL(k_hash_elems, dim3(grid_of(n)), dim3(kRelThreads), 0, c->stream,
                       static_cast<const uint64_t*>(d), n, (int)(dtype == DT_F64), b.a1, b.v1);
    if (!hip_ok(hipGetLastError(), "hash launch")) return false;
    if (sort_pairs_u64(b.a1, b.v1, b.a2, b.v2, n, c) < 0) return false;
    FirstOfHash src{b.a2, b.v2, t.h, t.row};
    return run_compact(src, n, c, distinct);
}

bool relw_hash_probe(const HashTab& t, uint64_t ntab, uint64_t h, uint64_t* row, RelBufs& b, StreamCtx* c) {
    *row = ~0ull;
    if (ntab == 0) return true;
    if (!b.ensure(0)) return false;
    hipLaunchKernelGGL(k_hash_probe, dim3(1), dim3(64), 0, c->stream, (const uint64_t*)t.h, (const uint64_t*)t.row,
                       ntab, h, b.dev + 2);
    if (!hip_ok(hipGetLastError(), "probe launch") || !read_slots(b, c)) return false;
    *row = b.host[2];
    return true;
}

}  // namespace vx
