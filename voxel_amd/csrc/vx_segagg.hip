// vx_segagg.hip -- device helpers of the segment aggregates (vx_segagg.hpp).
//
// Moments: every thread runs Welford's update over its rows (grid stride),
// the 256 threads of a workgroup are merged pairwise in thread order with
// Chan's formula, and the host merges the workgroups' partials in workgroup
// order, so the result does not depend on scheduling.  Order keys and the
// distinct set use the radix sort (vx_radix.hip) and the stream compaction
// (vx_compact.hpp).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>

#include "vx_acc.hpp"
#include "vx_compact.hpp"
#include "vx_expr_ops.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"
#include "vx_segagg.hpp"

#pragma clang fp contract(off)

namespace vx {

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

namespace {
using namespace xo;

constexpr int kMomThreads = 256;
constexpr unsigned kMomBlocksMax = 1024;

__host__ __device__ inline Moments chan_merge(const Moments& a, const Moments& b) {
    if (a.n == 0.0) return b;
    if (b.n == 0.0) return a;
    const double n = a.n + b.n;
    const double d = b.mean - a.mean;
    Moments r;
    r.n = n;
    r.mean = a.mean + d * (b.n / n);
    r.m2 = a.m2 + b.m2 + d * d * (a.n * (b.n / n));
    return r;
}

template <typename T>
__global__ __launch_bounds__(kMomThreads) void k_moments(const T* __restrict__ x, uint64_t n, Moments* __restrict__ parts) {
    __shared__ Moments red[kMomThreads];
    Moments m{0.0, 0.0, 0.0};
    const uint64_t stride = (uint64_t)gridDim.x * kMomThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kMomThreads + threadIdx.x; i < n; i += stride) {
        const double v = (double)x[i];
        m.n += 1.0;
        const double d = v - m.mean;
        m.mean += d / m.n;
        m.m2 += d * (v - m.mean);
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = kMomThreads / 2; s > 0; s >>= 1) {          // pairs (t, t + s): thread order preserved
        if ((int)threadIdx.x < s) red[threadIdx.x] = chan_merge(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(256) void k_i64_keys(const int64_t* __restrict__ x, uint64_t n, uint64_t* __restrict__ keys,
                                                  uint32_t* __restrict__ idx) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        keys[i] = (uint64_t)x[i] ^ 0x8000000000000000ull;
        idx[i] = (uint32_t)i;
    }
}

// first row of each run of equal keys -> its key
struct UniqueSrc {
    const uint64_t* k;
    uint64_t* out;
    __device__ __forceinline__ bool pass(uint64_t i) const { return i == 0 || k[i] != k[i - 1]; }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t j) const { out[j] = k[i]; }
};

// f64 order keys -> the values' bits
__global__ __launch_bounds__(256) void k_unkey_f64(const uint64_t* __restrict__ k, uint64_t n, uint64_t* __restrict__ out) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const uint64_t key = k[i];
        out[i] = (key >> 63) ? (key & 0x7FFFFFFFFFFFFFFFull) : ~key;
    }
}

// An f64 order key of a NaN: above the key of +inf or below that of -inf.
__device__ __forceinline__ bool nan_key(uint64_t k) {
    return k > 0xFFF0000000000000ull || k < 0x000FFFFFFFFFFFFFull;
}

// Map entries of sorted keys (AGG_MODE): a run of equal keys is one entry,
// each NaN row an entry of its own; emits each entry's first position.
struct EntrySrc {
    const uint64_t* k;
    uint32_t* out;
    int f64;
    __device__ __forceinline__ bool pass(uint64_t i) const {
        return i == 0 || k[i] != k[i - 1] || (f64 && nan_key(k[i]));
    }
    __device__ __forceinline__ void emit(uint64_t i, uint64_t j) const { out[j] = (uint32_t)i; }
};

// Largest entry size, how many entries have it, and the first position of
// the earliest of them.
struct ModePart {
    uint64_t best, ties, first;
};
__device__ __forceinline__ ModePart mode_merge(ModePart a, ModePart b) {
    if (a.best != b.best) return a.best > b.best ? a : b;
    return ModePart{a.best, a.ties + b.ties, a.first < b.first ? a.first : b.first};
}

__global__ __launch_bounds__(256) void k_mode_entries(const uint32_t* __restrict__ starts, uint64_t ne, uint64_t n,
                                                      ModePart* __restrict__ parts) {
    __shared__ ModePart red[256];
    ModePart m{0, 0, ~0ull};
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < ne; j += stride) {
        const uint64_t s = starts[j], e = j + 1 < ne ? starts[j + 1] : n;
        m = mode_merge(m, ModePart{e - s, 1, s});
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) red[threadIdx.x] = mode_merge(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
    }
    if (threadIdx.x == 0) parts[blockIdx.x] = red[0];
}

unsigned gridn(uint64_t n, unsigned cap) {
    const uint64_t b = (n + 255) / 256;
    return (unsigned)(b ? (b < cap ? b : cap) : 1);
}

// expr_row runs the words on one row (registers v): each word dispatches to
// apply_op.  v[prog.vload] holds the loaded value on entry; every register
// the words read was written earlier in the same iteration (the lowering
// checks it), so nothing else needs a start value.
template <typename T>
__device__ __forceinline__ void expr_row(const ExprProg& prog, T (&v)[16], int lane, bool& trapped) {
    uint32_t msk = 0;                                              // this lane's bit of masks 0..7
    for (int k = 0; k < prog.nops; ++k) {
        const ExprOp& o = prog.ops[k];
        T s;
        __builtin_memcpy(&s, &o.s, sizeof(T));
        switch (o.op) {
#define VX_EXPR_OP(c) \
    case c: v[o.dst] = apply_op<T, c>(o, prog, v[o.a], v[o.b], s, lane, msk, trapped); break;
            VX_EXPR_CASES(VX_EXPR_OP)
#undef VX_EXPR_OP
        default: v[o.dst] = apply_op<T, 0xC0>(o, prog, v[o.a], v[o.b], s, lane, msk, trapped); break;   // VFILTER_*
        }
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_expr_eval(const T* __restrict__ col, uint64_t count, uint64_t off,
                                                   uint64_t n, ExprProg prog, T* __restrict__ out,
                                                   int* __restrict__ trap) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;     // a multiple of 4: rows keep their lane
    const int lane = (int)(threadIdx.x & 3);                       // row off + i is lane i mod 4 of its group
    bool trapped = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        T v[16];
        v[prog.vload] = off + i < count ? col[off + i] : T(0);
        expr_row<T>(prog, v, lane, trapped);
        out[i] = v[prog.target];
    }
    if (trapped) *trap = 1;
}

// The words and up to four VSUM / VCOUNT / VRED_MIN / VRED_MAX reductions
// in one pass: no evaluated column is written.  A workgroup takes tiles of
// 256 * kExprPf rows (row base + j * 256 + t), issuing the tile's loads
// before evaluating any of its rows.  Reduction q keeps a compensated f64
// pair or a wrapping i64 sum (VSUM) or a count of lanes that are non-zero and
// not NaN (VCOUNT) per thread, merged afterwards by k_scan_final, or the
// signed and unsigned extremes of its group values (VRED_*: each group
// reduced in lane order, its value's bits kept by the group's lane 0),
// merged by k_gacc_final; the workgroup's record goes to
// parts[q * gridDim.x + blockIdx.x].  The last group of four rows is reduced
// as the reference reduces one group into last[q], the loop's partial
// register.
constexpr int kExprPf = 4;

template <typename T>
__global__ __launch_bounds__(256) void k_expr_reduce(const T* __restrict__ col, uint64_t count, uint64_t off,
                                                     uint64_t n, ExprProg prog, ExprRed rd,
                                                     uint64_t* __restrict__ parts, uint64_t* __restrict__ last,
                                                     int* __restrict__ trap) {
    __shared__ XAcc<T> lds[256 / 64];
    __shared__ GAccX glds[256 / 64];
    XAcc<T> acc[4];
    GAccX gacc[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) { xacc_init(acc[q]); gacc_init(gacc[q]); }
    const int lane = (int)(threadIdx.x & 3);
    bool trapped = false;
    const uint64_t tile = 256ull * kExprPf;
    for (uint64_t base = (uint64_t)blockIdx.x * tile; base < n; base += (uint64_t)gridDim.x * tile) {
        T x[kExprPf];
#pragma unroll
        for (int j = 0; j < kExprPf; ++j) {
            const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
            x[j] = i < n && off + i < count ? col[off + i] : T(0);
        }
#pragma unroll
        for (int j = 0; j < kExprPf; ++j) {
            const uint64_t i = base + (uint64_t)j * 256 + threadIdx.x;
            if (i >= n) break;                                     // n is whole groups: quads leave together
            T v[16];
            v[prog.vload] = x[j];
            expr_row<T>(prog, v, lane, trapped);
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                if (q >= rd.n) break;
                const T y = v[rd.vreg[q]];
                if (rd.red[q] == 0) {
                    xacc_sum(acc[q], y);
                } else if (rd.red[q] == 1) {
                    xacc_count(acc[q], y);
                } else {
                    const uint64_t g = group_minmax(y, rd.red[q]);        // the quad's group value
                    if (lane == 0) gacc_add(gacc[q], g);
                }
                if (i + 4 >= n) group_partial(y, rd.red[q], lane, &last[q]);   // the last group
            }
        }
    }
    if (trapped) *trap = 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        if (q >= rd.n) break;
        uint64_t* part = parts + 8 * ((uint64_t)q * gridDim.x + blockIdx.x);
        if (rd.red[q] <= 1) xacc_block_store(acc[q], lds, part);
        else gacc_block_store(gacc[q], glds, part);
    }
}

// One workgroup: the signed / unsigned extremes of nb workgroup records
// (VRED_MIN / VRED_MAX reductions) -> out (the same four words).
__global__ __launch_bounds__(256) void k_gacc_final(const uint64_t* __restrict__ parts, uint32_t nb,
                                                    uint64_t* __restrict__ out) {
    __shared__ GAccX lds[256 / 64];
    GAccX a;
    gacc_init(a);
    for (uint32_t i = threadIdx.x; i < nb; i += 256) {
        GAccX o;
        o.smin = (int64_t)parts[8 * (uint64_t)i]; o.smax = (int64_t)parts[8 * (uint64_t)i + 1];
        o.umin = parts[8 * (uint64_t)i + 2]; o.umax = parts[8 * (uint64_t)i + 3];
        gacc_merge(a, o);
    }
    gacc_block_store(a, lds, out);
}

__global__ __launch_bounds__(256) void k_expr_store(const uint64_t* __restrict__ src, uint64_t n, uint32_t count,
                                                    uint64_t* __restrict__ dst, uint64_t off, uint64_t seg_count) {
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        if ((i & 3) < count && off + i < seg_count) dst[off + i] = src[i];
}

}  // namespace

bool expr_store(int dtype, const void* src, uint64_t n, uint32_t count, void* dst, uint64_t off, uint64_t seg_count,
                StreamCtx* c) {
    (void)dtype;                                       // 8-byte lanes either way
    if (n == 0) return true;
    hipLaunchKernelGGL(k_expr_store, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, (const uint64_t*)src, n, count,
                       (uint64_t*)dst, off, seg_count);
    return hip_ok(hipGetLastError(), "store launch");
}

bool expr_reduce(int dtype, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog,
                 const ExprRed& rd, ScanOut* res, uint64_t* last, int* trap, StreamCtx* c) {
    const uint64_t tile = 256ull * kExprPf;
    uint64_t nb = (n + tile - 1) / tile;
    // eight workgroups per CU, as many as the partials buffer holds for rd.n reductions
    const uint64_t cap = std::min<uint64_t>((uint64_t)device_cu_count() * 8,
                                            c->scr.max_blocks * 64 / ((uint64_t)(rd.n > 0 ? rd.n : 1) * sizeof(ScanOut)));
    if (nb > cap) nb = cap;
    if (nb < 1) nb = 1;
    ScanOut* parts = static_cast<ScanOut*>(c->scr.partials);
    static_assert(sizeof(ScanOut) == 64, "a scan partial is eight words");
    bool compiled = false;
    if (!jit_expr_launch(dtype == DT_F64, col, count, off, n, prog, rd, reinterpret_cast<uint64_t*>(parts),
                         (unsigned)nb, last, trap, c->stream, &compiled))
        return false;
    if (compiled) {
    } else if (dtype == DT_F64)
        hipLaunchKernelGGL(k_expr_reduce<double>, dim3((unsigned)nb), dim3(256), 0, c->stream, (const double*)col, count,
                           off, n, prog, rd, reinterpret_cast<uint64_t*>(parts), last, trap);
    else
        hipLaunchKernelGGL(k_expr_reduce<int64_t>, dim3((unsigned)nb), dim3(256), 0, c->stream, (const int64_t*)col,
                           count, off, n, prog, rd, reinterpret_cast<uint64_t*>(parts), last, trap);
    for (int q = 0; q < rd.n; ++q) {
        if (rd.red[q] >= 2)
            hipLaunchKernelGGL(k_gacc_final, dim3(1), dim3(256), 0, c->stream,
                               reinterpret_cast<const uint64_t*>((const ScanOut*)parts + q * nb), (uint32_t)nb,
                               reinterpret_cast<uint64_t*>(res + q));
        else if (dtype == DT_F64)
            hipLaunchKernelGGL(k_scan_final<double>, dim3(1), dim3(kBlock), 0, c->stream, (const ScanOut*)parts + q * nb,
                               (uint32_t)nb, res + q);
        else
            hipLaunchKernelGGL(k_scan_final<int64_t>, dim3(1), dim3(kBlock), 0, c->stream, (const ScanOut*)parts + q * nb,
                               (uint32_t)nb, res + q);
    }
    return hip_ok(hipGetLastError(), "fused vector expression launch");
}

bool expr_eval(int dtype, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog, void* out,
               int* trap, StreamCtx* c) {
    if (n == 0) return true;
    if (dtype == DT_F64)
        hipLaunchKernelGGL(k_expr_eval<double>, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, (const double*)col, count,
                           off, n, prog, (double*)out, trap);
    else
        hipLaunchKernelGGL(k_expr_eval<int64_t>, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, (const int64_t*)col,
                           count, off, n, prog, (int64_t*)out, trap);
    return hip_ok(hipGetLastError(), "vector expression launch");
}

SegaggBufs::~SegaggBufs() {
    for (void* p : {(void*)k1, (void*)k2, (void*)v1, (void*)v2, (void*)parts})
        if (p) (void)hipFree(p);
    if (parts_host) (void)hipHostFree(parts_host);
}

bool SegaggBufs::ensure(uint64_t n) {
    if (!parts) {
        if (!hip_ok(hipMalloc((void**)&parts, kMomBlocksMax * sizeof(Moments)), "hipMalloc") ||
            !hip_ok(hipHostMalloc((void**)&parts_host, kMomBlocksMax * sizeof(Moments), hipHostMallocDefault),
                    "hipHostMalloc"))
            return false;
    }
    if (n <= cap) return true;
    for (void* p : {(void*)k1, (void*)k2, (void*)v1, (void*)v2})
        if (p) (void)hipFree(p);
    k1 = k2 = nullptr;
    v1 = v2 = nullptr;
    cap = 0;
    if (!hip_ok(hipMalloc((void**)&k1, n * 8), "hipMalloc(aggregate keys)") ||
        !hip_ok(hipMalloc((void**)&k2, n * 8), "hipMalloc(aggregate keys)") ||
        !hip_ok(hipMalloc((void**)&v1, n * 4), "hipMalloc(aggregate rows)") ||
        !hip_ok(hipMalloc((void**)&v2, n * 4), "hipMalloc(aggregate rows)"))
        return false;
    cap = n;
    return true;
}

bool segagg_moments(int dtype, const void* x, uint64_t n, Moments* out, SegaggBufs& b, StreamCtx* c) {
    *out = Moments{0.0, 0.0, 0.0};
    if (n == 0) return true;
    if (!b.ensure(0)) return false;
    unsigned blocks = (unsigned)device_cu_count() * 2;
    if (blocks > kMomBlocksMax) blocks = kMomBlocksMax;
    const uint64_t need = (n + kMomThreads - 1) / kMomThreads;
    if (need < blocks) blocks = (unsigned)need;
    if (dtype == DT_F64)
        hipLaunchKernelGGL(k_moments<double>, dim3(blocks), dim3(kMomThreads), 0, c->stream, (const double*)x, n, b.parts);
    else
        hipLaunchKernelGGL(k_moments<int64_t>, dim3(blocks), dim3(kMomThreads), 0, c->stream, (const int64_t*)x, n, b.parts);
    if (!hip_ok(hipGetLastError(), "moments launch") ||
        !hip_ok(hipMemcpyAsync(b.parts_host, b.parts, blocks * sizeof(Moments), hipMemcpyDeviceToHost, c->stream),
                "read moments") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return false;
    Moments m{0.0, 0.0, 0.0};
    for (unsigned i = 0; i < blocks; ++i) m = chan_merge(m, b.parts_host[i]);
    *out = m;
    return true;
}

bool segagg_sorted_keys(int dtype, const void* x, uint64_t n, SegaggBufs& b, StreamCtx* c) {
    if (n >= (1ull << 30)) { set_error("aggregate: %llu rows exceed the sort's 2^30 limit", (unsigned long long)n); return false; }
    if (!b.ensure(n)) return false;
    if (dtype == DT_F64) {
        if (!hip_ok(launch_f64_keys((const double*)x, n, 0, b.k1, b.v1, c->stream), "key launch")) return false;
    } else {
        hipLaunchKernelGGL(k_i64_keys, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, (const int64_t*)x, n, b.k1, b.v1);
        if (!hip_ok(hipGetLastError(), "key launch")) return false;
    }
    return sort_pairs_u64(b.k1, b.v1, b.k2, b.v2, n, c) >= 0;
}

bool segagg_unkey_f64(const uint64_t* keys, uint64_t n, uint64_t* out, StreamCtx* c) {
    if (n == 0) return true;
    hipLaunchKernelGGL(k_unkey_f64, dim3(gridn(n, 16384)), dim3(256), 0, c->stream, keys, n, out);
    return hip_ok(hipGetLastError(), "unkey launch");
}

bool segagg_mode(int dtype, const void* x, uint64_t n, SegaggBufs& b, uint64_t* row, uint64_t* ties, StreamCtx* c) {
    *row = 0;
    *ties = 0;
    if (n == 0) return true;
    if (!segagg_sorted_keys(dtype, x, n, b, c)) return false;   // k2: keys, v2: rows (stable); k1 / v1 free
    EntrySrc src{b.k2, b.v1, dtype == DT_F64 ? 1 : 0};
    uint64_t ne = 0;
    if (!run_compact(src, n, c, &ne)) return false;
    unsigned blocks = gridn(ne, kMomBlocksMax);
    static_assert(sizeof(ModePart) == sizeof(Moments), "ModePart reuses the moments partials");
    ModePart* parts = reinterpret_cast<ModePart*>(b.parts);
    hipLaunchKernelGGL(k_mode_entries, dim3(blocks), dim3(256), 0, c->stream, (const uint32_t*)b.v1, ne, n, parts);
    if (!hip_ok(hipGetLastError(), "mode launch") ||
        !hip_ok(hipMemcpyAsync(b.parts_host, parts, blocks * sizeof(ModePart), hipMemcpyDeviceToHost, c->stream),
                "read mode") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return false;
    const ModePart* ph = reinterpret_cast<const ModePart*>(b.parts_host);
    ModePart m{0, 0, ~0ull};
    for (unsigned i = 0; i < blocks; ++i) {
        const ModePart p = ph[i];
        if (p.best != m.best) m = p.best > m.best ? p : m;
        else m = ModePart{m.best, m.ties + p.ties, m.first < p.first ? m.first : p.first};
    }
    *ties = m.ties;
    uint32_t r32 = 0;                                           // the entry's first row in row order
    if (!hip_ok(hipMemcpyAsync(b.parts_host, b.v2 + m.first, 4, hipMemcpyDeviceToHost, c->stream), "read row") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return false;
    __builtin_memcpy(&r32, b.parts_host, 4);
    *row = r32;
    return true;
}

bool segagg_unique(const uint64_t* k, uint64_t n, uint64_t* out, uint64_t* nout, StreamCtx* c) {
    *nout = 0;
    if (n == 0) return true;
    UniqueSrc src{k, out};
    return run_compact(src, n, c, nout);
}

uint64_t order_key(uint64_t bits, int dtype) {
    if (dtype != DT_F64) return bits ^ 0x8000000000000000ull;
    if (bits == 0x8000000000000000ull) bits = 0;
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ull);
}

uint64_t order_unkey(uint64_t key, int dtype) {
    if (dtype != DT_F64) return key ^ 0x8000000000000000ull;
    return (key >> 63) ? (key & 0x7FFFFFFFFFFFFFFFull) : ~key;
}

}  // namespace vx
