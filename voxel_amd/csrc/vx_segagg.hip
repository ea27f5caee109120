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

#include <cmath>

#include "vx_compact.hpp"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"
#include "vx_segagg.hpp"

#pragma clang fp contract(off)

namespace vx {

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

namespace {

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

template <typename T>
__device__ __forceinline__ bool expr_pass(int cmp, T x, T t) {
    switch (cmp) {                            // VFilterImpl predicates (engine.hpp:2896-2924)
    case 0: return x == t;
    case 1: return x != t;
    case 2: return x < t;
    case 3: return x <= t;
    case 4: return x > t;
    case 5: return x >= t;
    default: return false;
    }
}

// A lane's 64 bits, and a lane from 64 bits (the logical words act on the
// bit pattern of either lane type).
template <typename T>
__device__ __forceinline__ uint64_t bits_of(T x) {
    uint64_t u;
    __builtin_memcpy(&u, &x, 8);
    return u;
}
template <typename T>
__device__ __forceinline__ T lane_of(uint64_t u) {
    T x;
    __builtin_memcpy(&x, &u, 8);
    return x;
}
// (int) of a double as the reference's x86-64 build converts it (cvttsd2si):
// NaN and values outside int32's range give INT32_MIN; the f64 shift words
// scale by 2^(int)b and 2^-(int)b, and negating INT32_MIN keeps INT32_MIN.
__device__ __forceinline__ int cvtt_i32(double b) {
    return (b != b || b >= 2147483648.0 || b <= -2147483649.0) ? INT32_MIN : (int)b;
}
__device__ __forceinline__ int neg_i32(int k) { return k == INT32_MIN ? INT32_MIN : -k; }

// The NaN an f64 operation of the reference's x86-64 build leaves in a lane:
// the first NaN operand, quieted, or the default NaN (sign bit set) when the
// operation itself is invalid (inf - inf, 0 * inf, fmod(x, 0), sqrt(-1)).
// The GPU's default NaN has the sign bit clear, and the lane's bits reach
// the CMP-select accumulators, which order registers as signed integers.
__device__ __forceinline__ double x86_nan(double r, double x, double y, double z) {
    if (r == r) return r;
    const uint64_t b = x != x ? bits_of(x) : y != y ? bits_of(y) : z != z ? bits_of(z) : 0xFFF8000000000000ull;
    return lane_of<double>(b | 0x0008000000000000ull);
}
__device__ __forceinline__ double x86_nan(double r, double x, double y) { return x86_nan(r, x, y, y); }

// A lane of the row's group: the value the group's thread `src` (0..3)
// holds (the four rows of a group sit in four consecutive threads of one
// wave, all of them active).
template <typename T>
__device__ __forceinline__ T quad_lane(T v, int src) {
    const int l = (int)(threadIdx.x & 63);
    const int from = (l & ~3) + src;
    const uint64_t u = bits_of(v);
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)u, from, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(u >> 32), from, 64);
    return lane_of<T>(((uint64_t)hi << 32) | lo);
}

// Engine<T>::IntPow (engine.hpp:443-454): square-and-multiply in wrapping
// i64, 0 for a negative exponent.
__device__ __forceinline__ int64_t int_pow(int64_t b, int64_t e) {
    if (e < 0) return 0;
    uint64_t r = 1, bb = (uint64_t)b;
    while (e) {
        if (e & 1) r *= bb;
        bb *= bb;
        e >>= 1;
    }
    return (int64_t)r;
}

// Row-wise evaluation of the body's vector words (engine.hpp:2221-2466,
// 2473-2590, 2671-2865, 1090-1110): an arithmetic lane depends on the same
// lane of its operands only; the lane words (1987-2006, 2124-2217,
// 2926-2936) read lanes of the same group through quad_lane.  Each row
// keeps its lane's bit of the eight mask registers the filters and
// compares write and VBLEND reads (the lowering admits only masks written
// earlier in the same iteration).
template <typename T>
__global__ __launch_bounds__(256) void k_expr_eval(const T* __restrict__ col, uint64_t count, uint64_t off,
                                                   uint64_t n, ExprProg prog, T* __restrict__ out,
                                                   int* __restrict__ trap) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;     // a multiple of 4: rows keep their lane
    const int lane = (int)(threadIdx.x & 3);                       // row off + i is lane i mod 4 of its group
    bool trapped = false;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        T v[16];
#pragma unroll
        for (int r = 0; r < 16; ++r) v[r] = T(0);
        v[prog.vload] = off + i < count ? col[off + i] : T(0);
        uint32_t msk = 0;                                          // this lane's bit of masks 0..7
        for (int k = 0; k < prog.nops; ++k) {
            const ExprOp& o = prog.ops[k];
            const T x = v[o.a], y = v[o.b];
            T s;
            __builtin_memcpy(&s, &o.s, sizeof(T));
            T r;
            switch (o.op) {
            case 0x7D: r = quad_lane(x, lane ^ 3); break;                                       // VREVERSE
            case 0x7E: r = quad_lane(x, (lane - (int)(o.imm & 3)) & 3); break;                  // VROTATE
            case 0x7F: {                                                                        // VSLIDE
                const uint64_t si = (uint64_t)lane + o.s;                 // size_t arithmetic (wraps)
                const T t = quad_lane(x, (int)(si & 3));
                r = si < 4 ? t : T(0);
                break;
            }
            case 0x7C: {                                                                        // VSHUFFLE
                const int m = o.imm & 3, h = lane & 1;
                const int j = m == 0 ? 2 * h : m == 1 ? 2 * h + 1 : m == 2 ? h : h + 2;
                const T ta = quad_lane(x, j), tb = quad_lane(y, j);
                r = lane < 2 ? ta : tb;
                break;
            }
            case 0x7B: {                                                                        // VPERMUTE
                const uint64_t ib = bits_of(quad_lane(y, 0));             // index bytes: the first lane's
                r = quad_lane(x, (int)((ib >> (8 * lane)) & 0xFF) & 3);
                break;
            }
            case 0x7A: r = lane == (int)(o.imm & 3) ? s : x; break;                             // VINSERT
            case 0xC7: r = ((msk >> (o.imm & 7)) & 1) ? y : x; break;                           // VBLEND
            case 0x72: {                                                                        // VGATHER
                // i32 index `lane` of the register's bytes: half (lane & 1) of lane (lane >> 1)
                const uint64_t ib = bits_of(quad_lane(x, lane >> 1));
                const uint64_t gi = (uint64_t)(int64_t)(int32_t)(uint32_t)(ib >> (32 * (lane & 1)));
                const T* const gcol = static_cast<const T*>(prog.gcol[o.cmp & 1]);
                r = (gcol && gi < prog.gcount[o.cmp & 1]) ? gcol[gi] : T(0);
                break;
            }
            case 0x8F:                                                                          // VPOW
                if constexpr (F) r = x86_nan(::pow((double)x, (double)y), x, y);   // the device libm pow
                else r = (T)int_pow((int64_t)x, (int64_t)y);
                break;
            case 0x80:                                                                          // VADD
                if constexpr (F) r = x86_nan(x + y, x, y);
                else r = (T)((uint64_t)x + (uint64_t)y);
                break;
            case 0x81:                                                                          // VSUB
                if constexpr (F) r = x86_nan(x - y, x, y);
                else r = (T)((uint64_t)x - (uint64_t)y);
                break;
            case 0x82:                                                                          // VMUL
                if constexpr (F) r = x86_nan(x * y, x, y);
                else r = (T)((uint64_t)x * (uint64_t)y);
                break;
            case 0x83:                                                                          // VDIV
                if constexpr (F) r = y == T(0) ? (T)__builtin_nan("") : x86_nan(x / y, x, y);
                else {
                    trapped |= y == T(-1) && x == (T)INT64_MIN;                                   // x86 idiv traps
                    r = y == T(0) ? (T)INT64_MAX : y == T(-1) ? (T)(0 - (uint64_t)x) : x / y;
                }
                break;
            case 0x84:                                                                          // VMOD
                if constexpr (F) r = x86_nan(__builtin_fmod(x, y), x, y);
                else {
                    trapped |= y == T(-1) && x == (T)INT64_MIN;
                    r = (y == T(0) || y == T(-1)) ? T(0) : x % y;
                }
                break;
            case 0x85: r = lane_of<T>(F ? bits_of(x) ^ (1ull << 63) : 0 - bits_of(x)); break;   // VNEG
            case 0x89:                                                                          // VAVG
                if constexpr (F) r = x86_nan((x + y) * T(0.5), x, y);
                else r = (x >> 1) + (y >> 1) + ((x & 1) + (y & 1)) / 2;
                break;
            case 0x8A:                                                                          // VFMA, c = s
                if constexpr (F) r = x86_nan(__builtin_fma(x, y, s), x, y, s);
                else r = (T)((uint64_t)x * (uint64_t)y + (uint64_t)s);
                break;
            case 0x8B:                                                                          // VFMS
                if constexpr (F) r = x86_nan(x86_nan(x * y, x, y) - s, x86_nan(x * y, x, y), s);
                else r = (T)((uint64_t)x * (uint64_t)y - (uint64_t)s);
                break;
            case 0x8C:                                                                          // VSQRT
                if constexpr (F) r = x86_nan(__builtin_sqrt(x), x, x);
                else r = (double)x < 0 ? T(0) : (T)__builtin_sqrt((double)x);
                break;
            case 0x8D:                                                                          // VRSQRT
                if constexpr (F) r = x86_nan(T(1) / __builtin_sqrt(x), x, x);
                else r = (double)x <= 0 ? T(0) : (T)(1.0 / __builtin_sqrt((double)x));
                break;
            case 0x8E:                                                                          // VRCP
                if constexpr (F) r = x86_nan(T(1) / x, x, x);
                else r = x == T(0) ? T(0) : (T)(1.0 / (double)x);
                break;
            case 0x86:                                                                          // VABS
                if constexpr (F) r = lane_of<T>(bits_of(x) & ~(1ull << 63));
                else r = x < T(0) ? (T)(0 - (uint64_t)x) : x;
                break;
            case 0x87: r = F ? ((x < y || x != x) ? x : y) : (x < y ? x : y); break;             // VMIN
            case 0x88: r = F ? ((x > y || y != y) ? x : y) : (x > y ? x : y); break;             // VMAX
            case 0x90:                                                                          // VSADD
                if constexpr (F) r = x86_nan(x + s, x, s);
                else r = (T)((uint64_t)x + (uint64_t)s);
                break;
            case 0x91:                                                                          // VSSUB
                if constexpr (F) r = x86_nan(x - s, x, s);
                else r = (T)((uint64_t)x - (uint64_t)s);
                break;
            case 0x92:                                                                          // VSMUL
                if constexpr (F) r = x86_nan(x * s, x, s);
                else r = (T)((uint64_t)x * (uint64_t)s);
                break;
            case 0x93:                                                                          // VSDIV
                if constexpr (F) r = s == T(0) ? (T)__builtin_nan("") : x86_nan(x / s, x, s);
                else {
                    trapped |= s == T(-1) && x == (T)INT64_MIN;
                    r = s == T(0) ? T(0) : s == T(-1) ? (T)(0 - (uint64_t)x) : x / s;
                }
                break;
            case 0x94:                                                                          // VSMOD
                if constexpr (F) r = x86_nan(__builtin_fmod(x, s), x, s);
                else {
                    trapped |= s == T(-1) && x == (T)INT64_MIN;
                    r = (s == T(0) || s == T(-1)) ? T(0) : x % s;
                }
                break;
            case 0x95: r = F ? ((x < s || x != x) ? x : s) : (x < s ? x : s); break;             // VSMIN
            case 0x96: r = F ? ((x > s || s != s) ? x : s) : (x > s ? x : s); break;             // VSMAX
            case 0x78: r = s; break;                                                             // VSPLAT
            case 0xA0: case 0xA1: case 0xA2: case 0xA3: case 0xA4: case 0xA5: {                   // VCMPxx: mask 0
                const bool p = expr_pass<T>(o.op - 0xA0, x, y);
                r = p ? lane_of<T>(~0ull) : T(0);
                msk = (msk & ~1u) | (p ? 1u : 0u);
                break;
            }
            case 0xA6:                                                                          // VCMPNULL
            case 0xA7: {                                                                        // VCMPNOTNULL
                const bool nul = F ? x != x : x == T(0);
                const bool p = o.op == 0xA6 ? nul : !nul;
                r = p ? lane_of<T>(~0ull) : T(0);
                msk = (msk & ~(1u << (o.imm & 7))) | ((p ? 1u : 0u) << (o.imm & 7));
                break;
            }
            case 0xB0: r = lane_of<T>(bits_of(x) & bits_of(y)); break;                            // VAND
            case 0xB1: r = lane_of<T>(bits_of(x) | bits_of(y)); break;                            // VOR
            case 0xB2: r = lane_of<T>(bits_of(x) ^ bits_of(y)); break;                            // VXOR
            case 0xB3: r = lane_of<T>(~bits_of(x)); break;                                        // VNOT
            case 0xB4: r = lane_of<T>(~bits_of(x) & bits_of(y)); break;                           // VANDN
            case 0xB5:                                                                          // VSHL
                if constexpr (F) r = x86_nan(__builtin_ldexp(x, cvtt_i32(y)), x, x);
                else r = (T)(bits_of(x) << (bits_of(y) & 63));
                break;
            case 0xB6:                                                                          // VSHR
            case 0xB7:                                                                          // VSAR
                if constexpr (F) r = x86_nan(__builtin_ldexp(x, neg_i32(cvtt_i32(y))), x, x);
                else r = o.op == 0xB6 ? (T)(bits_of(x) >> (bits_of(y) & 63)) : (T)(x >> (bits_of(y) & 63));
                break;
            default: {                                                                          // VFILTER_*
                const bool p = expr_pass<T>(o.cmp, x, s);
                r = p ? x : T(0);
                msk = (msk & ~(1u << (o.imm & 7))) | ((p ? 1u : 0u) << (o.imm & 7));
                break;
            }
            }
            v[o.dst] = r;
        }
        out[i] = v[prog.target];
    }
    if (trapped) *trap = 1;
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
