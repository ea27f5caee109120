// vx_expr_ops.hpp -- one row of a Tier-0 loop body's vector words on the GPU
// (engine.hpp:1987-2006, 2124-2217, 2221-2466, 2473-2590, 2671-2865,
// 2926-2936, 1090-1110): apply_op<T, OP> for every word, with the helpers it
// needs, and the per-thread sum / count accumulator of the fused reductions.
// Compiled into the library's interpreter kernels (vx_segagg.hip) and handed,
// as text, to the run-time compiler for programs compiled whole (vx_jit.cpp),
// so it uses no header but the ABI one.
#pragma once
#include "vx_expr_abi.hpp"

// The words round each operation on its own (no FMA contraction), as the
// reference's scalar loops do.
#pragma clang fp contract(off)

namespace vx {
namespace xo {

constexpr int64_t kI64Min = (int64_t)0x8000000000000000ull;
constexpr int64_t kI64Max = (int64_t)0x7FFFFFFFFFFFFFFFull;
constexpr int kI32Min = (int)0x80000000u;

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
// NaN and values outside int32's range give kI32Min; the f64 shift words
// scale by 2^(int)b and 2^-(int)b, and negating kI32Min keeps kI32Min.
__device__ __forceinline__ int cvtt_i32(double b) {
    return (b != b || b >= 2147483648.0 || b <= -2147483649.0) ? kI32Min : (int)b;
}
__device__ __forceinline__ int neg_i32(int k) { return k == kI32Min ? kI32Min : -k; }

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
//
// apply_op<T, OP>: the word OP on one row (x = R[a], y = R[b], s = the
// word's scalar operand); msk is the row's bits of mask registers 0..7.
// lane is the row's lane in its group: a lane word needs the group's four
// threads active together.  OP not listed: a VFILTER_* (its kCmp in o.cmp).
template <typename T, int OP>
__device__ __forceinline__ T apply_op(const ExprOp& o, const ExprProg& prog, T x, T y, T s, int lane, uint32_t& msk,
                                      bool& trapped) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    T r;
    switch (OP) {
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
            trapped |= y == T(-1) && x == (T)kI64Min;                                   // x86 idiv traps
            r = y == T(0) ? (T)kI64Max : y == T(-1) ? (T)(0 - (uint64_t)x) : x / y;
        }
        break;
    case 0x84:                                                                          // VMOD
        if constexpr (F) r = x86_nan(::fmod(x, y), x, y);        // the device libm fmod: exact
        else {
            trapped |= y == T(-1) && x == (T)kI64Min;
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
            trapped |= s == T(-1) && x == (T)kI64Min;
            r = s == T(0) ? T(0) : s == T(-1) ? (T)(0 - (uint64_t)x) : x / s;
        }
        break;
    case 0x94:                                                                          // VSMOD
        if constexpr (F) r = x86_nan(::fmod(x, s), x, s);        // (a compiler's own remainder expansion may not be)
        else {
            trapped |= s == T(-1) && x == (T)kI64Min;
            r = (s == T(0) || s == T(-1)) ? T(0) : x % s;
        }
        break;
    case 0x95: r = F ? ((x < s || x != x) ? x : s) : (x < s ? x : s); break;             // VSMIN
    case 0x96: r = F ? ((x > s || s != s) ? x : s) : (x > s ? x : s); break;             // VSMAX
    case 0x78: r = s; break;                                                             // VSPLAT
    case 0xA0: case 0xA1: case 0xA2: case 0xA3: case 0xA4: case 0xA5: {                   // VCMPxx: mask 0
        const bool p = expr_pass<T>(OP - 0xA0, x, y);
        r = p ? lane_of<T>(~0ull) : T(0);
        msk = (msk & ~1u) | (p ? 1u : 0u);
        break;
    }
    case 0xA6:                                                                          // VCMPNULL
    case 0xA7: {                                                                        // VCMPNOTNULL
        const bool nul = F ? x != x : x == T(0);
        const bool p = OP == 0xA6 ? nul : !nul;
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
        else r = OP == 0xB6 ? (T)(bits_of(x) >> (bits_of(y) & 63)) : (T)(x >> (bits_of(y) & 63));
        break;
    default: {                                                                          // VFILTER_*
        const bool p = expr_pass<T>(o.cmp, x, s);
        r = p ? x : T(0);
        msk = (msk & ~(1u << (o.imm & 7))) | ((p ? 1u : 0u) << (o.imm & 7));
        break;
    }
    }
    return r;
}

// Per-thread VSUM / VCOUNT accumulator of a fused reduction: a compensated
// f64 pair or a wrapping i64 sum, and a count of lanes that are non-zero and
// not NaN.
template <typename T>
struct XAcc {
    T s, c;
    uint64_t n;
};
template <typename T>
__device__ __forceinline__ void xacc_init(XAcc<T>& a) {
    a.s = T(0); a.c = T(0); a.n = 0;
}
template <typename T>
__device__ __forceinline__ void xacc_sum(XAcc<T>& a, T y) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    if constexpr (F) {                                    // Knuth TwoSum: s + c tracks the exact sum
        const T t = a.s + y, bp = t - a.s;
        a.c += (a.s - (t - bp)) + (y - bp);
        a.s = t;
    } else {
        a.s = (T)((uint64_t)a.s + (uint64_t)y);
    }
}
template <typename T>
__device__ __forceinline__ void xacc_count(XAcc<T>& a, T y) {
    a.n += (y == y && y != T(0)) ? 1u : 0u;
}
template <typename T>
__device__ __forceinline__ void xacc_merge(XAcc<T>& a, const XAcc<T>& b) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    if constexpr (F) {
        T cc = a.c + b.c;
        const T t = a.s + b.s, bp = t - a.s;
        cc += (a.s - (t - bp)) + (b.s - bp);
        a.s = t;
        a.c = cc;
    } else {
        a.s = (T)((uint64_t)a.s + (uint64_t)b.s);
    }
    a.n += b.n;
}
// The 256-thread workgroup's total (waves merged in order) written by thread
// 0 as a scan partial (eight words): s, c, count, then MIN / MAX left at
// "none".
template <typename T>
__device__ __forceinline__ void xacc_block_store(XAcc<T> a, XAcc<T>* lds, uint64_t* part) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    for (int m = 32; m > 0; m >>= 1) {
        XAcc<T> o;
        o.s = __shfl_xor(a.s, m, 64);
        o.c = __shfl_xor(a.c, m, 64);
        o.n = __shfl_xor(a.n, m, 64);
        xacc_merge(a, o);
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) xacc_merge(a, lds[w]);
        part[0] = bits_of(a.s);
        part[1] = bits_of(a.c);
        part[2] = a.n;
        part[3] = F ? 0x7FF0000000000000ull : (uint64_t)kI64Max;   // MIN: +inf / INT64_MAX
        part[4] = F ? 0xFFF0000000000000ull : (uint64_t)kI64Min;   // MAX: -inf / INT64_MIN
    }
    __syncthreads();
}
// VRED_MIN (red 2) / VRED_MAX (red 3) of one group's four lanes in lane
// order, as the reference reduces them: a lane replaces the running value
// when it is below (above) it, a NaN lane always replaces the running MIN, a
// NaN running MAX is always replaced.  All four threads of the group call it
// and get the group's value bits.
template <typename T>
__device__ __forceinline__ uint64_t group_minmax(T y, int red) {
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    const T l[4] = {quad_lane(y, 0), quad_lane(y, 1), quad_lane(y, 2), quad_lane(y, 3)};
    T m = l[0];
#pragma unroll
    for (int i = 1; i < 4; ++i) {
        bool take;
        if (red == 2) take = l[i] < m || (F && l[i] != l[i]);
        else take = l[i] > m || (F && m != m);
        m = take ? l[i] : m;
    }
    return bits_of(m);
}

// Extremes of the group values of a VRED_MIN / VRED_MAX reduction in the
// signed and the unsigned order of their bits (the CMP-select accumulator
// compares register bits as signed integers; whether the wrapping difference
// is the signed order is decided on the host from these four).
struct GAccX {
    int64_t smin, smax;
    uint64_t umin, umax;
};
__device__ __forceinline__ void gacc_init(GAccX& a) {
    a.smin = kI64Max; a.smax = kI64Min; a.umin = ~0ull; a.umax = 0ull;
}
__device__ __forceinline__ void gacc_add(GAccX& a, uint64_t v) {
    const int64_t s = (int64_t)v;
    a.smin = s < a.smin ? s : a.smin; a.smax = s > a.smax ? s : a.smax;
    a.umin = v < a.umin ? v : a.umin; a.umax = v > a.umax ? v : a.umax;
}
__device__ __forceinline__ void gacc_merge(GAccX& a, const GAccX& b) {
    a.smin = b.smin < a.smin ? b.smin : a.smin; a.smax = b.smax > a.smax ? b.smax : a.smax;
    a.umin = b.umin < a.umin ? b.umin : a.umin; a.umax = b.umax > a.umax ? b.umax : a.umax;
}
// The workgroup's extremes written by thread 0 into a scan-partial slot
// (eight words: smin, smax, umin, umax, then unused).
__device__ __forceinline__ void gacc_block_store(GAccX a, GAccX* lds, uint64_t* part) {
    for (int m = 32; m > 0; m >>= 1) {
        GAccX o;
        o.smin = __shfl_xor(a.smin, m, 64); o.smax = __shfl_xor(a.smax, m, 64);
        o.umin = __shfl_xor(a.umin, m, 64); o.umax = __shfl_xor(a.umax, m, 64);
        gacc_merge(a, o);
    }
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 4; ++w) gacc_merge(a, lds[w]);
        part[0] = (uint64_t)a.smin;
        part[1] = (uint64_t)a.smax;
        part[2] = a.umin;
        part[3] = a.umax;
    }
    __syncthreads();
}

// The loop's partial register from the last group's four lanes (one
// reduction, as the reference reduces one group): VSUM 0 + l0 + l1 + l2 + l3
// in lane order, VCOUNT lanes non-zero and not NaN, VRED_MIN / VRED_MAX as
// group_minmax.  All four threads of the group call it; the group's lane 0
// writes *out.
template <typename T>
__device__ __forceinline__ void group_partial(T y, int red, int lane, uint64_t* out) {
    if (red >= 2) {
        const uint64_t g = group_minmax(y, red);
        if (lane == 0) *out = g;
        return;
    }
    const T l0 = quad_lane(y, 0), l1 = quad_lane(y, 1), l2 = quad_lane(y, 2), l3 = quad_lane(y, 3);
    if (lane != 0) return;
    constexpr bool F = sizeof(T) == 8 && (T)0.5 != (T)0;
    uint64_t g;
    if (red == 0) {
        if constexpr (F) {
            T sm = T(0);
            sm = sm + l0; sm = sm + l1; sm = sm + l2; sm = sm + l3;
            g = bits_of(sm);
        } else {
            g = (uint64_t)l0 + (uint64_t)l1 + (uint64_t)l2 + (uint64_t)l3;
        }
    } else {
        g = (uint64_t)((l0 == l0 && l0 != T(0)) + (l1 == l1 && l1 != T(0)) + (l2 == l2 && l2 != T(0)) +
                       (l3 == l3 && l3 != T(0)));
    }
    *out = g;
}

}  // namespace xo
}  // namespace vx
