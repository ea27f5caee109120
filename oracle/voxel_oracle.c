/*
 * voxel_oracle.c -- CPU restatement of the VoxelVM hot-path semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: it may be loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing else.  The product library (voxel_amd/libvoxel_c.so) never links,
 * loads or calls it.
 *
 * Every function restates one piece of the reference algorithm from its
 * source, cited as file:line relative to the reference repository root.
 * Functions are plain C99 (+ unsigned __int128 from gcc) so that they are easy
 * to read next to the reference.
 *
 * Pinning (what checks each function against something other than itself):
 *   vo_mt64_*            first outputs of std::mt19937_64(42) quoted by the
 *                        reference (docs/Benchmarks.md:7 ground truth 500,720)
 *   vo_acc_*             exact: Python Fraction arithmetic on small inputs,
 *                        cancellation cases, math.fsum
 *   vo_fastpath_*        reference golden values (test_suite.cpp:254-321,
 *                        c_test.cpp:8-36) + oracle/_ref Engine<T>::Run
 *   vo_tier0_*           voxel_query golden values (sum/count/min/max bits)
 *                        + oracle/_ref Engine<f64> Tier-0
 *   vo_jit_f64           oracle/_ref JIT (jit_run) bit patterns
 *   vo_run_parallel_*    oracle/_ref RunParallel bit pattern
 *   vo_bitmap_*          test_suite.cpp:410-427 bits {2,3,6,7}, numpy, _ref
 *   vo_radix_*           std-library stable argsort (numpy kind="stable"),
 *                        golden keys {5,3,5,1,3,2^64-1,0}, _ref RadixSort
 *   vo_mergesort_*       numpy stable argsort, c_test idx [1 3 0 2 4], _ref
 *   vo_hashjoin_*        brute-force dict join, golden pairs, _ref HashJoin
 *   vo_groupby_*         brute-force per-key loops, golden groups, _ref
 */
#include <math.h>
#include <stdint.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

typedef unsigned __int128 u128;

/* ------------------------------------------------------------------------ */
/* Random generators used to make the reference's own example inputs.      */
/* ------------------------------------------------------------------------ */

/* mt19937_64 as published by Matsumoto & Nishimura (2004); this is the
 * std::mt19937_64 used by examples/validate.cpp:17-29, jit_test.cpp:15-30. */
typedef struct { uint64_t mt[312]; int idx; } vo_mt64;

void vo_mt64_seed(vo_mt64* s, uint64_t seed) {
    s->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        s->mt[i] = 6364136223846793005ULL * (s->mt[i - 1] ^ (s->mt[i - 1] >> 62)) + (uint64_t)i;
    s->idx = 312;
}

static void vo_mt64_twist(vo_mt64* s) {
    const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
    for (int i = 0; i < 312; ++i) {
        uint64_t x = (s->mt[i] & upper) | (s->mt[(i + 1) % 312] & lower);
        uint64_t xa = x >> 1;
        if (x & 1) xa ^= 0xB5026F5AA96619E9ULL;
        s->mt[i] = s->mt[(i + 156) % 312] ^ xa;
    }
    s->idx = 0;
}

uint64_t vo_mt64_next(vo_mt64* s) {
    if (s->idx >= 312) vo_mt64_twist(s);
    uint64_t x = s->mt[s->idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= x >> 43;
    return x;
}

/* uniform_real_distribution<double>(lo, hi) over a 64-bit engine as the
 * GNU C++ library computes it: one draw, converted to double and divided by
 * 2^64 (generate_canonical<double,53>), clamped below 1, then lo + u*(hi-lo). */
void vo_mt64_uniform_f64(uint64_t seed, size_t n, double lo, double hi, double* out) {
    vo_mt64 s; vo_mt64_seed(&s, seed);
    for (size_t i = 0; i < n; ++i) {
        double u = (double)vo_mt64_next(&s) / 18446744073709551616.0;
        if (u >= 1.0) u = nextafter(1.0, 0.0);
        out[i] = u * (hi - lo) + lo;
    }
}

/* uniform_int_distribution<int64_t/int32_t>(lo, hi) over a 64-bit engine:
 * Lemire's nearly-divisionless bounded draw (ACM TOMACS 29(1), 2019) with a
 * 128-bit product, which is what the GNU C++ library does for a 64-bit engine. */
static uint64_t vo_lemire64(vo_mt64* s, uint64_t range) {
    u128 prod = (u128)vo_mt64_next(s) * range;
    uint64_t low = (uint64_t)prod;
    if (low < range) {
        uint64_t thresh = (0 - range) % range;
        while (low < thresh) {
            prod = (u128)vo_mt64_next(s) * range;
            low = (uint64_t)prod;
        }
    }
    return (uint64_t)(prod >> 64);
}

void vo_mt64_uniform_i64(uint64_t seed, size_t n, int64_t lo, int64_t hi, int64_t* out) {
    vo_mt64 s; vo_mt64_seed(&s, seed);
    uint64_t urange = (uint64_t)hi - (uint64_t)lo;
    for (size_t i = 0; i < n; ++i) {
        uint64_t r = (urange == UINT64_MAX) ? vo_mt64_next(&s) : vo_lemire64(&s, urange + 1);
        out[i] = (int64_t)(r + (uint64_t)lo);
    }
}

void vo_mt64_uniform_i32(uint64_t seed, size_t n, int32_t lo, int32_t hi, int32_t* out) {
    vo_mt64 s; vo_mt64_seed(&s, seed);
    uint64_t urange = (uint64_t)(uint32_t)hi - (uint64_t)(uint32_t)lo;
    urange &= 0xFFFFFFFFULL;
    for (size_t i = 0; i < n; ++i) {
        uint64_t r = vo_lemire64(&s, urange + 1);
        out[i] = (int32_t)(uint32_t)(r + (uint64_t)(uint32_t)lo);
    }
}

/* ------------------------------------------------------------------------ */
/* Exact long accumulator for float sums (Kulisch-style fixed point).       */
/* A double is m * 2^(e), |m| < 2^53; bit position p = e + 1074 >= 0.       */
/* Digits are 32-bit values held in int64 with lazy carry propagation.      */
/* ------------------------------------------------------------------------ */
#define VO_ACC_DIGITS 72
typedef struct { int64_t d[VO_ACC_DIGITS]; uint32_t pending; int32_t nonfinite; } vo_acc;

void vo_acc_init(vo_acc* a) { memset(a, 0, sizeof(*a)); }

static void vo_acc_norm(vo_acc* a) {
    for (int i = 0; i < VO_ACC_DIGITS - 1; ++i) {
        int64_t carry = a->d[i] >> 32;           /* arithmetic shift: floor division */
        a->d[i] -= carry * 4294967296LL;
        a->d[i + 1] += carry;
    }
    a->pending = 0;
}

void vo_acc_add(vo_acc* a, double x) {
    if (x == 0.0) return;
    if (!isfinite(x)) { a->nonfinite = 1; return; }
    int e;
    double m = frexp(x, &e);                      /* x = m * 2^e, 0.5 <= |m| < 1 */
    int64_t mi = (int64_t)ldexp(m, 53);           /* exact: x = mi * 2^(e-53)     */
    int pos = e - 53 + 1074;
    uint64_t mag = (uint64_t)(mi < 0 ? -mi : mi);
    if (pos < 0) { mag >>= -pos; pos = 0; }       /* subnormal: trailing zeros only */
    u128 v = (u128)mag << (pos & 31);
    int dig = pos >> 5;
    for (int k = 0; k < 3; ++k) {
        int64_t chunk = (int64_t)(uint32_t)(v >> (32 * k));
        a->d[dig + k] += (mi < 0) ? -chunk : chunk;
    }
    if (++a->pending >= (1u << 29)) vo_acc_norm(a);
}

/* Canonical digits: d[i] in [0, 2^32) for i < top, top digit signed.
 * Value = sum d[i] * 2^(32 i) * 2^-1074. */
void vo_acc_export(vo_acc* a, int64_t* out) {
    vo_acc_norm(a);
    memcpy(out, a->d, sizeof(a->d));
}

int vo_acc_digits(void) { return VO_ACC_DIGITS; }

/* ------------------------------------------------------------------------ */
/* Comparison predicates: kCmp* codes of bytecode/opcodes.hpp:415-420       */
/* (EQ 0, NE 1, LT 2, LE 3, GT 4, GE 5); 6 = no predicate.                 */
/* C++ relational operators: NaN fails all but != (engine.hpp:124-131).    */
/* ------------------------------------------------------------------------ */
#define VO_PRED(op, x, t) \
    ((op) == 0 ? ((x) == (t)) : (op) == 1 ? ((x) != (t)) : (op) == 2 ? ((x) < (t)) : \
     (op) == 3 ? ((x) <= (t)) : (op) == 4 ? ((x) > (t)) : (op) == 5 ? ((x) >= (t)) : 1)

/* Relational (SQL-style) filter+aggregate over one or two columns:
 * rows i with pred_a(a[i], ta) && pred_b(b[i], tb); COUNT, exact SUM(a),
 * MIN/MAX(a) over passing non-NaN values.  This is what
 * VectorFilter::Apply* + ColumnFilter::AndWith (ops/filter.hpp:42-288,
 * 540-547) select, aggregated without the lane/tail rules of the VM. */
#define VO_DEF_REL(NAME, T, ISF)                                                        \
void NAME(const T* a, const T* b, size_t n, int opa, T ta, int opb, T tb,              \
          uint64_t* count, int64_t* digits, T* mn, T* mx, int64_t* isum, int* any) {    \
    vo_acc acc; vo_acc_init(&acc);                                                      \
    uint64_t c = 0; int64_t s = 0; int have = 0; T lo = 0, hi = 0;                      \
    for (size_t i = 0; i < n; ++i) {                                                    \
        T x = a[i];                                                                     \
        if (!VO_PRED(opa, x, ta)) continue;                                             \
        if (b && !VO_PRED(opb, b[i], tb)) continue;                                     \
        ++c;                                                                            \
        if (ISF) { vo_acc_add(&acc, (double)x); if (x != x) continue; }                 \
        else s = (int64_t)((uint64_t)s + (uint64_t)(int64_t)x);                         \
        if (!have) { lo = hi = x; have = 1; }                                           \
        else { if (x < lo) lo = x; if (x > hi) hi = x; }                                \
    }                                                                                   \
    *count = c; if (digits) vo_acc_export(&acc, digits);                                \
    *mn = lo; *mx = hi; *isum = s; *any = have;                                         \
}
VO_DEF_REL(vo_rel_f64, double, 1)
VO_DEF_REL(vo_rel_f32, float, 1)
VO_DEF_REL(vo_rel_i64, int64_t, 0)
VO_DEF_REL(vo_rel_i32, int32_t, 0)

/* Exact sum of a plain array (no predicate). */
void vo_exact_sum_f64(const double* x, size_t n, int64_t* digits) {
    vo_acc acc; vo_acc_init(&acc);
    for (size_t i = 0; i < n; ++i) vo_acc_add(&acc, x[i]);
    vo_acc_export(&acc, digits);
}

/* ------------------------------------------------------------------------ */
/* Interpreter fast path: Engine<T>::RunFastPath, exec/engine.hpp:109-141.  */
/* filter opcode 0xC1..0xC6 (EQ NE LT LE GT GE), any other -> GT (line 130).*/
/* total = min(seg.Count, (sz)(i64)R[total]); groups of kL = 32/sizeof(T)   */
/* while off + kL <= total; lanes summed ((v0+v1)+v2)+... in T; acc += s.   */
/* Integer T wraps (computed in unsigned arithmetic).  Returns final off.   */
/* ------------------------------------------------------------------------ */
static int vo_fast_pass_code(uint8_t fop) {
    switch (fop) { case 0xC1: return 0; case 0xC2: return 1; case 0xC3: return 2;
                   case 0xC4: return 3; case 0xC5: return 4; case 0xC6: return 5; default: return 4; }
}

#define VO_DEF_FAST(NAME, T, UT, KL, ISF)                                                 \
uint64_t NAME(const T* data, uint64_t count, uint64_t reg_total, uint64_t reg_off,       \
              T thresh, T acc0, uint8_t fop, T* acc_out) {                               \
    uint64_t total = reg_total; if (count < total) total = count;                        \
    uint64_t off = reg_off;                                                              \
    int op = vo_fast_pass_code(fop);                                                     \
    T acc = acc0;                                                                        \
    while (off + KL <= total) {                                                          \
        T s = 0;                                                                         \
        for (int i = 0; i < KL; ++i) {                                                   \
            T x = data[off + i];                                                         \
            T v = VO_PRED(op, x, thresh) ? x : (T)0;                                     \
            if (ISF) s = s + v; else s = (T)(UT)((UT)s + (UT)v);                         \
        }                                                                                \
        if (ISF) acc = acc + s; else acc = (T)(UT)((UT)acc + (UT)s);                     \
        off += KL;                                                                       \
    }                                                                                    \
    *acc_out = acc;                                                                      \
    return off;                                                                          \
}
VO_DEF_FAST(vo_fastpath_f64, double, uint64_t, 4, 1)
VO_DEF_FAST(vo_fastpath_f32, float, uint32_t, 8, 1)
VO_DEF_FAST(vo_fastpath_i64, int64_t, uint64_t, 4, 0)
VO_DEF_FAST(vo_fastpath_i32, int32_t, uint32_t, 8, 0)

/* ------------------------------------------------------------------------ */
/* Tier-0 loop programs of the voxel_query shapes (bindings/voxel_query.py: */
/* 74-135,155-171) and the MOV-prefixed test programs, Engine<f64/i64>.     */
/* Handlers: VLOAD engine.hpp:1947-1969 (zero past Count and past the VLOAD */
/* count field), VFILTER 2870-2924 (dst = pass ? src : 0), VSUM 2980-2987,  */
/* VCOUNT 3097-3111 (lanes != 0 and not NaN), VRED_MIN/MAX 3071-3095,       */
/* ADDF 1394-1401 (f64 on register bits), ADD 1286-1292 (i64), CMP          */
/* 1603-1608 (flags of i64 wrapping difference), JLE/JGE 1758-1776, MOVR.   */
/* The loop runs until off == total (JNZ): only called with                 */
/* (total - off) a positive multiple of kL (otherwise it never ends).       */
/* red: 0 VSUM, 1 VCOUNT, 2 VRED_MIN, 3 VRED_MAX.                           */
/* accop: 0 ADDF, 1 ADD, 2 CMP/JLE/MOVR, 3 CMP/JGE/MOVR.                    */
/* Returns R[acc] bits; *part_out = last group's R[part] bits.              */
/* ------------------------------------------------------------------------ */
static uint64_t vo_f64bits(double d) { uint64_t u; memcpy(&u, &d, 8); return u; }
static double vo_bitsf64(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

static uint64_t vo_tier0_accum(uint64_t acc, uint64_t part, int accop) {
    if (accop == 0) return vo_f64bits(vo_bitsf64(acc) + vo_bitsf64(part));
    if (accop == 1) return acc + part;
    int64_t diff = (int64_t)(acc - part);                   /* CMP acc, part */
    int z = diff == 0, sgn = diff < 0;
    if (accop == 2) return (z || sgn) ? acc : part;         /* JLE +2 skips MOVR */
    return (z || !sgn) ? acc : part;                        /* JGE +2 skips MOVR */
}

uint64_t vo_tier0_f64(const double* data, uint64_t count, uint64_t off, uint64_t total,
                      int filt, double thresh, int vcount, int red, int accop,
                      uint64_t acc, uint64_t* part_out) {
    uint64_t part = 0;
    if (vcount == 0) vcount = 4;
    do {
        double v[4];
        for (int i = 0; i < 4; ++i) {
            uint64_t idx = off + (uint64_t)i;
            v[i] = (i < vcount && idx < count) ? data[idx] : 0.0;
            if (filt != 6) v[i] = VO_PRED(filt, v[i], thresh) ? v[i] : 0.0;
        }
        if (red == 0) { double s = 0.0; for (int i = 0; i < 4; ++i) s += v[i]; part = vo_f64bits(s); }
        else if (red == 1) { uint64_t c = 0; for (int i = 0; i < 4; ++i) if (!isnan(v[i]) && v[i] != 0.0) ++c; part = c; }
        else if (red == 2) { double m = v[0]; for (int i = 1; i < 4; ++i) if (v[i] < m || isnan(v[i])) m = v[i]; part = vo_f64bits(m); }
        else { double m = v[0]; for (int i = 1; i < 4; ++i) if (v[i] > m || isnan(m)) m = v[i]; part = vo_f64bits(m); }
        acc = vo_tier0_accum(acc, part, accop);
        off += 4;
    } while (off != total);
    *part_out = part;
    return acc;
}

uint64_t vo_tier0_i64(const int64_t* data, uint64_t count, uint64_t off, uint64_t total,
                      int filt, int64_t thresh, int vcount, int red, int accop,
                      uint64_t acc, uint64_t* part_out) {
    uint64_t part = 0;
    if (vcount == 0) vcount = 4;
    do {
        int64_t v[4];
        for (int i = 0; i < 4; ++i) {
            uint64_t idx = off + (uint64_t)i;
            v[i] = (i < vcount && idx < count) ? data[idx] : 0;
            if (filt != 6) v[i] = VO_PRED(filt, v[i], thresh) ? v[i] : 0;
        }
        if (red == 0) { uint64_t s = 0; for (int i = 0; i < 4; ++i) s += (uint64_t)v[i]; part = s; }
        else if (red == 1) { uint64_t c = 0; for (int i = 0; i < 4; ++i) if (v[i] != 0) ++c; part = c; }
        else if (red == 2) { int64_t m = v[0]; for (int i = 1; i < 4; ++i) if (v[i] < m) m = v[i]; part = (uint64_t)m; }
        else { int64_t m = v[0]; for (int i = 1; i < 4; ++i) if (v[i] > m) m = v[i]; part = (uint64_t)m; }
        acc = vo_tier0_accum(acc, part, accop);
        off += 4;
    } while (off != total);
    *part_out = part;
    return acc;
}

/* ------------------------------------------------------------------------ */
/* Tier-0 loops with several filters and reductions, Engine<f64/i64>: per   */
/* kL = 4 group, VLOAD as above, then filters in program order, filter j    */
/* reading the loaded vector (fsrc[j] < 0) or filter fsrc[j]'s output       */
/* (VFilterImpl, engine.hpp:1090-1110: lane = pred(lane, t) ? lane : 0;     */
/* fcmp 7 = never passes), then reduction block b of the loaded vector      */
/* (bsrc[b] < 0) or a filter's output, folded into acc[b] as above.        */
/* The handlers are the ones vo_tier0_* restates; registers are distinct.   */
/* acc[] in/out (register bits); part[] = last group's partials.            */
/* ------------------------------------------------------------------------ */
static uint64_t vo_red_f64(const double* v, int red) {
    if (red == 0) { double s = 0.0; for (int i = 0; i < 4; ++i) s += v[i]; return vo_f64bits(s); }
    if (red == 1) { uint64_t c = 0; for (int i = 0; i < 4; ++i) if (!isnan(v[i]) && v[i] != 0.0) ++c; return c; }
    double m = v[0];
    for (int i = 1; i < 4; ++i)
        if (red == 2 ? (v[i] < m || isnan(v[i])) : (v[i] > m || isnan(m))) m = v[i];
    return vo_f64bits(m);
}

static uint64_t vo_red_i64(const int64_t* v, int red) {
    if (red == 0) { uint64_t s = 0; for (int i = 0; i < 4; ++i) s += (uint64_t)v[i]; return s; }
    if (red == 1) { uint64_t c = 0; for (int i = 0; i < 4; ++i) if (v[i] != 0) ++c; return c; }
    int64_t m = v[0];
    for (int i = 1; i < 4; ++i)
        if (red == 2 ? v[i] < m : v[i] > m) m = v[i];
    return (uint64_t)m;
}

#define VO_DEF_TIER0_MULTI(NAME, T, RED)                                                    \
void NAME(const T* data, uint64_t count, uint64_t off, uint64_t total, int vcount,          \
          int nf, const int* fcmp, const T* fthr, const int* fsrc,                          \
          int nb, const int* bsrc, const int* red, const int* accop,                        \
          uint64_t* acc, uint64_t* part) {                                                   \
    T ld[4], fv[4][4];                                                                       \
    if (vcount == 0) vcount = 4;                                                             \
    do {                                                                                     \
        for (int i = 0; i < 4; ++i) {                                                        \
            uint64_t idx = off + (uint64_t)i;                                                \
            ld[i] = (i < vcount && idx < count) ? data[idx] : (T)0;                          \
        }                                                                                    \
        for (int j = 0; j < nf; ++j) {                                                       \
            const T* src = fsrc[j] < 0 ? ld : fv[fsrc[j]];                                   \
            for (int i = 0; i < 4; ++i)                                                      \
                fv[j][i] = (fcmp[j] != 7 && VO_PRED(fcmp[j], src[i], fthr[j])) ? src[i] : (T)0; \
        }                                                                                    \
        for (int b = 0; b < nb; ++b) {                                                       \
            const T* v = bsrc[b] < 0 ? ld : fv[bsrc[b]];                                     \
            part[b] = RED(v, red[b]);                                                        \
            acc[b] = vo_tier0_accum(acc[b], part[b], accop[b]);                              \
        }                                                                                    \
        off += 4;                                                                            \
    } while (off != total);                                                                  \
}
VO_DEF_TIER0_MULTI(vo_tier0_multi_f64, double, vo_red_f64)
VO_DEF_TIER0_MULTI(vo_tier0_multi_i64, int64_t, vo_red_i64)

/* ------------------------------------------------------------------------ */
/* Tier-0 loops with vector arithmetic (engine.hpp:2221-2450 VADD VSUB VMUL */
/* VDIV VMOD VNEG VABS VMIN VMAX VAVG VFMA VFMS VSQRT VRSQRT VRCP,          */
/* 2473-2590 VSADD VSSUB VSMUL VSDIV VSMOD VSMIN VSMAX; VFMA/VFMS take the  */
/* scalar register imm12 & 0xF as c; 2660-2866 VCMPxx / VCMP(NOT)NULL       */
/* (all-ones lanes), VAND VOR VXOR VNOT VANDN on lane bits, VSHL VSHR VSAR  */
/* (f64: ldexp by the lane's (int) value))                                  */
/* and filters, per kL = 4 group: VLOAD into register vload_dst (zero past  */
/* Count), then ops in program order on 16 vector registers; op word        */
/* k: code[5k..5k+4] = opcode, dst, a, b, filter kCmp; sval[k] = the scalar */
/* operand / threshold (register value as T).  Reduction block b reduces    */
/* register bvreg[b] and folds into acc[b] as in vo_tier0_*.  f64 division  */
/* by zero gives a quiet NaN (VDIV per lane, VSDIV for every lane), f64 MOD  */
/* is fmod.  Integers: VDIV by a zero lane gives INT64_MAX, VSDIV by zero 0, */
/* VMOD / VSMOD by zero 0 (engine.hpp:2263-2291, 2516-2549); x / -1 is the   */
/* wrapping negation here (the reference's x86 build traps on INT64_MIN/-1). */
/* ------------------------------------------------------------------------ */
/* lane bits of a vector register and x86's cvttsd2si (NaN and out-of-range
   give INT_MIN), as the reference's build computes (int) of a double */
static uint64_t vo_lb_f64(double x) { uint64_t u; memcpy(&u, &x, 8); return u; }
static double vo_bl_f64(uint64_t u) { double x; memcpy(&x, &u, 8); return x; }
static uint64_t vo_lb_i64(int64_t x) { return (uint64_t)x; }
static int64_t vo_bl_i64(uint64_t u) { return (int64_t)u; }
static int vo_cvtt(double b) {
    return (b != b || b >= 2147483648.0 || b <= -2147483649.0) ? INT32_MIN : (int)b;
}
static int vo_neg32(int k) { return k == INT32_MIN ? INT32_MIN : -k; }

#define VO_DEF_EXPR_LANE(NAME, T, ISF, LB, BL)                                              \
static T NAME(int op, int cmp, T x, T y, T s) {                                              \
    switch (op) {                                                                            \
    case 0x80: return ISF ? x + y : (T)((uint64_t)x + (uint64_t)y);                          \
    case 0x81: return ISF ? x - y : (T)((uint64_t)x - (uint64_t)y);                          \
    case 0x82: return ISF ? x * y : (T)((uint64_t)x * (uint64_t)y);                          \
    case 0x83: return ISF ? ((y == (T)0) ? (T)NAN : x / y)                                    \
                          : (y == (T)0 ? (T)INT64_MAX : y == (T)-1 ? (T)(0 - (uint64_t)x) : (T)((int64_t)x / (int64_t)y)); \
    case 0x84: return ISF ? (T)fmod((double)x, (double)y)                                    \
                          : (y == (T)0 || y == (T)-1 ? (T)0 : (T)((int64_t)x % (int64_t)y));  \
    case 0x85: return ISF ? -x : (T)(0 - (uint64_t)x);                                       \
    case 0x89: return ISF ? (x + y) * (T)0.5                                                 \
                          : (T)(((int64_t)x >> 1) + ((int64_t)y >> 1) + (((int64_t)x & 1) + ((int64_t)y & 1)) / 2); \
    case 0x8A: return ISF ? (T)fma((double)x, (double)y, (double)s)                          \
                          : (T)((uint64_t)x * (uint64_t)y + (uint64_t)s);                    \
    case 0x8B: return ISF ? x * y - s : (T)((uint64_t)x * (uint64_t)y - (uint64_t)s);         \
    case 0x8C: return ISF ? (T)sqrt((double)x) : ((double)x < 0 ? (T)0 : (T)sqrt((double)x)); \
    case 0x8D: return ISF ? (T)((T)1 / (T)sqrt((double)x))                                   \
                          : ((double)x <= 0 ? (T)0 : (T)(1.0 / sqrt((double)x)));            \
    case 0x8E: return ISF ? (T)1 / x : (x == (T)0 ? (T)0 : (T)(1.0 / (double)x));            \
    case 0x78: return s;                                   /* VSPLAT: register bits as T */  \
    case 0xA0: case 0xA1: case 0xA2: case 0xA3: case 0xA4: case 0xA5:                        \
        return VO_PRED(op - 0xA0, x, y) ? BL(~0ull) : (T)0;       /* VCMPxx: all-ones lane */ \
    case 0xA6: return (ISF ? x != x : x == (T)0) ? BL(~0ull) : (T)0;   /* VCMPNULL */        \
    case 0xA7: return (ISF ? x == x : x != (T)0) ? BL(~0ull) : (T)0;   /* VCMPNOTNULL */     \
    case 0xB0: return BL(LB(x) & LB(y));                                                     \
    case 0xB1: return BL(LB(x) | LB(y));                                                     \
    case 0xB2: return BL(LB(x) ^ LB(y));                                                     \
    case 0xB3: return BL(~LB(x));                                                            \
    case 0xB4: return BL(~LB(x) & LB(y));                                                    \
    case 0xB5: return ISF ? (T)ldexp((double)x, vo_cvtt((double)y))                          \
                          : (T)(LB(x) << (LB(y) & 63));                                      \
    case 0xB6: return ISF ? (T)ldexp((double)x, vo_neg32(vo_cvtt((double)y)))                \
                          : (T)(LB(x) >> (LB(y) & 63));                                      \
    case 0xB7: return ISF ? (T)ldexp((double)x, vo_neg32(vo_cvtt((double)y)))                \
                          : (T)((int64_t)LB(x) >> (LB(y) & 63));                             \
    case 0x86: return ISF ? (T)fabs((double)x) : (x < (T)0 ? (T)(0 - (uint64_t)x) : x);      \
    case 0x87: return ISF ? ((x < y || x != x) ? x : y) : (x < y ? x : y);                   \
    case 0x88: return ISF ? ((x > y || y != y) ? x : y) : (x > y ? x : y);                   \
    case 0x90: return ISF ? x + s : (T)((uint64_t)x + (uint64_t)s);                          \
    case 0x91: return ISF ? x - s : (T)((uint64_t)x - (uint64_t)s);                          \
    case 0x92: return ISF ? x * s : (T)((uint64_t)x * (uint64_t)s);                          \
    case 0x93: return ISF ? ((s == (T)0) ? (T)NAN : x / s)                                    \
                          : (s == (T)0 ? (T)0 : s == (T)-1 ? (T)(0 - (uint64_t)x) : (T)((int64_t)x / (int64_t)s)); \
    case 0x94: return ISF ? (T)fmod((double)x, (double)s)                                    \
                          : (s == (T)0 || s == (T)-1 ? (T)0 : (T)((int64_t)x % (int64_t)s));  \
    case 0x95: return ISF ? ((x < s || x != x) ? x : s) : (x < s ? x : s);                   \
    case 0x96: return ISF ? ((x > s || s != s) ? x : s) : (x > s ? x : s);                   \
    case 0x8F: return ISF ? (T)pow((double)x, (double)y) : (T)vo_ipow((int64_t)x, (int64_t)y); \
    default: return ((cmp & 0xF) != 7 && VO_PRED(cmp & 0xF, x, s)) ? x : (T)0;              \
    }                                                                                        \
}
/* Engine<T>::IntPow (engine.hpp:443-454): square-and-multiply in wrapping
   i64, 0 for a negative exponent */
static int64_t vo_ipow(int64_t b, int64_t e) {
    if (e < 0) return 0;
    uint64_t r = 1, bb = (uint64_t)b;
    while (e) {
        if (e & 1) r *= bb;
        bb *= bb;
        e >>= 1;
    }
    return (int64_t)r;
}
VO_DEF_EXPR_LANE(vo_expr_lane_f64, double, 1, vo_lb_f64, vo_bl_f64)
VO_DEF_EXPR_LANE(vo_expr_lane_i64, int64_t, 0, vo_lb_i64, vo_bl_i64)

/* The engine's eight mask registers (RegFile masks): the filters, VCMPxx
   (mask 0) and VCMPNULL / VCMPNOTNULL (mask imm12 & 7) write lane bits that
   VBLEND reads (engine.hpp:1090-1110, 480-493, 2700-2750, 2926-2936). */
static uint32_t vo_masks[8];
void vo_expr_reset_masks(void) { memset(vo_masks, 0, sizeof(vo_masks)); }

/* The segment VGATHER reads (engine.hpp:1987-2006), one segment id at a
   time (test aid): other ids are segments the engine does not have. */
static const void* vo_gdata;
static uint64_t vo_gcount;
static int vo_gseg = -1;
void vo_expr_set_gather(int seg, const void* data, uint64_t count) {
    vo_gseg = seg;
    vo_gdata = data;
    vo_gcount = count;
}

/* The segment VSTORE writes (engine.hpp:1971-1985), one segment id at a
   time (test aid; it may be the loaded column itself): lanes [0, count) to
   rows off + lane below the segment's count.  Other ids: nothing written. */
static void* vo_sdata;
static uint64_t vo_scount;
static int vo_sseg = -1;
void vo_expr_set_store(int seg, void* data, uint64_t count) {
    vo_sseg = seg;
    vo_sdata = data;
    vo_scount = count;
}

/* one kL = 4 group of the body's vector words on the 16 vector registers.
   Op word k: code[5k..5k+4] = opcode, dst, a, b, x; x = the filter's kCmp
   | mask register << 4 for VFILTER*, imm12 for the lane words (VROTATE
   shift, VSHUFFLE mode, VINSERT lane, VBLEND and VCMP(NOT)NULL mask,
   VGATHER segment << 8 | count).  Lane words (engine.hpp:2124-2217):
   VINSERT lane = R[rb], VPERMUTE by the bytes of rb's first lane, VSHUFFLE
   halves, VREVERSE, VROTATE, VSLIDE by R[rb] (size_t arithmetic). */
#define VO_DEF_EXPR_GROUP(NAME, T, ISF, LANE, LB, BL)                                        \
static void NAME(T v[16][4], const T* data, uint64_t count, uint64_t off, int vload_dst,    \
                 int nops, const int* code, const T* sval) {                                 \
    for (int i = 0; i < 4; ++i) {                                                            \
        uint64_t idx = off + (uint64_t)i;                                                    \
        v[vload_dst][i] = idx < count ? data[idx] : (T)0;                                    \
    }                                                                                        \
    for (int k = 0; k < nops; ++k) {                                                         \
        const int* c = code + 5 * k;                                                         \
        const int op = c[0], a = c[2], b = c[3], x = c[4];                                   \
        T r[4];                                                                              \
        switch (op) {                                                                        \
        case 0x71:                                                              /* VSTORE */  \
            if (((x >> 8) & 0xF) == vo_sseg) {                                               \
                const int cnt = (x & 0xFF) ? (x & 0xFF) : 4;                                 \
                for (int i = 0; i < cnt && off + (uint64_t)i < vo_scount; ++i)               \
                    ((T*)vo_sdata)[off + (uint64_t)i] = v[a][i];                             \
            }                                                                                \
            continue;                                                   /* no register written */ \
        case 0x7D: for (int i = 0; i < 4; ++i) r[i] = v[a][3 - i]; break;      /* VREVERSE */ \
        case 0x7E: for (int i = 0; i < 4; ++i) r[(i + (x & 3)) & 3] = v[a][i]; break;  /* VROTATE */ \
        case 0x7F: {                                                            /* VSLIDE */  \
            const uint64_t sh = LB(sval[k]);                                                 \
            for (int i = 0; i < 4; ++i) { uint64_t si = (uint64_t)i + sh; r[i] = si < 4 ? v[a][si] : (T)0; } \
            break;                                                                           \
        }                                                                                    \
        case 0x7C:                                                              /* VSHUFFLE */ \
            for (int i = 0; i < 2; ++i) {                                                    \
                const int j = (x & 3) == 0 ? 2 * i : (x & 3) == 1 ? 2 * i + 1 : (x & 3) == 2 ? i : i + 2; \
                r[i] = v[a][j];                                                              \
                r[i + 2] = v[b][j];                                                          \
            }                                                                                \
            break;                                                                           \
        case 0x7B: {                                                            /* VPERMUTE */ \
            const uint64_t ib = LB(v[b][0]);                                                 \
            for (int i = 0; i < 4; ++i) r[i] = v[a][((ib >> (8 * i)) & 0xFF) % 4];           \
            break;                                                                           \
        }                                                                                    \
        case 0x7A: for (int i = 0; i < 4; ++i) r[i] = i == (x & 3) ? sval[k] : v[a][i]; break; /* VINSERT */ \
        case 0xC7:                                                              /* VBLEND */  \
            for (int i = 0; i < 4; ++i) r[i] = ((vo_masks[x & 7] >> i) & 1) ? v[b][i] : v[a][i]; \
            break;                                                                           \
        case 0x72: {                                                            /* VGATHER */ \
            const int cnt = (x & 0xFF) ? (x & 0xFF) : 4;                                     \
            if (((x >> 8) & 0xF) != vo_gseg) { for (int i = 0; i < 4; ++i) r[i] = (T)0; break; } \
            for (int i = 0; i < 4; ++i) r[i] = v[c[1]][i];                                   \
            for (int i = 0; i < 4 && i < cnt; ++i) {                                         \
                const int32_t i32 = (int32_t)(uint32_t)(LB(v[a][i >> 1]) >> (32 * (i & 1))); \
                const uint64_t gi = (uint64_t)(int64_t)i32;                                  \
                r[i] = gi < vo_gcount ? ((const T*)vo_gdata)[gi] : (T)0;                     \
            }                                                                                \
            break;                                                                           \
        }                                                                                    \
        default: {                                                                           \
            uint32_t m = 0;                                                                  \
            for (int i = 0; i < 4; ++i) {                                                    \
                r[i] = LANE(op, x, v[a][i], v[b][i], sval[k]);                               \
                if (op >= 0xC0 && op <= 0xC6)                                                \
                    m |= ((x & 0xF) != 7 && VO_PRED(x & 0xF, v[a][i], sval[k])) ? 1u << i : 0u; \
                else                                                                         \
                    m |= LB(r[i]) ? 1u << i : 0u;                                            \
            }                                                                                \
            if (op >= 0xC0 && op <= 0xC6) vo_masks[(x >> 4) & 7] = m;                        \
            else if (op >= 0xA0 && op <= 0xA5) vo_masks[0] = m;                              \
            else if (op == 0xA6 || op == 0xA7) vo_masks[x & 7] = m;                          \
            break;                                                                           \
        }                                                                                    \
        }                                                                                    \
        for (int i = 0; i < 4; ++i) v[c[1]][i] = r[i];                                       \
    }                                                                                        \
}
VO_DEF_EXPR_GROUP(vo_expr_group_f64, double, 1, vo_expr_lane_f64, vo_lb_f64, vo_bl_f64)
VO_DEF_EXPR_GROUP(vo_expr_group_i64, int64_t, 0, vo_expr_lane_i64, vo_lb_i64, vo_bl_i64)

#define VO_DEF_TIER0_EXPR(NAME, COLNAME, T, GROUP, RED)                                      \
void NAME(const T* data, uint64_t count, uint64_t off, uint64_t total, int vload_dst,       \
          int nops, const int* code, const T* sval, int nb, const int* bvreg,                \
          const int* red, const int* accop, uint64_t* acc, uint64_t* part) {                  \
    T v[16][4];                                                                              \
    memset(v, 0, sizeof(v));                                                                 \
    do {                                                                                     \
        GROUP(v, data, count, off, vload_dst, nops, code, sval);                             \
        for (int b = 0; b < nb; ++b) {                                                       \
            part[b] = RED(v[bvreg[b]], red[b]);                                              \
            acc[b] = vo_tier0_accum(acc[b], part[b], accop[b]);                              \
        }                                                                                    \
        off += 4;                                                                            \
    } while (off != total);                                                                  \
}                                                                                            \
/* the lanes of vector register `target` over rows [off, total) (test aid) */               \
void COLNAME(const T* data, uint64_t count, uint64_t off, uint64_t total, int vload_dst,    \
             int nops, const int* code, const T* sval, int target, T* out) {                 \
    T v[16][4];                                                                              \
    memset(v, 0, sizeof(v));                                                                 \
    for (uint64_t g = 0; off + 4 * g < total; ++g) {                                         \
        GROUP(v, data, count, off + 4 * g, vload_dst, nops, code, sval);                     \
        for (int i = 0; i < 4; ++i) out[4 * g + (uint64_t)i] = v[target][i];                \
    }                                                                                        \
}
VO_DEF_TIER0_EXPR(vo_tier0_expr_f64, vo_tier0_expr_col_f64, double, vo_expr_group_f64, vo_red_f64)
VO_DEF_TIER0_EXPR(vo_tier0_expr_i64, vo_tier0_expr_col_i64, int64_t, vo_expr_group_i64, vo_red_i64)

/* ------------------------------------------------------------------------ */
/* Straight-line scalar words (engine.hpp:1167-1178 MOV/MOVR, 1286-1380     */
/* integer arithmetic on register bits as i64, 1394-1446 f64 arithmetic,    */
/* 1457-1505 bitwise; opcodes.hpp:24-60).  Returns 1 when every word ran,    */
/* 0 when DIV/MOD/DIVF by zero stopped the program (Running_ = false),      */
/* -1 for INT64_MIN / -1 (a trap in the reference's host) or an opcode      */
/* outside this set.                                                        */
/* ------------------------------------------------------------------------ */
/* x86 SSE/AVX scalar arithmetic returns the first NaN operand, quieted,
   when an operand is NaN (the reference binary's ADDF..DIVF). */
static uint64_t vo_f64_op(unsigned op, uint64_t a, uint64_t b) {
    double fa = vo_bitsf64(a), fb = vo_bitsf64(b), r;
    if (isnan(fa)) return a | 0x0008000000000000ull;
    if (isnan(fb)) return b | 0x0008000000000000ull;
    r = op == 0x2A ? fa + fb : op == 0x2B ? fa - fb : op == 0x2C ? fa * fb : fa / fb;
    return vo_f64bits(r);
}

int vo_scalar_run(const uint32_t* w, size_t n, uint64_t* r) {
    for (size_t k = 0; k < n; ++k) {
        uint32_t x = w[k];
        unsigned op = x & 0xFF, rd = (x >> 8) & 0xF, ra = (x >> 12) & 0xF, rb = (x >> 16) & 0xF;
        int imm = (int)((x >> 20) & 0xFFF);
        if (imm & 0x800) imm -= 0x1000;
        uint64_t a = r[ra], b = r[rb];
        int64_t sa = (int64_t)a, sb = (int64_t)b;
        double fa = vo_bitsf64(a), fb = vo_bitsf64(b);
        switch (op) {
        case 0x00: break;
        case 0x10: r[rd] = (uint64_t)(int64_t)imm; break;
        case 0x11: r[rd] = a; break;
        case 0x20: r[rd] = a + b; break;
        case 0x21: r[rd] = a - b; break;
        case 0x22: r[rd] = a * b; break;
        case 0x23: case 0x24:
            if (sb == 0) return 0;
            if (sa == INT64_MIN && sb == -1) return -1;
            r[rd] = (uint64_t)(op == 0x23 ? sa / sb : sa % sb);
            break;
        case 0x25: r[rd] = 0 - a; break;
        case 0x26: r[rd] = sa < 0 ? 0 - a : a; break;
        case 0x27: r[rd] = (uint64_t)(sa < sb ? sa : sb); break;
        case 0x28: r[rd] = (uint64_t)(sa > sb ? sa : sb); break;
        case 0x29: r[rd] = (uint64_t)((sa >> 1) + (sb >> 1) + ((sa & 1) + (sb & 1)) / 2); break;
        case 0x2A: case 0x2B: case 0x2C: r[rd] = vo_f64_op(op, a, b); break;
        case 0x2D: if (fb == 0.0) return 0; r[rd] = vo_f64_op(op, a, b); break;
        case 0x2E: r[rd] = a ^ 0x8000000000000000ull; break;
        case 0x2F: r[rd] = a & 0x7FFFFFFFFFFFFFFFull; break;
        case 0x30: r[rd] = a & b; break;
        case 0x31: r[rd] = a | b; break;
        case 0x32: r[rd] = a ^ b; break;
        case 0x33: r[rd] = ~a; break;
        case 0x34: r[rd] = a << (b & 0x3F); break;
        case 0x35: r[rd] = a >> (b & 0x3F); break;
        case 0x36: r[rd] = (uint64_t)(sa >> (b & 0x3F)); break;
        /* conversions, engine.hpp:1822-1937 */
        case 0x60: r[rd] = (uint64_t)(int64_t)(int8_t)(uint8_t)a; break;
        case 0x61: r[rd] = (uint64_t)(int64_t)(int16_t)(uint16_t)a; break;
        case 0x62: r[rd] = (uint64_t)(int64_t)(int32_t)(uint32_t)a; break;
        case 0x64: r[rd] = vo_f64bits((double)(float)fa); break;
        case 0x66: r[rd] = a & 0xFF; break;
        case 0x67: r[rd] = a & 0xFFFF; break;
        case 0x68: r[rd] = a & 0xFFFFFFFFull; break;
        case 0x63: case 0x65: case 0x69: case 0x6A: case 0x6B: r[rd] = a; break;
        case 0x6C: {           /* (i64)trunc(x): cvttsd2si, INT64_MIN for NaN / out of range */
            double t = trunc(fa);
            r[rd] = (t != t || t >= 9223372036854775808.0 || t < -9223372036854775808.0)
                        ? 0x8000000000000000ull : (uint64_t)(int64_t)t;
            break;
        }
        case 0x6D: r[rd] = vo_f64bits(round(fa)); break;
        case 0x6E: r[rd] = vo_f64bits(ceil(fa)); break;
        case 0x6F: r[rd] = vo_f64bits(floor(fa)); break;
        default: return -1;
        }
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Segment aggregates AGG_COUNT..AGG_PERCENTILE (engine.hpp:499-698,        */
/* handlers 3229-3283) on Engine<f64/i64>, with the engine's 16 AggState    */
/* slots (engine.hpp:347-372), inside straight-line programs of scalar      */
/* words (vo_scalar_run) and AGG words.  AGG word: rd = slot and result     */
/* register, ra = segment (one here: 0), rb = register holding the start    */
/* row, imm12 = row count (0: all; AGG_PERCENTILE also reads the percent    */
/* from imm12 & 0xFF).  Sums and Welford run in row order as in the         */
/* reference.  The distinct set is a sorted array (its size is what the     */
/* reference returns).  AGG_SUM_DISTINCT sums the set in ascending order  */
/* (the reference: unordered_map order; integer sums are exact, f64 sums    */
/* agree within rounding); AGG_MODE is restated where the largest count is  */
/* unique (with ties the reference's pick follows unordered_map order).     */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint64_t count; double sumf; int64_t sumi;
    uint64_t minv, maxv, firstv, lastv;          /* T bits */
    double m2, mean;
    int init, first_set, has_set;
    uint64_t* set; size_t nset, capset;          /* sorted distinct T bits (f64: +0 for -0) */
} vo_aggslot;

typedef struct { vo_aggslot s[16]; int is_f64; } vo_aggstate;

static void vo_slot_reset(vo_aggslot* a, int is_f64) {
    free(a->set);
    memset(a, 0, sizeof(*a));
    a->minv = is_f64 ? vo_f64bits(1.7976931348623157e308) : (uint64_t)INT64_MAX;
    a->maxv = is_f64 ? vo_f64bits(-1.7976931348623157e308) : (uint64_t)INT64_MIN;
}

void* vo_agg_create(int is_f64) {
    vo_aggstate* st = (vo_aggstate*)calloc(1, sizeof(vo_aggstate));
    st->is_f64 = is_f64;
    for (int i = 0; i < 16; ++i) vo_slot_reset(&st->s[i], is_f64);
    return st;
}

void vo_agg_destroy(void* p) {
    vo_aggstate* st = (vo_aggstate*)p;
    if (!st) return;
    for (int i = 0; i < 16; ++i) free(st->s[i].set);
    free(st);
}

static int vo_cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : x > y;
}

/* (key, row) pairs: by key, then row */
static int vo_cmp_pair(const void* a, const void* b) {
    const uint64_t* x = (const uint64_t*)a;
    const uint64_t* y = (const uint64_t*)b;
    if (x[0] != y[0]) return x[0] < y[0] ? -1 : 1;
    return x[1] < y[1] ? -1 : x[1] > y[1];
}

/* order key of a T value (f64: -0 as +0) */
static uint64_t vo_okey(uint64_t bits, int is_f64) {
    if (!is_f64) return bits ^ 0x8000000000000000ull;
    if (bits == 0x8000000000000000ull) bits = 0;
    return (bits >> 63) ? ~bits : (bits | 0x8000000000000000ull);
}
static uint64_t vo_unkey(uint64_t k, int is_f64) {
    if (!is_f64) return k ^ 0x8000000000000000ull;
    return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
}

/* returns 1 ran, -1 not restated (or NaN where the reference's order is unspecified) */
static int vo_agg_word(vo_aggstate* st, const uint64_t* data, uint64_t count, uint32_t w, uint64_t* r) {
    const int F = st->is_f64;
    unsigned op = w & 0xFF, rd = (w >> 8) & 0xF, ra = (w >> 12) & 0xF, rb = (w >> 16) & 0xF, imm = (w >> 20) & 0xFFF;
    if (ra >= 1) return 1;                                   /* segment id past the engine's segments: PC++ only */
    uint64_t start = r[rb];
    uint64_t cnt = imm == 0 ? 0xFFFFFFFFull : (uint64_t)(imm & 0xFFF);
    uint64_t end = start + cnt;
    if (end > count) end = count;
    vo_aggslot* s = &st->s[rd & 0xF];
    #define X(i) (F ? vo_bitsf64(data[i]) : (double)(int64_t)data[i])
    switch (op) {
    case 0xE0:                                                           /* COUNT */
        if (!s->init) { s->count = 0; s->init = 1; }
        for (uint64_t i = start; i < end; ++i)
            if (F ? (vo_bitsf64(data[i]) != 0.0) : (data[i] != 0)) s->count++;
        r[rd] = s->count;
        return 1;
    case 0xE1:                                                           /* SUM */
        if (!s->init) { s->sumf = 0; s->sumi = 0; s->init = 1; }
        if (F) { for (uint64_t i = start; i < end; ++i) s->sumf += vo_bitsf64(data[i]); r[rd] = vo_f64bits(s->sumf); }
        else { for (uint64_t i = start; i < end; ++i) s->sumi = (int64_t)((uint64_t)s->sumi + data[i]); r[rd] = (uint64_t)s->sumi; }
        return 1;
    case 0xE2: {                                                         /* AVG */
        if (!s->init) { s->count = 0; s->sumf = 0; s->sumi = 0; s->init = 1; }
        uint64_t n = 0;
        if (F) {
            for (uint64_t i = start; i < end; ++i) { s->sumf += vo_bitsf64(data[i]); n++; }
            s->count += n;
            r[rd] = vo_f64bits(s->count ? s->sumf / (double)s->count : 0.0);
        } else {
            for (uint64_t i = start; i < end; ++i) { s->sumi = (int64_t)((uint64_t)s->sumi + data[i]); n++; }
            s->count += n;
            r[rd] = (uint64_t)(s->count ? s->sumi / (int64_t)s->count : 0);
        }
        return 1;
    }
    case 0xE3: case 0xE4:                                                /* MIN, MAX */
        if (!s->init) {
            if (op == 0xE3) s->minv = F ? vo_f64bits(1.7976931348623157e308) : (uint64_t)INT64_MAX;
            else s->maxv = F ? vo_f64bits(-1.7976931348623157e308) : (uint64_t)INT64_MIN;
            s->init = 1;
        }
        for (uint64_t i = start; i < end; ++i) {
            if (op == 0xE3) {
                if (F ? vo_bitsf64(data[i]) < vo_bitsf64(s->minv) : (int64_t)data[i] < (int64_t)s->minv) s->minv = data[i];
            } else {
                if (F ? vo_bitsf64(data[i]) > vo_bitsf64(s->maxv) : (int64_t)data[i] > (int64_t)s->maxv) s->maxv = data[i];
            }
        }
        r[rd] = op == 0xE3 ? s->minv : s->maxv;
        return 1;
    case 0xE5:                                                           /* FIRST */
        if (!s->init) { s->first_set = 0; s->init = 1; }
        if (!s->first_set && end > start) { s->firstv = data[start]; s->first_set = 1; }
        r[rd] = s->firstv;
        return 1;
    case 0xE6:                                                           /* LAST */
        if (!s->init) s->init = 1;
        if (end > start) s->lastv = data[end - 1];
        r[rd] = s->lastv;
        return 1;
    case 0xE7: case 0xE8: {                                              /* STDDEV, VARIANCE */
        if (!s->init) { s->count = 0; s->sumf = 0; s->sumi = 0; s->mean = 0; s->m2 = 0; s->init = 1; }
        for (uint64_t i = start; i < end; ++i) {
            double v = X(i);
            s->count++;
            double d = v - s->mean;
            s->mean += d / (double)s->count;
            double d2 = v - s->mean;
            s->m2 += d * d2;
        }
        double var = s->count > 1 ? s->m2 / (double)s->count : 0.0;
        double res = op == 0xE7 ? sqrt(var) : var;
        r[rd] = F ? vo_f64bits(res) : (uint64_t)(int64_t)res;
        return 1;
    }
    case 0xE9: case 0xEA: {                                              /* COUNT_DISTINCT, SUM_DISTINCT */
        if (!s->init) { s->has_set = 1; s->nset = 0; s->init = 1; }
        if (!s->has_set) return -1;                                      /* the reference dereferences a null set */
        for (uint64_t i = start; i < end; ++i) {
            if (F && isnan(vo_bitsf64(data[i]))) return -1;              /* NaN keys: unordered_map keeps each */
            uint64_t k = vo_okey(data[i], F);
            size_t lo = 0, hi = s->nset;
            while (lo < hi) { size_t mid = (lo + hi) / 2; if (s->set[mid] < k) lo = mid + 1; else hi = mid; }
            if (lo < s->nset && s->set[lo] == k) continue;
            if (s->nset == s->capset) {
                s->capset = s->capset ? 2 * s->capset : 64;
                s->set = (uint64_t*)realloc(s->set, s->capset * 8);
            }
            memmove(s->set + lo + 1, s->set + lo, (s->nset - lo) * 8);
            s->set[lo] = k;
            s->nset++;
        }
        if (op == 0xE9) { r[rd] = s->nset; return 1; }
        /* SUM_DISTINCT (engine.hpp:629-647): the set's values summed in T's
           accumulator (f64 / wrapping i64); the reference adds them in
           unordered_map order, restated here in ascending order (f64 sums
           agree within rounding, integer sums exactly) */
        if (F) {
            double sum = 0;
            for (size_t i = 0; i < s->nset; ++i) sum += vo_bitsf64(vo_unkey(s->set[i], F));
            r[rd] = vo_f64bits(sum);
        } else {
            uint64_t sum = 0;
            for (size_t i = 0; i < s->nset; ++i) sum += vo_unkey(s->set[i], F);
            r[rd] = sum;
        }
        return 1;
    }
    case 0xEC: {                                                         /* MODE (engine.hpp:669-681) */
        uint64_t n = end > start ? end - start : 0;
        if (n == 0) { r[rd] = 0; return 1; }
        /* (key, row) pairs sorted by key then row: each run of equal keys is
           one map entry whose stored key is its first row's value; a NaN row
           is an entry of its own (count 1).  The reference returns the
           entry with the largest count met first in unordered_map order,
           which is restated only when that largest count is unique (-1
           otherwise). */
        uint64_t* kr = (uint64_t*)malloc(2 * n * 8);
        for (uint64_t i = 0; i < n; ++i) { kr[2 * i] = vo_okey(data[start + i], F); kr[2 * i + 1] = i; }
        qsort(kr, n, 16, vo_cmp_pair);
        uint64_t best = 0, best_row = 0, ties = 0;
        for (uint64_t i = 0; i < n;) {
            uint64_t j = i + 1;
            const int nan = F && isnan(vo_bitsf64(data[start + kr[2 * i + 1]]));
            if (!nan) while (j < n && kr[2 * j] == kr[2 * i]) ++j;
            const uint64_t len = j - i;
            if (len > best) { best = len; best_row = kr[2 * i + 1]; ties = 1; }
            else if (len == best) ties++;
            i = j;
        }
        free(kr);
        if (ties != 1) return -1;
        r[rd] = data[start + best_row];
        return 1;
    }
    case 0xEB: case 0xED: {                                              /* MEDIAN, PERCENTILE */
        uint64_t n = end > start ? end - start : 0;
        if (n == 0) { r[rd] = 0; return 1; }
        uint64_t* k = (uint64_t*)malloc(n * 8);
        for (uint64_t i = 0; i < n; ++i) {
            if (F && isnan(vo_bitsf64(data[start + i]))) { free(k); return -1; }   /* nth_element order unspecified */
            k[i] = vo_okey(data[start + i], F);
        }
        qsort(k, n, 8, vo_cmp_u64);
        if (op == 0xED) {
            double pct = (double)(imm & 0xFF) / 100.0;
            uint64_t idx = (uint64_t)(pct * (double)n);
            if (idx >= n) idx = n - 1;
            r[rd] = vo_unkey(k[idx], F);
        } else if (n % 2 == 0) {
            uint64_t a = vo_unkey(k[n / 2], F), b = vo_unkey(k[n / 2 - 1], F);
            r[rd] = F ? vo_f64bits((vo_bitsf64(a) + vo_bitsf64(b)) * 0.5) : (uint64_t)(((int64_t)a + (int64_t)b) / 2);
        } else {
            r[rd] = vo_unkey(k[n / 2], F);
        }
        free(k);
        return 1;
    }
    default:
        return -1;
    }
    #undef X
}

/* Straight-line program of scalar and AGG words (stops at HALT or at a
   division by zero).  Returns 1 ran, 0 stopped by a division by zero,
   -1 a word outside the restated set. */
int vo_straight_run(void* state, const uint64_t* data, uint64_t count, const uint32_t* w, size_t n, uint64_t* r) {
    vo_aggstate* st = (vo_aggstate*)state;
    for (size_t k = 0; k < n; ++k) {
        unsigned op = w[k] & 0xFF;
        if (op == 0x01) return 1;
        if (op >= 0xE0 && op <= 0xED) {
            if (vo_agg_word(st, data, count, w[k], r) < 0) return -1;
            continue;
        }
        int rc = vo_scalar_run(&w[k], 1, r);
        if (rc != 1) return rc;
    }
    return 1;
}

/* ------------------------------------------------------------------------ */
/* JIT fused loop, codegen/x86_64.hpp:1632-1716, as entered by voxel_jit_run */
/* (bindings/voxel_c.cpp:80-101): iters = (R2 - off + 7) >> 3; each         */
/* iteration adds masked x[8k+j] into lane j of accumulator A (j<4) or B;   */
/* vcmppd predicates are ordered (NaN fails every one, NE included,         */
/* line 1606); result = (L0+L1)+(L2+L3) with L = A+B; the initial R[acc] is */
/* ignored.  Rows past data_len read as 0 here (the reference over-reads).  */
/* ------------------------------------------------------------------------ */
double vo_jit_f64(const double* data, uint64_t count, uint64_t r2, uint64_t off, uint8_t fop, double thresh) {
    int op = vo_fast_pass_code(fop);
    uint64_t iters = (r2 - off + 7) >> 3;
    double A[4] = {0, 0, 0, 0}, B[4] = {0, 0, 0, 0};
    for (uint64_t k = 0; k < iters; ++k) {
        for (int j = 0; j < 8; ++j) {
            uint64_t idx = off + 8 * k + (uint64_t)j;
            double x = idx < count ? data[idx] : 0.0;
            double v = (x == x && VO_PRED(op, x, thresh)) ? x : 0.0;
            if (j < 4) A[j] += v; else B[j - 4] += v;
        }
    }
    double L[4];
    for (int j = 0; j < 4; ++j) L[j] = A[j] + B[j];
    return (L[0] + L[1]) + (L[2] + L[3]);
}

/* ------------------------------------------------------------------------ */
/* Engine<T>::RunParallel, engine.hpp:143-175, for fast-path programs:      */
/* chunk = ceil(N/threads); worker t runs the fast path on rows             */
/* [t*chunk, t*chunk+count) with every register copied (so the same R[off]  */
/* and R[total], total clipped to the chunk's Count); combined = sum over   */
/* workers of R[result] in T, in worker order, written to R[result].        */
/* Only the result register's value is restated here.                       */
/* ------------------------------------------------------------------------ */
double vo_run_parallel_fast_f64(const double* data, uint64_t n, uint32_t threads,
                                uint64_t reg_total, uint64_t reg_off, double thresh,
                                double acc0, uint8_t fop) {
    if (threads < 1) threads = 1;
    uint64_t chunk = (n + threads - 1) / threads;
    double combined = 0.0;
    for (uint32_t t = 0; t < threads; ++t) {
        uint64_t start = (uint64_t)t * chunk;
        if (start >= n) continue;                    /* count == 0: worker never ran */
        uint64_t cnt = (start + chunk > n) ? (n - start) : chunk;
        double acc;
        vo_fastpath_f64(data + start, cnt, reg_total, reg_off, thresh, acc0, fop, &acc);
        combined += acc;
    }
    return combined;
}

/* ------------------------------------------------------------------------ */
/* Bitmap filters: VectorFilter<T>::Apply* (ops/filter.hpp:42-288): bit     */
/* (i & 63) of word i/64 = pred(x_i); tail bits zero; PassCount = popcount. */
/* op 6 = Between(lo=t, hi=t2).                                             */
/* ------------------------------------------------------------------------ */
#define VO_DEF_BITMAP(NAME, T)                                                        \
uint64_t NAME(const T* x, size_t n, int op, T t, T t2, uint64_t* words) {            \
    size_t nw = (n + 63) / 64; memset(words, 0, nw * 8);                              \
    uint64_t pc = 0;                                                                  \
    for (size_t i = 0; i < n; ++i) {                                                  \
        int p = op == 6 ? (x[i] >= t && x[i] <= t2) : VO_PRED(op, x[i], t);           \
        if (p) { words[i >> 6] |= 1ULL << (i & 63); ++pc; }                           \
    }                                                                                 \
    return pc;                                                                        \
}
VO_DEF_BITMAP(vo_bitmap_f64, double)
VO_DEF_BITMAP(vo_bitmap_f32, float)
VO_DEF_BITMAP(vo_bitmap_i64, int64_t)
VO_DEF_BITMAP(vo_bitmap_i32, int32_t)

/* ColumnFilter::Init/AndWith/OrWith/AndNotWith (ops/filter.hpp:528-570):   */
/* combine = 0 AND, 1 OR, 2 ANDNOT; Init's all-ones bitmap is tail-masked;  */
/* returns the active count.                                                */
uint64_t vo_bitmap_combine(uint64_t* active, const uint64_t* other, size_t rows, int combine) {
    size_t nw = (rows + 63) / 64;
    uint64_t pc = 0;
    for (size_t w = 0; w < nw; ++w) {
        uint64_t v = combine == 0 ? (active[w] & other[w]) : combine == 1 ? (active[w] | other[w])
                                                                          : (active[w] & ~other[w]);
        if (w == nw - 1 && (rows & 63)) v &= (1ULL << (rows & 63)) - 1;
        active[w] = v;
        pc += (uint64_t)__builtin_popcountll(v);
    }
    return pc;
}

/* SelectionVector::BuildFromBitmap (ops/filter.hpp:378-399): ascending    */
/* u32 row ids of set bits.                                                 */
uint64_t vo_select_from_bitmap(const uint64_t* words, size_t rows, uint32_t* out) {
    size_t nw = (rows + 63) / 64;
    uint64_t k = 0;
    for (size_t w = 0; w < nw; ++w) {
        uint64_t v = words[w];
        while (v) { out[k++] = (uint32_t)(w * 64 + (uint64_t)__builtin_ctzll(v)); v &= v - 1; }
    }
    return k;
}

/* scalar_filter_gt_* (simd/scalar.hpp:155,214,260,306): passing values in  */
/* row order (op generalised to the six kCmp codes).                        */
#define VO_DEF_COMPACT(NAME, T)                                                \
uint64_t NAME(const T* x, size_t n, int op, T t, T* out) {                    \
    uint64_t k = 0;                                                            \
    for (size_t i = 0; i < n; ++i) if (VO_PRED(op, x[i], t)) out[k++] = x[i];  \
    return k;                                                                  \
}
VO_DEF_COMPACT(vo_compact_f64, double)
VO_DEF_COMPACT(vo_compact_f32, float)
VO_DEF_COMPACT(vo_compact_i64, int64_t)
VO_DEF_COMPACT(vo_compact_i32, int32_t)

/* ------------------------------------------------------------------------ */
/* RadixSort<u64>::Sort(data, indices, count), ops/sort.hpp:83-160: eight   */
/* LSD passes over bytes 0..7: histogram, exclusive prefix, stable scatter  */
/* of (key, payload).  Unsigned keys: no sign pass.                         */
/* ------------------------------------------------------------------------ */
void vo_radix_sort_u64(uint64_t* keys, uint32_t* idx, size_t n) {
    if (n <= 1) return;
    uint64_t* tk = (uint64_t*)malloc(n * 8);
    uint32_t* ti = (uint32_t*)malloc(n * 4);
    uint64_t *sk = keys, *dk = tk; uint32_t *si = idx, *di = ti;
    for (int b = 0; b < 8; ++b) {
        size_t hist[256] = {0};
        int sh = 8 * b;
        for (size_t i = 0; i < n; ++i) ++hist[(sk[i] >> sh) & 0xFF];
        size_t pre[256]; pre[0] = 0;
        for (int d = 1; d < 256; ++d) pre[d] = pre[d - 1] + hist[d - 1];
        for (size_t i = 0; i < n; ++i) {
            size_t p = pre[(sk[i] >> sh) & 0xFF]++;
            dk[p] = sk[i]; di[p] = si[i];
        }
        uint64_t* t1 = sk; sk = dk; dk = t1;
        uint32_t* t2 = si; si = di; di = t2;
    }
    /* eight passes: the data is back in the caller's buffers */
    free(tk); free(ti);
}

/* MergeSort<f64>::Sort(data, idx, n) + SortOperator (ops/sort.hpp:196-213, */
/* 303-338, 506-541): bottom-up stable merge, "<=" takes left (ascending),  */
/* ">=" takes left (descending).  Writes the permutation of 0..n-1.         */
void vo_mergesort_f64(const double* data, size_t n, int descending, uint32_t* idx_out) {
    for (size_t i = 0; i < n; ++i) idx_out[i] = (uint32_t)i;
    if (n <= 1) return;
    double* v = (double*)malloc(n * 8); memcpy(v, data, n * 8);
    double* tv = (double*)malloc(n * 8); uint32_t* ti = (uint32_t*)malloc(n * 4);
    for (size_t w = 1; w < n; w *= 2) {
        for (size_t left = 0; left < n; left += 2 * w) {
            size_t mid = left + w < n ? left + w : n, right = left + 2 * w < n ? left + 2 * w : n;
            if (mid >= right) continue;
            size_t ll = mid - left;
            memcpy(tv, v + left, ll * 8); memcpy(ti, idx_out + left, ll * 4);
            size_t a = 0, b = mid, o = left;
            while (a < ll && b < right) {
                int takeLeft = descending ? (tv[a] >= v[b]) : (tv[a] <= v[b]);
                if (takeLeft) { v[o] = tv[a]; idx_out[o++] = ti[a++]; }
                else { v[o] = v[b]; idx_out[o++] = idx_out[b++]; }
            }
            while (a < ll) { v[o] = tv[a]; idx_out[o++] = ti[a++]; }
        }
    }
    free(v); free(tv); free(ti);
}

/* ------------------------------------------------------------------------ */
/* HashValue<i64> (ops/hash.hpp:328-333): the splitmix64 finaliser.          */
/* HashJoin<i64> (ops/join.hpp:36-80,150-165): slots = nextpow2(max(2n,64)), */
/* chains pushed at the head (so the first equal key met by Probe is the    */
/* highest build row id with that key); Probe emits (i, row) for the first  */
/* equal key in probe order.  Also returns the wrapping i64 sum of          */
/* payload[row] over the matches (the join+SUM harness of the configs).     */
/* ------------------------------------------------------------------------ */
uint64_t vo_hash_i64(int64_t key) {
    uint64_t x = (uint64_t)key;
    x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ULL;
    x = (x ^ (x >> 27)) * 0x94d049bb133111ebULL;
    return x ^ (x >> 31);
}

static size_t vo_next_pow2(size_t v) {
    v--; v |= v >> 1; v |= v >> 2; v |= v >> 4; v |= v >> 8; v |= v >> 16; v |= v >> 32; return v + 1;
}

uint64_t vo_hashjoin_i64(const int64_t* build, size_t nb, const int64_t* payload,
                         const int64_t* probe, size_t np,
                         uint32_t* left_out, uint32_t* right_out, int64_t* payload_sum) {
    size_t slots = vo_next_pow2((nb * 2) > 64 ? nb * 2 : 64);
    int64_t* head = (int64_t*)malloc(slots * 8);
    int64_t* next = (int64_t*)malloc((nb ? nb : 1) * 8);
    for (size_t s = 0; s < slots; ++s) head[s] = -1;
    for (size_t i = 0; i < nb; ++i) {
        size_t s = (size_t)(vo_hash_i64(build[i]) & (slots - 1));
        next[i] = head[s]; head[s] = (int64_t)i;
    }
    uint64_t m = 0, sum = 0;
    for (size_t i = 0; i < np; ++i) {
        size_t s = (size_t)(vo_hash_i64(probe[i]) & (slots - 1));
        for (int64_t e = head[s]; e >= 0; e = next[e]) {
            if (build[e] == probe[i]) {
                if (left_out) left_out[m] = (uint32_t)i;
                if (right_out) right_out[m] = (uint32_t)e;
                if (payload) sum += (uint64_t)payload[e];
                ++m;
                break;
            }
        }
    }
    free(head); free(next);
    if (payload_sum) *payload_sum = (int64_t)sum;
    return m;
}

/* ------------------------------------------------------------------------ */
/* HashAggregator<TKey,f64> per-group state (ops/aggregate.hpp:76-140,       */
/* 176-239): Sum in row order, Min/Max with strict compares (first of equal */
/* values kept), Count, Welford Mean/M2.  The table layout only decides the */
/* output order (slot order, Finalize 249-261); the oracle emits groups in  */
/* ascending key order.  Also an exact SUM per group (accumulator digits).  */
/* Keys are i32/u32 bit patterns given as uint32.  Returns group count.     */
/* ------------------------------------------------------------------------ */
typedef struct { uint32_t key; uint64_t count; double sum, mn, mx, mean, m2; } vo_group;

static int vo_cmp_u32(const void* a, const void* b) {
    uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b; return x < y ? -1 : x > y;
}

size_t vo_groupby_f64(const uint32_t* keys, const double* vals, size_t n, int keys_signed,
                      size_t max_groups, uint32_t* out_key, uint64_t* out_count,
                      double* out_sum, double* out_min, double* out_max,
                      double* out_mean, double* out_m2, int64_t* out_digits) {
    /* distinct keys, sorted */
    uint32_t* uk = (uint32_t*)malloc((n ? n : 1) * 4);
    memcpy(uk, keys, n * 4);
    for (size_t i = 0; i < n && keys_signed; ++i) uk[i] ^= 0x80000000u;
    qsort(uk, n, 4, vo_cmp_u32);
    size_t g = 0;
    for (size_t i = 0; i < n; ++i) if (i == 0 || uk[i] != uk[i - 1]) uk[g++] = uk[i];
    if (g > max_groups) { free(uk); return g; }
    vo_group* st = (vo_group*)calloc(g ? g : 1, sizeof(vo_group));
    vo_acc* acc = (vo_acc*)calloc(g ? g : 1, sizeof(vo_acc));
    for (size_t i = 0; i < n; ++i) {
        uint32_t k = keys_signed ? (keys[i] ^ 0x80000000u) : keys[i];
        size_t lo = 0, hi = g;
        while (lo + 1 < hi) { size_t mid = (lo + hi) / 2; if (uk[mid] <= k) lo = mid; else hi = mid; }
        vo_group* s = &st[lo];
        double v = vals[i];
        if (s->count == 0) { s->sum = s->mn = s->mx = v; s->count = 1; s->mean = v; s->m2 = 0.0; }
        else {
            s->sum += v;
            if (v < s->mn) s->mn = v;
            if (v > s->mx) s->mx = v;
            s->count++;
            double d = v - s->mean; s->mean += d / (double)s->count; s->m2 += d * (v - s->mean);
        }
        vo_acc_add(&acc[lo], v);
    }
    for (size_t j = 0; j < g; ++j) {
        out_key[j] = keys_signed ? (uk[j] ^ 0x80000000u) : uk[j];
        out_count[j] = st[j].count; out_sum[j] = st[j].sum; out_min[j] = st[j].mn;
        out_max[j] = st[j].mx; out_mean[j] = st[j].mean; out_m2[j] = st[j].m2;
        if (out_digits) vo_acc_export(&acc[j], out_digits + j * VO_ACC_DIGITS);
    }
    free(uk); free(st); free(acc);
    return g;
}

/* splitmix64 (Steele, Lea, Flood 2014), used for the synthetic columns.     */
uint64_t vo_splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

/* ------------------------------------------------------------------------ */
/* Window family, the static methods of Engine<f64> (engine.hpp:183-300):   */
/* kind 0 WindowDelta, 1 WindowSum, 2 WindowMean, 3 WindowStdDev,           */
/* 4 WindowVariance, 5 WindowQuantile.  Returns the number of outputs       */
/* written; each window is summed left to right as in the reference.        */
/* Quantile: the lo-th and (lo+1)-th smallest of the first min(window, 128) */
/* values (nth_element then the minimum right of it); inputs without NaN.   */
/* ------------------------------------------------------------------------ */
static int vo_cmp_f64(const void* a, const void* b) {
    double x = *(const double*)a, y = *(const double*)b;
    return x < y ? -1 : x > y;
}

int64_t vo_window_f64(int kind, const double* d, size_t n, double* out, size_t out_len, int window, double q) {
    size_t w = (size_t)window, N, j, k;
    if (kind == 0) {
        if (n < 2) return 0;
        N = n - 1 < out_len ? n - 1 : out_len;
        for (j = 0; j < N; ++j) out[j] = d[j + 1] - d[j];
        return (int64_t)N;
    }
    if (kind == 1 || kind == 2) { if (w == 0 || n < w) return 0; }
    else if (kind == 3 || kind == 4) { if (w < 2 || n < w) return 0; }
    else if (w < 2 || n < w || q < 0.0 || q > 1.0) return 0;
    N = n - w + 1 < out_len ? n - w + 1 : out_len;
    for (j = 0; j < N; ++j) {
        if (kind <= 4) {
            double s = 0.0;
            for (k = 0; k < w; ++k) s += d[j + k];
            if (kind == 1) { out[j] = s; continue; }
            if (kind == 2) { out[j] = s / (double)w; continue; }
            double mean = s / (double)w, ssq = 0.0;
            for (k = 0; k < w; ++k) { double e = d[j + k] - mean; ssq += e * e; }
            out[j] = kind == 3 ? sqrt(ssq / (double)w) : ssq / (double)w;
        } else {
            double buf[128];
            size_t ww = w < 128 ? w : 128, lo, hi;
            double pos, vlo, vhi;
            memcpy(buf, d + j, ww * sizeof(double));
            qsort(buf, ww, sizeof(double), vo_cmp_f64);
            pos = q * (double)(ww - 1);
            lo = (size_t)pos;
            if (lo >= ww) lo = ww - 1;
            hi = lo + 1 < ww ? lo + 1 : lo;
            vlo = buf[lo];
            if (lo == hi) { out[j] = vlo; continue; }
            vhi = buf[lo + 1];
            out[j] = vlo + (vhi - vlo) * (pos - (double)lo);
        }
    }
    return (int64_t)N;
}

/* TimeWindowSum (engine.hpp:229-241): bucket = (ts - start) / duration     */
/* (truncating, so rows up to one duration before start fall in bucket 0); */
/* rows in [0, out_len) are added onto out in row order.                    */
int64_t vo_time_window_sum_f64(const int64_t* ts, const double* v, size_t n, int64_t start, int64_t dur, double* out,
                               size_t out_len) {
    size_t i, maxb = 0;
    if (dur <= 0 || n == 0 || out_len == 0) return 0;
    for (i = 0; i < n; ++i) {
        int64_t b = (int64_t)((uint64_t)ts[i] - (uint64_t)start) / dur;
        if (b >= 0 && (size_t)b < out_len) {
            out[b] += v[i];
            if ((size_t)b > maxb) maxb = (size_t)b;
        }
    }
    return (int64_t)(maxb + 1);
}

/* ------------------------------------------------------------------------ */
/* NullableSegment<T> (data/nulls.hpp:241-420) over data and NullBitmap      */
/* words (bit r & 63 of word r >> 6, 1 = valid; rows >= n ignored):          */
/* res = {SumValid (row order, in T), MinValid, MaxValid (first valid value, */
/* then replaced only by a smaller / larger one; NullSentinel = NaN / lowest */
/* when no row is valid)}, counts = {ValidCount, NullCount, rows kept by     */
/* CompactValid}, compact = the valid rows in order.                         */
/* ------------------------------------------------------------------------ */
#define VO_DEF_NULLABLE(NAME, T, SENT)                                                       \
void NAME(const T* d, size_t n, const uint64_t* w, T* res, uint64_t* counts, T* compact) {  \
    T sum = 0, mn = SENT, mx = SENT;                                                         \
    size_t valid = 0, i;                                                                     \
    for (i = 0; i < n; ++i) {                                                                \
        if (!((w[i >> 6] >> (i & 63)) & 1)) continue;                                        \
        sum += d[i];                                                                         \
        if (valid == 0) { mn = d[i]; mx = d[i]; }                                            \
        else { if (d[i] < mn) mn = d[i]; if (d[i] > mx) mx = d[i]; }                         \
        compact[valid++] = d[i];                                                             \
    }                                                                                        \
    res[0] = sum; res[1] = mn; res[2] = mx;                                                  \
    counts[0] = valid; counts[1] = n - valid; counts[2] = valid;                             \
}
VO_DEF_NULLABLE(vo_nullable_f64, double, (double)NAN)
VO_DEF_NULLABLE(vo_nullable_i64, int64_t, INT64_MIN)

/* ------------------------------------------------------------------------ */
/* Columnar encodings (data/encoding.hpp:104-527): Build and DecodeBatch.   */
/* Dictionary: the distinct values (== equality, so -0 and +0 are one entry */
/* holding the first one seen; NaN-free input) sorted ascending, one u32     */
/* index per row.  RLE: runs of == equal neighbours, u32 run lengths;        */
/* DecodeBatch writes T{} for rows past the last run and counts rows in a    */
/* u32 as the reference does.  Delta (integers): deltas from the previous    */
/* row (the first from 0), wrapping; decode = BaseValue + running sum.       */
/* BitPacked: u32 values in BitWidth-bit fields (width of the maximum, at    */
/* least 1), little-endian across u64 words.  FrameOfReference: Base = min,  */
/* Scale = (max - min) / INT32_MAX (1 when max <= min), i32 codes =          */
/* llround((x - Base) / Scale); decode = Base + (f64)code * Scale.           */
/* ------------------------------------------------------------------------ */
typedef struct { double v; size_t row; } vo_vr_f64;
typedef struct { int64_t v; size_t row; } vo_vr_i64;
static int vo_cmp_vr_f64(const void* a, const void* b) {
    const vo_vr_f64 *x = (const vo_vr_f64*)a, *y = (const vo_vr_f64*)b;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return x->row < y->row ? -1 : x->row > y->row;
}
static int vo_cmp_vr_i64(const void* a, const void* b) {
    const vo_vr_i64 *x = (const vo_vr_i64*)a, *y = (const vo_vr_i64*)b;
    if (x->v < y->v) return -1;
    if (x->v > y->v) return 1;
    return x->row < y->row ? -1 : x->row > y->row;
}
#define VO_DEF_DICT(SUF, T, VR, CMP)                                                         \
size_t vo_dict_build_##SUF(const T* x, size_t n, T* dict, uint32_t* idx) {                 \
    VR* p = (VR*)malloc((n ? n : 1) * sizeof(VR));                                           \
    size_t i, g = 0;                                                                         \
    for (i = 0; i < n; ++i) { p[i].v = x[i]; p[i].row = i; }                                 \
    qsort(p, n, sizeof(VR), CMP);                                                            \
    for (i = 0; i < n; ++i) {                                                                \
        if (i > 0 && p[i].v == p[i - 1].v) { idx[p[i].row] = (uint32_t)(g - 1); continue; } \
        dict[g] = p[i].v;                          /* smallest row of an equal group */     \
        idx[p[i].row] = (uint32_t)g++;                                                       \
    }                                                                                        \
    free(p);                                                                                 \
    return g;                                                                                \
}                                                                                            \
void vo_dict_decode_##SUF(const T* dict, const uint32_t* idx, size_t rows, size_t off,      \
                          size_t cnt, T* out) {                                              \
    size_t i;                                                                                \
    for (i = 0; i < cnt && off + i < rows; ++i) out[i] = dict[idx[off + i]];                 \
}                                                                                            \
size_t vo_rle_build_##SUF(const T* x, size_t n, T* vals, uint32_t* lens) {                 \
    size_t i, r = 0;                                                                         \
    if (n == 0) return 0;                                                                    \
    vals[0] = x[0]; lens[0] = 1;                                                             \
    for (i = 1; i < n; ++i) {                                                                \
        if (x[i] == vals[r]) { lens[r]++; continue; }                                        \
        ++r; vals[r] = x[i]; lens[r] = 1;                                                    \
    }                                                                                        \
    return r + 1;                                                                            \
}                                                                                            \
void vo_rle_decode_##SUF(const T* vals, const uint32_t* lens, size_t nr, size_t off,        \
                         size_t cnt, T* out) {                                               \
    uint32_t cum = 0;                                                                        \
    size_t run = 0, row = off, i;                                                            \
    while (run < nr) { cum += lens[run]; if (off < (size_t)cum) break; ++run; }              \
    for (i = 0; i < cnt; ++i) {                                                              \
        if (run >= nr) { out[i] = (T)0; continue; }                                          \
        out[i] = vals[run];                                                                  \
        ++row;                                                                               \
        if (row >= (size_t)cum) { ++run; if (run < nr) cum += lens[run]; }                   \
    }                                                                                        \
}
VO_DEF_DICT(f64, double, vo_vr_f64, vo_cmp_vr_f64)
VO_DEF_DICT(i64, int64_t, vo_vr_i64, vo_cmp_vr_i64)

#define VO_DEF_DELTA(SUF, T, U)                                                              \
void vo_delta_build_##SUF(const T* x, size_t n, T* deltas) {                                \
    U prev = 0;                                                                              \
    size_t i;                                                                                \
    for (i = 0; i < n; ++i) { deltas[i] = (T)((U)x[i] - prev); prev = (U)x[i]; }             \
}                                                                                            \
void vo_delta_decode_##SUF(T base, const T* deltas, size_t rows, size_t off, size_t cnt,     \
                           T* out) {                                                         \
    U run = (U)base;                                                                         \
    size_t i;                                                                                \
    for (i = 0; i < off && i < rows; ++i) run += (U)deltas[i];                               \
    for (i = 0; i < cnt && off + i < rows; ++i) { run += (U)deltas[off + i]; out[i] = (T)run; } \
}
VO_DEF_DELTA(i64, int64_t, uint64_t)
VO_DEF_DELTA(i32, int32_t, uint32_t)

size_t vo_bitpack_build(const uint32_t* x, size_t n, uint64_t* words, uint32_t* bw_out) {
    uint32_t mx = 0, bw = 0;
    size_t i, nw;
    if (n == 0) { *bw_out = 0; return 0; }
    for (i = 0; i < n; ++i) if (x[i] > mx) mx = x[i];
    while (mx) { ++bw; mx >>= 1; }
    if (bw == 0) bw = 1;
    nw = (n * bw + 63) / 64;
    memset(words, 0, nw * sizeof(uint64_t));
    for (i = 0; i < n; ++i) {
        uint64_t pos = (uint64_t)i * bw, v = (uint64_t)x[i] & ((1ull << bw) - 1);
        size_t w = pos / 64;
        unsigned sh = pos % 64;
        words[w] |= v << sh;
        if (sh + bw > 64 && w + 1 < nw) words[w + 1] |= v >> (64 - sh);
    }
    *bw_out = bw;
    return nw;
}
void vo_bitpack_decode(const uint64_t* words, size_t nw, uint32_t bw, size_t rows, size_t off, size_t cnt,
                       uint32_t* out) {
    size_t i;
    for (i = 0; i < cnt && off + i < rows; ++i) {
        uint64_t pos = (uint64_t)(off + i) * bw, r;
        unsigned sh = pos % 64;
        r = words[pos / 64] >> sh;
        if (sh + bw > 64 && pos / 64 + 1 < nw) r |= words[pos / 64 + 1] << (64 - sh);
        out[i] = (uint32_t)(r & ((1ull << bw) - 1));
    }
}

void vo_for_build(const double* x, size_t n, double* base, double* scale, int32_t* enc) {
    double mn, mx, range;
    size_t i;
    *base = 0.0; *scale = 1.0;
    if (n == 0) return;
    mn = mx = x[0];
    for (i = 1; i < n; ++i) { if (x[i] < mn) mn = x[i]; if (x[i] > mx) mx = x[i]; }
    *base = mn;
    range = mx - mn;
    *scale = range <= 0.0 ? 1.0 : range / (double)INT32_MAX;
    for (i = 0; i < n; ++i) enc[i] = (int32_t)llround((x[i] - *base) / *scale);
}
void vo_for_decode(double base, double scale, const int32_t* enc, size_t rows, size_t off, size_t cnt, double* out) {
    size_t i;
    for (i = 0; i < cnt && off + i < rows; ++i) out[i] = base + (double)enc[off + i] * scale;
}

/* ------------------------------------------------------------------------ */
/* Tier-0 loops whose load is a window word on Engine<f64> (handlers        */
/* engine.hpp:2589-2640 WSTD / WVARIANCE_W, 3167-3226 WDELTA / WINDOW_SUM / */
/* WINDOW_MEAN): wkind 1 WINDOW_SUM, 2 WINDOW_MEAN (sum * (1 / w)), 3 WSTD,  */
/* 4 WVARIANCE_W, 5 WDELTA.  Lane j of a group at offset o is the window    */
/* value at row o + j, 0 when the window passes Count; the word itself      */
/* advances the offset register (min(4, Count - o); WDELTA: the lanes it    */
/* wrote), and the loop ends when the offset equals total.  The body's      */
/* vector words and reductions are those of vo_tier0_expr_f64.  Returns 1,  */
/* or 0 when the reference would not terminate (a no-op window word, or an  */
/* offset that steps past total or stops short of it).                      */
/* ------------------------------------------------------------------------ */
static double vo_window_lane(int wkind, const double* d, uint64_t count, uint64_t start, int w) {
    double s = 0.0, mean, q = 0.0;
    int k;
    if (start + (uint64_t)w > count) return 0.0;
    for (k = 0; k < w; ++k) s += d[start + k];
    if (wkind == 1) return s;
    if (wkind == 2) return s * (1.0 / (double)w);
    mean = s / (double)w;
    for (k = 0; k < w; ++k) { double e = d[start + k] - mean; q += e * e; }
    return wkind == 3 ? sqrt(q / (double)w) : q / (double)w;
}

/* WQUANTILE (engine.hpp:2642-2664): element idx of the sorted window; 0 past Count */
static double vo_nth_lane(const double* d, uint64_t count, uint64_t start, int w, int idx) {
    double buf[32];
    int k, j;
    if (start + (uint64_t)w > count) return 0.0;
    for (k = 0; k < w; ++k) {
        double x = d[start + k];
        for (j = k; j > 0 && buf[j - 1] > x; --j) buf[j] = buf[j - 1];
        buf[j] = x;
    }
    return buf[idx];
}

/* wkind 6 (WQUANTILE): `window` is the word's imm12 & 0xFF (5-bit window,
   0 meaning 16; 3-bit quantile q selecting element (q * w) / 7). */
int vo_tier0_window_f64(const double* data, uint64_t count, uint64_t off, uint64_t total, int wkind, int window,
                        int vload_dst, int nops, const int* code, const double* sval, int nb, const int* bvreg,
                        const int* red, const int* accop, uint64_t* acc, uint64_t* part) {
    double v[16][4];
    int i, k, b;
    int qidx = 0;
    memset(v, 0, sizeof(v));
    if (wkind == 6) {
        int imm = window;
        window = imm & 0x1F;
        if (window == 0) window = 16;
        qidx = (((imm >> 5) & 7) * window) / 7;
        if (qidx >= window) qidx = window - 1;
    }
    if (wkind != 5 && (window == 0 || ((wkind == 3 || wkind == 4 || wkind == 6) && window < 2))) return 0;
    do {
        uint64_t adv;
        if (wkind == 5) {
            uint64_t cnt = 4;
            if (off + 4 >= count) cnt = count > off + 1 ? count - off - 1 : 0;
            for (i = 0; i < 4; ++i) v[vload_dst][i] = (uint64_t)i < cnt ? data[off + i + 1] - data[off + i] : 0.0;
            adv = cnt;
        } else {
            for (i = 0; i < 4; ++i)
                v[vload_dst][i] = wkind == 6 ? vo_nth_lane(data, count, off + (uint64_t)i, window, qidx)
                                             : vo_window_lane(wkind, data, count, off + (uint64_t)i, window);
            adv = count > off ? (count - off < 4 ? count - off : 4) : 0;
        }
        for (k = 0; k < nops; ++k) {
            const int* c = code + 5 * k;
            double r[4];
            for (i = 0; i < 4; ++i) r[i] = vo_expr_lane_f64(c[0], c[4], v[c[2]][i], v[c[3]][i], sval[k]);
            for (i = 0; i < 4; ++i) v[c[1]][i] = r[i];
        }
        for (b = 0; b < nb; ++b) {
            part[b] = vo_red_f64(v[bvreg[b]], red[b]);
            acc[b] = vo_tier0_accum(acc[b], part[b], accop[b]);
        }
        /* the offset stays when the word is past the column's end (advance 0):
           CMP then falls through only if it already equals the total */
        if (adv == 0) return off == total;
        if (off + adv > total) return 0;
        off += adv;
    } while (off != total);
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Hash / sort / join words of straight-line programs (Engine<T>,           */
/* engine.hpp:704-842 handlers, 1051-1071 HashElement/HashU64; dispatch of  */
/* 0xEE-0xFA, 0xFD at engine.hpp:3286-3355).  Segments are u64 bit arrays   */
/* of the engine type (f64 or i64) and are written in place by the sort     */
/* words, as the reference writes its borrowed segments.                    */
/*   HASH_INIT/HASH_BUILD (0xEE/0xF0): slot rd & 15 := the distinct          */
/*     HashElement values of segment ra with the first row of each;         */
/*     R[rd] = their number (unordered_map size).                           */
/*   HASH_PROBE/HASH_LOOKUP (0xEF/0xF1): R[rd] = first row whose hash        */
/*     equals HashElement(R[rb] as T) in slot ra, else 2^64 - 1.            */
/*   SORT_ASC/SORT_DESC (0xF2/0xF3): std::sort / std::sort(greater) of the   */
/*     segment; R[rd] = Count.  Values are sorted by the total order of     */
/*     their bits (-0.0 before +0.0 ascending; the reference leaves equal   */
/*     elements in unspecified order); NaN is refused (-1).                 */
/*   SORT_TOPK/SORT_BOTTOMK (0xF4/0xF5): std::partial_sort with greater /    */
/*     less over k = min((size_t)R[rb], imm & 0xFFF) rows, restated step by */
/*     step from libstdc++ (GCC 11, bits/stl_heap.h __make_heap,            */
/*     __adjust_heap, __push_heap, __pop_heap, __sort_heap; bits/stl_algo.h */
/*     __heap_select, __partial_sort), so rows past k keep libstdc++'s      */
/*     order; k == 0 does nothing; k > Count is refused (-1: the reference  */
/*     reads past the segment there).  R[rd] = k.                           */
/*   JOIN_HASH/JOIN_NESTED (0xF6/0xF8): number of (left, right) pairs with   */
/*     equal values (== : NaN never equal, -0.0 == +0.0).                   */
/*   JOIN_MERGE (0xF7): sum over values of min(left count, right count) --  */
/*     the merge walk over sorted copies; NaN refused (-1).                 */
/*   JOIN_SEMI/JOIN_ANTI (0xFA/0xF9): left rows whose value is / is not      */
/*     equal to some right value.                                           */
/*   PARTITION_HASH (0xFD): R[rd] = Count.                                  */
/* A word naming a missing segment does nothing (PC++ only).                */
/* ------------------------------------------------------------------------ */
uint64_t vo_hash_u64(uint64_t x) {
    x ^= x >> 33;
    x *= 0xFF51AFD7ED558CCDull;
    x ^= x >> 33;
    x *= 0xC4CEB9FE1A85EC53ull;
    x ^= x >> 33;
    return x;
}

static int vo_bits_nan(uint64_t b) { return ((b >> 52) & 0x7FF) == 0x7FF && (b & 0xFFFFFFFFFFFFFull) != 0; }

uint64_t vo_hash_element(uint64_t bits, int is_f64) {
    if (is_f64 && vo_bits_nan(bits)) return 0;
    return vo_hash_u64(bits);
}

typedef struct {
    uint64_t* h;      /* distinct hashes, ascending */
    uint64_t* row;    /* first row of each */
    uint64_t n;
    int valid;
} vo_htab;

typedef struct {
    vo_htab t[16];
} vo_relstate;

void* vo_relw_create(void) { return calloc(1, sizeof(vo_relstate)); }

void vo_relw_destroy(void* p) {
    vo_relstate* s = (vo_relstate*)p;
    int i;
    if (!s) return;
    for (i = 0; i < 16; ++i) { free(s->t[i].h); free(s->t[i].row); }
    free(s);
}

typedef struct { uint64_t h, row; } vo_hrow;
static int vo_hrow_cmp(const void* a, const void* b) {
    const vo_hrow* x = (const vo_hrow*)a; const vo_hrow* y = (const vo_hrow*)b;
    if (x->h != y->h) return x->h < y->h ? -1 : 1;
    return x->row < y->row ? -1 : (x->row > y->row);
}

/* Value order of T as an unsigned key: f64 total order of the bits (-0.0
 * below +0.0, NaN at the ends), i64 signed order.  canon: -0.0 -> +0.0. */
uint64_t vo_value_key(uint64_t b, int is_f64, int canon) {
    if (!is_f64) return b ^ 0x8000000000000000ull;
    if (canon && b == 0x8000000000000000ull) b = 0;
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
static uint64_t vo_value_unkey(uint64_t k, int is_f64) {
    if (!is_f64) return k ^ 0x8000000000000000ull;
    return (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
}
static int vo_u64_cmp(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return x < y ? -1 : (x > y);
}

/* comparator of the partial sort: desc = std::greater<T>, else std::less<T> */
static int vo_cmp_t(uint64_t a, uint64_t b, int is_f64, int desc) {
    if (is_f64) {
        double x, y;
        memcpy(&x, &a, 8); memcpy(&y, &b, 8);
        return desc ? x > y : x < y;
    }
    return desc ? (int64_t)a > (int64_t)b : (int64_t)a < (int64_t)b;
}

static void vo_push_heap(uint64_t* f, long hole, long top, uint64_t v, int F, int desc) {
    long parent = (hole - 1) / 2;
    while (hole > top && vo_cmp_t(f[parent], v, F, desc)) {
        f[hole] = f[parent];
        hole = parent;
        parent = (hole - 1) / 2;
    }
    f[hole] = v;
}

static void vo_adjust_heap(uint64_t* f, long hole, long len, uint64_t v, int F, int desc) {
    const long top = hole;
    long child = hole;
    while (child < (len - 1) / 2) {
        child = 2 * (child + 1);
        if (vo_cmp_t(f[child], f[child - 1], F, desc)) child--;
        f[hole] = f[child];
        hole = child;
    }
    if ((len & 1) == 0 && child == (len - 2) / 2) {
        child = 2 * (child + 1);
        f[hole] = f[child - 1];
        hole = child - 1;
    }
    vo_push_heap(f, hole, top, v, F, desc);
}

void vo_partial_sort(uint64_t* f, uint64_t k, uint64_t n, int is_f64, int desc) {
    long len = (long)k, parent;
    uint64_t i;
    if (len >= 2) {                                   /* __make_heap */
        parent = (len - 2) / 2;
        for (;;) {
            vo_adjust_heap(f, parent, len, f[parent], is_f64, desc);
            if (parent == 0) break;
            parent--;
        }
    }
    for (i = k; i < n; ++i)                           /* __heap_select */
        if (vo_cmp_t(f[i], f[0], is_f64, desc)) {     /* __pop_heap(first, middle, i) */
            uint64_t v = f[i];
            f[i] = f[0];
            vo_adjust_heap(f, 0, len, v, is_f64, desc);
        }
    while (len > 1) {                                 /* __sort_heap */
        uint64_t v;
        --len;
        v = f[len];
        f[len] = f[0];
        vo_adjust_heap(f, 0, len, v, is_f64, desc);
    }
}

/* number of rows of sorted keys s[0, n) equal to key: [lo, hi) */
static void vo_equal_range(const uint64_t* s, uint64_t n, uint64_t key, uint64_t* lo, uint64_t* hi) {
    uint64_t a = 0, b = n;
    while (a < b) { uint64_t m = a + (b - a) / 2; if (s[m] < key) a = m + 1; else b = m; }
    *lo = a;
    b = n;
    while (a < b) { uint64_t m = a + (b - a) / 2; if (s[m] <= key) a = m + 1; else b = m; }
    *hi = a;
}

/* kind: 0 pairs (JOIN_HASH / JOIN_NESTED), 1 merge, 2 semi, 3 anti.
 * Returns -1 for JOIN_MERGE over NaN. */
int64_t vo_join_count(int kind, const uint64_t* l, uint64_t nl, const uint64_t* r, uint64_t nr, int is_f64) {
    uint64_t* rs;
    uint64_t* ls = NULL;
    uint64_t i, m = 0, total = 0;
    if (nl == 0 || nr == 0) return kind == 3 ? (int64_t)nl : 0;   /* no pair to test: defined even with NaN */
    rs = (uint64_t*)malloc(nr * 8);
    for (i = 0; i < nr; ++i) {
        if (is_f64 && vo_bits_nan(r[i])) { if (kind == 1) { free(rs); return -1; } continue; }
        rs[m++] = vo_value_key(r[i], is_f64, 1);
    }
    qsort(rs, m, 8, vo_u64_cmp);
    if (kind == 1) {
        uint64_t nls = 0;
        ls = (uint64_t*)malloc((nl ? nl : 1) * 8);
        for (i = 0; i < nl; ++i) {
            if (is_f64 && vo_bits_nan(l[i])) { free(rs); free(ls); return -1; }
            ls[nls++] = vo_value_key(l[i], is_f64, 1);
        }
        qsort(ls, nls, 8, vo_u64_cmp);
        for (i = 0; i < nls; ++i) {             /* the i-th of its run matches while the right run lasts */
            uint64_t llo, lhi, rlo, rhi;
            vo_equal_range(ls, nls, ls[i], &llo, &lhi);
            vo_equal_range(rs, m, ls[i], &rlo, &rhi);
            if (i - llo < rhi - rlo) total++;
        }
        free(ls);
        free(rs);
        return (int64_t)total;
    }
    for (i = 0; i < nl; ++i) {
        uint64_t lo, hi;
        if (is_f64 && vo_bits_nan(l[i])) { if (kind == 3) total++; continue; }
        vo_equal_range(rs, m, vo_value_key(l[i], is_f64, 1), &lo, &hi);
        if (kind == 0) total += hi - lo;
        else if (kind == 2) total += hi > lo;
        else total += hi == lo;
    }
    free(rs);
    return (int64_t)total;
}

/* Straight-line program of scalar words and the words above over nseg
 * segments; returns 1 at HALT / end, 0 when a DIV/MOD by zero stopped it,
 * -1 for a refused word (segments may already be sorted by earlier words). */
int vo_relw_run(void* state, uint64_t* const* segs, const uint64_t* counts, int nseg, int is_f64, const uint32_t* w,
                size_t n, uint64_t* r) {
    vo_relstate* st = (vo_relstate*)state;
    size_t k;
    for (k = 0; k < n; ++k) {
        const uint32_t x = w[k];
        const unsigned op = x & 0xFF, rd = (x >> 8) & 0xF, ra = (x >> 12) & 0xF, rb = (x >> 16) & 0xF;
        const unsigned imm = (x >> 20) & 0xFFF;
        if (op == 0x01) return 1;
        switch (op) {
        case 0xEE: case 0xF0: {                                  /* HASH_INIT / HASH_BUILD */
            vo_htab* t;
            vo_hrow* hr;
            uint64_t i, c, m = 0;
            if ((int)ra >= nseg) break;
            t = &st->t[rd & 0xF];
            c = counts[ra];
            hr = (vo_hrow*)malloc((c ? c : 1) * sizeof(vo_hrow));
            for (i = 0; i < c; ++i) { hr[i].h = vo_hash_element(segs[ra][i], is_f64); hr[i].row = i; }
            qsort(hr, c, sizeof(vo_hrow), vo_hrow_cmp);
            free(t->h); free(t->row);
            t->h = (uint64_t*)malloc((c ? c : 1) * 8);
            t->row = (uint64_t*)malloc((c ? c : 1) * 8);
            for (i = 0; i < c; ++i)
                if (i == 0 || hr[i].h != hr[i - 1].h) { t->h[m] = hr[i].h; t->row[m] = hr[i].row; m++; }
            t->n = m;
            t->valid = 1;
            free(hr);
            r[rd] = m;
            break;
        }
        case 0xEF: case 0xF1: {                                  /* HASH_PROBE / HASH_LOOKUP */
            const vo_htab* t = &st->t[ra & 0xF];
            uint64_t res = ~0ull;
            if (t->valid) {
                const uint64_t h = vo_hash_element(r[rb], is_f64);
                uint64_t lo, hi;
                vo_equal_range(t->h, t->n, h, &lo, &hi);
                if (hi > lo) res = t->row[lo];
            }
            r[rd] = res;
            break;
        }
        case 0xF2: case 0xF3: {                                  /* SORT_ASC / SORT_DESC */
            uint64_t i, c;
            if ((int)ra >= nseg) break;
            c = counts[ra];
            for (i = 0; i < c; ++i) {
                if (is_f64 && vo_bits_nan(segs[ra][i])) return -1;
                segs[ra][i] = vo_value_key(segs[ra][i], is_f64, 0);
            }
            qsort(segs[ra], c, 8, vo_u64_cmp);
            for (i = 0; i < c; ++i) segs[ra][i] = vo_value_unkey(segs[ra][i], is_f64);
            if (op == 0xF3)
                for (i = 0; i < c / 2; ++i) { uint64_t tmp = segs[ra][i]; segs[ra][i] = segs[ra][c - 1 - i]; segs[ra][c - 1 - i] = tmp; }
            r[rd] = c;
            break;
        }
        case 0xF4: case 0xF5: {                                  /* SORT_TOPK (greater) / SORT_BOTTOMK (less) */
            uint64_t kk = r[rb] < (uint64_t)imm ? r[rb] : (uint64_t)imm;
            if ((int)ra >= nseg || kk == 0) break;
            if (kk > counts[ra]) return -1;
            vo_partial_sort(segs[ra], kk, counts[ra], is_f64, op == 0xF4);
            r[rd] = kk;
            break;
        }
        case 0xF6: case 0xF7: case 0xF8: case 0xF9: case 0xFA: { /* JOIN_* */
            static const int kind_of[5] = {0, 1, 0, 3, 2};
            int64_t m;
            if ((int)ra >= nseg || (int)rb >= nseg) break;
            m = vo_join_count(kind_of[op - 0xF6], segs[ra], counts[ra], segs[rb], counts[rb], is_f64);
            if (m < 0) return -1;
            r[rd] = (uint64_t)m;
            break;
        }
        case 0xFD:                                               /* PARTITION_HASH */
            if ((int)ra < nseg) r[rd] = counts[ra];
            break;
        default: {
            int rc = vo_scalar_run(&w[k], 1, r);
            if (rc != 1) return rc;
        }
        }
    }
    return 1;
}
