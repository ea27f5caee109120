// vx_segagg.hpp -- device helpers of the segment aggregates AGG_* (the
// reference's Engine<T>::HandleAggregate, exec/engine.hpp:499-698): Welford
// moments, order keys, sort and distinct keys over a device column range.
// The slot logic (AggState semantics) is host code in vx_capi.cpp.
#pragma once
#include <stdint.h>

#include "vx_expr_abi.hpp"

namespace vx {

struct StreamCtx;
struct ScanOut;

// Moments of rows as f64: count, mean, sum of squared deviations.
struct Moments {
    double n, mean, m2;
};

// Device work buffers of one engine's aggregates (grown as needed).
struct SegaggBufs {
    uint64_t* k1 = nullptr;
    uint64_t* k2 = nullptr;
    uint32_t* v1 = nullptr;
    uint32_t* v2 = nullptr;
    uint64_t cap = 0;
    Moments* parts = nullptr;        // per-workgroup moments
    Moments* parts_host = nullptr;
    ~SegaggBufs();
    bool ensure(uint64_t n);
};

// x[0, n) (dtype DT_F64 or DT_I64) -> moments, merged in workgroup order.
bool segagg_moments(int dtype, const void* x, uint64_t n, Moments* out, SegaggBufs& b, StreamCtx* c);

// x[0, n) -> its order keys, ascending, in b.k2 (f64: -0 as +0; i64: sign
// bit flipped).  n < 2^30.
bool segagg_sorted_keys(int dtype, const void* x, uint64_t n, SegaggBufs& b, StreamCtx* c);

// Ascending keys k[0, n) -> the distinct ones in out (device), *nout.
bool segagg_unique(const uint64_t* k, uint64_t n, uint64_t* out, uint64_t* nout, StreamCtx* c);

// f64 order keys k[0, n) -> the values' bits in out (device).
bool segagg_unkey_f64(const uint64_t* k, uint64_t n, uint64_t* out, StreamCtx* c);

// AGG_MODE over x[0, n): entries of equal values (-0.0 with +0.0; each NaN
// row its own entry); *ties = entries with the largest row count, *row =
// first row (row order) of the earliest-sorting of them.  n < 2^30.
bool segagg_mode(int dtype, const void* x, uint64_t n, SegaggBufs& b, uint64_t* row, uint64_t* ties, StreamCtx* c);

// The words over rows off + i, i < n (as expr_eval), reduced without writing
// the evaluated values: res[q] (device) = result of reduction q over every
// row (VSUM: sum pair / wrapping sum; VCOUNT: cnt; VRED_MIN / VRED_MAX: the
// signed min and max, then the unsigned min and max of the group values'
// bits in the first four words), last[q] (device) = the reduction of the
// last group of four rows (the loop's partial register; not written when n
// = 0).  Uses the stream's scan partials.
bool expr_reduce(int dtype, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog,
                 const ExprRed& rd, ScanOut* res, uint64_t* last, int* trap, StreamCtx* c);

// The compiled form of a loop body (vx_jit.cpp): launches it over rows off +
// i, i < n, with nb workgroups writing 8-word scan partials (reduction q's
// at parts + 8 * q * nb) and the last group's values into last; *launched =
// false when the loop is left to the interpreter kernels.
bool jit_expr_launch(bool f64, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog,
                     const ExprRed& rd, uint64_t* parts, unsigned nb, uint64_t* last, int* trap, void* stream,
                     bool* launched);

// out[i] = register `prog.target` after the words, for rows off + i, i < n
// (the loaded value is 0 for rows at or past count; n a multiple of 4, the
// rows off + 4g .. off + 4g + 3 one group).  dtype DT_F64 / DT_I64.
// *trap (device) is set to 1 when an i64 lane divides INT64_MIN by -1.
bool expr_eval(int dtype, const void* col, uint64_t count, uint64_t off, uint64_t n, const ExprProg& prog, void* out,
               int* trap, StreamCtx* c);

// VSTORE of an evaluated vector: src[i] (i < n, lane i mod 4 below count)
// -> dst[off + i] for off + i < seg_count.  dtype DT_F64 / DT_I64.
bool expr_store(int dtype, const void* src, uint64_t n, uint32_t count, void* dst, uint64_t off, uint64_t seg_count,
                StreamCtx* c);

// The window family's kernels on a device f64 column (vx_window.hip; the
// caller holds the library lock): kind 0 delta, 1 sum, 2 mean (sum / w),
// 3 stddev, 4 variance, 5 quantile q, 6 mean as sum * (1 / w).  Writes
// out[0, N) and returns N (the reference's output count), -1 on error.
int64_t window_launch(int kind, const double* d, uint64_t count, uint64_t out_len, int window, double* out,
                      StreamCtx* c, double q = 0.0);

// Order key <-> value bits (host).
uint64_t order_key(uint64_t bits, int dtype);
uint64_t order_unkey(uint64_t key, int dtype);

}  // namespace vx
