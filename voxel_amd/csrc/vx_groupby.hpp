// vx_groupby.hpp -- group-by records shared by vx_groupby.hip and the C ABI.
#pragma once
#include <stdint.h>

namespace vx {

// Largest key range handled by the LDS-privatised dense path: eight wave
// copies of kMaxDense + 1 36-byte bins (144 KB) fit in a CU's 160 KB of LDS.
constexpr int kMaxDense = 512;

// Aggregate state of one group: compensated sum pair, count, min, max.
struct GBin {
    double s, c;
    double mn, mx;
    uint64_t cnt;
};

struct StreamCtx;

// Group-by of f64 values by i32/u32/i64 keys over device columns.  On success
// returns the number of groups and points *okey / *orec at internal device
// arrays holding them in ascending key order (valid until the next call on
// the stream).  force_general skips the dense path.
int64_t groupby_f64(int key_dtype, const void* keys, const double* vals, uint64_t n, int64_t** okey, GBin** orec,
                    StreamCtx* c, int force_general);
// The dense path of an i32/u32 key column decided on the device (bins by
// key mod kMaxDense, key range checked afterwards), finished straight into
// the caller's output arrays (any may be NULL; at most max_groups entries
// written) with one host synchronisation.  Returns the group count, -2 when
// the key range is wider than kMaxDense (the caller runs groupby_f64), -1
// on error.
int64_t groupby_dense_device(int key_dtype, const void* keys, const double* vals, uint64_t n, uint64_t max_groups,
                             int64_t* out_keys, uint64_t* out_count, double* out_sum, double* out_avg,
                             double* out_min, double* out_max, StreamCtx* c);
// Merges two ascending record sets; result in internal arrays as above.
int64_t groupby_merge(const int64_t* ka, const GBin* ra, uint64_t na, const int64_t* kb, const GBin* rb,
                      uint64_t nb, int64_t** okey, GBin** orec, StreamCtx* c);
// Finishes records into user arrays (device): sum = s + c, avg = sum / count.
int groupby_finish(const int64_t* keys, const GBin* rec, uint64_t n, int64_t* out_keys, uint64_t* out_count,
                   double* out_sum, double* out_avg, double* out_min, double* out_max, void* stream_ctx);

}  // namespace vx
