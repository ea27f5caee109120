/*
 * voxel_c.h -- C ABI of libvoxel_c.so, the MI355X (gfx950) execution backend
 * for VoxelVM.
 *
 * Part 1 declares the 64 symbols of the reference's C ABI with identical
 * names and signatures (reference: bindings/voxel_c.h:14-114), so Go (cgo),
 * Java (FFM) and C callers link unchanged.  Each block cites the reference
 * declaration it replaces.  Differences in behaviour are listed per function:
 * the reference's CPU loops are replaced by GPU kernels, and programs the GPU
 * lowering does not recognise set voxel_last_error() instead of running.
 *
 * Part 2 adds voxel_gpu_* functions (no reference counterpart) for what the
 * reference ABI cannot express: device-resident columns, f32/i32 columns,
 * two-column predicates, u64 key sort with payload, i64 join with payload
 * sum, per-group results, and batched/served queries.
 *
 * Conventions (all functions):
 *  - no C++ exceptions cross this boundary;
 *  - errors set a thread-local message readable with voxel_last_error();
 *    functions returning int return 0 on success and -1 on error;
 *  - "device pointer" means memory on the current HIP device (hipMalloc'ed,
 *    or a torch CUDA tensor's data_ptr());
 *  - a `void* stream` argument is a hipStream_t; NULL means the device's
 *    default stream (also used by every Part 1 function).
 */
#pragma once

#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ====================================================================== */
/* Part 1 -- reference ABI                                                 */
/* ====================================================================== */

/* Engine (f64) -- reference bindings/voxel_c.h:14-24.
 * add_segment borrows the caller's pointer exactly as the reference does
 * (voxel_c.cpp:18-20); the rows a run needs are copied to HBM at each run
 * (the caller may have changed them).  voxel_gpu_engine_set_resident() opts
 * out of the re-copy.  run: fast-path programs (engine.hpp:84-141), Tier-0
 * loops (filters, reductions, vector arithmetic, window words, scalar words
 * around the loop) and straight-line programs of AGG_*, HASH_*, SORT_*,
 * TOPK / BOTTOMK, JOIN_* and scalar words run on the GPU with the
 * reference's register effects (sort words sort the caller's rows in place,
 * as the reference does); any other program, and loops that would not
 * terminate in the reference, set voxel_last_error() and leave the
 * registers untouched.  run_parallel (engine.hpp:143-175): fast-path
 * programs as chunked scans, other lowered programs as one fresh engine per
 * chunk with copied registers, R[result] summed in T in worker order. */
void*     voxel_engine_create_f64(void);
void      voxel_engine_destroy(void* engine);
size_t    voxel_engine_add_segment(void* engine, const double* data, size_t count);
void      voxel_engine_load_program(void* engine, const uint32_t* code, size_t len);
void      voxel_engine_run(void* engine);
void      voxel_engine_run_parallel(void* engine, size_t seg_id, uint32_t threads, uint8_t result_reg);
void      voxel_engine_set_scalar(void* engine, size_t reg, uint64_t value);
uint64_t  voxel_engine_get_scalar(void* engine, size_t reg);
void      voxel_engine_set_scalar_f64(void* engine, size_t reg, double value);
double    voxel_engine_get_scalar_f64(void* engine, size_t reg);
size_t    voxel_engine_k_lanes(void);

/* Engine (i64) -- reference bindings/voxel_c.h:29-35. */
void*     voxel_engine_create_i64(void);
void      voxel_engine_destroy_i64(void* engine);
size_t    voxel_engine_add_segment_i64(void* engine, const int64_t* data, size_t count);
void      voxel_engine_load_program_i64(void* engine, const uint32_t* code, size_t len);
void      voxel_engine_run_i64(void* engine);
void      voxel_engine_set_scalar_i64(void* engine, size_t reg, uint64_t value);
uint64_t  voxel_engine_get_scalar_i64(void* engine, size_t reg);

/* JIT -- reference bindings/voxel_c.h:40-41 (voxel_c.cpp:80-101).  The fused
 * VLOAD/VFILTER/VSUM/ADDF/ADD/CMP/JNZ loop (codegen/x86_64.hpp:1600-1716) is
 * lowered to the GPU scan with the JIT's semantics (8-row iterations from
 * R2 = data_len, ordered compares, initial accumulator ignored).  Returns 0.0
 * with voxel_last_error() set where the reference returns 0.0 (no fusable
 * loop) and where it would run away (data_len == 0). */
double    voxel_jit_run(const uint32_t* code, size_t code_len,
                        const double* data, size_t data_len, double threshold);

/* Instruction encoding -- reference bindings/voxel_c.h:46-65 (pure). */
uint32_t  voxel_instr_vload(uint8_t vd, uint8_t ra, uint8_t seg_id, uint8_t count);
uint32_t  voxel_instr_vfilter_gt(uint8_t vd, uint8_t va, uint8_t rb);
uint32_t  voxel_instr_vfilter_ge(uint8_t vd, uint8_t va, uint8_t rb);
uint32_t  voxel_instr_vfilter_lt(uint8_t vd, uint8_t va, uint8_t rb);
uint32_t  voxel_instr_vfilter_le(uint8_t vd, uint8_t va, uint8_t rb);
uint32_t  voxel_instr_vsum(uint8_t rd, uint8_t va);
uint32_t  voxel_instr_vred_min(uint8_t rd, uint8_t va);
uint32_t  voxel_instr_vred_max(uint8_t rd, uint8_t va);
uint32_t  voxel_instr_vcount(uint8_t rd, uint8_t va);
uint32_t  voxel_instr_addf(uint8_t rd, uint8_t ra, uint8_t rb);
uint32_t  voxel_instr_add(uint8_t rd, uint8_t ra, uint8_t rb);
uint32_t  voxel_instr_sub(uint8_t rd, uint8_t ra, uint8_t rb);
uint32_t  voxel_instr_mul(uint8_t rd, uint8_t ra, uint8_t rb);
uint32_t  voxel_instr_mov(uint8_t rd, int16_t imm);
uint32_t  voxel_instr_movr(uint8_t rd, uint8_t ra);
uint32_t  voxel_instr_cmp(uint8_t ra, uint8_t rb);
uint32_t  voxel_instr_jnz(int16_t offset);
uint32_t  voxel_instr_jz(int16_t offset);
uint32_t  voxel_instr_jmp(int16_t offset);
uint32_t  voxel_instr_halt(void);

/* NullBitmap -- reference bindings/voxel_c.h:70-76 (host-side utility). */
void*     voxel_null_create(size_t count);
void      voxel_null_destroy(void* bitmap);
int       voxel_null_is_null(const void* bitmap, size_t row);
void      voxel_null_set_null(void* bitmap, size_t row);
void      voxel_null_set_valid(void* bitmap, size_t row);
size_t    voxel_null_null_count(const void* bitmap);
size_t    voxel_null_valid_count(const void* bitmap);

/* HashAggregator<u32,f64> -- reference bindings/voxel_c.h:81-85.  Rows are
 * accumulated into per-key states on the GPU (groupby kernels); group_count
 * is the number of distinct keys seen, as in ops/aggregate.hpp. */
void*     voxel_agg_create(void);
void      voxel_agg_destroy(void* agg);
void      voxel_agg_init(void* agg, size_t expected_groups);
void      voxel_agg_accumulate(void* agg, const uint32_t* keys, const double* values, size_t count);
size_t    voxel_agg_group_count(const void* agg);

/* Sort & TopK -- reference bindings/voxel_c.h:90-92.  voxel_sort_* is the
 * stable argsort of SortOperator (ops/sort.hpp:506-541), computed by the GPU
 * radix sort on order-preserving f64 keys (-0.0 ordered equal to +0.0). */
void      voxel_sort_ascending(const double* data, size_t count, uint32_t* indices_out);
void      voxel_sort_descending(const double* data, size_t count, uint32_t* indices_out);
void      voxel_topk(const double* data, size_t count, size_t k, double* out, int largest);

/* DictionaryEncoding<f64> -- reference bindings/voxel_c.h:97-101 (host). */
void*     voxel_dict_create(void);
void      voxel_dict_destroy(void* dict);
void      voxel_dict_build(void* dict, const double* data, size_t count);
size_t    voxel_dict_cardinality(const void* dict);
void      voxel_dict_decode_batch(void* dict, double* out, size_t offset, size_t count);

/* ThreadPool -- reference bindings/voxel_c.h:106-109 (host). */
void*     voxel_pool_create(uint32_t threads);
void      voxel_pool_destroy(void* pool);
void      voxel_pool_wait_all(void* pool);
size_t    voxel_pool_thread_count(const void* pool);

/* Error handling -- reference bindings/voxel_c.h:114.  Unlike the reference
 * (whose buffer is never written, voxel_c.cpp:7,227) the message is set by
 * every failing call; run, run_parallel, jit_run and the voxel_gpu_* calls
 * clear it when they start, so it is empty after a successful call. */
const char* voxel_last_error(void);

/* ====================================================================== */
/* Part 2 -- GPU extensions (no reference counterpart)                     */
/* ====================================================================== */

enum {
    VOXEL_GPU_F64 = 0, VOXEL_GPU_F32 = 1, VOXEL_GPU_I64 = 2, VOXEL_GPU_I32 = 3,
    VOXEL_GPU_U64 = 4, VOXEL_GPU_U32 = 5
};
/* Predicate codes: the reference's kCmp* values (opcodes.hpp:415-420). */
enum {
    VOXEL_CMP_EQ = 0, VOXEL_CMP_NE = 1, VOXEL_CMP_LT = 2, VOXEL_CMP_LE = 3,
    VOXEL_CMP_GT = 4, VOXEL_CMP_GE = 5, VOXEL_CMP_TRUE = 6
};

/* 1 when a HIP device is usable, 0 otherwise (message in voxel_last_error). */
int       voxel_gpu_available(void);
/* Device memory helpers (for callers without another allocator). */
void*     voxel_gpu_malloc(size_t bytes);
void      voxel_gpu_free(void* dev);
int       voxel_gpu_memcpy_h2d(void* dev, const void* host, size_t bytes);
int       voxel_gpu_memcpy_d2h(void* host, const void* dev, size_t bytes);
int       voxel_gpu_synchronize(void);

/* Fills a device column with deterministic uniform values: row i uses the
 * i-th output u of SplitMix64 seeded with `seed` (z = seed + (i+1) *
 * 0x9E3779B97F4A7C15, then the SplitMix64 finaliser); floats get
 * lo + (hi - lo) * ((u >> 11) * 2^-53) (f32: rounded), integers get
 * lo + floor(u * (hi - lo) / 2^64).  stream may be NULL. */
int       voxel_gpu_fill_uniform(int dtype, void* dev, size_t n, uint64_t seed, double lo, double hi,
                                 void* stream);

/* Typed engine for i32 columns (Engine<i32>, fast path only). dtype is
 * VOXEL_GPU_I32, VOXEL_GPU_I64 or VOXEL_GPU_F64; add_segment_raw borrows. */
void*     voxel_gpu_engine_create(int dtype);
size_t    voxel_gpu_engine_add_segment_raw(void* engine, const void* data, size_t count);
/* Declares a segment immutable until the engine is destroyed: it is copied
 * to HBM once and not re-copied at each run. */
int       voxel_gpu_engine_set_resident(void* engine, size_t seg_id, int resident);
/* Attaches a device-resident column (no copies at all). */
size_t    voxel_gpu_engine_add_device_segment(void* engine, const void* dev, size_t count);

/* 1 when voxel_jit_run would lower the program, else 0 (reason in
 * voxel_last_error); backs voxel_py.jit_compile. */
int       voxel_gpu_jit_lowerable(const uint32_t* code, size_t len);
/* voxel_jit_run with R2 given separately from the row count (the pybind
 * jit_run form, voxel_bindings.cpp:342-366: R2 = total_count). */
double    voxel_gpu_jit_run_r2(const uint32_t* code, size_t code_len, const double* data,
                               size_t data_len, uint64_t r2, double threshold);
/* How load_program lowered the engine's program: 1 fast path, 2 Tier-0
 * loop, 3 straight-line scalar and segment-aggregate (AGG_*) words, 0 not
 * lowerable (run will fail with the reason). */
int       voxel_gpu_engine_lowered(const void* engine);
/* Tier-0 loops with vector words whose reductions are all VSUM / VCOUNT run
 * as one fused pass.  mode 1 (the default): a loop of at least 2^20 rows is
 * compiled whole at run time (hipRTC, once per shape of its words, kept for
 * the process); smaller loops, and any loop whose compile fails, run on the
 * library's interpreter kernels.  mode 2: always compiled (a compile failure
 * fails the run); mode 0: never.  Returns the previous mode, -1 for a bad
 * one.  No counterpart in the reference (a tuning switch of this backend,
 * as jit_compile picks the reference's JIT). */
int       voxel_gpu_set_expr_jit(int mode);

/* Stable sort of (u64 key, u32 payload) pairs by key (RadixSort<u64>::Sort
 * with indices, ops/sort.hpp:83-160) over device buffers; vals_in NULL
 * means payload = row index.  Inputs are not modified.  n < 2^30. */
int       voxel_gpu_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                                   uint32_t* vals_out, size_t n, void* stream);
/* out[i] = x[idx[i]] for i < k (device arrays): the rows an argsort
 * ordered (no reference counterpart; the SQL front end's ORDER BY). */
int       voxel_gpu_gather_f64(const double* x_dev, const uint32_t* idx_dev, size_t k, double* out_dev,
                               void* stream);
/* out[i] = (int64) x[i], toward zero; NaN and values outside int64 give
 * INT64_MIN (x86-64's cvttsd2si, as the reference's casts compile).  The
 * SQL front end's GROUP BY of an f64 key column. */
int       voxel_gpu_f64_to_i64(const double* x_dev, size_t n, int64_t* out_dev, void* stream);
/* Stable f64 argsort over a device column (SortOperator order). */
int       voxel_gpu_argsort_f64(const double* x_dev, size_t n, int descending, uint32_t* idx_out_dev,
                                void* stream);

/* Relational filter + aggregate over device columns:
 * rows with op_a(a[i], t_a) AND op_b(b[i], t_b); COUNT, SUM(a), MIN(a),
 * MAX(a).  Thresholds are bit patterns of the column type (f32/i32 in the
 * low 32 bits).  Float sums are compensated in the column's own type; MIN/MAX
 * skip NaN.  b == a applies both predicates to column a (read once). */
typedef struct voxel_gpu_scan_desc {
    int32_t  dtype;
    int32_t  op_a, op_b;       /* op_b = VOXEL_CMP_TRUE when b is NULL */
    int32_t  blocks;           /* workgroups; 0 = library default */
    int32_t  unroll;           /* 16-byte loads in flight per lane and column: 2, 4, 8; 0 = default */
    int32_t  nontemporal;      /* 1 = non-temporal (streaming) loads, -1 = cached loads, 0 = default */
    int32_t  layout;           /* 0 = default (grid-stride tiles, partials combined by a second kernel),
                                  1 = contiguous tiles per workgroup + last-workgroup combine,
                                  2 = grid-stride + second kernel, 3 = contiguous + second kernel,
                                  4 = grid-stride + last-workgroup combine */
    int32_t  pad_;
    const void* a;
    const void* b;
    uint64_t n;
    uint64_t t_a, t_b;
} voxel_gpu_scan_desc;

typedef struct voxel_gpu_agg {
    uint64_t count;
    double   sum;              /* floats: s + c rounded to the column type */
    double   sum_hi, sum_lo;   /* floats: the compensated pair (s, c) */
    int64_t  isum;             /* ints: wrapping 64-bit sum (i32 widened) */
    double   fmin, fmax;       /* floats */
    int64_t  imin, imax;       /* ints */
    int32_t  has_minmax;
    int32_t  pad;
} voxel_gpu_agg;

/* Launch + wait + read back. */
int       voxel_gpu_scan(const voxel_gpu_scan_desc* d, void* stream, voxel_gpu_agg* out);
/* Enqueue only (result stays on the device until voxel_gpu_scan_result). */
int       voxel_gpu_scan_launch(const voxel_gpu_scan_desc* d, void* stream);
int       voxel_gpu_scan_result(void* stream, voxel_gpu_agg* out);

/* Allocation-free voxel_gpu_scan_launch.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The same
 * scan kernels are enqueued on `stream` (NULL = the default stream) with
 * their per-workgroup partials in a fixed static device buffer of
 * VOXEL_GPU_NOALLOC_MAX_BLOCKS x 64 bytes per GPU, held until the process
 * ends, instead of library-allocated scratch; a last one-thread kernel writes
 * the result as a voxel_gpu_agg (the same fields voxel_gpu_scan returns) into
 * out_dev, caller device memory of sizeof(voxel_gpu_agg) bytes.  Nothing is
 * allocated and nothing waits: read out_dev after the stream's work is done.
 * Fails (-1, voxel_last_error) for d->blocks above VOXEL_GPU_NOALLOC_MAX_BLOCKS
 * and for layouts 1 and 4 (their last-workgroup combine needs a ticket).
 * Every call on one GPU shares that GPU's buffer: calls that overlap on one
 * GPU (on two streams, or from two host threads) race; calls in order on one
 * stream do not. */
#define VOXEL_GPU_NOALLOC_MAX_BLOCKS 4096
int       voxel_gpu_scan_noalloc(const voxel_gpu_scan_desc* d, voxel_gpu_agg* out_dev, void* stream);

/* ---- bitmap filters, ColumnFilter algebra, compaction ------------------ */
/* Predicate code 7 = Between: t <= x <= t2 (VectorFilter::ApplyBetween). */
enum { VOXEL_CMP_BETWEEN = 7 };

/* VectorFilter<T>::Apply* (ops/filter.hpp:42-288) over a device column:
 * words_dev[i/64] bit (i%64) = pred(x_i), (n+63)/64 words, tail bits zero;
 * *pass_count = number of set bits (FilterResult::PassCount). */
int       voxel_gpu_filter_bitmap(int dtype, const void* col, size_t n, int op, uint64_t t, uint64_t t2,
                                  uint64_t* words_dev, uint64_t* pass_count, void* stream);
/* The compound predicate op_a(a, t_a) AND op_b(b, t_b) fused into one bitmap
 * (= ApplyXX(a) then ColumnFilter::AndWith(ApplyYY(b))). */
int       voxel_gpu_filter_bitmap2(int dtype, const void* a, const void* b, size_t n, int op_a, uint64_t t_a,
                                   int op_b, uint64_t t_b, uint64_t* words_dev, uint64_t* pass_count,
                                   void* stream);
/* ColumnFilter::AndWith / OrWith / AndNotWith (combine 0 / 1 / 2) in place on
 * active_dev, tail bits past `rows` cleared; *active_count = set bits. */
int       voxel_gpu_bitmap_combine(uint64_t* active_dev, const uint64_t* other_dev, size_t rows, int combine,
                                   uint64_t* active_count, void* stream);
/* SelectionVector::BuildFromBitmap (ops/filter.hpp:378-399): ascending u32
 * row ids of the set bits. */
int       voxel_gpu_select_from_bitmap(const uint64_t* words_dev, size_t rows, uint32_t* idx_out_dev,
                                       uint64_t* count, void* stream);
/* scalar_filter_* (simd/scalar.hpp:155-333) generalised to every predicate:
 * passing values in row order into values_out_dev and/or their row ids into
 * idx_out_dev (either may be NULL). */
int       voxel_gpu_compact(int dtype, const void* col, size_t n, int op, uint64_t t, void* values_out_dev,
                            uint32_t* idx_out_dev, uint64_t* count, void* stream);

/* ---- hash join (HashJoin<i64>, ops/join.hpp:36-80) ------------------------ */
/* Builds a table over device keys (payload_dev may be NULL: payload = row). */
void*     voxel_gpu_join_build(const int64_t* build_keys_dev, const int64_t* payload_dev, size_t nb, void* stream);
void      voxel_gpu_join_destroy(void* join);
/* Inner join: number of probe rows with a match, and the wrapping i64 sum of
 * the matched build rows' payloads (first match = highest build row id among
 * equal keys, as the reference's LIFO chains give). */
int       voxel_gpu_join_probe_sum(void* join, const int64_t* probe_dev, size_t np, uint64_t* match_count,
                                   int64_t* payload_sum, void* stream);
int       voxel_gpu_join_probe_sum_launch(void* join, const int64_t* probe_dev, size_t np, void* stream);
int       voxel_gpu_join_probe_sum_result(void* join, uint64_t* match_count, int64_t* payload_sum, void* stream);
/* HashJoin::Probe's output: (probe row, build row) pairs in probe order. */
int       voxel_gpu_join_probe_pairs(void* join, const int64_t* probe_dev, size_t np, uint32_t* left_dev,
                                     uint32_t* right_dev, uint64_t* match_count, void* stream);

/* ---- group-by (HashAggregator<TKey,f64>, ops/aggregate.hpp:71-301) ----- */
/* GROUP BY key: COUNT, SUM (compensated), AVG = SUM/COUNT, MIN, MAX of an f64
 * column.  key_dtype VOXEL_GPU_I32, VOXEL_GPU_U32 or VOXEL_GPU_I64; output
 * device arrays (any may be NULL) of max_groups entries, ascending key order.
 * *ngroups is set even when it exceeds max_groups (then -1 is returned). */
int       voxel_gpu_groupby_f64(int key_dtype, const void* keys_dev, const double* vals_dev, size_t n,
                                size_t max_groups, int64_t* out_keys, uint64_t* out_count, double* out_sum,
                                double* out_avg, double* out_min, double* out_max, uint64_t* ngroups,
                                void* stream);
/* Same result through the sort-based path whatever the key range. */
int       voxel_gpu_groupby_f64_general(int key_dtype, const void* keys_dev, const double* vals_dev, size_t n,
                                        size_t max_groups, int64_t* out_keys, uint64_t* out_count,
                                        double* out_sum, double* out_avg, double* out_min, double* out_max,
                                        uint64_t* ngroups, void* stream);
/* Per-group results of a voxel_agg_* aggregator, into host arrays. */
int       voxel_gpu_agg_results(void* agg, int64_t* keys, uint64_t* counts, double* sums, double* avgs,
                                double* mins, double* maxs, size_t max_groups);

/* ---- validity bitmaps (NullBitmap, data/nulls.hpp:16-235; NullableSegment<T>,
 * nulls.hpp:241-420) over device columns ------------------------------------ */
/* Validity words in NullBitmap's layout: bit (r & 63) of word r >> 6, 1 = valid,
 * (n + 63) / 64 words (NullBitmap::Bits copied to HBM as is).  NULL = all rows
 * valid. */
/* voxel_gpu_scan over nullable columns: a row takes part only when its bit in
 * valid_a (and, for a predicate on column b, in valid_b) is 1.  The filter
 * bitmap ANDed with the validity words (ColumnFilter over the same layout,
 * ops/filter.hpp:540-547), then COUNT/SUM/MIN/MAX, in one pass. */
int       voxel_gpu_scan_valid(const voxel_gpu_scan_desc* d, const uint64_t* valid_a, const uint64_t* valid_b,
                               void* stream, voxel_gpu_agg* out);
typedef struct voxel_gpu_nullable_stats {
    uint64_t valid_count, null_count;        /* ValidCount / NullCount (rows < n) */
    double   sum, min, max;                  /* f64/f32: SumValid (compensated, rounded to the column type),
                                                MinValid, MaxValid (NaN = NullSentinel when no row is valid) */
    int64_t  isum, imin, imax;               /* i64/i32: SumValid (wrapping in the column type), MinValid,
                                                MaxValid (lowest() when no row is valid) */
} voxel_gpu_nullable_stats;
/* NullableSegment<T>::SumValid / MinValid / MaxValid / ValidCount / NullCountVal
 * of a device column (dtype VOXEL_GPU_F64/F32/I64/I32). */
int       voxel_gpu_nullable_reduce(int dtype, const void* data_dev, const uint64_t* valid_dev, size_t n,
                                    voxel_gpu_nullable_stats* out, void* stream);
/* NullableSegment<T>::CompactValid: the valid rows' values in row order into
 * out_dev (n entries of room), *count = how many (4- or 8-byte dtypes). */
int       voxel_gpu_nullable_compact(int dtype, const void* data_dev, const uint64_t* valid_dev, size_t n,
                                     void* out_dev, uint64_t* count, void* stream);

/* ---- columnar encodings (data/encoding.hpp:104-527) ---------------------- */
/* An encoded column on the device, in the reference's member layout:
 *   Dictionary  values = Dictionary (sorted distinct values), index = Indices (u32 per row), nvalues = dictionary size
 *   RLE         values = Values, index = RunLengths (u32, none 0), nvalues = runs; rows = sum of the run lengths
 *   Delta       values = Deltas (i64 / i32), delta_base = BaseValue; rows = Deltas.size()
 *   BitPacked   values = PackedData (u64 words), nvalues = words, bit_width = BitWidth; values are u32
 *   FrameOfReference  values = EncodedValues (i32), base = BaseValue, scale = Scale; values are f64
 * kind takes EncodingKind's values. */
enum { VOXEL_ENC_DICTIONARY = 1, VOXEL_ENC_RLE = 2, VOXEL_ENC_DELTA = 3, VOXEL_ENC_BITPACKED = 4, VOXEL_ENC_FOR = 6 };
typedef struct voxel_gpu_encoded_col {
    int32_t  kind;
    int32_t  dtype;            /* value type: VOXEL_GPU_F64/F32/I64/I32 (Dictionary, RLE), I64/I32 (Delta);
                                  ignored for BitPacked (u32) and FrameOfReference (f64) */
    uint64_t rows;
    const void* values;
    const uint32_t* index;
    uint64_t nvalues;
    uint32_t bit_width;
    uint32_t pad;
    double   base, scale;
    int64_t  delta_base;
} voxel_gpu_encoded_col;
/* IEncoding::DecodeBatch(out, offset, count) into device memory: returns the
 * rows written (RLE writes count rows, 0 past the last run, as the reference
 * does; the others stop at the last row), -1 on error.  Waits for the stream. */
int64_t   voxel_gpu_decode(const voxel_gpu_encoded_col* col, uint64_t offset, uint64_t count, void* out_dev,
                           void* stream);
/* Filter op(value, t) + COUNT/SUM/MIN/MAX over the decoded rows, reading the
 * encoded bytes (Delta: decoded first).  t holds the value type's bits
 * (BitPacked: a u32; results as VOXEL_GPU_I64). */
int       voxel_gpu_scan_encoded(const voxel_gpu_encoded_col* col, int op, uint64_t t, void* stream,
                                 voxel_gpu_agg* out);

/* ---- window family (Engine<f64>::WindowDelta .. WindowQuantile and
 * TimeWindowSum, exec/engine.hpp:183-300; voxel_py EngineF64.window_*,
 * bindings/voxel_bindings.cpp:127-180) ------------------------------------- */
/* Device columns in and out.  Each returns what the reference method returns
 * (the number of outputs written, 0 in its argument-check cases) and writes
 * out[0, N) only; -1 on error (window outside 0..255; window_quantile with a
 * NaN inside a window).  Sums, means, variances and standard deviations are
 * bit-identical to the reference's left-to-right window loops.  The first
 * five are stream-ordered (no host wait); window_quantile and
 * time_window_sum wait for the stream. */
int64_t   voxel_gpu_window_delta_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                     void* stream);
int64_t   voxel_gpu_window_sum_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                   int window, void* stream);
int64_t   voxel_gpu_window_mean_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                    int window, void* stream);
int64_t   voxel_gpu_window_stddev_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                      int window, void* stream);
int64_t   voxel_gpu_window_variance_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                        int window, void* stream);
int64_t   voxel_gpu_window_quantile_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                        int window, double quantile, void* stream);
/* Adds each row's value onto out[(ts - start_time) / duration_sec] for buckets
 * in [0, out_len) (out is read, not cleared); returns the largest bucket
 * written + 1 (1 when none).  Bucket sums are compensated (within a few ulp
 * of the exact sum; the reference adds in row order). */
int64_t   voxel_gpu_time_window_sum_f64(const int64_t* ts_dev, const double* vals_dev, size_t count,
                                        int64_t start_time, int64_t duration_sec, double* out_dev, size_t out_len,
                                        void* stream);

/* ---- learned-CDF sort (no counterpart in the reference; same output as
 * voxel_gpu_sort_pairs_u64, ops/sort.hpp:83-160) ---------------------------- */
/* FP32 MLP 1 -> 32 -> 32 -> 1 (ReLU), parameters packed as
 * W1[32] b1[32] W2[32][32] b2[32] W3[32] b3 (voxel_gpu_mlp_param_count()
 * floats, device memory). */
int       voxel_gpu_mlp_param_count(void);
int       voxel_gpu_mlp_init(float* params_dev, uint64_t seed);
int       voxel_gpu_mlp_forward(const float* params_dev, const float* x_dev, size_t n, float* y_dev, void* stream);
/* Gradient of the mean squared error over n points (grad_dev: param_count
 * floats) and the loss itself (loss_dev: one float); either may be NULL. */
int       voxel_gpu_mlp_grad(const float* params_dev, const float* x_dev, const float* y_dev, size_t n,
                             float* grad_dev, float* loss_dev, void* stream);
/* steps of Adam (beta 0.9/0.999, eps 1e-8) at learning rate lr; the loss
 * before the last step into *loss_out_host (may be NULL). */
int       voxel_gpu_mlp_train(float* params_dev, const float* x_dev, const float* y_dev, size_t n, int steps,
                              float lr, float* loss_out_host, void* stream);

typedef struct voxel_gpu_learned_stats {
    int32_t  model;             /* 0 MLP CDF, 1 equal-width (linear in the key) */
    int32_t  log_grid;          /* model input: 1 = float bits of key - min */
    int32_t  fell_back;         /* 1 = whole input sorted by the plain radix sort (small input, or the
                                   model did not split the keys); 2 = the plain radix sort chosen up
                                   front: the sampled keys differ in at most 5 bytes, so it skips the
                                   other passes and costs less than the learned partition */
    int32_t  partition_passes;  /* digit passes over the bucket id */
    int32_t  train_steps;
    float    final_loss;        /* mean squared error of the logit CDF fit */
    uint64_t buckets, cells, groups;
    uint64_t longest_bucket;    /* rows of the largest bucket sorted by a wave */
    uint64_t widest_group;      /* rows of the largest group of buckets */
    uint64_t big_groups, big_rows, big_unsorted;  /* groups over a wave's capacity */
    float    sample_ms, train_ms, table_ms, partition_ms, bucket_ms, fixup_ms;
} voxel_gpu_learned_stats;
/* Stable sort of (u64 key, u32 payload) pairs, bit-identical to
 * voxel_gpu_sort_pairs_u64: keys are partitioned by a monotone bucket id
 * derived from a learned CDF (model 0: the MLP, trained on the GPU on the
 * logit of a sample's CDF, routing to 4096 linear leaf models) or a linear
 * one (model 1, no model), each bucket is sorted in LDS.  vals_in NULL = row
 * index.  n < 2^30.  stats may be NULL. */
int       voxel_gpu_sort_pairs_u64_learned(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                                           uint32_t* vals_out, size_t n, int model,
                                           voxel_gpu_learned_stats* stats, void* stream);
/* Tuning of the learned sort: log2 sample size (20), training points
 * (4096), Adam steps (60; -1 = none), log2 of the largest average bucket
 * (8, range 5..8); 0 keeps the default. */
int       voxel_gpu_learned_tune(int sample_log2, int train_points, int train_steps, int rows_per_bucket_log2);

/* ---- partitioned hash join (HashJoin<i64>, ops/join.hpp:36-80, with the
 * table cut into sub-tables; same results as voxel_gpu_join_build) ------- */
typedef struct voxel_gpu_join_stats {
    int32_t  mode;                   /* layout in effect: 1 one table, 2 hash, 3 learned range, 4 equal-width range */
    int32_t  fell_back;              /* mode whose sub-tables overflowed (the table was rebuilt as one), 0 none */
    uint32_t partitions;             /* sub-tables */
    uint32_t pad;
    uint64_t slots_per_partition;
    uint64_t max_partition_keys;     /* distinct keys in the fullest / emptiest sub-table */
    uint64_t min_partition_keys;
    float    build_ms;
    float    pad2;
    voxel_gpu_learned_stats model;   /* range modes: the CDF model that partitions the keys */
} voxel_gpu_join_stats;
/* mode 0: auto (one table: measured fastest on MI355X); 1: one table;
 * partitioned layouts, sub-tables of 2^17 slots (at most 256): 2: sub-table
 * = top bits of the key's hash; 3: learned range partition (the learned
 * sort's MLP CDF fitted to the build keys); 4: equal-width range partition
 * (no model).  A range partition whose sub-table overflows falls back to one
 * table (stats->fell_back).  With a partitioned layout, probe batches of 2^20
 * rows or more are partitioned by sub-table before probing (one stable digit
 * pass) unless voxel_gpu_join_set_probe_partitioning(j, 0). */
void*     voxel_gpu_join_build_ex(const int64_t* build_keys_dev, const int64_t* payload_dev, size_t nb, int mode,
                                  voxel_gpu_join_stats* stats, void* stream);
int       voxel_gpu_join_set_probe_partitioning(void* join, int on);

/* ---- batched queries (no counterpart in the reference: it runs one program
 * per Engine::Run, exec/engine.hpp:109-141) ------------------------------- */
/* A batch engine runs up to max_queries scan descriptors (voxel_gpu_scan's,
 * blocks/unroll/layout fields ignored) in one persistent kernel launched
 * through a HIP graph on the engine's own stream.  Each query is reduced in
 * units of 8192 rows (one compensated partial per unit, merged in unit
 * order), so its result depends only on its rows: bit-identical alone, in
 * any mix and under any configuration.  A launch cuts the queries into
 * tiles of whole units, at most tile_rows rows, fewer while the batch would
 * give under two tiles per resident workgroup.  One launch in flight per
 * engine; an engine is not thread-safe. */
void*     voxel_gpu_batch_create(size_t max_queries);
void      voxel_gpu_batch_destroy(void* batch);
/* unroll 2/4/8, tile_rows multiple of 8192, blocks_per_cu 1..8; 0 keeps. */
int       voxel_gpu_batch_config(void* batch, int unroll, uint32_t tile_rows, int blocks_per_cu);
int       voxel_gpu_batch_get_config(void* batch, int* unroll, uint32_t* tile_rows, int* blocks_per_cu);
int       voxel_gpu_batch_launch(void* batch, const voxel_gpu_scan_desc* q, size_t nq);
int       voxel_gpu_batch_done(void* batch);            /* 1 when the launched batch finished */
/* Waits and writes nq results; *ms (may be NULL) = device time of the batch. */
int       voxel_gpu_batch_wait(void* batch, voxel_gpu_agg* out, float* ms);
int       voxel_gpu_batch_run(void* batch, const voxel_gpu_scan_desc* q, size_t nq, voxel_gpu_agg* out,
                              float* ms);
/* Times every configuration of unroll {2,4,8} x tile_rows {16K..256K} x
 * blocks_per_cu {1,2,4} on these queries (median of reps), checks that each
 * gives the same per-query results bit for bit, and keeps the fastest. */
int       voxel_gpu_batch_autotune(void* batch, const voxel_gpu_scan_desc* q, size_t nq, int reps,
                                   double* best_ms, int* configs_tried);

/* ---- serving: continuous batching over resident columns ----------------- */
typedef struct voxel_gpu_serve_stats {
    uint64_t queries, batches, bytes, max_batch_seen, admission_waits;
    double   makespan_s, qps, gbps, mean_batch;
    double   p50_s, p99_s, mean_s, max_s;   /* arrival -> result on the host */
} voxel_gpu_serve_stats;
/* inflight batch engines (1..8); a query is admitted while the scratch of the
 * queries on the device (64 B per 8192-row unit + result + descriptor)
 * stays within memory_budget_bytes (a query larger than the budget runs
 * alone). */
void*     voxel_gpu_server_create(size_t max_batch, uint64_t memory_budget_bytes, int inflight);
/* Stops the request loop if it runs, then frees the server. */
void      voxel_gpu_server_destroy(void* server);
int       voxel_gpu_server_config(void* server, int unroll, uint32_t tile_rows, int blocks_per_cu);
int       voxel_gpu_server_autotune(void* server, const voxel_gpu_scan_desc* q, size_t nq, int reps,
                                    double* best_ms, int* configs_tried);

/* Live serving.  start() runs the request loop on a thread of its own: it
 * takes submitted queries in arrival order, admits them into the next free
 * batch engine while the memory budget allows, launches, and completes
 * earlier batches as they finish, so queries that arrive while others are
 * on the GPU join the next batch.  submit() may be called from any thread
 * and returns a ticket (>= 0; -1 with voxel_last_error set for a bad
 * descriptor or a server that is not running); the columns must stay
 * resident until the ticket completes.  poll() returns 1 and the result when
 * the ticket has completed, 0 while it is pending, -1 for an unknown ticket
 * or a failed query; wait() blocks until it completes (latency_s may be
 * NULL: seconds from submit to completion).  stop() lets every submitted
 * query finish, ends the loop and reports the queries served since start()
 * (stats may be NULL); results stay readable until the next start(). */
int       voxel_gpu_server_start(void* server);
int64_t   voxel_gpu_server_submit(void* server, const voxel_gpu_scan_desc* q);
int       voxel_gpu_server_poll(void* server, int64_t ticket, voxel_gpu_agg* out);
int       voxel_gpu_server_wait(void* server, int64_t ticket, voxel_gpu_agg* out, double* latency_s);
int       voxel_gpu_server_stop(void* server, voxel_gpu_serve_stats* stats);

/* Replays n queries arriving at arrival_s[i] seconds (non-decreasing) after
 * the call starts; latency is measured from the scheduled arrival.
 * mode 0: continuous batching (start, one submit per arrival from the
 * calling thread, the live request loop serving them, stop); mode 1: one
 * synchronous voxel_gpu_scan per query; mode 2: one launch per query,
 * queued back to back on a stream of the server's own without waiting,
 * each result read back through its own event as it completes.  out[i],
 * latency_s[i] (may be NULL), stats.  Not while the live loop runs. */
int       voxel_gpu_server_replay(void* server, const voxel_gpu_scan_desc* q, const double* arrival_s, size_t n,
                                  int mode, voxel_gpu_agg* out, double* latency_s, voxel_gpu_serve_stats* stats);

#ifdef __cplusplus
}
#endif
