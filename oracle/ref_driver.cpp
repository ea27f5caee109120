// ref_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// A thin extern "C" driver, compiled together with the reference's own
// header-only C++ (include/voxel/...) straight from the reference checkout by
// oracle/Makefile into oracle/_ref/libvoxel_ref.so.  It only instantiates the
// reference's classes and calls them; it contains no algorithm of its own.
// Used to pin the oracle restatement (tests/) and as the timed CPU baseline
// (bench.py cpu_baseline, kind "reference").  The product never loads it.
#include "voxel/voxel.hpp"
#include <algorithm>
#include <bit>
#include <cstring>
#include <random>
#include <vector>

using namespace voxel;

namespace {
template <typename T>
int run_engine_segs(T* const* data, const size_t* n, int nseg, const uint32_t* code, size_t len, uint64_t* regs) {
    Engine<T> e;
    for (int s = 0; s < nseg; ++s) e.AddSegment(data[s], n[s]);
    for (size_t r = 0; r < 64; ++r) e.ScalarReg(r) = regs[r];
    std::vector<u32> v(code, code + len);
    e.LoadProgram(v);
    e.Run();
    for (size_t r = 0; r < 64; ++r) regs[r] = e.ScalarReg(r);
    return 0;
}
template <typename T>
int run_engine(T* data, size_t n, const uint32_t* code, size_t len, uint64_t* regs) {
    Engine<T> e;
    e.AddSegment(data, n);
    for (size_t r = 0; r < 64; ++r) e.ScalarReg(r) = regs[r];
    std::vector<u32> v(code, code + len);
    e.LoadProgram(v);
    e.Run();
    for (size_t r = 0; r < 64; ++r) regs[r] = e.ScalarReg(r);
    return 0;
}
// NullableSegment<T> (data/nulls.hpp:241-420) over caller data and
// NullBitmap words: res = {SumValid, MinValid, MaxValid}, counts =
// {ValidCount, NullCountVal, rows after CompactValid}, compact = those rows.
template <typename T>
void nullable_run(T* data, size_t n, const uint64_t* words, T* res, uint64_t* counts, T* compact) {
    NullableSegment<T> s(data, n);
    s.Nulls.CopyFrom(words, (n + 63) / 64);
    res[0] = s.SumValid();
    res[1] = s.MinValid();
    res[2] = s.MaxValid();
    counts[0] = s.ValidCount();
    counts[1] = s.NullCountVal();
    Arena arena;
    s.CompactValid(arena);
    counts[2] = s.Data.Count;
    if (s.Data.Count) std::memcpy(compact, s.Data.Data, s.Data.Count * sizeof(T));
}
// Encodings (data/encoding.hpp:104-527): Build from raw rows, and
// DecodeBatch over member arrays set by the caller.
template <typename T>
size_t dict_build(const T* x, size_t n, T* dict, uint32_t* idx) {
    encoding::DictionaryEncoding<T> e;
    e.Build(x, n);
    std::copy(e.Dictionary.begin(), e.Dictionary.end(), dict);
    std::copy(e.Indices.begin(), e.Indices.end(), idx);
    return e.Dictionary.size();
}
template <typename T>
void dict_decode(const T* dict, size_t nd, const uint32_t* idx, size_t rows, size_t off, size_t cnt, T* out) {
    encoding::DictionaryEncoding<T> e;
    e.Dictionary.assign(dict, dict + nd);
    e.Indices.assign(idx, idx + rows);
    e.DecodeBatch(out, off, cnt);
}
template <typename T>
size_t rle_build(const T* x, size_t n, T* vals, uint32_t* lens) {
    encoding::RLEEncoding<T> e;
    e.Build(x, n);
    std::copy(e.Values.begin(), e.Values.end(), vals);
    std::copy(e.RunLengths.begin(), e.RunLengths.end(), lens);
    return e.Values.size();
}
template <typename T>
void rle_decode(const T* vals, const uint32_t* lens, size_t nr, size_t off, size_t cnt, T* out) {
    encoding::RLEEncoding<T> e;
    e.Values.assign(vals, vals + nr);
    e.RunLengths.assign(lens, lens + nr);
    e.DecodeBatch(out, off, cnt);
}
template <typename T>
void delta_build(const T* x, size_t n, T* deltas, T* base) {
    encoding::DeltaEncoding<T> e;
    e.Build(x, n);
    std::copy(e.Deltas.begin(), e.Deltas.end(), deltas);
    *base = e.BaseValue;
}
template <typename T>
void delta_decode(T base, const T* deltas, size_t rows, size_t off, size_t cnt, T* out) {
    encoding::DeltaEncoding<T> e;
    e.BaseValue = base;
    e.Deltas.assign(deltas, deltas + rows);
    e.DecodeBatch(out, off, cnt);
}
}  // namespace

extern "C" {

// Input generators of the reference's examples (validate.cpp:17-29).
void vr_mt_uniform_f64(uint64_t seed, size_t n, double lo, double hi, double* out) {
    std::mt19937_64 g(seed);
    std::uniform_real_distribution<double> d(lo, hi);
    for (size_t i = 0; i < n; ++i) out[i] = d(g);
}
void vr_mt_uniform_i64(uint64_t seed, size_t n, int64_t lo, int64_t hi, int64_t* out) {
    std::mt19937_64 g(seed);
    std::uniform_int_distribution<int64_t> d(lo, hi);
    for (size_t i = 0; i < n; ++i) out[i] = d(g);
}
void vr_mt_uniform_i32(uint64_t seed, size_t n, int32_t lo, int32_t hi, int32_t* out) {
    std::mt19937_64 g(seed);
    std::uniform_int_distribution<int32_t> d(lo, hi);
    for (size_t i = 0; i < n; ++i) out[i] = d(g);
}

// Engine<T>::LoadProgram + Run (fast path or Tier-0), registers in/out.
int vr_engine_run_f64(double* d, size_t n, const uint32_t* c, size_t l, uint64_t* r) { return run_engine<f64>(d, n, c, l, r); }
int vr_engine_run_i64(int64_t* d, size_t n, const uint32_t* c, size_t l, uint64_t* r) { return run_engine<i64>(d, n, c, l, r); }
int vr_engine_run_i32(int32_t* d, size_t n, const uint32_t* c, size_t l, uint64_t* r) { return run_engine<i32>(d, n, c, l, r); }
// The same over several segments (segment s = data[s], n[s] rows); the
// engine writes the segments in place as its sort words do.
int vr_engine_run_segs_f64(double* const* d, const size_t* n, int ns, const uint32_t* c, size_t l, uint64_t* r) {
    return run_engine_segs<f64>(d, n, ns, c, l, r);
}
int vr_engine_run_segs_i64(int64_t* const* d, const size_t* n, int ns, const uint32_t* c, size_t l, uint64_t* r) {
    return run_engine_segs<i64>(d, n, ns, c, l, r);
}

// Engine<f64>::RunParallel, registers in/out.
void vr_engine_run_parallel_f64(double* d, size_t n, const uint32_t* c, size_t l, uint64_t* regs,
                                uint32_t threads, uint8_t result_reg) {
    Engine<f64> e;
    e.AddSegment(d, n);
    for (size_t r = 0; r < 64; ++r) e.ScalarReg(r) = regs[r];
    std::vector<u32> v(c, c + l);
    e.LoadProgram(v);
    e.RunParallel(0, threads, result_reg);
    for (size_t r = 0; r < 64; ++r) regs[r] = e.ScalarReg(r);
}

// The voxel_jit_run sequence (bindings/voxel_c.cpp:80-101), compile and
// entry split so that a caller can time the entry alone.
void* vr_jit_compile(const uint32_t* code, size_t len, size_t* code_size) {
    auto compiler = codegen::CreateJitCompiler();
    if (!compiler) return nullptr;
    auto* f = new codegen::JitFunction();
    if (!compiler->Compile(code, len, *f) || !f->IsValid()) { compiler->Release(*f); delete f; return nullptr; }
    if (code_size) *code_size = f->CodeSize;
    return f;
}
double vr_jit_entry(void* fn, const double* data, size_t n, uint64_t r2, double threshold) {
    auto* f = static_cast<codegen::JitFunction*>(fn);
    RegFile rf;
    rf.Scalar(0) = 0;
    rf.Scalar(1) = 0;
    rf.Scalar(2) = r2;
    rf.Scalar(3) = std::bit_cast<u64>(threshold);
    f64* seg = const_cast<f64*>(data);
    u64 cnt = n;
    f->Entry(&rf, &seg, &cnt);
    return std::bit_cast<f64>(rf.Scalar(0));
}
void vr_jit_release(void* fn) {
    auto compiler = codegen::CreateJitCompiler();
    auto* f = static_cast<codegen::JitFunction*>(fn);
    if (compiler) compiler->Release(*f);
    delete f;
}

// VectorFilter<T>::Apply{EQ,NE,LT,LE,GT,GE,Between}; op = kCmp code, 6 = Between.
#define VR_BITMAP(NAME, T)                                                              \
uint64_t NAME(const T* x, size_t n, int op, T t, T t2, uint64_t* words) {              \
    ops::VectorFilter<T> f;                                                             \
    typename ops::VectorFilter<T>::FilterResult r;                                      \
    switch (op) {                                                                       \
    case 0: r = f.ApplyEQ(x, n, t); break; case 1: r = f.ApplyNE(x, n, t); break;       \
    case 2: r = f.ApplyLT(x, n, t); break; case 3: r = f.ApplyLE(x, n, t); break;       \
    case 4: r = f.ApplyGT(x, n, t); break; case 5: r = f.ApplyGE(x, n, t); break;       \
    default: r = f.ApplyBetween(x, n, t, t2); break; }                                  \
    if (!r.Bitmap.empty()) std::memcpy(words, r.Bitmap.data(), r.Bitmap.size() * 8);    \
    return r.PassCount;                                                                 \
}
VR_BITMAP(vr_bitmap_f64, double)
VR_BITMAP(vr_bitmap_f32, float)
VR_BITMAP(vr_bitmap_i64, int64_t)
VR_BITMAP(vr_bitmap_i32, int32_t)

// ColumnFilter::Init + AndWith/OrWith/AndNotWith; active is replaced.
uint64_t vr_column_filter(uint64_t* active, const uint64_t* other, size_t rows, int combine) {
    ops::ColumnFilter<f64> cf;
    cf.Init(rows);
    size_t nw = (rows + 63) / 64;
    cf.AndWith(active, nw);
    if (combine == 0) cf.AndWith(other, nw);
    else if (combine == 1) cf.OrWith(other, nw);
    else cf.AndNotWith(other, nw);
    std::memcpy(active, cf.GetBitmap().data(), nw * 8);
    return cf.GetActiveCount();
}

// SelectionVector::BuildFromBitmap.
uint64_t vr_select_from_bitmap(const uint64_t* words, size_t rows, uint32_t* out) {
    ops::SelectionVector<f64> sv;
    sv.BuildFromBitmap(words, rows);
    if (sv.Count) std::memcpy(out, sv.Indices.data(), sv.Count * 4);
    return sv.Count;
}

// simd::scalar::scalar_filter_gt_* / scalar_sum_*.
uint64_t vr_scalar_filter_gt_f64(double* dst, const double* src, double t, size_t n) { return simd::scalar::scalar_filter_gt_f64(dst, src, t, n); }
uint64_t vr_scalar_filter_gt_f32(float* dst, const float* src, float t, size_t n) { return simd::scalar::scalar_filter_gt_f32(dst, src, t, n); }
double vr_scalar_sum_f64(const double* src, size_t n) { return simd::scalar::scalar_sum_f64(src, n); }
float vr_scalar_sum_f32(const float* src, size_t n) { return simd::scalar::scalar_sum_f32(src, n); }

// RadixSort<u64>::Sort(data, indices, count).
void vr_radix_sort_u64(uint64_t* keys, uint32_t* idx, size_t n) {
    ops::RadixSort<u64> s;
    s.Sort(keys, idx, n);
}

// SortOperator::SortAscending/Descending (what voxel_sort_* call).
void vr_sort_f64(const double* data, size_t n, int descending, uint32_t* idx) {
    Segment<f64> seg(const_cast<f64*>(data), n);
    if (descending) ops::SortOperator::SortDescending(seg, idx);
    else ops::SortOperator::SortAscending(seg, idx);
}

// HashJoin<i64>::Build + Probe, plus the harness's wrapping payload sum.
uint64_t vr_hash_join_i64(const int64_t* build, size_t nb, const int64_t* payload,
                          const int64_t* probe, size_t np, uint32_t* left, uint32_t* right,
                          int64_t* payload_sum) {
    Arena arena;
    ops::HashJoin<i64> j(arena);
    j.Build(build, nb);
    std::vector<u32> l(left ? 0 : np), r(right ? 0 : np);
    u32* lo = left ? left : l.data();
    u32* ro = right ? right : r.data();
    sz m = 0;
    j.Probe(probe, np, lo, ro, m);
    u64 s = 0;
    if (payload) for (sz i = 0; i < m; ++i) s += static_cast<u64>(payload[ro[i]]);
    if (payload_sum) *payload_sum = static_cast<i64>(s);
    return m;
}

// HashAggregator<i32|u32, f64>::Init + Accumulate + Finalize (slot order).
#define VR_AGG(NAME, K)                                                                  \
size_t NAME(const K* keys, const double* vals, size_t n, size_t expected, size_t maxg,    \
            K* out_key, uint64_t* out_count, double* out_sum, double* out_min,            \
            double* out_max, double* out_mean, double* out_m2) {                          \
    Arena arena;                                                                          \
    ops::HashAggregator<K, f64> a(arena);                                                 \
    a.Init(expected);                                                                     \
    a.Accumulate(keys, vals, n);                                                          \
    size_t g = a.GroupCount();                                                            \
    if (g > maxg) return g;                                                               \
    std::vector<typename ops::HashAggregator<K, f64>::AggregateState> st(g);              \
    a.Finalize(out_key, st.data(), g);                                                    \
    for (size_t i = 0; i < g; ++i) {                                                      \
        out_count[i] = st[i].Count; out_sum[i] = st[i].Sum; out_min[i] = st[i].Min;       \
        out_max[i] = st[i].Max; out_mean[i] = st[i].Mean; out_m2[i] = st[i].M2;           \
    }                                                                                     \
    return g;                                                                             \
}
VR_AGG(vr_hash_agg_i32_f64, int32_t)
VR_AGG(vr_hash_agg_u32_f64, uint32_t)

// Compound predicate a > ta AND b < tb the reference's way (the path its
// operators offer for config 2: VectorFilter<f64>::ApplyGT / ApplyLT,
// ColumnFilter::Init + AndWith twice, SelectionVector::BuildFromBitmap), then
// a row-order loop over the selected rows for COUNT / SUM / MIN / MAX of a.
// out = {count, sum, min, max}.
void vr_compound_f64(const double* a, const double* b, size_t n, double ta, double tb, double* out) {
    ops::VectorFilter<f64> fa, fb;
    auto ra = fa.ApplyGT(a, n, ta);
    auto rb = fb.ApplyLT(b, n, tb);
    ops::ColumnFilter<f64> cf;
    cf.Init(n);
    cf.AndWith(ra.Bitmap.data(), ra.Bitmap.size());
    cf.AndWith(rb.Bitmap.data(), rb.Bitmap.size());
    ops::SelectionVector<f64> sel;
    cf.BuildSelectionVector(sel);
    double s = 0.0, mn = INFINITY, mx = -INFINITY;
    for (size_t i = 0; i < sel.Count; ++i) {
        const double v = a[sel.Indices[i]];
        s += v;
        mn = v < mn ? v : mn;
        mx = v > mx ? v : mx;
    }
    out[0] = (double)sel.Count;
    out[1] = s;
    out[2] = mn;
    out[3] = mx;
}

// TopKSelect (voxel_topk).
void vr_topk_f64(const double* data, size_t n, size_t k, double* out, int largest) {
    ops::TopKSelect(data, n, k, out, largest != 0);
}

void vr_nullable_f64(double* d, size_t n, const uint64_t* w, double* res, uint64_t* counts, double* compact) {
    nullable_run<f64>(d, n, w, res, counts, compact);
}
void vr_nullable_i64(int64_t* d, size_t n, const uint64_t* w, int64_t* res, uint64_t* counts, int64_t* compact) {
    nullable_run<i64>(d, n, w, res, counts, compact);
}

size_t vr_dict_build_f64(const double* x, size_t n, double* d, uint32_t* i) { return dict_build<f64>(x, n, d, i); }
size_t vr_dict_build_i64(const int64_t* x, size_t n, int64_t* d, uint32_t* i) { return dict_build<i64>(x, n, d, i); }
void vr_dict_decode_f64(const double* d, size_t nd, const uint32_t* i, size_t rows, size_t off, size_t cnt, double* o) {
    dict_decode<f64>(d, nd, i, rows, off, cnt, o);
}
void vr_dict_decode_i64(const int64_t* d, size_t nd, const uint32_t* i, size_t rows, size_t off, size_t cnt, int64_t* o) {
    dict_decode<i64>(d, nd, i, rows, off, cnt, o);
}
size_t vr_rle_build_f64(const double* x, size_t n, double* v, uint32_t* l) { return rle_build<f64>(x, n, v, l); }
size_t vr_rle_build_i64(const int64_t* x, size_t n, int64_t* v, uint32_t* l) { return rle_build<i64>(x, n, v, l); }
void vr_rle_decode_f64(const double* v, const uint32_t* l, size_t nr, size_t off, size_t cnt, double* o) {
    rle_decode<f64>(v, l, nr, off, cnt, o);
}
void vr_rle_decode_i64(const int64_t* v, const uint32_t* l, size_t nr, size_t off, size_t cnt, int64_t* o) {
    rle_decode<i64>(v, l, nr, off, cnt, o);
}
void vr_delta_build_i64(const int64_t* x, size_t n, int64_t* d, int64_t* b) { delta_build<i64>(x, n, d, b); }
void vr_delta_build_i32(const int32_t* x, size_t n, int32_t* d, int32_t* b) { delta_build<i32>(x, n, d, b); }
void vr_delta_decode_i64(int64_t b, const int64_t* d, size_t rows, size_t off, size_t cnt, int64_t* o) {
    delta_decode<i64>(b, d, rows, off, cnt, o);
}
void vr_delta_decode_i32(int32_t b, const int32_t* d, size_t rows, size_t off, size_t cnt, int32_t* o) {
    delta_decode<i32>(b, d, rows, off, cnt, o);
}
size_t vr_bitpack_build(const uint32_t* x, size_t n, uint64_t* words, uint32_t* bw) {
    encoding::BitPackedEncoding e;
    e.Build(x, n);
    std::copy(e.PackedData.begin(), e.PackedData.end(), words);
    *bw = e.BitWidth;
    return e.PackedData.size();
}
void vr_bitpack_decode(const uint64_t* words, size_t nw, uint32_t bw, size_t rows, size_t off, size_t cnt, uint32_t* o) {
    encoding::BitPackedEncoding e;
    e.PackedData.assign(words, words + nw);
    e.BitWidth = bw;
    e.ValueCount = rows;
    e.DecodeBatch(o, off, cnt);
}
void vr_for_build(const double* x, size_t n, double* base, double* scale, int32_t* enc) {
    encoding::FrameOfReferenceEncoding e;
    e.Build(x, n);
    *base = e.BaseValue;
    *scale = e.Scale;
    std::copy(e.EncodedValues.begin(), e.EncodedValues.end(), enc);
}
void vr_for_decode(double base, double scale, const int32_t* enc, size_t rows, size_t off, size_t cnt, double* o) {
    encoding::FrameOfReferenceEncoding e;
    e.BaseValue = base;
    e.Scale = scale;
    e.EncodedValues.assign(enc, enc + rows);
    e.DecodeBatch(o, off, cnt);
}

// Window family, static methods of Engine<f64> (exec/engine.hpp:183-300).
int64_t vr_window_f64(int kind, const double* d, size_t n, double* out, size_t out_len, int window, double q) {
    const u8 w = static_cast<u8>(window);
    switch (kind) {
    case 0: return (int64_t)Engine<f64>::WindowDelta(d, n, out, out_len);
    case 1: return (int64_t)Engine<f64>::WindowSum(d, n, out, out_len, w);
    case 2: return (int64_t)Engine<f64>::WindowMean(d, n, out, out_len, w);
    case 3: return (int64_t)Engine<f64>::WindowStdDev(d, n, out, out_len, w);
    case 4: return (int64_t)Engine<f64>::WindowVariance(d, n, out, out_len, w);
    default: return (int64_t)Engine<f64>::WindowQuantile(d, n, out, out_len, w, q);
    }
}
int64_t vr_time_window_sum_f64(const int64_t* ts, const double* v, size_t n, int64_t start, int64_t dur, double* out,
                               size_t out_len) {
    return (int64_t)Engine<f64>::TimeWindowSum(ts, v, n, start, dur, out, out_len);
}

// Instruction encoders (bytecode/instruction.hpp) for program-word fixtures.
uint32_t vr_encode(uint8_t op, uint8_t rd, uint8_t ra, uint8_t rb, uint16_t imm) {
    return Instruction::Encode(static_cast<Opcode>(op), rd, ra, rb, imm);
}

}  // extern "C"
