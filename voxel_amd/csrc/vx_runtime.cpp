// vx_runtime.cpp -- host runtime of libvoxel_c (see vx_runtime.hpp).
#include "vx_runtime.hpp"

#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <type_traits>
#include <unordered_map>

namespace vx {

namespace {
thread_local char g_err[512] = "";

struct Runtime {
    bool ok = false;
    int device = 0;
    StreamCtx own;
    std::unordered_map<hipStream_t, StreamCtx*> others;
};

Runtime& rt() {
    static Runtime r;
    return r;
}

std::mutex g_init_mutex;

bool init_ctx(StreamCtx* c, hipStream_t s) {
    c->stream = s;
    c->scr.max_blocks = (size_t)device_cu_count() * 16;
    if (!hip_ok(hipMalloc(&c->scr.partials, c->scr.max_blocks * 64), "hipMalloc(partials)")) return false;
    if (!hip_ok(hipMalloc((void**)&c->scr.ticket, 256), "hipMalloc(ticket)")) return false;
    if (!hip_ok(hipMemset(c->scr.ticket, 0, 256), "hipMemset(ticket)")) return false;
    if (!hip_ok(hipMalloc((void**)&c->dev, sizeof(ResultSlab)), "hipMalloc(result)")) return false;
    if (!hip_ok(hipMemset(c->dev, 0, sizeof(ResultSlab)), "hipMemset(result)")) return false;
    if (!hip_ok(hipHostMalloc((void**)&c->host, sizeof(ResultSlab), hipHostMallocDefault), "hipHostMalloc(result)"))
        return false;
    std::memset(c->host, 0, sizeof(ResultSlab));
    return true;
}
}  // namespace

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

const char* last_error() { return g_err; }

bool hip_ok(hipError_t e, const char* what) {
    if (e == hipSuccess) return true;
    set_error("%s failed: %s", what, hipGetErrorString(e));
    return false;
}

std::mutex& api_mutex() {
    static std::mutex m;
    return m;
}

bool runtime_ok() {
    Runtime& r = rt();
    std::lock_guard<std::mutex> g(g_init_mutex);
    if (r.ok) return true;
    // (a failed bring-up is retried on the next call)
    hipError_t ie = hipInit(0);
    int n = 0;
    hipError_t ce = hipGetDeviceCount(&n);
    if (ie != hipSuccess || ce != hipSuccess || n < 1) {
        set_error("no usable HIP device (voxel_amd needs an MI355X/gfx950 GPU): hipInit %s, hipGetDeviceCount %s, %d devices",
                  hipGetErrorString(ie), hipGetErrorString(ce), n);
        return false;
    }
    if (!hip_ok(hipGetDevice(&r.device), "hipGetDevice")) return false;
    // NULL stream = the device's default stream, so that work a caller queued
    // there (e.g. torch on its default stream) is ordered before ours.
    if (!init_ctx(&r.own, nullptr)) return false;
    r.ok = true;
    return true;
}

StreamCtx* ctx_for(void* stream) {
    if (!runtime_ok()) return nullptr;
    Runtime& r = rt();
    if (stream == nullptr) return &r.own;
    hipStream_t s = static_cast<hipStream_t>(stream);
    std::lock_guard<std::mutex> g(g_init_mutex);
    auto it = r.others.find(s);
    if (it != r.others.end()) return it->second;
    auto* c = new StreamCtx();
    if (!init_ctx(c, s)) { delete c; return nullptr; }
    r.others.emplace(s, c);
    return c;
}

// ---------------------------------------------------------------------------
// Predicate lowering.  A VM comparison x OP t becomes either a closed range
// [lo, hi] (optionally inverted) or a truth table over the compare state.
// Range bounds use the neighbouring representable value for strict compares,
// so that x > t is x >= next_up(t) for every non-NaN x, and NaN fails every
// range test (inverted ones pass it, as C++ != does).
// ---------------------------------------------------------------------------
namespace {
template <typename T> uint64_t bits_of(T v) {
    if constexpr (sizeof(T) == 8) { uint64_t u; std::memcpy(&u, &v, 8); return u; }
    else { uint32_t u; std::memcpy(&u, &v, 4); return u; }
}
template <typename T> T of_bits(uint64_t b) {
    T v;
    if constexpr (sizeof(T) == 8) std::memcpy(&v, &b, 8);
    else { uint32_t u = (uint32_t)b; std::memcpy(&v, &u, 4); }
    return v;
}

template <typename T>
PredDev range_pred(int cmp, T t) {
    PredDev p{};
    T lo, hi;
    bool empty = false;
    const T tmin = std::numeric_limits<T>::is_integer ? std::numeric_limits<T>::min() : -std::numeric_limits<T>::infinity();
    const T tmax = std::numeric_limits<T>::is_integer ? std::numeric_limits<T>::max() : std::numeric_limits<T>::infinity();
    bool tnan = !std::numeric_limits<T>::is_integer && (t != t);
    p.inv = 0;
    switch (cmp) {
    case CMP_EQ: lo = t; hi = t; break;
    case CMP_NE: lo = t; hi = t; p.inv = 1; break;
    case CMP_LE: lo = tmin; hi = t; empty = tnan; break;
    case CMP_GE: lo = t; hi = tmax; empty = tnan; break;
    case CMP_LT:
        lo = tmin;
        if (tnan || t == tmin) { empty = true; hi = tmin; }
        else if constexpr (std::numeric_limits<T>::is_integer) hi = t - 1;
        else hi = std::nextafter(t, -std::numeric_limits<T>::infinity());
        break;
    case CMP_GT:
        hi = tmax;
        if (tnan || t == tmax) { empty = true; lo = tmax; }
        else if constexpr (std::numeric_limits<T>::is_integer) lo = t + 1;
        else lo = std::nextafter(t, std::numeric_limits<T>::infinity());
        break;
    case CMP_TRUE: empty = true; p.inv = 1; lo = hi = t; break;
    default: empty = true; lo = hi = t; break;   // never passes
    }
    if (empty) { lo = tmax; hi = tmin; }
    p.lo = bits_of<T>(lo);
    p.hi = bits_of<T>(hi);
    p.t = bits_of<T>(t);
    return p;
}

// Truth-table bit per compare state: bit0 unordered, bit1 LT, bit2 EQ, bit4 GT.
uint32_t table_for(int cmp, bool ordered) {
    switch (cmp) {
    case CMP_EQ: return 0x04;
    case CMP_NE: return ordered ? 0x12 : 0x13;
    case CMP_LT: return 0x02;
    case CMP_LE: return 0x06;
    case CMP_GT: return 0x10;
    case CMP_GE: return 0x14;
    case CMP_TRUE: return 0x17;
    default: return 0x00;
    }
}
}  // namespace

PredDev make_pred(int dtype, int cmp, uint64_t tbits, bool ordered) {
    PredDev p{};
    switch (dtype) {
    case DT_F64: p = range_pred<double>(cmp, of_bits<double>(tbits)); break;
    case DT_F32: p = range_pred<float>(cmp, of_bits<float>(tbits)); break;
    case DT_I64: p = range_pred<int64_t>(cmp, of_bits<int64_t>(tbits)); break;
    case DT_I32: p = range_pred<int32_t>(cmp, of_bits<int32_t>(tbits)); break;
    default: break;
    }
    p.table = table_for(cmp, ordered);
    return p;
}

PredDev make_nonzero_pred(int dtype) {
    PredDev p = make_pred(dtype, CMP_NE, 0, true);
    p.t = 0;
    p.table = 0x12;
    return p;
}

PredDev make_never_pred(int dtype) { return make_pred(dtype, -1, 0); }

namespace {
template <typename T>
bool merge_range(const PredDev& a, const PredDev& b, PredDev* out) {
    const T alo = of_bits<T>(a.lo), ahi = of_bits<T>(a.hi), blo = of_bits<T>(b.lo), bhi = of_bits<T>(b.hi);
    const bool a_empty = !(alo <= ahi), b_empty = !(blo <= bhi);
    const bool a_true = a.inv && a_empty, b_true = b.inv && b_empty;
    if (a_true) { *out = b; return true; }
    if (b_true) { *out = a; return true; }
    if (a.inv || b.inv) return false;                 // x != t: the conjunction is no single range
    PredDev p = a;
    const T lo = alo < blo ? blo : alo, hi = ahi < bhi ? ahi : bhi;
    p.lo = bits_of<T>(lo);
    p.hi = bits_of<T>(hi);
    p.inv = 0;
    if (a_empty || b_empty || !(lo <= hi)) p = make_pred(std::is_same<T, double>::value ? DT_F64 :
                                                         std::is_same<T, float>::value ? DT_F32 :
                                                         std::is_same<T, int64_t>::value ? DT_I64 : DT_I32, -1, 0);
    p.table = 0;
    *out = p;
    return true;
}
}  // namespace

bool merge_preds(int dtype, const PredDev& a, const PredDev& b, PredDev* out) {
    switch (dtype) {
    case DT_F64: return merge_range<double>(a, b, out);
    case DT_F32: return merge_range<float>(a, b, out);
    case DT_I64: return merge_range<int64_t>(a, b, out);
    case DT_I32: return merge_range<int32_t>(a, b, out);
    default: return false;
    }
}

double finish_sum_f64(double acc0, double s, double c) {
    if (!std::isfinite(s)) return acc0 + s;
    // (acc0 + s) exactly as a pair, then fold in c.
    double t = acc0 + s;
    double bp = t - acc0;
    double e = (acc0 - (t - bp)) + (s - bp);
    return t + (e + c);
}

float finish_sum_f32(float acc0, float s, float c) {
    if (!std::isfinite(s)) return acc0 + s;
    float t = acc0 + s;
    float bp = t - acc0;
    float e = (acc0 - (t - bp)) + (s - bp);
    return t + (e + c);
}

DevColumn::~DevColumn() {
    if (dev) (void)hipFree(dev);
}

bool DevColumn::ensure(size_t need) {
    if (need <= bytes && dev) return true;
    if (dev) { (void)hipFree(dev); dev = nullptr; bytes = 0; valid = false; }
    size_t alloc = need < 256 ? 256 : need;
    if (!hip_ok(hipMalloc(&dev, alloc), "hipMalloc(column)")) { dev = nullptr; return false; }
    bytes = alloc;
    valid = false;
    return true;
}

}  // namespace vx
