// vx_nulls.hip -- validity bitmaps on the GPU: NullBitmap (data/nulls.hpp:16-235)
// and NullableSegment<T> (nulls.hpp:241-420) over device columns.
//
// Validity words use the reference's layout (bit r & 63 of word r >> 6,
// 1 = valid), so a bitmap built by NullBitmap can be copied to HBM as is.
//   * scans: the predicate scan with the validity bits of each column ANDed
//     into the row test (k_scan_valid, vx_scan.hip), one pass;
//   * NullableSegment SumValid / MinValid / MaxValid / ValidCount: the same
//     scan with a true predicate, plus the first valid row (k_first_valid) that
//     seeds the reference's MinValid/MaxValid loops;
//   * CompactValid: order-preserving stream compaction (vx_compact.hpp) with
//     the validity bit as the predicate.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <cstring>
#include <limits>
#include <mutex>

#include "../../include/voxel_c.h"
#include "vx_compact.hpp"
#include "vx_kernels.hpp"
#include "vx_runtime.hpp"

namespace vx {
namespace {

// Smallest row < n whose validity bit is set (UINT64_MAX when none): each
// thread takes the first set bit of its grid-stride words, the wave keeps the
// smallest, and one atomic per wave publishes it.
__global__ __launch_bounds__(256) void k_first_valid(const uint64_t* __restrict__ w, uint64_t n,
                                                     unsigned long long* __restrict__ out) {
    const uint64_t nw = (n + 63) >> 6;
    const uint64_t stride = (uint64_t)gridDim.x * 256;
    unsigned long long best = ~0ull;
    for (uint64_t i = (uint64_t)blockIdx.x * 256 + threadIdx.x; i < nw; i += stride) {
        uint64_t x = w[i];
        if (i == nw - 1 && (n & 63)) x &= (1ull << (n & 63)) - 1;
        if (x) {                                          // later words of this thread hold larger rows
            best = i * 64 + (uint64_t)__builtin_ctzll(x);
            break;
        }
    }
    for (int m = 32; m > 0; m >>= 1) {
        const unsigned long long o = __shfl_xor(best, m, 64);
        best = o < best ? o : best;
    }
    if ((threadIdx.x & 63) == 0 && best != ~0ull) atomicMin(out, best);
}

__global__ void k_set_u64(unsigned long long* p, unsigned long long v) { *p = v; }

// CompactValid: rows whose validity bit is set, in row order.
template <typename U>
struct ValidSrc {
    using V = U;
    const U* data;
    const uint64_t* valid;
    U* out;
    __device__ __forceinline__ bool pass(uint64_t i) const { return (valid[i >> 6] >> (i & 63)) & 1ull; }
    __device__ __forceinline__ U fetch(uint64_t i) const { return data[i]; }
    __device__ __forceinline__ void put(uint64_t, uint64_t j, U v) const { out[j] = v; }
};

struct NullBufs {
    unsigned long long* first_dev = nullptr;
    unsigned long long* host = nullptr;      // [0] first valid row, [1] the value's bits
    bool ensure() {
        if (first_dev) return true;
        return hip_ok(hipMalloc((void**)&first_dev, 8), "hipMalloc") &&
               hip_ok(hipHostMalloc((void**)&host, 16, hipHostMallocDefault), "hipHostMalloc");
    }
};
NullBufs& null_bufs() {                             // kept to process end (no hipFree at exit)
    static NullBufs* b = new NullBufs();
    return *b;
}

bool wait(StreamCtx* c) { return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"); }

// Scan with validity, result read back into *r.
bool scan_valid_once(StreamCtx* c, const ScanArgs& g, const uint64_t* va, const uint64_t* vb, ScanOut* r) {
    if (g.end == g.begin) {
        std::memset(r, 0, sizeof(*r));
        return true;
    }
    if (!hip_ok(launch_scan_valid(g, va, vb, c->scr, &c->dev->scan, c->stream), "nullable scan launch")) return false;
    if (!hip_ok(hipMemcpyAsync(&c->host->scan, &c->dev->scan, sizeof(ScanOut), hipMemcpyDeviceToHost, c->stream),
                "read back result") ||
        !wait(c))
        return false;
    *r = c->host->scan;
    return true;
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

int voxel_gpu_scan_valid(const voxel_gpu_scan_desc* d, const uint64_t* valid_a, const uint64_t* valid_b, void* stream,
                         voxel_gpu_agg* out) {
    set_error("%s", "");
    ScanArgs g;
    if (!make_scan_args(d, g)) return -1;
    if (!out) { set_error("scan_valid: null result"); return -1; }
    if (valid_b && g.mode != 2) { set_error("scan_valid: valid_b given without a predicate on column b"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    ScanOut r;
    if (!scan_valid_once(c, g, valid_a, valid_b, &r)) return -1;
    fill_agg(d->dtype, r, out);
    return 0;
}

int voxel_gpu_nullable_reduce(int dtype, const void* data_dev, const uint64_t* valid_dev, size_t n,
                             voxel_gpu_nullable_stats* out, void* stream) {
    set_error("%s", "");
    if (dtype < DT_F64 || dtype > DT_I32) { set_error("nullable_reduce: dtype %d not supported", dtype); return -1; }
    if (!out || (n && !data_dev)) { set_error("nullable_reduce: null argument"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    ScanArgs g{};
    g.dtype = dtype; g.a = data_dev; g.begin = 0; g.end = n; g.mode = 0;
    g.pa = make_pred(dtype, CMP_TRUE, 0);
    ScanOut r;
    if (!scan_valid_once(c, g, valid_dev, nullptr, &r)) return -1;
    voxel_gpu_agg a;
    fill_agg(dtype, r, &a);
    std::memset(out, 0, sizeof(*out));
    out->valid_count = r.cnt;
    out->null_count = (uint64_t)n - r.cnt;
    const bool flt = dtype == DT_F64 || dtype == DT_F32;
    if (flt) {
        out->sum = a.sum;
        out->min = a.fmin;
        out->max = a.fmax;
    } else {
        out->isum = dtype == DT_I32 ? (int64_t)(int32_t)(uint32_t)(uint64_t)a.isum : a.isum;   // T total
        out->imin = a.imin;
        out->imax = a.imax;
    }
    if (r.cnt == 0) {                                   // MinValid/MaxValid return NullSentinel
        out->min = out->max = std::numeric_limits<double>::quiet_NaN();
        const int64_t lowest = dtype == DT_I32 ? (int64_t)INT32_MIN : INT64_MIN;
        out->imin = out->imax = lowest;
        return 0;
    }
    if (flt) {
        // MinValid/MaxValid start from the first valid value and only replace
        // it by a smaller / larger one: a NaN there is the result.
        NullBufs& b = null_bufs();
        if (!b.ensure()) return -1;
        uint64_t row = 0;
        if (valid_dev) {
            hipLaunchKernelGGL(k_set_u64, dim3(1), dim3(1), 0, c->stream, b.first_dev, ~0ull);
            const uint64_t nw = ((uint64_t)n + 63) / 64;
            const unsigned grid = (unsigned)std::min<uint64_t>((nw + 255) / 256, 1024);
            hipLaunchKernelGGL(k_first_valid, dim3(grid), dim3(256), 0, c->stream, valid_dev, (uint64_t)n, b.first_dev);
            if (!hip_ok(hipGetLastError(), "first valid row launch")) return -1;
            if (!hip_ok(hipMemcpyAsync(b.host, b.first_dev, 8, hipMemcpyDeviceToHost, c->stream), "read back") ||
                !wait(c))
                return -1;
            row = b.host[0];
        }
        const size_t es = dtype == DT_F64 ? 8 : 4;
        b.host[1] = 0;
        if (!hip_ok(hipMemcpyAsync(&b.host[1], static_cast<const char*>(data_dev) + row * es, es,
                                   hipMemcpyDeviceToHost, c->stream), "read back") ||
            !wait(c))
            return -1;
        double first;
        if (dtype == DT_F64) std::memcpy(&first, &b.host[1], 8);
        else { float f; std::memcpy(&f, &b.host[1], 4); first = f; }
        if (std::isnan(first)) out->min = out->max = first;
    }
    return 0;
}

int voxel_gpu_nullable_compact(int dtype, const void* data_dev, const uint64_t* valid_dev, size_t n, void* out_dev,
                               uint64_t* count, void* stream) {
    set_error("%s", "");
    if (dtype < DT_F64 || dtype > DT_U32) { set_error("nullable_compact: dtype %d not supported", dtype); return -1; }
    if (!count || (n && (!data_dev || !valid_dev || !out_dev))) { set_error("nullable_compact: null argument"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    bool ok;
    if (dtype_size(dtype) == 8)
        ok = run_compact(ValidSrc<uint64_t>{static_cast<const uint64_t*>(data_dev), valid_dev,
                                            static_cast<uint64_t*>(out_dev)}, n, c, count);
    else
        ok = run_compact(ValidSrc<uint32_t>{static_cast<const uint32_t*>(data_dev), valid_dev,
                                            static_cast<uint32_t*>(out_dev)}, n, c, count);
    return ok ? 0 : -1;
}

}  // extern "C"
