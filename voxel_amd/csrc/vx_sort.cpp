// vx_sort.cpp -- host orchestration of the GPU radix sort: the reference's
// voxel_sort_ascending/descending (bindings/voxel_c.cpp:187-198 ->
// SortOperator, ops/sort.hpp:506-541), RadixSort<u64>::Sort(data, idx, n)
// (ops/sort.hpp:83-160) as voxel_gpu_sort_pairs_u64, and voxel_topk.
#include <hip/hip_runtime_api.h>

#include <cstring>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

namespace vx {

namespace {
// Grow-only device buffers of the sort, one set per stream context.
struct SortBufs {
    uint64_t* ktmp = nullptr;
    uint32_t* vtmp = nullptr;
    uint64_t cap = 0;               // elements
    RadixWork w;
    uint32_t* skip_host = nullptr;  // pinned
    // host-API staging
    double* x = nullptr;
    uint64_t* k = nullptr;
    uint32_t* v = nullptr;
    uint64_t* k2 = nullptr;
    uint32_t* v2 = nullptr;
    uint64_t stage_cap = 0;
};

std::mutex g_sort_mutex;

SortBufs& bufs_for(StreamCtx* c) {
    static std::vector<std::pair<StreamCtx*, SortBufs*>> all;
    for (auto& p : all)
        if (p.first == c) return *p.second;
    all.emplace_back(c, new SortBufs());
    return *all.back().second;
}

bool ensure_work(SortBufs& b, uint64_t n) {
    RadixWork& w = b.w;
    if (!w.hist) {
        if (!hip_ok(hipMalloc((void**)&w.hist, 8 * 256 * 8), "hipMalloc(hist)")) return false;
        if (!hip_ok(hipMalloc((void**)&w.skip, 64), "hipMalloc(skip)")) return false;
        if (!hip_ok(hipHostMalloc((void**)&b.skip_host, 64, hipHostMallocDefault), "hipHostMalloc(skip)")) return false;
    }
    uint64_t words = radix_work_words(n);
    if (words > w.status_cap) {
        if (w.status) (void)hipFree(w.status);
        w.status = nullptr;
        if (!hip_ok(hipMalloc((void**)&w.status, words * 4), "hipMalloc(status)")) { w.status_cap = 0; return false; }
        w.status_cap = words;
    }
    if (n > b.cap) {
        if (b.ktmp) (void)hipFree(b.ktmp);
        if (b.vtmp) (void)hipFree(b.vtmp);
        b.ktmp = nullptr; b.vtmp = nullptr; b.cap = 0;
        if (!hip_ok(hipMalloc((void**)&b.ktmp, n * 8), "hipMalloc(sort keys)")) return false;
        if (!hip_ok(hipMalloc((void**)&b.vtmp, n * 4), "hipMalloc(sort payload)")) return false;
        b.cap = n;
    }
    return true;
}

bool ensure_stage(SortBufs& b, uint64_t n) {
    if (n <= b.stage_cap) return true;
    for (void* p : {(void*)b.x, (void*)b.k, (void*)b.v, (void*)b.k2, (void*)b.v2}) if (p) (void)hipFree(p);
    b.x = nullptr; b.k = nullptr; b.v = nullptr; b.k2 = nullptr; b.v2 = nullptr; b.stage_cap = 0;
    if (!hip_ok(hipMalloc((void**)&b.x, n * 8), "hipMalloc")) return false;
    if (!hip_ok(hipMalloc((void**)&b.k, n * 8), "hipMalloc")) return false;
    if (!hip_ok(hipMalloc((void**)&b.v, n * 4), "hipMalloc")) return false;
    if (!hip_ok(hipMalloc((void**)&b.k2, n * 8), "hipMalloc")) return false;
    if (!hip_ok(hipMalloc((void**)&b.v2, n * 4), "hipMalloc")) return false;
    b.stage_cap = n;
    return true;
}
}  // namespace

// Stable sort of (key, payload) pairs by key; output in kout/vout (vin and
// vout both NULL: keys only; row_index: the payload of row i is i, made by
// the first pass).  The input is never written.  Returns the number of digit
// passes executed.
static int sort_pairs_impl(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                           StreamCtx* c, bool row_index) {
    if (n >= (1ull << 30)) { set_error("sort: %llu keys exceed the sort's 2^30-row limit",
                                       (unsigned long long)n); return -1; }
    SortBufs& b = bufs_for(c);
    if (!ensure_work(b, n)) return -1;
    hipStream_t st = c->stream;
    LsdDigit dig;
    if (!hip_ok(launch_constant_bytes(kin, n, b.w, st), "key bytes")) return -1;
    if (!hip_ok(hipMemcpyAsync(b.skip_host, b.w.skip, 32, hipMemcpyDeviceToHost, st), "read skip flags")) return -1;
    if (!hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize")) return -1;
    int passes[8], np = 0;
    for (int p = 0; p < 8; ++p)
        if (!b.skip_host[p]) passes[np++] = p;
    if (np == 0) {
        if (kout != kin && !hip_ok(hipMemcpyAsync(kout, kin, n * 8, hipMemcpyDeviceToDevice, st), "copy")) return -1;
        if (row_index) return hip_ok(launch_iota(vout, n, st), "iota") ? 0 : -1;
        if (vout != vin && !hip_ok(hipMemcpyAsync(vout, vin, n * 4, hipMemcpyDeviceToDevice, st), "copy")) return -1;
        return 0;
    }
    // the last pass writes kout; passes alternate between kout and the temporaries
    const uint64_t* ks = kin;
    const uint32_t* vs = vin;
    for (int i = 0; i < np; ++i) {
        bool to_out = ((np - 1 - i) % 2) == 0;
        uint64_t* kd = to_out ? kout : b.ktmp;
        uint32_t* vd = (vin || row_index) ? (to_out ? vout : b.vtmp) : nullptr;    // keys only without a payload
        if (!hip_ok(radix_pass(ks, vs, kd, vd, n, dig, passes[i], b.w, st, nullptr, nullptr, row_index && i == 0),
                    "radix pass"))
            return -1;
        ks = kd;
        vs = vd;
    }
    return np;
}

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n,
                   StreamCtx* c) {
    return sort_pairs_impl(kin, vin, kout, vout, n, c, false);
}

// The same with the row index as payload (vout required).
int sort_pairs_u64_rowidx(const uint64_t* kin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c) {
    return sort_pairs_impl(kin, nullptr, kout, vout, n, c, true);
}

}  // namespace vx

using namespace vx;

namespace {
void argsort_f64_host(const double* data, size_t count, int descending, uint32_t* out) {
    set_error("%s", "");
    if (count == 0) return;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return;
    SortBufs& b = bufs_for(c);
    if (!ensure_stage(b, count)) return;
    hipStream_t st = c->stream;
    if (!hip_ok(hipMemcpyAsync(b.x, data, count * 8, hipMemcpyHostToDevice, st), "copy data to HBM")) return;
    if (!hip_ok(launch_f64_keys(b.x, count, descending, b.k, b.v, st), "key transform")) return;
    if (sort_pairs_u64(b.k, b.v, b.k2, b.v2, count, c) < 0) return;
    if (!hip_ok(hipMemcpyAsync(out, b.v2, count * 4, hipMemcpyDeviceToHost, st), "read back indices")) return;
    (void)hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}
}  // namespace

extern "C" {

// SortOperator::SortAscending/Descending (ops/sort.hpp:506-541): the stable
// argsort of the f64 column; equal values (including -0.0 and +0.0) keep
// their input order, descending too (MergeSort's ">=" takes the left run).
void voxel_sort_ascending(const double* data, size_t count, uint32_t* indices_out) {
    argsort_f64_host(data, count, 0, indices_out);
}
void voxel_sort_descending(const double* data, size_t count, uint32_t* indices_out) {
    argsort_f64_host(data, count, 1, indices_out);
}

// TopKSelect (ops/sort.hpp:490-502): the k largest (smallest) values in
// descending (ascending) order; k >= count copies the input unchanged.
void voxel_topk(const double* data, size_t count, size_t k, double* out, int largest) {
    set_error("%s", "");
    if (k >= count) { if (count) std::memcpy(out, data, count * 8); return; }
    if (k == 0) return;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return;
    SortBufs& b = bufs_for(c);
    if (!ensure_stage(b, count)) return;
    hipStream_t st = c->stream;
    if (!hip_ok(hipMemcpyAsync(b.x, data, count * 8, hipMemcpyHostToDevice, st), "copy data to HBM")) return;
    if (!hip_ok(launch_f64_keys(b.x, count, largest ? 1 : 0, b.k, b.v, st), "key transform")) return;
    if (sort_pairs_u64(b.k, b.v, b.k2, b.v2, count, c) < 0) return;
    double* vals = reinterpret_cast<double*>(b.k);          // free after the sort
    if (!hip_ok(launch_gather_f64(b.x, b.v2, k, vals, st), "gather")) return;
    if (!hip_ok(hipMemcpyAsync(out, vals, k * 8, hipMemcpyDeviceToHost, st), "read back")) return;
    (void)hip_ok(hipStreamSynchronize(st), "hipStreamSynchronize");
}

int voxel_gpu_sort_pairs_u64(const uint64_t* keys_in, const uint32_t* vals_in, uint64_t* keys_out,
                             uint32_t* vals_out, size_t n, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (n == 0) return 0;
    if (!vals_in)       // payload = row index (RadixSort's usual indices argument)
        return sort_pairs_u64_rowidx(keys_in, keys_out, vals_out, n, c) < 0 ? -1 : 0;
    return sort_pairs_u64(keys_in, vals_in, keys_out, vals_out, n, c) < 0 ? -1 : 0;
}

int voxel_gpu_gather_f64(const double* x_dev, const uint32_t* idx_dev, size_t k, double* out_dev, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (k && (!x_dev || !idx_dev || !out_dev)) { set_error("gather: null pointer"); return -1; }
    return hip_ok(launch_gather_f64(x_dev, idx_dev, k, out_dev, c->stream), "gather launch") ? 0 : -1;
}

int voxel_gpu_f64_to_i64(const double* x_dev, size_t n, int64_t* out_dev, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (n && (!x_dev || !out_dev)) { set_error("f64_to_i64: null pointer"); return -1; }
    return hip_ok(launch_f64_to_i64(x_dev, n, out_dev, c->stream), "convert launch") ? 0 : -1;
}

int voxel_gpu_argsort_f64(const double* x_dev, size_t n, int descending, uint32_t* idx_out_dev, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (n == 0) return 0;
    SortBufs& b = bufs_for(c);
    if (!ensure_stage(b, n)) return -1;
    if (!hip_ok(launch_f64_keys(x_dev, n, descending, b.k, b.v, c->stream), "key transform")) return -1;
    return sort_pairs_u64(b.k, b.v, b.k2, idx_out_dev, n, c) < 0 ? -1 : 0;
}

}  // extern "C"
