// vx_agg.cpp -- the reference's voxel_agg_* C ABI (HashAggregator<u32,f64>,
// bindings/voxel_c.cpp:167-182) and voxel_gpu_groupby_f64, on the GPU
// group-by of vx_groupby.hip.  An aggregator keeps its groups in HBM as
// records sorted by key; each accumulate call groups its batch on the GPU and
// merges the batch's records into the kept ones.
#include <hip/hip_runtime_api.h>

#include <cstring>

#include "../../include/voxel_c.h"
#include "vx_groupby.hpp"
#include "vx_kernels.hpp"
#include "vx_runtime.hpp"

using namespace vx;

namespace {

struct Agg {
    int64_t* keys = nullptr;
    GBin* recs = nullptr;
    uint64_t n = 0, cap = 0;
    uint32_t* kstage = nullptr;
    double* vstage = nullptr;
    uint64_t scap = 0;
    ~Agg() {
        if (keys) (void)hipFree(keys);
        if (recs) (void)hipFree(recs);
        if (kstage) (void)hipFree(kstage);
        if (vstage) (void)hipFree(vstage);
    }
};

bool keep(Agg& a, const int64_t* k, const GBin* r, uint64_t n, StreamCtx* c) {
    if (n > a.cap) {
        if (a.keys) (void)hipFree(a.keys);
        if (a.recs) (void)hipFree(a.recs);
        a.keys = nullptr; a.recs = nullptr; a.cap = 0;
        uint64_t cap = n * 2;
        if (!hip_ok(hipMalloc((void**)&a.keys, cap * 8), "hipMalloc") ||
            !hip_ok(hipMalloc((void**)&a.recs, cap * sizeof(GBin)), "hipMalloc"))
            return false;
        a.cap = cap;
    }
    if (n) {
        if (!hip_ok(hipMemcpyAsync(a.keys, k, n * 8, hipMemcpyDeviceToDevice, c->stream), "copy") ||
            !hip_ok(hipMemcpyAsync(a.recs, r, n * sizeof(GBin), hipMemcpyDeviceToDevice, c->stream), "copy"))
            return false;
    }
    a.n = n;
    return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
}

}  // namespace

extern "C" {

void* voxel_agg_create(void) { return new Agg(); }
void voxel_agg_destroy(void* a) { delete static_cast<Agg*>(a); }

// HashAggregator::Init (aggregate.hpp:159-166) starts an empty table.
void voxel_agg_init(void* a, size_t) { static_cast<Agg*>(a)->n = 0; }

// HashAggregator::Accumulate (aggregate.hpp:168-174) over host arrays.
void voxel_agg_accumulate(void* ah, const uint32_t* keys, const double* values, size_t count) {
    set_error("%s", "");
    Agg& a = *static_cast<Agg*>(ah);
    if (count == 0) return;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return;
    if (count > a.scap) {
        if (a.kstage) (void)hipFree(a.kstage);
        if (a.vstage) (void)hipFree(a.vstage);
        a.kstage = nullptr; a.vstage = nullptr; a.scap = 0;
        if (!hip_ok(hipMalloc((void**)&a.kstage, count * 4), "hipMalloc") ||
            !hip_ok(hipMalloc((void**)&a.vstage, count * 8), "hipMalloc"))
            return;
        a.scap = count;
    }
    if (!hip_ok(hipMemcpyAsync(a.kstage, keys, count * 4, hipMemcpyHostToDevice, c->stream), "copy keys") ||
        !hip_ok(hipMemcpyAsync(a.vstage, values, count * 8, hipMemcpyHostToDevice, c->stream), "copy values"))
        return;
    int64_t* bk;
    GBin* br;
    int64_t nb = groupby_f64(DT_U32, a.kstage, a.vstage, count, &bk, &br, c, 0);
    if (nb < 0) return;
    if (a.n == 0) { keep(a, bk, br, (uint64_t)nb, c); return; }
    int64_t* mk;
    GBin* mr;
    int64_t nm = groupby_merge(a.keys, a.recs, a.n, bk, br, (uint64_t)nb, &mk, &mr, c);
    if (nm < 0) return;
    keep(a, mk, mr, (uint64_t)nm, c);
}

// HashAggregator::GroupCount: distinct keys accumulated since Init.
size_t voxel_agg_group_count(const void* a) { return static_cast<const Agg*>(a)->n; }

// Finalize-like read-out of an aggregator (ascending key order; the
// reference emits slot order, ops/aggregate.hpp:249-261).  Host arrays of
// at least max_groups entries; any array may be NULL.
int voxel_gpu_agg_results(void* ah, int64_t* keys, uint64_t* counts, double* sums, double* avgs, double* mins,
                          double* maxs, size_t max_groups) {
    set_error("%s", "");
    Agg& a = *static_cast<Agg*>(ah);
    if (a.n > max_groups) { set_error("agg results: %llu groups > max_groups", (unsigned long long)a.n); return -1; }
    if (a.n == 0) return 0;
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(nullptr);
    if (!c) return -1;
    uint64_t n = a.n;
    char* tmp = nullptr;
    if (!hip_ok(hipMalloc((void**)&tmp, n * 48), "hipMalloc")) return -1;
    int64_t* dk = (int64_t*)tmp;
    uint64_t* dc = (uint64_t*)(tmp + n * 8);
    double* ds = (double*)(tmp + n * 16);
    double* da = (double*)(tmp + n * 24);
    double* dmn = (double*)(tmp + n * 32);
    double* dmx = (double*)(tmp + n * 40);
    bool ok = groupby_finish(a.keys, a.recs, n, dk, dc, ds, da, dmn, dmx, c) == 0;
    auto d2h = [&](void* h, const void* d) {
        if (h && ok) ok = hip_ok(hipMemcpyAsync(h, d, n * 8, hipMemcpyDeviceToHost, c->stream), "read back");
    };
    d2h(keys, dk); d2h(counts, dc); d2h(sums, ds); d2h(avgs, da); d2h(mins, dmn); d2h(maxs, dmx);
    if (ok) ok = hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    (void)hipFree(tmp);
    return ok ? 0 : -1;
}

int voxel_gpu_groupby_f64(int key_dtype, const void* keys_dev, const double* vals_dev, size_t n, size_t max_groups,
                          int64_t* out_keys, uint64_t* out_count, double* out_sum, double* out_avg, double* out_min,
                          double* out_max, uint64_t* ngroups, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    // 32-bit keys: the residue-binned dense path decides on the device and
    // writes the caller's arrays (one host synchronisation); a key range
    // wider than 512 values falls back to groupby_f64 (sampled window, then
    // the sort-based path)
    if ((key_dtype == DT_I32 || key_dtype == DT_U32) && n > 0) {
        const int64_t g = groupby_dense_device(key_dtype, keys_dev, vals_dev, n, max_groups, out_keys, out_count,
                                               out_sum, out_avg, out_min, out_max, c);
        if (g == -1) return -1;
        if (g >= 0) {
            if (ngroups) *ngroups = (uint64_t)g;
            if ((uint64_t)g > max_groups) { set_error("group-by: %lld groups > max_groups %zu", (long long)g, max_groups); return -1; }
            return 0;
        }
    }
    int64_t* k;
    GBin* r;
    int64_t g = groupby_f64(key_dtype, keys_dev, vals_dev, n, &k, &r, c, 0);
    if (g < 0) return -1;
    if (ngroups) *ngroups = (uint64_t)g;
    if ((uint64_t)g > max_groups) { set_error("group-by: %lld groups > max_groups %zu", (long long)g, max_groups); return -1; }
    if (groupby_finish(k, r, (uint64_t)g, out_keys, out_count, out_sum, out_avg, out_min, out_max, c) != 0) return -1;
    return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? 0 : -1;
}

// Forces the sort-based path (tests compare both paths).
int voxel_gpu_groupby_f64_general(int key_dtype, const void* keys_dev, const double* vals_dev, size_t n,
                                  size_t max_groups, int64_t* out_keys, uint64_t* out_count, double* out_sum,
                                  double* out_avg, double* out_min, double* out_max, uint64_t* ngroups, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    int64_t* k;
    GBin* r;
    int64_t g = groupby_f64(key_dtype, keys_dev, vals_dev, n, &k, &r, c, 1);
    if (g < 0) return -1;
    if (ngroups) *ngroups = (uint64_t)g;
    if ((uint64_t)g > max_groups) { set_error("group-by: %lld groups > max_groups %zu", (long long)g, max_groups); return -1; }
    if (groupby_finish(k, r, (uint64_t)g, out_keys, out_count, out_sum, out_avg, out_min, out_max, c) != 0) return -1;
    return hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize") ? 0 : -1;
}

}  // extern "C"
