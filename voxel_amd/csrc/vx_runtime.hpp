// vx_runtime.hpp -- host runtime of libvoxel_c: device bring-up, per-stream
// scratch, error reporting, predicate lowering.  Host C++ only.
#pragma once
#include <hip/hip_runtime_api.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"

namespace vx {

// Thread-local last error (voxel_last_error).
void set_error(const char* fmt, ...) __attribute__((format(printf, 1, 2)));
const char* last_error();

// Device-side slab a launch writes its result into, mirrored in pinned host
// memory for the read-back.
struct ResultSlab {
    ScanOut scan;
    GroupOut group;
    uint64_t fold[2];
    uint64_t last;
    uint64_t pad[5];
};

// Scratch and result slab bound to one stream; launches on one stream are
// ordered, so they may share it.
struct StreamCtx {
    hipStream_t stream = nullptr;
    Scratch scr{};
    ResultSlab* dev = nullptr;
    ResultSlab* host = nullptr;
    int last_dtype = -1;     // dtype of the last scan launched on this stream
};

// Returns false (and sets the error) when no usable HIP device exists.
bool runtime_ok();
// Context of a caller stream; NULL selects the device's default stream.
StreamCtx* ctx_for(void* stream);
// Serialises library calls that share scratch (all C ABI entry points).
std::mutex& api_mutex();

// Checks a HIP status; on failure sets the error and returns false.
bool hip_ok(hipError_t e, const char* what);

// Lowers (dtype, cmp, threshold bits) to the device predicate form.
//   ordered: NaN fails every compare, NE included (the JIT's vcmppd).
PredDev make_pred(int dtype, int cmp, uint64_t tbits, bool ordered = false);
// Truth-table predicate "x != 0 and not NaN" (VCOUNT's lane rule).
PredDev make_nonzero_pred(int dtype);
// A predicate that never passes (generic VFILTER modes 6 and 7).
PredDev make_never_pred(int dtype);
// The conjunction of two range predicates (non-table form) as one range
// predicate; false when it is not one (a "!= t" with anything but "true").
bool merge_preds(int dtype, const PredDev& a, const PredDev& b, PredDev* out);

// Combines a compensated pair with an initial value and rounds to T.
double finish_sum_f64(double acc0, double s, double c);
float finish_sum_f32(float acc0, float s, float c);

// Scan result (device form) -> voxel_gpu_agg for a column dtype.
void fill_agg(int dtype, const ScanOut& r, voxel_gpu_agg* out);
// Checks a scan descriptor and lowers it; false with the error set.
bool make_scan_args(const voxel_gpu_scan_desc* d, ScanArgs& g);

// Device buffer that mirrors a borrowed host column.
struct DevColumn {
    void* dev = nullptr;
    size_t bytes = 0;
    bool valid = false;      // holds the host data (resident mode)
    ~DevColumn();
    bool ensure(size_t bytes_needed);
};

}  // namespace vx
