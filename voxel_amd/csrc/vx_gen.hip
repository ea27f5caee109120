// vx_gen.hip -- deterministic synthetic columns generated in HBM.
//
// Row i of a column with seed s draws u = SplitMix64 output number i of the
// generator seeded with s:  z = s + (i + 1) * 0x9E3779B97F4A7C15, then the
// SplitMix64 finaliser (Steele, Lea, Flood 2014).  From u:
//   floats: lo + (hi - lo) * ((u >> 11) * 2^-53)   (f32: that f64 rounded)
//   ints:   lo + floor(u * (hi - lo) / 2^64)        (128-bit product)
// so any row can be recomputed on the host from (seed, i) alone.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

template <typename T>
__global__ __launch_bounds__(256) void k_fill(T* __restrict__ out, uint64_t n, uint64_t seed, double lo, double hi,
                                              int64_t ilo, uint64_t irange) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        uint64_t u = mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ull);
        if constexpr (std::is_floating_point<T>::value) {
            out[i] = (T)(lo + (hi - lo) * ((double)(u >> 11) * 0x1.0p-53));
        } else {
            uint64_t r = __umul64hi(u, irange);
            out[i] = (T)(ilo + (int64_t)r);
        }
    }
}

}  // namespace
}  // namespace vx

using namespace vx;

extern "C" int voxel_gpu_fill_uniform(int dtype, void* dev, size_t n, uint64_t seed, double lo, double hi,
                                      void* stream) {
    set_error("%s", "");
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (n == 0) return 0;
    unsigned blocks = (unsigned)((n + 255) / 256 < 65536 ? (n + 255) / 256 : 65536);
    int64_t ilo = (int64_t)lo;
    uint64_t irange = (uint64_t)((int64_t)hi - (int64_t)lo);
    switch (dtype) {
    case DT_F64: hipLaunchKernelGGL(k_fill<double>, dim3(blocks), dim3(256), 0, c->stream, (double*)dev, n, seed, lo, hi, ilo, irange); break;
    case DT_F32: hipLaunchKernelGGL(k_fill<float>, dim3(blocks), dim3(256), 0, c->stream, (float*)dev, n, seed, lo, hi, ilo, irange); break;
    case DT_I64: case DT_U64: hipLaunchKernelGGL(k_fill<int64_t>, dim3(blocks), dim3(256), 0, c->stream, (int64_t*)dev, n, seed, lo, hi, ilo, irange); break;
    case DT_I32: case DT_U32: hipLaunchKernelGGL(k_fill<int32_t>, dim3(blocks), dim3(256), 0, c->stream, (int32_t*)dev, n, seed, lo, hi, ilo, irange); break;
    default: set_error("fill: dtype %d not supported", dtype); return -1;
    }
    return hip_ok(hipGetLastError(), "fill launch") ? 0 : -1;
}
