// vx_window.hip -- the window family of Engine<f64> (exec/engine.hpp:183-300,
// exposed by voxel_py as EngineF64.window_*; bindings/voxel_bindings.cpp:127-180).
//
// Sliding windows: one workgroup owns a tile of consecutive outputs and stages
// the tile's rows plus the window overhang in LDS once; each thread then
// computes its outputs exactly as the reference does, summing every window
// left to right (so sums, means, variances and standard deviations are
// bit-identical to the reference's loops), for five consecutive outputs per
// thread that share the window rows held in registers.  Reads per output stay
// in LDS and registers, so HBM traffic is one read of the column (plus the
// overhang, served by L2) and one write of the outputs.
//
// Quantile: the lo-th and (lo+1)-th smallest of a window are found by rank
// counting over the LDS-resident window (w^2 compares, w <= 128), which
// selects the same values as the reference's nth_element followed by the
// minimum right of it.
//
// Time windows: rows are mapped to their bucket, the buckets are summed by
// the group-by (compensated sums, vx_groupby.hip) and added onto out.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <cmath>
#include <mutex>

#include "../../include/voxel_c.h"
#include "vx_groupby.hpp"
#include "vx_kernels.hpp"
#include "vx_runtime.hpp"
#include "vx_segagg.hpp"

#pragma clang fp contract(off)

namespace vx {
namespace {

constexpr int kWThreads = 256;
constexpr int kWRows = 5;                         // consecutive outputs per thread (sum family; odd)
constexpr int kWTile = kWThreads * kWRows;        // outputs per workgroup
constexpr int kMaxWindow = 255;                   // the window is a u8 (engine.hpp:197)
constexpr int kQuantMax = 128;                    // WindowQuantile's buffer (engine.hpp:284)

// K_NTH: the q-th smallest value of each window (q an index, as the
// WQUANTILE word selects it).
enum WinKind { K_DELTA = 0, K_SUM = 1, K_MEAN = 2, K_STD = 3, K_VAR = 4, K_QUANT = 5, K_MEAN_MUL = 6, K_NTH = 7 };

// a / w, correctly rounded, for an integer window length w (1..255) and y =
// RN(1 / w): Markstein's division with a correctly rounded reciprocal (q0 =
// a * y, the residual r = fma(-q0, w, a) exact, q = fma(r, y, q0)); checked
// equal to the IEEE quotient on 10^9 random operands for every w.  Outside
// the range where that holds (zero, subnormal, huge, inf, NaN) the plain
// division, a branch that almost no wave takes.
__device__ __forceinline__ double div_window(double a, double w, double y) {
    const double m = __builtin_fabs(a);
    if (m >= 0x1p-1000 && m <= 0x1p1020) {
        const double q0 = a * y;
        const double r = __builtin_fma(-q0, w, a);
        return __builtin_fma(r, y, q0);
    }
    return a / w;
}

// Thread t owns the kWRows consecutive outputs o = t*kWRows + r.  The window
// rows are walked in chunks of kWRows: each chunk reads kWRows new values
// from LDS (lanes kWRows doubles apart: an odd stride, so the 64-bit reads
// are bank-conflict-free) and does kWRows^2 adds from registers, so LDS
// reads per output fall from w to about w / kWRows.  Every output's sum still
// runs left to right over its own window.
template <int KIND>
__global__ __launch_bounds__(kWThreads) void k_window(const double* __restrict__ d, uint64_t n,
                                                      double* __restrict__ out, uint64_t nout, int w) {
    constexpr int R = kWRows;
    __shared__ double x[kWTile + kMaxWindow + R];
    const uint64_t base = (uint64_t)blockIdx.x * kWTile;
    const int span = kWTile + w - 1 + R;
    for (int i = threadIdx.x; i < span; i += kWThreads) {
        const uint64_t g = base + (uint64_t)i;
        x[i] = g < n ? d[g] : 0.0;
    }
    __syncthreads();
    const double* xo = x + threadIdx.x * R;
    const int wfull = w - w % R;
    double s[R];
#pragma unroll
    for (int r = 0; r < R; ++r) s[r] = 0.0;
    {
        double v[2 * R];
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = xo[i];
#pragma unroll 2
        for (int k0 = 0; k0 < wfull; k0 += R) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[R + i] = xo[k0 + R + i];
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int r = 0; r < R; ++r) s[r] += v[i + r];
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = v[R + i];
        }
        for (int k = wfull; k < w; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) s[r] += xo[k + r];
    }
    const double wd = (double)w;
    double res[R];
    if constexpr (KIND == K_SUM) {
#pragma unroll
        for (int r = 0; r < R; ++r) res[r] = s[r];
    } else if constexpr (KIND == K_MEAN) {
        const double wy = 1.0 / wd;
#pragma unroll
        for (int r = 0; r < R; ++r) res[r] = div_window(s[r], wd, wy);   // WindowMean divides (engine.hpp:221)
    } else if constexpr (KIND == K_MEAN_MUL) {
        const double winv = 1.0 / wd;                             // WINDOW_MEAN multiplies (engine.hpp:3214)
#pragma unroll
        for (int r = 0; r < R; ++r) res[r] = s[r] * winv;
    } else {
        const double wy = 1.0 / wd;
        double mean[R], q[R];
#pragma unroll
        for (int r = 0; r < R; ++r) { mean[r] = div_window(s[r], wd, wy); q[r] = 0.0; }
        double v[2 * R];
#pragma unroll
        for (int i = 0; i < R; ++i) v[i] = xo[i];
#pragma unroll 2
        for (int k0 = 0; k0 < wfull; k0 += R) {
#pragma unroll
            for (int i = 0; i < R; ++i) v[R + i] = xo[k0 + R + i];
#pragma unroll
            for (int i = 0; i < R; ++i)
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const double e = v[i + r] - mean[r];
                    q[r] += e * e;
                }
#pragma unroll
            for (int i = 0; i < R; ++i) v[i] = v[R + i];
        }
        for (int k = wfull; k < w; ++k)
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const double e = xo[k + r] - mean[r];
                q[r] += e * e;
            }
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const double v = div_window(q[r], wd, wy);
            res[r] = KIND == K_STD ? __builtin_sqrt(v) : v;
        }
    }
    // stage the outputs through LDS so the global stores are coalesced
    __syncthreads();
#pragma unroll
    for (int r = 0; r < R; ++r) x[threadIdx.x * R + r] = res[r];
    __syncthreads();
    for (int i = threadIdx.x; i < kWTile; i += kWThreads) {
        const uint64_t j = base + (uint64_t)i;
        if (j < nout) out[j] = x[i];
    }
}

__global__ __launch_bounds__(kWThreads) void k_window_delta(const double* __restrict__ d, double* __restrict__ out,
                                                            uint64_t nout) {
    const uint64_t stride = (uint64_t)gridDim.x * kWThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kWThreads + threadIdx.x; i < nout; i += stride) out[i] = d[i + 1] - d[i];
}

// One output per thread: rank counting over the first ww rows of its window.
__global__ __launch_bounds__(kWThreads) void k_window_quantile(const double* __restrict__ d, uint64_t n,
                                                               double* __restrict__ out, uint64_t nout, int ww,
                                                               int lo, int hi, double frac, int* __restrict__ nan_seen) {
    __shared__ double x[kWThreads + kQuantMax];
    const uint64_t base = (uint64_t)blockIdx.x * kWThreads;
    const int span = kWThreads + ww - 1;
    for (int i = threadIdx.x; i < span; i += kWThreads) {
        const uint64_t g = base + (uint64_t)i;
        x[i] = g < n ? d[g] : 0.0;
    }
    __syncthreads();
    const uint64_t j = base + threadIdx.x;
    if (j >= nout) return;
    const double* win = x + threadIdx.x;
    double vlo = 0.0, vhi = 0.0;
    bool nan = false;
    for (int i = 0; i < ww; ++i) {
        const double xi = win[i];
        nan |= xi != xi;
        int lt = 0, le = 0;
        for (int k = 0; k < ww; ++k) {
            const double v = win[k];
            lt += v < xi;
            le += v <= xi;
        }
        if (lt <= lo && lo < le) vlo = xi;                 // xi is the lo-th smallest
        if (lt <= hi && hi < le) vhi = xi;
    }
    if (nan) *nan_seen = 1;
    out[j] = lo == hi ? vlo : vlo + (vhi - vlo) * frac;
}

__global__ __launch_bounds__(kWThreads) void k_time_bucket(const int64_t* __restrict__ ts, uint64_t n, int64_t start,
                                                           int64_t dur, uint64_t nbuckets, int64_t* __restrict__ keys) {
    const uint64_t stride = (uint64_t)gridDim.x * kWThreads;
    for (uint64_t i = (uint64_t)blockIdx.x * kWThreads + threadIdx.x; i < n; i += stride) {
        const int64_t b = (int64_t)((uint64_t)ts[i] - (uint64_t)start) / dur;   // truncating, as C++ does
        keys[i] = (b >= 0 && (uint64_t)b < nbuckets) ? b : -1;
    }
}

// out[key] += the group's compensated sum (out folded in by TwoSum).
__global__ __launch_bounds__(kWThreads) void k_time_apply(const int64_t* __restrict__ keys, const GBin* __restrict__ rec,
                                                          uint64_t ng, double* __restrict__ out) {
    const uint64_t g = (uint64_t)blockIdx.x * kWThreads + threadIdx.x;
    if (g >= ng) return;
    const int64_t k = keys[g];
    if (k < 0) return;
    const double o = out[k], s = rec[g].s, c = rec[g].c;
    if (!__builtin_isfinite(s) || !__builtin_isfinite(o)) {
        out[k] = o + s;
        return;
    }
    const double t = o + s, bp = t - o;
    const double e = (o - (t - bp)) + (s - bp);
    out[k] = t + (e + c);
}

__global__ void k_clear_flag(int* f) { *f = 0; }

unsigned blocks_for(uint64_t n, unsigned per) { return (unsigned)((n + per - 1) / per); }

struct WindowBufs {
    DevColumn keys;
    int* flag_dev = nullptr;
    int* flag_host = nullptr;
    bool ensure_flag() {
        if (flag_dev) return true;
        return hip_ok(hipMalloc((void**)&flag_dev, sizeof(int)), "hipMalloc") &&
               hip_ok(hipHostMalloc((void**)&flag_host, sizeof(int), hipHostMallocDefault), "hipHostMalloc");
    }
};
WindowBufs& window_bufs() {                       // kept to process end (no hipFree at exit)
    static WindowBufs* b = new WindowBufs();
    return *b;
}

// Host side of the reference's argument checks; returns the output count N.
int64_t window_count(int kind, uint64_t count, uint64_t out_len, int w, double q) {
    if (kind == K_DELTA) return count < 2 ? 0 : (int64_t)std::min<uint64_t>(count - 1, out_len);
    if (kind == K_SUM || kind == K_MEAN) {
        if (w == 0 || count < (uint64_t)w) return 0;
    } else if (kind == K_STD || kind == K_VAR) {
        if (w < 2 || count < (uint64_t)w) return 0;
    } else if (kind == K_NTH) {
        if (w < 2 || w > kQuantMax || count < (uint64_t)w || !(q >= 0.0 && q < (double)w)) return 0;
    } else if (w < 2 || count < (uint64_t)w || !(q >= 0.0 && q <= 1.0)) {
        return 0;
    }
    return (int64_t)std::min<uint64_t>(count - (uint64_t)w + 1, out_len);
}

}  // namespace

// The window family's launches (caller holds the library lock): returns N.
int64_t window_launch(int kind, const double* d, uint64_t count, uint64_t out_len, int window, double* out,
                      StreamCtx* c, double q) {
    const int64_t N = window_count(kind == K_MEAN_MUL ? K_MEAN : kind, count, out_len, window, q);
    if (N <= 0) return 0;
    const uint64_t nout = (uint64_t)N;
    switch (kind) {
    case K_DELTA:
        hipLaunchKernelGGL(k_window_delta, dim3(std::min<uint64_t>(blocks_for(nout, kWThreads), 65536)),
                           dim3(kWThreads), 0, c->stream, d, out, nout);
        break;
    case K_SUM:
        hipLaunchKernelGGL(k_window<K_SUM>, dim3(blocks_for(nout, kWTile)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, window);
        break;
    case K_MEAN:
        hipLaunchKernelGGL(k_window<K_MEAN>, dim3(blocks_for(nout, kWTile)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, window);
        break;
    case K_STD:
        hipLaunchKernelGGL(k_window<K_STD>, dim3(blocks_for(nout, kWTile)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, window);
        break;
    case K_VAR:
        hipLaunchKernelGGL(k_window<K_VAR>, dim3(blocks_for(nout, kWTile)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, window);
        break;
    case K_MEAN_MUL:
        hipLaunchKernelGGL(k_window<K_MEAN_MUL>, dim3(blocks_for(nout, kWTile)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, window);
        break;
    default: {
        WindowBufs& b = window_bufs();
        if (!b.ensure_flag()) return -1;
        const int ww = std::min(window, kQuantMax);
        int lo, hi;
        double frac;
        if (kind == K_NTH) {                              // one order statistic, no interpolation
            lo = hi = (int)q;
            frac = 0.0;
        } else {
            const double pos = q * (double)(ww - 1);      // engine.hpp:286-289
            lo = (int)pos;
            if (lo >= ww) lo = ww - 1;
            hi = lo + 1 < ww ? lo + 1 : lo;
            frac = pos - (double)lo;
        }
        hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c->stream, b.flag_dev);
        hipLaunchKernelGGL(k_window_quantile, dim3(blocks_for(nout, kWThreads)), dim3(kWThreads), 0, c->stream, d,
                           (uint64_t)count, out, nout, ww, lo, hi, frac, b.flag_dev);
        if (!hip_ok(hipGetLastError(), "window quantile launch")) return -1;
        if (!hip_ok(hipMemcpyAsync(b.flag_host, b.flag_dev, sizeof(int), hipMemcpyDeviceToHost, c->stream),
                    "read back NaN flag") ||
            !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
            return -1;
        if (*b.flag_host) {
            set_error("window_quantile: NaN inside a window (the reference's nth_element order is unspecified); "
                      "out is unspecified");
            return -1;
        }
        return N;
    }
    }
    if (!hip_ok(hipGetLastError(), "window launch")) return -1;
    return N;
}

namespace {
int64_t window_run(int kind, const double* d, size_t count, double* out, size_t out_len, int window, double q,
                   void* stream) {
    set_error("%s", "");
    if (window < 0 || window > kMaxWindow) { set_error("window %d is not a u8 (0..255)", window); return -1; }
    if ((count && !d) || (out_len && !out)) { set_error("window: null device pointer"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    return window_launch(kind, d, count, out_len, window, out, c, q);
}
}  // namespace
}  // namespace vx

using namespace vx;

extern "C" {

int64_t voxel_gpu_window_delta_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                   void* stream) {
    return window_run(K_DELTA, data_dev, count, out_dev, out_len, 0, 0.0, stream);
}
int64_t voxel_gpu_window_sum_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len, int window,
                                 void* stream) {
    return window_run(K_SUM, data_dev, count, out_dev, out_len, window, 0.0, stream);
}
int64_t voxel_gpu_window_mean_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len, int window,
                                  void* stream) {
    return window_run(K_MEAN, data_dev, count, out_dev, out_len, window, 0.0, stream);
}
int64_t voxel_gpu_window_stddev_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                    int window, void* stream) {
    return window_run(K_STD, data_dev, count, out_dev, out_len, window, 0.0, stream);
}
int64_t voxel_gpu_window_variance_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                      int window, void* stream) {
    return window_run(K_VAR, data_dev, count, out_dev, out_len, window, 0.0, stream);
}
int64_t voxel_gpu_window_quantile_f64(const double* data_dev, size_t count, double* out_dev, size_t out_len,
                                      int window, double quantile, void* stream) {
    return window_run(K_QUANT, data_dev, count, out_dev, out_len, window, quantile, stream);
}

int64_t voxel_gpu_time_window_sum_f64(const int64_t* ts_dev, const double* vals_dev, size_t count,
                                      int64_t start_time, int64_t duration_sec, double* out_dev, size_t out_len,
                                      void* stream) {
    set_error("%s", "");
    if (duration_sec <= 0 || count == 0 || out_len == 0) return 0;     // engine.hpp:232
    if (!ts_dev || !vals_dev || !out_dev) { set_error("time_window_sum: null device pointer"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    WindowBufs& b = window_bufs();
    if (!b.keys.ensure(count * sizeof(int64_t))) return -1;
    int64_t* keys = static_cast<int64_t*>(b.keys.dev);
    hipLaunchKernelGGL(k_time_bucket, dim3(std::min<uint64_t>(blocks_for(count, kWThreads), 65536)), dim3(kWThreads), 0,
                       c->stream, ts_dev, (uint64_t)count, start_time, duration_sec, (uint64_t)out_len, keys);
    if (!hip_ok(hipGetLastError(), "time bucket launch")) return -1;
    int64_t* gk;
    GBin* gr;
    const int64_t ng = groupby_f64(DT_I64, keys, vals_dev, count, &gk, &gr, c, 0);
    if (ng < 0) return -1;
    if (ng == 0) return 1;
    hipLaunchKernelGGL(k_time_apply, dim3(blocks_for((uint64_t)ng, kWThreads)), dim3(kWThreads), 0, c->stream, gk, gr,
                       (uint64_t)ng, out_dev);
    if (!hip_ok(hipGetLastError(), "time window launch")) return -1;
    int64_t last = -1;
    if (!hip_ok(hipMemcpyAsync(&last, gk + (ng - 1), sizeof(int64_t), hipMemcpyDeviceToHost, c->stream),
                "read back last bucket") ||
        !hip_ok(hipStreamSynchronize(c->stream), "hipStreamSynchronize"))
        return -1;
    return (last >= 0 ? last : 0) + 1;                                 // maxBucket + 1 (engine.hpp:240)
}

}  // extern "C"
