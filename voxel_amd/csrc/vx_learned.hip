// vx_learned.hip -- learned-CDF partitioned sort of (u64 key, u32 payload)
// on gfx950.  Same output as the stable LSD radix sort (RadixSort<u64>,
// ops/sort.hpp:83-160); fewer passes over HBM when the key distribution can
// be learned.
//
//  1. sample S keys at SplitMix64-random positions, sort the sample (radix);
//  2. pick the grid (linear offsets, or float bit patterns = logarithmic)
//     under which the sample CDF is closest to a straight line;
//  3. train a 1->32->32->1 FP32 ReLU MLP on (grid position, rank) pairs of the
//     sorted sample with this file's own forward/backward kernels and Adam;
//  4. evaluate the MLP at every grid cell start, make it non-decreasing
//     (prefix max), and turn it into the fine-bucket table of TableDigit
//     (vx_radix.hpp): fb(k) is non-decreasing in k by construction;
//  5. stable-partition the pairs by coarse bucket fb >> 8 with the onesweep
//     digit passes (vx_onesweep.hpp) -- ceil(log2(buckets) / 8) passes instead
//     of eight;
//  6. one workgroup per bucket (<= kCap rows): rank by fb & 255 in LDS
//     (stable), then insertion-sort each fine run by key (stable), write back;
//  7. buckets larger than kCap are checked for a single key value (then they
//     are already in stable order); any other one is sorted by the LSD radix
//     sort, and when those are too many the whole input is sorted by the LSD
//     radix sort instead (reported in the statistics).
// The equal-width variant (model = 1) replaces step 3-4 by a CDF that is
// linear in the key over [min, max] (the method without the network).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/voxel_c.h"
#include "vx_kernels.hpp"
#include "vx_onesweep.hpp"
#include "vx_radix.hpp"
#include "vx_runtime.hpp"

#pragma clang fp contract(off)

namespace vx {

int sort_pairs_u64(const uint64_t* kin, const uint32_t* vin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

constexpr int kH = 32;                       // hidden width
constexpr int kParams = 32 + 32 + 32 * 32 + 32 + 32 + 1;   // 1153
constexpr int oW1 = 0, oB1 = 32, oW2 = 64, oB2 = 64 + 1024, oW3 = oB2 + 32, oB3 = oW3 + 32;
constexpr int kGradThreads = 64;
constexpr int kCap = 4096;                   // rows a workgroup sorts in LDS
constexpr uint32_t kLeaves = 4096;           // linear models under the MLP

namespace {

__device__ __forceinline__ float mlp_fwd(const float* __restrict__ P, float x, float* h1, float* h2) {
#pragma unroll
    for (int j = 0; j < kH; ++j) h1[j] = fmaxf(P[oW1 + j] * x + P[oB1 + j], 0.0f);
#pragma unroll
    for (int i = 0; i < kH; ++i) {
        float s = P[oB2 + i];
#pragma unroll
        for (int j = 0; j < kH; ++j) s += P[oW2 + i * kH + j] * h1[j];
        h2[i] = fmaxf(s, 0.0f);
    }
    float y = P[oB3];
#pragma unroll
    for (int i = 0; i < kH; ++i) y += P[oW3 + i] * h2[i];
    return y;
}

__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ Pg, const float* __restrict__ x,
                                                     uint64_t n, float* __restrict__ y) {
    __shared__ float P[kParams];
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) P[i] = Pg[i];
    __syncthreads();
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        float h1[kH], h2[kH];
        y[i] = mlp_fwd(P, x[i], h1, h2);
    }
}

// Mean-squared-error gradient of one 64-sample slice per workgroup:
// part[block][0..1152] = sum over the slice of dL/dparam, part[block][1153] = sum of squared errors.
__global__ __launch_bounds__(kGradThreads) void k_mlp_grad(const float* __restrict__ Pg, const float* __restrict__ x,
                                                           const float* __restrict__ yt, uint64_t n, float inv_n,
                                                           float* __restrict__ part) {
    __shared__ float P[kParams];
    __shared__ float H1[kGradThreads][kH + 1], DH1[kGradThreads][kH + 1], DH2[kGradThreads][kH + 1];
    __shared__ float H2[kGradThreads][kH + 1];
    __shared__ float DY[kGradThreads], X[kGradThreads], E2[kGradThreads];
    for (int i = threadIdx.x; i < kParams; i += kGradThreads) P[i] = Pg[i];
    __syncthreads();
    const int t = threadIdx.x;
    const uint64_t i = (uint64_t)blockIdx.x * kGradThreads + t;
    float h1[kH], h2[kH];
    float xi = 0.0f, dy = 0.0f, e = 0.0f;
    if (i < n) {
        xi = x[i];
        float y = mlp_fwd(P, xi, h1, h2);
        e = y - yt[i];
        dy = 2.0f * e * inv_n;
    } else {
#pragma unroll
        for (int j = 0; j < kH; ++j) { h1[j] = 0.0f; h2[j] = 0.0f; }
    }
    float dh2[kH], dh1[kH];
#pragma unroll
    for (int k = 0; k < kH; ++k) dh2[k] = h2[k] > 0.0f ? dy * P[oW3 + k] : 0.0f;
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int k = 0; k < kH; ++k) s += dh2[k] * P[oW2 + k * kH + j];
        dh1[j] = h1[j] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int j = 0; j < kH; ++j) { H1[t][j] = h1[j]; H2[t][j] = h2[j]; DH1[t][j] = dh1[j]; DH2[t][j] = dh2[j]; }
    DY[t] = dy; X[t] = xi; E2[t] = e * e;
    __syncthreads();
    float* out = part + (uint64_t)blockIdx.x * (kParams + 1);
    // dW2[k][j] = sum_s DH2[s][k] * H1[s][j]: 16 entries per thread, samples in order
    for (int q = 0; q < 16; ++q) {
        int idx = t * 16 + q, k = idx / kH, j = idx % kH;
        float s = 0.0f;
        for (int r = 0; r < kGradThreads; ++r) s += DH2[r][k] * H1[r][j];
        out[oW2 + idx] = s;
    }
    if (t < kH) {
        float w1 = 0.0f, b1 = 0.0f, b2 = 0.0f, w3 = 0.0f;
        for (int r = 0; r < kGradThreads; ++r) {
            w1 += DH1[r][t] * X[r];
            b1 += DH1[r][t];
            b2 += DH2[r][t];
            w3 += DY[r] * H2[r][t];
        }
        out[oW1 + t] = w1; out[oB1 + t] = b1; out[oB2 + t] = b2; out[oW3 + t] = w3;
    } else if (t == kH) {
        float b3 = 0.0f, l = 0.0f;
        for (int r = 0; r < kGradThreads; ++r) { b3 += DY[r]; l += E2[r]; }
        out[oB3] = b3;
        out[kParams] = l;
    }
}

// Sums the workgroup gradients into grad (fixed order: wave w adds partials
// w, w + 4, ..., then the four wave sums in order); Adam step when adam != 0.
// 64 parameters per workgroup, one per lane.
__global__ __launch_bounds__(256) void k_mlp_reduce(float* __restrict__ Pg, float* __restrict__ m, float* __restrict__ v,
                                                    const float* __restrict__ part, uint32_t nblocks, int adam,
                                                    float lr, float b1c, float b2c, float* __restrict__ grad_out,
                                                    float* __restrict__ loss_out, float inv_n) {
    __shared__ float ws[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int p = blockIdx.x * 64 + lane;
    float g = 0.0f;
    if (p <= kParams) {
        uint32_t b = w;
        for (; b + 12 < nblocks; b += 16) {
            float a0 = part[(uint64_t)b * (kParams + 1) + p], a1 = part[(uint64_t)(b + 4) * (kParams + 1) + p];
            float a2 = part[(uint64_t)(b + 8) * (kParams + 1) + p], a3 = part[(uint64_t)(b + 12) * (kParams + 1) + p];
            g += a0; g += a1; g += a2; g += a3;
        }
        for (; b < nblocks; b += 4) g += part[(uint64_t)b * (kParams + 1) + p];
    }
    ws[w][lane] = g;
    __syncthreads();
    if (w != 0 || p > kParams) return;
    g = ((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane];
    if (p == kParams) {
        if (loss_out) *loss_out = g * inv_n;
        return;
    }
    if (grad_out) grad_out[p] = g;
    if (adam) {
        const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
        float mm = beta1 * m[p] + (1.0f - beta1) * g;
        float vv = beta2 * v[p] + (1.0f - beta2) * g * g;
        m[p] = mm;
        v[p] = vv;
        Pg[p] -= lr * (mm / b1c) / (sqrtf(vv / b2c) + eps);
    }
}

__device__ __forceinline__ uint64_t sm_mix(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__global__ __launch_bounds__(256) void k_sample(const uint64_t* __restrict__ keys, uint64_t n, uint64_t s,
                                                uint64_t seed, uint64_t* __restrict__ out, uint32_t* __restrict__ idx) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s; i += stride) {
        uint64_t pos = __umul64hi(sm_mix(seed + (i + 1) * 0x9E3779B97F4A7C15ull), n);
        out[i] = keys[pos];
        idx[i] = (uint32_t)i;
    }
}

__global__ __launch_bounds__(256) void k_quantiles(const uint64_t* __restrict__ ss, uint64_t s, uint64_t* __restrict__ q) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= 1024) {
        uint64_t r = (uint64_t)i * (s - 1) / 1024;
        q[i] = ss[r < s ? r : s - 1];
    }
}

__device__ __forceinline__ uint64_t grid_of(uint64_t k, uint64_t kmin, uint64_t kmax, int logmode) {
    uint64_t kc = k < kmin ? kmin : (k > kmax ? kmax : k);
    return TableDigit::grid(kc - kmin, logmode);
}

// Training pairs: T points evenly spaced in rank over the sorted sample,
// (grid position / grid max, logit of the rank fraction).  Fitting the logit
// keeps the relative error of the CDF small in both tails, where a fit of
// the CDF itself leaves flat stretches that pile keys into a few buckets.
__global__ __launch_bounds__(256) void k_train_data(const uint64_t* __restrict__ ss, uint64_t s, uint64_t t,
                                                    uint64_t kmin, uint64_t kmax, int logmode, double gmax,
                                                    float* __restrict__ x, float* __restrict__ y) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; j < t; j += stride) {
        uint64_t r = j * s / t;
        uint64_t g = grid_of(ss[r], kmin, kmax, logmode);
        x[j] = gmax > 0 ? (float)((double)g / gmax) : 0.0f;
        const double q = ((double)r + 0.5) / (double)s;
        y[j] = (float)log(q / (1.0 - q));
    }
}

// Root model at every cell start c (grid c << cshift), c = 0..ncells-1, and
// R[ncells] = 1.  model 0: sigmoid of the MLP, rescaled so that its values at
// the sample minimum and maximum are 0 and 1; model 1: linear in the key.
// Clamped to [0, 1].
__device__ __forceinline__ float sigmoidf(float z) { return 1.0f / (1.0f + expf(-z)); }

__global__ __launch_bounds__(256) void k_root_cells(const float* __restrict__ Pg, uint32_t ncells, int cshift,
                                                    double gmax, int model, float* __restrict__ R) {
    __shared__ float P[kParams];
    __shared__ float ends[2];
    for (int i = threadIdx.x; i < kParams; i += blockDim.x) P[i] = Pg[i];
    __syncthreads();
    if (model == 0 && threadIdx.x < 2) {
        float h1[kH], h2[kH];
        ends[threadIdx.x] = sigmoidf(mlp_fwd(P, threadIdx.x ? 1.0f : 0.0f, h1, h2));
    }
    __syncthreads();
    const float y0 = ends[0], inv = model == 0 && ends[1] > ends[0] ? 1.0f / (ends[1] - ends[0]) : 1.0f;
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c <= ncells; c += stride) {
        float y = 1.0f;
        if (c < ncells) {
            float x = gmax > 0 ? (float)((double)(c << cshift) / gmax) : 0.0f;
            if (model == 0) {
                float h1[kH], h2[kH];
                y = (sigmoidf(mlp_fwd(P, x, h1, h2)) - y0) * inv;
            } else {
                y = x;
            }
        }
        R[c] = fminf(fmaxf(y, 0.0f), 1.0f);
    }
}

// One workgroup: in-place inclusive prefix max of y[0..n).
template <class T>
__global__ __launch_bounds__(1024) void k_prefix_max(T* __restrict__ y, uint32_t n) {
    __shared__ T part[1024];
    const uint32_t per = (n + 1023) / 1024;
    const uint32_t lo = min(threadIdx.x * per, n), hi = min(lo + per, n);
    T m = 0;
    for (uint32_t c = lo; c < hi; ++c) m = y[c] > m ? y[c] : m;
    part[threadIdx.x] = m;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        T t = threadIdx.x >= (unsigned)off ? part[threadIdx.x - off] : (T)0;
        __syncthreads();
        part[threadIdx.x] = part[threadIdx.x] > t ? part[threadIdx.x] : t;
        __syncthreads();
    }
    T run = threadIdx.x ? part[threadIdx.x - 1] : (T)0;
    for (uint32_t c = lo; c < hi; ++c) {
        run = y[c] > run ? y[c] : run;
        y[c] = run;
    }
}

// Root value at grid position g: linear between the cell-start values.
__device__ __forceinline__ float root_at(const float* __restrict__ R, uint32_t ncells, int cshift, uint64_t g) {
    uint64_t c = g >> cshift;
    if (c >= ncells) c = ncells - 1;
    float f = (float)(g - (c << cshift)) * __uint_as_float((uint32_t)(127 - cshift) << 23);
    f = f > 1.0f ? 1.0f : f;
    return R[c] + (R[c + 1] - R[c]) * f;
}

__device__ __forceinline__ uint32_t leaf_of(float r, uint32_t L) {
    float f = floorf(r * (float)L);
    return f <= 0.0f ? 0u : (f >= (float)(L - 1) ? L - 1 : (uint32_t)f);
}

// lo[l] = first sample (sorted) routed to a leaf >= l, l = 0..L
__global__ __launch_bounds__(256) void k_leaf_lo(const uint64_t* __restrict__ ss, uint64_t s, const float* __restrict__ R,
                                                 uint32_t ncells, int cshift, uint64_t kmin, uint64_t kmax, int logmode,
                                                 uint32_t L, uint32_t* __restrict__ lo) {
    uint32_t l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l > L) return;
    uint64_t a = 0, z = s;
    while (a < z) {
        uint64_t mid = a + (z - a) / 2;
        if (leaf_of(root_at(R, ncells, cshift, grid_of(ss[mid], kmin, kmax, logmode)), L) < l) a = mid + 1;
        else z = mid;
    }
    lo[l] = (uint32_t)a;
}

// CDF at every cell start from the leaf the root routes it to: the leaf's
// samples [lo, hi) are joined by a line from (grid of its first sample,
// lo / s) to (grid of its last sample, hi / s).
__global__ __launch_bounds__(256) void k_leaf_cells(const uint64_t* __restrict__ ss, uint64_t s, const float* __restrict__ R,
                                                    uint32_t ncells, int cshift, uint64_t kmin, uint64_t kmax, int logmode,
                                                    uint32_t L, const uint32_t* __restrict__ lo, double* __restrict__ y) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < ncells; c += stride) {
        const uint64_t g = c << cshift;
        const uint32_t l = leaf_of(R[c], L);
        const uint32_t a = lo[l], b = lo[l + 1];
        const double y0 = (double)a / (double)s, y1 = (double)b / (double)s;
        double v = y0;
        if (b > a) {
            const uint64_t g0 = grid_of(ss[a], kmin, kmax, logmode), g1 = grid_of(ss[b - 1], kmin, kmax, logmode);
            if (g <= g0) v = y0;
            else if (g >= g1) v = y1;
            else v = y0 + (y1 - y0) * ((double)(g - g0) / (double)(g1 - g0));
        }
        y[c] = v;
    }
}

__global__ __launch_bounds__(256) void k_widen(const float* __restrict__ R, uint32_t n, double* __restrict__ y) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) y[c] = (double)R[c];
}

// cell[c] = {first fine bucket, fine buckets in the cell} from the
// non-decreasing CDF y at the cell starts (first bucket of cell 0 is 0, the
// last cell ends at nfine)
__global__ __launch_bounds__(256) void k_table(const double* __restrict__ y, uint32_t ncells, uint32_t nfine,
                                               uint2* __restrict__ cell) {
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= ncells) return;
    auto start = [&](uint32_t i) -> uint32_t {
        if (i == 0) return 0u;
        if (i >= ncells) return nfine;
        double f = floor(y[i] * (double)nfine);
        return f <= 0.0 ? 0u : (f >= (double)nfine ? nfine : (uint32_t)f);
    };
    uint32_t a = start(c);
    cell[c] = make_uint2(a, start(c + 1) - a);
}

// start[b] = first row whose coarse bucket is >= b (rows are in bucket order)
__global__ __launch_bounds__(256) void k_bounds(const uint64_t* __restrict__ keys, uint64_t n, TableDigit dig,
                                                uint32_t nb, uint64_t* __restrict__ start) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t b = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; b <= nb; b += stride) {
        uint64_t lo = 0, hi = n;
        while (lo < hi) {
            uint64_t mid = lo + (hi - lo) / 2;
            if ((dig.fine(keys[mid]) >> 8) < b) lo = mid + 1;
            else hi = mid;
        }
        start[b] = lo;
    }
}

// Per-bucket sort of buckets of at most CAP rows: stable rank by the low 8
// bits of the fine bucket in LDS, then every fine run sorted by key (stable):
// runs of up to kShortRun rows by one lane with insertion sort, longer runs
// by the whole workgroup, each row's rank counted over the run.
constexpr uint32_t kShortRun = 32;

template <int CAP>
__global__ __launch_bounds__(256) void k_bucket_sort(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                     uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                     const uint64_t* __restrict__ start, uint32_t nb, TableDigit dig,
                                                     uint32_t* __restrict__ big, unsigned long long* __restrict__ stat) {
    constexpr int R = CAP / 256;
    __shared__ uint64_t kb[CAP];
    __shared__ uint32_t vb[CAP];
    __shared__ uint32_t wcnt[4][256];
    __shared__ uint32_t sst[257], scn[256], longrun[256];
    __shared__ uint32_t nlong;
    __shared__ uint8_t dig8[CAP];                  // fine digit of each LDS slot
    __shared__ uint8_t islong[256];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, t = threadIdx.x;
    uint32_t longest = 0, nlong_runs = 0;
    for (uint32_t b = blockIdx.x; b < nb; b += gridDim.x) {
        const uint64_t s = start[b], e = start[b + 1], m = e - s;
        if (m == 0) continue;
        if (m > CAP || m == 1) {
            for (uint64_t i = t; i < m; i += 256) { kout[s + i] = kin[s + i]; vout[s + i] = vin[s + i]; }
            if (m > 1 && t == 0) big[b] = 1;
            continue;
        }
        for (int i = t; i < 4 * 256; i += 256) (&wcnt[0][0])[i] = 0;
        if (t == 0) nlong = 0;
        __syncthreads();
        uint64_t k[R];
        uint32_t v[R], rk[R], dd[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint64_t idx = (uint64_t)w * (64 * R) + r * 64 + lane;
            bool valid = idx < m;
            k[r] = valid ? kin[s + idx] : 0;
            v[r] = valid ? vin[s + idx] : 0;
            dd[r] = valid ? (dig.fine(k[r]) & 0xFFu) : 0u;
            uint64_t peers = match_digit(dd[r], valid);
            uint32_t before = valid ? wcnt[w][dd[r]] : 0;
            uint32_t below = (uint32_t)__popcll(peers & lanemask_lt());
            if (valid && below == 0) wcnt[w][dd[r]] = before + (uint32_t)__popcll(peers);
            rk[r] = before + below;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int ww = 0; ww < 4; ++ww) { uint32_t c = wcnt[ww][t]; wcnt[ww][t] = tot; tot += c; }
        scn[t] = tot;
        __syncthreads();
        for (int off = 1; off < 256; off <<= 1) {
            uint32_t x = t >= off ? scn[t - off] : 0;
            __syncthreads();
            scn[t] += x;
            __syncthreads();
        }
        sst[t] = scn[t] - tot;
        if (t == 255) sst[256] = scn[255];
        __syncthreads();
#pragma unroll
        for (int r = 0; r < R; ++r) {
            uint64_t idx = (uint64_t)w * (64 * R) + r * 64 + lane;
            if (idx < m) {
                uint32_t pos = sst[dd[r]] + wcnt[w][dd[r]] + rk[r];
                kb[pos] = k[r];
                vb[pos] = v[r];
                dig8[pos] = (uint8_t)dd[r];
            }
        }
        __syncthreads();
        const uint32_t lo = sst[t], hi = lo + tot;
        if (tot > longest) longest = tot;
        if (tot <= kShortRun) {
            for (uint32_t i = lo + 1; i < hi; ++i) {
                uint64_t key = kb[i];
                uint32_t val = vb[i];
                uint32_t j = i;
                while (j > lo && kb[j - 1] > key) { kb[j] = kb[j - 1]; vb[j] = vb[j - 1]; --j; }
                kb[j] = key;
                vb[j] = val;
            }
        } else {
            longrun[atomicAdd(&nlong, 1u)] = t;
        }
        islong[t] = tot > kShortRun;
        __syncthreads();
        const uint32_t nl = nlong;
        nlong_runs += nl;
        // rows of short runs are final in LDS; rows of long runs are ranked below
        for (uint32_t i = t; i < m; i += 256)
            if (!islong[dig8[i]]) { kout[s + i] = kb[i]; vout[s + i] = vb[i]; }
        for (uint32_t q = 0; q < nl; ++q) {
            const uint32_t d = longrun[q], rlo = sst[d], rhi = sst[d + 1];
            for (uint32_t i = rlo + t; i < rhi; i += 256) {
                const uint64_t key = kb[i];
                uint32_t rank = 0;
                for (uint32_t j = rlo; j < rhi; ++j) {
                    uint64_t kj = kb[j];
                    rank += (kj < key) || (kj == key && j < i);
                }
                kout[s + rlo + rank] = key;
                vout[s + rlo + rank] = vb[i];
            }
        }
        __syncthreads();
    }
    for (int msk = 32; msk > 0; msk >>= 1) {
        uint32_t o = __shfl_xor(longest, msk, 64);
        longest = o > longest ? o : longest;
    }
    if (lane == 0) atomicMax(&stat[0], (unsigned long long)longest);
    if (t == 0 && nlong_runs) atomicAdd(&stat[1], (unsigned long long)nlong_runs);
}

// chunk list entries: (row start, row end, bucket first row); flags[j] = 1 when
// a row of the chunk differs from the bucket's first key
__global__ __launch_bounds__(256) void k_big_check(const uint64_t* __restrict__ keys, const uint64_t* __restrict__ chunks,
                                                   uint64_t nchunks, uint32_t* __restrict__ flag) {
    for (uint64_t j = blockIdx.x; j < nchunks; j += gridDim.x) {
        uint64_t s = chunks[3 * j], e = chunks[3 * j + 1], f = chunks[3 * j + 2];
        uint64_t k0 = keys[f];
        int diff = 0;
        for (uint64_t i = s + threadIdx.x; i < e; i += 256) diff |= keys[i] != k0;
        if (__syncthreads_or(diff) && threadIdx.x == 0) flag[j] = 1;
    }
}

// Row moves (src row, dst row, rows) per list entry: dir 0 copies
// a[src..] -> b[dst..], dir 1 copies a[dst..] -> b[src..].
__global__ __launch_bounds__(256) void k_move(const uint64_t* __restrict__ ka, const uint32_t* __restrict__ va,
                                              uint64_t* __restrict__ kb, uint32_t* __restrict__ vb,
                                              const uint64_t* __restrict__ mv, uint64_t nmv, int dir) {
    for (uint64_t j = blockIdx.x; j < nmv; j += gridDim.x) {
        const uint64_t s = mv[3 * j + (dir ? 1 : 0)], d = mv[3 * j + (dir ? 0 : 1)], len = mv[3 * j + 2];
        for (uint64_t i = threadIdx.x; i < len; i += 256) {
            kb[d + i] = ka[s + i];
            vb[d + i] = va[s + i];
        }
    }
}

unsigned gridn(uint64_t n, unsigned cap) {
    uint64_t b = (n + 255) / 256;
    return (unsigned)(b ? (b < cap ? b : cap) : 1);
}

}  // namespace

struct LearnedBufs {
    float* P = nullptr; float* m = nullptr; float* v = nullptr;
    float* part = nullptr; uint64_t part_cap = 0;
    float* tx = nullptr; float* ty = nullptr; uint64_t t_cap = 0;
    float* grad = nullptr; float* loss = nullptr; float* loss_host = nullptr;
    uint64_t* samp = nullptr; uint64_t* samp2 = nullptr; uint32_t* sidx = nullptr; uint32_t* sidx2 = nullptr; uint64_t s_cap = 0;
    float* R = nullptr; double* yd = nullptr; uint32_t* leaf_lo = nullptr;
    uint2* cell = nullptr;
    uint64_t* k1 = nullptr; uint64_t* k2 = nullptr; uint32_t* v1 = nullptr; uint32_t* v2 = nullptr; uint64_t cap = 0;
    uint64_t* start = nullptr; uint32_t* big = nullptr; uint64_t nb_cap = 0;
    unsigned long long* stat = nullptr;
    RadixWork w;
    uint32_t* skip_host = nullptr;
    uint64_t* qdev = nullptr; uint64_t* qhost = nullptr;
    unsigned long long* stat_host = nullptr;
    float* phost = nullptr;
    uint64_t* cdev = nullptr; uint64_t* chost = nullptr; uint32_t* cflag = nullptr; uint32_t* cflag_host = nullptr;
    uint64_t c_cap = 0;                       // chunk list entries
    hipEvent_t ev[8] = {};
    // training loop as one graph (rebuilt when its buffers or shape change)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    const void* graph_key[7] = {};
    uint64_t graph_T = 0;
    int graph_steps = 0;
};

struct LearnedTune {
    int sample_log2 = 20, train_points = 4096, train_steps = 100, bucket_log2 = 11;
};
static LearnedTune g_tune;

static LearnedBufs& lbufs(StreamCtx* c) {
    static std::vector<std::pair<StreamCtx*, LearnedBufs*>> all;
    for (auto& p : all) if (p.first == c) return *p.second;
    auto* b = new LearnedBufs();
    all.emplace_back(c, b);
    return *b;
}

static bool galloc(void** p, size_t bytes) { return hip_ok(hipMalloc(p, bytes ? bytes : 16), "hipMalloc(learned)"); }

static bool ensure_model(LearnedBufs& b) {
    if (b.P) return true;
    bool ok = galloc((void**)&b.P, kParams * 4) && galloc((void**)&b.m, kParams * 4) && galloc((void**)&b.v, kParams * 4) &&
              galloc((void**)&b.grad, kParams * 4) && galloc((void**)&b.loss, 64) &&
              galloc((void**)&b.R, (65536 + 2) * 4) && galloc((void**)&b.yd, (65536 + 2) * 8) &&
              galloc((void**)&b.leaf_lo, (kLeaves + 2) * 4) && galloc((void**)&b.cell, (65536 + 2) * 8) && galloc((void**)&b.stat, 64) &&
              hip_ok(hipHostMalloc((void**)&b.loss_host, 64, hipHostMallocDefault), "hipHostMalloc") &&
              hip_ok(hipHostMalloc((void**)&b.skip_host, 64, hipHostMallocDefault), "hipHostMalloc") &&
              hip_ok(hipHostMalloc((void**)&b.qhost, 1025 * 8, hipHostMallocDefault), "hipHostMalloc") &&
              hip_ok(hipHostMalloc((void**)&b.stat_host, 64, hipHostMallocDefault), "hipHostMalloc") &&
              hip_ok(hipHostMalloc((void**)&b.phost, kParams * 4, hipHostMallocDefault), "hipHostMalloc") &&
              galloc((void**)&b.qdev, 1025 * 8) &&
              galloc((void**)&b.w.hist, 8 * 256 * 8) && galloc((void**)&b.w.offsets, 8 * 256 * 8) &&
              galloc((void**)&b.w.skip, 64);
    for (int i = 0; ok && i < 8; ++i) ok = hip_ok(hipEventCreate(&b.ev[i]), "hipEventCreate");
    return ok;
}

constexpr uint64_t kChunk = 65536;           // rows per chunk of the fix-up lists

static bool ensure_chunks(LearnedBufs& b, uint64_t nch) {
    if (nch <= b.c_cap) return true;
    if (b.cdev) (void)hipFree(b.cdev);
    if (b.cflag) (void)hipFree(b.cflag);
    if (b.chost) (void)hipHostFree(b.chost);
    if (b.cflag_host) (void)hipHostFree(b.cflag_host);
    b.cdev = nullptr; b.cflag = nullptr; b.chost = nullptr; b.cflag_host = nullptr; b.c_cap = 0;
    uint64_t cap = nch < 1024 ? 1024 : nch * 2;
    if (!galloc((void**)&b.cdev, cap * 24) || !galloc((void**)&b.cflag, cap * 4) ||
        !hip_ok(hipHostMalloc((void**)&b.chost, cap * 24, hipHostMallocDefault), "hipHostMalloc") ||
        !hip_ok(hipHostMalloc((void**)&b.cflag_host, cap * 4, hipHostMallocDefault), "hipHostMalloc"))
        return false;
    b.c_cap = cap;
    return true;
}

static bool ensure_train(LearnedBufs& b, uint64_t t) {
    uint64_t blocks = (t + kGradThreads - 1) / kGradThreads;
    if (t > b.t_cap) {
        if (b.tx) (void)hipFree(b.tx);
        if (b.ty) (void)hipFree(b.ty);
        b.tx = b.ty = nullptr; b.t_cap = 0;
        if (!galloc((void**)&b.tx, t * 4) || !galloc((void**)&b.ty, t * 4)) return false;
        b.t_cap = t;
    }
    if (blocks > b.part_cap) {
        if (b.part) (void)hipFree(b.part);
        b.part = nullptr; b.part_cap = 0;
        if (!galloc((void**)&b.part, blocks * (kParams + 1) * 4)) return false;
        b.part_cap = blocks;
    }
    return true;
}

// Deterministic initialisation: ReLU breakpoints of the first layer spread
// over the input range (so that the ends are not flat), He-scaled second
// layer, small output layer.
static void init_params(std::vector<float>& P, uint64_t seed) {
    P.assign(kParams, 0.0f);
    uint64_t z = seed;
    auto next = [&]() {
        z += 0x9E3779B97F4A7C15ull;
        uint64_t x = z;
        x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
        x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
        x ^= x >> 31;
        return (double)(x >> 11) * 0x1.0p-53;
    };
    auto normal = [&]() {
        double u1 = next(), u2 = next();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    };
    for (int j = 0; j < kH; ++j) {
        // rising and falling units alternately, breakpoints spread over [-0.05, 1.05]
        double w = (j % 2 ? -1.0 : 1.0) * (1.0 + next() * 3.0);
        double brk = -0.05 + 1.1 * (j + 0.5) / kH;
        P[oW1 + j] = (float)w;
        P[oB1 + j] = (float)(-w * brk);
    }
    for (int i = 0; i < kH * kH; ++i) P[oW2 + i] = (float)(normal() * std::sqrt(2.0 / kH));
    for (int i = 0; i < kH; ++i) P[oB2 + i] = 0.01f;
    for (int i = 0; i < kH; ++i) P[oW3 + i] = (float)(normal() * std::sqrt(1.0 / kH));
    P[oB3] = 0.0f;
}

// One gradient evaluation (mean squared error over n points) into grad/loss.
static bool mlp_grad(LearnedBufs& b, const float* P, const float* x, const float* y, uint64_t n, float* grad,
                     float* loss, int adam, float lr, int step, hipStream_t st) {
    uint64_t blocks = (n + kGradThreads - 1) / kGradThreads;
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(kGradThreads), 0, st, P, x, y, n, 1.0f / (float)n, b.part);
    float b1c = 1.0f - std::pow(0.9f, (float)step), b2c = 1.0f - std::pow(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + 63) / 64), dim3(256), 0, st, const_cast<float*>(P), b.m, b.v,
                       b.part, (uint32_t)blocks, adam, lr, b1c, b2c, grad, loss, 1.0f / (float)n);
    return hip_ok(hipGetLastError(), "mlp launch");
}

static float train_lr(int s, int steps) { return s <= (steps * 2) / 3 ? 0.02f : 0.004f; }

// The Adam loop of the learned sort (steps x {gradient, reduce + update}) as
// a graph of kernel nodes: built once per buffer set, then one launch.
static bool train_graph(LearnedBufs& b, uint64_t T, int steps, hipStream_t st) {
    const void* key[7] = {b.P, b.tx, b.ty, b.part, b.m, b.v, b.loss};
    bool same = b.graph_exec && b.graph_T == T && b.graph_steps == steps;
    for (int i = 0; same && i < 7; ++i) same = key[i] == b.graph_key[i];
    if (!same) {
        if (b.graph_exec) (void)hipGraphExecDestroy(b.graph_exec);
        if (b.graph) (void)hipGraphDestroy(b.graph);
        b.graph_exec = nullptr;
        b.graph = nullptr;
        if (!hip_ok(hipGraphCreate(&b.graph, 0), "hipGraphCreate")) return false;
        const float* Pc = b.P;
        float* Pm = b.P;
        const float* x = b.tx;
        const float* y = b.ty;
        uint64_t n = T;
        float inv_n = 1.0f / (float)T;
        float* part = b.part;
        const float* partc = b.part;
        uint32_t nblocks = (uint32_t)((T + kGradThreads - 1) / kGradThreads);
        int adam = 1;
        float* gnull = nullptr;
        hipGraphNode_t prev = nullptr;
        for (int s = 1; s <= steps; ++s) {
            float lr = train_lr(s, steps);
            float b1c = 1.0f - std::pow(0.9f, (float)s), b2c = 1.0f - std::pow(0.999f, (float)s);
            void* ga[] = {(void*)&Pc, (void*)&x, (void*)&y, (void*)&n, (void*)&inv_n, (void*)&part};
            hipKernelNodeParams kp = {};
            kp.func = reinterpret_cast<void*>(&k_mlp_grad);
            kp.gridDim = dim3(nblocks);
            kp.blockDim = dim3(kGradThreads);
            kp.sharedMemBytes = 0;
            kp.kernelParams = ga;
            kp.extra = nullptr;
            hipGraphNode_t gn, rn;
            if (!hip_ok(hipGraphAddKernelNode(&gn, b.graph, prev ? &prev : nullptr, prev ? 1 : 0, &kp), "graph node"))
                return false;
            void* ra[] = {(void*)&Pm, (void*)&b.m, (void*)&b.v, (void*)&partc, (void*)&nblocks, (void*)&adam,
                          (void*)&lr, (void*)&b1c, (void*)&b2c, (void*)&gnull, (void*)&b.loss, (void*)&inv_n};
            kp.func = reinterpret_cast<void*>(&k_mlp_reduce);
            kp.gridDim = dim3((kParams + 1 + 63) / 64);
            kp.blockDim = dim3(256);
            kp.kernelParams = ra;
            if (!hip_ok(hipGraphAddKernelNode(&rn, b.graph, &gn, 1, &kp), "graph node")) return false;
            prev = rn;
        }
        if (!hip_ok(hipGraphInstantiate(&b.graph_exec, b.graph, nullptr, nullptr, 0), "hipGraphInstantiate"))
            return false;
        for (int i = 0; i < 7; ++i) b.graph_key[i] = key[i];
        b.graph_T = T;
        b.graph_steps = steps;
    }
    return hip_ok(hipGraphLaunch(b.graph_exec, st), "hipGraphLaunch");
}

}  // namespace vx

using namespace vx;

extern "C" {

// ---- model kernels exposed for tests and callers ----------------------------
int voxel_gpu_mlp_forward(const float* params_dev, const float* x_dev, size_t n, float* y_dev, void* stream) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    if (n) hipLaunchKernelGGL(k_mlp_forward, dim3(gridn(n, 4096)), dim3(256), 0, c->stream, params_dev, x_dev, (uint64_t)n, y_dev);
    return hip_ok(hipGetLastError(), "mlp forward") ? 0 : -1;
}

int voxel_gpu_mlp_grad(const float* params_dev, const float* x_dev, const float* y_dev, size_t n, float* grad_dev,
                       float* loss_dev, void* stream) {
    set_error("%s", "");
    if (n == 0) { set_error("mlp grad: no points"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    LearnedBufs& b = lbufs(c);
    if (!ensure_model(b) || !ensure_train(b, n)) return -1;
    return mlp_grad(b, params_dev, x_dev, y_dev, n, grad_dev, loss_dev, 0, 0.0f, 1, c->stream) ? 0 : -1;
}

int voxel_gpu_mlp_train(float* params_dev, const float* x_dev, const float* y_dev, size_t n, int steps, float lr,
                        float* loss_out_host, void* stream) {
    set_error("%s", "");
    if (n == 0) { set_error("mlp train: no points"); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    LearnedBufs& b = lbufs(c);
    if (!ensure_model(b) || !ensure_train(b, n)) return -1;
    hipStream_t st = c->stream;
    if (!hip_ok(launch_zero32((uint32_t*)b.m, kParams, st), "zero") || !hip_ok(launch_zero32((uint32_t*)b.v, kParams, st), "zero"))
        return -1;
    for (int s = 1; s <= steps; ++s)
        if (!mlp_grad(b, params_dev, x_dev, y_dev, n, nullptr, b.loss, 1, lr, s, st)) return -1;
    if (loss_out_host) {
        if (!hip_ok(hipMemcpyAsync(b.loss_host, b.loss, 4, hipMemcpyDeviceToHost, st), "read loss") ||
            !hip_ok(hipStreamSynchronize(st), "sync"))
            return -1;
        *loss_out_host = *b.loss_host;
    }
    return 0;
}

int voxel_gpu_mlp_init(float* params_dev, uint64_t seed) {
    set_error("%s", "");
    std::lock_guard<std::mutex> lk(api_mutex());
    if (!ctx_for(nullptr)) return -1;
    std::vector<float> P;
    init_params(P, seed);
    return hip_ok(hipMemcpy(params_dev, P.data(), kParams * 4, hipMemcpyHostToDevice), "copy params") ? 0 : -1;
}

int voxel_gpu_mlp_param_count(void) { return kParams; }

// ---- learned sort -----------------------------------------------------------
// Tuning knobs of the learned sort (0 keeps a default).
int voxel_gpu_learned_tune(int sample_log2, int train_points, int train_steps, int rows_per_bucket_log2) {
    std::lock_guard<std::mutex> lk(api_mutex());
    g_tune.sample_log2 = sample_log2 > 0 ? sample_log2 : 20;
    g_tune.train_points = train_points > 0 ? train_points : 4096;
    g_tune.train_steps = train_steps >= 0 ? (train_steps ? train_steps : 100) : 0;
    g_tune.bucket_log2 = rows_per_bucket_log2 > 0 ? rows_per_bucket_log2 : 11;
    if (g_tune.bucket_log2 > 11) g_tune.bucket_log2 = 11;     // a bucket must fit the LDS sort on average
    return 0;
}

int voxel_gpu_sort_pairs_u64_learned(const uint64_t* kin, const uint32_t* vin_user, uint64_t* kout, uint32_t* vout,
                                     size_t n, int model, voxel_gpu_learned_stats* stats, void* stream) {
    set_error("%s", "");
    voxel_gpu_learned_stats S;
    std::memset(&S, 0, sizeof(S));
    S.model = model;
    if (model != 0 && model != 1) { set_error("learned sort: model must be 0 (MLP) or 1 (equal width)"); return -1; }
    if (n >= (1ull << 30)) { set_error("learned sort: %zu keys exceed the 2^30 limit", n); return -1; }
    std::lock_guard<std::mutex> lk(api_mutex());
    StreamCtx* c = ctx_for(stream);
    if (!c) return -1;
    hipStream_t st = c->stream;
    LearnedBufs& b = lbufs(c);
    if (!ensure_model(b)) return -1;
    if (n > b.cap) {
        for (void* p : {(void*)b.k1, (void*)b.k2, (void*)b.v1, (void*)b.v2})
            if (p) (void)hipFree(p);
        b.k1 = b.k2 = nullptr; b.v1 = b.v2 = nullptr; b.cap = 0;
        if (!galloc((void**)&b.k1, n * 8) || !galloc((void**)&b.k2, n * 8) || !galloc((void**)&b.v1, n * 4) ||
            !galloc((void**)&b.v2, n * 4))
            return -1;
        b.cap = n;
    }
    // payload: the caller's, or the row index (kept in v2 until the partition overwrites it)
    auto payload = [&]() -> const uint32_t* {
        if (vin_user) return vin_user;
        return hip_ok(launch_iota(b.v2, n, st), "iota") ? b.v2 : nullptr;
    };
    if (n < (1u << 16)) {
        // small inputs: the plain radix sort (sampling would dominate)
        S.fell_back = 1;
        const uint32_t* vin = n ? payload() : vin_user;
        if (n && (!vin || sort_pairs_u64(kin, vin, kout, vout, n, c) < 0)) return -1;
        if (!hip_ok(hipStreamSynchronize(st), "sync")) return -1;
        if (stats) *stats = S;
        return 0;
    }
    uint64_t words = radix_work_words(n);
    if (words > b.w.status_cap) {
        if (b.w.status) (void)hipFree(b.w.status);
        b.w.status = nullptr; b.w.status_cap = 0;
        if (!galloc((void**)&b.w.status, words * 4)) return -1;
        b.w.status_cap = words;
    }
    uint64_t SS = 1ull << g_tune.sample_log2;                        // sample size
    if (SS > n / 4) SS = n / 4;
    uint64_t T = std::min<uint64_t>(SS, (uint64_t)g_tune.train_points);   // training points
    if (SS > b.s_cap) {
        for (void* p : {(void*)b.samp, (void*)b.samp2, (void*)b.sidx, (void*)b.sidx2})
            if (p) (void)hipFree(p);
        b.samp = b.samp2 = nullptr; b.sidx = b.sidx2 = nullptr; b.s_cap = 0;
        if (!galloc((void**)&b.samp, SS * 8) || !galloc((void**)&b.samp2, SS * 8) || !galloc((void**)&b.sidx, SS * 4) ||
            !galloc((void**)&b.sidx2, SS * 4))
            return -1;
        b.s_cap = SS;
    }
    if (!ensure_train(b, T)) return -1;
    uint64_t nb = 1;
    while ((nb << g_tune.bucket_log2) < n) nb <<= 1;                 // ~2^bucket_log2 rows per bucket
    const uint32_t nfine = (uint32_t)(nb * 256);
    int cbits = 0;
    while ((1ull << cbits) < nb) ++cbits;
    const int parts = (cbits + 7) / 8;
    if (nb + 1 > b.nb_cap) {
        if (b.start) (void)hipFree(b.start);
        if (b.big) (void)hipFree(b.big);
        b.start = nullptr; b.big = nullptr; b.nb_cap = 0;
        if (!galloc((void**)&b.start, (nb + 1) * 8) || !galloc((void**)&b.big, (nb + 1) * 4)) return -1;
        b.nb_cap = nb + 1;
    }
    S.buckets = nb;

    hipEventRecord(b.ev[0], st);
    // 1. sample, sort it, read 1025 quantiles
    hipLaunchKernelGGL(k_sample, dim3(gridn(SS, 4096)), dim3(256), 0, st, kin, (uint64_t)n, SS, 0x5EEDull, b.samp, b.sidx);
    if (!hip_ok(hipGetLastError(), "sample")) return -1;
    if (sort_pairs_u64(b.samp, b.sidx, b.samp2, b.sidx2, SS, c) < 0) return -1;
    hipLaunchKernelGGL(k_quantiles, dim3(5), dim3(256), 0, st, b.samp2, SS, b.qdev);
    if (!hip_ok(hipGetLastError(), "quantiles") ||
        !hip_ok(hipMemcpyAsync(b.qhost, b.qdev, 1025 * 8, hipMemcpyDeviceToHost, st), "read quantiles") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return -1;
    // 2. grid: linear or logarithmic, whichever the sample CDF is straighter on
    const uint64_t* q = b.qhost;
    const uint64_t kmin = q[0], kmax = q[1024];
    const uint64_t span = kmax - kmin;
    auto logg = [&](uint64_t k) { return (double)TableDigit::grid(k - kmin, 1); };
    double dev_lin = 0, dev_log = 0;
    const double glog = logg(kmax);
    for (int i = 0; i <= 1024; ++i) {
        double target = i / 1024.0;
        if (span) dev_lin = std::max(dev_lin, std::fabs((double)(q[i] - kmin) / (double)span - target));
        if (glog > 0) dev_log = std::max(dev_log, std::fabs(logg(q[i]) / glog - target));
    }
    const int logmode = (model == 0 && dev_log < dev_lin) ? 1 : 0;
    S.log_grid = logmode;
    const uint64_t gmax_i = logmode ? (uint64_t)glog : span;
    int gbits = 0;
    while (gbits < 64 && (gmax_i >> gbits) != 0) ++gbits;
    const int cshift = gbits > 16 ? gbits - 16 : 0;
    const uint32_t ncells = (uint32_t)(gmax_i >> cshift) + 1;
    S.cells = ncells;
    hipEventRecord(b.ev[1], st);
    // 3. train the MLP on (grid position, sample rank)
    if (model == 0) {
        std::vector<float> P;
        init_params(P, 0xC0FFEEull);
        std::memcpy(b.phost, P.data(), kParams * 4);
        if (!hip_ok(hipMemcpyAsync(b.P, b.phost, kParams * 4, hipMemcpyHostToDevice, st), "copy params")) return -1;
        hipLaunchKernelGGL(k_train_data, dim3(gridn(T, 1024)), dim3(256), 0, st, b.samp2, SS, T, kmin, kmax, logmode,
                           (double)gmax_i, b.tx, b.ty);
        if (!hip_ok(launch_zero32((uint32_t*)b.m, kParams, st), "zero") || !hip_ok(launch_zero32((uint32_t*)b.v, kParams, st), "zero"))
            return -1;
        const int steps = g_tune.train_steps;
        if (steps > 0 && !train_graph(b, T, steps, st)) return -1;
        S.train_steps = steps;
    }
    hipEventRecord(b.ev[2], st);
    // 4. CDF at the cell starts -> fine-bucket table.  model 0: MLP root routes
    //    each cell to one of kLeaves linear models fitted to the sample;
    //    model 1: linear in the key (no model).
    hipLaunchKernelGGL(k_root_cells, dim3(gridn(ncells + 1, 256)), dim3(256), 0, st, b.P, ncells, cshift,
                       (double)gmax_i, model, b.R);
    hipLaunchKernelGGL(k_prefix_max<float>, dim3(1), dim3(1024), 0, st, b.R, ncells + 1);
    if (model == 0) {
        hipLaunchKernelGGL(k_leaf_lo, dim3((kLeaves + 256) / 256), dim3(256), 0, st, b.samp2, SS, b.R, ncells, cshift,
                           kmin, kmax, logmode, kLeaves, b.leaf_lo);
        hipLaunchKernelGGL(k_leaf_cells, dim3(gridn(ncells, 256)), dim3(256), 0, st, b.samp2, SS, b.R, ncells, cshift,
                           kmin, kmax, logmode, kLeaves, b.leaf_lo, b.yd);
    } else {
        hipLaunchKernelGGL(k_widen, dim3((ncells + 255) / 256), dim3(256), 0, st, b.R, ncells, b.yd);
    }
    hipLaunchKernelGGL(k_prefix_max<double>, dim3(1), dim3(1024), 0, st, b.yd, ncells);
    hipLaunchKernelGGL(k_table, dim3((ncells + 255) / 256), dim3(256), 0, st, b.yd, ncells, nfine, b.cell);
    if (!hip_ok(hipGetLastError(), "table")) return -1;
    TableDigit dig{b.cell, kmin, kmax, ncells, cshift, logmode, nfine, std::ldexp(1.0f, -cshift)};
    hipEventRecord(b.ev[3], st);
    // 5. stable partition by coarse bucket (LSD over its bytes)
    const uint32_t* vin = payload();
    if (!vin) return -1;
    if (!hip_ok(radix_histogram(kin, n, dig, parts, b.w, st), "model histogram")) return -1;
    if (!hip_ok(hipMemcpyAsync(b.skip_host, b.w.skip, 32, hipMemcpyDeviceToHost, st), "read skip") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return -1;
    const uint64_t* ks = kin;
    const uint32_t* vs = vin;
    int executed = 0;
    for (int p = 0; p < parts; ++p) {
        if (b.skip_host[p]) continue;
        uint64_t* kd = (ks == b.k1) ? b.k2 : b.k1;
        uint32_t* vd = (vs == b.v1) ? b.v2 : b.v1;          // v2 may hold the row index: read once, then free
        if (!hip_ok(radix_pass(ks, vs, kd, vd, n, dig, p, b.w, st), "model pass")) return -1;
        ks = kd;
        vs = vd;
        ++executed;
    }
    S.partition_passes = executed;
    hipEventRecord(b.ev[4], st);
    // 6. bucket bounds and per-bucket sort into kout/vout
    hipLaunchKernelGGL(k_bounds, dim3(gridn(nb + 1, 4096)), dim3(256), 0, st, ks, (uint64_t)n, dig, (uint32_t)nb, b.start);
    if (!hip_ok(launch_zero32(b.big, nb + 1, st), "zero") || !hip_ok(launch_zero32((uint32_t*)b.stat, 4, st), "zero"))
        return -1;
    unsigned sblocks = (unsigned)std::min<uint64_t>(nb, (uint64_t)device_cu_count() * 8);
    hipLaunchKernelGGL((k_bucket_sort<kCap>), dim3(sblocks), dim3(256), 0, st, ks, vs, kout, vout, b.start,
                       (uint32_t)nb, dig, b.big, b.stat);
    if (!hip_ok(hipGetLastError(), "bucket sort")) return -1;
    hipEventRecord(b.ev[5], st);
    // 7. buckets over the LDS capacity: those holding one key value are in
    //    stable order already; the others are gathered (in bucket order),
    //    sorted together by the radix sort -- key ranges of buckets do not
    //    overlap, so one stable sort of the concatenation sorts each bucket --
    //    and scattered back.
    std::vector<uint64_t> starts(nb + 1);
    std::vector<uint32_t> bigf(nb);
    if (!hip_ok(hipMemcpyAsync(starts.data(), b.start, (nb + 1) * 8, hipMemcpyDeviceToHost, st), "read bounds") ||
        !hip_ok(hipMemcpyAsync(bigf.data(), b.big, nb * 4, hipMemcpyDeviceToHost, st), "read flags") ||
        !hip_ok(hipMemcpyAsync(b.stat_host, b.stat, 16, hipMemcpyDeviceToHost, st), "read stats") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return -1;
    S.longest_fine_run = b.stat_host[0];
    S.long_runs = b.stat_host[1];
    std::vector<uint64_t> chunks;            // (row start, row end, bucket first row)
    std::vector<uint32_t> chunk_bucket;
    uint64_t maxb = 0;
    for (uint64_t i = 0; i < nb; ++i) {
        uint64_t m = starts[i + 1] - starts[i];
        maxb = std::max(maxb, m);
        if (!bigf[i]) continue;
        S.big_buckets++;
        S.big_rows += m;
        for (uint64_t r = starts[i]; r < starts[i + 1]; r += kChunk) {
            chunks.push_back(r);
            chunks.push_back(std::min(r + kChunk, starts[i + 1]));
            chunks.push_back(starts[i]);
            chunk_bucket.push_back((uint32_t)i);
        }
    }
    S.max_bucket = maxb;
    std::vector<uint32_t> nonuniform;
    if (!chunk_bucket.empty()) {
        const uint64_t nch = chunk_bucket.size();
        if (!ensure_chunks(b, nch)) return -1;
        std::memcpy(b.chost, chunks.data(), nch * 24);
        bool ok = hip_ok(hipMemcpyAsync(b.cdev, b.chost, nch * 24, hipMemcpyHostToDevice, st), "copy chunks") &&
                  hip_ok(launch_zero32(b.cflag, nch, st), "zero");
        if (ok) {
            hipLaunchKernelGGL(k_big_check, dim3((unsigned)std::min<uint64_t>(nch, 8192)), dim3(256), 0, st, kout,
                               b.cdev, nch, b.cflag);
            ok = hip_ok(hipGetLastError(), "big check") &&
                 hip_ok(hipMemcpyAsync(b.cflag_host, b.cflag, nch * 4, hipMemcpyDeviceToHost, st), "read") &&
                 hip_ok(hipStreamSynchronize(st), "sync");
        }
        if (!ok) return -1;
        for (uint64_t j = 0; j < nch; ++j)
            if (b.cflag_host[j] && (nonuniform.empty() || nonuniform.back() != chunk_bucket[j]))
                nonuniform.push_back(chunk_bucket[j]);
    }
    S.big_unsorted = nonuniform.size();
    uint64_t unsorted_rows = 0;
    for (uint32_t bi : nonuniform) unsorted_rows += starts[bi + 1] - starts[bi];
    if (unsorted_rows > n / 2) {
        // the model did not split the keys: plain radix sort of the input
        S.fell_back = 1;
        const uint32_t* v0 = payload();
        if (!v0 || sort_pairs_u64(kin, v0, kout, vout, n, c) < 0) return -1;
    } else if (unsorted_rows) {
        // move list (source row, destination row, rows) in kChunk pieces
        std::vector<uint64_t> mv;
        uint64_t at = 0;
        for (uint32_t bi : nonuniform)
            for (uint64_t r = starts[bi]; r < starts[bi + 1]; r += kChunk) {
                uint64_t len = std::min<uint64_t>(kChunk, starts[bi + 1] - r);
                mv.push_back(r); mv.push_back(at); mv.push_back(len);
                at += len;
            }
        const uint64_t nmv = mv.size() / 3;
        if (!ensure_chunks(b, nmv)) return -1;
        std::memcpy(b.chost, mv.data(), nmv * 24);
        if (!hip_ok(hipMemcpyAsync(b.cdev, b.chost, nmv * 24, hipMemcpyHostToDevice, st), "copy moves")) return -1;
        const unsigned mb = (unsigned)std::min<uint64_t>(nmv, 65535);
        hipLaunchKernelGGL(k_move, dim3(mb), dim3(256), 0, st, kout, vout, b.k1, b.v1, b.cdev, nmv, 0);
        if (!hip_ok(hipGetLastError(), "gather") || sort_pairs_u64(b.k1, b.v1, b.k2, b.v2, at, c) < 0) return -1;
        hipLaunchKernelGGL(k_move, dim3(mb), dim3(256), 0, st, b.k2, b.v2, kout, vout, b.cdev, nmv, 1);
        if (!hip_ok(hipGetLastError(), "scatter")) return -1;
    }
    hipEventRecord(b.ev[6], st);
    if (!hip_ok(hipEventSynchronize(b.ev[6]), "sync")) return -1;
    float ms[6] = {};
    for (int i = 0; i < 6; ++i) (void)hipEventElapsedTime(&ms[i], b.ev[i], b.ev[i + 1]);
    S.sample_ms = ms[0]; S.train_ms = ms[1]; S.table_ms = ms[2]; S.partition_ms = ms[3]; S.bucket_ms = ms[4];
    S.fixup_ms = ms[5];
    if (model == 0 && S.train_steps > 0 &&
        hip_ok(hipMemcpy(b.loss_host, b.loss, 4, hipMemcpyDeviceToHost), "read loss"))
        S.final_loss = *b.loss_host;
    if (stats) *stats = S;
    return 0;
}

}  // extern "C"
