// vx_learned.hip -- learned-CDF partitioned sort of (u64 key, u32 payload)
// on gfx950.  Same output as the stable LSD radix sort (RadixSort<u64>,
// ops/sort.hpp:83-160); fewer passes over HBM when the key distribution can
// be learned.
//
//  1. sample S keys at SplitMix64-random positions, sort the sample (radix),
//     read 1025 quantiles;
//  2. pick the grid (linear offsets, or float bit patterns = logarithmic)
//     under which the sample CDF is closest to a straight line;
//  3. train a 1->32->32->1 FP32 ReLU MLP on (grid position, logit of rank)
//     pairs of the sorted sample with this file's own forward/backward
//     kernels and Adam (the step loop is one graph launch);
//  4. CDF at every grid cell start: the MLP (sigmoid, rescaled to [0, 1] over
//     the sample range) routes the cell to one of 4096 linear leaf models
//     fitted to the sample (a two-stage recursive model); made non-decreasing
//     and turned into the bucket table of TableDigit (vx_radix.hpp): the
//     bucket id fb(k) is non-decreasing in k by construction;
//  5. stable partition of the pairs by fb with the digit passes of
//     vx_onesweep.hpp -- ceil(log2(buckets) / 8) passes instead of eight;
//     buckets are ~32-256 rows on average;
//  6. bucket starts (marks + suffix minimum), then one wave per group of
//     consecutive buckets (<= 512 rows) ranks every row inside its bucket in
//     LDS (key, then row order) and writes it to its place;
//  7. larger groups are checked for a single key value (then they are in
//     stable order already); the others are gathered, sorted together by the
//     LSD radix sort and scattered back, and when they hold most of the rows
//     the whole input is sorted by the LSD radix sort instead (reported in
//     the statistics).
// The equal-width variant (model = 1) replaces steps 3-4 by a CDF that is
// linear in the key over the sample's [min, max] (the method without a model).
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
int sort_pairs_u64_rowidx(const uint64_t* kin, uint64_t* kout, uint32_t* vout, uint64_t n, StreamCtx* c);

constexpr int kH = 32;                       // hidden width
constexpr int kParams = 32 + 32 + 32 * 32 + 32 + 32 + 1;   // 1153
constexpr int oW1 = 0, oB1 = 32, oW2 = 64, oB2 = 64 + 1024, oW3 = oB2 + 32, oB3 = oW3 + 32;
constexpr int kGradThreads = 64;             // samples per gradient workgroup
constexpr int kGradLanes = 4;                // lanes per sample
constexpr int kGradBlock = kGradThreads * kGradLanes;
constexpr uint32_t kLeaves = 4096;           // linear models under the MLP
constexpr int kMinBucketLog2 = 5;            // average bucket rows are kept at or above 32
constexpr uint32_t kBigCap = 65536;          // groups over a wave's capacity that are fixed up
constexpr int kTrainSteps = 60;              // Adam steps of the root MLP (fewer leave its logit-CDF fit
                                             // of lognormal keys above 0.02 mean squared error)

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
// Four lanes per sample (kGradLanes): lane g of a sample holds hidden units
// 8g..8g+7 of the second layer, so a sample's forward and backward pass is
// a quarter of the serial work; the output and the first layer's gradient
// are summed over the four lanes by quad shuffles (every lane of the quad
// gets the same sum).  Then the 1153 parameter gradients are summed over the
// 64 samples in sample order from LDS.
__global__ __launch_bounds__(kGradBlock) void k_mlp_grad(const float* __restrict__ Pg, const float* __restrict__ x,
                                                         const float* __restrict__ yt, uint64_t n, float inv_n,
                                                         float* __restrict__ part) {
    constexpr int Q = kH / kGradLanes;                   // second-layer units per lane
    __shared__ float P[kParams];
    __shared__ float H1[kGradThreads][kH + 1], DH1[kGradThreads][kH + 1], DH2[kGradThreads][kH + 1];
    __shared__ float H2[kGradThreads][kH + 1];
    __shared__ float DY[kGradThreads], X[kGradThreads], E2[kGradThreads];
    // every global load of the step issued before any is used (one memory
    // round trip instead of one per parameter chunk)
    constexpr int kPLoads = (kParams + kGradBlock - 1) / kGradBlock;
    float pv[kPLoads];
#pragma unroll
    for (int u = 0; u < kPLoads; ++u) {
        const int pi = (int)threadIdx.x + u * kGradBlock;
        pv[u] = pi < kParams ? Pg[pi] : 0.0f;
    }
    const int t = threadIdx.x, sp = t / kGradLanes, g = t % kGradLanes;
    const uint64_t i = (uint64_t)blockIdx.x * kGradThreads + sp;
    const bool live = i < n;
    const float xi = live ? x[i] : 0.0f;
    const float yti = live ? yt[i] : 0.0f;
#pragma unroll
    for (int u = 0; u < kPLoads; ++u) {
        const int pi = (int)threadIdx.x + u * kGradBlock;
        if (pi < kParams) P[pi] = pv[u];
    }
    __syncthreads();
    float h1[kH], h2[Q];
#pragma unroll
    for (int j = 0; j < kH; ++j) h1[j] = live ? fmaxf(P[oW1 + j] * xi + P[oB1 + j], 0.0f) : 0.0f;
    float yp = 0.0f;
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        const int k = g * Q + q;
        float s = P[oB2 + k];
#pragma unroll
        for (int j = 0; j < kH; ++j) s += P[oW2 + k * kH + j] * h1[j];
        h2[q] = live ? fmaxf(s, 0.0f) : 0.0f;
        yp += P[oW3 + k] * h2[q];
    }
    yp += __shfl_xor(yp, 1, kGradLanes);
    yp += __shfl_xor(yp, 2, kGradLanes);
    const float e = live ? (P[oB3] + yp) - yti : 0.0f;
    const float dy = 2.0f * e * inv_n;
    float dh2[Q];
#pragma unroll
    for (int q = 0; q < Q; ++q) dh2[q] = h2[q] > 0.0f ? dy * P[oW3 + g * Q + q] : 0.0f;
#pragma unroll
    for (int j = 0; j < kH; ++j) {
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < Q; ++q) s += dh2[q] * P[oW2 + (g * Q + q) * kH + j];
        s += __shfl_xor(s, 1, kGradLanes);
        s += __shfl_xor(s, 2, kGradLanes);
        if (j / Q == g) DH1[sp][j] = h1[j] > 0.0f ? s : 0.0f;
    }
#pragma unroll
    for (int q = 0; q < Q; ++q) {
        H1[sp][g * Q + q] = h1[g * Q + q];
        H2[sp][g * Q + q] = h2[q];
        DH2[sp][g * Q + q] = dh2[q];
    }
    if (g == 0) { DY[sp] = dy; X[sp] = xi; E2[sp] = e * e; }
    __syncthreads();
    float* out = part + (uint64_t)blockIdx.x * (kParams + 1);
    // dW2[k][j] = sum_s DH2[s][k] * H1[s][j]: four entries per thread (one k,
    // four j), samples in order
    {
        const int idx = t * 4, k = idx / kH, j = idx % kH;
        float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f, s3 = 0.0f;
#pragma unroll 8
        for (int r = 0; r < kGradThreads; ++r) {
            const float d = DH2[r][k];
            s0 += d * H1[r][j]; s1 += d * H1[r][j + 1]; s2 += d * H1[r][j + 2]; s3 += d * H1[r][j + 3];
        }
        out[oW2 + idx] = s0; out[oW2 + idx + 1] = s1; out[oW2 + idx + 2] = s2; out[oW2 + idx + 3] = s3;
    }
    if (t < 4 * kH) {
        // one of W1, B1, B2, W3 (by t / kH) for unit t % kH
        const int u = t % kH, which = t / kH;
        // a = the per-sample factor, b = the unit's value: (DH1, X), (DH1, 1), (DH2, 1), (H2, DY)
        const float (*va)[kH + 1] = which <= 1 ? DH1 : which == 2 ? DH2 : H2;
        float s = 0.0f;
#pragma unroll 8
        for (int r = 0; r < kGradThreads; ++r) {
            const float f = which == 0 ? X[r] : which == 3 ? DY[r] : 1.0f;
            s += va[r][u] * f;
        }
        out[(which == 0 ? oW1 : which == 1 ? oB1 : which == 2 ? oB2 : oW3) + u] = s;
    } else if (t == 4 * kH) {
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
        for (; b + 60 < nblocks; b += 64) {                  // sixteen loads in flight, added in order
            float a[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) a[u] = part[(uint64_t)(b + 4 * u) * (kParams + 1) + p];
#pragma unroll
            for (int u = 0; u < 16; ++u) g += a[u];
        }
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

// s keys at SplitMix64 positions, as order bits key ^ flip
__global__ __launch_bounds__(256) void k_sample(const uint64_t* __restrict__ keys, uint64_t n, uint64_t s,
                                                uint64_t seed, uint64_t flip, uint64_t* __restrict__ out,
                                                uint32_t* __restrict__ idx) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < s; i += stride) {
        uint64_t pos = __umul64hi(sm_mix(seed + (i + 1) * 0x9E3779B97F4A7C15ull), n);
        out[i] = keys[pos] ^ flip;
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

// start[f] = first row of bucket f, for every non-empty bucket (rows are
// in bucket order after the partition, fbs = their bucket ids); other
// entries keep the fill value n.
constexpr int kMarkRows = 8;
__global__ __launch_bounds__(256) void k_mark_starts(const uint32_t* __restrict__ fbs, uint64_t n,
                                                     uint32_t* __restrict__ start) {
    const uint64_t r0 = ((uint64_t)blockIdx.x * blockDim.x + threadIdx.x) * kMarkRows;
    if (r0 >= n) return;
    uint32_t f[kMarkRows];
    if (r0 + kMarkRows <= n) {                       // 32-byte aligned: two 16-byte loads
        const uint4 a = *reinterpret_cast<const uint4*>(fbs + r0);
        const uint4 b = *reinterpret_cast<const uint4*>(fbs + r0 + 4);
        f[0] = a.x; f[1] = a.y; f[2] = a.z; f[3] = a.w; f[4] = b.x; f[5] = b.y; f[6] = b.z; f[7] = b.w;
    } else {
#pragma unroll
        for (int u = 0; u < kMarkRows; ++u) f[u] = r0 + u < n ? fbs[r0 + u] : 0u;
    }
    uint32_t prev = r0 ? fbs[r0 - 1] : 0xFFFFFFFFu;
#pragma unroll
    for (int u = 0; u < kMarkRows; ++u) {
        if (r0 + u < n && f[u] != prev) start[f[u]] = (uint32_t)(r0 + u);
        prev = f[u];
    }
}

// Suffix minimum over start[0..m): empty buckets take the start of the next
// non-empty one.  Three steps over chunks of kMinChunk entries.
constexpr uint32_t kMinChunk = 8192;
__global__ __launch_bounds__(256) void k_min_chunks(const uint32_t* __restrict__ a, uint64_t m, uint32_t* __restrict__ cmin) {
    __shared__ uint32_t sm[256];
    const uint64_t s = (uint64_t)blockIdx.x * kMinChunk;
    uint32_t v = 0xFFFFFFFFu;
    for (uint64_t i = s + threadIdx.x; i < s + kMinChunk && i < m; i += 256) v = min(v, a[i]);
    sm[threadIdx.x] = v;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < (unsigned)o) sm[threadIdx.x] = min(sm[threadIdx.x], sm[threadIdx.x + o]);
        __syncthreads();
    }
    if (threadIdx.x == 0) cmin[blockIdx.x] = sm[0];
}

// one workgroup: carry[c] = min of cmin over chunks after c
__global__ __launch_bounds__(256) void k_min_carry(const uint32_t* __restrict__ cmin, uint32_t nc, uint32_t* __restrict__ carry) {
    __shared__ uint32_t sm[256];
    const uint32_t t = threadIdx.x, per = (nc + 255) / 256;
    const uint32_t lo = min(t * per, nc), hi = min(lo + per, nc);
    uint32_t v = 0xFFFFFFFFu;
    for (uint32_t c = lo; c < hi; ++c) v = min(v, cmin[c]);
    sm[t] = v;
    __syncthreads();
    for (int o = 1; o < 256; o <<= 1) {                 // inclusive suffix min over threads
        uint32_t x = t + o < 256 ? sm[t + o] : 0xFFFFFFFFu;
        __syncthreads();
        sm[t] = min(sm[t], x);
        __syncthreads();
    }
    uint32_t run = t + 1 < 256 ? sm[t + 1] : 0xFFFFFFFFu;
    for (uint32_t c = hi; c-- > lo;) {
        carry[c] = run;
        run = min(run, cmin[c]);
    }
}

__global__ __launch_bounds__(256) void k_min_apply(uint32_t* __restrict__ a, uint64_t m, const uint32_t* __restrict__ carry) {
    __shared__ uint32_t sm[256];
    constexpr uint32_t per = kMinChunk / 256;
    const uint64_t s = (uint64_t)blockIdx.x * kMinChunk + (uint64_t)threadIdx.x * per;
    uint32_t v[per];
    uint32_t run = 0xFFFFFFFFu;
#pragma unroll
    for (int j = per - 1; j >= 0; --j) {
        v[j] = s + j < m ? a[s + j] : 0xFFFFFFFFu;
        run = min(run, v[j]);
    }
    sm[threadIdx.x] = run;
    __syncthreads();
    const uint32_t t = threadIdx.x;
    for (int o = 1; o < 256; o <<= 1) {
        uint32_t x = t + o < 256 ? sm[t + o] : 0xFFFFFFFFu;
        __syncthreads();
        sm[t] = min(sm[t], x);
        __syncthreads();
    }
    run = min(carry[blockIdx.x], t + 1 < 256 ? sm[t + 1] : 0xFFFFFFFFu);
#pragma unroll
    for (int j = per - 1; j >= 0; --j) {
        run = min(run, v[j]);
        if (s + j < m) a[s + j] = run;
    }
}

// rank += 1 when (k, row j) orders before (key, row i): the borrow out of
// the 96-bit subtraction (k hi : k lo : jl) - (key hi : key lo : il), whose
// low words carry the row order (jl < il exactly when j < i; both offset so
// that neither wraps).  Four VALU instructions instead of two 64-bit
// compares, a row compare and their combination.
__device__ __forceinline__ void rank_step(uint32_t& rank, uint32_t jl, uint32_t il, uint64_t k, uint32_t klo,
                                          uint32_t khi) {
    uint32_t t0, t1, t2;
    asm("v_sub_co_u32 %0, vcc, %4, %5\n\t"
        "v_subb_co_u32 %1, vcc, %6, %7, vcc\n\t"
        "v_subb_co_u32 %2, vcc, %8, %9, vcc\n\t"
        "v_addc_co_u32 %3, vcc, 0, %3, vcc"
        : "=&v"(t0), "=&v"(t1), "=&v"(t2), "+v"(rank)
        : "v"(jl), "v"(il), "v"((uint32_t)k), "v"(klo), "v"((uint32_t)(k >> 32)), "v"(khi)
        : "vcc");
}

constexpr int kWaveRows = 512;
constexpr uint32_t kMaxGroup = 63;           // buckets per wave group (their starts sit in LDS)

// Rank sort of rows [s, s + m) (m <= kWaveRows) by one wave: the rows go to
// the wave's LDS slice, every row's rank inside its bucket is counted over
// the bucket's rows (key, then row order: stable) and the row is written to
// its place.  The nbk buckets of the range start at sb[0..nbk) (relative to
// s, sb[0] = 0, sb[nbk] = m; in LDS).
__device__ __forceinline__ void wave_rank_sort(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                               uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                               const uint32_t* sb, uint32_t nbk, uint32_t s, uint32_t m,
                                               uint64_t* kb, uint32_t* vb, int lane, uint32_t& longest) {
    for (uint32_t i = lane; i < m; i += 64) { kb[i] = kin[s + i]; vb[i] = vin[s + i]; }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    uint32_t a = 0;                     // bucket of the lane's previous row (a lane's rows increase)
    for (uint32_t i = lane; i < m; i += 64) {
        // bucket of row i: last q with sb[q] <= i, stepping on from the
        // previous row's (64 rows on: about two buckets); sb[nbk] = m > i
        // ends the walk
        while (sb[a + 1] <= i) ++a;
        const uint32_t lo = sb[a], hi = sb[a + 1];
        const uint64_t key = kb[i];
        const uint32_t klo = (uint32_t)key, khi = (uint32_t)(key >> 32);
        // rank = rows j of the bucket with (kb[j], j) < (key, i): row j + u
        // compares j + 4 with i + 4 - u
        const uint32_t i4 = i + 4;
        uint32_t rank = 0;
        uint32_t j = lo;
        for (; j + 4 <= hi; j += 4) {
            const uint64_t k0 = kb[j], k1 = kb[j + 1], k2 = kb[j + 2], k3 = kb[j + 3];
            const uint32_t j4 = j + 4;
            rank_step(rank, j4, i4, k0, klo, khi);
            rank_step(rank, j4, i4 - 1, k1, klo, khi);
            rank_step(rank, j4, i4 - 2, k2, klo, khi);
            rank_step(rank, j4, i4 - 3, k3, klo, khi);
        }
        for (; j < hi; ++j) rank_step(rank, j + 4, i4, kb[j], klo, khi);
        if (hi - lo > longest) longest = hi - lo;
        kout[s + lo + rank] = key;
        vout[s + lo + rank] = vb[i];
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One wave per group of G consecutive buckets.  A group of at most
// kWaveRows rows is rank-sorted at once; a larger group bucket by bucket,
// and buckets over kWaveRows rows are left in place and listed in biglist
// (the fix-up handles them with parallel kernels).
__global__ __launch_bounds__(256) void k_group_sort(const uint64_t* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                    uint64_t* __restrict__ kout, uint32_t* __restrict__ vout,
                                                    const uint32_t* __restrict__ start, uint32_t nb, uint32_t G,
                                                    uint32_t ngroups, uint32_t* __restrict__ biglist, uint32_t big_cap,
                                                    unsigned long long* __restrict__ stat) {
    __shared__ uint64_t kb[4][kWaveRows];
    __shared__ uint32_t vb[4][kWaveRows];
    __shared__ uint32_t sbs[4][kMaxGroup + 1];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t waves = gridDim.x * 4;
    uint32_t longest = 0, widest = 0;
    uint32_t* sb = sbs[w];
    for (uint32_t g = blockIdx.x * 4 + w; g < ngroups; g += waves) {
        const uint32_t b0 = g * G, b1 = min(b0 + G, nb), nbk = b1 - b0;
        const uint32_t s = start[b0];
        if (lane <= (int)nbk) sb[lane] = start[b0 + lane] - s;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const uint32_t m = sb[nbk];
        if (m == 0) continue;
        if (m <= kWaveRows) {
            widest = m > widest ? m : widest;
            if (m == 1) {
                if (lane == 0) { kout[s] = kin[s]; vout[s] = vin[s]; }
            } else {
                wave_rank_sort(kin, vin, kout, vout, sb, nbk, s, m, kb[w], vb[w], lane, longest);
            }
            continue;
        }
        for (uint32_t q = 0; q < nbk; ++q) {
            const uint32_t bs = s + sb[q], mb = sb[q + 1] - sb[q];
            widest = mb > widest ? mb : widest;
            if (mb == 0) continue;
            if (mb == 1) {
                if (lane == 0) { kout[bs] = kin[bs]; vout[bs] = vin[bs]; }
            } else if (mb <= kWaveRows) {
                const uint32_t one[2] = {0u, mb};
                wave_rank_sort(kin, vin, kout, vout, one, 1, bs, mb, kb[w], vb[w], lane, longest);
            } else if (lane == 0) {
                const uint32_t slot = atomicAdd((uint32_t*)&stat[2], 1u);
                if (slot < big_cap) biglist[slot] = b0 + q;
            }
        }
    }
    for (int msk = 32; msk > 0; msk >>= 1) {
        uint32_t o = __shfl_xor(longest, msk, 64);
        longest = o > longest ? o : longest;
        o = __shfl_xor(widest, msk, 64);
        widest = o > widest ? o : widest;
    }
    if (lane == 0) {
        atomicMax(&stat[0], (unsigned long long)longest);
        atomicMax(&stat[1], (unsigned long long)widest);
    }
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

__global__ __launch_bounds__(256) void k_fill32(uint32_t* __restrict__ p, uint64_t n, uint32_t v) {
    uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// list entries (first bucket, end bucket, -) -> (first row, end row, -)
__global__ __launch_bounds__(256) void k_gather_starts(const uint32_t* __restrict__ start, uint64_t* __restrict__ list,
                                                       uint64_t nl) {
    uint64_t j = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= nl) return;
    list[3 * j] = start[list[3 * j]];
    list[3 * j + 1] = start[list[3 * j + 1]];
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
    uint32_t* f1 = nullptr; uint32_t* f2 = nullptr;      // carried bucket ids
    uint32_t* start = nullptr; uint32_t* cmin = nullptr; uint64_t nb_cap = 0;
    uint32_t* biglist = nullptr; uint32_t* biglist_host = nullptr;
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
    int sample_log2 = 20, train_points = 4096, train_steps = kTrainSteps, bucket_log2 = 8;
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
              galloc((void**)&b.qdev, 1025 * 8) && galloc((void**)&b.biglist, kBigCap * 4) &&
              hip_ok(hipHostMalloc((void**)&b.biglist_host, kBigCap * 4, hipHostMallocDefault), "hipHostMalloc") &&
              galloc((void**)&b.w.hist, 8 * 256 * 8) &&
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
    hipLaunchKernelGGL(k_mlp_grad, dim3((unsigned)blocks), dim3(kGradBlock), 0, st, P, x, y, n, 1.0f / (float)n, b.part);
    float b1c = 1.0f - std::pow(0.9f, (float)step), b2c = 1.0f - std::pow(0.999f, (float)step);
    hipLaunchKernelGGL(k_mlp_reduce, dim3((kParams + 1 + 63) / 64), dim3(256), 0, st, const_cast<float*>(P), b.m, b.v,
                       b.part, (uint32_t)blocks, adam, lr, b1c, b2c, grad, loss, 1.0f / (float)n);
    return hip_ok(hipGetLastError(), "mlp launch");
}

static float train_lr(int s, int steps) { return s <= (steps * 2) / 3 ? 0.03f : 0.006f; }

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
            kp.blockDim = dim3(kGradBlock);
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

// Steps 1-4 of the learned sort: sample n keys (as order bits key ^ flip),
// pick the grid, train the MLP (model 0) and build the table that maps a key
// to one of nfine buckets, non-decreasing in the key.  Records b.ev[1..3]
// (b.ev[0] is the caller's); the table lives in b.cell.
// The plain radix sort skips every byte on which all keys agree; when the
// sample says at most this many bytes differ, it costs less than the
// learned partition (three passes plus the bucket sort, measured as about
// five and a half plain passes at 2^29 keys) and the sort takes it.
constexpr int kPlainPassesMax = 5;

static bool learned_table(LearnedBufs& b, const uint64_t* kin, uint64_t n, uint64_t flip, int model, uint32_t nfine,
                          voxel_gpu_learned_stats& S, TableDigit& dig, StreamCtx* c, bool* prefer_plain = nullptr) {
    hipStream_t st = c->stream;
    uint64_t SS = 1ull << g_tune.sample_log2;                        // sample size
    if (SS > n / 4) SS = n / 4;
    uint64_t T = std::min<uint64_t>(SS, (uint64_t)g_tune.train_points);   // training points
    if (SS > b.s_cap) {
        for (void* p : {(void*)b.samp, (void*)b.samp2, (void*)b.sidx, (void*)b.sidx2})
            if (p) (void)hipFree(p);
        b.samp = b.samp2 = nullptr; b.sidx = b.sidx2 = nullptr; b.s_cap = 0;
        if (!galloc((void**)&b.samp, SS * 8) || !galloc((void**)&b.samp2, SS * 8) || !galloc((void**)&b.sidx, SS * 4) ||
            !galloc((void**)&b.sidx2, SS * 4))
            return false;
        b.s_cap = SS;
    }
    if (!ensure_train(b, T)) return false;
    // 1. sample, sort it, read 1025 quantiles
    hipLaunchKernelGGL(k_sample, dim3(gridn(SS, 4096)), dim3(256), 0, st, kin, (uint64_t)n, SS, 0x5EEDull, flip, b.samp,
                       b.sidx);
    if (!hip_ok(hipGetLastError(), "sample")) return false;
    if (sort_pairs_u64(b.samp, b.sidx, b.samp2, b.sidx2, SS, c) < 0) return false;
    hipLaunchKernelGGL(k_quantiles, dim3(5), dim3(256), 0, st, b.samp2, SS, b.qdev);
    if (!hip_ok(hipGetLastError(), "quantiles") ||
        !hip_ok(hipMemcpyAsync(b.qhost, b.qdev, 1025 * 8, hipMemcpyDeviceToHost, st), "read quantiles") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return false;
    const uint64_t* q = b.qhost;
    if (prefer_plain) {                                // bytes on which the sampled keys differ
        uint64_t diff = 0;
        for (int i = 1; i <= 1024; ++i) diff |= q[i] ^ q[0];
        int bytes = 0;
        for (int j = 0; j < 8; ++j) bytes += ((diff >> (8 * j)) & 0xFF) != 0;
        *prefer_plain = bytes <= kPlainPassesMax;
        if (*prefer_plain) return true;
    }
    // 2. grid: linear or logarithmic, whichever the sample CDF is straighter on
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
    // at most 2^14 cells: the table (8 B a cell) fits in LDS for pass 0
    const int cshift = gbits > 14 ? gbits - 14 : 0;
    const uint32_t ncells = (uint32_t)(gmax_i >> cshift) + 1;
    S.cells = ncells;
    hipEventRecord(b.ev[1], st);
    // 3. train the MLP on (grid position, sample rank)
    if (model == 0) {
        std::vector<float> P;
        init_params(P, 0xC0FFEEull);
        std::memcpy(b.phost, P.data(), kParams * 4);
        if (!hip_ok(hipMemcpyAsync(b.P, b.phost, kParams * 4, hipMemcpyHostToDevice, st), "copy params")) return false;
        hipLaunchKernelGGL(k_train_data, dim3(gridn(T, 1024)), dim3(256), 0, st, b.samp2, SS, T, kmin, kmax, logmode,
                           (double)gmax_i, b.tx, b.ty);
        if (!hip_ok(launch_zero32((uint32_t*)b.m, kParams, st), "zero") || !hip_ok(launch_zero32((uint32_t*)b.v, kParams, st), "zero"))
            return false;
        const int steps = g_tune.train_steps;
        if (steps > 0 && !train_graph(b, T, steps, st)) return false;
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
    if (!hip_ok(hipGetLastError(), "table")) return false;
    dig = TableDigit{b.cell, kmin, kmax, ncells, cshift, logmode, nfine};
    hipEventRecord(b.ev[3], st);
    return true;
}

// The learned CDF of n keys as a partitioner into nfine buckets (for the
// range-partitioned join): the table (at most 2^14 + 1 entries) is copied
// to cell_out, a buffer of 65538 entries.
bool learned_partitioner(const uint64_t* kin, uint64_t n, uint64_t flip, int model, uint32_t nfine, uint2* cell_out,
                         TableDigit* dig, voxel_gpu_learned_stats* stats, StreamCtx* c) {
    voxel_gpu_learned_stats S;
    std::memset(&S, 0, sizeof(S));
    S.model = model;
    S.buckets = nfine;
    LearnedBufs& b = lbufs(c);
    if (!ensure_model(b)) return false;
    hipEventRecord(b.ev[0], c->stream);
    if (!learned_table(b, kin, n, flip, model, nfine, S, *dig, c)) return false;
    if (!hip_ok(hipMemcpyAsync(cell_out, b.cell, (size_t)(dig->ncells + 1) * sizeof(uint2), hipMemcpyDeviceToDevice,
                               c->stream), "copy partition table"))
        return false;
    dig->cell = cell_out;
    if (!hip_ok(hipEventSynchronize(b.ev[3]), "sync")) return false;
    float ms[3] = {};
    for (int i = 0; i < 3; ++i) (void)hipEventElapsedTime(&ms[i], b.ev[i], b.ev[i + 1]);
    S.sample_ms = ms[0]; S.train_ms = ms[1]; S.table_ms = ms[2];
    if (model == 0 && S.train_steps > 0 && hip_ok(hipMemcpy(b.loss_host, b.loss, 4, hipMemcpyDeviceToHost), "read loss"))
        S.final_loss = *b.loss_host;
    if (stats) *stats = S;
    return true;
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
    g_tune.train_steps = train_steps >= 0 ? (train_steps ? train_steps : kTrainSteps) : 0;
    g_tune.bucket_log2 = rows_per_bucket_log2 > 0 ? rows_per_bucket_log2 : 8;
    if (g_tune.bucket_log2 > 8) g_tune.bucket_log2 = 8;       // a bucket group must fit one wave's LDS on average
    if (g_tune.bucket_log2 < kMinBucketLog2) g_tune.bucket_log2 = kMinBucketLog2;
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
        for (void* p : {(void*)b.k1, (void*)b.k2, (void*)b.v1, (void*)b.v2, (void*)b.f1, (void*)b.f2})
            if (p) (void)hipFree(p);
        b.k1 = b.k2 = nullptr; b.v1 = b.v2 = nullptr; b.f1 = b.f2 = nullptr; b.cap = 0;
        if (!galloc((void**)&b.k1, n * 8) || !galloc((void**)&b.k2, n * 8) || !galloc((void**)&b.v1, n * 4) ||
            !galloc((void**)&b.v2, n * 4) || !galloc((void**)&b.f1, n * 4) || !galloc((void**)&b.f2, n * 4))
            return -1;
        b.cap = n;
    }
    // the plain radix sort of the input, with the caller's payload or the row index
    auto plain_sort = [&]() -> bool {
        if (vin_user) return sort_pairs_u64(kin, vin_user, kout, vout, n, c) >= 0;
        return sort_pairs_u64_rowidx(kin, kout, vout, n, c) >= 0;
    };
    if (n < (1u << 16)) {
        // small inputs: the plain radix sort (sampling would dominate)
        S.fell_back = 1;
        if (n && !plain_sort()) return -1;
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
    // Buckets: as few partition passes as keep the average bucket at or
    // below 2^max_log2 rows, then as many buckets as those passes address
    // without going below 2^min_log2 rows on average.
    auto clog2 = [](uint64_t x) { int b = 0; while ((1ull << b) < x) ++b; return b; };
    const int bits_min = std::max(1, clog2((n + (1ull << g_tune.bucket_log2) - 1) >> g_tune.bucket_log2));
    const int parts = (bits_min + 7) / 8;
    const int bits = std::max(bits_min, std::min(8 * parts, clog2(n >> kMinBucketLog2)));
    const uint64_t nb = 1ull << bits;
    const uint32_t nfine = (uint32_t)nb;
    // a wave takes ~256 rows at a time (loads of 4 rows per lane in flight)
    const uint32_t G = (uint32_t)std::min<uint64_t>(kMaxGroup, std::max<uint64_t>(1, 256 / std::max<uint64_t>(1, n >> bits)));
    const uint32_t ngroups = (uint32_t)((nb + G - 1) / G);
    const uint64_t nchunks_min = (nb + 1 + kMinChunk - 1) / kMinChunk;
    if (nb + 1 > b.nb_cap) {
        if (b.start) (void)hipFree(b.start);
        if (b.cmin) (void)hipFree(b.cmin);
        b.start = nullptr; b.cmin = nullptr; b.nb_cap = 0;
        if (!galloc((void**)&b.start, (nb + 1) * 4) || !galloc((void**)&b.cmin, nchunks_min * 8)) return -1;
        b.nb_cap = nb + 1;
    }
    S.buckets = nb;

    hipEventRecord(b.ev[0], st);
    TableDigit dig;
    bool plain = false;
    if (!learned_table(b, kin, n, 0, model, nfine, S, dig, c, &plain)) return -1;
    if (plain) {
        // few distinct key bytes: the plain radix sort's skipped passes win
        S.fell_back = 2;
        if (!plain_sort()) return -1;
        if (!hip_ok(hipStreamSynchronize(st), "sync")) return -1;
        if (stats) *stats = S;
        return 0;
    }
    // 5. stable partition by bucket id (LSD over its bytes); pass 0's
    //    counting kernel computes the id once per row (into the carried-id
    //    buffer that pass 0 does not write), its scatter reads it, and every
    //    pass carries it next to the rows; without a payload, pass 0 writes
    //    the row index
    const uint64_t* ks = kin;
    const uint32_t* vs = vin_user;
    const uint32_t* fs = nullptr;
    int executed = 0;
    for (int p = 0; p < parts; ++p) {
        uint64_t* kd = (ks == b.k1) ? b.k2 : b.k1;
        uint32_t* vd = (vs == b.v1) ? b.v2 : b.v1;
        uint32_t* fd = (fs == b.f1) ? b.f2 : b.f1;
        uint32_t* pre = p == 0 ? (fd == b.f1 ? b.f2 : b.f1) : nullptr;     // free during pass 0
        if (!hip_ok(radix_pass(ks, vs, kd, vd, n, dig, p, b.w, st, fs, fd, p == 0 && !vin_user, pre), "model pass"))
            return -1;
        ks = kd;
        vs = vd;
        fs = fd;
        ++executed;
    }
    S.partition_passes = executed;
    hipEventRecord(b.ev[4], st);
    // 6. bucket starts and per-group sort into kout/vout
    hipLaunchKernelGGL(k_fill32, dim3(gridn(nb + 1, 8192)), dim3(256), 0, st, b.start, nb + 1, (uint32_t)n);
    hipLaunchKernelGGL(k_mark_starts, dim3((unsigned)((n + 256 * kMarkRows - 1) / (256 * kMarkRows))), dim3(256), 0, st,
                       fs, (uint64_t)n, b.start);
    {
        const uint32_t nc = (uint32_t)nchunks_min;
        uint32_t* carry = b.cmin + nc;
        hipLaunchKernelGGL(k_min_chunks, dim3(nc), dim3(256), 0, st, b.start, nb + 1, b.cmin);
        hipLaunchKernelGGL(k_min_carry, dim3(1), dim3(256), 0, st, b.cmin, nc, carry);
        hipLaunchKernelGGL(k_min_apply, dim3(nc), dim3(256), 0, st, b.start, nb + 1, carry);
    }
    if (!hip_ok(launch_zero32((uint32_t*)b.stat, 16, st), "zero")) return -1;
    const unsigned gblocks = (unsigned)std::max<uint64_t>(1, std::min<uint64_t>((ngroups + 3) / 4, (uint64_t)device_cu_count() * 16));
    hipLaunchKernelGGL(k_group_sort, dim3(gblocks), dim3(256), 0, st, ks, vs, kout, vout, b.start, (uint32_t)nb, G,
                       ngroups, b.biglist, kBigCap, b.stat);
    if (!hip_ok(hipGetLastError(), "bucket sort")) return -1;
    hipEventRecord(b.ev[5], st);
    // 7. buckets over a wave's capacity (still in partition order in ks/vs):
    //    those holding one key value are in stable order already and are
    //    copied; the others are gathered (in bucket order), sorted together
    //    by the radix sort -- key ranges of buckets do not overlap, so one
    //    stable sort of the concatenation sorts each bucket -- and scattered
    //    back.
    if (!hip_ok(hipMemcpyAsync(b.stat_host, b.stat, 24, hipMemcpyDeviceToHost, st), "read stats") ||
        !hip_ok(hipStreamSynchronize(st), "sync"))
        return -1;
    S.longest_bucket = b.stat_host[0];
    S.widest_group = b.stat_host[1];
    S.groups = ngroups;
    const uint64_t nbig = b.stat_host[2] & 0xFFFFFFFFull;
    const bool overflow = nbig > kBigCap;
    std::vector<uint64_t> rng;                    // (first row, end row) per big bucket, bucket order
    if (nbig && !overflow) {
        if (!hip_ok(hipMemcpyAsync(b.biglist_host, b.biglist, nbig * 4, hipMemcpyDeviceToHost, st), "read list") ||
            !hip_ok(hipStreamSynchronize(st), "sync"))
            return -1;
        std::vector<uint32_t> bigs(b.biglist_host, b.biglist_host + nbig);
        std::sort(bigs.begin(), bigs.end());
        if (!ensure_chunks(b, nbig)) return -1;
        for (uint64_t j = 0; j < nbig; ++j) {
            b.chost[3 * j] = bigs[j];
            b.chost[3 * j + 1] = (uint64_t)bigs[j] + 1;
            b.chost[3 * j + 2] = 0;
        }
        if (!hip_ok(hipMemcpyAsync(b.cdev, b.chost, nbig * 24, hipMemcpyHostToDevice, st), "copy list")) return -1;
        hipLaunchKernelGGL(k_gather_starts, dim3((unsigned)((nbig + 255) / 256)), dim3(256), 0, st, b.start, b.cdev, nbig);
        if (!hip_ok(hipGetLastError(), "starts") ||
            !hip_ok(hipMemcpyAsync(b.chost, b.cdev, nbig * 24, hipMemcpyDeviceToHost, st), "read starts") ||
            !hip_ok(hipStreamSynchronize(st), "sync"))
            return -1;
        rng.resize(2 * nbig);
        for (uint64_t j = 0; j < nbig; ++j) { rng[2 * j] = b.chost[3 * j]; rng[2 * j + 1] = b.chost[3 * j + 1]; }
    }
    std::vector<uint64_t> chunks;                 // (row start, row end, bucket first row)
    std::vector<uint32_t> chunk_bucket;
    for (uint64_t j = 0; j < rng.size() / 2; ++j) {
        const uint64_t r0 = rng[2 * j], r1 = rng[2 * j + 1];
        S.big_groups++;
        S.big_rows += r1 - r0;
        for (uint64_t r = r0; r < r1; r += kChunk) {
            chunks.push_back(r);
            chunks.push_back(std::min(r + kChunk, r1));
            chunks.push_back(r0);
            chunk_bucket.push_back((uint32_t)j);
        }
    }
    std::vector<char> mixed(rng.size() / 2, 0);
    if (!chunk_bucket.empty()) {
        const uint64_t nch = chunk_bucket.size();
        if (!ensure_chunks(b, nch)) return -1;
        std::memcpy(b.chost, chunks.data(), nch * 24);
        bool ok = hip_ok(hipMemcpyAsync(b.cdev, b.chost, nch * 24, hipMemcpyHostToDevice, st), "copy chunks") &&
                  hip_ok(launch_zero32(b.cflag, nch, st), "zero");
        if (ok) {
            hipLaunchKernelGGL(k_big_check, dim3((unsigned)std::min<uint64_t>(nch, 8192)), dim3(256), 0, st, ks,
                               b.cdev, nch, b.cflag);
            ok = hip_ok(hipGetLastError(), "big check") &&
                 hip_ok(hipMemcpyAsync(b.cflag_host, b.cflag, nch * 4, hipMemcpyDeviceToHost, st), "read") &&
                 hip_ok(hipStreamSynchronize(st), "sync");
        }
        if (!ok) return -1;
        for (uint64_t j = 0; j < nch; ++j)
            if (b.cflag_host[j]) mixed[chunk_bucket[j]] = 1;
    }
    uint64_t unsorted_rows = 0;
    for (size_t j = 0; j < mixed.size(); ++j)
        if (mixed[j]) { S.big_unsorted++; unsorted_rows += rng[2 * j + 1] - rng[2 * j]; }
    if (overflow || unsorted_rows > n / 2) {
        // the model did not split the keys: plain radix sort of the input
        S.fell_back = 1;
        if (!plain_sort()) return -1;
    } else if (!rng.empty()) {
        // moves (source row, destination row, rows) in kChunk pieces:
        // single-key buckets straight to kout, the others packed into the free
        // ping-pong buffers
        uint64_t* ks_w = const_cast<uint64_t*>(ks);
        uint32_t* vs_w = const_cast<uint32_t*>(vs);
        uint64_t* kt = ks_w == b.k1 ? b.k2 : b.k1;
        uint32_t* vt = vs_w == b.v1 ? b.v2 : b.v1;
        std::vector<uint64_t> same, pack;
        uint64_t at = 0;
        for (size_t j = 0; j < mixed.size(); ++j)
            for (uint64_t r = rng[2 * j]; r < rng[2 * j + 1]; r += kChunk) {
                const uint64_t len = std::min<uint64_t>(kChunk, rng[2 * j + 1] - r);
                if (mixed[j]) { pack.push_back(r); pack.push_back(at); pack.push_back(len); at += len; }
                else { same.push_back(r); same.push_back(r); same.push_back(len); }
            }
        const uint64_t ns = same.size() / 3, np_ = pack.size() / 3;
        if (!ensure_chunks(b, ns + np_)) return -1;
        std::memcpy(b.chost, same.data(), ns * 24);
        std::memcpy(b.chost + 3 * ns, pack.data(), np_ * 24);
        if (!hip_ok(hipMemcpyAsync(b.cdev, b.chost, (ns + np_) * 24, hipMemcpyHostToDevice, st), "copy moves")) return -1;
        if (ns) {
            hipLaunchKernelGGL(k_move, dim3((unsigned)std::min<uint64_t>(ns, 65535)), dim3(256), 0, st, ks, vs, kout, vout,
                               b.cdev, ns, 0);
            if (!hip_ok(hipGetLastError(), "copy single-key buckets")) return -1;
        }
        if (np_) {
            const uint64_t* mv = b.cdev + 3 * ns;
            const unsigned mb = (unsigned)std::min<uint64_t>(np_, 65535);
            hipLaunchKernelGGL(k_move, dim3(mb), dim3(256), 0, st, ks, vs, kt, vt, mv, np_, 0);
            if (!hip_ok(hipGetLastError(), "gather") || sort_pairs_u64(kt, vt, ks_w, vs_w, at, c) < 0) return -1;
            hipLaunchKernelGGL(k_move, dim3(mb), dim3(256), 0, st, ks, vs, kout, vout, mv, np_, 1);
            if (!hip_ok(hipGetLastError(), "scatter")) return -1;
        }
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
