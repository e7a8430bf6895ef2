// sched.hip -- bucketed counting sort of cells by their last Newton-iteration
// count (see sched.hpp).
#include "sched.hpp"

#include <algorithm>

namespace ohcs {
namespace {

constexpr int kTB = 256;

// Bucket of a cost: exact below 256, then 32 log-spaced buckets per octave.
__device__ __forceinline__ int cost_bin(int c) {
    if (c < 256) return c < 0 ? 0 : c;
    const int lg = 31 - __clz(c);  // 8..30
    const int b = 256 + (lg - 8) * 32 + ((c >> (lg - 5)) & 31);
    return b < kBins ? b : kBins - 1;
}

__global__ __launch_bounds__(256) void bins_zero(unsigned *bins) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i < 2 * kBins) bins[i] = 0u;
}

__global__ __launch_bounds__(kTB) void bins_count(const int *cost, unsigned *bins, int64_t n) {
    __shared__ unsigned h[kBins];
    for (int i = threadIdx.x; i < kBins; i += kTB) h[i] = 0u;
    __syncthreads();
    for (int64_t i = (int64_t)blockIdx.x * kTB + threadIdx.x; i < n;
         i += (int64_t)gridDim.x * kTB)
        atomicAdd(&h[cost_bin(cost[i])], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < kBins; i += kTB)
        if (h[i]) atomicAdd(&bins[i], h[i]);
}

// off[b] = number of cells in buckets above b (descending order), one block.
__global__ __launch_bounds__(kBins) void bins_scan(unsigned *bins) {
    __shared__ unsigned s[kBins];
    const int t = threadIdx.x, b = kBins - 1 - t;
    const unsigned cnt = bins[b];
    s[t] = cnt;
    __syncthreads();
    for (int o = 1; o < kBins; o <<= 1) {
        const unsigned v = t >= o ? s[t - o] : 0u;
        __syncthreads();
        s[t] += v;
        __syncthreads();
    }
    bins[kBins + b] = s[t] - cnt;
}

// order[off[b] + rank] = cell: each block ranks its 256 cells per bucket in
// LDS and reserves one range per bucket.
__global__ __launch_bounds__(kTB) void bins_scatter(const int *cost, int *order, unsigned *bins,
                                                    int64_t n) {
    __shared__ unsigned h[kBins], base[kBins];
    unsigned *off = bins + kBins;
    for (int64_t start = (int64_t)blockIdx.x * kTB; start < n; start += (int64_t)gridDim.x * kTB) {
        for (int k = threadIdx.x; k < kBins; k += kTB) h[k] = 0u;
        __syncthreads();
        const int64_t i = start + threadIdx.x;
        int b = -1;
        unsigned pos = 0u;
        if (i < n) {
            b = cost_bin(cost[i]);
            pos = atomicAdd(&h[b], 1u);
        }
        __syncthreads();
        for (int k = threadIdx.x; k < kBins; k += kTB)
            if (h[k]) base[k] = atomicAdd(&off[k], h[k]);
        __syncthreads();
        if (b >= 0) order[base[b] + pos] = (int)i;
        __syncthreads();
    }
}

}  // namespace

hipError_t order_by_cost(const int *cost, int *order, unsigned *bins, int64_t n,
                         hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (n > (int64_t)0x7fffffff) return hipErrorInvalidValue;
    const unsigned g = (unsigned)std::min<int64_t>((n + kTB - 1) / kTB, 2048);
    hipLaunchKernelGGL(bins_zero, dim3((2 * kBins + 255) / 256), dim3(256), 0, s, bins);
    hipLaunchKernelGGL(bins_count, dim3(g), dim3(kTB), 0, s, cost, bins, n);
    hipLaunchKernelGGL(bins_scan, dim3(1), dim3(kBins), 0, s, bins);
    hipLaunchKernelGGL(bins_scatter, dim3(g), dim3(kTB), 0, s, cost, order, bins, n);
    return hipGetLastError();
}

}  // namespace ohcs
