// prims.hip — device-wide exclusive scan used by generators and pipelines.
#include "prims.hpp"

namespace bolap {

namespace {
constexpr int kScanThreads = 256;
constexpr int kScanItems = 16;  // per thread
constexpr int64_t kScanTile = int64_t(kScanThreads) * kScanItems;

template <typename T>
__device__ T block_exclusive_scan(T v, T* s_tmp, T* total) {
    const int lane = threadIdx.x & (kWave - 1);
    const int wave = threadIdx.x / kWave;
    // inclusive wave scan
    T x = v;
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        const T y = __shfl_up(x, off, kWave);
        if (lane >= off) x += y;
    }
    if (lane == kWave - 1) s_tmp[wave] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
        T run = 0;
        for (int w = 0; w < kScanThreads / kWave; ++w) {
            const T t = s_tmp[w];
            s_tmp[w] = run;
            run += t;
        }
        *total = run;
    }
    __syncthreads();
    return x - v + s_tmp[wave];
}

__global__ __launch_bounds__(kScanThreads) void scan_tile_sums(const int64_t* __restrict__ d,
                                                               int64_t n,
                                                               int64_t* __restrict__ sums) {
    __shared__ int64_t s_tmp[kScanThreads / kWave];
    __shared__ int64_t s_total;
    const int64_t base = int64_t(blockIdx.x) * kScanTile;
    int64_t acc = 0;
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + int64_t(i) * kScanThreads + threadIdx.x;
        if (idx < n) acc += d[idx];
    }
    (void)block_exclusive_scan<int64_t>(acc, s_tmp, &s_total);
    if (threadIdx.x == 0) sums[blockIdx.x] = s_total;
}

// Single workgroup exclusive scan of the tile sums (any count).
__global__ __launch_bounds__(kScanThreads) void scan_sums(int64_t* __restrict__ sums,
                                                          int64_t m) {
    __shared__ int64_t s_tmp[kScanThreads / kWave];
    __shared__ int64_t s_total;
    int64_t carry = 0;
    for (int64_t base = 0; base < m; base += kScanThreads) {
        const int64_t idx = base + threadIdx.x;
        const int64_t v = idx < m ? sums[idx] : 0;
        const int64_t ex = block_exclusive_scan<int64_t>(v, s_tmp, &s_total);
        if (idx < m) sums[idx] = ex + carry;
        carry += s_total;
        __syncthreads();
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply(int64_t* __restrict__ d, int64_t n,
                                                           const int64_t* __restrict__ sums) {
    __shared__ int64_t s_tmp[kScanThreads / kWave];
    __shared__ int64_t s_total;
    // Each thread scans kScanItems consecutive elements; threads are ordered.
    const int64_t base = int64_t(blockIdx.x) * kScanTile + int64_t(threadIdx.x) * kScanItems;
    int64_t v[kScanItems];
    int64_t acc = 0;
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + i;
        v[i] = idx < n ? d[idx] : 0;
        acc += v[i];
    }
    int64_t run = block_exclusive_scan<int64_t>(acc, s_tmp, &s_total) + sums[blockIdx.x];
#pragma unroll
    for (int i = 0; i < kScanItems; ++i) {
        const int64_t idx = base + i;
        if (idx < n) d[idx] = run;
        run += v[i];
    }
}
}  // namespace

size_t scan_scratch_bytes(int64_t n) {
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    return size_t(tiles > 0 ? tiles : 1) * sizeof(int64_t);
}

int exclusive_scan_i64(bolap_ctx* ctx, int64_t* data, int64_t n, void* tmp) {
    if (n <= 0) return BOLAP_OK;
    const int64_t tiles = (n + kScanTile - 1) / kScanTile;
    auto* sums = static_cast<int64_t*>(tmp);
    hipLaunchKernelGGL(scan_tile_sums, dim3(unsigned(tiles)), dim3(kScanThreads), 0,
                       ctx->stream, data, n, sums);
    hipLaunchKernelGGL(scan_sums, dim3(1), dim3(kScanThreads), 0, ctx->stream, sums, tiles);
    hipLaunchKernelGGL(scan_apply, dim3(unsigned(tiles)), dim3(kScanThreads), 0, ctx->stream,
                       data, n, sums);
    return launch_check("exclusive_scan_i64");
}

}  // namespace bolap
