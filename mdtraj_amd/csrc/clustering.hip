// Consumers of the all-pairs RMSD output (gfx950): centroid scores
// (examples/centroids.ipynb, cells 3-4):
//   scores[i] = sum_j exp(-beta d_ij / std(d)),   centroid = argmax_i,
// over the full symmetric n x n matrix d (zero diagonal) held as its
// condensed upper triangle, c(i, j) = i n - i (i + 1) / 2 + j - i - 1, i < j.
// Every sum is FP64 in a fixed order, so the scores do not depend on the
// launch.
//   stats    1,024 workgroups sum contiguous slices of d and d^2; one
//            workgroup adds the slices in order and forms the population
//            std of the n^2 entries: mean = 2 S1 / n^2, var = 2 S2 / n^2 -
//            mean^2.
//   rows     one wave per row i sums its contiguous slice c(i, i+1 .. n-1).
//   columns  the entries above the diagonal in column i, c(j, i) for j < i,
//            are consecutive in i for a fixed j: a workgroup takes 256
//            consecutive columns and a chunk of 2,048 rows j, so each load
//            instruction reads 256 consecutive floats; chunks write their
//            own partial row of the workspace.
//   finish   scores[i] = 1 + row[i] + the column partials in chunk order.
#include "common.h"

namespace hr {
namespace {

constexpr int kCBlock = 256;
constexpr int kStatBlocks = 1024;
constexpr int kColChunk = 2048;

__host__ __device__ __forceinline__ long long cidx(long long i, long long j, long long n) {
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Sum of v over the workgroup, fixed tree; result in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < kCBlock / kWave; ++k) t += sh[k];
    __syncthreads();
    return t;
}

__global__ __launch_bounds__(kCBlock) void centroid_stats_kernel(const float* __restrict__ d,
                                                                 long long n_pairs,
                                                                 double* __restrict__ partial) {
    __shared__ double sh[kCBlock / kWave];
    const long long b0 = n_pairs * blockIdx.x / kStatBlocks;
    const long long b1 = n_pairs * (blockIdx.x + 1) / kStatBlocks;
    double s1 = 0.0, s2 = 0.0;
    for (long long k = b0 + threadIdx.x; k < b1; k += kCBlock) {
        const double x = d[k];
        s1 += x;
        s2 = fma(x, x, s2);
    }
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = s1;
        partial[2 * blockIdx.x + 1] = s2;
    }
}

__global__ __launch_bounds__(kCBlock) void centroid_stats_final_kernel(
    const double* __restrict__ partial, long long n, double* __restrict__ stats) {
    __shared__ double sh[kCBlock / kWave];
    double s1 = 0.0, s2 = 0.0;
    for (int k = threadIdx.x; k < kStatBlocks; k += kCBlock) {
        s1 += partial[2 * k];
        s2 += partial[2 * k + 1];
    }
    s1 = block_sum(s1, sh);
    s2 = block_sum(s2, sh);
    if (threadIdx.x == 0) {
        const double n2 = (double)n * (double)n;
        const double mean = 2.0 * s1 / n2;
        const double var = 2.0 * s2 / n2 - mean * mean;
        stats[0] = s1;
        stats[1] = s2;
        stats[2] = sqrt(var > 0.0 ? var : 0.0);
    }
}

// exp((-beta d) / std), the notebook's order of operations
__device__ __forceinline__ double affinity(float dij, double neg_beta, double sd) {
    return exp((neg_beta * (double)dij) / sd);
}

__global__ __launch_bounds__(kCBlock) void centroid_rows_kernel(const float* __restrict__ d,
                                                                long long n, double neg_beta,
                                                                const double* __restrict__ stats,
                                                                double* __restrict__ rows) {
    const long long i = (long long)blockIdx.x * (kCBlock / kWave) + threadIdx.x / kWave;
    const int lane = threadIdx.x & (kWave - 1);
    if (i >= n) return;  // whole wave
    const double sd = stats[2];
    const float* row = d + (i + 1 < n ? cidx(i, i + 1, n) : 0);
    const long long len = n - 1 - i;
    double s = 0.0;
    for (long long k = lane; k < len; k += kWave) s += affinity(row[k], neg_beta, sd);
    s = wave_sum(s);
    if (lane == 0) rows[i] = s;
}

__global__ __launch_bounds__(kCBlock) void centroid_cols_kernel(const float* __restrict__ d,
                                                                long long n, double neg_beta,
                                                                const double* __restrict__ stats,
                                                                double* __restrict__ cols) {
    const long long i = (long long)blockIdx.x * kCBlock + threadIdx.x;
    const long long j0 = (long long)blockIdx.y * kColChunk;
    if (i >= n) return;
    const double sd = stats[2];
    long long j1 = j0 + kColChunk;
    if (j1 > i) j1 = i;
    double s = 0.0;
    for (long long j = j0; j < j1; ++j) s += affinity(d[cidx(j, i, n)], neg_beta, sd);
    cols[blockIdx.y * n + i] = s;
}

__global__ __launch_bounds__(kCBlock) void centroid_finish_kernel(
    const double* __restrict__ rows, const double* __restrict__ cols, long long n,
    int n_chunks, double* __restrict__ scores) {
    const long long i = (long long)blockIdx.x * kCBlock + threadIdx.x;
    if (i >= n) return;
    double s = 1.0 + rows[i];  // the diagonal: exp(0)
    for (int c = 0; c < n_chunks; ++c) s += cols[(long long)c * n + i];
    scores[i] = s;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int col_chunks(long long n) { return (int)((n + kColChunk - 1) / kColChunk); }

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_centroid_workspace_bytes(int64_t n_frames) {
    const long long n = n_frames > 0 ? n_frames : 1;
    return align256(2 * kStatBlocks * sizeof(double)) + align256(4 * sizeof(double)) +
           align256((size_t)n * sizeof(double)) +
           align256((size_t)col_chunks(n) * n * sizeof(double));
}

int hr_centroid_scores(const float* condensed, int64_t n_frames, double beta, double* scores,
                       double* stats, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0, "invalid n_frames=%lld", (long long)n_frames);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(scores && workspace && (condensed || n_frames == 1), "NULL pointer argument");
    HR_REQUIRE(col_chunks(n_frames) < 65536, "too many frames (%lld)", (long long)n_frames);
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const long long n = n_frames;
    const long long n_pairs = n * (n - 1) / 2;
    char* ws = static_cast<char*>(workspace);
    double* partial = reinterpret_cast<double*>(ws);
    ws += align256(2 * kStatBlocks * sizeof(double));
    double* st = reinterpret_cast<double*>(ws);
    ws += align256(4 * sizeof(double));
    double* rows = reinterpret_cast<double*>(ws);
    ws += align256((size_t)n * sizeof(double));
    double* cols = reinterpret_cast<double*>(ws);
    const int chunks = col_chunks(n);
    hipLaunchKernelGGL(centroid_stats_kernel, dim3(kStatBlocks), dim3(kCBlock), 0, s, condensed,
                       n_pairs, partial);
    int rc = launch_status("centroid_stats_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(centroid_stats_final_kernel, dim3(1), dim3(kCBlock), 0, s, partial, n, st);
    if ((rc = launch_status("centroid_stats_final_kernel"))) return rc;
    const long long row_blocks = (n + kCBlock / kWave - 1) / (kCBlock / kWave);
    hipLaunchKernelGGL(centroid_rows_kernel, dim3((unsigned)row_blocks), dim3(kCBlock), 0, s,
                       condensed, n, -beta, st, rows);
    if ((rc = launch_status("centroid_rows_kernel"))) return rc;
    const unsigned col_blocks = (unsigned)((n + kCBlock - 1) / kCBlock);
    hipLaunchKernelGGL(centroid_cols_kernel, dim3(col_blocks, (unsigned)chunks), dim3(kCBlock), 0,
                       s, condensed, n, -beta, st, cols);
    if ((rc = launch_status("centroid_cols_kernel"))) return rc;
    hipLaunchKernelGGL(centroid_finish_kernel, dim3(col_blocks), dim3(kCBlock), 0, s, rows, cols,
                       n, chunks, scores);
    if ((rc = launch_status("centroid_finish_kernel"))) return rc;
    if (stats) {
        HR_CHECK_HIP(hipMemcpyAsync(stats, st, 3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return HR_OK;
}

}  // extern "C"
