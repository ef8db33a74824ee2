// Consumers of the all-pairs RMSD output (gfx950): centroid scores
// (examples/centroids.ipynb, cells 3-4):
//   scores[i] = sum_j exp(-beta d_ij / std(d)),   centroid = argmax_i,
// over the full symmetric n x n matrix d (zero diagonal) held as its
// condensed upper triangle, c(i, j) = i n - i (i + 1) / 2 + j - i - 1, i < j.
// Every sum is FP64 in a fixed order, so the scores do not depend on the
// launch.
//   stats    4,096 workgroups sum contiguous slices of d and d^2 (eight
//            16-byte loads in flight per lane); one workgroup adds them in order and
//            forms the population std of the n^2 entries: mean = 2 S1 / n^2,
//            var = 2 S2 / n^2 - mean^2.
//   tiles    one pass over 256 x 256 tiles of the upper triangle: the lane of
//            column j walks the tile's rows i (each row's 256 entries are
//            consecutive in the condensed layout), so every entry's affinity
//            is computed once and added to its column in the lane's register
//            and to its row by a wave reduction of eight rows at a time; a
//            tile writes one partial per row and one per column.
//   finish   scores[i] = 1 + the row partials of the tiles right of i's row
//            block + the column partials of the tiles above i's column block,
//            in tile order.
#include "common.h"

namespace hr {
namespace {

constexpr int kCBlock = 256;
constexpr int kStatBlocks = 4096;
constexpr int kStatUnroll = 8;
constexpr int kCTile = 256;    // rows and columns of a score tile (one lane per column)
constexpr int kRowBatch = 8;   // rows whose loads are issued together (rows8_sum)

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

// Workgroup b sums the float4 groups [n4 b / kStatBlocks, n4 (b + 1) /
// kStatBlocks) of the matrix; the last workgroup also takes the n_pairs % 4
// entries after them.  The matrix starts 16-byte aligned.
__global__ __launch_bounds__(kCBlock) void centroid_stats_kernel(const float* __restrict__ d,
                                                                 long long n_pairs,
                                                                 double* __restrict__ partial) {
    __shared__ double sh[kCBlock / kWave];
    const long long n4 = n_pairs / 4;
    const long long b0 = n4 * blockIdx.x / kStatBlocks;
    const long long b1 = n4 * (blockIdx.x + 1) / kStatBlocks;
    const float4* __restrict__ d4 = reinterpret_cast<const float4*>(d);
    double s1 = 0.0, s2 = 0.0;
    for (long long k0 = b0 + threadIdx.x; k0 < b1; k0 += (long long)kCBlock * kStatUnroll) {
        float4 x[kStatUnroll];
#pragma unroll
        for (int u = 0; u < kStatUnroll; ++u) {
            const long long k = k0 + (long long)u * kCBlock;
            x[u] = k < b1 ? d4[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        }
#pragma unroll
        for (int u = 0; u < kStatUnroll; ++u) {
            const double v[4] = {x[u].x, x[u].y, x[u].z, x[u].w};
#pragma unroll
            for (int c = 0; c < 4; ++c) {
                s1 += v[c];
                s2 = fma(v[c], v[c], s2);
            }
        }
    }
    if (blockIdx.x == kStatBlocks - 1 && threadIdx.x == 0)
        for (long long k = 4 * n4; k < n_pairs; ++k) {
            const double v = d[k];
            s1 += v;
            s2 = fma(v, v, s2);
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

// exp(-beta d / std) with k = -beta / std formed once per lane (the
// notebook divides each product by std: the two differ by an ulp or so)
__device__ __forceinline__ double affinity(float dij, double k) { return exp(k * (double)dij); }

// Sums over the 64 lanes of a wave of the kRowBatch values every lane holds,
// one per row: three halving exchanges leave each lane one row's sum over 8
// lanes, three more finish it.  Lane l ends with the sum for row
// 4 (l & 1) + 2 ((l >> 1) & 1) + ((l >> 2) & 1); a fixed order, 10 exchanges
// for 8 rows where 8 wave sums would take 48.
__device__ __forceinline__ double rows8_sum(const double (&a)[8], int lane) {
    const bool b0 = lane & 1, b1 = lane & 2, b2 = lane & 4;
    double b[4], c[2];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const double send = b0 ? a[k] : a[k + 4], keep = b0 ? a[k + 4] : a[k];
        b[k] = keep + __shfl_xor(send, 1, kWave);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const double send = b1 ? b[k] : b[k + 2], keep = b1 ? b[k + 2] : b[k];
        c[k] = keep + __shfl_xor(send, 2, kWave);
    }
    const double send = b2 ? c[0] : c[1], keep = b2 ? c[1] : c[0];
    double e = keep + __shfl_xor(send, 4, kWave);
    e += __shfl_xor(e, 8, kWave);
    e += __shfl_xor(e, 16, kWave);
    e += __shfl_xor(e, 32, kWave);
    return e;
}

// Tile (bi, bj), bj >= bi, of the score matrix: rows bi*kCTile.., columns
// bj*kCTile..; row_part[bj * n + i] and col_part[bi * n + j].
__global__ __launch_bounds__(kCBlock) void centroid_tile_kernel(
    const float* __restrict__ d, long long n, double neg_beta, const double* __restrict__ stats,
    double* __restrict__ row_part, double* __restrict__ col_part) {
    const int bi = blockIdx.y, bj = blockIdx.x;
    if (bj < bi) return;  // whole workgroup: the lower triangle of tiles
    __shared__ double rows[kCBlock / kWave][kCTile];
    const double k_aff = neg_beta / stats[2];
    const int w = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const long long i0 = (long long)bi * kCTile;
    const long long j = (long long)bj * kCTile + threadIdx.x;
    const int n_rows = (int)(n - i0 < kCTile ? n - i0 : kCTile);
    double col = 0.0;
    if (n < 2) {  // no entries (and maybe no matrix): one tile of zero partials
        if (threadIdx.x < n_rows) row_part[(long long)bj * n + i0 + threadIdx.x] = 0.0;
        if (j < n) col_part[(long long)bi * n + j] = 0.0;
        return;
    }
    for (int r0 = 0; r0 < n_rows; r0 += kRowBatch) {
        float x[kRowBatch];
        bool on[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            const long long i = i0 + r0 + u;
            on[u] = r0 + u < n_rows && j > i && j < n;
            // every lane loads, an entry of its own or entry 0, so that the
            // batch's loads are in flight together rather than one
            // divergent branch (and one wait) at a time
            x[u] = d[on[u] ? cidx(i, j, n) : 0];
        }
        double a[kRowBatch];
#pragma unroll
        for (int u = 0; u < kRowBatch; ++u) {
            a[u] = on[u] ? affinity(x[u], k_aff) : 0.0;
            col += a[u];
        }
        const double rs = rows8_sum(a, lane);
        const int r = r0 + 4 * (lane & 1) + 2 * ((lane >> 1) & 1) + ((lane >> 2) & 1);
        if (lane < kRowBatch && r < n_rows) rows[w][r] = rs;
    }
    __syncthreads();
    if (threadIdx.x < n_rows) {
        const int t = threadIdx.x;
        double rsum = 0.0;
#pragma unroll
        for (int k = 0; k < kCBlock / kWave; ++k) rsum += rows[k][t];
        row_part[(long long)bj * n + i0 + t] = rsum;
    }
    if (j < n) col_part[(long long)bi * n + j] = col;
}

__global__ __launch_bounds__(kCBlock) void centroid_finish_kernel(
    const double* __restrict__ row_part, const double* __restrict__ col_part, long long n,
    int n_tiles, double* __restrict__ scores) {
    const long long i = (long long)blockIdx.x * kCBlock + threadIdx.x;
    if (i >= n) return;
    const int b = (int)(i / kCTile);
    double s = 1.0;  // the diagonal: exp(0)
    for (int t = b; t < n_tiles; ++t) s += row_part[(long long)t * n + i];
    for (int t = 0; t <= b; ++t) s += col_part[(long long)t * n + i];
    scores[i] = s;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int score_tiles(long long n) { return (int)((n + kCTile - 1) / kCTile); }

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_centroid_workspace_bytes(int64_t n_frames) {
    const long long n = n_frames > 0 ? n_frames : 1;
    return align256(2 * kStatBlocks * sizeof(double)) + align256(4 * sizeof(double)) +
           2 * align256((size_t)score_tiles(n) * n * sizeof(double));
}

int hr_centroid_scores(const float* condensed, int64_t n_frames, double beta, double* scores,
                       double* stats, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0, "invalid n_frames=%lld", (long long)n_frames);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(scores && workspace && (condensed || n_frames == 1), "NULL pointer argument");
    HR_REQUIRE((reinterpret_cast<uintptr_t>(condensed) & 15) == 0,
               "condensed must be 16-byte aligned");
    HR_REQUIRE(score_tiles(n_frames) < 65536, "too many frames (%lld)", (long long)n_frames);
    if (!current_device()) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const long long n = n_frames;
    const long long n_pairs = n * (n - 1) / 2;
    const int tiles = score_tiles(n);
    char* ws = static_cast<char*>(workspace);
    double* partial = reinterpret_cast<double*>(ws);
    ws += align256(2 * kStatBlocks * sizeof(double));
    double* st = reinterpret_cast<double*>(ws);
    ws += align256(4 * sizeof(double));
    double* row_part = reinterpret_cast<double*>(ws);
    ws += align256((size_t)tiles * n * sizeof(double));
    double* col_part = reinterpret_cast<double*>(ws);
    hipLaunchKernelGGL(centroid_stats_kernel, dim3(kStatBlocks), dim3(kCBlock), 0, s, condensed,
                       n_pairs, partial);
    int rc = launch_status("centroid_stats_kernel");
    if (rc) return rc;
    hipLaunchKernelGGL(centroid_stats_final_kernel, dim3(1), dim3(kCBlock), 0, s, partial, n, st);
    if ((rc = launch_status("centroid_stats_final_kernel"))) return rc;
    hipLaunchKernelGGL(centroid_tile_kernel, dim3((unsigned)tiles, (unsigned)tiles), dim3(kCBlock),
                       0, s, condensed, n, -beta, st, row_part, col_part);
    if ((rc = launch_status("centroid_tile_kernel"))) return rc;
    const unsigned blocks = (unsigned)((n + kCBlock - 1) / kCBlock);
    hipLaunchKernelGGL(centroid_finish_kernel, dim3(blocks), dim3(kCBlock), 0, s, row_part,
                       col_part, n, tiles, scores);
    if ((rc = launch_status("centroid_finish_kernel"))) return rc;
    if (stats) {
        HR_CHECK_HIP(hipMemcpyAsync(stats, st, 3 * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    return HR_OK;
}

}  // extern "C"
