// k nearest neighbours from an RMSD matrix (gfx950): for every row of a
// float32 matrix, its k smallest entries in ascending order, ties going to
// the lower column, NaN last -- np.argsort(row, kind="stable")[:k].
//
// One wave per row.  Entries become 64-bit keys (order-preserving float
// bits << 32 | column), so the order is total and ties break by column.  The
// wave takes k rounds: each round every lane finds its smallest key above
// the previous round's and a wave minimum picks the round's entry.  Rows of
// up to 64 x kRegs columns are read once into registers; longer rows are read
// from memory (cached) every round.  k is small in the uses this serves
// (nearest-leader lists, examples/two-pass-clustering.ipynb cell 6), so the
// k passes cost less than a sort.
#include <algorithm>

#include "common.h"

namespace hr {
namespace {

constexpr int kTBlock = 256;
constexpr int kRegs = 16;  // columns per lane held in registers

// Order-preserving map of a float's bits: negative below positive, NaN last
// (one canonical NaN), -0 as +0, so equal values tie and break by column.
__device__ __forceinline__ unsigned long long topk_key(float v, unsigned col) {
    unsigned u = __float_as_uint(v);
    if (v != v) u = 0x7FC00000u;
    if (v == 0.0f) u = 0u;
    u ^= (u >> 31) ? 0xFFFFFFFFu : 0x80000000u;
    return ((unsigned long long)u << 32) | col;
}

__device__ __forceinline__ float topk_value(unsigned long long key) {
    unsigned u = (unsigned)(key >> 32);
    u = (u & 0x80000000u) ? (u ^ 0x80000000u) : ~u;
    return __uint_as_float(u);
}

__device__ __forceinline__ unsigned long long wave_min_u64(unsigned long long v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        const unsigned hi = __shfl_xor((unsigned)(v >> 32), m, kWave);
        const unsigned lo = __shfl_xor((unsigned)v, m, kWave);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        v = o < v ? o : v;
    }
    return v;
}

template <bool kInRegs>
__global__ __launch_bounds__(kTBlock) void topk_rows_kernel(const float* __restrict__ m,
                                                            long long n_rows, long long n_cols,
                                                            long long ld, int k,
                                                            int* __restrict__ index,
                                                            float* __restrict__ value) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (kTBlock / kWave) + threadIdx.x / kWave;
    const long long n_waves = (long long)gridDim.x * (kTBlock / kWave);
    for (long long i = wave; i < n_rows; i += n_waves) {
        const GlobalFloat* row = global_ptr(m + i * ld);
        unsigned long long keys[kInRegs ? kRegs : 1];
        if constexpr (kInRegs) {
#pragma unroll
            for (int r = 0; r < kRegs; ++r) {
                const long long j = lane + (long long)r * kWave;
                keys[r] = j < n_cols ? topk_key(row[j], (unsigned)j) : ~0ull;
            }
        }
        unsigned long long last = 0;
        for (int t = 0; t < k; ++t) {
            unsigned long long best = ~0ull;
            if constexpr (kInRegs) {
#pragma unroll
                for (int r = 0; r < kRegs; ++r) {
                    const unsigned long long c = keys[r];
                    if ((t == 0 || c > last) && c < best) best = c;
                }
            } else {
                for (long long j = lane; j < n_cols; j += kWave) {
                    const unsigned long long c = topk_key(row[j], (unsigned)j);
                    if ((t == 0 || c > last) && c < best) best = c;
                }
            }
            best = wave_min_u64(best);
            last = best;
            if (lane == 0) {
                global_out(index)[i * k + t] = (int)(unsigned)(best & 0xFFFFFFFFull);
                global_out(value)[i * k + t] = topk_value(best);
            }
        }
    }
}

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

int hr_topk_rows(const float* matrix, int64_t n_rows, int64_t n_cols, int64_t ld, int32_t k,
                 int32_t* index, float* value, void* stream) {
    clear_error();
    HR_REQUIRE(n_rows >= 0 && n_cols > 0 && n_cols <= 0x7FFFFFFF && ld >= n_cols,
               "invalid sizes n_rows=%lld n_cols=%lld ld=%lld", (long long)n_rows,
               (long long)n_cols, (long long)ld);
    HR_REQUIRE(k >= 1 && k <= n_cols, "k must be in [1, n_cols] (k=%d, n_cols=%lld)", k,
               (long long)n_cols);
    if (n_rows == 0) return HR_OK;
    HR_REQUIRE(matrix && index && value, "NULL pointer argument");
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    const long long waves = std::min<long long>(n_rows, (long long)d->cu_count * 64);
    const unsigned blocks = (unsigned)((waves + kTBlock / kWave - 1) / (kTBlock / kWave));
    hipStream_t s = (hipStream_t)stream;
    if (n_cols <= (long long)kWave * kRegs)
        hipLaunchKernelGGL(topk_rows_kernel<true>, dim3(blocks), dim3(kTBlock), 0, s, matrix,
                           n_rows, n_cols, ld, k, index, value);
    else
        hipLaunchKernelGGL(topk_rows_kernel<false>, dim3(blocks), dim3(kTBlock), 0, s, matrix,
                           n_rows, n_cols, ld, k, index, value);
    return launch_status("topk_rows_kernel");
}

}  // extern "C"
