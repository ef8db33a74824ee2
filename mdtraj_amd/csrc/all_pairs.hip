// All-pairs RMSD matrix (gfx950).
//
// Layout in HBM (workspace): every frame's selection is centred in FP64 and
// rounded once to float32, stored axis-major per frame block:
//   planes[f][axis][k]  f < F, axis < 3, k < K_pad (zero padded)
// and its trace G_f = sum of squares of the ROUNDED centred coordinates
// (FP64), so that M and G describe the same structure.  The correlation
// matrix of a pair is M_ij[r][c] = sum_k P_r[i,k] P_c[j,k]: nine plane
// products, i.e. the 3F x 3F Gram matrix of the stacked planes.  Only tiles on
// or above the diagonal are computed; the QCP solve (FP64) runs in the tile
// epilogue, one lane per pair, and the result is written into the condensed
// upper triangle.
#include "common.h"

namespace hr {
namespace {

constexpr int kPrepBlock = 256;

// Gather + centre + round + trace, one wave per frame.
__global__ __launch_bounds__(kPrepBlock) void pairs_prep_kernel(
    const float* __restrict__ xyz, long long n_frames, int n_atoms_total,
    const int* __restrict__ idx, int n_sel, int k_pad, float* __restrict__ planes,
    double* __restrict__ traces) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * (kPrepBlock / kWave) + threadIdx.x / kWave;
    const long long n_waves = (long long)gridDim.x * (kPrepBlock / kWave);
    for (long long f = wave; f < n_frames; f += n_waves) {
        const float* fr = xyz + f * 3LL * n_atoms_total;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = lane; j < n_sel; j += kWave) {
            const long long a = idx ? idx[j] : j;
            sx += (double)fr[3 * a];
            sy += (double)fr[3 * a + 1];
            sz += (double)fr[3 * a + 2];
        }
        sx = wave_sum(sx) / n_sel;
        sy = wave_sum(sy) / n_sel;
        sz = wave_sum(sz) / n_sel;
        float* px = planes + f * 3LL * k_pad;
        float* py = px + k_pad;
        float* pz = py + k_pad;
        double g = 0.0;
        for (int j = lane; j < k_pad; j += kWave) {
            float x = 0.0f, y = 0.0f, z = 0.0f;
            if (j < n_sel) {
                const long long a = idx ? idx[j] : j;
                x = (float)((double)fr[3 * a] - sx);
                y = (float)((double)fr[3 * a + 1] - sy);
                z = (float)((double)fr[3 * a + 2] - sz);
            }
            px[j] = x;
            py[j] = y;
            pz[j] = z;
            const double dx = x, dy = y, dz = z;
            g = fma(dx, dx, fma(dy, dy, fma(dz, dz, g)));
        }
        g = wave_sum(g);
        if (lane == 0) traces[f] = g;
    }
}

__device__ __forceinline__ long long condensed_index(long long i, long long j, long long n) {
    // i < j; row-major upper triangle without the diagonal
    return i * n - i * (i + 1) / 2 + (j - i - 1);
}

// Tile (ti, tj), ti <= tj, from a linear index over the upper triangle of a
// T x T tile grid, row by row.
__device__ __forceinline__ void triangle_tile(long long lin, long long T, long long& ti,
                                              long long& tj) {
    // row r starts at r*T - r*(r-1)/2
    double t = (double)T + 0.5;
    long long r = (long long)floor(t - sqrt(t * t - 2.0 * (double)lin));
    if (r < 0) r = 0;
    if (r > T - 1) r = T - 1;
    while (r > 0 && r * T - r * (r - 1) / 2 > lin) --r;
    while (r + 1 < T && (r + 1) * T - (r + 1) * r / 2 <= lin) ++r;
    ti = r;
    tj = r + (lin - (r * T - r * (r - 1) / 2));
}

// XCD-aware block order: the hardware deals consecutive workgroups to the 8
// XCDs round-robin; remap so that each XCD walks a contiguous run of tiles
// (neighbouring tiles share a row block, which then stays in that XCD's L2).
__device__ __forceinline__ long long xcd_remap(long long b, long long nb) {
    constexpr int kXcd = 8;
    const long long per = (nb + kXcd - 1) / kXcd;
    const long long x = b % kXcd, q = b / kXcd;
    const long long full = nb % kXcd == 0 ? kXcd : nb % kXcd;  // XCDs with `per` blocks
    long long r;
    if (x < full) r = x * per + q;
    else r = full * per + (x - full) * (per - 1) + q;
    return r;
}

// ---------------------------------------------------------------- FP32 kernel
constexpr int kF32Tile = 64;   // frames per tile side
constexpr int kF32Block = 256; // 16 x 16 threads, 4 x 4 pairs each
constexpr int kF32K = 16;      // atoms per LDS stage

__global__ __launch_bounds__(kF32Block) void pairs_fp32_kernel(
    const float* __restrict__ planes, const double* __restrict__ traces, long long n_frames,
    int n_sel, int k_pad, long long n_tiles, float* __restrict__ out) {
    __shared__ float sA[3][kF32K][kF32Tile + 1];
    __shared__ float sB[3][kF32K][kF32Tile + 1];
    const long long T = (n_frames + kF32Tile - 1) / kF32Tile;
    const long long lin = xcd_remap(blockIdx.x, n_tiles);
    long long ti, tj;
    triangle_tile(lin, T, ti, tj);
    const long long i0 = ti * kF32Tile, j0 = tj * kF32Tile;
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;

    float acc[4][4][9];
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int v = 0; v < 4; ++v)
#pragma unroll
            for (int e = 0; e < 9; ++e) acc[u][v][e] = 0.0f;

    for (int k0 = 0; k0 < k_pad; k0 += kF32K) {
        // stage 64 frames x 16 atoms x 3 axes for both tiles
        for (int t = threadIdx.x; t < 3 * kF32K * kF32Tile; t += kF32Block) {
            const int k = t % kF32K;
            const int fr = (t / kF32K) % kF32Tile;
            const int ax = t / (kF32K * kF32Tile);
            const long long fi = i0 + fr, fj = j0 + fr;
            sA[ax][k][fr] = fi < n_frames ? planes[(fi * 3 + ax) * k_pad + k0 + k] : 0.0f;
            sB[ax][k][fr] = fj < n_frames ? planes[(fj * 3 + ax) * k_pad + k0 + k] : 0.0f;
        }
        __syncthreads();
#pragma unroll 4
        for (int k = 0; k < kF32K; ++k) {
            float a[4][3], b[4][3];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int ax = 0; ax < 3; ++ax) {
                    a[u][ax] = sA[ax][k][ty + 16 * u];
                    b[u][ax] = sB[ax][k][tx + 16 * u];
                }
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int v = 0; v < 4; ++v)
#pragma unroll
                    for (int r = 0; r < 3; ++r)
#pragma unroll
                        for (int c = 0; c < 3; ++c)
                            acc[u][v][3 * r + c] = fmaf(a[u][r], b[v][c], acc[u][v][3 * r + c]);
        }
        __syncthreads();
    }
    const double n = (double)n_sel;
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const long long i = i0 + ty + 16 * u;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const long long j = j0 + tx + 16 * v;
            if (i < n_frames && j < n_frames && i < j) {
                double M[9];
#pragma unroll
                for (int e = 0; e < 9; ++e) M[e] = (double)acc[u][v][e];
                const double msd = qcp_msd(M, traces[i], traces[j], n);
                out[condensed_index(i, j, n_frames)] = (float)sqrt(msd);
            }
        }
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int k_pad_for(int n_sel) { return (n_sel + 63) / 64 * 64; }

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_rmsd_all_pairs_workspace_bytes(int64_t n_frames, int32_t n_sel, int32_t mode) {
    if (n_frames <= 0 || n_sel <= 0) return 256;
    const size_t F = (size_t)n_frames, K = (size_t)k_pad_for(n_sel);
    size_t b = align256(F * sizeof(double));    // traces
    b += align256(F * 3 * K * sizeof(float));   // centred planes
    if (mode != HR_PAIRS_FP32) b += align256(F * 3 * K * 2 * sizeof(uint16_t));  // hi/lo
    return b;
}

int hr_rmsd_all_pairs(const float* xyz, int64_t n_frames, int32_t n_atoms_total,
                      const int32_t* atom_idx, int32_t n_sel, int32_t mode, float* out,
                      void* workspace, size_t workspace_bytes, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    HR_REQUIRE(mode == HR_PAIRS_FP32 || mode == HR_PAIRS_F16X3 || mode == HR_PAIRS_BF16X3,
               "unknown mode %d", mode);
    if (n_frames < 2) return HR_OK;
    HR_REQUIRE(xyz && out && workspace, "NULL pointer argument");
    HR_REQUIRE(workspace_bytes >= hr_rmsd_all_pairs_workspace_bytes(n_frames, n_sel, mode),
               "workspace too small: %zu < %zu", workspace_bytes,
               hr_rmsd_all_pairs_workspace_bytes(n_frames, n_sel, mode));
    if (mode != HR_PAIRS_FP32) {
        set_error("split-precision MFMA mode %d not built yet", mode);
        return HR_ERR_UNSUPPORTED;
    }
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const int K = k_pad_for(n_sel);
    char* ws = static_cast<char*>(workspace);
    double* traces = reinterpret_cast<double*>(ws);
    float* planes = reinterpret_cast<float*>(ws + align256((size_t)n_frames * sizeof(double)));

    long long pw = (long long)d->cu_count * 16;
    if (pw > n_frames) pw = n_frames;
    const int prep_blocks = (int)((pw + 3) / 4);
    hipLaunchKernelGGL(pairs_prep_kernel, dim3(prep_blocks), dim3(kPrepBlock), 0, s, xyz,
                       (long long)n_frames, n_atoms_total, atom_idx, n_sel, K, planes, traces);
    int rc = launch_status("pairs_prep_kernel");
    if (rc) return rc;

    const long long T = (n_frames + kF32Tile - 1) / kF32Tile;
    const long long n_tiles = T * (T + 1) / 2;
    HR_REQUIRE(n_tiles < (1LL << 31), "too many tiles (%lld)", n_tiles);
    hipLaunchKernelGGL(pairs_fp32_kernel, dim3((unsigned)n_tiles), dim3(kF32Block), 0, s, planes,
                       traces, (long long)n_frames, n_sel, K, n_tiles, out);
    return launch_status("pairs_fp32_kernel");
}

}  // extern "C"
