// One-vs-all RMSD, centering, no-superpose RMSD and superposition kernels
// (gfx950).  Layout in HBM: coordinates stay as the caller holds them
// (atom-major (F, N, 3) float32 or axis-major (F, 3, rowstride)); nothing is
// re-laid-out.  A row of 4 - 64 lanes of a wavefront owns one frame at a time
// and sweeps its selected atoms with lane-strided loads; per-frame sums are
// accumulated in FP64 in registers and reduced across the row; the QCP
// solve then runs one lane per frame (lane k finishes the k-th frame of the
// wave's current group of 64).  The reference selection and the atom indices
// are staged once per workgroup in LDS.
#include <algorithm>

#include "frames.h"

namespace hr {
namespace {

// Per-frame kernel.  One workgroup per CU holds all of the CU's waves
// (kMinWaves per SIMD).  Each wave walks its items (frame set, atom block)
// through a ring of kRing register buffers, so the loads of the next
// kRing - 1 items are in flight while one is accumulated in FP64.  A frame
// set is kWave / kRL frames, one per row of kRL lanes.  At the end of a set
// each row's sums are reduced across its lanes (DPP within 16 lanes,
// shuffles across) and parked in LDS; every 64 frames the QCP solve runs one
// lane per frame.  The workgroup owns 1/G of the frame sets: each wave takes
// a static share of the first static_pct percent of them, and the rest goes
// out in chunks of `chunk` frames by an LDS ticket, so that the waves the
// SIMD issues less often take fewer frames (process_cu).
//   kAxis   target is axis-major (F, 3, rowstride) instead of atom-major
//   kGather atom-major target with an index list
//   kStage  what the workgroup stages in LDS: kStageAll the reference
//           selection and the indices; kStageIdx the indices only (a
//           gathered selection whose reference does not fit; the reference
//           then rides the load ring); kStageNone nothing
template <Kind kKind, bool kAxis, bool kGather, int kStage, int kMinWaves, int kRL>
__global__ __launch_bounds__(CuGroup<kMinWaves>::kThreads, kMinWaves) void frames_kernel(
    const float* __restrict__ xyz, long long n_frames, long long frame_stride, int rowstride,
    const int* __restrict__ idx, int n_sel, const float* __restrict__ ref_sel,
    const double* __restrict__ ref_stats, long long self_frame, float* __restrict__ out,
    float* __restrict__ rot_out, double* __restrict__ mu_out, int static_pct, long long chunk) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    constexpr int kWaves = CuGroup<kMinWaves>::kWaves;
    constexpr int kThreads = CuGroup<kMinWaves>::kThreads;
    extern __shared__ double smem_d[];
    __shared__ unsigned s_ticket;
    const int lane = threadIdx.x & (kWave - 1);
    const int wv = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    double* s_res = smem_d + wv * kWave * kNS;
    float* smem = reinterpret_cast<float*>(smem_d + kWaves * kWave * kNS);
    const float* rx = ref_sel;
    const float* ry = ref_sel + n_sel;
    const float* rz = ref_sel + 2 * n_sel;
    const int* ix = idx;
    // ref_stats == nullptr (Kind::Rmsd, kStageAll only): ref_sel is the
    // reference selection itself (n_sel atoms, xyz interleaved), and every
    // workgroup prepares it, with the prep kernel's arithmetic, into its LDS
    __shared__ double s_stats[4];
    if (threadIdx.x == 0) s_ticket = 0u;
    if constexpr (kStage != kStageNone) {
        // LDS after the result area: [reference planes (kStageAll)] [indices]
        float* s_rx = smem;
        float* s_ry = smem + n_sel;
        float* s_rz = smem + 2 * n_sel;
        int* s_idx = reinterpret_cast<int*>(kStage == kStageAll ? smem + 3 * n_sel : smem);
        if constexpr (kStage == kStageAll) {
            if (kKind == Kind::Rmsd && ref_stats == nullptr)  // uniform over the grid
                ref_prep_body(ref_sel, nullptr, n_sel, 0, s_rx, s_stats);
            else
                stage_words<kThreads>(s_rx, ref_sel, 3 * n_sel);
        }
        if constexpr (kGather) stage_words<kThreads>(s_idx, idx, n_sel);
        if constexpr (kStage == kStageAll) {
            rx = s_rx;
            ry = s_ry;
            rz = s_rz;
        }
        if constexpr (kGather) ix = s_idx;
    }
    __syncthreads();
    double rs[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (kKind != Kind::NoSuperpose) {
        const double* st = ref_stats;
        if constexpr (kStage == kStageAll && kKind == Kind::Rmsd) {
            if (ref_stats == nullptr) st = s_stats;
        }
        rs[0] = st[0];
        rs[1] = st[1];
        rs[2] = st[2];
        rs[3] = st[3];
    }
    // whole frame sets (a set fills the wave's kWave / kRL rows): the
    // workgroup's sets b0 .. b1, the first n_static of them split evenly
    // over its waves, the rest ticketed
    constexpr int kRows = kWave / kRL;
    const long long n_sets = (n_frames + kRows - 1) / kRows;
    const long long g = blockIdx.x, n_groups = gridDim.x;
    const long long b0 = g * n_sets / n_groups, b1 = (g + 1) * n_sets / n_groups;
    const long long n_static = (b1 - b0) * static_pct / 100;
    const long long s0 = b0 + wv * n_static / kWaves, s1 = b0 + (wv + 1) * n_static / kWaves;
    process_cu<kKind, kAxis, kGather, kRL, kStage != kStageAll>(
        xyz, min(n_frames, kRows * s0), min(n_frames, kRows * s1),
        min(n_frames, kRows * (b0 + n_static)), min(n_frames, kRows * b1), chunk, &s_ticket,
        frame_stride, rowstride, ix, n_sel, rx, ry, rz, rs, self_frame, out, rot_out, mu_out, s_res,
        lane);
}

// In-place centering and traces, one wave per frame (two passes over the
// frame; the second usually hits in L2).  Rounding as the reference:
// double sums, float mean, float subtraction, float squares summed in double.
__global__ __launch_bounds__(kBlock) void center_kernel(float* __restrict__ xyz,
                                                        float* __restrict__ traces,
                                                        long long n_frames, int n_atoms) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long wave = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave;
    const long long n_waves = (long long)gridDim.x * kWavesPerBlock;
    for (long long f = wave; f < n_frames; f += n_waves) {
        float* fr = xyz + f * 3LL * n_atoms;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int a = lane; a < n_atoms; a += kWave) {
            sx += (double)fr[3 * a];
            sy += (double)fr[3 * a + 1];
            sz += (double)fr[3 * a + 2];
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sz = wave_sum(sz);
        const float mx = (float)(sx / n_atoms), my = (float)(sy / n_atoms),
                    mz = (float)(sz / n_atoms);
        double tr = 0.0;
        for (int a = lane; a < n_atoms; a += kWave) {
            const float x = fr[3 * a] - mx, y = fr[3 * a + 1] - my, z = fr[3 * a + 2] - mz;
            fr[3 * a] = x;
            fr[3 * a + 1] = y;
            fr[3 * a + 2] = z;
            tr += (double)(x * x);
            tr += (double)(y * y);
            tr += (double)(z * z);
        }
        tr = wave_sum(tr);
        if (lane == 0 && traces) traces[f] = (float)tr;
    }
}

// x' = float(float(x - mu_f) R + mu_ref) over every atom of every frame,
// rounding as Trajectory.superpose does (float64 offsets applied to float32
// coordinates, float32 rotation, core/trajectory.py:1135-1171).
__global__ __launch_bounds__(kBlock) void apply_kernel(float* __restrict__ xyz, long long n_frames,
                                                       int n_atoms, const float* __restrict__ rot,
                                                       const double* __restrict__ mu,
                                                       const double* __restrict__ ref_stats,
                                                       double n_ref, int translate) {
    const double rmx = translate ? ref_stats[0] / n_ref : 0.0;
    const double rmy = translate ? ref_stats[1] / n_ref : 0.0;
    const double rmz = translate ? ref_stats[2] / n_ref : 0.0;
    for (long long f = blockIdx.x; f < n_frames; f += gridDim.x) {
        float R[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = rot[9 * f + i];
        const double mx = translate ? mu[3 * f] : 0.0;
        const double my = translate ? mu[3 * f + 1] : 0.0;
        const double mz = translate ? mu[3 * f + 2] : 0.0;
        float* fr = xyz + f * 3LL * n_atoms;
        for (int a = threadIdx.x; a < n_atoms; a += kBlock) {
            const float x = translate ? (float)((double)fr[3 * a] - mx) : fr[3 * a];
            const float y = translate ? (float)((double)fr[3 * a + 1] - my) : fr[3 * a + 1];
            const float z = translate ? (float)((double)fr[3 * a + 2] - mz) : fr[3 * a + 2];
            const float tx = x * R[0] + y * R[3] + z * R[6];
            const float ty = x * R[1] + y * R[4] + z * R[7];
            const float tz = x * R[2] + y * R[5] + z * R[8];
            fr[3 * a] = translate ? (float)((double)tx + rmx) : tx;
            fr[3 * a + 1] = translate ? (float)((double)ty + rmy) : ty;
            fr[3 * a + 2] = translate ? (float)((double)tz + rmz) : tz;
        }
    }
}

struct Launch {
    size_t lds;
    int stage;  // kStageAll, kStageIdx or kStageNone
};

// Waves per SIMD of the per-frame kernel (HR_TUNE_FRAMES_WAVES_PER_SIMD).
int frames_waves_per_simd() { return tuning(HR_TUNE_FRAMES_WAVES_PER_SIMD) == 4 ? 4 : 3; }

// LDS plan of the one workgroup per CU: the result area of its waves
// always; the reference selection and the indices when they fit in the CU's
// LDS together with it, else the indices alone (gathered selections), else
// nothing.  At 3 waves/SIMD (80 KB of results) that stages selections of up
// to ~6,900 contiguous or ~5,200 gathered atoms.
template <Kind kKind>
int plan(int n_sel, bool gather, Launch* l) {
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    const size_t res = result_lds_bytes<kKind>(4 * frames_waves_per_simd());
    const size_t ref = (size_t)n_sel * 3 * sizeof(float);
    const size_t ids = gather ? (size_t)n_sel * sizeof(int) : 0;
    const size_t cap = std::min(d->lds_per_block, d->lds_per_cu) - 256;  // 256: static LDS
    if (res + ref + ids <= cap) {
        l->stage = kStageAll;
        l->lds = res + ref + ids;
    } else if (gather && res + ids <= cap) {
        l->stage = kStageIdx;
        l->lds = res + ids;
    } else {
        l->stage = kStageNone;
        l->lds = res;
    }
    return HR_OK;
}

// One workgroup per CU (occupancy API: the workgroup holds all of a CU's
// waves), capped by the work: no more waves than frame sets.
template <typename Kernel>
int frames_grid(Kernel kernel, int threads, size_t lds, long long n_frames, int rows,
                int* blocks) {
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    int per_cu = 0;
    HR_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, reinterpret_cast<const void*>(kernel), threads, lds));
    if (per_cu < 1) per_cu = 1;
    long long b = (long long)per_cu * d->cu_count;
    const long long n_sets = (n_frames + rows - 1) / rows;
    const int waves = threads / kWave;
    const long long by_work = (n_sets + waves - 1) / waves;
    if (b > by_work) b = by_work;
    *blocks = (int)(b < 1 ? 1 : b);
    return HR_OK;
}

template <Kind kKind, bool kAxis, bool kGather, int kStage, int kMinWaves, int kRL>
int launch_frames_v(const Launch& l, hipStream_t s, const float* xyz, long long n_frames,
                    long long frame_stride, int rowstride, const int* idx, int n_sel,
                    const float* ref_sel, const double* ref_stats, long long self_frame,
                    float* out, float* rot_out, double* mu_out) {
    int blocks = 0, rc;
    auto k = frames_kernel<kKind, kAxis, kGather, kStage, kMinWaves, kRL>;
    constexpr int kThreads = CuGroup<kMinWaves>::kThreads;
    constexpr int kRows = kWave / kRL;
    static bool lds_attr_set = false;  // dynamic LDS beyond 64 KB needs the attribute
    if (!lds_attr_set) {
        const DeviceInfo* d = current_device();
        if (!d) return HR_ERR_HIP;
        HR_CHECK_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k),
                                         hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)(std::min(d->lds_per_block, d->lds_per_cu) - 256)));
        lds_attr_set = true;
    }
    if ((rc = frames_grid(k, kThreads, l.lds, n_frames, kRows, &blocks))) return rc;
    // With at most one frame set per wave there is nothing to balance: every
    // set is assigned statically, so that no wave takes a chunk of several
    // sets in series while others idle (1,000 frames x 22 atoms: 18 us).
    const bool one_each = (n_frames + kRows - 1) / kRows <= (long long)blocks * (kThreads / kWave);
    const int pct = tuning(HR_TUNE_FRAMES_STATIC_PCT) ? tuning(HR_TUNE_FRAMES_STATIC_PCT)
                    : one_each                        ? 100
                                                      : kDefaultStaticPct;
    const long long kib = tuning(HR_TUNE_FRAMES_CHUNK_KIB) ? tuning(HR_TUNE_FRAMES_CHUNK_KIB)
                                                          : kDefaultChunkKiB;
    const long long frame_bytes = 12LL * (n_sel > 0 ? n_sel : 1);
    long long chunk = (kib * 1024 + frame_bytes - 1) / frame_bytes;
    chunk = std::max<long long>(kRows, std::min<long long>((chunk + kRows - 1) / kRows * kRows,
                                                           kWave));
    hipLaunchKernelGGL(k, dim3(blocks), dim3(kThreads), l.lds, s, xyz, n_frames, frame_stride,
                       rowstride, idx, n_sel, ref_sel, ref_stats, self_frame, out, rot_out,
                       mu_out, pct, chunk);
    return launch_status("frames_kernel");
}

// auto_row_lanes (frames.h) picks the row width; HR_TUNE_FRAMES_ROW_LANES
// overrides it.
// Register budget of the per-frame kernel: 3 waves per SIMD (no spills) or 4
// (more loads in flight, a few spilled registers); HR_TUNE_FRAMES_WAVES_PER_SIMD.
template <Kind kKind, bool kAxis, bool kGather>
int launch_frames(const Launch& l, hipStream_t s, const float* xyz, long long n_frames,
                  long long frame_stride, int rowstride, const int* idx, int n_sel,
                  const float* ref_sel, const double* ref_stats, long long self_frame, float* out,
                  float* rot_out, double* mu_out) {
    const bool four = frames_waves_per_simd() == 4;
    const int rl = tuning(HR_TUNE_FRAMES_ROW_LANES) ? tuning(HR_TUNE_FRAMES_ROW_LANES)
                                                   : auto_row_lanes(n_sel, kGather);
#define HR_FRAMES_ARGS l, s, xyz, n_frames, frame_stride, rowstride, idx, n_sel, ref_sel, \
                       ref_stats, self_frame, out, rot_out, mu_out
#define HR_FRAMES_RL(LDS, W)                                                               \
    (rl == 64   ? launch_frames_v<kKind, kAxis, kGather, LDS, W, 64>(HR_FRAMES_ARGS)       \
     : rl == 32 ? launch_frames_v<kKind, kAxis, kGather, LDS, W, 32>(HR_FRAMES_ARGS)       \
     : rl == 16 ? launch_frames_v<kKind, kAxis, kGather, LDS, W, 16>(HR_FRAMES_ARGS)       \
     : rl == 8  ? launch_frames_v<kKind, kAxis, kGather, LDS, W, 8>(HR_FRAMES_ARGS)        \
                : launch_frames_v<kKind, kAxis, kGather, LDS, W, 4>(HR_FRAMES_ARGS))
    if (l.stage == kStageAll)
        return four ? HR_FRAMES_RL(kStageAll, 4) : HR_FRAMES_RL(kStageAll, 3);
    if constexpr (kGather) {
        if (l.stage == kStageIdx)
            return four ? HR_FRAMES_RL(kStageIdx, 4) : HR_FRAMES_RL(kStageIdx, 3);
    }
    return four ? HR_FRAMES_RL(kStageNone, 4) : HR_FRAMES_RL(kStageNone, 3);
#undef HR_FRAMES_RL
#undef HR_FRAMES_ARGS
}

}  // namespace

// One launch for a small call: the per-frame kernel prepares the reference
// itself (see frames_kernel).  Returns HR_OK with *done false, launching
// nothing, when the selection does not fit the kernel's LDS whole.
int one_vs_all_fused_ref(hipStream_t s, const float* frames, long long n_frames, int n_sel,
                         const float* ref, long long self_frame, float* out, bool* done) {
    *done = false;
    Launch l;
    int rc = plan<Kind::Rmsd>(n_sel, false, &l);
    if (rc || l.stage != kStageAll) return rc;
    *done = true;
    return launch_frames<Kind::Rmsd, false, false>(l, s, frames, n_frames, 3LL * n_sel, 0,
                                                   nullptr, n_sel, ref, nullptr, self_frame, out,
                                                   nullptr, nullptr);
}

namespace {

int prep_ref(hipStream_t s, const float* ref, const int* ref_idx, int n_sel, int axis_rowstride,
             void* workspace, float** ref_sel, double** stats) {
    char* ws = static_cast<char*>(workspace);
    *stats = reinterpret_cast<double*>(ws);
    *ref_sel = reinterpret_cast<float*>(ws + kRefOffset);
    hipLaunchKernelGGL(ref_prep_kernel, dim3(1), dim3(kBlock), 0, s, ref, ref_idx, n_sel,
                       axis_rowstride, *ref_sel, *stats);
    return launch_status("ref_prep_kernel");
}


// Align/displace: RMS displacement of every frame's displacement set,
// rotated by its fitting rotation R_f (x R_f, about the origin), from the
// reference displacement set, in FP64; one wave per frame.  Also writes
// R_f^T, the rotation taking the reference onto frame f, as the reference's
// getMultipleAlignDisplaceRMSDs_atom_major returns it (rmsd/_rmsd.pyx:744-758
// with rot_msd_atom_major, rmsd/src/rotation_generic.h:47-102).
__global__ __launch_bounds__(kBlock) void rot_msd_kernel(const float* __restrict__ displ,
                                                         long long n_frames, int displ_stride,
                                                         int n_displ,
                                                         const float* __restrict__ displ_ref,
                                                         const float* __restrict__ rot,
                                                         float* __restrict__ out,
                                                         float* __restrict__ rot_t) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long n_waves = (long long)gridDim.x * kWavesPerBlock;
    for (long long f = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; f < n_frames;
         f += n_waves) {
        double R[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) R[i] = (double)rot[9 * f + i];
        const float* fr = displ + f * 3LL * displ_stride;
        double acc = 0.0;
        for (int a = lane; a < n_displ; a += kWave) {
            const double x = fr[3 * a], y = fr[3 * a + 1], z = fr[3 * a + 2];
            const double dx = x * R[0] + y * R[3] + z * R[6] - (double)displ_ref[3 * a];
            const double dy = x * R[1] + y * R[4] + z * R[7] - (double)displ_ref[3 * a + 1];
            const double dz = x * R[2] + y * R[5] + z * R[8] - (double)displ_ref[3 * a + 2];
            acc += dx * dx + dy * dy + dz * dz;
        }
        acc = wave_sum(acc);
        if (lane == 0) out[f] = (float)sqrt(acc / n_displ);
        if (rot_t && lane < 9) rot_t[9 * f + lane] = rot[9 * f + 3 * (lane % 3) + lane / 3];
    }
}

// Centroid (FP64) of every frame's selection; one wave per frame.
__global__ __launch_bounds__(kBlock) void centroid_kernel(const float* __restrict__ xyz,
                                                          long long n_frames, int n_atoms,
                                                          const int* __restrict__ idx, int n_sel,
                                                          double* __restrict__ mu) {
    const int lane = threadIdx.x & (kWave - 1);
    const long long n_waves = (long long)gridDim.x * kWavesPerBlock;
    for (long long f = (long long)blockIdx.x * kWavesPerBlock + threadIdx.x / kWave; f < n_frames;
         f += n_waves) {
        const float* fr = xyz + f * 3LL * n_atoms;
        double sx = 0.0, sy = 0.0, sz = 0.0;
        for (int j = lane; j < n_sel; j += kWave) {
            const int a = idx ? idx[j] : j;
            sx += fr[3 * a];
            sy += fr[3 * a + 1];
            sz += fr[3 * a + 2];
        }
        sx = wave_sum(sx);
        sy = wave_sum(sy);
        sz = wave_sum(sz);
        if (lane == 0) {
            mu[3 * f] = sx / n_sel;
            mu[3 * f + 1] = sy / n_sel;
            mu[3 * f + 2] = sz / n_sel;
        }
    }
}

// RMSF, pass 1: thread = selected atom, block row = chunk of frames.  The
// displaced coordinate of frame f is (x - mu_f) R_f, with mu_f = 0 unless
// the selections are centred and R_f = I for frames already aligned.  Sums
// of d = displaced - pivot (the atom's displaced coordinate in frame 0) and
// of |d|^2 in FP64, so E|d|^2 - |E d|^2 does not cancel far from the origin.
__global__ __launch_bounds__(kBlock) void rmsf_accum_kernel(
    const float* __restrict__ xyz, long long n_frames, int n_atoms, const int* __restrict__ idx,
    int n_sel, const float* __restrict__ rot, const double* __restrict__ mu, int n_chunks,
    double* __restrict__ partial) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n_sel) return;
    const int a = idx ? idx[j] : j;
    const long long f0 = (long long)blockIdx.y * n_frames / n_chunks;
    const long long f1 = (long long)(blockIdx.y + 1) * n_frames / n_chunks;
    auto displaced = [&](long long f, double& x, double& y, double& z) {
        const float* p = xyz + f * 3LL * n_atoms + 3LL * a;
        double px = p[0], py = p[1], pz = p[2];
        if (mu) {
            px -= mu[3 * f];
            py -= mu[3 * f + 1];
            pz -= mu[3 * f + 2];
        }
        if (rot) {
            const float* R = rot + 9 * f;
            x = px * R[0] + py * R[3] + pz * R[6];
            y = px * R[1] + py * R[4] + pz * R[7];
            z = px * R[2] + py * R[5] + pz * R[8];
        } else {
            x = px;
            y = py;
            z = pz;
        }
    };
    double qx, qy, qz;
    displaced(0, qx, qy, qz);
    double sx = 0.0, sy = 0.0, sz = 0.0, ss = 0.0;
    for (long long f = f0; f < f1; ++f) {
        double x, y, z;
        displaced(f, x, y, z);
        x -= qx;
        y -= qy;
        z -= qz;
        sx += x;
        sy += y;
        sz += z;
        ss += x * x + y * y + z * z;
    }
    double* o = partial + ((size_t)blockIdx.y * n_sel + j) * 4;
    o[0] = sx;
    o[1] = sy;
    o[2] = sz;
    o[3] = ss;
}

// RMSF, pass 2: per atom, sum the chunks; sqrt(E|d|^2 - |E d|^2).
__global__ __launch_bounds__(kBlock) void rmsf_finish_kernel(const double* __restrict__ partial,
                                                             int n_chunks, int n_sel,
                                                             long long n_frames,
                                                             float* __restrict__ out) {
    const int j = blockIdx.x * kBlock + threadIdx.x;
    if (j >= n_sel) return;
    double sx = 0.0, sy = 0.0, sz = 0.0, ss = 0.0;
    for (int c = 0; c < n_chunks; ++c) {
        const double* p = partial + ((size_t)c * n_sel + j) * 4;
        sx += p[0];
        sy += p[1];
        sz += p[2];
        ss += p[3];
    }
    const double inv = 1.0 / (double)n_frames;
    const double mx = sx * inv, my = sy * inv, mz = sz * inv;
    const double msf = ss * inv - (mx * mx + my * my + mz * mz);
    out[j] = (float)sqrt(msf > 0.0 ? msf : 0.0);
}

// Frame chunks of the RMSF accumulation: about 4,096 workgroups in all, at
// most 1,024 chunks and never more chunks than frames (sizes the workspace,
// so it depends on the problem only).
int rmsf_chunks(long long n_frames, int n_sel) {
    const long long atom_blocks = (n_sel + kBlock - 1) / kBlock;
    long long c = (4096 + atom_blocks - 1) / atom_blocks;
    if (c > 1024) c = 1024;
    if (c > n_frames) c = n_frames;
    return (int)(c < 1 ? 1 : c);
}

// Default chunk of hr_superpose (MiB of coordinates); HR_TUNE_SUPERPOSE_CHUNK_MIB
// overrides it.  On 100,000 frames x 2,000 atoms (2.4 GB), chunks of 1 GiB
// took 1.428 ms per call against 1.475 ms for one pass of each kernel;
// chunks of 256 MiB or less are slower (each chunk pays the two kernels'
// ramp and drain), so the rotation pass gains nothing measurable from
// re-reading a chunk that is still cached.
constexpr int kSuperposeChunkMiB = 1024;

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_rmsd_workspace_bytes(int32_t n_sel) {
    return kRefOffset + (size_t)(n_sel > 0 ? n_sel : 0) * 3 * sizeof(float);
}

int hr_center_and_trace(float* xyz, float* traces, int64_t n_frames, int32_t n_atoms,
                        void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0, "invalid sizes n_frames=%lld n_atoms=%d",
               (long long)n_frames, n_atoms);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(xyz != nullptr, "xyz is NULL");
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    long long waves = (long long)d->cu_count * kDefaultWavesPerCu;
    if (waves > n_frames) waves = n_frames;
    const int blocks = (int)((waves + kWavesPerBlock - 1) / kWavesPerBlock);
    hipLaunchKernelGGL(center_kernel, dim3(blocks), dim3(kBlock), 0, (hipStream_t)stream, xyz,
                       traces, (long long)n_frames, n_atoms);
    return launch_status("center_kernel");
}

int hr_rmsd_prepare_reference(const float* ref_frame, const int32_t* ref_idx, int32_t n_sel,
                              void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_sel > 0, "invalid n_sel=%d", n_sel);
    HR_REQUIRE(ref_frame && workspace, "NULL pointer argument");
    float* ref_sel;
    double* stats;
    return prep_ref((hipStream_t)stream, ref_frame, ref_idx, n_sel, 0, workspace, &ref_sel,
                    &stats);
}

int hr_rmsd_one_vs_all_prepared(const float* target, int64_t n_frames, int32_t n_atoms_total,
                                const int32_t* atom_idx, int32_t n_sel, int64_t self_frame,
                                float* out, const void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(target && out && workspace, "NULL pointer argument");
    const char* ws = static_cast<const char*>(workspace);
    const double* stats = reinterpret_cast<const double*>(ws);
    const float* ref_sel = reinterpret_cast<const float*>(ws + kRefOffset);
    Launch l;
    int rc = plan<Kind::Rmsd>(n_sel, atom_idx != nullptr, &l);
    if (rc) return rc;
    hipStream_t s = (hipStream_t)stream;
    const long long stride = 3LL * n_atoms_total;
    if (atom_idx)
        return launch_frames<Kind::Rmsd, false, true>(l, s, target, n_frames, stride, 0, atom_idx,
                                                      n_sel, ref_sel, stats, self_frame, out,
                                                      nullptr, nullptr);
    return launch_frames<Kind::Rmsd, false, false>(l, s, target, n_frames, stride, 0, nullptr,
                                                   n_sel, ref_sel, stats, self_frame, out, nullptr,
                                                   nullptr);
}

int hr_rmsd_one_vs_all(const float* target, int64_t n_frames, int32_t n_atoms_total,
                       const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                       const int32_t* ref_idx, int64_t self_frame, float* out, void* workspace,
                       void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(target && ref_frame && out && workspace, "NULL pointer argument");
    int rc = hr_rmsd_prepare_reference(ref_frame, ref_idx, n_sel, workspace, stream);
    if (rc) return rc;
    return hr_rmsd_one_vs_all_prepared(target, n_frames, n_atoms_total, atom_idx, n_sel,
                                       self_frame, out, workspace, stream);
}

int hr_rmsd_one_vs_all_axis_major(const float* target, int64_t n_frames, int32_t n_atoms,
                                  int32_t rowstride, const float* ref_frame_axis,
                                  int32_t ref_rowstride, int64_t self_frame, float* out,
                                  void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0 && rowstride >= n_atoms && ref_rowstride >= n_atoms,
               "invalid sizes n_frames=%lld n_atoms=%d rowstride=%d ref_rowstride=%d",
               (long long)n_frames, n_atoms, rowstride, ref_rowstride);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(target && ref_frame_axis && out && workspace, "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    float* ref_sel;
    double* stats;
    int rc = prep_ref(s, ref_frame_axis, nullptr, n_atoms, ref_rowstride, workspace, &ref_sel,
                      &stats);
    if (rc) return rc;
    Launch l;
    if ((rc = plan<Kind::Rmsd>(n_atoms, false, &l))) return rc;
    return launch_frames<Kind::Rmsd, true, false>(l, s, target, n_frames, 3LL * rowstride,
                                                  rowstride, nullptr, n_atoms, ref_sel, stats,
                                                  self_frame, out, nullptr, nullptr);
}

int hr_rmsd_nosuperpose(const float* target, int64_t n_frames, int32_t n_atoms_total,
                        const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                        const int32_t* ref_idx, float* out, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms_total > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames,
               n_atoms_total, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms_total,
               "n_sel (%d) must equal n_atoms_total (%d) without atom_idx", n_sel, n_atoms_total);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(target && ref_frame && out && workspace, "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    float* ref_sel;
    double* stats;
    int rc = prep_ref(s, ref_frame, ref_idx, n_sel, 0, workspace, &ref_sel, &stats);
    if (rc) return rc;
    Launch l;
    if ((rc = plan<Kind::NoSuperpose>(n_sel, atom_idx != nullptr, &l))) return rc;
    const long long stride = 3LL * n_atoms_total;
    if (atom_idx)
        return launch_frames<Kind::NoSuperpose, false, true>(l, s, target, n_frames, stride, 0,
                                                             atom_idx, n_sel, ref_sel, stats, -1,
                                                             out, nullptr, nullptr);
    return launch_frames<Kind::NoSuperpose, false, false>(l, s, target, n_frames, stride, 0,
                                                          nullptr, n_sel, ref_sel, stats, -1, out,
                                                          nullptr, nullptr);
}

size_t hr_superpose_workspace_bytes(int64_t n_frames, int32_t n_sel) {
    const size_t nf = (size_t)(n_frames > 0 ? n_frames : 0);
    size_t b = hr_rmsd_workspace_bytes(n_sel);
    b = (b + 255) & ~(size_t)255;
    b += nf * 3 * sizeof(double);  // per-frame centroids
    b = (b + 255) & ~(size_t)255;
    b += nf * 9 * sizeof(float);  // rotations when the caller passes none
    return b;
}

int hr_superpose(float* xyz, int64_t n_frames, int32_t n_atoms, const float* align_xyz,
                 int32_t n_align_atoms, const int32_t* atom_idx, int32_t n_sel,
                 const float* ref_frame, const int32_t* ref_idx, int32_t translate,
                 float* rot_out, float* rmsd_out, void* workspace, void* stream) {
    clear_error();
    if (!align_xyz) n_align_atoms = n_atoms;
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0 && n_align_atoms > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_align_atoms=%d n_sel=%d",
               (long long)n_frames, n_atoms, n_align_atoms, n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_align_atoms,
               "n_sel (%d) must equal n_align_atoms (%d) without atom_idx", n_sel, n_align_atoms);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(xyz && ref_frame && workspace, "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(workspace);
    size_t off = (hr_rmsd_workspace_bytes(n_sel) + 255) & ~(size_t)255;
    double* mu = reinterpret_cast<double*>(ws + off);
    off += (size_t)n_frames * 3 * sizeof(double);
    off = (off + 255) & ~(size_t)255;
    float* rot = rot_out ? rot_out : reinterpret_cast<float*>(ws + off);
    float* ref_sel;
    double* stats;
    int rc = prep_ref(s, ref_frame, ref_idx, n_sel, 0, workspace, &ref_sel, &stats);
    if (rc) return rc;
    Launch l;
    if ((rc = plan<Kind::Fit>(n_sel, atom_idx != nullptr, &l))) return rc;
    const float* fit = align_xyz ? align_xyz : xyz;
    const long long stride = 3LL * n_align_atoms;
    const DeviceInfo* d = current_device();
    // Fit and rotate chunk by chunk (kSuperposeChunkMiB); every frame's
    // rotation depends on that frame alone, so results equal one pass.
    const long long frame_bytes = 12LL * n_atoms;
    const int mib = tuning(HR_TUNE_SUPERPOSE_CHUNK_MIB) ? tuning(HR_TUNE_SUPERPOSE_CHUNK_MIB)
                                                        : kSuperposeChunkMiB;
    long long chunk = ((long long)mib << 20) / frame_bytes;
    if (chunk < 1) chunk = 1;
    for (long long f0 = 0; f0 < n_frames; f0 += chunk) {
        const long long nf = n_frames - f0 < chunk ? n_frames - f0 : chunk;
        float* out_f = rmsd_out ? rmsd_out + f0 : nullptr;
        if (atom_idx)
            rc = launch_frames<Kind::Fit, false, true>(l, s, fit + f0 * stride, nf, stride, 0,
                                                       atom_idx, n_sel, ref_sel, stats, -1, out_f,
                                                       rot + 9 * f0, mu + 3 * f0);
        else
            rc = launch_frames<Kind::Fit, false, false>(l, s, fit + f0 * stride, nf, stride, 0,
                                                        nullptr, n_sel, ref_sel, stats, -1, out_f,
                                                        rot + 9 * f0, mu + 3 * f0);
        if (rc) return rc;
        long long blocks = (long long)d->cu_count * 8;
        if (blocks > nf) blocks = nf;
        hipLaunchKernelGGL(apply_kernel, dim3((int)blocks), dim3(kBlock), 0, s,
                           xyz + f0 * 3LL * n_atoms, nf, n_atoms, rot + 9 * f0, mu + 3 * f0, stats,
                           (double)n_sel, (int)(translate != 0));
        if ((rc = launch_status("apply_kernel"))) return rc;
    }
    return HR_OK;
}

size_t hr_align_displace_workspace_bytes(int64_t n_frames, int32_t n_align) {
    return hr_superpose_workspace_bytes(n_frames, n_align);
}

int hr_rmsd_align_displace(const float* align, int64_t n_frames, int32_t align_stride,
                           int32_t n_align, const float* align_ref, const float* displ,
                           int32_t displ_stride, int32_t n_displ, const float* displ_ref,
                           float* rmsd_out, float* rot_out, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_align > 0 && n_displ > 0 && align_stride >= n_align &&
                   displ_stride >= n_displ,
               "invalid sizes n_frames=%lld n_align=%d (stride %d) n_displ=%d (stride %d)",
               (long long)n_frames, n_align, align_stride, n_displ, displ_stride);
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(align && align_ref && displ && displ_ref && rmsd_out && workspace,
               "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(workspace);
    size_t off = (hr_rmsd_workspace_bytes(n_align) + 255) & ~(size_t)255;
    double* mu = reinterpret_cast<double*>(ws + off);
    off += (size_t)n_frames * 3 * sizeof(double);
    off = (off + 255) & ~(size_t)255;
    float* rot = reinterpret_cast<float*>(ws + off);
    float* ref_sel;
    double* stats;
    int rc = prep_ref(s, align_ref, nullptr, n_align, 0, workspace, &ref_sel, &stats);
    if (rc) return rc;
    Launch l;
    if ((rc = plan<Kind::Fit>(n_align, false, &l))) return rc;
    rc = launch_frames<Kind::Fit, false, false>(l, s, align, n_frames, 3LL * align_stride, 0,
                                                nullptr, n_align, ref_sel, stats, -1, nullptr,
                                                rot, mu);
    if (rc) return rc;
    const DeviceInfo* d = current_device();
    long long blocks = (long long)d->cu_count * 8;
    const long long by_work = (n_frames + kWavesPerBlock - 1) / kWavesPerBlock;
    if (blocks > by_work) blocks = by_work;
    hipLaunchKernelGGL(rot_msd_kernel, dim3((int)blocks), dim3(kBlock), 0, s, displ,
                       (long long)n_frames, displ_stride, n_displ, displ_ref, rot, rmsd_out,
                       rot_out);
    return launch_status("rot_msd_kernel");
}

size_t hr_rmsf_workspace_bytes(int64_t n_frames, int32_t n_sel) {
    size_t b = hr_superpose_workspace_bytes(n_frames, n_sel);
    b = (b + 255) & ~(size_t)255;
    b += (size_t)rmsf_chunks(n_frames > 0 ? n_frames : 1, n_sel > 0 ? n_sel : 1) *
         (size_t)(n_sel > 0 ? n_sel : 0) * 4 * sizeof(double);
    return b;
}

int hr_rmsf(const float* xyz, int64_t n_frames, int32_t n_atoms, const int32_t* atom_idx,
            int32_t n_sel, const float* ref_frame, const int32_t* ref_idx, int32_t center,
            float* out, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames > 0 && n_atoms > 0 && n_sel > 0,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d", (long long)n_frames, n_atoms,
               n_sel);
    HR_REQUIRE(atom_idx != nullptr || n_sel == n_atoms,
               "n_sel (%d) must equal n_atoms (%d) without atom_idx", n_sel, n_atoms);
    HR_REQUIRE(xyz && out && workspace, "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(workspace);
    size_t off = (hr_rmsd_workspace_bytes(n_sel) + 255) & ~(size_t)255;
    double* mu = reinterpret_cast<double*>(ws + off);
    off += (size_t)n_frames * 3 * sizeof(double);
    off = (off + 255) & ~(size_t)255;
    float* rot = reinterpret_cast<float*>(ws + off);
    off += (size_t)n_frames * 9 * sizeof(float);
    off = (off + 255) & ~(size_t)255;
    double* partial = reinterpret_cast<double*>(ws + off);
    const DeviceInfo* d = current_device();
    if (!d) return HR_ERR_HIP;
    int rc;
    if (ref_frame) {
        float* ref_sel;
        double* stats;
        if ((rc = prep_ref(s, ref_frame, ref_idx, n_sel, 0, workspace, &ref_sel, &stats)))
            return rc;
        Launch l;
        if ((rc = plan<Kind::Fit>(n_sel, atom_idx != nullptr, &l))) return rc;
        if (atom_idx)
            rc = launch_frames<Kind::Fit, false, true>(l, s, xyz, n_frames, 3LL * n_atoms, 0,
                                                       atom_idx, n_sel, ref_sel, stats, -1,
                                                       nullptr, rot, mu);
        else
            rc = launch_frames<Kind::Fit, false, false>(l, s, xyz, n_frames, 3LL * n_atoms, 0,
                                                        nullptr, n_sel, ref_sel, stats, -1,
                                                        nullptr, rot, mu);
        if (rc) return rc;
    } else if (center) {
        long long blocks = (long long)d->cu_count * 8;
        const long long by_work = (n_frames + kWavesPerBlock - 1) / kWavesPerBlock;
        if (blocks > by_work) blocks = by_work;
        hipLaunchKernelGGL(centroid_kernel, dim3((int)blocks), dim3(kBlock), 0, s, xyz,
                           (long long)n_frames, n_atoms, atom_idx, n_sel, mu);
        if ((rc = launch_status("centroid_kernel"))) return rc;
    }
    const int n_chunks = rmsf_chunks(n_frames, n_sel);
    const int atom_blocks = (n_sel + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(rmsf_accum_kernel, dim3(atom_blocks, n_chunks), dim3(kBlock), 0, s, xyz,
                       (long long)n_frames, n_atoms, atom_idx, n_sel, ref_frame ? rot : nullptr,
                       center ? mu : nullptr, n_chunks, partial);
    if ((rc = launch_status("rmsf_accum_kernel"))) return rc;
    hipLaunchKernelGGL(rmsf_finish_kernel, dim3(atom_blocks), dim3(kBlock), 0, s, partial,
                       n_chunks, n_sel, (long long)n_frames, out);
    return launch_status("rmsf_finish_kernel");
}

}  // extern "C"
