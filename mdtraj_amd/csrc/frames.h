// Device code shared by the per-frame kernels (one-vs-all, superpose fit,
// no-superpose, batched systems): FP64 per-frame sums, the row item loop
// (kRL lanes per frame) with its register ring, DPP row reductions, and the
// per-frame finish (QCP or no-superpose) one lane per frame.
#pragma once

#include "common.h"

namespace hr {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kDefaultWavesPerCu = 16;

// The per-frame kernels run one workgroup per CU holding all of its waves
// (kMinWaves per SIMD), so that those waves can share work through LDS.
template <int kMinWaves>
struct CuGroup {
    static constexpr int kWaves = 4 * kMinWaves;
    static constexpr int kThreads = kWave * kWaves;
};

// Share of each workgroup's frame sets split statically over its waves (the
// rest goes out in chunks by an LDS ticket), and the bytes of frames per
// chunk; HR_TUNE_FRAMES_STATIC_PCT and HR_TUNE_FRAMES_CHUNK_KIB override.
constexpr int kDefaultStaticPct = 45;
constexpr int kDefaultChunkKiB = 32;

// One-vs-all workspace: [0, 64) reference sums (double sx, sy, sz, q);
// [256, 256 + 12 n_sel) gathered reference, axis-major (x[], y[], z[]).
constexpr size_t kRefOffset = 256;

enum class Kind { Rmsd, NoSuperpose, Fit };

// Lanes per frame (row width of the per-frame item loop) by selection size,
// from a sweep over 8 - 3,000 selected atoms at ~2.4 GB per launch on MI355X:
// narrow rows keep the lanes of short frames busy and amortise the per-frame
// reduction over more frames; wide rows make each load instruction one
// contiguous span of a frame (768 bytes at 64 lanes) instead of several short
// ones.  Gathered selections of 1,400+ atoms stream best one frame per wave,
// contiguous ones two.
__host__ __device__ __forceinline__ int auto_row_lanes(int n_sel, bool gather) {
    if (n_sel < 28) return 4;
    if (n_sel < 112) return 8;
    if (n_sel < 720) return 16;
    if (n_sel < 1400 || !gather) return 32;
    return 64;
}

// What a per-frame workgroup stages in LDS (see plan in one_vs_all.hip).
constexpr int kStageNone = 0;
constexpr int kStageIdx = 1;
constexpr int kStageAll = 2;

// Gather the reference selection into axis-major planes and sum it in FP64.
// The first kBlock threads of the workgroup (which may be larger) do the
// work, so the sums come out the same in any kernel that calls it.  Each
// thread loads kPrepBatch atoms before it uses any (indices clamped into the
// selection, so no load sits under a branch); the per-thread order of the
// sums is that of a plain strided loop.
constexpr int kPrepBatch = 8;

__device__ __forceinline__ void ref_prep_body(const float* __restrict__ ref,
                                              const int* __restrict__ ref_idx, int n_sel,
                                              int axis_rowstride, float* __restrict__ ref_out,
                                              double* __restrict__ stats) {
    double sx = 0.0, sy = 0.0, sz = 0.0, q = 0.0;
    const int t0 = threadIdx.x < kBlock ? (int)threadIdx.x : n_sel;  // the rest: no atoms
    for (int j0 = t0; j0 < n_sel; j0 += kPrepBatch * kBlock) {
        float x[kPrepBatch], y[kPrepBatch], z[kPrepBatch];
#pragma unroll
        for (int u = 0; u < kPrepBatch; ++u) {
            const int j = min(j0 + u * kBlock, n_sel - 1);
            if (axis_rowstride > 0) {
                x[u] = ref[j];
                y[u] = ref[axis_rowstride + j];
                z[u] = ref[2 * axis_rowstride + j];
            } else {
                const long long a = ref_idx ? ref_idx[j] : j;
                x[u] = ref[3 * a];
                y[u] = ref[3 * a + 1];
                z[u] = ref[3 * a + 2];
            }
        }
#pragma unroll
        for (int u = 0; u < kPrepBatch; ++u) {
            const int j = j0 + u * kBlock;
            if (j < n_sel) {
                ref_out[j] = x[u];
                ref_out[n_sel + j] = y[u];
                ref_out[2 * n_sel + j] = z[u];
                const double dx = x[u], dy = y[u], dz = z[u];
                sx += dx;
                sy += dy;
                sz += dz;
                q = fma(dx, dx, fma(dy, dy, fma(dz, dz, q)));
            }
        }
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    q = wave_sum(q);
    __shared__ double part[kWavesPerBlock][4];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (lane == 0 && w < kWavesPerBlock) {
        part[w][0] = sx;
        part[w][1] = sy;
        part[w][2] = sz;
        part[w][3] = q;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int i = 0; i < kWavesPerBlock; ++i) s += part[i][threadIdx.x];
        stats[threadIdx.x] = s;
    }
}

__global__ __launch_bounds__(kBlock) void ref_prep_kernel(const float* __restrict__ ref,
                                                          const int* __restrict__ ref_idx,
                                                          int n_sel, int axis_rowstride,
                                                          float* __restrict__ ref_out,
                                                          double* __restrict__ stats) {
    ref_prep_body(ref, ref_idx, n_sel, axis_rowstride, ref_out, stats);
}

struct Sums {
    double sx, sy, sz, q;
    double p[9];
};

template <Kind kKind>
__device__ __forceinline__ void accumulate(Sums& s, float ax, float ay, float az, float rx,
                                           float ry, float rz) {
    const double x = ax, y = ay, z = az;
    if constexpr (kKind == Kind::NoSuperpose) {
        const double dx = x - (double)rx, dy = y - (double)ry, dz = z - (double)rz;
        s.q = fma(dx, dx, fma(dy, dy, fma(dz, dz, s.q)));
    } else {
        const double bx = rx, by = ry, bz = rz;
        s.sx += x;
        s.sy += y;
        s.sz += z;
        s.q = fma(x, x, fma(y, y, fma(z, z, s.q)));
        s.p[0] = fma(x, bx, s.p[0]);
        s.p[1] = fma(x, by, s.p[1]);
        s.p[2] = fma(x, bz, s.p[2]);
        s.p[3] = fma(y, bx, s.p[3]);
        s.p[4] = fma(y, by, s.p[4]);
        s.p[5] = fma(y, bz, s.p[5]);
        s.p[6] = fma(z, bx, s.p[6]);
        s.p[7] = fma(z, by, s.p[7]);
        s.p[8] = fma(z, bz, s.p[8]);
    }
}

template <Kind kKind>
__device__ __forceinline__ void wave_reduce(Sums& s) {
    s.q = wave_sum(s.q);
    if constexpr (kKind != Kind::NoSuperpose) {
        s.sx = wave_sum(s.sx);
        s.sy = wave_sum(s.sy);
        s.sz = wave_sum(s.sz);
#pragma unroll
        for (int i = 0; i < 9; ++i) s.p[i] = wave_sum(s.p[i]);
    }
}

template <Kind kKind>
__device__ __forceinline__ void take(Sums& dst, const Sums& src, bool pred) {
    dst.q = pred ? src.q : dst.q;
    if constexpr (kKind != Kind::NoSuperpose) {
        dst.sx = pred ? src.sx : dst.sx;
        dst.sy = pred ? src.sy : dst.sy;
        dst.sz = pred ? src.sz : dst.sz;
#pragma unroll
        for (int i = 0; i < 9; ++i) dst.p[i] = pred ? src.p[i] : dst.p[i];
    }
}

// Centred correlation matrix and traces from raw FP64 sums:
//   M = sum a b^T - S_a S_b^T / N,  G = Q - |S|^2 / N.
__device__ __forceinline__ void centred(const Sums& s, const double* rs, double n, double M[9],
                                        double& ga, double& gb) {
    const double inv = 1.0 / n;
    const double sa[3] = {s.sx, s.sy, s.sz};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) M[3 * r + c] = s.p[3 * r + c] - sa[r] * rs[c] * inv;
    ga = s.q - (sa[0] * sa[0] + sa[1] * sa[1] + sa[2] * sa[2]) * inv;
    gb = rs[3] - (rs[0] * rs[0] + rs[1] * rs[1] + rs[2] * rs[2]) * inv;
    ga = ga > 0.0 ? ga : 0.0;
    gb = gb > 0.0 ? gb : 0.0;
}

template <Kind kKind>
struct SumsPerFrame {
    static constexpr int value = kKind == Kind::NoSuperpose ? 1 : 13;
};

template <Kind kKind>
__device__ __forceinline__ void store_sums(double* dst, const Sums& s) {
    dst[0] = s.q;
    if constexpr (kKind != Kind::NoSuperpose) {
        dst[1] = s.sx;
        dst[2] = s.sy;
        dst[3] = s.sz;
#pragma unroll
        for (int e = 0; e < 9; ++e) dst[4 + e] = s.p[e];
    }
}

template <Kind kKind>
__device__ __forceinline__ void load_sums(Sums& s, const double* src) {
    s.q = src[0];
    if constexpr (kKind != Kind::NoSuperpose) {
        s.sx = src[1];
        s.sy = src[2];
        s.sz = src[3];
#pragma unroll
        for (int e = 0; e < 9; ++e) s.p[e] = src[4 + e];
    }
}

// LDS bytes of the per-wave result area (64 frames of sums per wave) of a
// workgroup of `waves` waves.
template <Kind kKind>
constexpr size_t result_lds_bytes(int waves = kWavesPerBlock) {
    return (size_t)waves * kWave * SumsPerFrame<kKind>::value * sizeof(double);
}

// Copies n 32-bit words from global memory to LDS, the workgroup's kThreads
// lanes each issuing up to kStageBatch loads before the first LDS store, so that a
// reference of thousands of atoms costs a couple of load latencies rather
// than one per 256 words (a plain copy loop waits for each load before its
// store).  16-byte loads when both pointers allow them.  No barrier.
constexpr int kStageBatch = 8;

template <int kThreads = kBlock, typename T>
__device__ __forceinline__ void stage_words(T* __restrict__ dst, const T* __restrict__ src,
                                            int n) {
    static_assert(sizeof(T) == 4, "32-bit words");
    typedef float v4f __attribute__((ext_vector_type(4)));
    const float* s = reinterpret_cast<const float*>(src);
    float* d = reinterpret_cast<float*>(dst);
    int done = 0;
    if (((reinterpret_cast<uintptr_t>(src) | reinterpret_cast<uintptr_t>(dst)) & 15) == 0) {
        const int n4 = n / 4;
        const __attribute__((address_space(1))) v4f* s4 =
            reinterpret_cast<const __attribute__((address_space(1))) v4f*>(global_ptr(s));
        v4f* d4 = reinterpret_cast<v4f*>(d);
        for (int j0 = threadIdx.x; j0 < n4; j0 += kThreads * kStageBatch) {
            v4f v[kStageBatch];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u) v[u] = s4[min(j0 + u * kThreads, n4 - 1)];
#pragma unroll
            for (int u = 0; u < kStageBatch; ++u)
                if (j0 + u * kThreads < n4) d4[j0 + u * kThreads] = v[u];
        }
        done = 4 * n4;
    }
    const GlobalFloat* g = global_ptr(s);
    for (int j0 = done + (int)threadIdx.x; j0 < n; j0 += kThreads * kStageBatch) {
        float v[kStageBatch];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u) v[u] = g[min(j0 + u * kThreads, n - 1)];
#pragma unroll
        for (int u = 0; u < kStageBatch; ++u)
            if (j0 + u * kThreads < n) d[j0 + u * kThreads] = v[u];
    }
}

// Row layout: a row of kRL lanes (4, 8, 16, 32 or 64) owns one frame, so a
// wave works on kWave / kRL frames at once (row r = lane / kRL holds frame
// base + r).  An item is (frame set q, atom block b): lane t of a row takes
// the selected atoms b * kRL * kItemAtoms + t + kRL u, u < kItemAtoms, so
// one load instruction reads kRL consecutive selected atoms of each row's
// frame.  kRowLanes is the batched kernel's row width.
constexpr int kRowLanes = 16;
constexpr int kItemAtoms = 4;   // atoms per lane per item
constexpr int kRing = 3;        // items in the register ring (kRing - 1 in flight)

// Frame coordinates are read once per launch: nontemporal loads (one
// global_load_dwordx3 ... nt per atom) keep them from displacing reusable
// lines in the caches; on MI355X they stream 2-5% faster than plain loads
// (100,000 frames x 2,000 atoms: 0.419 ms against 0.426-0.442 ms).
// Slots past the selection (and rows past the group's last frame, for which
// fr points at a valid frame) are zeroed where they are used (valid_atom).
//   kBranchFree false: such slots skip their loads (a whole instruction when
//     no lane needs it), the cheaper form for short frames.
//   kBranchFree true: every slot loads, past-the-end ones the last selected
//     atom.  The batched kernel needs this form: with its ring carried from
//     one work unit to the next, a load under a branch is merged with the
//     branch's zero by a register copy, and that copy makes the wave wait
//     for each load as soon as it is issued (one load in flight, not two
//     items).
template <bool kAxis, bool kGather, int kRL = kRowLanes, bool kBranchFree = false>
__device__ __forceinline__ void load_item(const float* __restrict__ fr, int rowstride,
                                          const int* __restrict__ ix, int n_sel, int t, int b,
                                          float (&ax)[kItemAtoms], float (&ay)[kItemAtoms],
                                          float (&az)[kItemAtoms]) {
    const int j0 = b * kRL * kItemAtoms + t;
    const GlobalFloat* g = global_ptr(fr);
#pragma unroll
    for (int u = 0; u < kItemAtoms; ++u) {
        const int jr = j0 + u * kRL;
        const int j = kBranchFree ? min(jr, n_sel - 1) : jr;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (kBranchFree || j < n_sel) {
            if constexpr (kAxis) {
                x = __builtin_nontemporal_load(g + j);
                y = __builtin_nontemporal_load(g + rowstride + j);
                z = __builtin_nontemporal_load(g + 2 * rowstride + j);
            } else {
                const int a = kGather ? ix[j] : j;
                x = __builtin_nontemporal_load(g + 3 * a);
                y = __builtin_nontemporal_load(g + 3 * a + 1);
                z = __builtin_nontemporal_load(g + 3 * a + 2);
            }
        }
        ax[u] = x;
        ay[u] = y;
        az[u] = z;
    }
}

// Reference atoms of the same slots as load_item, for a reference that is
// read from global memory (not staged in LDS): loaded with the frame data so
// that one wait covers both (a reference read at accumulation time would
// wait for every frame load issued before it).
template <int kRL>
__device__ __forceinline__ void load_ref_item(const float* __restrict__ rx,
                                              const float* __restrict__ ry,
                                              const float* __restrict__ rz, int n_sel, int t,
                                              int b, float (&qx)[kItemAtoms],
                                              float (&qy)[kItemAtoms], float (&qz)[kItemAtoms]) {
    const int j0 = b * kRL * kItemAtoms + t;
#pragma unroll
    for (int u = 0; u < kItemAtoms; ++u) {
        const int j = j0 + u * kRL;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (j < n_sel) {
            x = global_ptr(rx)[j];
            y = global_ptr(ry)[j];
            z = global_ptr(rz)[j];
        }
        qx[u] = x;
        qy[u] = y;
        qz[u] = z;
    }
}

// Coordinates of a ring slot as accumulated: zero for slots past the
// selection or past the group's last frame, so they add nothing.
__device__ __forceinline__ float valid_atom(bool ok, float v) { return ok ? v : 0.0f; }

// Sum over each row of kRL lanes (4, 8, 16, 32 or 64), result in every lane
// of the row: the DPP steps of row16_sum up to the row width (xor 1, xor 2,
// half-row mirror, row mirror), then lane-xor shuffles across rows of 16.
template <int kRL>
__device__ __forceinline__ double row_sum(double v) {
    static_assert(kRL == 4 || kRL == 8 || kRL == 16 || kRL == 32 || kRL == 64, "row width");
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    if constexpr (kRL >= 8) v += dpp_move<0x141>(v);
    if constexpr (kRL >= 16) v += dpp_move<0x140>(v);
    if constexpr (kRL >= 32) v += __shfl_xor(v, 16, kWave);
    if constexpr (kRL >= 64) v += __shfl_xor(v, 32, kWave);
    return v;
}

template <Kind kKind, int kRL = kRowLanes>
__device__ __forceinline__ void row_reduce(Sums& s) {
    s.q = row_sum<kRL>(s.q);
    if constexpr (kKind != Kind::NoSuperpose) {
        s.sx = row_sum<kRL>(s.sx);
        s.sy = row_sum<kRL>(s.sy);
        s.sz = row_sum<kRL>(s.sz);
#pragma unroll
        for (int e = 0; e < 9; ++e) s.p[e] = row_sum<kRL>(s.p[e]);
    }
}

// Finish a group of cnt frames whose row sums are parked in s_res: one lane
// per frame computes the QCP MSD (or the fit, or the no-superpose MSD).
template <Kind kKind>
__device__ __forceinline__ void finish_group_at(long long f, int cnt, double n,
                                                const double (&rs)[4], long long self_frame,
                                                float* __restrict__ out,
                                                float* __restrict__ rot_out,
                                                double* __restrict__ mu_out,
                                                double* __restrict__ s_res, int lane);

template <Kind kKind>
__device__ __forceinline__ void finish_group(long long base, int cnt, double n,
                                             const double (&rs)[4], long long self_frame,
                                             float* __restrict__ out, float* __restrict__ rot_out,
                                             double* __restrict__ mu_out,
                                             double* __restrict__ s_res, int lane) {
    finish_group_at<kKind>(base + lane, cnt, n, rs, self_frame, out, rot_out, mu_out, s_res,
                           lane);
}

// As finish_group, lane k < cnt finishing frame f (its own frame index: the
// frames of a group need not be consecutive).
template <Kind kKind>
__device__ __forceinline__ void finish_group_at(long long f, int cnt, double n,
                                                const double (&rs)[4], long long self_frame,
                                                float* __restrict__ out,
                                                float* __restrict__ rot_out,
                                                double* __restrict__ mu_out,
                                                double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    // the row leaders' LDS stores must be visible to every lane of the wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < cnt) {
        Sums mine;
        load_sums<kKind>(mine, s_res + lane * kNS);
        if constexpr (kKind == Kind::NoSuperpose) {
            global_out(out)[f] = (float)sqrt(mine.q / n);
        } else {
            double M[9], ga, gb;
            centred(mine, rs, n, M, ga, gb);
            if constexpr (kKind == Kind::Rmsd) {
                const double msd = qcp_msd(M, ga, gb, n);
                global_out(out)[f] = (f == self_frame) ? 0.0f : (float)sqrt(msd);
            } else {
                double R[9];
                const double lam = qcp_solve<true>(M, 0.5 * (ga + gb), R);
#pragma unroll
                for (int e = 0; e < 9; ++e) global_out(rot_out)[9 * f + e] = (float)R[e];
                global_out(mu_out)[3 * f + 0] = mine.sx / n;
                global_out(mu_out)[3 * f + 1] = mine.sy / n;
                global_out(mu_out)[3 * f + 2] = mine.sz / n;
                if (out) {
                    const double msd = (ga + gb - 2.0 * lam) / n;
                    global_out(out)[f] = (float)sqrt(msd > 0.0 ? msd : 0.0);
                }
            }
        }
    }
    // the next group overwrites s_res: all lanes must have read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Loads of the first kRing - 1 items of a group (base .. base + cnt) into
// ring slots 0 .. kRing - 2 (with their reference atoms when kRefRing).
template <bool kAxis, bool kGather, int kRL, bool kRefRing>
__device__ __forceinline__ void prime_ring(
    const float* __restrict__ xyz, long long base, int cnt, long long frame_stride,
    int rowstride, const int* __restrict__ ix, int n_sel, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, int n_blk, int row, int t,
    float (&bx)[kRing][kItemAtoms], float (&by)[kRing][kItemAtoms],
    float (&bz)[kRing][kItemAtoms], float (&qx)[kRing][kItemAtoms],
    float (&qy)[kRing][kItemAtoms], float (&qz)[kRing][kItemAtoms]) {
    constexpr int kRows = kWave / kRL;
    const int n_items = (cnt + kRows - 1) / kRows * n_blk;
#pragma unroll
    for (int r = 0; r < kRing - 1; ++r) {
        if (r < n_items) {
            const int q = r / n_blk, b = r - q * n_blk;
            const int k = kRows * q + row;
            load_item<kAxis, kGather, kRL>(xyz + (base + (k < cnt ? k : 0)) * frame_stride,
                                           rowstride, ix, n_sel, t, b, bx[r], by[r], bz[r]);
            if constexpr (kRefRing) load_ref_item<kRL>(rx, ry, rz, n_sel, t, b, qx[r], qy[r], qz[r]);
        }
    }
}

// The wave's frames f_begin .. f_end of one trajectory, in groups of up to
// 64: items (frame set, atom block) stream through the register ring, each
// set's row sums are reduced (DPP) and parked in the wave's LDS slots, and
// at the end of a group one lane per frame finishes (QCP or no-superpose).
// The ring is empty when a group's last item has been accumulated, so the
// next group's first kRing - 1 items are loaded before the finish: their
// latency hides behind the QCP solves instead of following them.  The fit
// finish (rotation and centroid) needs the registers those loads would hold,
// so that kind primes the ring after its finish.  kRefRing: the reference is
// in global memory and rides the ring with the frame data; a ring that large
// would not fit beside the QCP solves either, and it serves selections too
// large for LDS, whose long frames make the early start matter little.
template <Kind kKind, bool kAxis, bool kGather, int kRL, bool kRefRing>
__device__ __forceinline__ void process_chunk(
    const float* __restrict__ xyz, long long f_begin, long long f_end, long long frame_stride,
    int rowstride, const int* __restrict__ ix, int n_sel, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, const double (&rs)[4],
    long long self_frame, float* __restrict__ out, float* __restrict__ rot_out,
    double* __restrict__ mu_out, double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    constexpr int kRows = kWave / kRL;  // frames side by side in the wave
    const int row = lane / kRL, t = lane % kRL;
    const double n = (double)n_sel;
    const int n_blk = (n_sel + kRL * kItemAtoms - 1) / (kRL * kItemAtoms);
    float bx[kRing][kItemAtoms], by[kRing][kItemAtoms], bz[kRing][kItemAtoms];
    float qx[kRing][kItemAtoms], qy[kRing][kItemAtoms], qz[kRing][kItemAtoms];  // kRefRing
    if (f_begin < f_end)
        prime_ring<kAxis, kGather, kRL, kRefRing>(
            xyz, f_begin, (int)min((long long)kWave, f_end - f_begin), frame_stride, rowstride, ix,
            n_sel, rx, ry, rz, n_blk, row, t, bx, by, bz, qx, qy, qz);
    for (long long base = f_begin; base < f_end; base += kWave) {
        const int cnt = (int)min((long long)kWave, f_end - base);
        const int n_items = (cnt + kRows - 1) / kRows * n_blk;
        Sums s = {};
        for (int it0 = 0; it0 < n_items; it0 += kRing) {
#pragma unroll
            for (int r = 0; r < kRing; ++r) {
                const int it = it0 + r;
                if (it < n_items) {
                    // refill the slot freed one step ago with item it + kRing - 1
                    const int nxt = it + kRing - 1;
                    if (nxt < n_items) {
                        const int q = nxt / n_blk, b = nxt - q * n_blk;
                        const int k = kRows * q + row;
                        const int slot = (r + kRing - 1) % kRing;  // static once unrolled
                        load_item<kAxis, kGather, kRL>(
                            xyz + (base + (k < cnt ? k : 0)) * frame_stride, rowstride, ix,
                            n_sel, t, b, bx[slot], by[slot], bz[slot]);
                        if constexpr (kRefRing)
                            load_ref_item<kRL>(rx, ry, rz, n_sel, t, b, qx[slot], qy[slot],
                                               qz[slot]);
                    }
                    const int q = it / n_blk, b = it - q * n_blk;
                    const int j0 = b * kRL * kItemAtoms + t;
                    const bool row_ok = kRows * q + row < cnt;
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int j = j0 + u * kRL;
                        const bool ok = row_ok && j < n_sel;
                        const int jc = min(j, n_sel - 1);
                        // products with a zeroed frame atom vanish; the
                        // no-superpose difference needs the reference zeroed
                        constexpr bool kRef = kKind == Kind::NoSuperpose;
                        const float ux = kRefRing ? qx[r][u] : rx[jc];
                        const float uy = kRefRing ? qy[r][u] : ry[jc];
                        const float uz = kRefRing ? qz[r][u] : rz[jc];
                        accumulate<kKind>(s, valid_atom(ok, bx[r][u]), valid_atom(ok, by[r][u]),
                                          valid_atom(ok, bz[r][u]), kRef ? valid_atom(ok, ux) : ux,
                                          kRef ? valid_atom(ok, uy) : uy,
                                          kRef ? valid_atom(ok, uz) : uz);
                    }
                    if (b == n_blk - 1) {
                        row_reduce<kKind, kRL>(s);
                        const int k = kRows * q + row;
                        if (t == 0 && k < cnt) store_sums<kKind>(s_res + k * kNS, s);
                        s = Sums{};
                    }
                }
            }
        }
        constexpr bool kEarly = kKind != Kind::Fit && !kRefRing;
        const long long next = base + kWave;
        if (kEarly && next < f_end)
            prime_ring<kAxis, kGather, kRL, kRefRing>(
                xyz, next, (int)min((long long)kWave, f_end - next), frame_stride, rowstride, ix,
                n_sel, rx, ry, rz, n_blk, row, t, bx, by, bz, qx, qy, qz);
        finish_group<kKind>(base, cnt, n, rs, self_frame, out, rot_out, mu_out, s_res, lane);
        if (!kEarly && next < f_end)
            prime_ring<kAxis, kGather, kRL, kRefRing>(
                xyz, next, (int)min((long long)kWave, f_end - next), frame_stride, rowstride, ix,
                n_sel, rx, ry, rz, n_blk, row, t, bx, by, bz, qx, qy, qz);
    }
}

// The per-frame item loop over a workgroup that fills a CU.  The waves
// sharing a SIMD do not progress alike (the oldest is issued first): with a
// static split of the frames over 3 waves/SIMD (MI355X, 100,000 x 2,000
// frames) they finish in three bands, at about 270, 315 and 375 us, and the
// last band sets the kernel's time.  One workgroup per CU holds all of a
// CU's waves, so they can share its work through an LDS counter: each wave
// first takes its static range [sa, sb), then chunks of `chunk` frames from
// [da, db) by the ticket *s_ticket (chunk c = frames da + c chunk ...), until
// the workgroup's frames run out.  Chunks fill the wave's group of 64 frame
// slots in turn, lane k keeping the index of the frame in slot k, so the QCP
// finish still runs once per 64 frames.  The item loop is process_chunk's.
template <Kind kKind, bool kAxis, bool kGather, int kRL, bool kRefRing>
__device__ __forceinline__ void process_cu(
    const float* __restrict__ xyz, long long sa, long long sb, long long da, long long db,
    long long chunk, unsigned* s_ticket, long long frame_stride, int rowstride,
    const int* __restrict__ ix, int n_sel, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, const double (&rs)[4],
    long long self_frame, float* __restrict__ out, float* __restrict__ rot_out,
    double* __restrict__ mu_out, double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    constexpr int kRows = kWave / kRL;  // frames side by side in the wave
    constexpr bool kEarly = kKind != Kind::Fit && !kRefRing;
    const int row = lane / kRL, t = lane % kRL;
    const double n = (double)n_sel;
    const int n_blk = (n_sel + kRL * kItemAtoms - 1) / (kRL * kItemAtoms);
    float bx[kRing][kItemAtoms], by[kRing][kItemAtoms], bz[kRing][kItemAtoms];
    float qx[kRing][kItemAtoms], qy[kRing][kItemAtoms], qz[kRing][kItemAtoms];  // kRefRing
    long long ra = sa, rb = sb;
    // the next chunk of the workgroup's shared frames (empty when none is left)
    auto take_chunk = [&]() {
        unsigned c = 0;
        if (lane == 0) c = atomicAdd(s_ticket, 1u);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, kWave));
        ra = da + (long long)c * chunk;
        rb = ra < db ? min(db, ra + chunk) : ra;
    };
    if (ra >= rb) take_chunk();
    if (ra < rb)
        prime_ring<kAxis, kGather, kRL, kRefRing>(
            xyz, ra, (int)min((long long)kWave, rb - ra), frame_stride, rowstride, ix, n_sel, rx,
            ry, rz, n_blk, row, t, bx, by, bz, qx, qy, qz);
    int k0 = 0;           // slots of the group filled so far
    long long my_f = 0;   // frame of this lane's slot
    while (ra < rb) {
        const long long base = ra;
        const int cnt = (int)min((long long)(kWave - k0), rb - ra);
        double* s_dst = s_res + k0 * kNS;
        const int n_items = (cnt + kRows - 1) / kRows * n_blk;
        Sums s = {};
        for (int it0 = 0; it0 < n_items; it0 += kRing) {
#pragma unroll
            for (int r = 0; r < kRing; ++r) {
                const int it = it0 + r;
                if (it < n_items) {
                    // refill the slot freed one step ago with item it + kRing - 1
                    const int nxt = it + kRing - 1;
                    if (nxt < n_items) {
                        const int q = nxt / n_blk, b = nxt - q * n_blk;
                        const int k = kRows * q + row;
                        const int slot = (r + kRing - 1) % kRing;  // static once unrolled
                        load_item<kAxis, kGather, kRL>(
                            xyz + (base + (k < cnt ? k : 0)) * frame_stride, rowstride, ix,
                            n_sel, t, b, bx[slot], by[slot], bz[slot]);
                        if constexpr (kRefRing)
                            load_ref_item<kRL>(rx, ry, rz, n_sel, t, b, qx[slot], qy[slot],
                                               qz[slot]);
                    }
                    const int q = it / n_blk, b = it - q * n_blk;
                    const int j0 = b * kRL * kItemAtoms + t;
                    const bool row_ok = kRows * q + row < cnt;
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int j = j0 + u * kRL;
                        const bool ok = row_ok && j < n_sel;
                        const int jc = min(j, n_sel - 1);
                        constexpr bool kRef = kKind == Kind::NoSuperpose;
                        const float ux = kRefRing ? qx[r][u] : rx[jc];
                        const float uy = kRefRing ? qy[r][u] : ry[jc];
                        const float uz = kRefRing ? qz[r][u] : rz[jc];
                        accumulate<kKind>(s, valid_atom(ok, bx[r][u]), valid_atom(ok, by[r][u]),
                                          valid_atom(ok, bz[r][u]),
                                          kRef ? valid_atom(ok, ux) : ux,
                                          kRef ? valid_atom(ok, uy) : uy,
                                          kRef ? valid_atom(ok, uz) : uz);
                    }
                    if (b == n_blk - 1) {
                        row_reduce<kKind, kRL>(s);
                        const int k = kRows * q + row;
                        if (t == 0 && k < cnt) store_sums<kKind>(s_dst + k * kNS, s);
                        s = Sums{};
                    }
                }
            }
        }
        if (lane >= k0 && lane < k0 + cnt) my_f = base + (lane - k0);
        k0 += cnt;
        ra += cnt;
        if (ra >= rb) take_chunk();
        const bool more = ra < rb;
        if (k0 == kWave || !more) {
            // the group is full (or the last): its first loads, then its finish
            const int ncnt = (int)min((long long)kWave, rb - ra);
            if (kEarly && more)
                prime_ring<kAxis, kGather, kRL, kRefRing>(
                    xyz, ra, ncnt, frame_stride, rowstride, ix, n_sel, rx, ry, rz, n_blk, row, t,
                    bx, by, bz, qx, qy, qz);
            finish_group_at<kKind>(my_f, k0, n, rs, self_frame, out, rot_out, mu_out, s_res,
                                   lane);
            if (!kEarly && more)
                prime_ring<kAxis, kGather, kRL, kRefRing>(
                    xyz, ra, ncnt, frame_stride, rowstride, ix, n_sel, rx, ry, rz, n_blk, row, t,
                    bx, by, bz, qx, qy, qz);
            k0 = 0;
        } else {
            prime_ring<kAxis, kGather, kRL, kRefRing>(
                xyz, ra, (int)min((long long)(kWave - k0), rb - ra), frame_stride, rowstride, ix,
                n_sel, rx, ry, rz, n_blk, row, t, bx, by, bz, qx, qy, qz);
        }
    }
}

// process_chunk's item ring for RMSDs of whole frames (atom-major, no
// gather), with each group's frame sums kept in registers instead of parked
// in LDS: after frame set q's row reduction, lane k of the wave takes frame
// base + k's sums from the first lane of row k mod kRows (13 doubles by
// permutes), so lane k finishes frame base + k from its own registers.  The
// workgroup's LDS then holds only the reference, which lets the batched
// kernel stage references of up to ~4,500 atoms at 3 waves/SIMD.  The
// reference is read at accumulation time from LDS (planes lx, ly, lz).
// kPartial: only atoms below n_lds are in LDS (a reference larger than the
// workgroup's share), and items past them read gx, gy, gz in global memory,
// where each read waits for the frame loads issued before it (a reference
// riding the load ring as in process_chunk would spill registers beside the
// sums kept here).
template <int kRL>
__device__ __forceinline__ void prime_ring_regs(const float* __restrict__ xyz, long long base,
                                                int cnt, int n_sel, int n_blk, int row, int t,
                                                float (&bx)[kRing][kItemAtoms],
                                                float (&by)[kRing][kItemAtoms],
                                                float (&bz)[kRing][kItemAtoms]) {
    constexpr int kRows = kWave / kRL;
    const int n_items = (cnt + kRows - 1) / kRows * n_blk;
#pragma unroll
    for (int r = 0; r < kRing - 1; ++r) {
        const int rc = min(r, n_items - 1);
        const int q = rc / n_blk, b = rc - q * n_blk;
        const int k = kRows * q + row;
        load_item<false, false, kRL, true>(xyz + (base + (k < cnt ? k : 0)) * 3LL * n_sel, 0,
                                           nullptr, n_sel, t, b, bx[r], by[r], bz[r]);
    }
}

// Issue priority for workgroups that share a CU.  The SIMD arbiter issues
// the oldest ready wave first, so workgroups given equal shares of work
// finish in bands ordered by age, and the CU runs its last part with fewer
// waves to cover the load latency.  A wave of such a workgroup steps its
// priority through 0, 1, 2 every kPrioSteps ring rounds, starting from its
// workgroup's slot on the CU, so that the workgroups take turns at the
// front: on the 512-system mix, 0.4285 ms against 0.440 ms (same bits).
constexpr int kPrioSteps = 4;

// The workgroup's slot on its CU: the TG_ID field (bits 16-19) of HW_ID.
__device__ __forceinline__ int cu_slot() {
    return (int)(__builtin_amdgcn_s_getreg((3 << 11) | (16 << 6) | 4) & 15);
}

__device__ __forceinline__ void rotate_priority(int step, int slot) {
    const int phase = (step / kPrioSteps + slot) % 3;
    if (phase == 0)
        __builtin_amdgcn_s_setprio(0);
    else if (phase == 1)
        __builtin_amdgcn_s_setprio(1);
    else
        __builtin_amdgcn_s_setprio(2);
}

// The item loop runs a multiple of kRing steps with no branch around a load
// (as run_group_batch): every step refills the slot freed one step ago, an
// item past the group's end reloading the last one, and steps past the last
// item accumulate zeros.  Loads under a branch are merged at its join by
// register copies that wait for them at once, which leaves one item in
// flight instead of kRing - 1.
template <int kRL, bool kPartial, typename Stage>
__device__ __forceinline__ void process_chunk_regs(
    const float* __restrict__ xyz, long long f_begin, long long f_end, int n_sel,
    const float* __restrict__ lx, const float* __restrict__ ly, const float* __restrict__ lz,
    int n_lds, const GlobalFloat* __restrict__ gx, const GlobalFloat* __restrict__ gy,
    const GlobalFloat* __restrict__ gz, const double (&rs)[4], long long self_frame,
    float* __restrict__ out, int lane, Stage&& stage) {
    constexpr int kRows = kWave / kRL;
    const int row = lane / kRL, t = lane % kRL;
    const long long frame_stride = 3LL * n_sel;
    const double n = (double)n_sel;
    const int n_blk = (n_sel + kRL * kItemAtoms - 1) / (kRL * kItemAtoms);
    float bx[kRing][kItemAtoms], by[kRing][kItemAtoms], bz[kRing][kItemAtoms];
    if (f_begin < f_end)
        prime_ring_regs<kRL>(xyz, f_begin, (int)min((long long)kWave, f_end - f_begin), n_sel,
                             n_blk, row, t, bx, by, bz);
    // the caller's staging of the reference (every wave, barriers included),
    // while the first frame loads are in flight
    stage();
    const int slot = cu_slot();
    int round = 0;
    for (long long base = f_begin; base < f_end; base += kWave) {
        const int cnt = (int)min((long long)kWave, f_end - base);
        const int n_items = (cnt + kRows - 1) / kRows * n_blk;
        const int n_steps = (n_items + kRing - 1) / kRing * kRing;
        Sums s = {}, mine = {};
        for (int it0 = 0; it0 < n_steps; it0 += kRing) {
            rotate_priority(round++, slot);
#pragma unroll
            for (int r = 0; r < kRing; ++r) {
                const int it = it0 + r;
                {
                    const int nxt = min(it + kRing - 1, n_items - 1);
                    const int q = nxt / n_blk, b = nxt - q * n_blk;
                    const int k = kRows * q + row;
                    const int slot = (r + kRing - 1) % kRing;
                    load_item<false, false, kRL, true>(
                        xyz + (base + (k < cnt ? k : 0)) * frame_stride, 0, nullptr, n_sel, t, b,
                        bx[slot], by[slot], bz[slot]);
                }
                const bool live = it < n_items;  // wave-uniform
                const int itc = live ? it : n_items - 1;
                const int q = itc / n_blk, b = itc - q * n_blk;
                const int j0 = b * kRL * kItemAtoms + t;
                const bool row_ok = live && kRows * q + row < cnt;
                // reference atoms of the item: from LDS, or past the staged
                // part (kPartial: atoms n_lds and up) from global memory
                float ux[kItemAtoms], uy[kItemAtoms], uz[kItemAtoms];
                const bool from_lds = !kPartial || (b + 1) * kRL * kItemAtoms <= n_lds;
                if (from_lds) {
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int jc = min(j0 + u * kRL, n_sel - 1);
                        ux[u] = lx[jc];
                        uy[u] = ly[jc];
                        uz[u] = lz[jc];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int jc = min(j0 + u * kRL, n_sel - 1);
                        ux[u] = gx[jc];
                        uy[u] = gy[jc];
                        uz[u] = gz[jc];
                    }
                }
#pragma unroll
                for (int u = 0; u < kItemAtoms; ++u) {
                    const int j = j0 + u * kRL;
                    const bool ok = row_ok && j < n_sel;
                    accumulate<Kind::Rmsd>(s, valid_atom(ok, bx[r][u]), valid_atom(ok, by[r][u]),
                                           valid_atom(ok, bz[r][u]), ux[u], uy[u], uz[u]);
                }
                if (live && b == n_blk - 1) {  // wave-uniform: the set's last block
                    row_reduce<Kind::Rmsd, kRL>(s);
                    const int src = (lane % kRows) * kRL;
                    Sums got;
                    got.q = __shfl(s.q, src, kWave);
                    got.sx = __shfl(s.sx, src, kWave);
                    got.sy = __shfl(s.sy, src, kWave);
                    got.sz = __shfl(s.sz, src, kWave);
#pragma unroll
                    for (int e = 0; e < 9; ++e) got.p[e] = __shfl(s.p[e], src, kWave);
                    take<Kind::Rmsd>(mine, got, lane / kRows == q);
                    s = Sums{};
                }
            }
        }
        // the next group's first loads go out before this group's QCP solves
        // (unconditional: past the range the last group is loaded again)
        const long long next = base + kWave;
        const long long nb = next < f_end ? next : base;
        prime_ring_regs<kRL>(xyz, nb, (int)min((long long)kWave, f_end - nb), n_sel, n_blk, row,
                             t, bx, by, bz);
        if (lane < cnt) {
            const long long f = base + lane;
            double M[9], ga, gb;
            centred(mine, rs, n, M, ga, gb);
            const double msd = qcp_msd(M, ga, gb, n);
            global_out(out)[f] = (f == self_frame) ? 0.0f : (float)sqrt(msd);
        }
    }
}

// The batched per-CU workgroup's share of one system: process_chunk_regs's
// branch-free item loop and register sums, with the frames [sa, sb) first
// and then chunks of `chunk` frames from [da, db) by the LDS ticket
// *s_ticket, shared by the workgroup's waves (so that the waves the SIMD
// issues less often take fewer frames).  Chunks fill the wave's group of 64
// frame slots in turn, lane k keeping the index of the frame in slot k.
template <int kRL, bool kPartial, typename Stage>
__device__ __forceinline__ void process_segment_cu(
    const float* __restrict__ xyz, long long sa, long long sb, long long da, long long db,
    long long chunk, unsigned* s_ticket, int n_sel, const float* __restrict__ lx,
    const float* __restrict__ ly, const float* __restrict__ lz, int n_lds,
    const GlobalFloat* __restrict__ gx, const GlobalFloat* __restrict__ gy,
    const GlobalFloat* __restrict__ gz, const double (&rs)[4], long long self_frame,
    float* __restrict__ out, int lane, Stage&& stage) {
    constexpr int kRows = kWave / kRL;
    const int row = lane / kRL, t = lane % kRL;
    const long long frame_stride = 3LL * n_sel;
    const double n = (double)n_sel;
    const int n_blk = (n_sel + kRL * kItemAtoms - 1) / (kRL * kItemAtoms);
    float bx[kRing][kItemAtoms], by[kRing][kItemAtoms], bz[kRing][kItemAtoms];
    long long ra = sa, rb = sb;
    auto take_chunk = [&]() {
        unsigned c = 0;
        if (lane == 0) c = atomicAdd(s_ticket, 1u);
        c = __builtin_amdgcn_readfirstlane(__shfl(c, 0, kWave));
        ra = da + (long long)c * chunk;
        rb = ra < db ? min(db, ra + chunk) : ra;
    };
    // the static range's first loads go out before the caller stages the
    // reference (barriers included; it also clears the ticket)
    if (ra < rb)
        prime_ring_regs<kRL>(xyz, ra, (int)min((long long)kWave, rb - ra), n_sel, n_blk, row, t,
                             bx, by, bz);
    stage();
    if (ra >= rb) {
        take_chunk();
        if (ra >= rb) return;
        prime_ring_regs<kRL>(xyz, ra, (int)min((long long)kWave, rb - ra), n_sel, n_blk, row, t,
                             bx, by, bz);
    }
    int k0 = 0;          // slots of the group filled so far
    long long my_f = 0;  // frame of this lane's slot
    Sums mine = {};
    while (ra < rb) {
        const long long base = ra;
        const int cnt = (int)min((long long)(kWave - k0), rb - ra);
        const int n_items = (cnt + kRows - 1) / kRows * n_blk;
        const int n_steps = (n_items + kRing - 1) / kRing * kRing;
        // lanes k0 + kRows q .. take set q's sums from row (lane - k0) mod kRows
        const int rel = lane - k0;
        const int src = ((rel > 0 ? rel : 0) % kRows) * kRL;
        Sums s = {};
        for (int it0 = 0; it0 < n_steps; it0 += kRing) {
#pragma unroll
            for (int r = 0; r < kRing; ++r) {
                const int it = it0 + r;
                {
                    const int nxt = min(it + kRing - 1, n_items - 1);
                    const int q = nxt / n_blk, b = nxt - q * n_blk;
                    const int k = kRows * q + row;
                    const int slot = (r + kRing - 1) % kRing;
                    load_item<false, false, kRL, true>(
                        xyz + (base + (k < cnt ? k : 0)) * frame_stride, 0, nullptr, n_sel, t, b,
                        bx[slot], by[slot], bz[slot]);
                }
                const bool live = it < n_items;  // wave-uniform
                const int itc = live ? it : n_items - 1;
                const int q = itc / n_blk, b = itc - q * n_blk;
                const int j0 = b * kRL * kItemAtoms + t;
                const bool row_ok = live && kRows * q + row < cnt;
                float ux[kItemAtoms], uy[kItemAtoms], uz[kItemAtoms];
                const bool from_lds = !kPartial || (b + 1) * kRL * kItemAtoms <= n_lds;
                if (from_lds) {
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int jc = min(j0 + u * kRL, n_sel - 1);
                        ux[u] = lx[jc];
                        uy[u] = ly[jc];
                        uz[u] = lz[jc];
                    }
                } else {
#pragma unroll
                    for (int u = 0; u < kItemAtoms; ++u) {
                        const int jc = min(j0 + u * kRL, n_sel - 1);
                        ux[u] = gx[jc];
                        uy[u] = gy[jc];
                        uz[u] = gz[jc];
                    }
                }
#pragma unroll
                for (int u = 0; u < kItemAtoms; ++u) {
                    const int j = j0 + u * kRL;
                    const bool ok = row_ok && j < n_sel;
                    accumulate<Kind::Rmsd>(s, valid_atom(ok, bx[r][u]), valid_atom(ok, by[r][u]),
                                           valid_atom(ok, bz[r][u]), ux[u], uy[u], uz[u]);
                }
                if (live && b == n_blk - 1) {  // wave-uniform: the set's last block
                    row_reduce<Kind::Rmsd, kRL>(s);
                    Sums got;
                    got.q = __shfl(s.q, src, kWave);
                    got.sx = __shfl(s.sx, src, kWave);
                    got.sy = __shfl(s.sy, src, kWave);
                    got.sz = __shfl(s.sz, src, kWave);
#pragma unroll
                    for (int e = 0; e < 9; ++e) got.p[e] = __shfl(s.p[e], src, kWave);
                    take<Kind::Rmsd>(mine, got, rel >= kRows * q && rel < kRows * q + kRows);
                    s = Sums{};
                }
            }
        }
        if (rel >= 0 && rel < cnt) my_f = base + rel;
        k0 += cnt;
        ra += cnt;
        if (ra >= rb) take_chunk();
        const bool more = ra < rb;
        const bool full = k0 == kWave || !more;
        // the next items' first loads before the QCP solves (unconditional:
        // with no more work the last group is loaded again and not used)
        const long long pb = more ? ra : base;
        const int pc = more ? (int)min((long long)(full ? kWave : kWave - k0), rb - ra) : cnt;
        prime_ring_regs<kRL>(xyz, pb, pc, n_sel, n_blk, row, t, bx, by, bz);
        if (full) {
            if (lane < k0) {
                double M[9], ga, gb;
                centred(mine, rs, n, M, ga, gb);
                const double msd = qcp_msd(M, ga, gb, n);
                global_out(out)[my_f] = (my_f == self_frame) ? 0.0f : (float)sqrt(msd);
            }
            k0 = 0;
            mine = Sums{};
        }
    }
}

// Batched variant of the item loop (one system's frames, all atoms, no
// gather).  Each wave works on its own system, so the reference is not
// staged in LDS up front: every item's block of 64 reference atoms rides in
// the load ring with the frame data (lane l loads atom 64 b + l of the three
// planes, 2 items ahead) and each lane fetches the four atoms it uses from
// the lanes that loaded them.  Same arithmetic as process_chunk.
constexpr int kRingB = 3;

__device__ __forceinline__ void load_item_batch(const float* __restrict__ fr,
                                                int n, int t, int b, int lane,
                                                const float* __restrict__ rx,
                                                const float* __restrict__ ry,
                                                const float* __restrict__ rz,
                                                float (&ax)[kItemAtoms], float (&ay)[kItemAtoms],
                                                float (&az)[kItemAtoms], float& qx, float& qy,
                                                float& qz) {
    load_item<false, false, kRowLanes, true>(fr, 0, nullptr, n, t, b, ax, ay, az);
    const int jr = min(b * kRowLanes * kItemAtoms + lane, n - 1);
    qx = global_ptr(rx)[jr];
    qy = global_ptr(ry)[jr];
    qz = global_ptr(rz)[jr];
}

// Loads of the first kRingB - 1 items of a batched group into ring slots
// 0 .. kRingB - 2.  Unconditional: an item past the group's end loads the
// last item again (cached) and is never accumulated.  A load under a branch
// would be merged at the join by a register copy, which waits for it at once.
__device__ __forceinline__ void prime_ring_batch(
    const float* __restrict__ xyz, long long base, int cnt, int n_atoms,
    const float* __restrict__ rx, const float* __restrict__ ry, const float* __restrict__ rz,
    int lane, float (&bx)[kRingB][kItemAtoms], float (&by)[kRingB][kItemAtoms],
    float (&bz)[kRingB][kItemAtoms], float (&qx)[kRingB], float (&qy)[kRingB],
    float (&qz)[kRingB]) {
    constexpr int kBlk = kRowLanes * kItemAtoms;
    const int row = lane / kRowLanes, t = lane % kRowLanes;
    const long long frame_stride = 3LL * n_atoms;
    const int n_blk = (n_atoms + kBlk - 1) / kBlk;
    const int n_items = (cnt + 3) / 4 * n_blk;
#pragma unroll
    for (int r = 0; r < kRingB - 1; ++r) {
        const int rc = min(r, n_items - 1);
        const int q = rc / n_blk, b = rc - q * n_blk;
        const int k = 4 * q + row;
        load_item_batch(xyz + (base + (k < cnt ? k : 0)) * frame_stride, n_atoms, t, b, lane, rx,
                        ry, rz, bx[r], by[r], bz[r], qx[r], qy[r], qz[r]);
    }
}

// The items of one batched group (ring primed by prime_ring_batch); on
// return the group's row sums are parked in s_res.  The item loop runs a
// multiple of kRingB steps with no branch around a load: every step refills
// the slot freed one step ago (an item past the end reloads the last one),
// and steps past the last item accumulate zeros.  So each slot's loads stay
// in flight for kRingB - 1 steps.  The ring holds loads of the group's last
// item when it returns.
__device__ __forceinline__ void run_group_batch(
    const float* __restrict__ xyz, long long base, int cnt, int n_atoms,
    const float* __restrict__ rx, const float* __restrict__ ry, const float* __restrict__ rz,
    double* __restrict__ s_res, int lane, float (&bx)[kRingB][kItemAtoms],
    float (&by)[kRingB][kItemAtoms], float (&bz)[kRingB][kItemAtoms], float (&qx)[kRingB],
    float (&qy)[kRingB], float (&qz)[kRingB]) {
    constexpr int kNS = SumsPerFrame<Kind::Rmsd>::value;
    constexpr int kBlk = kRowLanes * kItemAtoms;  // atoms per item block
    const int row = lane / kRowLanes, t = lane % kRowLanes;
    const long long frame_stride = 3LL * n_atoms;
    const int n_blk = (n_atoms + kBlk - 1) / kBlk;
    const int n_items = (cnt + 3) / 4 * n_blk;
    const int n_steps = (n_items + kRingB - 1) / kRingB * kRingB;
    Sums s = {};
    for (int it0 = 0; it0 < n_steps; it0 += kRingB) {
#pragma unroll
        for (int r = 0; r < kRingB; ++r) {
            const int it = it0 + r;
            {
                const int nxt = min(it + kRingB - 1, n_items - 1);
                const int q = nxt / n_blk, b = nxt - q * n_blk;
                const int k = 4 * q + row;
                const int slot = (r + kRingB - 1) % kRingB;
                load_item_batch(xyz + (base + (k < cnt ? k : 0)) * frame_stride, n_atoms, t, b,
                                lane, rx, ry, rz, bx[slot], by[slot], bz[slot], qx[slot],
                                qy[slot], qz[slot]);
            }
            // atom 64 b + t + 16 u of the reference was loaded by lane
            // 16 u + t: fetch it from there (cross-lane permute)
            const bool live = it < n_items;  // wave-uniform
            const int itc = live ? it : n_items - 1;
            const int q = itc / n_blk, b = itc - q * n_blk;
            const bool row_ok = live && 4 * q + row < cnt;
#pragma unroll
            for (int u = 0; u < kItemAtoms; ++u) {
                const int src = u * kRowLanes + t;
                const bool ok = row_ok && b * kBlk + src < n_atoms;
                accumulate<Kind::Rmsd>(s, valid_atom(ok, bx[r][u]), valid_atom(ok, by[r][u]),
                                       valid_atom(ok, bz[r][u]), __shfl(qx[r], src, kWave),
                                       __shfl(qy[r], src, kWave), __shfl(qz[r], src, kWave));
            }
            if (live && b == n_blk - 1) {
                row_reduce<Kind::Rmsd>(s);
                const int k = 4 * q + row;
                if (t == 0 && k < cnt) store_sums<Kind::Rmsd>(s_res + k * kNS, s);
                s = Sums{};
            }
        }
    }
}


}  // namespace
}  // namespace hr
