// Device code shared by the per-frame kernels (one-vs-all, superpose fit,
// no-superpose, batched systems): FP64 per-frame sums, the quarter-wave
// (16 lanes per frame) item loop with its register ring, DPP row reductions,
// and the per-frame finish (QCP or no-superpose) one lane per frame.
#pragma once

#include "common.h"

namespace hr {
namespace {

constexpr int kBlock = 256;
constexpr int kWavesPerBlock = kBlock / kWave;
constexpr int kDefaultWavesPerCu = 16;

// One-vs-all workspace: [0, 64) reference sums (double sx, sy, sz, q);
// [256, 256 + 12 n_sel) gathered reference, axis-major (x[], y[], z[]).
constexpr size_t kRefOffset = 256;

enum class Kind { Rmsd, NoSuperpose, Fit };

// Gather the reference selection into axis-major planes and sum it in FP64
// (one workgroup of kBlock threads).
__device__ __forceinline__ void ref_prep_body(const float* __restrict__ ref,
                                              const int* __restrict__ ref_idx, int n_sel,
                                              int axis_rowstride, float* __restrict__ ref_out,
                                              double* __restrict__ stats) {
    double sx = 0.0, sy = 0.0, sz = 0.0, q = 0.0;
    for (int j = threadIdx.x; j < n_sel; j += kBlock) {
        float x, y, z;
        if (axis_rowstride > 0) {
            x = ref[j];
            y = ref[axis_rowstride + j];
            z = ref[2 * axis_rowstride + j];
        } else {
            const long long a = ref_idx ? ref_idx[j] : j;
            x = ref[3 * a];
            y = ref[3 * a + 1];
            z = ref[3 * a + 2];
        }
        ref_out[j] = x;
        ref_out[n_sel + j] = y;
        ref_out[2 * n_sel + j] = z;
        const double dx = x, dy = y, dz = z;
        sx += dx;
        sy += dy;
        sz += dz;
        q = fma(dx, dx, fma(dy, dy, fma(dz, dz, q)));
    }
    sx = wave_sum(sx);
    sy = wave_sum(sy);
    sz = wave_sum(sz);
    q = wave_sum(q);
    __shared__ double part[kWavesPerBlock][4];
    const int lane = threadIdx.x & (kWave - 1);
    const int w = threadIdx.x / kWave;
    if (lane == 0) {
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

// LDS bytes of the per-wave result area (64 frames of sums per wave).
template <Kind kKind>
constexpr size_t result_lds_bytes() {
    return (size_t)kWavesPerBlock * kWave * SumsPerFrame<kKind>::value * sizeof(double);
}

// Quarter-wave layout: 16 lanes own one frame, a wave works on 4 frames at
// once (row r = lane / 16 holds frame base + r).  An item is (frame quad q,
// atom block b): lane t of a row takes the selected atoms
// b * 16 * kItemAtoms + t + 16 u, u < kItemAtoms.
constexpr int kRowLanes = 16;
constexpr int kItemAtoms = 4;   // atoms per lane per item
constexpr int kRing = 4;        // items in the register ring (kRing - 1 in flight)

// Frame coordinates are read once per launch: nontemporal loads (one
// global_load_dwordx3 ... nt per atom) keep them from displacing reusable
// lines in the caches; on MI355X they stream 2-5% faster than plain loads
// (100,000 frames x 2,000 atoms: 0.419 ms against 0.426-0.442 ms).
template <bool kAxis, bool kGather>
__device__ __forceinline__ void load_item(const float* __restrict__ fr, bool row_ok,
                                          int rowstride, const int* __restrict__ ix, int n_sel,
                                          int t, int b, float (&ax)[kItemAtoms],
                                          float (&ay)[kItemAtoms], float (&az)[kItemAtoms]) {
    const int j0 = b * kRowLanes * kItemAtoms + t;
#pragma unroll
    for (int u = 0; u < kItemAtoms; ++u) {
        const int j = j0 + u * kRowLanes;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (row_ok && j < n_sel) {
            if constexpr (kAxis) {
                x = __builtin_nontemporal_load(fr + j);
                y = __builtin_nontemporal_load(fr + rowstride + j);
                z = __builtin_nontemporal_load(fr + 2 * rowstride + j);
            } else {
                const int a = kGather ? ix[j] : j;
                x = __builtin_nontemporal_load(fr + 3 * a);
                y = __builtin_nontemporal_load(fr + 3 * a + 1);
                z = __builtin_nontemporal_load(fr + 3 * a + 2);
            }
        }
        ax[u] = x;
        ay[u] = y;
        az[u] = z;
    }
}

template <Kind kKind>
__device__ __forceinline__ void row_reduce(Sums& s) {
    s.q = row16_sum(s.q);
    if constexpr (kKind != Kind::NoSuperpose) {
        s.sx = row16_sum(s.sx);
        s.sy = row16_sum(s.sy);
        s.sz = row16_sum(s.sz);
#pragma unroll
        for (int e = 0; e < 9; ++e) s.p[e] = row16_sum(s.p[e]);
    }
}

// Finish a group of cnt frames whose row sums are parked in s_res: one lane
// per frame computes the QCP MSD (or the fit, or the no-superpose MSD).
template <Kind kKind>
__device__ __forceinline__ void finish_group(long long base, int cnt, double n,
                                             const double (&rs)[4], long long self_frame,
                                             float* __restrict__ out, float* __restrict__ rot_out,
                                             double* __restrict__ mu_out,
                                             double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    // the row leaders' LDS stores must be visible to every lane of the wave
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane < cnt) {
        const long long f = base + lane;
        Sums mine;
        load_sums<kKind>(mine, s_res + lane * kNS);
        if constexpr (kKind == Kind::NoSuperpose) {
            out[f] = (float)sqrt(mine.q / n);
        } else {
            double M[9], ga, gb;
            centred(mine, rs, n, M, ga, gb);
            if constexpr (kKind == Kind::Rmsd) {
                const double msd = qcp_msd(M, ga, gb, n);
                out[f] = (f == self_frame) ? 0.0f : (float)sqrt(msd);
            } else {
                double R[9];
                const double lam = qcp_solve<true>(M, 0.5 * (ga + gb), R);
#pragma unroll
                for (int e = 0; e < 9; ++e) rot_out[9 * f + e] = (float)R[e];
                mu_out[3 * f + 0] = mine.sx / n;
                mu_out[3 * f + 1] = mine.sy / n;
                mu_out[3 * f + 2] = mine.sz / n;
                if (out) {
                    const double msd = (ga + gb - 2.0 * lam) / n;
                    out[f] = (float)sqrt(msd > 0.0 ? msd : 0.0);
                }
            }
        }
    }
    // the next group overwrites s_res: all lanes must have read it
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One group of up to 64 consecutive frames (base .. base + cnt) of one
// trajectory through the wave: items (frame quad, atom block) stream through
// the register ring, each quad's row sums are reduced (DPP) and parked in the
// wave's LDS slots, then one lane per frame finishes (QCP or no-superpose).
template <Kind kKind, bool kAxis, bool kGather>
__device__ __forceinline__ void process_group(
    const float* __restrict__ xyz, long long base, int cnt, long long frame_stride,
    int rowstride, const int* __restrict__ ix, int n_sel, const float* __restrict__ rx,
    const float* __restrict__ ry, const float* __restrict__ rz, const double (&rs)[4],
    long long self_frame, float* __restrict__ out, float* __restrict__ rot_out,
    double* __restrict__ mu_out, double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<kKind>::value;
    const int row = lane / kRowLanes, t = lane % kRowLanes;
    const double n = (double)n_sel;
    const int n_blk = (n_sel + kRowLanes * kItemAtoms - 1) / (kRowLanes * kItemAtoms);
    const int n_quads = (cnt + 3) / 4;
    const int n_items = n_quads * n_blk;
    // item -> (quad, block); loads of `item` into ring slot `slot`
    float bx[kRing][kItemAtoms], by[kRing][kItemAtoms], bz[kRing][kItemAtoms];
#pragma unroll
    for (int r = 0; r < kRing - 1; ++r) {
        if (r < n_items) {
            const int q = r / n_blk, b = r - q * n_blk;
            const int k = 4 * q + row;
            load_item<kAxis, kGather>(xyz + (base + k) * frame_stride, k < cnt, rowstride, ix,
                                      n_sel, t, b, bx[r], by[r], bz[r]);
        }
    }
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
                    const int k = 4 * q + row;
                    const int slot = (r + kRing - 1) % kRing;  // static once unrolled
                    load_item<kAxis, kGather>(xyz + (base + k) * frame_stride, k < cnt,
                                              rowstride, ix, n_sel, t, b, bx[slot], by[slot],
                                              bz[slot]);
                }
                const int q = it / n_blk, b = it - q * n_blk;
                const int j0 = b * kRowLanes * kItemAtoms + t;
#pragma unroll
                for (int u = 0; u < kItemAtoms; ++u) {
                    const int j = j0 + u * kRowLanes;
                    float ux = 0.0f, uy = 0.0f, uz = 0.0f;
                    if (j < n_sel) {
                        ux = rx[j];
                        uy = ry[j];
                        uz = rz[j];
                    }
                    accumulate<kKind>(s, bx[r][u], by[r][u], bz[r][u], ux, uy, uz);
                }
                if (b == n_blk - 1) {
                    row_reduce<kKind>(s);
                    const int k = 4 * q + row;
                    if (t == 0 && k < cnt) store_sums<kKind>(s_res + k * kNS, s);
                    s = Sums{};
                }
            }
        }
    }
    finish_group<kKind>(base, cnt, n, rs, self_frame, out, rot_out, mu_out, s_res, lane);
}

// Batched variant of the item loop (one system's frames, all atoms, no
// gather).  Each wave works on its own system, so the reference is not
// staged in LDS up front: every item's block of 64 reference atoms rides in
// the load ring with the frame data (lane l loads atom 64 b + l of the three
// planes, 2 items ahead) and each lane fetches the four atoms it uses from
// the lanes that loaded them.  Same arithmetic as process_group.
constexpr int kRingB = 3;

__device__ __forceinline__ void load_item_batch(const float* __restrict__ fr, bool row_ok,
                                                int n, int t, int b, int lane,
                                                const float* __restrict__ rx,
                                                const float* __restrict__ ry,
                                                const float* __restrict__ rz,
                                                float (&ax)[kItemAtoms], float (&ay)[kItemAtoms],
                                                float (&az)[kItemAtoms], float& qx, float& qy,
                                                float& qz) {
    load_item<false, false>(fr, row_ok, 0, nullptr, n, t, b, ax, ay, az);
    const int jr = b * kRowLanes * kItemAtoms + lane;
    qx = 0.0f;
    qy = 0.0f;
    qz = 0.0f;
    if (jr < n) {
        qx = rx[jr];
        qy = ry[jr];
        qz = rz[jr];
    }
}

__device__ __forceinline__ void process_group_batch(
    const float* __restrict__ xyz, long long base, int cnt, int n_atoms,
    const float* __restrict__ rx, const float* __restrict__ ry, const float* __restrict__ rz,
    const double (&rs)[4], long long self_frame, float* __restrict__ out,
    double* __restrict__ s_res, int lane) {
    constexpr int kNS = SumsPerFrame<Kind::Rmsd>::value;
    constexpr int kBlk = kRowLanes * kItemAtoms;  // atoms per item block
    const int row = lane / kRowLanes, t = lane % kRowLanes;
    const long long frame_stride = 3LL * n_atoms;
    const int n_blk = (n_atoms + kBlk - 1) / kBlk;
    const int n_quads = (cnt + 3) / 4;
    const int n_items = n_quads * n_blk;
    float bx[kRingB][kItemAtoms], by[kRingB][kItemAtoms], bz[kRingB][kItemAtoms];
    float qx[kRingB], qy[kRingB], qz[kRingB];
#pragma unroll
    for (int r = 0; r < kRingB - 1; ++r) {
        if (r < n_items) {
            const int q = r / n_blk, b = r - q * n_blk;
            const int k = 4 * q + row;
            load_item_batch(xyz + (base + k) * frame_stride, k < cnt, n_atoms, t, b, lane, rx,
                            ry, rz, bx[r], by[r], bz[r], qx[r], qy[r], qz[r]);
        }
    }
    Sums s = {};
    for (int it0 = 0; it0 < n_items; it0 += kRingB) {
#pragma unroll
        for (int r = 0; r < kRingB; ++r) {
            const int it = it0 + r;
            if (it < n_items) {
                const int nxt = it + kRingB - 1;
                if (nxt < n_items) {
                    const int q = nxt / n_blk, b = nxt - q * n_blk;
                    const int k = 4 * q + row;
                    const int slot = (r + kRingB - 1) % kRingB;
                    load_item_batch(xyz + (base + k) * frame_stride, k < cnt, n_atoms, t, b, lane,
                                    rx, ry, rz, bx[slot], by[slot], bz[slot], qx[slot], qy[slot],
                                    qz[slot]);
                }
                // atom 64 b + t + 16 u of the reference was loaded by lane
                // 16 u + t: fetch it from there (cross-lane permute)
                const int q = it / n_blk, b = it - q * n_blk;
#pragma unroll
                for (int u = 0; u < kItemAtoms; ++u) {
                    const int src = u * kRowLanes + t;
                    accumulate<Kind::Rmsd>(s, bx[r][u], by[r][u], bz[r][u],
                                           __shfl(qx[r], src, kWave), __shfl(qy[r], src, kWave),
                                           __shfl(qz[r], src, kWave));
                }
                if (b == n_blk - 1) {
                    row_reduce<Kind::Rmsd>(s);
                    const int k = 4 * q + row;
                    if (t == 0 && k < cnt) store_sums<Kind::Rmsd>(s_res + k * kNS, s);
                    s = Sums{};
                }
            }
        }
    }
    finish_group<Kind::Rmsd>(base, cnt, (double)n_atoms, rs, self_frame, out, nullptr, nullptr,
                             s_res, lane);
}


}  // namespace
}  // namespace hr
