// LP-RMSD (gfx950): the RMSD minimised over rotations and over relabellings
// of atoms inside permute groups (rmsd/_lprmsd.pyx:72-228).  Per target frame:
//   (1) gather the selection and centre it (FP64 centroid, float coordinates
//       rounded once);
//   (2) if some selected atoms are in no group ("distinguishable"), fit them
//       alone, each set centred on its own centroid, and rotate the whole
//       centred selection by that rotation, x' = x R1 in float as
//       rot_atom_major;
//   (3) for each group, the assignment of reference atoms (rows) to target
//       atoms (columns) minimising the summed squared distances, the costs
//       formed as euclidean_permutation.cpp:28-35 does (float differences
//       and squares, summed in double), solved exactly by the Hungarian
//       method in its shortest-augmenting-path form (row potentials u,
//       column potentials v, one Dijkstra-like search per row);
//   (4) the QCP RMSD of the relabelled selection against the reference; with
//       superpose, the frame (every atom, centred) is rotated by R1 R2.
// One wave per frame: the columns of a search are spread over the lanes,
// the search's minimum is a wave reduction (ties to the lowest column), and
// the group's reference coordinates, potentials and search state sit in
// LDS (sized to keep eight waves per CU for groups of a few hundred atoms).
#include "common.h"

namespace hr {
namespace {

constexpr int kLpBlock = kWave;  // one wave per frame
constexpr size_t kLpLdsMax = 64 * 1024;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

__host__ __device__ inline size_t lp_lds_bytes(int n_sel, int kmax) {
    const size_t k1 = (size_t)kmax + 1;
    return 3 * 8 * k1      // u, v, minv
           + 2 * 4 * k1    // p, way
           + 4 * (size_t)kmax  // the group's atoms
           + 12 * (size_t)kmax  // the group's reference coordinates
           + 12 * (size_t)n_sel  // target selection, x, y, z
           + k1;           // used
}

// The workgroup is one wave, whose LDS accesses complete in program order:
// ordering them needs only a wavefront fence that keeps the compiler from
// moving them, not a workgroup barrier (which would also wait for memory).
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Index arrays live on the device, where the host cannot check them: every
// index is clamped into its array, so malformed input gives wrong numbers,
// never an access outside the buffers.
__device__ __forceinline__ int clamp_index(int i, int n) { return (unsigned)i < (unsigned)n ? i : 0; }

__device__ __forceinline__ int ref_atom(const int* idx, const int* dis, int part, int a, int n_sel,
                                        int n_atoms) {
    const int s = part == 0 ? a : clamp_index(dis[a], n_sel);
    return clamp_index(idx[s], n_atoms);
}

// Prepared reference: the selection centred (x, y, z planes of n_sel
// floats), the distinguishable subset centred on its own centroid (planes of
// n_dis), and their traces.
__global__ __launch_bounds__(kLpBlock) void lp_ref_kernel(const float* __restrict__ ref,
                                                          int n_atoms,
                                                          const int* __restrict__ idx, int n_sel,
                                                          const int* __restrict__ dis, int n_dis,
                                                          float* __restrict__ rsel,
                                                          float* __restrict__ rdis,
                                                          double* __restrict__ traces) {
    const int lane = threadIdx.x;
    for (int part = 0; part < 2; ++part) {
        const int n = part == 0 ? n_sel : n_dis;
        if (n == 0) {
            if (lane == 0) traces[part] = 0.0;
            continue;
        }
        float* out = part == 0 ? rsel : rdis;
        double s[3] = {0.0, 0.0, 0.0};
        for (int a = lane; a < n; a += kWave) {
            const int at = ref_atom(idx, dis, part, a, n_sel, n_atoms);
#pragma unroll
            for (int c = 0; c < 3; ++c) s[c] += (double)ref[3LL * at + c];
        }
        double mu[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) mu[c] = wave_sum(s[c]) / n;
        double g = 0.0;
        for (int a = lane; a < n; a += kWave) {
            const int at = ref_atom(idx, dis, part, a, n_sel, n_atoms);
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const float v = (float)((double)ref[3LL * at + c] - mu[c]);
                out[(size_t)c * n + a] = v;
                g += (double)v * v;
            }
        }
        g = wave_sum(g);
        if (lane == 0) traces[part] = g;
    }
}

struct LpArgs {
    float* xyz;  // (n_frames, n_atoms, 3); rotated in place with superpose
    long long n_frames;
    int n_atoms;
    const int* idx;
    int n_sel;
    const int* dis;  // selection positions of the distinguishable atoms
    int n_dis;
    const int* group_atoms;  // selection positions, group g at [off[g], off[g + 1])
    const int* group_off;
    int n_groups;
    int kmax;
    const float* rsel;  // prepared reference (lp_ref_kernel)
    const float* rdis;
    const double* traces;
    int superpose;
    float* out;
};

// (value, column) pairs ordered by value, then by the lower column; a NaN
// never beats a number, and any valid column beats none (INT_MAX).  The
// order is total on the pairs that occur, so any reduction tree gives the
// same minimum.
__device__ __forceinline__ void take_min(double& best, int& col, double ob, int oc) {
    if (ob < best || (!(ob > best) && oc < col)) {
        best = ob;
        col = oc;
    }
}

template <int kCtrl>
__device__ __forceinline__ void argmin_dpp(double& best, int& col) {
    const double ob = dpp_move<kCtrl>(best);
    const int oc = __builtin_amdgcn_mov_dpp(col, kCtrl, 0xF, 0xF, false);
    take_min(best, col, ob, oc);
}

// Minimum of the wave, in every lane: within rows of 16 lanes by DPP moves
// (xor 1, xor 2, half mirror, mirror: a few cycles each), then across rows
// by two lane exchanges.
__device__ __forceinline__ void wave_argmin(double& best, int& col) {
    argmin_dpp<0xB1>(best, col);
    argmin_dpp<0x4E>(best, col);
    argmin_dpp<0x141>(best, col);
    argmin_dpp<0x140>(best, col);
#pragma unroll
    for (int m = 16; m <= 32; m <<= 1) take_min(best, col, __shfl_xor(best, m, kWave),
                                                __shfl_xor(col, m, kWave));
}

__global__ __launch_bounds__(kLpBlock) void lp_frame_kernel(LpArgs A) {
    extern __shared__ __align__(16) unsigned char lds[];
    const int k1 = A.kmax + 1;
    double* u = reinterpret_cast<double*>(lds);
    double* v = u + k1;
    double* minv = v + k1;
    int* p = reinterpret_cast<int*>(minv + k1);
    int* way = p + k1;
    int* gat = way + k1;
    float* grx = reinterpret_cast<float*>(gat + A.kmax);  // the group's reference coordinates
    float* gry = grx + A.kmax;
    float* grz = gry + A.kmax;
    float* tx = grz + A.kmax;
    float* ty = tx + A.n_sel;
    float* tz = ty + A.n_sel;
    unsigned char* used = reinterpret_cast<unsigned char*>(tz + A.n_sel);

    const int lane = threadIdx.x;
    const long long f = blockIdx.x;
    float* fr = A.xyz + f * 3LL * A.n_atoms;
    const int n = A.n_sel;
    const float* rx = A.rsel;
    const float* ry = rx + n;
    const float* rz = ry + n;

    // (1) gather and centre the selection
    double s[3] = {0.0, 0.0, 0.0};
    for (int a = lane; a < n; a += kWave) {
        const long long at = clamp_index(A.idx[a], A.n_atoms);
        const float x = fr[3 * at], y = fr[3 * at + 1], z = fr[3 * at + 2];
        tx[a] = x;
        ty[a] = y;
        tz[a] = z;
        s[0] += x;
        s[1] += y;
        s[2] += z;
    }
    double mu[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) mu[c] = wave_sum(s[c]) / n;
    double tg = 0.0;
    for (int a = lane; a < n; a += kWave) {
        const float x = (float)((double)tx[a] - mu[0]);
        const float y = (float)((double)ty[a] - mu[1]);
        const float z = (float)((double)tz[a] - mu[2]);
        tx[a] = x;
        ty[a] = y;
        tz[a] = z;
        tg += (double)x * x + (double)y * y + (double)z * z;
    }
    tg = wave_sum(tg);
    __syncthreads();

    // (2) fit on the distinguishable atoms
    float R1[9] = {1.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 1.f};
    if (A.n_dis > 0) {
        const int nd = A.n_dis;
        double d[3] = {0.0, 0.0, 0.0};
        for (int k = lane; k < nd; k += kWave) {
            const int a = clamp_index(A.dis[k], n);
            d[0] += tx[a];
            d[1] += ty[a];
            d[2] += tz[a];
        }
        double dm[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) dm[c] = wave_sum(d[c]) / nd;
        double M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0}, ga = 0.0;
        const float* qx = A.rdis;
        const float* qy = qx + nd;
        const float* qz = qy + nd;
        for (int k = lane; k < nd; k += kWave) {
            const int a = clamp_index(A.dis[k], n);
            const double t[3] = {tx[a] - dm[0], ty[a] - dm[1], tz[a] - dm[2]};
            const double r[3] = {qx[k], qy[k], qz[k]};
#pragma unroll
            for (int i = 0; i < 3; ++i) {
#pragma unroll
                for (int j = 0; j < 3; ++j) M[3 * i + j] = fma(t[i], r[j], M[3 * i + j]);
                ga = fma(t[i], t[i], ga);
            }
        }
#pragma unroll
        for (int e = 0; e < 9; ++e) M[e] = wave_sum(M[e]);
        ga = wave_sum(ga);
        double R[9];
        qcp_solve<true>(M, 0.5 * (ga + A.traces[1]), R);
#pragma unroll
        for (int e = 0; e < 9; ++e) R1[e] = (float)R[e];
        for (int a = lane; a < n; a += kWave) {
            const float x = tx[a], y = ty[a], z = tz[a];
            tx[a] = x * R1[0] + y * R1[3] + z * R1[6];
            ty[a] = x * R1[1] + y * R1[4] + z * R1[7];
            tz[a] = x * R1[2] + y * R1[5] + z * R1[8];
        }
        __syncthreads();
    }

    // correlation of the relabelled selection with the reference: the
    // distinguishable atoms keep their labels
    double M2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = lane; k < A.n_dis; k += kWave) {
        const int a = clamp_index(A.dis[k], n);
        const double t[3] = {tx[a], ty[a], tz[a]};
        const double r[3] = {rx[a], ry[a], rz[a]};
#pragma unroll
        for (int i = 0; i < 3; ++i)
#pragma unroll
            for (int j = 0; j < 3; ++j) M2[3 * i + j] = fma(t[i], r[j], M2[3 * i + j]);
    }

    // (3) the assignment inside each group
    for (int g = 0; g < A.n_groups; ++g) {
        const int g0 = A.group_off[g];
        const int k = A.group_off[g + 1] - g0;
        if (k <= 0) continue;
        if (k > A.kmax) {  // larger than the LDS was sized for: no result
            if (lane == 0) A.out[f] = __builtin_nanf("");
            return;
        }
        for (int j = lane; j < k; j += kWave) {
            const int a = clamp_index(A.group_atoms[g0 + j], n);
            gat[j] = a;
            grx[j] = rx[a];
            gry[j] = ry[a];
            grz[j] = rz[a];
        }
        for (int j = lane; j <= k; j += kWave) {
            u[j] = 0.0;
            v[j] = 0.0;
            p[j] = 0;
            way[j] = 0;
        }
        wave_sync();
        for (int i = 1; i <= k; ++i) {
            for (int j = lane; j <= k; j += kWave) {
                minv[j] = __builtin_huge_val();
                used[j] = 0;
            }
            if (lane == 0) p[0] = i;
            wave_sync();
            int j0 = 0;
            // each pass marks one more column used; a free column (p = 0) is
            // reached within i passes, so the loop ends for any costs
            for (int pass = 0; pass <= k; ++pass) {
                if (lane == 0) used[j0] = 1;
                wave_sync();
                const int i0 = p[j0];
                const float qx = grx[i0 - 1], qy = gry[i0 - 1], qz = grz[i0 - 1];
                const double ui = u[i0];
                double best = __builtin_huge_val();
                int col = INT_MAX;
                for (int j = 1 + lane; j <= k; j += kWave) {
                    if (used[j]) continue;
                    // euclidean_permutation.cpp: squares of float differences, summed in double
                    const int ta = gat[j - 1];
                    const float dx = qx - tx[ta], dy = qy - ty[ta], dz = qz - tz[ta];
                    const double c = (double)__fmul_rn(dx, dx) + (double)__fmul_rn(dy, dy) +
                                     (double)__fmul_rn(dz, dz);
                    const double cur = c - ui - v[j];
                    double mv = minv[j];
                    if (cur < mv) {
                        mv = cur;
                        minv[j] = cur;
                        way[j] = j0;
                    }
                    if (mv < best || col == INT_MAX) {
                        best = mv;
                        col = j;
                    }
                }
                wave_argmin(best, col);
                const double delta = best;
                wave_sync();  // every lane has read u[i0] before it changes
                for (int j = lane; j <= k; j += kWave) {
                    if (used[j]) {
                        u[p[j]] += delta;
                        v[j] -= delta;
                    } else {
                        minv[j] -= delta;
                    }
                }
                wave_sync();
                j0 = col;
                if (p[j0] == 0) break;
            }
            if (lane == 0) {
                int jj = j0;
                for (int step = 0; step <= k && jj != 0; ++step) {
                    const int j1 = way[jj];
                    p[jj] = p[j1];
                    jj = j1;
                }
            }
            wave_sync();
        }
        // column j (target atom gat[j - 1]) takes row p[j] (reference atom gat[p[j] - 1])
        for (int j = 1 + lane; j <= k; j += kWave) {
            const int ta = gat[j - 1], ra = p[j] - 1;
            const double t[3] = {tx[ta], ty[ta], tz[ta]};
            const double r[3] = {grx[ra], gry[ra], grz[ra]};
#pragma unroll
            for (int a = 0; a < 3; ++a)
#pragma unroll
                for (int b = 0; b < 3; ++b) M2[3 * a + b] = fma(t[a], r[b], M2[3 * a + b]);
        }
        wave_sync();  // the next group overwrites the group's arrays
    }

    // (4) QCP on the relabelled selection
#pragma unroll
    for (int e = 0; e < 9; ++e) M2[e] = wave_sum(M2[e]);
    const double rg = A.traces[0];
    double R2[9];
    double lam;
    if (A.superpose)
        lam = qcp_solve<true>(M2, 0.5 * (tg + rg), R2);
    else
        lam = qcp_solve<false>(M2, 0.5 * (tg + rg), nullptr);
    const double msd = (tg + rg - 2.0 * lam) / n;
    if (lane == 0) A.out[f] = (float)sqrt(msd > 0.0 ? msd : 0.0);
    if (!A.superpose) return;

    // superpose: every atom of the frame, centred, times R1 R2 (float, as sgemm33)
    float R3[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            float acc = 0.f;
#pragma unroll
            for (int q = 0; q < 3; ++q) acc += R1[3 * i + q] * (float)R2[3 * q + j];
            R3[3 * i + j] = acc;
        }
    double w[3] = {0.0, 0.0, 0.0};
    for (int a = lane; a < A.n_atoms; a += kWave) {
        w[0] += fr[3LL * a];
        w[1] += fr[3LL * a + 1];
        w[2] += fr[3LL * a + 2];
    }
    double wm[3];
#pragma unroll
    for (int c = 0; c < 3; ++c) wm[c] = wave_sum(w[c]) / A.n_atoms;
    for (int a = lane; a < A.n_atoms; a += kWave) {
        const float x = (float)((double)fr[3LL * a] - wm[0]);
        const float y = (float)((double)fr[3LL * a + 1] - wm[1]);
        const float z = (float)((double)fr[3LL * a + 2] - wm[2]);
        fr[3LL * a] = x * R3[0] + y * R3[3] + z * R3[6];
        fr[3LL * a + 1] = x * R3[1] + y * R3[4] + z * R3[7];
        fr[3LL * a + 2] = x * R3[2] + y * R3[5] + z * R3[8];
    }
}

}  // namespace
}  // namespace hr

using namespace hr;

extern "C" {

size_t hr_lprmsd_workspace_bytes(int32_t n_sel, int32_t n_dis) {
    const size_t ns = n_sel > 0 ? (size_t)n_sel : 1, nd = n_dis > 0 ? (size_t)n_dis : 1;
    return align256(12 * ns) + align256(12 * nd) + align256(2 * sizeof(double));
}

int hr_lprmsd(float* xyz, int64_t n_frames, int32_t n_atoms, const float* ref_frame,
              const int32_t* atom_indices, int32_t n_sel, const int32_t* dis, int32_t n_dis,
              const int32_t* group_atoms, const int32_t* group_offsets, int32_t n_groups,
              int32_t max_group, int32_t superpose, float* out, void* workspace, void* stream) {
    clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0 && n_sel > 0 && n_sel <= n_atoms && n_dis >= 0 &&
                   n_dis <= n_sel && n_groups >= 0 && max_group >= 0 && max_group <= n_sel,
               "invalid sizes n_frames=%lld n_atoms=%d n_sel=%d n_dis=%d n_groups=%d "
               "max_group=%d",
               (long long)n_frames, n_atoms, n_sel, n_dis, n_groups, max_group);
    HR_REQUIRE(n_frames < (1LL << 31), "too many frames (%lld)", (long long)n_frames);
    const size_t lds = lp_lds_bytes(n_sel, max_group);
    if (lds > kLpLdsMax) {
        set_error("a selection of %d atoms with a permute group of %d needs %zu bytes of LDS "
                  "per frame (at most %zu)", n_sel, max_group, lds, kLpLdsMax);
        return HR_ERR_UNSUPPORTED;
    }
    if (n_frames == 0) return HR_OK;
    HR_REQUIRE(xyz && ref_frame && atom_indices && out && workspace &&
                   (n_dis == 0 || dis) && (n_groups == 0 || (group_atoms && group_offsets)),
               "NULL pointer argument");
    hipStream_t s = (hipStream_t)stream;
    char* ws = static_cast<char*>(workspace);
    float* rsel = reinterpret_cast<float*>(ws);
    ws += align256(12 * (size_t)n_sel);
    float* rdis = reinterpret_cast<float*>(ws);
    ws += align256(12 * (size_t)(n_dis > 0 ? n_dis : 1));
    double* traces = reinterpret_cast<double*>(ws);
    hipLaunchKernelGGL(lp_ref_kernel, dim3(1), dim3(kLpBlock), 0, s, ref_frame, n_atoms, atom_indices,
                       n_sel, dis, n_dis, rsel, rdis, traces);
    int rc = launch_status("lp_ref_kernel");
    if (rc) return rc;
    LpArgs A;
    A.xyz = xyz;
    A.n_frames = n_frames;
    A.n_atoms = n_atoms;
    A.idx = atom_indices;
    A.n_sel = n_sel;
    A.dis = dis;
    A.n_dis = n_dis;
    A.group_atoms = group_atoms;
    A.group_off = group_offsets;
    A.n_groups = n_groups;
    A.kmax = max_group;
    A.rsel = rsel;
    A.rdis = rdis;
    A.traces = traces;
    A.superpose = superpose ? 1 : 0;
    A.out = out;
    hipLaunchKernelGGL(lp_frame_kernel, dim3((unsigned)n_frames), dim3(kLpBlock), lds, s, A);
    return launch_status("lp_frame_kernel");
}

}  // extern "C"
