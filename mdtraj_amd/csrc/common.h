// Shared host/device helpers of the hiprmsd library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdint.h>
#include <stdio.h>

#include "hiprmsd.h"

namespace hr {

// ----------------------------------------------------------------- host side
void set_error(const char* fmt, ...);
void clear_error();

struct DeviceInfo {
    int device = -1;
    int cu_count = 0;
    size_t lds_per_block = 0;
    size_t lds_per_cu = 0;
};
// Properties of the current device (initialised on first use per device).
const DeviceInfo* current_device();

#define HR_CHECK_HIP(expr)                                                         \
    do {                                                                           \
        hipError_t e_ = (expr);                                                    \
        if (e_ != hipSuccess) {                                                    \
            ::hr::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                            __FILE__, __LINE__);                                   \
            return HR_ERR_HIP;                                                     \
        }                                                                          \
    } while (0)

#define HR_REQUIRE(cond, ...)            \
    do {                                 \
        if (!(cond)) {                   \
            ::hr::set_error(__VA_ARGS__); \
            return HR_ERR_INVALID;       \
        }                                \
    } while (0)

// Process-wide tuning knobs (HR_TUNE_*), set by hr_set_tuning; 0 = default.
int tuning(int knob);

// One-vs-all over contiguous frames of n_sel atoms against a reference
// selection given as n_sel interleaved atoms, in one launch: the per-frame
// kernel prepares the reference itself (one_vs_all.hip).  *done is false, and
// nothing is launched, when the selection does not fit that kernel's LDS.
int one_vs_all_fused_ref(hipStream_t s, const float* frames, long long n_frames, int n_sel,
                         const float* ref, long long self_frame, float* out, bool* done);

// Launch status of the last kernel launch.
inline int launch_status(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        set_error("launch of %s failed: %s", what, hipGetErrorString(e));
        return HR_ERR_HIP;
    }
    return HR_OK;
}

// --------------------------------------------------------------- device side
constexpr int kWave = 64;

// Coordinates, references and indices always live in global memory, but a
// pointer the kernel reads from a descriptor table is generic to the
// compiler, which then emits flat loads.  Flat loads count against lgkmcnt as
// well as vmcnt, so every LDS access or cross-lane permute that waits on
// lgkmcnt would also drain the frame loads in flight.  Loads through these
// casts are global_load instructions whatever the pointer's origin.
typedef __attribute__((address_space(1))) const float GlobalFloat;
typedef __attribute__((address_space(1))) const int GlobalInt;
__device__ __forceinline__ const GlobalFloat* global_ptr(const float* p) {
    return (const GlobalFloat*)p;
}
__device__ __forceinline__ const GlobalInt* global_ptr(const int* p) {
    return (const GlobalInt*)p;
}
typedef __attribute__((address_space(1))) const double GlobalDouble;
__device__ __forceinline__ const GlobalDouble* global_ptr(const double* p) {
    return (const GlobalDouble*)p;
}
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T* global_out(T* p) {
    return (__attribute__((address_space(1))) T*)p;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// Cross-lane move of a double by a DPP pattern (both 32-bit halves).
template <int kCtrl>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), kCtrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), kCtrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sum over each row of 16 lanes, result in every lane of the row, all on the
// VALU (no LDS traffic): quad_perm [1,0,3,2] and [2,3,0,1] (xor 1, xor 2),
// then row_half_mirror (lane i <- 7 - i) and row_mirror (lane i <- 15 - i),
// each of which pairs every lane with one holding the other half's sum.
__device__ __forceinline__ double row16_sum(double v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return v;
}

__device__ __forceinline__ double seg_sum(double v, int width) {
    for (int m = width >> 1; m >= 1; m >>= 1) v += __shfl_xor(v, m, kWave);
    return v;
}

// 1/d.  On the GPU: the hardware estimate (v_rcp_f64) refined by one
// Newton-Raphson step, r + r (1 - d r), in place of the IEEE division
// sequence; the Newton iteration below corrects any last-ulp difference.
// The host build of this code (used by the CPU tests) divides.
__host__ __device__ __forceinline__ double recip(double d) {
#if defined(__HIP_DEVICE_COMPILE__)
    const double r = __builtin_amdgcn_rcp(d);
    return fma(r, fma(-d, r, 1.0), r);
#else
    return 1.0 / d;
#endif
}

// Largest real root of P(x) = x^4 + c2 x^2 + c1 x + c0 whose roots are all
// real (P is the characteristic polynomial of the symmetric key matrix K).
// Newton from a start that lies above every root converges monotonically to
// the largest root (Theobald, Acta Cryst. A61:478, 2005).  A point lies
// above every root iff P, its first and second derivatives and x (the third
// derivative is 24 x) are all positive there (Budan-Fourier); the start
// e0 = (G_a + G_b) / 2 is such a point in exact arithmetic, and when rounding
// says otherwise the Frobenius bound ||K||_F = 2 ||M||_F is used.  Every
// Newton iterate is checked the same way and the iteration stops at the last
// one that passes, so rounding noise near a (near-)double root can never
// carry it below lambda_max.
//
// The result is one more Newton step from the last such point x,
// y = x - P/P'.  With h = x - lambda_max: P'/P = sum_i 1/(x - lambda_i)
// <= 4/h gives h <= 4 P/P'; Taylor at x with P'' increasing on x > 0 (its
// derivative is 24 x) gives P' h <= P + P''(x) h^2 / 2, and together
// lambda_max lies in [y - 8 P'' P^2 / P'^3, y].  *resolved says whether that
// bracket, widened by the rounding bound of P, is within 1e-11 x.  For a
// simple root it is (the bracket is quadratic in the last step); it is not at
// (near-)double roots -- the two largest eigenvalues of K (nearly) coincide,
// as for a centred two-atom (collinear) structure -- where P's values
// determine the root only to about sqrt(eps) lambda.
// Coefficients of P(x) = x^4 + c2 x^2 + c1 x + c0 and |M|_F^2, whose root
// gives the Frobenius bound (taken only when the start e0 fails its check,
// so the square root is not paid on the common path).
struct Quartic {
    double c2, c1, c0, mm;
};

// Newton state: the last point verified above every root, P, P', P'' there;
// run: the iteration continues.
struct NewtonState {
    double x, p, d1, d2;
    bool run;
};

__host__ __device__ __forceinline__ void quartic_eval(const Quartic& q, double t, double& p,
                                                     double& d1, double& d2) {
    const double t2 = t * t;
    p = (t2 + q.c2) * t2 + q.c1 * t + q.c0;
    d1 = (4.0 * t2 + 2.0 * q.c2) * t + q.c1;
    d2 = 12.0 * t2 + 2.0 * q.c2;
}

// Start at the smaller of two upper bounds of lambda_max: e0, and
// sqrt(3) ||M||_F (lambda_max = s1 + s2 +- s3 over the singular values of M,
// at most sqrt(3) times their root sum of squares).  The second is the
// tighter one for pairs far apart (one Newton step fewer per solve on a
// 20,000-frame walk: 2.93 -> 2.00); its square root is taken in float and
// widened by 1e-6, which keeps it an upper bound.  Else start at e0, else
// at the Frobenius bound of K; run is false when none is verified above the
// roots (then the root is left to the Jacobi fallback).
__host__ __device__ __forceinline__ NewtonState newton_start(const Quartic& q, double e0) {
    NewtonState s;
    const double nb = (double)sqrtf((float)(3.0 * q.mm)) * (1.0 + 1e-6);
    s.x = nb < e0 ? nb : e0;
    quartic_eval(q, s.x, s.p, s.d1, s.d2);
    s.run = s.p > 0.0 && s.d1 > 0.0 && s.d2 > 0.0 && s.x > 0.0;
    if (!s.run && s.x != e0) {
        s.x = e0;
        quartic_eval(q, s.x, s.p, s.d1, s.d2);
        s.run = s.p > 0.0 && s.d1 > 0.0 && s.d2 > 0.0 && s.x > 0.0;
    }
    if (!s.run) {
        s.x = 2.0 * sqrt(q.mm) * (1.0 + 1e-12) + 1e-300;
        quartic_eval(q, s.x, s.p, s.d1, s.d2);
        s.run = s.p > 0.0 && s.d1 > 0.0 && s.d2 > 0.0 && s.x > 0.0;
    }
    return s;
}

// One Newton step, branch-free (so that two solves can be interleaved): the
// new point is taken only if it is still verified above the roots; the
// iteration stops once a step falls below 1e-4 x -- quadratic convergence
// then leaves x about 1e-8 x above a simple root and the final estimate
// x - P/P' about 1e-16 x from it, which the bracket of newton_finish
// certifies (or not: a near-double root converges slowly, fails the
// bracket and takes the Jacobi fallback).  A 1e-8 stop costs one more
// step on nearly every solve for no accuracy on a certified one.
__host__ __device__ __forceinline__ void newton_step(const Quartic& q, NewtonState& s) {
    const double nx = s.x - s.p * recip(s.d1);
    double np, nd1, nd2;
    quartic_eval(q, nx, np, nd1, nd2);
    const bool take = s.run && nx < s.x && np > 0.0 && nd1 > 0.0 && nd2 > 0.0 && nx > 0.0;
    const double step = s.x - nx;
    s.x = take ? nx : s.x;
    s.p = take ? np : s.p;
    s.d1 = take ? nd1 : s.d1;
    s.d2 = take ? nd2 : s.d2;
    s.run = take && step > 1e-4 * nx;
}

// Final estimate y = x - P/P' and whether its bracket certifies it:
// width 8 P'' q^2 / P'^3 + 2 perr / P' (q = P + perr, perr the Horner
// rounding bound of P) within 1e-11 x, multiplied through by P'^3 (an
// underflowing P'^3 counts as unresolved).
__host__ __device__ __forceinline__ double newton_finish(const Quartic& q, const NewtonState& s,
                                                        bool started, bool* resolved) {
    const double x = s.x, x2 = x * x;
    const double perr = 8.9e-16 * (x2 * x2 + fabs(q.c2) * x2 + fabs(q.c1) * x + fabs(q.c0));
    const double qq = s.p + perr;
    const double d13 = s.d1 * s.d1 * s.d1;
    *resolved = started && d13 > 0.0 &&
                8.0 * s.d2 * qq * qq + 2.0 * perr * s.d1 * s.d1 <= 1e-11 * x * d13;
    return x - s.p * recip(s.d1);
}

__host__ __device__ inline double qcp_largest_root(const Quartic& q, double e0, bool* resolved) {
    NewtonState s = newton_start(q, e0);
    const bool started = s.run;
    for (int it = 0; it < 64 && s.run; ++it) newton_step(q, s);
    return newton_finish(q, s, started, resolved);
}

// The same from the coefficients and |M|_F.
__host__ __device__ inline double qcp_largest_root(double c2, double c1, double c0, double e0,
                                                   double m_frob, bool* resolved) {
    return qcp_largest_root(Quartic{c2, c1, c0, m_frob * m_frob}, e0, resolved);
}

// Key-matrix entries of Horn/Theobald for M[r*3+c] = sum a_r b_c, in the
// orientation of rmsd/src/theobald_rmsd.cpp:227-236 (which reads M
// column-major), so that the top eigenvector is the quaternion rotating a
// onto b.
struct KeyMatrix {
    double k00, k01, k02, k03, k11, k12, k13, k22, k23, k33;
};

__host__ __device__ __forceinline__ KeyMatrix key_matrix(const double M[9]) {
    KeyMatrix k;
    k.k00 = M[0] + M[4] + M[8];
    k.k01 = M[7] - M[5];
    k.k02 = M[2] - M[6];
    k.k03 = M[3] - M[1];
    k.k11 = M[0] - M[4] - M[8];
    k.k12 = M[3] + M[1];
    k.k13 = M[2] + M[6];
    k.k22 = -M[0] + M[4] - M[8];
    k.k23 = M[7] + M[5];
    k.k33 = -M[0] - M[4] + M[8];
    return k;
}

// One Jacobi rotation of the symmetric 4x4 a in the (p, r) plane that zeroes
// a[p][r] (Golub & Van Loan, Matrix Computations, 4th ed., sec. 8.5.2, with
// the smaller rotation angle); v accumulates the rotations when kVec.
template <int p, int r, bool kVec>
__host__ __device__ __forceinline__ void jacobi_rotate(double (&a)[4][4], double (&v)[4][4]) {
    const double apr = a[p][r];
    if (apr == 0.0) return;
    const double theta = (a[r][r] - a[p][p]) / (2.0 * apr);
    const double t = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(fma(theta, theta, 1.0)));
    const double c = 1.0 / sqrt(fma(t, t, 1.0));
    const double s = t * c;
    a[p][p] -= t * apr;
    a[r][r] += t * apr;
    a[p][r] = 0.0;
    a[r][p] = 0.0;
#pragma unroll
    for (int m = 0; m < 4; ++m) {
        if (m == p || m == r) continue;
        const double g = a[m][p], h = a[m][r];
        a[m][p] = a[p][m] = c * g - s * h;
        a[m][r] = a[r][m] = s * g + c * h;
    }
    if constexpr (kVec) {
#pragma unroll
        for (int m = 0; m < 4; ++m) {
            const double g = v[m][p], h = v[m][r];
            v[m][p] = c * g - s * h;
            v[m][r] = s * g + c * h;
        }
    }
}

// lambda_max of K by cyclic Jacobi: absolute error a few eps ||K|| whatever
// the eigenvalue spacing.  With kVec, q receives a unit eigenvector of it.
// Used only where Newton on P cannot resolve the root (see above).
template <bool kVec>
__host__ __device__ inline double qcp_jacobi(const KeyMatrix& k, double q[4]) {
    double a[4][4] = {{k.k00, k.k01, k.k02, k.k03},
                      {k.k01, k.k11, k.k12, k.k13},
                      {k.k02, k.k12, k.k22, k.k23},
                      {k.k03, k.k13, k.k23, k.k33}};
    double v[4][4] = {{1, 0, 0, 0}, {0, 1, 0, 0}, {0, 0, 1, 0}, {0, 0, 0, 1}};
    for (int sweep = 0; sweep < 16; ++sweep) {
        const double off = a[0][1] * a[0][1] + a[0][2] * a[0][2] + a[0][3] * a[0][3] +
                           a[1][2] * a[1][2] + a[1][3] * a[1][3] + a[2][3] * a[2][3];
        const double diag =
            a[0][0] * a[0][0] + a[1][1] * a[1][1] + a[2][2] * a[2][2] + a[3][3] * a[3][3];
        if (!(off > 1e-34 * diag)) break;
        jacobi_rotate<0, 1, kVec>(a, v);
        jacobi_rotate<0, 2, kVec>(a, v);
        jacobi_rotate<0, 3, kVec>(a, v);
        jacobi_rotate<1, 2, kVec>(a, v);
        jacobi_rotate<1, 3, kVec>(a, v);
        jacobi_rotate<2, 3, kVec>(a, v);
    }
    int top = 0;
    double lam = a[0][0];
#pragma unroll
    for (int i = 1; i < 4; ++i)
        if (a[i][i] > lam) { lam = a[i][i]; top = i; }
    if constexpr (kVec) {
#pragma unroll
        for (int m = 0; m < 4; ++m)
            q[m] = top == 0 ? v[m][0] : top == 1 ? v[m][1] : top == 2 ? v[m][2] : v[m][3];
    }
    return lam;
}

__host__ __device__ __forceinline__ double det3(double a, double b, double c, double d, double e,
                                                double f, double g, double h, double i) {
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Rotation matrix (row-major, applied as x' = x R) of the unit quaternion
// (w, x, y, z) in the basis order of K.
__host__ __device__ __forceinline__ void quat_rotation(double w, double x, double y, double z,
                                                       double R[9]) {
    R[0] = w * w + x * x - y * y - z * z;
    R[1] = 2.0 * (x * y - w * z);
    R[2] = 2.0 * (x * z + w * y);
    R[3] = 2.0 * (x * y + w * z);
    R[4] = w * w - x * x + y * y - z * z;
    R[5] = 2.0 * (y * z - w * x);
    R[6] = 2.0 * (x * z - w * y);
    R[7] = 2.0 * (y * z + w * x);
    R[8] = w * w - x * x - y * y + z * z;
}

// Optimal rotation from K and a simple lambda_max: the top eigenvector of K is
// any non-zero column of adj(K - lambda I); the column with the largest
// diagonal cofactor is used (the reference always takes the first,
// rmsd/src/theobald_rmsd.cpp:279-331, and falls back to the identity when it
// vanishes).  Returns false (R untouched) if K - lambda I has rank < 3.
__host__ __device__ inline bool qcp_rotation(const KeyMatrix& k, double lam, double R[9]) {
    const double a00 = k.k00 - lam, a11 = k.k11 - lam, a22 = k.k22 - lam, a33 = k.k33 - lam;
    const double a01 = k.k01, a02 = k.k02, a03 = k.k03, a12 = k.k12, a13 = k.k13, a23 = k.k23;
    // Full adjugate of the symmetric 4x4 A (cofactors C_ij = (-1)^(i+j) det minor_ij).
    const double c00 = det3(a11, a12, a13, a12, a22, a23, a13, a23, a33);
    const double c01 = -det3(a01, a12, a13, a02, a22, a23, a03, a23, a33);
    const double c02 = det3(a01, a11, a13, a02, a12, a23, a03, a13, a33);
    const double c03 = -det3(a01, a11, a12, a02, a12, a22, a03, a13, a23);
    const double c11 = det3(a00, a02, a03, a02, a22, a23, a03, a23, a33);
    const double c12 = -det3(a00, a01, a03, a02, a12, a23, a03, a13, a33);
    const double c13 = det3(a00, a01, a02, a02, a12, a22, a03, a13, a23);
    const double c22 = det3(a00, a01, a03, a01, a11, a13, a03, a13, a33);
    const double c23 = -det3(a00, a01, a02, a01, a11, a12, a03, a13, a23);
    const double c33 = det3(a00, a01, a02, a01, a11, a12, a02, a12, a22);
    double w = c00, x = c01, y = c02, z = c03, best = fabs(c00);
    if (fabs(c11) > best) { w = c01; x = c11; y = c12; z = c13; best = fabs(c11); }
    if (fabs(c22) > best) { w = c02; x = c12; y = c22; z = c23; best = fabs(c22); }
    if (fabs(c33) > best) { w = c03; x = c13; y = c23; z = c33; best = fabs(c33); }
    const double qq = w * w + x * x + y * y + z * z;
    const double scale = fabs(a00) + fabs(a11) + fabs(a22) + fabs(a33) + fabs(a01) + fabs(a02) +
                         fabs(a03) + fabs(a12) + fabs(a13) + fabs(a23);
    const double s3 = scale * scale * scale;
    if (!(qq > 1e-28 * s3 * s3) || !(qq > 1e-300)) return false;
    const double inv = 1.0 / sqrt(qq);
    quat_rotation(w * inv, x * inv, y * inv, z * inv, R);
    return true;
}

// lambda_max of K(M) in FP64, and with kRot the optimal rotation R (identity
// when K = 0).  Coefficients of the characteristic polynomial:
// C2 = -2 |M|_F^2, C1 = -8 det M, C0 = det K.  Newton on P is the fast path;
// an unresolved (near-double) root, or a rank-deficient K - lambda I for the
// rotation, is settled by Jacobi on K itself.
// Characteristic polynomial of K(M): C2 = -2 |M|_F^2, C1 = -8 det M,
// C0 = det K (cofactor expansion along the first row).
__host__ __device__ __forceinline__ Quartic qcp_quartic(const double M[9], const KeyMatrix& k) {
    double mm = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) mm = fma(M[i], M[i], mm);
    const double det_m = M[0] * (M[4] * M[8] - M[5] * M[7]) -
                         M[1] * (M[3] * M[8] - M[5] * M[6]) +
                         M[2] * (M[3] * M[7] - M[4] * M[6]);
    const double a = k.k22 * k.k33 - k.k23 * k.k23;
    const double b = k.k12 * k.k33 - k.k13 * k.k23;
    const double c = k.k12 * k.k23 - k.k13 * k.k22;
    const double d = k.k02 * k.k33 - k.k03 * k.k23;
    const double e = k.k02 * k.k23 - k.k03 * k.k22;
    const double f = k.k02 * k.k13 - k.k03 * k.k12;
    const double c0 = k.k00 * (k.k11 * a - k.k12 * b + k.k13 * c) -
                      k.k01 * (k.k01 * a - k.k12 * d + k.k13 * e) +
                      k.k02 * (k.k01 * b - k.k11 * d + k.k13 * f) -
                      k.k03 * (k.k01 * c - k.k11 * e + k.k12 * f);
    return Quartic{-2.0 * mm, -8.0 * det_m, c0, mm};
}

template <bool kRot>
__host__ __device__ inline double qcp_solve(const double M[9], double e0, double R[9]) {
    const KeyMatrix k = key_matrix(M);
    const Quartic q = qcp_quartic(M, k);
    bool resolved = true;
    double lam = qcp_largest_root(q, e0, &resolved);
    if constexpr (kRot) {
        if (resolved && qcp_rotation(k, lam, R)) return lam;
        if (!(q.mm > 0.0)) {
            quat_rotation(1.0, 0.0, 0.0, 0.0, R);
            return 0.0;
        }
        double qv[4];
        lam = qcp_jacobi<true>(k, qv);
        const double qq = qv[0] * qv[0] + qv[1] * qv[1] + qv[2] * qv[2] + qv[3] * qv[3];
        const double inv = 1.0 / sqrt(qq);
        quat_rotation(qv[0] * inv, qv[1] * inv, qv[2] * inv, qv[3] * inv, R);
        return lam;
    } else {
        if (!resolved) lam = qcp_jacobi<false>(k, nullptr);
        return lam;
    }
}

// MSD from the centred correlation matrix and traces, all FP64.
__host__ __device__ __forceinline__ double qcp_msd(const double M[9], double ga, double gb,
                                                   double n) {
    const double lam = qcp_solve<false>(M, 0.5 * (ga + gb), nullptr);
    const double msd = (ga + gb - 2.0 * lam) / n;
    return msd > 0.0 ? msd : 0.0;
}

// Two independent MSDs with their Newton iterations interleaved step by step
// (the branch-free steps of the two solves overlap in the pipeline, where a
// single solve is a chain of dependent operations); same results as two
// qcp_msd calls.  An unresolved root takes the Jacobi fallback on its own.
__host__ __device__ inline void qcp_msd2(const double Ma[9], double gaa, double gba,
                                         const double Mb[9], double gab, double gbb, double n,
                                         double& msd_a, double& msd_b) {
    const KeyMatrix ka = key_matrix(Ma), kb = key_matrix(Mb);
    const Quartic qa = qcp_quartic(Ma, ka), qb = qcp_quartic(Mb, kb);
    NewtonState sa = newton_start(qa, 0.5 * (gaa + gba));
    NewtonState sb = newton_start(qb, 0.5 * (gab + gbb));
    const bool started_a = sa.run, started_b = sb.run;
    for (int it = 0; it < 64 && (sa.run || sb.run); ++it) {
        newton_step(qa, sa);
        newton_step(qb, sb);
    }
    bool res_a, res_b;
    double la = newton_finish(qa, sa, started_a, &res_a);
    double lb = newton_finish(qb, sb, started_b, &res_b);
    if (!res_a) la = qcp_jacobi<false>(ka, nullptr);
    if (!res_b) lb = qcp_jacobi<false>(kb, nullptr);
    const double ma = (gaa + gba - 2.0 * la) / n, mb = (gab + gbb - 2.0 * lb) / n;
    msd_a = ma > 0.0 ? ma : 0.0;
    msd_b = mb > 0.0 ? mb : 0.0;
}

}  // namespace hr
