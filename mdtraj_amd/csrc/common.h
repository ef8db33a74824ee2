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

// Largest real root of P(x) = x^4 + c2 x^2 + c1 x + c0 whose roots are all
// real (P is the characteristic polynomial of the symmetric key matrix K).
// Newton from a start that lies above every root converges monotonically to
// the largest root (Theobald, Acta Cryst. A61:478, 2005).  The start is
// e0 = (G_a + G_b) / 2, an upper bound of lambda_max; if rounding puts it
// below (some derivative of P negative there), the Frobenius bound
// ||K||_F = 2 ||M||_F is used instead.
__device__ inline double qcp_largest_root(double c2, double c1, double c0, double e0,
                                          double m_frob) {
    double x = e0;
    {
        const double x2 = x * x;
        const double p = (x2 + c2) * x2 + c1 * x + c0;
        const double d1 = (4.0 * x2 + 2.0 * c2) * x + c1;
        const double d2 = 12.0 * x2 + 2.0 * c2;
        if (!(p >= 0.0 && d1 >= 0.0 && d2 >= 0.0 && x >= 0.0))
            x = 2.0 * m_frob * (1.0 + 1e-12) + 1e-300;
    }
    for (int it = 0; it < 64; ++it) {
        const double x2 = x * x;
        const double p = (x2 + c2) * x2 + c1 * x + c0;
        const double d1 = (4.0 * x2 + 2.0 * c2) * x + c1;
        if (!(p > 0.0) || !(d1 > 0.0)) break;
        const double nx = x - p / d1;
        if (!(nx < x)) break;
        const double step = x - nx;
        x = nx;
        if (step <= 1e-15 * x) break;
    }
    return x;
}

// Key-matrix entries of Horn/Theobald for M[r*3+c] = sum a_r b_c, in the
// orientation of rmsd/src/theobald_rmsd.cpp:227-236 (which reads M
// column-major), so that the top eigenvector is the quaternion rotating a
// onto b.
struct KeyMatrix {
    double k00, k01, k02, k03, k11, k12, k13, k22, k23, k33;
};

__device__ __forceinline__ KeyMatrix key_matrix(const double M[9]) {
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

// lambda_max of K(M) in FP64.  Coefficients of the characteristic
// polynomial: C2 = -2 |M|_F^2, C1 = -8 det M, C0 = det K.
__device__ inline double qcp_lambda(const double M[9], double e0, KeyMatrix* kout) {
    const KeyMatrix k = key_matrix(M);
    double mm = 0.0;
#pragma unroll
    for (int i = 0; i < 9; ++i) mm = fma(M[i], M[i], mm);
    const double c2 = -2.0 * mm;
    const double det_m = M[0] * (M[4] * M[8] - M[5] * M[7]) -
                         M[1] * (M[3] * M[8] - M[5] * M[6]) +
                         M[2] * (M[3] * M[7] - M[4] * M[6]);
    const double c1 = -8.0 * det_m;
    // det K by cofactor expansion along the first row
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
    if (kout) *kout = k;
    return qcp_largest_root(c2, c1, c0, e0, sqrt(mm));
}

// MSD from the centred correlation matrix and traces, all FP64.
__device__ __forceinline__ double qcp_msd(const double M[9], double ga, double gb, double n) {
    const double e0 = 0.5 * (ga + gb);
    const double lam = qcp_lambda(M, e0, nullptr);
    const double msd = (ga + gb - 2.0 * lam) / n;
    return msd > 0.0 ? msd : 0.0;
}

__device__ __forceinline__ double det3(double a, double b, double c, double d, double e, double f,
                                       double g, double h, double i) {
    return a * (e * i - f * h) - b * (d * i - f * g) + c * (d * h - e * g);
}

// Optimal rotation (row-major, applied as x' = x R) from K and lambda_max:
// the top eigenvector of K is any non-zero column of adj(K - lambda I); the
// column with the largest diagonal cofactor is used (the reference always
// takes the first, rmsd/src/theobald_rmsd.cpp:279-331, and falls back to the
// identity when it vanishes).  Identity if K - lambda I has rank < 3.
__device__ inline void qcp_rotation(const KeyMatrix& k, double lam, double R[9]) {
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
    if (!(qq > 1e-28 * s3 * s3) || !(qq > 1e-300)) {
        R[0] = 1; R[1] = 0; R[2] = 0;
        R[3] = 0; R[4] = 1; R[5] = 0;
        R[6] = 0; R[7] = 0; R[8] = 1;
        return;
    }
    const double inv = 1.0 / sqrt(qq);
    w *= inv; x *= inv; y *= inv; z *= inv;
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

}  // namespace hr
