// Host build of the device QCP solver (mdtraj_amd/csrc/common.h), for the
// CPU tests: the same source the kernels inline, compiled for the host.
#include "common.h"

extern "C" {

double qcp_msd_host(const double* M, double ga, double gb, double n) {
    return hr::qcp_msd(M, ga, gb, n);
}

void qcp_msd2_host(const double* Ma, double gaa, double gba, const double* Mb, double gab,
                   double gbb, double n, double* out) {
    hr::qcp_msd2(Ma, gaa, gba, Mb, gab, gbb, n, out[0], out[1]);
}

double qcp_rot_host(const double* M, double ga, double gb, double* R) {
    return hr::qcp_solve<true>(M, 0.5 * (ga + gb), R);
}

// 1 when Newton on the characteristic polynomial alone certifies lambda_max
int qcp_newton_resolved_host(const double* M, double ga, double gb) {
    const hr::KeyMatrix k = hr::key_matrix(M);
    double mm = 0.0;
    for (int i = 0; i < 9; ++i) mm += M[i] * M[i];
    const double det_m = M[0] * (M[4] * M[8] - M[5] * M[7]) - M[1] * (M[3] * M[8] - M[5] * M[6]) +
                         M[2] * (M[3] * M[7] - M[4] * M[6]);
    const double a = k.k22 * k.k33 - k.k23 * k.k23, b = k.k12 * k.k33 - k.k13 * k.k23;
    const double c = k.k12 * k.k23 - k.k13 * k.k22, d = k.k02 * k.k33 - k.k03 * k.k23;
    const double e = k.k02 * k.k23 - k.k03 * k.k22, f = k.k02 * k.k13 - k.k03 * k.k12;
    const double c0 = k.k00 * (k.k11 * a - k.k12 * b + k.k13 * c) -
                      k.k01 * (k.k01 * a - k.k12 * d + k.k13 * e) +
                      k.k02 * (k.k01 * b - k.k11 * d + k.k13 * f) -
                      k.k03 * (k.k01 * c - k.k11 * e + k.k12 * f);
    bool resolved = false;
    hr::qcp_largest_root(-2.0 * mm, -8.0 * det_m, c0, 0.5 * (ga + gb), sqrt(mm), &resolved);
    return resolved ? 1 : 0;
}

}  // extern "C"
