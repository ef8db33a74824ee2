/*
 * Float32 restatement of the reference RMSD algorithm — test infrastructure.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load
 * this library, as the checker or as the timed CPU baseline; the product
 * (mdtraj_amd) never calls it.
 *
 * It models the reference's float32 behaviour (paths relative to the mdtraj
 * source tree):
 *   - centering: double sums, float mean, float subtraction, double trace of
 *     float squares, cast to float            (rmsd/src/center_generic.h:3-44)
 *   - M accumulated in float, atom by atom     (rmsd/src/theobald_rmsd_generic.h:118-174)
 *   - K, C2, C1, C0 in float                   (rmsd/src/theobald_rmsd.cpp:227-271)
 *   - the largest root of the quartic in double, cast to float
 *       (the reference uses a closed-form resolvent-cubic solve,
 *        rmsd/src/theobald_rmsd.cpp:81-193; here a safeguarded Newton
 *        iteration written from the mathematics — both return the largest
 *        real root of the same polynomial)
 *   - msd = max(0, (G_a + G_b - 2 lambda) / N) in float (theobald_rmsd.cpp:273-277)
 *   - rotation from the adjugate column of K - lambda I, identity below
 *     |q|^2 < 1e-11                            (theobald_rmsd.cpp:279-331)
 *   - the driver loop with OpenMP over frames  (rmsd/_rmsd.pyx:197-224)
 *
 * Parity: pinned against the compiled reference core (oracle/_ref) through
 * the golden vectors in tests/golden/ (tests/test_oracle32.py).  The SSE
 * build of the reference sums M in four lanes, so results agree to a few
 * float ulps, not bit for bit.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* Largest real root of x^4 + c2 x^2 + c1 x + c0 (double), all roots real. */
static double largest_root(double c2, double c1, double c0, double start, double mfrob) {
    /* Start above the largest root: every derivative non-negative at x
     * means no root lies above x (Descartes on P(x + t)). */
    double x = start;
    for (int pass = 0; pass < 2; ++pass) {
        double x2 = x * x;
        double p = (x2 + c2) * x2 + c1 * x + c0;
        double d1 = (4.0 * x2 + 2.0 * c2) * x + c1;
        double d2 = 12.0 * x2 + 2.0 * c2;
        if (p >= 0.0 && d1 >= 0.0 && d2 >= 0.0 && x >= 0.0) break;
        /* Frobenius bound: lambda_max <= ||K||_F = 2 ||M||_F */
        x = 2.0 * mfrob * (1.0 + 1e-12) + 1e-300;
    }
    for (int it = 0; it < 100; ++it) {
        double x2 = x * x;
        double p = (x2 + c2) * x2 + c1 * x + c0;
        double d1 = (4.0 * x2 + 2.0 * c2) * x + c1;
        if (p <= 0.0 || d1 <= 0.0) break;
        double nx = x - p / d1;
        if (!(nx < x)) break;
        double step = x - nx;
        x = nx;
        if (step <= 1e-15 * fabs(x)) break;
    }
    return x;
}

/* theobald_rmsd.cpp:217-334 restated. */
static float msd_from_m_and_g(const float M[9], float Ga, float Gb, int n_atoms,
                              int compute_rot, float rot[9]) {
    float k00 = M[0] + M[4] + M[8];
    /* The reference reads M column-major (M[i + 3 j]) although it is
     * stored row-major, i.e. K is built from M^T; the eigenvalue is the same
     * either way, the quaternion below is the one that rotates a onto b. */
    float k01 = M[7] - M[5];
    float k02 = M[2] - M[6];
    float k03 = M[3] - M[1];
    float k11 = M[0] - M[4] - M[8];
    float k12 = M[3] + M[1];
    float k13 = M[2] + M[6];
    float k22 = -M[0] + M[4] - M[8];
    float k23 = M[7] + M[5];
    float k33 = -M[0] - M[4] + M[8];

    float c2 = 0.0f;
    for (int i = 0; i < 9; ++i) c2 += M[i] * M[i];
    float mfrob = sqrtf(c2);
    c2 *= -2.0f;
    float det_m = M[0] * (M[4] * M[8] - M[5] * M[7])
                + M[3] * (M[7] * M[2] - M[8] * M[1])
                + M[6] * (M[1] * M[5] - M[2] * M[4]);
    /* det K by cofactor expansion along the first row */
    float a = k22 * k33 - k23 * k23;
    float b = k12 * k33 - k13 * k23;
    float c = k12 * k23 - k13 * k22;
    float d = k02 * k33 - k03 * k23;
    float e = k02 * k23 - k03 * k22;
    float f = k02 * k13 - k03 * k12;
    float det_k = k00 * (k11 * a - k12 * b + k13 * c)
                - k01 * (k01 * a - k12 * d + k13 * e)
                + k02 * (k01 * b - k11 * d + k13 * f)
                - k03 * (k01 * c - k11 * e + k12 * f);
    float c1 = -8.0f * det_m;
    float c0 = det_k;

    float e0 = (Ga + Gb) / 2.0f;
    float lambda = (float)largest_root((double)c2, (double)c1, (double)c0, (double)e0, (double)mfrob);

    float msd = (Ga + Gb - 2.0f * lambda) / n_atoms;
    if (!(msd > 0.0f)) msd = 0.0f;

    if (compute_rot) {
        float q00 = k00 - lambda, q11 = k11 - lambda, q22 = k22 - lambda, q33 = k33 - lambda;
        float s2233 = q22 * q33 - k23 * k23;
        float s1233 = k12 * q33 - k13 * k23;
        float s1223 = k12 * k23 - k13 * q22;
        float s0223 = k02 * k23 - k03 * q22;
        float s0233 = k02 * q33 - k03 * k23;
        float s0213 = k02 * k13 - k03 * k12;
        float w = q11 * s2233 - k12 * s1233 + k13 * s1223;
        float x = -k01 * s2233 + k12 * s0233 - k13 * s0223;
        float y = k01 * s1233 - q11 * s0233 + k13 * s0213;
        float z = -k01 * s1223 + q11 * s0223 - k12 * s0213;
        (void)q00;
        float qq = w * w + x * x + y * y + z * z;
        if (qq < 1e-11f) {
            memset(rot, 0, 9 * sizeof(float));
            rot[0] = rot[4] = rot[8] = 1.0f;
        } else {
            float nq = sqrtf(qq);
            w /= nq; x /= nq; y /= nq; z /= nq;
            rot[0] = w * w + x * x - y * y - z * z;
            rot[1] = 2 * (x * y - w * z);
            rot[2] = 2 * (z * x + w * y);
            rot[3] = 2 * (x * y + w * z);
            rot[4] = w * w - x * x + y * y - z * z;
            rot[5] = 2 * (y * z - w * x);
            rot[6] = 2 * (z * x - w * y);
            rot[7] = 2 * (y * z + w * x);
            rot[8] = w * w - x * x - y * y + z * z;
        }
    }
    return msd;
}

/* center_generic.h:3-44 restated. */
void o32_center_and_trace(float* xyz, float* traces, int n_frames, int n_atoms) {
    #pragma omp parallel for schedule(static)
    for (long long f = 0; f < n_frames; ++f) {
        float* p = xyz + f * (long long)n_atoms * 3;
        double sx = 0, sy = 0, sz = 0, tr = 0;
        for (int i = 0; i < n_atoms; ++i) {
            sx += p[3 * i]; sy += p[3 * i + 1]; sz += p[3 * i + 2];
        }
        float mx = (float)(sx / n_atoms), my = (float)(sy / n_atoms), mz = (float)(sz / n_atoms);
        for (int i = 0; i < n_atoms; ++i) {
            p[3 * i] -= mx; p[3 * i + 1] -= my; p[3 * i + 2] -= mz;
            tr += p[3 * i] * p[3 * i] + p[3 * i + 1] * p[3 * i + 1] + p[3 * i + 2] * p[3 * i + 2];
        }
        if (traces) traces[f] = (float)tr;
    }
}

/* theobald_rmsd_generic.h:118-174 restated (M[r*3+c] = sum a_r b_c). */
float o32_msd_atom_major(int n_atoms, const float* a, const float* b, float Ga, float Gb,
                         int compute_rot, float* rot) {
    if (a == b && Ga == Gb) {
        if (compute_rot) {
            memset(rot, 0, 9 * sizeof(float));
            rot[0] = rot[4] = rot[8] = 1.0f;
        }
        return 0.0f;
    }
    float M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < n_atoms; ++k) {
        float ax = a[3 * k], ay = a[3 * k + 1], az = a[3 * k + 2];
        float bx = b[3 * k], by = b[3 * k + 1], bz = b[3 * k + 2];
        M[0] += ax * bx; M[1] += ax * by; M[2] += ax * bz;
        M[3] += ay * bx; M[4] += ay * by; M[5] += ay * bz;
        M[6] += az * bx; M[7] += az * by; M[8] += az * bz;
    }
    return msd_from_m_and_g(M, Ga, Gb, n_atoms, compute_rot, rot);
}

/* theobald_rmsd_generic.h:62-115 restated: planes of length rowstride. */
float o32_msd_axis_major(int n_atoms, int rowstride, const float* aT, const float* bT,
                         float Ga, float Gb) {
    if (aT == bT && Ga == Gb) return 0.0f;
    float M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int k = 0; k < n_atoms; ++k) {
        float a3[3] = {aT[k], aT[rowstride + k], aT[2 * rowstride + k]};
        float b3[3] = {bT[k], bT[rowstride + k], bT[2 * rowstride + k]};
        for (int r = 0; r < 3; ++r)
            for (int c = 0; c < 3; ++c) M[3 * r + c] += a3[r] * b3[c];
    }
    return msd_from_m_and_g(M, Ga, Gb, n_atoms, 0, NULL);
}

/* rotation_generic.h:29-44 restated: x' = x R (row vector times R). */
void o32_rot_atom_major(int n_atoms, float* a, const float* rot) {
    for (int k = 0; k < n_atoms; ++k) {
        float x = a[3 * k], y = a[3 * k + 1], z = a[3 * k + 2];
        a[3 * k] = x * rot[0] + y * rot[3] + z * rot[6];
        a[3 * k + 1] = x * rot[1] + y * rot[4] + z * rot[7];
        a[3 * k + 2] = x * rot[2] + y * rot[5] + z * rot[8];
    }
}

/* rotation_generic.h:47-102 restated: MSD after a given rotation. */
float o32_rot_msd_atom_major(int n_atoms, const float* a, const float* b, const float* rot) {
    double s = 0.0;
    for (int k = 0; k < n_atoms; ++k) {
        float x = a[3 * k], y = a[3 * k + 1], z = a[3 * k + 2];
        float tx = x * rot[0] + y * rot[3] + z * rot[6];
        float ty = x * rot[1] + y * rot[4] + z * rot[7];
        float tz = x * rot[2] + y * rot[5] + z * rot[8];
        float dx = b[3 * k] - tx, dy = b[3 * k + 1] - ty, dz = b[3 * k + 2] - tz;
        s += dx * dx + dy * dy + dz * dz;
    }
    return (float)(s / (double)n_atoms);
}

/* _rmsd.pyx:765-793 restated: plain float MSD, no centering. */
float o32_msd_nosuperpose(int n_atoms, const float* a, const float* b) {
    float s = 0.0f;
    for (int k = 0; k < n_atoms; ++k)
        for (int d = 0; d < 3; ++d) {
            float t = a[3 * k + d] - b[3 * k + d];
            s += t * t;
        }
    return s / n_atoms;
}

/* _rmsd.pyx:197-224 restated on contiguous, already-gathered copies: both
 * arrays are centred in place, then one QCP per frame, OpenMP over frames. */
void o32_rmsd(float* target, int n_frames, int n_atoms, float* ref, int parallel, float* out) {
    float* g = (float*)malloc(sizeof(float) * (size_t)(n_frames > 0 ? n_frames : 1));
    float rg = 0.0f;
    o32_center_and_trace(target, g, n_frames, n_atoms);
    o32_center_and_trace(ref, &rg, 1, n_atoms);
    #pragma omp parallel for schedule(static) if (parallel)
    for (long long i = 0; i < n_frames; ++i) {
        float msd = o32_msd_atom_major(n_atoms, target + i * (long long)n_atoms * 3, ref, g[i], rg, 0, NULL);
        out[i] = sqrtf(msd);
    }
    free(g);
}

/* _rmsd.pyx:620-674 restated: align mobile frames onto target frame
 * (already centred, traces supplied), rotate the displaced coordinates. */
void o32_superpose_atom_major(const float* align_target, const float* align_mobile,
                              const float* g_target, const float* g_mobile,
                              float* displace, int n_frames, int n_align, int n_displace,
                              int target_frame, float* rot_out) {
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < n_frames; ++i) {
        float rot[9];
        o32_msd_atom_major(n_align, align_mobile + i * (long long)n_align * 3,
                           align_target + (long long)target_frame * n_align * 3,
                           g_target[target_frame], g_mobile[i], 1, rot);
        o32_rot_atom_major(n_displace, displace + i * (long long)n_displace * 3, rot);
        if (rot_out) memcpy(rot_out + 9 * i, rot, sizeof(rot));
    }
}
