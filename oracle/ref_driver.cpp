/*
 * Driver loops around the reference's compiled C++ core — test infrastructure.
 *
 * oracle/Makefile compiles this file together with the reference sources
 * rmsd/src/{theobald_rmsd,center,rotation}.cpp (read in place from the
 * reference tree, never copied) into oracle/_ref/libref_rmsd.so.  The loops
 * below restate the Cython drivers that call that core; the arithmetic is the
 * reference's own.  Only tests/ and bench.py's cpu_baseline leg load it.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "center.h"
#include "rotation.h"
#include "theobald_rmsd.h"

extern "C" {

/* rmsd/_rmsd.pyx:197-224 with separate (gathered, contiguous) target and
 * reference buffers: both are centred in place. */
void ref_rmsd(float* target, int n_frames, int n_atoms, float* ref, int parallel, float* out) {
    float* g = (float*)malloc(sizeof(float) * (size_t)(n_frames > 0 ? n_frames : 1));
    float rg = 0.0f;
    inplace_center_and_trace_atom_major(target, g, n_frames, n_atoms);
    inplace_center_and_trace_atom_major(ref, &rg, 1, n_atoms);
    #pragma omp parallel for schedule(static) if (parallel)
    for (long long i = 0; i < n_frames; ++i) {
        float msd = msd_atom_major(n_atoms, n_atoms, target + i * (long long)n_atoms * 3, ref,
                                   g[i], rg, 0, NULL);
        out[i] = sqrtf(msd);
    }
    free(g);
}

/* rmsd/_rmsd.pyx:197-224 for md.rmsd(t, t, frame) with atom_indices=None:
 * the reference frame is a view into the same buffer, so it is centred twice
 * and the pointer-equality early-out of msd_atom_major can fire. */
void ref_rmsd_self(float* xyz, int n_frames, int n_atoms, int frame, int parallel, float* out) {
    float* g = (float*)malloc(sizeof(float) * (size_t)(n_frames > 0 ? n_frames : 1));
    float rg = 0.0f;
    float* ref = xyz + (long long)frame * n_atoms * 3;
    inplace_center_and_trace_atom_major(xyz, g, n_frames, n_atoms);
    inplace_center_and_trace_atom_major(ref, &rg, 1, n_atoms);
    #pragma omp parallel for schedule(static) if (parallel)
    for (long long i = 0; i < n_frames; ++i) {
        float msd = msd_atom_major(n_atoms, n_atoms, xyz + i * (long long)n_atoms * 3, ref,
                                   g[i], rg, 0, NULL);
        out[i] = sqrtf(msd);
    }
    free(g);
}

/* rmsd/_rmsd.pyx:487-491 */
void ref_center(float* xyz, int n_frames, int n_atoms, float* traces) {
    inplace_center_and_trace_atom_major(xyz, traces, n_frames, n_atoms);
}

/* rmsd/_rmsd.pyx:620-674: centred align arrays and traces supplied. */
void ref_superpose_atom_major(float* align_target, float* align_mobile, float* g_target,
                              float* g_mobile, float* displace, int n_frames, int n_align,
                              int n_displace, int target_frame, float* rot_out) {
    #pragma omp parallel for schedule(static)
    for (long long i = 0; i < n_frames; ++i) {
        float rot[9];
        msd_atom_major(n_align, n_align, align_mobile + i * (long long)n_align * 3,
                       align_target + (long long)target_frame * n_align * 3,
                       g_target[target_frame], g_mobile[i], 1, rot);
        rot_atom_major(n_displace, displace + i * (long long)n_displace * 3, rot);
        if (rot_out) memcpy(rot_out + 9 * i, rot, sizeof(rot));
    }
}

/* One call of the reference QCP core on a supplied M (test of the solver). */
float ref_msd_from_pair(int n_atoms, const float* a, const float* b, float Ga, float Gb,
                        int compute_rot, float* rot) {
    return msd_atom_major(n_atoms, n_atoms, a, b, Ga, Gb, compute_rot, rot);
}

float ref_msd_axis_major(int n_atoms, int rowstride, const float* aT, const float* bT, float Ga,
                         float Gb) {
    return msd_axis_major(n_atoms, n_atoms, rowstride, aT, bT, Ga, Gb);
}

}  // extern "C"
