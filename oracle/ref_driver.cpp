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

#include <vector>

#include "euclidean_permutation.hpp"

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

extern "C" {

/* rmsd/_lprmsd.pyx:137-228: LP-RMSD of every frame of target (n_frames,
 * n_total, 3) to ref_frame (n_total, 3) over atom_indices[0..n_sel) (sorted,
 * unique), with permute groups given as indices into the selection
 * (group_rel[group_off[g] .. group_off[g + 1]]).  With superpose, every frame
 * of target is centred (all atoms) and rotated in place, as the reference.
 * The arithmetic is the reference's: inplace_center_and_trace_atom_major,
 * msd_atom_major, rot_atom_major, euclidean_permutation (Munkres), sgemm33. */
void ref_lprmsd(float* target, int n_frames, int n_total, const float* ref_frame,
                const int* atom_indices, int n_sel, const int* group_rel, const int* group_off,
                int n_groups, int superpose, float* out) {
    std::vector<std::vector<int> > groups(n_groups);
    std::vector<int> in_group(n_sel, 0);
    for (int g = 0; g < n_groups; ++g)
        for (int k = group_off[g]; k < group_off[g + 1]; ++k) {
            groups[g].push_back(group_rel[k]);
            in_group[group_rel[k]] = 1;
        }
    std::vector<int> dis;
    for (int i = 0; i < n_sel; ++i)
        if (!in_group[i]) dis.push_back(i);
    const int n_dis = (int)dis.size();
    std::vector<float> ref(3 * n_sel), tf(3 * n_sel), tf2(3 * n_sel), td(3 * (n_dis + 1)),
        refd(3 * (n_dis + 1));
    for (int i = 0; i < n_sel; ++i)
        for (int d = 0; d < 3; ++d) ref[3 * i + d] = ref_frame[3 * atom_indices[i] + d];
    float ref_g = 0.0f, ref_g_dis = 0.0f;
    inplace_center_and_trace_atom_major(ref.data(), &ref_g, 1, n_sel);
    if (n_dis > 0) {
        for (int k = 0; k < n_dis; ++k)
            for (int d = 0; d < 3; ++d) refd[3 * k + d] = ref_frame[3 * atom_indices[dis[k]] + d];
        inplace_center_and_trace_atom_major(refd.data(), &ref_g_dis, 1, n_dis);
    }
    for (int i = 0; i < n_frames; ++i) {
        float* fr = target + (size_t)i * n_total * 3;
        for (int a = 0; a < n_sel; ++a)
            for (int d = 0; d < 3; ++d) tf[3 * a + d] = fr[3 * atom_indices[a] + d];
        float tg = 0.0f;
        inplace_center_and_trace_atom_major(tf.data(), &tg, 1, n_sel);
        float rot1[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, rot2[9], rot3[9];
        if (n_dis > 0) {
            for (int k = 0; k < n_dis; ++k)
                for (int d = 0; d < 3; ++d) td[3 * k + d] = tf[3 * dis[k] + d];
            float tgd = 0.0f;
            inplace_center_and_trace_atom_major(td.data(), &tgd, 1, n_dis);
            msd_atom_major(n_dis, n_dis, td.data(), refd.data(), tgd, ref_g_dis, 1, rot1);
            rot_atom_major(n_sel, tf.data(), rot1);
        }
        std::vector<int> mapping = euclidean_permutation(ref.data(), tf.data(), n_sel, 3, groups);
        for (int a = 0; a < n_sel; ++a)
            for (int d = 0; d < 3; ++d) tf2[3 * a + d] = tf[3 * mapping[a] + d];
        float msd;
        if (superpose) {
            msd = msd_atom_major(n_sel, n_sel, tf2.data(), ref.data(), tg, ref_g, 1, rot2);
            inplace_center_and_trace_atom_major(fr, NULL, 1, n_total);
            sgemm33(rot1, rot2, rot3);
            rot_atom_major(n_total, fr, rot3);
        } else {
            msd = msd_atom_major(n_sel, n_sel, ref.data(), tf2.data(), tg, ref_g, 0, NULL);
        }
        out[i] = sqrtf(msd);
    }
}

}  // extern "C" (LP-RMSD)
