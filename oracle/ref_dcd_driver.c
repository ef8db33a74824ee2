/* Driver over the reference's own DCD code (dcdplugin.c, compiled from its
 * source into oracle/_ref/libref_dcd.so).  Test infrastructure: it reads a
 * whole file with read_next_timestep and writes one with write_timestep,
 * the calls the reference's DCDTrajectoryFile makes
 * (formats/dcd/dcd.pyx:498-516, 619-636). */
#include <stdlib.h>

#include "dcdplugin.h"

/* Atom and frame counts as open_dcd_read reports them; 0, or -1 when the
 * file does not open. */
int refdcd_info(const char* fn, int* natoms, int* nsets) {
    dcdhandle* h = open_dcd_read(fn, "dcd", natoms, nsets);
    if (!h) return -1;
    close_file_read(h);
    return 0;
}

/* Read frames start, start + stride, ... (at most max_frames) the way
 * DCDTrajectoryFile.read does: skip `start` frames, then read one frame and
 * skip stride - 1.  xyz is (max_frames, natoms, 3); cell is (max_frames, 6)
 * as A, B, C, alpha, beta, gamma.  Returns the frames read (a failed read
 * ends the loop, as end of file does), or -1 when the file does not open. */
int refdcd_read(const char* fn, int start, int stride, int max_frames, float* xyz, float* cell) {
    int natoms = 0, nsets = 0;
    dcdhandle* h = open_dcd_read(fn, "dcd", &natoms, &nsets);
    if (!h) return -1;
    for (int i = 0; i < start; ++i) read_next_timestep(h, natoms, NULL);
    int n = 0;
    molfile_timestep_t ts;
    while (n < max_frames) {
        ts.coords = xyz + (size_t)3 * natoms * n;
        ts.velocities = NULL;
        if (read_next_timestep(h, natoms, &ts) != MOLFILE_SUCCESS) break;
        float* c = cell + 6 * n;
        c[0] = ts.A;
        c[1] = ts.B;
        c[2] = ts.C;
        c[3] = ts.alpha;
        c[4] = ts.beta;
        c[5] = ts.gamma;
        ++n;
        for (int j = 0; j < stride - 1; ++j)
            if (read_next_timestep(h, natoms, NULL) != MOLFILE_SUCCESS) break;
    }
    close_file_read(h);
    return n;
}

/* Write n_frames frames (CHARMM format, with a unit cell when lengths is
 * not NULL); returns 0, or -1 on failure. */
int refdcd_write(const char* fn, int natoms, int n_frames, const float* xyz, const float* lengths,
                 const float* angles) {
    dcdhandle* h = open_dcd_write(fn, "dcd", natoms, lengths != NULL);
    if (!h) return -1;
    molfile_timestep_t ts;
    for (int f = 0; f < n_frames; ++f) {
        ts.coords = (float*)xyz + (size_t)3 * natoms * f;
        ts.velocities = NULL;
        ts.A = ts.B = ts.C = 0.0f;
        ts.alpha = ts.beta = ts.gamma = 90.0f;
        if (lengths) {
            ts.A = lengths[3 * f];
            ts.B = lengths[3 * f + 1];
            ts.C = lengths[3 * f + 2];
            ts.alpha = angles[3 * f];
            ts.beta = angles[3 * f + 1];
            ts.gamma = angles[3 * f + 2];
        }
        if (write_timestep(h, &ts) != MOLFILE_SUCCESS) {
            close_file_write(h);
            return -1;
        }
    }
    close_file_write(h);
    return 0;
}
