/* Driver over the reference's own XTC code (xdrfile.c, xdrfile_xtc.c,
 * compiled from their sources into oracle/_ref/libref_xtc.so).  Test
 * infrastructure: it reads a whole file frame by frame with read_xtc
 * (xdrfile_xtc.c) and writes one with write_xtc, the calls the reference's
 * XTCTrajectoryFile makes (formats/xtc/xtc.pyx:476-480, 595). */
#include <stdlib.h>

#include "xdrfile.h"
#include "xdrfile_xtc.h"

/* Number of atoms of the file's first frame, or -1. */
int refxtc_natoms(const char* fn) {
    int natoms = 0;
    if (read_xtc_natoms((char*)fn, &natoms) != exdrOK) return -1;
    return natoms;
}

/* Decode up to max_frames frames; returns the frame count, or -(status)
 * of the first failing read (other than end of file). */
int refxtc_read(const char* fn, int natoms, int max_frames, float* xyz, float* time, int* step,
                float* box, float* prec) {
    XDRFILE* xd = xdrfile_open(fn, "r");
    if (!xd) return -1000;
    int n = 0;
    while (n < max_frames) {
        int st = read_xtc(xd, natoms, step + n, time + n, (rvec*)(box + 9 * n),
                          (rvec*)(xyz + (size_t)3 * natoms * n), prec + n);
        if (st == exdrENDOFFILE) break;
        if (st != exdrOK) {
            xdrfile_close(xd);
            return -st;
        }
        ++n;
    }
    xdrfile_close(xd);
    return n;
}

/* Write n_frames frames at one precision; returns 0 or the failing status. */
int refxtc_write(const char* fn, int natoms, int n_frames, const float* xyz, const float* time,
                 const int* step, const float* box, float prec) {
    XDRFILE* xd = xdrfile_open(fn, "w");
    if (!xd) return 1000;
    for (int i = 0; i < n_frames; ++i) {
        int st = write_xtc(xd, natoms, step[i], time[i], (rvec*)(box + 9 * i),
                           (rvec*)(xyz + (size_t)3 * natoms * i), prec);
        if (st != exdrOK) {
            xdrfile_close(xd);
            return st;
        }
    }
    xdrfile_close(xd);
    return 0;
}
