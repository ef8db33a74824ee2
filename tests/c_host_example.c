/* A C caller of libhiprmsd with no HIP headers and no torch: the reference's
 * C ABI takes host pointers (rmsd/include/theobald_rmsd.h:16-18), and so do
 * hr_rmsd_one_vs_all_host and the hr_alloc / hr_memcpy_* helpers used here.
 *
 *   c_host_example IN OUT
 * IN  holds int64 F, int32 N, int32 n_sel, then float32 xyz[F][N][3] and
 *     int32 idx[n_sel].
 * OUT receives float32 host_path[F] (hr_rmsd_one_vs_all_host, frame 0 as the
 *     reference) and float32 device_path[F] (hr_alloc + hr_memcpy_h2d +
 *     hr_rmsd_one_vs_all + hr_memcpy_d2h on a library stream).
 * Exit status 0 on success; a library message on stderr otherwise. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "hiprmsd.h"

#define CHECK(call)                                                    \
    do {                                                               \
        int rc_ = (call);                                              \
        if (rc_ != HR_OK) {                                            \
            fprintf(stderr, "%s: status %d: %s\n", #call, rc_, hr_last_error()); \
            return 1;                                                  \
        }                                                              \
    } while (0)

int main(int argc, char** argv) {
    if (argc != 3) return 2;
    FILE* in = fopen(argv[1], "rb");
    if (!in) return 2;
    int64_t F;
    int32_t N, n_sel;
    if (fread(&F, 8, 1, in) != 1 || fread(&N, 4, 1, in) != 1 || fread(&n_sel, 4, 1, in) != 1)
        return 2;
    size_t n_xyz = (size_t)F * N * 3;
    float* xyz = malloc(n_xyz * sizeof(float));
    int32_t* idx = malloc((size_t)n_sel * sizeof(int32_t));
    float* out_host = malloc((size_t)F * sizeof(float));
    float* out_dev = malloc((size_t)F * sizeof(float));
    if (!xyz || !idx || !out_host || !out_dev) return 2;
    if (fread(xyz, sizeof(float), n_xyz, in) != n_xyz ||
        fread(idx, sizeof(int32_t), (size_t)n_sel, in) != (size_t)n_sel)
        return 2;
    fclose(in);

    int n_gpus = 0;
    CHECK(hr_device_count(&n_gpus));
    if (n_gpus < 1) {
        fprintf(stderr, "no GPU visible\n");
        return 3;
    }

    /* 1. host pointers in and out, one synchronous call */
    CHECK(hr_rmsd_one_vs_all_host(xyz, F, N, idx, n_sel, xyz /* frame 0 */, N, idx,
                                  0 /* frame 0 is the reference */, out_host));

    /* 2. device buffers from the library, on a library stream */
    void *d_xyz = NULL, *d_idx = NULL, *d_out = NULL, *d_ws = NULL, *s = NULL;
    CHECK(hr_stream_create(&s));
    CHECK(hr_alloc(&d_xyz, n_xyz * sizeof(float)));
    CHECK(hr_alloc(&d_idx, (size_t)n_sel * sizeof(int32_t)));
    CHECK(hr_alloc(&d_out, (size_t)F * sizeof(float)));
    CHECK(hr_alloc(&d_ws, hr_rmsd_workspace_bytes(n_sel)));
    CHECK(hr_memcpy_h2d(d_xyz, xyz, n_xyz * sizeof(float), s));
    CHECK(hr_memcpy_h2d(d_idx, idx, (size_t)n_sel * sizeof(int32_t), s));
    CHECK(hr_rmsd_one_vs_all((const float*)d_xyz, F, N, (const int32_t*)d_idx, n_sel,
                             (const float*)d_xyz, (const int32_t*)d_idx, 0, (float*)d_out, d_ws,
                             s));
    CHECK(hr_memcpy_d2h(out_dev, d_out, (size_t)F * sizeof(float), s));
    CHECK(hr_stream_synchronize(s));
    CHECK(hr_free(d_ws));
    CHECK(hr_free(d_out));
    CHECK(hr_free(d_idx));
    CHECK(hr_free(d_xyz));
    CHECK(hr_stream_destroy(s));

    FILE* out = fopen(argv[2], "wb");
    if (!out) return 2;
    fwrite(out_host, sizeof(float), (size_t)F, out);
    fwrite(out_dev, sizeof(float), (size_t)F, out);
    fclose(out);
    free(xyz);
    free(idx);
    free(out_host);
    free(out_dev);
    return 0;
}
