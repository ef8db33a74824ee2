/*
 * hiprmsd — C ABI of the MI355X (gfx950) RMSD library.
 *
 * The reference ships its RMSD core as a C ABI over host pointers, one frame
 * pair per call (rmsd/include/theobald_rmsd.h:7-18, rmsd/include/center.h:7-8,
 * rmsd/include/rotation.h:7-12, installed as libtheobald.a via
 * mdtraj/__init__.py:217-225).  This library replaces it with one call per
 * BATCH over device pointers: every pointer argument below is device memory
 * (hipMalloc'd or a torch CUDA tensor's data_ptr()) unless noted, every call
 * is asynchronous on the given HIP stream (NULL = the default stream), and
 * every entry point returns HR_OK (0) or an error code, with a message from
 * hr_last_error() (thread-local).  Nothing is printed by the library.
 *
 * Coordinates are float32, atom-major ("AoS", shape (n_frames, n_atoms, 3),
 * C order) exactly as mdtraj's Trajectory.xyz, or axis-major where stated.
 * Atom selections are int32 device arrays (NULL = all atoms).
 *
 * Precision: per-frame sums (centroid, trace, correlation matrix) are
 * accumulated in FP64 on the GPU and the QCP eigenvalue and the final
 * MSD = (G_a + G_b - 2 lambda) / N are evaluated in FP64; results are
 * rounded to float32 once at the end.
 */
#ifndef HIPRMSD_H_
#define HIPRMSD_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HR_OK 0
#define HR_ERR_INVALID 1     /* bad argument (size, pointer, index range) */
#define HR_ERR_HIP 2         /* a HIP runtime call failed */
#define HR_ERR_UNSUPPORTED 3 /* a configuration the library does not handle */

/* Library version (major*10000 + minor*100 + patch). */
int hr_version(void);

/* Message of the last failed call on this thread ("" if none). */
const char* hr_last_error(void);

/* Select `device` for this thread and cache its properties (CU count, LDS
 * size).  Optional: every entry point initialises the current device lazily. */
int hr_init(int device);

/* Number of compute units of the current device (after hr_init). */
int hr_device_cu_count(void);

/* Tuning knobs (no counterpart in the reference; set by the autotuner in
 * mdtraj_amd/autotune.py).  Values are process-wide; 0 restores the default.
 *   HR_TUNE_FRAMES_WAVES_PER_SIMD  register budget of the per-frame kernel:
 *                                  3 (default, no spills) or 4 waves/SIMD
 *   HR_TUNE_FRAMES_GRID            accepted for compatibility (0, 1, 2);
 *                                  the per-frame kernel's grid is now one
 *                                  workgroup per CU
 *   HR_TUNE_FRAMES_ROW_LANES       lanes per frame of the per-frame kernel:
 *                                  0 (default) chosen by selection size,
 *                                  or 4, 8, 16, 32, 64 (16 down to 1 frames
 *                                  per wave)
 *   HR_TUNE_BATCH_SCHEDULE         work split of batch plans created after
 *                                  the call: 3 (default) a static split of
 *                                  all systems' frames into one contiguous,
 *                                  cost-balanced range per workgroup, each
 *                                  system's reference staged in LDS (up to
 *                                  ~4,400 atoms; the rest of a larger one
 *                                  read from global memory) and the frame
 *                                  sums kept in registers; 1 work units
 *                                  largest first, each wave's first unit its
 *                                  wave index and the rest from an atomic
 *                                  ticket (the tile shape applies to this
 *                                  one only); 2 the static split per wave,
 *                                  the reference read from global memory;
 *                                  4 as 3 with one workgroup per CU (up to
 *                                  ~13,600 reference atoms in LDS) whose
 *                                  waves share each system's frames through
 *                                  an LDS counter
 *   HR_TUNE_SUPERPOSE_CHUNK_MIB    hr_superpose fits and rotates the frames
 *                                  in chunks of about this many MiB of
 *                                  coordinates (a fit pass, then a rotation
 *                                  pass, per chunk); 0 the default (1024), a
 *                                  value at least the trajectory's size one
 *                                  fit pass and one rotation pass over all
 *                                  frames.  Results do not depend on it.
 *   HR_TUNE_HOST_DIRECT_KIB        hr_rmsd_one_vs_all_host: selections of up
 *                                  to this many KiB in all (0: the default,
 *                                  1024) are read by the kernels straight
 *                                  from the pinned staging buffer and the
 *                                  RMSDs written straight to pinned memory,
 *                                  with no device copy either way; -1 never.
 *                                  Results do not depend on it.
 *   HR_TUNE_PAIRS_TILE             split-precision Gram kernels (all-pairs,
 *                                  cross, nearest column): 1 64 x 64 pair
 *                                  tiles, 8 waves, one workgroup per CU;
 *                                  2-5 64 x 32 tiles, 4 waves, two
 *                                  workgroups per CU: 25-slot LDS rows and
 *                                  an epilogue through LDS (2), 26-slot rows
 *                                  (3), the QCP solved from registers (4),
 *                                  both (5, the default).  Results do not
 *                                  depend on it.
 *   HR_TUNE_FRAMES_STATIC_PCT      per-frame kernel (one workgroup per CU):
 *                                  percent of each workgroup's frame sets
 *                                  split evenly over its waves up front (0:
 *                                  the default, 45); the rest goes out in
 *                                  chunks by a counter in LDS, so that the
 *                                  waves a SIMD issues less often take fewer
 *                                  frames; 100: all of them up front.
 *                                  Results do not depend on it.
 *   HR_TUNE_FRAMES_CHUNK_KIB       per-frame kernel: about this many KiB of
 *                                  frames per chunk (0: the default, 32), in
 *                                  whole frame sets of at most 64 frames.
 *                                  Results do not depend on it. */
#define HR_TUNE_FRAMES_WAVES_PER_SIMD 0
#define HR_TUNE_FRAMES_GRID 1
#define HR_TUNE_FRAMES_ROW_LANES 2
#define HR_TUNE_BATCH_SCHEDULE 3
#define HR_TUNE_SUPERPOSE_CHUNK_MIB 4
#define HR_TUNE_HOST_DIRECT_KIB 5
#define HR_TUNE_PAIRS_TILE 6
#define HR_TUNE_FRAMES_STATIC_PCT 7
#define HR_TUNE_FRAMES_CHUNK_KIB 8
#define HR_TUNE_COUNT 9
int hr_set_tuning(int knob, int value);
int hr_get_tuning(int knob);

/* ---------------------------------------------------------------------------
 * Centering (replaces inplace_center_and_trace_atom_major,
 * rmsd/include/center.h:7-8; rmsd/src/center_generic.h:3-44).
 * In place on xyz (n_frames, n_atoms, 3): per frame, mean accumulated in
 * double and cast to float, subtracted in float; traces[f] = sum of squares of
 * the centred float coordinates accumulated in double, cast to float.
 * traces may be NULL.
 * ------------------------------------------------------------------------- */
int hr_center_and_trace(float* xyz, float* traces, int64_t n_frames, int32_t n_atoms,
                        void* stream);

/* ---------------------------------------------------------------------------
 * One-vs-all RMSD (replaces the frame loop of md.rmsd, rmsd/_rmsd.pyx:186-224,
 * which calls msd_atom_major (rmsd/include/theobald_rmsd.h:16-18) per frame
 * after inplace_center_and_trace_atom_major on copies of the data).
 *
 *   target      (n_frames, n_atoms_total, 3) float32, not modified
 *   atom_idx    int32[n_sel] into the target's atoms, or NULL (all atoms,
 *               then n_sel must equal n_atoms_total)
 *   ref_frame   (n_ref_atoms_total, 3) float32: the reference conformation
 *   ref_idx     int32[n_sel] into the reference's atoms, or NULL (then the
 *               reference has n_sel atoms used in order)
 *   self_frame  index of a target frame that IS the reference frame (same
 *               buffer, same selection): its RMSD is written as exactly 0, as
 *               the reference's pointer-equality early-out does
 *               (rmsd/src/theobald_rmsd_sse.h:256-262); -1 for none
 *   out         float32[n_frames] RMSDs in nm
 *   workspace   device buffer of hr_rmsd_workspace_bytes(n_sel) bytes
 *
 * Targets are not centred in place (the reference mutates its inputs; this
 * library never does).
 * ------------------------------------------------------------------------- */
size_t hr_rmsd_workspace_bytes(int32_t n_sel);
int hr_rmsd_one_vs_all(const float* target, int64_t n_frames, int32_t n_atoms_total,
                       const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                       const int32_t* ref_idx, int64_t self_frame, float* out, void* workspace,
                       void* stream);

/* The two halves of hr_rmsd_one_vs_all, for callers that score many target
 * batches against one reference: prepare gathers the reference selection into
 * `workspace` and sums it in FP64 (one small kernel); the prepared call then
 * runs the per-frame kernel only, reading the reference from `workspace`. */
int hr_rmsd_prepare_reference(const float* ref_frame, const int32_t* ref_idx, int32_t n_sel,
                              void* workspace, void* stream);
int hr_rmsd_one_vs_all_prepared(const float* target, int64_t n_frames, int32_t n_atoms_total,
                                const int32_t* atom_idx, int32_t n_sel, int64_t self_frame,
                                float* out, const void* workspace, void* stream);

/* hr_rmsd_one_vs_all with no allocation and no workspace argument.  No
 * counterpart in the reference: it is for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The workspace lives in a static per-device array
 * of 786,688 bytes, so selections are limited to
 * hr_rmsd_one_vs_all_noalloc_max_sel() = 65,536 atoms (larger ones fail with
 * HR_ERR_INVALID).  Every call on one GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, or two threads) race. */
size_t hr_rmsd_one_vs_all_noalloc_max_sel(void);
int hr_rmsd_one_vs_all_noalloc(const float* target, int64_t n_frames, int32_t n_atoms_total,
                               const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                               const int32_t* ref_idx, int64_t self_frame, float* out,
                               void* stream);

/* Same, for axis-major targets (n_frames, 3, rowstride) — replaces
 * msd_axis_major (rmsd/include/theobald_rmsd.h:13-14) as driven by
 * getMultipleRMSDs_axis_major (rmsd/_rmsd.pyx:500-557).  ref_frame_axis is
 * (3, ref_rowstride).  No selection: the first n_atoms of each row are used. */
int hr_rmsd_one_vs_all_axis_major(const float* target, int64_t n_frames, int32_t n_atoms,
                                  int32_t rowstride, const float* ref_frame_axis,
                                  int32_t ref_rowstride, int64_t self_frame, float* out,
                                  void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * RMSD without superposition (md.rmsd(..., superpose=False);
 * rmsd/_rmsd.pyx:228-241,765-793): sqrt(mean |a_i - b_i|^2), no centering,
 * accumulated in FP64.  Arguments as hr_rmsd_one_vs_all (workspace of
 * hr_rmsd_workspace_bytes(n_sel) bytes).
 * ------------------------------------------------------------------------- */
int hr_rmsd_nosuperpose(const float* target, int64_t n_frames, int32_t n_atoms_total,
                        const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                        const int32_t* ref_idx, float* out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Superposition (replaces Trajectory.superpose, core/trajectory.py:1083-1173,
 * and the low-level superpose_atom_major, rmsd/_rmsd.pyx:620-674, which call
 * msd_atom_major(..., computeRot=1) and rot_atom_major,
 * rmsd/include/rotation.h:7).
 *   xyz         (n_frames, n_atoms, 3) float32, moved in place
 *   align_xyz   (n_frames, n_align_atoms, 3) frames the fit is computed on,
 *               or NULL to fit on xyz itself (then n_align_atoms = n_atoms)
 *   atom_idx    int32[n_sel] into align_xyz's atoms, or NULL (all)
 *   ref_frame   (n_ref_atoms, 3), ref_idx int32[n_sel] or NULL
 *   translate   1: x' = (x - c_f) R + c_ref with c_f the centroid of frame f's
 *               fitted subset and c_ref the reference subset's centroid
 *               (Trajectory.superpose); 0: x' = x R (superpose_atom_major,
 *               which expects centred inputs)
 *   rot_out     float32 (n_frames, 3, 3) row-major, applied as x' = x R, or NULL
 *   rmsd_out    float32[n_frames] fit RMSDs, or NULL
 *   workspace   hr_superpose_workspace_bytes(n_frames, n_sel) bytes
 * ------------------------------------------------------------------------- */
size_t hr_superpose_workspace_bytes(int64_t n_frames, int32_t n_sel);
int hr_superpose(float* xyz, int64_t n_frames, int32_t n_atoms, const float* align_xyz,
                 int32_t n_align_atoms, const int32_t* atom_idx, int32_t n_sel,
                 const float* ref_frame, const int32_t* ref_idx, int32_t translate,
                 float* rot_out, float* rmsd_out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Align/displace RMSD (replaces getMultipleAlignDisplaceRMSDs_atom_major,
 * rmsd/_rmsd.pyx:677-759, which calls msd_atom_major(..., computeRot=1) and
 * rot_msd_atom_major, rmsd/include/rotation.h:9-10): for every frame f, the
 * rotation R_f that best fits align_ref onto align[f] (centroids removed;
 * the reference expects centred alignment sets), then the RMS displacement
 * of displ_ref R_f from displ[f], about the origin (no centering), FP64.
 *   align       (n_frames, align_stride, 3), atoms 0..n_align-1 used
 *   align_ref   (n_align, 3)
 *   displ       (n_frames, displ_stride, 3), atoms 0..n_displ-1 used
 *   displ_ref   (n_displ, 3)
 *   rmsd_out    float32[n_frames]
 *   rot_out     float32 (n_frames, 3, 3) R_f row-major (x' = x R), or NULL
 *   workspace   hr_align_displace_workspace_bytes(n_frames, n_align) bytes
 * ------------------------------------------------------------------------- */
size_t hr_align_displace_workspace_bytes(int64_t n_frames, int32_t n_align);
int hr_rmsd_align_displace(const float* align, int64_t n_frames, int32_t align_stride,
                           int32_t n_align, const float* align_ref, const float* displ,
                           int32_t displ_stride, int32_t n_displ, const float* displ_ref,
                           float* rmsd_out, float* rot_out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Per-atom RMS fluctuation (replaces the kernel loops of md.rmsf,
 * rmsd/_rmsd.pyx:246-443: msd_atom_major + rot_atom_major per frame, then
 * the mean structure and the fluctuation about it).
 *   xyz         (n_frames, n_atoms, 3); atom_idx int32[n_sel] or NULL (all)
 *   ref_frame   (n_ref_atoms, 3) with ref_idx int32[n_sel] or NULL: every
 *               frame's selection is fitted onto it and rotated by the fit;
 *               ref_frame NULL: frames already aligned (no rotation)
 *   center      1: the fluctuating coordinates are each frame's selection
 *               centred on its centroid; 0: the selection as stored (rotated
 *               about the origin when fitted), as md.rmsf does when it is
 *               given atom_indices
 *   out         float32[n_sel]: sqrt(mean_f |x_f - <x>|^2), FP64 sums
 *   workspace   hr_rmsf_workspace_bytes(n_frames, n_sel) bytes
 * ------------------------------------------------------------------------- */
size_t hr_rmsf_workspace_bytes(int64_t n_frames, int32_t n_sel);
int hr_rmsf(const float* xyz, int64_t n_frames, int32_t n_atoms, const int32_t* atom_idx,
            int32_t n_sel, const float* ref_frame, const int32_t* ref_idx, int32_t center,
            float* out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * All-pairs RMSD matrix.  The reference has no such routine; its users loop
 * md.rmsd(traj, traj, i) over i (examples/clustering.ipynb:80-81), so row i
 * of this matrix equals one-vs-all with reference frame i.
 *   mode        HR_PAIRS_FP32       plain FP32 FMA Gram
 *               HR_PAIRS_F16X3      FP16 hi+lo split, three MFMA products
 *                                   (hi.hi + hi.lo + lo.hi) in one FP32 accumulator
 *               HR_PAIRS_BF16X3     the same with BF16
 *               HR_PAIRS_F16_EXACT  FP16 split with a fixed-point integer hi whose
 *                                   hi.hi products accumulate exactly in FP32, and
 *                                   hi.lo + lo.hi + lo.lo in a second, small FP32
 *                                   accumulator (four MFMA products; the accurate
 *                                   mode: <= 1e-5 nm vs FP64)
 *               HR_PAIRS_BF16_EXACT the same with BF16 lo parts
 *   out         float32 condensed upper triangle (i < j, row-major, the layout
 *               of scipy.spatial.distance.squareform), n_frames*(n_frames-1)/2
 * ------------------------------------------------------------------------- */
#define HR_PAIRS_FP32 0
#define HR_PAIRS_F16X3 1
#define HR_PAIRS_BF16X3 2
#define HR_PAIRS_F16_EXACT 3
#define HR_PAIRS_BF16_EXACT 4
size_t hr_rmsd_all_pairs_workspace_bytes(int64_t n_frames, int32_t n_sel, int32_t mode);
int hr_rmsd_all_pairs(const float* xyz, int64_t n_frames, int32_t n_atoms_total,
                      const int32_t* atom_idx, int32_t n_sel, int32_t mode, float* out,
                      void* workspace, size_t workspace_bytes, void* stream);

/* ---------------------------------------------------------------------------
 * Cross RMSD between two frame sets through the split-precision Gram (no
 * counterpart in the reference; it is md.rmsd(target, queries, q) for every
 * query frame q, in one pass over the target).  A set is first prepared into
 * a self-contained block (centred, split, traces and scales) of
 * hr_split_bytes(n_frames, n_sel, mode) bytes; prepared blocks can be kept
 * resident (serving pools) and reused by any number of cross calls.
 *   out[i * out_ld + j] = RMSD(set A frame i, set B frame j), float32
 * Each entry depends only on its two frames, never on the other frames in
 * the call, so results do not change with how queries are batched.
 * ------------------------------------------------------------------------- */
size_t hr_split_bytes(int64_t n_frames, int32_t n_sel, int32_t mode);
int hr_split_prepare(const float* xyz, int64_t n_frames, int32_t n_atoms_total,
                     const int32_t* atom_idx, int32_t n_sel, int32_t mode, void* block,
                     size_t block_bytes, void* stream);
int hr_rmsd_cross(const void* block_a, int64_t n_a, const void* block_b, int64_t n_b,
                  int32_t n_sel, int32_t mode, float* out, int64_t out_ld, void* stream);

/* ---------------------------------------------------------------------------
 * Consumers of the RMSD matrix (clustering), fused or on the condensed output
 *
 * hr_rmsd_cross_argmin replaces the reference's leader assignment loop
 * `leader_labels[np.argmin(md.rmsd(leaders, frame, 0))]` over every frame
 * (examples/two-pass-clustering.ipynb, cell 6): for each frame i of set A,
 * the nearest frame of set B, index[i] = argmin_j RMSD(A_i, B_j) (the lowest
 * j on ties, as np.argmin) and value[i] = that RMSD, with the minimum taken
 * inside the cross kernel (no n_a x n_b matrix).  The workspace holds 8 n_a
 * bytes (hr_rmsd_cross_argmin_workspace_bytes); no allocation.
 *
 * hr_centroid_scores replaces examples/centroids.ipynb, cells 3-4:
 *   scores[i] = sum_j exp(-beta d_ij / std(d)),
 * d the full symmetric n x n RMSD matrix given as its condensed upper
 * triangle (hr_rmsd_all_pairs' output, device, 16-byte aligned), d_ii = 0,
 * std over all n^2 entries
 * (numpy's population std); FP64 throughout; the centroid is the argmax.
 * stats (optional, 3 doubles) receives sum(d_ij, i<j), sum(d_ij^2, i<j) and
 * the std.  Workspace: hr_centroid_workspace_bytes(n).
 * ------------------------------------------------------------------------- */
size_t hr_rmsd_cross_argmin_workspace_bytes(int64_t n_a);
int hr_rmsd_cross_argmin(const void* block_a, int64_t n_a, const void* block_b, int64_t n_b,
                         int32_t n_sel, int32_t mode, int32_t* index, float* value,
                         void* workspace, void* stream);
size_t hr_centroid_workspace_bytes(int64_t n_frames);
int hr_centroid_scores(const float* condensed, int64_t n_frames, double beta, double* scores,
                       double* stats, void* workspace, void* stream);

/* hr_topk_rows: the k nearest columns of every row of a float32 matrix (an
 * hr_rmsd_cross output: frames x leaders), generalising the argmin of the
 * reference's leader assignment (examples/two-pass-clustering.ipynb, cell 6,
 * `np.argmin(md.rmsd(leaders, frame, 0))`) to its k smallest entries:
 * index[i k + t], value[i k + t] for t = 0 .. k-1 in ascending order, ties to
 * the lower column, NaN last -- np.argsort(row, kind="stable")[:k].  Device
 * pointers; row i starts at matrix + i ld; 1 <= k <= n_cols < 2^31.  No
 * allocation. */
int hr_topk_rows(const float* matrix, int64_t n_rows, int64_t n_cols, int64_t ld, int32_t k,
                 int32_t* index, float* value, void* stream);

/* ---------------------------------------------------------------------------
 * Batched one-vs-all over many independent systems of uneven size (no
 * counterpart in the reference, whose users loop md.rmsd per trajectory).
 * A plan captures, once, a hipGraph: reference prep (one workgroup per
 * system), then the scoring kernel of the work split HR_TUNE_BATCH_SCHEDULE
 * selects when the plan is created (default: one cost-balanced range of all
 * systems' frames per workgroup, each system's reference staged in LDS; the
 * ticket schedule adds a ticket reset kernel first and has its waves pull
 * (system, frame chunk) units, largest first, from an atomic ticket).  Each
 * launch is one graph launch.  Plan creation allocates and copies the
 * descriptor, unit and range tables; launches allocate nothing.  The
 * systems' buffers must stay valid for the plan's lifetime.  Do not launch
 * one plan on two streams at once.
 *   shape            work-unit shape of the ticket schedule (NULL =
 *                    defaults; waves_per_simd applies to every schedule);
 *                    the autotuner in mdtraj_amd/autotune.py picks the
 *                    schedule and shape per workload:
 *     frames_per_unit  frames per work unit, 1..64 (0 = 64)
 *     unit_kib         if > 0, a system's units hold fewer frames (a multiple
 *                      of 4, at least 4) when frames_per_unit frames would
 *                      exceed unit_kib KiB of coordinates
 *     waves_per_simd   register budget of the scoring kernel: 3 (0 = 3) or 4
 * ------------------------------------------------------------------------- */
typedef struct hr_batch_shape {
    int32_t frames_per_unit;
    int32_t unit_kib;
    int32_t waves_per_simd;
} hr_batch_shape;
typedef struct hr_system_desc {
    const float* xyz;  /* (n_frames, n_atoms, 3) float32, device */
    int64_t n_frames;
    int32_t n_atoms;
    int32_t ref_frame; /* scored against this frame of the same system (RMSD 0 there) */
    float* out;        /* float32[n_frames], device */
} hr_system_desc;
typedef struct hr_batch_plan hr_batch_plan;
int hr_batch_plan_create(const hr_system_desc* systems, int32_t n_systems,
                         const hr_batch_shape* shape, hr_batch_plan** plan);
int hr_batch_plan_launch(hr_batch_plan* plan, void* stream);
int hr_batch_plan_destroy(hr_batch_plan* plan);
int64_t hr_batch_plan_units(const hr_batch_plan* plan);

/* ---------------------------------------------------------------------------
 * XTC ingestion: trajectory files decoded on the device
 *
 * Replaces the reference's frame loop over read_xtc
 * (formats/xtc/xtc.pyx:424-517 -> formats/xtc/src/xdrfile_xtc.c:41-90, the
 * header and box, and formats/xtc/src/xdrfile.c:778-1000,
 * xdrfile_decompress_coord_float): the file's bytes are copied to HBM once
 * and every frame is decoded in parallel into float32 (F, N, 3).
 * ------------------------------------------------------------------------- */
/* Per-frame status words of hr_xtc_decode. */
#define HR_XTC_OK 0
#define HR_XTC_BAD_MAGIC 1  /* not an XTC frame header */
#define HR_XTC_ATOMS 2      /* atom count differs from n_atoms */
#define HR_XTC_TRUNCATED 3  /* the frame's bytes end early */
#define HR_XTC_STREAM 4     /* malformed bit stream (the reference would misread) */

/* Frame index of XTC bytes in HOST memory: walks the frame headers
 * (xdrfile_xtc.c:41-90), writes the byte offset of the first max_frames
 * frames (offsets may be NULL with max_frames 0 to count them), and returns
 * the frame count and the atom count.  A file that is not XTC, changes its
 * atom count or ends inside a frame fails with HR_ERR_INVALID. */
int hr_xtc_scan(const void* data, int64_t n_bytes, int64_t* offsets, int64_t max_frames,
                int64_t* n_frames, int32_t* n_atoms);
/* Decode n_frames frames of XTC bytes in DEVICE memory (4-byte aligned)
 * starting at the device offsets[] (any subset of the frames, e.g. a
 * stride): xyz (n_frames, n_atoms, 3), time, step, box (n_frames, 3, 3),
 * precision (0 for uncompressed frames of <= 9 atoms) and status
 * (n_frames HR_XTC_* words), all device arrays.  Results are bit-identical
 * to the reference's reader. */
int hr_xtc_decode(const void* data, int64_t n_bytes, const int64_t* offsets, int64_t n_frames,
                  int32_t n_atoms, float* xyz, float* time, int32_t* step, float* box,
                  float* prec, int32_t* status, void* stream);

/* ---------------------------------------------------------------------------
 * DCD ingestion: CHARMM / NAMD / X-PLOR trajectory files decoded on the device
 *
 * Replaces the reference's frame loop over read_next_timestep
 * (formats/dcd/dcd.pyx:463-551 -> formats/dcd/src/dcdplugin.c:935-1013,
 * read_dcdstep :481-623): the file's bytes are copied to HBM once and every
 * selected atom of every requested frame is written as float32 (F, n_sel, 3).
 * ------------------------------------------------------------------------- */
/* hr_dcd_header.flags, the reference's feature flags (dcdplugin.h) */
#define HR_DCD_CHARMM 0x01      /* CHARMM-style header (a version number is set) */
#define HR_DCD_FOUR_DIMS 0x02   /* a fourth coordinate record follows Z (skipped) */
#define HR_DCD_EXTRA_BLOCK 0x04 /* every frame starts with a unit-cell record */
/* Per-frame status words of hr_dcd_decode. */
#define HR_DCD_OK 0
#define HR_DCD_BAD_RECORD 1 /* a record length marker is not what the header implies */
#define HR_DCD_TRUNCATED 2  /* the frame lies outside the bytes, or is not in the file */

typedef struct hr_dcd_header {
    int32_t n_atoms;           /* N */
    int32_t n_fixed;           /* NAMNF: atoms stored only in frame 0 */
    int64_t n_frames;          /* frames in the file, from its size (dcdplugin.c:870-905) */
    int64_t first_frame;       /* byte offset of frame 0 */
    int64_t first_frame_bytes; /* bytes of frame 0 (every atom) */
    int64_t frame_bytes;       /* bytes of every later frame (the free atoms) */
    int32_t flags;             /* HR_DCD_* */
    int32_t reverse_endian;    /* 1 when the file's byte order is not the host's */
    int32_t n_set;             /* NSET as the header states it */
    int32_t istart;            /* ISTART */
    int32_t nsavc;             /* NSAVC */
    int32_t pad;
    double delta;              /* DELTA (a float in CHARMM headers, a double in X-PLOR ones) */
} hr_dcd_header;

/* Header of DCD bytes in HOST memory (read_dcdheader, dcdplugin.c:98-374):
 * fills *header, and the 1-based free-atom indices (n_atoms - n_fixed of
 * them; free_indexes may be NULL with max_free 0).  n_bytes is the whole
 * file's size: the frame count follows from it, as in the reference.  Files
 * that are not DCD, or hold no frame, fail with HR_ERR_INVALID; CHARMM -i8
 * files (64-bit record markers) with HR_ERR_UNSUPPORTED. */
int hr_dcd_scan(const void* data, int64_t n_bytes, hr_dcd_header* header, int32_t* free_indexes,
                int32_t max_free);
/* Decode frames[0..n_frames) (device int64 frame numbers, any subset of the
 * file's frames) of DCD bytes in DEVICE memory (4-byte aligned):
 * xyz (n_frames, n_sel, 3) = the atoms atom_indices[0..n_sel) (device int32;
 * NULL for all atoms, n_sel == n_atoms) times `scale` in float32 (1 keeps
 * the file's angstroms; 0.1f gives nanometres exactly as the reference's
 * in-place conversion); cell (n_frames, 6) float32, may be NULL: the raw
 * record (A, cos gamma, B, cos beta, cos alpha, C) cast to float, or
 * (0, 90, 0, 90, 90, 0) without one; status (n_frames HR_DCD_* words).
 * With fixed atoms, free_slot (device int32[n_atoms]) gives each atom's
 * position in the free list, or -1 for a fixed atom.  Atom indices outside
 * [0, n_atoms) give NaN coordinates. */
int hr_dcd_decode(const void* data, int64_t n_bytes, const hr_dcd_header* header,
                  const int64_t* frames, int64_t n_frames, const int32_t* atom_indices,
                  int32_t n_sel, const int32_t* free_slot, float scale, float* xyz, float* cell,
                  int32_t* status, void* stream);

/* Atom selection of a decoded device trajectory (the reader's copy of the
 * selected atoms after decoding whole frames, formats/xtc/xtc.pyx:477-482):
 * out (n_frames, n_sel, 3) = xyz (n_frames, n_atoms, 3) at atoms
 * atom_idx[0..n_sel) (device int32; indices are clamped into range). */
int hr_gather_atoms(const float* xyz, int64_t n_frames, int32_t n_atoms, const int32_t* atom_idx,
                    int32_t n_sel, float* out, void* stream);

/* ---------------------------------------------------------------------------
 * LP-RMSD: the RMSD minimised over rotations and over relabellings of atoms
 * inside permute groups
 *
 * Replaces the frame loop of md.lprmsd (rmsd/_lprmsd.pyx:137-228) and the
 * assignment of euclidean_permutation (rmsd/src/euclidean_permutation.cpp:
 * 17-71, Munkres): per frame, the selection is centred; when some selected
 * atoms are in no group, they alone are fitted and the selection rotated;
 * each group's labels are assigned by an exact minimum-cost assignment of
 * squared distances; the RMSD is the QCP RMSD of the relabelled selection.
 * atom_indices (n_sel, sorted, unique), dis (the n_dis positions in the
 * selection of atoms in no group) and the groups (positions in the
 * selection, group g at group_atoms[group_offsets[g] .. group_offsets[g+1]),
 * max_group the largest) are device arrays; ref_frame is the reference
 * frame (n_atoms, 3), device.  With superpose, every frame of xyz is centred
 * (all atoms) and rotated in place, as the reference.  A selection and group
 * too large for one wave's LDS (64 KiB) fail with HR_ERR_UNSUPPORTED.
 * ------------------------------------------------------------------------- */
size_t hr_lprmsd_workspace_bytes(int32_t n_sel, int32_t n_dis);
int hr_lprmsd(float* xyz, int64_t n_frames, int32_t n_atoms, const float* ref_frame,
              const int32_t* atom_indices, int32_t n_sel, const int32_t* dis, int32_t n_dis,
              const int32_t* group_atoms, const int32_t* group_offsets, int32_t n_groups,
              int32_t max_group, int32_t superpose, float* out, void* workspace, void* stream);

/* ---------------------------------------------------------------------------
 * Memory, copies and streams, for callers without a device runtime of their
 * own.  The reference's C ABI takes host pointers and allocates nothing
 * (rmsd/include/theobald_rmsd.h:7-18); the device-pointer entry points above
 * need device buffers, which these provide without HIP headers or torch.
 *   hr_device_count   number of visible GPUs (0, and HR_OK, when none)
 *   hr_alloc/hr_free  device memory (hipMalloc/hipFree; hr_free waits for
 *                     the device, as hipFree does)
 *   hr_alloc_host/hr_free_host  pinned host memory (hipHostMalloc)
 *   hr_memcpy_*       copies on `stream` (NULL = the default stream):
 *                     asynchronous between device and pinned memory; with
 *                     pageable host memory they return when the copy is done
 *   hr_memset         asynchronous byte fill on `stream`
 *   hr_stream_*       a non-blocking stream of default priority
 * ------------------------------------------------------------------------- */
int hr_device_count(int* count);
int hr_alloc(void** ptr, size_t nbytes);
int hr_free(void* ptr);
int hr_alloc_host(void** ptr, size_t nbytes);
int hr_free_host(void* ptr);
int hr_memcpy_h2d(void* dst, const void* src, size_t nbytes, void* stream);
int hr_memcpy_d2h(void* dst, const void* src, size_t nbytes, void* stream);
int hr_memcpy_d2d(void* dst, const void* src, size_t nbytes, void* stream);
int hr_memset(void* dst, int value, size_t nbytes, void* stream);
int hr_stream_create(void** stream);
int hr_stream_destroy(void* stream);
int hr_stream_synchronize(void* stream);

/* ---------------------------------------------------------------------------
 * One-vs-all RMSD over HOST memory (replaces the reference's host-pointer
 * path: the gather of rmsd/_rmsd.pyx:197,199 and its frame loop
 * rmsd/_rmsd.pyx:186-224 over msd_atom_major, rmsd/include/theobald_rmsd.h:
 * 16-18).  Every pointer is host memory owned by the caller; the call is
 * synchronous and returns when out holds the RMSDs.
 *   target      (n_frames, n_atoms_total, 3) float32, not modified
 *   atom_idx    int32[n_sel] into the target's atoms, or NULL (all atoms)
 *   ref_frame   (n_ref_atoms_total, 3) float32
 *   ref_idx     int32[n_sel] into the reference's atoms, or NULL (its first
 *               n_sel atoms)
 *   self_frame  target frame that IS the reference frame (exact 0), or -1
 *   out         float32[n_frames]
 * Indices are checked against the atom counts (HR_ERR_INVALID).  Only the
 * selected atoms are copied to the device, in pinned chunks gathered by a
 * worker pool while earlier chunks are copied and scored; staging buffers are
 * kept per device between calls.  Calls on one device run one at a time.
 * ------------------------------------------------------------------------- */
int hr_rmsd_one_vs_all_host(const float* target, int64_t n_frames, int32_t n_atoms_total,
                            const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                            int32_t n_ref_atoms_total, const int32_t* ref_idx,
                            int64_t self_frame, float* out);

/* ---------------------------------------------------------------------------
 * Diagnostics (no counterpart in the reference): read `bytes` of device
 * memory (16-byte aligned, a multiple of 16 bytes) with a float4 streaming
 * kernel of `blocks` x 256 threads, with plain or nontemporal loads; the
 * achieved rate is the HBM read ceiling the RMSD kernels are compared with.
 * ------------------------------------------------------------------------- */
int hr_hbm_read_probe(const void* data, size_t bytes, float* sink, int blocks, void* stream);
int hr_hbm_read_probe_nt(const void* data, size_t bytes, float* sink, int blocks, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* HIPRMSD_H_ */
