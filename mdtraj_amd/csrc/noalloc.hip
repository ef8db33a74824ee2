// One-vs-all RMSD with no allocation and no caller workspace, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It runs the kernels of
// hr_rmsd_one_vs_all with their workspace (FP64 reference sums and the
// gathered reference planes) in `scratch`, a static device array: each GPU
// holds its own copy, at its own address, from the first call on that GPU
// until the process ends.  The array is reached only through the address
// hipGetSymbolAddress returns for the current device, never through its
// host-side name.  Every call on one GPU shares that GPU's array, and the
// reference preparation rewrites all of it that the per-frame kernel reads,
// so nothing must start at zero; calls that overlap on one GPU (different
// streams, or threads) race on it.
#include <mutex>

#include "common.h"

// Selections of up to kScratchMaxSel atoms: 256 bytes of reference sums plus
// 12 bytes per selected atom, 786,688 bytes in all.
static constexpr int kScratchMaxSel = 65536;
static constexpr size_t kScratchSumsBytes = 256;
static constexpr size_t kScratchBytes =
    kScratchSumsBytes + (size_t)kScratchMaxSel * 3 * sizeof(float);

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's `scratch`, looked up once per device.
int scratch_address(void** out) {
    int dev = 0;
    HR_CHECK_HIP(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        hr::set_error("device index %d out of range", dev);
        return HR_ERR_HIP;
    }
    std::lock_guard<std::mutex> g(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HR_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            hr::set_error("hipGetSymbolAddress returned NULL for the scratch array");
            return HR_ERR_HIP;
        }
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return HR_OK;
}

}  // namespace

extern "C" {

size_t hr_rmsd_one_vs_all_noalloc_max_sel(void) { return (size_t)kScratchMaxSel; }

int hr_rmsd_one_vs_all_noalloc(const float* target, int64_t n_frames, int32_t n_atoms_total,
                               const int32_t* atom_idx, int32_t n_sel, const float* ref_frame,
                               const int32_t* ref_idx, int64_t self_frame, float* out,
                               void* stream) {
    hr::clear_error();
    HR_REQUIRE(n_sel > 0 && n_sel <= kScratchMaxSel,
               "hr_rmsd_one_vs_all_noalloc handles 1..%d selected atoms (got %d); use "
               "hr_rmsd_one_vs_all with a workspace for larger selections", kScratchMaxSel,
               n_sel);
    HR_REQUIRE(hr_rmsd_workspace_bytes(n_sel) <= kScratchBytes,
               "workspace of %zu bytes exceeds the %zu-byte scratch array",
               hr_rmsd_workspace_bytes(n_sel), kScratchBytes);
    void* ws = nullptr;
    int rc = scratch_address(&ws);
    if (rc) return rc;
    return hr_rmsd_one_vs_all(target, n_frames, n_atoms_total, atom_idx, n_sel, ref_frame, ref_idx,
                              self_frame, out, ws, stream);
}

}  // extern "C"
