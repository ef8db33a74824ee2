// Atom selection of whole device trajectories: out[f, i] = xyz[f, idx[i]].
// Used by the XTC reader, whose decoder must expand every atom of a frame
// (the compressed stream is sequential) before a selection can be taken,
// as the reference's reader does (formats/xtc/xtc.pyx:477-482 decodes each
// frame into a buffer of all atoms, then copies the selected ones).
#include "common.h"

namespace hr {
namespace {

constexpr int kGatherBlock = 256;

// One thread per selected atom of a frame: consecutive threads read atoms in
// index order and write consecutive 12-byte atoms.
__global__ __launch_bounds__(kGatherBlock) void gather_atoms_kernel(
    const float* __restrict__ xyz, long long n_frames, int n_atoms, const int* __restrict__ idx,
    int n_sel, float* __restrict__ out) {
    const long long total = n_frames * (long long)n_sel;
    const long long stride = (long long)gridDim.x * kGatherBlock;
    for (long long t = blockIdx.x * (long long)kGatherBlock + threadIdx.x; t < total; t += stride) {
        const long long f = t / n_sel;
        const int i = (int)(t - f * n_sel);
        int a = idx[i];
        a = a < 0 ? 0 : (a >= n_atoms ? n_atoms - 1 : a);   // callers check; never read outside
        const float* src = xyz + (f * n_atoms + a) * 3;
        float* dst = out + t * 3;
        dst[0] = src[0];
        dst[1] = src[1];
        dst[2] = src[2];
    }
}

}  // namespace
}  // namespace hr

extern "C" int hr_gather_atoms(const float* xyz, int64_t n_frames, int32_t n_atoms,
                               const int32_t* atom_idx, int32_t n_sel, float* out, void* stream) {
    hr::clear_error();
    HR_REQUIRE(n_frames >= 0 && n_atoms > 0 && n_sel >= 0, "invalid sizes n_frames=%lld "
               "n_atoms=%d n_sel=%d", (long long)n_frames, n_atoms, n_sel);
    if (n_frames == 0 || n_sel == 0) return HR_OK;
    HR_REQUIRE(xyz && atom_idx && out, "NULL pointer argument");
    const hr::DeviceInfo* d = hr::current_device();
    if (!d) return HR_ERR_HIP;
    const long long total = n_frames * (long long)n_sel;
    long long blocks = (total + hr::kGatherBlock - 1) / hr::kGatherBlock;
    const long long cap = (long long)d->cu_count * 16;
    if (blocks > cap) blocks = cap;
    hipLaunchKernelGGL(hr::gather_atoms_kernel, dim3((unsigned)blocks), dim3(hr::kGatherBlock), 0,
                       (hipStream_t)stream, xyz, (long long)n_frames, n_atoms, atom_idx, n_sel,
                       out);
    return hr::launch_status("gather_atoms_kernel");
}
