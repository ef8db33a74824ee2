// Diagnostics: a plain streaming-read kernel that measures the HBM read
// bandwidth this device actually delivers (the ceiling the HBM-bound RMSD
// kernels are compared with).  Not part of the RMSD path.
#include "common.h"

namespace hr {
namespace {

__global__ __launch_bounds__(256) void stream_read_kernel(const float4* __restrict__ p,
                                                          long long n4, float* __restrict__ sink) {
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const long long stride = (long long)gridDim.x * 256;
#pragma unroll 4
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += stride) {
        const float4 v = p[i];
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
    }
    // never true for finite data; keeps the loads alive
    if (acc.x + acc.y + acc.z + acc.w == -1.2345e-38f) sink[0] = acc.x;
}

}  // namespace
}  // namespace hr

extern "C" int hr_hbm_read_probe(const void* data, size_t bytes, float* sink, int blocks,
                                 void* stream) {
    hr::clear_error();
    HR_REQUIRE(data && sink && blocks > 0 && bytes % 16 == 0,
               "invalid probe arguments (bytes must be a multiple of 16)");
    hipLaunchKernelGGL(hr::stream_read_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const float4*>(data), (long long)(bytes / 16), sink);
    return hr::launch_status("stream_read_kernel");
}
