// Diagnostics: streaming-read kernels (plain and nontemporal float4 loads)
// that measure the HBM read bandwidth this device actually delivers, the
// ceiling the HBM-bound RMSD kernels are compared with.  Not part of the
// RMSD path.
#include "common.h"

namespace hr {
namespace {

template <bool kNT>
__global__ __launch_bounds__(256) void stream_read_kernel(const float4* __restrict__ p,
                                                          long long n4, float* __restrict__ sink) {
    typedef float v4f __attribute__((ext_vector_type(4)));
    float4 acc = make_float4(0.f, 0.f, 0.f, 0.f);
    const long long stride = (long long)gridDim.x * 256;
#pragma unroll 4
    for (long long i = blockIdx.x * 256LL + threadIdx.x; i < n4; i += stride) {
        float4 v;
        if constexpr (kNT) {
            const v4f t = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p) + i);
            v = make_float4(t.x, t.y, t.z, t.w);
        } else {
            v = p[i];
        }
        acc.x += v.x;
        acc.y += v.y;
        acc.z += v.z;
        acc.w += v.w;
    }
    // never true for finite data; keeps the loads alive
    if (acc.x + acc.y + acc.z + acc.w == -1.2345e-38f) sink[0] = acc.x;
}

template <bool kNT>
int read_probe(const void* data, size_t bytes, float* sink, int blocks, void* stream) {
    clear_error();
    HR_REQUIRE(data && sink && blocks > 0 && bytes % 16 == 0,
               "invalid probe arguments (bytes must be a multiple of 16)");
    HR_REQUIRE(reinterpret_cast<uintptr_t>(data) % 16 == 0, "probe data must be 16-byte aligned");
    hipLaunchKernelGGL(stream_read_kernel<kNT>, dim3(blocks), dim3(256), 0, (hipStream_t)stream,
                       static_cast<const float4*>(data), (long long)(bytes / 16), sink);
    return launch_status("stream_read_kernel");
}

}  // namespace
}  // namespace hr

extern "C" int hr_hbm_read_probe(const void* data, size_t bytes, float* sink, int blocks,
                                 void* stream) {
    return hr::read_probe<false>(data, bytes, sink, blocks, stream);
}

extern "C" int hr_hbm_read_probe_nt(const void* data, size_t bytes, float* sink, int blocks,
                                    void* stream) {
    return hr::read_probe<true>(data, bytes, sink, blocks, stream);
}
