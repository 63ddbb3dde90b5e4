// copy.hip -- plain device copy: the bandwidth / counter reference that the
// profiler figures of the other kernels are scaled against.
#include "common.hpp"

namespace th {

__global__ __launch_bounds__(256) void k_copy16(float4* __restrict__ dst, const float4* __restrict__ src, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

int launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(n, 256), 8192);
    ProfScope prof("copy16", st);
    k_copy16<<<grid, 256, 0, st>>>((float4*)dst, (const float4*)src, n);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

}  // namespace th
