// copy.hip -- calibration kernels: a plain device copy (the bandwidth and
// profiler-counter reference the other kernels are scaled against) and an
// FP32 FMA loop (the achievable unpacked FP32 VALU rate).
#include "common.hpp"

namespace th {

// Each thread moves kCopyUnroll 16-byte vectors (block-strided, so every
// load instruction of a wave reads 1 KiB contiguous), all loads issued
// before the first store.
constexpr int kCopyUnroll = 4;
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy16(f32x4* __restrict__ dst, const f32x4* __restrict__ src, size_t n) {
    const size_t base = (size_t)blockIdx.x * (256 * kCopyUnroll) + threadIdx.x;
    f32x4 r[kCopyUnroll];
#pragma unroll
    for (int j = 0; j < kCopyUnroll; j++) {
        const size_t i = base + (size_t)j * 256;
        if (i < n) r[j] = __builtin_nontemporal_load(&src[i]);
    }
#pragma unroll
    for (int j = 0; j < kCopyUnroll; j++) {
        const size_t i = base + (size_t)j * 256;
        if (i < n) __builtin_nontemporal_store(r[j], &dst[i]);
    }
}

int launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
    const size_t n = bytes / 16;
    if (!n) return TIRA_OK;
    const size_t blocks = (n + 256 * kCopyUnroll - 1) / (256 * kCopyUnroll);
    if (blocks > 0x7fffffffu) {
        set_error("copy of %zu bytes exceeds one launch", bytes);
        return TIRA_ERR_INVALID_ARG;
    }
    ProfScope prof("copy16", st);
    k_copy16<<<(unsigned)blocks, 256, 0, st>>>((f32x4*)dst, (const f32x4*)src, n);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// 8 independent FMA chains per thread, `iters` steps: 16 * iters FLOP per thread.
__global__ __launch_bounds__(256) void k_fma_loop(float* __restrict__ out, int iters, float a, float b) {
    float x[8];
#pragma unroll
    for (int k = 0; k < 8; k++) x[k] = (float)(threadIdx.x + k);
#pragma unroll 4
    for (int i = 0; i < iters; i++) {
#pragma unroll
        for (int k = 0; k < 8; k++) x[k] = fmaf(x[k], a, b);
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; k++) s += x[k];
    if (s == 12345.678f) out[blockIdx.x] = s;  // keeps the chains live
}

int launch_fma_loop(float* out, unsigned blocks, int iters, hipStream_t st) {
    ProfScope prof("fma_loop", st);
    k_fma_loop<<<blocks, 256, 0, st>>>(out, iters, 0.999999f, 1e-7f);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

}  // namespace th
