// pipeline.hip -- the front end of 3D tensor voting and the whole pipeline on
// device buffers:
//   intensity -> gradient -> structure tensor g g^T (one fused kernel)
//             -> separable Gaussian of the 6 components ("same", zero padded)
//             -> stick + plate vote (closed form, literal sum or learned plate)
//             -> eig3 saliency maps (l2 - l1, l1 - l0, l0).
// The gradient follows tira::volume::gradient (volume.h:974-1021): forward
// difference at the first voxel of an axis, backward at the last, central
// inside, each divided by the axis spacing in double and stored as float.
#include "common.hpp"
#include "mlp.hpp"

#include <cmath>

namespace th {

// g g^T of voxel (x, y, z) at linear index i into o[0..5]. x = axis 0
// (fastest, s2 voxels), y = axis 1 (s1), z = axis 2 (s0); output packed-6
// (xx, xy, xz, yy, yz, zz) with tensor axis 0 along x.
__device__ __forceinline__ void structure_tensor_voxel(const float* __restrict__ I, unsigned x, unsigned y, unsigned z,
                                                       size_t i, unsigned s0, unsigned s1, unsigned s2, double dx,
                                                       double dy, double dz, float* __restrict__ o) {
    const size_t plane = (size_t)s1 * s2;
    float g[3];
    const unsigned pos[3] = {x, y, z}, len[3] = {s2, s1, s0};
    const size_t stride[3] = {1, s2, plane};
    const double sp[3] = {dx, dy, dz};
#pragma unroll
    for (int a = 0; a < 3; a++) {
        if (pos[a] == 0) g[a] = (float)((double)(I[i + stride[a]] - I[i]) / sp[a]);
        else if (pos[a] == len[a] - 1) g[a] = (float)((double)(I[i] - I[i - stride[a]]) / sp[a]);
        else g[a] = (float)((double)(I[i + stride[a]] - I[i - stride[a]]) / (2.0 * sp[a]));
    }
    o[0] = g[0] * g[0]; o[1] = g[0] * g[1]; o[2] = g[0] * g[2];
    o[3] = g[1] * g[1]; o[4] = g[1] * g[2]; o[5] = g[2] * g[2];
}

constexpr unsigned kStX = 64, kStY = 4;

// One voxel per thread, blocks of 64 (x) x 4 (y) voxels of one plane: the
// coordinates come from the block index (a grid-stride loop splits the linear
// index with three 64-bit divisions per voxel: 0.92 -> 0.83 ms at
// 512 x 512 x 384). Staging a wave's 64 x 6 outputs in LDS for 8-byte stores
// of consecutive addresses measured 10% slower than the six 4-byte stores.
__global__ __launch_bounds__(kStX * kStY) void k_structure_tensor(const float* __restrict__ I, unsigned s0,
                                                                  unsigned s1, unsigned s2, double dx, double dy,
                                                                  double dz, float* __restrict__ out) {
    const unsigned lane = threadIdx.x % kStX, wave = threadIdx.x / kStX;
    const unsigned x0 = blockIdx.x * kStX, x = x0 + lane, y = blockIdx.y * kStY + wave;
    const unsigned z = blockIdx.z;
    if (y >= s1) return;  // whole waves
    const unsigned nv = min(kStX, s2 - x0);  // voxels of this wave's run
    const size_t i0 = ((size_t)z * s1 + y) * s2 + x0;
    if (lane < nv) structure_tensor_voxel(I, x, y, z, i0 + lane, s0, s1, s2, dx, dy, dz, out + (i0 + lane) * 6);
}

// Grid-stride form for shapes beyond the 3D grid's y and z limits (65535).
__global__ __launch_bounds__(256) void k_structure_tensor_flat(const float* __restrict__ I, unsigned s0, unsigned s1,
                                                               unsigned s2, double dx, double dy, double dz,
                                                               float* __restrict__ out) {
    const size_t plane = (size_t)s1 * s2;
    const size_t n = plane * s0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const unsigned x = (unsigned)(i % s2), y = (unsigned)((i / s2) % s1), z = (unsigned)(i / plane);
        structure_tensor_voxel(I, x, y, z, i, s0, s1, s2, dx, dy, dz, out + i * 6);
    }
}

static int launch_structure_tensor(const float* I, unsigned s0, unsigned s1, unsigned s2, double dx, double dy,
                                   double dz, float* out, hipStream_t st) {
    const size_t n = (size_t)s0 * s1 * s2;
    ProfScope prof("structure_tensor", st);
    const unsigned gy = (unsigned)ceil_div(s1, kStY);
    if (gy <= 65535 && s0 <= 65535) {
        const dim3 grid((unsigned)ceil_div(s2, kStX), gy, s0);
        k_structure_tensor<<<grid, kStX * kStY, 0, st>>>(I, s0, s1, s2, dx, dy, dz, out);
    } else {
        const unsigned grid = (unsigned)std::min<size_t>(ceil_div(n, 256), 65536);
        k_structure_tensor_flat<<<grid, 256, 0, st>>>(I, s0, s1, s2, dx, dy, dz, out);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_structure_tensor3_async(const float* intensity, unsigned s0, unsigned s1, unsigned s2, double spacing_x,
                                 double spacing_y, double spacing_z, float* out6) {
    TH_ARG(intensity && out6, "NULL pointer");
    TH_ARG(s0 >= 2 && s1 >= 2 && s2 >= 2, "every axis needs at least 2 voxels (got %u x %u x %u)", s0, s1, s2);
    TH_ARG(spacing_x > 0 && spacing_y > 0 && spacing_z > 0, "spacings must be > 0");
    TH_ARG(is_device(intensity) && is_device(out6), "the structure tensor takes device pointers");
    return launch_structure_tensor(intensity, s0, s1, s2, spacing_x, spacing_y, spacing_z, out6, stream());
}

int tira_tensorvote3_pipeline_async(const float* intensity, unsigned s0, unsigned s1, unsigned s2, double spacing_x,
                                    double spacing_y, double spacing_z, float smooth_sigma, float vote_sigma1,
                                    float vote_sigma2, unsigned w, unsigned power, int stick, int plate,
                                    unsigned samples, int plate_mode, void* mlp, float mlp_tol, float* saliency,
                                    unsigned flags) {
    TH_ARG(intensity && saliency, "NULL pointer");
    TH_ARG(s0 >= 2 && s1 >= 2 && s2 >= 2, "every axis needs at least 2 voxels (got %u x %u x %u)", s0, s1, s2);
    TH_ARG(is_device(intensity) && is_device(saliency), "the pipeline takes device pointers");
    TH_ARG(smooth_sigma >= 0.0f, "smoothing sigma must be >= 0");
    TH_ARG(plate_mode >= TIRA_PLATE_AUTO && plate_mode <= TIRA_VOTE_SPECTRAL, "bad plate mode %d", plate_mode);
    TH_ARG(plate_mode != TIRA_PLATE_MLP || mlp, "plate mode MLP needs a network handle (its sigmas and power apply)");
    const size_t n = (size_t)s0 * s1 * s2;
    float *a = nullptr, *b = nullptr;
    TH_TRY(ws_get(WS_PIPE0, n * 24, (void**)&a));
    TH_TRY(ws_get(WS_PIPE1, n * 24, (void**)&b));
    hipStream_t st = stream();
    TH_TRY(launch_structure_tensor(intensity, s0, s1, s2, spacing_x, spacing_y, spacing_z, a, st));
    const float* voted_in = a;
    if (smooth_sigma > 0.0f) {
        // kernel length ceil(6 sigma) forced odd (filter.h:405-410)
        unsigned k = (unsigned)std::ceil(6.0f * smooth_sigma);
        if (k % 2 == 0) k++;
        TH_ARG(k <= 13, "smoothing sigma %g needs %u taps; the 6-component filter takes at most 13", smooth_sigma, k);
        float taps[13];
        TH_TRY(tira_gaussian_taps(k, smooth_sigma, taps));
        TH_TRY(tira_gaussian_filter3_tensor_async(a, b, s0, s1, s2, 6, taps, k, 1));
        voted_in = b;
    }
    float* voted = voted_in == a ? b : a;
    if (plate_mode == TIRA_PLATE_MLP) {
        TH_TRY(tira_tensorvote3_mlp(mlp, voted_in, TIRA_LAYOUT_PACKED6, voted, TIRA_LAYOUT_PACKED6, s0, s1, s2, w,
                                    stick, mlp_tol, nullptr));
    } else {
        TH_TRY(tira_tensorvote3_async(voted_in, TIRA_LAYOUT_PACKED6, voted, TIRA_LAYOUT_PACKED6, s0, s1, s2,
                                      vote_sigma1, vote_sigma2, w, power, stick, plate, samples, plate_mode));
    }
    return tira_eig3_async(voted, n, TIRA_LAYOUT_PACKED6, saliency, nullptr, nullptr, flags | TIRA_EIG_SALIENCY,
                           nullptr);
}

}  // extern "C"
