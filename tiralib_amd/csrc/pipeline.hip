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
constexpr int kStZ = 2;  // planes per thread

// One (x, y) column of kStZ planes per thread, blocks of 64 (x) x 4 (y)
// columns. Every load of the column (its kStZ + 2 voxels along z and the four
// in-plane neighbours of each plane) is issued before any arithmetic: more
// reads in flight per wave than one voxel per thread, 0.83 -> 0.78 ms at
// 512 x 512 x 384 with unit spacings (a kernel that only writes the same 24 B
// per voxel takes 0.47 ms). The differences and quotients are those of
// structure_tensor_voxel, so the two kernels give the same bits. Every axis
// holds at least 2 voxels (the entry points check it), so a voxel is either
// first (forward difference), last (backward) or inside (central) along each.
// A grid-stride loop splitting the linear index (three 64-bit divisions per
// voxel) took 0.92 ms; LDS-staged 8-byte output stores measured 10% slower
// than per-thread stores.
// POW2: every spacing is a power of two and (ax, ay, az) are their
// reciprocals, which are exact: each quotient (double)d / q equals the product
// with 1 / q (the same real number, the same rounded bits) without the FP64
// division sequence. Otherwise (ax, ay, az) are the spacings.
template <bool POW2>
__global__ __launch_bounds__(kStX * kStY) void k_structure_tensor(const float* __restrict__ I, unsigned s0,
                                                                  unsigned s1, unsigned s2, double ax, double ay,
                                                                  double az, float* __restrict__ out) {
    const unsigned lane = threadIdx.x % kStX, wave = threadIdx.x / kStX;
    const unsigned x = blockIdx.x * kStX + lane, y = blockIdx.y * kStY + wave;
    if (y >= s1 || x >= s2) return;
    const size_t plane = (size_t)s1 * s2;
    const int z0 = (int)blockIdx.z * kStZ, nz = min(kStZ, (int)s0 - z0);
    // in-plane neighbour offsets: zero on the side where the voxel is first or last
    const long xl = x > 0 ? -1 : 0, xr = x + 1 < s2 ? 1 : 0;
    const long yl = y > 0 ? -(long)s2 : 0, yr = y + 1 < s1 ? (long)s2 : 0;
    // central differences divide by twice the spacing (multiply by half its reciprocal)
    const double half = POW2 ? 0.5 : 2.0;
    const double qx = (xl && xr) ? half * ax : ax, qy = (yl && yr) ? half * ay : ay, qz = half * az;
    auto quot = [](float d, double q) { return POW2 ? (float)((double)d * q) : (float)((double)d / q); };
    const size_t base = (size_t)y * s2 + x;
    float c[kStZ + 2], xa[kStZ], xb[kStZ], ya[kStZ], yb[kStZ];
#pragma unroll
    for (int k = 0; k < kStZ + 2; k++) {  // c[k]: plane z0 - 1 + k of this column
        const int z = z0 - 1 + k;
        c[k] = (z >= 0 && z < (int)s0) ? I[(size_t)z * plane + base] : 0.0f;
    }
#pragma unroll
    for (int k = 0; k < kStZ; k++) {
        const size_t i = (size_t)(z0 + min(k, nz - 1)) * plane + base;
        xa[k] = I[i + xr]; xb[k] = I[i + xl]; ya[k] = I[i + yr]; yb[k] = I[i + yl];
    }
#pragma unroll
    for (int k = 0; k < kStZ; k++) {
        if (k < nz) {
            const int z = z0 + k;
            const float gx = quot(xa[k] - xb[k], qx);
            const float gy = quot(ya[k] - yb[k], qy);
            float gz;
            if (z == 0) gz = quot(c[k + 2] - c[k + 1], az);
            else if (z == (int)s0 - 1) gz = quot(c[k + 1] - c[k], az);
            else gz = quot(c[k + 2] - c[k], qz);
            float* o = out + ((size_t)z * plane + base) * 6;
            o[0] = gx * gx; o[1] = gx * gy; o[2] = gx * gz;
            o[3] = gy * gy; o[4] = gy * gz; o[5] = gz * gz;
        }
    }
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
    const unsigned gy = (unsigned)ceil_div(s1, kStY), gz = (unsigned)ceil_div(s0, kStZ);
    if (gy <= 65535 && gz <= 65535) {
        const dim3 grid((unsigned)ceil_div(s2, kStX), gy, gz);
        auto pow2 = [](double v) { int e; return std::frexp(v, &e) == 0.5; };
        if (pow2(dx) && pow2(dy) && pow2(dz))
            k_structure_tensor<true><<<grid, kStX * kStY, 0, st>>>(I, s0, s1, s2, 1.0 / dx, 1.0 / dy, 1.0 / dz, out);
        else
            k_structure_tensor<false><<<grid, kStX * kStY, 0, st>>>(I, s0, s1, s2, dx, dy, dz, out);
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
