// filter.hip -- separable Gaussian filtering on gfx950, replacing
// tira::cuda::gaussian_filter3d (filter.h:400-534) and
// tira::cuda::gaussian_convolve2 (filter.h:304-397), plus the native
// zero-padded ("same") filter of packed 6-component tensor fields.
//
// 3D design (k_gauss3): one kernel does all three passes. A 256-thread block
// owns a 16 x 16 tile of output columns and a chunk of output z-planes. For
// every input z-plane it
//   1. stages the (16+2R) x (16+2R) x C input patch in LDS (zero outside the
//      volume: zero padding for "same", never reached for "valid"),
//   2. filters along x (LDS -> LDS) and along y (LDS -> registers),
//   3. adds the plane into a register ring of 2R+1 pending output planes
//      (z-pass), retiring one finished output plane per input plane.
// Each input voxel is read from HBM about once and each output written once:
// 2 * C * 4 bytes per voxel instead of 6 * C * 4 for three separate passes.
// The ring slot of every tap is a compile-time constant because the plane
// loop is unrolled by 2R+1.
#include "common.hpp"
#include "gauss_tile.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace th {


// ---------------------------------------------------------------- taps
// normaldist (filter.h:14-22) + gaussian1d (filter.h:37-64): K samples at
// x = -K/2 + 1/2 + i, normalised to sum 1. Evaluated like the reference's
// host code (the unqualified exp of its float code is the double overload).
void gaussian_taps(unsigned ksize, float sigma, std::vector<float>& k) {
    k.assign(ksize, 0.0f);
    const float hwid = (float)ksize * 1.0f / 2.0f;
    const float startx = -hwid + 0.5f;
    float sum = 0.0f;
    for (unsigned i = 0; i < ksize; i++) {
        const float x = startx + (float)i;
        const float ssq = sigma * sigma;
        const float nrm = 1.0f / std::sqrt(2.0f * (float)M_PI * ssq);
        const float ex = -(x * x) / (2.0f * ssq);
        k[i] = (float)((double)nrm * std::exp((double)ex));
        sum += k[i];
    }
    if (sum > 0.0f) {
        const float inv = 1.0f / sum;
        for (auto& v : k) v *= inv;
    }
}

// ceil(6 sigma) forced odd (filter.h:405-410)
unsigned gaussian_ksize3(float sigma) {
    unsigned k = (unsigned)std::ceil(6.0f * sigma);
    if (k % 2 == 0) k++;
    return k;
}

// 1D valid-mode pass of gaussian_convolve2 (cuda_convolve2y / 2x, filter.h:209-240).
__global__ __launch_bounds__(256) void k_conv2_valid(const float* __restrict__ src, float* __restrict__ dst,
                                                     const float* __restrict__ k, int K, int in_w, int out_w,
                                                     int out_h, int axis) {
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    if (x >= out_w || y >= out_h) return;
    float s = 0.0f;
    if (axis == 1) for (int i = 0; i < K; i++) s = fmaf(k[i], src[(size_t)(y + i) * in_w + x], s);
    else for (int i = 0; i < K; i++) s = fmaf(k[i], src[(size_t)y * in_w + x + i], s);
    dst[(size_t)y * out_w + x] = s;
}

template <int C>
static int launch_gauss3_R(int R, const float* in, float* out, const float* taps, const Gauss3Geom& g,
                           hipStream_t st) {
    const unsigned grid = g.nbx * g.nby * g.nbz;
// 16 x 16 tiles: 16 x 8 tiles (more blocks per CU) measured 10-15% slower here
#define TH_G3(RR, CC) \
    k_gauss3<RR, CC><<<grid, GaussThreads<CC>::N, gauss_tile_lds_floats<RR, CC, kGT>() * 4, st>>>(in, out, taps, g)
    switch (R) {
        case 1: TH_G3(1, C); break;
        case 2: TH_G3(2, C); break;
        case 3: TH_G3(3, C); break;
        case 4: TH_G3(4, C); break;
        case 5: TH_G3(5, C); break;
        case 6: TH_G3(6, C); break;
        case 7: if (C == 1) { TH_G3(7, 1); break; } [[fallthrough]];
        case 8: if (C == 1) { TH_G3(8, 1); break; } [[fallthrough]];
        default:
            // 6-component fields keep a (2R+1) x 6 register ring: R <= 6 (13 taps) stays spill-free
            set_error("Gaussian radius %d outside 1..%d for %d components", R, C == 1 ? 8 : 6, C);
            return TIRA_ERR_INVALID_ARG;
    }
#undef TH_G3
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// Uploads three centred, zero-padded tap arrays of length 2R+1 (x, y, z).
static int upload_taps(const std::vector<float>& kx, const std::vector<float>& ky, const std::vector<float>& kz,
                       int R, int slot, hipStream_t st, float** dtaps) {
    const int K = 2 * R + 1;
    std::vector<float> t(3 * K, 0.0f);
    const std::vector<float>* ks[3] = {&kx, &ky, &kz};
    for (int a = 0; a < 3; a++) {
        const int n = (int)ks[a]->size(), pad = (K - n) / 2;
        for (int i = 0; i < n; i++) t[a * K + pad + i] = (*ks[a])[i];
    }
    TH_TRY(ws_get(slot, t.size() * 4, (void**)dtaps));
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(*dtaps, t.data(), t.size() * 4, hipMemcpyHostToDevice));
    return TIRA_OK;
}

int gauss3_run(const float* in, float* out, int C, unsigned s0, unsigned s1, unsigned s2,
               const std::vector<float>& kx, const std::vector<float>& ky, const std::vector<float>& kz, bool same,
               hipStream_t st) {
    const int rx = (int)kx.size() / 2, ry = (int)ky.size() / 2, rz = (int)kz.size() / 2;
    const int R = std::max(rx, std::max(ry, rz));
    Gauss3Geom g;
    g.s0 = (int)s0; g.s1 = (int)s1; g.s2 = (int)s2;
    if (same) {
        g.o0 = g.s0; g.o1 = g.s1; g.o2 = g.s2;
        g.c0 = g.c1 = g.c2 = 0;
    } else {
        g.o0 = g.s0 - (int)kz.size() + 1; g.o1 = g.s1 - (int)ky.size() + 1; g.o2 = g.s2 - (int)kx.size() + 1;
        g.c0 = rz; g.c1 = ry; g.c2 = rx;
    }
    if (g.o0 <= 0 || g.o1 <= 0 || g.o2 <= 0) {
        set_error("valid-mode output is empty (kernel longer than the volume)");
        return TIRA_ERR_INVALID_ARG;
    }
    g.nbx = ceil_div(g.o2, kGT);
    g.nby = ceil_div(g.o1, kGT);
    // z chunk: a block re-filters 2R halo planes per chunk, so the deepest of
    // 128, 64, 32, 16 planes that still gives two blocks per CU (512) and whose
    // last chunk is more than half full; 8 when none does. Measured on
    // 6-component fields (K = 11 and 13) at 192x256x256, 256^3, 384x512x512,
    // 128^3 and 96x320x288, this picks the fastest chunk or one within 1% of
    // it (e.g. 192x256x256, K = 13: 0.255 ms against 0.356 ms for 16 planes).
    const unsigned xy = g.nbx * g.nby;
    int zc = 8;
    for (int c = 128; c >= 16; c /= 2) {
        const int nch = ceil_div(g.o0, c), last = g.o0 - (nch - 1) * c;
        if ((size_t)xy * nch >= 512 && 2 * last > c) { zc = c; break; }
    }
    g.zchunk = zc;
    g.nbz = ceil_div(g.o0, zc);
    float* dt = nullptr;
    TH_TRY(upload_taps(kx, ky, kz, R, WS_SCRATCH2, st, &dt));
    ProfScope prof(C == 6 ? "gauss3_tensor" : "gauss3", st);
    if (C == 6) return launch_gauss3_R<6>(R, in, out, dt, g, st);
    if (C == 1) return launch_gauss3_R<1>(R, in, out, dt, g, st);
    set_error("unsupported component count %d", C);
    return TIRA_ERR_INVALID_ARG;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_gaussian_filter3d(const float* src, unsigned width, unsigned height, unsigned depth, float sigma_w,
                           float sigma_h, float sigma_d, float* dst, unsigned* out_width, unsigned* out_height,
                           unsigned* out_depth) {
    TH_ARG(src && out_width && out_height && out_depth, "NULL pointer");
    TH_ARG(sigma_w > 0 && sigma_h > 0 && sigma_d > 0, "sigma must be > 0");
    const unsigned kw = gaussian_ksize3(sigma_w), kh = gaussian_ksize3(sigma_h), kd = gaussian_ksize3(sigma_d);
    TH_ARG(kw <= width && kh <= height && kd <= depth, "kernel (%u, %u, %u) longer than the volume", kw, kh, kd);
    TH_ARG(kw <= 17 && kh <= 17 && kd <= 17, "sigma too large for the fused filter (kernel > 17 taps)");
    *out_width = width - kw + 1;
    *out_height = height - kh + 1;
    *out_depth = depth - kd + 1;
    if (!dst) return TIRA_OK;  // size query
    std::vector<float> gw, gh, gd;
    gaussian_taps(kw, sigma_w, gw);
    gaussian_taps(kh, sigma_h, gh);
    gaussian_taps(kd, sigma_d, gd);
    const size_t nin = (size_t)width * height * depth;
    const size_t nout = (size_t)(*out_width) * (*out_height) * (*out_depth);
    const float* din; void* dout;
    TH_TRY(stage_in(src, nin * 4, WS_IN0, (const void**)&din));
    TH_TRY(stage_out(dst, nout * 4, WS_OUT0, &dout));
    TH_TRY(gauss3_run(din, (float*)dout, 1, depth, height, width, gw, gh, gd, false, stream()));
    return finish_out(dst, dout, nout * 4);
}

int tira_gaussian_convolve2(const float* src, unsigned in_sx, unsigned in_sy, float sigma1, float sigma2, float* dst,
                            unsigned* out_width, unsigned* out_height) {
    TH_ARG(src && out_width && out_height, "NULL pointer");
    // window = (unsigned)(6 sigma), not forced odd (filter.h:308-309)
    const unsigned kw = (unsigned)(sigma1 * 6), kh = (unsigned)(sigma2 * 6);
    TH_ARG(kw >= 1 && kh >= 1 && kw <= in_sx && kh <= in_sy, "kernel longer than the image or empty");
    *out_width = in_sx - (kw - 1);
    *out_height = in_sy - (kh - 1);
    if (!dst) return TIRA_OK;
    std::vector<float> gw, gh;
    gaussian_taps(kw, sigma1, gw);
    gaussian_taps(kh, sigma2, gh);
    hipStream_t st = stream();
    const float* din; void* dout; float* tmp; float* dk;
    TH_TRY(stage_in(src, (size_t)in_sx * in_sy * 4, WS_IN0, (const void**)&din));
    TH_TRY(stage_out(dst, (size_t)(*out_width) * (*out_height) * 4, WS_OUT0, &dout));
    TH_TRY(ws_get(WS_SCRATCH0, (size_t)in_sx * (*out_height) * 4, (void**)&tmp));
    TH_TRY(ws_get(WS_SCRATCH2, (kw + kh) * 4, (void**)&dk));
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(dk, gh.data(), kh * 4, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(dk + kh, gw.data(), kw * 4, hipMemcpyHostToDevice));
    // y pass (image -> in_sx x out_height), then x pass (filter.h:363-374)
    k_conv2_valid<<<dim3(ceil_div(in_sx, 64), ceil_div(*out_height, 4)), 256, 0, st>>>(din, tmp, dk, (int)kh,
                                                                                         (int)in_sx, (int)in_sx,
                                                                                         (int)*out_height, 1);
    k_conv2_valid<<<dim3(ceil_div(*out_width, 64), ceil_div(*out_height, 4)), 256, 0, st>>>(
        tmp, (float*)dout, dk + kh, (int)kw, (int)in_sx, (int)*out_width, (int)*out_height, 0);
    TH_HIP(hipGetLastError());
    return finish_out(dst, dout, (size_t)(*out_width) * (*out_height) * 4);
}

int tira_gaussian_filter3_tensor_async(const float* in, float* out, unsigned s0, unsigned s1, unsigned s2,
                                       unsigned components, const float* taps, unsigned ntaps, int same) {
    TH_ARG(in && out && taps, "NULL pointer");
    TH_ARG(ntaps % 2 == 1 && ntaps >= 3 && ntaps <= (components == 1 ? 17u : 13u),
           "taps must be odd, 3..%u (got %u)", components == 1 ? 17u : 13u, ntaps);
    TH_ARG(components == 1 || components == 6, "components must be 1 or 6");
    std::vector<float> k(taps, taps + ntaps);
    return gauss3_run(in, out, (int)components, s0, s1, s2, k, k, k, same != 0, stream());
}

int tira_gaussian_taps(unsigned ksize, float sigma, float* out) {
    TH_ARG(out && ksize >= 1, "bad arguments");
    std::vector<float> k;
    gaussian_taps(ksize, sigma, k);
    std::memcpy(out, k.data(), ksize * 4);
    return TIRA_OK;
}

}  // extern "C"
