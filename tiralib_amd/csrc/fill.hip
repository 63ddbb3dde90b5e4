// fill.hip -- the fill=True branch of the 2D saliency reconstruction
// (reconstruct.py:89-116, _fill_enclosed) on gfx950: an optional
// morphological closing of the thresholded boundary mask with a
// (2r + 1) x (2r + 1) square, then hole filling.
//
//   * closing = dilation then erosion with the square, both separable
//     (a square is the product of a row and a column segment); as in
//     scipy.ndimage.binary_closing with border_value 0, pixels outside the
//     image count as 0: dilation ignores them, erosion clears every pixel
//     whose square reaches outside;
//   * hole filling (scipy.ndimage.binary_fill_holes, 4-connected): a
//     background pixel stays background exactly when it is 4-connected to
//     the outside through background. That reachability is propagated from
//     the border by sweeps over 32 x 32 tiles: each tile relaxes its pixels
//     in LDS until nothing changes locally (reading its neighbours' state at
//     the sweep's start through a one-pixel halo), and sweeps repeat until no
//     pixel changes. Reachability only grows, so concurrent tiles cannot undo
//     each other and the fixed point -- the reachable set -- is unique.
#include "common.hpp"

namespace th {

typedef unsigned char u8;

// One axis of dilation (DIL) or erosion of a 0/1 mask with a segment of
// radius r; `dx`, `dy` select the axis.
template <bool DIL>
__global__ __launch_bounds__(256) void k_morph(const u8* __restrict__ in, u8* __restrict__ out, int h, int w, int r,
                                               int dx, int dy) {
    const size_t n = (size_t)h * w;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(p / w), x = (int)(p % w);
        u8 v = DIL ? 0 : 1;
        for (int k = -r; k <= r; k++) {
            const int xx = x + k * dx, yy = y + k * dy;
            const bool inside = xx >= 0 && yy >= 0 && xx < w && yy < h;
            const u8 s = inside ? (in[(size_t)yy * w + xx] ? 1 : 0) : 0;
            v = DIL ? (u8)(v | s) : (u8)(v & s);
        }
        out[p] = v;
    }
}

__global__ void k_fill_zero8(u8* __restrict__ x, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) x[i] = 0;
}

__global__ void k_clear_flags(unsigned* __restrict__ f, int n) {
    if ((int)threadIdx.x < n) f[threadIdx.x] = 0u;
}

constexpr int kFillBatch = 4;  // sweeps per read-back

constexpr int kFillT = 32;  // tile edge; one thread per pixel

// One sweep of the reachability flood: reach[p] = 1 for background pixels
// (mask 0) 4-connected to the outside through background. Cells outside the
// image count as reached. *changed becomes nonzero when the sweep reached a
// new pixel.
// prev (optional): the previous sweep's flag; when it is zero the reachable
// set is already complete and the sweep returns at once.
__global__ __launch_bounds__(kFillT * kFillT) void k_fill_sweep(const u8* __restrict__ mask, u8* __restrict__ reach,
                                                                 int h, int w, unsigned* __restrict__ changed,
                                                                 const unsigned* __restrict__ prev) {
    if (prev && *prev == 0u) return;
    __shared__ u8 s_r[kFillT + 2][kFillT + 2];
    const int tx = (int)threadIdx.x % kFillT, ty = (int)threadIdx.x / kFillT;
    const int x0 = (int)blockIdx.x * kFillT, y0 = (int)blockIdx.y * kFillT;
    const int x = x0 + tx, y = y0 + ty;
    const bool in = x < w && y < h;
    const bool bg = in && mask[(size_t)y * w + x] == 0;
    for (int i = (int)threadIdx.x; i < (kFillT + 2) * (kFillT + 2); i += kFillT * kFillT) {
        const int hy = i / (kFillT + 2), hx = i % (kFillT + 2);
        const int gy = y0 + hy - 1, gx = x0 + hx - 1;
        const bool inside = gx >= 0 && gy >= 0 && gx < w && gy < h;
        s_r[hy][hx] = inside ? reach[(size_t)gy * w + gx] : (u8)1;
    }
    __syncthreads();
    const u8 r0 = s_r[ty + 1][tx + 1];
    u8 r = r0;
    for (int it = 0; it < kFillT * kFillT; it++) {
        int grew = 0;
        if (bg && !r && (s_r[ty][tx + 1] | s_r[ty + 2][tx + 1] | s_r[ty + 1][tx] | s_r[ty + 1][tx + 2])) {
            r = 1;
            s_r[ty + 1][tx + 1] = 1;
            grew = 1;
        }
        if (!__syncthreads_or(grew)) break;
    }
    if (in && r != r0) {
        reach[(size_t)y * w + x] = 1;
        atomicOr(changed, 1u);
    }
}

// out = mask OR (background not reached from the outside).
__global__ void k_fill_out(const u8* __restrict__ mask, const u8* __restrict__ reach, size_t n, u8* __restrict__ out) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (mask[i] || !reach[i]) ? 1 : 0;
}

}  // namespace th

using namespace th;

extern "C" {

// _fill_enclosed(boundary, close_radius), reconstruct.py:98-116, on a 0/1
// byte mask (h x w, w fastest; any nonzero byte counts as 1).
int tira_fill_holes2(const unsigned char* mask, unsigned h, unsigned w, unsigned close_radius, unsigned char* out,
                     unsigned* sweeps) {
    TH_ARG(mask && out, "NULL pointer");
    TH_ARG(close_radius <= 1024, "close radius %u too large", close_radius);
    const size_t n = (size_t)h * w;
    if (sweeps) *sweeps = 0;
    if (!n) return TIRA_OK;
    TH_ARG(h < (1u << 30) && w < (1u << 30), "mask too large");
    hipStream_t st = stream();
    const u8* dm;
    void* dout;
    TH_TRY(stage_in(mask, n, WS_IN0, (const void**)&dm));
    TH_TRY(stage_out(out, n, WS_OUT0, &dout));
    u8 *a = nullptr, *b = nullptr, *reach = nullptr;
    unsigned* flag = nullptr;
    TH_TRY(ws_get(WS_SCRATCH0, n, (void**)&a));
    TH_TRY(ws_get(WS_SCRATCH1, n, (void**)&b));
    TH_TRY(ws_get(WS_SCRATCH2, n, (void**)&reach));
    TH_TRY(ws_get(WS_COUNTER, 64, (void**)&flag));
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(n, 256), 65536);
    const u8* boundary = dm;
    if (close_radius > 0) {
        const int r = (int)close_radius;
        k_morph<true><<<grid, 256, 0, st>>>(dm, a, (int)h, (int)w, r, 1, 0);
        k_morph<true><<<grid, 256, 0, st>>>(a, b, (int)h, (int)w, r, 0, 1);
        k_morph<false><<<grid, 256, 0, st>>>(b, a, (int)h, (int)w, r, 1, 0);
        k_morph<false><<<grid, 256, 0, st>>>(a, b, (int)h, (int)w, r, 0, 1);
        TH_HIP(hipGetLastError());
        boundary = b;
    }
    k_fill_zero8<<<grid, 256, 0, st>>>(reach, n);
    const dim3 tiles(ceil_div(w, kFillT), ceil_div(h, kFillT));
    // every sweep but the last reaches a new pixel: at most n + 1 sweeps. They
    // run kFillBatch to a read-back; a sweep after one that reached nothing
    // returns at once, and `sweeps` counts those that ran.
    unsigned host_flags[kFillBatch], count = 0;
    bool done = false;
    while (!done && count <= n) {
        k_clear_flags<<<1, 64, 0, st>>>(flag, kFillBatch);
        for (int b = 0; b < kFillBatch; b++)
            k_fill_sweep<<<tiles, kFillT * kFillT, 0, st>>>(boundary, reach, (int)h, (int)w, flag + b,
                                                            b ? flag + b - 1 : nullptr);
        TH_HIP(hipGetLastError());
        TH_HIP(hipMemcpyAsync(host_flags, flag, sizeof(host_flags), hipMemcpyDeviceToHost, st));
        TH_HIP(hipStreamSynchronize(st));
        for (int b = 0; b < kFillBatch && !done; b++) {
            count++;
            done = host_flags[b] == 0;
        }
    }
    k_fill_out<<<grid, 256, 0, st>>>(boundary, reach, n, (u8*)dout);
    TH_HIP(hipGetLastError());
    if (sweeps) *sweeps = count;
    return finish_out(out, dout, n);
}

}  // extern "C"
