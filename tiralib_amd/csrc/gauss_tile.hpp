// gauss_tile.hpp -- the fused separable-Gaussian tile shared by the single-
// volume filter (filter.hip) and the persistent batched launch (batch.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "eig_device.hpp"

namespace th {

constexpr int kGT = 16;  // tile width along x (tile height TY along y: 8, 16 or 32)

// LDS row of a staged patch of edge pw: room for the x-pass's 16-byte reads
// past the edge, rounded up to 16 floats past a multiple of 32.
__host__ __device__ constexpr int gauss_row_stride(int pw) { return ((pw + 6 + 15) / 32) * 32 + 16; }

struct Gauss3Geom {
    int s0, s1, s2;     // input shape (z, y, x)
    int o0, o1, o2;     // output shape
    int c0, c1, c2;     // input coordinate of output (0,0,0)'s centre
    int zchunk;
    unsigned nbx, nby, nbz;
};

// Eigenvalues (ascending, FP32 with deflation) of the 6-component tensor at
// LDS position `pos` of a component-major staged plane (component stride cs).
__device__ __forceinline__ void tile_eig3(const float* s_raw, int pos, int cs, float* __restrict__ dst) {
    float a = s_raw[pos], b = s_raw[pos + cs], d = s_raw[pos + 2 * cs];
    float c = s_raw[pos + 3 * cs], e = s_raw[pos + 4 * cs], f = s_raw[pos + 5 * cs];
    const float m = fmaxf(fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fabsf(d))), fmaxf(fabsf(e), fabsf(f)));
    int ex = 0;
    if (m > 0.0f && isfinite(m)) {
        frexpf(m, &ex);
        a = ldexpf(a, -ex); b = ldexpf(b, -ex); c = ldexpf(c, -ex);
        d = ldexpf(d, -ex); e = ldexpf(e, -ex); f = ldexpf(f, -ex);
    }
    float lam[3];
    eig3_solve<float>(a, b, c, d, e, f, 1e-30f, lam, nullptr, nullptr);
    dst[0] = ldexpf(lam[0], ex);
    dst[1] = ldexpf(lam[1], ex);
    dst[2] = ldexpf(lam[2], ex);
}

// taps: 3 x (2R+1) floats (x, y, z), each centred and zero-padded.
//
// LDS: the staged plane is component-major, so the x-pass reads 16-byte
// vectors and slides a register window over 4 outputs at a time; the
// x-filtered plane is component-major too. Rows are padded to 16 floats past
// a multiple of 32, so the 16-byte window reads of neighbouring rows fall on
// disjoint banks. Each thread stages whole voxels (all components, read with
// 8-byte loads for tensors) and writes one component at a time, so the lanes
// of a wave write consecutive x of one component: no bank conflicts (one lane
// per element of a component-interleaved voxel put the six components of a
// voxel on one bank). A thread's staged voxels sit at the same in-plane
// offsets for every z-plane, so their global offsets and LDS slots are
// computed once, and the next plane is prefetched into registers while the
// current one is filtered.
// A tile is 16 (x) by TY (y) outputs, one thread per output of a plane;
// 6-component fields use two such thread sets: the halves split the
// components (3 each) for the y- and z-passes, which keeps the register ring
// at (2R+1) x 3 floats per thread.
template <int C, int TY = kGT> struct GaussThreads {
    static constexpr int CG = C == 6 ? 2 : 1, N = kGT * TY * CG;
};

// LDS floats of a tile: the staged patch (component-major, padded rows), then
// the x-filtered patch.
template <int R, int C, int TY>
__host__ __device__ constexpr int gauss_tile_lds_floats() {
    return C * (TY + 2 * R) * gauss_row_stride(kGT + 2 * R) + C * (TY + 2 * R) * kGT;
}

template <int R, int C, bool EIG, int TY = kGT>
__device__ __forceinline__ void gauss3_tile(const float* __restrict__ in, float* __restrict__ out,
                                            const float* __restrict__ taps, const Gauss3Geom& g, unsigned b,
                                            float* __restrict__ evals, float* s_raw, float* s_xf) {
    constexpr int K = 2 * R + 1;
    constexpr int NT = GaussThreads<C, TY>::N;
    constexpr int CP = C / GaussThreads<C, TY>::CG;     // components per thread in the y/z passes
    constexpr int PW = kGT + 2 * R;                     // staged patch width (x)
    constexpr int PH = TY + 2 * R;                      // staged patch height (y)
    constexpr int RS = gauss_row_stride(PW);            // padded LDS row
    constexpr int NV = PW * PH;                         // staged voxels per plane
    constexpr int NS = (NV + NT - 1) / NT;              // staged voxels per thread
    constexpr int NSEG = kGT / 4;                       // 4-output segments per x-filtered row
    constexpr int NXT = C * PH * NSEG;                  // x-pass tasks per plane
    constexpr int NX = (NXT + NT - 1) / NT;
    constexpr int NWIN = 4 + 2 * R;                     // inputs per x-pass task
    constexpr int NWIN4 = (NWIN + 3) / 4;

    const int tid = threadIdx.x;
    const int tx = tid % kGT, ty = (tid / kGT) % TY, c0 = (tid / (kGT * TY)) * CP;
    const int bx = (int)(b % g.nbx), by = (int)((b / g.nbx) % g.nby), bz = (int)(b / (g.nbx * g.nby));
    const int ox0 = bx * kGT, oy0 = by * TY, oz0 = bz * g.zchunk;
    const int nz = min(g.zchunk, g.o0 - oz0);
    const int nplanes = nz + 2 * R;
    const float* gx = taps;
    const float* gy = taps + K;
    const float* gz = taps + 2 * K;
    const int ix0 = g.c2 + ox0 - R, iy0 = g.c1 + oy0 - R, iz0 = g.c0 + oz0 - R;
    const size_t plane = (size_t)g.s1 * g.s2 * C;

    // staged voxels of this thread: voxel (yy, xx) of the patch, LDS slot
    // [c][yy][xx] = lslot + c * PH * RS; goff < 0 marks voxels outside the volume
    int goff[NS], lslot[NS];
#pragma unroll
    for (int s = 0; s < NS; s++) {
        const int v = tid + NT * s;
        goff[s] = -1;
        lslot[s] = -1;
        if (v < NV) {
            const int xx = v % PW, yy = v / PW;
            const int x = ix0 + xx, y = iy0 + yy;
            lslot[s] = yy * RS + xx;
            if (x >= 0 && x < g.s2 && y >= 0 && y < g.s1) goff[s] = (y * g.s2 + x) * C;
        }
    }

    float pre[NS][C];
    auto prefetch = [&](int i) {
        const int iz = iz0 + i;
        const bool zin = i < nplanes && iz >= 0 && iz < g.s0;
        const float* pl = in + (zin ? (size_t)iz * plane : 0);
#pragma unroll
        for (int s = 0; s < NS; s++) {
            const bool ok = zin && goff[s] >= 0;
            if (C % 2 == 0) {
                // a voxel's components are 8-byte aligned (the field is)
                const float2* v2 = reinterpret_cast<const float2*>(pl + (ok ? goff[s] : 0));
#pragma unroll
                for (int c = 0; c < C; c += 2) {
                    const float2 q = ok ? v2[c / 2] : make_float2(0.0f, 0.0f);
                    pre[s][c] = q.x;
                    pre[s][c + 1] = q.y;
                }
            } else {
#pragma unroll
                for (int c = 0; c < C; c++) pre[s][c] = ok ? pl[goff[s] + c] : 0.0f;
            }
        }
    };

    float acc[K][CP];
#pragma unroll
    for (int s = 0; s < K; s++)
#pragma unroll
        for (int c = 0; c < CP; c++) acc[s][c] = 0.0f;

    const int ox = ox0 + tx, oy = oy0 + ty;
    const bool live = ox < g.o2 && oy < g.o1;
    prefetch(0);

    for (int i0 = 0; i0 < nplanes; i0 += K) {
#pragma unroll
        for (int m = 0; m < K; m++) {
            const int i = i0 + m;
            if (i < nplanes) {  // uniform
#pragma unroll
                for (int s = 0; s < NS; s++)
                    if (lslot[s] >= 0) {
#pragma unroll
                        for (int c = 0; c < C; c++) s_raw[lslot[s] + c * PH * RS] = pre[s][c];
                    }
                __syncthreads();
                prefetch(i + 1);  // in flight while this plane is filtered
                // x-pass: task = (component, row, 4 consecutive outputs)
#pragma unroll
                for (int s = 0; s < NX; s++) {
                    const int tsk = tid + NT * s;
                    if (tsk < NXT) {
                        const int seg = tsk % NSEG, row = tsk / NSEG;  // row = c * PH + yy
                        const float4* src = reinterpret_cast<const float4*>(s_raw + row * RS + seg * 4);
                        float win[NWIN4 * 4];
#pragma unroll
                        for (int q = 0; q < NWIN4; q++) {
                            const float4 v4 = src[q];
                            win[4 * q] = v4.x; win[4 * q + 1] = v4.y; win[4 * q + 2] = v4.z; win[4 * q + 3] = v4.w;
                        }
                        float o4[4];
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            float sacc = 0.0f;
#pragma unroll
                            for (int k = 0; k < K; k++) sacc = fmaf(gx[k], win[u + k], sacc);
                            o4[u] = sacc;
                        }
                        *reinterpret_cast<float4*>(s_xf + row * kGT + seg * 4) = make_float4(o4[0], o4[1], o4[2], o4[3]);
                    }
                }
                __syncthreads();
                // y-pass for this thread's column, then the z ring
#pragma unroll
                for (int c = 0; c < CP; c++) {
                    const float* col = s_xf + ((c0 + c) * PH + ty) * kGT + tx;
                    float v = 0.0f;
#pragma unroll
                    for (int k = 0; k < K; k++) v = fmaf(gy[k], col[k * kGT], v);
                    // plane i feeds outputs t = i - 2R + j (j = 0..2R) with weight gz[2R - j];
                    // output t lives in ring slot t mod K = (m + 1 + j) mod K.
#pragma unroll
                    for (int j = 0; j < K; j++) acc[(m + 1 + j) % K][c] = fmaf(gz[2 * R - j], v, acc[(m + 1 + j) % K][c]);
                }
                // output t = i - 2R is complete (slot (m + 1) % K); the slot is
                // cleared either way -- before the first real output it only
                // collected contributions of outputs t < 0
                const int t = i - 2 * R;
                if (live && t >= 0 && t < nz) {
                    float* o = out + (((size_t)(oz0 + t) * g.o1 + oy) * g.o2 + ox) * C + c0;
#pragma unroll
                    for (int c = 0; c < CP; c++) o[c] = acc[(m + 1) % K][c];
                }
#pragma unroll
                for (int c = 0; c < CP; c++) acc[(m + 1) % K][c] = 0.0f;
            }
        }
    }
    if (EIG) {
        // Eigenvalues of the tile's own voxels (same mode only: input and
        // output grids coincide), after the z loop so that the solver's
        // registers do not add to the ring's: the tile's input columns were
        // read moments ago and come from L2.
        const int ntile = kGT * TY * nz;
        for (int k = tid; k < ntile; k += NT) {
            const int t = k / (kGT * TY), rem = k % (kGT * TY);
            const int ex = ox0 + rem % kGT, ey = oy0 + rem / kGT;
            if (ex < g.o2 && ey < g.o1) {
                const size_t v = ((size_t)(oz0 + t) * g.o1 + ey) * g.o2 + ex;
                tile_eig3(in + v * 6, 0, 1, evals + v * 3);
            }
        }
    }
}

// Tensor fields up to R = 5 fit 80 VGPRs: six waves per SIMD, three blocks
// per CU (4% faster at R = 5); R = 6 would spill at that bound (30% slower).
// LDS: gauss_tile_lds_floats<R, C, TY>() floats, sized at launch.
template <int R, int C, int TY> struct Gauss3Bound { static constexpr int W = (C == 6 && TY == kGT && R <= 5) ? 6 : 1; };
template <int R, int C, int TY = kGT>
__global__ __launch_bounds__((GaussThreads<C, TY>::N), (Gauss3Bound<R, C, TY>::W)) void k_gauss3(
    const float* __restrict__ in, float* __restrict__ out, const float* __restrict__ taps, Gauss3Geom g) {
    constexpr int PW = kGT + 2 * R;
    extern __shared__ __attribute__((aligned(16))) float s_gauss[];
    float* s_raw = s_gauss;
    float* s_xf = s_gauss + C * (TY + 2 * R) * gauss_row_stride(PW);
    // XCD-major tile order: neighbouring tiles (x fastest) run on one XCD at
    // the same time, so the halo columns they share come from its L2
    gauss3_tile<R, C, false, TY>(in, out, taps, g, xcd_remap(blockIdx.x, gridDim.x), nullptr, s_raw, s_xf);
}

}  // namespace th
