// eigen.hip -- batched eigen-solvers for symmetric 2x2 / 3x3 tensor fields
// on gfx950, and the C-ABI entry points that replace tira::cuda::evals2/3 and
// evecs2polar / evecs3spherical (eigen.h:700-1012).
//
// eig3 kernel design (HBM-bound: 36 B per tensor for packed-6 in / 3 floats
// out, 48 B in the reference 9-float layout):
//  * a 256-thread block owns 512 consecutive tensors; the AoS input is read
//    with fully coalesced 16-byte loads into LDS and each lane then picks its
//    tensors out of LDS (the 9-float stride is odd, so no bank conflicts);
//  * FP32 trigonometric solve for every tensor;
//  * tensors whose relative eigengap is below 1e-4 are appended to a
//    block-local list in LDS and re-solved in FP64 after a barrier, so the
//    FP64 path runs on ~ceil(flagged/64) waves of the block instead of on
//    every wave (10% scattered near-degenerate tensors would otherwise put
//    the FP64 path in 99.9% of all waves);
//  * results are staged in LDS and written back with 16-byte stores.
#include "common.hpp"
#include "eig_device.hpp"
#include "eig_store.hpp"

#include <cfloat>
#include <cmath>

namespace th {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kEigThreads = 256;
constexpr int kEigPerThread = 2;
constexpr int kEigTile = kEigThreads * kEigPerThread;  // tensors per block

// Reads tensor `i` of a staged LDS tile as (a, b, c, d, e, f) = (xx, xy, yy, xz, yz, zz).
template <int LAYOUT>
__device__ __forceinline__ void lds_tensor(const float* s, int i, float& a, float& b, float& c, float& d, float& e,
                                           float& f) {
    if (LAYOUT == TIRA_LAYOUT_MAT9) {
        const float* m = s + i * 9;
        a = m[0]; b = m[1]; d = m[2]; c = m[4]; e = m[5]; f = m[8];
    } else {
        const float* m = s + i * 6;
        a = m[0]; b = m[1]; d = m[2]; c = m[3]; e = m[4]; f = m[5];
    }
}

template <int LAYOUT>
__device__ __forceinline__ void global_tensor(const float* g, size_t i, float& a, float& b, float& c, float& d,
                                              float& e, float& f) {
    if (LAYOUT == TIRA_LAYOUT_MAT9) {
        const float* m = g + i * 9;
        a = m[0]; b = m[1]; d = m[2]; c = m[4]; e = m[5]; f = m[8];
    } else {
        const float* m = g + i * 6;
        a = m[0]; b = m[1]; d = m[2]; c = m[3]; e = m[4]; f = m[5];
    }
}

// Segment table: one launch covers up to kEigMaxSeg independent tensor
// arrays. It travels in the kernel arguments (no upload, capturable); tile t
// belongs to the last segment whose tile0 <= t.
struct EigSeg {
    const float* in;
    float* evals;
    float* evec0;
    float* evec2;
    unsigned long long n;
    unsigned tile0;
    unsigned pad;
};
constexpr int kEigMaxSeg = 64;
struct EigBatchArgs {
    EigSeg seg[kEigMaxSeg];
    int nseg;
};

struct EigTile {
    const float* in;
    float* out;
    float* v0;
    float* v2;
    int cnt;
    bool vec;  // full tile with a 16-byte aligned input: staged with float4 loads
};

template <int W>
__device__ __forceinline__ EigTile eig_tile_at(const EigBatchArgs& args, unsigned t) {
    int lo = 0, hi = args.nseg - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (args.seg[mid].tile0 <= t) lo = mid;
        else hi = mid - 1;
    }
    const EigSeg& sg = args.seg[lo];
    const size_t base = (size_t)(t - sg.tile0) * kEigTile;
    const size_t left = sg.n - base;
    EigTile r;
    r.in = sg.in + base * W;
    r.out = sg.evals + base * 3;
    r.v0 = sg.evec0 ? sg.evec0 + base * 3 : nullptr;
    r.v2 = sg.evec2 ? sg.evec2 + base * 3 : nullptr;
    r.cnt = left < (size_t)kEigTile ? (int)left : kEigTile;
    r.vec = (r.cnt == kEigTile) && ((((uintptr_t)r.in) & 15u) == 0);
    return r;
}

// One block per tile of kEigTile consecutive tensors of one array: staged in
// LDS with 16-byte loads, solved in FP32 (deflation, eig3_solve); tensors
// whose relative eigengap is below 1e-4 are compacted into a block-local list
// and re-solved in FP64 when TIRA_EIG_REFINE_FP64 is set; results are staged
// in LDS and written with 16-byte stores. evec0 / evec2 are written when VECS.
// (The kernel is VALU-bound near the HBM rate: a persistent variant that
// prefetches the next tile into registers lost more to the occupancy its
// extra registers cost than it gained in overlap.)
template <int LAYOUT, bool VECS, bool REFINE>
__global__ __launch_bounds__(kEigThreads) void k_eig3(const EigBatchArgs args, unsigned ntiles, unsigned flags,
                                                      unsigned long long* __restrict__ n_refined) {
    constexpr int W = (LAYOUT == TIRA_LAYOUT_MAT9) ? 9 : 6;
    constexpr int TOT4 = kEigTile * W / 4;
    static_assert((kEigTile * W) % 4 == 0, "tile must be a whole number of float4");
    __shared__ f32x4 s_in4[TOT4];
    __shared__ f32x4 s_out4[kEigTile * 3 / 4];
    __shared__ short s_list[kEigTile];
    __shared__ int s_count;
    float* s_in = reinterpret_cast<float*>(s_in4);
    float* s_out = reinterpret_cast<float*>(s_out4);

    const int tid = threadIdx.x;
    const unsigned t = blockIdx.x;
    if (t >= ntiles) return;
    const EigTile cur = eig_tile_at<W>(args, t);
    const int cnt = cur.cnt;
    if (tid == 0) s_count = 0;

    // ---- stage the tile
    if (cur.vec) {
        const f32x4* g4 = reinterpret_cast<const f32x4*>(cur.in);
#pragma unroll
        for (int k = tid; k < TOT4; k += kEigThreads) s_in4[k] = __builtin_nontemporal_load(&g4[k]);
    } else {
        for (int k = tid; k < cnt * W; k += kEigThreads) s_in[k] = cur.in[k];
    }
    __syncthreads();

    const bool scale_inv = (flags & TIRA_EIG_SCALE_INVARIANT) != 0;
    const int saliency = (flags & TIRA_EIG_SALIENCY) ? ((flags & TIRA_EIG_MAGNITUDE) ? 2 : 1) : 0;
    const float deps = scale_inv ? 1e-30f : 1e-6f;

    // ---- FP32 pass
#pragma unroll
    for (int j = 0; j < kEigPerThread; j++) {
        const int i = j * kEigThreads + tid;
        if (i >= cnt) break;
        float a, b, c, d, e, f;
        lds_tensor<LAYOUT>(s_in, i, a, b, c, d, e, f);
        int ex = 0;
        if (scale_inv) ex = prescale6(a, b, c, d, e, f);
        float lam[3], v0[3], v2[3];
        eig3_solve<float>(a, b, c, d, e, f, deps, lam, VECS ? v0 : nullptr, v2);
        if (REFINE && eig3_near_degenerate(lam[0], lam[1], lam[2])) {
            const int slot = atomicAdd(&s_count, 1);
            s_list[slot] = (short)i;
        }
        if (VECS) {
            if (cur.v0) { cur.v0[i * 3 + 0] = v0[0]; cur.v0[i * 3 + 1] = v0[1]; cur.v0[i * 3 + 2] = v0[2]; }
            if (cur.v2) { cur.v2[i * 3 + 0] = v2[0]; cur.v2[i * 3 + 1] = v2[1]; cur.v2[i * 3 + 2] = v2[2]; }
        }
        store_eig(s_out + i * 3, ldexpf(lam[0], ex), ldexpf(lam[1], ex), ldexpf(lam[2], ex), saliency);
    }
    __syncthreads();

    // ---- FP64 re-solve of the compacted near-degenerate list
    const int nflag = REFINE ? s_count : 0;
    for (int k = tid; k < nflag; k += kEigThreads) {
        const int i = s_list[k];
        float af, bf, cf, df, ef, ff;
        lds_tensor<LAYOUT>(s_in, i, af, bf, cf, df, ef, ff);
        int ex = 0;
        if (scale_inv) ex = prescale6(af, bf, cf, df, ef, ff);
        double lam[3], v0[3], v2[3];
        eig3_solve<double>(af, bf, cf, df, ef, ff, scale_inv ? 1e-30 : 1e-6, lam, VECS ? v0 : nullptr, v2);
        store_eig(s_out + i * 3, ldexpf((float)lam[0], ex), ldexpf((float)lam[1], ex), ldexpf((float)lam[2], ex),
                  saliency);
        if (VECS) {
            if (cur.v0) { cur.v0[i * 3 + 0] = (float)v0[0]; cur.v0[i * 3 + 1] = (float)v0[1]; cur.v0[i * 3 + 2] = (float)v0[2]; }
            if (cur.v2) { cur.v2[i * 3 + 0] = (float)v2[0]; cur.v2[i * 3 + 1] = (float)v2[1]; cur.v2[i * 3 + 2] = (float)v2[2]; }
        }
    }
    if (n_refined && tid == 0 && nflag) atomicAdd(n_refined, (unsigned long long)nflag);
    __syncthreads();

    // ---- write back
    if (cnt == kEigTile && ((((uintptr_t)cur.out) & 15u) == 0)) {
        f32x4* g4 = reinterpret_cast<f32x4*>(cur.out);
#pragma unroll
        for (int k = tid; k < (kEigTile * 3) / 4; k += kEigThreads) __builtin_nontemporal_store(s_out4[k], &g4[k]);
    } else {
        for (int k = tid; k < cnt * 3; k += kEigThreads) cur.out[k] = s_out[k];
    }
}

// evecs3spherical_symmetric (eigen.h:744-755): (theta, phi) of e0, e1, e2.
__global__ __launch_bounds__(256) void k_evecs3spherical(const float* __restrict__ mats,
                                                         const float* __restrict__ lam, size_t n,
                                                         float* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float a, b, c, d, e, f;
    global_tensor<TIRA_LAYOUT_MAT9>(mats, i, a, b, c, d, e, f);
    float l0 = lam[i * 3], l1 = lam[i * 3 + 1], l2 = lam[i * 3 + 2];
    float v[3][3];
    if (eig3_near_degenerate(l0, l1, l2)) {
        double dv[3][3];
        eig3_vecs<double>(a, b, c, d, e, f, l0, l1, l2, 1e-6, dv[0], dv[1], dv[2]);
        for (int k = 0; k < 3; k++)
            for (int r = 0; r < 3; r++) v[k][r] = (float)dv[k][r];
    } else {
        eig3_vecs<float>(a, b, c, d, e, f, l0, l1, l2, 1e-6f, v[0], v[1], v[2]);
    }
    for (int k = 0; k < 3; k++) {
        const float z = fminf(1.0f, fmaxf(-1.0f, v[k][2]));
        out[i * 6 + 2 * k + 0] = atan2f(v[k][1], v[k][0]);
        out[i * 6 + 2 * k + 1] = acosf(z);
    }
}

// kernel_eval2_symmetric (eigen.h:700-712)
__global__ __launch_bounds__(256) void k_evals2(const float* __restrict__ m, size_t n, float* __restrict__ ev) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 t = reinterpret_cast<const float4*>(m)[i];
    float e0, e1;
    eig2_vals(t.x, t.y, t.w, e0, e1);
    reinterpret_cast<float2*>(ev)[i] = make_float2(e0, e1);
}

// kernel_evec2polar_symmetric (eigen.h:714-722)
__global__ __launch_bounds__(256) void k_evecs2polar(const float* __restrict__ m, const float* __restrict__ ev,
                                                     size_t n, float* __restrict__ th) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 t = reinterpret_cast<const float4*>(m)[i];
    float vx, vy, ux, uy;
    eig2_vecs(t.x, t.y, t.w, ev[i * 2], vx, vy, ux, uy);
    float t0 = atan2f(vy, vx), t1 = atan2f(uy, ux);
    const float pi = 3.14159265358979323846f;
    if (t0 <= -pi) t0 += 2.0f * pi;
    if (t1 <= -pi) t1 += 2.0f * pi;
    if (t0 > pi) t0 -= 2.0f * pi;
    if (t1 > pi) t1 -= 2.0f * pi;
    reinterpret_cast<float2*>(th)[i] = make_float2(t0, t1);
}

template <int LAYOUT, bool VECS>
static void launch_eig3_args(const EigBatchArgs& a, unsigned tiles, unsigned flags, unsigned long long* n_refined,
                             hipStream_t st) {
    if (flags & TIRA_EIG_REFINE_FP64) k_eig3<LAYOUT, VECS, true><<<tiles, kEigThreads, 0, st>>>(a, tiles, flags, n_refined);
    else k_eig3<LAYOUT, VECS, false><<<tiles, kEigThreads, 0, st>>>(a, tiles, flags, n_refined);
}

static int launch_eig3_table(const EigBatchArgs& a, unsigned tiles, int layout, bool vecs, unsigned flags,
                             unsigned long long* n_refined, hipStream_t st) {
    if (layout == TIRA_LAYOUT_MAT9) {
        if (vecs) launch_eig3_args<TIRA_LAYOUT_MAT9, true>(a, tiles, flags, n_refined, st);
        else launch_eig3_args<TIRA_LAYOUT_MAT9, false>(a, tiles, flags, n_refined, st);
    } else {
        if (vecs) launch_eig3_args<TIRA_LAYOUT_PACKED6, true>(a, tiles, flags, n_refined, st);
        else launch_eig3_args<TIRA_LAYOUT_PACKED6, false>(a, tiles, flags, n_refined, st);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// Largest tensor count one launch addresses (32-bit tile index).
constexpr size_t kEigMaxTensors = (size_t)0x7fffffffu * kEigTile;

int launch_eig3(const float* in, size_t n, int layout, float* evals, float* evec0, float* evec2, unsigned flags,
                unsigned long long* n_refined, hipStream_t st) {
    if (n == 0) return TIRA_OK;
    if (layout != TIRA_LAYOUT_MAT9 && layout != TIRA_LAYOUT_PACKED6) {
        set_error("unknown tensor layout %d", layout);
        return TIRA_ERR_INVALID_ARG;
    }
    if (n > kEigMaxTensors) {
        set_error("eig3: %zu tensors exceed the %zu one launch addresses", n, kEigMaxTensors);
        return TIRA_ERR_INVALID_ARG;
    }
    EigBatchArgs a{};
    a.nseg = 1;
    a.seg[0] = EigSeg{in, evals, evec0, evec2, (unsigned long long)n, 0u, 0u};
    ProfScope prof("eig3", st);
    return launch_eig3_table(a, ceil_div(n, kEigTile), layout, evec0 || evec2, flags, n_refined, st);
}

// Up to kEigMaxSeg non-empty arrays per launch.
int launch_eig3_batch(const float* const* in, const size_t* counts, int nvol, int layout, float* const* evals,
                      float* const* evec0, float* const* evec2, unsigned flags, unsigned long long* n_refined,
                      hipStream_t st) {
    if (layout != TIRA_LAYOUT_MAT9 && layout != TIRA_LAYOUT_PACKED6) {
        set_error("unknown tensor layout %d", layout);
        return TIRA_ERR_INVALID_ARG;
    }
    const bool vecs = evec0 || evec2;
    ProfScope prof("eig3_batch", st);
    int v = 0;
    while (v < nvol) {
        EigBatchArgs a{};
        size_t tiles = 0;
        for (; v < nvol && a.nseg < kEigMaxSeg; v++) {
            if (!counts[v]) continue;
            if (counts[v] > kEigMaxTensors) {
                set_error("eig3: array %d holds %zu tensors, above the %zu one launch addresses", v, counts[v],
                          kEigMaxTensors);
                return TIRA_ERR_INVALID_ARG;
            }
            const size_t t = ceil_div(counts[v], kEigTile);
            if (tiles + t > 0x7fffffffu) break;
            EigSeg& s = a.seg[a.nseg++];
            s.in = in[v];
            s.evals = evals[v];
            s.evec0 = evec0 ? evec0[v] : nullptr;
            s.evec2 = evec2 ? evec2[v] : nullptr;
            s.n = counts[v];
            s.tile0 = (unsigned)tiles;
            tiles += t;
        }
        if (!a.nseg) break;
        TH_TRY(launch_eig3_table(a, (unsigned)tiles, layout, vecs, flags, n_refined, st));
    }
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_evals2_symmetric(const float* mats, size_t n, float* evals, int device) {
    TH_TRY(use_device(device));
    TH_ARG(mats && evals, "NULL pointer");
    if (!n) return TIRA_OK;
    const float* dm; void* de;
    TH_TRY(stage_in(mats, n * 16, WS_IN0, (const void**)&dm));
    TH_TRY(stage_out(evals, n * 8, WS_OUT0, &de));
    k_evals2<<<ceil_div(n, 256), 256, 0, stream()>>>(dm, n, (float*)de);
    TH_HIP(hipGetLastError());
    return finish_out(evals, de, n * 8);
}

int tira_evecs2polar_symmetric(const float* mats, const float* evals, size_t n, float* thetas, int device) {
    TH_TRY(use_device(device));
    TH_ARG(mats && evals && thetas, "NULL pointer");
    if (!n) return TIRA_OK;
    const float *dm, *dl; void* dt;
    TH_TRY(stage_in(mats, n * 16, WS_IN0, (const void**)&dm));
    TH_TRY(stage_in(evals, n * 8, WS_IN1, (const void**)&dl));
    TH_TRY(stage_out(thetas, n * 8, WS_OUT0, &dt));
    k_evecs2polar<<<ceil_div(n, 256), 256, 0, stream()>>>(dm, dl, n, (float*)dt);
    TH_HIP(hipGetLastError());
    return finish_out(thetas, dt, n * 8);
}

int tira_evals3_symmetric(const float* mats, size_t n, float* evals, int device) {
    TH_TRY(use_device(device));
    TH_ARG(mats && evals, "NULL pointer");
    if (!n) return TIRA_OK;
    const float* dm; void* de;
    TH_TRY(stage_in(mats, n * 36, WS_IN0, (const void**)&dm));
    TH_TRY(stage_out(evals, n * 12, WS_OUT0, &de));
    TH_TRY(launch_eig3(dm, n, TIRA_LAYOUT_MAT9, (float*)de, nullptr, nullptr,
                       TIRA_EIG_REFINE_FP64 | TIRA_EIG_SCALE_INVARIANT, nullptr, stream()));
    return finish_out(evals, de, n * 12);
}

int tira_evecs3spherical_symmetric(const float* mats, const float* lambdas, size_t n, float* evecs, int device) {
    TH_TRY(use_device(device));
    TH_ARG(mats && lambdas && evecs, "NULL pointer");
    if (!n) return TIRA_OK;
    const float *dm, *dl; void* dv;
    TH_TRY(stage_in(mats, n * 36, WS_IN0, (const void**)&dm));
    TH_TRY(stage_in(lambdas, n * 12, WS_IN1, (const void**)&dl));
    TH_TRY(stage_out(evecs, n * 24, WS_OUT0, &dv));
    k_evecs3spherical<<<ceil_div(n, 256), 256, 0, stream()>>>(dm, dl, n, (float*)dv);
    TH_HIP(hipGetLastError());
    return finish_out(evecs, dv, n * 24);
}

int tira_eig3_async(const float* tensors, size_t n, int layout, float* evals, float* evec0, float* evec2,
                    unsigned flags, unsigned long long* n_refined_dev) {
    TH_ARG(tensors && evals, "NULL pointer");
    return launch_eig3(tensors, n, layout, evals, evec0, evec2, flags, n_refined_dev, stream());
}

int tira_eig3_batch_async(const float* const* tensors, const size_t* counts, int nvol, int layout,
                          float* const* evals, float* const* evec0, float* const* evec2, unsigned flags,
                          unsigned long long* n_refined_dev) {
    TH_ARG(nvol >= 0, "negative volume count");
    if (!nvol) return TIRA_OK;
    TH_ARG(tensors && counts && evals, "NULL pointer");
    // (pointer-attribute queries cost ~10 us each, more than the launch for
    // many small arrays: only NULLs are rejected; device memory is the contract)
    for (int v = 0; v < nvol; v++)
        TH_ARG(!counts[v] || (tensors[v] && evals[v]), "NULL tensor or eigenvalue pointer in volume %d", v);
    return launch_eig3_batch(tensors, counts, nvol, layout, evals, evec0, evec2, flags, n_refined_dev, stream());
}

}  // extern "C"
