// voxel_payload.hpp -- per-voxel eigen-decomposition into the voter payload
// of the 3D vote, shared by the gather kernels (vote3.hip) and the spectral
// vote (spectral.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "eig_device.hpp"
#include "tira_hip.h"

namespace th {

__device__ __forceinline__ void load_tensor(const float* __restrict__ in, int layout, size_t v, float& a, float& b,
                                            float& c, float& d, float& e, float& f) {
    if (layout == TIRA_LAYOUT_MAT9) {
        const float* m = in + v * 9;
        a = m[0]; b = m[1]; d = m[2]; c = m[4]; e = m[5]; f = m[8];
    } else {
        const float* m = in + v * 6;
        a = m[0]; b = m[1]; d = m[2]; c = m[3]; e = m[4]; f = m[5];
    }
}

// Eigen-decomposition of voxel v for the payload: power-of-two prescale
// (exact; makes every threshold relative to |A|), FP32 deflated solve, and
// with REFINE an FP64 re-solve when the relative eigengap is below 1e-4.
// Without REFINE the close pair's eigenvectors carry FP32 errors of order
// eps |A| / gap; the vote weighs e2 by |l2| - |l1| and e0 by |l1| - |l0|,
// so a vector of an ill-separated pair enters multiplied by that pair's gap
// and the product stays of order eps |A|.
template <bool REFINE = true>
__device__ __forceinline__ void voxel_eig(const float* __restrict__ in, int layout, size_t v, float& l0, float& l1,
                                          float& l2, float e0[3], float e2[3]) {
    float a, b, c, d, e, f;
    load_tensor(in, layout, v, a, b, c, d, e, f);
    const float m = fmaxf(fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fabsf(d))), fmaxf(fabsf(e), fabsf(f)));
    int ex = 0;
    if (m > 0.0f && isfinite(m)) {
        frexpf(m, &ex);
        a = ldexpf(a, -ex); b = ldexpf(b, -ex); c = ldexpf(c, -ex);
        d = ldexpf(d, -ex); e = ldexpf(e, -ex); f = ldexpf(f, -ex);
    }
    float lam[3];
    eig3_solve<float>(a, b, c, d, e, f, 1e-30f, lam, e0, e2);
    if (REFINE && eig3_near_degenerate(lam[0], lam[1], lam[2])) {
        double dl[3], de0[3], de2[3];
        eig3_solve<double>(a, b, c, d, e, f, 1e-30, dl, de0, de2);
        for (int k = 0; k < 3; k++) { lam[k] = (float)dl[k]; e0[k] = (float)de0[k]; e2[k] = (float)de2[k]; }
    }
    l0 = ldexpf(lam[0], ex); l1 = ldexpf(lam[1], ex); l2 = ldexpf(lam[2], ex);
}

// Stick payload (e2, copysign(|l2| - |l1|, l2)) and plate payload
// (e0, copysign(|l1| - |l0|, l1)) (tensorvote.h:2438-2442, 2609-2616). The
// reference skips plate voters whose normal has |n|^2 <= 1e-12 and
// normalises the rest (tensorvote.h:2464-2467).
__device__ __forceinline__ void stick_plate_payload(float l0, float l1, float l2, const float* e0, const float* e2,
                                                    float4& S, float4& N) {
    S = make_float4(e2[0], e2[1], e2[2], copysignf(fabsf(l2) - fabsf(l1), l2));
    float n[3] = {e0[0], e0[1], e0[2]};
    const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
    float sp = copysignf(fabsf(l1) - fabsf(l0), l1);
    if (!(nn > 1e-12f)) { sp = 0.0f; n[0] = n[1] = n[2] = 0.0f; }
    else { const float inv = rsqrtf(nn); n[0] *= inv; n[1] *= inv; n[2] *= inv; }
    N = make_float4(n[0], n[1], n[2], sp);
}

}  // namespace th
