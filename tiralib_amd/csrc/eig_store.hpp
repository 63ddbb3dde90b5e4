// eig_store.hpp -- per-tensor eigenvalue steps shared by the eig3 kernels
// and the kernels that produce a field and its eigenvalues in one pass.
#pragma once
#include <hip/hip_runtime.h>

#include "eig_device.hpp"
#include "tira_hip.h"

namespace th {

// Power-of-two prescale so that the largest |entry| lies in [0.5, 1): exact
// in floating point, and it turns the reference's absolute thresholds into
// relative ones. Returns the exponent to scale eigenvalues back with.
__device__ __forceinline__ int prescale6(float& a, float& b, float& c, float& d, float& e, float& f) {
    float m = fmaxf(fmaxf(fmaxf(fabsf(a), fabsf(b)), fmaxf(fabsf(c), fabsf(d))), fmaxf(fabsf(e), fabsf(f)));
    if (!(m > 0.0f) || !isfinite(m)) return 0;
    int ex;
    frexpf(m, &ex);
    a = ldexpf(a, -ex); b = ldexpf(b, -ex); c = ldexpf(c, -ex);
    d = ldexpf(d, -ex); e = ldexpf(e, -ex); f = ldexpf(f, -ex);
    return ex;
}

// Eigenvalues l0 <= l1 <= l2 (mode 0), the saliency maps (l2 - l1, l1 - l0,
// l0) -- stick, plate and ball saliency of a voted field (mode 1) -- or the
// same with the eigenvalues ranked by magnitude, m0 <= m1 <= m2 of |l|
// (mode 2, the ranking of reconstruct.py:14-34).
__device__ __forceinline__ void store_eig(float* o, float l0, float l1, float l2, int mode) {
    if (mode == 2) {
        const float a = fabsf(l0), b = fabsf(l1), c = fabsf(l2);
        const float m2 = fmaxf(a, fmaxf(b, c)), m0 = fminf(a, fminf(b, c));
        const float m1 = __builtin_amdgcn_fmed3f(a, b, c);
        o[0] = m2 - m1; o[1] = m1 - m0; o[2] = m0;
    } else if (mode == 1) { o[0] = l2 - l1; o[1] = l1 - l0; o[2] = l0; }
    else { o[0] = l0; o[1] = l1; o[2] = l2; }
}

// Eigenvalues (or saliency, store_eig's modes) of one tensor (a, b, c, d, e, f)
// = (xx, xy, yy, xz, yz, zz) under TIRA_EIG_* `flags`, the steps of k_eig3 for
// one tensor: optional prescale, FP32 solve, and an FP64 re-solve when the
// FP32 eigenvalues are near-degenerate and TIRA_EIG_REFINE_FP64 is set.
__device__ __forceinline__ void eig3_tensor_values(float a, float b, float c, float d, float e, float f,
                                                   unsigned flags, float* o,
                                                   unsigned long long* n_refined) {
    const bool scale_inv = (flags & TIRA_EIG_SCALE_INVARIANT) != 0;
    const int saliency = (flags & TIRA_EIG_SALIENCY) ? ((flags & TIRA_EIG_MAGNITUDE) ? 2 : 1) : 0;
    int ex = 0;
    if (scale_inv) ex = prescale6(a, b, c, d, e, f);
    float lam[3], v2[3];
    eig3_solve<float>(a, b, c, d, e, f, scale_inv ? 1e-30f : 1e-6f, lam, nullptr, v2);
    if ((flags & TIRA_EIG_REFINE_FP64) && eig3_near_degenerate(lam[0], lam[1], lam[2])) {
        double dl[3], dv2[3];
        eig3_solve<double>(a, b, c, d, e, f, scale_inv ? 1e-30 : 1e-6, dl, nullptr, dv2);
        store_eig(o, ldexpf((float)dl[0], ex), ldexpf((float)dl[1], ex), ldexpf((float)dl[2], ex), saliency);
        if (n_refined) atomicAdd(n_refined, 1ull);
        return;
    }
    store_eig(o, ldexpf(lam[0], ex), ldexpf(lam[1], ex), ldexpf(lam[2], ex), saliency);
}

}  // namespace th
