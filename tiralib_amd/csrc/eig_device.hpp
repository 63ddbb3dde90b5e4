// eig_device.hpp -- eigen-solvers for symmetric 2x2 / 3x3 tensors (host and
// device).
//
// Same mathematics as TIRA's eigen.h, written for CDNA4:
//  * eigenvalues of 3x3: Smith's trigonometric closed form (eigen.h:391-446),
//    with cos(phi + 2pi/3) taken from a single sincos(phi) by angle addition,
//    then repaired by deflation (eig3_solve): the trigonometric values alone
//    lose accuracy as 1 / (relative gap of the closest pair) -- up to 2e-4
//    normwise in FP32 for eigenvalues spread over five decades -- while the
//    deflated FP32 result stays within ~6e-7 of FP64 normwise;
//  * eigenvectors of 3x3: cofactor row of the best-separated eigenvalue, the
//    other two from the 2x2 problem on its orthogonal complement
//    (eigen.h:217-370, 452-512), canonical sign (largest |component|
//    positive, eigen.h:525-537);
//  * 2x2: stable quadratic and the larger-row null vector (eigen.h:89-212).
// Every routine is a template on the arithmetic type: float carries the bulk
// of the work; double re-solves tensors whose relative eigengap is below
// kRefineGap, where the eigenvectors of the close pair are ill-conditioned.
#pragma once
#include <hip/hip_runtime.h>

#ifndef TH_HD
#define TH_HD __host__ __device__ __forceinline__
#endif

namespace th {

// Absolute thresholds of the reference (TIRA_EIGEN_EPSILON, eigen.h:13).
template <typename T> TH_HD T eig_eps() { return T(1e-6); }

constexpr float kRefineGap = 1e-4f;  // relative eigengap that sends a tensor to FP64

TH_HD float d_acos(float x) { return acosf(x); }
TH_HD double d_acos(double x) { return acos(x); }
TH_HD void d_sincos(float x, float* s, float* c) { sincosf(x, s, c); }
TH_HD void d_sincos(double x, double* s, double* c) { sincos(x, s, c); }
TH_HD float d_sqrt(float x) { return sqrtf(x); }
TH_HD double d_sqrt(double x) { return sqrt(x); }
TH_HD float d_atan2(float y, float x) { return atan2f(y, x); }
TH_HD double d_atan2(double y, double x) { return atan2(y, x); }
template <typename T> TH_HD T d_abs(T x) { return x < T(0) ? -x : x; }

template <typename T>
TH_HD void sort3(T& x, T& y, T& z) {
    T t;
    if (x > y) { t = x; x = y; y = t; }
    if (y > z) { t = y; y = z; z = t; }
    if (x > y) { t = x; x = y; y = t; }
}
#ifdef __HIP_DEVICE_COMPILE__
// min3 / med3 / max3 (one instruction each on the device)
template <>
TH_HD void sort3<float>(float& x, float& y, float& z) {
    const float lo = fminf(fminf(x, y), z), hi = fmaxf(fmaxf(x, y), z), mid = __builtin_amdgcn_fmed3f(x, y, z);
    x = lo; y = mid; z = hi;
}
#endif

// Eigenvalues (ascending) of | a b d ; b c e ; d e f |. `diag_eps` is the
// off-diagonal energy below which the diagonal is returned sorted (the
// reference uses an absolute 1e-6; scale-invariant callers pass a value
// relative to the tensor norm).
template <typename T>
TH_HD void eig3_vals(T a, T b, T c, T d, T e, T f, T diag_eps, T& l0, T& l1, T& l2) {
    const T p1 = b * b + d * d + e * e;
    if (p1 < diag_eps) {
        l0 = a; l1 = c; l2 = f;
        sort3(l0, l1, l2);
        return;
    }
    const T tr = a + c + f;
    const T q = tr * T(1.0 / 3.0);
    const T aq = a - q, cq = c - q, fq = f - q;
    const T p2 = aq * aq + cq * cq + fq * fq + T(2) * p1;
    const T p = d_sqrt(p2 * T(1.0 / 6.0));
    const T pinv = T(1) / p;
    const T Ba = pinv * aq, Bb = pinv * b, Bc = pinv * cq, Bd = pinv * d, Be = pinv * e, Bf = pinv * fq;
    const T detB = Ba * (Bc * Bf - Be * Be) - Bb * (Bb * Bf - Bd * Be) + Bd * (Bb * Be - Bd * Bc);
    const T r = detB * T(0.5);
    T phi;
    if (r <= T(-1)) phi = T(3.14159265358979323846 / 3.0);
    else if (r >= T(1)) phi = T(0);
    else phi = d_acos(r) * T(1.0 / 3.0);
    T sp, cp;
    d_sincos(phi, &sp, &cp);
    l2 = q + T(2) * p * cp;
    // cos(phi + 2pi/3) = -cos(phi)/2 - sqrt(3)/2 sin(phi)
    l0 = q - p * (cp + T(1.7320508075688772935) * sp);
    l1 = tr - l0 - l2;
}

// True when the relative gap between adjacent eigenvalues is below kRefineGap
// (and the tensor is not exactly diagonal-handled / zero).
template <typename T>
TH_HD bool eig3_near_degenerate(T l0, T l1, T l2) {
    const T scale = d_abs(l0) > d_abs(l2) ? d_abs(l0) : d_abs(l2);
    const T g = (l1 - l0) < (l2 - l1) ? (l1 - l0) : (l2 - l1);
    return scale > T(0) && g < T(kRefineGap) * scale;
}

template <typename T>
TH_HD void cross3(const T* a, const T* b, T* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
TH_HD void normalize3_eps(T* v) {  // normalize3, eigen.h:56-65
    const T n = d_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > eig_eps<T>()) {
        const T inv = T(1) / n;
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    }
}

// Largest cofactor row of (A - lambda I) (unnormalised); returns its squared
// norm. adj(A - lambda I) is symmetric (6 distinct cofactors), and for a
// simple eigenvalue it is k v v^T, so the row with the largest diagonal
// entry |k| v_i^2 is the row with the largest norm.
template <typename T>
TH_HD T cofactor_row(T a, T b, T c, T d, T e, T f, T lam, T* v) {
    const T m00 = a - lam, m11 = c - lam, m22 = f - lam;
    const T c00 = m11 * m22 - e * e, c11 = m00 * m22 - d * d, c22 = m00 * m11 - b * b;
    const T c01 = d * e - b * m22, c02 = b * e - d * m11, c12 = b * d - m00 * e;
    const T a0 = d_abs(c00), a1 = d_abs(c11), a2 = d_abs(c22);
    if (a0 >= a1 && a0 >= a2) { v[0] = c00; v[1] = c01; v[2] = c02; }
    else if (a1 >= a2) { v[0] = c01; v[1] = c11; v[2] = c12; }
    else { v[0] = c02; v[1] = c12; v[2] = c22; }
    return v[0] * v[0] + v[1] * v[1] + v[2] * v[2];
}

// Null vector of (A - lambda I) from its largest cofactor row (eigen.h:247-288).
template <typename T>
TH_HD void evec_cofactor(T a, T b, T c, T d, T e, T f, T lam, T* v) {
    const T m00 = a - lam, m11 = c - lam, m22 = f - lam;
    const T c00 = m11 * m22 - e * e, c01 = e * d - b * m22, c02 = b * e - m11 * d;
    const T c10 = d * e - b * m22, c11 = m00 * m22 - d * d, c12 = b * d - m00 * e;
    const T c20 = b * e - d * m11, c21 = d * b - m00 * e, c22 = m00 * m11 - b * b;
    const T r0 = c00 * c00 + c01 * c01 + c02 * c02;
    const T r1 = c10 * c10 + c11 * c11 + c12 * c12;
    const T r2 = c20 * c20 + c21 * c21 + c22 * c22;
    if (r0 >= r1 && r0 >= r2) { v[0] = c00; v[1] = c01; v[2] = c02; }
    else if (r1 >= r0 && r1 >= r2) { v[0] = c10; v[1] = c11; v[2] = c12; }
    else { v[0] = c20; v[1] = c21; v[2] = c22; }
    normalize3_eps(v);
}

// Eigenvector of lam1 in the orthogonal complement of `ev` (eigen.h:217-245, 290-370).
template <typename T>
TH_HD void evec_complement(T a, T b, T c, T d, T e, T f, T lam1, const T* ev, T* out) {
    const T eps = eig_eps<T>();
    const T ax = d_abs(ev[0]), ay = d_abs(ev[1]), az = d_abs(ev[2]);
    T t[3] = {T(0), T(0), T(1)};
    if (ay - ax > eps) {
        if (az - ax > eps) { t[0] = T(1); t[2] = T(0); }
    } else if (az - ay > eps) {
        t[1] = T(1); t[2] = T(0);
    }
    T U[3], V[3];
    cross3(ev, t, U);
    normalize3_eps(U);
    cross3(ev, U, V);
    const T AU0 = a * U[0] + b * U[1] + d * U[2], AU1 = b * U[0] + c * U[1] + e * U[2],
            AU2 = d * U[0] + e * U[1] + f * U[2];
    const T AV0 = a * V[0] + b * V[1] + d * V[2], AV1 = b * V[0] + c * V[1] + e * V[2],
            AV2 = d * V[0] + e * V[1] + f * V[2];
    T m00 = U[0] * AU0 + U[1] * AU1 + U[2] * AU2 - lam1;
    T m01 = U[0] * AV0 + U[1] * AV1 + U[2] * AV2;
    T m11 = V[0] * AV0 + V[1] * AV1 + V[2] * AV2 - lam1;
    const T a00 = d_abs(m00), a01 = d_abs(m01), a11 = d_abs(m11);
    T s0, s1;  // out = s0 * U - s1 * V
    bool use_u = false;
    if (a00 >= a11) {
        if (a00 > eps || a01 > eps) {
            if (a00 >= a01) { m01 /= m00; const T il = T(1) / d_sqrt(T(1) + m01 * m01); s1 = il; s0 = m01 * il; }
            else { m00 /= m01; const T il = T(1) / d_sqrt(T(1) + m00 * m00); s0 = il; s1 = m00 * il; }
        } else use_u = true;
    } else {
        if (a11 > eps || a01 > eps) {
            if (a11 >= a01) { m01 /= m11; const T il = T(1) / d_sqrt(T(1) + m01 * m01); s0 = il; s1 = m01 * il; }
            else { m11 /= m01; const T il = T(1) / d_sqrt(T(1) + m11 * m11); s1 = il; s0 = m11 * il; }
        } else use_u = true;
    }
    if (use_u) { out[0] = U[0]; out[1] = U[1]; out[2] = U[2]; }
    else { for (int k = 0; k < 3; k++) out[k] = s0 * U[k] - s1 * V[k]; }
}

template <typename T>
TH_HD void canonical_sign(T* v) {  // eigen.h:525-537
    const T ax = d_abs(v[0]), ay = d_abs(v[1]), az = d_abs(v[2]);
    int k = 0;
    T am = ax;
    if (ay > am) { k = 1; am = ay; }
    if (az > am) { k = 2; }
    const T s = (v[k] < T(0)) ? T(-1) : T(1);
    v[0] *= s; v[1] *= s; v[2] *= s;
}

// Eigenvectors e0 (smallest), e1, e2 (largest) of the tensor with eigenvalues
// l0 <= l1 <= l2 (eigen.h:452-512), canonical signs applied.
template <typename T>
TH_HD void eig3_vecs(T a, T b, T c, T d, T e, T f, T l0, T l1, T l2, T diag_eps,
                                          T* e0, T* e1, T* e2) {
    const T ud = b * b + d * d + e * e;
    if (ud < diag_eps) {
        // diagonal: basis vectors ordered like the sorted diagonal (eigen.h:461-488)
        T v0 = a, v1 = c, v2 = f;
        int i0 = 0, i1 = 1, i2 = 2, ti;
        T tv;
        if (v0 > v1) { tv = v0; v0 = v1; v1 = tv; ti = i0; i0 = i1; i1 = ti; }
        if (v1 > v2) { tv = v1; v1 = v2; v2 = tv; ti = i1; i1 = i2; i2 = ti; }
        if (v0 > v1) { tv = v0; v0 = v1; v1 = tv; ti = i0; i0 = i1; i1 = ti; }
        for (int k = 0; k < 3; k++) {
            e0[k] = (k == i0) ? T(1) : T(0);
            e1[k] = (k == i1) ? T(1) : T(0);
            e2[k] = (k == i2) ? T(1) : T(0);
        }
        return;
    }
    const T q = (a + c + f) * T(1.0 / 3.0);
    const T aq = a - q, cq = c - q, fq = f - q;
    const T p2 = aq * aq + cq * cq + fq * fq + T(2) * ud;
    const T p = d_sqrt(p2 * T(1.0 / 6.0));
    T hd = T(0);
    if (!(p < eig_eps<T>())) {
        const T det = aq * (cq * fq - e * e) - b * (b * fq - e * d) + d * (b * e - cq * d);
        hd = det / (T(2) * p * p * p);
    }
    if (hd >= T(0)) {
        evec_cofactor(a, b, c, d, e, f, l2, e2);
        evec_complement(a, b, c, d, e, f, l1, e2, e1);
        cross3(e1, e2, e0);
    } else {
        evec_cofactor(a, b, c, d, e, f, l0, e0);
        evec_complement(a, b, c, d, e, f, l1, e0, e1);
        cross3(e0, e1, e2);
        normalize3_eps(e2);
    }
    canonical_sign(e0);
    canonical_sign(e1);
    canonical_sign(e2);
}

// Fast reciprocal / reciprocal square root: hardware approximations in FP32
// device code (used only inside self-correcting or normalising steps), exact
// division elsewhere. FP64 device code takes the hardware approximation and
// two Newton steps (relative error at the level of the last bit instead of
// the correctly rounded division and square-root sequences, a third of their
// instructions); the operands are positive and normal wherever these are
// called (a spread p > 0, 3 (b^2 - 1) >= 6, squared norms of nonzero
// vectors).
TH_HD float d_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
TH_HD double d_rcp(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double r = __builtin_amdgcn_rcp(x);
    r = fma(r, fma(-x, r, 1.0), r);
    return fma(r, fma(-x, r, 1.0), r);
#else
    return 1.0 / x;
#endif
}
TH_HD float d_rsqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rsqf(x);
#else
    return 1.0f / sqrtf(x);
#endif
}
TH_HD double d_rsqrt(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    double y = __builtin_amdgcn_rsq(x);
    y = fma(0.5 * y, fma(-x * y, y, 1.0), y);
    return fma(0.5 * y, fma(-x * y, y, 1.0), y);
#else
    return 1.0 / sqrt(x);
#endif
}
// Square root without the correctly-rounded fix-up sequence (hardware
// v_sqrt_f32, about 1 ulp) in FP32 device code; exact sqrt elsewhere.
TH_HD float d_sqrt_fast(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
TH_HD double d_sqrt_fast(double x) {
#ifdef __HIP_DEVICE_COMPILE__
    // x rsqrt(x) with a final Newton correction of the root itself; 0 stays 0
    const double y = d_rsqrt(x), s = x * y;
    return x > 0.0 ? fma(0.5 * y, fma(-s, s, x), s) : 0.0;
#else
    return sqrt(x);
#endif
}

// Largest root of beta^3 - 3 beta - 2 r = 0 for r in [0, 1], i.e.
// beta = 2 cos(acos(r) / 3) in [sqrt 3, 2]: a cubic fit (max error 1.3e-4)
// refined by Newton (derivative >= 6 on the interval, so one step reaches
// FP32 accuracy and two reach FP64 accuracy).
template <typename T>
TH_HD T cubic_top_root(T r) {
    T b = ((T(0.01874897) * r - T(0.08126101)) * r + T(0.33042943)) * r + T(1.73218504);
    constexpr int kIt = sizeof(T) == 4 ? 1 : 2;
#pragma unroll
    for (int i = 0; i < kIt; i++) {
        const T b2 = b * b;
        const T f = (b2 - T(3)) * b - T(2) * r;
        b -= f * d_rcp(T(3) * (b2 - T(1)));
    }
    return b;
}

// Eigenvalues with normwise accuracy. Smith's closed form (eigen.h:391-446)
// writes the eigenvalues as q + p beta_k, beta_k the roots of
// beta^3 - 3 beta - 2 r = 0 with r = det((A - qI)/p) / 2; its trigonometric
// evaluation loses accuracy as 1 / (relative gap of the closest pair), up to
// 2e-4 normwise in FP32 for eigenvalues spread over five decades. Here only
// the best-separated root is taken from the cubic (largest when r >= 0,
// smallest otherwise -- the reference's halfDet test, eigen.h:497); its
// cofactor eigenvector v (eigen.h:247-288) deflates A to the plane orthogonal
// to v, where the close pair is a symmetric 2x2 problem (perfectly
// conditioned), and the separated eigenvalue becomes the Rayleigh quotient
// v.Av (taken from the trace). Optional outputs: unit eigenvectors e0 (smallest) and e2 (largest)
// with canonical sign.
template <typename T>
TH_HD void eig3_solve(T a, T b, T c, T d, T e, T f, T diag_eps, T* lam, T* e0, T* e2) {
    const T p1 = b * b + d * d + e * e;
    if (p1 < diag_eps) {
        T l0 = a, l1 = c, l2 = f;
        sort3(l0, l1, l2);
        lam[0] = l0; lam[1] = l1; lam[2] = l2;
        if (e0) {
            T v0[3], v1[3], v2[3];
            eig3_vecs<T>(a, b, c, d, e, f, l0, l1, l2, diag_eps, v0, v1, v2);
            for (int k = 0; k < 3; k++) { e0[k] = v0[k]; e2[k] = v2[k]; }
        }
        return;
    }
    const T q = (a + c + f) * T(1.0 / 3.0);
    const T aq = a - q, cq = c - q, fq = f - q;
    const T p = d_sqrt_fast((aq * aq + cq * cq + fq * fq + T(2) * p1) * T(1.0 / 6.0));
    const T pinv = d_rcp(p);
    const T det = aq * (cq * fq - e * e) - b * (b * fq - e * d) + d * (b * e - cq * d);
    T r = det * (pinv * pinv * pinv) * T(0.5);
    r = r > T(1) ? T(1) : (r < T(-1) ? T(-1) : r);
    const bool top = r >= T(0);
    const T beta = cubic_top_root<T>(top ? r : -r);
    const T lsep = top ? q + p * beta : q - p * beta;
    T v[3];
    // cofactors scale with the square of the eigenvalue spread: with callers
    // prescaling |A| into [0.5, 1), a row below 1e-24 carries no direction
    const T rr = cofactor_row(a, b, c, d, e, f, lsep, v);
    if (!(rr > T(1e-24))) {
        // near-isotropic: the other two roots by Vieta (sum 0, pairwise products -3)
        const T h = T(0.5) * d_sqrt(fmax(T(12) - T(3) * beta * beta, T(0)));
        T x0, x1, x2;
        if (top) { x2 = lsep; x0 = q + p * (T(-0.5) * beta - h); x1 = q + p * (T(-0.5) * beta + h); }
        else { x0 = lsep; x1 = q + p * (T(0.5) * beta - h); x2 = q + p * (T(0.5) * beta + h); }
        sort3(x0, x1, x2);
        lam[0] = x0; lam[1] = x1; lam[2] = x2;
        if (e0) {
            T v0[3], v1[3], v2[3];
            eig3_vecs<T>(a, b, c, d, e, f, x0, x1, x2, diag_eps, v0, v1, v2);
            for (int k = 0; k < 3; k++) { e0[k] = v0[k]; e2[k] = v2[k]; }
        }
        return;
    }
    {
        const T iv = d_rsqrt(rr);
        v[0] *= iv; v[1] *= iv; v[2] *= iv;
    }
    // orthonormal complement {U, W} of v: U = v x e_k for the axis k of the
    // smallest |v_k|
    T U[3], W[3];
    {
        const T ax = d_abs(v[0]), ay = d_abs(v[1]), az = d_abs(v[2]);
        if (ax <= ay && ax <= az) { U[0] = T(0); U[1] = v[2]; U[2] = -v[1]; }
        else if (ay <= az) { U[0] = -v[2]; U[1] = T(0); U[2] = v[0]; }
        else { U[0] = v[1]; U[1] = -v[0]; U[2] = T(0); }
    }
    const T iu = d_rsqrt(U[0] * U[0] + U[1] * U[1] + U[2] * U[2]);
    U[0] *= iu; U[1] *= iu; U[2] *= iu;
    cross3(v, U, W);
    const T AU0 = a * U[0] + b * U[1] + d * U[2], AU1 = b * U[0] + c * U[1] + e * U[2],
            AU2 = d * U[0] + e * U[1] + f * U[2];
    const T AW0 = a * W[0] + b * W[1] + d * W[2], AW1 = b * W[0] + c * W[1] + e * W[2],
            AW2 = d * W[0] + e * W[1] + f * W[2];
    const T m00 = U[0] * AU0 + U[1] * AU1 + U[2] * AU2;
    const T m01 = U[0] * AW0 + U[1] * AW1 + U[2] * AW2;
    const T m11 = W[0] * AW0 + W[1] * AW1 + W[2] * AW2;
    // Rayleigh quotient v.Av through the trace: {v, U, W} is orthonormal, so
    // tr(A) = v.Av + U.AU + W.AW exactly
    const T ls = (a + c + f) - (m00 + m11);
    const T mean = T(0.5) * (m00 + m11), hd = T(0.5) * (m00 - m11);
    const T h = d_sqrt_fast(hd * hd + m01 * m01);
    const T lo = mean - h, hi = mean + h;
    T x0, x1, x2;
    if (top) { x0 = lo; x1 = hi; x2 = ls; } else { x0 = ls; x1 = lo; x2 = hi; }
    // keep the ascending order when rounding crosses the separated value
    sort3(x0, x1, x2);
    lam[0] = x0; lam[1] = x1; lam[2] = x2;
    if (e0) {
        // eigenvector of the 2x2 problem for eigenvalue `lo` (unit, in the U/W plane)
        T cx = m01, cy = lo - m00;          // row 0 null vector of (M - lo I)
        T dx = lo - m11, dy = m01;          // row 1 null vector
        const T n0 = cx * cx + cy * cy, n1 = dx * dx + dy * dy;
        T px = n0 >= n1 ? cx : dx, py = n0 >= n1 ? cy : dy;
        T pn = px * px + py * py;
        if (!(pn > T(0))) { px = T(1); py = T(0); pn = T(1); }
        const T ip = d_rsqrt(pn);
        px *= ip; py *= ip;
        T vlo[3], vhi[3];
        for (int k = 0; k < 3; k++) {
            vlo[k] = px * U[k] + py * W[k];
            vhi[k] = -py * U[k] + px * W[k];
        }
        const T* s0 = top ? vlo : v;
        const T* s2 = top ? v : vhi;
        for (int k = 0; k < 3; k++) { e0[k] = s0[k]; e2[k] = s2[k]; }
        canonical_sign(e0);
        canonical_sign(e2);
    }
}

// ---------------------------------------------------------------- 2x2
// Eigenvalues of [a b; b c] (eval2_symmetric, eigen.h:118-133) in the stable
// form its documentation states, mean -+ hypot((a - c)/2, b): the quadratic-
// root form the reference evaluates (quad_root, eigen.h:89-98) loses sqrt(eps)
// when the eigenvalues nearly coincide (tr^2 - 4 det cancels). Tiny negatives
// are clamped as the reference does (eigen.h:127-132).
TH_HD void eig2_vals(float a, float b, float c, float& e0, float& e1) {
    const float mean = 0.5f * (a + c);
    const float h = hypotf(0.5f * (a - c), b);
    float r0 = mean - h, r1 = mean + h;
    const float ar1 = fabsf(r1);
    const float tol = 1e-6f * (ar1 > 1.0f ? ar1 : 1.0f);
    if (r0 < 0.0f && r0 > -tol) r0 = 0.0f;
    if (r1 < 0.0f && r1 > -tol) r1 = 0.0f;
    e0 = r0; e1 = r1;
}

// Unit eigenvectors (v for l0, u = perp(v) for l1) with the reference's sign
// convention (evec2polar_symmetric, eigen.h:149-212) -- cartesian, no angles.
TH_HD void eig2_vecs(float a, float b, float c, float l0, float& vx, float& vy,
                                          float& ux, float& uy) {
    const float vx0 = -b, vy0 = a - l0, vx1 = c - l0, vy1 = -b;
    const float n0 = vx0 * vx0 + vy0 * vy0, n1 = vx1 * vx1 + vy1 * vy1;
    float x = (n0 >= n1) ? vx0 : vx1, y = (n0 >= n1) ? vy0 : vy1;
    const float n = sqrtf(x * x + y * y);
    if (n < 1e-6f) { x = 1.0f; y = 0.0f; }
    else { const float inv = 1.0f / n; x *= inv; y *= inv; }
    if (fabsf(x) >= fabsf(y)) { if (x < 0.0f) { x = -x; y = -y; } }
    else if (y < 0.0f) { x = -x; y = -y; }
    vx = x; vy = y;
    float px = -y, py = x;
    if (fabsf(px) >= fabsf(py)) { if (px < 0.0f) { px = -px; py = -py; } }
    else if (py < 0.0f) { px = -px; py = -py; }
    ux = px; uy = py;
}

}  // namespace th
