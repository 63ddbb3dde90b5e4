// eig_device.hpp -- device eigen-solvers for symmetric 2x2 / 3x3 tensors.
//
// Same mathematics as TIRA's eigen.h, written for CDNA4:
//  * eigenvalues of 3x3: Smith's trigonometric closed form (eigen.h:391-446),
//    with cos(phi + 2pi/3) taken from a single sincos(phi) by angle addition;
//  * eigenvectors of 3x3: cofactor row of the best-separated eigenvalue, the
//    middle one from the 2x2 problem on its orthogonal complement, the third
//    by a cross product (eigen.h:217-370, 452-512), canonical sign
//    (largest |component| positive, eigen.h:525-537);
//  * 2x2: stable quadratic and the larger-row null vector (eigen.h:89-212).
// Every routine is a template on the arithmetic type: float carries the bulk
// of the work, double re-solves near-degenerate tensors (relative eigengap
// below kRefineGap) so that the cubic-root / trigonometric step and the
// cross-product step keep their accuracy where FP32 loses it.
#pragma once
#include <hip/hip_runtime.h>

namespace th {

// Absolute thresholds of the reference (TIRA_EIGEN_EPSILON, eigen.h:13).
template <typename T> __device__ __forceinline__ T eig_eps() { return T(1e-6); }

constexpr float kRefineGap = 1e-4f;  // relative eigengap that sends a tensor to FP64

__device__ __forceinline__ float d_acos(float x) { return acosf(x); }
__device__ __forceinline__ double d_acos(double x) { return acos(x); }
__device__ __forceinline__ void d_sincos(float x, float* s, float* c) { sincosf(x, s, c); }
__device__ __forceinline__ void d_sincos(double x, double* s, double* c) { sincos(x, s, c); }
__device__ __forceinline__ float d_sqrt(float x) { return sqrtf(x); }
__device__ __forceinline__ double d_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ float d_atan2(float y, float x) { return atan2f(y, x); }
__device__ __forceinline__ double d_atan2(double y, double x) { return atan2(y, x); }
template <typename T> __device__ __forceinline__ T d_abs(T x) { return x < T(0) ? -x : x; }

template <typename T>
__device__ __forceinline__ void sort3(T& x, T& y, T& z) {
    T t;
    if (x > y) { t = x; x = y; y = t; }
    if (y > z) { t = y; y = z; z = t; }
    if (x > y) { t = x; x = y; y = t; }
}

// Eigenvalues (ascending) of | a b d ; b c e ; d e f |. `diag_eps` is the
// off-diagonal energy below which the diagonal is returned sorted (the
// reference uses an absolute 1e-6; scale-invariant callers pass a value
// relative to the tensor norm).
template <typename T>
__device__ __forceinline__ void eig3_vals(T a, T b, T c, T d, T e, T f, T diag_eps, T& l0, T& l1, T& l2) {
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
__device__ __forceinline__ bool eig3_near_degenerate(T l0, T l1, T l2) {
    const T scale = d_abs(l0) > d_abs(l2) ? d_abs(l0) : d_abs(l2);
    const T g = (l1 - l0) < (l2 - l1) ? (l1 - l0) : (l2 - l1);
    return scale > T(0) && g < T(kRefineGap) * scale;
}

template <typename T>
__device__ __forceinline__ void cross3(const T* a, const T* b, T* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

template <typename T>
__device__ __forceinline__ void normalize3_eps(T* v) {  // normalize3, eigen.h:56-65
    const T n = d_sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (n > eig_eps<T>()) {
        const T inv = T(1) / n;
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    }
}

// Null vector of (A - lambda I) from its largest cofactor row (eigen.h:247-288).
template <typename T>
__device__ __forceinline__ void evec_cofactor(T a, T b, T c, T d, T e, T f, T lam, T* v) {
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
__device__ __forceinline__ void evec_complement(T a, T b, T c, T d, T e, T f, T lam1, const T* ev, T* out) {
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
__device__ __forceinline__ void canonical_sign(T* v) {  // eigen.h:525-537
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
__device__ __forceinline__ void eig3_vecs(T a, T b, T c, T d, T e, T f, T l0, T l1, T l2, T diag_eps,
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

// ---------------------------------------------------------------- 2x2
// eval2_symmetric (eigen.h:118-133) with quad_root (89-98).
__device__ __forceinline__ void eig2_vals(float a, float b, float c, float& e0, float& e1) {
    const float tr = a + c, det = a * c - b * b;
    const float bq = -tr;
    float disc = bq * bq - 4.0f * det;
    if (disc < 0.0f) disc = 0.0f;
    const float sg = (bq < 0.0f) ? -1.0f : 1.0f;
    const float qq = -0.5f * (bq + sg * sqrtf(disc));
    float r1 = qq, r0 = (qq == 0.0f) ? 0.0f : det / qq;
    if (r0 > r1) { float t = r0; r0 = r1; r1 = t; }
    const float ar1 = fabsf(r1);
    const float tol = 1e-6f * (ar1 > 1.0f ? ar1 : 1.0f);
    if (r0 < 0.0f && r0 > -tol) r0 = 0.0f;
    if (r1 < 0.0f && r1 > -tol) r1 = 0.0f;
    e0 = r0; e1 = r1;
}

// Unit eigenvectors (v for l0, u = perp(v) for l1) with the reference's sign
// convention (evec2polar_symmetric, eigen.h:149-212) -- cartesian, no angles.
__device__ __forceinline__ void eig2_vecs(float a, float b, float c, float l0, float& vx, float& vy,
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
