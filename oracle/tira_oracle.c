/*
 * tira_oracle.c -- CPU oracle for the tensor-voting hot path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path, never the thing
 * measured or shipped. Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg load it.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the TIRA source tree). FP32 functions keep the reference's operation order,
 * its absolute thresholds (TIRA_EIGEN_EPSILON = 1e-6, TV_EPSILON = 1e-12) and
 * its quirks; FP64 functions are independent routes used to state error.
 *
 * Pinning (how each function is checked, see tests/test_oracle_*.py):
 *   eigen 2x2 / 3x3 FP32     bit-exact against the reference eigen.h compiled
 *                            from source (oracle/_ref, same flags) and the
 *                            golden vectors it wrote (tests/golden/); FP64
 *                            Jacobi and numpy.linalg.eigh; trace/det/Frobenius
 *                            invariants.
 *   Gaussian kernel + conv   bit-exact against filter.h compiled from source
 *                            and its golden vectors; scipy.ndimage.correlate1d.
 *   2D stick / plate vote    single-voter brute force (table vs direct kernel),
 *                            p=1 closed-form 2D plate, symmetry d -> -d.
 *   3D stick vote            single-voter brute force against an independent
 *                            FP64 pair kernel, rotation / axis-permutation
 *                            equivariance, symmetry.
 *   3D integrated plate      three independent formulations agree: FP64
 *                            literal N-sample sum, the reference's analytic
 *                            hypergeometric form restated, and the exact
 *                            moment closed form (N >= p+2 makes the Riemann
 *                            sum exact).
 *   3D analytic plate        its Boost.Math terms (beta, 2F1 with -p) are
 *                            restated from their terminating series / Gamma
 *                            closed forms; pinned by the numerical sum and,
 *                            term by term, against scipy.special's hyp2f1
 *                            and beta (independent implementations of the
 *                            same functions; Boost itself is not in the
 *                            image).
 */
#include "tira_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define O_PI 3.14159265358979323846
#define O_EIG_EPS_F 1e-6f
#define O_TV_EPS 1e-12

/* ======================================================================
 * Eigen 2x2 (eigen.h:89-212)
 * ====================================================================== */

/* quad_root, eigen.h:89-98 */
static void o_quad_root(float b, float c, float* r0, float* r1) {
    float disc = (b * b) - 4.0f * c;
    if (disc < 0.0f) disc = 0.0f;
    float sg = (b < 0.0f) ? -1.0f : 1.0f; /* sgn_plus, eigen.h:29-32 */
    /* unqualified sqrt(float) resolves to ::sqrt(double) under g++: the bracket is evaluated in double */
    float q = (float)(-0.5 * ((double)b + (double)sg * sqrt((double)disc)));
    *r1 = q;
    if (q == 0.0f) *r0 = 0.0f;
    else *r0 = c / q;
    if (*r0 > *r1) { float t = *r0; *r0 = *r1; *r1 = t; }
}

/* eval2_symmetric, eigen.h:118-133 */
void oracle_eval2_f32(float a, float b, float c, float* e0, float* e1) {
    float tr = a + c;
    float det = a * c - b * b;
    o_quad_root(-tr, det, e0, e1);
    const float ae1 = (*e1 < 0.0f) ? -*e1 : *e1;
    const float scale = (ae1 > 1.0f) ? ae1 : 1.0f;
    const float tol = O_EIG_EPS_F * scale;
    if (*e0 < 0.0f && *e0 > -tol) *e0 = 0.0f;
    if (*e1 < 0.0f && *e1 > -tol) *e1 = 0.0f;
}

/* evec2polar_symmetric, eigen.h:149-212 */
void oracle_evec2polar_f32(float a, float b, float c, const float* lambdas, float* th0, float* th1) {
    const float pi = (float)O_PI;
    float l0 = lambdas[0];
    float vx0 = -b, vy0 = a - l0;
    float vx1 = c - l0, vy1 = -b;
    const float n0 = vx0 * vx0 + vy0 * vy0;
    const float n1 = vx1 * vx1 + vy1 * vy1;
    float vx = (n0 >= n1) ? vx0 : vx1;
    float vy = (n0 >= n1) ? vy0 : vy1;
    float n = sqrtf(vx * vx + vy * vy);
    if (n < O_EIG_EPS_F) {
        vx = 1.0f; vy = 0.0f;
    } else {
        const float inv = 1.0f / n;
        vx *= inv; vy *= inv;
    }
    if (fabsf(vx) >= fabsf(vy)) {
        if (vx < 0.0f) { vx = -vx; vy = -vy; }
    } else {
        if (vy < 0.0f) { vx = -vx; vy = -vy; }
    }
    float t0 = (float)atan2((double)vy, (double)vx); /* ::atan2(double, double) */
    float ux = -vy, uy = vx;
    if (fabsf(ux) >= fabsf(uy)) {
        if (ux < 0.0f) { ux = -ux; uy = -uy; }
    } else {
        if (uy < 0.0f) { ux = -ux; uy = -uy; }
    }
    float t1 = (float)atan2((double)uy, (double)ux);
    if (t0 <= -pi) t0 += 2.0f * pi;
    if (t1 <= -pi) t1 += 2.0f * pi;
    if (t0 > pi) t0 -= 2.0f * pi;
    if (t1 > pi) t1 -= 2.0f * pi;
    *th0 = t0; *th1 = t1;
}

/* cpu::evals2_symmetric, eigen.h:572-586 */
void oracle_evals2_symmetric(const float* m, size_t n, float* ev) {
    for (size_t i = 0; i < n; i++)
        oracle_eval2_f32(m[i * 4 + 0], m[i * 4 + 1], m[i * 4 + 3], &ev[i * 2 + 0], &ev[i * 2 + 1]);
}

/* cpu::evecs2polar_symmetric, eigen.h:597-611 */
void oracle_evecs2polar_symmetric(const float* m, const float* ev, size_t n, float* th) {
    for (size_t i = 0; i < n; i++)
        oracle_evec2polar_f32(m[i * 4 + 0], m[i * 4 + 1], m[i * 4 + 3], &ev[i * 2], &th[i * 2 + 0], &th[i * 2 + 1]);
}

/* ======================================================================
 * Eigen 3x3 (eigen.h:40-557)
 * ====================================================================== */

static float o_dot3(const float* a, const float* b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* normalize3, eigen.h:56-65 (absolute threshold on the norm) */
static void o_normalize3(float* v) {
    float n = sqrtf(o_dot3(v, v));
    if (n > O_EIG_EPS_F) {
        float inv = 1.0f / n;
        v[0] *= inv; v[1] *= inv; v[2] *= inv;
    }
}

/* cross3, eigen.h:68-75 */
static void o_cross3(const float* a, const float* b, float* r) {
    r[0] = a[1] * b[2] - a[2] * b[1];
    r[1] = a[2] * b[0] - a[0] * b[2];
    r[2] = a[0] * b[1] - a[1] * b[0];
}

/* determinant3, eigen.h:47-53 */
static float o_det3(float a, float b, float c, float d, float e, float f) {
    return a * (c * f - e * e) - b * (b * f - e * d) + d * (b * e - c * d);
}

/* ComputeOrthogonalComplement, eigen.h:217-245 */
static void o_orth_complement(const float* ev, float* U, float* V) {
    const float ax = fabsf(ev[0]), ay = fabsf(ev[1]), az = fabsf(ev[2]);
    float t[3];
    if (ay - ax > O_EIG_EPS_F) {
        if (az - ax > O_EIG_EPS_F) { t[0] = 1; t[1] = 0; t[2] = 0; }
        else { t[0] = 0; t[1] = 0; t[2] = 1; }
    } else {
        if (az - ay > O_EIG_EPS_F) { t[0] = 0; t[1] = 1; t[2] = 0; }
        else { t[0] = 0; t[1] = 0; t[2] = 1; }
    }
    o_cross3(ev, t, U);
    o_normalize3(U);
    o_cross3(ev, U, V);
}

/* evec3_symmetric_0, eigen.h:247-288 */
static void o_evec3_0(float a, float b, float c, float d, float e, float f, float ev, float* out) {
    float m00 = a - ev, m01 = b, m02 = d;
    float m10 = b, m11 = c - ev, m12 = e;
    float m20 = d, m21 = e, m22 = f - ev;
    float c00 = m11 * m22 - m12 * m21;
    float c01 = m12 * m20 - m10 * m22;
    float c02 = m10 * m21 - m11 * m20;
    float c10 = m02 * m21 - m01 * m22;
    float c11 = m00 * m22 - m02 * m20;
    float c12 = m01 * m20 - m00 * m21;
    float c20 = m01 * m12 - m02 * m11;
    float c21 = m02 * m10 - m00 * m12;
    float c22 = m00 * m11 - m01 * m10;
    float r0 = c00 * c00 + c01 * c01 + c02 * c02;
    float r1 = c10 * c10 + c11 * c11 + c12 * c12;
    float r2 = c20 * c20 + c21 * c21 + c22 * c22;
    if (r0 >= r1 && r0 >= r2) { out[0] = c00; out[1] = c01; out[2] = c02; }
    else if (r1 >= r0 && r1 >= r2) { out[0] = c10; out[1] = c11; out[2] = c12; }
    else { out[0] = c20; out[1] = c21; out[2] = c22; }
    o_normalize3(out);
}

/* evec3_symmetric_1, eigen.h:290-370 */
static void o_evec3_1(float a, float b, float c, float d, float e, float f, float ev1,
                      const float* e0, float* out) {
    float U[3], V[3];
    o_orth_complement(e0, U, V);
    const float AU[3] = {a * U[0] + b * U[1] + d * U[2], b * U[0] + c * U[1] + e * U[2],
                         d * U[0] + e * U[1] + f * U[2]};
    const float AV[3] = {a * V[0] + b * V[1] + d * V[2], b * V[0] + c * V[1] + e * V[2],
                         d * V[0] + e * V[1] + f * V[2]};
    float m00 = o_dot3(U, AU) - ev1;
    float m01 = o_dot3(U, AV);
    float m11 = o_dot3(V, AV) - ev1;
    float a00 = fabsf(m00), a01 = fabsf(m01), a11 = fabsf(m11);
    if (a00 >= a11) {
        if (a00 > O_EIG_EPS_F || a01 > O_EIG_EPS_F) {
            if (a00 >= a01) {
                m01 /= m00;
                float il = (float)(1.0 / sqrt((double)(1.0f + m01 * m01)));
                m00 = il; m01 *= il;
            } else {
                m00 /= m01;
                float il = (float)(1.0 / sqrt((double)(1.0f + m00 * m00)));
                m01 = il; m00 *= il;
            }
            for (int k = 0; k < 3; k++) out[k] = m01 * U[k] - m00 * V[k];
        } else {
            out[0] = U[0]; out[1] = U[1]; out[2] = U[2];
        }
    } else {
        if (a11 > O_EIG_EPS_F || a01 > O_EIG_EPS_F) {
            if (a11 >= a01) {
                m01 /= m11;
                float il = (float)(1.0 / sqrt((double)(1.0f + m01 * m01)));
                m11 = il; m01 *= il;
            } else {
                m11 /= m01;
                float il = (float)(1.0 / sqrt((double)(1.0f + m11 * m11)));
                m01 = il; m11 *= il;
            }
            for (int k = 0; k < 3; k++) out[k] = m11 * U[k] - m01 * V[k];
        } else {
            out[0] = U[0]; out[1] = U[1]; out[2] = U[2];
        }
    }
}

/* eval3_symmetric (Smith 1961 trigonometric), eigen.h:391-446 */
void oracle_eval3_f32(float a, float b, float c, float d, float e, float f,
                      float* l0, float* l1, float* l2) {
    const float p1 = b * b + d * d + e * e;
    if (p1 < O_EIG_EPS_F) {
        float e0 = a, e1 = c, e2 = f, t;
        if (e0 > e1) { t = e0; e0 = e1; e1 = t; }
        if (e1 > e2) { t = e1; e1 = e2; e2 = t; }
        if (e0 > e1) { t = e0; e0 = e1; e1 = t; }
        *l0 = e0; *l1 = e1; *l2 = e2;
        return;
    }
    const float tr = a + c + f;
    const float q = tr / 3.0f;
    const float aq = a - q, cq = c - q, fq = f - q;
    const float p2 = aq * aq + cq * cq + fq * fq + 2.0f * p1;
    float p = sqrtf(p2 / 6.0f);
    const float pinv = 1.0f / p;
    const float Ba = pinv * aq, Bb = pinv * b, Bc = pinv * cq;
    const float Bd = pinv * d, Be = pinv * e, Bf = pinv * fq;
    const float detB = Ba * (Bc * Bf - Be * Be) - Bb * (Bb * Bf - Bd * Be) + Bd * (Bb * Be - Bd * Bc);
    float r = detB / 2.0f;
    float phi;
    if (r < -1.0f) phi = (float)O_PI / 3.0f;
    else if (r >= 1.0f) phi = 0.0f;
    else phi = (float)(acos((double)r) / 3.0); /* ::acos(double) / T(3) in double */
    *l2 = (float)((double)q + (double)(2.0f * p) * cos((double)phi));
    *l0 = (float)((double)q + (double)(2.0f * p) * cos((double)(phi + (2.0f * (float)O_PI / 3.0f))));
    *l1 = tr - *l0 - *l2;
}

/* evec3_symmetric, eigen.h:452-512 */
void oracle_evec3_f32(float a, float b, float c, float d, float e, float f,
                      const float* ev, float* v0, float* v1, float* v2) {
    const float ud = b * b + d * d + e * e;
    if (ud < O_EIG_EPS_F) {
        float val[3] = {a, c, f};
        int idx[3] = {0, 1, 2};
        /* bubble sort on (eval, index), eigen.h:474-476 */
#define O_SWP(i, j) if (val[i] > val[j]) { float tv = val[i]; val[i] = val[j]; val[j] = tv; int ti = idx[i]; idx[i] = idx[j]; idx[j] = ti; }
        O_SWP(0, 1) O_SWP(1, 2) O_SWP(0, 1)
#undef O_SWP
        float* outs[3] = {v0, v1, v2};
        for (int k = 0; k < 3; k++)
            for (int j = 0; j < 3; j++) outs[k][j] = (idx[k] == j) ? 1.0f : 0.0f;
        return;
    }
    const float q = (a + c + f) / 3.0f;
    const float p2 = (a - q) * (a - q) + (c - q) * (c - q) + (f - q) * (f - q) + 2.0f * ud;
    const float p = sqrtf(p2 / 6.0f);
    float hd;
    if (p < O_EIG_EPS_F) hd = 0.0f;
    else hd = o_det3(a - q, b, c - q, d, e, f - q) / (2.0f * (p * p * p));
    hd = (hd < -1.0f) ? -1.0f : ((hd > 1.0f) ? 1.0f : hd);
    if (hd >= 0.0f) {
        o_evec3_0(a, b, c, d, e, f, ev[2], v2);
        o_evec3_1(a, b, c, d, e, f, ev[1], v2, v1);
        o_cross3(v1, v2, v0);
    } else {
        o_evec3_0(a, b, c, d, e, f, ev[0], v0);
        o_evec3_1(a, b, c, d, e, f, ev[1], v0, v1);
        o_cross3(v0, v1, v2);
        o_normalize3(v2);
    }
}

/* canonical sign lambda of evec3spherical_symmetric, eigen.h:525-537 */
static void o_canonical_sign(float* v) {
    float ax = fabsf(v[0]), ay = fabsf(v[1]), az = fabsf(v[2]);
    int k = 0;
    float am = ax;
    if (ay > am) { k = 1; am = ay; }
    if (az > am) { k = 2; am = az; }
    if (v[k] < 0.0f) { v[0] = -v[0]; v[1] = -v[1]; v[2] = -v[2]; }
}

/* evec3spherical_symmetric, eigen.h:514-557 */
static void o_evec3spherical(float a, float b, float c, float d, float e, float f,
                             const float* ev, float* out6) {
    float c0[3], c1[3], c2[3];
    oracle_evec3_f32(a, b, c, d, e, f, ev, c0, c1, c2);
    o_canonical_sign(c0);
    o_canonical_sign(c1);
    o_canonical_sign(c2);
    float* cs[3] = {c0, c1, c2};
    for (int k = 0; k < 3; k++) {
        float z = cs[k][2];
        z = (z < -1.0f) ? -1.0f : ((z > 1.0f) ? 1.0f : z);
        out6[2 * k + 0] = (float)atan2((double)cs[k][1], (double)cs[k][0]);
        out6[2 * k + 1] = (float)acos((double)z);
    }
}

/* cpu::evals3_symmetric, eigen.h:620-640 */
void oracle_evals3_symmetric(const float* m, size_t n, float* ev) {
    for (size_t i = 0; i < n; i++) {
        const float* x = m + i * 9;
        oracle_eval3_f32(x[0], x[1], x[4], x[2], x[5], x[8], &ev[i * 3], &ev[i * 3 + 1], &ev[i * 3 + 2]);
    }
}

/* cpu::evecs3spherical_symmetric, eigen.h:651-680 */
void oracle_evecs3spherical_symmetric(const float* m, const float* lam, size_t n, float* out) {
    for (size_t i = 0; i < n; i++) {
        const float* x = m + i * 9;
        o_evec3spherical(x[0], x[1], x[4], x[2], x[5], x[8], lam + i * 3, out + i * 6);
    }
}

/* spherical_to_cartesian, tensor.h:26-30 */
void oracle_spherical_to_cartesian(float theta, float phi, float* xyz) {
    const float st = sinf(theta), ct = cosf(theta);
    const float sp = sinf(phi), cp = cosf(phi);
    xyz[0] = ct * sp; xyz[1] = st * sp; xyz[2] = cp;
}

/* ---------------- FP64 Jacobi (independent of the reference) ---------------- */
void oracle_eig3_jacobi_f64(const double* a6, double* evals, double* evecs) {
    double A[3][3] = {{a6[0], a6[1], a6[2]}, {a6[1], a6[3], a6[4]}, {a6[2], a6[4], a6[5]}};
    double V[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 64; sweep++) {
        double off = A[0][1] * A[0][1] + A[0][2] * A[0][2] + A[1][2] * A[1][2];
        double diag = A[0][0] * A[0][0] + A[1][1] * A[1][1] + A[2][2] * A[2][2];
        if (off == 0.0 || off <= 1e-40 * diag) break;
        for (int p = 0; p < 2; p++) {
            for (int q = p + 1; q < 3; q++) {
                double apq = A[p][q];
                if (apq == 0.0) continue;
                double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
                double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
                for (int k = 0; k < 3; k++) { /* A <- A J */
                    double akp = A[k][p], akq = A[k][q];
                    A[k][p] = c * akp - s * akq;
                    A[k][q] = s * akp + c * akq;
                }
                for (int k = 0; k < 3; k++) { /* A <- J^T A */
                    double apk = A[p][k], aqk = A[q][k];
                    A[p][k] = c * apk - s * aqk;
                    A[q][k] = s * apk + c * aqk;
                }
                for (int k = 0; k < 3; k++) {
                    double vkp = V[k][p], vkq = V[k][q];
                    V[k][p] = c * vkp - s * vkq;
                    V[k][q] = s * vkp + c * vkq;
                }
            }
        }
    }
    double w[3] = {A[0][0], A[1][1], A[2][2]};
    int o[3] = {0, 1, 2};
    for (int i = 0; i < 3; i++)
        for (int j = i + 1; j < 3; j++)
            if (w[o[j]] < w[o[i]]) { int t = o[i]; o[i] = o[j]; o[j] = t; }
    for (int k = 0; k < 3; k++) {
        evals[k] = w[o[k]];
        double v[3] = {V[0][o[k]], V[1][o[k]], V[2][o[k]]};
        int im = 0;
        for (int j = 1; j < 3; j++) if (fabs(v[j]) > fabs(v[im])) im = j;
        double sg = v[im] < 0 ? -1.0 : 1.0;
        for (int r = 0; r < 3; r++) evecs[r * 3 + k] = sg * v[r];
    }
}

void oracle_eig3_f64(const float* m, size_t n, double* ev, double* e0, double* e2) {
    for (size_t i = 0; i < n; i++) {
        const float* x = m + i * 9;
        double a6[6] = {x[0], x[1], x[2], x[4], x[5], x[8]};
        double w[3], V[9];
        oracle_eig3_jacobi_f64(a6, w, V);
        if (ev) { ev[i * 3] = w[0]; ev[i * 3 + 1] = w[1]; ev[i * 3 + 2] = w[2]; }
        for (int r = 0; r < 3; r++) {
            if (e0) e0[i * 3 + r] = V[r * 3 + 0];
            if (e2) e2[i * 3 + r] = V[r * 3 + 2];
        }
    }
}

void oracle_eig3_packed6_f64(const float* t6, size_t n, double* ev) {
    for (size_t i = 0; i < n; i++) {
        double a6[6];
        for (int k = 0; k < 6; k++) a6[k] = t6[i * 6 + k];
        double w[3], V[9];
        oracle_eig3_jacobi_f64(a6, w, V);
        ev[i * 3] = w[0]; ev[i * 3 + 1] = w[1]; ev[i * 3 + 2] = w[2];
    }
}

/* ======================================================================
 * 2D tensor voting (tensorvote.h:34-60, 776-1142, 1223-1399)
 * ====================================================================== */

/* atv_attenuation, tensorvote.h:34-40 (exp(-d^2/sigma^2), delta at sigma=0) */
static float o_atten_f(float dist, float sigma) {
    if (sigma == 0.0f) return (dist > (float)O_TV_EPS) ? 0.0f : 1.0f;
    return (float)exp((double)(-(dist * dist) / (sigma * sigma)));
}

/* tk_stick_eta_2d, tensorvote.h:803-824 */
static float o_tk_stick_eta_2d(float dist, float qTd, float s1, float s2, unsigned power) {
    if (power < 1) power = 1;
    const float c1 = o_atten_f(dist, s1), c2 = o_atten_f(dist, s2);
    const float q2 = qTd * qTd;
    const float t1 = 1.0f - q2, t2 = q2;
    if (power == 1) return c1 * t1 + c2 * t2;
    float e1 = t1, e2 = t2;
    for (unsigned i = 1; i < power; ++i) { e1 *= t1; e2 *= t2; }
    return c1 * e1 + c2 * e2;
}

/* tk_stick_norm_2d, tensorvote.h:842-850 */
float oracle_tk_stick_norm_2d(float s1, float s2, unsigned power) {
    if (power < 1) power = 1;
    float ratio = 1.0f;
    for (unsigned k = 1; k <= power; ++k) ratio *= (float)(2 * k) / (float)(2 * k - 1);
    return ratio / ((float)O_PI * (s1 * s1 + s2 * s2));
}

/* tk_stickvote_2d, tensorvote.h:871-914 */
void oracle_tk_stickvote_2d(float rx, float ry, float s1, float s2, unsigned power, float theta,
                            float* Va, float* Vb, float* Vc) {
    const float qx = (float)cos((double)theta), qy = (float)sin((double)theta); /* ::cos(double) */
    const float dist = sqrtf(rx * rx + ry * ry);
    float dx, dy;
    if (dist == 0.0f) { dx = 0.0f; dy = 0.0f; }
    else { dx = rx / dist; dy = ry / dist; }
    const float qTd = qx * dx + qy * dy;
    const float D = o_tk_stick_eta_2d(dist, qTd, s1, s2, power);
    const float eta_s = oracle_tk_stick_norm_2d(s1, s2, power);
    const float factor = D * eta_s;
    const float Ra = 1 - 2 * dx * dx, Rb = -2 * dx * dy, Rc = 1 - 2 * dy * dy; /* atv_rotation 55-60 */
    const float Rqx = Ra * qx + Rb * qy;
    const float Rqy = Rb * qx + Rc * qy;
    *Va = Rqx * Rqx * factor;
    *Vb = Rqx * Rqy * factor;
    *Vc = Rqy * Rqy * factor;
}

/* tk_plate_numerical (2D), tensorvote.h:1047-1071 */
void oracle_tk_plate_numerical_2d(float rx, float ry, float s1, float s2, unsigned power,
                                  float* Va, float* Vb, float* Vc, unsigned n) {
    float a_ = 0, b_ = 0, c_ = 0;
    *Va = *Vb = *Vc = 0.0f;
    if (n == 0) return;
    const float dth = (float)(O_PI / (double)(float)n);
    for (unsigned i = 0; i < n; i++) {
        const float th = dth * (float)i;
        float a, b, c;
        oracle_tk_stickvote_2d(rx, ry, s1, s2, power, th, &a, &b, &c);
        a_ += a; b_ += b; c_ += c;
    }
    const float wb = (float)O_PI / (float)n;
    *Va = a_ * wb; *Vb = b_ * wb; *Vc = c_ * wb;
}

/* cpu::build_neighbors2d, tensorvote.h:1223-1280 (dv outer, du inner) */
int oracle_build_neighbors2d(int w, float s1, float s2, unsigned power, unsigned samples,
                             int stick, int plate, oracle_neighbor2d* out) {
    if (w <= 0) return 0;
    const int hw = w / 2;
    const float is1 = (s1 > 0.0f) ? 1.0f / (s1 * s1) : 0.0f;
    const float is2 = (s2 > 0.0f) ? 1.0f / (s2 * s2) : 0.0f;
    int k = 0;
    for (int dv = -hw; dv <= hw; dv++) {
        for (int du = -hw; du <= hw; du++) {
            oracle_neighbor2d nb;
            memset(&nb, 0, sizeof(nb));
            nb.du = du; nb.dv = dv;
            const float l2 = (float)(du * du + dv * dv);
            if (stick) {
                if (l2 == 0.0f) { nb.dx = 0.0f; nb.dy = 0.0f; }
                else {
                    const float il = 1.0f / sqrtf(l2);
                    nb.dx = (float)du * il;
                    nb.dy = (float)dv * il;
                }
                nb.c1 = (s1 == 0.0f) ? (l2 > O_TV_EPS ? 0.0f : 1.0f) : expf(-l2 * is1);
                nb.c2 = (s2 == 0.0f) ? (l2 > O_TV_EPS ? 0.0f : 1.0f) : expf(-l2 * is2);
                nb.Ra = 1.0f - 2.0f * nb.dx * nb.dx;
                nb.Rb = 0.0f - 2.0f * nb.dx * nb.dy;
                nb.Rc = 1.0f - 2.0f * nb.dy * nb.dy;
            }
            if (plate && samples > 0) {
                float Va, Vb, Vc;
                oracle_tk_plate_numerical_2d((float)du, (float)dv, s1, s2, power, &Va, &Vb, &Vc, samples);
                nb.Va = Va; nb.Vb = Vb; nb.Vc = Vc;
            }
            out[k++] = nb;
        }
    }
    return k;
}

/* tk_stick_window fast path (938-1005) + tk_plate_window fast path
 * (1100-1115), driven by cpu::tensorvote_tk (1335-1378). */
void oracle_tensorvote2_tk(float* t_out, const float* lam, const float* th, float s1, float s2,
                           unsigned power, unsigned w, unsigned S0, unsigned S1, int stick, int plate,
                           unsigned samples, const oracle_neighbor2d* NB, int nbc, int nthreads) {
    (void)w;
    unsigned p = power < 1 ? 1 : power;
    const float eta_s = oracle_tk_stick_norm_2d(s1, s2, p);
    const int s0 = (int)S0, sz1 = (int)S1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int x0 = 0; x0 < s0; x0++) {
        for (int x1 = 0; x1 < sz1; x1++) {
            float oa = 0, ob = 0, oc = 0;
            if (stick) {
                float a = 0, b = 0, c = 0;
                for (int k = 0; k < nbc; ++k) {
                    const int r0 = x0 + NB[k].dv;
                    if ((unsigned)r0 >= (unsigned)s0) continue;
                    const int r1 = x1 + NB[k].du;
                    if ((unsigned)r1 >= (unsigned)sz1) continue;
                    const int idx = r0 * sz1 * 2 + r1 * 2;
                    const float theta = th[idx + 1];
                    const float qx = (float)cos((double)theta), qy = (float)sin((double)theta);
                    const float qTd = qx * NB[k].dx + qy * NB[k].dy;
                    const float q2 = qTd * qTd;
                    const float t1 = 1.0f - q2, t2 = q2;
                    float D;
                    if (p == 1) D = NB[k].c1 * t1 + NB[k].c2 * t2;
                    else {
                        float e1 = t1, e2 = t2;
                        for (unsigned i = 1; i < p; ++i) { e1 *= t1; e2 *= t2; }
                        D = NB[k].c1 * e1 + NB[k].c2 * e2;
                    }
                    const float Rqx = NB[k].Ra * qx + NB[k].Rb * qy;
                    const float Rqy = NB[k].Rb * qx + NB[k].Rc * qy;
                    const float factor = D * eta_s;
                    const float Va = Rqx * Rqx * factor, Vb = Rqx * Rqy * factor, Vc = Rqy * Rqy * factor;
                    const float l0 = lam[idx + 0], l1 = lam[idx + 1];
                    float scale = fabsf(l1) - fabsf(l0);
                    if (l1 < 0.0f) scale = scale * -1.0f;
                    a += scale * Va; b += scale * Vb; c += scale * Vc;
                }
                oa += a; ob += b; oc += c;
            }
            if (plate) {
                float a = 0, b = 0, c = 0;
                if (samples != 0) {
                    for (int k = 0; k < nbc; k++) {
                        const int r0 = x0 + NB[k].dv;
                        if ((unsigned)r0 >= (unsigned)s0) continue;
                        const int r1 = x1 + NB[k].du;
                        if ((unsigned)r1 >= (unsigned)sz1) continue;
                        const int idx = r0 * sz1 * 2 + r1 * 2;
                        const float l0 = lam[idx + 0];
                        if (l0 == 0.0f) continue;
                        a += l0 * NB[k].Va; b += l0 * NB[k].Vb; c += l0 * NB[k].Vc;
                    }
                }
                oa += a; ob += b; oc += c;
            }
            const size_t o = (size_t)x0 * S1 * 4 + (size_t)x1 * 4;
            t_out[o + 0] = oa; t_out[o + 1] = ob; t_out[o + 2] = ob; t_out[o + 3] = oc;
        }
    }
}

/* cpu::tensorvote (2D), tensorvote.h:1380-1399 */
void oracle_tensorvote2(float* t_out, const float* t_in, float s1, float s2, int power, unsigned w,
                        unsigned S0, unsigned S1, int stick, int plate, unsigned samples, int nthreads) {
    const size_t n = (size_t)S0 * S1;
    float* lam = (float*)malloc(n * 2 * sizeof(float));
    float* th = (float*)malloc(n * 2 * sizeof(float));
    oracle_evals2_symmetric(t_in, n, lam);
    oracle_evecs2polar_symmetric(t_in, lam, n, th);
    const int use_plate = plate && samples > 0;
    oracle_neighbor2d* NB = NULL;
    int nbc = 0;
    if (stick || use_plate) {
        NB = (oracle_neighbor2d*)malloc((size_t)w * w * sizeof(oracle_neighbor2d));
        nbc = oracle_build_neighbors2d((int)w, s1, s2, (unsigned)power, samples, stick, use_plate, NB);
    }
    oracle_tensorvote2_tk(t_out, lam, th, s1, s2, (unsigned)power, w, S0, S1, stick, plate, samples, NB, nbc,
                          nthreads);
    free(NB); free(lam); free(th);
}

/* FP64 2D route: eigenvector by FP64 closed form, direct kernel per pair
 * (no table), plate by the FP64 literal N-sample sum. */
static void o_stick2_f64(double rx, double ry, double s1, double s2, unsigned p, double qx, double qy,
                         double eta_s, double* V) {
    double dist = sqrt(rx * rx + ry * ry);
    double dx = 0, dy = 0;
    if (dist > 0) { dx = rx / dist; dy = ry / dist; }
    double c1 = (s1 == 0) ? (dist > O_TV_EPS ? 0 : 1) : exp(-(dist * dist) / (s1 * s1));
    double c2 = (s2 == 0) ? (dist > O_TV_EPS ? 0 : 1) : exp(-(dist * dist) / (s2 * s2));
    double t = qx * dx + qy * dy, t2 = t * t;
    double D = c1 * pow(1 - t2, (double)p) + c2 * pow(t2, (double)p);
    double rqx = qx - 2 * t * dx, rqy = qy - 2 * t * dy;
    V[0] = rqx * rqx * D * eta_s; V[1] = rqx * rqy * D * eta_s; V[2] = rqy * rqy * D * eta_s;
}

void oracle_tensorvote2_f64(double* t_out, const float* t_in, double s1, double s2, int power, unsigned w,
                            unsigned S0, unsigned S1, int stick, int plate, unsigned samples, int nthreads) {
    const size_t n = (size_t)S0 * S1;
    unsigned p = power < 1 ? 1 : (unsigned)power;
    double* lam = (double*)malloc(n * 2 * sizeof(double));
    double* q = (double*)malloc(n * 2 * sizeof(double));
    for (size_t i = 0; i < n; i++) {
        double a = t_in[i * 4], b = t_in[i * 4 + 1], c = t_in[i * 4 + 3];
        double m = 0.5 * (a + c), r = hypot(0.5 * (a - c), b);
        lam[i * 2] = m - r; lam[i * 2 + 1] = m + r;
        /* eigenvector of the larger eigenvalue: angle 0.5*atan2(2b, a-c) */
        double ang = 0.5 * atan2(2 * b, a - c);
        q[i * 2] = cos(ang); q[i * 2 + 1] = sin(ang);
    }
    double ratio = 1;
    for (unsigned k = 1; k <= p; ++k) ratio *= (double)(2 * k) / (double)(2 * k - 1);
    const double eta_s = ratio / (O_PI * (s1 * s1 + s2 * s2));
    const int hw = (int)w / 2;
    /* per-offset FP64 plate table */
    double* PV = (double*)calloc((size_t)w * w * 3, sizeof(double));
    if (plate && samples > 0) {
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++) {
                double acc[3] = {0, 0, 0};
                for (unsigned i = 0; i < samples; i++) {
                    double th = O_PI / samples * i, V[3];
                    o_stick2_f64(du, dv, s1, s2, p, cos(th), sin(th), eta_s, V);
                    acc[0] += V[0]; acc[1] += V[1]; acc[2] += V[2];
                }
                double* o = PV + ((size_t)(dv + hw) * w + (du + hw)) * 3;
                for (int k = 0; k < 3; k++) o[k] = acc[k] * O_PI / samples;
            }
    }
    const int s0 = (int)S0, sz1 = (int)S1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int x0 = 0; x0 < s0; x0++)
        for (int x1 = 0; x1 < sz1; x1++) {
            double a = 0, b = 0, c = 0;
            for (int dv = -hw; dv <= hw; dv++) {
                int r0 = x0 + dv;
                if (r0 < 0 || r0 >= s0) continue;
                for (int du = -hw; du <= hw; du++) {
                    int r1 = x1 + du;
                    if (r1 < 0 || r1 >= sz1) continue;
                    size_t vi = (size_t)r0 * S1 + r1;
                    double l0 = lam[vi * 2], l1 = lam[vi * 2 + 1];
                    if (stick) {
                        double scale = fabs(l1) - fabs(l0);
                        if (l1 < 0) scale = -scale;
                        double V[3];
                        o_stick2_f64(du, dv, s1, s2, p, q[vi * 2], q[vi * 2 + 1], eta_s, V);
                        a += scale * V[0]; b += scale * V[1]; c += scale * V[2];
                    }
                    if (plate && samples > 0) {
                        const double* o = PV + ((size_t)(dv + hw) * w + (du + hw)) * 3;
                        a += l0 * o[0]; b += l0 * o[1]; c += l0 * o[2];
                    }
                }
            }
            size_t o = ((size_t)x0 * S1 + x1) * 4;
            t_out[o] = a; t_out[o + 1] = b; t_out[o + 2] = b; t_out[o + 3] = c;
        }
    free(PV); free(lam); free(q);
}

/* ======================================================================
 * 3D tensor voting (tensorvote.h:1147-1213, 1411-1436, 1513-1859)
 * ====================================================================== */

/* cpu::build_neighbors3d, tensorvote.h:1411-1436 (dw outer, dv, du inner) */
int oracle_build_neighbors3d(int w, float s1, float s2, oracle_neighbor3d* out) {
    const int hw = w / 2;
    const float is1 = (s1 > 0) ? 1.0f / (s1 * s1) : 0.0f;
    const float is2 = (s2 > 0) ? 1.0f / (s2 * s2) : 0.0f;
    int k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++) {
                oracle_neighbor3d n;
                n.du = du; n.dv = dv; n.dw = dw;
                const float l2 = (float)(du * du + dv * dv + dw * dw);
                if (l2 == 0.0f) { n.d[0] = n.d[1] = n.d[2] = 0.0f; }
                else {
                    const float L = sqrtf(l2);
                    n.d[0] = (float)du / L; n.d[1] = (float)dv / L; n.d[2] = (float)dw / L;
                }
                n.c1 = s1 == 0.0f ? (l2 > O_TV_EPS ? 0.0f : 1.0f) : expf(-l2 * is1);
                n.c2 = s2 == 0.0f ? (l2 > O_TV_EPS ? 0.0f : 1.0f) : expf(-l2 * is2);
                out[k++] = n;
            }
    return k;
}

static double o_factorial(unsigned n) { /* tensor.h:19-24 */
    double f = 1;
    for (unsigned i = 1; i <= n; i++) f *= i;
    return f;
}

/* sticknorm3, tensorvote.h:1162-1172 (Type = float) */
float oracle_sticknorm3(float s1, float s2, unsigned p) {
    float pi_term = (float)(O_PI * sqrt(O_PI) / 2.0);
    float s13 = s1 * s1 * s1, s23 = s2 * s2 * s2;
    float num1 = (float)(pow(2, 2 * p + 1) * o_factorial(p) * o_factorial(p));
    float num2 = 2.0f;
    float den1 = (float)o_factorial(2 * p + 1);
    float den2 = (float)(2 * p + 1);
    return pi_term * ((s13 * num1 / den1) + (s23 * num2 / den2));
}

/* stickvote3_accumulate_kernel, tensorvote.h:1183-1213 */
static void o_stick3_acc(float* m, const float* d, float c1, float c2, const float* q, unsigned power,
                         float scale) {
    const float qTd = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
    const float q2 = qTd * qTd;
    float eta;
    if (power == 1) eta = c1 * (1.0f - q2) + c2 * q2;
    else {
        const float t1 = 1.0f - q2, t2 = q2;
        float e1 = t1, e2 = t2;
        for (unsigned i = 1; i < power; ++i) { e1 *= t1; e2 *= t2; }
        eta = c1 * e1 + c2 * e2;
    }
    const float rx = q[0] - 2.0f * qTd * d[0];
    const float ry = q[1] - 2.0f * qTd * d[1];
    const float rz = q[2] - 2.0f * qTd * d[2];
    const float term = scale * eta;
    m[0] += term * rx * rx; m[1] += term * rx * ry; m[2] += term * rx * rz;
    m[3] += term * ry * ry; m[4] += term * ry * rz; m[5] += term * rz * rz;
}

static void o_normalize_glm(float* v) { /* glm::normalize = v * inversesqrt(dot(v,v)) */
    float inv = 1.0f / sqrtf(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    v[0] *= inv; v[1] *= inv; v[2] *= inv;
}

static void o_plate_basis_f32(const float* n, float* u, float* v) {
    /* tensorvote.h:1667-1673: a = z unless |n.z| >= 0.999, u = normalize(a x n), v = n x u */
    float a[3];
    if (fabsf(n[2]) < 0.999f) { a[0] = 0; a[1] = 0; a[2] = 1; }
    else { a[0] = 1; a[1] = 0; a[2] = 0; }
    o_cross3(a, n, u);
    o_normalize_glm(u);
    o_cross3(n, u, v);
}

/* cpu::tk_plate_numerical (3D), tensorvote.h:1651-1695; out6 upper triangle */
void oracle_plate3_numerical_f32(const float* dn, float c1, float c2, unsigned power, const float* e0,
                                 unsigned samples, float* out6) {
    for (int k = 0; k < 6; k++) out6[k] = 0.0f;
    const float en2 = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
    if (samples == 0 || !(en2 > O_TV_EPS)) return;
    float u[3], v[3];
    o_plate_basis_f32(e0, u, v);
    const float db = (float)O_PI / (float)samples;
    float m[6] = {0, 0, 0, 0, 0, 0};
    for (unsigned i = 0; i < samples; ++i) {
        const float beta = db * (float)i;
        const float cb = cosf(beta), sb = sinf(beta);
        const float q[3] = {cb * u[0] + sb * v[0], cb * u[1] + sb * v[1], cb * u[2] + sb * v[2]};
        o_stick3_acc(m, dn, c1, c2, q, power, 1.0f);
    }
    for (int k = 0; k < 6; k++) out6[k] = m[k] * db;
}

/* Terminating Gauss series 2F1(-p, b; c; z) (the reference calls
 * boost::math::hypergeometric_pFq at tensorvote.h:1589-1592 and the same
 * series on the device at 2364-2376). */
static double o_hyp2f1_negp(unsigned p, double b, double c, double z) {
    double sum = 1.0, term = 1.0;
    for (unsigned k = 0; k < p; ++k) {
        term *= ((-(double)p + k) * (b + k)) / ((k + 1.0) * (c + k)) * z;
        sum += term;
    }
    return sum;
}

static double o_beta(double a, double b) { return exp(lgamma(a) + lgamma(b) - lgamma(a + b)); }

/* cpu::tk_plate_analytic, tensorvote.h:1564-1635, evaluated in FP64. out6 upper triangle. */
void oracle_plate3_analytic_f64(const double* dn, double c1, double c2, unsigned power, const double* n,
                                double* out6) {
    double a[3] = {0, 0, 1};
    if (!(fabs(n[2]) < 0.999)) { a[0] = 1; a[2] = 0; }
    double u[3] = {a[1] * n[2] - a[2] * n[1], a[2] * n[0] - a[0] * n[2], a[0] * n[1] - a[1] * n[0]};
    double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] /= ul; u[1] /= ul; u[2] /= ul;
    double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    const double dx = u[0] * dn[0] + u[1] * dn[1] + u[2] * dn[2];
    const double dy = v[0] * dn[0] + v[1] * dn[1] + v[2] * dn[2];
    const double dz = n[0] * dn[0] + n[1] * dn[1] + n[2] * dn[2];
    const double alpha = sqrt(dx * dx + dy * dy), a2 = alpha * alpha;
    const double phi = atan2(dy, dx);
    const double a2d = a2 < 1.0 ? a2 : 1.0;
    const double J0 = 0.5 * O_PI * o_hyp2f1_negp(power, 1.5, 2.0, a2d);
    const double J1 = O_PI * o_hyp2f1_negp(power, 0.5, 1.0, a2d);
    const double K0 = o_beta(0.5, power + 1.5), K1 = o_beta(0.5, power + 0.5);
    const double t = 1.0 - 2.0 * a2;
    /* local-frame matrices (row, col), symmetric */
    double A[3][3] = {{t * t * J0, 0, -2 * alpha * dz * t * J0}, {0, J1 - J0, 0},
                      {-2 * alpha * dz * t * J0, 0, 4 * a2 * dz * dz * J0}};
    const double a2p = pow(a2, (double)power);
    double B[3][3] = {{a2p * t * t * K0, 0, -2 * alpha * a2p * dz * t * K0}, {0, a2p * (K1 - K0), 0},
                      {-2 * alpha * a2p * dz * t * K0, 0, 4 * a2p * a2 * dz * dz * K0}};
    /* glm Rz built column-wise: col0 = (c, -s, 0), col1 = (s, c, 0) => Rz(r,c) */
    const double cph = cos(phi), sph = sin(phi);
    double Rz[3][3] = {{cph, sph, 0}, {-sph, cph, 0}, {0, 0, 1}};
    double Z[3][3] = {{u[0], v[0], n[0]}, {u[1], v[1], n[1]}, {u[2], v[2], n[2]}};
    double M[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) M[r][c] = c1 * A[r][c] + c2 * B[r][c];
    /* term = Rz^T M Rz ; global = Z term Z^T */
    double T1[3][3], T2[3][3], G[3][3];
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += M[r][k] * Rz[k][c];
            T1[r][c] = s;
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++) s += Rz[k][r] * T1[k][c];
            T2[r][c] = s;
        }
    for (int r = 0; r < 3; r++)
        for (int c = 0; c < 3; c++) {
            double s = 0;
            for (int k = 0; k < 3; k++)
                for (int l = 0; l < 3; l++) s += Z[r][k] * T2[k][l] * Z[c][l];
            G[r][c] = s;
        }
    out6[0] = G[0][0]; out6[1] = G[0][1]; out6[2] = G[0][2];
    out6[3] = G[1][1]; out6[4] = G[1][2]; out6[5] = G[2][2];
}

/* FP64 stick pair: eta * r r^T (tensorvote.h:1183-1213 semantics, scale 1) */
void oracle_stick3_pair_f64(const double* d, double c1, double c2, unsigned power, const double* q,
                            double* o) {
    const double t = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
    const double t2 = t * t;
    const unsigned p = power < 1 ? 1 : power;
    const double eta = c1 * pow(1.0 - t2, (double)p) + c2 * pow(t2, (double)p);
    const double r[3] = {q[0] - 2 * t * d[0], q[1] - 2 * t * d[1], q[2] - 2 * t * d[2]};
    o[0] = eta * r[0] * r[0]; o[1] = eta * r[0] * r[1]; o[2] = eta * r[0] * r[2];
    o[3] = eta * r[1] * r[1]; o[4] = eta * r[1] * r[2]; o[5] = eta * r[2] * r[2];
}

/* FP64 literal N-sample plate integral (1651-1695 in double). n need not be unit. */
void oracle_plate3_sum_f64(const double* d, double c1, double c2, unsigned power, const double* n0,
                           unsigned samples, double* o) {
    for (int k = 0; k < 6; k++) o[k] = 0;
    double nn = n0[0] * n0[0] + n0[1] * n0[1] + n0[2] * n0[2];
    if (samples == 0 || !(nn > O_TV_EPS)) return;
    double il = 1.0 / sqrt(nn);
    double n[3] = {n0[0] * il, n0[1] * il, n0[2] * il};
    double a[3] = {0, 0, 1};
    if (!(fabs(n[2]) < 0.999)) { a[0] = 1; a[2] = 0; }
    double u[3] = {a[1] * n[2] - a[2] * n[1], a[2] * n[0] - a[0] * n[2], a[0] * n[1] - a[1] * n[0]};
    double ul = sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
    u[0] /= ul; u[1] /= ul; u[2] /= ul;
    double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    const double db = O_PI / samples;
    for (unsigned i = 0; i < samples; i++) {
        double cb = cos(db * i), sb = sin(db * i), q[3], s[6];
        for (int k = 0; k < 3; k++) q[k] = cb * u[k] + sb * v[k];
        oracle_stick3_pair_f64(d, c1, c2, power, q, s);
        for (int k = 0; k < 6; k++) o[k] += s[k];
    }
    for (int k = 0; k < 6; k++) o[k] *= db;
}

/* Shared 3D receiver loop. mode 0: FP32 faithful. */
typedef struct {
    unsigned s0, s1, s2;
    const float* L;      /* n x 3 */
    const float* Q;      /* n x 3 largest evec (stick) */
    const float* P;      /* n x 3 smallest evec (plate) */
    const oracle_neighbor3d* NB;
    int nbc;
    unsigned power, samples;
    int stick, plate;
    float sticknorm, platenorm;
    int asym; /* reproduce the CPU plate's upper-triangle-only accumulation */
} o_v3ctx;

static void o_vote3_receiver_f32(const o_v3ctx* c, int x0, int x1, int x2, float* out9) {
    float sm[6] = {0, 0, 0, 0, 0, 0}, pm[6] = {0, 0, 0, 0, 0, 0};
    if (c->stick) { /* cpu::tk_stickvote, tensorvote.h:1513-1547 */
        for (int k = 0; k < c->nbc; k++) {
            const oracle_neighbor3d* n = &c->NB[k];
            const int r0 = x0 + n->dw; if ((unsigned)r0 >= c->s0) continue;
            const int r1 = x1 + n->dv; if ((unsigned)r1 >= c->s1) continue;
            const int r2 = x2 + n->du; if ((unsigned)r2 >= c->s2) continue;
            const size_t base = (size_t)r0 * c->s1 * c->s2 + (size_t)r1 * c->s2 + (size_t)r2;
            const float l1 = c->L[base * 3 + 1], l2 = c->L[base * 3 + 2];
            const float scale = copysignf(fabsf(l2) - fabsf(l1), l2);
            o_stick3_acc(sm, n->d, n->c1, n->c2, c->Q + base * 3, c->power, scale);
        }
    }
    if (c->plate) { /* cpu::tk_platevote, tensorvote.h:1714-1754 */
        double K0 = 0, K1 = 0;
        if (c->samples == 0) { K0 = o_beta(0.5, c->power + 1.5); K1 = o_beta(0.5, c->power + 0.5); }
        (void)K0; (void)K1;
        for (int k = 0; k < c->nbc; k++) {
            const oracle_neighbor3d* n = &c->NB[k];
            const int r0 = x0 + n->dw; if ((unsigned)r0 >= c->s0) continue;
            const int r1 = x1 + n->dv; if ((unsigned)r1 >= c->s1) continue;
            const int r2 = x2 + n->du; if ((unsigned)r2 >= c->s2) continue;
            const size_t base = (size_t)r0 * c->s1 * c->s2 + (size_t)r1 * c->s2 + (size_t)r2;
            const float l0 = c->L[base * 3 + 0], l1 = c->L[base * 3 + 1];
            const float scale = copysignf(fabsf(l1) - fabsf(l0), l1);
            float V[6];
            if (c->samples > 0) oracle_plate3_numerical_f32(n->d, n->c1, n->c2, c->power, c->P + base * 3, c->samples, V);
            else {
                /* analytic form restated in FP64 then rounded (the reference mixes FP32 geometry
                 * with FP64 J/K terms, so the two agree to FP32 rounding, not bit for bit). */
                const float* e = c->P + base * 3;
                double dd[3] = {n->d[0], n->d[1], n->d[2]}, ee[3] = {e[0], e[1], e[2]}, V6[6];
                oracle_plate3_analytic_f64(dd, n->c1, n->c2, c->power, ee, V6);
                for (int j = 0; j < 6; j++) V[j] = (float)V6[j];
            }
            for (int j = 0; j < 6; j++) pm[j] += scale * V[j];
        }
    }
    /* cpu::tensorvote_tk, tensorvote.h:1790-1796: Vote = sticknorm*stick + platenorm*plate.
     * glm::mat3 column-major flat slots: 0=00 1=01 2=02 3=10 4=11 5=12 6=20 7=21 8=22. */
    const float sn = c->sticknorm, pn = c->platenorm;
    float S[9] = {sm[0], sm[1], sm[2], sm[1], sm[3], sm[4], sm[2], sm[4], sm[5]};
    float Pl[9];
    if (c->asym) {
        float t[9] = {pm[0], pm[1], pm[2], 0, pm[3], pm[4], 0, 0, pm[5]};
        memcpy(Pl, t, sizeof(t));
    } else {
        float t[9] = {pm[0], pm[1], pm[2], pm[1], pm[3], pm[4], pm[2], pm[4], pm[5]};
        memcpy(Pl, t, sizeof(t));
    }
    for (int j = 0; j < 9; j++) {
        float v = 0.0f;
        if (c->stick) v += sn * S[j];
        if (c->plate) v += pn * Pl[j];
        out9[j] = v;
    }
}

static void o_vote3_driver_f32(const o_v3ctx* c, float* out9, const int64_t* recv, size_t nrecv, int nthreads) {
    const size_t nvox = (size_t)c->s0 * c->s1 * c->s2;
    const size_t nr = recv ? nrecv : nvox;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)nr; i++) {
        size_t v = recv ? (size_t)recv[i] : (size_t)i;
        int x2 = (int)(v % c->s2), x1 = (int)((v / c->s2) % c->s1), x0 = (int)(v / ((size_t)c->s1 * c->s2));
        o_vote3_receiver_f32(c, x0, x1, x2, out9 + (size_t)i * 9);
    }
}

/* cpu::tensorvote3, tensorvote.h:1819-1859 */
void oracle_tensorvote3_f32(const float* in9, float* out9, unsigned s0, unsigned s1, unsigned s2,
                            float sigma1, float sigma2, unsigned power, unsigned w,
                            int stick, int plate, unsigned samples,
                            const int64_t* recv, size_t nrecv, int nthreads) {
    const size_t n = (size_t)s0 * s1 * s2;
    float* L = (float*)malloc(n * 3 * sizeof(float));
    float* T = (float*)malloc(n * 6 * sizeof(float));
    float* Q = (float*)malloc(n * 3 * sizeof(float));
    float* P = (float*)malloc(n * 3 * sizeof(float));
    oracle_evals3_symmetric(in9, n, L);
    oracle_evecs3spherical_symmetric(in9, L, n, T);
    for (size_t i = 0; i < n; i++) {
        oracle_spherical_to_cartesian(T[i * 6 + 4], T[i * 6 + 5], Q + i * 3);
        oracle_spherical_to_cartesian(T[i * 6 + 0], T[i * 6 + 1], P + i * 3);
    }
    oracle_neighbor3d* NB = (oracle_neighbor3d*)malloc((size_t)w * w * w * sizeof(oracle_neighbor3d));
    o_v3ctx c;
    c.s0 = s0; c.s1 = s1; c.s2 = s2; c.L = L; c.Q = Q; c.P = P; c.NB = NB;
    c.nbc = oracle_build_neighbors3d((int)w, sigma1, sigma2, NB);
    c.power = power; c.samples = samples; c.stick = stick; c.plate = plate;
    c.sticknorm = 1.0f / oracle_sticknorm3(sigma1, sigma2, power);
    c.platenorm = 1.0f / (float)O_PI;
    c.asym = 1;
    o_vote3_driver_f32(&c, out9, recv, nrecv, nthreads);
    free(NB); free(L); free(T); free(Q); free(P);
}

/* _tensorvote3_tk semantics (tensorvote.h:2741-2809): symmetric output. */
void oracle_tensorvote3_tk_f32(const float* L, const float* e2, const float* e0, float* out9,
                               unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2,
                               unsigned power, unsigned w, int stick, int plate, unsigned samples,
                               const int64_t* recv, size_t nrecv, int nthreads) {
    oracle_neighbor3d* NB = (oracle_neighbor3d*)malloc((size_t)w * w * w * sizeof(oracle_neighbor3d));
    o_v3ctx c;
    c.s0 = s0; c.s1 = s1; c.s2 = s2; c.L = L; c.Q = e2; c.P = e0; c.NB = NB;
    c.nbc = oracle_build_neighbors3d((int)w, sigma1, sigma2, NB);
    c.power = power; c.samples = samples; c.stick = stick; c.plate = plate;
    c.sticknorm = 1.0f / oracle_sticknorm3(sigma1, sigma2, power);
    c.platenorm = 1.0f / (float)O_PI;
    c.asym = 0;
    o_vote3_driver_f32(&c, out9, recv, nrecv, nthreads);
    free(NB);
}

/* FP64 route */
void oracle_tensorvote3_f64(const float* in9, double* out9, unsigned s0, unsigned s1, unsigned s2,
                            double sigma1, double sigma2, unsigned power, unsigned w,
                            int stick, int plate, unsigned samples,
                            const int64_t* recv, size_t nrecv, int nthreads) {
    const size_t n = (size_t)s0 * s1 * s2;
    double* L = (double*)malloc(n * 3 * sizeof(double));
    double* E0 = (double*)malloc(n * 3 * sizeof(double));
    double* E2 = (double*)malloc(n * 3 * sizeof(double));
    oracle_eig3_f64(in9, n, L, E0, E2);
    const unsigned p = power < 1 ? 1 : power;
    /* sticknorm3 in FP64 */
    double sn = O_PI * sqrt(O_PI) / 2.0 *
                (sigma1 * sigma1 * sigma1 * pow(2, 2 * p + 1) * o_factorial(p) * o_factorial(p) / o_factorial(2 * p + 1) +
                 sigma2 * sigma2 * sigma2 * 2.0 / (2 * p + 1));
    sn = 1.0 / sn;
    const double pn = 1.0 / O_PI;
    const int hw = (int)w / 2;
    const size_t nr = recv ? nrecv : n;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)nr; i++) {
        size_t vx = recv ? (size_t)recv[i] : (size_t)i;
        int x2 = (int)(vx % s2), x1 = (int)((vx / s2) % s1), x0 = (int)(vx / ((size_t)s1 * s2));
        double S[6] = {0, 0, 0, 0, 0, 0}, Pm[6] = {0, 0, 0, 0, 0, 0};
        for (int dw = -hw; dw <= hw; dw++) {
            int r0 = x0 + dw;
            if (r0 < 0 || r0 >= (int)s0) continue;
            for (int dv = -hw; dv <= hw; dv++) {
                int r1 = x1 + dv;
                if (r1 < 0 || r1 >= (int)s1) continue;
                for (int du = -hw; du <= hw; du++) {
                    int r2 = x2 + du;
                    if (r2 < 0 || r2 >= (int)s2) continue;
                    size_t b = ((size_t)r0 * s1 + r1) * s2 + r2;
                    double l2sq = (double)du * du + (double)dv * dv + (double)dw * dw;
                    double d[3] = {0, 0, 0};
                    if (l2sq > 0) { double L1 = sqrt(l2sq); d[0] = du / L1; d[1] = dv / L1; d[2] = dw / L1; }
                    double c1 = sigma1 == 0 ? (l2sq > O_TV_EPS ? 0 : 1) : exp(-l2sq / (sigma1 * sigma1));
                    double c2 = sigma2 == 0 ? (l2sq > O_TV_EPS ? 0 : 1) : exp(-l2sq / (sigma2 * sigma2));
                    double o6[6];
                    if (stick) {
                        double sc = copysign(fabs(L[b * 3 + 2]) - fabs(L[b * 3 + 1]), L[b * 3 + 2]);
                        if (sc != 0) {
                            oracle_stick3_pair_f64(d, c1, c2, p, E2 + b * 3, o6);
                            for (int k = 0; k < 6; k++) S[k] += sc * o6[k];
                        }
                    }
                    if (plate) {
                        double sc = copysign(fabs(L[b * 3 + 1]) - fabs(L[b * 3 + 0]), L[b * 3 + 1]);
                        if (sc != 0) {
                            if (samples > 0) oracle_plate3_sum_f64(d, c1, c2, p, E0 + b * 3, samples, o6);
                            else oracle_plate3_analytic_f64(d, c1, c2, p, E0 + b * 3, o6);
                            for (int k = 0; k < 6; k++) Pm[k] += sc * o6[k];
                        }
                    }
                }
            }
        }
        double M[6];
        for (int k = 0; k < 6; k++) M[k] = (stick ? sn * S[k] : 0) + (plate ? pn * Pm[k] : 0);
        double* o = out9 + (size_t)i * 9;
        o[0] = M[0]; o[1] = M[1]; o[2] = M[2]; o[3] = M[1]; o[4] = M[3]; o[5] = M[4];
        o[6] = M[2]; o[7] = M[4]; o[8] = M[5];
    }
    free(L); free(E0); free(E2);
}

/* ======================================================================
 * Gaussian filter (filter.h:14-200)
 * ====================================================================== */

/* kernel length rule of gaussian_convolve3 / gaussian_filter3d, filter.h:178-183, 405-410 */
unsigned oracle_gaussian_ksize(float sigma) {
    unsigned k = (unsigned)ceil(6.0f * sigma);
    if (k % 2 == 0) k++;
    return k;
}

/* normaldist (filter.h:14-22) + gaussian1d (filter.h:37-64), Type = float */
void oracle_gaussian1d(unsigned ksize, float mu, float sigma, float dx, float* k) {
    const float hwid = (float)ksize * dx / 2.0f;
    const float startx = -hwid + (dx / 2.0f);
    float sum = 0.0f;
    for (unsigned i = 0; i < ksize; ++i) {
        const float x = startx + (float)i * dx;
        float ssq = sigma * sigma;
        float pi = (float)O_PI;
        float nrm = 1.0f / sqrtf(2.0f * pi * ssq);
        float ex = -((x - mu) * (x - mu)) / (2.0f * ssq);
        k[i] = (float)((double)nrm * exp((double)ex)); /* ::exp(double) */
        sum += k[i];
    }
    if (sum > 0.0f) {
        float inv = 1.0f / sum;
        for (unsigned i = 0; i < ksize; i++) k[i] *= inv;
    }
}

/* cpu::convolve3 (valid mode), filter.h:107-144 */
void oracle_convolve3(const float* in, unsigned i0, unsigned i1, unsigned i2, const float* ker,
                      unsigned k0, unsigned k1, unsigned k2, float* out, unsigned* o0, unsigned* o1,
                      unsigned* o2) {
    *o0 = i0 - (k0 - 1); *o1 = i1 - (k1 - 1); *o2 = i2 - (k2 - 1);
    for (unsigned zi = 0; zi < *o2; ++zi)
        for (unsigned yi = 0; yi < *o1; ++yi)
            for (unsigned xi = 0; xi < *o0; ++xi) {
                float s = 0.0f;
                for (unsigned wi = 0; wi < k2; ++wi)
                    for (unsigned vi = 0; vi < k1; ++vi)
                        for (unsigned ui = 0; ui < k0; ++ui) {
                            const size_t kidx = (size_t)wi * k1 * k0 + (size_t)vi * k0 + ui;
                            const size_t iidx = (size_t)(zi + wi) * i1 * i0 + (size_t)(yi + vi) * i0 + (xi + ui);
                            s += ker[kidx] * in[iidx];
                        }
                out[(size_t)zi * (*o1) * (*o0) + (size_t)yi * (*o0) + xi] = s;
            }
}

/* cpu::gaussian_convolve3, filter.h:174-200 (x, then y, then z) */
void oracle_gaussian_convolve3(const float* in, unsigned sx, unsigned sy, unsigned sz, float sw, float sh,
                               float sd, float* out, unsigned* ox, unsigned* oy, unsigned* oz) {
    unsigned kw = oracle_gaussian_ksize(sw), kh = oracle_gaussian_ksize(sh), kd = oracle_gaussian_ksize(sd);
    float* gw = (float*)malloc(kw * sizeof(float));
    float* gh = (float*)malloc(kh * sizeof(float));
    float* gd = (float*)malloc(kd * sizeof(float));
    oracle_gaussian1d(kw, 0, sw, 1, gw);
    oracle_gaussian1d(kh, 0, sh, 1, gh);
    oracle_gaussian1d(kd, 0, sd, 1, gd);
    unsigned tx, ty, tz;
    float* dx = (float*)malloc((size_t)sx * sy * sz * sizeof(float));
    oracle_convolve3(in, sx, sy, sz, gw, kw, 1, 1, dx, &tx, &ty, &tz);
    float* dxy = (float*)malloc((size_t)tx * ty * tz * sizeof(float));
    oracle_convolve3(dx, tx, ty, tz, gh, 1, kh, 1, dxy, &tx, &ty, &tz);
    oracle_convolve3(dxy, tx, ty, tz, gd, 1, 1, kd, out, ox, oy, oz);
    free(gw); free(gh); free(gd); free(dx); free(dxy);
}

void oracle_gaussian3_same_packed6_f64(const float* in6, unsigned s0, unsigned s1, unsigned s2,
                                       const float* ker, unsigned K, double* out6, int nthreads) {
    const size_t n = (size_t)s0 * s1 * s2;
    const int r = (int)K / 2;
    double* t1 = (double*)malloc(n * 6 * sizeof(double));
    double* t2 = (double*)malloc(n * 6 * sizeof(double));
    /* x (fastest axis, s2) */
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
    for (long long row = 0; row < (long long)s0 * s1; row++)
        for (int x = 0; x < (int)s2; x++)
            for (int c = 0; c < 6; c++) {
                double s = 0;
                for (int k = 0; k < (int)K; k++) {
                    int xx = x + k - r;
                    if (xx >= 0 && xx < (int)s2) s += (double)ker[k] * in6[((size_t)row * s2 + xx) * 6 + c];
                }
                t1[((size_t)row * s2 + x) * 6 + c] = s;
            }
    /* y (s1) */
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
    for (long long z = 0; z < (long long)s0; z++)
        for (int y = 0; y < (int)s1; y++)
            for (int x = 0; x < (int)s2; x++)
                for (int c = 0; c < 6; c++) {
                    double s = 0;
                    for (int k = 0; k < (int)K; k++) {
                        int yy = y + k - r;
                        if (yy >= 0 && yy < (int)s1) s += (double)ker[k] * t1[(((size_t)z * s1 + yy) * s2 + x) * 6 + c];
                    }
                    t2[(((size_t)z * s1 + y) * s2 + x) * 6 + c] = s;
                }
    /* z (s0) */
#pragma omp parallel for num_threads(nthreads > 0 ? nthreads : 1)
    for (long long z = 0; z < (long long)s0; z++)
        for (int y = 0; y < (int)s1; y++)
            for (int x = 0; x < (int)s2; x++)
                for (int c = 0; c < 6; c++) {
                    double s = 0;
                    for (int k = 0; k < (int)K; k++) {
                        long long zz = z + k - r;
                        if (zz >= 0 && zz < (long long)s0) s += (double)ker[k] * t2[(((size_t)zz * s1 + y) * s2 + x) * 6 + c];
                    }
                    out6[(((size_t)z * s1 + y) * s2 + x) * 6 + c] = s;
                }
    free(t1); free(t2);
}

/* ======================================================================
 * Analytical tensor voting (ATV), 2D and 3D (tensorvote.h:23-760,
 * 255-306, 1298-1316, 1380-1399, 1443-1495, 1819-1859)
 *
 * One generic body per precision: T = float is the reference-faithful FP32
 * restatement (operation order kept; the unqualified exp of the reference's
 * host code resolves to the double overload, so it is evaluated in double
 * and rounded), T = double is the FP64 route used to state error.
 * ====================================================================== */
#define O_ATV_DEFINE(T, SFX, SQRT)                                                                       \
    /* atv_attenuation, tensorvote.h:34-40 */                                                            \
    static T o_atv_atten_##SFX(T dist, T sigma) {                                                        \
        if (sigma == (T)0) return (dist > (T)O_TV_EPS) ? (T)0 : (T)1;                                    \
        return (T)exp((double)(-(dist * dist) / (sigma * sigma)));                                      \
    }                                                                                                    \
    /* atv (2D) with atv_rotation, atv_normalized_T, atv_vTvT, atv_TvvT, atv_H:                       \
     * tensorvote.h:54-60, 73-79, 100-106, 125-138, 159-253. V = (a, b, c). */                          \
    void oracle_atv2_##SFX(T l0, T l1, T ev0x, T ev0y, T ev1x, T ev1y, T rx, T ry, T sigma, T* V) {      \
        const T dist = SQRT(rx * rx + ry * ry);                                                          \
        const T c = o_atv_atten_##SFX(dist, sigma);                                                      \
        T dx, dy;                                                                                        \
        if (dist == 0) { dx = 0; dy = 0; }                                                               \
        else { dx = rx / dist; dy = ry / dist; }                                                         \
        const T R_a = 1 - 2 * dx * dx, R_b = -2 * dx * dy, R_c = 1 - 2 * dy * dy;                        \
        const T T1_a = ev1x * ev1x, T1_b = ev1x * ev1y, T1_c = ev1y * ev1y;                              \
        const T T0_a = ev0x * ev0x + T1_a, T0_b = ev0x * ev0y + T1_b, T0_c = ev0y * ev0y + T1_c;         \
        const T v0 = dx * (dx * T0_a + dy * T0_b) + dy * (dx * T0_b + dy * T0_c);                       \
        const T v1 = dx * (dx * T1_a + dy * T1_b) + dy * (dx * T1_b + dy * T1_c);                       \
        const T t0x = T0_a * dx + T0_b * dy, t0y = T0_b * dx + T0_c * dy;                                \
        const T t1x = T1_a * dx + T1_b * dy, t1y = T1_b * dx + T1_c * dy;                                \
        const T s1 = l1 - l0, inv3 = (T)1.0 / (T)3.0;                                                    \
        T H_a = s1 * (T1_a - inv3 * (T1_a * v1 + 2 * (t1x * t1x)));                                      \
        T H_b = s1 * (T1_b - inv3 * (T1_b * v1 + 2 * (t1x * t1y)));                                      \
        T H_c = s1 * (T1_c - inv3 * (T1_c * v1 + 2 * (t1y * t1y)));                                      \
        const T s0 = l0, inv4 = (T)1.0 / (T)4.0;                                                         \
        H_a += s0 * (T0_a - inv4 * (T0_a * v0 + 2 * (t0x * t0x)));                                       \
        H_b += s0 * (T0_b - inv4 * (T0_b * v0 + 2 * (t0x * t0y)));                                       \
        H_c += s0 * (T0_c - inv4 * (T0_c * v0 + 2 * (t0y * t0y)));                                       \
        const T Ra2 = R_a * R_a, Rb2 = R_b * R_b, Rc2 = R_c * R_c;                                       \
        const T Rab = R_a * R_b, Rbc = R_b * R_c, Rac = R_a * R_c;                                       \
        V[0] = c * (Ra2 * H_a + 2 * Rab * H_b + Rb2 * H_c);                                              \
        V[1] = c * (Rb2 * H_b + Rab * H_a + Rac * H_b + Rbc * H_c);                                      \
        V[2] = c * (Rb2 * H_a + 2 * Rbc * H_b + Rc2 * H_c);                                              \
    }                                                                                                    \
    /* atv (3D) with atv_rotation, atv_normalized_T, atv_vTvT, atv_TvvT, atv_H:                       \
     * tensorvote.h:454-464, 485-493, 518-537, 563-578, 602-683, 688-760.                             \
     * e1 middle, e2 largest eigenvector. V = (a, b, c, d, e, f) = (00, 01, 11, 02, 12, 22). */        \
    void oracle_atv3_##SFX(T l0, T l1, T l2, const T* e1, const T* e2, T rx, T ry, T rz, T sigma, T* V) { \
        const T dist = SQRT(rx * rx + ry * ry + rz * rz);                                                \
        const T c = o_atv_atten_##SFX(dist, sigma);                                                      \
        T dx, dy, dz;                                                                                    \
        if (dist == (T)0) { dx = dy = dz = (T)0; }                                                       \
        else { dx = rx / dist; dy = ry / dist; dz = rz / dist; }                                         \
        const T two = (T)2;                                                                              \
        const T R00 = (T)1 - two * dx * dx, R01 = (T)0 - two * dx * dy, R11 = (T)1 - two * dy * dy;      \
        const T R02 = (T)0 - two * dx * dz, R12 = (T)0 - two * dy * dz, R22 = (T)1 - two * dz * dz;      \
        T Tk[3][6];                                                                                      \
        Tk[0][0] = e2[0] * e2[0]; Tk[0][1] = e2[0] * e2[1]; Tk[0][2] = e2[1] * e2[1];                    \
        Tk[0][3] = e2[0] * e2[2]; Tk[0][4] = e2[1] * e2[2]; Tk[0][5] = e2[2] * e2[2];                    \
        Tk[1][0] = e1[0] * e1[0]; Tk[1][1] = e1[0] * e1[1]; Tk[1][2] = e1[1] * e1[1];                    \
        Tk[1][3] = e1[0] * e1[2]; Tk[1][4] = e1[1] * e1[2]; Tk[1][5] = e1[2] * e1[2];                    \
        for (int k = 0; k < 6; k++) Tk[1][k] += Tk[0][k];                                                \
        Tk[2][0] = 1; Tk[2][1] = 0; Tk[2][2] = 1; Tk[2][3] = 0; Tk[2][4] = 0; Tk[2][5] = 1;              \
        const T s[3] = {l2 - l1, l1 - l0, l0};                                                           \
        const T inv[3] = {(T)1 / (T)3, (T)1 / (T)4, (T)1 / (T)5};                                        \
        T term[3][6];                                                                                    \
        for (int k = 0; k < 3; k++) {                                                                    \
            const T* t = Tk[k];                                                                          \
            const T tvx = t[0] * dx + t[1] * dy + t[3] * dz;                                             \
            const T tvy = t[1] * dx + t[2] * dy + t[4] * dz;                                             \
            const T tvz = t[3] * dx + t[4] * dy + t[5] * dz;                                             \
            const T vTv = dx * tvx + dy * tvy + dz * tvz;                                                \
            const T tt[6] = {tvx * tvx, tvx * tvy, tvy * tvy, tvx * tvz, tvy * tvz, tvz * tvz};          \
            for (int j = 0; j < 6; j++) term[k][j] = s[k] * (t[j] - inv[k] * (t[j] * vTv + two * tt[j])); \
        }                                                                                                \
        T Hh[6];                                                                                         \
        for (int j = 0; j < 6; j++) Hh[j] = term[0][j] + term[1][j] + term[2][j];                        \
        const T H00 = Hh[0], H01 = Hh[1], H02 = Hh[3], H11 = Hh[2], H12 = Hh[4], H22 = Hh[5];            \
        const T H10 = H01, H20 = H02, H21 = H12, R10 = R01, R20 = R02, R21 = R12;                        \
        const T m00 = R00 * H00 + R01 * H10 + R02 * H20, m01 = R00 * H01 + R01 * H11 + R02 * H21;        \
        const T m02 = R00 * H02 + R01 * H12 + R02 * H22, m10 = R10 * H00 + R11 * H10 + R12 * H20;        \
        const T m11 = R10 * H01 + R11 * H11 + R12 * H21, m12 = R10 * H02 + R11 * H12 + R12 * H22;        \
        const T m20 = R20 * H00 + R21 * H10 + R22 * H20, m21 = R20 * H01 + R21 * H11 + R22 * H21;        \
        const T m22 = R20 * H02 + R21 * H12 + R22 * H22;                                                 \
        V[0] = c * (m00 * R00 + m01 * R10 + m02 * R20);                                                  \
        V[1] = c * (m00 * R01 + m01 * R11 + m02 * R21);                                                  \
        V[2] = c * (m10 * R01 + m11 * R11 + m12 * R21);                                                  \
        V[3] = c * (m00 * R02 + m01 * R12 + m02 * R22);                                                  \
        V[4] = c * (m10 * R02 + m11 * R12 + m12 * R22);                                                  \
        V[5] = c * (m20 * R02 + m21 * R12 + m22 * R22);                                                  \
    }

O_ATV_DEFINE(float, f32, sqrtf)
O_ATV_DEFINE(double, f64, sqrt)

/* atv_window (tensorvote.h:255-306) over every pixel, as cpu::tensorvote_atv
 * (1298-1316) and the CUDA kernel (1998-2013) write it: [a b b c]. */
void oracle_tensorvote2_atv_f32(float* t_out, const float* lam, const float* th, float sigma, unsigned w,
                                unsigned S0, unsigned S1, int nthreads) {
    const int hw = (int)w / 2, s0 = (int)S0, s1 = (int)S1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int x0 = 0; x0 < s0; x0++)
        for (int x1 = 0; x1 < s1; x1++) {
            float oa = 0, ob = 0, oc = 0;
            for (int v = -hw; v <= hw; v++) {
                const int r0 = x0 + v;
                if (r0 < 0 || r0 >= s0) continue;
                for (int u = -hw; u <= hw; u++) {
                    const int r1 = x1 + u;
                    if (r1 < 0 || r1 >= s1) continue;
                    const size_t idx = (size_t)r0 * s1 * 2 + (size_t)r1 * 2;
                    const float t0 = th[idx], t1 = th[idx + 1];
                    float V[3];
                    oracle_atv2_f32(lam[idx], lam[idx + 1], cosf(t0), sinf(t0), cosf(t1), sinf(t1), (float)u,
                                    (float)v, sigma, V);
                    oa += V[0]; ob += V[1]; oc += V[2];
                }
            }
            float* o = t_out + ((size_t)x0 * s1 + x1) * 4;
            o[0] = oa; o[1] = ob; o[2] = ob; o[3] = oc;
        }
}

/* cpu::tensorvote with is_atv (tensorvote.h:1380-1399): evals2 + evecs2polar + ATV window. */
void oracle_tensorvote2_atv_full_f32(float* t_out, const float* t_in, float sigma, unsigned w, unsigned S0,
                                     unsigned S1, int nthreads) {
    const size_t n = (size_t)S0 * S1;
    float* lam = (float*)malloc(n * 2 * sizeof(float));
    float* th = (float*)malloc(n * 2 * sizeof(float));
    oracle_evals2_symmetric(t_in, n, lam);
    oracle_evecs2polar_symmetric(t_in, lam, n, th);
    oracle_tensorvote2_atv_f32(t_out, lam, th, sigma, w, S0, S1, nthreads);
    free(lam); free(th);
}

/* FP64 2D ATV: closed-form FP64 eigenpairs, FP64 pair formula, same window. */
void oracle_tensorvote2_atv_f64(double* t_out, const float* t_in, double sigma, unsigned w, unsigned S0,
                                unsigned S1, int nthreads) {
    const size_t n = (size_t)S0 * S1;
    double* e = (double*)malloc(n * 6 * sizeof(double)); /* l0, l1, e0x, e0y, e1x, e1y */
    for (size_t i = 0; i < n; i++) {
        const double a = t_in[i * 4], b = t_in[i * 4 + 1], c = t_in[i * 4 + 3];
        const double m = 0.5 * (a + c), r = hypot(0.5 * (a - c), b);
        const double ang = 0.5 * atan2(2 * b, a - c); /* eigenvector of the larger eigenvalue */
        double* o = e + i * 6;
        o[0] = m - r; o[1] = m + r;
        o[4] = cos(ang); o[5] = sin(ang);
        o[2] = -o[5]; o[3] = o[4];
    }
    const int hw = (int)w / 2, s0 = (int)S0, s1 = (int)S1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1)
    for (int x0 = 0; x0 < s0; x0++)
        for (int x1 = 0; x1 < s1; x1++) {
            double acc[3] = {0, 0, 0};
            for (int v = -hw; v <= hw; v++) {
                const int r0 = x0 + v;
                if (r0 < 0 || r0 >= s0) continue;
                for (int u = -hw; u <= hw; u++) {
                    const int r1 = x1 + u;
                    if (r1 < 0 || r1 >= s1) continue;
                    const double* p = e + ((size_t)r0 * s1 + r1) * 6;
                    double V[3];
                    oracle_atv2_f64(p[0], p[1], p[2], p[3], p[4], p[5], u, v, sigma, V);
                    acc[0] += V[0]; acc[1] += V[1]; acc[2] += V[2];
                }
            }
            double* o = t_out + ((size_t)x0 * s1 + x1) * 4;
            o[0] = acc[0]; o[1] = acc[1]; o[2] = acc[1]; o[3] = acc[2];
        }
    free(e);
}

/* atv_window3 (tensorvote.h:1443-1482) from eigenpairs: L n x 3 ascending,
 * E1 middle and E2 largest eigenvectors n x 3 (the inputs of _tensorvote3_atv,
 * 2695-2739). Output glm::mat3 (9 floats, symmetric). */
void oracle_tensorvote3_atv_eig_f32(const float* L, const float* E1, const float* E2, float* out9, unsigned s0,
                                    unsigned s1, unsigned s2, float sigma, unsigned w, const int64_t* recv,
                                    size_t nrecv, int nthreads) {
    const int hw = (int)w / 2;
    const size_t nvox = (size_t)s0 * s1 * s2, nr = recv ? nrecv : nvox;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)nr; i++) {
        const size_t vx = recv ? (size_t)recv[i] : (size_t)i;
        const int x2 = (int)(vx % s2), x1 = (int)((vx / s2) % s1), x0 = (int)(vx / ((size_t)s1 * s2));
        float M[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        for (int dz = -hw; dz <= hw; ++dz) {
            const int r0 = x0 + dz;
            if (r0 < 0 || r0 >= (int)s0) continue;
            for (int dy = -hw; dy <= hw; ++dy) {
                const int r1 = x1 + dy;
                if (r1 < 0 || r1 >= (int)s1) continue;
                for (int dx = -hw; dx <= hw; ++dx) {
                    const int r2 = x2 + dx;
                    if (r2 < 0 || r2 >= (int)s2) continue;
                    const size_t b = ((size_t)r0 * s1 + r1) * s2 + r2;
                    float V[6];
                    oracle_atv3_f32(L[b * 3], L[b * 3 + 1], L[b * 3 + 2], E1 + b * 3, E2 + b * 3, (float)dx,
                                    (float)dy, (float)dz, sigma, V);
                    /* Receiver[c][r] += ..., column-major flat: 0=00 1=01 2=02 3=10 4=11 5=12 6=20 7=21 8=22 */
                    M[0] += V[0]; M[1] += V[1]; M[2] += V[3];
                    M[3] += V[1]; M[4] += V[2]; M[5] += V[4];
                    M[6] += V[3]; M[7] += V[4]; M[8] += V[5];
                }
            }
        }
        memcpy(out9 + (size_t)i * 9, M, sizeof(M));
    }
}

/* cpu::tensorvote3 with is_atv (tensorvote.h:1819-1852): evals3, evecs3spherical,
 * middle / largest eigenvectors back to cartesian, ATV window. */
void oracle_tensorvote3_atv_f32(const float* in9, float* out9, unsigned s0, unsigned s1, unsigned s2, float sigma,
                                unsigned w, const int64_t* recv, size_t nrecv, int nthreads) {
    const size_t n = (size_t)s0 * s1 * s2;
    float* L = (float*)malloc(n * 3 * sizeof(float));
    float* Th = (float*)malloc(n * 6 * sizeof(float));
    float* E1 = (float*)malloc(n * 3 * sizeof(float));
    float* E2 = (float*)malloc(n * 3 * sizeof(float));
    oracle_evals3_symmetric(in9, n, L);
    oracle_evecs3spherical_symmetric(in9, L, n, Th);
    for (size_t i = 0; i < n; i++) {
        oracle_spherical_to_cartesian(Th[i * 6 + 4], Th[i * 6 + 5], E2 + i * 3);
        oracle_spherical_to_cartesian(Th[i * 6 + 2], Th[i * 6 + 3], E1 + i * 3);
    }
    oracle_tensorvote3_atv_eig_f32(L, E1, E2, out9, s0, s1, s2, sigma, w, recv, nrecv, nthreads);
    free(L); free(Th); free(E1); free(E2);
}

/* FP64 3D ATV: Jacobi eigenpairs, e1 = e2 x e0, FP64 pair formula, same window. */
void oracle_tensorvote3_atv_f64(const float* in9, double* out9, unsigned s0, unsigned s1, unsigned s2,
                                double sigma, unsigned w, const int64_t* recv, size_t nrecv, int nthreads) {
    const size_t n = (size_t)s0 * s1 * s2;
    double* L = (double*)malloc(n * 3 * sizeof(double));
    double* E0 = (double*)malloc(n * 3 * sizeof(double));
    double* E1 = (double*)malloc(n * 3 * sizeof(double));
    double* E2 = (double*)malloc(n * 3 * sizeof(double));
    oracle_eig3_f64(in9, n, L, E0, E2);
    for (size_t i = 0; i < n; i++) {
        const double* a = E2 + i * 3; const double* b = E0 + i * 3;
        double* c = E1 + i * 3;
        c[0] = a[1] * b[2] - a[2] * b[1]; c[1] = a[2] * b[0] - a[0] * b[2]; c[2] = a[0] * b[1] - a[1] * b[0];
    }
    const int hw = (int)w / 2;
    const size_t nr = recv ? nrecv : n;
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long i = 0; i < (long long)nr; i++) {
        const size_t vx = recv ? (size_t)recv[i] : (size_t)i;
        const int x2 = (int)(vx % s2), x1 = (int)((vx / s2) % s1), x0 = (int)(vx / ((size_t)s1 * s2));
        double M[6] = {0, 0, 0, 0, 0, 0};
        for (int dz = -hw; dz <= hw; ++dz) {
            const int r0 = x0 + dz;
            if (r0 < 0 || r0 >= (int)s0) continue;
            for (int dy = -hw; dy <= hw; ++dy) {
                const int r1 = x1 + dy;
                if (r1 < 0 || r1 >= (int)s1) continue;
                for (int dx = -hw; dx <= hw; ++dx) {
                    const int r2 = x2 + dx;
                    if (r2 < 0 || r2 >= (int)s2) continue;
                    const size_t b = ((size_t)r0 * s1 + r1) * s2 + r2;
                    double V[6];
                    oracle_atv3_f64(L[b * 3], L[b * 3 + 1], L[b * 3 + 2], E1 + b * 3, E2 + b * 3, dx, dy, dz,
                                    sigma, V);
                    for (int k = 0; k < 6; k++) M[k] += V[k];
                }
            }
        }
        double* o = out9 + (size_t)i * 9;
        o[0] = M[0]; o[1] = M[1]; o[2] = M[3]; o[3] = M[1]; o[4] = M[2]; o[5] = M[4];
        o[6] = M[3]; o[7] = M[4]; o[8] = M[5];
    }
    free(L); free(E0); free(E1); free(E2);
}

/* ======================================================================
 * Saliency threshold (reconstruct.py:14-63) and Lee 3D thinning
 * (thinning/lee_thinning.h:112-139, 185-598, 600-706)
 *
 * A 3x3x3 neighbourhood is a 27-bit code, bit n = voxel
 * (dx, dy, dz) with n = 9 (dz + 1) + 3 (dy + 1) + (dx + 1), the order of
 * get_neighborhood (lee_thinning.h:52-91); bit 13 is the centre.
 * ====================================================================== */

/* Otsu threshold of nonnegative float32 values, reconstruct.py:37-63 with
 * numpy.histogram(values, 256, range=(0, max)) semantics for float32 input:
 * float32 edges (float64 linspace rounded), bin index from the float64
 * quotient with the +-1 edge corrections, float32 bin centres, float64
 * cumulative sums, first maximum of the between-class variance. */
float oracle_otsu_threshold_f32(const float* v, size_t n) {
    if (n == 0) return 0.0f;
    float vmax = v[0];
    for (size_t i = 1; i < n; i++) if (v[i] > vmax) vmax = v[i];
    if (!(vmax > 0.0f)) return 0.0f;
    float edges[257];
    const double step = (double)vmax / 256.0;
    for (int i = 0; i < 256; i++) edges[i] = (float)((double)i * step);
    edges[256] = vmax;
    double hist[256];
    for (int i = 0; i < 256; i++) hist[i] = 0.0;
    for (size_t k = 0; k < n; k++) {
        const float a = v[k];
        if (!(a >= 0.0f && a <= vmax)) continue;
        long idx = (long)(((double)a / (double)vmax) * 256.0);
        if (idx == 256) idx = 255;
        if (a < edges[idx]) idx--;
        if (idx != 255 && a >= edges[idx + 1]) idx++;
        hist[idx] += 1.0;
    }
    double total = 0.0;
    for (int i = 0; i < 256; i++) total += hist[i];
    if (total == 0.0) return 0.0f;
    float centers[256];
    for (int i = 0; i < 256; i++) centers[i] = (edges[i] + edges[i + 1]) / 2.0f;
    double cum[256], w0[256], acc = 0.0, wacc = 0.0;
    for (int i = 0; i < 256; i++) {
        acc += hist[i] * (double)centers[i];
        wacc += hist[i];
        cum[i] = acc; w0[i] = wacc;
    }
    const double mu_total = cum[255];
    int best = 0;
    double bv = 0.0;
    for (int i = 0; i < 256; i++) {
        const double w1 = total - w0[i];
        const double mu0 = w0[i] > 0 ? cum[i] / w0[i] : 0.0;
        const double mu1 = w1 > 0 ? (mu_total - cum[i]) / w1 : 0.0;
        const double d = mu0 - mu1;
        const double b = w0[i] * w1 * (d * d);
        if (i == 0 || b > bv) { bv = b; best = i; }
    }
    return centers[best];
}

/* Euler characteristic LUT of Lee et al. 1994 for 26-connectivity, indexed by
 * a 2x2x2 octant code (lee_thinning.h:112-139); only odd entries are used. */
static const signed char o_euler_lut[256] = {
    0, 1, 0, -1, 0, -1, 0, 1, 0, -3, 0, -1, 0, -1, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, -3, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, -3, 0, 3, 0, -1, 0, 1, 0, 1, 0, 3, 0, -1, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, 1, 0, 3, 0, 3, 0, 1, 0, 5, 0, 3, 0, 3, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, -7, 0, -1, 0, -1, 0, 1, 0, -3, 0, -1, 0, -1, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, -3, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, -3, 0, 3, 0, -1, 0, 1, 0, 1, 0, 3, 0, -1, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1,
    0, 1, 0, 3, 0, 3, 0, 1, 0, 5, 0, 3, 0, 3, 0, 1, 0, -1, 0, 1, 0, 1, 0, -1, 0, 3, 0, 1, 0, 1, 0, -1};

/* Octant index of the Euler test: for each octant the neighbourhood bits
 * that set code bits 7..1 (bit 0, the centre, is always set), in the order
 * of index_octant_SWU, SEU, NWU, NEU, SWB, SEB, NWB, NEB (lee_thinning.h:185-283). */
static const unsigned char o_octant_bits[8][7] = {
    {24, 25, 15, 16, 21, 22, 12}, {26, 23, 17, 14, 25, 22, 16},
    {18, 21, 9, 12, 19, 22, 10}, {20, 23, 19, 22, 11, 14, 10},
    {6, 15, 7, 16, 3, 12, 4}, {8, 7, 17, 16, 5, 4, 14},
    {0, 9, 3, 12, 1, 10, 4}, {2, 1, 11, 10, 5, 4, 14}};

/* is_euler_invariant (lee_thinning.h:288-299): sum of the octant LUT values. */
int oracle_thin_euler_lut_sum(unsigned code) {
    int s = 0;
    for (int o = 0; o < 8; o++) {
        unsigned idx = 1;
        for (int k = 0; k < 7; k++)
            if (code >> o_octant_bits[o][k] & 1u) idx |= 128u >> k;
        s += o_euler_lut[idx];
    }
    return s;
}

/* Independent route: change of the Euler characteristic V - E + F - C of the
 * union of closed unit cubes when the centre cube is removed. Only cells of
 * the centre cube that no other set voxel's cube contains disappear. */
int oracle_thin_delta_euler(unsigned code) {
    int d = -1; /* the centre cube itself */
    for (int sx = -1; sx <= 1; sx++)
        for (int sy = -1; sy <= 1; sy++)
            for (int sz = -1; sz <= 1; sz++) {
                /* a cell of the centre cube is identified by which axes it sits on a face of:
                 * s = 0 means the cell spans that axis, +-1 the side. Voxels sharing the cell
                 * are those offset by 0 or s on each axis. */
                const int dim = (sx == 0) + (sy == 0) + (sz == 0);
                if (dim == 3) continue;
                int covered = 0;
                for (int ox = (sx < 0 ? sx : 0); ox <= (sx > 0 ? sx : 0) && !covered; ox++)
                    for (int oy = (sy < 0 ? sy : 0); oy <= (sy > 0 ? sy : 0) && !covered; oy++)
                        for (int oz = (sz < 0 ? sz : 0); oz <= (sz > 0 ? sz : 0) && !covered; oz++) {
                            if (!ox && !oy && !oz) continue;
                            const int n = 9 * (oz + 1) + 3 * (oy + 1) + (ox + 1);
                            if (code >> n & 1u) covered = 1;
                        }
                if (!covered) d += (dim % 2 == 0) ? 1 : -1;
            }
    return d;
}

/* Octree labelling (lee_thinning.h:304-518) restated as a flood fill over
 * the octant membership it encodes: two neighbours are connected when an
 * octant (2x2x2 block with the centre) holds both. is_simple_point
 * (lee_thinning.h:530-579): at most one component among the 26 neighbours. */
static void o_octree_fill(int oct, int label, int* cube) {
    for (int n = 0; n < 27; n++) {
        if (n == 13) continue;
        const int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
        const int ox = oct & 1, oy = (oct >> 1) & 1, oz = (oct >> 2) & 1; /* 0: minus side, 1: plus side */
        if ((ox ? dx < 0 : dx > 0) || (oy ? dy < 0 : dy > 0) || (oz ? dz < 0 : dz > 0)) continue;
        const int c = n < 13 ? n : n - 1;
        if (cube[c] != 1) continue;
        cube[c] = label;
        for (int o2 = 0; o2 < 8; o2++) {
            if (o2 == oct) continue;
            const int px = o2 & 1, py = (o2 >> 1) & 1, pz = (o2 >> 2) & 1;
            if ((px ? dx < 0 : dx > 0) || (py ? dy < 0 : dy > 0) || (pz ? dz < 0 : dz > 0)) continue;
            o_octree_fill(o2, label, cube);
        }
    }
}

int oracle_thin_simple_octree(unsigned code) {
    int cube[26];
    for (int n = 0, c = 0; n < 27; n++)
        if (n != 13) cube[c++] = (int)(code >> n & 1u);
    int label = 2;
    for (int c = 0; c < 26; c++) {
        if (cube[c] != 1) continue;
        const int n = c < 13 ? c : c + 1;
        const int dx = n % 3 - 1, dy = (n / 3) % 3 - 1, dz = n / 9 - 1;
        const int oct = (dx > 0) | ((dy > 0) << 1) | ((dz > 0) << 2); /* an octant holding it */
        o_octree_fill(oct, label, cube);
        label++;
        if (label - 2 >= 2) return 0;
    }
    return 1;
}

/* Independent route: breadth-first 26-connected components. */
int oracle_thin_simple_bfs(unsigned code) {
    unsigned obj = code & ~(1u << 13) & 0x7FFFFFFu;
    int comps = 0;
    while (obj) {
        int stack[27], sp = 0;
        int s = __builtin_ctz(obj);
        obj &= ~(1u << s);
        stack[sp++] = s;
        while (sp) {
            const int n = stack[--sp];
            const int x = n % 3, y = (n / 3) % 3, z = n / 9;
            for (int m = 0; m < 27; m++) {
                if (!(obj >> m & 1u)) continue;
                const int mx = m % 3, my = (m / 3) % 3, mz = m / 9;
                if (abs(mx - x) <= 1 && abs(my - y) <= 1 && abs(mz - z) <= 1) {
                    obj &= ~(1u << m);
                    stack[sp++] = m;
                }
            }
        }
        if (++comps >= 2) return 0;
    }
    return 1;
}

/* Exhaustive comparison over all 2^26 neighbourhoods (centre set):
 * out[0] = #codes where (LUT sum == 0) != (delta Euler == 0),
 * out[1] = #codes where octree and BFS simple tests differ. */
void oracle_thin_exhaustive(unsigned long long* out, int nthreads) {
    unsigned long long m0 = 0, m1 = 0;
#pragma omp parallel for schedule(static, 65536) reduction(+ : m0, m1) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long k = 0; k < (1ll << 26); k++) {
        const unsigned low = (unsigned)(k & 0x1FFF), high = (unsigned)(k >> 13);
        const unsigned code = low | (1u << 13) | (high << 14);
        m0 += (oracle_thin_euler_lut_sum(code) == 0) != (oracle_thin_delta_euler(code) == 0);
        m1 += oracle_thin_simple_octree(code) != oracle_thin_simple_bfs(code);
    }
    out[0] = m0; out[1] = m1;
}

static unsigned o_thin_code(const unsigned char* v, int s0, int s1, int s2, int x, int y, int z) {
    unsigned code = 0;
    for (int dz = -1; dz <= 1; dz++)
        for (int dy = -1; dy <= 1; dy++)
            for (int dx = -1; dx <= 1; dx++) {
                const int X = x + dx, Y = y + dy, Z = z + dz;
                if (X < 0 || Y < 0 || Z < 0 || X >= s2 || Y >= s1 || Z >= s0) continue;
                if (v[((size_t)Z * s1 + Y) * s2 + X] == 1) code |= 1u << (9 * (dz + 1) + 3 * (dy + 1) + dx + 1);
            }
    return code;
}

/* computeThinImage (lee_thinning.h:600-680) on a 0/1 volume, s2 (x) fastest,
 * s1 = y, s0 = z. Border directions 1..6 = N (y-1), S (y+1), E (x+1),
 * W (x-1), U (z+1), B (z-1). Candidates of one direction are collected on
 * the unchanged volume, then re-checked for simplicity and deleted one by
 * one in raster order. Returns the number of outer iterations. */
int oracle_lee_thin3(unsigned char* v, unsigned S0, unsigned S1, unsigned S2) {
    const int s0 = (int)S0, s1 = (int)S1, s2 = (int)S2;
    const size_t n = (size_t)s0 * s1 * s2;
    for (size_t i = 0; i < n; i++) v[i] = v[i] != 0; /* lee(): binarise */
    size_t cap = 1024, cnt = 0;
    size_t* list = (size_t*)malloc(cap * sizeof(size_t));
    static const int dirs[6][3] = {{0, -1, 0}, {0, 1, 0}, {1, 0, 0}, {-1, 0, 0}, {0, 0, 1}, {0, 0, -1}};
    int unchanged = 0, iterations = 0;
    while (unchanged < 6) {
        unchanged = 0;
        iterations++;
        for (int b = 0; b < 6; b++) {
            int nochange = 1;
            cnt = 0;
            for (int z = 0; z < s0; z++)
                for (int y = 0; y < s1; y++)
                    for (int x = 0; x < s2; x++) {
                        const size_t i = ((size_t)z * s1 + y) * s2 + x;
                        if (v[i] != 1) continue;
                        const int X = x + dirs[b][0], Y = y + dirs[b][1], Z = z + dirs[b][2];
                        const int nb = (X < 0 || Y < 0 || Z < 0 || X >= s2 || Y >= s1 || Z >= s0)
                                           ? 0 : v[((size_t)Z * s1 + Y) * s2 + X];
                        if (nb > 0) continue;
                        const unsigned code = o_thin_code(v, s0, s1, s2, x, y, z);
                        if (__builtin_popcount(code) - 1 == 1) continue; /* is_endpoint */
                        if (oracle_thin_euler_lut_sum(code) != 0) continue;
                        if (!oracle_thin_simple_octree(code)) continue;
                        if (cnt == cap) { cap *= 2; list = (size_t*)realloc(list, cap * sizeof(size_t)); }
                        list[cnt++] = i;
                    }
            for (size_t k = 0; k < cnt; k++) {
                const size_t i = list[k];
                const int x = (int)(i % s2), y = (int)((i / s2) % s1), z = (int)(i / ((size_t)s1 * s2));
                if (oracle_thin_simple_octree(o_thin_code(v, s0, s1, s2, x, y, z))) {
                    v[i] = 0;
                    nochange = 0;
                }
            }
            if (nochange) unchanged++;
        }
    }
    free(list);
    return iterations;
}

/* Flags of codes first .. first+count-1, enumerated with the centre set
 * (code = low 13 bits | 1 << 13 | high << 14): bit 0 Euler invariant (LUT),
 * bit 1 simple (octree), bit 2 end point. */
void oracle_thin_classify(unsigned first, unsigned count, unsigned char* flags, int nthreads) {
#pragma omp parallel for schedule(static, 65536) num_threads(nthreads > 0 ? nthreads : 1)
    for (long long k = 0; k < (long long)count; k++) {
        const unsigned e = first + (unsigned)k;
        const unsigned code = (e & 0x1FFFu) | (1u << 13) | ((e >> 13) << 14);
        flags[k] = (unsigned char)((oracle_thin_euler_lut_sum(code) == 0 ? 1 : 0) |
                                   (oracle_thin_simple_octree(code) ? 2 : 0) |
                                   (__builtin_popcount(code) == 2 ? 4 : 0));
    }
}
