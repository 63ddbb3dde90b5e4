// vote3.hip -- 3D tensor voting (stick + numerically integrated plate, and
// analytical tensor voting) on gfx950, replacing tira::cuda::tensorvote3 /
// _tensorvote3_tk / _tensorvote3_atv (tensorvote.h:2412-2873).
//
// Pipeline of one call:
//  1. k_vote3_prep: eigen-decompose every voxel (FP32, FP64 for near-
//     degenerate tensors) and write a zero-padded voter payload:
//       stick  float4(e2, copysign(|l2|-|l1|, l2))    (tensorvote.h:2438-2442)
//       plate  float4(e0, copysign(|l1|-|l0|, l1))    (tensorvote.h:2609-2616)
//     The padding (r voxels plus tile round-up on every side) makes the vote
//     loop branch-free: out-of-volume voters carry scale 0 and contribute 0,
//     which is what the reference's bounds test skips.
//  2. one fused stick+plate gather kernel over receivers:
//       * a wave owns 64 consecutive x receivers; each lane keeps BZ
//         receivers along z in registers and reuses every voter payload load
//         for the BZ receivers it reaches (register blocking in z);
//       * the per-offset constants (direction, attenuation, plate
//         polynomial coefficients, both normalisations folded in) come from
//         a host-built table read with wave-uniform (scalar) loads;
//       * the plate vote is either the closed form of the reference's
//         N-sample integral (exact once N >= p+2, see plate_pair below) or
//         the literal N-sample sum (any N, the reference's algorithm);
//       * blocks are numbered XCD-major so neighbouring tiles -- which share
//         halo voters -- land on the same XCD's L2.
#include "common.hpp"
#include "eig_device.hpp"
#include "mlp.hpp"
#include "vote3.hpp"
#include "voxel_payload.hpp"

#include <cmath>
#include <vector>

namespace th {

constexpr int kTileX = 64;      // receivers along x per wave
constexpr int kTileY = 4;       // waves per block (along y)
constexpr int kThreads = kTileX * kTileY;
constexpr int kAtvStride = 24;  // floats per ATV offset entry

// ---------------------------------------------------------------- tables
static double binom(int n, int k) {
    double r = 1;
    for (int i = 1; i <= k; i++) r = r * (n - k + i) / i;
    return r;
}
static double gk(int k) { return binom(2 * k, k) / std::pow(4.0, k); }  // (2k-1)!!/(2k)!!

static double factorial(unsigned n) {
    double f = 1;
    for (unsigned i = 1; i <= n; i++) f *= i;
    return f;
}

// sticknorm3 (tensorvote.h:1162-1172), evaluated as the reference does (float).
float sticknorm3_host(float s1, float s2, unsigned p) {
    const float pi_term = (float)(M_PI * std::sqrt(M_PI) / 2.0);
    const float s13 = s1 * s1 * s1, s23 = s2 * s2 * s2;
    const float num1 = (float)(std::pow(2.0, 2 * p + 1) * factorial(p) * factorial(p));
    const float den1 = (float)factorial(2 * p + 1);
    const float den2 = (float)(2 * p + 1);
    return pi_term * ((s13 * num1 / den1) + (s23 * 2.0f / den2));
}

// Offset direction and attenuations exactly as cpu::build_neighbors3d
// (tensorvote.h:1411-1436) computes them in float.
static void offset_geometry(int du, int dv, int dw, float s1, float s2, float d[3], float& c1, float& c2) {
    const float is1 = (s1 > 0) ? 1.0f / (s1 * s1) : 0.0f;
    const float is2 = (s2 > 0) ? 1.0f / (s2 * s2) : 0.0f;
    const float l2 = (float)(du * du + dv * dv + dw * dw);
    if (l2 == 0.0f) { d[0] = d[1] = d[2] = 0.0f; }
    else {
        const float L = std::sqrt(l2);
        d[0] = (float)du / L; d[1] = (float)dv / L; d[2] = (float)dw / L;
    }
    c1 = s1 == 0.0f ? (l2 > 1e-12f ? 0.0f : 1.0f) : std::exp(-l2 * is1);
    c2 = s2 == 0.0f ? (l2 > 1e-12f ? 0.0f : 1.0f) : std::exp(-l2 * is2);
}

// Closed form of the plate integral: for a plate voter with unit normal n at
// unit offset d, s = n.d, x = 1 - s^2,
//   int_0^pi eta(q.d) r r^T dbeta = alpha(x) I + beta(x) n n^T
//                                   + s gamma(x) (n d^T + d n^T) + delta(x) d d^T
// with q on the unit circle orthogonal to n, r = q - 2 (q.d) d and
// eta = c1 (1-t^2)^p + c2 t^(2p). alpha..delta are polynomials in x whose
// coefficients follow from the circle moments int cos^(2k) = pi (2k-1)!!/(2k)!!.
// The N-point Riemann sum of the reference (pi/N sum over beta_i = i pi/N,
// tensorvote.h:1676-1689) integrates trigonometric polynomials of degree
// 2p+2 exactly once N >= p+2, so it equals this closed form there.
void plate_polys(int P, double c1, double c2, std::vector<double>& al, std::vector<double>& be,
                        std::vector<double>& ga, std::vector<double>& de) {
    std::vector<double> e(P + 1), A(P + 3, 0.0), B(P + 3, 0.0), C(P + 3, 0.0), D(P + 3, 0.0);
    for (int j = 0; j <= P; j++) e[j] = c1 * binom(P, j) * ((j & 1) ? -1.0 : 1.0) + (j == P ? c2 : 0.0);
    for (int j = 0; j <= P; j++) {
        A[j] += M_PI * e[j] * (gk(j) - gk(j + 1));
        if (j >= 1) B[j - 1] += M_PI * e[j] * (2 * gk(j + 1) - gk(j));
        C[j] += -2 * M_PI * e[j] * gk(j + 1);
        D[j + 1] += 4 * M_PI * e[j] * gk(j + 1);
    }
    al.assign(P + 1, 0.0); be.assign(P + 1, 0.0); ga.assign(P + 1, 0.0); de.assign(P + 2, 0.0);
    for (int k = 0; k <= P; k++) {
        al[k] = A[k];
        be[k] = B[k] - (k > 0 ? B[k - 1] : 0.0) - A[k];   // B (1 - x) - A
        ga[k] = -(B[k] + C[k]);
    }
    for (int k = 0; k <= P + 1; k++) de[k] = 0.5 * (B[k] + 2 * C[k] + D[k]);  // halved: see plate_pair
}

int closed_stride(int P) { return (4 * P + 13 + 3) & ~3; }

void build_closed_table(int P, int w, float s1, float s2, unsigned power, std::vector<float>& tab) {
    const int hw = w / 2, stride = closed_stride(P);
    tab.assign((size_t)w * w * w * stride, 0.0f);
    const double sn = 1.0 / (double)sticknorm3_host(s1, s2, power);
    const float snf = 1.0f / sticknorm3_host(s1, s2, power);
    const double pn = 1.0 / M_PI;
    std::vector<double> al, be, ga, de;
    size_t k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++, k++) {
                float d[3], c1, c2;
                offset_geometry(du, dv, dw, s1, s2, d, c1, c2);
                float* t = &tab[k * stride];
                t[0] = d[0]; t[1] = d[1]; t[2] = d[2];
                t[3] = 2 * d[0]; t[4] = 2 * d[1]; t[5] = 2 * d[2];
                if (P == 1) { t[6] = snf * c1; t[7] = (float)(sn * ((double)c2 - (double)c1)); }
                else { t[6] = snf * c1; t[7] = snf * c2; }
                if (du == 0 && dv == 0 && dw == 0) {
                    // Self offset: t = q.d = 0 for every sample, so the plate is
                    // c1 int q q^T = c1 (pi/2)(I - n n^T) -- constant polynomials
                    // (the general form assumes |d| = 1).
                    al.assign(P + 1, 0.0); be.assign(P + 1, 0.0); ga.assign(P + 1, 0.0); de.assign(P + 2, 0.0);
                    al[0] = 0.5 * M_PI * c1;
                    be[0] = -0.5 * M_PI * c1;
                } else {
                    plate_polys(P, c1, c2, al, be, ga, de);
                }
                int o = 8;
                for (int i = 0; i <= P; i++) t[o++] = (float)(pn * al[i]);
                for (int i = 0; i <= P; i++) t[o++] = (float)(pn * be[i]);
                for (int i = 0; i <= P; i++) t[o++] = (float)(pn * ga[i]);
                for (int i = 0; i <= P + 1; i++) t[o++] = (float)(pn * de[i]);
            }
}

// Offset table of the tabulated-network vote, kMlpTabStride floats per
// offset: d (3), 2 d (3), the stick constants as in the closed table (2),
// c1 / pi (c1 as cpu::build_neighbors3d computes it) and the offset's
// distance row of the network table (mlp_shell_row).
void build_mlptab_table(int P, int w, float s1, float s2, unsigned power, std::vector<float>& tab) {
    const int hw = w / 2;
    tab.assign((size_t)w * w * w * kMlpTabStride, 0.0f);
    const double sn = 1.0 / (double)sticknorm3_host(s1, s2, power);
    const float snf = 1.0f / sticknorm3_host(s1, s2, power);
    std::vector<int> shells;
    mlp_shells(s1, shells);
    size_t k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++, k++) {
                float d[3], c1, c2;
                offset_geometry(du, dv, dw, s1, s2, d, c1, c2);
                float* t = &tab[k * kMlpTabStride];
                t[0] = d[0]; t[1] = d[1]; t[2] = d[2];
                t[3] = 2 * d[0]; t[4] = 2 * d[1]; t[5] = 2 * d[2];
                if (P == 1) { t[6] = snf * c1; t[7] = (float)(sn * ((double)c2 - (double)c1)); }
                else { t[6] = snf * c1; t[7] = snf * c2; }
                t[8] = (float)(c1 / M_PI);
                t[9] = (float)mlp_shell_row(shells, du * du + dv * dv + dw * dw);
            }
}

// Literal-sum table: d, stick c1/c2 (sticknorm folded), plate c1/c2.
void build_sum_table(int w, float s1, float s2, unsigned power, std::vector<float>& tab) {
    const int hw = w / 2;
    tab.assign((size_t)w * w * w * kSumStride, 0.0f);
    const float snf = 1.0f / sticknorm3_host(s1, s2, power);
    size_t k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++, k++) {
                float d[3], c1, c2;
                offset_geometry(du, dv, dw, s1, s2, d, c1, c2);
                float* t = &tab[k * kSumStride];
                t[0] = d[0]; t[1] = d[1]; t[2] = d[2];
                t[3] = snf * c1; t[4] = snf * c2;
                t[5] = c1; t[6] = c2; t[7] = 0.0f;
            }
}

// ATV table: unit offset v, attenuation c = exp(-|d|^2 / sigma^2) (delta at
// sigma = 0, atv_attenuation tensorvote.h:34-40) and the c-scaled factors of
// the collapsed vote (see atv3_pair); offsets in the order dw, dv, du.
void build_atv3_table(int w, float sigma, std::vector<float>& tab) {
    const int hw = w / 2;
    tab.assign((size_t)w * w * w * kAtvStride, 0.0f);
    size_t k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++, k++) {
                const double l2 = (double)du * du + (double)dv * dv + (double)dw * dw;
                const double L = std::sqrt(l2);
                double v[3] = {0.0, 0.0, 0.0};
                if (l2 > 0.0) { v[0] = du / L; v[1] = dv / L; v[2] = dw / L; }
                double c;
                if (sigma == 0.0f) c = L > 1e-12 ? 0.0 : 1.0;
                else c = std::exp(-l2 / ((double)sigma * sigma));
                const double rho = l2 > 0.0 ? 1.0 : 0.0, k4 = 1.0 - rho / 4.0, k5 = 1.0 - rho / 5.0;
                const double f[21] = {v[0], v[1], v[2], c, -c, -0.75 * c, -c * k4, 0.25 * c, c * k4, c * k5,
                                      -3.0 * c, -0.5 * c, -0.4 * c, 3.0 * c, 2.0 * c,
                                      v[0] * v[0], v[1] * v[1], v[2] * v[2],
                                      2.0 * v[0] * v[1], 2.0 * v[0] * v[2], 2.0 * v[1] * v[2]};
                float* t = &tab[k * kAtvStride];
                for (int i = 0; i < 21; i++) t[i] = (float)f[i];
            }
}

// ---------------------------------------------------------------- device pair kernels
struct Acc {
    float xx, xy, xz, yy, yz, zz, ii;
};

template <int N>
__device__ __forceinline__ float ipow(float x) {
    float r = x;
#pragma unroll
    for (int i = 1; i < N; i++) r *= x;
    return r;
}

template <int N>  // c[0] + c[1] x + ... + c[N] x^N
__device__ __forceinline__ float horner(const float* c, float x) {
    float r = c[N];
#pragma unroll
    for (int i = N - 1; i >= 0; i--) r = fmaf(r, x, c[i]);
    return r;
}

// Stick vote of voter S = (q, scale) at offset entry e (stickvote3_accumulate_kernel,
// tensorvote.h:1183-1213, sticknorm folded into e[6..7]; e[3..5] = 2 d, so the
// reflection q - 2 t d takes one FMA per component).
template <int P>
__device__ __forceinline__ void stick_pair(Acc& a, const float4 S, const float* __restrict__ e) {
    const float d0 = e[0], d1 = e[1], d2 = e[2];
    const float t = fmaf(S.x, d0, fmaf(S.y, d1, S.z * d2));
    float eta;
    if (P == 1) eta = fmaf(t, t * e[7], e[6]);
    else {
        const float t2 = t * t;
        eta = fmaf(e[6], ipow<P>(1.0f - t2), e[7] * ipow<P>(t2));
    }
    const float f = eta * S.w;
    const float r0 = fmaf(-t, e[3], S.x), r1 = fmaf(-t, e[4], S.y), r2 = fmaf(-t, e[5], S.z);
    const float f0 = f * r0, f1 = f * r1, f2 = f * r2;
    a.xx = fmaf(f0, r0, a.xx); a.xy = fmaf(f0, r1, a.xy); a.xz = fmaf(f0, r2, a.xz);
    a.yy = fmaf(f1, r1, a.yy); a.yz = fmaf(f1, r2, a.yz); a.zz = fmaf(f2, r2, a.zz);
}

// Accumulates w [A I + B n n^T + G (n d^T + d n^T) + 2 D d d^T] for the voter
// normal n = N.xyz, scale w = N.w and the offset entry e (d = e[0..2],
// 2 d = e[3..5]): w G (n d^T + d n^T) + 2 w D d d^T = h d^T + d h^T with
// h = G nw + w D d. nw = w n is computed once per voter (voter_nw): the scale
// then costs two products per pair instead of five.
__device__ __forceinline__ float3 voter_nw(const float4 N) { return make_float3(N.w * N.x, N.w * N.y, N.w * N.z); }

__device__ __forceinline__ void plate_accumulate(Acc& a, const float4 N, const float3 nw, const float* __restrict__ e,
                                                 float A, float B, float G, float D) {
    const float d0 = e[0], d1 = e[1], d2 = e[2];
    a.ii = fmaf(N.w, A, a.ii);
    const float b0 = B * nw.x, b1 = B * nw.y, b2 = B * nw.z;
    a.xx = fmaf(b0, N.x, a.xx); a.xy = fmaf(b0, N.y, a.xy); a.xz = fmaf(b0, N.z, a.xz);
    a.yy = fmaf(b1, N.y, a.yy); a.yz = fmaf(b1, N.z, a.yz); a.zz = fmaf(b2, N.z, a.zz);
    const float Dw = N.w * D;
    const float h0 = fmaf(Dw, d0, G * nw.x), h1 = fmaf(Dw, d1, G * nw.y), h2 = fmaf(Dw, d2, G * nw.z);
    a.xx = fmaf(h0, e[3], a.xx); a.yy = fmaf(h1, e[4], a.yy); a.zz = fmaf(h2, e[5], a.zz);
    a.xy = fmaf(h0, d1, fmaf(h1, d0, a.xy));
    a.xz = fmaf(h0, d2, fmaf(h2, d0, a.xz));
    a.yz = fmaf(h1, d2, fmaf(h2, d1, a.yz));
}

// Plate vote of voter N = (n, scale) at offset entry e, closed form:
//   scale * [alpha I + beta n n^T + gamma (n d^T + d n^T) + delta d d^T]
// (the table holds delta / 2, see plate_accumulate); nw = voter_nw(N).
template <int P>
__device__ __forceinline__ void plate_pair(Acc& a, const float4 N, const float3 nw, const float* __restrict__ e) {
    const float s = fmaf(N.x, e[0], fmaf(N.y, e[1], N.z * e[2]));
    const float x = fmaf(-s, s, 1.0f);
    const float* al = e + 8;
    const float* be = al + (P + 1);
    const float* ga = be + (P + 1);
    const float* de = ga + (P + 1);
    plate_accumulate(a, N, nw, e, horner<P>(al, x), horner<P>(be, x), s * horner<P>(ga, x), horner<P + 1>(de, x));
}

template <int P>
__device__ __forceinline__ void plate_pair(Acc& a, const float4 N, const float* __restrict__ e) {
    plate_pair<P>(a, N, voter_nw(N), e);
}

// Network-table row as the tabulated vote stages it in LDS: kMlpTabPad cells
// of (value, difference to the next cell), cell p holding the stored cell
// p - 1 scaled by c1 / pi of the offset; the first and the last cell continue
// the end segments linearly, so that the lookup needs no clamp.
constexpr int kMlpTabPad = kMlpTabM + 1;

// Linear interpolation in a staged row at x = 1 - (n.d)^2 in [0, 1]:
// t = 64 x + 1/2 lies in [1/2, 64 + 1/2], its cell in [0, 64] (|n.d| <= 1
// within rounding keeps t above zero, where truncation is the floor).
__device__ __forceinline__ float4 mlp_tab_lookup_staged(const float4* __restrict__ row, float x) {
    const float t = fmaf(x, (float)kMlpTabM, 0.5f);
    const int p = (int)t;
    const float f = __builtin_amdgcn_fractf(t);
    const float4 v = row[2 * p], dl = row[2 * p + 1];
    return make_float4(fmaf(f, dl.x, v.x), fmaf(f, dl.y, v.y), fmaf(f, dl.z, v.z), fmaf(f, dl.w, v.w));
}

// Plate vote of voter N = (n, scale) from the tabulated network: the staged
// row of the offset's distance, looked up at x = 1 - (n.d)^2, gives
// c1 / pi (A, B, g, D/2) of the offset.
__device__ __forceinline__ void tab_plate_pair(Acc& a, const float4 N, const float3 nw, const float* __restrict__ e,
                                               const float4* row) {
    const float s = fmaf(N.x, e[0], fmaf(N.y, e[1], N.z * e[2]));
    const float x = fmaf(-s, s, 1.0f);
    const float4 v = mlp_tab_lookup_staged(row, x);
    plate_accumulate(a, N, nw, e, v.x, v.y, s * v.z, v.w);
}

// Literal N-sample plate sum (plate3_numerical_device, tensorvote.h:2457-2485):
// U = (u, scale/N), V = (v, 0) is the voter's plate basis. Samples are summed
// into a per-pair tensor first and scaled into the receiver afterwards, as the
// reference does: two-level summation keeps FP32 error at the level of the
// closed form instead of growing with N * w^3 terms.
__device__ __forceinline__ void plate_sum_pair(Acc& a, const float4 U, const float4 V, const float* __restrict__ e,
                                               const float2* __restrict__ cs, int nsamp, int power) {
    const float d0 = e[0], d1 = e[1], d2 = e[2], c1 = e[5], c2 = e[6];
    const float k = c2 - c1;
    float v00 = 0.0f, v01 = 0.0f, v02 = 0.0f, v11 = 0.0f, v12 = 0.0f, v22 = 0.0f;
    for (int i = 0; i < nsamp; i++) {
        const float2 sc = cs[i];  // (cos beta_i, sin beta_i)
        const float q0 = fmaf(sc.x, U.x, sc.y * V.x), q1 = fmaf(sc.x, U.y, sc.y * V.y),
                    q2 = fmaf(sc.x, U.z, sc.y * V.z);
        const float t = fmaf(q0, d0, fmaf(q1, d1, q2 * d2));
        const float t2 = t * t;
        float eta;
        if (power == 1) eta = fmaf(k, t2, c1);
        else {
            const float t1 = 1.0f - t2;
            float e1 = t1, e2 = t2;
            for (int j = 1; j < power; j++) { e1 *= t1; e2 *= t2; }
            eta = fmaf(c1, e1, c2 * e2);
        }
        const float m2t = -2.0f * t;
        const float r0 = fmaf(m2t, d0, q0), r1 = fmaf(m2t, d1, q1), r2 = fmaf(m2t, d2, q2);
        const float f0 = eta * r0, f1 = eta * r1, f2 = eta * r2;
        v00 = fmaf(f0, r0, v00); v01 = fmaf(f0, r1, v01); v02 = fmaf(f0, r2, v02);
        v11 = fmaf(f1, r1, v11); v12 = fmaf(f1, r2, v12); v22 = fmaf(f2, r2, v22);
    }
    const float w = U.w;
    a.xx = fmaf(w, v00, a.xx); a.xy = fmaf(w, v01, a.xy); a.xz = fmaf(w, v02, a.xz);
    a.yy = fmaf(w, v11, a.yy); a.yz = fmaf(w, v12, a.yz); a.zz = fmaf(w, v22, a.zz);
}

// Stick pair for the literal path (runtime power).
__device__ __forceinline__ void stick_pair_rt(Acc& a, const float4 S, const float* __restrict__ e, int power) {
    const float d0 = e[0], d1 = e[1], d2 = e[2];
    const float t = fmaf(S.x, d0, fmaf(S.y, d1, S.z * d2));
    const float t2 = t * t;
    const float t1 = 1.0f - t2;
    float e1 = t1, e2 = t2;
    for (int j = 1; j < power; j++) { e1 *= t1; e2 *= t2; }
    const float f = fmaf(e[3], e1, e[4] * e2) * S.w;
    const float m2t = -2.0f * t;
    const float r0 = fmaf(m2t, d0, S.x), r1 = fmaf(m2t, d1, S.y), r2 = fmaf(m2t, d2, S.z);
    const float f0 = f * r0, f1 = f * r1, f2 = f * r2;
    a.xx = fmaf(f0, r0, a.xx); a.xy = fmaf(f0, r1, a.xy); a.xz = fmaf(f0, r2, a.xz);
    a.yy = fmaf(f1, r1, a.yy); a.yz = fmaf(f1, r2, a.yz); a.zz = fmaf(f2, r2, a.zz);
}

__device__ __forceinline__ void write_out(float* __restrict__ out, int layout, size_t v, const Acc& a,
                                          bool add = false) {
    const float xx = a.xx + a.ii, yy = a.yy + a.ii, zz = a.zz + a.ii;
    if (layout == TIRA_LAYOUT_MAT9) {
        float* o = out + v * 9;
        const float t[9] = {xx, a.xy, a.xz, a.xy, yy, a.yz, a.xz, a.yz, zz};
        for (int k = 0; k < 9; k++) o[k] = add ? o[k] + t[k] : t[k];
    } else {
        float* o = out + v * 6;
        const float t[6] = {xx, a.xy, a.xz, yy, a.yz, zz};
        for (int k = 0; k < 6; k++) o[k] = add ? o[k] + t[k] : t[k];
    }
}

__device__ __forceinline__ void tile_coords(const VoteGeom& g, int bz_tile, unsigned& bx, unsigned& by,
                                            unsigned& bz) {
    const unsigned nb = g.nbx * g.nby * g.nbz;
    const unsigned b = xcd_remap(blockIdx.x, nb);
    bx = b % g.nbx;
    by = (b / g.nbx) % g.nby;
    bz = b / (g.nbx * g.nby);
    (void)bz_tile;
}

// The same with z fastest: the tiles stacked along z, which share most of
// their voter planes, run on one XCD at the same time.
__device__ __forceinline__ void tile_coords_zfast(const VoteGeom& g, unsigned& bx, unsigned& by, unsigned& bz) {
    const unsigned nb = g.nbx * g.nby * g.nbz;
    const unsigned b = xcd_remap(blockIdx.x, nb);
    bz = b % g.nbz;
    bx = (b / g.nbz) % g.nbx;
    by = b / (g.nbz * g.nbx);
}

// Closed-form stick + plate gather. Payloads are padded volumes of
// (P0, P1, P2) voxels; receiver (z, y, x) reads voter (z+dw+r, y+dv+r, x+du+r).
// gate (optional): the launch does nothing while *gate == 0; accumulate:
// add the votes to `out` instead of writing them.
template <int P, bool STICK, bool PLATE, int BZ>
__global__ __launch_bounds__(kThreads) void k_vote3_closed(const float4* __restrict__ pay_s,
                                                           const float4* __restrict__ pay_p,
                                                           const float* __restrict__ tab, VoteGeom g,
                                                           float* __restrict__ out, int out_layout,
                                                           const unsigned long long* __restrict__ gate,
                                                           int accumulate) {
    constexpr int STRIDE = (4 * P + 13 + 3) & ~3;
    if (gate && *gate == 0ull) return;
    unsigned bx, by, bz;
    tile_coords_zfast(g, bx, by, bz);
    const int x = (int)bx * kTileX + (int)(threadIdx.x % kTileX);
    const int y = (int)by * kTileY + (int)(threadIdx.x / kTileX);
    const int z0 = (int)bz * BZ;
    const int r = g.r, w = 2 * g.r + 1;
    const size_t plane = (size_t)g.P1 * g.P2;

    Acc acc[BZ];
#pragma unroll
    for (int j = 0; j < BZ; j++) acc[j] = Acc{0, 0, 0, 0, 0, 0, 0};

    // Offsets in the order (dv, voter plane, du): consecutive du steps read
    // the same voter row shifted by one voxel, so a wave's cache lines are
    // reused at once instead of after a whole z column of the window.
    for (int dv = -r; dv <= r; dv++) {
        const size_t rowb = ((size_t)z0 * g.P1 + (size_t)(y + r + dv)) * g.P2 + (size_t)(x + r);
        for (int zz = 0; zz < BZ + 2 * r; zz++) {
            const size_t zrow = rowb + (size_t)zz * plane;
            for (int du = -r; du <= r; du++) {
                float4 S, N;
                if (STICK) S = pay_s[zrow + du];
                if (PLATE) N = pay_p[zrow + du];
                const float3 nw = PLATE ? voter_nw(N) : make_float3(0.0f, 0.0f, 0.0f);
                const int ebase = (dv + r) * w + (du + r);
#pragma unroll
                for (int j = 0; j < BZ; j++) {
                    const int dw = zz - r - j;
                    if (dw >= -r && dw <= r) {
                        const float* e = tab + (size_t)((dw + r) * w * w + ebase) * STRIDE;
                        if (STICK) stick_pair<P>(acc[j], S, e);
                        if (PLATE) plate_pair<P>(acc[j], N, nw, e);
                    }
                }
            }
        }
    }
    if (x < g.s2 && y < g.s1) {
#pragma unroll
        for (int j = 0; j < BZ; j++) {
            const int z = z0 + j;
            if (z < g.s0) write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, acc[j], accumulate != 0);
        }
    }
}

// The closed-form gather with the voters staged in LDS (the reference's
// shared-memory neighbour tile, tensorvote.h:2412-2455): a wave's receivers
// (64 x of one row y, BZ planes) read, for each (dv, voter plane), the same
// voter row segment of 64 + 2r voxels at du = -r..r. The wave stages that
// segment (stick and plate payloads) in its own LDS slice and reads it from
// there; nothing is shared between waves, so the only synchronisation is the
// wave's own (its LDS writes precede its reads in program order, and the
// compiler is kept from moving them across each other). LDS: 4 waves x
// (64 + 2r) x 32 B, sized at launch.
template <int P, bool STICK, bool PLATE, int BZ>
__global__ __launch_bounds__(kThreads) void k_vote3_closed_lds(const float4* __restrict__ pay_s,
                                                               const float4* __restrict__ pay_p,
                                                               const float* __restrict__ tab, VoteGeom g,
                                                               float* __restrict__ out, int out_layout) {
    constexpr int STRIDE = (4 * P + 13 + 3) & ~3;
    extern __shared__ float4 s_vox[];
    unsigned bx, by, bz;
    tile_coords_zfast(g, bx, by, bz);
    const int lane = (int)(threadIdx.x % kTileX), wave = (int)(threadIdx.x / kTileX);
    const int x = (int)bx * kTileX + lane;
    const int y = (int)by * kTileY + wave;
    const int z0 = (int)bz * BZ;
    const int r = g.r, w = 2 * g.r + 1, seg = kTileX + 2 * g.r;
    const size_t plane = (size_t)g.P1 * g.P2;
    float4* sS = s_vox + (size_t)wave * 2 * seg;
    float4* sN = sS + seg;

    Acc acc[BZ];
#pragma unroll
    for (int j = 0; j < BZ; j++) acc[j] = Acc{0, 0, 0, 0, 0, 0, 0};

    for (int dv = -r; dv <= r; dv++) {
        // padded voter row y + r + dv, x from the tile's first receiver: voters x + r + du
        const size_t rowb = ((size_t)z0 * g.P1 + (size_t)(y + r + dv)) * g.P2 + (size_t)bx * kTileX;
        for (int zz = 0; zz < BZ + 2 * r; zz++) {
            const size_t zrow = rowb + (size_t)zz * plane;
            for (int k = lane; k < seg; k += kTileX) {
                if (STICK) sS[k] = pay_s[zrow + k];
                if (PLATE) sN[k] = pay_p[zrow + k];
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int du = -r; du <= r; du++) {
                float4 S, N;
                if (STICK) S = sS[lane + r + du];
                if (PLATE) N = sN[lane + r + du];
                const float3 nw = PLATE ? voter_nw(N) : make_float3(0.0f, 0.0f, 0.0f);
                const int ebase = (dv + r) * w + (du + r);
#pragma unroll
                for (int j = 0; j < BZ; j++) {
                    const int dw = zz - r - j;
                    if (dw >= -r && dw <= r) {
                        const float* e = tab + (size_t)((dw + r) * w * w + ebase) * STRIDE;
                        if (STICK) stick_pair<P>(acc[j], S, e);
                        if (PLATE) plate_pair<P>(acc[j], N, nw, e);
                    }
                }
            }
            // the segment is read before the next one overwrites it
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    if (x < g.s2 && y < g.s1) {
#pragma unroll
        for (int j = 0; j < BZ; j++) {
            const int z = z0 + j;
            if (z < g.s0) write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, acc[j]);
        }
    }
}

// Literal-sum stick + plate gather (any power, any N >= 1).
template <bool STICK, bool PLATE, int BZ>
__global__ __launch_bounds__(kThreads) void k_vote3_sum(const float4* __restrict__ pay_s,
                                                        const float4* __restrict__ pay_u,
                                                        const float4* __restrict__ pay_v,
                                                        const float* __restrict__ tab,
                                                        const float2* __restrict__ cs, int nsamp, int power,
                                                        VoteGeom g, float* __restrict__ out, int out_layout) {
    unsigned bx, by, bz;
    tile_coords(g, BZ, bx, by, bz);
    const int x = (int)bx * kTileX + (int)(threadIdx.x % kTileX);
    const int y = (int)by * kTileY + (int)(threadIdx.x / kTileX);
    const int z0 = (int)bz * BZ;
    const int r = g.r, w = 2 * g.r + 1;
    const size_t plane = (size_t)g.P1 * g.P2;

    Acc acc[BZ];
#pragma unroll
    for (int j = 0; j < BZ; j++) acc[j] = Acc{0, 0, 0, 0, 0, 0, 0};

    // Offsets in the order (dv, voter plane, du), as in k_vote3_closed.
    for (int dv = -r; dv <= r; dv++) {
        const size_t rowb = ((size_t)z0 * g.P1 + (size_t)(y + r + dv)) * g.P2 + (size_t)(x + r);
        for (int zz = 0; zz < BZ + 2 * r; zz++) {
            const size_t zrow = rowb + (size_t)zz * plane;
            for (int du = -r; du <= r; du++) {
                const int ebase = (dv + r) * w + (du + r);
                float4 S, U, V;
                if (STICK) S = pay_s[zrow + du];
                if (PLATE) { U = pay_u[zrow + du]; V = pay_v[zrow + du]; }
#pragma unroll
                for (int j = 0; j < BZ; j++) {
                    const int dw = zz - r - j;
                    if (dw >= -r && dw <= r) {
                        const float* e = tab + (size_t)((dw + r) * w * w + ebase) * kSumStride;
                        if (STICK) stick_pair_rt(acc[j], S, e, power);
                        if (PLATE) plate_sum_pair(acc[j], U, V, e, cs, nsamp, power);
                    }
                }
            }
        }
    }
    if (x < g.s2 && y < g.s1) {
#pragma unroll
        for (int j = 0; j < BZ; j++) {
            const int z = z0 + j;
            if (z < g.s0) write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, acc[j]);
        }
    }
}

// ---------------------------------------------------------------- payload prep
// Plate basis of tensorvote.h:1667-1673 / 2468-2470 for a unit normal n.
__device__ __forceinline__ void plate_basis(const float n[3], float u[3], float v[3]) {
    float a0 = 0.0f, a1 = 0.0f, a2 = 1.0f;
    if (!(fabsf(n[2]) < 0.999f)) { a0 = 1.0f; a2 = 0.0f; }
    u[0] = a1 * n[2] - a2 * n[1];
    u[1] = a2 * n[0] - a0 * n[2];
    u[2] = a0 * n[1] - a1 * n[0];
    const float l2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2];
    const float inv = l2 > 0.0f ? rsqrtf(l2) : 0.0f;
    u[0] *= inv; u[1] *= inv; u[2] *= inv;
    v[0] = n[1] * u[2] - n[2] * u[1];
    v[1] = n[2] * u[0] - n[0] * u[2];
    v[2] = n[0] * u[1] - n[1] * u[0];
}

// Writes the voter payload for one voxel given its eigen-decomposition.
__device__ __forceinline__ void write_payload(const VoteGeom& g, size_t pidx, float l0, float l1, float l2,
                                              const float* e0, const float* e2, float4* pay_s, float4* pay_p,
                                              float4* pay_v, float sum_inv_n) {
    float4 S, N;
    stick_plate_payload(l0, l1, l2, e0, e2, S, N);
    if (pay_s) pay_s[pidx] = S;
    if (pay_p) {
        if (!pay_v) {
            pay_p[pidx] = N;
        } else {
            const float n[3] = {N.x, N.y, N.z};
            float u[3], v[3];
            plate_basis(n, u, v);
            pay_p[pidx] = make_float4(u[0], u[1], u[2], N.w * sum_inv_n);
            pay_v[pidx] = make_float4(v[0], v[1], v[2], 0.0f);
        }
    }
}

// Grid-stride over the padded payload volume: voxels outside the field get a
// zero payload.
__global__ __launch_bounds__(256) void k_vote3_prep(const float* __restrict__ in, int layout, VoteGeom g,
                                                    float4* __restrict__ pay_s, float4* __restrict__ pay_p,
                                                    float4* __restrict__ pay_v, float sum_inv_n) {
    const size_t total = (size_t)g.P0 * g.P1 * g.P2;
    const float4 zero = make_float4(0, 0, 0, 0);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(p % g.P2), py = (int)((p / g.P2) % g.P1), pz = (int)(p / ((size_t)g.P1 * g.P2));
        const int x = px - g.r, y = py - g.r, z = pz - g.r;
        if (x < 0 || y < 0 || z < 0 || x >= g.s2 || y >= g.s1 || z >= g.s0) {
            if (pay_s) pay_s[p] = zero;
            if (pay_p) pay_p[p] = zero;
            if (pay_v) pay_v[p] = zero;
            continue;
        }
        const size_t v = ((size_t)z * g.s1 + y) * g.s2 + x;
        float l0, l1, l2, e0[3], e2[3];
        voxel_eig(in, layout, v, l0, l1, l2, e0, e2);
        write_payload(g, p, l0, l1, l2, e0, e2, pay_s, pay_p, pay_v, sum_inv_n);
    }
}

// Payload from caller eigenpairs (_tensorvote3_tk inputs).
__global__ __launch_bounds__(256) void k_vote3_prep_tk(const float* __restrict__ lam, const float* __restrict__ ev2,
                                                       const float* __restrict__ ev0, VoteGeom g,
                                                       float4* __restrict__ pay_s, float4* __restrict__ pay_p,
                                                       float4* __restrict__ pay_v, float sum_inv_n) {
    const size_t total = (size_t)g.P0 * g.P1 * g.P2;
    const float4 zero = make_float4(0, 0, 0, 0);
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(p % g.P2), py = (int)((p / g.P2) % g.P1), pz = (int)(p / ((size_t)g.P1 * g.P2));
        const int x = px - g.r, y = py - g.r, z = pz - g.r;
        if (x < 0 || y < 0 || z < 0 || x >= g.s2 || y >= g.s1 || z >= g.s0) {
            if (pay_s) pay_s[p] = zero;
            if (pay_p) pay_p[p] = zero;
            if (pay_v) pay_v[p] = zero;
            continue;
        }
        const size_t v = ((size_t)z * g.s1 + y) * g.s2 + x;
        const float z3[3] = {0, 0, 0};
        const float* e2 = ev2 ? ev2 + v * 3 : z3;
        const float* e0 = ev0 ? ev0 + v * 3 : z3;
        write_payload(g, p, lam[v * 3], lam[v * 3 + 1], lam[v * 3 + 2], e0, e2, pay_s, pay_p, pay_v, sum_inv_n);
    }
}

// ---------------------------------------------------------------- learned plate (MLP)
// Payload of the surrogate path, per voter: N = (n, plate scale), U = (u,
// fallback flag) with u the reference's plate-basis vector (used when an
// offset runs along n); `flag` marks voters whose orientation bin failed the
// surrogate's error check -- their plate votes come from the integrator (the
// closed form of the N-sample sum). Without pay_u (the tabulated network)
// a flagged voter gets plate scale 0 in pay_n and its (n, scale) in pay_f,
// which is zero for every other voter.
__global__ __launch_bounds__(256) void k_vote3_prep_mlp(const float* __restrict__ in, int layout, VoteGeom g,
                                                        float4* __restrict__ pay_s, float4* __restrict__ pay_n,
                                                        float4* __restrict__ pay_u, float4* __restrict__ pay_f,
                                                        const float* __restrict__ bin_err, float tol,
                                                        unsigned long long* __restrict__ n_fallback) {
    const size_t total = (size_t)g.P0 * g.P1 * g.P2;
    const float4 zero = make_float4(0, 0, 0, 0);
    unsigned nf = 0;
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(p % g.P2), py = (int)((p / g.P2) % g.P1), pz = (int)(p / ((size_t)g.P1 * g.P2));
        const int x = px - g.r, y = py - g.r, z = pz - g.r;
        if (x < 0 || y < 0 || z < 0 || x >= g.s2 || y >= g.s1 || z >= g.s0) {
            if (pay_s) pay_s[p] = zero;
            pay_n[p] = zero;
            if (pay_u) pay_u[p] = zero;
            if (pay_f) pay_f[p] = zero;
            continue;
        }
        const size_t v = ((size_t)z * g.s1 + y) * g.s2 + x;
        float l0, l1, l2, e0[3], e2[3];
        voxel_eig(in, layout, v, l0, l1, l2, e0, e2);
        if (pay_s) pay_s[p] = make_float4(e2[0], e2[1], e2[2], copysignf(fabsf(l2) - fabsf(l1), l2));
        float n[3] = {e0[0], e0[1], e0[2]};
        const float nn = n[0] * n[0] + n[1] * n[1] + n[2] * n[2];
        float sp = copysignf(fabsf(l1) - fabsf(l0), l1);
        if (!(nn > 1e-12f)) { sp = 0.0f; n[0] = n[1] = n[2] = 0.0f; n[2] = 1.0f; }
        else { const float inv = rsqrtf(nn); n[0] *= inv; n[1] *= inv; n[2] *= inv; }
        float u[3], w[3];
        plate_basis(n, u, w);
        const bool fb = sp != 0.0f && bin_err[mlp_bin(n[0], n[1], n[2])] > tol;
        nf += fb ? 1u : 0u;
        if (pay_u) {
            pay_n[p] = make_float4(n[0], n[1], n[2], sp);
            pay_u[p] = make_float4(u[0], u[1], u[2], fb ? 1.0f : 0.0f);
        } else {
            // tabulated network: flagged voters leave this payload and vote
            // through the fallback payload (the integrator's pass)
            pay_n[p] = make_float4(n[0], n[1], n[2], fb ? 0.0f : sp);
            pay_f[p] = fb ? make_float4(n[0], n[1], n[2], sp) : zero;
        }
    }
    if (n_fallback && nf) atomicAdd(n_fallback, (unsigned long long)nf);
}

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2_t __attribute__((ext_vector_type(2)));
typedef float f32x16_t __attribute__((ext_vector_type(16)));
typedef float f32x2_t __attribute__((ext_vector_type(2)));

__device__ __forceinline__ unsigned pack_bf16(float a, float b) {  // v_cvt_pk_bf16_f32, element 0 low
    const bf16x2_t h = __builtin_convertvector((f32x2_t){a, b}, bf16x2_t);
    return __builtin_bit_cast(unsigned, h);
}
__device__ __forceinline__ bf16x8_t op4(unsigned a, unsigned b, unsigned c, unsigned d) {
    return __builtin_bit_cast(bf16x8_t, make_uint4(a, b, c, d));
}
// ReLU of 8 accumulator registers as MFMA operands: the BF16 high parts and
// the BF16 residuals (the activation is their sum, ~16 mantissa bits).
__device__ __forceinline__ void act8(const f32x16_t& c, int o, bf16x8_t& hi, bf16x8_t& lo) {
    unsigned h[4], l[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float r0 = fmaxf(c[o + 2 * i], 0.0f), r1 = fmaxf(c[o + 2 * i + 1], 0.0f);
        h[i] = pack_bf16(r0, r1);
        l[i] = pack_bf16(r0 - __uint_as_float(h[i] << 16), r1 - __uint_as_float(h[i] & 0xffff0000u));
    }
    hi = op4(h[0], h[1], h[2], h[3]);
    lo = op4(l[0], l[1], l[2], l[3]);
}
#define TH_MFMA(a, b, c) __builtin_amdgcn_mfma_f32_32x32x16_bf16((a), (b), (c), 0, 0, 0)

// Hidden layer of 64 units from the previous layer's 4 K-steps of high parts
// and residuals: two 32-row output tiles (weights wl[0..3], wl[4..7]).
__device__ __forceinline__ void mlp_layer(const bf16x8_t* __restrict__ wl, const bf16x8_t h[4], const bf16x8_t l[4],
                                          f32x16_t& t0, f32x16_t& t1) {
    const f32x16_t z = {};
    t0 = TH_MFMA(wl[0], h[0], z);
    t1 = TH_MFMA(wl[4], h[0], z);
    t0 = TH_MFMA(wl[0], l[0], t0);
    t1 = TH_MFMA(wl[4], l[0], t1);
#pragma unroll
    for (int s = 1; s < 4; s++) {
        t0 = TH_MFMA(wl[s], h[s], t0);
        t1 = TH_MFMA(wl[4 + s], h[s], t1);
        t0 = TH_MFMA(wl[s], l[s], t0);
        t1 = TH_MFMA(wl[4 + s], l[s], t1);
    }
}

__device__ __forceinline__ void acts(const f32x16_t& t0, const f32x16_t& t1, bf16x8_t h[4], bf16x8_t l[4]) {
    act8(t0, 0, h[0], l[0]);
    act8(t0, 8, h[1], l[1]);
    act8(t1, 0, h[2], l[2]);
    act8(t1, 8, h[3], l[3]);
}

// The 64 pairs of a wave: lanes 0-31 carry pairs 0-31 as batch A, lanes
// 32-63 pairs 32-63 as batch B. The two batches are staggered by half a
// layer, so the activations (VALU) of one overlap the matrix products of the
// other within the wave. Outputs 0-3 of a pair sit in registers 0-3 of
// lanes 0-31, outputs 4-5 in registers 0-1 of lanes 32-63 (row
// 8 (r / 4) + 4 (l / 32) + r % 4 of the tile); returns the lane's own pair.
__device__ __forceinline__ void mlp_round(const bf16x8_t* __restrict__ w, bf16x8_t xa, bf16x8_t xb, bool lo,
                                          float o[6]) {
    const f32x16_t z = {};
    f32x16_t a0 = TH_MFMA(w[0], xa, z), a1 = TH_MFMA(w[1], xa, z);
    f32x16_t b0 = TH_MFMA(w[0], xb, z), b1 = TH_MFMA(w[1], xb, z);
    bf16x8_t ha[4], la[4], hb[4], lb[4];
    acts(a0, a1, ha, la);
#pragma unroll
    for (int layer = 0; layer < 2; layer++) {
        const bf16x8_t* wl = w + 2 + 8 * layer;
        mlp_layer(wl, ha, la, a0, a1);
        acts(b0, b1, hb, lb);
        mlp_layer(wl, hb, lb, b0, b1);
        acts(a0, a1, ha, la);
    }
    f32x16_t ya = TH_MFMA(w[18], ha[0], z);
    ya = TH_MFMA(w[18], la[0], ya);
#pragma unroll
    for (int s = 1; s < 4; s++) {
        ya = TH_MFMA(w[18 + s], ha[s], ya);
        ya = TH_MFMA(w[18 + s], la[s], ya);
    }
    acts(b0, b1, hb, lb);
    f32x16_t yb = TH_MFMA(w[18], hb[0], z);
    yb = TH_MFMA(w[18], lb[0], yb);
#pragma unroll
    for (int s = 1; s < 4; s++) {
        yb = TH_MFMA(w[18 + s], hb[s], yb);
        yb = TH_MFMA(w[18 + s], lb[s], yb);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) {
        const float t = __shfl_xor(yb[k], 32);
        o[k] = lo ? ya[k] : t;
    }
#pragma unroll
    for (int k = 0; k < 2; k++) {
        const float t = __shfl_xor(ya[k], 32);
        o[4 + k] = lo ? t : yb[k];
    }
}
#undef TH_MFMA

// Rotates the local-frame vote s * [o] (uu, uv, un, vv, vn, nn) to the field
// frame, R [o] R^T with R = [u v n], and accumulates it.
__device__ __forceinline__ void mlp_accumulate(Acc& a, const float o[6], float s, const float4& U, const float4& V,
                                               const float4& N) {
    const float o0 = s * o[0], o1 = s * o[1], o2 = s * o[2], o3 = s * o[3], o4 = s * o[4], o5 = s * o[5];
    const float u[3] = {U.x, U.y, U.z}, v[3] = {V.x, V.y, V.z}, n[3] = {N.x, N.y, N.z};
    float zz[3][3];
#pragma unroll
    for (int j = 0; j < 3; j++) {
        zz[j][0] = fmaf(o0, u[j], fmaf(o1, v[j], o2 * n[j]));
        zz[j][1] = fmaf(o1, u[j], fmaf(o3, v[j], o4 * n[j]));
        zz[j][2] = fmaf(o2, u[j], fmaf(o4, v[j], o5 * n[j]));
    }
    a.xx = fmaf(u[0], zz[0][0], fmaf(v[0], zz[0][1], fmaf(n[0], zz[0][2], a.xx)));
    a.xy = fmaf(u[0], zz[1][0], fmaf(v[0], zz[1][1], fmaf(n[0], zz[1][2], a.xy)));
    a.xz = fmaf(u[0], zz[2][0], fmaf(v[0], zz[2][1], fmaf(n[0], zz[2][2], a.xz)));
    a.yy = fmaf(u[1], zz[1][0], fmaf(v[1], zz[1][1], fmaf(n[1], zz[1][2], a.yy)));
    a.yz = fmaf(u[1], zz[2][0], fmaf(v[1], zz[2][1], fmaf(n[1], zz[2][2], a.yz)));
    a.zz = fmaf(u[2], zz[2][0], fmaf(v[2], zz[2][1], fmaf(n[2], zz[2][2], a.zz)));
}

// Stick (closed table) + learned plate gather: a wave owns 64 consecutive x
// receivers of one (y, z) row and walks every window offset; each offset is
// one network round over the wave's 64 (receiver, voter) pairs, in the plate
// frame turned about n so that the offset lies in the (u, n) half-plane (the
// frame the network was trained in). The self offset and voters flagged by
// the error check take the closed-form plate.
// mtab per offset: (d, c1 / pi), (|d| / sigma1 clamped, 0, 0, 0); the network's
// standardized outputs are mapped back with the FP32 affine `aff`.
template <int P, bool STICK>
__global__ __launch_bounds__(kThreads) void k_vote3_mlp(const float4* __restrict__ pay_s,
                                                        const float4* __restrict__ pay_n,
                                                        const float4* __restrict__ pay_u,
                                                        const float* __restrict__ ctab,
                                                        const float4* __restrict__ mtab,
                                                        const uint4* __restrict__ wreg, MlpAffine aff,
                                                        VoteGeom g, float* __restrict__ out, int out_layout) {
    constexpr int CS = (4 * P + 13 + 3) & ~3;
    const int lane = threadIdx.x & 63;
    bf16x8_t w[kMlpWreg];
#pragma unroll
    for (int q = 0; q < kMlpWreg; q++) w[q] = __builtin_bit_cast(bf16x8_t, wreg[q * 64 + lane]);
    unsigned bx, by, bz;
    tile_coords(g, 1, bx, by, bz);
    const int x = (int)bx * kTileX + lane;
    const int y = (int)by * kTileY + (int)(threadIdx.x / kTileX);
    const int z = (int)bz;
    const int r = g.r, wd = 2 * g.r + 1;
    const bool lo = lane < 32;
    const unsigned one_zero = pack_bf16(1.0f, 0.0f);
    const size_t plane = (size_t)g.P1 * g.P2;
    const size_t base = ((size_t)z * g.P1 + (size_t)y) * g.P2 + (size_t)x;  // voter of offset (-r, -r, -r)
    const int nk = wd * wd * wd, kself = (nk - 1) / 2;
    Acc acc = Acc{0, 0, 0, 0, 0, 0, 0};
    // payload of the next offset is loaded one round ahead
    int du = 0, dv = 0, dw = 0;  // offset + r of the next load
    float4 Sn = make_float4(0, 0, 0, 0), Nn, Un;
    if (STICK) Sn = pay_s[base];
    Nn = pay_n[base];
    Un = pay_u[base];
    for (int k = 0; k < nk; k++) {
        const float4 S = Sn, N = Nn, U = Un;
        if (++du == wd) {
            du = 0;
            if (++dv == wd) { dv = 0; ++dw; }
        }
        if (k + 1 < nk) {
            const size_t vi = base + (size_t)dw * plane + (size_t)dv * g.P2 + (size_t)du;
            if (STICK) Sn = pay_s[vi];
            Nn = pay_n[vi];
            Un = pay_u[vi];
        }
        const float* e = ctab + (size_t)k * CS;
        if (STICK) stick_pair<P>(acc, S, e);
        if (k == kself) {
            plate_pair<P>(acc, N, e);
            continue;
        }
        const float4 m0 = mtab[2 * k], m1 = mtab[2 * k + 1];
        // offset-aligned plate frame: u along the in-plane part of d
        const float f2 = fmaf(m0.x, N.x, fmaf(m0.y, N.y, m0.z * N.z));
        float4 Ua, Va;
        {
            const float w0 = fmaf(-f2, N.x, m0.x), w1 = fmaf(-f2, N.y, m0.y), w2 = fmaf(-f2, N.z, m0.z);
            const float ww = fmaf(w0, w0, fmaf(w1, w1, w2 * w2));
            if (ww > 1e-12f) {
                const float iw = rsqrtf(ww);
                Ua = make_float4(w0 * iw, w1 * iw, w2 * iw, 0.0f);
            } else {
                Ua = U;
            }
            Va = make_float4(N.y * Ua.z - N.z * Ua.y, N.z * Ua.x - N.x * Ua.z, N.x * Ua.y - N.y * Ua.x, 0.0f);
        }
        const float f0 = fmaf(m0.x, Ua.x, fmaf(m0.y, Ua.y, m0.z * Ua.z));
        const float f1 = fmaf(m0.x, Va.x, fmaf(m0.y, Va.y, m0.z * Va.z));
        const float f4 = fmaf(2.0f * f2, f2, -1.0f);
        // BF16 high parts and residuals (input slots 0-4, 5-9), constant slot 10
        const unsigned h01 = pack_bf16(f0, f1), h23 = pack_bf16(f2, m1.x);
        const float r0 = f0 - __uint_as_float(h01 << 16), r1 = f1 - __uint_as_float(h01 & 0xffff0000u);
        const float r2 = f2 - __uint_as_float(h23 << 16), r3 = m1.x - __uint_as_float(h23 & 0xffff0000u);
        const unsigned h4l0 = pack_bf16(f4, r0);
        const float r4 = f4 - __uint_as_float(h4l0 << 16);
        const unsigned l12 = pack_bf16(r1, r2), l34 = pack_bf16(r3, r4);
        // slots 0-7 of the lane's pair, and slots 8-15 of the partner lane's
        const unsigned s0 = __shfl_xor(h01, 32), s1 = __shfl_xor(h23, 32), s2 = __shfl_xor(h4l0, 32),
                       s3 = __shfl_xor(l12, 32), s4 = __shfl_xor(l34, 32);
        const bf16x8_t xa = lo ? op4(h01, h23, h4l0, l12) : op4(s4, one_zero, 0u, 0u);
        const bf16x8_t xb = lo ? op4(s0, s1, s2, s3) : op4(l34, one_zero, 0u, 0u);
        float o[6];
        mlp_round(w, xa, xb, lo, o);
#pragma unroll
        for (int q = 0; q < 6; q++) o[q] = fmaf(aff.sd[q], o[q], aff.mu[q]);
        if (U.w != 0.0f) plate_pair<P>(acc, N, e);
        else mlp_accumulate(acc, o, N.w * m0.w, Ua, Va, N);
    }
    if (x < g.s2 && y < g.s1 && z < g.s0) write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, acc);
}

// Stick + tabulated-network plate gather. Same tiling, padding, z register
// blocking and XCD order as k_vote3_closed. Per offset column (dv, du) the
// block stages in LDS the network-table rows of the column's distances
// (r + 1 rows: |dw| = 0..r, kMlpTabPad cells of two float4 each, scaled by
// the offset's c1 / pi), so every pair reads its row from LDS. tab:
// kMlpTabStride floats per offset (build_mlptab_table). Voters flagged by the
// error check carry plate scale 0 here; their plates come from the
// integrator's pass.
template <int P, bool STICK, int BZ>
__global__ __launch_bounds__(kThreads) void k_vote3_mlptab(const float4* __restrict__ pay_s,
                                                           const float4* __restrict__ pay_p,
                                                           const float* __restrict__ tab,
                                                           const float4* __restrict__ vtab, int nrows,
                                                           VoteGeom g, float* __restrict__ out, int out_layout) {
    constexpr int STRIDE = kMlpTabStride;
    constexpr int C1PI = 8, ROW = 9;
    constexpr int CELLS2 = 2 * kMlpTabPad;  // float4 per staged row
    extern __shared__ float4 s_rows[];
    unsigned bx, by, bz;
    tile_coords_zfast(g, bx, by, bz);
    const int x = (int)bx * kTileX + (int)(threadIdx.x % kTileX);
    const int y = (int)by * kTileY + (int)(threadIdx.x / kTileX);
    const int z0 = (int)bz * BZ;
    const int r = g.r, w = 2 * g.r + 1;
    const size_t plane = (size_t)g.P1 * g.P2;
    const int nst = (r + 1) * CELLS2;

    Acc acc[BZ];
#pragma unroll
    for (int j = 0; j < BZ; j++) acc[j] = Acc{0, 0, 0, 0, 0, 0, 0};

    for (int dv = -r; dv <= r; dv++) {
        for (int du = -r; du <= r; du++) {
            const int ebase = (dv + r) * w + (du + r);
            __syncthreads();  // the previous column's rows are no longer read
            for (int i = threadIdx.x; i < nst; i += kThreads) {
                const int slot = i / CELLS2;  // |dw|: offsets +-dw share the distance
                const int q = i - slot * CELLS2, p = q >> 1;
                const float* e = tab + (size_t)((slot + r) * w * w + ebase) * STRIDE;
                const int row = min(max((int)e[ROW], 0), nrows - 1);
                // staged cell p: stored cell o = p - 1, the end cells extended by +-1 step
                const int o = min(max(p - 1, 0), kMlpTabM - 2);
                const float4* src = vtab + ((size_t)row * kMlpTabM + o) * 2;
                const float4 dl = src[1];
                float4 v = dl;
                if (!(q & 1)) {
                    const float sh = (float)(p - 1 - o);
                    v = src[0];
                    v = make_float4(fmaf(sh, dl.x, v.x), fmaf(sh, dl.y, v.y), fmaf(sh, dl.z, v.z), fmaf(sh, dl.w, v.w));
                }
                const float c = e[C1PI];
                s_rows[i] = make_float4(c * v.x, c * v.y, c * v.z, c * v.w);
            }
            __syncthreads();
            const size_t col = ((size_t)z0 * g.P1 + (size_t)(y + r + dv)) * g.P2 + (size_t)(x + r + du);
            for (int zz = 0; zz < BZ + 2 * r; zz++) {
                float4 S;
                if (STICK) S = pay_s[col + (size_t)zz * plane];
                const float4 N = pay_p[col + (size_t)zz * plane];
                const float3 nw = voter_nw(N);
#pragma unroll
                for (int j = 0; j < BZ; j++) {
                    const int dw = zz - r - j;
                    if (dw >= -r && dw <= r) {
                        const float* e = tab + (size_t)((dw + r) * w * w + ebase) * STRIDE;
                        if (STICK) stick_pair<P>(acc[j], S, e);
                        tab_plate_pair(acc[j], N, nw, e, s_rows + (dw < 0 ? -dw : dw) * CELLS2);
                    }
                }
            }
        }
    }
    if (x < g.s2 && y < g.s1) {
#pragma unroll
        for (int j = 0; j < BZ; j++) {
            const int z = z0 + j;
            if (z < g.s0) write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, acc[j]);
        }
    }
}

// ---------------------------------------------------------------- analytical tensor voting (ATV)
// tira::atv (tensorvote.h:688-760) for voter eigenvalues l0 <= l1 <= l2 and an
// orthonormal eigenframe (e0, e1, e2) at unit offset v with attenuation c:
//   c R H R,  R = I - 2 v v^T,
//   H = sum_k s_k (T_k - (v^T T_k v T_k + 2 T_k v v^T T_k) / (k + 2)),
//   T_1 = e2 e2^T, T_2 = e1 e1^T + e2 e2^T = I - e0 e0^T, T_3 = I,
//   s = (l2 - l1, l1 - l0, l0).
// Every T_k is a projector, so with a = e2.v, b = e0.v the vote collapses to
//   KI I + K00 e0 e0^T + K22 e2 e2^T + Kvv v v^T + K0 (v e0^T + e0 v^T) + K2 (v e2^T + e2 v^T)
// with (S = c s, rho = |v|^2: 1, or 0 at the voter itself where the vote is c A)
//   K22 = S1 (1 - a^2)                 K2  = -2 a K22
//   K00 = -S2 (1 - rho/4 + 3 b^2/4)    K0  = S2 b (1 + 3 b^2 / 2)
//   KI  = S2 (1 - rho/4 + b^2/4) + S3 (1 - rho/5)
//   Kvv = -S2 (1/2 + b^2 + 3 b^4) - 2 S3 / 5 + 4 a^2 K22.
// The per-offset factors of c live in the table below; off-diagonal
// accumulators carry twice their value so that the symmetric products need
// one FMA per term.
__device__ __forceinline__ void atv3_pair(Acc& acc, const float4 A, const float4 B, const float s3,
                                          const float* __restrict__ t) {
    const float v0 = t[0], v1 = t[1], v2 = t[2];
    const float a = fmaf(B.x, v0, fmaf(B.y, v1, B.z * v2));  // e2 . v
    const float b = fmaf(A.x, v0, fmaf(A.y, v1, A.z * v2));  // e0 . v
    const float al = a * a, be = b * b;
    const float s1 = B.w, s2 = A.w;
    const float K22 = s1 * fmaf(t[4], al, t[3]);
    const float K00 = s2 * fmaf(t[5], be, t[6]);
    const float KI = fmaf(s2, fmaf(t[7], be, t[8]), s3 * t[9]);
    const float P = fmaf(fmaf(t[10], be, t[4]), be, t[11]);
    const float Kvv = fmaf(4.0f * al, K22, fmaf(s3, t[12], s2 * P));
    const float K0d = (s2 * b) * fmaf(t[13], be, t[14]);  // 2 K0
    const float K2d = (-4.0f * a) * K22;                   // 2 K2
    const float h00 = fmaf(K00, A.x, K0d * v0), h01 = fmaf(K00, A.y, K0d * v1), h02 = fmaf(K00, A.z, K0d * v2);
    const float h20 = fmaf(K22, B.x, K2d * v0), h21 = fmaf(K22, B.y, K2d * v1), h22 = fmaf(K22, B.z, K2d * v2);
    acc.xx = fmaf(A.x, h00, fmaf(B.x, h20, fmaf(Kvv, t[15], acc.xx)));
    acc.yy = fmaf(A.y, h01, fmaf(B.y, h21, fmaf(Kvv, t[16], acc.yy)));
    acc.zz = fmaf(A.z, h02, fmaf(B.z, h22, fmaf(Kvv, t[17], acc.zz)));
    acc.xy = fmaf(A.x, h01, fmaf(h00, A.y, fmaf(B.x, h21, fmaf(h20, B.y, fmaf(Kvv, t[18], acc.xy)))));
    acc.xz = fmaf(A.x, h02, fmaf(h00, A.z, fmaf(B.x, h22, fmaf(h20, B.z, fmaf(Kvv, t[19], acc.xz)))));
    acc.yz = fmaf(A.y, h02, fmaf(h01, A.z, fmaf(B.y, h22, fmaf(h21, B.z, fmaf(Kvv, t[20], acc.yz)))));
    acc.ii += KI;
}

// ATV gather: same tiling, padding and z register blocking as k_vote3_closed.
// Payloads: pay_a = (e0, l1 - l0), pay_b = (e2, l2 - l1), pay_c = l0.
// Same tiling, order and z blocking as k_vote3_closed (voters read through
// L1: staging the 36-byte ATV voters in LDS as k_vote3_closed_lds does was 7%
// slower at 128^3, w 13).
template <int BZ>
__global__ __launch_bounds__(kThreads) void k_vote3_atv(const float4* __restrict__ pay_a,
                                                        const float4* __restrict__ pay_b,
                                                        const float* __restrict__ pay_c,
                                                        const float* __restrict__ tab, VoteGeom g,
                                                        float* __restrict__ out, int out_layout) {
    unsigned bx, by, bz;
    tile_coords_zfast(g, bx, by, bz);
    const int x = (int)bx * kTileX + (int)(threadIdx.x % kTileX);
    const int y = (int)by * kTileY + (int)(threadIdx.x / kTileX);
    const int z0 = (int)bz * BZ;
    const int r = g.r, w = 2 * g.r + 1;
    const size_t plane = (size_t)g.P1 * g.P2;

    Acc acc[BZ];
#pragma unroll
    for (int j = 0; j < BZ; j++) acc[j] = Acc{0, 0, 0, 0, 0, 0, 0};

    // Offsets in the order (dv, voter plane, du), as in k_vote3_closed.
    for (int dv = -r; dv <= r; dv++) {
        const size_t rowb = ((size_t)z0 * g.P1 + (size_t)(y + r + dv)) * g.P2 + (size_t)(x + r);
        for (int zz = 0; zz < BZ + 2 * r; zz++) {
            for (int du = -r; du <= r; du++) {
                const int ebase = (dv + r) * w + (du + r);
                const size_t pi = rowb + (size_t)zz * plane + du;
                const float4 A = pay_a[pi], B = pay_b[pi];
                const float s3 = pay_c[pi];
#pragma unroll
                for (int j = 0; j < BZ; j++) {
                    const int dw = zz - r - j;
                    if (dw >= -r && dw <= r)
                        atv3_pair(acc[j], A, B, s3, tab + (size_t)((dw + r) * w * w + ebase) * kAtvStride);
                }
            }
        }
    }
    if (x < g.s2 && y < g.s1) {
#pragma unroll
        for (int j = 0; j < BZ; j++) {
            const int z = z0 + j;
            if (z < g.s0) {
                Acc o = acc[j];
                o.xy *= 0.5f; o.xz *= 0.5f; o.yz *= 0.5f;
                write_out(out, out_layout, ((size_t)z * g.s1 + y) * g.s2 + x, o);
            }
        }
    }
}

__device__ __forceinline__ void write_atv_payload(size_t p, float l0, float l1, float l2, const float e0[3],
                                                  const float e2[3], float4* pay_a, float4* pay_b, float* pay_c) {
    pay_a[p] = make_float4(e0[0], e0[1], e0[2], l1 - l0);
    pay_b[p] = make_float4(e2[0], e2[1], e2[2], l2 - l1);
    pay_c[p] = l0;
}

// ATV payload from the tensor field (one eigen-decomposition per voxel,
// padded with zero payloads like k_vote3_prep).
__global__ __launch_bounds__(256) void k_vote3_atv_prep(const float* __restrict__ in, int layout, VoteGeom g,
                                                        float4* __restrict__ pay_a, float4* __restrict__ pay_b,
                                                        float* __restrict__ pay_c) {
    const size_t total = (size_t)g.P0 * g.P1 * g.P2;
    const float z3[3] = {0, 0, 0};
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(p % g.P2), py = (int)((p / g.P2) % g.P1), pz = (int)(p / ((size_t)g.P1 * g.P2));
        const int x = px - g.r, y = py - g.r, z = pz - g.r;
        if (x < 0 || y < 0 || z < 0 || x >= g.s2 || y >= g.s1 || z >= g.s0) {
            write_atv_payload(p, 0.0f, 0.0f, 0.0f, z3, z3, pay_a, pay_b, pay_c);
            continue;
        }
        float l0, l1, l2, e0[3], e2[3];
        voxel_eig(in, layout, ((size_t)z * g.s1 + y) * g.s2 + x, l0, l1, l2, e0, e2);
        write_atv_payload(p, l0, l1, l2, e0, e2, pay_a, pay_b, pay_c);
    }
}

// ATV payload from caller eigenpairs (_tensorvote3_atv inputs: ascending
// eigenvalues, middle and largest eigenvectors). The frame is completed with
// e0 = e1 x e2, which is the smallest eigenvector for the orthonormal pairs
// evecs3spherical_symmetric returns.
__global__ __launch_bounds__(256) void k_vote3_atv_prep_eig(const float* __restrict__ lam,
                                                            const float* __restrict__ ev1,
                                                            const float* __restrict__ ev2, VoteGeom g,
                                                            float4* __restrict__ pay_a, float4* __restrict__ pay_b,
                                                            float* __restrict__ pay_c) {
    const size_t total = (size_t)g.P0 * g.P1 * g.P2;
    const float z3[3] = {0, 0, 0};
    for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < total; p += (size_t)gridDim.x * blockDim.x) {
        const int px = (int)(p % g.P2), py = (int)((p / g.P2) % g.P1), pz = (int)(p / ((size_t)g.P1 * g.P2));
        const int x = px - g.r, y = py - g.r, z = pz - g.r;
        if (x < 0 || y < 0 || z < 0 || x >= g.s2 || y >= g.s1 || z >= g.s0) {
            write_atv_payload(p, 0.0f, 0.0f, 0.0f, z3, z3, pay_a, pay_b, pay_c);
            continue;
        }
        const size_t v = ((size_t)z * g.s1 + y) * g.s2 + x;
        const float* m = ev1 + v * 3;
        const float e2[3] = {ev2[v * 3], ev2[v * 3 + 1], ev2[v * 3 + 2]};
        float e0[3] = {m[1] * e2[2] - m[2] * e2[1], m[2] * e2[0] - m[0] * e2[2], m[0] * e2[1] - m[1] * e2[0]};
        const float nn = e0[0] * e0[0] + e0[1] * e0[1] + e0[2] * e0[2];
        const float inv = nn > 0.0f ? rsqrtf(nn) : 0.0f;
        e0[0] *= inv; e0[1] *= inv; e0[2] *= inv;
        write_atv_payload(p, lam[v * 3], lam[v * 3 + 1], lam[v * 3 + 2], e0, e2, pay_a, pay_b, pay_c);
    }
}

__global__ void k_fill_zero(float* __restrict__ out, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = 0.0f;
}

// ---------------------------------------------------------------- host driver
VoteGeom make_geom(unsigned s0, unsigned s1, unsigned s2, unsigned w, int bz) {
    VoteGeom g;
    g.s0 = (int)s0; g.s1 = (int)s1; g.s2 = (int)s2;
    g.r = (int)(w / 2);
    g.nbx = ceil_div(s2, kTileX);
    g.nby = ceil_div(s1, kTileY);
    g.nbz = ceil_div(s0, bz);
    g.P2 = (int)(g.nbx * kTileX) + 2 * g.r;
    g.P1 = (int)(g.nby * kTileY) + 2 * g.r;
    g.P0 = (int)(g.nbz * bz) + 2 * g.r;
    return g;
}

// Unpadded voter payloads (stick (e2, scale), plate (e0, scale)) of every
// voxel, for the spectral vote; either pointer may be NULL.
int vote3_payload_flat(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0,
                       unsigned s0, unsigned s1, unsigned s2, float4* ps, float4* pp, hipStream_t st) {
    VoteGeom g{};
    g.s0 = (int)s0; g.s1 = (int)s1; g.s2 = (int)s2;
    g.r = 0;
    g.P0 = (int)s0; g.P1 = (int)s1; g.P2 = (int)s2;
    const size_t n = (size_t)s0 * s1 * s2;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(n, 256), 65536);
    if (tensors) k_vote3_prep<<<grid, 256, 0, st>>>(tensors, in_layout, g, ps, pp, nullptr, 1.0f);
    else k_vote3_prep_tk<<<grid, 256, 0, st>>>(lam, ev2, ev0, g, ps, pp, nullptr, 1.0f);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int resolve_plate_mode(int plate_mode, unsigned power, unsigned samples) {
    const unsigned p = power < 1 ? 1 : power;
    if (plate_mode == TIRA_PLATE_AUTO) {
        if (p <= 4 && (samples == 0 || samples >= p + 2)) return TIRA_PLATE_CLOSED;
        return TIRA_PLATE_SUM;
    }
    if (plate_mode == TIRA_PLATE_CLOSED && (p > 4 || (samples > 0 && samples < p + 2))) return TIRA_PLATE_SUM;
    return plate_mode;
}

struct TableKey {
    int sum;
    int w;
    float s1, s2;
    unsigned power, nsamp;
    bool operator==(const TableKey& o) const {
        return sum == o.sum && w == o.w && s1 == o.s1 && s2 == o.s2 && power == o.power && nsamp == o.nsamp;
    }
};

struct TableCache {
    TableKey key{-1, 0, 0, 0, 0, 0};
    int device = -1;
    void* ptr = nullptr;
};
static thread_local TableCache g_tab;

// Device copy of the offset table for `key` (plus the cos/sin table of the
// literal sum), rebuilt and uploaded only when the key changes.
static int table_for(const TableKey& key, int P, hipStream_t st, char** dtab, float2** dcs) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    std::vector<float> tab;
    std::vector<float2> cs;
    size_t tab_bytes, cs_bytes = 0;
    if (key.sum == 2) {
        tab_bytes = (size_t)key.w * key.w * key.w * kAtvStride * 4;
    } else if (key.sum == 1) {
        tab_bytes = (size_t)key.w * key.w * key.w * kSumStride * 4;
        cs_bytes = (size_t)key.nsamp * 8;
    } else {
        tab_bytes = (size_t)key.w * key.w * key.w * closed_stride(P) * 4;
    }
    const size_t cs_off = (tab_bytes + 255) & ~size_t(255);
    void* buf = nullptr;
    TH_TRY(ws_get(WS_TABLE, cs_off + cs_bytes + 256, &buf));
    *dtab = (char*)buf;
    *dcs = key.sum == 1 ? (float2*)((char*)buf + cs_off) : nullptr;
    if (g_tab.key == key && g_tab.device == dev && g_tab.ptr == buf) return TIRA_OK;
    if (key.sum == 2) {
        build_atv3_table(key.w, key.s1, tab);
    } else if (key.sum == 1) {
        build_sum_table(key.w, key.s1, key.s2, key.power, tab);
        cs.resize(key.nsamp);
        const float db = (float)M_PI / (float)key.nsamp;  // dbeta of tensorvote.h:2473
        for (unsigned i = 0; i < key.nsamp; i++) {
            const float beta = db * (float)i;
            cs[i] = make_float2(std::cos(beta), std::sin(beta));
        }
    } else {
        build_closed_table(P, key.w, key.s1, key.s2, key.power, tab);
    }
    // Synchronous upload from pageable memory: ordered after earlier work on
    // `st` that may still read the previous table.
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(buf, tab.data(), tab_bytes, hipMemcpyHostToDevice));
    if (key.sum == 1) TH_HIP(hipMemcpy((char*)buf + cs_off, cs.data(), cs_bytes, hipMemcpyHostToDevice));
    g_tab.key = key;
    g_tab.device = dev;
    g_tab.ptr = buf;
    return TIRA_OK;
}

constexpr int kBZ = 4;     // receivers per lane along z: closed, literal-sum and ATV gathers
constexpr int kBZMlp = 8;  // the tabulated-network gather (its LDS rows are staged per offset column)

template <int P>
static void launch_closed(bool stick, bool plate, unsigned grid, hipStream_t st, const float4* ps, const float4* pp,
                          const float* tab, const VoteGeom& g, float* out, int layout) {
    // voters staged in LDS: 3% faster than reading them through L1 at 256^3, w 25;
    // windows above 449 (more than 64 KB of segments per block) read them through L1
    const size_t sh = (size_t)kTileY * 2 * (kTileX + 2 * g.r) * sizeof(float4);
    if (sh > 65536) {
        if (stick && plate) k_vote3_closed<P, true, true, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, tab, g, out, layout, nullptr, 0);
        else if (stick) k_vote3_closed<P, true, false, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, tab, g, out, layout, nullptr, 0);
        else k_vote3_closed<P, false, true, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, tab, g, out, layout, nullptr, 0);
        return;
    }
    if (stick && plate) k_vote3_closed_lds<P, true, true, kBZ><<<grid, kThreads, sh, st>>>(ps, pp, tab, g, out, layout);
    else if (stick) k_vote3_closed_lds<P, true, false, kBZ><<<grid, kThreads, sh, st>>>(ps, pp, tab, g, out, layout);
    else k_vote3_closed_lds<P, false, true, kBZ><<<grid, kThreads, sh, st>>>(ps, pp, tab, g, out, layout);
}

// Runs the vote on device buffers. `eig_in`: tensors (layout) or NULL when
// lam/ev2/ev0 eigenpairs are given.
int vote3_run(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0, float* out,
              int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
              unsigned power, bool stick, bool plate, unsigned samples, int plate_mode, hipStream_t st,
              EigOut* eig) {
    const size_t nvox = (size_t)s0 * s1 * s2;
    if (nvox == 0) return TIRA_OK;
    if (!stick && !plate) {
        const size_t nf = nvox * (out_layout == TIRA_LAYOUT_MAT9 ? 9 : 6);
        k_fill_zero<<<std::min<size_t>(4096, ceil_div(nf, 256)), 256, 0, st>>>(out, nf);
        TH_HIP(hipGetLastError());
        return TIRA_OK;
    }
    if (plate_mode == TIRA_PLATE_AUTO && spectral_preferred(s0, s1, s2, w, power, stick, plate, samples))
        plate_mode = TIRA_VOTE_SPECTRAL;
    if (plate_mode == TIRA_VOTE_SPECTRAL) {
        const int rc = vote3_spectral_run(tensors, in_layout, lam, ev2, ev0, out, out_layout, s0, s1, s2, sigma1,
                                          sigma2, w, power, stick, plate, samples, false, st, eig);
        if (rc != kSpectralNotApplicable) return rc;
        plate_mode = TIRA_PLATE_AUTO;  // power, samples or size outside the spectral method: direct gather
    }
    const unsigned p = power < 1 ? 1 : power;
    const int mode = plate ? resolve_plate_mode(plate_mode, power, samples) : TIRA_PLATE_CLOSED;
    if (mode == TIRA_PLATE_MLP) {
        set_error("plate_mode MLP is served by tira_tensorvote3_mlp");
        return TIRA_ERR_INVALID_ARG;
    }
    const VoteGeom g = make_geom(s0, s1, s2, w, kBZ);
    const size_t ptotal = (size_t)g.P0 * g.P1 * g.P2;
    // Payload buffers.
    float4 *ps = nullptr, *pp = nullptr, *pv = nullptr;
    if (stick) TH_TRY(ws_get(WS_PAY0, ptotal * 16, (void**)&ps));
    if (plate) TH_TRY(ws_get(WS_PAY1, ptotal * 16, (void**)&pp));
    const bool sum = plate && mode == TIRA_PLATE_SUM;
    const unsigned nsamp = samples == 0 ? p + 2 : samples;  // samples == 0 (analytic) equals the exact integral
    if (sum) TH_TRY(ws_get(WS_SCRATCH0, ptotal * 16, (void**)&pv));
    const float inv_n = sum ? 1.0f / (float)nsamp : 1.0f;
    const unsigned pgrid = (unsigned)std::min<size_t>(ceil_div(ptotal, 256), 65536);
    {
        ProfScope prof("vote3_prep", st);
        if (tensors) k_vote3_prep<<<pgrid, 256, 0, st>>>(tensors, in_layout, g, ps, pp, pv, inv_n);
        else k_vote3_prep_tk<<<pgrid, 256, 0, st>>>(lam, ev2, ev0, g, ps, pp, pv, inv_n);
    }
    TH_HIP(hipGetLastError());

    // Offset table: host-built, uploaded once per (mode, w, sigmas, power, N)
    // and kept on the device, so repeated calls launch without a host sync.
    const bool sum_tab = sum;
    const int P = (int)p;
    const TableKey key{sum_tab ? 1 : 0, (int)w, sigma1, sigma2, power, sum_tab ? nsamp : 0u};
    char* dtab = nullptr;
    float2* dcs = nullptr;
    TH_TRY(table_for(key, P, st, &dtab, &dcs));
    const unsigned grid = g.nbx * g.nby * g.nbz;
    const float* t = (const float*)dtab;
    ProfScope prof_vote(sum ? "vote3_sum" : "vote3_closed", st);
    if (sum) {
        if (stick && plate) k_vote3_sum<true, true, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, pv, t, dcs, nsamp, p, g, out, out_layout);
        else if (stick) k_vote3_sum<true, false, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, pv, t, dcs, nsamp, p, g, out, out_layout);
        else k_vote3_sum<false, true, kBZ><<<grid, kThreads, 0, st>>>(ps, pp, pv, t, dcs, nsamp, p, g, out, out_layout);
    } else {
        switch (P) {
            case 1: launch_closed<1>(stick, plate, grid, st, ps, pp, t, g, out, out_layout); break;
            case 2: launch_closed<2>(stick, plate, grid, st, ps, pp, t, g, out, out_layout); break;
            case 3: launch_closed<3>(stick, plate, grid, st, ps, pp, t, g, out, out_layout); break;
            default: launch_closed<4>(stick, plate, grid, st, ps, pp, t, g, out, out_layout); break;
        }
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// ATV on device buffers (tensorvote.h:2635-2739 semantics): `tensors` in
// `in_layout`, or caller eigenpairs lam / ev1 (middle) / ev2 (largest).
// The window half-width is w / 2, as the reference's loops take it.
// method: TIRA_PLATE_AUTO (the spectral vote when it applies: tensor input,
// w >= 5, at least 32^3 voxels, x-rows of at most 256), TIRA_PLATE_CLOSED
// (always the gather) or TIRA_VOTE_SPECTRAL (spectral whenever the input
// allows it).
int vote3_atv_run(const float* tensors, int in_layout, const float* lam, const float* ev1, const float* ev2,
                  float* out, int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w,
                  hipStream_t st, int method) {
    const size_t nvox = (size_t)s0 * s1 * s2;
    if (nvox == 0) return TIRA_OK;
    if (tensors && (method == TIRA_VOTE_SPECTRAL ||
                    (method == TIRA_PLATE_AUTO && spectral_preferred(s0, s1, s2, w, 1, true, true, 0, true)))) {
        const int rc = vote3_spectral_run(tensors, in_layout, nullptr, nullptr, nullptr, out, out_layout, s0, s1, s2,
                                          sigma, 0.0f, w, 1, true, true, 0, true, st);
        if (rc != kSpectralNotApplicable) return rc;
    }
    const unsigned wo = 2 * (w / 2) + 1;
    const VoteGeom g = make_geom(s0, s1, s2, wo, kBZ);
    const size_t ptotal = (size_t)g.P0 * g.P1 * g.P2;
    float4 *pa = nullptr, *pb = nullptr;
    float* pc = nullptr;
    TH_TRY(ws_get(WS_PAY0, ptotal * 16, (void**)&pa));
    TH_TRY(ws_get(WS_PAY1, ptotal * 16, (void**)&pb));
    TH_TRY(ws_get(WS_SCRATCH0, ptotal * 4, (void**)&pc));
    const unsigned pgrid = (unsigned)std::min<size_t>(ceil_div(ptotal, 256), 65536);
    {
        ProfScope prof("vote3_atv_prep", st);
        if (tensors) k_vote3_atv_prep<<<pgrid, 256, 0, st>>>(tensors, in_layout, g, pa, pb, pc);
        else k_vote3_atv_prep_eig<<<pgrid, 256, 0, st>>>(lam, ev1, ev2, g, pa, pb, pc);
    }
    TH_HIP(hipGetLastError());
    const TableKey key{2, (int)wo, sigma, 0.0f, 0u, 0u};
    char* dtab = nullptr;
    float2* dcs = nullptr;
    TH_TRY(table_for(key, 1, st, &dtab, &dcs));
    const unsigned grid = g.nbx * g.nby * g.nbz;
    {
        ProfScope prof("vote3_atv", st);
        k_vote3_atv<kBZ><<<grid, kThreads, 0, st>>>(pa, pb, pc, (const float*)dtab, g, out, out_layout);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// ---------------------------------------------------------------- learned plate driver
// Per-offset table of the surrogate: (d, c1 / pi), (min(|d| / sigma1, 3), 0, 0, 0),
// d and c1 as cpu::build_neighbors3d computes them.
static void build_mlp_table(int w, float s1, float s2, std::vector<float4>& t) {
    const int hw = w / 2;
    t.assign(2 * (size_t)w * w * w, make_float4(0, 0, 0, 0));
    size_t k = 0;
    for (int dw = -hw; dw <= hw; dw++)
        for (int dv = -hw; dv <= hw; dv++)
            for (int du = -hw; du <= hw; du++, k++) {
                float d[3], c1, c2;
                offset_geometry(du, dv, dw, s1, s2, d, c1, c2);
                const float L = std::sqrt((float)(du * du + dv * dv + dw * dw));
                t[2 * k] = make_float4(d[0], d[1], d[2], (float)(c1 / M_PI));
                t[2 * k + 1] = make_float4(std::min(L / s1, kMlpRhoMax), 0.0f, 0.0f, 0.0f);
            }
}

struct MlpTabCache {
    int w = 0;
    float s1 = 0, s2 = 0;
    int device = -1;
    void* ptr = nullptr;
};
static thread_local MlpTabCache g_mtab;

static int mlp_table_for(int w, float s1, float s2, hipStream_t st, float4** dtab) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    const size_t bytes = 2 * (size_t)w * w * w * 16;
    void* buf = nullptr;
    TH_TRY(ws_get(WS_TABLE3, bytes, &buf));
    *dtab = (float4*)buf;
    if (g_mtab.w == w && g_mtab.s1 == s1 && g_mtab.s2 == s2 && g_mtab.device == dev && g_mtab.ptr == buf)
        return TIRA_OK;
    std::vector<float4> t;
    build_mlp_table(w, s1, s2, t);
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(buf, t.data(), bytes, hipMemcpyHostToDevice));
    g_mtab.w = w; g_mtab.s1 = s1; g_mtab.s2 = s2; g_mtab.device = dev; g_mtab.ptr = buf;
    return TIRA_OK;
}

struct MlpCompactCache {
    int P = 0, w = 0;
    float s1 = 0, s2 = 0;
    unsigned power = 0;
    int device = -1;
    void* ptr = nullptr;
};
static thread_local MlpCompactCache g_ctab;

// Device copy of the tabulated-network vote's offset table, rebuilt and
// uploaded only when its parameters change.
static int mlptab_table_for(int P, int w, float s1, float s2, unsigned power, hipStream_t st, float** dtab) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    const size_t bytes = (size_t)w * w * w * kMlpTabStride * 4;
    void* buf = nullptr;
    TH_TRY(ws_get(WS_TABLE4, bytes, &buf));
    *dtab = (float*)buf;
    if (g_ctab.P == P && g_ctab.w == w && g_ctab.s1 == s1 && g_ctab.s2 == s2 && g_ctab.power == power &&
        g_ctab.device == dev && g_ctab.ptr == buf)
        return TIRA_OK;
    std::vector<float> t;
    build_mlptab_table(P, w, s1, s2, power, t);
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(buf, t.data(), bytes, hipMemcpyHostToDevice));
    g_ctab.P = P; g_ctab.w = w; g_ctab.s1 = s1; g_ctab.s2 = s2; g_ctab.power = power; g_ctab.device = dev;
    g_ctab.ptr = buf;
    return TIRA_OK;
}

template <int P>
static void launch_mlp(bool stick, unsigned grid, hipStream_t st, const float4* ps, const float4* pn,
                       const float4* pu, const float* ct, const float4* mt, const uint4* wr, const MlpAffine& aff,
                       const VoteGeom& g, float* out, int layout) {
    if (stick) k_vote3_mlp<P, true><<<grid, kThreads, 0, st>>>(ps, pn, pu, ct, mt, wr, aff, g, out, layout);
    else k_vote3_mlp<P, false><<<grid, kThreads, 0, st>>>(ps, pn, pu, ct, mt, wr, aff, g, out, layout);
}

// Largest window of the tabulated-network vote: its LDS holds (w / 2 + 1)
// staged distance rows of 2 kMlpTabPad float4 (65 KB at w = 63).
constexpr unsigned kMlpTabMaxW = 63;

// Rows above 64 KB (w = 63) need the kernel's dynamic-LDS limit raised, once
// per process and kernel.
template <int P, bool STICK>
static int mlptab_lds_attr(size_t lds) {
    static bool set = false;
    if (lds > 65536 && !set) {
        TH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_vote3_mlptab<P, STICK, kBZMlp>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)((kMlpTabMaxW / 2 + 1) * 2 * kMlpTabPad * sizeof(float4))));
        set = true;
    }
    return TIRA_OK;
}

template <int P>
static int launch_mlptab(bool stick, unsigned grid, size_t lds, hipStream_t st, const float4* ps, const float4* pn,
                         const float* t, const float4* vt, int nrows, const VoteGeom& g, float* out, int layout) {
    if (stick) {
        TH_TRY((mlptab_lds_attr<P, true>(lds)));
        k_vote3_mlptab<P, true, kBZMlp><<<grid, kThreads, lds, st>>>(ps, pn, t, vt, nrows, g, out, layout);
    } else {
        TH_TRY((mlptab_lds_attr<P, false>(lds)));
        k_vote3_mlptab<P, false, kBZMlp><<<grid, kThreads, lds, st>>>(ps, pn, t, vt, nrows, g, out, layout);
    }
    return TIRA_OK;
}

// Stick + learned plate vote on device buffers with the network `m` (trained
// for its sigma1, sigma2 and power). Voters whose orientation bin's error
// exceeds `tol` use the integrator. n_fallback (host, optional) receives the
// number of such voters (synchronises). method: TIRA_MLP_TABLE (the network
// tabulated over distance rows and x = 1 - (n.d)^2; windows up to
// kMlpTabMaxW, larger ones evaluate per pair) or TIRA_MLP_PER_PAIR.
int vote3_mlp_run(PlateMlp* m, const float* tensors, int in_layout, float* out, int out_layout, unsigned s0,
                  unsigned s1, unsigned s2, unsigned w, bool stick, float tol, unsigned long long* n_fallback,
                  hipStream_t st, int method) {
    const size_t nvox = (size_t)s0 * s1 * s2;
    if (nvox == 0) return TIRA_OK;
    const int P = (int)m->power;
    if (P > 4) {
        set_error("the learned plate vote supports power 1..4 (network trained for %d)", P);
        return TIRA_ERR_UNSUPPORTED;
    }
    const bool table = method == TIRA_MLP_TABLE && w <= kMlpTabMaxW && m->vtab != nullptr;
    const VoteGeom g = make_geom(s0, s1, s2, w, table ? kBZMlp : 1);
    const size_t ptotal = (size_t)g.P0 * g.P1 * g.P2;
    // payloads: stick, plate; per pair: plate basis vector; table: the
    // flagged voters' plates
    float4 *ps = nullptr, *pn = nullptr, *pu = nullptr, *pf = nullptr;
    if (stick) TH_TRY(ws_get(WS_PAY0, ptotal * 16, (void**)&ps));
    TH_TRY(ws_get(WS_PAY1, ptotal * 16, (void**)&pn));
    TH_TRY(ws_get(WS_PAY2, ptotal * 16, table ? (void**)&pf : (void**)&pu));
    unsigned long long* cnt = nullptr;
    if (n_fallback || table) {
        TH_TRY(ws_get(WS_COUNTER, 8, (void**)&cnt));
        k_fill_zero<<<1, 64, 0, st>>>((float*)cnt, 2);
    }
    const unsigned pgrid = (unsigned)std::min<size_t>(ceil_div(ptotal, 256), 65536);
    {
        ProfScope prof("vote3_prep_mlp", st);
        k_vote3_prep_mlp<<<pgrid, 256, 0, st>>>(tensors, in_layout, g, ps, pn, pu, pf,
                                                table ? m->bin_err + kMlpBins : m->bin_err, tol, cnt);
    }
    TH_HIP(hipGetLastError());
    const TableKey key{0, (int)w, m->sigma1, m->sigma2, m->power, 0u};
    char* dtab = nullptr;
    float2* dcs = nullptr;
    TH_TRY(table_for(key, P, st, &dtab, &dcs));
    const unsigned grid = g.nbx * g.nby * g.nbz;
    const float* ct = (const float*)dtab;
    if (table) {
        float* mt = nullptr;
        TH_TRY(mlptab_table_for(P, (int)w, m->sigma1, m->sigma2, m->power, st, &mt));
        const size_t lds = (size_t)(w / 2 + 1) * 2 * kMlpTabPad * sizeof(float4);
        const float4* vt = m->vtab;
        const int nr = m->vtab_rows;
        {
            ProfScope prof("vote3_mlptab", st);
            switch (P) {
                case 1: TH_TRY(launch_mlptab<1>(stick, grid, lds, st, ps, pn, mt, vt, nr, g, out, out_layout)); break;
                case 2: TH_TRY(launch_mlptab<2>(stick, grid, lds, st, ps, pn, mt, vt, nr, g, out, out_layout)); break;
                case 3: TH_TRY(launch_mlptab<3>(stick, grid, lds, st, ps, pn, mt, vt, nr, g, out, out_layout)); break;
                default: TH_TRY(launch_mlptab<4>(stick, grid, lds, st, ps, pn, mt, vt, nr, g, out, out_layout)); break;
            }
        }
        TH_HIP(hipGetLastError());
        // integrator plates of the flagged voters, added to the output; the
        // launch returns at once when the error check flagged no voter
        ProfScope prof("vote3_mlp_fallback", st);
        switch (P) {
            case 1: k_vote3_closed<1, false, true, kBZMlp><<<grid, kThreads, 0, st>>>(nullptr, pf, ct, g, out, out_layout, cnt, 1); break;
            case 2: k_vote3_closed<2, false, true, kBZMlp><<<grid, kThreads, 0, st>>>(nullptr, pf, ct, g, out, out_layout, cnt, 1); break;
            case 3: k_vote3_closed<3, false, true, kBZMlp><<<grid, kThreads, 0, st>>>(nullptr, pf, ct, g, out, out_layout, cnt, 1); break;
            default: k_vote3_closed<4, false, true, kBZMlp><<<grid, kThreads, 0, st>>>(nullptr, pf, ct, g, out, out_layout, cnt, 1); break;
        }
    } else {
        float4* mt = nullptr;
        TH_TRY(mlp_table_for((int)w, m->sigma1, m->sigma2, st, &mt));
        const uint4* wr = (const uint4*)m->wreg;
        ProfScope prof("vote3_mlp", st);
        switch (P) {
            case 1: launch_mlp<1>(stick, grid, st, ps, pn, pu, ct, mt, wr, m->aff, g, out, out_layout); break;
            case 2: launch_mlp<2>(stick, grid, st, ps, pn, pu, ct, mt, wr, m->aff, g, out, out_layout); break;
            case 3: launch_mlp<3>(stick, grid, st, ps, pn, pu, ct, mt, wr, m->aff, g, out, out_layout); break;
            default: launch_mlp<4>(stick, grid, st, ps, pn, pu, ct, mt, wr, m->aff, g, out, out_layout); break;
        }
    }
    TH_HIP(hipGetLastError());
    if (n_fallback) {
        TH_HIP(hipMemcpyAsync(n_fallback, cnt, 8, hipMemcpyDeviceToHost, st));
        TH_HIP(hipStreamSynchronize(st));
    }
    return TIRA_OK;
}

static int check_vote_args(unsigned s0, unsigned s1, unsigned s2, unsigned w, float sigma1, float sigma2) {
    TH_ARG(w >= 1 && (w % 2) == 1, "w must be a positive odd integer (got %u)", w);
    TH_ARG(sigma1 >= 0.0f && sigma2 >= 0.0f, "sigma must be >= 0");
    TH_ARG((size_t)s0 * s1 * s2 < ((size_t)1 << 40), "volume too large");
    TH_ARG(w <= 129, "window %u exceeds the supported maximum (129)", w);
    (void)s0; (void)s1; (void)s2;
    return TIRA_OK;
}

static int check_atv_args(unsigned s0, unsigned s1, unsigned s2, unsigned w, float sigma) {
    TH_ARG(sigma >= 0.0f, "sigma must be >= 0");
    TH_ARG((size_t)s0 * s1 * s2 < ((size_t)1 << 40), "volume too large");
    TH_ARG(w <= 129, "window %u exceeds the supported maximum (129)", w);
    (void)s0; (void)s1; (void)s2;
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_tensorvote3(const float* input_field, float* output_field, unsigned s0, unsigned s1, unsigned s2,
                     float sigma, float sigma2, unsigned w, unsigned power, int device, int stick, int plate,
                     unsigned samples, int is_atv) {
    TH_TRY(use_device(device));
    TH_ARG(input_field && output_field, "NULL pointer");
    const size_t n = (size_t)s0 * s1 * s2;
    const float* din; void* dout;
    if (is_atv) {
        // ATV takes sigma only; stick, plate, sigma2, power and samples do not apply (tensorvote.h:2860-2861).
        TH_TRY(check_atv_args(s0, s1, s2, w, sigma));
        TH_TRY(stage_in(input_field, n * 36, WS_IN0, (const void**)&din));
        TH_TRY(stage_out(output_field, n * 36, WS_OUT0, &dout));
        TH_TRY(vote3_atv_run(din, TIRA_LAYOUT_MAT9, nullptr, nullptr, nullptr, (float*)dout, TIRA_LAYOUT_MAT9, s0,
                             s1, s2, sigma, w, stream(), TIRA_PLATE_AUTO));
        return finish_out(output_field, dout, n * 36);
    }
    TH_TRY(check_vote_args(s0, s1, s2, w, sigma, sigma2));
    TH_TRY(stage_in(input_field, n * 36, WS_IN0, (const void**)&din));
    TH_TRY(stage_out(output_field, n * 36, WS_OUT0, &dout));
    TH_TRY(vote3_run(din, TIRA_LAYOUT_MAT9, nullptr, nullptr, nullptr, (float*)dout, TIRA_LAYOUT_MAT9, s0, s1, s2,
                     sigma, sigma2, w, power, stick != 0, plate != 0, samples, TIRA_PLATE_AUTO, stream()));
    return finish_out(output_field, dout, n * 36);
}

int tira_tensorvote3_tk(float* output_field, const float* lambdas, const float* evec2, const float* evec0,
                        unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned power,
                        unsigned w, int stick, int plate, unsigned samples) {
    TH_ARG(output_field && lambdas, "NULL pointer");
    TH_ARG(!stick || evec2, "stick voting needs evec2");
    TH_ARG(!plate || evec0, "plate voting needs evec0");
    TH_TRY(check_vote_args(s0, s1, s2, w, sigma1, sigma2));
    const size_t n = (size_t)s0 * s1 * s2;
    const float *dl, *d2 = nullptr, *d0 = nullptr; void* dout;
    TH_TRY(stage_in(lambdas, n * 12, WS_IN0, (const void**)&dl));
    if (stick) TH_TRY(stage_in(evec2, n * 12, WS_IN1, (const void**)&d2));
    if (plate) TH_TRY(stage_in(evec0, n * 12, WS_IN2, (const void**)&d0));
    TH_TRY(stage_out(output_field, n * 36, WS_OUT0, &dout));
    TH_TRY(vote3_run(nullptr, 0, dl, d2, d0, (float*)dout, TIRA_LAYOUT_MAT9, s0, s1, s2, sigma1, sigma2, w, power,
                     stick != 0, plate != 0, samples, TIRA_PLATE_AUTO, stream()));
    return finish_out(output_field, dout, n * 36);
}

int tira_tensorvote3_mlp(void* mlp, const float* input_field, int in_layout, float* output_field, int out_layout,
                         unsigned s0, unsigned s1, unsigned s2, unsigned w, int stick, float tol,
                         unsigned long long* n_fallback) {
    return tira_tensorvote3_mlp_method(mlp, input_field, in_layout, output_field, out_layout, s0, s1, s2, w, stick,
                                       tol, n_fallback, TIRA_MLP_TABLE);
}

int tira_tensorvote3_mlp_method(void* mlp, const float* input_field, int in_layout, float* output_field,
                                int out_layout, unsigned s0, unsigned s1, unsigned s2, unsigned w, int stick,
                                float tol, unsigned long long* n_fallback, int method) {
    TH_ARG(mlp && input_field && output_field, "NULL pointer");
    TH_ARG(method == TIRA_MLP_TABLE || method == TIRA_MLP_PER_PAIR, "bad MLP evaluation method %d", method);
    PlateMlp* m = static_cast<PlateMlp*>(mlp);
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    TH_ARG(dev == m->device, "MLP handle belongs to device %d (current %d)", m->device, dev);
    TH_ARG(m->wreg != nullptr, "the plate network has no weights yet (train or set its parameters)");
    TH_ARG(in_layout == TIRA_LAYOUT_MAT9 || in_layout == TIRA_LAYOUT_PACKED6, "bad input layout");
    TH_ARG(out_layout == TIRA_LAYOUT_MAT9 || out_layout == TIRA_LAYOUT_PACKED6, "bad output layout");
    TH_ARG(tol >= 0.0f, "tolerance must be >= 0");
    TH_ARG(is_device(input_field) && is_device(output_field), "the learned plate vote takes device pointers");
    TH_TRY(check_vote_args(s0, s1, s2, w, m->sigma1, m->sigma2));
    return vote3_mlp_run(m, input_field, in_layout, output_field, out_layout, s0, s1, s2, w, stick != 0, tol,
                         n_fallback, stream(), method);
}

int tira_tensorvote3_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                           unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
                           unsigned power, int stick, int plate, unsigned samples, int plate_mode) {
    TH_ARG(input_field && output_field, "NULL pointer");
    TH_ARG(in_layout == TIRA_LAYOUT_MAT9 || in_layout == TIRA_LAYOUT_PACKED6, "bad input layout");
    TH_ARG(out_layout == TIRA_LAYOUT_MAT9 || out_layout == TIRA_LAYOUT_PACKED6, "bad output layout");
    TH_ARG((plate_mode >= TIRA_PLATE_AUTO && plate_mode <= TIRA_PLATE_CLOSED) || plate_mode == TIRA_VOTE_SPECTRAL, "bad plate mode %d", plate_mode);
    TH_TRY(check_vote_args(s0, s1, s2, w, sigma1, sigma2));
    return vote3_run(input_field, in_layout, nullptr, nullptr, nullptr, output_field, out_layout, s0, s1, s2, sigma1,
                     sigma2, w, power, stick != 0, plate != 0, samples, plate_mode, stream());
}

int tira_tensorvote3_eig_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                               unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
                               unsigned power, int stick, int plate, unsigned samples, int plate_mode, float* evals,
                               unsigned eig_flags, unsigned long long* n_refined_dev) {
    TH_ARG(input_field && output_field && evals, "NULL pointer");
    TH_ARG(in_layout == TIRA_LAYOUT_MAT9 || in_layout == TIRA_LAYOUT_PACKED6, "bad input layout");
    TH_ARG(out_layout == TIRA_LAYOUT_MAT9 || out_layout == TIRA_LAYOUT_PACKED6, "bad output layout");
    TH_ARG((plate_mode >= TIRA_PLATE_AUTO && plate_mode <= TIRA_PLATE_CLOSED) || plate_mode == TIRA_VOTE_SPECTRAL, "bad plate mode %d", plate_mode);
    TH_TRY(check_vote_args(s0, s1, s2, w, sigma1, sigma2));
    EigOut eo{evals, eig_flags, n_refined_dev, false};
    TH_TRY(vote3_run(input_field, in_layout, nullptr, nullptr, nullptr, output_field, out_layout, s0, s1, s2, sigma1,
                     sigma2, w, power, stick != 0, plate != 0, samples, plate_mode, stream(), &eo));
    if (eo.done) return TIRA_OK;
    return tira_eig3_async(output_field, (size_t)s0 * s1 * s2, out_layout, evals, nullptr, nullptr, eig_flags,
                           n_refined_dev);
}

// _tensorvote3_atv (tensorvote.h:2695-2739): ATV from eigenpairs.
int tira_tensorvote3_atv(float* output_field, const float* lambdas, const float* evec1, const float* evec2,
                         unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w) {
    TH_ARG(output_field && lambdas && evec1 && evec2, "NULL pointer");
    TH_TRY(check_atv_args(s0, s1, s2, w, sigma));
    const size_t n = (size_t)s0 * s1 * s2;
    const float *dl, *d1, *d2; void* dout;
    TH_TRY(stage_in(lambdas, n * 12, WS_IN0, (const void**)&dl));
    TH_TRY(stage_in(evec1, n * 12, WS_IN1, (const void**)&d1));
    TH_TRY(stage_in(evec2, n * 12, WS_IN2, (const void**)&d2));
    TH_TRY(stage_out(output_field, n * 36, WS_OUT0, &dout));
    TH_TRY(vote3_atv_run(nullptr, 0, dl, d1, d2, (float*)dout, TIRA_LAYOUT_MAT9, s0, s1, s2, sigma, w, stream(),
                         TIRA_PLATE_AUTO));
    return finish_out(output_field, dout, n * 36);
}

int tira_tensorvote3_atv_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                               unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w) {
    TH_ARG(input_field && output_field, "NULL pointer");
    TH_ARG(in_layout == TIRA_LAYOUT_MAT9 || in_layout == TIRA_LAYOUT_PACKED6, "bad input layout");
    TH_ARG(out_layout == TIRA_LAYOUT_MAT9 || out_layout == TIRA_LAYOUT_PACKED6, "bad output layout");
    TH_TRY(check_atv_args(s0, s1, s2, w, sigma));
    return vote3_atv_run(input_field, in_layout, nullptr, nullptr, nullptr, output_field, out_layout, s0, s1, s2,
                         sigma, w, stream(), TIRA_PLATE_AUTO);
}

int tira_tensorvote3_atv_method_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                                      unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w, int method) {
    TH_ARG(input_field && output_field, "NULL pointer");
    TH_ARG(in_layout == TIRA_LAYOUT_MAT9 || in_layout == TIRA_LAYOUT_PACKED6, "bad input layout");
    TH_ARG(out_layout == TIRA_LAYOUT_MAT9 || out_layout == TIRA_LAYOUT_PACKED6, "bad output layout");
    TH_ARG(method == TIRA_PLATE_AUTO || method == TIRA_PLATE_CLOSED || method == TIRA_VOTE_SPECTRAL,
           "method must be TIRA_PLATE_AUTO, TIRA_PLATE_CLOSED (gather) or TIRA_VOTE_SPECTRAL");
    TH_TRY(check_atv_args(s0, s1, s2, w, sigma));
    return vote3_atv_run(input_field, in_layout, nullptr, nullptr, nullptr, output_field, out_layout, s0, s1, s2,
                         sigma, w, stream(), method);
}

}  // extern "C"
