// mlp.hip -- training of the BF16 plate-vote surrogate on gfx950: targets
// from the FP64 N-sample plate integrator, MFMA forward / backward GEMMs,
// Adam, validation, and the per-orientation error table that decides where
// the vote falls back to the integrator.
//
// Training step on a batch of B samples (B a multiple of 256), activations
// row-major [sample][unit] in BF16:
//   H1 = relu(X W1^T), H2 = relu(H1 W2^T), H3 = relu(H2 W3^T), Y = H3 W4^T
//   dY = d loss / dY (Frobenius-weighted MSE on the 6 tensor components)
//   dW4 = dY^T H3,  dZ3 = [H3 > 0] (dY W4),  dW3 = dZ3^T H2, ... dW1 = dZ1^T X
//   Adam on the FP32 master weights, BF16 copies (and transposes) refreshed.
// "NT" products (activation x weight^T) run one 32x32 MFMA tile per wave with
// operands read straight from global memory; weight gradients ("TN",
// reducing over the batch) stage 256-sample chunks transposed in LDS and
// reduce the four waves' tiles before one atomic add per element.
#include "mlp.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <vector>

namespace th {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef unsigned short u16;

__device__ __forceinline__ u16 to_bf16(float x) { return __builtin_bit_cast(u16, (__bf16)x); }
__device__ __forceinline__ float from_bf16(u16 b) { return __uint_as_float((unsigned)b << 16); }

static u16 host_bf16(float x) {  // round to nearest even, as v_cvt_pk_bf16_f32
    unsigned u;
    std::memcpy(&u, &x, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (u16)((u >> 16) | 0x40);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (u16)(u >> 16);
}
static float host_from_bf16(u16 b) {
    unsigned u = (unsigned)b << 16;
    float f;
    std::memcpy(&f, &u, 4);
    return f;
}

// ---------------------------------------------------------------- RNG
__device__ __forceinline__ unsigned long long mlp_mix(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ double mlp_u01(unsigned long long seed, unsigned long long i, unsigned k) {
    const unsigned long long h = mlp_mix(seed ^ mlp_mix(i * 8ull + k));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// ---------------------------------------------------------------- targets
// Plate vote of a unit-normal voter n at offset vector d, divided by
// c1 = exp(-|d|^2/sigma1^2), in the offset-aligned plate frame (u, v, n):
// the literal N-sample sum of tensorvote.h:1651-1695 (q_i = cos(b_i) u +
// sin(b_i) v, b_i = i pi / N, eta = (1 - t^2)^p + (c2/c1) t^(2p)) in FP64,
// the (cos, sin) pairs by exact rotation recurrence. The integrand is a
// trigonometric polynomial of degree 2p + 2 in b, which N equispaced samples
// integrate exactly for any phase once N >= p + 2, so the sum does not
// depend on which orthonormal (u, v) spans the plate: the frame is turned
// about n so that the offset lies in the (u, n) half-plane, where the vote is
// a function of n.d and |d| alone. Writes the 5 features and the 6 local
// components.
__device__ void plate_target(const double d[3], const double n[3], double s1, double s2, unsigned power,
                             unsigned nsamp, float feat[kMlpFeat], double loc[6], double& c1) {
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    const double r = sqrt(r2);
    const double ir = r > 0 ? 1.0 / r : 0.0;
    const double dh[3] = {d[0] * ir, d[1] * ir, d[2] * ir};
    double u[3];
    {
        const double sn = dh[0] * n[0] + dh[1] * n[1] + dh[2] * n[2];
        const double w[3] = {dh[0] - sn * n[0], dh[1] - sn * n[1], dh[2] - sn * n[2]};
        const double ww = w[0] * w[0] + w[1] * w[1] + w[2] * w[2];
        if (ww > 1e-12) {
            const double iw = 1.0 / sqrt(ww);
            u[0] = w[0] * iw; u[1] = w[1] * iw; u[2] = w[2] * iw;
        } else {  // offset along the normal: the reference's basis (tensorvote.h:1667-1673)
            double a0 = 0, a1 = 0, a2 = 1;
            if (!(fabs(n[2]) < 0.999)) { a0 = 1; a2 = 0; }
            u[0] = a1 * n[2] - a2 * n[1]; u[1] = a2 * n[0] - a0 * n[2]; u[2] = a0 * n[1] - a1 * n[0];
            const double ul = 1.0 / sqrt(u[0] * u[0] + u[1] * u[1] + u[2] * u[2]);
            u[0] *= ul; u[1] *= ul; u[2] *= ul;
        }
    }
    const double v[3] = {n[1] * u[2] - n[2] * u[1], n[2] * u[0] - n[0] * u[2], n[0] * u[1] - n[1] * u[0]};
    c1 = s1 > 0 ? exp(-r2 / (s1 * s1)) : (r2 > 1e-12 ? 0.0 : 1.0);
    const double ratio = s2 > 0 ? exp(r2 / (s1 * s1) - r2 / (s2 * s2)) : 0.0;
    const unsigned p = power < 1 ? 1 : power;
    const double db = 3.14159265358979323846 / (double)nsamp;
    const double cd = cos(db), sd = sin(db);
    double cb = 1.0, sb = 0.0;
    double V[6] = {0, 0, 0, 0, 0, 0};
    for (unsigned i = 0; i < nsamp; i++) {
        const double q[3] = {cb * u[0] + sb * v[0], cb * u[1] + sb * v[1], cb * u[2] + sb * v[2]};
        const double t = q[0] * dh[0] + q[1] * dh[1] + q[2] * dh[2];
        const double t2 = t * t;
        double e1 = 1.0, e2 = 1.0;
        for (unsigned j = 0; j < p; j++) { e1 *= 1.0 - t2; e2 *= t2; }
        const double eta = e1 + ratio * e2;
        const double rr[3] = {q[0] - 2 * t * dh[0], q[1] - 2 * t * dh[1], q[2] - 2 * t * dh[2]};
        V[0] += eta * rr[0] * rr[0]; V[1] += eta * rr[0] * rr[1]; V[2] += eta * rr[0] * rr[2];
        V[3] += eta * rr[1] * rr[1]; V[4] += eta * rr[1] * rr[2]; V[5] += eta * rr[2] * rr[2];
        const double cn = cb * cd - sb * sd;
        sb = sb * cd + cb * sd;
        cb = cn;
    }
    for (int k = 0; k < 6; k++) V[k] *= db;
    // local components e_a^T V e_b, e = (u, v, n)
    const double* E[3] = {u, v, n};
    double M[3][3] = {{V[0], V[1], V[2]}, {V[1], V[3], V[4]}, {V[2], V[4], V[5]}};
    const int ia[6] = {0, 0, 0, 1, 1, 2}, ib[6] = {0, 1, 2, 1, 2, 2};
    for (int k = 0; k < 6; k++) {
        const double* x = E[ia[k]];
        const double* y = E[ib[k]];
        double s = 0;
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) s += x[i] * M[i][j] * y[j];
        loc[k] = s;
    }
    const double f2 = dh[0] * n[0] + dh[1] * n[1] + dh[2] * n[2];
    feat[0] = (float)(dh[0] * u[0] + dh[1] * u[1] + dh[2] * u[2]);
    feat[1] = (float)(dh[0] * v[0] + dh[1] * v[1] + dh[2] * v[2]);
    feat[2] = (float)f2;
    feat[3] = (float)fmin(r / s1, (double)kMlpRhoMax);
    feat[4] = (float)(2.0 * f2 * f2 - 1.0);
}

// One input row: BF16 high parts, residuals, constant slot, zeros.
__device__ __forceinline__ void write_input_row(u16* __restrict__ x, const float f[kMlpFeat]) {
    u16 row[kMlpIn];
    for (int k = 0; k < kMlpIn; k++) row[k] = 0;
    for (int k = 0; k < kMlpFeat; k++) {
        const u16 hi = to_bf16(f[k]);
        row[k] = hi;
        row[kMlpFeat + k] = to_bf16(f[k] - from_bf16(hi));
    }
    row[kMlpConstIn] = to_bf16(1.0f);
    uint4* dst = reinterpret_cast<uint4*>(x);
    const uint4* src = reinterpret_cast<const uint4*>(row);
    for (int k = 0; k < kMlpIn / 8; k++) dst[k] = src[k];
}

// Random training samples: offset uniform in the ball of radius `radius`,
// plate normal uniform on the sphere. raw (optional): d (3), n (3), c1, 0.
__global__ __launch_bounds__(256) void k_mlp_samples(size_t n, unsigned long long seed, double s1, double s2,
                                                     unsigned power, unsigned nsamp, double radius,
                                                     u16* __restrict__ X, float* __restrict__ T,
                                                     float* __restrict__ raw) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double two_pi = 6.283185307179586;
        const double r = radius * cbrt(mlp_u01(seed, i, 0));
        const double z = 2.0 * mlp_u01(seed, i, 1) - 1.0, ph = two_pi * mlp_u01(seed, i, 2);
        const double rs = sqrt(fmax(0.0, 1.0 - z * z));
        const double d[3] = {r * rs * cos(ph), r * rs * sin(ph), r * z};
        const double nz = 2.0 * mlp_u01(seed, i, 3) - 1.0, nph = two_pi * mlp_u01(seed, i, 4);
        const double ns = sqrt(fmax(0.0, 1.0 - nz * nz));
        const double nv[3] = {ns * cos(nph), ns * sin(nph), nz};
        float f[kMlpFeat];
        double loc[6], c1;
        plate_target(d, nv, s1, s2, power, nsamp, f, loc, c1);
        write_input_row(X + i * kMlpIn, f);
        float* t = T + i * 8;
        for (int k = 0; k < 6; k++) t[k] = (float)loc[k];
        t[6] = 0.0f; t[7] = 0.0f;
        if (raw) {
            float* w = raw + i * 8;
            w[0] = (float)d[0]; w[1] = (float)d[1]; w[2] = (float)d[2];
            w[3] = (float)nv[0]; w[4] = (float)nv[1]; w[5] = (float)nv[2];
            w[6] = (float)c1; w[7] = 0.0f;
        }
    }
}

// Frobenius weight of the 6 packed components (off-diagonals count twice).
__device__ __forceinline__ float comp_weight(int o) { return (o == 1 || o == 2 || o == 4) ? 2.0f : 1.0f; }

// Centre normal of orientation bin b (the bins of mlp_bin).
__device__ __forceinline__ void bin_normal(int b, double nv[3]) {
    const double pi = 3.14159265358979323846;
    const int it = b / kMlpBinsP, ip = b % kMlpBinsP;
    const double t = (it + 0.5) * (pi / kMlpBinsT);
    const double p = (ip + 0.5) * (2.0 * pi / kMlpBinsP) - pi;
    nv[0] = sin(t) * cos(p); nv[1] = sin(t) * sin(p); nv[2] = cos(t);
}

// Error-table inputs: row i = (bin b0 + i / noff, offset i % noff); rows past
// nrows_valid are zero inputs with zero weight. wgt[i] = c1^2.
__global__ __launch_bounds__(256) void k_mlp_bin_inputs(size_t nrows, size_t nvalid, int b0, int noff,
                                                        const float4* __restrict__ offs, double s1, double s2,
                                                        unsigned power, unsigned nsamp, u16* __restrict__ X,
                                                        float* __restrict__ T, float* __restrict__ wgt) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += (size_t)gridDim.x * blockDim.x) {
        float* t = T + i * 8;
        if (i >= nvalid) {
            const float z[kMlpFeat] = {0, 0, 0, 0, 0};
            write_input_row(X + i * kMlpIn, z);
            for (int k = 0; k < 8; k++) t[k] = 0.0f;
            wgt[i] = 0.0f;
            continue;
        }
        const int b = b0 + (int)(i / noff);
        const float4 o = offs[i % noff];
        double nv[3];
        bin_normal(b, nv);
        const double d[3] = {o.x, o.y, o.z};
        float f[kMlpFeat];
        double loc[6], c1;
        plate_target(d, nv, s1, s2, power, nsamp, f, loc, c1);
        write_input_row(X + i * kMlpIn, f);
        for (int k = 0; k < 6; k++) t[k] = (float)loc[k];
        t[6] = 0.0f; t[7] = 0.0f;
        wgt[i] = (float)(c1 * c1);
    }
}

// Per-row weighted errors of the error table against the FP64 integrator
// targets T: the network evaluated per pair (Y, plate-frame outputs) into
// rs_net, the tabulated network (lookup of the offset's distance row, offs.w,
// at s = n.d) into rs_tab; (sum w (pred - T)^2 c1^2, sum w T^2 c1^2) per row.
__global__ __launch_bounds__(256) void k_mlp_bin_rows(size_t nvalid, int b0, int noff, const float4* __restrict__ offs,
                                                      const float* __restrict__ Y, const float* __restrict__ T,
                                                      const float* __restrict__ wgt, MlpAffine aff,
                                                      const float4* __restrict__ vtab, float* __restrict__ rs_net,
                                                      float* __restrict__ rs_tab) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvalid; i += (size_t)gridDim.x * blockDim.x) {
        const int b = b0 + (int)(i / noff);
        const float4 o = offs[i % noff];
        double nv[3];
        bin_normal(b, nv);
        const float n0 = (float)nv[0], n1 = (float)nv[1], n2 = (float)nv[2];
        const float il = rsqrtf(o.x * o.x + o.y * o.y + o.z * o.z);
        const float s = fmaf(n0, o.x * il, fmaf(n1, o.y * il, n2 * (o.z * il)));
        const float x = fmaf(-s, s, 1.0f);
        const float4 v = mlp_tab_lookup(vtab + (size_t)o.w * 2 * kMlpTabM, x);
        const float sn = sqrtf(fmaxf(x, 0.0f));
        const float tab[kMlpOut] = {fmaf(2.0f * v.w, x, v.x), 0.0f, sn * s * (v.z + 2.0f * v.w), v.x, 0.0f,
                                    v.x + v.y + 2.0f * s * s * (v.z + v.w)};
        const float* t = T + i * 8;
        float en = 0.0f, et = 0.0f, tt = 0.0f;
        for (int c = 0; c < kMlpOut; c++) {
            const float w = comp_weight(c);
            const float yn = fmaf(aff.sd[c], Y[i * kMlpOutPad + c], aff.mu[c]);
            en += w * (yn - t[c]) * (yn - t[c]);
            et += w * (tab[c] - t[c]) * (tab[c] - t[c]);
            tt += w * t[c] * t[c];
        }
        const float rw = wgt[i];
        rs_net[2 * i] = en * rw;
        rs_net[2 * i + 1] = tt * rw;
        rs_tab[2 * i] = et * rw;
        rs_tab[2 * i + 1] = tt * rw;
    }
}

// ---------------------------------------------------------------- tabulated network
// Input rows of the vote table: row i = (distance row k, cell j, sign) with
// k = i / (2 M), j = (i / 2) % M, s = +-sqrt(1 - x_j), x_j = (j + 0.5) / M;
// the features of an offset at s in the offset-aligned frame (see
// plate_target): (sqrt(x_j), 0, s, rho_k, 2 s^2 - 1). Rows past nvalid are
// zero.
__global__ __launch_bounds__(256) void k_mlp_tab_inputs(size_t nrows, size_t nvalid, const float* __restrict__ rho,
                                                        u16* __restrict__ X) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nrows; i += (size_t)gridDim.x * blockDim.x) {
        float f[kMlpFeat] = {0, 0, 0, 0, 0};
        if (i < nvalid) {
            const size_t k = i / (2 * kMlpTabM);
            const int j = (int)((i / 2) % kMlpTabM);
            const double x = (j + 0.5) / kMlpTabM;
            const double s = (i & 1) ? -sqrt(1.0 - x) : sqrt(1.0 - x);
            f[0] = (float)sqrt(x);
            f[2] = (float)s;
            f[3] = rho[k];
            f[4] = (float)(2.0 * s * s - 1.0);
        }
        write_input_row(X + i * kMlpIn, f);
    }
}

// (A, B, g, D/2) of distance row k at cell j from the network outputs Y (the
// rows k_mlp_tab_inputs wrote), symmetrised under s -> -s.
__device__ double4 tab_coeffs(const float* __restrict__ Y, const MlpAffine& aff, int k, int j) {
    const size_t ip = ((size_t)k * kMlpTabM + j) * 2, im = ip + 1;
    double yp[kMlpOut], ym[kMlpOut];
    for (int c = 0; c < kMlpOut; c++) {
        yp[c] = (double)fmaf(aff.sd[c], Y[ip * kMlpOutPad + c], aff.mu[c]);
        ym[c] = (double)fmaf(aff.sd[c], Y[im * kMlpOutPad + c], aff.mu[c]);
    }
    const double uu = 0.5 * (yp[0] + ym[0]), un = 0.5 * (yp[2] - ym[2]);
    const double vv = 0.5 * (yp[3] + ym[3]), nn = 0.5 * (yp[5] + ym[5]);
    const double x = (j + 0.5) / kMlpTabM, s = sqrt(1.0 - x), sn = sqrt(x);
    const double D = (uu - vv) / x;
    const double g = -D + un / (s * sn);
    const double B = (nn - vv) - s * s * D - 2.0 * s * s * g;
    return make_double4(vv, B, g, 0.5 * D);
}

// Vote table rows 0..nrows-1: cell j holds (A, B, g, D/2) at x_j and the
// difference to x_(j+1). Row 0 (the voter itself, d = 0) is the integrator's
// self vote (pi / 2)(I - n n^T).
__global__ __launch_bounds__(256) void k_mlp_tab_convert(const float* __restrict__ Y, int nrows, MlpAffine aff,
                                                         float4* __restrict__ tab) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nrows * kMlpTabM) return;
    const int k = i / kMlpTabM, j = i % kMlpTabM;
    float4* cell = tab + 2 * (size_t)i;
    if (k == 0) {
        const float hp = (float)(0.5 * 3.14159265358979323846);
        cell[0] = make_float4(hp, -hp, 0.0f, 0.0f);
        cell[1] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        return;
    }
    const double4 v = tab_coeffs(Y, aff, k, j);
    double4 dl = make_double4(0.0, 0.0, 0.0, 0.0);
    if (j + 1 < kMlpTabM) {
        const double4 n = tab_coeffs(Y, aff, k, j + 1);
        dl = make_double4(n.x - v.x, n.y - v.y, n.z - v.z, n.w - v.w);
    }
    cell[0] = make_float4((float)v.x, (float)v.y, (float)v.z, (float)v.w);
    cell[1] = make_float4((float)dl.x, (float)dl.y, (float)dl.z, (float)dl.w);
}

// ---------------------------------------------------------------- GEMMs
enum { kEpiF32 = 0, kEpiReluSplit = 1, kEpiMaskBf16 = 2 };

// C[rows x cols] = A[rows x K] B[cols x K]^T (BF16 in, FP32 accumulate).
// One 32x32 tile per wave, four waves per block along rows. SPLIT_A: A holds
// hidden activations as a BF16 high part (columns 0..K-1) and a BF16
// residual (columns K..2K-1) of the same row; both meet the same B columns,
// so the product sees ~16-bit activations. Epilogues: FP32 store; ReLU
// split into high part (C[row][col]) and residual (C[row][kMlpHid + col]);
// BF16 masked by H > 0 (the ReLU derivative, on H's high part).
template <int EPI, bool SPLIT_A>
__global__ __launch_bounds__(256) void k_gemm_nt(const u16* __restrict__ A, int lda, const u16* __restrict__ B,
                                                 int ldb, int K, int rows, void* __restrict__ C, int ldc,
                                                 const u16* __restrict__ H, int ldh) {
    const int wave = threadIdx.x >> 6, l = threadIdx.x & 63;
    const int row0 = (blockIdx.x * 4 + wave) * 32;
    const int col0 = blockIdx.y * 32;
    if (row0 >= rows) return;
    const int m = l & 31, h = l >> 5;
    f32x16 acc = {};
    const u16* pa = A + (size_t)(row0 + m) * lda + 8 * h;
    const u16* pb = B + (size_t)(col0 + m) * ldb + 8 * h;
    for (int s = 0; s < K / 16; s++) {
        const bf16x8 b = *reinterpret_cast<const bf16x8*>(pb + 16 * s);
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(pa + 16 * s);
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        if (SPLIT_A) {
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(pa + K + 16 * s);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al, b, acc, 0, 0, 0);
        }
    }
#pragma unroll
    for (int r = 0; r < 16; r++) {
        const int row = row0 + 8 * (r >> 2) + 4 * h + (r & 3);
        const size_t o = (size_t)row * ldc + col0 + m;
        if (EPI == kEpiF32) {
            reinterpret_cast<float*>(C)[o] = acc[r];
        } else if (EPI == kEpiReluSplit) {
            const float v = fmaxf(acc[r], 0.0f);
            const u16 hi = to_bf16(v);
            reinterpret_cast<u16*>(C)[o] = hi;
            reinterpret_cast<u16*>(C)[o + kMlpHid] = to_bf16(v - from_bf16(hi));
        } else {
            const short hv = (short)H[(size_t)row * ldh + col0 + m];
            reinterpret_cast<u16*>(C)[o] = hv > 0 ? to_bf16(acc[r]) : (u16)0;
        }
    }
}

constexpr int kTnChunk = 256;
constexpr int kTnPad = kTnChunk + 8;
constexpr int kTnMaxSplit = 64;  // row splits of a weight-gradient product

// Weight-gradient product, split over the batch without atomics:
//   part[split][n][k] = sum over chunks c = split, split + nsplit, ... of
//                       sum_{b in chunk c} A[b, n] B[b, k]
// for the 32x32 output tile (blockIdx.y, blockIdx.z), split blockIdx.x.
// Every chunk is summed in a fixed order, so the result does not depend on
// how the blocks are scheduled; k_tn_reduce then adds the splits in order.
__global__ __launch_bounds__(256) void k_gemm_tn(const u16* __restrict__ A, int lda, const u16* __restrict__ B,
                                                 int ldb, int rows, float* __restrict__ part, int ldc,
                                                 size_t part_stride) {
    __shared__ __attribute__((aligned(16))) u16 sA[32 * kTnPad];
    __shared__ __attribute__((aligned(16))) u16 sB[32 * kTnPad];
    __shared__ float sC[4 * 32 * 32];
    const int tid = threadIdx.x;
    const int n0 = blockIdx.y * 32, k0 = blockIdx.z * 32;
    const int wave = tid >> 6, l = tid & 63, m = l & 31, h = l >> 5;
    f32x16 acc = {};
    for (int b0 = blockIdx.x * kTnChunk; b0 < rows; b0 += gridDim.x * kTnChunk) {
        for (int seg = tid; seg < kTnChunk * 4; seg += 256) {
            const int b = seg >> 2, c8 = (seg & 3) * 8;
            uint4 va = make_uint4(0, 0, 0, 0), vb = make_uint4(0, 0, 0, 0);
            if (b0 + b < rows) {
                va = *reinterpret_cast<const uint4*>(A + (size_t)(b0 + b) * lda + n0 + c8);
                vb = *reinterpret_cast<const uint4*>(B + (size_t)(b0 + b) * ldb + k0 + c8);
            }
            const u16* ea = reinterpret_cast<const u16*>(&va);
            const u16* eb = reinterpret_cast<const u16*>(&vb);
#pragma unroll
            for (int j = 0; j < 8; j++) {
                sA[(c8 + j) * kTnPad + b] = ea[j];
                sB[(c8 + j) * kTnPad + b] = eb[j];
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < kTnChunk / 64; s++) {
            const int bb = wave * (kTnChunk / 4) + 16 * s + 8 * h;
            const bf16x8 a = *reinterpret_cast<const bf16x8*>(&sA[m * kTnPad + bb]);
            const bf16x8 b = *reinterpret_cast<const bf16x8*>(&sB[m * kTnPad + bb]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < 16; r++) sC[wave * 1024 + (8 * (r >> 2) + 4 * h + (r & 3)) * 32 + m] = acc[r];
    __syncthreads();
    float* dst = part + (size_t)blockIdx.x * part_stride;
    for (int e = tid; e < 1024; e += 256) {
        const float v = ((sC[e] + sC[1024 + e]) + sC[2048 + e]) + sC[3072 + e];
        dst[(size_t)(n0 + (e >> 5)) * ldc + k0 + (e & 31)] = v;
    }
}

// C[e] = sum over s = 0 .. nsplit-1 (in order) of part[s][e].
__global__ __launch_bounds__(256) void k_tn_reduce(const float* __restrict__ part, int nsplit, size_t stride,
                                                   int n, float* __restrict__ C) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    float v = 0.0f;
    for (int s = 0; s < nsplit; s++) v += part[(size_t)s * stride + e];
    C[e] = v;
}

// ---------------------------------------------------------------- loss, Adam, fills

// Frobenius-weighted squared error on the 6 components (off-diagonals count
// twice) of mu + sd * Y against T. dY (optional) = gscale * d/dY. Block b
// writes its two sums, sum w (Y-T)^2 * rw and sum w T^2 * rw (rw = row
// weight, 1 when rowwgt is NULL), to part[2 (blk0 + b) .. + 1]; the block sum
// is a fixed tree, so the partials (and k_sum_partials over them) do not
// depend on scheduling. rowsums (optional): the same two sums per row.
__global__ __launch_bounds__(256) void k_mlp_loss(const float* __restrict__ Y, const float* __restrict__ T, int rows,
                                                  u16* __restrict__ dY, float gscale, float* __restrict__ part,
                                                  int blk0, const float* __restrict__ rowwgt,
                                                  float* __restrict__ rowsums, MlpAffine aff) {
    __shared__ float s_e[256], s_t[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    float e = 0.0f, t = 0.0f;
    if (i < rows) {
        const float rw = rowwgt ? rowwgt[i] : 1.0f;
        for (int o = 0; o < kMlpOutPad; o++) {
            if (o < kMlpOut) {
                const float w = comp_weight(o);
                const float y = fmaf(aff.sd[o], Y[(size_t)i * kMlpOutPad + o], aff.mu[o]);
                const float tt = T[(size_t)i * 8 + o];
                const float diff = y - tt;
                e += w * diff * diff * rw;
                t += w * tt * tt * rw;
                if (dY) dY[(size_t)i * kMlpOutPad + o] = to_bf16(gscale * 2.0f * w * diff * aff.sd[o]);
            } else if (dY) {
                dY[(size_t)i * kMlpOutPad + o] = 0;
            }
        }
        if (rowsums) {
            rowsums[2 * (size_t)i] = e;
            rowsums[2 * (size_t)i + 1] = t;
        }
    }
    s_e[threadIdx.x] = e;
    s_t[threadIdx.x] = t;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_e[threadIdx.x] += s_e[threadIdx.x + s];
            s_t[threadIdx.x] += s_t[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0 && part) {
        part[2 * (size_t)(blk0 + blockIdx.x)] = s_e[0];
        part[2 * (size_t)(blk0 + blockIdx.x) + 1] = s_t[0];
    }
}

// Sums pairs of partials in a fixed order: out[2 g + c] = sum over the
// `count` pairs i of group g (pairs g * count .. g * count + count - 1) of
// in[2 i + c]. One block per group; each thread adds a strided run in
// order, then a fixed tree combines the threads.
__global__ __launch_bounds__(256) void k_sum_partials(const float* __restrict__ in, size_t count,
                                                      float* __restrict__ out) {
    __shared__ float s_a[256], s_b[256];
    const float* src = in + 2 * (size_t)blockIdx.x * count;
    float a = 0.0f, b = 0.0f;
    for (size_t i = threadIdx.x; i < count; i += 256) {
        a += src[2 * i];
        b += src[2 * i + 1];
    }
    s_a[threadIdx.x] = a;
    s_b[threadIdx.x] = b;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            s_a[threadIdx.x] += s_a[threadIdx.x + s];
            s_b[threadIdx.x] += s_b[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * (size_t)blockIdx.x] = s_a[0];
        out[2 * (size_t)blockIdx.x + 1] = s_b[0];
    }
}

struct AdamArgs {
    float lr_t, b1, b2, eps_t;
};

// Adam on one layer [rows x cols] (FP32 master p), refreshing its BF16 copy
// wb and (if given) the BF16 transpose wtb. `tie`: layer 1, whose residual
// input columns kMlpFeat..2*kMlpFeat-1 share the weight of the high-part
// column: the shared weight's gradient is the sum of both columns.
__global__ __launch_bounds__(256) void k_mlp_adam(float* __restrict__ p, const float* __restrict__ g,
                                                  float* __restrict__ m, float* __restrict__ v,
                                                  const unsigned char* __restrict__ mask, int rows, int cols,
                                                  u16* __restrict__ wb, u16* __restrict__ wtb, AdamArgs a, int tie) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= rows * cols) return;
    const int row = idx / cols, col = idx % cols;
    if (tie && col >= kMlpFeat && col < 2 * kMlpFeat) return;
    float gg = g[idx];
    if (tie && col < kMlpFeat) gg += g[idx + kMlpFeat];
    float w = p[idx];
    if (mask[idx]) {
        const float mm = a.b1 * m[idx] + (1.0f - a.b1) * gg;
        const float vv = a.b2 * v[idx] + (1.0f - a.b2) * gg * gg;
        m[idx] = mm;
        v[idx] = vv;
        w -= a.lr_t * mm / (sqrtf(vv) + a.eps_t);
        p[idx] = w;
    }
    const u16 wbits = to_bf16(w);
    wb[idx] = wbits;
    if (wtb) wtb[(size_t)col * rows + row] = wbits;
    if (tie && col < kMlpFeat) {
        p[idx + kMlpFeat] = w;
        wb[idx + kMlpFeat] = wbits;
        if (wtb) wtb[(size_t)(col + kMlpFeat) * rows + row] = wbits;
    }
}

// Y[:, 0..5] = mu + sd * Y (the network's plate-frame components).
__global__ void k_mlp_affine(float* __restrict__ Y, size_t rows, MlpAffine aff) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < rows * kMlpOut;
         i += (size_t)gridDim.x * blockDim.x) {
        const size_t r = i / kMlpOut;
        const int o = (int)(i % kMlpOut);
        Y[r * kMlpOutPad + o] = fmaf(aff.sd[o], Y[r * kMlpOutPad + o], aff.mu[o]);
    }
}

__global__ void k_mlp_fill(float* __restrict__ x, size_t n, float v) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        x[i] = v;
}

// ---------------------------------------------------------------- host side
static unsigned grid_for(size_t n) { return (unsigned)std::min<size_t>(ceil_div(n, 256), 16384); }

static void fill(float* x, size_t n, float v, hipStream_t st) {
    k_mlp_fill<<<grid_for(n), 256, 0, st>>>(x, n, v);
}

struct Layer {
    int off, rows, cols;
    int wt_off;  // offset in wtb, or -1
};
static const Layer kLayers[4] = {
    {kMlpW1, kMlpHid, kMlpIn, -1},
    {kMlpW2, kMlpHid, kMlpHid, 0},
    {kMlpW3, kMlpHid, kMlpHid, kMlpHid * kMlpHid},
    {kMlpW4, kMlpOutPad, kMlpHid, 2 * kMlpHid * kMlpHid},
};
constexpr int kMlpNWt = 2 * kMlpHid * kMlpHid + kMlpHid * kMlpOutPad;

// Activations of one batch, plus the partial-sum buffers of the
// deterministic reductions: `part` holds kTnMaxSplit split tiles of every
// layer's weight gradient, `lpart` one (error, target) pair per loss block.
struct Acts {
    size_t cap = 0;
    u16 *h1 = nullptr, *h2 = nullptr, *h3 = nullptr, *dy = nullptr, *dz1 = nullptr, *dz2 = nullptr, *dz3 = nullptr;
    float* y = nullptr;
    float* sums = nullptr;
    float* part = nullptr;
    float* lpart = nullptr;
    ~Acts() { release(); hipFree(sums); hipFree(part); }
    void release() {
        hipFree(h1); hipFree(h2); hipFree(h3); hipFree(dy); hipFree(dz1); hipFree(dz2); hipFree(dz3);
        hipFree(y); hipFree(lpart);
        h1 = h2 = h3 = dy = dz1 = dz2 = dz3 = nullptr;
        y = lpart = nullptr;
    }
    int reserve(size_t rows) {
        if (rows <= cap) return TIRA_OK;
        release();
        cap = 0;
        const size_t hb = rows * kMlpHid * 2, hs = rows * kMlpHidSplit * 2;
        TH_HIP(hipMalloc(&h1, hs)); TH_HIP(hipMalloc(&h2, hs)); TH_HIP(hipMalloc(&h3, hs));
        TH_HIP(hipMalloc(&dz1, hb)); TH_HIP(hipMalloc(&dz2, hb)); TH_HIP(hipMalloc(&dz3, hb));
        TH_HIP(hipMalloc(&dy, rows * kMlpOutPad * 2));
        TH_HIP(hipMalloc(&y, rows * kMlpOutPad * 4));
        TH_HIP(hipMalloc(&lpart, ((rows + 255) / 256) * 8));
        if (!sums) TH_HIP(hipMalloc(&sums, 64));
        if (!part) TH_HIP(hipMalloc(&part, (size_t)kTnMaxSplit * kMlpNParam * 4));
        cap = rows;
        return TIRA_OK;
    }
};

static void launch_nt(int epi, bool split_a, const u16* A, int lda, const u16* B, int ldb, int K, int rows, int cols,
                      void* C, int ldc, const u16* H, int ldh, hipStream_t st) {
    const dim3 grid(ceil_div(rows, 128), cols / 32);
    if (epi == kEpiF32) {
        if (split_a) k_gemm_nt<kEpiF32, true><<<grid, 256, 0, st>>>(A, lda, B, ldb, K, rows, C, ldc, H, ldh);
        else k_gemm_nt<kEpiF32, false><<<grid, 256, 0, st>>>(A, lda, B, ldb, K, rows, C, ldc, H, ldh);
    } else if (epi == kEpiReluSplit) {
        if (split_a) k_gemm_nt<kEpiReluSplit, true><<<grid, 256, 0, st>>>(A, lda, B, ldb, K, rows, C, ldc, H, ldh);
        else k_gemm_nt<kEpiReluSplit, false><<<grid, 256, 0, st>>>(A, lda, B, ldb, K, rows, C, ldc, H, ldh);
    } else {
        k_gemm_nt<kEpiMaskBf16, false><<<grid, 256, 0, st>>>(A, lda, B, ldb, K, rows, C, ldc, H, ldh);
    }
}

// C[N x K] = A^T B over `rows` samples (N = ncols_a, K = ncols_b): split
// partial tiles into `part`, then their ordered sum into C.
static void launch_tn(const u16* A, int lda, int ncols_a, const u16* B, int ldb, int ncols_b, int rows, float* C,
                      int ldc, float* part, hipStream_t st) {
    const int nsplit = (int)std::min<size_t>(kTnMaxSplit, ceil_div(rows, kTnChunk));
    const size_t stride = (size_t)ncols_a * ldc;
    const dim3 grid(nsplit, ncols_a / 32, ncols_b / 32);
    k_gemm_tn<<<grid, 256, 0, st>>>(A, lda, B, ldb, rows, part, ldc, stride);
    const int n = ncols_a * ldc;
    k_tn_reduce<<<ceil_div(n, 256), 256, 0, st>>>(part, nsplit, stride, n, C);
}

// Loss sums of `rows` rows into sums[0..1] (deterministic: block partials,
// then their ordered sum).
static void loss_sums(const float* Y, const float* T, int rows, u16* dY, float gscale, const MlpAffine& aff,
                      Acts& a, hipStream_t st) {
    const unsigned nb = ceil_div(rows, 256);
    k_mlp_loss<<<nb, 256, 0, st>>>(Y, T, rows, dY, gscale, a.lpart, 0, nullptr, nullptr, aff);
    k_sum_partials<<<1, 256, 0, st>>>(a.lpart, nb, a.sums);
}

// Forward pass of `rows` samples (a multiple of 32): Y (rows x 32 FP32).
// Hidden activations H1..H3 are rows of kMlpHidSplit BF16: the high parts of
// the 64 units, then their residuals.
static void mlp_forward(PlateMlp* m, const u16* X, int rows, Acts& a, hipStream_t st) {
    const u16* W = m->wb;
    launch_nt(kEpiReluSplit, false, X, kMlpIn, W + kMlpW1, kMlpIn, kMlpIn, rows, kMlpHid, a.h1, kMlpHidSplit, nullptr,
              0, st);
    launch_nt(kEpiReluSplit, true, a.h1, kMlpHidSplit, W + kMlpW2, kMlpHid, kMlpHid, rows, kMlpHid, a.h2,
              kMlpHidSplit, nullptr, 0, st);
    launch_nt(kEpiReluSplit, true, a.h2, kMlpHidSplit, W + kMlpW3, kMlpHid, kMlpHid, rows, kMlpHid, a.h3,
              kMlpHidSplit, nullptr, 0, st);
    launch_nt(kEpiF32, true, a.h3, kMlpHidSplit, W + kMlpW4, kMlpHid, kMlpHid, rows, kMlpOutPad, a.y, kMlpOutPad,
              nullptr, 0, st);
}

// Forward + backward of one batch (rows a multiple of 256): gradients into
// m->grads, loss sums into a.sums. Weight gradients take the activations'
// high parts; every reduction runs in a fixed order, so the same weights and
// batch give bit-identical gradients on every run.
static void mlp_fwd_bwd(PlateMlp* m, const u16* X, const float* T, int rows, Acts& a, hipStream_t st) {
    mlp_forward(m, X, rows, a, st);
    loss_sums(a.y, T, rows, a.dy, 1.0f / (float)rows, m->aff, a, st);
    const u16* WT = m->wtb;
    float* G = m->grads;
    launch_tn(a.dy, kMlpOutPad, kMlpOutPad, a.h3, kMlpHidSplit, kMlpHid, rows, G + kMlpW4, kMlpHid, a.part, st);
    launch_nt(kEpiMaskBf16, false, a.dy, kMlpOutPad, WT + kLayers[3].wt_off, kMlpOutPad, kMlpOutPad, rows, kMlpHid,
              a.dz3, kMlpHid, a.h3, kMlpHidSplit, st);
    launch_tn(a.dz3, kMlpHid, kMlpHid, a.h2, kMlpHidSplit, kMlpHid, rows, G + kMlpW3, kMlpHid, a.part, st);
    launch_nt(kEpiMaskBf16, false, a.dz3, kMlpHid, WT + kLayers[2].wt_off, kMlpHid, kMlpHid, rows, kMlpHid, a.dz2,
              kMlpHid, a.h2, kMlpHidSplit, st);
    launch_tn(a.dz2, kMlpHid, kMlpHid, a.h1, kMlpHidSplit, kMlpHid, rows, G + kMlpW2, kMlpHid, a.part, st);
    launch_nt(kEpiMaskBf16, false, a.dz2, kMlpHid, WT + kLayers[1].wt_off, kMlpHid, kMlpHid, rows, kMlpHid, a.dz1,
              kMlpHid, a.h1, kMlpHidSplit, st);
    launch_tn(a.dz1, kMlpHid, kMlpHid, X, kMlpIn, kMlpIn, rows, G + kMlpW1, kMlpIn, a.part, st);
}

static void mlp_adam(PlateMlp* m, float lr, hipStream_t st) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    m->adam_t++;
    const double c1 = 1.0 - std::pow((double)b1, (double)m->adam_t);
    const double c2 = 1.0 - std::pow((double)b2, (double)m->adam_t);
    AdamArgs a{(float)(lr * std::sqrt(c2) / c1), b1, b2, (float)(eps * std::sqrt(c2))};
    for (int li = 0; li < 4; li++) {
        const Layer& L = kLayers[li];
        const int n = L.rows * L.cols;
        k_mlp_adam<<<ceil_div(n, 256), 256, 0, st>>>(m->params + L.off, m->grads + L.off, m->adam_m + L.off,
                                                     m->adam_v + L.off, m->mask + L.off, L.rows, L.cols,
                                                     m->wb + L.off, L.wt_off >= 0 ? m->wtb + L.wt_off : nullptr, a,
                                                     li == 0 ? 1 : 0);
    }
}

// Refreshes the BF16 copies from the FP32 master (host round trip).
static int mlp_sync_bf16(PlateMlp* m, const std::vector<float>& p, hipStream_t st) {
    std::vector<u16> wb(kMlpNParam), wt(kMlpNWt);
    for (int i = 0; i < kMlpNParam; i++) wb[i] = host_bf16(p[i]);
    for (int li = 1; li < 4; li++) {
        const Layer& L = kLayers[li];
        for (int r = 0; r < L.rows; r++)
            for (int c = 0; c < L.cols; c++) wt[L.wt_off + c * L.rows + r] = wb[L.off + r * L.cols + c];
    }
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(m->params, p.data(), kMlpNParam * 4, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(m->wb, wb.data(), kMlpNParam * 2, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(m->wtb, wt.data(), kMlpNWt * 2, hipMemcpyHostToDevice));
    return TIRA_OK;
}

static void init_params(std::vector<float>& p, std::vector<unsigned char>& mask, unsigned long long seed) {
    p.assign(kMlpNParam, 0.0f);
    mask.assign(kMlpNParam, 0);
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    for (int o = 0; o < kMlpHid; o++)
        for (int k = 0; k < kMlpIn; k++) {
            float& w = p[kMlpW1 + o * kMlpIn + k];
            if (o == kMlpBiasUnit) { w = k == kMlpConstIn ? 1.0f : 0.0f; continue; }
            if (k < kMlpFeat) { w = (float)(nd(rng) * std::sqrt(2.0 / kMlpFeat)); mask[kMlpW1 + o * kMlpIn + k] = 1; }
            else if (k == kMlpConstIn) mask[kMlpW1 + o * kMlpIn + k] = 1;
        }
    for (int o = 0; o < kMlpHid; o++)
        for (int k = 0; k < kMlpIn; k++)
            if (k >= kMlpFeat && k < 2 * kMlpFeat) p[kMlpW1 + o * kMlpIn + k] = p[kMlpW1 + o * kMlpIn + k - kMlpFeat];
    for (int base : {kMlpW2, kMlpW3})
        for (int o = 0; o < kMlpHid; o++)
            for (int k = 0; k < kMlpHid; k++) {
                float& w = p[base + o * kMlpHid + k];
                if (o == kMlpBiasUnit) { w = k == kMlpBiasUnit ? 1.0f : 0.0f; continue; }
                mask[base + o * kMlpHid + k] = 1;
                w = k == kMlpBiasUnit ? 0.0f : (float)(nd(rng) * std::sqrt(2.0 / (kMlpHid - 1)));
            }
    for (int o = 0; o < kMlpOut; o++)
        for (int k = 0; k < kMlpHid; k++) {
            mask[kMlpW4 + o * kMlpHid + k] = 1;
            p[kMlpW4 + o * kMlpHid + k] = k == kMlpBiasUnit ? 0.0f : (float)(nd(rng) * std::sqrt(1.0 / (kMlpHid - 1)));
        }
}

// ---------------------------------------------------------------- vote table
void mlp_shells(float sigma1, std::vector<int>& l2s) {
    l2s.clear();
    const int top = (int)std::ceil(9.0 * (double)sigma1 * sigma1) + 1;
    for (int l2 = 1; l2 <= top; l2++) {
        // Legendre: l2 is a sum of three squares unless l2 = 4^a (8 b + 7)
        int q = l2;
        while (q % 4 == 0) q /= 4;
        if (q % 8 == 7) continue;
        if (std::min(std::sqrt((float)l2) / sigma1, kMlpRhoMax) < kMlpRhoMax) l2s.push_back(l2);
    }
}

int mlp_shell_row(const std::vector<int>& l2s, int l2) {
    if (l2 == 0) return 0;
    const auto it = std::lower_bound(l2s.begin(), l2s.end(), l2);
    if (it != l2s.end() && *it == l2) return 1 + (int)(it - l2s.begin());
    return (int)l2s.size() + 1;
}

// Evaluates the network on the grid of the vote table (every distance row,
// kMlpTabM cells, +s and -s) with the MFMA forward pass and writes vtab.
static int mlp_build_vote_table(PlateMlp* m, hipStream_t st) {
    mlp_shells(m->sigma1, m->shells);
    const int nrows = (int)m->shells.size() + 2;
    std::vector<float> rho(nrows, 0.0f);
    for (size_t k = 0; k < m->shells.size(); k++) rho[k + 1] = std::sqrt((float)m->shells[k]) / m->sigma1;
    rho[nrows - 1] = kMlpRhoMax;
    const size_t nvalid = (size_t)nrows * kMlpTabM * 2, rows = (nvalid + 255) / 256 * 256;
    if (m->vtab_rows != nrows) {
        hipFree(m->vtab);
        m->vtab = nullptr;
        m->vtab_rows = 0;
        TH_HIP(hipMalloc(&m->vtab, (size_t)nrows * kMlpTabM * 32));
        m->vtab_rows = nrows;
    }
    float* drho = nullptr;
    u16* X = nullptr;
    Acts a;
    int rc = TIRA_OK;
    do {
        if (hipMalloc(&drho, nrows * 4) != hipSuccess || hipMalloc(&X, rows * kMlpIn * 2) != hipSuccess) {
            rc = TIRA_ERR_OUT_OF_MEMORY;
            set_error("MLP vote table: out of device memory");
            break;
        }
        if ((rc = a.reserve(rows)) != TIRA_OK) break;
        if (hipMemcpy(drho, rho.data(), nrows * 4, hipMemcpyHostToDevice) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
        k_mlp_tab_inputs<<<grid_for(rows), 256, 0, st>>>(rows, nvalid, drho, X);
        mlp_forward(m, X, (int)rows, a, st);
        k_mlp_tab_convert<<<ceil_div((size_t)nrows * kMlpTabM, 256), 256, 0, st>>>(a.y, nrows, m->aff, m->vtab);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) {
            rc = TIRA_ERR_HIP;
            set_error("MLP vote table: launch failed");
        }
    } while (0);
    hipFree(drho); hipFree(X);
    return rc;
}

// Per-lane MFMA operand layout of the vote kernel (v_mfma_f32_32x32x16_bf16:
// lane l holds A[l % 32][8 (l / 32) + j], j = 0..7). Layer 1 takes the input
// slots in natural order. Layers 2-4 take the previous layer's accumulator
// registers directly as their B operand: K step s, lane half h, element j is
// hidden unit 16 s + 8 (j / 4) + 4 h + j % 4 -- the row that accumulator
// register 8 (s % 2) + j of row tile s / 2 holds -- so the weight columns are
// stored in that order.
int mlp_pack_inference(PlateMlp* m, hipStream_t st) {
    std::vector<u16> wb(kMlpNParam);
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(wb.data(), m->wb, kMlpNParam * 2, hipMemcpyDeviceToHost));
    std::vector<u16> pk((size_t)kMlpWreg * 64 * 8, 0);
    auto put = [&](int reg, int lane, int j, u16 v) { pk[((size_t)reg * 64 + lane) * 8 + j] = v; };
    int reg = 0;
    for (int t = 0; t < 2; t++, reg++)
        for (int l = 0; l < 64; l++)
            for (int j = 0; j < 8; j++) put(reg, l, j, wb[kMlpW1 + (32 * t + l % 32) * kMlpIn + 8 * (l / 32) + j]);
    for (int base : {kMlpW2, kMlpW3})
        for (int t = 0; t < 2; t++)
            for (int s = 0; s < 4; s++, reg++)
                for (int l = 0; l < 64; l++)
                    for (int j = 0; j < 8; j++) {
                        const int u = 16 * s + 8 * (j / 4) + 4 * (l / 32) + (j % 4);
                        put(reg, l, j, wb[base + (32 * t + l % 32) * kMlpHid + u]);
                    }
    for (int s = 0; s < 4; s++, reg++)
        for (int l = 0; l < 64; l++)
            for (int j = 0; j < 8; j++) {
                const int u = 16 * s + 8 * (j / 4) + 4 * (l / 32) + (j % 4);
                put(reg, l, j, wb[kMlpW4 + (l % 32) * kMlpHid + u]);
            }
    if (!m->wreg) TH_HIP(hipMalloc(&m->wreg, pk.size() * 2));
    TH_HIP(hipMemcpy(m->wreg, pk.data(), pk.size() * 2, hipMemcpyHostToDevice));
    TH_TRY(mlp_build_vote_table(m, st));
    m->version++;
    return TIRA_OK;
}

static int mlp_alloc(PlateMlp* m) {
    TH_HIP(hipMalloc(&m->params, kMlpNParam * 4));
    TH_HIP(hipMalloc(&m->grads, kMlpNParam * 4));
    TH_HIP(hipMalloc(&m->adam_m, kMlpNParam * 4));
    TH_HIP(hipMalloc(&m->adam_v, kMlpNParam * 4));
    TH_HIP(hipMalloc(&m->mask, kMlpNParam));
    TH_HIP(hipMalloc(&m->wb, kMlpNParam * 2));
    TH_HIP(hipMalloc(&m->wtb, kMlpNWt * 2));
    TH_HIP(hipMalloc(&m->bin_err, 2 * kMlpBins * 4));
    return TIRA_OK;
}

static void mlp_free(PlateMlp* m) {
    hipFree(m->params); hipFree(m->grads); hipFree(m->adam_m); hipFree(m->adam_v); hipFree(m->mask);
    hipFree(m->wb); hipFree(m->wtb); hipFree(m->wreg); hipFree(m->bin_err); hipFree(m->vtab);
}

static int mlp_reset(PlateMlp* m, unsigned long long seed, hipStream_t st) {
    std::vector<float> p;
    std::vector<unsigned char> mask;
    init_params(p, mask, seed);
    TH_TRY(mlp_sync_bf16(m, p, st));
    TH_HIP(hipMemcpy(m->mask, mask.data(), kMlpNParam, hipMemcpyHostToDevice));
    std::vector<float> z(kMlpNParam, 0.0f), ones(2 * kMlpBins, 1.0f);
    TH_HIP(hipMemcpy(m->adam_m, z.data(), kMlpNParam * 4, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(m->adam_v, z.data(), kMlpNParam * 4, hipMemcpyHostToDevice));
    TH_HIP(hipMemcpy(m->bin_err, ones.data(), 2 * kMlpBins * 4, hipMemcpyHostToDevice));
    m->adam_t = 0;
    m->trained = false;
    m->aff = MlpAffine{{0, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 1}};
    return mlp_pack_inference(m, st);
}

// Relative Frobenius error against the integrator for every orientation bin
// of the plate normal: bin centre normal, every lattice offset
// 0 < |d| <= 3 sigma1, weighted by c1^2 (the squared attenuation the vote
// applies); for the network evaluated per pair (bin_err[0 .. kMlpBins)) and
// for the tabulated network (bin_err[kMlpBins .. 2 kMlpBins)). Every sum runs
// in a fixed order (per-row errors, then one ordered block sum per bin).
static int mlp_bin_table(PlateMlp* m, hipStream_t st) {
    std::vector<float4> offs;
    const double R = kMlpRhoMax * m->sigma1;
    const int ri = (int)std::floor(R);
    for (int dw = -ri; dw <= ri; dw++)
        for (int dv = -ri; dv <= ri; dv++)
            for (int du = -ri; du <= ri; du++) {
                const int l2 = du * du + dv * dv + dw * dw;
                const double r2 = (double)l2;
                if (r2 == 0.0 || r2 > R * R) continue;
                offs.push_back(make_float4((float)du, (float)dv, (float)dw, (float)mlp_shell_row(m->shells, l2)));
            }
    const int noff = (int)offs.size();
    if (noff == 0) return TIRA_OK;
    const int bins_per = std::max(1, std::min(kMlpBins, (int)((size_t)1 << 20) / noff));
    const size_t cap = ((size_t)bins_per * noff + 255) / 256 * 256;
    float4* doffs = nullptr;
    u16* X = nullptr;
    float *T = nullptr, *W = nullptr, *pb = nullptr, *rs = nullptr;
    Acts a;
    int rc = TIRA_OK;
    do {
        if (hipMalloc(&doffs, offs.size() * 16) != hipSuccess || hipMalloc(&X, cap * kMlpIn * 2) != hipSuccess ||
            hipMalloc(&T, cap * 32) != hipSuccess || hipMalloc(&W, cap * 4) != hipSuccess ||
            hipMalloc(&rs, cap * 16) != hipSuccess || hipMalloc(&pb, (size_t)kMlpBins * 16) != hipSuccess) {
            rc = TIRA_ERR_OUT_OF_MEMORY;
            set_error("MLP error table: out of device memory");
            break;
        }
        if ((rc = a.reserve(cap)) != TIRA_OK) break;
        if (hipMemcpy(doffs, offs.data(), offs.size() * 16, hipMemcpyHostToDevice) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
        float* rs_tab = rs + 2 * cap;
        for (int b0 = 0; b0 < kMlpBins; b0 += bins_per) {
            const int nb = std::min(bins_per, kMlpBins - b0);
            const size_t valid = (size_t)nb * noff, rows = (valid + 255) / 256 * 256;
            k_mlp_bin_inputs<<<grid_for(rows), 256, 0, st>>>(rows, valid, b0, noff, doffs, m->sigma1, m->sigma2,
                                                             m->power, m->nsamp, X, T, W);
            mlp_forward(m, X, (int)rows, a, st);
            k_mlp_bin_rows<<<grid_for(valid), 256, 0, st>>>(valid, b0, noff, doffs, a.y, T, W, m->aff, m->vtab, rs,
                                                            rs_tab);
            k_sum_partials<<<nb, 256, 0, st>>>(rs, noff, pb + 2 * (size_t)b0);
            k_sum_partials<<<nb, 256, 0, st>>>(rs_tab, noff, pb + 2 * ((size_t)kMlpBins + b0));
        }
        std::vector<float> h((size_t)kMlpBins * 4), e(2 * kMlpBins);
        if (hipStreamSynchronize(st) != hipSuccess ||
            hipMemcpy(h.data(), pb, h.size() * 4, hipMemcpyDeviceToHost) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
        for (int b = 0; b < 2 * kMlpBins; b++) e[b] = h[2 * b + 1] > 0 ? std::sqrt(h[2 * b] / h[2 * b + 1]) : 0.0f;
        if (hipMemcpy(m->bin_err, e.data(), 2 * kMlpBins * 4, hipMemcpyHostToDevice) != hipSuccess) rc = TIRA_ERR_HIP;
    } while (0);
    hipFree(doffs); hipFree(X); hipFree(T); hipFree(W); hipFree(pb); hipFree(rs);
    return rc;
}

static int mlp_train(PlateMlp* m, size_t nsamples, unsigned epochs, unsigned batch, float lr, float lr_final,
                     unsigned long long seed, float* loss_out, float* val_out, hipStream_t st) {
    const size_t B = batch;
    const size_t steps = nsamples / B;
    const size_t nval = 1 << 18, vchunk = 65536;
    u16 *X = nullptr, *XV = nullptr;
    float *T = nullptr, *TV = nullptr, *vpart = nullptr;
    Acts a;
    int rc = TIRA_OK;
    do {
        if (hipMalloc(&X, steps * B * kMlpIn * 2) != hipSuccess || hipMalloc(&T, steps * B * 32) != hipSuccess ||
            hipMalloc(&XV, nval * kMlpIn * 2) != hipSuccess || hipMalloc(&TV, nval * 32) != hipSuccess ||
            hipMalloc(&vpart, (nval / 256) * 8) != hipSuccess) {
            rc = TIRA_ERR_OUT_OF_MEMORY;
            set_error("MLP training: out of device memory for %zu samples", steps * B);
            break;
        }
        if ((rc = a.reserve(std::max(B, vchunk))) != TIRA_OK) break;
        const double radius = kMlpRhoMax * (double)m->sigma1;
        k_mlp_samples<<<grid_for(steps * B), 256, 0, st>>>(steps * B, seed, m->sigma1, m->sigma2, m->power,
                                                          m->nsamp, radius, X, T, nullptr);
        k_mlp_samples<<<grid_for(nval), 256, 0, st>>>(nval, seed ^ 0x5bd1e9955bd1e995ull, m->sigma1, m->sigma2,
                                                     m->power, m->nsamp, radius, XV, TV, nullptr);
        if (hipGetLastError() != hipSuccess) { rc = TIRA_ERR_HIP; set_error("MLP sample launch failed"); break; }
        {
            // output standardization from the first 65536 training targets
            const size_t ns = std::min<size_t>(steps * B, 65536);
            std::vector<float> th(ns * 8);
            if (hipStreamSynchronize(st) != hipSuccess ||
                hipMemcpy(th.data(), T, ns * 32, hipMemcpyDeviceToHost) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
            for (int o = 0; o < kMlpOut; o++) {
                double s1 = 0, s2 = 0;
                for (size_t i = 0; i < ns; i++) { s1 += th[i * 8 + o]; s2 += (double)th[i * 8 + o] * th[i * 8 + o]; }
                const double mu = s1 / ns, var = std::max(0.0, s2 / ns - mu * mu);
                m->aff.mu[o] = (float)mu;
                m->aff.sd[o] = (float)std::max(std::sqrt(var), 1e-3 * (std::fabs(mu) + 1e-3));
            }
        }
        const size_t total = (size_t)epochs * steps;
        const size_t warm = std::min<size_t>(200, total / 10 + 1);
        float loss_sum = 0.0f, tgt_sum = 0.0f;
        std::vector<float> hs(2);
        for (size_t it = 0; it < total; it++) {
            const size_t s = it % steps;
            const double fl = lr_final > 0.0f ? (double)lr_final / (double)lr : 0.01;
            double f = it < warm ? (double)(it + 1) / (double)warm
                                 : fl + (1.0 - fl) * 0.5 * (1.0 + std::cos(M_PI * (double)(it - warm) / (double)(total - warm)));
            mlp_fwd_bwd(m, X + s * B * kMlpIn, T + s * B * 8, (int)B, a, st);
            mlp_adam(m, (float)(lr * f), st);
            if (it + steps >= total && s + 1 == steps) {
                // loss of the last epoch's final batch
                if (hipMemcpyAsync(hs.data(), a.sums, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
                    hipStreamSynchronize(st) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
                loss_sum = hs[0];
                tgt_sum = hs[1];
            }
        }
        if (rc != TIRA_OK) break;
        if (hipGetLastError() != hipSuccess) { rc = TIRA_ERR_HIP; set_error("MLP training launch failed"); break; }
        // validation: block partials of every chunk, then one ordered sum
        for (size_t v0 = 0; v0 < nval; v0 += vchunk) {
            mlp_forward(m, XV + v0 * kMlpIn, (int)vchunk, a, st);
            k_mlp_loss<<<vchunk / 256, 256, 0, st>>>(a.y, TV + v0 * 8, (int)vchunk, nullptr, 0.0f, vpart,
                                                     (int)(v0 / 256), nullptr, nullptr, m->aff);
        }
        k_sum_partials<<<1, 256, 0, st>>>(vpart, nval / 256, a.sums);
        if (hipMemcpyAsync(hs.data(), a.sums, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) { rc = TIRA_ERR_HIP; break; }
        m->last_loss = tgt_sum > 0 ? std::sqrt(loss_sum / tgt_sum) : 0.0f;
        m->last_val = hs[1] > 0 ? std::sqrt(hs[0] / hs[1]) : 0.0f;
        if (loss_out) *loss_out = m->last_loss;
        if (val_out) *val_out = m->last_val;
    } while (0);
    hipFree(X); hipFree(T); hipFree(XV); hipFree(TV); hipFree(vpart);
    if (rc != TIRA_OK) return rc;
    TH_TRY(mlp_pack_inference(m, st));
    TH_TRY(mlp_bin_table(m, st));
    m->trained = true;
    return TIRA_OK;
}

static int mlp_check(void* h, PlateMlp** out) {
    TH_ARG(h, "NULL MLP handle");
    PlateMlp* m = static_cast<PlateMlp*>(h);
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    TH_ARG(dev == m->device, "MLP handle belongs to device %d (current %d)", m->device, dev);
    *out = m;
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_plate_mlp_create(void** handle, float sigma1, float sigma2, unsigned power, unsigned integrator_samples,
                          unsigned long long seed) {
    TH_ARG(handle, "NULL pointer");
    *handle = nullptr;
    TH_ARG(sigma1 > 0.0f && sigma2 >= 0.0f, "sigma1 must be > 0 and sigma2 >= 0");
    TH_ARG(power >= 1 && power <= 8, "power must be 1..8");
    TH_ARG(integrator_samples >= 1 && integrator_samples <= 4096, "integrator samples must be 1..4096");
    PlateMlp* m = new PlateMlp();
    m->sigma1 = sigma1;
    m->sigma2 = sigma2;
    m->power = power;
    m->nsamp = integrator_samples;
    int rc = TIRA_OK;
    if (hipGetDevice(&m->device) != hipSuccess) rc = TIRA_ERR_HIP;
    if (rc == TIRA_OK) rc = mlp_alloc(m);
    if (rc == TIRA_OK) rc = mlp_reset(m, seed, stream());
    if (rc != TIRA_OK) {
        mlp_free(m);
        delete m;
        return rc;
    }
    *handle = m;
    return TIRA_OK;
}

int tira_plate_mlp_destroy(void* handle) {
    if (!handle) return TIRA_OK;
    PlateMlp* m = static_cast<PlateMlp*>(handle);
    (void)hipStreamSynchronize(stream());
    mlp_free(m);
    delete m;
    return TIRA_OK;
}

int tira_plate_mlp_train(void* handle, size_t nsamples, unsigned epochs, unsigned batch, float lr, float lr_final,
                         unsigned long long seed, float* train_rel_err, float* val_rel_err) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(batch >= 256 && batch % 256 == 0, "batch must be a positive multiple of 256 (got %u)", batch);
    TH_ARG(nsamples >= batch, "need at least one batch of samples");
    TH_ARG(epochs >= 1, "epochs must be >= 1");
    TH_ARG(lr > 0.0f && lr_final >= 0.0f && lr_final <= lr, "need lr > 0 and 0 <= lr_final <= lr");
    return mlp_train(m, nsamples, epochs, batch, lr, lr_final, seed, train_rel_err, val_rel_err, stream());
}

int tira_plate_mlp_params(void* handle, float* get, const float* set) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    hipStream_t st = stream();
    if (set) {
        std::vector<float> p(set, set + kMlpNParam);
        TH_TRY(mlp_sync_bf16(m, p, st));
        TH_TRY(mlp_pack_inference(m, st));
    }
    if (get) {
        TH_HIP(hipStreamSynchronize(st));
        TH_HIP(hipMemcpy(get, m->params, kMlpNParam * 4, hipMemcpyDeviceToHost));
    }
    return TIRA_OK;
}

int tira_plate_mlp_samples(void* handle, size_t n, unsigned long long seed, void* X, float* T, float* raw) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(X && T, "NULL pointer");
    TH_ARG(is_device(X) && is_device(T) && (!raw || is_device(raw)), "MLP samples are written to device memory");
    if (!n) return TIRA_OK;
    k_mlp_samples<<<grid_for(n), 256, 0, stream()>>>(n, seed, m->sigma1, m->sigma2, m->power, m->nsamp,
                                                    kMlpRhoMax * (double)m->sigma1, (u16*)X, T, raw);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int tira_plate_mlp_forward(void* handle, const void* X, size_t n, float* Y) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(X && Y && is_device(X) && is_device(Y), "MLP forward takes device pointers");
    TH_ARG(n % 32 == 0, "row count must be a multiple of 32 (got %zu)", n);
    if (!n) return TIRA_OK;
    Acts a;
    TH_TRY(a.reserve(std::min<size_t>(n, 1 << 20)));
    hipStream_t st = stream();
    for (size_t r0 = 0; r0 < n; r0 += a.cap) {
        const size_t rows = std::min(a.cap, n - r0);
        mlp_forward(m, (const u16*)X + r0 * kMlpIn, (int)rows, a, st);
        k_mlp_affine<<<grid_for(rows * kMlpOut), 256, 0, st>>>(a.y, rows, m->aff);
        TH_HIP(hipMemcpyAsync(Y + r0 * kMlpOutPad, a.y, rows * kMlpOutPad * 4, hipMemcpyDeviceToDevice, st));
    }
    TH_HIP(hipStreamSynchronize(st));
    return TIRA_OK;
}

int tira_plate_mlp_gradients(void* handle, const void* X, const float* T, size_t n, float* grads, float* sums) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(X && T && grads && is_device(X) && is_device(T), "device X / T and a host gradient buffer expected");
    TH_ARG(n > 0 && n % 256 == 0 && n <= (1u << 24), "row count must be a positive multiple of 256");
    Acts a;
    TH_TRY(a.reserve(n));
    hipStream_t st = stream();
    mlp_fwd_bwd(m, (const u16*)X, T, (int)n, a, st);
    TH_HIP(hipGetLastError());
    TH_HIP(hipStreamSynchronize(st));
    TH_HIP(hipMemcpy(grads, m->grads, kMlpNParam * 4, hipMemcpyDeviceToHost));
    if (sums) TH_HIP(hipMemcpy(sums, a.sums, 8, hipMemcpyDeviceToHost));
    return TIRA_OK;
}

int tira_plate_mlp_step(void* handle, const void* X, const float* T, size_t n, float lr) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(X && T && is_device(X) && is_device(T), "MLP step takes device pointers");
    TH_ARG(n > 0 && n % 256 == 0 && n <= (1u << 24), "row count must be a positive multiple of 256");
    Acts a;
    TH_TRY(a.reserve(n));
    hipStream_t st = stream();
    mlp_fwd_bwd(m, (const u16*)X, T, (int)n, a, st);
    mlp_adam(m, lr, st);
    TH_HIP(hipGetLastError());
    TH_HIP(hipStreamSynchronize(st));
    return TIRA_OK;
}

int tira_plate_mlp_affine(void* handle, float* get, const float* set) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    if (set) {
        for (int o = 0; o < kMlpOut; o++)
            TH_ARG(std::isfinite(set[o]) && std::isfinite(set[kMlpOut + o]) && set[kMlpOut + o] != 0.0f,
                   "affine: finite means and non-zero scales expected");
        for (int o = 0; o < kMlpOut; o++) {
            m->aff.mu[o] = set[o];
            m->aff.sd[o] = set[kMlpOut + o];
        }
        TH_TRY(mlp_build_vote_table(m, stream()));
        m->version++;
    }
    if (get)
        for (int o = 0; o < kMlpOut; o++) {
            get[o] = m->aff.mu[o];
            get[kMlpOut + o] = m->aff.sd[o];
        }
    return TIRA_OK;
}

int tira_plate_mlp_bin_errors(void* handle, float* errs) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(errs, "NULL pointer");
    TH_HIP(hipStreamSynchronize(stream()));
    TH_HIP(hipMemcpy(errs, m->bin_err, kMlpBins * 4, hipMemcpyDeviceToHost));
    return TIRA_OK;
}

int tira_plate_mlp_table_errors(void* handle, float* errs) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(errs, "NULL pointer");
    TH_HIP(hipStreamSynchronize(stream()));
    TH_HIP(hipMemcpy(errs, m->bin_err + kMlpBins, kMlpBins * 4, hipMemcpyDeviceToHost));
    return TIRA_OK;
}

int tira_plate_mlp_set_vote_table(void* handle, const float* tab, size_t nfloats) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    TH_ARG(tab, "NULL pointer");
    const size_t n = (size_t)m->vtab_rows * kMlpTabM * 8;
    TH_ARG(nfloats == n, "vote table has %zu floats (got %zu)", n, nfloats);
    TH_HIP(hipStreamSynchronize(stream()));
    TH_HIP(hipMemcpy(m->vtab, tab, n * 4, hipMemcpyHostToDevice));
    m->version++;
    return TIRA_OK;
}

int tira_plate_mlp_vote_table(void* handle, float* tab, size_t cap_floats, int* rows, int* cells, int* shells) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    if (rows) *rows = m->vtab_rows;
    if (cells) *cells = kMlpTabM;
    if (shells)
        for (size_t k = 0; k < m->shells.size(); k++) shells[k] = m->shells[k];
    if (tab) {
        const size_t n = (size_t)m->vtab_rows * kMlpTabM * 8;
        TH_ARG(cap_floats >= n, "vote table needs %zu floats (got %zu)", n, cap_floats);
        TH_HIP(hipStreamSynchronize(stream()));
        TH_HIP(hipMemcpy(tab, m->vtab, n * 4, hipMemcpyDeviceToHost));
    }
    return TIRA_OK;
}

int tira_plate_mlp_refresh_bins(void* handle) {
    PlateMlp* m;
    TH_TRY(mlp_check(handle, &m));
    return mlp_bin_table(m, stream());
}

}  // extern "C"
