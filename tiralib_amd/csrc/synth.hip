// synth.hip -- seeded synthetic tensor generation on the device, for
// benchmarks whose inputs are too large to generate on the host in time
// (hundreds of millions of tensors). Counter-based: tensor i depends only on
// (seed, i), so any subset can be regenerated and checked on the host.
#include "common.hpp"

namespace th {

__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

// uniform in (0, 1): 53 random bits
__device__ __forceinline__ double u01(unsigned long long seed, unsigned long long i, unsigned k) {
    const unsigned long long h = splitmix64(seed ^ splitmix64(i * 16ull + k));
    return ((double)(h >> 11) + 0.5) * (1.0 / 9007199254740992.0);
}

// Packed-6 tensors R diag(l) R^T: l log-uniform in [lo, hi] (sorted), a
// fraction `deg` with one adjacent pair at relative gap 10^U(log gap_lo,
// log gap_hi), R Haar-uniform (normalised Gaussian quaternion). Built in FP64,
// stored in FP32; `lam` (optional) receives the construction eigenvalues.
__global__ __launch_bounds__(256) void k_synth_eig(float* __restrict__ out, size_t n, unsigned long long seed,
                                                   double llo, double lhi, double deg, double glo, double ghi,
                                                   float* __restrict__ lam) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double l[3];
        for (int k = 0; k < 3; k++) l[k] = exp(llo + (lhi - llo) * u01(seed, i, k));
        double t;
        if (l[0] > l[1]) { t = l[0]; l[0] = l[1]; l[1] = t; }
        if (l[1] > l[2]) { t = l[1]; l[1] = l[2]; l[2] = t; }
        if (l[0] > l[1]) { t = l[0]; l[0] = l[1]; l[1] = t; }
        if (u01(seed, i, 3) < deg) {
            const double gap = exp(glo + (ghi - glo) * u01(seed, i, 4));
            const int j = u01(seed, i, 5) < 0.5 ? 0 : 1;
            l[j + 1] = l[j] * (1.0 + gap);
            if (l[0] > l[1]) { t = l[0]; l[0] = l[1]; l[1] = t; }
            if (l[1] > l[2]) { t = l[1]; l[1] = l[2]; l[2] = t; }
            if (l[0] > l[1]) { t = l[0]; l[0] = l[1]; l[1] = t; }
        }
        // Gaussian quaternion by Box-Muller
        double q[4];
        for (int k = 0; k < 2; k++) {
            const double r = sqrt(-2.0 * log(u01(seed, i, 6 + 2 * k)));
            double s, c;
            sincos(6.283185307179586 * u01(seed, i, 7 + 2 * k), &s, &c);
            q[2 * k] = r * c;
            q[2 * k + 1] = r * s;
        }
        const double qn = 1.0 / sqrt(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]);
        const double w = q[0] * qn, x = q[1] * qn, y = q[2] * qn, z = q[3] * qn;
        const double R[3][3] = {{1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)},
                                {2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)},
                                {2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)}};
        double A[6] = {0, 0, 0, 0, 0, 0};
        const int ri[6] = {0, 0, 0, 1, 1, 2}, ci[6] = {0, 1, 2, 1, 2, 2};
        for (int m = 0; m < 6; m++)
            for (int k = 0; k < 3; k++) A[m] += R[ri[m]][k] * l[k] * R[ci[m]][k];
        float* o = out + i * 6;
        for (int m = 0; m < 6; m++) o[m] = (float)A[m];
        if (lam) { lam[i * 3] = (float)l[0]; lam[i * 3 + 1] = (float)l[1]; lam[i * 3 + 2] = (float)l[2]; }
    }
}

// Curved plate sheets: sheet k is the quadric height field
// q_z = a q_x^2 + b q_y^2 + c q_x q_y in local coordinates q = R_k^T (p - C_k),
// thickness 1 (|F| <= |grad F| / 2), patch radius `patch`; voxels on a sheet
// hold I - n n^T with n its unit normal (the last sheet wins). Other voxels
// hold, with probability `noise`, m (I - q q^T) with q uniform on the sphere
// and m ~ U(0, 0.5), drawn from (seed, voxel).
constexpr int kMaxSheets = 32;
struct SheetSet {
    float R[kMaxSheets][9];  // row-major, columns = local axes in world coordinates
    float C[kMaxSheets][3];
    float K[kMaxSheets][3];
    int n;
    float patch2;
};

__global__ __launch_bounds__(256) void k_synth_sheets(float* __restrict__ out, unsigned s0, unsigned s1, unsigned s2,
                                                      const SheetSet S, double noise, unsigned long long seed) {
    const size_t nv = (size_t)s0 * s1 * s2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (size_t)gridDim.x * blockDim.x) {
        const float px = (float)(i % s2), py = (float)((i / s2) % s1), pz = (float)(i / ((size_t)s1 * s2));
        float n0 = 0, n1 = 0, n2 = 0, mag = 0;
        for (int k = 0; k < S.n; k++) {
            const float dx = px - S.C[k][0], dy = py - S.C[k][1], dz = pz - S.C[k][2];
            const float* R = S.R[k];
            const float qx = dx * R[0] + dy * R[3] + dz * R[6];
            const float qy = dx * R[1] + dy * R[4] + dz * R[7];
            const float qz = dx * R[2] + dy * R[5] + dz * R[8];
            const float a = S.K[k][0], b = S.K[k][1], c = S.K[k][2];
            const float f = qz - (a * qx * qx + b * qy * qy + c * qx * qy);
            const float gx = -(2.0f * a * qx + c * qy), gy = -(2.0f * b * qy + c * qx);
            const float gn = sqrtf(gx * gx + gy * gy + 1.0f);
            if (fabsf(f) <= 0.5f * gn && qx * qx + qy * qy <= S.patch2) {
                const float ig = 1.0f / gn;
                n0 = (R[0] * gx + R[1] * gy + R[2]) * ig;
                n1 = (R[3] * gx + R[4] * gy + R[5]) * ig;
                n2 = (R[6] * gx + R[7] * gy + R[8]) * ig;
                mag = 1.0f;
            }
        }
        if (mag == 0.0f && u01(seed, i, 0) < noise) {
            const double z = 2.0 * u01(seed, i, 1) - 1.0, ph = 6.283185307179586 * u01(seed, i, 2);
            const double r = sqrt(fmax(0.0, 1.0 - z * z));
            n0 = (float)(r * cos(ph)); n1 = (float)(r * sin(ph)); n2 = (float)z;
            mag = (float)(0.5 * u01(seed, i, 3));
        }
        float* o = out + i * 6;
        o[0] = mag * (1.0f - n0 * n0); o[1] = -mag * n0 * n1; o[2] = -mag * n0 * n2;
        o[3] = mag * (1.0f - n1 * n1); o[4] = -mag * n1 * n2; o[5] = mag * (1.0f - n2 * n2);
    }
}

// Vascular tree intensity volume: 1 inside any tapered capsule segment
// (p0, p1, r0, r1), 0 outside, plus Gaussian noise N(0, sigma^2) drawn from
// (seed, voxel) by Box-Muller. seg: nseg x 8 floats (p0 xyz, p1 xyz, r0, r1),
// coordinates (x, y, z) = (axis 2, axis 1, axis 0) of the volume.
__global__ __launch_bounds__(256) void k_synth_vessels(float* __restrict__ out, unsigned s0, unsigned s1, unsigned s2,
                                                       const float* __restrict__ seg, int nseg, double sigma,
                                                       unsigned long long seed) {
    const size_t nv = (size_t)s0 * s1 * s2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nv; i += (size_t)gridDim.x * blockDim.x) {
        const float px = (float)(i % s2), py = (float)((i / s2) % s1), pz = (float)(i / ((size_t)s1 * s2));
        float v = 0.0f;
        for (int k = 0; k < nseg && v == 0.0f; k++) {
            const float* q = seg + 8 * k;
            const float rm = fmaxf(q[6], q[7]);
            if (px < fminf(q[0], q[3]) - rm || px > fmaxf(q[0], q[3]) + rm || py < fminf(q[1], q[4]) - rm ||
                py > fmaxf(q[1], q[4]) + rm || pz < fminf(q[2], q[5]) - rm || pz > fmaxf(q[2], q[5]) + rm)
                continue;
            const float ux = q[3] - q[0], uy = q[4] - q[1], uz = q[5] - q[2];
            const float wx = px - q[0], wy = py - q[1], wz = pz - q[2];
            const float uu = ux * ux + uy * uy + uz * uz;
            float t = uu > 0.0f ? (wx * ux + wy * uy + wz * uz) / uu : 0.0f;
            t = fminf(1.0f, fmaxf(0.0f, t));
            const float dx = wx - t * ux, dy = wy - t * uy, dz = wz - t * uz;
            const float r = q[6] + t * (q[7] - q[6]);
            if (dx * dx + dy * dy + dz * dz <= r * r) v = 1.0f;
        }
        if (sigma > 0.0) {
            const double g = sqrt(-2.0 * log(u01(seed, i, 0))) * cos(6.283185307179586 * u01(seed, i, 1));
            v += (float)(sigma * g);
        }
        out[i] = v;
    }
}

}  // namespace th

using namespace th;

extern "C" {

int tira_synth_eig_tensors_async(float* out6, size_t n, unsigned long long seed, float lo, float hi,
                                 float degenerate_frac, float gap_lo, float gap_hi, float* lambdas) {
    TH_ARG(out6, "NULL pointer");
    TH_ARG(lo > 0 && hi >= lo && gap_lo > 0 && gap_hi >= gap_lo, "bad ranges");
    if (!n) return TIRA_OK;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(n, 256), 16384);
    k_synth_eig<<<grid, 256, 0, stream()>>>(out6, n, seed, std::log((double)lo), std::log((double)hi),
                                            degenerate_frac, std::log((double)gap_lo), std::log((double)gap_hi),
                                            lambdas);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int tira_synth_sheets_async(float* out6, unsigned s0, unsigned s1, unsigned s2, int n_sheets, const float* params,
                            float patch, float noise_frac, unsigned long long seed) {
    TH_ARG(out6 && (n_sheets == 0 || params), "NULL pointer");
    TH_ARG(n_sheets >= 0 && n_sheets <= kMaxSheets, "at most %d sheets", kMaxSheets);
    TH_ARG(noise_frac >= 0.0f && noise_frac <= 1.0f, "noise fraction must be in [0, 1]");
    SheetSet S{};
    S.n = n_sheets;
    S.patch2 = patch * patch;
    for (int k = 0; k < n_sheets; k++) {
        const float* q = params + 15 * k;
        for (int j = 0; j < 9; j++) S.R[k][j] = q[j];
        for (int j = 0; j < 3; j++) { S.C[k][j] = q[9 + j]; S.K[k][j] = q[12 + j]; }
    }
    const size_t nv = (size_t)s0 * s1 * s2;
    if (!nv) return TIRA_OK;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(nv, 256), 65536);
    k_synth_sheets<<<grid, 256, 0, stream()>>>(out6, s0, s1, s2, S, (double)noise_frac, seed);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int tira_synth_vessels_async(float* out, unsigned s0, unsigned s1, unsigned s2, const float* segments, int nseg,
                             float noise_sigma, unsigned long long seed) {
    TH_ARG(out && (nseg == 0 || segments), "NULL pointer");
    TH_ARG(nseg >= 0 && noise_sigma >= 0.0f, "bad arguments");
    TH_ARG(is_device(out) && (nseg == 0 || is_device(segments)), "the vessel generator takes device pointers");
    const size_t nv = (size_t)s0 * s1 * s2;
    if (!nv) return TIRA_OK;
    const unsigned grid = (unsigned)std::min<size_t>(ceil_div(nv, 256), 65536);
    k_synth_vessels<<<grid, 256, 0, stream()>>>(out, s0, s1, s2, segments, nseg, (double)noise_sigma, seed);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

}  // extern "C"
