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

}  // extern "C"
