// vote3.hpp -- shared declarations of the 3D voting translation unit.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace th {

constexpr int kSumStride = 8;     // floats per offset entry of the literal-sum table
constexpr int kMlpTabStride = 12; // floats per offset entry of the tabulated-network vote

// Receiver / payload geometry of one 3D vote launch.
struct VoteGeom {
    int s0, s1, s2;      // field shape (s2 fastest)
    int r;               // window half-width, w = 2r + 1
    int P0, P1, P2;      // padded payload shape
    unsigned nbx, nby, nbz;  // receiver tiles
};

float sticknorm3_host(float s1, float s2, unsigned p);
// Closed-form plate polynomials alpha, beta, gamma, delta/2 in x = 1 - (n.d)^2.
void plate_polys(int P, double c1, double c2, std::vector<double>& al, std::vector<double>& be,
                 std::vector<double>& ga, std::vector<double>& de);
int vote3_payload_flat(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0,
                       unsigned s0, unsigned s1, unsigned s2, float4* ps, float4* pp, hipStream_t st);
bool spectral_preferred(unsigned s0, unsigned s1, unsigned s2, unsigned w, unsigned power, bool stick, bool plate,
                        unsigned samples, bool atv = false);
// Eigenvalues of the voted field requested with the vote: the spectral vote
// writes them from its last pass and sets `done`; otherwise the caller runs
// eig3 on the field.
struct EigOut {
    float* evals;
    unsigned flags;
    unsigned long long* n_refined;
    bool done;
};
// Status of vote3_spectral_run when the request is outside the spectral method.
constexpr int kSpectralNotApplicable = -1;
int vote3_spectral_run(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0,
                       float* out, int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma1,
                       float sigma2, unsigned w, unsigned power, bool stick, bool plate, unsigned samples,
                       bool atv, hipStream_t st, EigOut* eig = nullptr);
int resolve_plate_mode(int plate_mode, unsigned power, unsigned samples);
int vote3_run(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0, float* out,
              int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
              unsigned power, bool stick, bool plate, unsigned samples, int plate_mode, hipStream_t st,
              EigOut* eig = nullptr);

}  // namespace th
