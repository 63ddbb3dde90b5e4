// mlp.hpp -- the learned plate-vote surrogate shared by the trainer (mlp.hip)
// and the vote kernel (vote3.hip).
//
// Network: 5 features -> 64 -> 64 -> 64 -> 6, ReLU, BF16 weights, FP32
// accumulation (v_mfma_f32_32x32x16_bf16); every hidden activation enters
// the next layer as a BF16 high part plus a BF16 residual (two MFMA passes
// over the same weights, ~16 mantissa bits: single BF16 activations cost
// about 40% of the trained network's error). It maps the
// offset of a receiver from a plate voter, expressed in the voter's plate
// frame (u, v, n) of tensorvote.h:1667-1673, to the plate vote divided by
// the attenuation c1 = exp(-|d|^2 / sigma1^2), as the 6 upper-triangle
// components (uu, uv, un, vv, vn, nn) in that frame.
//
// Features (all cheap per pair): f0..f2 = plate-frame components of the unit
// offset d, f3 = |d| / sigma1 (clamped to 3), f4 = 2 (d.n)^2 - 1 (cosine of
// twice the angle between the offset and the plate normal). Each feature
// enters as a BF16 high part and a BF16 residual that share one weight
// column, so the input carries ~16 mantissa bits instead of 8.
//
// Input slots (training width 32, inference uses the first 16):
//   0..4 high parts of f0..f4, 5..9 residuals, 10 constant 1, rest 0.
// Biases ride on constant units: input slot 10 and hidden unit 63 of every
// hidden layer are pinned to 1 (their incoming weights are frozen), so each
// layer's bias is the weight column of that unit and no separate bias add is
// needed in the MFMA pipeline.
#pragma once
#include "common.hpp"

#include <vector>

namespace th {

constexpr int kMlpIn = 32;       // training input width
constexpr int kMlpHid = 64;
constexpr int kMlpHidSplit = 2 * kMlpHid;  // BF16 per activation row: high parts, then residuals
constexpr int kMlpOutPad = 32;   // output rows of the last MFMA tile
constexpr int kMlpOut = 6;
constexpr int kMlpFeat = 5;
constexpr int kMlpConstIn = 10;  // input slot fixed at 1
constexpr int kMlpBiasUnit = 63; // hidden unit fixed at 1
constexpr float kMlpRhoMax = 3.0f;

// Parameter layout (FP32 master copy, row-major [out][in]).
constexpr int kMlpW1 = 0;
constexpr int kMlpW2 = kMlpW1 + kMlpHid * kMlpIn;
constexpr int kMlpW3 = kMlpW2 + kMlpHid * kMlpHid;
constexpr int kMlpW4 = kMlpW3 + kMlpHid * kMlpHid;
constexpr int kMlpNParam = kMlpW4 + kMlpOutPad * kMlpHid;

// Inference operand registers per lane: layer 1 (2 row tiles x 1 K step),
// layers 2 and 3 (2 x 4), layer 4 (1 x 4).
constexpr int kMlpWreg = 2 + 8 + 8 + 4;

// Per-orientation error table: plate normals binned on the sphere.
constexpr int kMlpBinsT = 32;
constexpr int kMlpBinsP = 64;
constexpr int kMlpBins = kMlpBinsT * kMlpBinsP;

__device__ __forceinline__ int mlp_bin(float nx, float ny, float nz) {
    const float pi = 3.14159265358979323846f;
    const float t = acosf(fminf(1.0f, fmaxf(-1.0f, nz)));
    const float p = atan2f(ny, nx) + pi;
    int it = (int)(t * (kMlpBinsT / pi));
    int ip = (int)(p * (kMlpBinsP / (2.0f * pi)));
    it = it < 0 ? 0 : (it >= kMlpBinsT ? kMlpBinsT - 1 : it);
    ip = ip < 0 ? 0 : (ip >= kMlpBinsP ? kMlpBinsP - 1 : ip);
    return it * kMlpBinsP + ip;
}

// Tabulated network (the vote's default evaluation of the surrogate). In the
// offset-aligned plate frame the network's inputs are functions of
// s = n.d (unit offset) and rho = |d| / sigma1 alone: f0 = sqrt(1 - s^2),
// f1 = 0, f2 = s, f3 = rho, f4 = 2 s^2 - 1. The vote of a voter therefore
// needs the network only at the distance rows of the lattice offsets and at
// s in [-1, 1]: it is evaluated once per weight change on that grid (MFMA,
// every grid point two network evaluations, at +s and -s) and written in the
// invariant form of the plate vote,
//   V / c1 = A I + B n n^T + s g (n d^T + d n^T) + D d d^T,
// with A, B, g, D functions of x = 1 - s^2 (the closed-form polynomials
// alpha, beta, gamma, delta for the exact integral). The conversion from the
// network's plate-frame outputs (uu, uv, un, vv, vn, nn):
//   A = vv, D = (uu - vv) / x, g = -D + un / (s sqrt(x)),
//   B = (nn - vv) - s^2 D - 2 s^2 g,
// from the even parts of uu, vv, nn and the odd part of un under s -> -s
// (the plate with normal -n is the same plate; uv and vn vanish for it by
// the mirror symmetry through the (u, n) plane and are dropped).
// Row layout: kMlpTabM cells at x_j = (j + 0.5) / kMlpTabM, each two float4:
// (A, B, g, D/2) at x_j and its difference to x_(j+1). A lookup is linear in
// x (extrapolated over the half cells at both ends).
constexpr int kMlpTabM = 64;

// Distance rows of the table: row 0 is the voter itself (the integrator's
// constant self vote), rows 1..K the distinct squared lattice distances l2
// with sqrt(l2) / sigma1 < 3 in ascending order, row K + 1 every offset at or
// beyond 3 sigma1 (the network's distance input is clamped to 3 there).
void mlp_shells(float sigma1, std::vector<int>& l2s);
int mlp_shell_row(const std::vector<int>& l2s, int l2);

// (A, B, g, D/2) of the tabulated network at x = 1 - s^2 in one row.
__device__ __forceinline__ float4 mlp_tab_lookup(const float4* __restrict__ row, float x) {
    const float t = fmaf(x, (float)kMlpTabM, -0.5f);
    int j = (int)t;  // truncation: t in [-0.5, 0) gives 0, the first cell
    j = min(max(j, 0), kMlpTabM - 2);
    const float f = t - (float)j;
    const float4 v = row[2 * j], dl = row[2 * j + 1];
    return make_float4(fmaf(f, dl.x, v.x), fmaf(f, dl.y, v.y), fmaf(f, dl.z, v.z), fmaf(f, dl.w, v.w));
}

// Per-component FP32 output affine: the network predicts standardized
// components, the plate-frame vote is mu + sd * output (so the BF16 output
// layer resolves each component's spread, not its mean).
struct MlpAffine {
    float mu[kMlpOut];
    float sd[kMlpOut];
};

struct PlateMlp {
    int device = -1;
    float sigma1 = 0, sigma2 = 0;
    unsigned power = 1;
    unsigned nsamp = 64;         // integrator samples behind every training target
    bool trained = false;
    // device state
    float* params = nullptr;     // kMlpNParam FP32
    float* grads = nullptr;
    float* adam_m = nullptr;
    float* adam_v = nullptr;
    unsigned char* mask = nullptr;
    unsigned short* wb = nullptr;   // BF16 weights, row-major, same layout as params
    unsigned short* wtb = nullptr;  // BF16 transposes of W2, W3, W4 (for the backward pass)
    void* wreg = nullptr;           // packed inference operands, kMlpWreg x 64 lanes x 16 B
    float* bin_err = nullptr;       // 2 x kMlpBins relative errors (FP64-integrator judged): the
                                    // network evaluated per pair, then the tabulated network
    float4* vtab = nullptr;         // tabulated network, vtab_rows x kMlpTabM x 2 float4
    int vtab_rows = 0;
    std::vector<int> shells;        // squared distances of rows 1..K (mlp_shells)
    MlpAffine aff = {{0, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 1}};
    unsigned adam_t = 0;
    unsigned long long version = 0; // bumped on every weight change
    float last_loss = 0, last_val = 0;
};

// Packs the BF16 weights of `m` into the per-lane MFMA operand layout the
// vote kernel loads (wreg) and evaluates the network into the vote table
// (vtab). Synchronous.
int mlp_pack_inference(PlateMlp* m, hipStream_t st);

}  // namespace th
