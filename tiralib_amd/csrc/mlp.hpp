// mlp.hpp -- the learned plate-vote surrogate shared by the trainer (mlp.hip)
// and the vote kernel (vote3.hip).
//
// Network: 5 features -> 64 -> 64 -> 64 -> 6, ReLU, BF16 weights and
// activations, FP32 accumulation (v_mfma_f32_32x32x16_bf16). It maps the
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

namespace th {

constexpr int kMlpIn = 32;       // training input width
constexpr int kMlpHid = 64;
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
    float* bin_err = nullptr;       // kMlpBins relative errors (FP64-integrator judged)
    MlpAffine aff = {{0, 0, 0, 0, 0, 0}, {1, 1, 1, 1, 1, 1}};
    unsigned adam_t = 0;
    float last_loss = 0, last_val = 0;
};

// Packs the BF16 weights of `m` into the per-lane MFMA operand layout the
// vote kernel loads (wreg). Synchronous.
int mlp_pack_inference(PlateMlp* m, hipStream_t st);

}  // namespace th
