// vote3.hpp -- shared declarations of the 3D voting translation unit.
#pragma once
#include <hip/hip_runtime.h>

namespace th {

constexpr int kSumStride = 8;  // floats per offset entry of the literal-sum table

// Receiver / payload geometry of one 3D vote launch.
struct VoteGeom {
    int s0, s1, s2;      // field shape (s2 fastest)
    int r;               // window half-width, w = 2r + 1
    int P0, P1, P2;      // padded payload shape
    unsigned nbx, nby, nbz;  // receiver tiles
};

float sticknorm3_host(float s1, float s2, unsigned p);
int resolve_plate_mode(int plate_mode, unsigned power, unsigned samples);
int vote3_run(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0, float* out,
              int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
              unsigned power, bool stick, bool plate, unsigned samples, int plate_mode, hipStream_t st);

}  // namespace th
