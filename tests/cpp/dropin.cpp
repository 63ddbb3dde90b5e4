// Drop-in check of include/tira/hip/tira.hpp: reference-style call sites
// (tira::cuda::..., host buffers in, new[] / malloc'd buffers out) against
// the library, printing results the Python test compares with the oracle.
// Usage: dropin <case>; writes whitespace-separated floats to stdout.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "tira/hip/tira.hpp"

int main(int argc, char** argv) {
    const char* which = argc > 1 ? argv[1] : "eig3";
    if (!std::strcmp(which, "eig3")) {
        // two column-major 3x3 symmetric tensors
        const float m[18] = {4, 1, 0, 1, 3, 0, 0, 0, 2, 2, 0, 0, 0, 2, 0, 0, 0, 5};
        float* ev = tira::cuda::evals3_symmetric<float>(m, 2, 0);
        for (int i = 0; i < 6; i++) std::printf("%.9g ", ev[i]);
        delete[] ev;
    } else if (!std::strcmp(which, "gauss3")) {
        const unsigned W = 20, H = 18, D = 16;
        std::vector<float> src(W * H * D);
        for (size_t i = 0; i < src.size(); i++) src[i] = (float)((i * 2654435761u) % 1000) / 1000.0f;
        unsigned ow, oh, od;
        float* out = tira::cuda::gaussian_filter3d<float>(src.data(), W, H, D, 1.0f, 1.0f, 1.0f, ow, oh, od);
        std::printf("%u %u %u ", ow, oh, od);
        for (unsigned i = 0; i < ow * oh * od; i++) std::printf("%.9g ", out[i]);
        std::free(out);
    } else if (!std::strcmp(which, "vote3")) {
        const unsigned S = 12;
        std::vector<float> in(S * S * S * 9, 0.0f);
        for (unsigned i = 0; i < S * S * S; i += 7) { in[i * 9 + 0] = 1.0f; in[i * 9 + 4] = 0.5f; in[i * 9 + 8] = 0.25f; }
        std::vector<float> out(S * S * S * 9);
        tira::cuda::tensorvote3<float>(in.data(), out.data(), S, S, S, 1.5f, 0.0f, 7, 1, 0, true, true, 20, false);
        for (float v : out) std::printf("%.9g ", v);
    } else if (!std::strcmp(which, "atv3")) {
        const unsigned S = 12;
        std::vector<float> in(S * S * S * 9, 0.0f);
        for (unsigned i = 0; i < S * S * S; i += 5) { in[i * 9 + 0] = 1.0f; in[i * 9 + 4] = 0.5f; in[i * 9 + 8] = 0.25f; }
        std::vector<float> out(S * S * S * 9);
        tira::cuda::tensorvote3<float>(in.data(), out.data(), S, S, S, 1.5f, 0.0f, 7, 1, 0, true, true, 20, true);
        for (float v : out) std::printf("%.9g ", v);
    } else if (!std::strcmp(which, "atv2")) {
        const unsigned H = 10, W = 13;
        std::vector<float> in(H * W * 4), out(H * W * 4);
        for (unsigned i = 0; i < H * W; i++) {
            const float t = 0.37f * (float)i, l0 = 0.1f + 0.01f * (float)(i % 7), l1 = 1.0f;
            const float c = std::cos(t), sn = std::sin(t);
            in[i * 4 + 0] = l0 * sn * sn + l1 * c * c;
            in[i * 4 + 1] = in[i * 4 + 2] = (l1 - l0) * c * sn;
            in[i * 4 + 3] = l0 * c * c + l1 * sn * sn;
        }
        tira::cuda::tensorvote<float>(out.data(), in.data(), 2.0f, 0.0f, 1, 9, H, W, true, true, 10, true);
        for (float v : in) std::printf("%.9g ", v);
        for (float v : out) std::printf("%.9g ", v);
    } else if (!std::strcmp(which, "nb2")) {
        auto nb = tira::cpu::build_neighbors2d(7, 2.0f, 0.0f, 1, 10, true, true);
        std::printf("%zu ", nb.size());
        for (auto& e : nb) std::printf("%d %d %.9g %.9g ", e.du, e.dv, e.c1, e.Va);
    } else {
        try {
            tira::cuda::tensorvote3<float>(nullptr, nullptr, 4, 4, 4, 1.0f, 0.0f, 4, 1, 0, true, true, 20, false);
        } catch (const std::invalid_argument& e) {
            std::printf("invalid_argument");
            return 0;
        }
        return 1;
    }
    return 0;
}
