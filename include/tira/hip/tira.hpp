// tira.hpp -- C++ drop-in for TIRA's CUDA entry points on gfx950.
//
// Declares, in namespace tira::cuda, the same functions (names, template
// parameters, argument order and meaning, default arguments, ownership of
// returned buffers) as the reference headers, backed by the C ABI of
// libtira_hip.so (tira_hip.h):
//   eigen.h:768-1012    evals2_symmetric, evecs2polar_symmetric,
//                       evals3_symmetric, evecs3spherical_symmetric
//   tensorvote.h:2029-2305  tensorvote_atv, tensorvote_tk, tensorvote,
//                           tensorvote_prebuilt (2D)
//   tensorvote.h:2695-2873  _tensorvote3_atv, _tensorvote3_tk, tensorvote3 (3D)
//   filter.h:304-534    gaussian_convolve2, gaussian_filter3d
//   tensorvote.h:1223-1280  build_neighbors2d (tira::cpu in the reference;
//                       the library builds the same table on the host)
// Only float instantiations exist (the reference's callers use float);
// other types fail to compile.
//
// Ownership, as in the reference: eigen outputs are new T[] when the input
// is host memory and device memory (free with tira_hip_free) when it is
// device memory; Gaussian outputs are malloc'd (host input) or device memory
// (device input); vote outputs go into caller memory.
// Errors: invalid arguments throw std::invalid_argument (the reference's
// Python layer throws on them); runtime failures print the library's message
// and continue, like HANDLE_ERROR (error.h:11-15).
#pragma once

#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../../tira_hip.h"

namespace tira {

struct Neighbor2D {  // tensorvote.h:776-786, 48 bytes
    int du, dv;
    float dx, dy;
    float c1, c2;
    float Ra, Rb, Rc;
    float Va, Vb, Vc;
};
static_assert(sizeof(Neighbor2D) == 48, "Neighbor2D must be 48 bytes");

namespace hip_detail {

inline void check(int status, const char* what) {
    if (status == TIRA_OK) return;
    const std::string msg = std::string(what) + ": " + tira_hip_last_error();
    if (status == TIRA_ERR_INVALID_ARG || status == TIRA_ERR_UNSUPPORTED) throw std::invalid_argument(msg);
    std::printf("%s\n", msg.c_str());
}

inline bool on_device(const void* p) {
    int d = 0;
    check(tira_hip_pointer_is_device(p, &d), "pointer query");
    return d != 0;
}

// Output buffer of `count` floats placed like `like`: device memory when
// `like` is device memory, else new[] (eigen) or malloc (Gaussian).
inline float* alloc_like(const void* like, size_t count, bool use_malloc) {
    if (on_device(like)) {
        void* p = nullptr;
        check(tira_hip_malloc(&p, count * sizeof(float)), "tira_hip_malloc");
        return static_cast<float*>(p);
    }
    if (use_malloc) return static_cast<float*>(std::malloc(count * sizeof(float)));
    return new float[count];
}

template <typename T>
struct is_float : std::is_same<T, float> {};

}  // namespace hip_detail

namespace cuda {

// ---------------------------------------------------------------- eigen.h
template <typename T>
T* evals2_symmetric(const T* mats, size_t n, int device) {
    static_assert(hip_detail::is_float<T>::value, "libtira_hip serves float tensors");
    float* out = hip_detail::alloc_like(mats, n * 2, false);
    hip_detail::check(tira_evals2_symmetric(mats, n, out, device), "evals2_symmetric");
    return out;
}

template <typename T>
T* evecs2polar_symmetric(const T* mats, const T* evals, size_t n, int device) {
    static_assert(hip_detail::is_float<T>::value, "libtira_hip serves float tensors");
    float* out = hip_detail::alloc_like(mats, n * 2, false);
    hip_detail::check(tira_evecs2polar_symmetric(mats, evals, n, out, device), "evecs2polar_symmetric");
    return out;
}

template <typename T>
T* evals3_symmetric(const T* mats, size_t n, int device) {
    static_assert(hip_detail::is_float<T>::value, "libtira_hip serves float tensors");
    float* out = hip_detail::alloc_like(mats, n * 3, false);
    hip_detail::check(tira_evals3_symmetric(mats, n, out, device), "evals3_symmetric");
    return out;
}

template <typename T>
T* evecs3spherical_symmetric(const T* mats, const T* lambda, size_t n, int device) {
    static_assert(hip_detail::is_float<T>::value, "libtira_hip serves float tensors");
    float* out = hip_detail::alloc_like(mats, n * 6, false);
    hip_detail::check(tira_evecs3spherical_symmetric(mats, lambda, n, out, device), "evecs3spherical_symmetric");
    return out;
}

// ---------------------------------------------------------------- tensorvote.h, 2D
template <typename Type>
static void tensorvote_tk(Type* t_out, const Type* lambdas, const Type* thetas, Type sigma1, Type sigma2,
                          unsigned power, const unsigned w, const unsigned shape0, const unsigned shape1,
                          const bool stick = true, const bool plate = true, const unsigned samples = 10) {
    static_assert(hip_detail::is_float<Type>::value, "libtira_hip serves float tensors");
    hip_detail::check(tira_tensorvote2_tk(t_out, lambdas, thetas, sigma1, sigma2, power, w, shape0, shape1, stick,
                                          plate, samples),
                      "tensorvote_tk");
}

template <typename Type>
static void tensorvote_atv(Type* t_out, const Type* lambdas, const Type* thetas, Type sigma, const unsigned w,
                           const unsigned s0, const unsigned s1) {
    static_assert(hip_detail::is_float<Type>::value, "libtira_hip serves float tensors");
    hip_detail::check(tira_tensorvote2_atv(t_out, lambdas, thetas, sigma, w, s0, s1), "tensorvote_atv");
}

template <typename Type>
static void tensorvote(Type* t_out, const Type* t_in, const Type sigma1, const Type sigma2, const int power,
                       const unsigned w, const unsigned shape0, const unsigned shape1, const bool stick = true,
                       const bool plate = true, const unsigned samples = 10, const bool is_atv = false,
                       const bool debug = false) {
    static_assert(hip_detail::is_float<Type>::value, "libtira_hip serves float tensors");
    (void)debug;
    hip_detail::check(tira_tensorvote2(t_out, t_in, sigma1, sigma2, power, w, shape0, shape1, stick, plate, samples,
                                       is_atv),
                      "tensorvote");
}

template <typename Type>
static void tensorvote_prebuilt(Type* t_out, const Type* t_in, const Neighbor2D* host_nb, int nb_count,
                                const Type sigma1, const Type sigma2, const int power, const unsigned w,
                                const unsigned shape0, const unsigned shape1, const bool stick = true,
                                const bool plate = true, const unsigned samples = 10) {
    static_assert(hip_detail::is_float<Type>::value, "libtira_hip serves float tensors");
    hip_detail::check(tira_tensorvote2_prebuilt(t_out, t_in, host_nb, nb_count, sigma1, sigma2, power, w, shape0,
                                                shape1, stick, plate, samples),
                      "tensorvote_prebuilt");
}

// ---------------------------------------------------------------- tensorvote.h, 3D
static inline void _tensorvote3_tk(float* output_field, const float* lambdas, const float* evec2,
                                   const float* evec0, unsigned int s0, unsigned int s1, unsigned int s2,
                                   float sigma1, float sigma2, unsigned int power, unsigned int w, bool STICK,
                                   bool PLATE, unsigned samples) {
    hip_detail::check(tira_tensorvote3_tk(output_field, lambdas, evec2, evec0, s0, s1, s2, sigma1, sigma2, power, w,
                                          STICK, PLATE, samples),
                      "_tensorvote3_tk");
}

static inline void _tensorvote3_atv(float* output_field, const float* lambdas, const float* evec1,
                                    const float* evec2, unsigned int s0, unsigned int s1, unsigned int s2,
                                    float sigma, unsigned int w) {
    hip_detail::check(tira_tensorvote3_atv(output_field, lambdas, evec1, evec2, s0, s1, s2, sigma, w),
                      "_tensorvote3_atv");
}

template <typename Type>
static void tensorvote3(const Type* input_field, Type* output_field, unsigned int s0, unsigned int s1,
                        unsigned int s2, float sigma, float sigma2, unsigned int w, unsigned int power, int device,
                        bool stick, bool plate, unsigned samples, bool is_atv) {
    static_assert(hip_detail::is_float<Type>::value, "libtira_hip serves float tensors");
    hip_detail::check(tira_tensorvote3(input_field, output_field, s0, s1, s2, sigma, sigma2, w, power, device, stick,
                                       plate, samples, is_atv),
                      "tensorvote3");
}

// ---------------------------------------------------------------- filter.h
template <typename ImageType, typename KernelType = float>
ImageType* gaussian_convolve2(const ImageType* source, const unsigned in_sx, const unsigned in_sy, float sigma1,
                              float sigma2, unsigned int& out_width, unsigned int& out_height) {
    static_assert(hip_detail::is_float<ImageType>::value && hip_detail::is_float<KernelType>::value,
                  "libtira_hip serves float images");
    hip_detail::check(tira_gaussian_convolve2(source, in_sx, in_sy, sigma1, sigma2, nullptr, &out_width, &out_height),
                      "gaussian_convolve2");
    float* out = hip_detail::alloc_like(source, (size_t)out_width * out_height, true);
    hip_detail::check(tira_gaussian_convolve2(source, in_sx, in_sy, sigma1, sigma2, out, &out_width, &out_height),
                      "gaussian_convolve2");
    return out;
}

template <typename ImageType, typename KernelType = float>
ImageType* gaussian_filter3d(const ImageType* source, unsigned int width, unsigned int height, unsigned int depth,
                             float sigma_w, float sigma_h, float sigma_d, unsigned int& out_width,
                             unsigned int& out_height, unsigned int& out_depth) {
    static_assert(hip_detail::is_float<ImageType>::value && hip_detail::is_float<KernelType>::value,
                  "libtira_hip serves float images");
    hip_detail::check(tira_gaussian_filter3d(source, width, height, depth, sigma_w, sigma_h, sigma_d, nullptr,
                                             &out_width, &out_height, &out_depth),
                      "gaussian_filter3d");
    float* out = hip_detail::alloc_like(source, (size_t)out_width * out_height * out_depth, true);
    hip_detail::check(tira_gaussian_filter3d(source, width, height, depth, sigma_w, sigma_h, sigma_d, out,
                                             &out_width, &out_height, &out_depth),
                      "gaussian_filter3d");
    return out;
}

}  // namespace cuda

namespace hip = cuda;

namespace cpu {
// The 2D neighbour table of tensorvote.h:1223-1280, built by the library.
inline std::vector<Neighbor2D> build_neighbors2d(const int w, const float sigma1, const float sigma2,
                                                 const unsigned power, const unsigned samples, const bool stick,
                                                 const bool plate) {
    int count = 0;
    hip_detail::check(tira_build_neighbors2d(w, sigma1, sigma2, power, samples, stick, plate, nullptr, 0, &count),
                      "build_neighbors2d");
    std::vector<Neighbor2D> nb((size_t)count);
    hip_detail::check(tira_build_neighbors2d(w, sigma1, sigma2, power, samples, stick, plate, nb.data(), count,
                                             &count),
                      "build_neighbors2d");
    nb.resize((size_t)count);
    return nb;
}
}  // namespace cpu

}  // namespace tira
