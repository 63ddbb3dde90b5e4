// ref_shim.cpp -- exposes the reference's own CPU eigen / filter code with a
// C ABI so the oracle restatement can be checked against it.
//
// TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile (target `ref`) from the
// header-only sources where they lie in the TIRA tree (-I <tree>); output goes
// to oracle/_ref/ only. Nothing here is shipped. tensorvote.h / tensor.h are
// not built: they need GLM and Boost.Math, which this image does not have.
#include <cmath>    // filter.h calls std::sqrt without including <cmath>
#include <cstdlib>
#include <cstring>
#include <tira/functions/eigen.h>   // tira::cpu::evals2/3, evecs2polar, evecs3spherical
#include <tira/functions/filter.h>  // tira::cpu::gaussian1d, convolve3, gaussian_convolve3

extern "C" {

void ref_evals2_symmetric(const float* m, size_t n, float* out) {
    float* e = tira::cpu::evals2_symmetric<float>(m, n);
    std::memcpy(out, e, sizeof(float) * 2 * n);
    delete[] e;
}

void ref_evecs2polar_symmetric(const float* m, const float* ev, size_t n, float* out) {
    float* e = tira::cpu::evecs2polar_symmetric<float>(m, ev, n);
    std::memcpy(out, e, sizeof(float) * 2 * n);
    delete[] e;
}

void ref_evals3_symmetric(const float* m, size_t n, float* out) {
    float* e = tira::cpu::evals3_symmetric<float>(m, n);
    std::memcpy(out, e, sizeof(float) * 3 * n);
    delete[] e;
}

void ref_evecs3spherical_symmetric(const float* m, const float* lam, size_t n, float* out) {
    float* e = tira::cpu::evecs3spherical_symmetric<float>(m, lam, n);
    std::memcpy(out, e, sizeof(float) * 6 * n);
    delete[] e;
}

void ref_gaussian1d(unsigned k, float mu, float sigma, float dx, float* out) {
    float* g = tira::cpu::gaussian1d<float>(k, mu, sigma, dx);
    std::memcpy(out, g, sizeof(float) * k);
    delete[] g;
}

void ref_convolve3(const float* in, unsigned i0, unsigned i1, unsigned i2, const float* ker,
                   unsigned k0, unsigned k1, unsigned k2, float* out, unsigned* o0, unsigned* o1, unsigned* o2) {
    float* r = tira::cpu::convolve3<float, float>(in, i0, i1, i2, ker, k0, k1, k2, *o0, *o1, *o2);
    std::memcpy(out, r, sizeof(float) * (size_t)(*o0) * (*o1) * (*o2));
    delete[] r;
}

void ref_gaussian_convolve3(const float* in, unsigned sx, unsigned sy, unsigned sz, float sw, float sh, float sd,
                            float* out, unsigned* ox, unsigned* oy, unsigned* oz) {
    float* r = tira::cpu::gaussian_convolve3<float, float>(in, sx, sy, sz, sw, sh, sd, *ox, *oy, *oz);
    std::memcpy(out, r, sizeof(float) * (size_t)(*ox) * (*oy) * (*oz));
    delete[] r;
}

}  // extern "C"
