/*
 * tira_oracle.h -- CPU oracle for the tensor-voting hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the checker for the HIP product
 * path (tiralib_amd/libtira_hip.so). Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it. Nothing in the product links it.
 *
 * Two flavours:
 *   *_f32  reference-faithful FP32 restatements of TIRA's CPU path
 *          (same operation order, same thresholds and quirks);
 *   *_f64  independent FP64 routes (Jacobi eigen-solver, FP64 literal
 *          integration) used to state absolute error.
 *
 * Layouts follow the reference:
 *   2D field : H x W x 4 floats [a, b, b, c] per pixel (tensorvote_tk.cu:9-12)
 *   3D field : s0 x s1 x s2 x 9 floats, column-major glm::mat3, s2 fastest
 *              (tensorvote.h:1460, 2420)
 *
 * Pinning status per function is recorded in the header comment of
 * tira_oracle.c.
 */
#ifndef TIRA_ORACLE_H
#define TIRA_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* 48-byte 2D neighbour record, field order of tira::Neighbor2D
 * (tensorvote.h:776-786). */
typedef struct {
    int32_t du, dv;
    float dx, dy, c1, c2, Ra, Rb, Rc, Va, Vb, Vc;
} oracle_neighbor2d;

/* 3D neighbour record of tira::Neighbor3D (tensorvote.h:1147-1151). */
typedef struct {
    int32_t du, dv, dw;
    float d[3];
    float c1, c2;
} oracle_neighbor3d;

/* ---------------- eigen (eigen.h) ---------------- */
void oracle_eval2_f32(float a, float b, float c, float* e0, float* e1);
void oracle_evec2polar_f32(float a, float b, float c, const float* lambdas, float* th0, float* th1);
void oracle_evals2_symmetric(const float* mats4, size_t n, float* evals2);
void oracle_evecs2polar_symmetric(const float* mats4, const float* evals2, size_t n, float* thetas2);

void oracle_eval3_f32(float a, float b, float c, float d, float e, float f,
                      float* l0, float* l1, float* l2);
void oracle_evec3_f32(float a, float b, float c, float d, float e, float f,
                      const float* evals, float* v0, float* v1, float* v2);
void oracle_evals3_symmetric(const float* mats9, size_t n, float* evals3);
void oracle_evecs3spherical_symmetric(const float* mats9, const float* lambdas3, size_t n, float* evecs6);
void oracle_spherical_to_cartesian(float theta, float phi, float* xyz);

/* FP64 cyclic Jacobi on the packed upper triangle (xx, xy, xz, yy, yz, zz).
 * evals ascending; evecs column k = eigenvector of evals[k] (3x3 row-major,
 * evecs[r*3+k]); eigenvector sign canonicalised (largest |component| > 0). */
void oracle_eig3_jacobi_f64(const double* a6, double* evals, double* evecs);
/* Batch: mats9 (reference 9-float layout, slots 0,1,2,4,5,8 read). */
void oracle_eig3_f64(const float* mats9, size_t n, double* evals3, double* evec0, double* evec2);
/* Batch on packed-6 input (xx, xy, xz, yy, yz, zz). */
void oracle_eig3_packed6_f64(const float* t6, size_t n, double* evals3);

/* ---------------- 2D tensor voting (tensorvote.h) ---------------- */
int  oracle_build_neighbors2d(int w, float sigma1, float sigma2, unsigned power, unsigned samples,
                              int stick, int plate, oracle_neighbor2d* out);
float oracle_tk_stick_norm_2d(float sigma1, float sigma2, unsigned power);
void oracle_tk_stickvote_2d(float rx, float ry, float sigma1, float sigma2, unsigned power, float theta,
                            float* Va, float* Vb, float* Vc);
void oracle_tk_plate_numerical_2d(float rx, float ry, float sigma1, float sigma2, unsigned power,
                                  float* Va, float* Vb, float* Vc, unsigned n);
void oracle_tensorvote2_tk(float* t_out, const float* lambdas, const float* thetas,
                           float sigma1, float sigma2, unsigned power, unsigned w,
                           unsigned s0, unsigned s1, int stick, int plate, unsigned samples,
                           const oracle_neighbor2d* nb, int nb_count, int nthreads);
void oracle_tensorvote2(float* t_out, const float* t_in, float sigma1, float sigma2, int power, unsigned w,
                        unsigned s0, unsigned s1, int stick, int plate, unsigned samples, int nthreads);
/* FP64 2D vote: eigen in double, stick from the FP64 kernel, plate from the
 * FP64 literal N-sample integral, same window. */
void oracle_tensorvote2_f64(double* t_out, const float* t_in, double sigma1, double sigma2, int power,
                            unsigned w, unsigned s0, unsigned s1, int stick, int plate, unsigned samples,
                            int nthreads);

/* ---------------- 3D tensor voting (tensorvote.h) ---------------- */
int   oracle_build_neighbors3d(int w, float sigma1, float sigma2, oracle_neighbor3d* out);
float oracle_sticknorm3(float sigma1, float sigma2, unsigned p);
/* Reference-faithful FP32 3D vote. recv == NULL: every voxel, out9 holds
 * s0*s1*s2*9 floats. Otherwise out9 holds nrecv*9 floats for the listed
 * receivers (flat voxel indices). Reproduces the CPU path's asymmetric
 * plate accumulation (flat slots 3, 6, 7 carry the stick part only). */
void oracle_tensorvote3_f32(const float* in9, float* out9, unsigned s0, unsigned s1, unsigned s2,
                            float sigma1, float sigma2, unsigned power, unsigned w,
                            int stick, int plate, unsigned samples,
                            const int64_t* recv, size_t nrecv, int nthreads);
/* Same, from precomputed eigenpairs (reference _tensorvote3_tk inputs):
 * lambdas n x 3, evec2 / evec0 n x 3 cartesian. Output symmetric 9 floats. */
void oracle_tensorvote3_tk_f32(const float* lambdas, const float* evec2, const float* evec0, float* out9,
                               unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2,
                               unsigned power, unsigned w, int stick, int plate, unsigned samples,
                               const int64_t* recv, size_t nrecv, int nthreads);
/* FP64 oracle: Jacobi eigenpairs, FP64 stick, FP64 literal N-sample plate
 * (samples == 0: FP64 analytic plate). Output symmetric 9 doubles. */
void oracle_tensorvote3_f64(const float* in9, double* out9, unsigned s0, unsigned s1, unsigned s2,
                            double sigma1, double sigma2, unsigned power, unsigned w,
                            int stick, int plate, unsigned samples,
                            const int64_t* recv, size_t nrecv, int nthreads);

/* Single-pair kernels in FP64 (pinning checks). out6 = (xx, xy, xz, yy, yz, zz). */
void oracle_stick3_pair_f64(const double* d, double c1, double c2, unsigned power, const double* q, double* out6);
void oracle_plate3_sum_f64(const double* d, double c1, double c2, unsigned power, const double* n,
                           unsigned samples, double* out6);
void oracle_plate3_analytic_f64(const double* d, double c1, double c2, unsigned power, const double* n,
                                double* out6);
/* Single-pair FP32 reference formulation (tensorvote.h:1651-1695). */
void oracle_plate3_numerical_f32(const float* d, float c1, float c2, unsigned power, const float* n,
                                 unsigned samples, float* out6);

/* ---------------- Gaussian filter (filter.h) ---------------- */
unsigned oracle_gaussian_ksize(float sigma);
void oracle_gaussian1d(unsigned ksize, float mu, float sigma, float dx, float* out);
void oracle_convolve3(const float* in, unsigned in_s0, unsigned in_s1, unsigned in_s2,
                      const float* kernel, unsigned k_s0, unsigned k_s1, unsigned k_s2,
                      float* out, unsigned* out_s0, unsigned* out_s1, unsigned* out_s2);
void oracle_gaussian_convolve3(const float* in, unsigned sx, unsigned sy, unsigned sz,
                               float sigma_w, float sigma_h, float sigma_d,
                               float* out, unsigned* ox, unsigned* oy, unsigned* oz);
/* Zero-padded "same" separable filter of a packed-6 tensor field
 * (s0 x s1 x s2 x 6, s2 fastest) with one 1D kernel of K taps on all three
 * axes, accumulated in FP64. */
void oracle_gaussian3_same_packed6_f64(const float* in6, unsigned s0, unsigned s1, unsigned s2,
                                       const float* kernel, unsigned K, double* out6, int nthreads);

/* ---------------- analytical tensor voting (tensorvote.h:23-760) ---------------- */
/* Single pair: 2D V = (a, b, c); 3D V = (00, 01, 11, 02, 12, 22). rx.. = receiver
 * offset from the voter; e1 middle, e2 largest eigenvector (3D). */
void oracle_atv2_f32(float l0, float l1, float ev0x, float ev0y, float ev1x, float ev1y, float rx, float ry,
                     float sigma, float* V);
void oracle_atv2_f64(double l0, double l1, double ev0x, double ev0y, double ev1x, double ev1y, double rx,
                     double ry, double sigma, double* V);
void oracle_atv3_f32(float l0, float l1, float l2, const float* e1, const float* e2, float rx, float ry, float rz,
                     float sigma, float* V);
void oracle_atv3_f64(double l0, double l1, double l2, const double* e1, const double* e2, double rx, double ry,
                     double rz, double sigma, double* V);
/* 2D ATV window from eigenpairs (lambdas n x 2, thetas n x 2) / from the field. */
void oracle_tensorvote2_atv_f32(float* t_out, const float* lambdas, const float* thetas, float sigma, unsigned w,
                                unsigned s0, unsigned s1, int nthreads);
void oracle_tensorvote2_atv_full_f32(float* t_out, const float* t_in, float sigma, unsigned w, unsigned s0,
                                     unsigned s1, int nthreads);
void oracle_tensorvote2_atv_f64(double* t_out, const float* t_in, double sigma, unsigned w, unsigned s0,
                                unsigned s1, int nthreads);
/* 3D ATV window from eigenpairs (L n x 3, E1 middle, E2 largest) / from the field. */
void oracle_tensorvote3_atv_eig_f32(const float* L, const float* E1, const float* E2, float* out9, unsigned s0,
                                    unsigned s1, unsigned s2, float sigma, unsigned w, const int64_t* recv,
                                    size_t nrecv, int nthreads);
void oracle_tensorvote3_atv_f32(const float* in9, float* out9, unsigned s0, unsigned s1, unsigned s2, float sigma,
                                unsigned w, const int64_t* recv, size_t nrecv, int nthreads);
void oracle_tensorvote3_atv_f64(const float* in9, double* out9, unsigned s0, unsigned s1, unsigned s2,
                                double sigma, unsigned w, const int64_t* recv, size_t nrecv, int nthreads);

/* ---------------- saliency threshold and Lee thinning ---------------- */
/* Otsu threshold (reconstruct.py:37-63, numpy.histogram semantics for float32 values). */
float oracle_otsu_threshold_f32(const float* values, size_t n);
/* 27-bit neighbourhood codes (bit 9(dz+1)+3(dy+1)+(dx+1)); Euler LUT sum, the
 * independent Euler change, octree and BFS simple-point tests. */
int oracle_thin_euler_lut_sum(unsigned code);
int oracle_thin_delta_euler(unsigned code);
int oracle_thin_simple_octree(unsigned code);
int oracle_thin_simple_bfs(unsigned code);
void oracle_thin_exhaustive(unsigned long long* out2, int nthreads);
/* computeThinImage (lee_thinning.h:600-680) on a 0/1 byte volume in place. */
int oracle_lee_thin3(unsigned char* vol, unsigned s0, unsigned s1, unsigned s2);
void oracle_thin_classify(unsigned first, unsigned count, unsigned char* flags, int nthreads);

#ifdef __cplusplus
}
#endif
#endif
