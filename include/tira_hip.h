/*
 * tira_hip.h -- C ABI of the MI355X (gfx950) tensor-voting library
 * (tiralib_amd/libtira_hip.so).
 *
 * Every compute entry point replaces one TIRA CUDA entry point; the comment
 * above it names the reference declaration (paths relative to the TIRA tree).
 * Plain pointers and sizes only. Rules shared by all entry points:
 *
 *  - Pointers may be host or device memory (detected with
 *    hipPointerGetAttributes, as the reference does with
 *    cudaPointerGetAttributes, tensorvote.h:1880-1890). Host buffers are staged
 *    through device memory owned by the library.
 *  - Outputs are caller-allocated (the reference returns new[]/cudaMalloc
 *    buffers; the C++ header include/tira/hip/tira.hpp restores that
 *    convention on top of this ABI).
 *  - Every call returns a tira_status; TIRA_OK is 0. tira_hip_last_error()
 *    returns a thread-local message for the last failure.
 *  - Calls are blocking (they synchronise the library stream before
 *    returning) unless named *_async. The library stream is the null stream
 *    unless tira_hip_set_stream() installs another one.
 *  - No global mutable device symbols: calls are re-entrant across host
 *    threads that use different streams.
 */
#ifndef TIRA_HIP_H
#define TIRA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TIRA_HIP_ABI_VERSION 1

typedef enum tira_status {
    TIRA_OK = 0,
    TIRA_ERR_INVALID_ARG = 1,   /* bad shape / sigma / window (pybind checks, tensorvote_tk.cu:73-83, 162-178) */
    TIRA_ERR_HIP = 2,           /* a HIP runtime call failed */
    TIRA_ERR_OUT_OF_MEMORY = 3,
    TIRA_ERR_UNSUPPORTED = 4,   /* valid request the library does not serve */
    TIRA_ERR_NO_DEVICE = 5
} tira_status;

/* Tensor layouts for the native entry points. */
#define TIRA_LAYOUT_MAT9    0   /* 9 floats, column-major glm::mat3 (reference 3D layout) */
#define TIRA_LAYOUT_PACKED6 1   /* 6 floats: xx, xy, xz, yy, yz, zz */

/* Plate-vote evaluation for the 3D vote (samples > 0 is the reference's
 * numerically integrated plate, tensorvote.h:1651-1695 / 2457-2485). */
#define TIRA_PLATE_AUTO    0   /* the spectral vote when it applies (power 1, samples 0 or >= 3, w >= 5, at
                                  least 32^3 voxels, memory within 80% of the device); else the closed
                                  form when samples >= power+2 (the Riemann sum is then exact), else the
                                  literal sum */
#define TIRA_PLATE_SUM     1   /* literal N-sample Riemann sum, the reference's algorithm, any N */
#define TIRA_PLATE_CLOSED  2   /* closed form of the same integral (exact for N >= p+2 and for samples == 0) */
#define TIRA_PLATE_MLP     3   /* learned surrogate with per-voxel integrator fallback: served by tira_tensorvote3_mlp */
#define TIRA_VOTE_SPECTRAL 4   /* stick + exact (closed-form) plate vote as an FFT convolution: power 1 and
                                  samples == 0 or >= 3; other requests take the direct gather (AUTO) */

/* eig3 flags */
#define TIRA_EIG_REFINE_FP64   1u  /* FP64 re-solve of voxels with relative eigengap < 1e-4 */
#define TIRA_EIG_SCALE_INVARIANT 2u /* thresholds relative to the tensor norm (reference: absolute 1e-6) */
#define TIRA_EIG_SALIENCY 4u        /* write (l2 - l1, l1 - l0, l0): stick, plate, ball saliency */
#define TIRA_EIG_MAGNITUDE 8u       /* with SALIENCY: rank by |l| (reconstruct.py:14-34):
                                       (|l|max - |l|mid, |l|mid - |l|min, |l|min) */

/* ------------------------------------------------------------------ runtime */
int tira_hip_abi_version(void);
const char* tira_hip_last_error(void);
int tira_hip_device_count(int* count);
int tira_hip_set_device(int device);
int tira_hip_malloc(void** ptr, size_t bytes);
/* *is_device = 1 for device (or managed) memory, 0 for host memory. */
int tira_hip_pointer_is_device(const void* ptr, int* is_device);
int tira_hip_free(void* ptr);
/* kind: 0 = infer from pointer attributes */
int tira_hip_memcpy(void* dst, const void* src, size_t bytes);
int tira_hip_memset(void* dst, int value, size_t bytes);
int tira_hip_synchronize(void);
int tira_hip_stream_create(void** stream);
int tira_hip_stream_destroy(void* stream);
/* Install the stream every later call on this host thread launches on (NULL = null stream). */
int tira_hip_set_stream(void* stream);
int tira_hip_event_create(void** event);
int tira_hip_event_destroy(void* event);
int tira_hip_event_record(void* event, void* stream);
int tira_hip_event_elapsed_ms(void* start, void* stop, float* ms);
int tira_hip_stream_synchronize(void* stream);
/* Kernel timing: between begin and end, every library kernel launch on this
 * host thread is bracketed by HIP events on its own stream; end() waits for
 * them and aggregates per kernel name (e.g. "vote3_closed", "eig3"). */
int tira_hip_profile_begin(void);
int tira_hip_profile_end(void);
int tira_hip_profile_get(const char* kernel_name, double* total_ms, long* count);
/* Plain device-to-device copy kernel (bytes a multiple of 16): the reference
 * for bandwidth and profiler-counter calibration. Non-blocking. */
int tira_hip_copy_async(void* dst, const void* src, size_t bytes);
/* FP32 FMA loop (blocks x 256 threads x 16 x iters FLOP): the achievable
 * unpacked FP32 VALU rate. `out` (device, blocks floats) is written only in
 * a case that does not occur; it keeps the arithmetic live. Non-blocking. */
int tira_hip_fma_loop_async(float* out, unsigned blocks, int iters);
/* Release every cached device buffer held by the library on the current device. */
int tira_hip_release_workspace(void);

/* ------------------------------------------------------------------ eigen */
/* tira::cuda::evals2_symmetric(const T* mats, size_t n, int device), eigen.h:768-812.
 * mats n x 4 [a b b c]; evals n x 2 ascending. device < 0 is rejected (the
 * reference falls back to its CPU path there; use tira::cpu for that). */
int tira_evals2_symmetric(const float* mats, size_t n, float* evals, int device);
/* tira::cuda::evecs2polar_symmetric(mats, evals, n, device), eigen.h:827-893. */
int tira_evecs2polar_symmetric(const float* mats, const float* evals, size_t n, float* thetas, int device);
/* tira::cuda::evals3_symmetric(const T* mats, size_t n, int device), eigen.h:895-942.
 * mats n x 9 column-major (slots 0,1,2,4,5,8 read); evals n x 3 ascending.
 * FP32 with FP64 re-solve of near-degenerate tensors. */
int tira_evals3_symmetric(const float* mats, size_t n, float* evals, int device);
/* tira::cuda::evecs3spherical_symmetric(mats, lambda, n, device), eigen.h:944-1012.
 * evecs n x 6 = (theta, phi) of e0, e1, e2. */
int tira_evecs3spherical_symmetric(const float* mats, const float* lambdas, size_t n, float* evecs, int device);
/* Native batched eig3 (no reference counterpart beyond eigen.h:895-1012):
 * layout TIRA_LAYOUT_*, evals n x 3 ascending; evec0 / evec2 (n x 3,
 * cartesian, canonical sign) may be NULL. n_refined (may be NULL) receives
 * the number of tensors re-solved in FP64. Device pointers only. */
int tira_eig3_async(const float* tensors, size_t n, int layout, float* evals, float* evec0, float* evec2,
                    unsigned flags, unsigned long long* n_refined_dev);
/* tira_eig3_async over nvol independent tensor arrays in one launch (64
 * arrays per launch; the array table travels in the kernel arguments, so the
 * call uploads nothing and can be graph-captured). tensors[v] / evals[v] /
 * evec0[v] / evec2[v] are device pointers for counts[v] tensors; evec0 and
 * evec2 (the arrays of pointers) may be NULL. Results equal tira_eig3_async
 * per array bit for bit. */
int tira_eig3_batch_async(const float* const* tensors, const size_t* counts, int nvol, int layout,
                          float* const* evals, float* const* evec0, float* const* evec2, unsigned flags,
                          unsigned long long* n_refined_dev);

/* ------------------------------------------------------------------ 2D voting */
/* 48-byte neighbour record, field order of tira::Neighbor2D (tensorvote.h:776-786):
 * int32 du, dv; float dx, dy, c1, c2, Ra, Rb, Rc, Va, Vb, Vc. */
/* tira::cpu::build_neighbors2d(w, sigma1, sigma2, power, samples, stick, plate),
 * tensorvote.h:1223-1280 (host). Writes up to `capacity` records to `out`
 * (may be NULL to query), the record count to *count. */
int tira_build_neighbors2d(int w, float sigma1, float sigma2, unsigned power, unsigned samples, int stick, int plate,
                           void* out, int capacity, int* count);
/* tira::cuda::tensorvote<T>(t_out, t_in, sigma1, sigma2, power, w, shape0, shape1,
 * stick, plate, samples, is_atv), tensorvote.h:2200-2226. t_in / t_out:
 * shape0 x shape1 x 4 floats [a b b c], shape1 fastest. is_atv != 0: analytical
 * tensor voting (tensorvote.h:1993-2066) with sigma1 and half-window w / 2;
 * sigma2, power, stick, plate and samples do not apply, as in the reference. */
int tira_tensorvote2(float* t_out, const float* t_in, float sigma1, float sigma2, int power, unsigned w,
                     unsigned shape0, unsigned shape1, int stick, int plate, unsigned samples, int is_atv);
/* tira::cuda::tensorvote_prebuilt<T>(t_out, t_in, host_nb, nb_count, sigma1,
 * sigma2, power, w, shape0, shape1, stick, plate, samples), tensorvote.h:2236-2305.
 * host_nb: nb_count 48-byte records (host memory). */
int tira_tensorvote2_prebuilt(float* t_out, const float* t_in, const void* host_nb, int nb_count, float sigma1,
                              float sigma2, int power, unsigned w, unsigned shape0, unsigned shape1, int stick,
                              int plate, unsigned samples);
/* cpu::build_neighbors2d (tensorvote.h:1223-1280) on the device: the w * w
 * Neighbor2D entries (dv outer, du inner) into dev_out, non-blocking. The same
 * operations as tira_build_neighbors2d, FP64 cos/sin/exp rounded to float,
 * no contraction; c1 and c2 take the correctly rounded float exponential
 * where the host takes glibc's expf. */
int tira_build_neighbors2d_async(int w, float sigma1, float sigma2, unsigned power, unsigned samples, int stick,
                                 int plate, void* dev_out);
/* Search-loop session: tira::cuda::tensorvote_prebuilt (tensorvote.h:2236-2305,
 * bound as tensorvote_tk_prebuilt at tensorvote_tk.cu:239-253) with the field
 * bound once, for optimize_tk's repeated small calls (optimize.py:256-288).
 * create copies the s0 x s1 x 4 field (host or device) into the session;
 * vote votes it with a host Neighbor2D table (nb, nb_count) or, with nb ==
 * NULL, with the table built from the parameters on the device
 * (tira_build_neighbors2d_async). Each session keeps the device copies of its
 * last 64 tables (keyed by parameters or by the table's bytes) and runs each
 * vote as one replay of an instantiated graph. Results are identical to
 * tira_tensorvote2_prebuilt with the same table; with a device-built table
 * they are within a few FP32 ulps of tira_tensorvote2 (whose host table can
 * differ in c1, c2 by one ulp). t_out host or device; blocking.
 * A session belongs to the device current at create; calls on one session
 * must not overlap. */
int tira_vote2_session_create(void** session, const float* t_in, unsigned s0, unsigned s1);
int tira_vote2_session_vote(void* session, float* t_out, const void* nb, int nb_count, float sigma1, float sigma2,
                            int power, unsigned w, int stick, int plate, unsigned samples);
/* Graph replays, table cache hits and table uploads of a session so far. */
int tira_vote2_session_stats(void* session, unsigned long long* replays, unsigned long long* table_hits,
                             unsigned long long* table_builds);
int tira_vote2_session_destroy(void* session);
/* No reference counterpart. tira_tensorvote2_prebuilt for callers that cannot
 * allocate device memory (code under stream capture, tight serving loops,
 * harnesses that forbid hipMalloc): the same kernels and results, with the
 * voter payload (16 B per pixel) and the neighbour table (48 B per entry)
 * kept in a 32 MiB device array owned by the library (one per GPU, held until
 * the process ends) instead of allocated workspace. Device fields only; the
 * table may be host or device memory (a host table is copied into the
 * array). Rejects fields whose payload and table exceed 32 MiB
 * (TIRA_ERR_INVALID_ARG, nothing launched). Blocking. Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU (from several host
 * threads or streams) race. */
int tira_tensorvote2_prebuilt_noalloc(float* t_out, const float* t_in, const void* nb, int nb_count, float sigma1,
                                      float sigma2, int power, unsigned w, unsigned s0, unsigned s1, int stick,
                                      int plate, unsigned samples);
/* tira::cuda::tensorvote_tk<T>(t_out, lambdas, thetas, sigma1, sigma2, power, w,
 * s0, s1, stick, plate, samples), tensorvote.h:2125-2179. lambdas / thetas
 * s0 x s1 x 2 (ascending eigenvalues, eigenvector angles). */
int tira_tensorvote2_tk(float* t_out, const float* lambdas, const float* thetas, float sigma1, float sigma2,
                        unsigned power, unsigned w, unsigned s0, unsigned s1, int stick, int plate,
                        unsigned samples);
/* tira::cuda::tensorvote_atv<T>(t_out, lambdas, thetas, sigma, w, s0, s1),
 * tensorvote.h:2029-2066: 2D analytical tensor voting from eigenpairs
 * (lambdas / thetas s0 x s1 x 2 as evals2 / evecs2polar return them). */
int tira_tensorvote2_atv(float* t_out, const float* lambdas, const float* thetas, float sigma, unsigned w,
                         unsigned s0, unsigned s1);

/* ------------------------------------------------------------------ 3D voting */
/* tira::cuda::tensorvote3(input_field, output_field, s0, s1, s2, sigma, sigma2, w,
 * power, device, stick, plate, samples, is_atv), tensorvote.h:2816-2873.
 * in / out: s0*s1*s2 x 9 floats (glm::mat3, s2 fastest). is_atv != 0:
 * analytical tensor voting (tensorvote.h:2635-2739) with sigma and half-window
 * w / 2; sigma2, power, stick, plate and samples do not apply. */
int tira_tensorvote3(const float* input_field, float* output_field, unsigned s0, unsigned s1, unsigned s2,
                     float sigma, float sigma2, unsigned w, unsigned power, int device, int stick, int plate,
                     unsigned samples, int is_atv);
/* tira::cuda::_tensorvote3_tk(out, lambdas, evec2, evec0, s0, s1, s2, sigma1,
 * sigma2, power, w, STICK, PLATE, samples), tensorvote.h:2741-2809.
 * lambdas n x 3, evec2 / evec0 n x 3 cartesian. */
int tira_tensorvote3_tk(float* output_field, const float* lambdas, const float* evec2, const float* evec0,
                        unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned power,
                        unsigned w, int stick, int plate, unsigned samples);
/* Native 3D vote on device buffers, non-blocking on the library stream.
 * in_layout / out_layout: TIRA_LAYOUT_*; plate_mode: TIRA_PLATE_*. */
int tira_tensorvote3_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                           unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
                           unsigned power, int stick, int plate, unsigned samples, int plate_mode);
/* tira_tensorvote3_async followed by tira_eig3_async(output_field, s0 s1 s2,
 * out_layout, evals, NULL, NULL, eig_flags, n_refined_dev): the same field,
 * bit for bit, and eigenvalues (n x 3) of the same solver to FP32 rounding
 * (a few ulps of the field's largest entry). No reference counterpart: the
 * reference's callers vote, then decompose. When the spectral vote runs, its
 * last pass writes the eigenvalues of every voxel it writes (one read of the
 * field and one launch fewer); otherwise eig3 runs after the vote. */
int tira_tensorvote3_eig_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                               unsigned s0, unsigned s1, unsigned s2, float sigma1, float sigma2, unsigned w,
                               unsigned power, int stick, int plate, unsigned samples, int plate_mode, float* evals,
                               unsigned eig_flags, unsigned long long* n_refined_dev);
/* tira::cuda::_tensorvote3_atv(out, lambdas, evec1, evec2, s0, s1, s2, sigma, w),
 * tensorvote.h:2695-2739: 3D analytical tensor voting from eigenpairs
 * (lambdas n x 3 ascending, evec1 middle and evec2 largest eigenvectors n x 3,
 * orthonormal as evecs3spherical_symmetric returns them). */
int tira_tensorvote3_atv(float* output_field, const float* lambdas, const float* evec1, const float* evec2,
                         unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w);
/* Native 3D analytical tensor voting on device buffers, non-blocking on the
 * library stream (layouts TIRA_LAYOUT_*). The spectral form is taken when it
 * applies (w >= 5, at least 32^3 voxels, FFT lengths of at most 1024). */
int tira_tensorvote3_atv_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                               unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w);
/* The same with the method named: TIRA_PLATE_AUTO, TIRA_PLATE_CLOSED (the
 * per-pair gather) or TIRA_VOTE_SPECTRAL (the FFT form whenever the input
 * allows it). No reference counterpart (the reference has one algorithm). */
int tira_tensorvote3_atv_method_async(const float* input_field, int in_layout, float* output_field, int out_layout,
                                      unsigned s0, unsigned s1, unsigned s2, float sigma, unsigned w, int method);

/* Spectral vote internals, exported for tests (no reference counterpart).
 * Offset kernels K[ch][dw][dv][du] (ch = c * nf + f; c = xx, xy, xz, yy, yz,
 * zz; f < nf the voter features: the 15 quartic monomials q_x^a q_y^b q_z^c,
 * a = 4..0, b = 4-a..0, of the stick vector e2 (when stick) then of the plate
 * normal e0 (when plate)) such that the pair vote of tensorvote.h:1183-1213 /
 * 1651-1695 (normalised by sticknorm3 and pi, plate exact) at offset d equals
 * sum_f K[c,f](d) * scale * monomial_f. Host computation, FP64, no GPU; out
 * holds 6 * nf * (2(w/2)+1)^3 doubles; *nf_out receives nf. power must be 1. */
int tira_spectral_offset_kernels(unsigned w, float sigma1, float sigma2, unsigned power, int stick, int plate,
                                 double* out, unsigned* nf_out);
/* The same for analytical tensor voting (tensorvote.h:688-760, attenuation
 * exp(-|d|^2 / sigma^2), half-window w / 2): nf = 31 features, the 15
 * monomials of e2 times l2 - l1, of e0 times l1 - l0, and l0 itself. */
int tira_spectral_offset_kernels_atv(unsigned w, float sigma, double* out, unsigned* nf_out);
/* In-place 3D FFT of nchan complex volumes (interleaved float pairs, each
 * L0 x L1 x L2 with L2 fastest, volumes back to back); forward e^{-2 pi i k x / L},
 * inverse unnormalised. Every L must be <= 1024 with prime factors 2, 3, 5, 7.
 * Device pointer, non-blocking. */
int tira_fft3_async(float* data, unsigned nchan, unsigned L0, unsigned L1, unsigned L2, int inverse);
/* FFT length the spectral vote uses for an axis of n voxels and window w
 * (smallest 7-smooth L >= n + w/2), 0 when above 1024. */
unsigned tira_spectral_length(unsigned n, unsigned w);
/* The calling thread's cache of kernel spectra: entries held (the last 8
 * parameter sets at most, within 1/8 of device memory, least recently used
 * evicted first), their bytes, and the spectra built so far. */
int tira_spectral_cache_info(unsigned* entries, unsigned long long* bytes, unsigned long long* builds);

/* ------------------------------------------------------------------ Gaussian filter */
/* tira::cuda::gaussian_filter3d<float>(src, width, height, depth, sigma_w, sigma_h,
 * sigma_d, out_width, out_height, out_depth), filter.h:400-534. Valid-mode
 * separable Gaussian of a scalar volume (width fastest); kernel lengths
 * ceil(6 sigma) forced odd (at most 17). dst == NULL only reports the output
 * size; otherwise dst holds out_width*out_height*out_depth floats. */
int tira_gaussian_filter3d(const float* src, unsigned width, unsigned height, unsigned depth, float sigma_w,
                           float sigma_h, float sigma_d, float* dst, unsigned* out_width, unsigned* out_height,
                           unsigned* out_depth);
/* tira::cuda::gaussian_convolve2<float>(source, in_sx, in_sy, sigma1, sigma2,
 * out_width, out_height), filter.h:304-397. Kernel lengths (unsigned)(6 sigma). */
int tira_gaussian_convolve2(const float* src, unsigned in_sx, unsigned in_sy, float sigma1, float sigma2, float* dst,
                            unsigned* out_width, unsigned* out_height);
/* tira::cpu::gaussian1d(ksize, 0, sigma, 1), filter.h:37-64 (host): ksize taps. */
int tira_gaussian_taps(unsigned ksize, float sigma, float* out);
/* Native fused separable filter of an s0 x s1 x s2 field with `components`
 * (1 or 6, packed) floats per voxel, the same odd-length kernel on all axes;
 * same != 0: zero-padded output of the input size, else valid mode. Device
 * pointers, non-blocking. No reference counterpart beyond filter.h. */
int tira_gaussian_filter3_tensor_async(const float* in, float* out, unsigned s0, unsigned s1, unsigned s2,
                                       unsigned components, const float* taps, unsigned ntaps, int same);

/* ------------------------------------------------------------------ batched volumes */
/* Many uneven 6-component volumes in one persistent launch (no reference
 * counterpart): for every volume i, out[i] = zero-padded separable Gaussian of
 * in[i] (the same odd kernel of ntaps taps, 7..13, on all three axes) and
 * evals[i] = ascending eigenvalues of every input tensor. shapes: nvol x 3
 * (s0, s1, s2), all pointers device memory. autotune != 0 times the candidate
 * tile shapes (16 x 8, 16 x 16, 16 x 32 outputs), z-chunks (the tile's depth,
 * hence its z-halo share) and occupancies on this batch once per process and
 * batch shape; 0 takes 16 x 16 x 32 at two blocks per CU. The launch is
 * captured in a hipGraph at creation; run() replays it on the library stream
 * (non-blocking). Buffers must stay valid until destroy(). */
int tira_batch_gauss_eig3_create(void** handle, int nvol, const float* const* in, float* const* out,
                                 float* const* evals, const unsigned* shapes, const float* taps, unsigned ntaps,
                                 int autotune);
int tira_batch_gauss_eig3_run(void* handle);
int tira_batch_gauss_eig3_plan(void* handle, int* zchunk, int* blocks_per_cu, float* tuned_ms, unsigned* items);
/* The plan's tile shape: outputs per tile along x and y. */
int tira_batch_gauss_eig3_tile(void* handle, int* tile_x, int* tile_y);
/* Replaces the plan (tile height 8, 16 or 32; z-chunk; blocks per CU, at
 * most 2 for 32-row tiles) and captures the launch again. Blocking. */
int tira_batch_gauss_eig3_set_plan(void* handle, int tile_y, int zchunk, int blocks_per_cu);
int tira_batch_gauss_eig3_destroy(void* handle);

/* ------------------------------------------------------------------ synthetic inputs */
/* Seeded packed-6 test tensors generated on the device (benchmark inputs, no
 * reference counterpart): eigenvalues log-uniform in [lo, hi], a fraction
 * degenerate_frac with one adjacent pair at relative gap 10^U(log gap_lo,
 * log gap_hi), Haar-random rotations; tensor i depends only on (seed, i).
 * lambdas (optional, n x 3) receives the construction eigenvalues. */
int tira_synth_eig_tensors_async(float* out6, size_t n, unsigned long long seed, float lo, float hi,
                                 float degenerate_frac, float gap_lo, float gap_hi, float* lambdas);
/* Packed-6 plate volume s0 x s1 x s2 (device): n_sheets curved sheets, each
 * 15 floats of `params` (host): R (3x3 row-major; columns = local axes),
 * centre (x, y, z), curvatures (a, b, c) of q_z = a q_x^2 + b q_y^2 + c q_x q_y;
 * thickness 1, patch radius `patch`, tensor I - n n^T; elsewhere plate noise
 * m (I - q q^T) on a fraction `noise_frac` of the voxels. Non-blocking. */
int tira_synth_sheets_async(float* out6, unsigned s0, unsigned s1, unsigned s2, int n_sheets, const float* params,
                            float patch, float noise_frac, unsigned long long seed);
/* Intensity volume of a vascular tree: 1 inside any tapered capsule of
 * `segments` (device, nseg x 8 floats: p0 xyz, p1 xyz, r0, r1; x fastest),
 * 0 outside, plus N(0, noise_sigma^2) noise. Non-blocking. */
int tira_synth_vessels_async(float* out, unsigned s0, unsigned s1, unsigned s2, const float* segments, int nseg,
                             float noise_sigma, unsigned long long seed);

/* ------------------------------------------------------------------ front end and pipeline
 * tira::volume::gradient (volume.h:974-1021, forward / backward difference at
 * the ends of an axis, central inside, divided by the spacing in double)
 * along x (axis 0, fastest, s2 voxels), y (s1) and z (s0) with the volume's
 * double spacings (volume.h:22), fused with the
 * structure tensor g g^T: out6 is s0 x s1 x s2 x 6 packed (xx, xy, xz, yy,
 * yz, zz). Device pointers, non-blocking. */
int tira_structure_tensor3_async(const float* intensity, unsigned s0, unsigned s1, unsigned s2, double spacing_x,
                                 double spacing_y, double spacing_z, float* out6);
/* The whole 3D pipeline on device buffers: structure tensor -> Gaussian of
 * the 6 components (sigma smooth_sigma, ceil(6 sigma) odd taps, zero-padded
 * "same"; 0 skips it) -> stick + plate vote (plate_mode TIRA_PLATE_*; for
 * TIRA_PLATE_MLP the network `mlp` with its own sigmas and power, voters
 * whose error check exceeds mlp_tol taking the integrator) -> eig3 saliency
 * (l2 - l1, l1 - l0, l0) into saliency (n x 3). flags: TIRA_EIG_* for the
 * eig3 stage. Intermediates live in library workspace (48 B per voxel). */
int tira_tensorvote3_pipeline_async(const float* intensity, unsigned s0, unsigned s1, unsigned s2, double spacing_x,
                                    double spacing_y, double spacing_z, float smooth_sigma, float vote_sigma1,
                                    float vote_sigma2, unsigned w, unsigned power, int stick, int plate,
                                    unsigned samples, int plate_mode, void* mlp, float mlp_tol, float* saliency,
                                    unsigned flags);

/* ------------------------------------------------------------------ learned plate vote
 * No reference counterpart: a BF16 MLP (5 -> 64 -> 64 -> 64 -> 6, MFMA, FP32
 * accumulation) that stands in for the plate-vote integral of
 * tensorvote.h:1651-1695, trained on the device by this library's own forward,
 * backward and Adam kernels against the FP64 N-sample integrator. Handles are
 * bound to the device current at creation; calls are synchronous unless
 * noted. Parameters are FP32, row-major [out][in]: W1 64x32, W2 64x64,
 * W3 64x64, W4 32x64 (12288 floats); input slots 0-4 feature high parts,
 * 5-9 residuals, 10 constant 1; hidden unit 63 is a constant-1 unit. */
#define TIRA_MLP_NPARAM 12288
#define TIRA_MLP_NBINS 2048
int tira_plate_mlp_create(void** handle, float sigma1, float sigma2, unsigned power, unsigned integrator_samples,
                          unsigned long long seed);
int tira_plate_mlp_destroy(void* handle);
/* Trains on `nsamples` fresh samples (offset uniform in the 3 sigma1 ball,
 * normal uniform on the sphere), `epochs` passes in batches of `batch` (a
 * multiple of 256), Adam with warm-up and cosine decay from `lr` to
 * `lr_final` (0: lr / 100); then validates on 262144 held-out samples and
 * rebuilds the per-orientation error table. Outputs relative Frobenius
 * errors (last batch, validation). */
int tira_plate_mlp_train(void* handle, size_t nsamples, unsigned epochs, unsigned batch, float lr, float lr_final,
                         unsigned long long seed, float* train_rel_err, float* val_rel_err);
/* Host copies of the FP32 parameters: `get` (may be NULL) receives them,
 * `set` (may be NULL) replaces them first. */
int tira_plate_mlp_params(void* handle, float* get, const float* set);
/* n samples into device buffers: X n x 32 BF16 inputs, T n x 8 targets (6
 * plate-frame components / c1, then 0, 0), raw n x 8 (may be NULL): offset
 * vector, normal, c1, 0. */
int tira_plate_mlp_samples(void* handle, size_t n, unsigned long long seed, void* X, float* T, float* raw);
/* Y (device, n x 32 FP32; columns 0-5 meaningful) = network(X); n % 32 == 0. */
int tira_plate_mlp_forward(void* handle, const void* X, size_t n, float* Y);
/* One forward + backward pass over n samples (n % 256 == 0): host `grads`
 * (TIRA_MLP_NPARAM) receive d loss / d params; sums (host, 2 floats, may be
 * NULL) receive the summed weighted squared error and target energy. */
int tira_plate_mlp_gradients(void* handle, const void* X, const float* T, size_t n, float* grads, float* sums);
/* One training step (forward, backward, Adam) with learning rate lr. */
int tira_plate_mlp_step(void* handle, const void* X, const float* T, size_t n, float lr);
/* The FP32 output affine (12 floats: means, then scales of the 6 plate-frame
 * components; the network predicts standardized components). `get` (may be
 * NULL) receives it, `set` (may be NULL) replaces it first. Training sets it
 * from the training targets. */
int tira_plate_mlp_affine(void* handle, float* get, const float* set);
/* Per-orientation relative error of the network evaluated per pair
 * (TIRA_MLP_NBINS floats). */
int tira_plate_mlp_bin_errors(void* handle, float* errs);
/* The same for the tabulated network (the default vote method). */
int tira_plate_mlp_table_errors(void* handle, float* errs);
int tira_plate_mlp_refresh_bins(void* handle);
/* Host copy of the tabulated network the vote reads: *rows distance rows of
 * *cells cells, each cell 8 floats ((A, B, g, D/2) at x_j = (j + 0.5) / cells
 * and the difference to the next cell), the plate vote / c1 being
 * A I + B n n^T + s g (n d^T + d n^T) + D d d^T at s = n.d, x = 1 - s^2.
 * Row 0 is the voter itself, rows 1..K the squared lattice distances
 * shells[0..K-1] (|d| < 3 sigma1), row K + 1 every farther offset. Any
 * output may be NULL; `tab` needs rows x cells x 8 floats (cap_floats). */
int tira_plate_mlp_vote_table(void* handle, float* tab, size_t cap_floats, int* rows, int* cells, int* shells);
/* Replaces the vote table with a caller-built one of the same shape (host,
 * rows x cells x 8 floats, layout as above) until the next weight change:
 * a test hook that lets the table vote be judged with a table of the exact
 * integral. */
int tira_plate_mlp_set_vote_table(void* handle, const float* tab, size_t nfloats);

/* Evaluation of the network in the vote. */
#define TIRA_MLP_TABLE    0   /* tabulated once per weight change over (distance row, x = 1 - (n.d)^2) on
                                 MFMA, linear in x per pair (default) */
#define TIRA_MLP_PER_PAIR 1   /* one network evaluation per (receiver, voter) pair on MFMA */
/* 3D stick + learned plate vote (device pointers, non-blocking unless
 * n_fallback is given): the network's sigma1, sigma2 and power; voters whose
 * orientation bin's error exceeds `tol` take the integrator's plate vote.
 * n_fallback (host, may be NULL) receives the number of such voters. Uses
 * TIRA_MLP_TABLE (windows up to 63; larger windows evaluate per pair). */
int tira_tensorvote3_mlp(void* mlp, const float* input_field, int in_layout, float* output_field, int out_layout,
                         unsigned s0, unsigned s1, unsigned s2, unsigned w, int stick, float tol,
                         unsigned long long* n_fallback);
/* The same with the evaluation method chosen (TIRA_MLP_*). */
int tira_tensorvote3_mlp_method(void* mlp, const float* input_field, int in_layout, float* output_field,
                                int out_layout, unsigned s0, unsigned s1, unsigned s2, unsigned w, int stick,
                                float tol, unsigned long long* n_fallback, int method);

/* ------------------------------------------------------------------ saliency, threshold, thinning */
/* Stick saliency |l|max - |l|min of n 2x2 tensors [a b b c] (reconstruct.py:14-34),
 * device pointers, non-blocking. (3D: tira_eig3_async with TIRA_EIG_SALIENCY |
 * TIRA_EIG_MAGNITUDE.) */
int tira_stick_saliency2_async(const float* field4, size_t n, float* saliency);
/* Otsu threshold (reconstruct.py:37-63) of values[i * stride], i < n, nonnegative
 * float32: 256-bin histogram on [0, max] with numpy.histogram's float32 binning,
 * between-class variance in FP64, threshold = the winning bin's centre, written to
 * *threshold (device). Non-blocking, no host round trip. */
int tira_otsu_threshold_async(const float* values, size_t n, int stride, float* threshold);
/* mask[i] = values[i * stride] > *threshold (reconstruct.py:87), device pointers. */
int tira_threshold_async(const float* values, size_t n, int stride, const float* threshold, unsigned char* mask);
/* _fill_enclosed(boundary, close_radius) of saliency_threshold(fill=True)
 * (reconstruct.py:92-116): a 0/1 byte mask of h x w (w fastest; nonzero = 1),
 * closed with a (2r+1) x (2r+1) square when close_radius r > 0 (scipy's
 * binary_closing, pixels outside the image count as 0), then holes filled
 * (binary_fill_holes: background not 4-connected to the outside through
 * background becomes 1). out receives 0/1 bytes. Host or device pointers;
 * blocking. *sweeps (may be NULL) receives the number of flood sweeps. */
int tira_fill_holes2(const unsigned char* mask, unsigned h, unsigned w, unsigned close_radius, unsigned char* out,
                     unsigned* sweeps);
/* lee(in, out, x, y, z) / computeThinImage, thinning/lee_thinning.h:600-706:
 * Lee et al. 1994 3D thinning of an int volume (x fastest; nonzero = object),
 * bit-exact, output 0/1. Host or device pointers; blocking. *iterations (may
 * be NULL) receives the number of outer passes. */
int tira_lee_thin3(const int* in, int* out, unsigned x, unsigned y, unsigned z, unsigned* iterations);
/* The same on a device byte mask of s0 x s1 x s2 (s2 fastest), in place. */
int tira_lee_thin3_u8(unsigned char* mask, unsigned s0, unsigned s1, unsigned s2, unsigned* iterations);
/* Neighbourhood tests of the thinning for codes first .. first+count-1 (27-bit
 * code = low 13 bits | centre bit 13 | high bits << 14; bit 9(dz+1)+3(dy+1)+(dx+1)):
 * flags bit 0 Euler invariant (lee_thinning.h:288-299), bit 1 simple (530-579),
 * bit 2 end point (96-103). Device output, non-blocking. */
int tira_thin3_classify_async(unsigned first, unsigned count, unsigned char* flags);
/* Voted field (device, layout) -> saliency3 (n x 3 magnitude-ranked saliency,
 * device) -> Otsu threshold of the stick saliency (*threshold, device) ->
 * skeleton (device bytes) = Lee thinning of saliency > threshold. Blocking. */
int tira_skeleton3(const float* voted, int layout, unsigned s0, unsigned s1, unsigned s2, float* saliency3,
                   float* threshold, unsigned char* skeleton, unsigned* iterations);

#ifdef __cplusplus
}
#endif
#endif /* TIRA_HIP_H */
