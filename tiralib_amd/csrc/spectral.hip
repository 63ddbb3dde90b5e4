// spectral.hip -- the 3D stick + plate vote as an FFT convolution on gfx950.
//
// For power p = 1 the vote a voter casts at offset d is a homogeneous quartic
// in its eigenvector:
//   * stick (stickvote3_accumulate_kernel, tensorvote.h:1183-1213):
//     s * eta(q.d) r r^T with r = q - 2 (q.d) d and eta = c1 (1 - t^2) + c2 t^2,
//     where 1 = |q|^2 makes every term degree 4 in q;
//   * plate (tensorvote.h:1651-1695): the exact circle integral (the
//     reference's N-sample sum equals it for N >= p + 2 = 3), a quartic in the
//     normal n by the same argument.
// On the unit sphere the even polynomials of degree <= 4 span exactly the 15
// quartic monomials q_x^a q_y^b q_z^c (a + b + c = 4), so
//     V_c(x) = sum_f sum_d K_{c,f}(d) F_f(x + d),      c = xx, xy, xz, yy, yz, zz,
// with offset kernels K_{c,f}(d) of the stick vote fitted once on the host in
// FP64. The plate vote integrates the stick vote over the great circle
// perpendicular to n, and averaging a quartic over that circle is a fixed
// linear map on the 15 monomials (add_circle_average), so the plate kernels
// are the stick kernels times that map and both votes share them: the voter
// features are F_f = s m_f(e2) + sticknorm3 t (circle average of m(e0))_f.
// The whole field's vote is a 15-in / 6-out correlation with a w^3 window:
// O(N log N) through FFTs instead of O(N w^3) pair evaluations.
//
// Pipeline of one call (device buffers, zero padding to L >= n + w/2 per axis
// so the circular convolution equals the reference's bounded window):
//  1. payload: eigen-decomposition of every voxel, for eigenpair input only
//     (tensor input is eigen-solved in the x pass);
//  2. forward x pass fused with the features: rows of the field only, two
//     real features packed as one complex line;
//  3. forward y pass (planes z < n0), forward z pass (all columns);
//  4. pointwise product: one thread per octant group {(+-kx, +-ky, +-kz)}.
//     The kernels are real, even and of definite parity along each axis, so
//     their spectra are real and one octant (FP32, FP64-built) serves all
//     eight frequencies; the packed complex features are untangled with
//     Z(-k), and the six real outputs are packed back into three complex
//     spectra, computed at four of the eight frequencies and conjugated for
//     the other four;
//  5. inverse z pass (outputs z < n0), inverse y pass (outputs y < n1), and
//     the inverse x pass fused with the output write (either layout).
// FFTs are Stockham autosort, mixed radix 2/3/4/5/7/8/9, one block per batch of
// lines staged in LDS; the twiddle table is FP64-built and rounded once.
// The kernel spectra depend on (L0, L1, L2, w, sigma1, sigma2, stick, plate)
// only: built on the device (separable FP64 cosine/sine transforms of the
// offset kernels) and cached per host thread and device.
#include "common.hpp"
#include "vote3.hpp"
#include "voxel_payload.hpp"
#include "eig_store.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace th {

constexpr int kFT = 256;       // threads per FFT block
#ifndef TH_FFT_BL
#define TH_FFT_BL 4096
#endif
constexpr int kFftBL = TH_FFT_BL;  // complex values staged per block (32 KB of LDS)
// k_fft_cols staging + twiddles, bytes: four blocks per CU (B = 16 at L = 400
// took 54 KB, two blocks per CU, and ran its pass 1.4x slower than B = 8)
constexpr size_t kColsLds = 40960;

constexpr int kMaxL = 1024;    // longest transform
constexpr int kMaxStages = 16;
constexpr int kMono = 15;      // quartic monomials

// Exponents (a, b, c) of monomial m: a = 4..0, b = 4-a..0, c = 4-a-b.
__host__ __device__ constexpr int mono_a(int m) {
    return m < 1 ? 4 : m < 3 ? 3 : m < 6 ? 2 : m < 10 ? 1 : 0;
}
__host__ __device__ constexpr int mono_b(int m) {
    const int a = mono_a(m);
    const int first = (4 - a) * (5 - a) / 2;  // index of (a, 4-a, 0)
    return (4 - a) - (m - first);
}
__host__ __device__ constexpr int mono_c(int m) { return 4 - mono_a(m) - mono_b(m); }
// Number of x / y / z indices in output component c (xx, xy, xz, yy, yz, zz).
__host__ __device__ constexpr int comp_nx(int c) { return c == 0 ? 2 : (c == 1 || c == 2) ? 1 : 0; }
__host__ __device__ constexpr int comp_ny(int c) { return c == 3 ? 2 : (c == 1 || c == 4) ? 1 : 0; }
__host__ __device__ constexpr int comp_nz(int c) { return c == 5 ? 2 : (c == 2 || c == 4) ? 1 : 0; }

struct FftPlan {
    int L;
    int nst;
    int rad[kMaxStages];
};

// Radices of L (7s, 5s, 9s, a 3, then 8s and 4s, a final 2); false when L has a
// prime factor above 7.
static bool make_plan(int L, FftPlan& p) {
    p.L = L;
    p.nst = 0;
    int n = L;
    for (int f : {7, 5, 9, 3})  // 9 before 3: pairs of radix-3 stages become one radix-9 stage
        while (n % f == 0) { p.rad[p.nst++] = f; n /= f; }
    int t = 0;
    while (n % 2 == 0) { t++; n /= 2; }
    if (n != 1) return false;
    while (t >= 3 && t != 4) { p.rad[p.nst++] = 8; t -= 3; }
    while (t >= 2) { p.rad[p.nst++] = 4; t -= 2; }
    if (t) p.rad[p.nst++] = 2;
    return p.nst <= kMaxStages;
}

// Plan of a line transform in LDS (fft_lines): as make_plan, but a 9 and a
// lone 3 become one radix-27 stage, which takes a single round per thread.
// That stage goes first: at p = 1 its 26 twiddles per butterfly are all W^0
// and are skipped, while the radix after it reads its fewer twiddles once
// per stage (fft_stage's fixed mapping).
static bool make_lines_plan(int L, FftPlan& p) {
    if (!make_plan(L, p)) return false;
    int i9 = -1, i3 = -1;
    for (int i = 0; i < p.nst; i++) {
        if (p.rad[i] == 9 && i9 < 0) i9 = i;
        if (p.rad[i] == 3) i3 = i;
    }
    if (i9 < 0 || i3 < 0) return true;
    int k = 0;
    for (int i = 0; i < p.nst; i++)
        if (i != i9 && i != i3) p.rad[1 + k++] = p.rad[i];
    p.rad[0] = 27;
    p.nst = k + 1;
    return true;
}

// Plan of a column pass (k_fft_cols): as make_plan, but a 5 and a lone 3
// become one radix-15 stage placed last. The last stage of a column pass
// reads LDS and writes HBM one butterfly at a time, so radix 15 costs no
// more registers there, while as an in-place LDS stage it would hold two
// rounds of 15 values across the barrier.
static bool make_cols_plan(int L, FftPlan& p) {
    if (!make_plan(L, p)) return false;
    int i5 = -1, i3 = -1;
    for (int i = 0; i < p.nst; i++) {
        if (p.rad[i] == 5 && i5 < 0) i5 = i;
        if (p.rad[i] == 3) i3 = i;
    }
    if (i5 < 0 || i3 < 0) return true;
    int k = 0;
    for (int i = 0; i < p.nst; i++)
        if (i != i5 && i != i3) p.rad[k++] = p.rad[i];
    p.rad[k++] = 15;
    p.nst = k;
    return true;
}

// Whether `nlines` lines of plan `p` fit the register rounds of every stage
// (fft_stage holds at most MAXQ = ceil((kFftBL / R) / kFT) butterflies per
// thread); beyond that a stage would skip butterflies.
static bool lines_fit(const FftPlan& p, int nlines) {
    for (int i = 0; i < p.nst; i++) {
        const int R = p.rad[i];
        const long maxq = (kFftBL / R + kFT - 1) / kFT;
        if ((long)nlines * (p.L / R) > maxq * kFT) return false;
    }
    return true;
}

static unsigned spectral_length(unsigned n, unsigned w) {
    const unsigned need = n + w / 2;
    for (unsigned L = std::max(need, 1u); L <= (unsigned)kMaxL; L++) {
        FftPlan p;
        if (make_plan((int)L, p)) return L;
    }
    return 0;
}

// ---------------------------------------------------------------- device FFT
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }

// cos / sin of 2 pi n / R for the generic small DFTs.
template <int R> __device__ __forceinline__ float dcos(int n);
template <int R> __device__ __forceinline__ float dsin(int n);
template <> __device__ __forceinline__ float dcos<5>(int n) {
    constexpr float t[5] = {1.0f, 0.309016994374947424f, -0.809016994374947424f, -0.809016994374947424f,
                            0.309016994374947424f};
    return t[n];
}
template <> __device__ __forceinline__ float dsin<5>(int n) {
    constexpr float t[5] = {0.0f, 0.951056516295153572f, 0.587785252292473129f, -0.587785252292473129f,
                            -0.951056516295153572f};
    return t[n];
}
template <> __device__ __forceinline__ float dcos<7>(int n) {
    constexpr float t[7] = {1.0f, 0.623489801858733531f, -0.222520933956314404f, -0.900968867902419126f,
                            -0.900968867902419126f, -0.222520933956314404f, 0.623489801858733531f};
    return t[n];
}
template <> __device__ __forceinline__ float dsin<7>(int n) {
    constexpr float t[7] = {0.0f, 0.781831482468029809f, 0.974927912181823607f, 0.433883739117558120f,
                            -0.433883739117558120f, -0.974927912181823607f, -0.781831482468029809f};
    return t[n];
}
// In-register DFT of length R: y_k = sum_j u_j W^(jk), W = e^(-+2 pi i / R).
template <int R, bool INV>
__device__ __forceinline__ void dft_small(float2 (&u)[R]) {
    if constexpr (R == 2) {
        const float2 a = u[0], b = u[1];
        u[0] = cadd(a, b);
        u[1] = csub(a, b);
    } else if constexpr (R == 3) {
        constexpr float kS = 0.866025403784438647f;  // sin(2 pi / 3)
        const float2 s = cadd(u[1], u[2]), d = csub(u[1], u[2]);
        const float2 t = make_float2(fmaf(-0.5f, s.x, u[0].x), fmaf(-0.5f, s.y, u[0].y));
        const float sg = INV ? kS : -kS;  // i * sg * d
        u[0] = cadd(u[0], s);
        u[1] = make_float2(t.x - sg * d.y, t.y + sg * d.x);
        u[2] = make_float2(t.x + sg * d.y, t.y - sg * d.x);
    } else if constexpr (R == 4) {
        const float2 a = cadd(u[0], u[2]), b = csub(u[0], u[2]);
        const float2 c = cadd(u[1], u[3]), d = csub(u[1], u[3]);
        // forward: y1 = b - i d, y3 = b + i d
        const float2 mid = INV ? make_float2(b.x - d.y, b.y + d.x) : make_float2(b.x + d.y, b.y - d.x);
        const float2 oth = INV ? make_float2(b.x + d.y, b.y - d.x) : make_float2(b.x - d.y, b.y + d.x);
        u[0] = cadd(a, c);
        u[2] = csub(a, c);
        u[1] = mid;
        u[3] = oth;
    } else if constexpr (R == 9) {
        // 3 x 3 Cooley-Tukey: j = 3a + b, k = c + 3d, W9^(jk) = W3^(ac) W9^(bc) W3^(bd):
        // length-3 DFTs over a, twiddles W9^(bc), length-3 DFTs over b
        constexpr float c1 = 0.766044443118978035f, s1 = 0.642787609686539326f;   // 2 pi / 9
        constexpr float c2 = 0.173648177666930349f, s2 = 0.984807753012208059f;   // 4 pi / 9
        constexpr float c4 = -0.939692620785908384f, s4 = 0.342020143325668733f;  // 8 pi / 9
        const float sg = INV ? 1.0f : -1.0f;
        float2 v[3][3];
#pragma unroll
        for (int b = 0; b < 3; b++) {
            float2 t[3] = {u[b], u[3 + b], u[6 + b]};
            dft_small<3, INV>(t);
            v[b][0] = t[0];
            v[b][1] = t[1];
            v[b][2] = t[2];
        }
        v[1][1] = cmul(v[1][1], make_float2(c1, sg * s1));
        v[1][2] = cmul(v[1][2], make_float2(c2, sg * s2));
        v[2][1] = cmul(v[2][1], make_float2(c2, sg * s2));
        v[2][2] = cmul(v[2][2], make_float2(c4, sg * s4));
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float2 t[3] = {v[0][c], v[1][c], v[2][c]};
            dft_small<3, INV>(t);
            u[c] = t[0];
            u[c + 3] = t[1];
            u[c + 6] = t[2];
        }
    } else if constexpr (R == 27) {
        // 3 x 9 Cooley-Tukey: j = 9a + b, k = c + 3d, W27^(jk) = W3^(ac) W27^(bc) W9^(bd):
        // length-3 DFTs over a, twiddles W27^(bc), length-9 DFTs over b
        constexpr float tc[17] = {1.0f, 0.973044870579823806f, 0.893632640323412275f, 0.766044443118978013f,
                                  0.597158591702786179f, 0.396079766039156900f, 0.173648177666930414f,
                                  -0.058144828910475774f, -0.286803232711090206f, -0.5f,
                                  -0.686241637868733489f, -0.835487811412936265f, -0.939692620785908317f,
                                  -0.993238357741943023f, -0.993238357741943023f, -0.939692620785908539f,
                                  -0.835487811412936487f};
        constexpr float ts[17] = {0.0f, 0.230615870742440165f, 0.448799180200462167f, 0.642787609686539252f,
                                  0.802123192755043735f, 0.918216106880273997f, 0.984807753012208020f,
                                  0.998308158271268176f, 0.957989512315488900f, 0.866025403784438708f,
                                  0.727373641573048846f, 0.549508978070806231f, 0.342020143325668879f,
                                  0.116092914125229929f, -0.116092914125230123f, -0.342020143325668213f,
                                  -0.549508978070806009f};
        const float sg = INV ? 1.0f : -1.0f;
#pragma unroll
        for (int b = 0; b < 9; b++) {
            float2 t[3] = {u[b], u[9 + b], u[18 + b]};
            dft_small<3, INV>(t);
            u[b] = t[0];
            if (b == 0) {
                u[9] = t[1];
                u[18] = t[2];
            } else {
                u[9 + b] = cmul(t[1], make_float2(tc[b], sg * ts[b]));
                u[18 + b] = cmul(t[2], make_float2(tc[2 * b], sg * ts[2 * b]));
            }
        }
        float2 y[27];
#pragma unroll
        for (int c = 0; c < 3; c++) {
            float2 t[9];
#pragma unroll
            for (int b = 0; b < 9; b++) t[b] = u[9 * c + b];
            dft_small<9, INV>(t);
#pragma unroll
            for (int d = 0; d < 9; d++) y[c + 3 * d] = t[d];
        }
#pragma unroll
        for (int k = 0; k < 27; k++) u[k] = y[k];
    } else if constexpr (R == 15) {
        // Good-Thomas (3 and 5 coprime, no twiddles): input n = (5 n1 + 3 n2) mod 15,
        // output k = (10 k1 + 6 k2) mod 15, so W15^(nk) = W3^(n1 k1) W5^(n2 k2)
        float2 v[3][5];
#pragma unroll
        for (int a = 0; a < 3; a++) {
#pragma unroll
            for (int b = 0; b < 5; b++) v[a][b] = u[(5 * a + 3 * b) % 15];
            dft_small<5, INV>(v[a]);
        }
#pragma unroll
        for (int b = 0; b < 5; b++) {
            float2 t[3] = {v[0][b], v[1][b], v[2][b]};
            dft_small<3, INV>(t);
#pragma unroll
            for (int a = 0; a < 3; a++) u[(10 * a + 6 * b) % 15] = t[a];
        }
    } else if constexpr (R == 8) {
        // two length-4 DFTs (even and odd inputs), the odd ones turned by W8^k
        float2 e[4] = {u[0], u[2], u[4], u[6]}, o[4] = {u[1], u[3], u[5], u[7]};
        dft_small<4, INV>(e);
        dft_small<4, INV>(o);
        constexpr float r = 0.707106781186547524f;
        const float sg = INV ? 1.0f : -1.0f;  // W8 = r (1 + sg i)
        const float2 w1 = make_float2(r * (o[1].x - sg * o[1].y), r * (o[1].y + sg * o[1].x));
        const float2 w2 = make_float2(-sg * o[2].y, sg * o[2].x);
        const float2 w3 = make_float2(r * (-o[3].x - sg * o[3].y), r * (-o[3].y + sg * o[3].x));
        u[0] = cadd(e[0], o[0]);
        u[4] = csub(e[0], o[0]);
        u[1] = cadd(e[1], w1);
        u[5] = csub(e[1], w1);
        u[2] = cadd(e[2], w2);
        u[6] = csub(e[2], w2);
        u[3] = cadd(e[3], w3);
        u[7] = csub(e[3], w3);
    } else {
        // odd R (5, 7): with s_m = u_m + u_(R-m) and d_m = u_m - u_(R-m),
        // y_k = a_k -+ i b_k and y_(R-k) = a_k +- i b_k, where
        // a_k = u_0 + sum_m cos(2 pi m k / R) s_m and b_k = sum_m sin(2 pi m k / R) d_m
        static_assert(R % 2 == 1, "symmetric-pair DFT needs an odd length");
        constexpr int H = (R - 1) / 2;
        float2 s[H], d[H];
        float2 y0 = u[0];
#pragma unroll
        for (int m = 1; m <= H; m++) {
            s[m - 1] = cadd(u[m], u[R - m]);
            d[m - 1] = csub(u[m], u[R - m]);
            y0 = cadd(y0, s[m - 1]);
        }
        float2 y[R];
        y[0] = y0;
#pragma unroll
        for (int k = 1; k <= H; k++) {
            float2 a = u[0], b = make_float2(0.0f, 0.0f);
#pragma unroll
            for (int m = 1; m <= H; m++) {
                const int n = (m * k) % R;
                const float c = dcos<R>(n), sn = dsin<R>(n);
                a.x = fmaf(c, s[m - 1].x, a.x);
                a.y = fmaf(c, s[m - 1].y, a.y);
                b.x = fmaf(sn, d[m - 1].x, b.x);
                b.y = fmaf(sn, d[m - 1].y, b.y);
            }
            const float2 amib = make_float2(a.x + b.y, a.y - b.x);  // a - i b
            const float2 apib = make_float2(a.x - b.y, a.y + b.x);  // a + i b
            y[k] = INV ? apib : amib;
            y[R - k] = INV ? amib : apib;
        }
#pragma unroll
        for (int k = 0; k < R; k++) u[k] = y[k];
    }
}

// LDS pointers keep their address space across the (non-inlined) stage calls,
// so the stages use ds_read / ds_write.
typedef float __attribute__((ext_vector_type(2))) f2v;
typedef __attribute__((address_space(3))) f2v lds_f2;
__device__ __forceinline__ lds_f2* to_lds(float2* p) { return (lds_f2*)p; }
__device__ __forceinline__ float2 ld2(const lds_f2* p) {
    const f2v v = *p;
    return make_float2(v.x, v.y);
}
__device__ __forceinline__ void st2(lds_f2* p, float2 a) { *p = f2v{a.x, a.y}; }

// n / d for 0 <= n < 2^22 and d >= 1, with rd = 1 / d in FP32: the float
// quotient is within one of the exact one, which the remainder corrects.
__device__ __forceinline__ int fdiv(int n, int d, float rd) {
    int q = (int)((float)n * rd);
    const int r = n - q * d;
    q += (r >= d) - (r < 0);
    return q;
}

// One Stockham stage of radix R over `nlines` lines of length L in LDS
// (line stride L), p = product of the radices already applied: butterfly i
// of a line reads u_r = x[i + r L/R], twiddles u_r by W_L^(r k L/(pR)),
// k = i mod p, and writes y[(i - k) R + k + r p]. In place: every thread
// reads its butterflies into registers before any write.
// Each radix is its own function: inlined together, the six stage bodies
// shared one register allocation of more than twice any single one.
// Two thread mappings: when m = L/R <= kFT and it takes no more rounds, a
// thread keeps one butterfly index i for all its rounds (G = kFT/m lines per
// round), so its R - 1 twiddles are read once per stage and a round costs
// the butterfly alone; otherwise the butterflies of the block are dealt out
// in order, kFT per round. FIXED = false compiles the second mapping only
// (the column passes, whose B lines rarely fill whole rounds the first way).
// TW = false compiles no twiddles at all, for a stage that only ever runs
// first (p = 1). TAG gives one kernel stage functions of its own, so that its
// launch bound reaches them (a bound does not constrain a shared callee).
template <int R, bool INV, int BL = kFftBL, bool FIXED = true, bool TW = true, int TAG = 0>
__device__ __noinline__ void fft_stage(lds_f2* __restrict__ s, int L, int nlines, int p,
                                       const lds_f2* __restrict__ tw) {
    constexpr int MAXQ = (BL / R + kFT - 1) / kFT;
    const int m = L / R;
    const int nb = nlines * m;
    const int step = L / (p * R);
    const float rm = 1.0f / (float)m, rp = 1.0f / (float)p;
    float2 u[MAXQ][R];
    const int G = kFT / m;
    if (FIXED && G > 0 && (nlines + G - 1) / G <= (nb + kFT - 1) / kFT) {
        const int l0 = fdiv((int)threadIdx.x, m, rm), i = (int)threadIdx.x - l0 * m;
        const bool act = l0 < G;  // threadIdx.x < G m
        const int k = i - fdiv(i, p, rp) * p;
        float2 w[R];
#pragma unroll
        for (int r = 1; r < R; r++) {
            w[r] = (TW && act) ? ld2(tw + r * k * step) : make_float2(1.0f, 0.0f);  // r k step < L
            if (INV) w[r].y = -w[r].y;
        }
#pragma unroll
        for (int q = 0; q < MAXQ; q++) {  // rounds: ceil(nlines / G) <= MAXQ
            const int line = l0 + q * G;
            if (act && line < nlines) {
                const lds_f2* src = s + line * L + i;
#pragma unroll
                for (int r = 0; r < R; r++) u[q][r] = ld2(src + r * m);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < MAXQ; q++) {
            const int line = l0 + q * G;
            if (act && line < nlines) {
                if (TW && p > 1) {  // p = 1: every twiddle is W^0
#pragma unroll
                    for (int r = 1; r < R; r++) u[q][r] = cmul(u[q][r], w[r]);
                }
                dft_small<R, INV>(u[q]);
                lds_f2* dst = s + line * L + (i - k) * R + k;
#pragma unroll
                for (int r = 0; r < R; r++) st2(dst + r * p, u[q][r]);
            }
        }
        __syncthreads();
        return;
    }
#pragma unroll
    for (int q = 0; q < MAXQ; q++) {
        const int t = (int)threadIdx.x + q * kFT;
        if (t < nb) {
            const int line = fdiv(t, m, rm), i = t - line * m;
            const lds_f2* src = s + line * L + i;
#pragma unroll
            for (int r = 0; r < R; r++) u[q][r] = ld2(src + r * m);
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < MAXQ; q++) {
        const int t = (int)threadIdx.x + q * kFT;
        if (t < nb) {
            const int line = fdiv(t, m, rm), i = t - line * m;
            const int k = i - fdiv(i, p, rp) * p;
            if (TW && p > 1) {  // p = 1: every twiddle is W^0
#pragma unroll
                for (int r = 1; r < R; r++) {
                    float2 w = ld2(tw + r * k * step);  // r k step < L
                    if (INV) w.y = -w.y;
                    u[q][r] = cmul(u[q][r], w);
                }
            }
            dft_small<R, INV>(u[q]);
            lds_f2* dst = s + line * L + (i - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) st2(dst + r * p, u[q][r]);
        }
    }
    __syncthreads();
}

// Transform `nlines` lines staged in LDS (nlines * L <= kFftBL).
template <bool INV>
__device__ __forceinline__ void fft_lines(float2* __restrict__ s_, const FftPlan& pl, int nlines,
                                          const float2* __restrict__ tw_) {
    lds_f2* s = to_lds(s_);
    const lds_f2* tw = to_lds(const_cast<float2*>(tw_));
    int p = 1;
    for (int st = 0; st < pl.nst; st++) {
        const int R = pl.rad[st];
        switch (R) {
            case 2: fft_stage<2, INV>(s, pl.L, nlines, p, tw); break;
            case 3: fft_stage<3, INV>(s, pl.L, nlines, p, tw); break;
            case 4: fft_stage<4, INV>(s, pl.L, nlines, p, tw); break;
            case 5: fft_stage<5, INV>(s, pl.L, nlines, p, tw); break;
            case 7: fft_stage<7, INV>(s, pl.L, nlines, p, tw); break;
            case 9: fft_stage<9, INV>(s, pl.L, nlines, p, tw); break;
            // one round per thread at 4096 staged values; make_lines_plan puts it
            // first (p = 1), so it has no twiddles, and none hold registers
            case 27: fft_stage<27, INV, kFftBL, false, false>(s, pl.L, nlines, p, tw); break;
            default: fft_stage<8, INV>(s, pl.L, nlines, p, tw); break;
        }
        p *= R;
    }
}

__device__ __forceinline__ void load_twiddles(float2* __restrict__ tw, const float2* __restrict__ g, int L) {
    for (int i = (int)threadIdx.x; i < L; i += kFT) tw[i] = g[i];
}

// Lines of one axis of a stack of complex volumes. Line (a, b) of channel c
// starts at c * cstride + b * bstride + a * astride; its elements are
// estride apart. Elements at or beyond n_in read as zero; only the first
// n_out are written back. Blocks take `B` consecutive a (a batch), then b,
// then the channel.
struct LineGeom {
    size_t cstride, bstride, astride, estride;
    int na, nb, nchan, B, nab;
    int n_in, n_out;
    int lb;  // log2 B for k_fft_cols (B a power of two)
    int Bh;  // k_fft_cols: lines per group of the middle stages (a power of two dividing B)
};

// ROWS: elements contiguous (estride 1), loads walk a line first; otherwise
// the B lines of a batch are adjacent in memory (astride 1) and loads walk
// across them first, so both cases read contiguous runs.
template <bool INV, bool ROWS>
__global__ __launch_bounds__(kFT) void k_fft_lines(float2* __restrict__ data, LineGeom g, FftPlan pl,
                                                   const float2* __restrict__ twg) {
    __shared__ float2 s[kFftBL];
    __shared__ float2 tw[kMaxL];
    const int L = pl.L;
    unsigned bid = blockIdx.x;
    const int ab = (int)(bid % (unsigned)g.nab);
    bid /= (unsigned)g.nab;
    const int b = (int)(bid % (unsigned)g.nb);
    const int c = (int)(bid / (unsigned)g.nb);
    if (c >= g.nchan) return;
    const int a0 = ab * g.B;
    const int Bv = min(g.B, g.na - a0);
    load_twiddles(tw, twg, L);
    float2* base = data + (size_t)c * g.cstride + (size_t)b * g.bstride + (size_t)a0 * g.astride;
    const int nin = Bv * g.n_in;
    for (int idx = (int)threadIdx.x; idx < Bv * L; idx += kFT) {
        int e, l;
        if (ROWS) { l = idx / L; e = idx - l * L; }
        else { e = idx / Bv; l = idx - e * Bv; }
        float2 v = make_float2(0.0f, 0.0f);
        if (e < g.n_in) v = base[(size_t)l * g.astride + (size_t)e * g.estride];
        s[l * L + e] = v;
    }
    (void)nin;
    __syncthreads();
    fft_lines<INV>(s, pl, Bv, tw);
    for (int idx = (int)threadIdx.x; idx < Bv * g.n_out; idx += kFT) {
        int e, l;
        if (ROWS) { l = idx / g.n_out; e = idx - l * g.n_out; }
        else { e = idx / Bv; l = idx - e * Bv; }
        base[(size_t)l * g.astride + (size_t)e * g.estride] = s[l * L + e];
    }
}

// Strided passes (y and z): the B lines of a block are B adjacent columns
// (astride 1, B a power of two, lane-fastest in every global access). The
// first radix stage reads its butterfly inputs straight from global memory
// (p = 1: no twiddles), the last one writes its outputs straight back
// (p = L / R: output r of butterfly i is element i + r L/R), so the lines
// touch LDS only between stages.
template <int R, bool INV, int BL>
__device__ __forceinline__ void cols_first(const float2* __restrict__ base, lds_f2* __restrict__ s, int L, int B,
                                           int lb, int Bv, size_t estride, int n_in) {
    // every load of a round of MAXQ butterflies per thread is issued before
    // its first butterfly; batches above BL values (B L > BL) take more rounds
    constexpr int MAXQ = (BL / R + kFT - 1) / kFT;
    const int m = L / R;
    const int nb = B * m;
    float2 u[MAXQ][R];
    for (int t0 = 0; t0 < nb; t0 += MAXQ * kFT) {
#pragma unroll
        for (int q = 0; q < MAXQ; q++) {
            const int t = t0 + (int)threadIdx.x + q * kFT;
            const int l = t & (B - 1), i = t >> lb;
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = i + r * m;
                u[q][r] = (t < nb && l < Bv && e < n_in) ? base[(size_t)e * estride + l] : make_float2(0.0f, 0.0f);
            }
        }
#pragma unroll
        for (int q = 0; q < MAXQ; q++) {
            const int t = t0 + (int)threadIdx.x + q * kFT;
            if (t < nb) {
                const int l = t & (B - 1), i = t >> lb;
                dft_small<R, INV>(u[q]);
                lds_f2* d = s + l * L + i * R;
#pragma unroll
                for (int r = 0; r < R; r++) st2(d + r, u[q][r]);
            }
        }
    }
}

template <int R, bool INV>
__device__ __forceinline__ void cols_last(const lds_f2* __restrict__ s, float2* __restrict__ base,
                                          const lds_f2* __restrict__ tw, int L, int B, int lb, int Bv,
                                          size_t estride, int n_out) {
    const int m = L / R;
    for (int t = (int)threadIdx.x; t < B * m; t += kFT) {
        const int l = t & (B - 1), i = t >> lb;
        float2 u[R];
        const lds_f2* src = s + l * L + i;
#pragma unroll
        for (int r = 0; r < R; r++) u[r] = ld2(src + r * m);
#pragma unroll
        for (int r = 1; r < R; r++) {
            float2 w = ld2(tw + r * i);  // W_L^(r i), r i < L
            if (INV) w.y = -w.y;
            u[r] = cmul(u[r], w);
        }
        dft_small<R, INV>(u);
        if (l < Bv) {
#pragma unroll
            for (int r = 0; r < R; r++) {
                const int e = i + r * m;
                if (e < n_out) base[(size_t)e * estride + l] = u[r];
            }
        }
    }
}

// Whole transform of a length-R line (R <= 9) in registers.
template <int R, bool INV>
__device__ __forceinline__ void cols_single(float2* __restrict__ base, int B, int Bv, size_t estride, int n_in,
                                            int n_out) {
    for (int l = (int)threadIdx.x; l < B; l += kFT) {
        float2 u[R];
#pragma unroll
        for (int r = 0; r < R; r++)
            u[r] = (l < Bv && r < n_in) ? base[(size_t)r * estride + l] : make_float2(0.0f, 0.0f);
        dft_small<R, INV>(u);
        if (l < Bv) {
#pragma unroll
            for (int r = 0; r < R; r++)
                if (r < n_out) base[(size_t)r * estride + l] = u[r];
        }
    }
}

#define TH_RADIX_SWITCH(R, CALL)                          \
    switch (R) {                                           \
        case 2: { constexpr int RR = 2; CALL; } break;     \
        case 3: { constexpr int RR = 3; CALL; } break;     \
        case 4: { constexpr int RR = 4; CALL; } break;     \
        case 5: { constexpr int RR = 5; CALL; } break;     \
        case 7: { constexpr int RR = 7; CALL; } break;     \
        case 9: { constexpr int RR = 9; CALL; } break;     \
        default: { constexpr int RR = 8; CALL; } break;    \
    }
// The last stage of a column pass also takes radix 15 (make_cols_plan).
#define TH_RADIX_SWITCH_LAST(R, CALL)                     \
    switch (R) {                                           \
        case 15: { constexpr int RR = 15; CALL; } break;   \
        default: TH_RADIX_SWITCH(R, CALL);                 \
    }

// MINW: waves per SIMD the kernel is compiled for. Stage groups of 16 lines
// take 6: the bound reaches the stage functions that only this instantiation
// calls (TAG), which then fit 73 VGPRs without spilling (101 unbounded, five
// waves), 1% faster at L = 135; eight waves spill and run 25% slower. Smaller
// groups (L > 256) take no bound: there the bounded code ran 8% slower and
// fetched 1.3x the algorithmic bytes.
template <bool INV, int BL, int MINW>
__global__ __launch_bounds__(kFT, MINW) void k_fft_cols(float2* __restrict__ data, LineGeom g, FftPlan pl,
                                                  const float2* __restrict__ twg) {
    // B L staged values, then L twiddles: sized at launch (at most kColsLds
    // bytes), 18 KB at L = 135 where a static BL-value array took 40 KB and
    // held the pass to four blocks per CU
    extern __shared__ float2 cols_dyn[];
    const int L = pl.L;
    float2* s_ = cols_dyn;
    float2* tw_ = cols_dyn + g.B * L;
    // XCD-major order: neighbouring column batches (which share 128-byte
    // lines when B < 16) run on one XCD and meet in its L2
    unsigned bid = xcd_remap(blockIdx.x, gridDim.x);
    const int ab = (int)(bid % (unsigned)g.nab);
    bid /= (unsigned)g.nab;
    const int b = (int)(bid % (unsigned)g.nb);
    const int c = (int)(bid / (unsigned)g.nb);
    if (c >= g.nchan) return;
    const int B = g.B, lb = g.lb;
    const int a0 = ab * B;
    const int Bv = min(B, g.na - a0);
    float2* base = data + (size_t)c * g.cstride + (size_t)b * g.bstride + (size_t)a0;
    if (pl.nst == 0) return;  // length 1: the identity
    if (pl.nst == 1) {
        TH_RADIX_SWITCH_LAST(pl.rad[0], (cols_single<RR, INV>(base, B, Bv, g.estride, g.n_in, g.n_out)));
        return;
    }
    load_twiddles(tw_, twg, L);
    lds_f2* s = to_lds(s_);
    const lds_f2* tw = to_lds(tw_);
    TH_RADIX_SWITCH(pl.rad[0], (cols_first<RR, INV, BL>(base, s, L, B, lb, Bv, g.estride, g.n_in)));
    __syncthreads();
    // the middle stages take the batch in groups of Bh lines (Bh L <= BL, the
    // stages' register rounds), so that B can reach 16 (128-byte rows) for
    // L above 256
    for (int h0 = 0; h0 < B; h0 += g.Bh) {
        int p = pl.rad[0];
        for (int st = 1; st < pl.nst - 1; st++) {
            const int R = pl.rad[st];
            TH_RADIX_SWITCH(R, (fft_stage<RR, INV, BL, false, true, MINW>(s + h0 * L, L, g.Bh, p, tw)));
            p *= R;
        }
    }
    TH_RADIX_SWITCH_LAST(pl.rad[pl.nst - 1],
                         (cols_last<RR, INV>(s, base, tw, L, B, lb, Bv, g.estride, g.n_out)));
}

// ---------------------------------------------------------------- spectral vote kernels
struct SpecGeom {
    int n0, n1, n2;       // field (n2 fastest)
    int L0, L1, L2;       // FFT lengths
    int h0, h1, h2;       // octant extents L/2 + 1
    int nf, nz;           // real features of the product (the 15 combined quartic features), complex channels
    int stick, plate;     // which votes the features carry
    int atv;              // analytical tensor voting: s1 m(e2) + 2 s2 circle(e0) + 3 s3 sphere
    float alpha;          // scale of the plate part: sticknorm3 (vote) or 2 (ATV)
    int P2;               // row stride in complex values (L2 rounded up to 16)
    size_t cs;            // complex values per channel volume
    size_t ocs;           // kernel-spectrum values per channel (octant)
};

// The 15 quartic monomials of (x, y, z) in the order of mono_a / mono_b,
// times the scale w.
__device__ __forceinline__ void mono15(const float4 P, float* __restrict__ A) {
    const float x = P.x, y = P.y, z = P.z;
    const float sx = P.w * x, sy = P.w * y, sz = P.w * z;
    const float x2 = x * x, y2 = y * y, z2 = z * z, xy = x * y, xz = x * z, yz = y * z;
    A[0] = sx * x * x2;   A[1] = sx * x * xy;   A[2] = sx * x * xz;
    A[3] = sx * x * y2;   A[4] = sx * x * yz;   A[5] = sx * x * z2;
    A[6] = sx * y * y2;   A[7] = sx * y * yz;   A[8] = sx * y * z2;   A[9] = sx * z * z2;
    A[10] = sy * y * y2;  A[11] = sy * y * yz;  A[12] = sy * y * z2;  A[13] = sy * z * z2;
    A[14] = sz * z * z2;
}

// Circle average over the great circle perpendicular to n of each quartic
// monomial, as a quartic in n: avg_f(n) = sum_g A[f][g] n^g with exact
// rational A (42 nonzero terms). The plate vote of normal n integrates the
// stick vote over that circle, so its offset kernels are the stick kernels
// times sticknorm3 A and its features fold into the stick features:
// F_f += sticknorm3 t sum_g A[f][g] m_g(e0). The ATV's e0 part is likewise
// twice the circle average of its e2 part and its constant part three times
// the sphere average. N.w carries the scale (t times sticknorm3, or 2 s2).
struct CircleTerm {
    int f, g;
    float a;
};
__device__ __forceinline__ void add_circle_average(const float4 N, float* __restrict__ A) {
    float B[kMono];
    mono15(N, B);
    constexpr CircleTerm t[42] = {
        {0, 10, 0.375f}, {0, 12, 0.75f}, {0, 14, 0.375f}, {1, 6, -0.375f}, {1, 8, -0.375f}, {2, 7, -0.375f},
        {2, 9, -0.375f}, {3, 3, 0.375f}, {3, 5, 0.125f}, {3, 12, 0.125f}, {3, 14, 0.125f}, {4, 4, 0.25f},
        {4, 11, -0.125f}, {4, 13, -0.125f}, {5, 3, 0.125f}, {5, 5, 0.375f}, {5, 10, 0.125f},
        {5, 12, 0.125f}, {6, 1, -0.375f}, {6, 8, -0.375f}, {7, 2, -0.125f}, {7, 7, 0.25f}, {7, 9, -0.125f},
        {8, 1, -0.125f}, {8, 6, -0.125f}, {8, 8, 0.25f}, {9, 2, -0.375f}, {9, 7, -0.375f}, {10, 0, 0.375f},
        {10, 5, 0.75f}, {10, 14, 0.375f}, {11, 4, -0.375f}, {11, 13, -0.375f}, {12, 0, 0.125f},
        {12, 3, 0.125f}, {12, 5, 0.125f}, {12, 12, 0.375f}, {13, 4, -0.375f}, {13, 11, -0.375f},
        {14, 0, 0.375f}, {14, 3, 0.75f}, {14, 10, 0.375f}};
#pragma unroll
    for (int i = 0; i < 42; i++) A[t[i].f] = fmaf(t[i].a, B[t[i].g], A[t[i].f]);
}

// Forward x pass fused with the features. A block takes B rows (z, y) of the
// field and the complex channels in groups of C (all of them unless a row's
// lines exceed 64 KB of LDS): every element computes all features of its
// voxel from one payload read, channel j = features (2j, 2j+1) as (real,
// imaginary), and the lines (channel, row) are transformed FG at a time.
// Each channel group re-reads its voxels.
// FROM_TENSORS: each element eigen-decomposes its voxel of the input field
// (layout) instead of reading a payload, once per channel group (a second
// eigen-solve costs less than writing and re-reading a 32-byte payload per
// voxel).
// Four waves per SIMD (at most 128 VGPRs): unbounded, the eigen-solve and the
// features take 136 and three waves; bounded, they fit in 120 without a spill.
template <bool FROM_TENSORS>
__global__ __launch_bounds__(kFT, 4) void k_spec_xfwd(const float4* __restrict__ ps, const float4* __restrict__ pp,
                                                   const float* __restrict__ tensors, int layout, SpecGeom g, int B,
                                                   int C, int FG, FftPlan pl, const float2* __restrict__ twg,
                                                   float2* __restrict__ Z) {
    extern __shared__ float2 xs_dyn[];  // C B L staged values, then L twiddles
    const int L = pl.L;
    float2* s = xs_dyn;
    float2* tw = xs_dyn + C * B * L;
    const int row0 = (int)blockIdx.x * B;
    const int nrows = g.n0 * g.n1;
    if (row0 >= nrows) return;
    const int Bv = min(B, nrows - row0);
    load_twiddles(tw, twg, L);
    const int npad = L - g.n2;
    // index splits below use fdiv (a float reciprocal), not integer division
    const float rB = 1.0f / (float)Bv, rn1 = 1.0f / (float)g.n1, rn2 = 1.0f / (float)g.n2,
                rpad = 1.0f / (float)max(npad, 1);
    for (int j0 = 0; j0 < g.nz; j0 += C) {
        const int Cv = min(C, g.nz - j0);
        if (j0) __syncthreads();  // the previous group's stores have read s
        // the zero padding of every line, then one element per voxel (the
        // eigen-decompositions fill whole waves: 2 rows of 128 are one pass)
        for (int idx = (int)threadIdx.x; idx < Cv * Bv * npad; idx += kFT) {
            const int line = fdiv(idx, npad, rpad);
            s[line * L + g.n2 + (idx - line * npad)] = make_float2(0.0f, 0.0f);
        }
        for (int idx = (int)threadIdx.x; idx < Bv * g.n2; idx += kFT) {
            const int b = fdiv(idx, g.n2, rn2), e = idx - b * g.n2;
            // the 15 combined features and a zero (the last channel's imaginary
            // part); the stick features are written first, zero when the vote
            // has no stick part (zeroing A ahead of the eigen-solve kept 15
            // registers live through it and spilled 80 bytes per lane)
            float A[kMono + 1];
            A[kMono] = 0.0f;
            {
                const size_t vox = (size_t)(row0 + b) * g.n2 + e;
                float4 S = make_float4(0.0f, 0.0f, 0.0f, 0.0f), N = S;
                float s3 = 0.0f;
                if (FROM_TENSORS) {
                    float l0, l1, l2, e0[3], e2[3];
                    voxel_eig<false>(tensors, layout, vox, l0, l1, l2, e0, e2);
                    if (g.atv) {  // (e2, l2 - l1), (e0, l1 - l0), l0 (the ATV eigen-subspace weights)
                        S = make_float4(e2[0], e2[1], e2[2], l2 - l1);
                        N = make_float4(e0[0], e0[1], e0[2], l1 - l0);
                        s3 = l0;
                    } else {
                        stick_plate_payload(l0, l1, l2, e0, e2, S, N);
                    }
                } else {
                    if (g.stick) S = ps[vox];
                    if (g.plate) N = pp[vox];
                }
                mono15(g.stick ? S : make_float4(0.0f, 0.0f, 0.0f, 0.0f), A);
                if (g.plate) add_circle_average(make_float4(N.x, N.y, N.z, N.w * g.alpha), A);
                if (g.atv) {  // three times the sphere averages: 1/5 for x^4, 1/15 for x^2 y^2
                    A[0] = fmaf(0.6f, s3, A[0]);
                    A[3] = fmaf(0.2f, s3, A[3]);
                    A[5] = fmaf(0.2f, s3, A[5]);
                    A[10] = fmaf(0.6f, s3, A[10]);
                    A[12] = fmaf(0.2f, s3, A[12]);
                    A[14] = fmaf(0.6f, s3, A[14]);
                }
            }
#pragma unroll
            for (int j = 0; j < (kMono + 1) / 2; j++)
                if (j >= j0 && j < j0 + Cv) s[((j - j0) * Bv + b) * L + e] = make_float2(A[2 * j], A[2 * j + 1]);
        }
        __syncthreads();
        // the staged lines in groups of FG (each group fits the stages' register rounds)
        for (int l0 = 0; l0 < Cv * Bv; l0 += FG) fft_lines<false>(s + l0 * L, pl, min(FG, Cv * Bv - l0), tw);
        // each wave stores whole lines (their bases are wave-uniform)
        for (int line = (int)threadIdx.x >> 6; line < Cv * Bv; line += kFT / 64) {
            const int ln = __builtin_amdgcn_readfirstlane(line);
            const int c = fdiv(ln, Bv, rB), b = ln - c * Bv;
            const int row = row0 + b;
            const int z = fdiv(row, g.n1, rn1), y = row - z * g.n1;
            float2* dst = Z + (size_t)(j0 + c) * g.cs + ((size_t)z * g.L1 + y) * g.P2;
            for (int e = (int)threadIdx.x & 63; e < L; e += 64) dst[e] = s[ln * L + e];
        }
    }
}

// Pointwise product, one thread per octant group (kx, ky, kz) in
// [0, h2) x [0, h1) x [0, h0). Slot bits (x, y, z flip) pick the frequency
// (+-kx, +-ky, +-kz). For slots 0..3 (kz unflipped):
//   F_f(k) = (Z_j(k) + conj Z_j(-k)) / 2        (f = 2j, real part's spectrum)
//          = -i (Z_j(k) - conj Z_j(-k)) / 2     (f = 2j + 1)
//   Y_c(k) = sum_f Khat_{c,f}(k) F_f(k),  Khat(flipped axes) = parity sign x octant value
// (the 1/2 and the inverse-FFT 1/(L0 L1 L2) are folded into the octant values);
// then W_i(k) = Y_{2i}(k) + i Y_{2i+1}(k) and W_i(-k) = conj Y_{2i}(k) + i conj Y_{2i+1}(k)
// overwrite Z_0..Z_2 (each group's frequencies belong to this thread only).
__global__ __launch_bounds__(256) void k_spec_mul(float2* __restrict__ Z, const float* __restrict__ Kh,
                                                  SpecGeom g) {
    const size_t gid = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= g.ocs) return;
    const int kx = (int)(gid % (size_t)g.h2);
    const int ky = (int)((gid / (size_t)g.h2) % (size_t)g.h1);
    const int kz = (int)(gid / ((size_t)g.h2 * g.h1));
    const int fx[2] = {kx, kx ? g.L2 - kx : 0};
    const int fy[2] = {ky, ky ? g.L1 - ky : 0};
    const int fz[2] = {kz, kz ? g.L0 - kz : 0};
    size_t off[8];
#pragma unroll
    for (int sl = 0; sl < 8; sl++)
        off[sl] = ((size_t)fz[(sl >> 2) & 1] * g.L1 + fy[(sl >> 1) & 1]) * g.P2 + fx[sl & 1];
    float2 Y[4][6];
#pragma unroll
    for (int sl = 0; sl < 4; sl++)
#pragma unroll
        for (int c = 0; c < 6; c++) Y[sl][c] = make_float2(0.0f, 0.0f);

    // the next channel's eight values are in flight while this one is multiplied
    float2 zn[8];
#pragma unroll
    for (int sl = 0; sl < 8; sl++) zn[sl] = Z[off[sl]];
    for (int j = 0; j < g.nz; j++) {
        float2 z8[8];
#pragma unroll
        for (int sl = 0; sl < 8; sl++) z8[sl] = zn[sl];
        if (j + 1 < g.nz) {
            const float2* zq = Z + (size_t)(j + 1) * g.cs;
#pragma unroll
            for (int sl = 0; sl < 8; sl++) zn[sl] = zq[off[sl]];
        }
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int f = 2 * j + h;
            if (f >= g.nf) break;
            const int m = f % kMono;  // the constant feature (ATV) shares x^4's parities (all even)
            const int am = mono_a(m), bm = mono_b(m);
            float2 F[4];
#pragma unroll
            for (int sl = 0; sl < 4; sl++) {
                const float2 a = z8[sl], bc = make_float2(z8[sl ^ 7].x, -z8[sl ^ 7].y);
                F[sl] = h == 0 ? cadd(a, bc) : make_float2(a.y - bc.y, bc.x - a.x);
            }
#pragma unroll
            for (int c = 0; c < 6; c++) {
                const float k = Kh[(size_t)(c * g.nf + f) * g.ocs + gid];
                const bool ox = ((comp_nx(c) + am) & 1) != 0, oy = ((comp_ny(c) + bm) & 1) != 0;
                const float k1 = ox ? -k : k, k2 = oy ? -k : k, k3 = (ox != oy) ? -k : k;
                Y[0][c].x = fmaf(k, F[0].x, Y[0][c].x);  Y[0][c].y = fmaf(k, F[0].y, Y[0][c].y);
                Y[1][c].x = fmaf(k1, F[1].x, Y[1][c].x); Y[1][c].y = fmaf(k1, F[1].y, Y[1][c].y);
                Y[2][c].x = fmaf(k2, F[2].x, Y[2][c].x); Y[2][c].y = fmaf(k2, F[2].y, Y[2][c].y);
                Y[3][c].x = fmaf(k3, F[3].x, Y[3][c].x); Y[3][c].y = fmaf(k3, F[3].y, Y[3][c].y);
            }
        }
    }
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const float2 ya = Y[sl][2 * i], yb = Y[sl][2 * i + 1];
            float2* zc = Z + (size_t)i * g.cs;
            zc[off[sl]] = make_float2(ya.x - yb.y, ya.y + yb.x);
            zc[off[sl ^ 7]] = make_float2(ya.x + yb.y, yb.x - ya.y);
        }
    }
}

// Inverse x pass fused with the output: rows (z, y) of the field, the three
// output channels of B rows together; W_0 = (xx, xy), W_1 = (xz, yy),
// W_2 = (yz, zz) as (real, imaginary) parts.
__global__ __launch_bounds__(kFT) void k_spec_xinv(const float2* __restrict__ Z, SpecGeom g, int B, FftPlan pl,
                                                   const float2* __restrict__ twg, float* __restrict__ out,
                                                   int layout, float* __restrict__ evals, unsigned eflags,
                                                   unsigned long long* __restrict__ n_refined) {
    __shared__ float2 s[kFftBL];
    __shared__ float2 tw[kMaxL];
    const int L = pl.L;
    const int row0 = (int)blockIdx.x * B;
    const int nrows = g.n0 * g.n1;
    if (row0 >= nrows) return;
    const int Bv = min(B, nrows - row0);
    load_twiddles(tw, twg, L);
    // index splits use fdiv (a float reciprocal), not integer division
    const float rL = 1.0f / (float)L, rn1 = 1.0f / (float)g.n1, rn2 = 1.0f / (float)g.n2;
    for (int idx = (int)threadIdx.x; idx < 3 * Bv * L; idx += kFT) {
        const int line = fdiv(idx, L, rL), e = idx - line * L;  // line = 3 l + i
        const int l = fdiv(line, 3, 1.0f / 3.0f), i = line - 3 * l;
        const int row = row0 + l;
        const int z = fdiv(row, g.n1, rn1), y = row - z * g.n1;
        s[line * L + e] = Z[(size_t)i * g.cs + ((size_t)z * g.L1 + y) * g.P2 + e];
    }
    __syncthreads();
    fft_lines<true>(s, pl, 3 * Bv, tw);
    for (int idx = (int)threadIdx.x; idx < Bv * g.n2; idx += kFT) {
        const int l = fdiv(idx, g.n2, rn2), x = idx - l * g.n2;
        const float2 w0 = s[(3 * l + 0) * L + x], w1 = s[(3 * l + 1) * L + x], w2 = s[(3 * l + 2) * L + x];
        const size_t v = (size_t)(row0 + l) * g.n2 + x;
        if (layout == TIRA_LAYOUT_MAT9) {
            float* o = out + v * 9;
            o[0] = w0.x; o[1] = w0.y; o[2] = w1.x;
            o[3] = w0.y; o[4] = w1.y; o[5] = w2.x;
            o[6] = w1.x; o[7] = w2.x; o[8] = w2.y;
        } else {
            float* o = out + v * 6;
            o[0] = w0.x; o[1] = w0.y; o[2] = w1.x; o[3] = w1.y; o[4] = w2.x; o[5] = w2.y;
        }
        // eigenvalues of the tensor just written, from the same floats
        // (xx, xy, yy, xz, yz, zz) = (w0.x, w0.y, w1.y, w1.x, w2.x, w2.y)
        if (evals) eig3_tensor_values(w0.x, w0.y, w1.y, w1.x, w2.x, w2.y, eflags, evals + v * 3, n_refined);
    }
}

// ---------------------------------------------------------------- kernel spectra
// Separable FP64 transforms of the offset kernels K[ch][dw][dv][du] (parity
// of channel ch along x / y / z in bits 0 / 1 / 2 of par[ch]): cosine along
// even axes, sine along odd ones, trig tables trig[(odd * h + k) * w + d + r]
// = cos / sin(2 pi k d / L).
__global__ __launch_bounds__(256) void k_khat_x(const double* __restrict__ K, const double* __restrict__ trig,
                                                const unsigned char* __restrict__ par, double* __restrict__ T1,
                                                int w, int h2, int nch) {
    const size_t n = (size_t)nch * w * w * h2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % (size_t)h2);
        const size_t line = i / (size_t)h2;  // (ch, dw, dv)
        const int ch = (int)(line / ((size_t)w * w));
        const double* kr = K + line * w;
        const double* tr = trig + ((size_t)(par[ch] & 1) * h2 + kx) * w;
        double acc = 0.0;
        for (int d = 0; d < w; d++) acc = fma(kr[d], tr[d], acc);
        T1[i] = acc;
    }
}

__global__ __launch_bounds__(256) void k_khat_y(const double* __restrict__ T1, const double* __restrict__ trig,
                                                const unsigned char* __restrict__ par, double* __restrict__ T2,
                                                int w, int h1, int h2, int nch) {
    const size_t n = (size_t)nch * w * h1 * h2;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % (size_t)h2);
        const int ky = (int)((i / (size_t)h2) % (size_t)h1);
        const size_t cw = i / ((size_t)h2 * h1);  // (ch, dw)
        const int ch = (int)(cw / (size_t)w);
        const double* tr = trig + ((size_t)((par[ch] >> 1) & 1) * h1 + ky) * w;
        const double* src = T1 + cw * w * h2 + kx;  // T1[ch][dw][dv][kx]
        double acc = 0.0;
        for (int d = 0; d < w; d++) acc = fma(src[(size_t)d * h2], tr[d], acc);
        T2[i] = acc;
    }
}

__global__ __launch_bounds__(256) void k_khat_z(const double* __restrict__ T2, const double* __restrict__ trig,
                                                const unsigned char* __restrict__ par, float* __restrict__ Kh,
                                                int w, int h0, int h1, int h2, int nch, double scale) {
    const size_t plane = (size_t)h1 * h2;
    const size_t n = (size_t)nch * h0 * plane;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const size_t kyx = i % plane;
        const int kz = (int)((i / plane) % (size_t)h0);
        const int ch = (int)(i / (plane * h0));
        const unsigned pb = par[ch];
        const double* tr = trig + ((size_t)((pb >> 2) & 1) * h0 + kz) * w;
        const double* src = T2 + (size_t)ch * w * plane + kyx;  // T2[ch][dw][ky][kx]
        double acc = 0.0;
        for (int d = 0; d < w; d++) acc = fma(src[(size_t)d * plane], tr[d], acc);
        const int odd = (pb & 1) + ((pb >> 1) & 1) + ((pb >> 2) & 1);
        Kh[i] = (float)((odd == 2 ? -scale : scale) * acc);  // (-i)^2 for two sine axes
    }
}

// ---------------------------------------------------------------- host: offset kernels
// Fibonacci points on the unit sphere; the monomial matrix at these points
// has full rank 15 (even polynomials of degree <= 4 restricted to the sphere).
static void fit_points(std::vector<double>& Q, int npts) {
    Q.resize((size_t)npts * 3);
    const double ga = M_PI * (3.0 - std::sqrt(5.0));
    for (int i = 0; i < npts; i++) {
        const double z = 1.0 - (2.0 * i + 1.0) / npts;
        const double rr = std::sqrt(std::max(0.0, 1.0 - z * z));
        Q[i * 3 + 0] = rr * std::cos(ga * i);
        Q[i * 3 + 1] = rr * std::sin(ga * i);
        Q[i * 3 + 2] = z;
    }
}

static double mono_eval(const double* q, int m) {
    return std::pow(q[0], mono_a(m)) * std::pow(q[1], mono_b(m)) * std::pow(q[2], mono_c(m));
}

// Least-squares projector P (15 x npts): coefficients = P * values.
static bool fit_projector(const std::vector<double>& Q, int npts, std::vector<double>& P) {
    std::vector<double> A((size_t)npts * kMono), N(kMono * 2 * kMono, 0.0);
    for (int i = 0; i < npts; i++)
        for (int m = 0; m < kMono; m++) A[(size_t)i * kMono + m] = mono_eval(&Q[i * 3], m);
    // [A^T A | I] -> [I | (A^T A)^-1], Gauss-Jordan with partial pivoting.
    for (int r = 0; r < kMono; r++) {
        for (int c = 0; c < kMono; c++) {
            double s = 0.0;
            for (int i = 0; i < npts; i++) s += A[(size_t)i * kMono + r] * A[(size_t)i * kMono + c];
            N[r * 2 * kMono + c] = s;
        }
        N[r * 2 * kMono + kMono + r] = 1.0;
    }
    for (int c = 0; c < kMono; c++) {
        int piv = c;
        for (int r = c + 1; r < kMono; r++)
            if (std::fabs(N[r * 2 * kMono + c]) > std::fabs(N[piv * 2 * kMono + c])) piv = r;
        if (std::fabs(N[piv * 2 * kMono + c]) < 1e-12) return false;
        if (piv != c)
            for (int k = 0; k < 2 * kMono; k++) std::swap(N[c * 2 * kMono + k], N[piv * 2 * kMono + k]);
        const double inv = 1.0 / N[c * 2 * kMono + c];
        for (int k = 0; k < 2 * kMono; k++) N[c * 2 * kMono + k] *= inv;
        for (int r = 0; r < kMono; r++) {
            if (r == c) continue;
            const double f = N[r * 2 * kMono + c];
            if (f != 0.0)
                for (int k = 0; k < 2 * kMono; k++) N[r * 2 * kMono + k] -= f * N[c * 2 * kMono + k];
        }
    }
    P.assign((size_t)kMono * npts, 0.0);
    for (int m = 0; m < kMono; m++)
        for (int i = 0; i < npts; i++) {
            double s = 0.0;
            for (int k = 0; k < kMono; k++) s += N[m * 2 * kMono + kMono + k] * A[(size_t)i * kMono + k];
            P[(size_t)m * npts + i] = s;
        }
    return true;
}

static void sym_add(double V[6], double f, const double a[3], const double b[3]) {  // f (a b^T + b a^T) / 2
    V[0] += f * a[0] * b[0];
    V[1] += 0.5 * f * (a[0] * b[1] + a[1] * b[0]);
    V[2] += 0.5 * f * (a[0] * b[2] + a[2] * b[0]);
    V[3] += f * a[1] * b[1];
    V[4] += 0.5 * f * (a[1] * b[2] + a[2] * b[1]);
    V[5] += f * a[2] * b[2];
}

// Pair votes (FP64) of a unit voter vector q at unit offset d (d = 0 at the
// voter itself), attenuations c1, c2, power 1, normalised as the gather path
// normalises them (1 / sticknorm3, 1 / pi).
static void stick_vote64(const double q[3], const double d[3], double c1, double c2, double sn, double V[6]) {
    const double t = q[0] * d[0] + q[1] * d[1] + q[2] * d[2];
    const double eta = c1 * (1.0 - t * t) + c2 * t * t;
    const double r[3] = {q[0] - 2 * t * d[0], q[1] - 2 * t * d[1], q[2] - 2 * t * d[2]};
    for (int k = 0; k < 6; k++) V[k] = 0.0;
    sym_add(V, eta * sn, r, r);
}

static void plate_vote64(const double n[3], const double d[3], bool self, const std::vector<double>& al,
                         const std::vector<double>& be, const std::vector<double>& ga,
                         const std::vector<double>& de, double V[6]) {
    for (int k = 0; k < 6; k++) V[k] = 0.0;
    const double s = n[0] * d[0] + n[1] * d[1] + n[2] * d[2];
    const double x = 1.0 - s * s;
    auto poly = [x](const std::vector<double>& c) {
        double r = 0.0;
        for (size_t i = c.size(); i-- > 0;) r = r * x + c[i];
        return r;
    };
    const double A = poly(al) / M_PI, B = poly(be) / M_PI;
    V[0] += A; V[3] += A; V[5] += A;
    sym_add(V, B, n, n);
    if (!self) {
        const double G = s * poly(ga) / M_PI, D = 2.0 * poly(de) / M_PI;  // de holds delta / 2
        sym_add(V, 2.0 * G, n, d);
        sym_add(V, D, d, d);
    }
}

// Analytical tensor voting pair vote (tira::atv, tensorvote.h:688-760) in
// its collapsed form, FP64: voter eigen-subspace weights (s1, s2, s3) =
// (l2 - l1, l1 - l0, l0), eigenvectors e0, e2, unit offset v (rho = 1) or the
// voter itself (v = 0, rho = 0), attenuation c:
//   KI I + K00 e0 e0^T + K22 e2 e2^T + Kvv v v^T + K0 (v e0^T + e0 v^T) + K2 (v e2^T + e2 v^T).
static void atv_vote64(const double e0[3], const double e2[3], const double v[3], double rho, double c, double s1,
                       double s2, double s3, double V[6]) {
    const double a = e2[0] * v[0] + e2[1] * v[1] + e2[2] * v[2];
    const double b = e0[0] * v[0] + e0[1] * v[1] + e0[2] * v[2];
    const double S1 = c * s1, S2 = c * s2, S3 = c * s3;
    const double K22 = S1 * (1.0 - a * a);
    const double K2 = -2.0 * a * K22;
    const double K00 = -S2 * (1.0 - rho / 4.0 + 0.75 * b * b);
    const double K0 = S2 * b * (1.0 + 1.5 * b * b);
    const double KI = S2 * (1.0 - rho / 4.0 + 0.25 * b * b) + S3 * (1.0 - rho / 5.0);
    const double Kvv = -S2 * (0.5 + b * b + 3.0 * b * b * b * b) - 0.4 * S3 + 4.0 * a * a * K22;
    for (int k = 0; k < 6; k++) V[k] = 0.0;
    V[0] += KI; V[3] += KI; V[5] += KI;
    sym_add(V, K00, e0, e0);
    sym_add(V, K22, e2, e2);
    sym_add(V, Kvv, v, v);
    sym_add(V, 2.0 * K0, v, e0);
    sym_add(V, 2.0 * K2, v, e2);
}

// K[ch][dw][dv][du], ch = c * nf + f (see tira_spectral_offset_kernels). With
// `atv`, the features are s1 m(e2) (15), s2 m(e0) (15) and s3 (1), the window
// half-width w / 2 and the attenuation exp(-|d|^2 / s1^2) of ATV.
static int offset_kernels(unsigned w, float s1, float s2, bool stick, bool plate, bool atv, std::vector<double>& K,
                          int& nf) {
    const int r = (int)(w / 2), W = 2 * r + 1;
    if (atv) {
        stick = plate = true;
        s2 = 0.0f;
    }
    nf = (stick ? kMono : 0) + (plate ? kMono : 0) + (atv ? 1 : 0);
    const int npts = 32;
    std::vector<double> Q, P;
    fit_points(Q, npts);
    if (!fit_projector(Q, npts, P)) {
        set_error("spectral vote: degenerate fitting points (internal error)");
        return TIRA_ERR_INVALID_ARG;
    }
    const double sn = 1.0 / (double)sticknorm3_host(s1, s2, 1);
    const size_t noff = (size_t)W * W * W;
    K.assign((size_t)6 * nf * noff, 0.0);
    std::vector<double> al, be, ga, de, Vs((size_t)npts * 6), Vp((size_t)npts * 6);
    size_t o = 0;
    for (int dw = -r; dw <= r; dw++)
        for (int dv = -r; dv <= r; dv++)
            for (int du = -r; du <= r; du++, o++) {
                const double l2 = (double)du * du + (double)dv * dv + (double)dw * dw;
                double d[3] = {0.0, 0.0, 0.0};
                if (l2 > 0) { const double L = std::sqrt(l2); d[0] = du / L; d[1] = dv / L; d[2] = dw / L; }
                const double c1 = s1 > 0 ? std::exp(-l2 / ((double)s1 * s1)) : (l2 > 0 ? 0.0 : 1.0);
                const double c2 = s2 > 0 ? std::exp(-l2 / ((double)s2 * s2)) : (l2 > 0 ? 0.0 : 1.0);
                if (plate && !atv) {
                    if (l2 == 0) {
                        al.assign(2, 0.0); be.assign(2, 0.0); ga.assign(2, 0.0); de.assign(3, 0.0);
                        al[0] = 0.5 * M_PI * c1;
                        be[0] = -0.5 * M_PI * c1;
                    } else {
                        plate_polys(1, c1, c2, al, be, ga, de);
                    }
                }
                const double zero[3] = {0.0, 0.0, 0.0};
                const double rho = l2 > 0 ? 1.0 : 0.0;
                for (int i = 0; i < npts; i++) {
                    if (atv) {
                        atv_vote64(zero, &Q[i * 3], d, rho, c1, 1.0, 0.0, 0.0, &Vs[(size_t)i * 6]);
                        atv_vote64(&Q[i * 3], zero, d, rho, c1, 0.0, 1.0, 0.0, &Vp[(size_t)i * 6]);
                        continue;
                    }
                    if (stick) stick_vote64(&Q[i * 3], d, c1, c2, sn, &Vs[(size_t)i * 6]);
                    if (plate) plate_vote64(&Q[i * 3], d, l2 == 0, al, be, ga, de, &Vp[(size_t)i * 6]);
                }
                if (atv) {  // the ball weight s3: a constant kernel
                    double V3[6];
                    atv_vote64(zero, zero, d, rho, c1, 0.0, 0.0, 1.0, V3);
                    for (int c = 0; c < 6; c++) K[((size_t)c * nf + 2 * kMono) * noff + o] = V3[c];
                }
                for (int kind = 0; kind < 2; kind++) {
                    if ((kind == 0 && !stick) || (kind == 1 && !plate)) continue;
                    const std::vector<double>& V = kind == 0 ? Vs : Vp;
                    const int f0 = (kind == 1 && stick) ? kMono : 0;
                    for (int m = 0; m < kMono; m++)
                        for (int c = 0; c < 6; c++) {
                            double acc = 0.0;
                            for (int i = 0; i < npts; i++) acc += P[(size_t)m * npts + i] * V[(size_t)i * 6 + c];
                            K[((size_t)c * nf + f0 + m) * noff + o] = acc;
                        }
                }
            }
    return TIRA_OK;
}

// ---------------------------------------------------------------- host: spectra cache
struct SpecKey {
    int L0, L1, L2, w;
    float s1, s2;
    int stick, plate, atv;
    bool operator==(const SpecKey& o) const {
        return L0 == o.L0 && L1 == o.L1 && L2 == o.L2 && w == o.w && s1 == o.s1 && s2 == o.s2 &&
               stick == o.stick && plate == o.plate && atv == o.atv;
    }
};

// Kernel spectra of the last few parameter sets of this host thread (a
// caller alternating between shapes or scales rebuilds nothing), evicted
// least recently used beyond kSpecEntries or 1/8 of the device's memory.
struct SpecEntry {
    SpecKey key{0, 0, 0, 0, 0, 0, 0, 0, 0};
    int device = -1;
    void* buf = nullptr;  // twiddles (3 axes) then the octant spectra
    size_t bytes = 0;
    unsigned long long used = 0;
};
constexpr size_t kSpecEntries = 8;
struct SpecCache {
    std::vector<SpecEntry> e;
    unsigned long long clock = 0;
    unsigned long long builds = 0;
    ~SpecCache() {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return;
        for (auto& x : e)
            if (x.buf && hipSetDevice(x.device) == hipSuccess) (void)hipFree(x.buf);
        (void)hipSetDevice(cur);
    }
};
static thread_local SpecCache g_spec;

// Frees cache entry i (its device made current for the free).
static void spec_evict(size_t i) {
    SpecEntry& x = g_spec.e[i];
    if (x.buf) {
        int cur = 0;
        (void)hipGetDevice(&cur);
        if (x.device != cur) (void)hipSetDevice(x.device);
        (void)hipFree(x.buf);
        if (x.device != cur) (void)hipSetDevice(cur);
    }
    g_spec.e.erase(g_spec.e.begin() + (long)i);
}

static size_t spec_lru(int dev_only) {
    size_t best = g_spec.e.size();
    for (size_t i = 0; i < g_spec.e.size(); i++)
        if ((dev_only < 0 || g_spec.e[i].device == dev_only) &&
            (best == g_spec.e.size() || g_spec.e[i].used < g_spec.e[best].used))
            best = i;
    return best;
}

// Bytes the cache holds on device `dev`.
static size_t spec_held(int dev) {
    size_t b = 0;
    for (const auto& x : g_spec.e)
        if (x.device == dev) b += x.bytes;
    return b;
}

static void twiddles(int L, float2* t) {
    for (int j = 0; j < L; j++) {
        const double a = -2.0 * M_PI * (double)j / (double)L;
        t[j] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
}

static size_t tw_bytes(int L) { return ((size_t)L * 8 + 255) & ~size_t(255); }

// Kernel spectra for `key` (built and uploaded unless cached).
static int spectra_for(const SpecKey& key, const SpecGeom& g, hipStream_t st, const float2** tw0,
                       const float2** tw1, const float2** tw2, const float** kh) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    const size_t tb = tw_bytes(key.L0) + tw_bytes(key.L1) + tw_bytes(key.L2);
    const size_t kbytes = (size_t)6 * g.nf * g.ocs * 4;
    auto set_ptrs = [&](const SpecEntry& x) {
        char* b = (char*)x.buf;
        *tw0 = (const float2*)b;
        *tw1 = (const float2*)(b + tw_bytes(key.L0));
        *tw2 = (const float2*)(b + tw_bytes(key.L0) + tw_bytes(key.L1));
        *kh = (const float*)(b + tb);
    };
    g_spec.clock++;
    for (auto& x : g_spec.e)
        if (x.buf && x.device == dev && x.key == key) {
            x.used = g_spec.clock;
            set_ptrs(x);
            return TIRA_OK;
        }
    // evicted spectra may still be read by work queued on the stream
    TH_HIP(hipStreamSynchronize(st));
    size_t total_b = 0, free_b = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) { (void)hipGetLastError(); total_b = 0; }
    const size_t need = tb + kbytes, budget = total_b / 8;
    while (!g_spec.e.empty() &&
           (g_spec.e.size() >= kSpecEntries || (spec_held(dev) > 0 && spec_held(dev) + need > budget))) {
        size_t i = spec_lru(dev);
        if (i == g_spec.e.size()) i = spec_lru(-1);
        spec_evict(i);
    }
    void* buf = nullptr;
    while (hipMalloc(&buf, need) != hipSuccess) {
        (void)hipGetLastError();
        buf = nullptr;
        if (g_spec.e.empty()) {
            set_error("spectral vote: out of device memory for the kernel spectra (%zu bytes)", need);
            return TIRA_ERR_OUT_OF_MEMORY;
        }
        spec_evict(spec_lru(-1));  // make room: the older spectra go first
    }
    g_spec.e.push_back(SpecEntry{});
    SpecEntry& ent = g_spec.e.back();
    ent.buf = buf;
    ent.bytes = need;
    ent.device = dev;
    ent.used = g_spec.clock;
    set_ptrs(ent);
    // an entry whose build fails is dropped again
    struct Drop {
        bool armed = true;
        ~Drop() {
            if (armed) spec_evict(g_spec.e.size() - 1);
        }
    } drop;

    // twiddles
    {
        std::vector<float2> t(tb / 8);
        twiddles(key.L0, t.data());
        twiddles(key.L1, t.data() + tw_bytes(key.L0) / 8);
        twiddles(key.L2, t.data() + (tw_bytes(key.L0) + tw_bytes(key.L1)) / 8);
        TH_HIP(hipMemcpy(ent.buf, t.data(), tb, hipMemcpyHostToDevice));
    }
    // offset kernels, parities and trig tables
    std::vector<double> K;
    int nf = 0;
    // the stick (ATV: e2) part only: the plate part, and the ATV's e0 and
    // constant parts, are linear maps of it and fold into the features
    TH_TRY(offset_kernels((unsigned)key.w, key.s1, key.s2, true, false, key.atv != 0, K, nf));
    if (nf > kMono) {  // ATV: keep the first 15 features of every component
        const size_t noff = K.size() / ((size_t)6 * nf);
        std::vector<double> K15((size_t)6 * kMono * noff);
        for (int c = 0; c < 6; c++)
            std::copy(K.begin() + (size_t)c * nf * noff, K.begin() + ((size_t)c * nf + kMono) * noff,
                      K15.begin() + (size_t)c * kMono * noff);
        K.swap(K15);
        nf = kMono;
    }
    const int w = 2 * (key.w / 2) + 1, r = key.w / 2;
    const int nch = 6 * nf;
    std::vector<unsigned char> par(nch);
    for (int c = 0; c < 6; c++)
        for (int f = 0; f < nf; f++) {
            const int m = f % kMono;
            par[c * nf + f] = (unsigned char)(((comp_nx(c) + mono_a(m)) & 1) | (((comp_ny(c) + mono_b(m)) & 1) << 1) |
                                              (((comp_nz(c) + mono_c(m)) & 1) << 2));
        }
    const int hs[3] = {g.h0, g.h1, g.h2}, Ls[3] = {g.L0, g.L1, g.L2};
    std::vector<double> trig[3];
    for (int a = 0; a < 3; a++) {
        trig[a].resize((size_t)2 * hs[a] * w);
        for (int k = 0; k < hs[a]; k++)
            for (int d = -r; d <= r; d++) {
                const double ang = 2.0 * M_PI * (double)k * (double)d / (double)Ls[a];
                trig[a][(size_t)k * w + d + r] = std::cos(ang);
                trig[a][((size_t)hs[a] + k) * w + d + r] = std::sin(ang);
            }
    }
    const size_t t1n = (size_t)nch * w * w * g.h2, t2n = (size_t)nch * w * g.h1 * g.h2;
    const size_t trig_n = trig[0].size() + trig[1].size() + trig[2].size();
    const size_t tmp_bytes = (K.size() + t1n + t2n + trig_n) * 8 + nch + 256;
    char* tmp = nullptr;
    if (hipMalloc((void**)&tmp, tmp_bytes) != hipSuccess) {
        (void)hipGetLastError();
        set_error("spectral vote: out of device memory building the kernel spectra (%zu bytes)", tmp_bytes);
        return TIRA_ERR_OUT_OF_MEMORY;
    }
    double* dK = (double*)tmp;
    double* dT1 = dK + K.size();
    double* dT2 = dT1 + t1n;
    double* dtr[3];
    dtr[0] = dT2 + t2n;
    dtr[1] = dtr[0] + trig[0].size();
    dtr[2] = dtr[1] + trig[1].size();
    unsigned char* dpar = (unsigned char*)(dtr[2] + trig[2].size());
    int rc = TIRA_OK;
    do {
        if (hipMemcpy(dK, K.data(), K.size() * 8, hipMemcpyHostToDevice) != hipSuccess) { rc = -2; break; }
        bool ok = true;
        for (int a = 0; a < 3; a++)
            ok = ok && hipMemcpy(dtr[a], trig[a].data(), trig[a].size() * 8, hipMemcpyHostToDevice) == hipSuccess;
        ok = ok && hipMemcpy(dpar, par.data(), nch, hipMemcpyHostToDevice) == hipSuccess;
        if (!ok) { rc = -2; break; }
        auto grid = [](size_t n) { return (unsigned)std::min<size_t>(ceil_div(n, 256), 65536); };
        k_khat_x<<<grid(t1n), 256, 0, st>>>(dK, dtr[2], dpar, dT1, w, g.h2, nch);
        k_khat_y<<<grid(t2n), 256, 0, st>>>(dT1, dtr[1], dpar, dT2, w, g.h1, g.h2, nch);
        const double scale = 0.5 / ((double)g.L0 * g.L1 * g.L2);
        k_khat_z<<<grid((size_t)nch * g.ocs), 256, 0, st>>>(dT2, dtr[0], dpar, (float*)*kh, w, g.h0, g.h1, g.h2,
                                                           nch, scale);
        if (hipGetLastError() != hipSuccess || hipStreamSynchronize(st) != hipSuccess) rc = -2;
    } while (false);
    (void)hipFree(tmp);
    if (rc != TIRA_OK) {
        set_error("spectral vote: building the kernel spectra failed");
        return TIRA_ERR_HIP;
    }
    g_spec.e.back().key = key;
    g_spec.builds++;
    drop.armed = false;
    return TIRA_OK;
}

static int launch_lines(float2* data, const LineGeom& lg, const FftPlan& pl, const float2* tw, bool inv, bool rows,
                        hipStream_t st) {
    const size_t nblk = (size_t)lg.nab * lg.nb * lg.nchan;
    if (nblk == 0) return TIRA_OK;
    TH_ARG(nblk < (1ull << 31), "spectral vote: volume too large");
    const unsigned grid = (unsigned)nblk;
    const size_t clds = ((size_t)lg.B * pl.L + pl.L) * sizeof(float2);  // k_fft_cols staging + twiddles
    TH_ARG(rows || (clds <= kColsLds && lg.Bh >= 1 && lg.B % lg.Bh == 0 && (size_t)lg.Bh * pl.L <= (size_t)kFftBL),
           "spectral vote: column batch exceeds the staging");
    if (inv) {
        if (rows) k_fft_lines<true, true><<<grid, kFT, 0, st>>>(data, lg, pl, tw);
        else if (lg.Bh >= 16) k_fft_cols<true, kFftBL, 6><<<grid, kFT, clds, st>>>(data, lg, pl, tw);
        else k_fft_cols<true, kFftBL, 1><<<grid, kFT, clds, st>>>(data, lg, pl, tw);
    } else {
        if (rows) k_fft_lines<false, true><<<grid, kFT, 0, st>>>(data, lg, pl, tw);
        else if (lg.Bh >= 16) k_fft_cols<false, kFftBL, 6><<<grid, kFT, clds, st>>>(data, lg, pl, tw);
        else k_fft_cols<false, kFftBL, 1><<<grid, kFT, clds, st>>>(data, lg, pl, tw);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// Lines of the axis with stride `estride` over a (n0 x n1 x n2)-indexed
// stack: the batch dimension `a` is x (stride 1) for the y and z axes.
static LineGeom strided_geom(size_t cs, int nchan, int L, size_t estride, int na, int nb, size_t bstride, int n_in,
                             int n_out) {
    LineGeom g{};
    g.cstride = cs;
    g.bstride = bstride;
    g.astride = 1;
    g.estride = estride;
    g.na = na;
    g.nb = nb;
    g.nchan = nchan;
    // B: up to 16 adjacent lines (128-byte rows of complex values) within
    // kColsLds of staging; Bh: the largest group of them the middle stages
    // take at once (Bh L <= kFftBL)
    int B = 1, lb = 0;
    while (2 * B <= 16 && ((size_t)2 * B * L + L) * sizeof(float2) <= kColsLds) { B *= 2; lb++; }
    int Bh = B;
    while (Bh > 1 && Bh * L > kFftBL) Bh /= 2;
    g.B = B;
    g.lb = lb;
    g.Bh = Bh;
    g.nab = (int)ceil_div((size_t)na, (size_t)g.B);
    g.n_in = n_in;
    g.n_out = n_out;
    return g;
}

// Whether TIRA_PLATE_AUTO takes the spectral vote: power 1, an exact plate
// (samples 0 or >= 3), FFT lengths within range, a window of at least 5 and
// a volume of at least 32^3 voxels (below that the gather costs microseconds
// and the first call's spectra build would dominate), and device memory for
// the spectra, their build and the feature volumes within 80% of the device.
// With `atv`: analytical tensor voting (31 features, 16 complex channels),
// whose features come from the input tensors in the x pass only.
bool spectral_preferred(unsigned s0, unsigned s1, unsigned s2, unsigned w, unsigned power, bool stick, bool plate,
                        unsigned samples, bool atv) {
    const unsigned p = power < 1 ? 1 : power;
    if (!atv && (p != 1 || (plate && samples > 0 && samples < 3))) return false;
    if (w < 5 || (size_t)s0 * s1 * s2 < 32768) return false;
    const unsigned L0 = spectral_length(s0, w), L1 = spectral_length(s1, w), L2 = spectral_length(s2, w);
    if (!L0 || !L1 || !L2 || 3 * L2 > (unsigned)kFftBL) return false;
    const size_t nf = kMono, nz = (nf + 1) / 2;  // the combined features
    const size_t W = 2 * (w / 2) + 1;
    const size_t h0 = L0 / 2 + 1, h1 = L1 / 2 + 1, h2 = L2 / 2 + 1;
    const size_t need = nz * (size_t)L0 * L1 * L2 * 8 + 6 * nf * h0 * h1 * h2 * 4 + 6 * nf * W * h1 * h2 * 8 +
                        (size_t)s0 * s1 * s2 * 32;
    size_t free_b = 0, total_b = 0;
    int dev = 0;
    if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || hipGetDevice(&dev) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    // memory this thread's spectral vote already holds is reused, not allocated again
    size_t held = ws_capacity(WS_SPEC0) + ws_capacity(WS_PAY0) + ws_capacity(WS_PAY1);
    held += spec_held(dev);
    return need <= total_b / 10 * 8 && need <= (free_b + held) / 10 * 9;
}

int vote3_spectral_run(const float* tensors, int in_layout, const float* lam, const float* ev2, const float* ev0,
                       float* out, int out_layout, unsigned s0, unsigned s1, unsigned s2, float sigma1,
                       float sigma2, unsigned w, unsigned power, bool stick, bool plate, unsigned samples,
                       bool atv, hipStream_t st, EigOut* eig) {
    const unsigned p = power < 1 ? 1 : power;
    if (atv) {  // sigma2, power, stick, plate and samples do not apply to ATV
        stick = plate = true;
        sigma2 = 0.0f;
        if (!tensors) return kSpectralNotApplicable;
    } else if (p != 1 || (plate && samples > 0 && samples < 3)) {
        return kSpectralNotApplicable;
    }
    const unsigned L0 = spectral_length(s0, w), L1 = spectral_length(s1, w), L2 = spectral_length(s2, w);
    if (!L0 || !L1 || !L2 || 3 * L2 > (unsigned)kFftBL) return kSpectralNotApplicable;
    FftPlan p0, p1, p2;  // z and y: column passes; x: lines in LDS
    make_cols_plan((int)L0, p0);
    make_cols_plan((int)L1, p1);
    make_lines_plan((int)L2, p2);

    SpecGeom g{};
    g.n0 = (int)s0; g.n1 = (int)s1; g.n2 = (int)s2;
    g.L0 = (int)L0; g.L1 = (int)L1; g.L2 = (int)L2;
    g.h0 = g.L0 / 2 + 1; g.h1 = g.L1 / 2 + 1; g.h2 = g.L2 / 2 + 1;
    g.nf = kMono;  // stick, plate and ATV parts combined into 15 quartic features
    g.nz = (g.nf + 1) / 2;
    g.stick = stick ? 1 : 0;
    g.plate = plate ? 1 : 0;
    g.atv = atv ? 1 : 0;
    g.alpha = atv ? 2.0f : sticknorm3_host(sigma1, sigma2, 1);
    g.P2 = ((int)L2 + 15) & ~15;  // rows start on 128-byte lines
    g.cs = (size_t)L0 * L1 * g.P2;
    g.ocs = (size_t)g.h0 * g.h1 * g.h2;

    const size_t nvox = (size_t)s0 * s1 * s2;
    // Channels per group: all of them when the lines of one row fit 64 KB of
    // LDS, so that each voxel's eigen-solve and features are computed once;
    // else the fewest balanced groups that fit. The staged lines are
    // transformed in groups of FG that fit every stage's register rounds (at
    // L = 540 two channel groups had solved every voxel twice).
    int C = g.nz;
    for (int ng = 1; ng <= g.nz; ng++) {
        C = (g.nz + ng - 1) / ng;
        if (((size_t)C * g.L2 + g.L2) * 8 <= 65536) break;
    }
    // Rows per block: among the counts that fit 64 KB of dynamic LDS (and,
    // where one row's lines fit the stages' rounds, whose lines all do), the
    // one whose voxels best fill whole passes of the block's threads in the
    // eigen-solve (ties: more rows). 128-wide rows of 8 channels take 4 rows
    // (512 voxels, 4320 values), not the 3 that fit 4096 values (384 voxels:
    // half the threads solving twice).
    const bool row_fits = lines_fit(p2, C);
    int B = 1;
    {
        double best = -1.0;
        for (int b = 1; b <= 64; b++) {
            if (((size_t)C * b * g.L2 + g.L2) * 8 > 65536 || (row_fits && !lines_fit(p2, C * b))) break;
            const int vox = b * g.n2, passes = (vox + kFT - 1) / kFT;
            const double fill = (double)vox / ((double)passes * kFT);
            if (fill >= best) { best = fill; B = b; }
            if (vox >= 4 * kFT) break;
        }
    }
    int FG = C * B;
    for (int ng = 1; !lines_fit(p2, FG); ng++) FG = (C * B + ng) / (ng + 1);
    const size_t xlds = ((size_t)C * B * g.L2 + g.L2) * 8;
    // tensor input: the x pass decomposes the voxels itself, in every channel
    // group (no payload round trip); eigenpair input goes through the payload
    const bool fused = tensors != nullptr;
    float4 *ps = nullptr, *pp = nullptr;
    float2* Z = nullptr;
    // Out of device memory for the spectral buffers: clear the error and let
    // the caller take the direct gather, which needs far less.
    auto alloc = [&](int rc) {
        if (rc == TIRA_ERR_OUT_OF_MEMORY) (void)hipGetLastError();
        return rc;
    };
    int rc = TIRA_OK;
    if (stick && !fused && (rc = alloc(ws_get(WS_PAY0, nvox * 16, (void**)&ps))) != TIRA_OK)
        return rc == TIRA_ERR_OUT_OF_MEMORY ? kSpectralNotApplicable : rc;
    if (plate && !fused && (rc = alloc(ws_get(WS_PAY1, nvox * 16, (void**)&pp))) != TIRA_OK)
        return rc == TIRA_ERR_OUT_OF_MEMORY ? kSpectralNotApplicable : rc;
    if ((rc = alloc(ws_get(WS_SPEC0, (size_t)g.nz * g.cs * 8, (void**)&Z))) != TIRA_OK)
        return rc == TIRA_ERR_OUT_OF_MEMORY ? kSpectralNotApplicable : rc;
    // the spectra are those of the stick part whatever the stick / plate selection
    const SpecKey key{g.L0, g.L1, g.L2, (int)w, sigma1, sigma2, 1, 1, atv ? 1 : 0};
    const float2 *tw0, *tw1, *tw2;
    const float* kh;
    if ((rc = alloc(spectra_for(key, g, st, &tw0, &tw1, &tw2, &kh))) != TIRA_OK)
        return rc == TIRA_ERR_OUT_OF_MEMORY ? kSpectralNotApplicable : rc;

    if (!fused) {
        ProfScope prof("vote3_prep", st);
        TH_TRY(vote3_payload_flat(tensors, in_layout, lam, ev2, ev0, s0, s1, s2, ps, pp, st));
    }
    const size_t row = (size_t)g.P2, plane = (size_t)L1 * row;
    {
        ProfScope prof("spec_fwd_x", st);
        const size_t nrows = (size_t)s0 * s1;
        const size_t nblk = ceil_div(nrows, (size_t)B);
        TH_ARG(nblk < (1ull << 31), "spectral vote: volume too large");
        if (fused)
            k_spec_xfwd<true><<<(unsigned)nblk, kFT, xlds, st>>>(ps, pp, tensors, in_layout, g, B, C, FG, p2, tw2, Z);
        else
            k_spec_xfwd<false><<<(unsigned)nblk, kFT, xlds, st>>>(ps, pp, tensors, in_layout, g, B, C, FG, p2, tw2, Z);
        TH_HIP(hipGetLastError());
    }
    {
        ProfScope prof("spec_fwd_y", st);
        TH_TRY(launch_lines(Z, strided_geom(g.cs, g.nz, g.L1, row, g.L2, g.n0, plane, g.n1, g.L1), p1, tw1, false,
                            false, st));
    }
    {
        ProfScope prof("spec_fwd_z", st);
        TH_TRY(launch_lines(Z, strided_geom(g.cs, g.nz, g.L0, plane, g.L2, g.L1, row, g.n0, g.L0), p0, tw0, false,
                            false, st));
    }
    {
        ProfScope prof("spec_mul", st);
        k_spec_mul<<<(unsigned)ceil_div(g.ocs, 256), 256, 0, st>>>(Z, kh, g);
        TH_HIP(hipGetLastError());
    }
    {
        ProfScope prof("spec_inv_z", st);
        TH_TRY(launch_lines(Z, strided_geom(g.cs, 3, g.L0, plane, g.L2, g.L1, row, g.L0, g.n0), p0, tw0, true, false,
                            st));
    }
    {
        ProfScope prof("spec_inv_y", st);
        TH_TRY(launch_lines(Z, strided_geom(g.cs, 3, g.L1, row, g.L2, g.n0, plane, g.L1, g.n1), p1, tw1, true, false,
                            st));
    }
    {
        ProfScope prof("spec_inv_x", st);
        // Rows per block: up to the staging's 4096 values, the count whose grid
        // best fills whole rounds of the resident blocks (four per CU). 128^2
        // rows of L = 135 take 8 rows (2048 blocks, two full rounds), not 10
        // (1639 blocks: a second round 60% full).
        const size_t nrows = (size_t)s0 * s1;
        const int Bmax = std::max(1, kFftBL / (3 * g.L2));
        const double slots = 4.0 * cu_count();
        int B = Bmax;
        {
            double best = -1.0;
            for (int b = Bmax; b >= std::max(1, Bmax / 2); b--) {
                const double rounds = (double)ceil_div(nrows, (size_t)b) / slots;
                const double fill = rounds / std::ceil(rounds);
                if (fill > best + 0.02) { best = fill; B = b; }
            }
        }
        k_spec_xinv<<<ceil_div(nrows, (size_t)B), kFT, 0, st>>>(Z, g, B, p2, tw2, out, out_layout,
                                                                eig ? eig->evals : nullptr, eig ? eig->flags : 0u,
                                                                eig ? eig->n_refined : nullptr);
        TH_HIP(hipGetLastError());
        if (eig) eig->done = true;
    }
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_spectral_offset_kernels(unsigned w, float sigma1, float sigma2, unsigned power, int stick, int plate,
                                 double* out, unsigned* nf_out) {
    TH_ARG(out, "NULL pointer");
    TH_ARG(power == 1, "the spectral vote serves power 1 only");
    TH_ARG(stick || plate, "neither stick nor plate");
    TH_ARG(sigma1 >= 0.0f && sigma2 >= 0.0f, "sigmas must be >= 0");
    std::vector<double> K;
    int nf = 0;
    TH_TRY(offset_kernels(w, sigma1, sigma2, stick != 0, plate != 0, false, K, nf));
    std::copy(K.begin(), K.end(), out);
    if (nf_out) *nf_out = (unsigned)nf;
    return TIRA_OK;
}

int tira_spectral_offset_kernels_atv(unsigned w, float sigma, double* out, unsigned* nf_out) {
    TH_ARG(out, "NULL pointer");
    TH_ARG(sigma >= 0.0f, "sigma must be >= 0");
    std::vector<double> K;
    int nf = 0;
    TH_TRY(offset_kernels(w, sigma, 0.0f, true, true, true, K, nf));
    std::copy(K.begin(), K.end(), out);
    if (nf_out) *nf_out = (unsigned)nf;
    return TIRA_OK;
}

unsigned tira_spectral_length(unsigned n, unsigned w) { return spectral_length(n, w); }

int tira_spectral_cache_info(unsigned* entries, unsigned long long* bytes, unsigned long long* builds) {
    unsigned long long b = 0;
    for (const auto& x : g_spec.e) b += x.bytes;
    if (entries) *entries = (unsigned)g_spec.e.size();
    if (bytes) *bytes = b;
    if (builds) *builds = g_spec.builds;
    return TIRA_OK;
}

int tira_fft3_async(float* data, unsigned nchan, unsigned L0, unsigned L1, unsigned L2, int inverse) {
    TH_ARG(data, "NULL pointer");
    FftPlan p0, p1, p2;
    TH_ARG(L0 >= 1 && L1 >= 1 && L2 >= 1 && L0 <= (unsigned)kMaxL && L1 <= (unsigned)kMaxL &&
               L2 <= (unsigned)kMaxL,
           "FFT lengths must be in [1, %d]", kMaxL);
    TH_ARG(make_cols_plan((int)L0, p0) && make_cols_plan((int)L1, p1) && make_lines_plan((int)L2, p2),
           "FFT lengths must have prime factors 2, 3, 5, 7 only");
    TH_ARG(is_device(data), "tira_fft3_async takes a device pointer");
    if (!nchan) return TIRA_OK;
    hipStream_t st = stream();
    float2* d = (float2*)data;
    const size_t cs = (size_t)L0 * L1 * L2, plane = (size_t)L1 * L2;
    const bool inv = inverse != 0;
    // Twiddle tables of the three lengths in a workspace slot of this entry.
    const size_t tb = tw_bytes((int)L0) + tw_bytes((int)L1) + tw_bytes((int)L2);
    float2* tw = nullptr;
    TH_TRY(ws_get(WS_SPEC1, tb, (void**)&tw));
    std::vector<float2> t(tb / 8);
    twiddles((int)L0, t.data());
    twiddles((int)L1, t.data() + tw_bytes((int)L0) / 8);
    twiddles((int)L2, t.data() + (tw_bytes((int)L0) + tw_bytes((int)L1)) / 8);
    TH_HIP(hipStreamSynchronize(st));  // earlier calls may still read the previous tables
    TH_HIP(hipMemcpy(tw, t.data(), tb, hipMemcpyHostToDevice));
    const float2* t0 = tw;
    const float2* t1 = tw + tw_bytes((int)L0) / 8;
    const float2* t2 = tw + (tw_bytes((int)L0) + tw_bytes((int)L1)) / 8;
    // x: rows of L2 contiguous elements; the batch dimension is the row (stride L2)
    LineGeom gx{};
    gx.cstride = cs; gx.bstride = plane; gx.astride = L2; gx.estride = 1;
    gx.na = (int)L1; gx.nb = (int)L0; gx.nchan = (int)nchan;
    gx.B = std::max(1, std::min((int)L1, kFftBL / (int)L2));
    gx.nab = (int)ceil_div(L1, (size_t)gx.B);
    gx.n_in = (int)L2; gx.n_out = (int)L2;
    TH_TRY(launch_lines(d, gx, p2, t2, inv, true, st));
    // y: columns x of every plane z; z: columns x of every row y
    TH_TRY(launch_lines(d, strided_geom(cs, (int)nchan, (int)L1, L2, (int)L2, (int)L0, plane, (int)L1, (int)L1), p1,
                        t1, inv, false, st));
    TH_TRY(launch_lines(d, strided_geom(cs, (int)nchan, (int)L0, plane, (int)L2, (int)L1, L2, (int)L0, (int)L0), p0,
                        t0, inv, false, st));
    return TIRA_OK;
}

}  // extern "C"
