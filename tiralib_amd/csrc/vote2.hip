// vote2.hip -- 2D tensor-kernel (TK) and analytical (ATV) voting on gfx950,
// replacing tira::cuda::tensorvote / tensorvote_prebuilt / tensorvote_tk /
// tensorvote_atv (tensorvote.h:1993-2305) and the host table builder
// cpu::build_neighbors2d (tensorvote.h:1223-1280).
//
// Design: one prep kernel turns each pixel into a voter payload
//   float4(q.x, q.y, eta_s * stick scale, plate scale lambda0)
// (q = eigenvector of the larger eigenvalue; the per-pair cos/sin of the
// reference, tensorvote.h:967-969, become one eigenvector per voter), then a
// gather kernel walks the 48-byte Neighbor2D table with wave-uniform scalar
// loads. Lanes map to the fast image axis (x1), so voter loads coalesce (the
// reference maps threadIdx.x to the slow axis, tensorvote.h:2075).
#include "common.hpp"
#include "eig_device.hpp"

#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

namespace th {

struct Nb2 {  // tira::Neighbor2D, 48 bytes (tensorvote.h:776-786)
    int du, dv;
    float dx, dy, c1, c2, Ra, Rb, Rc, Va, Vb, Vc;
};
static_assert(sizeof(Nb2) == 48, "Neighbor2D must stay 48 bytes");

// ---------------------------------------------------------------- neighbour table
// Same arithmetic as the reference (tensorvote.h:34-40, 803-914, 1047-1071):
// the unqualified cos/sin/exp of the reference's float code resolve to the
// double overloads under the host compiler, so they are evaluated in double
// and rounded here too. The functions serve the host build and the device
// build (k_build_nb2) alike, with contraction off so that both round every
// operation as the reference's host code does; the only other difference is
// the float exponential of c1, c2 (glibc's expf on the host, the double
// exponential rounded to float on the device: equal wherever expf rounds
// correctly).
__host__ __device__ inline float exp_f32(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return (float)exp((double)x);
#else
    return std::exp(x);
#endif
}

__host__ __device__ inline float atten(float dist, float sigma) {
#pragma clang fp contract(off)
    if (sigma == 0.0f) return (dist > 1e-12f) ? 0.0f : 1.0f;
    return (float)exp((double)(-(dist * dist) / (sigma * sigma)));
}

__host__ __device__ float stick_norm_2d(float s1, float s2, unsigned power) {
#pragma clang fp contract(off)
    if (power < 1) power = 1;
    float ratio = 1.0f;
    for (unsigned k = 1; k <= power; ++k) ratio *= (float)(2 * k) / (float)(2 * k - 1);
    return ratio / ((float)M_PI * (s1 * s1 + s2 * s2));
}

__host__ __device__ inline void stickvote_2d(float rx, float ry, float s1, float s2, unsigned power, float theta,
                                             float& Va, float& Vb, float& Vc) {
#pragma clang fp contract(off)
    const float qx = (float)cos((double)theta), qy = (float)sin((double)theta);
    const float dist = sqrtf(rx * rx + ry * ry);
    float dx = 0.0f, dy = 0.0f;
    if (dist != 0.0f) { dx = rx / dist; dy = ry / dist; }
    const float qTd = qx * dx + qy * dy;
    unsigned p = power < 1 ? 1 : power;
    const float c1 = atten(dist, s1), c2 = atten(dist, s2);
    const float q2 = qTd * qTd, t1 = 1.0f - q2, t2 = q2;
    float D;
    if (p == 1) D = c1 * t1 + c2 * t2;
    else {
        float e1 = t1, e2 = t2;
        for (unsigned i = 1; i < p; ++i) { e1 *= t1; e2 *= t2; }
        D = c1 * e1 + c2 * e2;
    }
    const float factor = D * stick_norm_2d(s1, s2, power);
    const float Ra = 1 - 2 * dx * dx, Rb = -2 * dx * dy, Rc = 1 - 2 * dy * dy;
    const float Rqx = Ra * qx + Rb * qy, Rqy = Rb * qx + Rc * qy;
    Va = Rqx * Rqx * factor;
    Vb = Rqx * Rqy * factor;
    Vc = Rqy * Rqy * factor;
}

__host__ __device__ inline void plate_numerical_2d(float rx, float ry, float s1, float s2, unsigned power, float& a,
                                                   float& b, float& c, unsigned n) {
#pragma clang fp contract(off)
    a = b = c = 0.0f;
    if (n == 0) return;
    const float dth = (float)(M_PI / (double)(float)n);
    for (unsigned i = 0; i < n; i++) {
        float va, vb, vc;
        stickvote_2d(rx, ry, s1, s2, power, dth * (float)i, va, vb, vc);
        a += va; b += vb; c += vc;
    }
    const float wb = (float)M_PI / (float)n;
    a *= wb; b *= wb; c *= wb;
}

// Entry k of the w x w table (dv outer, du inner; tensorvote.h:1223-1280).
__host__ __device__ inline Nb2 neighbor2d(int k, int w, float s1, float s2, unsigned power, unsigned samples,
                                          bool stick, bool plate) {
#pragma clang fp contract(off)
    const int hw = w / 2;
    const float is1 = (s1 > 0.0f) ? 1.0f / (s1 * s1) : 0.0f;
    const float is2 = (s2 > 0.0f) ? 1.0f / (s2 * s2) : 0.0f;
    Nb2 n{};
    const int dv = k / w - hw, du = k % w - hw;
    n.du = du; n.dv = dv;
    const float l2 = (float)(du * du + dv * dv);
    if (stick) {
        if (l2 != 0.0f) {
            const float il = 1.0f / sqrtf(l2);
            n.dx = (float)du * il;
            n.dy = (float)dv * il;
        }
        n.c1 = (s1 == 0.0f) ? (l2 > 1e-12f ? 0.0f : 1.0f) : exp_f32(-l2 * is1);
        n.c2 = (s2 == 0.0f) ? (l2 > 1e-12f ? 0.0f : 1.0f) : exp_f32(-l2 * is2);
        n.Ra = 1.0f - 2.0f * n.dx * n.dx;
        n.Rb = 0.0f - 2.0f * n.dx * n.dy;
        n.Rc = 1.0f - 2.0f * n.dy * n.dy;
    }
    if (plate && samples > 0) plate_numerical_2d((float)du, (float)dv, s1, s2, power, n.Va, n.Vb, n.Vc, samples);
    return n;
}

int build_neighbors2d(int w, float s1, float s2, unsigned power, unsigned samples, bool stick, bool plate,
                      std::vector<Nb2>& nb) {
    nb.clear();
    if (w <= 0) return 0;
    nb.reserve((size_t)w * w);
    for (int k = 0; k < w * w; k++) nb.push_back(neighbor2d(k, w, s1, s2, power, samples, stick, plate));
    return (int)nb.size();
}

__global__ __launch_bounds__(256) void k_build_nb2(int w, float s1, float s2, unsigned power, unsigned samples,
                                                   int stick, int plate, Nb2* __restrict__ out) {
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= w * w) return;
    out[k] = neighbor2d(k, w, s1, s2, power, samples, stick != 0, plate != 0);
}

// Device copy of a host table in workspace `slot`, kept while its bytes and
// the buffer (same device, address and workspace generation) are unchanged:
// repeated calls with the same parameters skip the stream wait and the copy.
struct TableCache {
    int dev = -1;
    const void* buf = nullptr;
    unsigned long long gen = 0;
    std::vector<unsigned char> bytes;
};
static int table_to_device(int slot, const void* host, size_t nbytes, TableCache& c, hipStream_t st, void** out) {
    void* d = nullptr;
    TH_TRY(ws_get(slot, nbytes ? nbytes : 1, &d));
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    const bool same = nbytes > 0 && c.dev == dev && c.buf == d && c.gen == ws_generation() &&
                      c.bytes.size() == nbytes && std::memcmp(c.bytes.data(), host, nbytes) == 0;
    if (nbytes > 0 && !same) {
        // pageable source; the buffer may still be read by earlier work on the stream
        TH_HIP(hipStreamSynchronize(st));
        c.dev = -1;
        TH_HIP(hipMemcpy(d, host, nbytes, hipMemcpyHostToDevice));
        c.bytes.assign((const unsigned char*)host, (const unsigned char*)host + nbytes);
        c.buf = d;
        c.gen = ws_generation();
        c.dev = dev;
    }
    *out = d;
    return TIRA_OK;
}

// ---------------------------------------------------------------- kernels
// Voter payload from the tensor field [a b b c]: eigenvector of the larger
// eigenvalue, eta_s-scaled stick scale (|l1| - |l0|, negated when l1 < 0,
// tensorvote.h:995-998) and the signed smaller eigenvalue for the plate
// (tensorvote.h:1107-1113).
__global__ __launch_bounds__(256) void k_vote2_prep(const float4* __restrict__ t, size_t n, float eta_s,
                                                    float4* __restrict__ pay) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 m = t[i];
    float l0, l1, vx, vy, ux, uy;
    eig2_vals(m.x, m.y, m.w, l0, l1);
    eig2_vecs(m.x, m.y, m.w, l0, vx, vy, ux, uy);
    float scale = fabsf(l1) - fabsf(l0);
    if (l1 < 0.0f) scale = -scale;
    pay[i] = make_float4(ux, uy, scale * eta_s, l0);
}

// Payload from caller eigenpairs (tensorvote_tk inputs): lambdas n x 2,
// thetas n x 2 (the stick direction is theta1, tensorvote.h:966).
__global__ __launch_bounds__(256) void k_vote2_prep_tk(const float2* __restrict__ lam,
                                                       const float2* __restrict__ th, size_t n, float eta_s,
                                                       float4* __restrict__ pay) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 l = lam[i];
    float s, c;
    sincosf(th[i].y, &s, &c);
    float scale = fabsf(l.y) - fabsf(l.x);
    if (l.y < 0.0f) scale = -scale;
    pay[i] = make_float4(c, s, scale * eta_s, l.x);
}

// Table-driven gather: receiver (x0, x1) sums over the table entries
// (tk_stick_window 938-1005 + tk_plate_window 1100-1115). A small field does
// not fill the chip with one thread per receiver (256 x 256 pixels are 1024
// waves, one per SIMD), so the table is split over blockIdx.z into chunks of
// `per` entries: chunk z writes its partial sums to part[z * n + pixel] and
// k_vote2_reduce adds the chunks in order (deterministic). With part == NULL
// (one chunk) the sums go straight to out.
// The entries of one receiver, branch-free: a voter outside the field reads
// voxel 0 with zero weights (adds exactly nothing), so the loop unrolls and
// the table and voter reads of several entries are in flight together (a
// loop that skipped them waited on each entry's reads in turn: 1.39 -> 1.27
// ms per search-loop evaluation). Staging the table in LDS per block as well
// measured 11% slower there (68 VGPRs and a barrier per 256 entries).
template <bool STICK, bool PLATE, bool P1>
__device__ __forceinline__ void vote2_entries(const float4* __restrict__ pay, const Nb2* __restrict__ nb, int k0,
                                              int k1, int x0, int x1, int s0, int s1, int power, float& a, float& b,
                                              float& c) {
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const Nb2 e = nb[k];
        const int r0 = x0 + e.dv, r1 = x1 + e.du;
        const bool inb = (unsigned)r0 < (unsigned)s0 && (unsigned)r1 < (unsigned)s1;
        const float4 P = pay[inb ? (size_t)r0 * s1 + r1 : 0];
        if (STICK) {
            const float qTd = fmaf(P.x, e.dx, P.y * e.dy);
            const float q2 = qTd * qTd;
            float D;
            if (P1) D = fmaf(e.c2 - e.c1, q2, e.c1);
            else {
                const float t1 = 1.0f - q2;
                float e1 = t1, e2 = q2;
                for (int i = 1; i < power; i++) { e1 *= t1; e2 *= q2; }
                D = fmaf(e.c1, e1, e.c2 * e2);
            }
            const float Rqx = fmaf(e.Ra, P.x, e.Rb * P.y), Rqy = fmaf(e.Rb, P.x, e.Rc * P.y);
            const float f = D * (inb ? P.z : 0.0f);
            const float fx = f * Rqx;
            a = fmaf(fx, Rqx, a); b = fmaf(fx, Rqy, b); c = fmaf(f * Rqy, Rqy, c);
        }
        if (PLATE) {
            const float wp = inb ? P.w : 0.0f;
            a = fmaf(wp, e.Va, a); b = fmaf(wp, e.Vb, b); c = fmaf(wp, e.Vc, c);
        }
    }
}

__global__ __launch_bounds__(256) void k_vote2(const float4* __restrict__ pay, const Nb2* __restrict__ nb, int nbc,
                                               int per, int s0, int s1, int power, int stick, int plate,
                                               float* __restrict__ out, float4* __restrict__ part) {
    const int x1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k0 = (int)blockIdx.z * per, k1 = min(nbc, k0 + per);
    float a = 0.0f, b = 0.0f, c = 0.0f;
    const bool p1 = power == 1;
    if (stick && plate) {
        if (p1) vote2_entries<true, true, true>(pay, nb, k0, k1, x0, x1, s0, s1, power, a, b, c);
        else vote2_entries<true, true, false>(pay, nb, k0, k1, x0, x1, s0, s1, power, a, b, c);
    } else if (stick) {
        if (p1) vote2_entries<true, false, true>(pay, nb, k0, k1, x0, x1, s0, s1, power, a, b, c);
        else vote2_entries<true, false, false>(pay, nb, k0, k1, x0, x1, s0, s1, power, a, b, c);
    } else if (plate) {
        vote2_entries<false, true, true>(pay, nb, k0, k1, x0, x1, s0, s1, power, a, b, c);
    }
    if (x0 < s0 && x1 < s1) {
        const size_t p = (size_t)x0 * s1 + x1;
        if (part) part[(size_t)blockIdx.z * ((size_t)s0 * s1) + p] = make_float4(a, b, c, 0.0f);
        else reinterpret_cast<float4*>(out)[p] = make_float4(a, b, b, c);
    }
}

__global__ __launch_bounds__(256) void k_vote2_reduce(const float4* __restrict__ part, int chunks, size_t n,
                                                      float* __restrict__ out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float4 v = part[p];
    for (int z = 1; z < chunks; z++) {
        const float4 q = part[(size_t)z * n + p];
        v.x += q.x; v.y += q.y; v.z += q.z;
    }
    reinterpret_cast<float4*>(out)[p] = make_float4(v.x, v.y, v.y, v.z);
}

// Table chunks for an s0 x s1 field: enough blocks for about 8 waves per
// SIMD (2048 blocks of 4 waves), at least 64 entries per chunk, at most 32.
constexpr int kVote2TargetBlocks = 2048;
static int vote2_chunks(unsigned s0, unsigned s1, int nbc) {
    const size_t blocks = (size_t)ceil_div(s1, 64) * ceil_div(s0, 4);
    if (nbc < 128 || blocks >= (size_t)kVote2TargetBlocks) return 1;
    const int want = (int)ceil_div(kVote2TargetBlocks, blocks);
    return std::max(1, std::min(std::min(want, nbc / 64), 32));
}
static int vote2_per(int nbc, int chunks) { return chunks > 1 ? (nbc + chunks - 1) / chunks : nbc; }

// ---------------------------------------------------------------- analytical tensor voting (ATV)
// tira::atv 2D (tensorvote.h:159-253): c R H R with R = I - 2 v v^T and
//   H = (l1 - l0)(T1 - (v^T T1 v T1 + 2 T1 v v^T T1) / 3) + l0 (T0 - (v^T T0 v T0 + 2 T0 v v^T T0) / 4),
//   T1 = e1 e1^T, T0 = e0 e0^T + e1 e1^T = I.
// With a = e1.v (rho = |v|^2: 1, or 0 at the voter itself) this collapses to
//   K11 e1 e1^T + K1 (v e1^T + e1 v^T) + Kvv v v^T + KI I,
//   K11 = c s1 (1 - a^2), K1 = -2 a K11, Kvv = -c l0 / 2 + 4 a^2 K11, KI = c l0 (1 - rho/4).
struct Atv2 {  // one window offset, 48 bytes
    int du, dv;
    float vx, vy, c, nc, ck4, nhalfc, vxx, vyy, vxy2, pad;
};
static_assert(sizeof(Atv2) == 48, "ATV offset entry must stay 48 bytes");

// Window of the reference's ATV loops (tensorvote.h:268-279): hw = w / 2,
// v (axis 0) outer, u (axis 1) inner.
static void build_atv2_table(unsigned w, float sigma, std::vector<Atv2>& t) {
    const int hw = (int)w / 2;
    t.clear();
    for (int dv = -hw; dv <= hw; dv++)
        for (int du = -hw; du <= hw; du++) {
            const double l2 = (double)du * du + (double)dv * dv, L = std::sqrt(l2);
            double vx = 0.0, vy = 0.0;
            if (l2 > 0.0) { vx = du / L; vy = dv / L; }
            const double c = sigma == 0.0f ? (L > 1e-12 ? 0.0 : 1.0) : std::exp(-l2 / ((double)sigma * sigma));
            const double k4 = l2 > 0.0 ? 0.75 : 1.0;
            Atv2 e;
            e.du = du; e.dv = dv;
            e.vx = (float)vx; e.vy = (float)vy; e.c = (float)c; e.nc = (float)-c;
            e.ck4 = (float)(c * k4); e.nhalfc = (float)(-0.5 * c);
            e.vxx = (float)(vx * vx); e.vyy = (float)(vy * vy); e.vxy2 = (float)(2.0 * vx * vy); e.pad = 0.0f;
            t.push_back(e);
        }
}

// ATV payload from the field: (e1, l1 - l0, l0), e1 the eigenvector of the
// larger eigenvalue (evals2 / evecs2polar, tensorvote.h:1388-1393).
__global__ __launch_bounds__(256) void k_vote2_atv_prep(const float4* __restrict__ t, size_t n,
                                                        float4* __restrict__ pay) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float4 m = t[i];
    float l0, l1, vx, vy, ux, uy;
    eig2_vals(m.x, m.y, m.w, l0, l1);
    eig2_vecs(m.x, m.y, m.w, l0, vx, vy, ux, uy);
    pay[i] = make_float4(ux, uy, l1 - l0, l0);
}

// ATV payload from caller eigenpairs (lambdas, thetas n x 2; e1 = angle
// theta1, with theta0 perpendicular as evecs2polar_symmetric returns it).
__global__ __launch_bounds__(256) void k_vote2_atv_prep_eig(const float2* __restrict__ lam,
                                                            const float2* __restrict__ th, size_t n,
                                                            float4* __restrict__ pay) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float2 l = lam[i];
    float sn, cs;
    sincosf(th[i].y, &sn, &cs);
    pay[i] = make_float4(cs, sn, l.y - l.x, l.x);
}

// The table is split over blockIdx.z as in k_vote2: chunk z covers entries
// [z per, (z + 1) per) and, with several chunks, writes its partial sums
// (xx, 2 xy, yy, isotropic) to part[z * n + pixel] for k_vote2_atv_reduce.
__global__ __launch_bounds__(256) void k_vote2_atv(const float4* __restrict__ pay, const Atv2* __restrict__ tab,
                                                   int nt, int per, int s0, int s1, float* __restrict__ out,
                                                   float4* __restrict__ part) {
    const int x1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int k0 = (int)blockIdx.z * per, k1 = min(nt, k0 + per);
    float xx = 0.0f, xy2 = 0.0f, yy = 0.0f, ii = 0.0f;
    // branch-free as vote2_entries: a voter outside the field weighs zero
#pragma unroll 4
    for (int k = k0; k < k1; k++) {
        const Atv2 e = tab[k];
        const int r0 = x0 + e.dv, r1 = x1 + e.du;
        const bool inb = (unsigned)r0 < (unsigned)s0 && (unsigned)r1 < (unsigned)s1;
        float4 P = pay[inb ? (size_t)r0 * s1 + r1 : 0];
        P.z = inb ? P.z : 0.0f;
        P.w = inb ? P.w : 0.0f;
        const float a = fmaf(P.x, e.vx, P.y * e.vy);
        const float al = a * a;
        const float K11 = P.z * fmaf(e.nc, al, e.c);
        const float Kvv = fmaf(4.0f * al, K11, P.w * e.nhalfc);
        const float K1d = (-4.0f * a) * K11;  // 2 K1
        const float hx = fmaf(K11, P.x, K1d * e.vx), hy = fmaf(K11, P.y, K1d * e.vy);
        xx = fmaf(P.x, hx, fmaf(Kvv, e.vxx, xx));
        yy = fmaf(P.y, hy, fmaf(Kvv, e.vyy, yy));
        xy2 = fmaf(P.x, hy, fmaf(hx, P.y, fmaf(Kvv, e.vxy2, xy2)));
        ii = fmaf(P.w, e.ck4, ii);
    }
    if (x0 < s0 && x1 < s1) {
        const size_t p = (size_t)x0 * s1 + x1;
        if (part) {
            part[(size_t)blockIdx.z * ((size_t)s0 * s1) + p] = make_float4(xx, xy2, yy, ii);
        } else {
            const float b = 0.5f * xy2;
            reinterpret_cast<float4*>(out)[p] = make_float4(xx + ii, b, b, yy + ii);
        }
    }
}

__global__ __launch_bounds__(256) void k_vote2_atv_reduce(const float4* __restrict__ part, int chunks, size_t n,
                                                          float* __restrict__ out) {
    const size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    float4 v = part[p];
    for (int z = 1; z < chunks; z++) {
        const float4 q = part[(size_t)z * n + p];
        v.x += q.x; v.y += q.y; v.z += q.z; v.w += q.w;
    }
    const float b = 0.5f * v.y;
    reinterpret_cast<float4*>(out)[p] = make_float4(v.x + v.w, b, b, v.z + v.w);
}

static int vote2_atv_run(const float* tin, const float* lam, const float* th, float* out, float sigma, unsigned w,
                         unsigned s0, unsigned s1, hipStream_t st) {
    const size_t n = (size_t)s0 * s1;
    if (!n) return TIRA_OK;
    float4* pay = nullptr;
    TH_TRY(ws_get(WS_PAY0, n * 16, (void**)&pay));
    {
        ProfScope prof("vote2_atv_prep", st);
        if (tin) k_vote2_atv_prep<<<ceil_div(n, 256), 256, 0, st>>>((const float4*)tin, n, pay);
        else k_vote2_atv_prep_eig<<<ceil_div(n, 256), 256, 0, st>>>((const float2*)lam, (const float2*)th, n, pay);
    }
    TH_HIP(hipGetLastError());
    std::vector<Atv2> tab;
    build_atv2_table(w, sigma, tab);
    Atv2* dtab = nullptr;
    static thread_local TableCache atv_cache;
    TH_TRY(table_to_device(WS_TABLE2, tab.data(), tab.size() * sizeof(Atv2), atv_cache, st, (void**)&dtab));
    const int nt = (int)tab.size(), chunks = vote2_chunks(s0, s1, nt);
    float4* part = nullptr;
    if (chunks > 1) TH_TRY(ws_get(WS_PAY1, (size_t)chunks * n * 16, (void**)&part));
    dim3 grid(ceil_div(s1, 64), ceil_div(s0, 4), chunks);
    {
        ProfScope prof("vote2_atv", st);
        k_vote2_atv<<<grid, 256, 0, st>>>(pay, dtab, nt, vote2_per(nt, chunks), (int)s0, (int)s1, out, part);
        if (chunks > 1) k_vote2_atv_reduce<<<ceil_div(n, 256), 256, 0, st>>>(part, chunks, n, out);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

static int vote2_atv_entry(float* t_out, const float* t_in, const float* lam, const float* th, float sigma,
                           unsigned w, unsigned s0, unsigned s1) {
    TH_ARG(t_out && (t_in || (lam && th)), "NULL pointer");
    TH_ARG(sigma >= 0.0f, "sigma must be >= 0");
    TH_ARG(w <= 1025, "window %u exceeds the supported maximum (1025)", w);
    const size_t n = (size_t)s0 * s1;
    if (!n) return TIRA_OK;
    const float *din = nullptr, *dl = nullptr, *dt = nullptr;
    void* dout;
    if (t_in) TH_TRY(stage_in(t_in, n * 16, WS_IN0, (const void**)&din));
    else {
        TH_TRY(stage_in(lam, n * 8, WS_IN0, (const void**)&dl));
        TH_TRY(stage_in(th, n * 8, WS_IN1, (const void**)&dt));
    }
    TH_TRY(stage_out(t_out, n * 16, WS_OUT0, &dout));
    TH_TRY(vote2_atv_run(din, dl, dt, (float*)dout, sigma, w, s0, s1, stream()));
    return finish_out(t_out, dout, n * 16);
}

// ---------------------------------------------------------------- host driver
// The payload pass and the table-driven gather on device buffers: `pay` holds
// n float4, `dnb` the nbc-entry neighbour table. plate: the reference's kernel
// adds the plate term when plate && samples > 0 (tensorvote.h:2093); the
// table's V entries are zero otherwise.
static int vote2_launch(const float* tin, const float* lam, const float* th, float* out, float4* pay, const Nb2* dnb,
                        int nbc, float s1, float s2, int power, unsigned s0, unsigned s1n, bool stick, bool plate,
                        unsigned samples, float4* part, hipStream_t st) {
    const size_t n = (size_t)s0 * s1n;
    const unsigned p = power < 1 ? 1 : (unsigned)power;
    const float eta_s = stick_norm_2d(s1, s2, p);
    {
        ProfScope prof("vote2_prep", st);
        if (tin) k_vote2_prep<<<ceil_div(n, 256), 256, 0, st>>>((const float4*)tin, n, eta_s, pay);
        else k_vote2_prep_tk<<<ceil_div(n, 256), 256, 0, st>>>((const float2*)lam, (const float2*)th, n, eta_s, pay);
    }
    TH_HIP(hipGetLastError());
    const int chunks = vote2_chunks(s0, s1n, nbc);
    dim3 grid(ceil_div(s1n, 64), ceil_div(s0, 4), chunks);
    {
        ProfScope prof("vote2", st);
        k_vote2<<<grid, 256, 0, st>>>(pay, dnb, nbc, vote2_per(nbc, chunks), (int)s0, (int)s1n, (int)p,
                                      stick ? 1 : 0, (plate && samples > 0) ? 1 : 0, out,
                                      chunks > 1 ? part : nullptr);
        if (chunks > 1) k_vote2_reduce<<<ceil_div(n, 256), 256, 0, st>>>(part, chunks, n, out);
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// Bytes of partial sums vote2_launch needs for this field and table.
static size_t vote2_part_bytes(unsigned s0, unsigned s1, int nbc) {
    const int chunks = vote2_chunks(s0, s1, nbc);
    return chunks > 1 ? (size_t)chunks * s0 * s1 * 16 : 0;
}

static int vote2_run(const float* tin, const float* lam, const float* th, float* out, const Nb2* host_nb, int nbc,
                     float s1, float s2, int power, unsigned s0, unsigned s1n, bool stick, bool plate,
                     unsigned samples, hipStream_t st) {
    const size_t n = (size_t)s0 * s1n;
    if (!n) return TIRA_OK;
    float4* pay = nullptr;
    TH_TRY(ws_get(WS_PAY0, n * 16, (void**)&pay));
    Nb2* dnb = nullptr;
    static thread_local TableCache nb_cache;
    TH_TRY(table_to_device(WS_NB2, host_nb, (size_t)(nbc > 0 ? nbc : 0) * sizeof(Nb2), nb_cache, st, (void**)&dnb));
    float4* part = nullptr;
    const size_t pb = vote2_part_bytes(s0, s1n, nbc);
    if (pb) TH_TRY(ws_get(WS_PAY1, pb, (void**)&part));
    return vote2_launch(tin, lam, th, out, pay, dnb, nbc, s1, s2, power, s0, s1n, stick, plate, samples, part, st);
}

// ---------------------------------------------------------------- allocation-free 2D vote
// tira_tensorvote2_prebuilt_noalloc keeps the voter payload and the
// neighbour table in this file-scope device array instead of library
// workspace. Each device holds its own copy for the life of the process; its
// device address is resolved once per device (hipGetSymbolAddress) and kept
// in g_scratch: host code passes that address, never the array's name.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
constexpr int kNoAllocMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch[kNoAllocMaxDevices] = {};

static int scratch_address(void** out) {
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    TH_ARG(dev >= 0 && dev < kNoAllocMaxDevices, "device %d beyond the allocation-free vote's table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        TH_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        TH_ARG(p != nullptr, "no device address for the allocation-free vote's scratch array");
        g_scratch[dev] = p;
    }
    *out = g_scratch[dev];
    return TIRA_OK;
}

// ---------------------------------------------------------------- search-loop session
// A parameter search (differential evolution over sigma1, sigma2, power on
// one field) votes the same field thousands of times with small changes. The
// session keeps the field on the device, keeps the device copies of the last
// kSessionTables neighbour tables (by parameters, or by the bytes of a
// caller's table), and launches each vote as one replay of an instantiated
// graph (payload pass, gather, copy-out) whose kernel arguments are updated in
// place: one launch and one wait per call instead of an upload of the field,
// a table copy, two kernel launches and a download.
constexpr int kSessionTables = 64;

struct SessionTable {
    // key: parameters when built here, or the caller's bytes (tables of other
    // parameters differ within the first entries, so memcmp settles a miss early)
    bool built = false;
    int w = 0, power = 0, stick = 0, plate = 0;
    unsigned samples = 0;
    float s1 = 0.0f, s2 = 0.0f;
    std::vector<unsigned char> bytes;
    Nb2* dev = nullptr;
    int count = 0;
    unsigned long long used = 0;
};

struct Vote2Session {
    int device = -1;
    unsigned s0 = 0, s1 = 0;
    size_t n = 0;
    float4* field = nullptr;
    float4* pay = nullptr;
    float4* out = nullptr;      // device result when the caller's output is host memory
    float4* pinned = nullptr;   // page-locked staging of a host result
    float4* part = nullptr;     // partial sums of the table chunks
    size_t part_cap = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipGraphNode_t n_prep = nullptr, n_vote = nullptr, n_reduce = nullptr;
    int graph_host_out = -1;
    unsigned long long clock = 0, replays = 0, table_hits = 0, table_builds = 0;
    SessionTable tabs[kSessionTables];
};

static void session_free(Vote2Session* s) {
    if (!s) return;
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    if (s->graph) (void)hipGraphDestroy(s->graph);
    for (auto& t : s->tabs)
        if (t.dev) (void)hipFree(t.dev);
    if (s->field) (void)hipFree(s->field);
    if (s->pay) (void)hipFree(s->pay);
    if (s->out) (void)hipFree(s->out);
    if (s->part) (void)hipFree(s->part);
    if (s->pinned) (void)hipHostFree(s->pinned);
    delete s;
}

// The device table for this call: a cached entry, or a new one (built from the
// parameters when `nb` is NULL) in the least recently used slot.
static int session_table(Vote2Session& s, const Nb2* nb, int nbc, int w, float s1, float s2, unsigned power,
                         unsigned samples, bool stick, bool plate, const Nb2** dev, int* count) {
    const bool built = nb == nullptr;
    const size_t nbytes = built ? 0 : (size_t)nbc * sizeof(Nb2);
    s.clock++;
    SessionTable* lru = &s.tabs[0];
    for (auto& t : s.tabs) {
        const bool hit = t.used && t.built == built &&
                         (built ? (t.w == w && t.s1 == s1 && t.s2 == s2 && t.power == (int)power &&
                                   t.samples == samples && t.stick == (int)stick && t.plate == (int)plate)
                                : (t.bytes.size() == nbytes && std::memcmp(t.bytes.data(), nb, nbytes) == 0));
        if (hit) {
            t.used = s.clock;
            s.table_hits++;
            *dev = t.dev;
            *count = t.count;
            return TIRA_OK;
        }
        if (t.used < lru->used) lru = &t;
    }
    if (built) nbc = (stick || plate) && w > 0 ? w * w : 0;
    SessionTable& t = *lru;
    if (t.dev) {
        // the graph's last replay may still read the slot being replaced
        TH_HIP(hipStreamSynchronize(stream()));
        TH_HIP(hipFree(t.dev));
        t.dev = nullptr;
    }
    t.used = 0;
    TH_HIP(hipMalloc((void**)&t.dev, (size_t)(nbc > 0 ? nbc : 1) * sizeof(Nb2)));
    if (built && nbc > 0) {
        // built on the device, ahead of the vote on the same stream
        k_build_nb2<<<ceil_div((size_t)nbc, 256), 256, 0, stream()>>>(w, s1, s2, power, samples, stick ? 1 : 0,
                                                                      plate ? 1 : 0, t.dev);
        TH_HIP(hipGetLastError());
    } else if (nbc > 0) {
        TH_HIP(hipMemcpy(t.dev, nb, (size_t)nbc * sizeof(Nb2), hipMemcpyHostToDevice));
    }
    t.built = built;
    t.w = w; t.s1 = s1; t.s2 = s2; t.power = (int)power; t.samples = samples;
    t.stick = stick; t.plate = plate;
    if (built) t.bytes.clear();
    else t.bytes.assign((const unsigned char*)nb, (const unsigned char*)nb + nbytes);
    t.count = nbc;
    t.used = s.clock;
    s.table_builds++;
    *dev = t.dev;
    *count = nbc;
    return TIRA_OK;
}

// Kernel-node parameters of one call: the argument values live in `a`, which
// must outlive the graph update that reads them. The session always takes the
// chunked form (partial sums, then the ordered reduction), also with one chunk,
// so that the graph's shape does not depend on the table.
struct SessionArgs {
    const float4* field; size_t n; float eta_s; float4* pay;
    const Nb2* nb; int nbc, per, s0, s1, power, stick, plate; float* out; float4* part;
    int chunks;
    void* prep[4];
    void* vote[11];
    void* reduce[4];
    void bind() {
        prep[0] = &field; prep[1] = &n; prep[2] = &eta_s; prep[3] = &pay;
        vote[0] = &pay; vote[1] = &nb; vote[2] = &nbc; vote[3] = &per; vote[4] = &s0; vote[5] = &s1;
        vote[6] = &power; vote[7] = &stick; vote[8] = &plate; vote[9] = &out; vote[10] = &part;
        reduce[0] = &part; reduce[1] = &chunks; reduce[2] = &n; reduce[3] = &out;
    }
};

static hipKernelNodeParams kernel_node(void* func, dim3 grid, void** args) {
    hipKernelNodeParams p{};
    p.func = func;
    p.gridDim = grid;
    p.blockDim = dim3(256);
    p.sharedMemBytes = 0;
    p.kernelParams = args;
    p.extra = nullptr;
    return p;
}

struct SessionNodes {
    hipKernelNodeParams prep, vote, reduce;
    SessionNodes(const Vote2Session& s, SessionArgs& a)
        : prep(kernel_node(reinterpret_cast<void*>(k_vote2_prep), dim3(ceil_div(s.n, 256)), a.prep)),
          vote(kernel_node(reinterpret_cast<void*>(k_vote2), dim3(ceil_div(s.s1, 64), ceil_div(s.s0, 4), a.chunks),
                           a.vote)),
          reduce(kernel_node(reinterpret_cast<void*>(k_vote2_reduce), dim3(ceil_div(s.n, 256)), a.reduce)) {}
};

static int session_build_graph(Vote2Session& s, SessionArgs& a, bool host_out) {
    if (s.exec) { TH_HIP(hipGraphExecDestroy(s.exec)); s.exec = nullptr; }
    if (s.graph) { TH_HIP(hipGraphDestroy(s.graph)); s.graph = nullptr; }
    TH_HIP(hipGraphCreate(&s.graph, 0));
    const SessionNodes nodes(s, a);
    TH_HIP(hipGraphAddKernelNode(&s.n_prep, s.graph, nullptr, 0, &nodes.prep));
    TH_HIP(hipGraphAddKernelNode(&s.n_vote, s.graph, &s.n_prep, 1, &nodes.vote));
    TH_HIP(hipGraphAddKernelNode(&s.n_reduce, s.graph, &s.n_vote, 1, &nodes.reduce));
    if (host_out) {
        hipGraphNode_t n_copy = nullptr;
        TH_HIP(hipGraphAddMemcpyNode1D(&n_copy, s.graph, &s.n_reduce, 1, s.pinned, s.out, s.n * 16,
                                       hipMemcpyDeviceToHost));
    }
    TH_HIP(hipGraphInstantiate(&s.exec, s.graph, nullptr, nullptr, 0));
    s.graph_host_out = host_out ? 1 : 0;
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_build_neighbors2d(int w, float sigma1, float sigma2, unsigned power, unsigned samples, int stick, int plate,
                           void* out, int capacity, int* count) {
    TH_ARG(count, "count is NULL");
    std::vector<Nb2> nb;
    const int k = build_neighbors2d(w, sigma1, sigma2, power, samples, stick != 0, plate != 0, nb);
    *count = k;
    if (out) {
        TH_ARG(capacity >= k, "capacity %d < %d entries", capacity, k);
        std::memcpy(out, nb.data(), (size_t)k * sizeof(Nb2));
    }
    return TIRA_OK;
}

// cpu::build_neighbors2d (tensorvote.h:1223-1280) on the device: all w * w
// entries into dev_out (device memory), non-blocking.
int tira_build_neighbors2d_async(int w, float sigma1, float sigma2, unsigned power, unsigned samples, int stick,
                                 int plate, void* dev_out) {
    TH_ARG(dev_out, "NULL pointer");
    TH_ARG(w >= 1 && w <= 46339, "w must be 1..46339");
    TH_ARG(sigma1 >= 0.0f && sigma2 >= 0.0f, "sigma must be >= 0");
    const size_t n = (size_t)w * w;
    k_build_nb2<<<ceil_div(n, 256), 256, 0, stream()>>>(w, sigma1, sigma2, power < 1 ? 1 : power, samples,
                                                         stick ? 1 : 0, plate ? 1 : 0, (Nb2*)dev_out);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

static int vote2_entry(float* t_out, const float* t_in, const float* lam, const float* th, const Nb2* nb, int nbc,
                       float s1, float s2, int power, unsigned w, unsigned s0, unsigned s1n, int stick, int plate,
                       unsigned samples) {
    TH_ARG(t_out && (t_in || (lam && th)), "NULL pointer");
    TH_ARG(s1 >= 0.0f && s2 >= 0.0f, "sigma must be >= 0");
    (void)w;
    const size_t n = (size_t)s0 * s1n;
    if (!n) return TIRA_OK;
    const float *din = nullptr, *dl = nullptr, *dt = nullptr;
    void* dout;
    if (t_in) TH_TRY(stage_in(t_in, n * 16, WS_IN0, (const void**)&din));
    else {
        TH_TRY(stage_in(lam, n * 8, WS_IN0, (const void**)&dl));
        TH_TRY(stage_in(th, n * 8, WS_IN1, (const void**)&dt));
    }
    TH_TRY(stage_out(t_out, n * 16, WS_OUT0, &dout));
    TH_TRY(vote2_run(din, dl, dt, (float*)dout, nb, nbc, s1, s2, power, s0, s1n, stick != 0, plate != 0, samples,
                     stream()));
    return finish_out(t_out, dout, n * 16);
}

// tira::cuda::tensorvote (2D), tensorvote.h:2200-2226
int tira_tensorvote2(float* t_out, const float* t_in, float sigma1, float sigma2, int power, unsigned w, unsigned s0,
                     unsigned s1, int stick, int plate, unsigned samples, int is_atv) {
    // ATV takes sigma1 only; sigma2, power, stick, plate and samples do not apply (tensorvote.h:2211-2212).
    if (is_atv) return vote2_atv_entry(t_out, t_in, nullptr, nullptr, sigma1, w, s0, s1);
    TH_ARG(w >= 1, "w must be >= 1");
    std::vector<Nb2> nb;
    const bool use_plate = plate && samples > 0;
    int nbc = 0;
    if (stick || use_plate)
        nbc = build_neighbors2d((int)w, sigma1, sigma2, (unsigned)(power < 1 ? 1 : power), samples, stick != 0,
                                use_plate, nb);
    return vote2_entry(t_out, t_in, nullptr, nullptr, nb.data(), nbc, sigma1, sigma2, power, w, s0, s1, stick,
                       plate, samples);
}

// tira::cuda::tensorvote_prebuilt, tensorvote.h:2236-2305
int tira_tensorvote2_prebuilt(float* t_out, const float* t_in, const void* host_nb, int nb_count, float sigma1,
                              float sigma2, int power, unsigned w, unsigned s0, unsigned s1, int stick, int plate,
                              unsigned samples) {
    TH_ARG(nb_count >= 0 && (nb_count == 0 || host_nb), "bad neighbour table");
    return vote2_entry(t_out, t_in, nullptr, nullptr, (const Nb2*)host_nb, nb_count, sigma1, sigma2, power, w, s0,
                       s1, stick, plate, samples);
}

// No reference counterpart: tira_tensorvote2_prebuilt without any device
// allocation (scratch array above), for callers that cannot allocate.
int tira_tensorvote2_prebuilt_noalloc(float* t_out, const float* t_in, const void* nb, int nb_count,
                                       float sigma1, float sigma2, int power, unsigned w, unsigned s0, unsigned s1,
                                       int stick, int plate, unsigned samples) {
    TH_ARG(t_out && t_in, "NULL pointer");
    TH_ARG(nb_count >= 0 && (nb_count == 0 || nb), "bad neighbour table");
    TH_ARG(sigma1 >= 0.0f && sigma2 >= 0.0f, "sigma must be >= 0");
    (void)w;
    const size_t n = (size_t)s0 * s1;
    // scratch layout: payload | table | partial sums of the table chunks
    const size_t pay_bytes = (n * 16 + 255) & ~(size_t)255;
    const size_t tab_bytes = (size_t)nb_count * sizeof(Nb2);
    const size_t part_off = pay_bytes + ((tab_bytes + 255) & ~(size_t)255);
    const size_t need = part_off + vote2_part_bytes(s0, s1, nb_count);
    TH_ARG(need <= kNoAllocScratchBytes,
           "a %u x %u field with %d table entries needs %zu bytes of scratch; the allocation-free vote holds %zu "
           "(use tira_tensorvote2_prebuilt)", s0, s1, nb_count, need, kNoAllocScratchBytes);
    if (!n) return TIRA_OK;
    TH_ARG(is_device(t_in) && is_device(t_out), "the allocation-free vote takes device fields (it stages nothing)");
    void* base = nullptr;
    TH_TRY(scratch_address(&base));
    hipStream_t st = stream();
    float4* pay = (float4*)base;
    const Nb2* dnb = (const Nb2*)nb;
    if (nb_count > 0 && !is_device(nb)) {
        Nb2* tab = (Nb2*)((unsigned char*)base + pay_bytes);
        TH_HIP(hipMemcpyAsync(tab, nb, tab_bytes, hipMemcpyHostToDevice, st));
        dnb = tab;
    }
    float4* part = (float4*)((unsigned char*)base + part_off);
    TH_TRY(vote2_launch(t_in, nullptr, nullptr, t_out, pay, dnb, nb_count, sigma1, sigma2, power, s0, s1,
                        stick != 0, plate != 0, samples, part, st));
    TH_HIP(hipStreamSynchronize(st));
    return TIRA_OK;
}

// tira::cuda::tensorvote_tk (2D, from eigenpairs), tensorvote.h:2125-2179
int tira_tensorvote2_tk(float* t_out, const float* lambdas, const float* thetas, float sigma1, float sigma2,
                        unsigned power, unsigned w, unsigned s0, unsigned s1, int stick, int plate,
                        unsigned samples) {
    TH_ARG(w >= 1, "w must be >= 1");
    std::vector<Nb2> nb;
    const bool use_plate = plate && samples > 0;
    int nbc = 0;
    if (stick || use_plate)
        nbc = build_neighbors2d((int)w, sigma1, sigma2, power < 1 ? 1 : power, samples, stick != 0, use_plate, nb);
    return vote2_entry(t_out, nullptr, lambdas, thetas, nb.data(), nbc, sigma1, sigma2, (int)power, w, s0, s1,
                       stick, plate, samples);
}

// tira::cuda::tensorvote_atv (2D, from eigenpairs), tensorvote.h:2029-2066
int tira_tensorvote2_atv(float* t_out, const float* lambdas, const float* thetas, float sigma, unsigned w,
                         unsigned s0, unsigned s1) {
    return vote2_atv_entry(t_out, nullptr, lambdas, thetas, sigma, w, s0, s1);
}

// ---------------------------------------------------------------- search-loop session
// tensorvote_tk_prebuilt (tensorvote_tk.cu:239-253, tensorvote.h:2236-2305)
// with the field bound once, for optimize_tk's repeated calls
// (optimize.py:256-288).
int tira_vote2_session_create(void** session, const float* t_in, unsigned s0, unsigned s1) {
    TH_ARG(session && t_in, "NULL pointer");
    TH_ARG(s0 > 0 && s1 > 0, "empty field");
    *session = nullptr;
    Vote2Session* s = new Vote2Session();
    s->s0 = s0;
    s->s1 = s1;
    s->n = (size_t)s0 * s1;
    const size_t bytes = s->n * 16;
    int st = TIRA_OK;
    do {
        if (hipGetDevice(&s->device) != hipSuccess ||
            hipMalloc((void**)&s->field, bytes) != hipSuccess || hipMalloc((void**)&s->pay, bytes) != hipSuccess) {
            set_error("device allocation for the vote session failed (%zu bytes per buffer)", bytes);
            st = TIRA_ERR_OUT_OF_MEMORY;
            break;
        }
        if (hipMemcpy(s->field, t_in, bytes, hipMemcpyDefault) != hipSuccess) {
            set_error("copying the field into the vote session failed");
            st = TIRA_ERR_HIP;
        }
    } while (0);
    if (st != TIRA_OK) { session_free(s); (void)hipGetLastError(); return st; }
    *session = s;
    return TIRA_OK;
}

// One vote of the bound field: with `nb` (nb_count 48-byte Neighbor2D entries,
// host memory) like tensorvote_tk_prebuilt, or with nb == NULL from the
// parameters like tensorvote_tk (table built on the device by the same
// arithmetic as build_neighbors2d). t_out:
// s0 x s1 x 4 floats, host or device. Blocking.
int tira_vote2_session_vote(void* session, float* t_out, const void* nb, int nb_count, float sigma1, float sigma2,
                            int power, unsigned w, int stick, int plate, unsigned samples) {
    TH_ARG(session && t_out, "NULL pointer");
    TH_ARG(sigma1 >= 0.0f && sigma2 >= 0.0f, "sigma must be >= 0");
    TH_ARG(nb_count >= 0 && (nb_count == 0 || nb), "bad neighbour table");
    Vote2Session& s = *(Vote2Session*)session;
    int dev = 0;
    TH_HIP(hipGetDevice(&dev));
    TH_ARG(dev == s.device, "the session lives on device %d, the current device is %d", s.device, dev);
    TH_ARG(nb || w >= 1, "w must be >= 1");
    TH_ARG(!nb || !is_device(nb), "the session takes a host neighbour table");
    const unsigned p = power < 1 ? 1 : (unsigned)power;
    const bool use_plate = plate && samples > 0;
    const Nb2* dnb = nullptr;
    int nbc = 0;
    TH_TRY(session_table(s, (const Nb2*)nb, nb_count, (int)w, sigma1, sigma2, p, samples, stick != 0, use_plate,
                         &dnb, &nbc));
    const bool host_out = !is_device(t_out);
    if (host_out && !s.out) {
        TH_HIP(hipMalloc((void**)&s.out, s.n * 16));
        TH_HIP(hipHostMalloc((void**)&s.pinned, s.n * 16, hipHostMallocDefault));
    }
    const int chunks = vote2_chunks(s.s0, s.s1, nbc);
    const size_t part_bytes = (size_t)chunks * s.n * 16;
    if (part_bytes > s.part_cap) {
        if (s.part) TH_HIP(hipFree(s.part));  // the previous call has finished (blocking)
        s.part = nullptr;
        s.part_cap = 0;
        TH_HIP(hipMalloc((void**)&s.part, part_bytes));
        s.part_cap = part_bytes;
    }
    SessionArgs a;
    a.field = s.field; a.n = s.n; a.eta_s = stick_norm_2d(sigma1, sigma2, p); a.pay = s.pay;
    a.nb = dnb; a.nbc = nbc; a.per = vote2_per(nbc, chunks); a.s0 = (int)s.s0; a.s1 = (int)s.s1;
    a.power = (int)p; a.stick = stick ? 1 : 0; a.plate = use_plate ? 1 : 0;
    a.out = host_out ? (float*)s.out : t_out;
    a.part = s.part;
    a.chunks = chunks;
    a.bind();
    if (s.graph_host_out != (host_out ? 1 : 0)) {
        TH_TRY(session_build_graph(s, a, host_out));
    } else {
        const SessionNodes nodes(s, a);
        TH_HIP(hipGraphExecKernelNodeSetParams(s.exec, s.n_prep, &nodes.prep));
        TH_HIP(hipGraphExecKernelNodeSetParams(s.exec, s.n_vote, &nodes.vote));
        TH_HIP(hipGraphExecKernelNodeSetParams(s.exec, s.n_reduce, &nodes.reduce));
    }
    hipStream_t st = stream();
    {
        ProfScope prof("vote2_session", st);
        TH_HIP(hipGraphLaunch(s.exec, st));
    }
    TH_HIP(hipStreamSynchronize(st));
    s.replays++;
    if (host_out) std::memcpy(t_out, s.pinned, s.n * 16);
    return TIRA_OK;
}

// Counters of a session: graph replays, table cache hits and table uploads.
int tira_vote2_session_stats(void* session, unsigned long long* replays, unsigned long long* table_hits,
                             unsigned long long* table_builds) {
    TH_ARG(session, "NULL session");
    const Vote2Session& s = *(const Vote2Session*)session;
    if (replays) *replays = s.replays;
    if (table_hits) *table_hits = s.table_hits;
    if (table_builds) *table_builds = s.table_builds;
    return TIRA_OK;
}

int tira_vote2_session_destroy(void* session) {
    if (!session) return TIRA_OK;
    TH_HIP(hipStreamSynchronize(stream()));
    session_free((Vote2Session*)session);
    return TIRA_OK;
}

}  // extern "C"
