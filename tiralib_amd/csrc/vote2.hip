// vote2.hip -- 2D tensor-kernel (TK) voting on gfx950, replacing
// tira::cuda::tensorvote / tensorvote_prebuilt / tensorvote_tk
// (tensorvote.h:2071-2305) and the host table builder
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
#include <vector>

namespace th {

struct Nb2 {  // tira::Neighbor2D, 48 bytes (tensorvote.h:776-786)
    int du, dv;
    float dx, dy, c1, c2, Ra, Rb, Rc, Va, Vb, Vc;
};
static_assert(sizeof(Nb2) == 48, "Neighbor2D must stay 48 bytes");

// ---------------------------------------------------------------- host table
// Same arithmetic as the reference (tensorvote.h:34-40, 803-914, 1047-1071):
// the unqualified cos/sin/exp of the reference's float code resolve to the
// double overloads under the host compiler, so they are evaluated in double
// and rounded here too.
static float atten(float dist, float sigma) {
    if (sigma == 0.0f) return (dist > 1e-12f) ? 0.0f : 1.0f;
    return (float)std::exp((double)(-(dist * dist) / (sigma * sigma)));
}

float stick_norm_2d(float s1, float s2, unsigned power) {
    if (power < 1) power = 1;
    float ratio = 1.0f;
    for (unsigned k = 1; k <= power; ++k) ratio *= (float)(2 * k) / (float)(2 * k - 1);
    return ratio / ((float)M_PI * (s1 * s1 + s2 * s2));
}

static void stickvote_2d(float rx, float ry, float s1, float s2, unsigned power, float theta, float& Va, float& Vb,
                         float& Vc) {
    const float qx = (float)std::cos((double)theta), qy = (float)std::sin((double)theta);
    const float dist = std::sqrt(rx * rx + ry * ry);
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

static void plate_numerical_2d(float rx, float ry, float s1, float s2, unsigned power, float& a, float& b, float& c,
                               unsigned n) {
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

int build_neighbors2d(int w, float s1, float s2, unsigned power, unsigned samples, bool stick, bool plate,
                      std::vector<Nb2>& nb) {
    nb.clear();
    if (w <= 0) return 0;
    const int hw = w / 2;
    const float is1 = (s1 > 0.0f) ? 1.0f / (s1 * s1) : 0.0f;
    const float is2 = (s2 > 0.0f) ? 1.0f / (s2 * s2) : 0.0f;
    nb.reserve((size_t)w * w);
    for (int dv = -hw; dv <= hw; dv++)
        for (int du = -hw; du <= hw; du++) {
            Nb2 n{};
            n.du = du; n.dv = dv;
            const float l2 = (float)(du * du + dv * dv);
            if (stick) {
                if (l2 != 0.0f) {
                    const float il = 1.0f / std::sqrt(l2);
                    n.dx = (float)du * il;
                    n.dy = (float)dv * il;
                }
                n.c1 = (s1 == 0.0f) ? (l2 > 1e-12f ? 0.0f : 1.0f) : std::exp(-l2 * is1);
                n.c2 = (s2 == 0.0f) ? (l2 > 1e-12f ? 0.0f : 1.0f) : std::exp(-l2 * is2);
                n.Ra = 1.0f - 2.0f * n.dx * n.dx;
                n.Rb = 0.0f - 2.0f * n.dx * n.dy;
                n.Rc = 1.0f - 2.0f * n.dy * n.dy;
            }
            if (plate && samples > 0) plate_numerical_2d((float)du, (float)dv, s1, s2, power, n.Va, n.Vb, n.Vc, samples);
            nb.push_back(n);
        }
    return (int)nb.size();
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
// (tk_stick_window 938-1005 + tk_plate_window 1100-1115).
__global__ __launch_bounds__(256) void k_vote2(const float4* __restrict__ pay, const Nb2* __restrict__ nb, int nbc,
                                               int s0, int s1, int power, int stick, int plate,
                                               float* __restrict__ out) {
    const int x1 = blockIdx.x * 64 + (threadIdx.x & 63);
    const int x0 = blockIdx.y * 4 + (threadIdx.x >> 6);
    float a = 0.0f, b = 0.0f, c = 0.0f;
    for (int k = 0; k < nbc; k++) {
        const Nb2 e = nb[k];
        const int r0 = x0 + e.dv, r1 = x1 + e.du;
        if ((unsigned)r0 >= (unsigned)s0 || (unsigned)r1 >= (unsigned)s1) continue;
        const float4 P = pay[(size_t)r0 * s1 + r1];
        if (stick) {
            const float qTd = fmaf(P.x, e.dx, P.y * e.dy);
            const float q2 = qTd * qTd;
            float D;
            if (power == 1) D = fmaf(e.c2 - e.c1, q2, e.c1);
            else {
                const float t1 = 1.0f - q2;
                float e1 = t1, e2 = q2;
                for (int i = 1; i < power; i++) { e1 *= t1; e2 *= q2; }
                D = fmaf(e.c1, e1, e.c2 * e2);
            }
            const float Rqx = fmaf(e.Ra, P.x, e.Rb * P.y), Rqy = fmaf(e.Rb, P.x, e.Rc * P.y);
            const float f = D * P.z;
            const float fx = f * Rqx;
            a = fmaf(fx, Rqx, a); b = fmaf(fx, Rqy, b); c = fmaf(f * Rqy, Rqy, c);
        }
        if (plate) {
            a = fmaf(P.w, e.Va, a); b = fmaf(P.w, e.Vb, b); c = fmaf(P.w, e.Vc, c);
        }
    }
    if (x0 < s0 && x1 < s1) {
        float4* o = reinterpret_cast<float4*>(out) + (size_t)x0 * s1 + x1;
        *o = make_float4(a, b, b, c);
    }
}

// ---------------------------------------------------------------- host driver
// plate: the reference's kernel adds the plate term when plate && samples > 0
// (tensorvote.h:2093); the table's V entries are zero otherwise.
static int vote2_run(const float* tin, const float* lam, const float* th, float* out, const Nb2* host_nb, int nbc,
                     float s1, float s2, int power, unsigned s0, unsigned s1n, bool stick, bool plate,
                     unsigned samples, hipStream_t st) {
    const size_t n = (size_t)s0 * s1n;
    if (!n) return TIRA_OK;
    const unsigned p = power < 1 ? 1 : (unsigned)power;
    const float eta_s = stick_norm_2d(s1, s2, p);
    float4* pay = nullptr;
    TH_TRY(ws_get(WS_PAY0, n * 16, (void**)&pay));
    {
        ProfScope prof("vote2_prep", st);
        if (tin) k_vote2_prep<<<ceil_div(n, 256), 256, 0, st>>>((const float4*)tin, n, eta_s, pay);
        else k_vote2_prep_tk<<<ceil_div(n, 256), 256, 0, st>>>((const float2*)lam, (const float2*)th, n, eta_s, pay);
    }
    TH_HIP(hipGetLastError());
    Nb2* dnb = nullptr;
    TH_TRY(ws_get(WS_TABLE2, (size_t)(nbc > 0 ? nbc : 1) * sizeof(Nb2), (void**)&dnb));
    if (nbc > 0) {
        // pageable source: copy synchronously after prior work on the stream
        TH_HIP(hipStreamSynchronize(st));
        TH_HIP(hipMemcpy(dnb, host_nb, (size_t)nbc * sizeof(Nb2), hipMemcpyHostToDevice));
    }
    dim3 grid(ceil_div(s1n, 64), ceil_div(s0, 4));
    {
        ProfScope prof("vote2", st);
        k_vote2<<<grid, 256, 0, st>>>(pay, dnb, nbc, (int)s0, (int)s1n, (int)p, stick ? 1 : 0,
                                      (plate && samples > 0) ? 1 : 0, out);
    }
    TH_HIP(hipGetLastError());
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
    if (is_atv) {
        set_error("analytical tensor voting (is_atv) is not provided by this library");
        return TIRA_ERR_UNSUPPORTED;
    }
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

}  // extern "C"
