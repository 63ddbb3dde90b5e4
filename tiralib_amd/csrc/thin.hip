// thin.hip -- the consumer end of tensor voting on gfx950: stick saliency,
// Otsu threshold and Lee's 3D thinning.
//
//   * 2D stick saliency |l|max - |l|min (reconstruct.py:14-34); the 3D
//     magnitude-ranked saliency is an eig3 output mode (TIRA_EIG_MAGNITUDE);
//   * Otsu's threshold over a 256-bin histogram on [0, max]
//     (reconstruct.py:37-63) with numpy.histogram's float32 binning rules,
//     all on the device (max reduction, LDS-privatised histogram, one-thread
//     scan in FP64), then the binary mask `s > t`;
//   * Lee et al. 1994 thinning as tira::lee / computeThinImage implement it
//     (thinning/lee_thinning.h:600-706), bit-exact. Per border direction the
//     reference collects candidates on the unchanged volume, then re-checks
//     them one by one in raster order, so a candidate sees the deletions of
//     the candidates before it. Here candidates are collected in parallel and
//     the sequential re-check is solved as the fixed point of
//       keep(p) = not simple(N_p with earlier deleted candidates removed),
//     iterated in parallel (Jacobi) until no decision changes: the decisions
//     form a DAG along raster order, whose solution is unique, so any fixed
//     point is the sequential result.
// Neighbourhoods are 27-bit codes, bit 9(dz+1) + 3(dy+1) + (dx+1), the order
// of get_neighborhood (lee_thinning.h:52-91); bit 13 is the voxel itself.
#include "common.hpp"
#include "eig_device.hpp"

#include <vector>

namespace th {

// ---------------------------------------------------------------- neighbourhood tests
// Bits with x = 0 / x = 2 / y = 0 / y = 2 inside the 3x3x3 code.
constexpr unsigned kX0 = 0x1249249u, kX2 = 0x4924924u, kY0 = 0x01C0E07u, kY2 = 0x70381C0u, kAll = 0x7FFFFFFu;

// 3x3x3 dilation of a set of neighbourhood bits, one axis at a time; each
// step is clipped to the 27 bits, or a left shift past bit 26 would come back
// into range with the next right shift.
TH_HD inline unsigned dilate26(unsigned s) {
    s = (s | ((s << 1) & ~kX0) | ((s >> 1) & ~kX2)) & kAll;
    s = (s | ((s << 3) & ~kY0) | ((s >> 3) & ~kY2)) & kAll;
    return (s | (s << 9) | (s >> 9)) & kAll;
}

// Simple-point test of is_simple_point (lee_thinning.h:530-579): the 26
// neighbours form at most one 26-connected component. Flood fill by repeated
// 3x3x3 dilation inside the neighbour set.
TH_HD inline bool thin_simple(unsigned code) {
    const unsigned obj = code & ~(1u << 13) & kAll;
    if (!obj) return true;
    unsigned c = obj & (0u - obj);
    for (;;) {
        const unsigned g = dilate26(c) & obj;
        if (g == c) break;
        c = g;
    }
    return c == obj;
}

// Mask of the voxels other than the centre that share the centre cube's cell
// (sx, sy, sz): s = 0 spans an axis, +-1 lies on that side (a vertex has
// three nonzero s, an edge two, a face one).
TH_HD constexpr unsigned cell_mask(int sx, int sy, int sz) {
    unsigned m = 0;
    for (int ox = (sx < 0 ? sx : 0); ox <= (sx > 0 ? sx : 0); ox++)
        for (int oy = (sy < 0 ? sy : 0); oy <= (sy > 0 ? sy : 0); oy++)
            for (int oz = (sz < 0 ? sz : 0); oz <= (sz > 0 ? sz : 0); oz++)
                if (ox || oy || oz) m |= 1u << (9 * (oz + 1) + 3 * (oy + 1) + (ox + 1));
    return m;
}

// Euler invariance of is_euler_invariant (lee_thinning.h:288-299), computed as
// the change of V - E + F - C of the union of closed voxel cubes when the
// centre cube is removed: only centre-cube cells that no other voxel's cube
// contains disappear. Equivalent to the octant-LUT sum being zero over all
// 2^26 neighbourhoods.
TH_HD inline bool thin_euler_invariant(unsigned code) {
    int d = -1;
#pragma unroll
    for (int sx = -1; sx <= 1; sx++)
#pragma unroll
        for (int sy = -1; sy <= 1; sy++)
#pragma unroll
            for (int sz = -1; sz <= 1; sz++) {
                const int dim = (sx == 0) + (sy == 0) + (sz == 0);
                if (dim == 3) continue;
                if ((code & cell_mask(sx, sy, sz)) == 0) d += (dim % 2 == 0) ? 1 : -1;
            }
    return d == 0;
}

// is_endpoint (lee_thinning.h:96-103): exactly one set neighbour.
TH_HD inline bool thin_endpoint(unsigned code) { return __builtin_popcount(code) == 2; }

// Code of voxel (x, y, z) in a 0/1 byte volume (s2 = x fastest), zero outside.
__device__ __forceinline__ unsigned thin_code(const unsigned char* __restrict__ v, int s0, int s1, int s2, int x,
                                              int y, int z) {
    unsigned code = 0;
#pragma unroll
    for (int dz = -1; dz <= 1; dz++) {
        const int Z = z + dz;
        if ((unsigned)Z >= (unsigned)s0) continue;
#pragma unroll
        for (int dy = -1; dy <= 1; dy++) {
            const int Y = y + dy;
            if ((unsigned)Y >= (unsigned)s1) continue;
            const unsigned char* row = v + ((size_t)Z * s1 + Y) * s2;
#pragma unroll
            for (int dx = -1; dx <= 1; dx++) {
                const int X = x + dx;
                if ((unsigned)X < (unsigned)s2 && row[X] == 1) code |= 1u << (9 * (dz + 1) + 3 * (dy + 1) + dx + 1);
            }
        }
    }
    return code;
}

// Wave-aggregated append of `take` lanes to list[*count...].
__device__ __forceinline__ void wave_append(bool take, unsigned value, unsigned* __restrict__ list,
                                            unsigned* __restrict__ count) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(m & ((1ull << lane) - 1ull))] = value;
}

// As wave_append, with a second value per entry appended to `list2` at the
// same slot.
__device__ __forceinline__ void wave_append2(bool take, unsigned value, unsigned value2, unsigned* __restrict__ list,
                                             unsigned* __restrict__ list2, unsigned* __restrict__ count) {
    const unsigned long long m = __ballot(take);
    if (!m) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll((long long)m) - 1;
    unsigned base = 0;
    if (lane == leader) base = atomicAdd(count, (unsigned)__popcll(m));
    base = __shfl(base, leader);
    if (take) {
        const unsigned slot = base + __popcll(m & ((1ull << lane) - 1ull));
        list[slot] = value;
        list2[slot] = value2;
    }
}

__device__ __forceinline__ void wave_count(bool hit, unsigned* __restrict__ count) {
    const unsigned long long m = __ballot(hit);
    if (m && (threadIdx.x & 63) == (unsigned)(__ffsll((long long)m) - 1)) atomicAdd(count, (unsigned)__popcll(m));
}

// ---------------------------------------------------------------- thinning kernels
// Binarise (lee(), lee_thinning.h:690-699): any nonzero value becomes 1.
template <typename T>
__global__ __launch_bounds__(256) void k_thin_binarise(const T* in, unsigned char* v, size_t n) {  // may alias
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        v[i] = in[i] != (T)0 ? 1 : 0;
}

template <typename T>
__global__ __launch_bounds__(256) void k_thin_export(const unsigned char* __restrict__ v, T* __restrict__ out,
                                                     size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = (T)v[i];
}

// Object voxels (value 1) of the volume, or of a previous list, appended to
// `list`; every lane of a wave runs the same number of iterations.
__global__ __launch_bounds__(256) void k_thin_objects(const unsigned char* __restrict__ v, size_t n,
                                                      const unsigned* __restrict__ from, unsigned nfrom,
                                                      unsigned* __restrict__ list, unsigned* __restrict__ count) {
    const size_t total = from ? (size_t)nfrom : n;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    const size_t iters = (total + stride - 1) / stride;
    for (size_t it = 0; it < iters; it++) {
        const size_t k = it * stride + (size_t)blockIdx.x * blockDim.x + threadIdx.x;
        unsigned i = 0;
        bool take = false;
        if (k < total) {
            i = from ? from[k] : (unsigned)k;
            take = v[i] == 1;
        }
        if (list) wave_append(take, i, list, count);
        else wave_count(take, count);
    }
}

// Candidates of border direction `dir` (0..5 = N y-1, S y+1, E x+1, W x-1,
// U z+1, B z-1; lee_thinning.h:142-147, 623-628) among the listed object
// voxels: voxels whose neighbour in that direction is 0 (or outside), that are
// not end points, are Euler invariant and simple (lee_thinning.h:614-646).
// Each candidate's 26-neighbourhood code goes to codes[] beside its index:
// the volume does not change until the direction's deletions are applied, so
// the re-check sweeps read it instead of the 26 neighbours.
__global__ __launch_bounds__(256) void k_thin_candidates_list(const unsigned char* __restrict__ v, int s0, int s1,
                                                              int s2, int dir, const unsigned* __restrict__ obj,
                                                              unsigned nobj, unsigned char* __restrict__ mark,
                                                              unsigned* __restrict__ list,
                                                              unsigned* __restrict__ codes,
                                                              unsigned* __restrict__ count) {
    const int ddx = dir == 2 ? 1 : dir == 3 ? -1 : 0;
    const int ddy = dir == 0 ? -1 : dir == 1 ? 1 : 0;
    const int ddz = dir == 4 ? 1 : dir == 5 ? -1 : 0;
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned iters = (nobj + stride - 1) / stride;
    for (unsigned it = 0; it < iters; it++) {
        const unsigned k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool take = false;
        unsigned i = 0, code = 0;
        if (k < nobj) {
            i = obj[k];
            if (v[i] == 1) {
                const int x = (int)(i % (unsigned)s2), y = (int)((i / (unsigned)s2) % (unsigned)s1);
                const int z = (int)(i / ((unsigned)s1 * (unsigned)s2));
                const int X = x + ddx, Y = y + ddy, Z = z + ddz;
                const bool border = (unsigned)X >= (unsigned)s2 || (unsigned)Y >= (unsigned)s1 ||
                                    (unsigned)Z >= (unsigned)s0 || v[((size_t)Z * s1 + Y) * s2 + X] == 0;
                if (border) {
                    code = thin_code(v, s0, s1, s2, x, y, z);
                    take = !thin_endpoint(code) && thin_euler_invariant(code) && thin_simple(code);
                }
            }
        }
        if (take) mark[i] = 1;
        wave_append2(take, i, code, list, codes, count);
    }
}

// One Jacobi sweep of the sequential re-check (lee_thinning.h:651-661):
// candidate p is deleted (mark 1) when it is simple in the volume where the
// candidates before it in raster order (neighbour bits 0..12) that are
// currently marked for deletion are removed; otherwise kept (mark 2).
// The candidate count is read from device memory (the candidate pass wrote
// it), so the host launches over an upper bound without reading it back.
// codes[k]: candidate k's neighbourhood code (k_thin_candidates_list); a set
// bit is an object neighbour, inside the volume, so its index is the
// candidate's plus the bit's offset.
// prev (optional): the previous sweep's change count; zero means the marks
// are already a fixed point (mark_in equals mark_out on every candidate), and
// the sweep returns at once.
__global__ __launch_bounds__(256) void k_thin_resolve(int s1, int s2, const unsigned* __restrict__ list,
                                                      const unsigned* __restrict__ codes,
                                                      const unsigned* __restrict__ count,
                                                      const unsigned char* __restrict__ mark_in,
                                                      unsigned char* __restrict__ mark_out,
                                                      unsigned* __restrict__ changes,
                                                      const unsigned* __restrict__ prev) {
    if (prev && *prev == 0u) return;
    const unsigned cnt = *count;
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned iters = (cnt + stride - 1) / stride;
    const long long plane = (long long)s1 * s2;
    for (unsigned it = 0; it < iters; it++) {
        const unsigned k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool changed = false;
        if (k < cnt) {
            const size_t i = list[k];
            unsigned code = codes[k];
            unsigned earlier = code & 0x1FFFu;
            while (earlier) {
                const int b = __builtin_ctz(earlier);
                earlier &= earlier - 1;
                const int dx = b % 3 - 1, dy = (b / 3) % 3 - 1, dz = b / 9 - 1;
                const size_t q = (size_t)((long long)i + dz * plane + (long long)dy * s2 + dx);
                if (mark_in[q] == 1) code &= ~(1u << b);
            }
            const unsigned char m = thin_simple(code) ? 1 : 2;
            mark_out[i] = m;
            changed = m != mark_in[i];
        }
        wave_count(changed, changes);
    }
}

// Apply the decisions: delete marked candidates, clear both mark buffers.
__global__ __launch_bounds__(256) void k_thin_apply(unsigned char* __restrict__ v, const unsigned* __restrict__ list,
                                                    const unsigned* __restrict__ count,
                                                    unsigned char* __restrict__ mark_a,
                                                    unsigned char* __restrict__ mark_b, const int final_is_a,
                                                    unsigned* __restrict__ deleted) {
    const unsigned cnt = *count;
    const unsigned stride = gridDim.x * blockDim.x;
    const unsigned iters = (cnt + stride - 1) / stride;
    for (unsigned it = 0; it < iters; it++) {
        const unsigned k = it * stride + blockIdx.x * blockDim.x + threadIdx.x;
        bool del = false;
        if (k < cnt) {
            const size_t i = list[k];
            del = (final_is_a ? mark_a[i] : mark_b[i]) == 1;
            if (del) v[i] = 0;
            mark_a[i] = 0;
            mark_b[i] = 0;
        }
        wave_count(del, deleted);
    }
}

// Per-code flags of the three tests (bit 0 Euler invariant, bit 1 simple,
// bit 2 end point) for codes `first + k` enumerated with the centre set:
// code = low 13 bits | 1 << 13 | high bits << 14.
__global__ __launch_bounds__(256) void k_thin_classify(unsigned first, unsigned count,
                                                       unsigned char* __restrict__ flags) {
    for (unsigned k = blockIdx.x * blockDim.x + threadIdx.x; k < count; k += gridDim.x * blockDim.x) {
        const unsigned e = first + k;
        const unsigned code = (e & 0x1FFFu) | (1u << 13) | ((e >> 13) << 14);
        flags[k] = (unsigned char)((thin_euler_invariant(code) ? 1 : 0) | (thin_simple(code) ? 2 : 0) |
                                   (thin_endpoint(code) ? 4 : 0));
    }
}

__global__ __launch_bounds__(256) void k_thin_zero_n(unsigned* __restrict__ c, int n) {
    for (int i = (int)threadIdx.x; i < n; i += (int)blockDim.x) c[i] = 0u;
}

__global__ void k_thin_zero(unsigned* __restrict__ c, int n) {
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(256) void k_fill_u8(unsigned char* __restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0;
}

__global__ void k_fill_u64(unsigned long long* __restrict__ p, int n) {
    if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------- saliency and Otsu
// 2D stick saliency |l|max - |l|min of [a b; b c] (reconstruct.py:14-34).
__global__ __launch_bounds__(256) void k_stick_saliency2(const float4* __restrict__ t, size_t n,
                                                         float* __restrict__ s) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const float4 m = t[i];
        float l0, l1;
        eig2_vals(m.x, m.y, m.w, l0, l1);
        const float a0 = fabsf(l0), a1 = fabsf(l1);
        s[i] = fmaxf(a0, a1) - fminf(a0, a1);
    }
}

// Order-preserving float <-> unsigned key (for atomicMax on any finite float).
__device__ __forceinline__ unsigned fkey(float f) {
    const unsigned b = __float_as_uint(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}
__device__ __forceinline__ float fkey_inv(unsigned k) {
    return __uint_as_float((k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k);
}

__global__ __launch_bounds__(256) void k_otsu_max(const float* __restrict__ v, size_t n, int stride,
                                                  unsigned* __restrict__ key) {
    unsigned m = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        m = max(m, fkey(v[i * stride]));
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) m = max(m, (unsigned)__shfl_xor((int)m, o));
    if ((threadIdx.x & 63) == 0) atomicMax(key, m);
}

// numpy.histogram(values, 256, range=(0, vmax)) for float32 values: float32
// edges = float32(i * vmax / 256), index from the float64 quotient, then the
// +-1 corrections at the edges; the last bin is closed.
__global__ __launch_bounds__(256) void k_otsu_hist(const float* __restrict__ v, size_t n, int stride,
                                                   const unsigned* __restrict__ key,
                                                   unsigned long long* __restrict__ hist) {
    __shared__ unsigned h[256];
    h[threadIdx.x] = 0;
    __syncthreads();
    const float vmax = fkey_inv(*key);
    if (vmax > 0.0f) {
        const double dv = (double)vmax, step = dv / 256.0;
        for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
            const float a = v[i * stride];
            if (!(a >= 0.0f && a <= vmax)) continue;
            int idx = (int)(((double)a / dv) * 256.0);
            if (idx == 256) idx = 255;
            if (a < (float)((double)idx * step)) idx--;
            if (idx != 255 && a >= (idx + 1 == 256 ? vmax : (float)((double)(idx + 1) * step))) idx++;
            atomicAdd(&h[idx], 1u);
        }
    }
    __syncthreads();
    if (h[threadIdx.x]) atomicAdd(&hist[threadIdx.x], (unsigned long long)h[threadIdx.x]);
}

// Otsu's between-class variance over the 256 bins, FP64, first maximum
// (reconstruct.py:46-63); threshold = that bin's float32 centre.
__global__ void k_otsu_select(const unsigned* __restrict__ key, const unsigned long long* __restrict__ hist,
                              float* __restrict__ thr) {
    if (threadIdx.x != 0) return;
    const float vmax = fkey_inv(*key);
    if (!(vmax > 0.0f)) { *thr = 0.0f; return; }
    const double step = (double)vmax / 256.0;
    double total = 0.0;
    for (int i = 0; i < 256; i++) total += (double)hist[i];
    if (total == 0.0) { *thr = 0.0f; return; }
    double cum = 0.0, w0 = 0.0, best = 0.0;
    // mu_total is the last cumulative sum: a first pass in the same order
    for (int i = 0; i < 256; i++) {
        const float e0 = (float)((double)i * step), e1 = i == 255 ? vmax : (float)((double)(i + 1) * step);
        cum += (double)hist[i] * (double)((e0 + e1) / 2.0f);
    }
    const double mu_total = cum;
    cum = 0.0;
    float bc = 0.0f;
    for (int i = 0; i < 256; i++) {
        const float e0 = (float)((double)i * step), e1 = i == 255 ? vmax : (float)((double)(i + 1) * step);
        const float c = (e0 + e1) / 2.0f;
        cum += (double)hist[i] * (double)c;
        w0 += (double)hist[i];
        const double w1 = total - w0;
        const double mu0 = w0 > 0 ? cum / w0 : 0.0;
        const double mu1 = w1 > 0 ? (mu_total - cum) / w1 : 0.0;
        const double d = mu0 - mu1;
        const double b = w0 * w1 * (d * d);
        if (i == 0 || b > best) { best = b; bc = c; }
    }
    *thr = bc;
}

__global__ __launch_bounds__(256) void k_threshold(const float* __restrict__ v, size_t n, int stride,
                                                   const float* __restrict__ thr, unsigned char* __restrict__ mask) {
    const float t = *thr;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        mask[i] = v[i * stride] > t ? 1 : 0;
}

static unsigned grid_for(size_t n, unsigned per_block = 256, unsigned cap = 8192) {
    return (unsigned)std::max<size_t>(1, std::min<size_t>(ceil_div(n, per_block), cap));
}

// ---------------------------------------------------------------- host drivers
int otsu_run(const float* v, size_t n, int stride, float* thr, hipStream_t st) {
    unsigned* key = nullptr;
    TH_TRY(ws_get(WS_COUNTER, 8 + 256 * 8 + 64, (void**)&key));
    unsigned long long* hist = (unsigned long long*)((char*)key + 64);
    k_fill_u64<<<1, 256, 0, st>>>(hist, 256);
    k_thin_zero<<<1, 64, 0, st>>>(key, 1);
    ProfScope prof("otsu", st);
    k_otsu_max<<<grid_for(n, 256, 4096), 256, 0, st>>>(v, n, stride, key);
    k_otsu_hist<<<grid_for(n, 256 * 16, 2048), 256, 0, st>>>(v, n, stride, key, hist);
    k_otsu_select<<<1, 64, 0, st>>>(key, hist, thr);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

constexpr int kThinSlots = 34;  // counters per direction in an outer pass (1 + 33 sweeps)
constexpr unsigned kThinFirst = 4;  // re-check sweeps of a direction before its first read-back

// Lee thinning of the 0/1 byte volume `v` (device) in place; blocking. The
// object voxels are compacted into a list once and the list is refiltered
// after every outer pass, so each border direction examines the remaining
// object voxels only instead of the whole volume.
int lee_run(unsigned char* v, unsigned s0, unsigned s1, unsigned s2, unsigned* iterations, hipStream_t st) {
    const size_t n = (size_t)s0 * s1 * s2;
    unsigned char *ma = nullptr, *mb = nullptr;
    unsigned *list = nullptr, *cnt = nullptr, *objA = nullptr, *objB = nullptr;
    TH_TRY(ws_get(WS_SCRATCH0, n, (void**)&ma));
    TH_TRY(ws_get(WS_SCRATCH1, n, (void**)&mb));
    TH_TRY(ws_get(WS_COUNTER, (16 + 6 * kThinSlots) * 4, (void**)&cnt));
    k_fill_u8<<<grid_for(n), 256, 0, st>>>(ma, n);
    k_fill_u8<<<grid_for(n), 256, 0, st>>>(mb, n);
    TH_HIP(hipGetLastError());
    unsigned host[4];
    ProfScope prof("lee_thin", st);
    // object list: count, size the buffers, gather
    k_thin_zero<<<1, 64, 0, st>>>(cnt, 4);
    k_thin_objects<<<grid_for(n), 256, 0, st>>>(v, n, nullptr, 0, nullptr, cnt + 3);
    TH_HIP(hipGetLastError());
    TH_HIP(hipMemcpyAsync(host + 3, cnt + 3, 4, hipMemcpyDeviceToHost, st));
    TH_HIP(hipStreamSynchronize(st));
    unsigned nobj = host[3];
    if (nobj == 0) {
        if (iterations) *iterations = 1;  // one outer pass finds nothing to delete
        return TIRA_OK;
    }
    TH_TRY(ws_get(WS_SCRATCH2, (size_t)nobj * 4, (void**)&list));
    TH_TRY(ws_get(WS_THIN0, (size_t)nobj * 4, (void**)&objA));
    TH_TRY(ws_get(WS_THIN1, (size_t)nobj * 4, (void**)&objB));
    k_thin_zero<<<1, 64, 0, st>>>(cnt + 3, 1);
    k_thin_objects<<<grid_for(n), 256, 0, st>>>(v, n, nullptr, 0, objA, cnt + 3);
    TH_HIP(hipGetLastError());
    // Counters: [1] [2] changes of the two sweeps of a batch beyond the
    // per-direction slots, [3] object voxels, [8 + dir] deletions of this
    // outer pass, then per direction kThinSlots slots: [0] its candidates,
    // [1 + s] the changes of its re-check sweep s. One launch clears a pass's
    // counters; the re-check and apply passes read the candidate count on the
    // device (launched over the object count, an upper bound), so a direction
    // needs one read-back for its first kThinFirst sweeps and one per two
    // after, and an outer pass one more. The
    // candidates' neighbourhood codes go to objB, which is free until the
    // pass's object list is rebuilt.
    int iters = 0;
    unsigned host_pass[6];
    for (;;) {
        iters++;
        k_thin_zero_n<<<1, 256, 0, st>>>(cnt + 8, 8 + 6 * kThinSlots);
        for (int dir = 0; dir < 6; dir++) {
            unsigned* dc = cnt + 16 + dir * kThinSlots;
            k_thin_candidates_list<<<grid_for(nobj), 256, 0, st>>>(v, (int)s0, (int)s1, (int)s2, dir, objA, nobj, ma,
                                                                   list, objB, dc);
            TH_HIP(hipGetLastError());
            unsigned char *in = ma, *out = mb;
            // the first batch runs kThinFirst sweeps (most directions settle
            // within them; a sweep after a fixed point returns at once), later
            // batches two, each batch ending with one read-back
            for (unsigned sweep = 0;;) {
                // bounded: each sweep settles at least one more level of the raster-order DAG, and
                // there are at most nobj candidates; a sweep that changes nothing is a fixed point,
                // so the later sweeps of a batch may repeat a settled state harmlessly
                TH_ARG(sweep <= nobj + kThinFirst, "thinning re-check did not settle (internal error)");
                const unsigned nb = sweep == 0 ? kThinFirst : 2;
                unsigned* ch = dc + 1 + sweep;
                if (sweep + nb > (unsigned)kThinSlots - 1) {  // beyond the cleared slots
                    ch = cnt + 1;
                    k_thin_zero<<<1, 64, 0, st>>>(ch, 2);
                }
                for (unsigned b = 0; b < nb; b++) {
                    k_thin_resolve<<<grid_for(nobj), 256, 0, st>>>((int)s1, (int)s2, list, objB, dc, in, out, ch + b,
                                                                   b ? ch + b - 1 : nullptr);
                    std::swap(in, out);
                }
                TH_HIP(hipGetLastError());
                TH_HIP(hipMemcpyAsync(host + 2, ch + nb - 1, 4, hipMemcpyDeviceToHost, st));
                TH_HIP(hipStreamSynchronize(st));
                if (host[2] == 0) break;
                sweep += nb;
            }
            k_thin_apply<<<grid_for(nobj), 256, 0, st>>>(v, list, dc, ma, mb, in == ma ? 1 : 0, cnt + 8 + dir);
            TH_HIP(hipGetLastError());
        }
        // drop deleted voxels from the object list; read it with this pass's deletions
        k_thin_zero<<<1, 64, 0, st>>>(cnt + 3, 1);
        k_thin_objects<<<grid_for(nobj), 256, 0, st>>>(v, n, objA, nobj, objB, cnt + 3);
        TH_HIP(hipGetLastError());
        TH_HIP(hipMemcpyAsync(host + 3, cnt + 3, 4, hipMemcpyDeviceToHost, st));
        TH_HIP(hipMemcpyAsync(host_pass, cnt + 8, 6 * 4, hipMemcpyDeviceToHost, st));
        TH_HIP(hipStreamSynchronize(st));
        bool deleted_any = false;
        for (int d = 0; d < 6; d++) deleted_any = deleted_any || host_pass[d] != 0;
        if (!deleted_any) break;  // a whole pass of six directions deleted nothing
        nobj = host[3];
        std::swap(objA, objB);
        if (nobj == 0) {  // the next outer pass would find no candidate in any direction
            iters++;
            break;
        }
    }
    if (iterations) *iterations = (unsigned)iters;
    return TIRA_OK;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_stick_saliency2_async(const float* field4, size_t n, float* saliency) {
    TH_ARG(field4 && saliency, "NULL pointer");
    if (!n) return TIRA_OK;
    hipStream_t st = stream();
    k_stick_saliency2<<<grid_for(n), 256, 0, st>>>((const float4*)field4, n, saliency);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int tira_otsu_threshold_async(const float* values, size_t n, int stride, float* threshold) {
    TH_ARG(values && threshold, "NULL pointer");
    TH_ARG(stride >= 1, "stride must be >= 1");
    TH_ARG(n >= 1, "Otsu needs at least one value");
    return otsu_run(values, n, stride, threshold, stream());
}

int tira_threshold_async(const float* values, size_t n, int stride, const float* threshold, unsigned char* mask) {
    TH_ARG(values && threshold && mask, "NULL pointer");
    TH_ARG(stride >= 1, "stride must be >= 1");
    if (!n) return TIRA_OK;
    hipStream_t st = stream();
    k_threshold<<<grid_for(n), 256, 0, st>>>(values, n, stride, threshold, mask);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

int tira_lee_thin3_u8(unsigned char* mask, unsigned s0, unsigned s1, unsigned s2, unsigned* iterations) {
    TH_ARG(mask, "NULL pointer");
    TH_ARG((size_t)s0 * s1 * s2 < ((size_t)1 << 32), "volume too large for 32-bit voxel indices");
    TH_ARG(is_device(mask), "tira_lee_thin3_u8 takes a device pointer");
    const size_t n = (size_t)s0 * s1 * s2;
    if (!n) { if (iterations) *iterations = 0; return TIRA_OK; }
    hipStream_t st = stream();
    k_thin_binarise<unsigned char><<<grid_for(n), 256, 0, st>>>(mask, mask, n);
    TH_HIP(hipGetLastError());
    return lee_run(mask, s0, s1, s2, iterations, st);
}

int tira_lee_thin3(const int* in, int* out, unsigned x, unsigned y, unsigned z, unsigned* iterations) {
    TH_ARG(in && out, "NULL pointer");
    const size_t n = (size_t)x * y * z;
    TH_ARG(n < ((size_t)1 << 32), "volume too large for 32-bit voxel indices");
    if (!n) { if (iterations) *iterations = 0; return TIRA_OK; }
    const int* din; void* dout;
    TH_TRY(stage_in(in, n * 4, WS_IN0, (const void**)&din));
    TH_TRY(stage_out(out, n * 4, WS_OUT0, &dout));
    unsigned char* v = nullptr;
    TH_TRY(ws_get(WS_PIPE0, n, (void**)&v));
    hipStream_t st = stream();
    k_thin_binarise<int><<<grid_for(n), 256, 0, st>>>(din, v, n);
    TH_HIP(hipGetLastError());
    TH_TRY(lee_run(v, z, y, x, iterations, st));
    k_thin_export<int><<<grid_for(n), 256, 0, st>>>(v, (int*)dout, n);
    TH_HIP(hipGetLastError());
    return finish_out(out, dout, n * 4);
}

int tira_thin3_classify_async(unsigned first, unsigned count, unsigned char* flags) {
    TH_ARG(flags, "NULL pointer");
    TH_ARG((unsigned long long)first + count <= (1ull << 26), "codes beyond 2^26");
    if (!count) return TIRA_OK;
    hipStream_t st = stream();
    k_thin_classify<<<grid_for(count), 256, 0, st>>>(first, count, flags);
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

// Voted field -> magnitude-ranked saliency (eig3) -> Otsu threshold of the
// stick saliency -> mask -> Lee thinning. Device pointers; blocking (the
// thinning loop reads counters back).
int tira_skeleton3(const float* voted, int layout, unsigned s0, unsigned s1, unsigned s2, float* saliency3,
                   float* threshold, unsigned char* skeleton, unsigned* iterations) {
    TH_ARG(voted && saliency3 && threshold && skeleton, "NULL pointer");
    TH_ARG(layout == TIRA_LAYOUT_MAT9 || layout == TIRA_LAYOUT_PACKED6, "bad layout");
    TH_ARG(is_device(voted) && is_device(saliency3) && is_device(threshold) && is_device(skeleton),
           "tira_skeleton3 takes device pointers");
    const size_t n = (size_t)s0 * s1 * s2;
    TH_ARG(n >= 1, "empty volume");
    TH_ARG(n < ((size_t)1 << 32), "volume too large for 32-bit voxel indices");
    TH_TRY(tira_eig3_async(voted, n, layout, saliency3, nullptr, nullptr,
                           TIRA_EIG_SALIENCY | TIRA_EIG_MAGNITUDE | TIRA_EIG_REFINE_FP64 | TIRA_EIG_SCALE_INVARIANT,
                           nullptr));
    hipStream_t st = stream();
    TH_TRY(otsu_run(saliency3, n, 3, threshold, st));
    k_threshold<<<grid_for(n), 256, 0, st>>>(saliency3, n, 3, threshold, skeleton);
    TH_HIP(hipGetLastError());
    return lee_run(skeleton, s0, s1, s2, iterations, st);
}

}  // extern "C"
