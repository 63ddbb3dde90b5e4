// batch.hip -- many uneven tensor volumes in one persistent launch: the
// zero-padded separable Gaussian of every 6-component field and the
// eigenvalues of every input tensor, captured in a hipGraph.
//
// Work items are (volume, 16 x 16 tile, z-chunk) triples, ordered largest
// volume first, and handed out to a persistent grid of blocks (blocks/CU x
// CUs). The item range is cut into one contiguous part per XCD, each with its
// own atomic counter: block b runs on XCD b % 8 and takes the items of its
// XCD's part in order, so neighbouring tiles are in flight on one XCD together
// and share their halo columns through its L2; a block whose part is used up
// takes items from the other parts. Each item runs the fused Gaussian tile; because
// the tile stages every input plane of its columns in LDS anyway, the same
// item also eigen-decomposes those voxels from LDS, so each input tensor is
// read from HBM once for both results (24 B in, 24 B smoothed + 12 B
// eigenvalues out per voxel). An autotuner times the candidate z-chunks and
// occupancies on the caller's batch the first time a batch shape is seen and
// caches the winner per process.
//
// No reference counterpart: TIRA filters / decomposes one volume per call
// (filter.h:400-534, eigen.h:895-942).
#include "common.hpp"
#include "gauss_tile.hpp"

#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

namespace th {

struct BatchVolume {
    const float* in;     // s0 x s1 x s2 x 6
    float* out;          // smoothed, same shape
    float* evals;        // s0 x s1 x s2 x 3
    Gauss3Geom g;
    unsigned first_item;
    unsigned nitems;
};

constexpr int kBatchR = 5;  // 11 taps; other radii dispatch below
constexpr int kMaxVolumes = 4096;

constexpr unsigned kBatchParts = 8;  // one item part (and counter) per XCD

__device__ __forceinline__ unsigned batch_part_begin(unsigned total, unsigned q) {
    return (unsigned)(((unsigned long long)total * q) / kBatchParts);
}

// Waves per SIMD the batch kernel is bounded to: six for 16-row tiles up to
// R = 5 (80 VGPRs, no spill), none otherwise.
template <int R, int TY> struct BatchBound { static constexpr int W = (TY == 16 && R <= 5) ? 6 : 1; };

// TY: tile height (8, 16 or 32 rows of 16 outputs); LDS
// gauss_tile_lds_floats<R, 6, TY>() floats, sized at launch.
template <int R, int TY>
__global__ __launch_bounds__((GaussThreads<6, TY>::N), (BatchBound<R, TY>::W)) void k_batch_gauss_eig(
    const BatchVolume* __restrict__ vols, int nvol, unsigned total, unsigned* __restrict__ counter,
    const float* __restrict__ taps) {
    constexpr int PW = kGT + 2 * R;
    constexpr int PH = TY + 2 * R;
    extern __shared__ __attribute__((aligned(16))) float s_tile[];
    float* s_raw = s_tile;
    float* s_xf = s_tile + 6 * PH * gauss_row_stride(PW);
    __shared__ unsigned s_item;
    const unsigned home = blockIdx.x % kBatchParts;
    unsigned spent = 0;  // parts this block found used up (thread 0 only)
    while (true) {
        if (threadIdx.x == 0) {
            // the first part, from the home XCD's on, that still has an item;
            // each counter passes its part's end at most once per block, so the
            // search ends and every block leaves once all parts are used up
            unsigned item = total;
            for (unsigned t = 0; t < kBatchParts; t++) {
                const unsigned q = (home + t) % kBatchParts;
                if (spent & (1u << q)) continue;
                const unsigned lo = batch_part_begin(total, q), hi = batch_part_begin(total, q + 1);
                const unsigned c = atomicAdd(&counter[q], 1u);
                if (c < hi - lo) { item = lo + c; break; }
                spent |= 1u << q;
            }
            s_item = item;
        }
        __syncthreads();
        const unsigned item = s_item;
        __syncthreads();  // every lane has read s_item before it is overwritten
        if (item >= total) break;  // uniform: every block drains the counters and exits
        int v = 0;
        while (v + 1 < nvol && vols[v + 1].first_item <= item) v++;
        const BatchVolume& bv = vols[v];
        gauss3_tile<R, 6, true, TY>(bv.in, bv.out, taps, bv.g, item - bv.first_item, bv.evals, s_raw, s_xf);
    }
}

__global__ void k_reset_counter(unsigned* c) {
    if (threadIdx.x < kBatchParts) c[threadIdx.x] = 0u;
}

struct BatchPlan {
    int zchunk = 32;
    int blocks_per_cu = 2;
    int tile_y = 16;  // tile height (x width kGT)
    float tuned_ms = 0.0f;
};

struct Batch {
    int device = 0;
    int nvol = 0;
    int R = kBatchR;
    std::vector<const float*> in;
    std::vector<float*> out, evals;
    std::vector<unsigned> shapes;
    BatchVolume* d_vols = nullptr;
    unsigned* d_counter = nullptr;
    float* d_taps = nullptr;
    unsigned total = 0;
    BatchPlan plan;
    int ncu = 256;
    hipStream_t cap_stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static std::mutex g_plan_mu;
static std::map<std::string, BatchPlan> g_plan_cache;

static std::string plan_key(const Batch& b) {
    std::string k = std::to_string(b.device) + ":" + std::to_string(b.R);
    for (unsigned s : b.shapes) k += "," + std::to_string(s);
    return k;
}

// Builds the item table for a z-chunk (largest volumes first) and uploads it.
static int build_items(Batch& b, int zchunk, int tile_y, hipStream_t st) {
    std::vector<int> order(b.nvol);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const size_t vx = (size_t)b.shapes[3 * x] * b.shapes[3 * x + 1] * b.shapes[3 * x + 2];
        const size_t vy = (size_t)b.shapes[3 * y] * b.shapes[3 * y + 1] * b.shapes[3 * y + 2];
        return vx > vy;
    });
    std::vector<BatchVolume> vols(b.nvol);
    unsigned first = 0;
    for (int k = 0; k < b.nvol; k++) {
        const int i = order[k];
        BatchVolume& v = vols[k];
        v.in = b.in[i]; v.out = b.out[i]; v.evals = b.evals[i];
        Gauss3Geom& g = v.g;
        g.s0 = g.o0 = (int)b.shapes[3 * i];
        g.s1 = g.o1 = (int)b.shapes[3 * i + 1];
        g.s2 = g.o2 = (int)b.shapes[3 * i + 2];
        g.c0 = g.c1 = g.c2 = 0;
        g.zchunk = zchunk;
        g.nbx = ceil_div(g.o2, kGT);
        g.nby = ceil_div(g.o1, tile_y);
        g.nbz = ceil_div(g.o0, zchunk);
        v.first_item = first;
        v.nitems = g.nbx * g.nby * g.nbz;
        first += v.nitems;
    }
    b.total = first;
    TH_HIP(hipMemcpyAsync(b.d_vols, vols.data(), vols.size() * sizeof(BatchVolume), hipMemcpyHostToDevice, st));
    TH_HIP(hipStreamSynchronize(st));
    return TIRA_OK;
}

// Tiles above 64 KB of LDS (32 rows at R = 6) need the kernel's dynamic-LDS
// limit raised, once per process, before any launch (and before a capture).
template <int R, int TY>
static int batch_lds_attr() {
    const size_t lds = (size_t)gauss_tile_lds_floats<R, 6, TY>() * 4;
    static bool set = false;
    if (lds > 65536 && !set) {
        TH_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(k_batch_gauss_eig<R, TY>),
                                   hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
        set = true;
    }
    return TIRA_OK;
}

template <int R, int TY>
static int launch_batch_rt(Batch& b, unsigned grid, hipStream_t st, bool prepare_only) {
    TH_TRY((batch_lds_attr<R, TY>()));
    if (prepare_only) return TIRA_OK;
    const size_t lds = (size_t)gauss_tile_lds_floats<R, 6, TY>() * 4;
    k_batch_gauss_eig<R, TY><<<grid, GaussThreads<6, TY>::N, lds, st>>>(b.d_vols, b.nvol, b.total, b.d_counter,
                                                                       b.d_taps);
    return TIRA_OK;
}

template <int R>
static int launch_batch_r(Batch& b, unsigned grid, hipStream_t st, bool prepare_only) {
    switch (b.plan.tile_y) {
        case 8: return launch_batch_rt<R, 8>(b, grid, st, prepare_only);
        case 32: return launch_batch_rt<R, 32>(b, grid, st, prepare_only);
        default: return launch_batch_rt<R, 16>(b, grid, st, prepare_only);
    }
}

// prepare_only: set the kernel attributes the plan needs and launch nothing.
static int launch_batch(Batch& b, hipStream_t st, bool prepare_only = false) {
    const unsigned grid = (unsigned)(b.ncu * b.plan.blocks_per_cu);
    if (!prepare_only) k_reset_counter<<<1, 64, 0, st>>>(b.d_counter);
    switch (b.R) {
        case 3: TH_TRY(launch_batch_r<3>(b, grid, st, prepare_only)); break;
        case 4: TH_TRY(launch_batch_r<4>(b, grid, st, prepare_only)); break;
        case 5: TH_TRY(launch_batch_r<5>(b, grid, st, prepare_only)); break;
        case 6: TH_TRY(launch_batch_r<6>(b, grid, st, prepare_only)); break;
        default: set_error("batched filter radius %d outside 3..6", b.R); return TIRA_ERR_INVALID_ARG;
    }
    TH_HIP(hipGetLastError());
    return TIRA_OK;
}

static int autotune(Batch& b, hipStream_t st) {
    const std::string key = plan_key(b);
    {
        std::lock_guard<std::mutex> lk(g_plan_mu);
        auto it = g_plan_cache.find(key);
        if (it != g_plan_cache.end()) {
            b.plan = it->second;
            return build_items(b, b.plan.zchunk, b.plan.tile_y, st);
        }
    }
    hipEvent_t e0, e1;
    TH_HIP(hipEventCreate(&e0));
    TH_HIP(hipEventCreate(&e1));
    BatchPlan best;
    best.tuned_ms = 1e30f;
    // tile height x z-chunk (the tile's shape and halo share) x resident
    // blocks (256-, 512- and 1024-thread blocks for 8-, 16- and 32-row tiles)
    const int tys[] = {8, 16, 32};
    const int zcs[] = {8, 16, 32, 64};
    const int occ_by_ty[3][4] = {{2, 4, 6, 8}, {1, 2, 3, 4}, {1, 2, 0, 0}};
    for (int ti = 0; ti < 3; ti++)
    for (int zc : zcs) {
        const int ty = tys[ti];
        TH_TRY(build_items(b, zc, ty, st));
        for (int o : occ_by_ty[ti]) {
            if (o == 0) continue;
            b.plan.tile_y = ty;
            b.plan.zchunk = zc;
            b.plan.blocks_per_cu = o;
            TH_TRY(launch_batch(b, st));  // warm
            TH_HIP(hipEventRecord(e0, st));
            TH_TRY(launch_batch(b, st));
            TH_HIP(hipEventRecord(e1, st));
            TH_HIP(hipEventSynchronize(e1));
            float ms = 0;
            TH_HIP(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best.tuned_ms) { best = b.plan; best.tuned_ms = ms; }
        }
    }
    TH_HIP(hipEventDestroy(e0));
    TH_HIP(hipEventDestroy(e1));
    b.plan = best;
    TH_TRY(build_items(b, best.zchunk, best.tile_y, st));
    std::lock_guard<std::mutex> lk(g_plan_mu);
    g_plan_cache[key] = best;
    return TIRA_OK;
}

static int capture_graph(Batch& b) {
    if (b.exec) { TH_HIP(hipGraphExecDestroy(b.exec)); b.exec = nullptr; }
    if (b.graph) { TH_HIP(hipGraphDestroy(b.graph)); b.graph = nullptr; }
    TH_TRY(launch_batch(b, b.cap_stream, true));  // kernel attributes outside the capture
    TH_HIP(hipStreamBeginCapture(b.cap_stream, hipStreamCaptureModeThreadLocal));
    const int st = launch_batch(b, b.cap_stream);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(b.cap_stream, &g);
    TH_TRY(st);
    TH_HIP(ee);
    b.graph = g;
    TH_HIP(hipGraphInstantiate(&b.exec, b.graph, nullptr, nullptr, 0));
    return TIRA_OK;
}

static void destroy_batch(Batch* b) {
    if (!b) return;
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->cap_stream) (void)hipStreamDestroy(b->cap_stream);
    if (b->d_vols) (void)hipFree(b->d_vols);
    if (b->d_counter) (void)hipFree(b->d_counter);
    if (b->d_taps) (void)hipFree(b->d_taps);
    delete b;
}

}  // namespace th

using namespace th;

extern "C" {

int tira_batch_gauss_eig3_create(void** handle, int nvol, const float* const* in, float* const* out,
                                 float* const* evals, const unsigned* shapes, const float* taps, unsigned ntaps,
                                 int autotune_plan) {
    TH_ARG(handle && in && out && evals && shapes && taps, "NULL pointer");
    TH_ARG(nvol >= 1 && nvol <= kMaxVolumes, "nvol must be 1..%d", kMaxVolumes);
    TH_ARG(ntaps % 2 == 1 && ntaps >= 7 && ntaps <= 13, "taps must be odd, 7..13 (got %u)", ntaps);
    *handle = nullptr;
    Batch* b = new Batch();
    int st_ = TIRA_OK;
    do {
        if ((st_ = (hipGetDevice(&b->device) == hipSuccess) ? TIRA_OK : TIRA_ERR_HIP) != TIRA_OK) break;
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, b->device) != hipSuccess) { st_ = TIRA_ERR_HIP; break; }
        b->ncu = prop.multiProcessorCount;
        b->nvol = nvol;
        b->R = (int)ntaps / 2;
        b->in.assign(in, in + nvol);
        b->out.assign(out, out + nvol);
        b->evals.assign(evals, evals + nvol);
        b->shapes.assign(shapes, shapes + 3 * (size_t)nvol);
        bool ok = true;
        for (int i = 0; i < nvol; i++) {
            if (!in[i] || !out[i] || !evals[i] || !shapes[3 * i] || !shapes[3 * i + 1] || !shapes[3 * i + 2]) ok = false;
            if (!is_device(in[i]) || !is_device(out[i]) || !is_device(evals[i])) ok = false;
        }
        if (!ok) { set_error("every volume needs device in/out/evals pointers and a non-empty shape"); st_ = TIRA_ERR_INVALID_ARG; break; }
        if (hipMalloc(&b->d_vols, sizeof(BatchVolume) * nvol) != hipSuccess ||
            hipMalloc(&b->d_counter, 256) != hipSuccess || hipMalloc(&b->d_taps, 3 * 13 * 4) != hipSuccess) {
            set_error("device allocation for the batch failed"); st_ = TIRA_ERR_OUT_OF_MEMORY; break;
        }
        std::vector<float> t(3 * ntaps);
        for (int a = 0; a < 3; a++) std::memcpy(&t[a * ntaps], taps, ntaps * 4);
        if (hipMemcpy(b->d_taps, t.data(), t.size() * 4, hipMemcpyHostToDevice) != hipSuccess) { st_ = TIRA_ERR_HIP; break; }
        if (hipStreamCreateWithFlags(&b->cap_stream, hipStreamNonBlocking) != hipSuccess) { st_ = TIRA_ERR_HIP; break; }
        if (autotune_plan) st_ = autotune(*b, b->cap_stream);
        else st_ = build_items(*b, b->plan.zchunk, b->plan.tile_y, b->cap_stream);
        if (st_ != TIRA_OK) break;
        st_ = capture_graph(*b);
    } while (0);
    if (st_ != TIRA_OK) { destroy_batch(b); return st_; }
    *handle = b;
    return TIRA_OK;
}

int tira_batch_gauss_eig3_run(void* handle) {
    TH_ARG(handle, "NULL handle");
    Batch* b = (Batch*)handle;
    ProfScope prof("batch_gauss_eig3", stream());
    TH_HIP(hipGraphLaunch(b->exec, stream()));
    return TIRA_OK;
}

int tira_batch_gauss_eig3_set_plan(void* handle, int tile_y, int zchunk, int blocks_per_cu) {
    TH_ARG(handle, "NULL handle");
    TH_ARG(tile_y == 8 || tile_y == 16 || tile_y == 32, "tile height must be 8, 16 or 32 (got %d)", tile_y);
    TH_ARG(zchunk >= 1 && zchunk <= 4096, "z-chunk must be 1..4096 (got %d)", zchunk);
    TH_ARG(blocks_per_cu >= 1 && blocks_per_cu <= (tile_y == 32 ? 2 : 8), "blocks per CU out of range (got %d)",
           blocks_per_cu);
    Batch* b = (Batch*)handle;
    TH_HIP(hipStreamSynchronize(stream()));  // a replay on the library stream may still read the items
    b->plan.tile_y = tile_y;
    b->plan.zchunk = zchunk;
    b->plan.blocks_per_cu = blocks_per_cu;
    b->plan.tuned_ms = 0.0f;
    TH_TRY(build_items(*b, zchunk, tile_y, b->cap_stream));
    return capture_graph(*b);
}

int tira_batch_gauss_eig3_tile(void* handle, int* tile_x, int* tile_y) {
    TH_ARG(handle, "NULL handle");
    const Batch* b = (const Batch*)handle;
    if (tile_x) *tile_x = kGT;
    if (tile_y) *tile_y = b->plan.tile_y;
    return TIRA_OK;
}

int tira_batch_gauss_eig3_plan(void* handle, int* zchunk, int* blocks_per_cu, float* tuned_ms, unsigned* items) {
    TH_ARG(handle, "NULL handle");
    Batch* b = (Batch*)handle;
    if (zchunk) *zchunk = b->plan.zchunk;
    if (blocks_per_cu) *blocks_per_cu = b->plan.blocks_per_cu;
    if (tuned_ms) *tuned_ms = b->plan.tuned_ms;
    if (items) *items = b->total;
    return TIRA_OK;
}

int tira_batch_gauss_eig3_destroy(void* handle) {
    if (handle) {
        Batch* b = (Batch*)handle;
        TH_HIP(hipStreamSynchronize(b->cap_stream));
        destroy_batch(b);
    }
    return TIRA_OK;
}

}  // extern "C"
