// common.hpp -- runtime plumbing shared by every translation unit of
// libtira_hip.so: error reporting, the per-thread launch stream, host/device
// pointer detection and a per-thread, per-device grow-only workspace.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "tira_hip.h"

namespace th {

void set_error(const char* fmt, ...);
int fail_hip(hipError_t e, const char* what, const char* file, int line);

#define TH_HIP(expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) return th::fail_hip(e_, #expr, __FILE__, __LINE__);    \
    } while (0)
#define TH_ARG(cond, ...)                                                            \
    do {                                                                             \
        if (!(cond)) {                                                               \
            th::set_error(__VA_ARGS__);                                              \
            return TIRA_ERR_INVALID_ARG;                                             \
        }                                                                            \
    } while (0)
#define TH_TRY(expr)                                                                 \
    do {                                                                             \
        int s_ = (expr);                                                             \
        if (s_ != TIRA_OK) return s_;                                                \
    } while (0)

// Stream all launches of the calling host thread go to.
hipStream_t stream();
// True for device (or managed) memory; host / unregistered memory is false.
bool is_device(const void* p);
// Compute units of the current device (cached per device).
int cu_count();
// Select `device` when >= 0; rejects negative devices.
int use_device(int device);

// Workspace slots: each is a grow-only device buffer owned by the calling
// thread on the current device.
enum Slot {
    WS_IN0 = 0, WS_IN1, WS_IN2, WS_OUT0, WS_OUT1,
    WS_PAY0, WS_PAY1, WS_TABLE, WS_TABLE2, WS_SCRATCH0, WS_SCRATCH1, WS_SCRATCH2, WS_COUNTER,
    WS_TABLE3, WS_PAY2, WS_PIPE0, WS_PIPE1, WS_SPEC0, WS_SPEC1, WS_THIN0, WS_THIN1,
    WS_NB2,  // the 2D neighbour table of the last stick/plate vote (kept while unchanged)
    WS_TABLE4,  // offset table of the tabulated-network vote
    WS_NSLOTS
};
int ws_get(int slot, size_t bytes, void** p);
// Bytes held by `slot` for this thread on the current device (0 if none).
size_t ws_capacity(int slot);
// Changes whenever any workspace buffer of this thread is allocated or released.
unsigned long long ws_generation();

// Device view of a caller buffer: `src` itself when it is device memory,
// else a workspace copy (H2D on the library stream).
int stage_in(const void* src, size_t bytes, int slot, const void** dev);
// Device buffer to write an output into: `dst` when device, else workspace.
int stage_out(void* dst, size_t bytes, int slot, void** dev);
// Copy a staged output back when `dst` is host memory; synchronises.
int finish_out(void* dst, const void* dev, size_t bytes);

// Per-kernel timing with HIP events on the launch stream, active between
// tira_hip_profile_begin() and tira_hip_profile_end().
class ProfScope {
  public:
    ProfScope(const char* name, hipStream_t st);
    ~ProfScope();
  private:
    int slot_ = -1;
    hipStream_t st_;
};

// Plain 16-byte-vector copy kernel (bandwidth reference for profiler counters).
int launch_copy(void* dst, const void* src, size_t bytes, hipStream_t st);
// FP32 FMA loop: blocks x 256 threads x 16 x iters FLOP (peak VALU rate reference).
int launch_fma_loop(float* out, unsigned blocks, int iters, hipStream_t st);

inline unsigned ceil_div(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }

// XCD-major block numbering: hardware hands consecutive workgroup ids to the
// 8 XCDs round-robin; remap so each XCD walks a contiguous run of tiles.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    const unsigned per = (nblocks + kXcd - 1) / kXcd;
    const unsigned xcd = b % kXcd, idx = b / kXcd;
    const unsigned full = nblocks % kXcd == 0 ? kXcd : nblocks % kXcd;  // XCDs that own `per` blocks
    const unsigned start = xcd < full ? xcd * per : full * per + (xcd - full) * (per - 1);
    return start + idx;
}


}  // namespace th
