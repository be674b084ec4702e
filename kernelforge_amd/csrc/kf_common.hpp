// kf_common.hpp -- shared host/device helpers for the gfx950 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <mutex>
#include <vector>

#include "../../include/kf_hip.h"

namespace kf {

// ---------------------------------------------------------------------------
// Error plumbing: every C-ABI entry point returns an int status.  HIP failures
// map to KF_ERR_HIP and the message is kept for kf_last_error().
// ---------------------------------------------------------------------------
void set_last_error(const char *fmt, ...);

#define KF_HIP_TRY(expr)                                                              \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::kf::set_last_error("%s:%d %s -> %s", __FILE__, __LINE__, #expr,         \
                                 hipGetErrorString(_e));                              \
            return KF_ERR_HIP;                                                        \
        }                                                                             \
    } while (0)

#define KF_REQUIRE(cond, code, ...)                                                   \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::kf::set_last_error(__VA_ARGS__);                                        \
            return (code);                                                            \
        }                                                                             \
    } while (0)

// ---------------------------------------------------------------------------
// Grow-only device workspace, one set of slots per device.  Slots are named by
// small integers so independent scratch buffers of one call do not alias.
// Calls on one device are serialised by the caller's stream; the cache itself is
// guarded by a mutex for the host-side bookkeeping.
// ---------------------------------------------------------------------------
enum WsSlot {
    WS_PACK = 0,   // packed per-label descriptor rows
    WS_PACK2,      // second set (rectangular kernels)
    WS_PLAN,       // plan tables (row offsets, source indices, tiles)
    WS_PLAN2,
    WS_AUX0,
    WS_AUX1,
    WS_AUX2,
    WS_AUX3,
    WS_SOLVE0,     // iterative-solver vectors (kept apart from the operator's slots)
    WS_SOLVE1,     // solver reductions
    WS_SOLVE2,     // solver metadata
    WS_NET,        // preconditioner-network activations
    WS_NSLOTS
};
void *workspace(int slot, size_t bytes, hipStream_t stream);

// Per-device auxiliary stream and events (created once, kept for the process) for calls
// that fork work onto a second stream and join back through events.
constexpr int kAuxEvents = 4;
hipStream_t aux_stream();
hipEvent_t aux_event(int i);

// Pinned host staging for small metadata round trips.
void *host_staging(int slot, size_t bytes);

// Q (nm x max_atoms) and N (nm) device arrays -> host vectors, through staging slot
// `slot`; synchronises `stream`.
int read_metadata(const int *Q, const int *N, int nm, int ma, std::vector<int> &hQ,
                  std::vector<int> &hN, hipStream_t stream, int slot);

inline hipStream_t as_stream(void *s) { return reinterpret_cast<hipStream_t>(s); }

// Kernel timers: when enabled (kf_timer_enable), entry points bracket their main
// kernel with HIP events recorded on the launch stream; kf_timer_read sums them.
enum TimerId {
    TM_FCHL19 = 0,
    TM_LOCAL_SYMM,
    TM_LOCAL_RECT,
    TM_POTRF,
    TM_POTRS,
    TM_FULL,
    TM_MATVEC,
    TM_COUNT
};
struct ScopedTimer {
    int id;
    hipStream_t stream;
    void *handle;
    ScopedTimer(int id_, hipStream_t s);
    void stop();  // record the end event now (idempotent)
    ~ScopedTimer() { stop(); }
};

// ---------------------------------------------------------------------------
// RFP addressing, TRANSR='N', UPLO='U' (the layout LAPACK dpftrf('N','U') and
// the reference kernels use).  Element (r, c) of the LOWER triangle of the full
// symmetric matrix (r >= c) -- equivalently the upper element (c, r) -- lives at
//   r <  k : (k + 1 + r) + c * s      (column-major block, leading dim s)
//   r >= k : (r - k) * s + c          (row-major block, leading dim s)
// with k = n / 2 and s = n + 1 (n even) or n (n odd).
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t rfp_lower(size_t n, size_t r, size_t c) {
    const size_t k = n >> 1;
    const size_t s = (n & 1) ? n : n + 1;
    return (r < k) ? (k + 1 + r) + c * s : (r - k) * s + c;
}
// Upper element (i, j), i <= j.
__host__ __device__ inline size_t rfp_upper(size_t n, size_t i, size_t j) {
    return rfp_lower(n, j, i);
}

__host__ __device__ inline int ceil_div(int a, int b) { return (a + b - 1) / b; }

// Row of result register v held by `lane` in a 16x16x4 MFMA (column = lane & 15).
// Both take A[i][k] from lane i + 16k and B[k][j] from lane j + 16k, but they place
// the 16x16 result differently: the FP64 form interleaves the four registers across
// row groups (row = 4v + lane/16), the FP32 form keeps them consecutive
// (row = 4*(lane/16) + v).
__device__ inline int mfma16_row_f64(int lane, int v) { return 4 * v + (lane >> 4); }
__device__ inline int mfma16_row_f32(int lane, int v) { return 4 * (lane >> 4) + v; }
template <typename T> __device__ inline int mfma16_row(int lane, int v);
template <> __device__ inline int mfma16_row<double>(int lane, int v) { return mfma16_row_f64(lane, v); }
template <> __device__ inline int mfma16_row<float>(int lane, int v) { return mfma16_row_f32(lane, v); }

}  // namespace kf
