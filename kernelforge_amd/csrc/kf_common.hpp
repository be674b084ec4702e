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

// Serialises the C-ABI entry points of one device.  They share the device workspace
// slots and pinned staging above, so every compute entry point holds this guard for
// its whole host-side body (a per-device recursive mutex: entry points may call one
// another).  When the outermost call on a device uses a different stream from the
// previous call, its stream first waits for an event recorded at the end of that
// call, so the GPU work of calls on different streams never overlaps in the
// workspace either.  Not applied while the stream is being captured into a graph.
class EntryGuard {
  public:
    explicit EntryGuard(hipStream_t s);
    ~EntryGuard();
    EntryGuard(const EntryGuard &) = delete;
    EntryGuard &operator=(const EntryGuard &) = delete;

  private:
    int dev_ = -1;
    hipStream_t stream_ = nullptr;
};

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

// ---------------------------------------------------------------------------
// FP64 exp of the kernels' Gaussian exponents, by the table-driven method of
// P. T. P. Tang, "Table-driven implementation of the exponential function in IEEE
// floating-point arithmetic", ACM TOMS 15(2), 1989, with 64 table entries:
//   exp(x) = 2^y/64 for y = x * 64 / ln 2;  y = 64 m + j + f, j in [0, 64), |f| <= 1/2
//   exp(x) = 2^m * 2^(j/64) * exp(f ln2 / 64)
// and exp(f ln2/64) - 1 by its Taylor polynomial of degree 5 (|f ln2/64| <= 0.0055, the
// remainder is below 4e-17).  The caller passes y directly (the exponent's scale folded
// into one constant, kExpScale).  Error <= 1.3 ulp over results in the normal range.  y is
// clamped to [-70000, 70000] first: below 2^-1093 the result is 0, above 2^1093 it is
// +inf (as exp underflows and overflows), and a NaN y gives 0.
// ---------------------------------------------------------------------------
constexpr double kExpScale = 64.0 / 0.693147180559945309417232121458176568;  // 64 / ln 2

// 2^(j/64), j = 0..63, correctly rounded.
static __constant__ double kExp2Tab64[64] = {
    0x1.0000000000000p+0, 0x1.02c9a3e778061p+0, 0x1.059b0d3158574p+0, 0x1.0874518759bc8p+0,
    0x1.0b5586cf9890fp+0, 0x1.0e3ec32d3d1a2p+0, 0x1.11301d0125b51p+0, 0x1.1429aaea92de0p+0,
    0x1.172b83c7d517bp+0, 0x1.1a35beb6fcb75p+0, 0x1.1d4873168b9aap+0, 0x1.2063b88628cd6p+0,
    0x1.2387a6e756238p+0, 0x1.26b4565e27cddp+0, 0x1.29e9df51fdee1p+0, 0x1.2d285a6e4030bp+0,
    0x1.306fe0a31b715p+0, 0x1.33c08b26416ffp+0, 0x1.371a7373aa9cbp+0, 0x1.3a7db34e59ff7p+0,
    0x1.3dea64c123422p+0, 0x1.4160a21f72e2ap+0, 0x1.44e086061892dp+0, 0x1.486a2b5c13cd0p+0,
    0x1.4bfdad5362a27p+0, 0x1.4f9b2769d2ca7p+0, 0x1.5342b569d4f82p+0, 0x1.56f4736b527dap+0,
    0x1.5ab07dd485429p+0, 0x1.5e76f15ad2148p+0, 0x1.6247eb03a5585p+0, 0x1.6623882552225p+0,
    0x1.6a09e667f3bcdp+0, 0x1.6dfb23c651a2fp+0, 0x1.71f75e8ec5f74p+0, 0x1.75feb564267c9p+0,
    0x1.7a11473eb0187p+0, 0x1.7e2f336cf4e62p+0, 0x1.82589994cce13p+0, 0x1.868d99b4492edp+0,
    0x1.8ace5422aa0dbp+0, 0x1.8f1ae99157736p+0, 0x1.93737b0cdc5e5p+0, 0x1.97d829fde4e50p+0,
    0x1.9c49182a3f090p+0, 0x1.a0c667b5de565p+0, 0x1.a5503b23e255dp+0, 0x1.a9e6b5579fdbfp+0,
    0x1.ae89f995ad3adp+0, 0x1.b33a2b84f15fbp+0, 0x1.b7f76f2fb5e47p+0, 0x1.bcc1e904bc1d2p+0,
    0x1.c199bdd85529cp+0, 0x1.c67f12e57d14bp+0, 0x1.cb720dcef9069p+0, 0x1.d072d4a07897cp+0,
    0x1.d5818dcfba487p+0, 0x1.da9e603db3285p+0, 0x1.dfc97337b9b5fp+0, 0x1.e502ee78b3ff6p+0,
    0x1.ea4afa2a490dap+0, 0x1.efa1bee615a27p+0, 0x1.f50765b6e4540p+0, 0x1.fa7c1819e90d8p+0,
};

// exp(y ln2 / 64) with tab = a copy of kExp2Tab64 (in LDS for per-lane lookups).
__device__ __forceinline__ double exp_scaled(double y, const double *tab) {
    // (ln2/64)^n / n!, n = 1..5
    constexpr double c1 = 0x1.62e42fefa39efp-7, c2 = 0x1.ebfbdff82c58fp-15,
                     c3 = 0x1.c6b08d704a0c0p-23, c4 = 0x1.3b2ab6fba4e77p-31,
                     c5 = 0x1.5d87fe78a6731p-40;
    y = fmin(fmax(y, -70000.0), 70000.0);
    const double k = __builtin_rint(y);
    const double f = y - k;  // exact
    double p = fma(f, c5, c4);
    p = fma(f, p, c3);
    p = fma(f, p, c2);
    p = fma(f, p, c1);
    p *= f;
    const int ki = static_cast<int>(k);
    const double t = tab[ki & 63];
    return ldexp(fma(t, p, t), ki >> 6);
}

}  // namespace kf
