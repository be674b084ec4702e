// rfp_chol.hip -- Cholesky factorisation and solve of an RFP-packed SPD matrix on gfx950.
//
// Reference: kf::math::solve_cholesky_rfp (src/math.cpp:134-171) = add l2 to the RFP
// diagonal, LAPACK dpftrf('N','U') then dpftrs('N','U').
//
// Design: the RFP array (TRANSR='N', UPLO='U') is addressed as the LOWER triangle of
// the full matrix (rfp_lower in kf_common.hpp): rows < n/2 form a column-major block,
// rows >= n/2 a row-major block, both with leading dimension s.  The panel partition
// is aligned at n/2, so every 64x64 tile lies in one region and is a plain strided
// matrix.  A right-looking blocked Cholesky then runs in place on the packed array:
//   per panel J:  POTRF of the 64x64 diagonal tile + its triangular inverse (one WG,
//                 LDS), panel TRSM as an MFMA product with the inverse, and the
//                 trailing SYRK/GEMM update on MFMA (v_mfma_f64_16x16x4 / f32_16x16x4).
// The factor lands exactly where LAPACK's dpftrf leaves U (= L^T of this layout).
// Solves run forward (L) and backward (L^T) over the same partition in FP64.
#include <cmath>

#include "kf_common.hpp"

namespace kf {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NB = 64;        // panel width / tile edge
constexpr int TP = NB + 1;    // LDS pitch

__device__ inline v4d mfma4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef v4d type; };
template <> struct Vec4<float> { typedef v4f type; };

// Partition of [0, n): 64-wide blocks inside [0, k) then inside [k, n).
struct Part {
    int n, k, nlo, nblk;  // nlo = number of blocks in [0, k)
    __host__ __device__ int start(int J) const {
        return J < nlo ? J * NB : k + (J - nlo) * NB;
    }
    __host__ __device__ int width(int J) const {
        const int s = start(J);
        const int end = J < nlo ? k : n;
        return (end - s) < NB ? (end - s) : NB;
    }
};
inline Part make_part(int n) {
    Part p;
    p.n = n;
    p.k = n / 2;
    p.nlo = (p.k + NB - 1) / NB;
    p.nblk = p.nlo + (n - p.k + NB - 1) / NB;
    return p;
}

// Tile view: element (r, c) of the lower triangle at base + r*rs + c*cs (tile-local).
struct Tile {
    size_t base;
    long rs, cs;
};
__device__ inline Tile tile_at(const Part &P, int r0, int c0) {
    const size_t s = (P.n & 1) ? P.n : P.n + 1;
    Tile t;
    if (r0 < P.k) {  // column-major region
        t.base = (size_t)(P.k + 1 + r0) + (size_t)c0 * s;
        t.rs = 1;
        t.cs = (long)s;
    } else {  // row-major region
        t.base = (size_t)(r0 - P.k) * s + c0;
        t.rs = (long)s;
        t.cs = 1;
    }
    return t;
}

// Load a (rn x cn) tile into LDS dst[r*TP + c]; coalesced along the contiguous axis.
template <typename T, bool LOWER_ONLY>
__device__ inline void load_tile(const T *__restrict__ a, const Tile &t, int rn, int cn, T *dst) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        T v = T(0);
        if (r < rn && c < cn && (!LOWER_ONLY || r >= c)) v = a[t.base + r * t.rs + c * t.cs];
        dst[r * TP + c] = v;
    }
}
template <typename T, bool LOWER_ONLY>
__device__ inline void store_tile(T *__restrict__ a, const Tile &t, int rn, int cn, const T *src) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        if (r < rn && c < cn && (!LOWER_ONLY || r >= c)) a[t.base + r * t.rs + c * t.cs] = src[r * TP + c];
    }
}

template <typename T>
__global__ void add_diag(T *a, int n, double l2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) a[rfp_lower((size_t)n, j, j)] += T(l2);
}

// POTRF of diagonal tile J plus its inverse (lower) into inv[NB*NB].
template <typename T>
__global__ __launch_bounds__(256) void potrf_diag(T *a, Part P, int J, T *inv, int *info) {
    __shared__ T L[NB * TP];
    __shared__ T Li[NB * TP];
    __shared__ int fail;
    if (*info) return;
    const int s0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, s0, s0);
    load_tile<T, true>(a, t, w, w, L);
    if (threadIdx.x == 0) fail = 0;
    __syncthreads();
    for (int j = 0; j < w; ++j) {
        const T piv = L[j * TP + j];
        if (!(piv > T(0))) {  // also catches NaN
            if (threadIdx.x == 0) fail = s0 + j + 1;
            break;
        }
        const T d = sqrt(piv);
        __syncthreads();
        for (int i = j + 1 + threadIdx.x; i < w; i += blockDim.x) L[i * TP + j] /= d;
        if (threadIdx.x == 0) L[j * TP + j] = d;
        __syncthreads();
        const int m = w - j - 1;
        for (int idx = threadIdx.x; idx < m * m; idx += blockDim.x) {
            const int i = j + 1 + idx / m, c = j + 1 + idx % m;
            if (c <= i) L[i * TP + c] -= L[i * TP + j] * L[c * TP + j];
        }
        __syncthreads();
    }
    __syncthreads();
    if (fail) {
        if (threadIdx.x == 0) atomicCAS(info, 0, fail);
        return;
    }
    // inverse of the lower factor: thread c owns column c
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) Li[idx / NB * TP + idx % NB] = T(0);
    __syncthreads();
    if (threadIdx.x < w) {
        const int c = threadIdx.x;
        for (int r = c; r < w; ++r) {
            T acc = (r == c) ? T(1) : T(0);
            for (int q = c; q < r; ++q) acc -= L[r * TP + q] * Li[q * TP + c];
            Li[r * TP + c] = acc / L[r * TP + r];
        }
    }
    __syncthreads();
    store_tile<T, true>(a, t, w, w, L);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) inv[idx] = Li[(idx / NB) * TP + idx % NB];
}

// Panel: A_IJ <- A_IJ * inv^T for every block row I > J (one WG per I).
template <typename T>
__global__ __launch_bounds__(256) void trsm_panel(T *a, Part P, int J, const T *inv, const int *info) {
    __shared__ T Pn[NB * TP];
    __shared__ T Iv[NB * TP];
    if (*info) return;
    const int I = J + 1 + blockIdx.x;
    const int r0 = P.start(I), rn = P.width(I), c0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, r0, c0);
    load_tile<T, false>(a, t, rn, w, Pn);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) Iv[(idx / NB) * TP + idx % NB] = inv[idx];
    __syncthreads();
    // out[r][c] = sum_q Pn[r][q] * Iv[c][q]   (q <= c)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    typename Vec4<T>::type acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = typename Vec4<T>::type{0, 0, 0, 0};
    for (int k0 = 0; k0 < NB; k0 += 4) {
        const int kk = k0 + (lane >> 4);
        T av[2], bv[2];
        for (int i = 0; i < 2; ++i) av[i] = Pn[(wr * 32 + i * 16 + (lane & 15)) * TP + kk];
        for (int j = 0; j < 2; ++j) bv[j] = Iv[(wc * 32 + j * 16 + (lane & 15)) * TP + kk];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 4; ++v)
                Pn[(wr * 32 + i * 16 + mfma16_row<T>(lane, v)) * TP + wc * 32 + j * 16 + (lane & 15)] = acc[i][j][v];
    __syncthreads();
    store_tile<T, false>(a, t, rn, w, Pn);
}

// Trailing update: A_IK -= A_IJ A_KJ^T for J < K <= I (one WG per lower tile pair).
template <typename T>
__global__ __launch_bounds__(256) void syrk_update(T *a, Part P, int J, const int *info) {
    __shared__ T Ai[NB * TP];
    __shared__ T Ak[NB * TP];
    __shared__ T C[NB * TP];
    if (*info) return;
    // decode lower-triangular pair index over the trailing blocks
    const int m = P.nblk - J - 1;
    const int t = blockIdx.x;
    int ii = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while ((long)ii * (ii + 1) / 2 > t) --ii;
    while ((long)(ii + 1) * (ii + 2) / 2 <= t) ++ii;
    const int kk_ = t - ii * (ii + 1) / 2;
    if (ii >= m) return;
    const int I = J + 1 + ii, K = J + 1 + kk_;
    const int c0 = P.start(J), w = P.width(J);
    const int ri = P.start(I), rni = P.width(I), rk = P.start(K), rnk = P.width(K);
    load_tile<T, false>(a, tile_at(P, ri, c0), rni, w, Ai);
    load_tile<T, false>(a, tile_at(P, rk, c0), rnk, w, Ak);
    const bool diagt = (I == K);
    const Tile tc = tile_at(P, ri, rk);
    if (diagt) load_tile<T, true>(a, tc, rni, rnk, C); else load_tile<T, false>(a, tc, rni, rnk, C);
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    typename Vec4<T>::type acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = typename Vec4<T>::type{0, 0, 0, 0};
    for (int k0 = 0; k0 < NB; k0 += 4) {
        const int kk = k0 + (lane >> 4);
        T av[2], bv[2];
        for (int i = 0; i < 2; ++i) av[i] = Ai[(wr * 32 + i * 16 + (lane & 15)) * TP + kk];
        for (int j = 0; j < 2; ++j) bv[j] = Ak[(wc * 32 + j * 16 + (lane & 15)) * TP + kk];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(av[i], bv[j], acc[i][j]);
    }
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 4; ++v)
                C[(wr * 32 + i * 16 + mfma16_row<T>(lane, v)) * TP + wc * 32 + j * 16 + (lane & 15)] -= acc[i][j][v];
    __syncthreads();
    if (diagt) store_tile<T, true>(a, tc, rni, rnk, C); else store_tile<T, false>(a, tc, rni, rnk, C);
}

// ---------------------------------------------------------------------------
// Triangular solves, FP64 accumulation.  b is (nrhs, n) row-major.
// ---------------------------------------------------------------------------
template <typename T, bool TRANS>
__global__ __launch_bounds__(64) void diag_solve(const T *a, Part P, int J, double *b, int n, int nrhs) {
    __shared__ double L[NB * TP];
    __shared__ double y[NB];
    const int s0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, s0, s0);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        double v = 0.0;
        if (r < w && c < w && r >= c) v = (double)a[t.base + r * t.rs + c * t.cs];
        L[r * TP + c] = v;
    }
    __syncthreads();
    const int lane = threadIdx.x;
    for (int q = 0; q < nrhs; ++q) {
        double *bq = b + (size_t)q * n + s0;
        if (lane < w) y[lane] = bq[lane];
        __syncthreads();
        if (!TRANS) {  // L y = b, forward
            for (int j = 0; j < w; ++j) {
                double part = lane < j ? L[j * TP + lane] * y[lane] : 0.0;
                for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
                if (lane == 0) y[j] = (y[j] - part) / L[j * TP + j];
                __syncthreads();
            }
        } else {  // L^T x = y, backward
            for (int j = w - 1; j >= 0; --j) {
                double part = (lane > j && lane < w) ? L[lane * TP + j] * y[lane] : 0.0;
                for (int off = 32; off > 0; off >>= 1) part += __shfl_xor(part, off);
                if (lane == 0) y[j] = (y[j] - part) / L[j * TP + j];
                __syncthreads();
            }
        }
        if (lane < w) bq[lane] = y[lane];
        __syncthreads();
    }
}

// forward update: b_I -= L_IJ y_J for blocks I > J (one WG per I, 64 threads = rows)
template <typename T>
__global__ __launch_bounds__(256) void fwd_update(const T *a, Part P, int J, double *b, int n, int nrhs) {
    __shared__ double yj[NB];
    __shared__ double Lt[NB * TP];
    const int I = J + 1 + blockIdx.x;
    const int r0 = P.start(I), rn = P.width(I), c0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, r0, c0);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        Lt[r * TP + c] = (r < rn && c < w) ? (double)a[t.base + r * t.rs + c * t.cs] : 0.0;
    }
    for (int q = 0; q < nrhs; ++q) {
        __syncthreads();
        if (threadIdx.x < w) yj[threadIdx.x] = b[(size_t)q * n + c0 + threadIdx.x];
        __syncthreads();
        // 4 threads per row, each a quarter of the columns
        const int r = threadIdx.x >> 2, qq = threadIdx.x & 3;
        double s = 0.0;
        for (int c = qq; c < w; c += 4) s += Lt[r * TP + c] * yj[c];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (qq == 0 && r < rn) b[(size_t)q * n + r0 + r] -= s;
    }
}

// backward update: x_c -= sum_{r in J} L[r][c] x_r for columns c in block I < J
template <typename T>
__global__ __launch_bounds__(256) void bwd_update(const T *a, Part P, int J, double *b, int n, int nrhs) {
    __shared__ double xj[NB];
    __shared__ double Lt[NB * TP];
    const int I = blockIdx.x;  // I < J
    const int r0 = P.start(J), w = P.width(J), c0 = P.start(I), cn = P.width(I);
    const Tile t = tile_at(P, r0, c0);
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        Lt[r * TP + c] = (r < w && c < cn) ? (double)a[t.base + r * t.rs + c * t.cs] : 0.0;
    }
    for (int q = 0; q < nrhs; ++q) {
        __syncthreads();
        if (threadIdx.x < w) xj[threadIdx.x] = b[(size_t)q * n + r0 + threadIdx.x];
        __syncthreads();
        const int c = threadIdx.x >> 2, qq = threadIdx.x & 3;
        double s = 0.0;
        for (int r = qq; r < w; r += 4) s += Lt[r * TP + c] * xj[r];
        s += __shfl_xor(s, 1);
        s += __shfl_xor(s, 2);
        if (qq == 0 && c < cn) b[(size_t)q * n + c0 + c] -= s;
    }
}

template <typename T>
int potrf_impl(T *a, int n, double l2, hipStream_t stream) {
    KF_REQUIRE(a && n > 0, KF_ERR_ARG, "n must be > 0");
    const Part P = make_part(n);
    char *ws = static_cast<char *>(workspace(WS_AUX2, sizeof(T) * NB * NB + 256, stream));
    KF_REQUIRE(ws, KF_ERR_NOMEM, "workspace allocation failed");
    int *info = reinterpret_cast<int *>(ws);
    T *inv = reinterpret_cast<T *>(ws + 256);
    KF_HIP_TRY(hipMemsetAsync(info, 0, sizeof(int), stream));
    ScopedTimer timer(TM_POTRF, stream);
    if (l2 != 0.0) {
        hipLaunchKernelGGL(add_diag<T>, dim3((n + 255) / 256), dim3(256), 0, stream, a, n, l2);
    }
    for (int J = 0; J < P.nblk; ++J) {
        hipLaunchKernelGGL(potrf_diag<T>, dim3(1), dim3(256), 0, stream, a, P, J, inv, info);
        const int m = P.nblk - J - 1;
        if (m > 0) {
            hipLaunchKernelGGL(trsm_panel<T>, dim3(m), dim3(256), 0, stream, a, P, J, inv, info);
            hipLaunchKernelGGL(syrk_update<T>, dim3(m * (m + 1) / 2), dim3(256), 0, stream, a, P,
                               J, info);
        }
    }
    KF_HIP_TRY(hipGetLastError());
    timer.stop();
    int h = 0;
    KF_HIP_TRY(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    return h;
}

template <typename T>
int potrs_impl(const T *a, int n, double *b, int nrhs, hipStream_t stream) {
    KF_REQUIRE(a && b && n > 0 && nrhs >= 1, KF_ERR_ARG, "bad potrs arguments");
    const Part P = make_part(n);
    ScopedTimer timer(TM_POTRS, stream);
    for (int J = 0; J < P.nblk; ++J) {
        hipLaunchKernelGGL((diag_solve<T, false>), dim3(1), dim3(64), 0, stream, a, P, J, b, n, nrhs);
        const int m = P.nblk - J - 1;
        if (m > 0)
            hipLaunchKernelGGL(fwd_update<T>, dim3(m), dim3(256), 0, stream, a, P, J, b, n, nrhs);
    }
    for (int J = P.nblk - 1; J >= 0; --J) {
        hipLaunchKernelGGL((diag_solve<T, true>), dim3(1), dim3(64), 0, stream, a, P, J, b, n, nrhs);
        if (J > 0)
            hipLaunchKernelGGL(bwd_update<T>, dim3(J), dim3(256), 0, stream, a, P, J, b, n, nrhs);
    }
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_rfp_potrf(int precision, void *arf, int n, double l2, void *stream) {
    using namespace kf;
    if (precision == KF_PREC_F64) return potrf_impl(static_cast<double *>(arf), n, l2, as_stream(stream));
    if (precision == KF_PREC_F32) return potrf_impl(static_cast<float *>(arf), n, l2, as_stream(stream));
    set_last_error("precision %d not supported by kf_rfp_potrf", precision);
    return KF_ERR_ARG;
}

extern "C" int kf_rfp_potrs(int precision, const void *arf, int n, double *b, int nrhs, void *stream) {
    using namespace kf;
    if (precision == KF_PREC_F64)
        return potrs_impl(static_cast<const double *>(arf), n, b, nrhs, as_stream(stream));
    if (precision == KF_PREC_F32)
        return potrs_impl(static_cast<const float *>(arf), n, b, nrhs, as_stream(stream));
    set_last_error("precision %d not supported by kf_rfp_potrs", precision);
    return KF_ERR_ARG;
}
