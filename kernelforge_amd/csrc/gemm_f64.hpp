// gemm_f64.hpp -- FP64 MFMA GEMM tile kernel for gfx950 (v_mfma_f64_16x16x4).
//
//   C (M x N, ldc) = alpha * A (M x K, lda) * op(B) + beta * C,   all row-major
//   op(B) = B^T with B stored N x K (TRANS_B = true), or B stored K x N (false).
//
// 64 x 64 output tile per 256-thread work-group (2 x 2 waves of 32 x 32, each a
// 2 x 2 grid of 16 x 16 MFMA blocks), K staged through LDS in chunks of 16 with the
// next chunk prefetched into registers while the current one is multiplied.
#pragma once

#include "kf_common.hpp"

namespace kf {
namespace gemm {

typedef double v4d_t __attribute__((ext_vector_type(4)));

constexpr int TILE = 64;
constexpr int KC = 16;
constexpr int LDP = TILE + 1;  // LDS row pitch (k-major tiles), breaks bank aliasing

template <bool TRANS_B>
__global__ __launch_bounds__(256) void gemm_f64_kernel(int M, int N, int K, double alpha,
                                                       const double *__restrict__ A, size_t lda,
                                                       const double *__restrict__ B, size_t ldb,
                                                       double beta, double *__restrict__ C,
                                                       size_t ldc) {
    __shared__ double As[KC][LDP];
    __shared__ double Bs[KC][LDP];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int wr = wv >> 1, wc = wv & 1;
    const int m0 = blockIdx.y * TILE, n0 = blockIdx.x * TILE;
    // loader coordinates: A (and B^T) tiles are 64 rows x 16 k; thread -> (row, 4 k)
    const int lr = tid >> 2, lk = (tid & 3) * 4;
    // B (K x N) tiles are 16 k x 64 columns; thread -> (k, 4 columns)
    const int bk = tid >> 4, bn = (tid & 15) * 4;
    double ra[4], rb[4];
    // Loads are unconditional, of clamped (valid) elements, into registers; the padding
    // is applied when the chunk is stored to LDS, by a select on a condition formed with
    // non-short-circuit '&'.  A guarded load (or a '&&' condition, whose first comparison
    // becomes a branch the load is sunk into) is waited for before the next one is
    // issued, and a select right after the loads waits for them before the products of
    // the current chunk.
    const int am = min(m0 + lr, M - 1);
    const bool okm = m0 + lr < M;
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) ra[u] = A[(size_t)am * lda + min(k0 + lk + u, K - 1)];
        if (TRANS_B) {
            const size_t brow = (size_t)min(n0 + lr, N - 1) * ldb;
#pragma unroll
            for (int u = 0; u < 4; ++u) rb[u] = B[brow + min(k0 + lk + u, K - 1)];
        } else {
            const size_t brow = (size_t)min(k0 + bk, K - 1) * ldb;
#pragma unroll
            for (int u = 0; u < 4; ++u) rb[u] = B[brow + min(n0 + bn + u, N - 1)];
        }
    };
    auto store = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) As[lk + u][lr] = (okm & (k0 + lk + u < K)) ? ra[u] : 0.0;
        if (TRANS_B) {
            const bool okn = n0 + lr < N;
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[lk + u][lr] = (okn & (k0 + lk + u < K)) ? rb[u] : 0.0;
        } else {
            const bool okk = k0 + bk < K;
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[bk][bn + u] = ((n0 + bn + u < N) & okk) ? rb[u] : 0.0;
        }
    };
    v4d_t c[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) c[x][y] = (v4d_t){0.0, 0.0, 0.0, 0.0};
    if (K > 0) load(0);  // the clamped addresses need K >= 1 (M, N >= 1 by the launcher)
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();
        store(k0);
        __syncthreads();
        if (k0 + KC < K) load(k0 + KC);
#pragma unroll
        for (int kk = 0; kk < KC; kk += 4) {
            const int kr = kk + (lane >> 4), col = lane & 15;
            const double a0 = As[kr][wr * 32 + col], a1 = As[kr][wr * 32 + 16 + col];
            const double b0 = Bs[kr][wc * 32 + col], b1 = Bs[kr][wc * 32 + 16 + col];
            c[0][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b0, c[0][0], 0, 0, 0);
            c[0][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a0, b1, c[0][1], 0, 0, 0);
            c[1][0] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b0, c[1][0], 0, 0, 0);
            c[1][1] = __builtin_amdgcn_mfma_f64_16x16x4f64(a1, b1, c[1][1], 0, 0, 0);
        }
    }
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int row = m0 + wr * 32 + x * 16 + mfma16_row_f64(lane, v);
                const int col = n0 + wc * 32 + y * 16 + (lane & 15);
                if (row < M && col < N) {
                    double *cp = C + (size_t)row * ldc + col;
                    const double r = alpha * c[x][y][v];
                    *cp = beta == 0.0 ? r : r + beta * *cp;
                }
            }
}

// Host launcher on `stream`; M, N, K may be anything >= 0.
template <bool TRANS_B>
inline int gemm_f64(int M, int N, int K, double alpha, const double *A, size_t lda,
                    const double *B, size_t ldb, double beta, double *C, size_t ldc,
                    hipStream_t stream) {
    if (M <= 0 || N <= 0) return KF_OK;
    dim3 grid((N + TILE - 1) / TILE, (M + TILE - 1) / TILE);
    hipLaunchKernelGGL(gemm_f64_kernel<TRANS_B>, grid, dim3(256), 0, stream, M, N, K, alpha, A,
                       lda, B, ldb, beta, C, ldc);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace gemm
}  // namespace kf
