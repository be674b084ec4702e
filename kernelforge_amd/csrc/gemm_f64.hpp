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
    auto load = [&](int k0) {
        const int gm = m0 + lr;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int gk = k0 + lk + u;
            ra[u] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : 0.0;
        }
        if (TRANS_B) {
            const int gn = n0 + lr;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gk = k0 + lk + u;
                rb[u] = (gn < N && gk < K) ? B[(size_t)gn * ldb + gk] : 0.0;
            }
        } else {
            const int gk = k0 + bk;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int gn = n0 + bn + u;
                rb[u] = (gn < N && gk < K) ? B[(size_t)gk * ldb + gn] : 0.0;
            }
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int u = 0; u < 4; ++u) As[lk + u][lr] = ra[u];
        if (TRANS_B) {
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[lk + u][lr] = rb[u];
        } else {
#pragma unroll
            for (int u = 0; u < 4; ++u) Bs[bk][bn + u] = rb[u];
        }
    };
    v4d_t c[2][2];
#pragma unroll
    for (int x = 0; x < 2; ++x)
#pragma unroll
        for (int y = 0; y < 2; ++y) c[x][y] = (v4d_t){0.0, 0.0, 0.0, 0.0};
    load(0);
    for (int k0 = 0; k0 < K; k0 += KC) {
        __syncthreads();
        store();
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
