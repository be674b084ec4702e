// rfp_convert.hip -- full <-> RFP conversions on the device (LAPACK dtrttf / dtfttr
// semantics, all four TRANSR/UPLO variants), used by kernelmath.full_to_rfp /
// rfp_to_full (reference src/math.cpp:173-211, src/math_bindings.cpp:85-165).
//
// Position of element (i, j) of the stored triangle of an n x n column-major matrix
// (UPLO='U': i <= j, UPLO='L': i >= j), k = n / 2:
//   n odd : U: j >= k -> (i, j-k)      else (j+k+1, i)
//           L: j <= k -> (i, j)        else (j-k-1, i-k)
//           TRANSR='N': r + c*n          TRANSR='T': c + r*(k+1)
//   n even: U: j >= k -> (i, j-k)      else (j+k+1, i)
//           L: j <  k -> (i+1, j)      else (j-k, i-k)
//           TRANSR='N': r + c*(n+1)      TRANSR='T': c + r*k
#include "kf_common.hpp"

namespace kf {

__host__ __device__ inline size_t rfp_pos(int n, bool trans, bool upper, int i, int j) {
    const int k = n / 2;
    int r, c;
    if (n & 1) {
        if (upper) {
            if (j >= k) { r = i; c = j - k; } else { r = j + k + 1; c = i; }
        } else {
            if (j <= k) { r = i; c = j; } else { r = j - k - 1; c = i - k; }
        }
        return trans ? (size_t)c + (size_t)r * (k + 1) : (size_t)r + (size_t)c * n;
    }
    if (upper) {
        if (j >= k) { r = i; c = j - k; } else { r = j + k + 1; c = i; }
    } else {
        if (j < k) { r = i + 1; c = j; } else { r = j - k; c = i - k; }
    }
    return trans ? (size_t)c + (size_t)r * k : (size_t)r + (size_t)c * (n + 1);
}

namespace {

__global__ void trttf_kernel(int n, bool trans, bool upper, const double *A, int lda, double *arf) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const int i = (int)(t % n), j = (int)(t / n);  // column-major walk: coalesced reads
    if (upper ? i > j : i < j) return;
    arf[rfp_pos(n, trans, upper, i, j)] = A[i + (size_t)j * lda];
}

__global__ void tfttr_kernel(int n, bool trans, bool upper, const double *arf, double *A, int lda,
                             int symmetrize) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const int i = (int)(t % n), j = (int)(t / n);
    const bool in_tri = upper ? i <= j : i >= j;
    if (in_tri) {
        A[i + (size_t)j * lda] = arf[rfp_pos(n, trans, upper, i, j)];
    } else if (symmetrize) {
        A[i + (size_t)j * lda] = arf[rfp_pos(n, trans, upper, j, i)];
    }
}

inline int flags(char transr, char uplo, bool &trans, bool &upper) {
    KF_REQUIRE(transr == 'N' || transr == 'n' || transr == 'T' || transr == 't', KF_ERR_ARG,
               "transr must be 'N' or 'T'");
    KF_REQUIRE(uplo == 'U' || uplo == 'u' || uplo == 'L' || uplo == 'l', KF_ERR_ARG,
               "uplo must be 'U' or 'L'");
    trans = transr == 'T' || transr == 't';
    upper = uplo == 'U' || uplo == 'u';
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_trttf(char transr, char uplo, int n, const double *A, int lda, double *arf,
                        void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    bool trans, upper;
    int rc = flags(transr, uplo, trans, upper);
    if (rc) return rc;
    KF_REQUIRE(n >= 0 && lda >= (n > 1 ? n : 1), KF_ERR_ARG, "bad n / lda");
    if (n == 0) return KF_OK;
    KF_REQUIRE(A && arf, KF_ERR_ARG, "null pointer");
    const size_t tot = (size_t)n * n;
    hipLaunchKernelGGL(trttf_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       as_stream(stream), n, trans, upper, A, lda, arf);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_tfttr(char transr, char uplo, int n, const double *arf, double *A, int lda,
                        int symmetrize, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    bool trans, upper;
    int rc = flags(transr, uplo, trans, upper);
    if (rc) return rc;
    KF_REQUIRE(n >= 0 && lda >= (n > 1 ? n : 1), KF_ERR_ARG, "bad n / lda");
    if (n == 0) return KF_OK;
    KF_REQUIRE(A && arf, KF_ERR_ARG, "null pointer");
    const size_t tot = (size_t)n * n;
    hipLaunchKernelGGL(tfttr_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0,
                       as_stream(stream), n, trans, upper, arf, A, lda, symmetrize);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}
