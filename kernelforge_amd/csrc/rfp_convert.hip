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

// Symmetric matrix in one RFP variant -> another (elements (i, j), i <= j, of the
// column-major upper triangle; the lower-triangle variants store (j, i)).
template <typename Tin, typename Tout>
__global__ void relayout_kernel(int n, bool tin, bool uin, bool tout, bool uout,
                                const Tin *__restrict__ in, Tout *__restrict__ out) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)n * n) return;
    const int i = (int)(t % n), j = (int)(t / n);
    if (i > j) return;
    const size_t src = uin ? rfp_pos(n, tin, true, i, j) : rfp_pos(n, tin, false, j, i);
    const size_t dst = uout ? rfp_pos(n, tout, true, i, j) : rfp_pos(n, tout, false, j, i);
    out[dst] = (Tout)in[src];
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

// Symmetric RFP re-layout: a matrix stored as RFP (transr_in, uplo_in) rewritten as
// (transr_out, uplo_out), optionally narrowed from FP64 to FP32 (out_precision
// KF_PREC_F32); in and out must not overlap.  Used to hand the kernels' native packing
// (TRANSR='N', UPLO='U', the reference CPU kernels' own) to callers of the reference's
// GPU kernels, which pack TRANSR='N', UPLO='L' (src/cuda_local_kernels.hpp:52).
extern "C" int kf_rfp_relayout(char transr_in, char uplo_in, char transr_out, char uplo_out,
                               int n, const double *in, int out_precision, void *out,
                               void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    bool tin, uin, tout, uout;
    int rc = flags(transr_in, uplo_in, tin, uin);
    if (rc) return rc;
    if ((rc = flags(transr_out, uplo_out, tout, uout))) return rc;
    KF_REQUIRE(out_precision == KF_PREC_F64 || out_precision == KF_PREC_F32, KF_ERR_ARG,
               "out_precision must be F64 or F32");
    KF_REQUIRE(n >= 0, KF_ERR_ARG, "n must be >= 0");
    if (n == 0) return KF_OK;
    KF_REQUIRE(in && out && in != out, KF_ERR_ARG, "null or aliased pointer");
    const size_t tot = (size_t)n * n;
    const dim3 grid((unsigned)((tot + 255) / 256));
    if (out_precision == KF_PREC_F64)
        hipLaunchKernelGGL((relayout_kernel<double, double>), grid, dim3(256), 0, as_stream(stream), n,
                           tin, uin, tout, uout, in, static_cast<double *>(out));
    else
        hipLaunchKernelGGL((relayout_kernel<double, float>), grid, dim3(256), 0, as_stream(stream), n,
                           tin, uin, tout, uout, in, static_cast<float *>(out));
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}
