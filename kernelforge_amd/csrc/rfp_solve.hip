// rfp_solve.hip -- matrix-vector products on RFP matrices and the refined mixed-precision
// KRR solve on gfx950.
//
// The reference solves (K + l2 I) alpha = y with LAPACK dpftrf/dpftrs in FP64
// (src/math.cpp:134-171).  The mixed path factors K in FP32 and recovers FP64-level
// residuals by preconditioned conjugate gradients: the FP32 RFP Cholesky factor is the
// preconditioner, every matrix-vector product and vector update is FP64.  When
// kappa(K) * eps_fp32 ~ 1 (lambda = 1e-8 on a 20k kernel), stationary refinement
// stalls but PCG with this preconditioner still converges in a few iterations.
//
// RFP symmetric matvec: one workgroup per 64-row block I computes
//   y_I = sum_K A_IK x_K (+ lambda x_I)
// reading lower tiles (I, K <= I) directly and upper ones as (K, I)^T, so every
// output has one writer and a fixed summation order (deterministic).
#include <cmath>

#include "kf_common.hpp"

namespace kf {
namespace {

constexpr int NB = 64;

struct SPart {
    int n, k, nlo, nblk;
    __host__ __device__ int start(int J) const { return J < nlo ? J * NB : k + (J - nlo) * NB; }
    __host__ __device__ int width(int J) const {
        const int s = start(J);
        const int end = J < nlo ? k : n;
        return (end - s) < NB ? (end - s) : NB;
    }
};
inline SPart make_spart(int n) {
    SPart p;
    p.n = n;
    p.k = n / 2;
    p.nlo = (p.k + NB - 1) / NB;
    p.nblk = p.nlo + (n - p.k + NB - 1) / NB;
    return p;
}

// y = (A + lambda I) x, A symmetric RFP (N, U), FP64 accumulation.  Each 64x64 block
// is staged through LDS, read along its contiguous axis (coalesced) and transposed
// in LDS when it lies above the diagonal.
template <typename T>
__global__ __launch_bounds__(256) void rfp_symv(const T *__restrict__ a, SPart P, double lambda,
                                                const double *__restrict__ x,
                                                double *__restrict__ y) {
    __shared__ double xs[NB];
    __shared__ double tile[NB * (NB + 1)];
    __shared__ double part[4][NB];
    const int I = blockIdx.x;
    const int r0 = P.start(I), rn = P.width(I);
    const size_t s = (P.n & 1) ? P.n : P.n + 1;
    const int tid = threadIdx.x, row = tid & 63, qq = tid >> 6;
    double acc = 0.0;
    for (int K = 0; K < P.nblk; ++K) {
        const int c0 = P.start(K), cn = P.width(K);
        // stored lower block (hi, lo) = (I, K) or (K, I)
        const bool upper = K > I;
        const int hr0 = upper ? c0 : r0, hrn = upper ? cn : rn;
        const int lc0 = upper ? r0 : c0, lcn = upper ? rn : cn;
        const bool colmajor = hr0 < P.k;
        const size_t base = colmajor ? (size_t)(P.k + 1 + hr0) + (size_t)lc0 * s
                                     : (size_t)(hr0 - P.k) * s + lc0;
        const size_t rs = colmajor ? 1 : s, cs = colmajor ? s : 1;
        __syncthreads();
        if (tid < NB) xs[tid] = tid < cn ? x[c0 + tid] : 0.0;
        for (int idx = tid; idx < NB * NB; idx += 256) {
            int rr, cc;  // element (rr, cc) of the stored block
            if (colmajor) { rr = idx % NB; cc = idx / NB; } else { cc = idx % NB; rr = idx / NB; }
            double v = 0.0;
            if (rr < hrn && cc < lcn && (I != K || rr >= cc)) v = (double)a[base + rr * rs + cc * cs];
            if (upper) tile[cc * (NB + 1) + rr] = v;   // A_IK = (stored)^T
            else tile[rr * (NB + 1) + cc] = v;
        }
        __syncthreads();
        if (I == K) {  // complete the diagonal block from its lower half
            for (int idx = tid; idx < NB * NB; idx += 256) {
                const int rr = idx / NB, cc = idx % NB;
                if (cc > rr) tile[rr * (NB + 1) + cc] = tile[cc * (NB + 1) + rr];
            }
            __syncthreads();
        }
        for (int c = qq; c < cn; c += 4) acc += tile[row * (NB + 1) + c] * xs[c];
    }
    part[qq][row] = acc;
    __syncthreads();
    if (tid < rn) {
        const double v = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        y[r0 + tid] = v + lambda * x[r0 + tid];
    }
}

// Deterministic dot products: per-block partials then one block folds them in order.
__global__ __launch_bounds__(256) void dot_partial(const double *__restrict__ a,
                                                   const double *__restrict__ b, int n,
                                                   double *__restrict__ part) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) s += a[i] * b[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}
__global__ __launch_bounds__(256) void dot_final(const double *__restrict__ part, int np,
                                                 double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += 256) s += part[i];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}

// y = a*x + b*y
__global__ void axpby(int n, double a, const double *__restrict__ x, double b,
                      double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = a * x[i] + b * y[i];
}
// r = y - q
__global__ void sub_into(int n, const double *__restrict__ y, const double *__restrict__ q,
                         double *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) r[i] = y[i] - q[i];
}

constexpr int kDotBlocks = 240;

int dot(const double *a, const double *b, int n, double *scratch, double *hres,
        hipStream_t stream) {
    hipLaunchKernelGGL(dot_partial, dim3(kDotBlocks), dim3(256), 0, stream, a, b, n, scratch);
    hipLaunchKernelGGL(dot_final, dim3(1), dim3(256), 0, stream, scratch, kDotBlocks,
                       scratch + kDotBlocks);
    KF_HIP_TRY(hipMemcpyAsync(hres, scratch + kDotBlocks, sizeof(double), hipMemcpyDeviceToHost,
                              stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    return KF_OK;
}

template <typename T>
int symv(const T *a, int n, double lambda, const double *x, double *y, hipStream_t stream) {
    const SPart P = make_spart(n);
    hipLaunchKernelGGL(rfp_symv<T>, dim3(P.nblk), dim3(256), 0, stream, a, P, lambda, x, y);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_rfp_symv(int precision, const void *arf, int n, double lambda, const double *x,
                           double *y, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    KF_REQUIRE(arf && x && y && n > 0, KF_ERR_ARG, "bad kf_rfp_symv arguments");
    if (precision == KF_PREC_F64)
        return symv(static_cast<const double *>(arf), n, lambda, x, y, as_stream(stream));
    if (precision == KF_PREC_F32)
        return symv(static_cast<const float *>(arf), n, lambda, x, y, as_stream(stream));
    set_last_error("precision %d not supported by kf_rfp_symv", precision);
    return KF_ERR_ARG;
}

// Preconditioned CG on (A + lambda I) alpha = y, A an RFP matrix (F64 or F32), with the
// preconditioner M = L L^T given by an RFP Cholesky factor (F64 or F32) of (A + mu I).
extern "C" int kf_rfp_pcg(int a_precision, const void *arf, int m_precision, const void *lfac,
                          int n, double lambda, const double *y, double *alpha, double tol,
                          int max_iter, int *iters, double *rel_residual, void *stream_) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream_));
    using namespace kf;
    hipStream_t stream = as_stream(stream_);
    KF_REQUIRE(arf && lfac && y && alpha && n > 0 && max_iter >= 0, KF_ERR_ARG,
               "bad kf_rfp_pcg arguments");
    KF_REQUIRE(a_precision == KF_PREC_F64 || a_precision == KF_PREC_F32, KF_ERR_ARG,
               "a_precision must be F64 or F32");
    KF_REQUIRE(m_precision == KF_PREC_F64 || m_precision == KF_PREC_F32, KF_ERR_ARG,
               "m_precision must be F64 or F32");
    double *ws = static_cast<double *>(
        workspace(WS_AUX0, sizeof(double) * ((size_t)4 * n + 2 * kDotBlocks + 16), stream));
    KF_REQUIRE(ws, KF_ERR_NOMEM, "workspace allocation failed");
    double *r = ws, *z = ws + n, *p = ws + 2 * (size_t)n, *q = ws + 3 * (size_t)n;
    double *scratch = ws + 4 * (size_t)n;
    auto A = [&](const double *xin, double *yout) {
        return a_precision == KF_PREC_F64
                   ? kf_rfp_symv(KF_PREC_F64, arf, n, lambda, xin, yout, stream)
                   : kf_rfp_symv(KF_PREC_F32, arf, n, lambda, xin, yout, stream);
    };
    auto M = [&](double *v) { return kf_rfp_potrs(m_precision, lfac, n, v, 1, stream); };
    const int nb = (n + 255) / 256;
    double yy = 0.0, rz = 0.0, rr = 0.0, pq = 0.0;
    int rc;
    ScopedTimer timer(TM_MATVEC, stream);
    if ((rc = dot(y, y, n, scratch, &yy, stream))) return rc;
    const double ynorm = sqrt(yy);
    if (ynorm == 0.0) {
        KF_HIP_TRY(hipMemsetAsync(alpha, 0, sizeof(double) * n, stream));
        if (iters) *iters = 0;
        if (rel_residual) *rel_residual = 0.0;
        return KF_OK;
    }
    // r = y - A alpha
    if ((rc = A(alpha, q))) return rc;
    hipLaunchKernelGGL(sub_into, dim3(nb), dim3(256), 0, stream, n, y, q, r);
    if ((rc = dot(r, r, n, scratch, &rr, stream))) return rc;
    int it = 0;
    double rel = sqrt(rr) / ynorm;
    if (rel > tol) {
        KF_HIP_TRY(hipMemcpyAsync(z, r, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
        if ((rc = M(z))) return rc;
        KF_HIP_TRY(hipMemcpyAsync(p, z, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
        if ((rc = dot(r, z, n, scratch, &rz, stream))) return rc;
        while (it < max_iter) {
            ++it;
            if ((rc = A(p, q))) return rc;
            if ((rc = dot(p, q, n, scratch, &pq, stream))) return rc;
            if (!(pq > 0.0)) break;  // breakdown (operator not SPD to working precision)
            const double step = rz / pq;
            hipLaunchKernelGGL(axpby, dim3(nb), dim3(256), 0, stream, n, step, p, 1.0, alpha);
            hipLaunchKernelGGL(axpby, dim3(nb), dim3(256), 0, stream, n, -step, q, 1.0, r);
            if (it % 8 == 0) {  // replace the recurrence residual by the true one
                if ((rc = A(alpha, q))) return rc;
                hipLaunchKernelGGL(sub_into, dim3(nb), dim3(256), 0, stream, n, y, q, r);
            }
            if ((rc = dot(r, r, n, scratch, &rr, stream))) return rc;
            rel = sqrt(rr) / ynorm;
            if (rel <= tol) break;
            KF_HIP_TRY(hipMemcpyAsync(z, r, sizeof(double) * n, hipMemcpyDeviceToDevice, stream));
            if ((rc = M(z))) return rc;
            double rz_new = 0.0;
            if ((rc = dot(r, z, n, scratch, &rz_new, stream))) return rc;
            const double beta = rz_new / rz;
            rz = rz_new;
            hipLaunchKernelGGL(axpby, dim3(nb), dim3(256), 0, stream, n, 1.0, z, beta, p);
        }
        // final true residual
        if ((rc = A(alpha, q))) return rc;
        hipLaunchKernelGGL(sub_into, dim3(nb), dim3(256), 0, stream, n, y, q, r);
        if ((rc = dot(r, r, n, scratch, &rr, stream))) return rc;
        rel = sqrt(rr) / ynorm;
    }
    KF_HIP_TRY(hipGetLastError());
    if (iters) *iters = it;
    if (rel_residual) *rel_residual = rel;
    if (rel > tol) {
        set_last_error("PCG stopped at relative residual %.3e after %d iterations", rel, it);
        return KF_ERR_NOCONV;
    }
    return KF_OK;
}

namespace kf {
namespace {
template <typename T>
__global__ void diag_sum(const T *__restrict__ a, int n, double *__restrict__ out) {
    __shared__ double sh[256];
    double s = 0.0;
    for (int j = threadIdx.x; j < n; j += 256) s += (double)a[rfp_lower((size_t)n, j, j)];
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = sh[0];
}
}  // namespace
}  // namespace kf

// Mixed-precision solve of (K + lambda I) alpha = y for an FP32 RFP kernel matrix K:
// factor K + mu I in FP32 into `lwork` (mu = lambda first; if the FP32 factorisation
// breaks down, mu grows by 10x from 1e-7 * mean(diag K) until it succeeds), then PCG
// with FP64 residuals on (K + lambda I) preconditioned by that factor.  The system
// solved is always the requested one: when PCG does not reach the tolerance (or breaks
// down because K + lambda I is not positive definite in FP32 storage) the call returns
// KF_ERR_NOCONV and the caller falls back to the FP64 factorisation.
extern "C" int kf_krr_solve_mixed(const float *K32, float *lwork, int n, double lambda,
                                  const double *y, double *alpha, double tol, int max_iter,
                                  int *iters, double *rel_residual, double *lambda_used,
                                  void *stream_) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream_));
    using namespace kf;
    hipStream_t stream = as_stream(stream_);
    KF_REQUIRE(K32 && lwork && y && alpha && n > 0, KF_ERR_ARG, "bad kf_krr_solve_mixed arguments");
    const size_t nt = (size_t)n * (n + 1) / 2;
    double *scr = static_cast<double *>(workspace(WS_AUX1, 64, stream));
    KF_REQUIRE(scr, KF_ERR_NOMEM, "workspace allocation failed");
    hipLaunchKernelGGL(diag_sum<float>, dim3(1), dim3(256), 0, stream, K32, n, scr);
    double dsum = 0.0;
    KF_HIP_TRY(hipMemcpyAsync(&dsum, scr, sizeof(double), hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    const double dmean = dsum / n;
    double mu = lambda;
    int info = 1;
    for (int attempt = 0; attempt < 12 && info != 0; ++attempt) {
        KF_HIP_TRY(hipMemcpyAsync(lwork, K32, nt * sizeof(float), hipMemcpyDeviceToDevice, stream));
        info = kf_rfp_potrf(KF_PREC_F32, lwork, n, mu, stream_);
        if (info < 0) return info;
        if (info > 0) mu = (attempt == 0 ? (1e-7 * dmean > lambda ? 1e-7 * dmean : 10 * lambda)
                                         : mu * 10.0);
    }
    if (info != 0) {
        set_last_error("FP32 factorisation failed for every shift up to %.3e", mu);
        return info;
    }
    KF_HIP_TRY(hipMemsetAsync(alpha, 0, sizeof(double) * n, stream));
    if (lambda_used) *lambda_used = mu;  // the shift of the preconditioner's factor
    return kf_rfp_pcg(KF_PREC_F32, K32, KF_PREC_F32, lwork, n, lambda, y, alpha, tol, max_iter,
                      iters, rel_residual, stream_);
}
