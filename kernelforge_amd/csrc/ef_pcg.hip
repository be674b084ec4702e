// ef_pcg.hip -- preconditioned CG for the energy+force kernel system
// (K_full + lambda I) alpha = [E; -F], with a block-Jacobi preconditioner whose
// per-molecule blocks can carry a learned diagonal correction, and the pieces that
// train that correction: the reverse-mode gradient of an unrolled PCG residual and
// a small MLP (BF16 weights, FP32 accumulation) with forward / backward / Adam kernels.
//
// Vector layout of the E+F system (the reference's full-kernel order,
// src/local_kernels.cpp:2511-2815): [E_0 .. E_{nm-1}, F_0 (3 N_0), F_1, ...].
// Block b gathers local index 0 -> E_b and 1 + r -> nm + offs_b + r, bs = 1 + 3 max_atoms.
// Preconditioner block: P_b = (K_bb + lambda I + diag(s_b))^-1 on the live 1 + 3 N_b
// rows, kept as an explicit inverse (identity on padded rows).
#include <algorithm>
#include <cmath>
#include <vector>

#include "feature_compact.hpp"
#include "kf_common.hpp"
#include "local_plan.hpp"

extern "C" int kf_rfp_symv(int precision, const void *arf, int n, double lambda, const double *x,
                           double *y, void *stream);
extern "C" int kf_local_alpha_desc(const double *dX, const int *Q, const int *N, int nm,
                                   int max_atoms, int rep, const double *alpha, double *out,
                                   void *stream);
extern "C" int kf_local_full_matvec(const double *X1, const double *dX1, const double *X2,
                                    const double *alpha_desc, const double *alpha_E,
                                    const int *Q1, const int *Q2, const int *N1, const int *N2,
                                    int nm1, int nm2, int max_atoms1, int max_atoms2, int rep,
                                    double sigma, double *E, double *F, void *stream);

namespace kf {
namespace {

constexpr int NT = 256;
constexpr int MAXBS = 97;  // 1 + 3 * 32

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

// global index of local position l of block b
__device__ inline size_t blk_index(int b, int l, int nm, const int *offs) {
    return l == 0 ? (size_t)b : (size_t)nm + offs[b] + (l - 1);
}

// ---------------------------------------------------------------------------
// Batched block factor: inv_b = (A_b + lambda I + diag(shift_b))^-1 on the live
// 1 + 3 N_b rows (Cholesky in LDS, triangular inverse, L^-T L^-1); identity on
// padded rows.  info[b] = 0, or the 1-based failing pivot.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(NT) void block_factor_kernel(const double *__restrict__ blocks,
                                                          int bs, const int *__restrict__ N,
                                                          int ma, double lambda,
                                                          const double *__restrict__ shift,
                                                          double *__restrict__ inv,
                                                          int *__restrict__ info) {
    __shared__ double L[MAXBS][MAXBS + 1];
    __shared__ double Li[MAXBS][MAXBS + 1];
    __shared__ int s_bad;
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n = 1 + 3 * clampi(N[b], 0, ma);
    const double *A = blocks + (size_t)b * bs * bs;
    for (int e = tid; e < n * n; e += NT) {
        const int r = e / n, c = e - r * n;
        double v = A[(size_t)r * bs + c];
        if (r == c) v += lambda + (shift ? shift[(size_t)b * bs + r] : 0.0);
        L[r][c] = v;
        Li[r][c] = 0.0;
    }
    if (tid == 0) s_bad = 0;
    __syncthreads();
    // right-looking Cholesky, lower triangle
    for (int j = 0; j < n; ++j) {
        if (tid == 0) {
            const double d = L[j][j];
            if (!(d > 0.0) && !s_bad) s_bad = j + 1;
            L[j][j] = d > 0.0 ? sqrt(d) : 1.0;
        }
        __syncthreads();
        const double djj = L[j][j];
        for (int r = j + 1 + tid; r < n; r += NT) L[r][j] /= djj;
        __syncthreads();
        const int m = n - j - 1;
        for (int e = tid; e < m * m; e += NT) {
            const int r = j + 1 + e / m, c = j + 1 + e % m;
            if (c <= r) L[r][c] -= L[r][j] * L[c][j];
        }
        __syncthreads();
    }
    // Li = L^-1 (lower), row by row: Li[r][c] = (delta - sum_{c<=q<r} L[r][q] Li[q][c]) / L[r][r]
    for (int r = 0; r < n; ++r) {
        for (int c = tid; c <= r; c += NT) {
            double s = (r == c) ? 1.0 : 0.0;
            for (int q = c; q < r; ++q) s -= L[r][q] * Li[q][c];
            Li[r][c] = s / L[r][r];
        }
        __syncthreads();
    }
    // inv = Li^T Li
    double *o = inv + (size_t)b * bs * bs;
    for (int e = tid; e < bs * bs; e += NT) {
        const int r = e / bs, c = e - r * bs;
        double v;
        if (r < n && c < n) {
            v = 0.0;
            for (int q = max(r, c); q < n; ++q) v += Li[q][r] * Li[q][c];
        } else {
            v = (r == c) ? 1.0 : 0.0;
        }
        o[e] = v;
    }
    if (tid == 0 && info) info[b] = s_bad;
}

// z = blockdiag(inv) r (one work-group per block; bs <= MAXBS)
__global__ __launch_bounds__(128) void block_apply_kernel(const double *__restrict__ inv, int bs,
                                                          const int *__restrict__ N, int ma,
                                                          const int *__restrict__ offs, int nm,
                                                          const double *__restrict__ r,
                                                          double *__restrict__ z) {
    __shared__ double rv[MAXBS];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n = 1 + 3 * clampi(N[b], 0, ma);
    for (int l = tid; l < n; l += blockDim.x) rv[l] = r[blk_index(b, l, nm, offs)];
    __syncthreads();
    const double *M = inv + (size_t)b * bs * bs;
    for (int l = tid; l < n; l += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s = fma(M[(size_t)l * bs + c], rv[c], s);
        z[blk_index(b, l, nm, offs)] = s;
    }
}

// grad[b, l] += -(inv_b v)_l * z_l  (d/d shift of v . P r with z = P r), per block
__global__ __launch_bounds__(128) void shift_grad_kernel(const double *__restrict__ inv, int bs,
                                                         const int *__restrict__ N, int ma,
                                                         const int *__restrict__ offs, int nm,
                                                         const double *__restrict__ v,
                                                         const double *__restrict__ z,
                                                         double *__restrict__ grad) {
    __shared__ double vv[MAXBS];
    const int b = blockIdx.x, tid = threadIdx.x;
    const int n = 1 + 3 * clampi(N[b], 0, ma);
    for (int l = tid; l < n; l += blockDim.x) vv[l] = v[blk_index(b, l, nm, offs)];
    __syncthreads();
    const double *M = inv + (size_t)b * bs * bs;
    for (int l = tid; l < n; l += blockDim.x) {
        double s = 0.0;
        for (int c = 0; c < n; ++c) s = fma(M[(size_t)l * bs + c], vv[c], s);
        grad[(size_t)b * bs + l] -= s * z[blk_index(b, l, nm, offs)];
    }
}

// ---- vector helpers (deterministic two-stage dot) ----
constexpr int DOT_BLOCKS = 512;
__global__ __launch_bounds__(NT) void dot_stage1(const double *__restrict__ a,
                                                 const double *__restrict__ b, size_t n,
                                                 double *__restrict__ part) {
    __shared__ double red[NT];
    double s = 0.0;
    for (size_t i = (size_t)blockIdx.x * NT + threadIdx.x; i < n; i += (size_t)gridDim.x * NT)
        s = fma(a[i], b[i], s);
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}
__global__ void dot_stage2(const double *__restrict__ part, int np, double *__restrict__ out) {
    __shared__ double red[NT];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += NT) s += part[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = NT / 2; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}
// y = a x + b y
__global__ void axpby_kernel(size_t n, double a, const double *__restrict__ x, double b,
                             double *__restrict__ y) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        y[i] = fma(a, x[i], b * y[i]);
}

struct Vec {
    hipStream_t st;
    double *part = nullptr, *res = nullptr, *host = nullptr;
    int init(hipStream_t s) {
        st = s;
        part = static_cast<double *>(workspace(WS_SOLVE1, sizeof(double) * (DOT_BLOCKS + 8), s));
        host = static_cast<double *>(host_staging(WS_SOLVE1, sizeof(double) * 8));
        KF_REQUIRE(part && host, KF_ERR_NOMEM, "workspace allocation failed");
        res = part + DOT_BLOCKS;
        return KF_OK;
    }
    int dot(const double *a, const double *b, size_t n, double *out) {
        hipLaunchKernelGGL(dot_stage1, dim3(DOT_BLOCKS), dim3(NT), 0, st, a, b, n, part);
        hipLaunchKernelGGL(dot_stage2, dim3(1), dim3(NT), 0, st, part, DOT_BLOCKS, res);
        KF_HIP_TRY(hipMemcpyAsync(host, res, sizeof(double), hipMemcpyDeviceToHost, st));
        KF_HIP_TRY(hipStreamSynchronize(st));
        *out = host[0];
        return KF_OK;
    }
    int axpby(size_t n, double a, const double *x, double b, double *y) {
        hipLaunchKernelGGL(axpby_kernel, dim3(1024), dim3(NT), 0, st, n, a, x, b, y);
        KF_HIP_TRY(hipGetLastError());
        return KF_OK;
    }
    int copy(double *dst, const double *src, size_t n) {
        KF_HIP_TRY(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, st));
        return KF_OK;
    }
};

// ---- the operator y = (K + lambda I) x ----
struct Operator {
    virtual ~Operator() = default;
    virtual int apply(const double *x, double *y) = 0;
};

struct RfpOp : Operator {
    const double *arf;
    int n;
    double lambda;
    hipStream_t st;
    int apply(const double *x, double *y) override {
        return kf_rfp_symv(KF_PREC_F64, arf, n, lambda, x, y, st);
    }
};

// Matrix-free: [E; F] = K_full [x_E; x_F] via alpha_desc(x_F) and the full matvec.
struct MatFreeOp : Operator {
    const double *X, *dX;
    const int *Q, *N;
    int nm, ma, rep, n;
    double sigma, lambda;
    double *ad;
    Vec *vec;
    hipStream_t st;
    int apply(const double *x, double *y) override {
        int rc = kf_local_alpha_desc(dX, Q, N, nm, ma, rep, x + nm, ad, st);
        if (rc) return rc;
        rc = kf_local_full_matvec(X, dX, X, ad, x, Q, Q, N, N, nm, nm, ma, ma, rep, sigma, y,
                                  y + nm, st);
        if (rc) return rc;
        return vec->axpby((size_t)n, lambda, x, 1.0, y);
    }
};

struct BlockPrec {
    const double *inv = nullptr;  // null: identity
    const int *N, *offs;
    int nb, bs, ma, nm;
    hipStream_t st;
    int apply(const double *r, double *z, size_t n, Vec &v) {
        if (!inv) return v.copy(z, r, n);
        hipLaunchKernelGGL(block_apply_kernel, dim3(nb), dim3(128), 0, st, inv, bs, N, ma, offs,
                           nm, r, z);
        KF_HIP_TRY(hipGetLastError());
        return KF_OK;
    }
};

// offs (nm + 1) of the force blocks on the device, from N
int upload_offs(const int *N, int nm, int ma, hipStream_t st, int **d_offs) {
    int *stage = static_cast<int *>(host_staging(WS_SOLVE2, sizeof(int) * (2 * nm + 2)));
    KF_REQUIRE(stage, KF_ERR_NOMEM, "staging allocation failed");
    KF_HIP_TRY(hipMemcpyAsync(stage, N, sizeof(int) * nm, hipMemcpyDeviceToHost, st));
    KF_HIP_TRY(hipStreamSynchronize(st));
    int *offs = stage + nm;
    int acc = 0;
    for (int a = 0; a < nm; ++a) {
        offs[a] = acc;
        acc += 3 * clamp_atoms(stage[a], ma);
    }
    offs[nm] = acc;
    *d_offs = static_cast<int *>(workspace(WS_SOLVE2, sizeof(int) * (nm + 1), st));
    KF_REQUIRE(*d_offs, KF_ERR_NOMEM, "workspace allocation failed");
    KF_HIP_TRY(hipMemcpyAsync(*d_offs, offs, sizeof(int) * (nm + 1), hipMemcpyHostToDevice, st));
    KF_HIP_TRY(hipStreamSynchronize(st));
    return acc;  // naq
}

// Standard PCG from x = 0; stops at ||r|| / ||y|| <= tol (recurrence residual, the
// true residual is recomputed at the end).  history[k] = ||r_k|| / ||y|| if given.
int pcg_run(Operator &A, BlockPrec &P, const double *y, double *x, size_t n, double tol,
            int max_iter, int *iters, double *rel, double *history, hipStream_t st) {
    Vec v;
    int rc = v.init(st);
    if (rc) return rc;
    double *w = static_cast<double *>(workspace(WS_SOLVE0, sizeof(double) * 4 * n, st));
    KF_REQUIRE(w, KF_ERR_NOMEM, "workspace allocation failed");
    double *r = w, *z = w + n, *p = w + 2 * n, *q = w + 3 * n;
    KF_HIP_TRY(hipMemsetAsync(x, 0, n * sizeof(double), st));
    if ((rc = v.copy(r, y, n))) return rc;
    double yy;
    if ((rc = v.dot(y, y, n, &yy))) return rc;
    const double ny = std::sqrt(yy);
    if (ny == 0.0) {
        if (iters) *iters = 0;
        if (rel) *rel = 0.0;
        return KF_OK;
    }
    if ((rc = P.apply(r, z, n, v))) return rc;
    if ((rc = v.copy(p, z, n))) return rc;
    double rz;
    if ((rc = v.dot(r, z, n, &rz))) return rc;
    double rr = yy;
    int k = 0;
    if (history) history[0] = 1.0;
    for (; k < max_iter && std::sqrt(rr) / ny > tol; ++k) {
        if ((rc = A.apply(p, q))) return rc;
        double pq;
        if ((rc = v.dot(p, q, n, &pq))) return rc;
        KF_REQUIRE(pq > 0.0 && std::isfinite(pq), KF_ERR_NOCONV,
                   "PCG breakdown at iteration %d (p.Ap = %g)", k, pq);
        const double al = rz / pq;
        if ((rc = v.axpby(n, al, p, 1.0, x))) return rc;
        if ((rc = v.axpby(n, -al, q, 1.0, r))) return rc;
        if ((rc = v.dot(r, r, n, &rr))) return rc;
        if (history) history[k + 1] = std::sqrt(rr) / ny;
        if ((rc = P.apply(r, z, n, v))) return rc;
        double rz2;
        if ((rc = v.dot(r, z, n, &rz2))) return rc;
        if ((rc = v.axpby(n, 1.0, z, rz2 / rz, p))) return rc;
        rz = rz2;
    }
    // true residual
    if ((rc = A.apply(x, q))) return rc;
    if ((rc = v.axpby(n, 1.0, y, -1.0, q))) return rc;
    double tr;
    if ((rc = v.dot(q, q, n, &tr))) return rc;
    if (iters) *iters = k;
    if (rel) *rel = std::sqrt(tr) / ny;
    return KF_OK;
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_block_factor(const double *blocks, int nb, int bs, const int *N, int max_atoms,
                               double lambda, const double *shift, double *inv, int *info,
                               void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(nb >= 0 && bs == 1 + 3 * max_atoms && bs <= MAXBS, KF_ERR_ARG,
               "block size must be 1 + 3 max_atoms <= %d", MAXBS);
    if (nb == 0) return KF_OK;
    KF_REQUIRE(blocks && N && inv, KF_ERR_ARG, "null pointer");
    hipStream_t st = as_stream(stream);
    hipLaunchKernelGGL(block_factor_kernel, dim3(nb), dim3(NT), 0, st, blocks, bs, N, max_atoms,
                       lambda, shift, inv, info);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_block_apply(const double *inv, int nb, int bs, const int *N, int max_atoms,
                              const double *r, double *z, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(nb >= 0 && bs == 1 + 3 * max_atoms && bs <= MAXBS, KF_ERR_ARG, "bad block size");
    if (nb == 0) return KF_OK;
    hipStream_t st = as_stream(stream);
    int *d_offs = nullptr;
    const int naq = upload_offs(N, nb, max_atoms, st, &d_offs);
    if (naq < 0) return naq;
    hipLaunchKernelGGL(block_apply_kernel, dim3(nb), dim3(128), 0, st, inv, bs, N, max_atoms,
                       d_offs, nb, r, z);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_ef_pcg_rfp(const double *arf, int n, double lambda, const double *inv,
                             int nb, int bs, const int *N, int max_atoms, const double *y,
                             double *x, double tol, int max_iter, int *iters, double *rel,
                             double *history, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(n >= 0 && nb >= 0 && max_iter >= 0, KF_ERR_ARG, "bad sizes");
    KF_REQUIRE(!inv || (bs == 1 + 3 * max_atoms && bs <= MAXBS), KF_ERR_ARG,
               "block size must be 1 + 3 max_atoms <= %d", MAXBS);
    if (n == 0) return KF_OK;
    hipStream_t st = as_stream(stream);
    int *d_offs = nullptr;
    const int naq = upload_offs(N, nb, max_atoms, st, &d_offs);
    if (naq < 0) return naq;
    KF_REQUIRE(nb + naq == n, KF_ERR_ARG, "n (%d) != nm + naq (%d)", n, nb + naq);
    RfpOp A;
    A.arf = arf; A.n = n; A.lambda = lambda; A.st = st;
    BlockPrec P;
    P.inv = inv; P.N = N; P.offs = d_offs; P.nb = nb; P.bs = bs; P.ma = max_atoms; P.nm = nb;
    P.st = st;
    return pcg_run(A, P, y, x, (size_t)n, tol, max_iter, iters, rel, history, st);
}

extern "C" int kf_ef_pcg_matfree(const double *X, const double *dX, const int *Q, const int *N,
                                 int nm, int max_atoms, int rep, double sigma, double lambda,
                                 const double *inv, int bs, const double *y, double *x,
                                 double tol, int max_iter, int *iters, double *rel,
                                 double *history, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(nm >= 0 && max_iter >= 0, KF_ERR_ARG, "bad sizes");
    KF_REQUIRE(!inv || (bs == 1 + 3 * max_atoms && bs <= MAXBS), KF_ERR_ARG,
               "block size must be 1 + 3 max_atoms <= %d", MAXBS);
    if (nm == 0) return KF_OK;
    hipStream_t st = as_stream(stream);
    int *d_offs = nullptr;
    const int naq = upload_offs(N, nm, max_atoms, st, &d_offs);
    if (naq < 0) return naq;
    const size_t n = (size_t)nm + naq;
    Vec vec;
    int rc = vec.init(st);
    if (rc) return rc;
    // every matvec runs on the descriptor columns the set uses (exact: the others are zero)
    CompactSets cs;
    cs.x1 = cs.x2 = X; cs.dx1 = cs.dx2 = dX; cs.n1 = cs.n2 = N;
    cs.nm1 = cs.nm2 = nm; cs.ma1 = cs.ma2 = max_atoms;
    int nf = rep;
    if ((rc = compact_columns(cs, rep, WS_PACK, WS_PACK, st, &nf))) return rc;
    MatFreeOp A;
    A.X = cs.x1; A.dX = cs.dx1; A.Q = Q; A.N = N; A.nm = nm; A.ma = max_atoms; A.rep = nf;
    A.n = (int)n; A.sigma = sigma; A.lambda = lambda; A.vec = &vec; A.st = st;
    A.ad = static_cast<double *>(
        workspace(WS_PACK2, sizeof(double) * (size_t)nm * max_atoms * nf, st));
    KF_REQUIRE(A.ad, KF_ERR_NOMEM, "workspace allocation failed");
    BlockPrec P;
    P.inv = inv; P.N = N; P.offs = d_offs; P.nb = nm; P.bs = bs; P.ma = max_atoms; P.nm = nm;
    P.st = st;
    return pcg_run(A, P, y, x, n, tol, max_iter, iters, rel, history, st);
}

// Unrolled PCG (x0 = 0, `steps` iterations) with the explicit RFP operator and the
// block preconditioner inv (= (K_bb + lambda + diag(s))^-1); returns
// loss = log(||r_steps|| / ||y||) and adds d loss / d s to grad_shift (nb x bs), by
// reverse-mode differentiation of the recurrences.
extern "C" int kf_pcg_unrolled_grad_rfp(const double *arf, int n, double lambda,
                                        const double *inv, int nb, int bs, const int *N,
                                        int max_atoms, const double *y, int steps, double *loss,
                                        double *grad_shift, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(n > 0 && steps >= 1 && inv && grad_shift && loss, KF_ERR_ARG, "bad arguments");
    KF_REQUIRE(bs == 1 + 3 * max_atoms && bs <= MAXBS, KF_ERR_ARG,
               "block size must be 1 + 3 max_atoms <= %d", MAXBS);
    hipStream_t st = as_stream(stream);
    int *d_offs = nullptr;
    const int naq = upload_offs(N, nb, max_atoms, st, &d_offs);
    if (naq < 0) return naq;
    KF_REQUIRE(nb + naq == n, KF_ERR_ARG, "n (%d) != nm + naq (%d)", n, nb + naq);
    Vec v;
    int rc = v.init(st);
    if (rc) return rc;
    const size_t N_ = (size_t)n;
    const int K = steps;
    // storage: r_0..r_K, z_0..z_{K-1}, p_0..p_{K-1}, q_0..q_{K-1}, plus 5 adjoint vectors
    const size_t nvec = (size_t)(K + 1) + 3 * (size_t)K + 5;
    double *w = static_cast<double *>(workspace(WS_SOLVE0, sizeof(double) * nvec * N_, st));
    KF_REQUIRE(w, KF_ERR_NOMEM, "workspace allocation failed");
    auto R = [&](int k) { return w + (size_t)k * N_; };
    auto Z = [&](int k) { return w + (size_t)(K + 1 + k) * N_; };
    auto Pv = [&](int k) { return w + (size_t)(2 * K + 1 + k) * N_; };
    auto Qv = [&](int k) { return w + (size_t)(3 * K + 1 + k) * N_; };
    double *rb = w + (size_t)(4 * K + 1) * N_, *pb = rb + N_, *zb = pb + N_, *qb = zb + N_,
           *tmp = qb + N_;
    RfpOp A;
    A.arf = arf; A.n = n; A.lambda = lambda; A.st = st;
    BlockPrec P;
    P.inv = inv; P.N = N; P.offs = d_offs; P.nb = nb; P.bs = bs; P.ma = max_atoms; P.nm = nb;
    P.st = st;
    std::vector<double> rho(K + 1), sig(K), alp(K), bet(K);
    // forward
    if ((rc = v.copy(R(0), y, N_))) return rc;
    double yy;
    if ((rc = v.dot(y, y, N_, &yy))) return rc;
    KF_REQUIRE(yy > 0.0, KF_ERR_ARG, "zero right-hand side");
    if ((rc = P.apply(R(0), Z(0), N_, v))) return rc;
    if ((rc = v.copy(Pv(0), Z(0), N_))) return rc;
    if ((rc = v.dot(R(0), Z(0), N_, &rho[0]))) return rc;
    for (int k = 0; k < K; ++k) {
        if ((rc = A.apply(Pv(k), Qv(k)))) return rc;
        if ((rc = v.dot(Pv(k), Qv(k), N_, &sig[k]))) return rc;
        KF_REQUIRE(sig[k] > 0.0 && std::isfinite(sig[k]), KF_ERR_NOCONV, "breakdown");
        alp[k] = rho[k] / sig[k];
        if ((rc = v.copy(R(k + 1), R(k), N_))) return rc;
        if ((rc = v.axpby(N_, -alp[k], Qv(k), 1.0, R(k + 1)))) return rc;
        if (k + 1 < K) {
            if ((rc = P.apply(R(k + 1), Z(k + 1), N_, v))) return rc;
            if ((rc = v.dot(R(k + 1), Z(k + 1), N_, &rho[k + 1]))) return rc;
            bet[k] = rho[k + 1] / rho[k];
            if ((rc = v.copy(Pv(k + 1), Z(k + 1), N_))) return rc;
            if ((rc = v.axpby(N_, bet[k], Pv(k), 1.0, Pv(k + 1)))) return rc;
        }
    }
    double rr;
    if ((rc = v.dot(R(K), R(K), N_, &rr))) return rc;
    *loss = 0.5 * std::log(rr / yy);
    // reverse: rb = dL/dr_K = r_K / rr
    KF_HIP_TRY(hipMemsetAsync(pb, 0, sizeof(double) * N_, st));
    if ((rc = v.copy(rb, R(K), N_))) return rc;
    if ((rc = v.axpby(N_, 0.0, rb, 1.0 / rr, rb))) return rc;
    std::vector<double> rho_b(K + 1, 0.0);
    auto prec_adjoint = [&](int k) -> int {
        // z_k = P r_k with adjoint zb: rb += P zb ; grad += -(P zb) * z_k
        int e = P.apply(zb, tmp, N_, v);
        if (e) return e;
        if ((e = v.axpby(N_, 1.0, tmp, 1.0, rb))) return e;
        hipLaunchKernelGGL(shift_grad_kernel, dim3(nb), dim3(128), 0, st, inv, bs, N, max_atoms,
                           d_offs, nb, zb, Z(k), grad_shift);
        KF_HIP_TRY(hipGetLastError());
        return KF_OK;
    };
    for (int k = K - 1; k >= 0; --k) {
        if (k + 1 < K) {
            // p_{k+1} = z_{k+1} + bet_k p_k
            if ((rc = v.copy(zb, pb, N_))) return rc;
            double betb;
            if ((rc = v.dot(pb, Pv(k), N_, &betb))) return rc;
            // pb becomes the adjoint of p_k: bet_k * pb (the other terms are added below)
            if ((rc = v.axpby(N_, 0.0, pb, bet[k], pb))) return rc;
            // bet_k = rho_{k+1} / rho_k
            rho_b[k + 1] += betb / rho[k];
            rho_b[k] += -betb * rho[k + 1] / (rho[k] * rho[k]);
            // rho_{k+1} = r_{k+1} . z_{k+1}
            if ((rc = v.axpby(N_, rho_b[k + 1], Z(k + 1), 1.0, rb))) return rc;
            if ((rc = v.axpby(N_, rho_b[k + 1], R(k + 1), 1.0, zb))) return rc;
            // z_{k+1} = P r_{k+1}
            if ((rc = prec_adjoint(k + 1))) return rc;
        }
        // r_{k+1} = r_k - alp_k q_k   (rb holds dL/dr_{k+1} and continues as dL/dr_k)
        double alpb;
        if ((rc = v.dot(rb, Qv(k), N_, &alpb))) return rc;
        alpb = -alpb;
        // qb = -alp_k rb
        if ((rc = v.copy(qb, rb, N_))) return rc;
        if ((rc = v.axpby(N_, 0.0, qb, -alp[k], qb))) return rc;
        // alp_k = rho_k / sig_k
        rho_b[k] += alpb / sig[k];
        const double sigb = -alpb * rho[k] / (sig[k] * sig[k]);
        // sig_k = p_k . q_k
        if ((rc = v.axpby(N_, sigb, Qv(k), 1.0, pb))) return rc;
        if ((rc = v.axpby(N_, sigb, Pv(k), 1.0, qb))) return rc;
        // q_k = A p_k  ->  pb += A qb
        if ((rc = A.apply(qb, tmp))) return rc;
        if ((rc = v.axpby(N_, 1.0, tmp, 1.0, pb))) return rc;
    }
    // p_0 = z_0 ; rho_0 = r_0 . z_0 ; z_0 = P r_0 (r_0 = y is a constant)
    if ((rc = v.copy(zb, pb, N_))) return rc;
    if ((rc = v.axpby(N_, rho_b[0], R(0), 1.0, zb))) return rc;
    if ((rc = prec_adjoint(0))) return rc;
    KF_HIP_TRY(hipStreamSynchronize(st));
    return KF_OK;
}
