// local_matvec.hip -- matrix-free products with the energy+force local kernel:
// alpha_desc, the Jacobian-transposed (energy) matvec, the Hessian (force) matvec and
// the full [E; F] matvec that drives PCG on the energy+force system.
//
// With alpha_desc[b,j,:] = J_bj alpha_F[b] (kf_local_alpha_desc) and, for every
// query atom (a,i) and label-matched train atom (b,j):
//   d = x_ai - x_bj,  e = exp(-|d|^2/2s^2),  c = e/s^2,  inner = d . alpha_desc[b,j]
//   E[a]  += e alpha_E[b] + c inner
//   G_ai  += c alpha_desc[b,j] - (e/s^4) inner d + c alpha_E[b] x_bj - c alpha_E[b] x_ai
//   F[a]  += J_ai^T G_ai
// (src/local_kernels.cpp:2895-3425).  Collecting terms per label, with
//   beta_ij = (e/s^4) inner_ij + c_ij alpha_E[b(j)]
//   G = C * ALPHA_DESC + B * X2 - diag(B 1) * X1
// so each label is two FP64 MFMA GEMMs over packed rows:
//   T = X1 [X2; ALPHA_DESC]^T            (query block x 2 R2, K = rep)
//   elementwise: T -> [B | C] plus the per-row E and B 1 (one wave per query row)
//   G = T [X2; ALPHA_DESC]               (query block x rep, K = 2 R2)
// then one pass over dX1 forms J_ai^T (G_ai - w_ai x_ai) per query atom, and per-
// molecule sums in a fixed order (deterministic, no atomics).
#include <algorithm>
#include <new>
#include <cmath>
#include <vector>

#include "gemm_f64.hpp"
#include "kf_common.hpp"
#include "local_plan.hpp"

namespace kf {
namespace {

constexpr int NT = 256;

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// alpha_desc[b,j,f] = sum_c dX[b,j,f,c] alpha[offs[b] + c] for j < N[b]; 0 otherwise.
// One wave per (molecule, atom), lanes over Jacobian columns, wave sum per feature.
__global__ __launch_bounds__(NT) void alpha_desc_kernel(const double *__restrict__ dX,
                                                        const int *__restrict__ N,
                                                        const int *__restrict__ offs, int nm,
                                                        int ma, int rep,
                                                        const double *__restrict__ alpha,
                                                        double *__restrict__ out) {
    const int lane = threadIdx.x & 63;
    const long atom = (long)blockIdx.x * 4 + (threadIdx.x >> 6);
    if (atom >= (long)nm * ma) return;
    const int b = (int)(atom / ma), j = (int)(atom - (long)b * ma);
    const int nb = min(max(N[b], 0), ma);
    double *o = out + atom * rep;
    if (j >= nb) {
        for (int f = lane; f < rep; f += 64) o[f] = 0.0;
        return;
    }
    const int ld = 3 * ma, nc = 3 * nb;
    const double *J = dX + atom * rep * ld;
    // Jacobian columns in groups of 128 (two per lane); one group for up to 42 atoms
    for (int c0 = 0; c0 < nc; c0 += 128) {
        const double a0 = c0 + lane < nc ? alpha[offs[b] + c0 + lane] : 0.0;
        const double a1 = c0 + lane + 64 < nc ? alpha[offs[b] + c0 + lane + 64] : 0.0;
        for (int f = 0; f < rep; ++f) {
            const double *row = J + (size_t)f * ld + c0;
            double s = c0 + lane < nc ? row[lane] * a0 : 0.0;
            if (c0 + lane + 64 < nc) s = fma(row[lane + 64], a1, s);
            s = wave_sum(s);
            if (lane == 0) o[f] = c0 == 0 ? s : o[f] + s;
        }
    }
}

// Gather packed rows: dst[r, :] = src[idx[r], :] for r < rows, with the squared
// norm (n2) and optionally the dot with a second gathered matrix (dot).
__global__ __launch_bounds__(NT) void gather_rows(const double *__restrict__ src,
                                                  const double *__restrict__ src2,
                                                  const int *__restrict__ idx, int rows, int rep,
                                                  double *__restrict__ dst,
                                                  double *__restrict__ dst2,
                                                  double *__restrict__ n2,
                                                  double *__restrict__ dot) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const double *s = src + (size_t)idx[r] * rep;
    const double *s2 = src2 ? src2 + (size_t)idx[r] * rep : nullptr;
    double nn = 0.0, dd = 0.0;
    for (int f = lane; f < rep; f += 64) {
        const double v = s[f];
        dst[(size_t)r * rep + f] = v;
        nn = fma(v, v, nn);
        if (s2) {
            const double w = s2[f];
            dst2[(size_t)r * rep + f] = w;
            dd = fma(v, w, dd);
        }
    }
    nn = wave_sum(nn);
    dd = wave_sum(dd);
    if (lane == 0) {
        n2[r] = nn;
        if (dot) dot[r] = dd;
    }
}

// T (rows x 2 R2) = [S | P] -> [B | C] in place, one wave per query row:
//   d2 = n1 + n2 - 2 S, e = exp(cexp d2), c = e / s^2, inner = P - rdot
//   B = (e / s^4) inner + c aE,  C = c
// Row outputs: erow = sum(e aE + c inner) (when wanted), wsum = sum(B).
__global__ __launch_bounds__(NT) void coef_kernel(double *__restrict__ T, int rows, int R2,
                                                  const double *__restrict__ n1,
                                                  const double *__restrict__ n2,
                                                  const double *__restrict__ rdot,
                                                  const double *__restrict__ aE, double cexp,
                                                  double inv_s2, const int *__restrict__ qsrc,
                                                  double *__restrict__ erow,
                                                  double *__restrict__ wsum) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    double *S = T + (size_t)r * 2 * R2, *P = S + R2;
    const double nq = n1[r];
    double es = 0.0, ws = 0.0;
    for (int t = lane; t < R2; t += 64) {
        const double d2 = fmax(nq + n2[t] - 2.0 * S[t], 0.0);
        const double e = exp(cexp * d2), c = e * inv_s2;
        const double inner = P[t] - rdot[t];
        const double ae = aE ? aE[t] : 0.0;
        const double bco = c * inv_s2 * inner + c * ae;
        es += e * ae + c * inner;
        ws += bco;
        S[t] = bco;
        P[t] = c;
    }
    es = wave_sum(es);
    ws = wave_sum(ws);
    if (lane == 0) {
        if (erow) erow[qsrc[r]] = es;
        wsum[r] = ws;
    }
}

// Fat[qsrc[r], c] = sum_f dX1[qsrc[r], f, c] (G[r, f] - wsum[r] X1p[r, f]).
__global__ __launch_bounds__(NT) void backproj_kernel(const double *__restrict__ G,
                                                      const double *__restrict__ X1p,
                                                      const double *__restrict__ wsum,
                                                      const double *__restrict__ dX1,
                                                      const int *__restrict__ qsrc,
                                                      const int *__restrict__ N1, int rows,
                                                      int ma, int rep,
                                                      double *__restrict__ Fat) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int src = qsrc[r], a = src / ma;
    const int nc = 3 * min(max(N1[a], 0), ma), ld = 3 * ma;
    const double *J = dX1 + (size_t)src * rep * ld;
    const double *g = G + (size_t)r * rep, *x = X1p + (size_t)r * rep;
    const double w = wsum[r];
    double *o = Fat + (size_t)src * ld;
    // Jacobian columns in groups of 128 (two per lane); one group for up to 42 atoms
    for (int c0 = 0; c0 < nc; c0 += 128) {
        const int ca = c0 + lane, cb = c0 + lane + 64;
        double s0 = 0.0, s1 = 0.0;
        for (int f = 0; f < rep; ++f) {
            const double gv = g[f] - w * x[f];
            const double *row = J + (size_t)f * ld;
            if (ca < nc) s0 = fma(row[ca], gv, s0);
            if (cb < nc) s1 = fma(row[cb], gv, s1);
        }
        if (ca < nc) o[ca] = s0;
        if (cb < nc) o[cb] = s1;
    }
}

// E[a] = sum_i erow[a*ma+i]; F[offs[a] + c] = sum_i Fat[a*ma+i, c] (fixed order).
__global__ __launch_bounds__(NT) void reduce_mol_kernel(const double *__restrict__ erow,
                                                        const double *__restrict__ Fat,
                                                        const int *__restrict__ N,
                                                        const int *__restrict__ offs, int nm,
                                                        int ma, double *__restrict__ E,
                                                        double *__restrict__ F) {
    const int a = blockIdx.x;
    if (a >= nm) return;
    const int na = min(max(N[a], 0), ma), ld = 3 * ma;
    if (E && threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < na; ++i) s += erow[(size_t)a * ma + i];
        E[a] = s;
    }
    if (F)
        for (int c = threadIdx.x; c < 3 * na; c += blockDim.x) {
            double s = 0.0;
            for (int i = 0; i < na; ++i) s += Fat[((size_t)a * ma + i) * ld + c];
            F[offs[a] + c] = s;
        }
}

// backproj_kernel with the molecule's atom count taken from the force offsets
// (3 N[a] = offs[a + 1] - offs[a]), so a plan needs no N array at execution.
__global__ __launch_bounds__(NT) void backproj_plan_kernel(const double *__restrict__ G,
                                                           const double *__restrict__ X1p,
                                                           const double *__restrict__ wsum,
                                                           const double *__restrict__ dX1,
                                                           const int *__restrict__ qsrc,
                                                           const int *__restrict__ offs, int rows,
                                                           int ma, int rep,
                                                           double *__restrict__ Fat) {
    const int lane = threadIdx.x & 63;
    const int r = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (r >= rows) return;
    const int src = qsrc[r], a = src / ma;
    const int nc = offs[a + 1] - offs[a], ld = 3 * ma;
    const double *J = dX1 + (size_t)src * rep * ld;
    const double *g = G + (size_t)r * rep, *x = X1p + (size_t)r * rep;
    const double w = wsum[r];
    double *o = Fat + (size_t)src * ld;
    // Jacobian columns in groups of 128 (two per lane); one group for up to 42 atoms
    for (int c0 = 0; c0 < nc; c0 += 128) {
        const int ca = c0 + lane, cb = c0 + lane + 64;
        double s0 = 0.0, s1 = 0.0;
        for (int f = 0; f < rep; ++f) {
            const double gv = g[f] - w * x[f];
            const double *row = J + (size_t)f * ld;
            if (ca < nc) s0 = fma(row[ca], gv, s0);
            if (cb < nc) s1 = fma(row[cb], gv, s1);
        }
        if (ca < nc) o[ca] = s0;
        if (cb < nc) o[cb] = s1;
    }
}

// reduce_mol_kernel with atom counts from the force offsets.
__global__ __launch_bounds__(NT) void reduce_plan_kernel(const double *__restrict__ erow,
                                                         const double *__restrict__ Fat,
                                                         const int *__restrict__ offs, int nm,
                                                         int ma, double *__restrict__ E,
                                                         double *__restrict__ F) {
    const int a = blockIdx.x;
    if (a >= nm) return;
    const int nc = offs[a + 1] - offs[a], na = nc / 3, ld = 3 * ma;
    if (threadIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < na; ++i) s += erow[(size_t)a * ma + i];
        E[a] = s;
    }
    for (int c = threadIdx.x; c < nc; c += blockDim.x) {
        double s = 0.0;
        for (int i = 0; i < na; ++i) s += Fat[((size_t)a * ma + i) * ld + c];
        F[offs[a] + c] = s;
    }
}

// Expand alpha_E (per molecule) to the packed train rows.
__global__ void expand_mol(const double *__restrict__ aE, const int *__restrict__ src, int rows,
                           int ma, double *__restrict__ out) {
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < rows) out[r] = aE[src[r] / ma];
}

// modes: bit 1 E, bit 2 F, bit 4 alpha_E terms
int matvec_impl(const double *X1, const double *dX1, const double *X2, const double *ad,
                const double *aE, const int *Q1, const int *Q2, const int *N1, const int *N2,
                int nm1, int nm2, int ma1, int ma2, int rep, double sigma, double *E, double *F,
                int mode, hipStream_t stream) {
    KF_REQUIRE(std::isfinite(sigma) && sigma > 0.0, KF_ERR_SIGMA, "sigma must be > 0");
    KF_REQUIRE(nm1 >= 0 && nm2 >= 0 && ma1 >= 0 && ma2 >= 0 && rep >= 1, KF_ERR_ARG,
               "bad dimensions");
    if (nm1 == 0) return KF_OK;
    KF_REQUIRE(X1 && Q1 && N1 && (nm2 == 0 || (X2 && ad && Q2 && N2)), KF_ERR_ARG,
               "null pointer");
    KF_REQUIRE(!(mode & 2) || dX1, KF_ERR_ARG, "null dX1");  // F may be null when naq1 == 0
    KF_REQUIRE(!(mode & 1) || E, KF_ERR_ARG, "null E");
    KF_REQUIRE(!(mode & 4) || nm2 == 0 || aE, KF_ERR_ARG, "null alpha_E");
    std::vector<int> q1, n1, q2, n2;
    int rc;
    if ((rc = read_metadata(Q1, N1, nm1, ma1, q1, n1, stream, WS_AUX0))) return rc;
    if (nm2 > 0 && (rc = read_metadata(Q2, N2, nm2, ma2, q2, n2, stream, WS_AUX1))) return rc;
    std::vector<int> labels = distinct_labels(q1.data(), n1.data(), nm1, ma1);
    LocalPlan p1, p2;
    build_plan(q1.data(), n1.data(), nm1, ma1, labels, p1);
    if (nm2 > 0) build_plan(q2.data(), n2.data(), nm2, ma2, labels, p2);
    std::vector<int> offs(nm1 + 1, 0);
    for (int a = 0; a < nm1; ++a) offs[a + 1] = offs[a] + 3 * clamp_atoms(n1[a], ma1);

    // device metadata: offs | src1 | src2
    const size_t n_meta = offs.size() + p1.src.size() + (nm2 > 0 ? p2.src.size() : 0);
    int *meta = static_cast<int *>(workspace(WS_PLAN, sizeof(int) * n_meta, stream));
    int *stage = static_cast<int *>(host_staging(WS_PLAN, sizeof(int) * n_meta));
    KF_REQUIRE(meta && stage, KF_ERR_NOMEM, "workspace allocation failed");
    std::copy(offs.begin(), offs.end(), stage);
    std::copy(p1.src.begin(), p1.src.end(), stage + offs.size());
    if (nm2 > 0) std::copy(p2.src.begin(), p2.src.end(), stage + offs.size() + p1.src.size());
    KF_HIP_TRY(hipMemcpyAsync(meta, stage, sizeof(int) * n_meta, hipMemcpyHostToDevice, stream));
    const int *d_offs = meta, *d_src1 = meta + offs.size(),
              *d_src2 = d_src1 + p1.src.size();

    // per-atom outputs (zero for atoms without matches)
    const size_t natom1 = (size_t)nm1 * ma1;
    double *erow = nullptr, *Fat = nullptr;
    {
        const size_t bytes = sizeof(double) * (natom1 + ((mode & 2) ? natom1 * 3 * ma1 : 0));
        double *w = static_cast<double *>(workspace(WS_AUX3, bytes, stream));
        KF_REQUIRE(w, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", bytes);
        KF_HIP_TRY(hipMemsetAsync(w, 0, bytes, stream));
        erow = w;
        if (mode & 2) Fat = w + natom1;
    }
    ScopedTimer timer(TM_MATVEC, stream);
    const double cexp = -1.0 / (2.0 * sigma * sigma), inv_s2 = 1.0 / (sigma * sigma);
    for (int l = 0; l < p1.nlabels && nm2 > 0; ++l) {
        const int r1b = p1.mol_row[(size_t)l * (nm1 + 1)], R1 = p1.mol_row[(size_t)l * (nm1 + 1) + nm1] - r1b;
        const int r2b = p2.mol_row[(size_t)l * (nm2 + 1)], R2 = p2.mol_row[(size_t)l * (nm2 + 1) + nm2] - r2b;
        if (R1 == 0 || R2 == 0) continue;
        // packed train block: [X2 (R2 x rep); AD (R2 x rep)], n2, rdot, aE per row
        const size_t train_bytes = sizeof(double) * ((size_t)2 * R2 * rep + 3 * (size_t)R2);
        // query block size: T = QB x 2 R2 and G = QB x rep within ~1 GiB
        const size_t budget = size_t(1) << 30;
        int QB = (int)std::min<size_t>(R1, std::max<size_t>(64, budget / (sizeof(double) *
                                                                             ((size_t)2 * R2 + 2 * rep + 2))));
        const size_t q_bytes = sizeof(double) * ((size_t)QB * (2 * R2 + 2 * rep) + 2 * (size_t)QB);
        char *ws = static_cast<char *>(workspace(WS_AUX2, train_bytes + q_bytes, stream));
        KF_REQUIRE(ws, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", train_bytes + q_bytes);
        double *P2 = reinterpret_cast<double *>(ws);
        double *AD2 = P2 + (size_t)R2 * rep;
        double *nrm2 = AD2 + (size_t)R2 * rep, *rdot = nrm2 + R2, *aEr = rdot + R2;
        double *T = aEr + R2;
        double *Xq = T + (size_t)QB * 2 * R2, *G = Xq + (size_t)QB * rep;
        double *nq = G + (size_t)QB * rep, *wsum = nq + QB;
        hipLaunchKernelGGL(gather_rows, dim3((R2 + 3) / 4), dim3(NT), 0, stream, X2, ad,
                           d_src2 + r2b, R2, rep, P2, AD2, nrm2, rdot);
        KF_HIP_TRY(hipGetLastError());
        if (mode & 4) {
            hipLaunchKernelGGL(expand_mol, dim3((R2 + 255) / 256), dim3(256), 0, stream, aE,
                               d_src2 + r2b, R2, ma2, aEr);
            KF_HIP_TRY(hipGetLastError());
        }
        for (int q0 = 0; q0 < R1; q0 += QB) {
            const int qb = std::min(QB, R1 - q0);
            const int *qsrc = d_src1 + r1b + q0;
            hipLaunchKernelGGL(gather_rows, dim3((qb + 3) / 4), dim3(NT), 0, stream, X1,
                               (const double *)nullptr, qsrc, qb, rep, Xq, (double *)nullptr, nq,
                               (double *)nullptr);
            KF_HIP_TRY(hipGetLastError());
            // T = Xq [X2; AD]^T
            if ((rc = gemm::gemm_f64<true>(qb, 2 * R2, rep, 1.0, Xq, rep, P2, rep, 0.0, T,
                                           (size_t)2 * R2, stream)))
                return rc;
            hipLaunchKernelGGL(coef_kernel, dim3((qb + 3) / 4), dim3(NT), 0, stream, T, qb, R2,
                               nq, nrm2, rdot, (mode & 4) ? aEr : (const double *)nullptr, cexp,
                               inv_s2, qsrc, (mode & 1) ? erow : (double *)nullptr, wsum);
            KF_HIP_TRY(hipGetLastError());
            if (mode & 2) {
                // G = [B | C] [X2; AD]
                if ((rc = gemm::gemm_f64<false>(qb, rep, 2 * R2, 1.0, T, (size_t)2 * R2, P2, rep,
                                                0.0, G, rep, stream)))
                    return rc;
                hipLaunchKernelGGL(backproj_kernel, dim3((qb + 3) / 4), dim3(NT), 0, stream, G, Xq,
                                   wsum, dX1, qsrc, N1, qb, ma1, rep, Fat);
                KF_HIP_TRY(hipGetLastError());
            }
        }
    }
    hipLaunchKernelGGL(reduce_mol_kernel, dim3(nm1), dim3(128), 0, stream, erow, Fat, N1, d_offs,
                       nm1, ma1, (mode & 1) ? E : (double *)nullptr,
                       (mode & 2) ? F : (double *)nullptr);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

// Fill a buffer with zeros by a kernel (stream-ordered, and safe inside a captured graph).
__global__ void zero_fill(double *__restrict__ p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0.0;
}

}  // namespace
}  // namespace kf

// A full matvec planned once for fixed labels of the query set and a fixed training side
// (descriptors, alpha_desc, alpha_E): the label plans, the packed training rows of every
// label with their norms and dot products, and the query workspace all live in device
// memory owned by the plan, so an execution neither reads anything back nor allocates.
struct kf_matvec_plan {
    int device = 0;
    int nm1 = 0, ma1 = 0, nm2 = 0, ma2 = 0, rep = 0;
    double sigma = 1.0;
    struct Label {
        int r1b, R1, R2, QB;
        size_t train_off;  // doubles into `train`
    };
    std::vector<Label> labels;
    int QBmax = 0;
    size_t q_doubles = 0;  // per-execution query workspace
    int *meta = nullptr;   // offs (nm1 + 1) | src1
    double *train = nullptr, *query = nullptr, *acc = nullptr;  // acc: erow | Fat
    size_t acc_doubles = 0;
    ~kf_matvec_plan() {
        (void)hipFree(meta);
        (void)hipFree(train);
        (void)hipFree(query);
        (void)hipFree(acc);
    }
};

namespace kf {
namespace {

int plan_alloc(void **p, size_t bytes) {
    if (hipMalloc(p, bytes ? bytes : 256) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        set_last_error("hipMalloc of %zu bytes for a matvec plan failed", bytes);
        return KF_ERR_NOMEM;
    }
    return KF_OK;
}

int plan_create(const int *Q1, const int *N1, int nm1, int ma1, const double *X2,
                const double *ad, const double *aE, const int *Q2, const int *N2, int nm2,
                int ma2, int rep, double sigma, kf_matvec_plan *P, hipStream_t stream) {
    KF_REQUIRE(std::isfinite(sigma) && sigma > 0.0, KF_ERR_SIGMA, "sigma must be > 0");
    KF_REQUIRE(nm1 >= 1 && nm2 >= 0 && ma1 >= 1 && ma2 >= 0 && rep >= 1, KF_ERR_ARG,
               "bad dimensions");
    KF_REQUIRE(Q1 && N1 && (nm2 == 0 || (X2 && ad && aE && Q2 && N2)), KF_ERR_ARG,
               "null pointer");
    KF_HIP_TRY(hipGetDevice(&P->device));
    P->nm1 = nm1; P->ma1 = ma1; P->nm2 = nm2; P->ma2 = ma2; P->rep = rep; P->sigma = sigma;
    std::vector<int> q1, n1, q2, n2;
    int rc;
    if ((rc = read_metadata(Q1, N1, nm1, ma1, q1, n1, stream, WS_AUX0))) return rc;
    if (nm2 > 0 && (rc = read_metadata(Q2, N2, nm2, ma2, q2, n2, stream, WS_AUX1))) return rc;
    std::vector<int> labels = distinct_labels(q1.data(), n1.data(), nm1, ma1);
    LocalPlan p1, p2;
    build_plan(q1.data(), n1.data(), nm1, ma1, labels, p1);
    if (nm2 > 0) build_plan(q2.data(), n2.data(), nm2, ma2, labels, p2);
    std::vector<int> meta(nm1 + 1 + p1.src.size(), 0);
    for (int a = 0; a < nm1; ++a) meta[a + 1] = meta[a] + 3 * clamp_atoms(n1[a], ma1);
    std::copy(p1.src.begin(), p1.src.end(), meta.begin() + nm1 + 1);
    std::vector<int> src2_all;
    size_t train_doubles = 0;
    for (int l = 0; l < p1.nlabels && nm2 > 0; ++l) {
        const int r1b = p1.mol_row[(size_t)l * (nm1 + 1)];
        const int R1 = p1.mol_row[(size_t)l * (nm1 + 1) + nm1] - r1b;
        const int r2b = p2.mol_row[(size_t)l * (nm2 + 1)];
        const int R2 = p2.mol_row[(size_t)l * (nm2 + 1) + nm2] - r2b;
        if (R1 == 0 || R2 == 0) continue;
        const size_t budget = size_t(1) << 30;
        const int QB = (int)std::min<size_t>(
            R1, std::max<size_t>(64, budget / (sizeof(double) * ((size_t)2 * R2 + 2 * rep + 2))));
        P->labels.push_back({r1b, R1, R2, QB, train_doubles});
        train_doubles += (size_t)2 * R2 * rep + 3 * (size_t)R2;
        P->q_doubles = std::max(P->q_doubles,
                                (size_t)QB * (2 * (size_t)R2 + 2 * (size_t)rep) + 2 * (size_t)QB);
        P->QBmax = std::max(P->QBmax, QB);
        src2_all.insert(src2_all.end(), p2.src.begin() + r2b, p2.src.begin() + r2b + R2);
    }
    const size_t natom1 = (size_t)nm1 * ma1;
    P->acc_doubles = natom1 + natom1 * 3 * ma1;
    if ((rc = plan_alloc((void **)&P->meta, sizeof(int) * (meta.size() + src2_all.size()))))
        return rc;
    if ((rc = plan_alloc((void **)&P->train, sizeof(double) * train_doubles))) return rc;
    if ((rc = plan_alloc((void **)&P->query, sizeof(double) * P->q_doubles))) return rc;
    if ((rc = plan_alloc((void **)&P->acc, sizeof(double) * P->acc_doubles))) return rc;
    meta.insert(meta.end(), src2_all.begin(), src2_all.end());
    KF_HIP_TRY(hipMemcpyAsync(P->meta, meta.data(), sizeof(int) * meta.size(),
                              hipMemcpyHostToDevice, stream));
    // pack the training side of every label once
    const int *d_src2 = P->meta + nm1 + 1 + p1.src.size();
    size_t s2 = 0;
    for (const auto &L : P->labels) {
        double *P2 = P->train + L.train_off;
        double *AD2 = P2 + (size_t)L.R2 * rep;
        double *nrm2 = AD2 + (size_t)L.R2 * rep, *rdot = nrm2 + L.R2, *aEr = rdot + L.R2;
        hipLaunchKernelGGL(gather_rows, dim3((L.R2 + 3) / 4), dim3(NT), 0, stream, X2, ad,
                           d_src2 + s2, L.R2, rep, P2, AD2, nrm2, rdot);
        hipLaunchKernelGGL(expand_mol, dim3((L.R2 + 255) / 256), dim3(256), 0, stream, aE,
                           d_src2 + s2, L.R2, ma2, aEr);
        KF_HIP_TRY(hipGetLastError());
        s2 += L.R2;
    }
    // the host vectors above are released on return: wait for the copy that reads them
    KF_HIP_TRY(hipStreamSynchronize(stream));
    return KF_OK;
}

int plan_execute(const kf_matvec_plan *P, const double *X1, const double *dX1, double *E,
                 double *F, hipStream_t stream) {
    KF_REQUIRE(P && X1 && dX1 && E && F, KF_ERR_ARG, "null pointer");
    int dev = -1;
    KF_HIP_TRY(hipGetDevice(&dev));
    KF_REQUIRE(dev == P->device, KF_ERR_ARG, "plan was created on device %d, current is %d",
               P->device, dev);
    const int nm1 = P->nm1, ma1 = P->ma1, rep = P->rep;
    const int *d_offs = P->meta, *d_src1 = P->meta + nm1 + 1;
    const int *N1 = nullptr;  // atom counts are baked into the plan (offs, row lists)
    double *erow = P->acc, *Fat = P->acc + (size_t)nm1 * ma1;
    hipLaunchKernelGGL(zero_fill, dim3(256), dim3(256), 0, stream, P->acc, P->acc_doubles);
    const double cexp = -1.0 / (2.0 * P->sigma * P->sigma), inv_s2 = 1.0 / (P->sigma * P->sigma);
    int rc;
    for (const auto &L : P->labels) {
        const double *P2 = P->train + L.train_off;
        const double *nrm2 = P2 + (size_t)2 * L.R2 * rep, *rdot = nrm2 + L.R2, *aEr = rdot + L.R2;
        double *T = P->query;
        double *Xq = T + (size_t)L.QB * 2 * L.R2, *G = Xq + (size_t)L.QB * rep;
        double *nq = G + (size_t)L.QB * rep, *wsum = nq + L.QB;
        for (int q0 = 0; q0 < L.R1; q0 += L.QB) {
            const int qb = std::min(L.QB, L.R1 - q0);
            const int *qsrc = d_src1 + L.r1b + q0;
            hipLaunchKernelGGL(gather_rows, dim3((qb + 3) / 4), dim3(NT), 0, stream, X1,
                               (const double *)nullptr, qsrc, qb, rep, Xq, (double *)nullptr, nq,
                               (double *)nullptr);
            if ((rc = gemm::gemm_f64<true>(qb, 2 * L.R2, rep, 1.0, Xq, rep, P2, rep, 0.0, T,
                                           (size_t)2 * L.R2, stream)))
                return rc;
            hipLaunchKernelGGL(coef_kernel, dim3((qb + 3) / 4), dim3(NT), 0, stream, T, qb, L.R2,
                               nq, nrm2, rdot, aEr, cexp, inv_s2, qsrc, erow, wsum);
            if ((rc = gemm::gemm_f64<false>(qb, rep, 2 * L.R2, 1.0, T, (size_t)2 * L.R2, P2, rep,
                                            0.0, G, rep, stream)))
                return rc;
            hipLaunchKernelGGL(backproj_plan_kernel, dim3((qb + 3) / 4), dim3(NT), 0, stream, G,
                               Xq, wsum, dX1, qsrc, d_offs, qb, ma1, rep, Fat);
            KF_HIP_TRY(hipGetLastError());
        }
    }
    (void)N1;
    hipLaunchKernelGGL(reduce_plan_kernel, dim3(nm1), dim3(128), 0, stream, erow, Fat, d_offs,
                       nm1, ma1, E, F);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_matvec_plan_create(const int *Q1, const int *N1, int nm1, int max_atoms1,
                                     const double *X2, const double *alpha_desc,
                                     const double *alpha_E, const int *Q2, const int *N2,
                                     int nm2, int max_atoms2, int rep, double sigma,
                                     kf_matvec_plan **plan, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    KF_REQUIRE(plan, KF_ERR_ARG, "null plan out-pointer");
    *plan = nullptr;
    auto *P = new (std::nothrow) kf_matvec_plan();
    KF_REQUIRE(P, KF_ERR_NOMEM, "plan allocation failed");
    const int rc = plan_create(Q1, N1, nm1, max_atoms1, X2, alpha_desc, alpha_E, Q2, N2, nm2,
                               max_atoms2, rep, sigma, P, as_stream(stream));
    if (rc) {
        delete P;
        return rc;
    }
    *plan = P;
    return KF_OK;
}

extern "C" int kf_matvec_plan_execute(const kf_matvec_plan *plan, const double *X1,
                                      const double *dX1, double *E, double *F, void *stream) {
    return plan_execute(plan, X1, dX1, E, F, as_stream(stream));
}

extern "C" int kf_matvec_plan_destroy(kf_matvec_plan *plan) {
    if (plan) {
        (void)hipDeviceSynchronize();  // no execution may still read the plan's memory
        delete plan;
    }
    return KF_OK;
}

extern "C" int kf_local_alpha_desc(const double *dX, const int *Q, const int *N, int nm,
                                   int max_atoms, int rep, const double *alpha, double *out,
                                   void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    (void)Q;
    KF_REQUIRE(nm >= 0 && max_atoms >= 0 && rep >= 1, KF_ERR_ARG, "bad dimensions");
    if (nm == 0 || max_atoms == 0) return KF_OK;
    KF_REQUIRE(dX && N && alpha && out, KF_ERR_ARG, "null pointer");
    hipStream_t st = as_stream(stream);
    std::vector<int> hN(nm);
    int *stage = static_cast<int *>(host_staging(WS_AUX0, sizeof(int) * (nm + 1)));
    KF_REQUIRE(stage, KF_ERR_NOMEM, "staging allocation failed");
    KF_HIP_TRY(hipMemcpyAsync(stage, N, sizeof(int) * nm, hipMemcpyDeviceToHost, st));
    KF_HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> offs(nm + 1, 0);
    for (int a = 0; a < nm; ++a) offs[a + 1] = offs[a] + 3 * clamp_atoms(stage[a], max_atoms);
    int *d_offs = static_cast<int *>(workspace(WS_PLAN, sizeof(int) * (nm + 1), st));
    int *st2 = static_cast<int *>(host_staging(WS_PLAN, sizeof(int) * (nm + 1)));
    KF_REQUIRE(d_offs && st2, KF_ERR_NOMEM, "workspace allocation failed");
    std::copy(offs.begin(), offs.end(), st2);
    KF_HIP_TRY(hipMemcpyAsync(d_offs, st2, sizeof(int) * (nm + 1), hipMemcpyHostToDevice, st));
    const long atoms = (long)nm * max_atoms;
    hipLaunchKernelGGL(alpha_desc_kernel, dim3((unsigned)((atoms + 3) / 4)), dim3(NT), 0, st, dX,
                       N, d_offs, nm, max_atoms, rep, alpha, out);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

extern "C" int kf_local_full_matvec(const double *X1, const double *dX1, const double *X2,
                                    const double *alpha_desc, const double *alpha_E,
                                    const int *Q1, const int *Q2, const int *N1, const int *N2,
                                    int nm1, int nm2, int max_atoms1, int max_atoms2, int rep,
                                    double sigma, double *E, double *F, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return matvec_impl(X1, dX1, X2, alpha_desc, alpha_E, Q1, Q2, N1, N2, nm1, nm2, max_atoms1,
                       max_atoms2, rep, sigma, E, F, 1 | 2 | 4, as_stream(stream));
}

extern "C" int kf_local_hessian_matvec(const double *X1, const double *dX1, const double *X2,
                                       const double *alpha_desc, const int *Q1, const int *Q2,
                                       const int *N1, const int *N2, int nm1, int nm2,
                                       int max_atoms1, int max_atoms2, int rep, double sigma,
                                       double *F, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return matvec_impl(X1, dX1, X2, alpha_desc, nullptr, Q1, Q2, N1, N2, nm1, nm2, max_atoms1,
                       max_atoms2, rep, sigma, nullptr, F, 2, as_stream(stream));
}

extern "C" int kf_local_jacobian_t_matvec(const double *X1, const double *X2,
                                          const double *alpha_desc, const int *Q1,
                                          const int *Q2, const int *N1, const int *N2, int nm1,
                                          int nm2, int max_atoms1, int max_atoms2, int rep,
                                          double sigma, double *E, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return matvec_impl(X1, nullptr, X2, alpha_desc, nullptr, Q1, Q2, N1, N2, nm1, nm2,
                       max_atoms1, max_atoms2, rep, sigma, E, nullptr, 1, as_stream(stream));
}
