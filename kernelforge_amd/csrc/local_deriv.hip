// local_deriv.hip -- energy+force forms of the element-matched local Gaussian
// kernel on gfx950: Jacobian (dK/dR2), transposed Jacobian (-dK/dR1), Hessian and
// the full [[EE, EF], [FE, FF]] kernel, dense or symmetric RFP (TRANSR='N', UPLO='U').
//
// For a molecule pair (a, b) and every label-matched atom pair m = (i, j):
//   d_m = x_ai - x_bj,  e_m = exp(-|d_m|^2 / 2 s^2),  c_m = e_m / s^2,  w_m = e_m / s^4
//   EE = sum_m e_m
//   FE = -sum_m c_m J_ai^T d_m                       (dK/dR_a, 3 n_a)
//   EF = +sum_m c_m J_bj^T d_m                       (dK/dR_b, 3 n_b)
//   FF = sum_i J_ai^T T_i - sum_m w_m v_m u_m^T      (3 n_a x 3 n_b)
//        T_i = sum_{j matched to i} c_m J_bj,  v_m = J_ai^T d_m,  u_m = J_bj^T d_m
// which is the reference's per-pair DGEMV/DGER + S_sum DGEMM scheme
// (src/local_kernels.cpp:1171-1419 derivation, 2511-2815 full RFP) regrouped so
// that the dominant term is one long contraction over K = (feature, atom).
//
// Two kernels per batch of molecule pairs, one 256-thread work-group per pair:
//  * deriv_prep_kernel (FP64 VALU): the matched list, e_m, and the projections of
//    J onto difference vectors streamed over the features (v_m, u_m for the
//    Hessian forms; s_i = sum_j c_m d_m and t_j = sum_i c_m d_m for the E-only
//    Jacobian forms).  Writes EE / EF / FE, and e_m, v_m, u_m to a per-pair slot.
//  * deriv_ff_kernel (FP64 MFMA 16x16x4): per feature f, T_f = C J_b[:,f,:] and
//    FF += J_a[:,f,:]^T T_f, both on MFMA with operands streamed from HBM into
//    registers (T_f's accumulators are the B operands of the second product), then
//    the rank-1 sum as a second MFMA pass over the slot; writes FF.
// Every output entry belongs to exactly one molecule pair, so no atomics are used
// and results are deterministic.
#include <algorithm>
#include <cmath>
#include <vector>

#include "feature_compact.hpp"
#include "kf_common.hpp"
#include "local_plan.hpp"

namespace kf {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int NT = 256;            // threads per work-group (4 waves)
constexpr int MAXA = 32;           // atoms per molecule supported here (3*MAXA <= 96 columns)
constexpr int MAXM = MAXA * MAXA;  // matched atom pairs per molecule pair
constexpr int VCH = 12;            // difference vectors per projection pass
constexpr int FCH = 128;           // features per staged projection chunk

enum DerivMode {
    M_JAC = 0,         // (nm1, naq2)        EF
    M_JAC_T,           // (naq1, nm2)        -FE
    M_HESS,            // (naq1, naq2)       FF
    M_FULL,            // (nm1+naq1, nm2+naq2)
    M_HESS_SYMM,       // (naq, naq) dense, pairs a <= b mirrored
    M_HESS_SYMM_RFP,   // RFP of order naq
    M_FULL_SYMM,       // (nm+naq)^2 dense
    M_FULL_SYMM_RFP,   // RFP of order nm+naq
    M_DIAG_BLOCKS,     // per molecule b: the (E_b, F_b) diagonal block, bs x bs, bs = 1 + 3 ma
};
__host__ __device__ constexpr bool has_ff(int m) { return m >= M_HESS; }
__host__ __device__ constexpr bool has_ee(int m) {
    return m == M_FULL || m == M_FULL_SYMM || m == M_FULL_SYMM_RFP || m == M_DIAG_BLOCKS;
}
__host__ __device__ constexpr bool has_ef(int m) { return m == M_JAC || has_ee(m); }
__host__ __device__ constexpr bool has_fe(int m) { return m == M_JAC_T || has_ee(m); }
__host__ __device__ constexpr bool is_symm(int m) { return m >= M_HESS_SYMM && m != M_DIAG_BLOCKS; }
__host__ __device__ constexpr bool is_rfp(int m) {
    return m == M_HESS_SYMM_RFP || m == M_FULL_SYMM_RFP;
}

struct DerivArgs {
    const double *x1, *dx1, *x2, *dx2;
    const int *q1, *q2, *n1, *n2;
    const int *offs1, *offs2;  // 3 * prefix sums of clamped atom counts
    const int2 *pairs;         // (a, b) per work-group
    int ma1, ma2, rep;
    double cexp, inv_s2, inv_s4;
    double *out;
    size_t ld;     // dense row stride, or the RFP order
    int ebase1;    // energy rows preceding the force rows (nm1 for the full forms)
    int ebase2;    // energy columns preceding the force columns
    double *slot;  // per-pair scratch: e[maxm] | V[maxm][ldv] | U[maxm][ldu]
    size_t slot_len;
    int maxm, ldv, ldu;
};

__device__ inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }


__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Work-group -> pair index.  Consecutive pairs share their first molecule; the
// hardware deals work-groups round-robin over the 8 XCDs, so this keeps runs of
// consecutive pairs (and their J_a) on one XCD's L2.
__device__ inline int pair_slot() {
    const int g = blockIdx.x, G = gridDim.x;
    return (G % 8 == 0) ? (g % 8) * (G / 8) + g / 8 : g;
}

// ---------------------------------------------------------------------------
// Output addressing per form.  (a, b) is the pair, oa/ob the force offsets.
// ---------------------------------------------------------------------------
template <int MODE>
struct Out {
    double *o;
    size_t ld;
    int a, b, oa, ob, e1, e2;
    __device__ void put(size_t r, size_t c, double v) const {
        if (MODE == M_DIAG_BLOCKS)
            o[((size_t)a * ld + r) * ld + c] = v;  // block a, local (r, c)
        else if (is_rfp(MODE))
            o[rfp_upper(ld, r, c)] = v;  // caller guarantees r <= c
        else
            o[r * ld + c] = v;
    }
    __device__ void ee(double v) const {
        if (MODE == M_DIAG_BLOCKS) put(0, 0, v);
        if (MODE == M_FULL) put(a, b, v);
        if (MODE == M_FULL_SYMM) { put(a, b, v); put(b, a, v); }
        if (MODE == M_FULL_SYMM_RFP) put(a, b, v);
    }
    // EF[r] = dK/dR_b; row a, force column ob + r.
    __device__ void ef(int r, double v) const {
        const size_t col = (size_t)e2 + ob + r;
        if (MODE == M_DIAG_BLOCKS) put(0, 1 + r, v);
        if (MODE == M_JAC || MODE == M_FULL || MODE == M_FULL_SYMM_RFP) put(a, col, v);
        if (MODE == M_FULL_SYMM) {
            put(a, col, v);
            if (a != b) put(col, a, v);
        }
    }
    // FE[r] = dK/dR_a; force row oa + r, column b.
    __device__ void fe(int r, double v) const {
        const size_t row = (size_t)e1 + oa + r;
        if (MODE == M_DIAG_BLOCKS) put(1 + r, 0, v);
        if (MODE == M_JAC_T) put(oa + r, b, -v);
        if (MODE == M_FULL) put(row, b, v);
        if (MODE == M_FULL_SYMM) {
            put(row, b, v);
            if (a != b) put(b, row, v);
        }
        if (MODE == M_FULL_SYMM_RFP && a != b) put(b, row, v);
    }
    __device__ void ff(int r, int t, double v) const {
        const size_t row = (size_t)e1 + oa + r, col = (size_t)e2 + ob + t;
        if (MODE == M_DIAG_BLOCKS) put(1 + r, 1 + t, v);
        if (MODE == M_HESS || MODE == M_FULL) put(row, col, v);
        if (MODE == M_HESS_SYMM || MODE == M_FULL_SYMM) {
            put(row, col, v);
            if (a != b) put(col, row, v);
        }
        if (MODE == M_HESS_SYMM_RFP || MODE == M_FULL_SYMM_RFP) {
            if (a != b || r <= t) put(row, col, v);
        }
    }
};

// Matched list of one pair, built by thread 0 in i-major order (m -> (i, j)),
// with per-i ranges and a j-major permutation.
struct PairLists {
    int ml[MAXM];        // (i << 8) | j
    int ist[MAXA + 1];   // i-major ranges
    int jst[MAXA + 1];   // j-major ranges into jl
    int jl[MAXM];        // m indices grouped by j (ascending i inside a group)
    int qa[MAXA], qb[MAXA];
    int cj[MAXA];        // per-j fill cursors
    int np;
};

__device__ void build_lists(PairLists &L, const int *q1, const int *q2, int a, int b, int ma1,
                            int ma2, int na, int nb) {
    const int tid = threadIdx.x;
    if (tid < na) L.qa[tid] = q1[(size_t)a * ma1 + tid];
    if (tid < nb) L.qb[tid] = q2[(size_t)b * ma2 + tid];
    __syncthreads();
    if (tid == 0) {
        int cnt = 0;
        int *cj = L.cj;
        for (int j = 0; j < nb; ++j) cj[j] = 0;
        for (int i = 0; i < na; ++i) {
            L.ist[i] = cnt;
            const int qi = L.qa[i];
            for (int j = 0; j < nb; ++j)
                if (L.qb[j] == qi) { L.ml[cnt++] = (i << 8) | j; cj[j]++; }
        }
        L.ist[na] = cnt;
        int acc = 0;
        for (int j = 0; j < nb; ++j) { L.jst[j] = acc; acc += cj[j]; cj[j] = L.jst[j]; }
        L.jst[nb] = acc;
        for (int m = 0; m < cnt; ++m) L.jl[cj[L.ml[m] & 255]++] = m;
        L.np = cnt;
    }
    __syncthreads();
}

// acc[v] (for this thread's column p and feature slice) += sum_f J[f, p] vec_v[f],
// vectors produced chunk-wise by mk(v, f); emit(v, p, pb, value) after the
// cross-slice reduction.  J rows have stride ldj; ncol columns are live.
template <int NV, class MakeVec, class Emit>
__device__ void project(const double *J, int ldj, int ncol, int rep, int nv, MakeVec mk,
                        Emit emit, double (*dv)[VCH], double (*red)[VCH][64]) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int pb = 0; pb * 64 < ncol; ++pb) {
        const int p = pb * 64 + lane;
        double acc[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[v] = 0.0;
        for (int f0 = 0; f0 < rep; f0 += FCH) {
            const int fn = min(FCH, rep - f0);
            __syncthreads();
            for (int e = tid; e < nv * FCH; e += NT) {
                const int v = e / FCH, f = e - v * FCH;
                dv[f][v] = f < fn ? mk(v, f0 + f) : 0.0;
            }
            __syncthreads();
            if (p < ncol) {
                // four independent row loads in flight per thread
                const double *jp = J + (size_t)f0 * ldj + p;
                int f = wv;
                for (; f + 12 < fn; f += 16) {
                    const double j0 = jp[(size_t)f * ldj], j1 = jp[(size_t)(f + 4) * ldj],
                                 j2 = jp[(size_t)(f + 8) * ldj], j3 = jp[(size_t)(f + 12) * ldj];
#pragma unroll
                    for (int v = 0; v < NV; ++v) {
                        double t = fma(j0, dv[f][v], acc[v]);
                        t = fma(j1, dv[f + 4][v], t);
                        t = fma(j2, dv[f + 8][v], t);
                        acc[v] = fma(j3, dv[f + 12][v], t);
                    }
                }
                for (; f < fn; f += 4) {
                    const double jv = jp[(size_t)f * ldj];
#pragma unroll
                    for (int v = 0; v < NV; ++v) acc[v] = fma(jv, dv[f][v], acc[v]);
                }
            }
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) red[wv][v][lane] = acc[v];
        __syncthreads();
        for (int v = wv; v < nv; v += 4)
            if (p < ncol)
                emit(v, p, pb, red[0][v][lane] + red[1][v][lane] + red[2][v][lane] +
                                   red[3][v][lane]);
        __syncthreads();
    }
}

template <int MODE>
__global__ __launch_bounds__(NT) void deriv_prep_kernel(DerivArgs A, int pair0) {
    __shared__ PairLists L;
    __shared__ double em[MAXM];
    __shared__ __align__(16) double dv[FCH][VCH];  // feature-major: one row per feature
    __shared__ double red[4][VCH][64];
    __shared__ double part[4][2 * 64];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int t = pair0 + pair_slot();
    const int a = A.pairs[t].x, b = A.pairs[t].y;
    const int na = clampi(A.n1[a], 0, A.ma1), nb = clampi(A.n2[b], 0, A.ma2);
    const int rep = A.rep, nca = 3 * na, ncb = 3 * nb;
    build_lists(L, A.q1, A.q2, a, b, A.ma1, A.ma2, na, nb);
    const int np = L.np;
    const double *xa0 = A.x1 + (size_t)a * A.ma1 * rep;
    const double *xb0 = A.x2 + (size_t)b * A.ma2 * rep;

    // e_m from the exact squared distance (one wave per matched pair).
    for (int m = wv; m < np; m += 4) {
        const double *xa = xa0 + (size_t)(L.ml[m] >> 8) * rep;
        const double *xb = xb0 + (size_t)(L.ml[m] & 255) * rep;
        double s = 0.0;
        for (int f = lane; f < rep; f += 64) {
            const double d = xa[f] - xb[f];
            s = fma(d, d, s);
        }
        s = wave_sum(s);
        if (lane == 0) em[m] = exp(s * A.cexp);
    }
    __syncthreads();

    Out<MODE> out{A.out, A.ld, a, b, A.offs1[a], A.offs2[b], A.ebase1, A.ebase2};
    double *slot = has_ff(MODE) ? A.slot + (size_t)(t - pair0) * A.slot_len : nullptr;
    if (has_ff(MODE))
        for (int m = tid; m < np; m += NT) slot[m] = em[m];
    if (has_ee(MODE) && tid == 0) {
        double s = 0.0;
        for (int m = 0; m < np; ++m) s += em[m];
        out.ee(s);
    }
    const double inv_s2 = A.inv_s2;
    // Per-thread partial sums of FE (pb 0/1 for column pb*64+lane).
    double fe_acc[2] = {0.0, 0.0};
    double ef_acc[2] = {0.0, 0.0};

    // ---------------- A side: J_ai projections ----------------
    if (has_fe(MODE) || has_ff(MODE)) {
        const size_t lda = 3 * (size_t)A.ma1;
        for (int i = 0; i < na; ++i) {
            const int m0 = L.ist[i], m1 = L.ist[i + 1];
            if (m1 == m0) continue;
            const double *J = A.dx1 + ((size_t)a * A.ma1 + i) * rep * lda;
            const double *xa = xa0 + (size_t)i * rep;
            if (has_ff(MODE)) {
                double *V = slot + A.maxm;
                for (int c0 = m0; c0 < m1; c0 += VCH) {
                    const int nv = min(VCH, m1 - c0);
                    project<VCH>(
                        J, (int)lda, nca, rep, nv,
                        [&](int v, int f) {
                            return xa[f] - xb0[(size_t)(L.ml[c0 + v] & 255) * rep + f];
                        },
                        [&](int v, int p, int pb, double val) {
                            V[(size_t)(c0 + v) * A.ldv + p] = val;
                            const double c = em[c0 + v] * inv_s2 * val;
                            if (pb == 0) fe_acc[0] -= c; else fe_acc[1] -= c;
                        },
                        dv, red);
                }
            } else {
                project<1>(
                    J, (int)lda, nca, rep, 1,
                    [&](int, int f) {
                        double s = 0.0;
                        for (int m = m0; m < m1; ++m)
                            s += em[m] * (xa[f] - xb0[(size_t)(L.ml[m] & 255) * rep + f]);
                        return s * inv_s2;
                    },
                    [&](int, int, int pb, double val) {
                        if (pb == 0) fe_acc[0] -= val; else fe_acc[1] -= val;
                    },
                    dv, red);
            }
        }
    }
    // ---------------- B side: J_bj projections ----------------
    if (has_ff(MODE) || has_ef(MODE)) {
        const size_t ldb = 3 * (size_t)A.ma2;
        for (int j = 0; j < nb; ++j) {
            const int g0 = L.jst[j], g1 = L.jst[j + 1];
            if (g1 == g0) continue;
            const double *J = A.dx2 + ((size_t)b * A.ma2 + j) * rep * ldb;
            const double *xb = xb0 + (size_t)j * rep;
            if (has_ff(MODE)) {
                double *U = slot + A.maxm + (size_t)A.maxm * A.ldv;
                for (int c0 = g0; c0 < g1; c0 += VCH) {
                    const int nv = min(VCH, g1 - c0);
                    project<VCH>(
                        J, (int)ldb, ncb, rep, nv,
                        [&](int v, int f) {
                            return xa0[(size_t)(L.ml[L.jl[c0 + v]] >> 8) * rep + f] - xb[f];
                        },
                        [&](int v, int p, int pb, double val) {
                            const int m = L.jl[c0 + v];
                            U[(size_t)m * A.ldu + p] = val;
                            const double c = em[m] * inv_s2 * val;
                            if (pb == 0) ef_acc[0] += c; else ef_acc[1] += c;
                        },
                        dv, red);
                }
            } else {
                project<1>(
                    J, (int)ldb, ncb, rep, 1,
                    [&](int, int f) {
                        double s = 0.0;
                        for (int g = g0; g < g1; ++g) {
                            const int m = L.jl[g];
                            s += em[m] * (xa0[(size_t)(L.ml[m] >> 8) * rep + f] - xb[f]);
                        }
                        return s * inv_s2;
                    },
                    [&](int, int, int pb, double val) {
                        if (pb == 0) ef_acc[0] += val; else ef_acc[1] += val;
                    },
                    dv, red);
            }
        }
    }
    // ---------------- reduce the partial FE / EF and write ----------------
    if (has_fe(MODE)) {
        part[wv][lane] = fe_acc[0];
        part[wv][64 + lane] = fe_acc[1];
        __syncthreads();
        for (int r = tid; r < nca; r += NT)
            out.fe(r, part[0][r] + part[1][r] + part[2][r] + part[3][r]);
        __syncthreads();
    }
    if (has_ef(MODE)) {
        part[wv][lane] = ef_acc[0];
        part[wv][64 + lane] = ef_acc[1];
        __syncthreads();
        for (int r = tid; r < ncb; r += NT)
            out.ef(r, part[0][r] + part[1][r] + part[2][r] + part[3][r]);
    }
}

// ---------------------------------------------------------------------------
// Projections for the Hessian forms on FP64 MFMA, pairs grouped by a shared molecule.
//   v_m = J_ai^T d_m,  u_m = J_bj^T d_m,  d_m = x_ai - x_bj   (63-vectors for 21 atoms)
// u of pair (a, b) at (i, j) is -v of the ordered pair (b, a) at (j, i), so one kernel
// computes v for ordered pairs (x, y): a group is a molecule x and up to PG partners
// y_k.  Side 0 groups the stored pairs (x, y) by their first molecule and writes V (and
// e_m); side 1 groups them by their second molecule (x = b, y = a) and writes U = -v.
// For one atom i of x the columns are every label-matched (k, j) of the group:
//   V_i (3 n_x x ncol) = J_xi^T (3 n_x x rep) * D (rep x ncol),  D[f][c] = x_xi[f] - x_yj[f]
// which streams each J_xi row once for the whole group (the one-pair form streamed it
// once per pair).  The columns depend only on the label of atom i, so they are listed
// once per label of x, and the (atom, pair of 16-column fragments) items of a label are
// dealt over the four waves; A operands (J rows, shared by the two fragments) and B
// operands (differences, formed in registers) come straight from HBM, 16 features ahead.
// ---------------------------------------------------------------------------
constexpr int PG = 8;  // partners per group

struct ProjGroup {
    int x, n, side, pad;
    int y[PG];     // partner molecules
    int slot[PG];  // batch-local slot of the stored pair
};

template <int NPF>
__global__ __launch_bounds__(NT) void deriv_vproj_kernel(DerivArgs A, const ProjGroup *groups) {
    __shared__ int qx[MAXA], qy[PG][MAXA], ny[PG], ist[PG][MAXA], rnk[PG][MAXA], rx[MAXA];
    __shared__ int cols[PG * MAXA], wbase[5], gy[PG], gslot[PG], xl[MAXA], s_nl;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, col = lane & 15, kr = lane >> 4;
    const ProjGroup *gp = groups + blockIdx.x;
    const bool s0 = gp->side == 0;
    // side 0: x from set 1, partners from set 2; side 1 the other way round
    const double *Xx = s0 ? A.x1 : A.x2, *Xy = s0 ? A.x2 : A.x1;
    const double *Jx = s0 ? A.dx1 : A.dx2;
    const int *Qx = s0 ? A.q1 : A.q2, *Qy = s0 ? A.q2 : A.q1;
    const int *Nx = s0 ? A.n1 : A.n2, *Ny = s0 ? A.n2 : A.n1;
    const int max_ = s0 ? A.ma1 : A.ma2, may = s0 ? A.ma2 : A.ma1;
    const int rep = A.rep, x = gp->x, n = gp->n;
    const int nx = clampi(Nx[x], 0, max_), ncx = 3 * nx;
    const int lda = 3 * max_;
    if (tid < MAXA) qx[tid] = tid < nx ? Qx[(size_t)x * max_ + tid] : -1;
    if (tid < PG) {
        gy[tid] = tid < n ? gp->y[tid] : 0;
        gslot[tid] = tid < n ? gp->slot[tid] : 0;
        ny[tid] = tid < n ? clampi(Ny[gp->y[tid]], 0, may) : 0;
    }
    __syncthreads();
    {
        const int k = tid >> 5, j = tid & 31;  // 8 x 32 = NT
        qy[k][j] = (k < n && j < ny[k]) ? Qy[(size_t)gy[k] * may + j] : -2;
    }
    __syncthreads();
    {
        // m-offset tables.  Side 0, pair (x, y): m(i, j) = ist[k][i] + rnk[k][j], with
        // ist = sum over atoms i' < i of x of the partner's count of label q(i') and rnk =
        // rank of j among the partner's atoms of its label.  Side 1, pair (y, x):
        // m(j, i) = ist[k][j] + rx[i], ist summing x's counts over the partner's atoms j' < j.
        const int k = tid >> 5, t = tid & 31;
        if (k < n) {
            int s = 0;
            if (s0) {
                for (int i2 = 0; i2 < t && i2 < nx; ++i2)
                    for (int j2 = 0; j2 < ny[k]; ++j2) s += qy[k][j2] == qx[i2];
            } else {
                for (int j2 = 0; j2 < t && j2 < ny[k]; ++j2)
                    for (int i2 = 0; i2 < nx; ++i2) s += qx[i2] == qy[k][j2];
            }
            ist[k][t] = s;
            int r = 0;
            for (int j2 = 0; j2 < t; ++j2) r += qy[k][j2] == qy[k][t];
            rnk[k][t] = r;
        }
        if (tid < MAXA) {
            int r = 0;
            for (int i2 = 0; i2 < tid; ++i2) r += qx[i2] == qx[tid];
            rx[tid] = r;
        }
    }
    const double *xx0 = Xx + (size_t)x * max_ * rep;
    const double *Jx0 = Jx + (size_t)x * max_ * rep * lda;
    int pofs[NPF];
#pragma unroll
    for (int y = 0; y < NPF; ++y) pofs[y] = min(16 * y + col, max(ncx, 1) - 1);

    // one pass per label of x (taken at its first atom, rx == 0): the columns are every
    // label-matched (k, j) of the group, the same for all atoms of x with that label, and
    // the (atom, column fragment) items of the label are dealt over the four waves
    __syncthreads();  // tables ready (rx decides the passes below, uniformly)
    for (int i0 = 0; i0 < nx; ++i0) {
        if (rx[i0] != 0) continue;  // not the first atom of its label (uniform)
        __syncthreads();  // the previous label's columns and atoms consumed
        const int l = qx[i0];
        {
            // columns: every (k, j) with q_y[k][j] == l, k-major
            const int k = tid >> 5, j = tid & 31;
            const bool v = qy[k][j] == l;
            const unsigned long long bal = __ballot(v);
            const int pre = __popcll(bal & ((1ull << lane) - 1));
            if (lane == 0) wbase[wv + 1] = __popcll(bal);
            // atoms of x with label l, in order (rx is the rank among them)
            if (tid < MAXA && tid < nx && qx[tid] == l) xl[rx[tid]] = tid;
            if (wv == 0) {
                const unsigned long long bx = __ballot(tid < MAXA && tid < nx && qx[tid] == l);
                if (lane == 0) s_nl = __popcll(bx);
            }
            __syncthreads();
            if (tid == 0) {
                wbase[0] = 0;
                for (int w = 1; w <= 4; ++w) wbase[w] += wbase[w - 1];
            }
            __syncthreads();
            if (v) cols[wbase[wv] + pre] = (k << 5) | j;
            __syncthreads();
        }
        const int ncol = wbase[4];
        const int ncf = (ncol + 15) >> 4;
        // items: an atom i of the label and a pair of column fragments (2 zp, 2 zp + 1); the
        // two fragments share the streamed J_xi rows (eight MFMAs per k-step on one load of
        // the A operands).  A second fragment past the last is computed on clamped columns
        // and not stored.
        const int npf2 = (ncf + 1) >> 1;
        const int nitems = s_nl * npf2;
        for (int it = wv; it < nitems; it += 4) {
            const int i = xl[it / npf2], zp = it % npf2;
            const double *xi = xx0 + (size_t)i * rep;
            const double *Ji = Jx0 + (size_t)i * rep * lda;
            const double *xj[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = min(16 * (2 * zp + h) + col, ncol - 1);
                const int kc = cols[c] >> 5, jc = cols[c] & 31;
                xj[h] = Xy + ((size_t)gy[kc] * may + jc) * rep;
            }
            v4d acc[2][NPF];
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int y = 0; y < NPF; ++y) acc[h][y] = (v4d){0.0, 0.0, 0.0, 0.0};
            double d2[2] = {0.0, 0.0};
            // 16 features (4 k-steps) per block; the next block is loaded during this one
            double ja[2][4][NPF], da[2][4], db[2][2][4];
#define KF_VP_LOAD(S, F0)                                                                      \
    do {                                                                                       \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                        \
            const int f_ = min((F0) + 4 * u + kr, rep - 1);                                    \
            da[S][u] = xi[f_];                                                                 \
            db[S][0][u] = xj[0][f_];                                                           \
            db[S][1][u] = xj[1][f_];                                                           \
            _Pragma("unroll") for (int y = 0; y < NPF; ++y)                                    \
                ja[S][u][y] = Ji[(size_t)f_ * lda + pofs[y]];                                  \
        }                                                                                      \
    } while (0)
#define KF_VP_MMA(S, F0)                                                                       \
    do {                                                                                       \
        _Pragma("unroll") for (int u = 0; u < 4; ++u) {                                        \
            const bool in_ = (F0) + 4 * u + kr < rep;                                          \
            _Pragma("unroll") for (int h = 0; h < 2; ++h) {                                    \
                const double d_ = in_ ? da[S][u] - db[S][h][u] : 0.0;                          \
                d2[h] = fma(d_, d_, d2[h]);                                                    \
                _Pragma("unroll") for (int y = 0; y < NPF; ++y)                                \
                    acc[h][y] = __builtin_amdgcn_mfma_f64_16x16x4f64(ja[S][u][y], d_, acc[h][y], 0, 0, 0); \
            }                                                                                  \
        }                                                                                      \
    } while (0)
            KF_VP_LOAD(0, 0);
            for (int f0 = 0; f0 < rep; f0 += 32) {
                KF_VP_LOAD(1, f0 + 16);
                KF_VP_MMA(0, f0);
                if (f0 + 16 >= rep) break;
                KF_VP_LOAD(0, f0 + 32);
                KF_VP_MMA(1, f0 + 16);
            }
#undef KF_VP_LOAD
#undef KF_VP_MMA
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                // |d|^2 of column c: the four lanes kr = 0..3 hold quarter sums
                double q2 = d2[h];
                q2 += __shfl_xor(q2, 16, 64);
                q2 += __shfl_xor(q2, 32, 64);
                const int cc = 16 * (2 * zp + h) + col;
                if (cc >= ncol) continue;
                const int k = cols[cc] >> 5, j = cols[cc] & 31;
                double *slot = A.slot + (size_t)gslot[k] * A.slot_len;
                if (s0) {
                    const int m = ist[k][i] + rnk[k][j];
                    double *V = slot + A.maxm + (size_t)m * A.ldv;
                    if (kr == 0) slot[m] = exp(q2 * A.cexp);
                    // acc[h][y][v] is row p = 16y + 4v + kr of column cc
#pragma unroll
                    for (int y = 0; y < NPF; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int p = 16 * y + 4 * v + kr;
                            if (p < ncx) V[p] = acc[h][y][v];
                        }
                } else {
                    const int m = ist[k][j] + rx[i];
                    double *U = slot + A.maxm + (size_t)A.maxm * A.ldv + (size_t)m * A.ldu;
#pragma unroll
                    for (int y = 0; y < NPF; ++y)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            const int p = 16 * y + 4 * v + kr;
                            if (p < ncx) U[p] = -acc[h][y][v];
                        }
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// FF = sum_f J_a[:,f,:]^T (C J_b[:,f,:])  -  sum_m w_m v_m u_m^T   on FP64 MFMA,
// C[i][j] = c_ij for label-matched atoms (0 otherwise).  Per feature f both factors
// are MFMA products:
//   T_f = C x J_b[:,f,q-block]      (A = C from LDS, B = J_b rows straight from HBM)
//   FF += J_a[:,f,:]^T x T_f        (A = J_a[:,f,:] staged in LDS, B = T_f)
// The f64 16x16x4 accumulator layout (register v of lane l = row 4v + l/16, column
// l%16) is exactly the B-operand layout of k-step v (lane l supplies row 4v + l/16,
// column l%16), so T_f's accumulators feed the second product without a shuffle.
// Wave w owns q-block w (16 force columns of b) and all p-blocks of a.  J_a[:,f,:] is
// shared by the four waves: it is staged through two LDS buffers (one barrier per
// feature); T for f+1 is formed while the product of f runs and the operands of f+2
// are loaded during f.  Atoms are padded to 4*NK (rows of C past na / columns past nb
// are zero, the matching Jacobian rows are clamped reads), force columns to 16*NP
// (clamped reads, never stored).  Every global load is unconditional.
// ---------------------------------------------------------------------------
template <int MODE, int NK, int NP>
__device__ __forceinline__ void ff_pair(const DerivArgs &A, int a, int b, int na, int nb,
                                        const double *Cm, double *Jbuf, int np,
                                        const double *mc, const double *slot) {
    constexpr int NX = (NK + 3) / 4;         // 16-row fragments of C / T (atoms of a)
    constexpr int FFR = 4 * NK, FFP = 16 * NP;  // staged J_a feature: rows x columns
    constexpr int NEL = FFR * FFP, NST = (NEL + NT - 1) / NT;  // elements, per thread
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, col = lane & 15,
              kr = lane >> 4;
    const int rep = A.rep, nca = 3 * na, ncb = 3 * nb;
    const int nqb = (ncb + 15) >> 4;
    const int lda = 3 * A.ma1, ldb = 3 * A.ma2;
    const double *dxa = A.dx1 + (size_t)a * A.ma1 * rep * lda;
    const double *dxb = A.dx2 + (size_t)b * A.ma2 * rep * ldb;
    // staging map: element e = tid + NT*u of a feature -> row e / FFP, column e % FFP
    int soff[NST];
#pragma unroll
    for (int u = 0; u < NST; ++u) {
        const int e = min(tid + NT * u, NEL - 1), r = e / FFP, c = e % FFP;
        soff[u] = min(r, na - 1) * rep * lda + min(c, nca - 1);
    }
    // column-operand rows 4s + kr of b: affine in s except the last k-step (clamped)
    const int sb = 4 * rep * ldb, rb0 = kr * rep * ldb;
    const int rbL = min(4 * (NK - 1) + kr, nb - 1) * rep * ldb;
    Out<MODE> out{A.out, A.ld, a, b, A.offs1[a], A.offs2[b], A.ebase1, A.ebase2};
    // rounds of four q-blocks (more than one only past 21 atoms of b)
    for (int qr = 0; 4 * qr < nqb; ++qr) {
    const int qb = 4 * qr + wv;
    const bool busy = qb < nqb;
    const int q = min(16 * qb + col, ncb - 1);

    v4d acc[NP];
#pragma unroll
    for (int y = 0; y < NP; ++y) acc[y] = (v4d){0.0, 0.0, 0.0, 0.0};
    double st[NST], jb[2][NK];
    v4d tt[2][NX];
#define KF_FF_STAGE_LOAD(F)                                                                   \
    do {                                                                                      \
        const double *pa_ = dxa + (size_t)(F) * lda;                                          \
        _Pragma("unroll") for (int u = 0; u < NST; ++u) st[u] = pa_[soff[u]];                 \
    } while (0)
#define KF_FF_STAGE_STORE(BUF)                                                                \
    do {                                                                                      \
        _Pragma("unroll") for (int u = 0; u < NST; ++u)                                       \
            if (NEL % NT == 0 || tid + NT * u < NEL) (BUF)[tid + NT * u] = st[u];             \
    } while (0)
#define KF_FF_LOAD_B(S, F)                                                                    \
    do {                                                                                      \
        const double *pb_ = dxb + (size_t)(F) * ldb + q;                                      \
        _Pragma("unroll") for (int s = 0; s < NK; ++s)                                        \
            jb[S][s] = pb_[s < NK - 1 ? rb0 + s * sb : rbL];                                  \
    } while (0)
#define KF_FF_T(S)                                                                            \
    do {                                                                                      \
        _Pragma("unroll") for (int x = 0; x < NX; ++x) tt[S][x] = (v4d){0.0, 0.0, 0.0, 0.0};  \
        _Pragma("unroll") for (int s = 0; s < NK; ++s)                                        \
            _Pragma("unroll") for (int x = 0; x < NX; ++x)                                    \
                tt[S][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(                              \
                    Cm[(16 * x + col) * 33 + 4 * s + kr], jb[S][s], tt[S][x], 0, 0, 0);       \
    } while (0)
#define KF_FF_MAIN(S, BUF)                                                                    \
    do {                                                                                      \
        _Pragma("unroll") for (int s = 0; s < NK; ++s)                                        \
            _Pragma("unroll") for (int y = 0; y < NP; ++y)                                    \
                acc[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(                                \
                    (BUF)[(4 * s + kr) * FFP + 16 * y + col], tt[S][s >> 2][s & 3], acc[y],   \
                    0, 0, 0);                                                                 \
    } while (0)
    // prologue: J_a(0) in buffer 0, J_a(1) in registers, T(0), J_b(1)
    double *const J0 = Jbuf, *const J1 = Jbuf + NEL;
    KF_FF_STAGE_LOAD(0);
    KF_FF_STAGE_STORE(J0);
    KF_FF_STAGE_LOAD(min(1, rep - 1));
    KF_FF_LOAD_B(0, 0);
    KF_FF_T(0);
    KF_FF_LOAD_B(1, min(1, rep - 1));
    __syncthreads();
    for (int f = 0; f < rep; f += 2) {
        // feature f: buffer J0, T slot 0; prepare f+1 (J1, T slot 1) and loads for f+2
        KF_FF_STAGE_STORE(J1);
        KF_FF_STAGE_LOAD(min(f + 2, rep - 1));
        if (busy) {
            KF_FF_T(1);
            KF_FF_MAIN(0, J0);
        }
        KF_FF_LOAD_B(0, min(f + 2, rep - 1));
        __syncthreads();
        if (f + 1 >= rep) break;
        // feature f+1: buffer J1, T slot 1; prepare f+2 (J0, T slot 0) and loads for f+3
        KF_FF_STAGE_STORE(J0);
        KF_FF_STAGE_LOAD(min(f + 3, rep - 1));
        if (busy) {
            KF_FF_T(0);
            KF_FF_MAIN(1, J1);
        }
        KF_FF_LOAD_B(1, min(f + 3, rep - 1));
        __syncthreads();
    }
#undef KF_FF_STAGE_LOAD
#undef KF_FF_STAGE_STORE
#undef KF_FF_LOAD_B
#undef KF_FF_T
#undef KF_FF_MAIN
    if (!busy) continue;
    // rank-1 part: acc -= sum_m (w_m v_m) u_m^T, K = matched pairs
    {
        const double *V = slot + A.maxm, *U = V + (size_t)A.maxm * A.ldv;
        // loads of clamped (valid) entries, unconditional, masked by selects on '&'
        // conditions ('&&' turns into a branch around each load, which is then waited
        // for before the next one); np >= 1 and nca, ncb >= 1 inside the loop
        for (int m0 = 0; m0 < np; m0 += 4) {
            const int m = m0 + kr;
            const bool ok = m < np;
            const int mm = min(m, np - 1);
            const double w = ok ? -mc[mm] * A.inv_s2 : 0.0;  // -e/s^4
            const int qq = 16 * qb + col;
            const double u = U[(size_t)mm * A.ldu + min(qq, ncb - 1)];
            const double bv = (ok & (qq < ncb)) ? u : 0.0;
            double vv[NP];
#pragma unroll
            for (int y = 0; y < NP; ++y) vv[y] = V[(size_t)mm * A.ldv + min(16 * y + col, nca - 1)];
#pragma unroll
            for (int y = 0; y < NP; ++y) {
                const int p = 16 * y + col;
                const double av = (ok & (p < nca)) ? w * vv[y] : 0.0;
                acc[y] = __builtin_amdgcn_mfma_f64_16x16x4f64(av, bv, acc[y], 0, 0, 0);
            }
        }
    }
    const int tc = 16 * qb + col;
#pragma unroll
    for (int y = 0; y < NP; ++y)
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int r = 16 * y + mfma16_row_f64(lane, v);
            if (r < nca && tc < ncb) out.ff(r, tc, acc[y][v]);
        }
    }
}

// Shape buckets by m = max(atoms of a, atoms of b): (NK, NP) and the LDS they stage.
inline void ff_bucket(int m, int &nk, int &np) {
    if (m <= 5) { nk = 2; np = 1; }
    else if (m <= 10) { nk = 3; np = 2; }
    else if (m <= 16) { nk = 4; np = 3; }
    else if (m <= 21) { nk = 6; np = 4; }
    else if (m <= 24) { nk = 6; np = 5; }
    else { nk = 8; np = 6; }
}
inline size_t ff_lds_bytes(int ma1, int ma2) {
    int nk, np;
    ff_bucket(ma1 > ma2 ? ma1 : ma2, nk, np);
    return 2 * sizeof(double) * (size_t)(4 * nk) * (16 * np);
}

// Shape buckets by m = max(atoms of a, atoms of b): NK k-steps of 4 atoms, NP 16-column
// force blocks (3 m columns).
template <int MODE>
__global__ __launch_bounds__(NT, 2) void deriv_ff_kernel(DerivArgs A, int pair0) {
    __shared__ double Cm[32 * 33];
    extern __shared__ double Jbuf[];  // two staged J_a features of the largest bucket
    __shared__ int qa[MAXA], qb_[MAXA], ist[MAXA + 1], mi[MAXM], mj[MAXM];
    __shared__ double mc[MAXM];
    __shared__ int s_np;
    const int tid = threadIdx.x;
    const int t = pair0 + pair_slot();
    const int a = A.pairs[t].x, b = A.pairs[t].y;
    const int na = clampi(A.n1[a], 0, A.ma1), nb = clampi(A.n2[b], 0, A.ma2);
    // a molecule without atoms still gets its zero EE / FE / EF entries written
    const double *slot = A.slot + (size_t)(t - pair0) * A.slot_len;
    if (tid < na) qa[tid] = A.q1[(size_t)a * A.ma1 + tid];
    if (tid < nb) qb_[tid] = A.q2[(size_t)b * A.ma2 + tid];
    for (int e = tid; e < 32 * 33; e += NT) Cm[e] = 0.0;
    __syncthreads();
    if (tid == 0) {  // matched list, i-major: the order of the slot entries
        int cnt = 0;
        for (int i = 0; i < na; ++i) {
            ist[i] = cnt;
            for (int j = 0; j < nb; ++j)
                if (qb_[j] == qa[i]) { mi[cnt] = i; mj[cnt] = j; ++cnt; }
        }
        ist[na] = cnt;
        s_np = cnt;
    }
    __syncthreads();
    const int np = s_np;
    for (int m = tid; m < np; m += NT) {
        const double c = slot[m] * A.inv_s2;
        mc[m] = c;
        Cm[mi[m] * 33 + mj[m]] = c;
    }
    __syncthreads();
    {
        // EE = sum_m e_m, FE = -sum_m c_m v_m, EF = sum_m c_m u_m from the slot
        Out<MODE> out{A.out, A.ld, a, b, A.offs1[a], A.offs2[b], A.ebase1, A.ebase2};
        const double *V = slot + A.maxm, *U = V + (size_t)A.maxm * A.ldv;
        if (has_ee(MODE) && tid == 0) {
            double e = 0.0;
            for (int m = 0; m < np; ++m) e += slot[m];
            out.ee(e);
        }
        if (has_fe(MODE) && tid < 3 * na) {
            double sfe = 0.0;
            for (int m = 0; m < np; ++m) sfe -= mc[m] * V[(size_t)m * A.ldv + tid];
            out.fe(tid, sfe);
        }
        if (has_ef(MODE) && tid >= 128 && tid - 128 < 3 * nb) {
            double sef = 0.0;
            for (int m = 0; m < np; ++m) sef += mc[m] * U[(size_t)m * A.ldu + tid - 128];
            out.ef(tid - 128, sef);
        }
    }
    if (na == 0 || nb == 0) return;
    const int mx = na > nb ? na : nb;
    if (mx <= 5) ff_pair<MODE, 2, 1>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
    else if (mx <= 10) ff_pair<MODE, 3, 2>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
    else if (mx <= 16) ff_pair<MODE, 4, 3>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
    else if (mx <= 21) ff_pair<MODE, 6, 4>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
    else if (mx <= 24) ff_pair<MODE, 6, 5>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
    else ff_pair<MODE, 8, 6>(A, a, b, na, nb, Cm, Jbuf, np, mc, slot);
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct SideMeta {
    std::vector<int> q, n, offs;  // offs: nm + 1
    std::vector<int> hist;        // per molecule label histogram over `labels`
    int naq = 0;
};

void side_meta(const std::vector<int> &hQ, const std::vector<int> &hN, int nm, int ma,
               const std::vector<int> &labels, SideMeta &s) {
    s.q = hQ;
    s.n = hN;
    s.offs.assign(nm + 1, 0);
    const int L = (int)labels.size();
    s.hist.assign((size_t)nm * L, 0);
    int acc = 0;
    for (int m = 0; m < nm; ++m) {
        const int na = std::min(std::max(hN[m], 0), ma);
        s.offs[m] = acc;
        acc += 3 * na;
        for (int i = 0; i < na; ++i) {
            const int l = int(std::lower_bound(labels.begin(), labels.end(), hQ[(size_t)m * ma + i]) -
                              labels.begin());
            if (l < L && labels[l] == hQ[(size_t)m * ma + i]) s.hist[(size_t)m * L + l]++;
        }
    }
    s.offs[nm] = acc;
    s.naq = acc;
}

template <int MODE>
int deriv_launch(DerivArgs A, const std::vector<int2> &pairs, const SideMeta &s1,
                 const SideMeta &s2, int nlab, hipStream_t stream) {
    const size_t npairs = pairs.size();
    if (npairs == 0) return KF_OK;
    // offsets + pair list to the device (one staging copy)
    const size_t n_off = s1.offs.size() + s2.offs.size();
    const size_t pair_off = (sizeof(int) * n_off + 15) / 16 * 16;
    const size_t bytes = pair_off + sizeof(int2) * npairs;
    char *dev = static_cast<char *>(workspace(WS_PLAN, bytes, stream));
    char *stage = static_cast<char *>(host_staging(WS_PLAN, bytes));
    KF_REQUIRE(dev && stage, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", bytes);
    std::copy(s1.offs.begin(), s1.offs.end(), reinterpret_cast<int *>(stage));
    std::copy(s2.offs.begin(), s2.offs.end(), reinterpret_cast<int *>(stage) + s1.offs.size());
    std::copy(pairs.begin(), pairs.end(), reinterpret_cast<int2 *>(stage + pair_off));
    KF_HIP_TRY(hipMemcpyAsync(dev, stage, bytes, hipMemcpyHostToDevice, stream));
    A.offs1 = reinterpret_cast<int *>(dev);
    A.offs2 = A.offs1 + s1.offs.size();
    A.pairs = reinterpret_cast<const int2 *>(dev + pair_off);

    size_t batch = npairs;
    if (has_ff(MODE)) {
        // slot layout and the largest matched count over the pairs
        int maxm = 1;
        for (const int2 &p : pairs) {
            int c = 0;
            for (int l = 0; l < nlab; ++l)
                c += s1.hist[(size_t)p.x * nlab + l] * s2.hist[(size_t)p.y * nlab + l];
            maxm = std::max(maxm, c);
        }
        A.maxm = maxm;
        A.ldv = (3 * A.ma1 + 15) / 16 * 16;
        A.ldu = (3 * A.ma2 + 15) / 16 * 16;
        A.slot_len = (size_t)maxm * (1 + A.ldv + A.ldu);
        const size_t budget = size_t(1) << 30;
        batch = std::max<size_t>(1, std::min(npairs, budget / (A.slot_len * sizeof(double))));
        A.slot = static_cast<double *>(
            workspace(WS_AUX2, batch * A.slot_len * sizeof(double), stream));
        KF_REQUIRE(A.slot, KF_ERR_NOMEM, "scratch allocation failed");
    } else {
        batch = std::min<size_t>(npairs, size_t(1) << 20);
    }
    // Hessian forms: the projection groups of every batch, built once and uploaded once
    // (side 0: runs of stored pairs with the same first molecule; side 1: the same pairs
    // grouped by their second molecule)
    std::vector<size_t> gstart;
    const ProjGroup *dgroups = nullptr;
    if (has_ff(MODE)) {
        std::vector<ProjGroup> groups;
        std::vector<std::pair<int, int>> byb;
        for (size_t p0 = 0; p0 < npairs; p0 += batch) {
            const int nbat = (int)std::min(batch, npairs - p0);
            gstart.push_back(groups.size());
            for (int k = 0; k < nbat;) {
                ProjGroup g{};
                g.x = pairs[p0 + k].x;
                g.side = 0;
                while (k < nbat && pairs[p0 + k].x == g.x && g.n < PG) {
                    g.y[g.n] = pairs[p0 + k].y;
                    g.slot[g.n++] = k++;
                }
                groups.push_back(g);
            }
            byb.clear();
            for (int k = 0; k < nbat; ++k) byb.emplace_back(pairs[p0 + k].y, k);
            std::stable_sort(byb.begin(), byb.end(),
                             [](const std::pair<int, int> &u, const std::pair<int, int> &v) {
                                 return u.first < v.first;
                             });
            for (size_t k = 0; k < byb.size();) {
                ProjGroup g{};
                g.x = byb[k].first;
                g.side = 1;
                while (k < byb.size() && byb[k].first == g.x && g.n < PG) {
                    g.y[g.n] = pairs[p0 + byb[k].second].x;
                    g.slot[g.n++] = byb[k].second;
                    ++k;
                }
                groups.push_back(g);
            }
        }
        gstart.push_back(groups.size());
        const size_t gb = sizeof(ProjGroup) * groups.size();
        ProjGroup *dg = static_cast<ProjGroup *>(workspace(WS_AUX3, gb, stream));
        ProjGroup *hg = static_cast<ProjGroup *>(host_staging(WS_AUX3, gb));
        KF_REQUIRE(dg && hg, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", gb);
        std::copy(groups.begin(), groups.end(), hg);
        KF_HIP_TRY(hipMemcpyAsync(dg, hg, gb, hipMemcpyHostToDevice, stream));
        dgroups = dg;
    }
    const int npf = (3 * std::max(A.ma1, A.ma2) + 15) / 16;
    size_t bi = 0;
    for (size_t p0 = 0; p0 < npairs; p0 += batch, ++bi) {
        const unsigned nb = (unsigned)std::min(batch, npairs - p0);
        if (!has_ff(MODE)) {
            hipLaunchKernelGGL(deriv_prep_kernel<MODE>, dim3(nb), dim3(NT), 0, stream, A, (int)p0);
            KF_HIP_TRY(hipGetLastError());
        }
        if (has_ff(MODE)) {
            const unsigned ng = (unsigned)(gstart[bi + 1] - gstart[bi]);
            const ProjGroup *gp = dgroups + gstart[bi];
            switch (npf) {
                case 1: hipLaunchKernelGGL(deriv_vproj_kernel<1>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
                case 2: hipLaunchKernelGGL(deriv_vproj_kernel<2>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
                case 3: hipLaunchKernelGGL(deriv_vproj_kernel<3>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
                case 4: hipLaunchKernelGGL(deriv_vproj_kernel<4>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
                case 5: hipLaunchKernelGGL(deriv_vproj_kernel<5>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
                default: hipLaunchKernelGGL(deriv_vproj_kernel<6>, dim3(ng), dim3(NT), 0, stream, A, gp); break;
            }
            KF_HIP_TRY(hipGetLastError());
            hipLaunchKernelGGL(deriv_ff_kernel<MODE>, dim3(nb), dim3(NT), ff_lds_bytes(A.ma1, A.ma2),
                               stream, A, (int)p0);
            KF_HIP_TRY(hipGetLastError());
        }
    }
    return KF_OK;
}

// ---------------------------------------------------------------------------
// Any molecule size: the derivative forms for sets with more than MAXA atoms per molecule.
// One work-group per molecule pair, every pair-size-dependent list in global scratch:
//   matched list (i, j) of equal labels, i-major; e_m = exp(-|d_m|^2 / 2 s^2),
//   d_m = x_ai - x_bj;  v_m = J_ai^T d_m (3 na), u_m = J_bj^T d_m (3 nb);
//   EE = sum e_m, FE = -sum c_m v_m, EF = sum c_m u_m (c = e / s^2), and
//   FF = sum_f J_a[:, f, :]^T T_f - sum_m (e_m / s^4) v_m u_m^T with
//   T_f[i, :] = sum_{m in row i} c_m J_b[j_m, f, :],
// FF in 64 x 64 output tiles, T_f staged in LDS per (feature, 64-row block of i).  The
// blocks land through the same Out<MODE> addressing as the fixed-size kernels.  VALU
// FP64: it carries the sizes the MFMA kernels above do not take, not their speed.
// ---------------------------------------------------------------------------
constexpr int GT = 64;  // output tile edge / staged i-rows

template <int MODE>
__global__ __launch_bounds__(NT) void deriv_generic_kernel(DerivArgs A, int pair0,
                                                           int *__restrict__ iscr,
                                                           size_t iscr_len) {
    __shared__ double T[GT][GT + 1];
    __shared__ int s_np;
    const int tid = threadIdx.x;
    const int t = pair0 + blockIdx.x;
    const int a = A.pairs[t].x, b = A.pairs[t].y;
    const int na = clampi(A.n1[a], 0, A.ma1), nb = clampi(A.n2[b], 0, A.ma2);
    const int ld1 = 3 * A.ma1, ld2 = 3 * A.ma2, rep = A.rep;
    double *slot = A.slot + (size_t)blockIdx.x * A.slot_len;
    double *em = slot, *V = slot + A.maxm, *U = V + (size_t)A.maxm * A.ldv;
    int *mi = iscr + (size_t)blockIdx.x * iscr_len, *mj = mi + A.maxm, *ist = mj + A.maxm;
    if (tid == 0) {  // matched list, i-major
        int cnt = 0;
        for (int i = 0; i < na; ++i) {
            ist[i] = cnt;
            const int qi = A.q1[(size_t)a * A.ma1 + i];
            for (int j = 0; j < nb; ++j)
                if (A.q2[(size_t)b * A.ma2 + j] == qi) { mi[cnt] = i; mj[cnt] = j; ++cnt; }
        }
        ist[na] = cnt;
        s_np = cnt;
    }
    __syncthreads();
    const int np = s_np;
    const double *x1 = A.x1 + (size_t)a * A.ma1 * rep, *x2 = A.x2 + (size_t)b * A.ma2 * rep;
    for (int m = tid; m < np; m += NT) {
        const double *p = x1 + (size_t)mi[m] * rep, *q = x2 + (size_t)mj[m] * rep;
        double s2 = 0.0;
        for (int f = 0; f < rep; ++f) {
            const double d = p[f] - q[f];
            s2 = fma(d, d, s2);
        }
        em[m] = exp(A.cexp * s2);
    }
    const bool want_v = has_fe(MODE) || has_ff(MODE), want_u = has_ef(MODE) || has_ff(MODE);
    if (want_v)
        for (long idx = tid; idx < (long)np * 3 * na; idx += NT) {
            const int m = (int)(idx / (3 * na)), r = (int)(idx - (long)m * 3 * na);
            const double *p = x1 + (size_t)mi[m] * rep, *q = x2 + (size_t)mj[m] * rep;
            const double *J = A.dx1 + ((size_t)a * A.ma1 + mi[m]) * rep * ld1 + r;
            double s = 0.0;
            for (int f = 0; f < rep; ++f) s = fma(J[(size_t)f * ld1], p[f] - q[f], s);
            V[(size_t)m * A.ldv + r] = s;
        }
    if (want_u)
        for (long idx = tid; idx < (long)np * 3 * nb; idx += NT) {
            const int m = (int)(idx / (3 * nb)), c = (int)(idx - (long)m * 3 * nb);
            const double *p = x1 + (size_t)mi[m] * rep, *q = x2 + (size_t)mj[m] * rep;
            const double *J = A.dx2 + ((size_t)b * A.ma2 + mj[m]) * rep * ld2 + c;
            double s = 0.0;
            for (int f = 0; f < rep; ++f) s = fma(J[(size_t)f * ld2], p[f] - q[f], s);
            U[(size_t)m * A.ldu + c] = s;
        }
    __syncthreads();
    Out<MODE> out{A.out, A.ld, a, b, A.offs1[a], A.offs2[b], A.ebase1, A.ebase2};
    if (has_ee(MODE) && tid == 0) {
        double e = 0.0;
        for (int m = 0; m < np; ++m) e += em[m];
        out.ee(e);
    }
    if (has_fe(MODE))
        for (int r = tid; r < 3 * na; r += NT) {
            double s = 0.0;
            for (int m = 0; m < np; ++m) s -= em[m] * A.inv_s2 * V[(size_t)m * A.ldv + r];
            out.fe(r, s);
        }
    if (has_ef(MODE))
        for (int c = tid; c < 3 * nb; c += NT) {
            double s = 0.0;
            for (int m = 0; m < np; ++m) s += em[m] * A.inv_s2 * U[(size_t)m * A.ldu + c];
            out.ef(c, s);
        }
    if (!has_ff(MODE) || na == 0 || nb == 0) return;
    const int tr = tid >> 4, tc = tid & 15;
    const double *J1 = A.dx1 + (size_t)a * A.ma1 * rep * ld1;
    const double *J2 = A.dx2 + (size_t)b * A.ma2 * rep * ld2;
    for (int r0 = 0; r0 < 3 * na; r0 += GT)
        for (int c0 = 0; c0 < 3 * nb; c0 += GT) {
            double acc[4][4];
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) acc[u][w] = 0.0;
            for (int f = 0; f < rep; ++f)
                for (int i0 = 0; i0 < na; i0 += GT) {
                    __syncthreads();  // the previous T tile is no longer read
                    for (int e = tid; e < GT * GT; e += NT) {
                        const int ii = e / GT, cc = e - ii * GT;
                        const int i = i0 + ii, c = c0 + cc;
                        double s = 0.0;
                        if (i < na && c < 3 * nb)
                            for (int m = ist[i]; m < ist[i + 1]; ++m)
                                s = fma(em[m] * A.inv_s2,
                                        J2[((size_t)mj[m] * rep + f) * ld2 + c], s);
                        T[ii][cc] = s;
                    }
                    __syncthreads();
                    const int iend = min(GT, na - i0);
                    for (int ii = 0; ii < iend; ++ii) {
                        const double *jr = J1 + ((size_t)(i0 + ii) * rep + f) * ld1;
                        double jv[4];
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int r = r0 + tr + 16 * u;
                            jv[u] = r < 3 * na ? jr[r] : 0.0;
                        }
#pragma unroll
                        for (int w = 0; w < 4; ++w) {
                            const double tv = T[ii][tc + 16 * w];
#pragma unroll
                            for (int u = 0; u < 4; ++u) acc[u][w] = fma(jv[u], tv, acc[u][w]);
                        }
                    }
                }
            // rank-one terms, then write the tile
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int w = 0; w < 4; ++w) {
                    const int r = r0 + tr + 16 * u, c = c0 + tc + 16 * w;
                    if (r >= 3 * na || c >= 3 * nb) continue;
                    double s = acc[u][w];
                    for (int m = 0; m < np; ++m)
                        s -= em[m] * A.inv_s4 * V[(size_t)m * A.ldv + r] * U[(size_t)m * A.ldu + c];
                    out.ff(r, c, s);
                }
        }
}

template <int MODE>
int deriv_launch_generic(DerivArgs A, const std::vector<int2> &pairs, const SideMeta &s1,
                         const SideMeta &s2, int nlab, hipStream_t stream) {
    const size_t npairs = pairs.size();
    if (npairs == 0) return KF_OK;
    const size_t n_off = s1.offs.size() + s2.offs.size();
    const size_t pair_off = (sizeof(int) * n_off + 15) / 16 * 16;
    const size_t bytes = pair_off + sizeof(int2) * npairs;
    char *dev = static_cast<char *>(workspace(WS_PLAN, bytes, stream));
    char *stage = static_cast<char *>(host_staging(WS_PLAN, bytes));
    KF_REQUIRE(dev && stage, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", bytes);
    std::copy(s1.offs.begin(), s1.offs.end(), reinterpret_cast<int *>(stage));
    std::copy(s2.offs.begin(), s2.offs.end(), reinterpret_cast<int *>(stage) + s1.offs.size());
    std::copy(pairs.begin(), pairs.end(), reinterpret_cast<int2 *>(stage + pair_off));
    KF_HIP_TRY(hipMemcpyAsync(dev, stage, bytes, hipMemcpyHostToDevice, stream));
    A.offs1 = reinterpret_cast<int *>(dev);
    A.offs2 = A.offs1 + s1.offs.size();
    A.pairs = reinterpret_cast<const int2 *>(dev + pair_off);
    int maxm = 1;
    for (const int2 &p : pairs) {
        int c = 0;
        for (int l = 0; l < nlab; ++l)
            c += s1.hist[(size_t)p.x * nlab + l] * s2.hist[(size_t)p.y * nlab + l];
        maxm = std::max(maxm, c);
    }
    A.maxm = maxm;
    A.ldv = 3 * A.ma1;
    A.ldu = 3 * A.ma2;
    A.slot_len = (size_t)maxm * (1 + A.ldv + A.ldu);
    const size_t iscr_len = 2 * (size_t)maxm + A.ma1 + 1;
    const size_t per_pair = A.slot_len * sizeof(double) + iscr_len * sizeof(int);
    const size_t budget = size_t(1) << 30;
    const size_t batch = std::max<size_t>(1, std::min(npairs, budget / per_pair));
    char *ws = static_cast<char *>(workspace(WS_AUX2, batch * per_pair + 256, stream));
    KF_REQUIRE(ws, KF_ERR_NOMEM, "scratch allocation failed");
    A.slot = reinterpret_cast<double *>(ws);
    int *iscr = reinterpret_cast<int *>(ws + batch * A.slot_len * sizeof(double));
    for (size_t p0 = 0; p0 < npairs; p0 += batch) {
        const unsigned nb = (unsigned)std::min(batch, npairs - p0);
        hipLaunchKernelGGL(deriv_generic_kernel<MODE>, dim3(nb), dim3(NT), 0, stream, A, (int)p0,
                           iscr, iscr_len);
        KF_HIP_TRY(hipGetLastError());
    }
    return KF_OK;
}

// Drop the descriptor columns that are zero for every atom of both sets (in X and in the
// Jacobians the mode reads): A then points at compacted copies.
int compact_features(DerivArgs &A, const int *N1, const int *N2, int nm1, int nm2,
                     hipStream_t stream) {
    CompactSets c;
    c.x1 = A.x1; c.dx1 = A.dx1; c.n1 = N1; c.nm1 = nm1; c.ma1 = A.ma1;
    c.x2 = A.x2; c.dx2 = A.dx2; c.n2 = N2; c.nm2 = nm2; c.ma2 = A.ma2;
    int nf = A.rep;
    const int rc = compact_columns(c, A.rep, WS_PACK, WS_PACK2, stream, &nf);
    if (rc) return rc;
    A.x1 = c.x1; A.dx1 = c.dx1; A.x2 = c.x2; A.dx2 = c.dx2;
    A.rep = nf;
    return KF_OK;
}

template <int MODE>
int deriv_impl(const double *X1, const double *dX1, const double *X2, const double *dX2,
               const int *Q1, const int *Q2, const int *N1, const int *N2, int nm1, int nm2,
               int ma1, int ma2, int rep, double sigma, double *out, hipStream_t stream) {
    KF_REQUIRE(std::isfinite(sigma) && sigma > 0.0, KF_ERR_SIGMA, "sigma must be > 0");
    KF_REQUIRE(nm1 >= 0 && nm2 >= 0 && ma1 >= 0 && ma2 >= 0 && rep >= 1, KF_ERR_ARG,
               "bad dimensions");
    // more than MAXA atoms per molecule: the size-general kernel
    const bool general = ma1 > MAXA || ma2 > MAXA;
    if (nm1 == 0 || nm2 == 0) return KF_OK;
    KF_REQUIRE(X1 && X2 && Q1 && Q2 && N1 && N2 && out, KF_ERR_ARG, "null pointer");
    const bool need1 = has_fe(MODE) || has_ff(MODE), need2 = has_ef(MODE) || has_ff(MODE);
    KF_REQUIRE((!need1 || dX1) && (!need2 || dX2), KF_ERR_ARG, "null Jacobian pointer");
    std::vector<int> q1, n1, q2, n2;
    int rc;
    if ((rc = read_metadata(Q1, N1, nm1, ma1, q1, n1, stream, WS_AUX0))) return rc;
    if (is_symm(MODE) || MODE == M_DIAG_BLOCKS) {
        q2 = q1;
        n2 = n1;
    } else if ((rc = read_metadata(Q2, N2, nm2, ma2, q2, n2, stream, WS_AUX1))) {
        return rc;
    }
    std::vector<int> labels;
    {
        std::vector<int> l1 = distinct_labels(q1.data(), n1.data(), nm1, ma1),
                         l2 = distinct_labels(q2.data(), n2.data(), nm2, ma2);
        labels.resize(l1.size() + l2.size());
        labels.erase(std::set_union(l1.begin(), l1.end(), l2.begin(), l2.end(), labels.begin()),
                     labels.end());
    }
    SideMeta s1, s2;
    side_meta(q1, n1, nm1, ma1, labels, s1);
    side_meta(q2, n2, nm2, ma2, labels, s2);

    DerivArgs A{};
    // a Jacobian the mode does not use is not read (nor compacted)
    A.x1 = X1; A.dx1 = need1 ? dX1 : nullptr; A.x2 = X2; A.dx2 = need2 ? dX2 : nullptr;
    A.q1 = Q1; A.q2 = Q2; A.n1 = N1; A.n2 = N2;
    A.ma1 = ma1; A.ma2 = ma2; A.rep = rep;
    A.cexp = -1.0 / (2.0 * sigma * sigma);
    A.inv_s2 = 1.0 / (sigma * sigma);
    A.inv_s4 = A.inv_s2 * A.inv_s2;
    A.out = out;
    switch (MODE) {
        case M_JAC: A.ld = s2.naq; break;
        case M_JAC_T: A.ld = nm2; break;
        case M_HESS: A.ld = s2.naq; break;
        case M_FULL: A.ld = (size_t)nm2 + s2.naq; A.ebase1 = nm1; A.ebase2 = nm2; break;
        case M_HESS_SYMM:
        case M_HESS_SYMM_RFP: A.ld = s1.naq; break;
        case M_DIAG_BLOCKS: A.ld = 1 + 3 * (size_t)ma1; break;
        default: A.ld = (size_t)nm1 + s1.naq; A.ebase1 = nm1; A.ebase2 = nm1; break;
    }
    std::vector<int2> pairs;
    if (MODE == M_DIAG_BLOCKS) {
        pairs.reserve(nm1);
        for (int a = 0; a < nm1; ++a) pairs.push_back(make_int2(a, a));
    } else if (is_symm(MODE)) {
        pairs.reserve((size_t)nm1 * (nm1 + 1) / 2);
        for (int a = 0; a < nm1; ++a)
            for (int b = a; b < nm1; ++b) pairs.push_back(make_int2(a, b));
    } else {
        pairs.reserve((size_t)nm1 * nm2);
        for (int a = 0; a < nm1; ++a)
            for (int b = 0; b < nm2; ++b) pairs.push_back(make_int2(a, b));
    }
    // the Hessian forms of zero-size sets have nothing to write
    if ((MODE == M_HESS || MODE == M_HESS_SYMM || MODE == M_HESS_SYMM_RFP) &&
        (s1.naq == 0 || s2.naq == 0))
        return KF_OK;
    if ((rc = compact_features(A, N1, N2, nm1, nm2, stream))) return rc;
    ScopedTimer timer(TM_FULL, stream);
    if (general) return deriv_launch_generic<MODE>(A, pairs, s1, s2, (int)labels.size(), stream);
    return deriv_launch<MODE>(A, pairs, s1, s2, (int)labels.size(), stream);
}

}  // namespace
}  // namespace kf

using namespace kf;

extern "C" int kf_local_kernel_jacobian(const double *X1, const double *X2, const double *dX2,
                                        const int *Q1, const int *Q2, const int *N1,
                                        const int *N2, int nm1, int nm2, int max_atoms1,
                                        int max_atoms2, int rep, double sigma, double *K,
                                        void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_JAC>(X1, nullptr, X2, dX2, Q1, Q2, N1, N2, nm1, nm2, max_atoms1,
                             max_atoms2, rep, sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_jacobian_t(const double *X1, const double *X2, const double *dX1,
                                          const int *Q1, const int *Q2, const int *N1,
                                          const int *N2, int nm1, int nm2, int max_atoms1,
                                          int max_atoms2, int rep, double sigma, double *K,
                                          void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_JAC_T>(X1, dX1, X2, nullptr, Q1, Q2, N1, N2, nm1, nm2, max_atoms1,
                               max_atoms2, rep, sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_hessian(const double *X1, const double *X2, const double *dX1,
                                       const double *dX2, const int *Q1, const int *Q2,
                                       const int *N1, const int *N2, int nm1, int nm2,
                                       int max_atoms1, int max_atoms2, int rep, double sigma,
                                       double *K, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_HESS>(X1, dX1, X2, dX2, Q1, Q2, N1, N2, nm1, nm2, max_atoms1, max_atoms2,
                              rep, sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_full(const double *X1, const double *X2, const double *dX1,
                                    const double *dX2, const int *Q1, const int *Q2,
                                    const int *N1, const int *N2, int nm1, int nm2,
                                    int max_atoms1, int max_atoms2, int rep, double sigma,
                                    double *K, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_FULL>(X1, dX1, X2, dX2, Q1, Q2, N1, N2, nm1, nm2, max_atoms1, max_atoms2,
                              rep, sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_hessian_symm(const double *X, const double *dX, const int *Q,
                                            const int *N, int nm, int max_atoms, int rep,
                                            double sigma, double *K, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_HESS_SYMM>(X, dX, X, dX, Q, Q, N, N, nm, nm, max_atoms, max_atoms, rep,
                                   sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_hessian_symm_rfp(const double *X, const double *dX, const int *Q,
                                                const int *N, int nm, int max_atoms, int rep,
                                                double sigma, double *arf, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_HESS_SYMM_RFP>(X, dX, X, dX, Q, Q, N, N, nm, nm, max_atoms, max_atoms,
                                       rep, sigma, arf, as_stream(stream));
}

extern "C" int kf_local_kernel_full_symm(const double *X, const double *dX, const int *Q,
                                         const int *N, int nm, int max_atoms, int rep,
                                         double sigma, double *K, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_FULL_SYMM>(X, dX, X, dX, Q, Q, N, N, nm, nm, max_atoms, max_atoms, rep,
                                   sigma, K, as_stream(stream));
}

extern "C" int kf_local_kernel_full_symm_rfp(const double *X, const double *dX, const int *Q,
                                             const int *N, int nm, int max_atoms, int rep,
                                             double sigma, double *arf, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return deriv_impl<M_FULL_SYMM_RFP>(X, dX, X, dX, Q, Q, N, N, nm, nm, max_atoms, max_atoms,
                                       rep, sigma, arf, as_stream(stream));
}

extern "C" int kf_local_full_diag_blocks(const double *X, const double *dX, const int *Q,
                                         const int *N, int nm, int max_atoms, int rep,
                                         double sigma, double *blocks, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    // entries beyond a molecule's 1 + 3 N[b] live rows/columns are not written
    return deriv_impl<M_DIAG_BLOCKS>(X, dX, X, dX, Q, Q, N, N, nm, nm, max_atoms, max_atoms, rep,
                                     sigma, blocks, as_stream(stream));
}
