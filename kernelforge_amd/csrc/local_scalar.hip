// local_scalar.hip -- element-matched local Gaussian kernel (scalar forms) on gfx950.
//
//   K[a,b] = sum_{i in a, j in b, Q_i == Q_j} exp(-(|x_i|^2 + |x_j|^2 - 2 x_i.x_j) / 2 sigma^2)
//
// Reference: kernel_gaussian (src/local_kernels.cpp:155-298), kernel_gaussian_symm_rfp
// (:403-577), kernel_gaussian_symm (:584-758).  The reference packs rows per label,
// runs DSYRK/DGEMM (alpha=-2) on 8192-row tiles and scatters an OpenMP exp-sum.
//
// gfx950 design
//  * Atoms are packed by (label, molecule) once (local_plan.hpp), so the rows of a
//    molecule block for one label are one contiguous range.
//  * The OUTPUT is tiled: one workgroup owns a 64x64 block of molecule pairs (upper
//    triangle of blocks for the symmetric forms).  For every label it runs the
//    rows(A-block) x rows(B-block) Gram product as 64x64 sub-tiles on FP64 MFMA
//    (v_mfma_f64_16x16x4, 2x2 waves of 32x32), with operands staged through LDS
//    and the next k-chunk prefetched into registers during the MFMAs.
//  * Epilogue fused per sub-tile: squared norms (FP64), exp, then a deterministic
//    two-stage segmented reduction (rows -> column-molecule runs in LDS, then row
//    molecules into per-thread register accumulators of the 64x64 block).  No
//    atomics; each K entry is written exactly once, straight into RFP (UPLO='U') or
//    dense layout.  64-molecule blocks keep the 64-row label sub-tiles ~81% full on
//    ragged molecules (32-molecule blocks: ~67%).
#include <algorithm>
#include <cmath>
#include <vector>

#include "local_mixed.hpp"

namespace kf {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int MB = 64;       // molecules per output block
constexpr int TS = 64;       // sub-tile edge (packed rows)
constexpr int KC = 16;       // k chunk staged in LDS
constexpr int LDA = KC + 1;  // LDS pitch of operand tiles (doubles)
constexpr int EP = TS + 1;   // LDS pitch of the exp tile
constexpr int NT = 256;      // threads per workgroup

enum Mode { SYMM_RFP = 0, SYMM_DENSE = 1, RECT = 2 };

// pack: Xp[row, k] = X[src[row], k] (k < rep), 0 for rep <= k < rep_pad;
// n2[row] = |x|^2 in FP64.  One wave per row.
__global__ __launch_bounds__(256) void pack_rows_f64(const double *__restrict__ X,
                                                     const int *__restrict__ src, int rows,
                                                     int rep, int rep_pad,
                                                     double *__restrict__ Xp,
                                                     double *__restrict__ n2) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const double *s = X + (size_t)src[row] * rep;
    double *d = Xp + (size_t)row * rep_pad;
    double acc = 0.0;
    for (int k = lane; k < rep_pad; k += 64) {
        const double v = k < rep ? s[k] : 0.0;
        d[k] = v;
        acc += v * v;
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) n2[row] = acc;
}


// One k-chunk (KC) of the wave's Gram fragments: NA 16-row fragments of the row tile
// times the wave's 16 columns of NTW column tiles, as v_mfma_f64_16x16x4 (A[i][k] from
// lane i + 16k, B[k][j] from lane j + 16k, D row 4v + lane/16, column lane % 16).
// The operands of k-step ks+1 are read from LDS while the MFMAs of ks issue.
template <int NA, int NTW>
__device__ __forceinline__ void mma_chunk(const double *__restrict__ As,
                                          const double *__restrict__ B0,
                                          const double *__restrict__ B1, v4d (&c)[2][4],
                                          int r16, int bcol, int kq) {
    if (NTW == 0) return;
    double a[2][NA], b[2][NTW];
    auto read = [&](int ks, int s) {
        const int kk = ks * 4 + kq;
#pragma unroll
        for (int x = 0; x < NA; ++x) a[s][x] = As[(x * 16 + r16) * LDA + kk];
        b[s][0] = B0[bcol * LDA + kk];
        if (NTW == 2) b[s][NTW - 1] = B1[bcol * LDA + kk];
    };
    read(0, 0);
#pragma unroll
    for (int ks = 0; ks < KC / 4; ++ks) {
        if (ks + 1 < KC / 4) read(ks + 1, (ks + 1) & 1);
#pragma unroll
        for (int u = 0; u < NTW; ++u)
#pragma unroll
            for (int x = 0; x < NA; ++x)
                c[u][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[ks & 1][x], b[ks & 1][u],
                                                               c[u][x], 0, 0, 0);
    }
}

// Operand pointers and predicates of one row tile and NTL column tiles.
struct GramArgs {
    const double *pa, *pb0, *pb1;
    bool va, vb0, vb1;
    int rep_pad, lofs, r16, bcol, kq;
};

// Gram products of one row sub-tile with NTL column sub-tiles over the full depth.
// Operands go global -> registers -> LDS with two LDS buffers: while the MFMAs run on
// chunk k, chunk k+1 is stored into the other buffer and chunk k+2 is loaded into
// registers, so each chunk costs one barrier.  Every wave stages its share of all NTL
// tiles but multiplies only its NA x NTW non-empty fragments.  On entry the caller's
// last barrier has freed both buffers; on exit a barrier has retired every read.
template <int NTL, int NA, int NTW>
__device__ void gram_tiles(double *__restrict__ Ops, const GramArgs &g, v4d (&c)[2][4]) {
    constexpr int LV = KC / 8, BUF = 3 * TS * LDA;
    double2 ra[LV], rb0[LV], rb1[LV];
    auto load = [&](int k) {
#pragma unroll
        for (int u = 0; u < LV; ++u) {
            ra[u] = g.va ? *(const double2 *)(g.pa + k + 2 * u) : double2{0, 0};
            rb0[u] = g.vb0 ? *(const double2 *)(g.pb0 + k + 2 * u) : double2{0, 0};
            if (NTL == 2) rb1[u] = g.vb1 ? *(const double2 *)(g.pb1 + k + 2 * u) : double2{0, 0};
        }
    };
    auto store = [&](double *buf) {
        double *as = buf + g.lofs, *bs0 = buf + TS * LDA + g.lofs;
        double *bs1 = buf + 2 * TS * LDA + g.lofs;
#pragma unroll
        for (int u = 0; u < LV; ++u) {
            as[2 * u] = ra[u].x;
            as[2 * u + 1] = ra[u].y;
            bs0[2 * u] = rb0[u].x;
            bs0[2 * u + 1] = rb0[u].y;
            if (NTL == 2) {
                bs1[2 * u] = rb1[u].x;
                bs1[2 * u + 1] = rb1[u].y;
            }
        }
    };
    const int nch = g.rep_pad / KC;
    load(0);
    store(Ops);
    if (nch > 1) load(KC);
    __syncthreads();
    for (int kc = 0; kc < nch; ++kc) {
        double *cur = Ops + (kc & 1) * BUF;
        if (kc + 1 < nch) {
            store(Ops + ((kc + 1) & 1) * BUF);
            if (kc + 2 < nch) load((kc + 2) * KC);
        }
        mma_chunk<NA, NTW>(cur, cur + TS * LDA, cur + 2 * TS * LDA, c, g.r16, g.bcol, g.kq);
        __syncthreads();
    }
}

template <int NTL, int NA>
__device__ __forceinline__ void gram_ntw(double *Ops, const GramArgs &g, int ntw,
                                         v4d (&c)[2][4]) {
    if (ntw == 0) gram_tiles<NTL, NA, 0>(Ops, g, c);
    else if (NTL == 1 || ntw == 1) gram_tiles<NTL, NA, 1>(Ops, g, c);
    else gram_tiles<NTL, NA, NTL>(Ops, g, c);
}

template <int NTL>
__device__ __forceinline__ void gram_dispatch(double *Ops, const GramArgs &g, int na16, int ntw,
                                              v4d (&c)[2][4]) {
    switch (na16) {
        case 1: gram_ntw<NTL, 1>(Ops, g, ntw, c); break;
        case 2: gram_ntw<NTL, 2>(Ops, g, ntw, c); break;
        case 3: gram_ntw<NTL, 3>(Ops, g, ntw, c); break;
        default: gram_ntw<NTL, 4>(Ops, g, ntw, c); break;
    }
}

// Output block: 64 x 64 molecule pairs per work-group.  The block's K entries live in
// registers: thread t owns column molecule b = t % 64 and row molecules
// 16 (t / 64) .. +15.  Per label, the row sub-tile sr is multiplied with two column
// sub-tiles at once (the row operand is staged once for both, 3 operand panels per 2
// output tiles instead of 4); wave w computes columns 16w .. 16w+15 of both tiles for
// every non-empty 16-row fragment.  Each tile's exp tile is then reduced to row-wise run
// sums over column molecules (a run is the columns of one molecule; four threads per
// row, the partial sums of a run that crosses a 16-column quarter carried across the
// four lanes), and each thread adds the rows of its 16 row molecules for its column
// molecule's run.  The molecule of a packed row is src[row] / max_atoms
// (src = a * max_atoms + i, see local_plan.hpp).
template <int MODE>
__global__ __launch_bounds__(NT, 2) void local_scalar_f64_kernel(
    const double *__restrict__ Xp1, const double *__restrict__ n2a, const int *__restrict__ mr1,
    const int *__restrict__ src1, int ma1, const double *__restrict__ Xp2,
    const double *__restrict__ n2b, const int *__restrict__ mr2, const int *__restrict__ src2,
    int ma2, int L, int nm1, int nm2, int rep_pad, double cexp, const int2 *__restrict__ order,
    double *__restrict__ out) {
    // two operand buffers (row tile + two column tiles each); the exp tile and the run
    // sums of the epilogue reuse the same memory once the Gram products are done
    constexpr int OPS = 2 * 3 * TS * LDA, EPI = 2 * TS * EP;
    __shared__ double Sh[OPS > EPI ? OPS : EPI];
    double *const Ops = Sh, *const Es = Sh, *const Rs = Sh + TS * EP;
    __shared__ double n2r[TS], n2c[2][TS];
    __shared__ int rowmol[TS], colmol[2][TS], colrun[2][TS + 1], run_of_mol[2][MB];
    __shared__ int mrA[MB + 1], mrB[MB + 1];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // work-group -> output block through the host-built order (8 x 8 super-blocks of
    // blocks); the hardware deals work-groups round-robin over the 8 XCDs, so logical
    // slot (g % 8) * (G / 8) + g / 8 gives each XCD a contiguous run of the order.
    const int gi = blockIdx.x, G = gridDim.x;
    const int2 ab = order[(G % 8 == 0) ? (gi % 8) * (G / 8) + gi / 8 : gi];
    const int A = ab.x, B = ab.y;
    const int a0 = A * MB, na = min(MB, nm1 - a0);
    const int b0 = B * MB, nb = min(MB, nm2 - b0);
    const bool diag = (MODE != RECT) && (A == B);
    const int ob = tid & (MB - 1), oa0 = (tid / MB) * (MB / 4);
    const int lrow = tid >> 2;
    GramArgs ga;
    ga.rep_pad = rep_pad;
    ga.lofs = lrow * LDA + (tid & 3) * (KC / 4);
    ga.r16 = lane & 15;
    ga.bcol = wave * 16 + (lane & 15);
    ga.kq = lane >> 4;
    double acc[MB / 4];
#pragma unroll
    for (int k = 0; k < MB / 4; ++k) acc[k] = 0.0;

    for (int l = 0; l < L; ++l) {
        __syncthreads();
        if (tid <= na) mrA[tid] = mr1[(size_t)l * (nm1 + 1) + a0 + tid];
        if (tid >= 128 && tid - 128 <= nb) mrB[tid - 128] = mr2[(size_t)l * (nm2 + 1) + b0 + tid - 128];
        __syncthreads();
        const int ra0 = mrA[0], ra1 = mrA[na], rb0 = mrB[0], rb1 = mrB[nb];
        if (ra1 == ra0 || rb1 == rb0) continue;
        const int nsr = (ra1 - ra0 + TS - 1) / TS, nsc = (rb1 - rb0 + TS - 1) / TS;
        for (int sr = 0; sr < nsr; ++sr) {
            for (int sc0 = diag ? sr : 0; sc0 < nsc; sc0 += 2) {
                const int r0 = ra0 + sr * TS, rn = min(TS, ra1 - r0);
                const int ntile = sc0 + 1 < nsc ? 2 : 1;
                int c0[2], cn[2];
                c0[0] = rb0 + sc0 * TS;
                cn[0] = min(TS, rb1 - c0[0]);
                c0[1] = c0[0] + TS;
                cn[1] = ntile == 2 ? min(TS, rb1 - c0[1]) : 0;
                __syncthreads();  // the previous tiles' epilogue is done with Sh and the maps
                if (wave == 0) {  // molecule and |x|^2 of every row
                    const bool v = lane < rn;
                    rowmol[lane] = v ? src1[r0 + lane] / ma1 - a0 : -1;
                    n2r[lane] = v ? n2a[r0 + lane] : 0.0;
                } else if (wave <= 2) {  // molecule, run and |x|^2 of every column of tile wave-1
                    const int u = wave - 1;
                    const bool v = lane < cn[u];
                    run_of_mol[u][lane] = -1;  // same wave: ordered before the writes below
                    const int m = v ? src2[c0[u] + lane] / ma2 - b0 : -1;
                    n2c[u][lane] = v ? n2b[c0[u] + lane] : 0.0;
                    const int prev = __shfl_up(m, 1, 64);
                    const bool head = v && (lane == 0 || prev != m);
                    const unsigned long long heads = __ballot(head);
                    const int run = __popcll(heads & ((2ull << lane) - 1)) - 1;
                    colmol[u][lane] = m;
                    colrun[u][lane] = run;
                    if (head) run_of_mol[u][m] = run;
                }

                // ---- Gram sub-tiles on MFMA, the row operand staged once ----
                v4d c[2][4];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int x = 0; x < 4; ++x) c[u][x] = (v4d){0, 0, 0, 0};
                ga.pa = Xp1 + (size_t)(r0 + lrow) * rep_pad + (tid & 3) * (KC / 4);
                ga.pb0 = Xp2 + (size_t)(c0[0] + lrow) * rep_pad + (tid & 3) * (KC / 4);
                ga.pb1 = ga.pb0 + (size_t)TS * rep_pad;
                ga.va = lrow < rn;
                ga.vb0 = lrow < cn[0];
                ga.vb1 = lrow < cn[1];
                const int na16 = (rn + 15) >> 4;
                const int ntw = (wave * 16 < cn[1]) ? 2 : (wave * 16 < cn[0] ? 1 : 0);
                if (ntile == 2)
                    gram_dispatch<2>(Ops, ga, na16, ntw, c);
                else
                    gram_dispatch<1>(Ops, ga, na16, ntw, c);

                // ---- per tile: exp into LDS, row runs, register accumulation ----
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && ntile == 1) break;
                    const int cnu = cn[u];
                    const bool dbl_self = diag && sc0 + u > sr;  // mirror sub-tile not visited
                    if (u == 1) __syncthreads();  // tile 0's stage 2 is done with Es and Rs
                    {
                        const int col = wave * 16 + (lane & 15);
                        const double nc = n2c[u][col];
                        const int cm = colmol[u][col];
#pragma unroll
                        for (int x = 0; x < 4; ++x) {
                            const v4d cv = c[u][x];
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int row = x * 16 + mfma16_row_f64(lane, v);
                                double e = 0.0;
                                if (row < rn && col < cnu) {
                                    e = exp((n2r[row] + nc - 2.0 * cv[v]) * cexp);
                                    if (dbl_self && rowmol[row] == cm) e *= 2.0;
                                }
                                Es[row * EP + col] = e;
                            }
                        }
                    }
                    __syncthreads();
                    // stage 1: run sums of row i over column molecules, four threads per
                    // row (16 columns each); a run's partial sum is carried to the next
                    // quarter while the run continues there
                    {
                        const int i = tid >> 2, q = tid & 3;
                        const int *cr = colrun[u];
                        const double *er = Es + i * EP;
                        double s = 0.0, first = 0.0;
                        int rfirst = -1;
#pragma unroll
                        for (int j = 0; j < 16; ++j) {
                            const int col = q * 16 + j;
                            if (col < cnu) {
                                s += er[col];
                                const int rj = cr[col];
                                if (col == cnu - 1 || cr[col + 1] != rj) {  // run ends here
                                    if (rfirst < 0) {
                                        first = s;
                                        rfirst = rj;
                                    } else if (i < rn) {
                                        Rs[i * EP + rj] = s;
                                    }
                                    s = 0.0;
                                }
                            }
                        }
                        // carry: the open run at the end of quarter q-1 (plus its own
                        // carry-in when it spans that whole quarter) continues here
                        double cin = 0.0, cout = s;
#pragma unroll
                        for (int step = 1; step < 4; ++step) {
                            const double prev = __shfl_up(cout, 1, 64);
                            if (q == step) {
                                cin = prev;
                                cout = rfirst < 0 ? cin + s : s;
                            }
                        }
                        if (rfirst >= 0 && i < rn) Rs[i * EP + rfirst] = first + cin;
                    }
                    __syncthreads();
                    // stage 2: each thread adds the rows of its row molecules for its column
                    const int run = ob < nb ? run_of_mol[u][ob] : -1;
                    if (run >= 0) {
#pragma unroll
                        for (int k = 0; k < MB / 4; ++k) {
                            const int a = oa0 + k;
                            if (a < na) {
                                const int rs = max(mrA[a], r0) - r0, re = min(mrA[a + 1], r0 + rn) - r0;
                                double sum = 0.0;
                                for (int i = rs; i < re; ++i) sum += Rs[i * EP + run];
                                acc[k] += sum;
                            }
                        }
                    }
                }
            }
        }
    }
    // ---- write the block ----
#pragma unroll
    for (int k = 0; k < MB / 4; ++k) {
        const int a = oa0 + k, b = ob;
        if (a >= na || b >= nb) continue;
        if (MODE == SYMM_RFP) {
            if (!diag || a <= b) out[rfp_upper((size_t)nm1, (size_t)(a0 + a), (size_t)(b0 + b))] = acc[k];
        } else if (MODE == RECT) {
            out[(size_t)(a0 + a) * nm2 + b0 + b] = acc[k];
        } else if (!diag || a <= b) {
            out[(size_t)(a0 + a) * nm1 + b0 + b] = acc[k];
            out[(size_t)(b0 + b) * nm1 + a0 + a] = acc[k];
        }
    }
}

// Device copy of a plan plus the packed rows.
struct DevicePacked {
    double *Xp = nullptr, *n2 = nullptr;
    int *mol_row = nullptr, *src = nullptr;
    int rows = 0, rep_pad = 0;
};

int upload_and_pack(const LocalPlan &p, const double *X, int rep, int ws_pack, int ws_plan,
                    hipStream_t stream, DevicePacked &d) {
    d.rows = p.rows;
    d.rep_pad = (rep + KC - 1) / KC * KC;
    const size_t mr_n = p.mol_row.size(), src_n = p.src.size();
    char *plan = static_cast<char *>(workspace(ws_plan, sizeof(int) * (mr_n + src_n), stream));
    const size_t pack_bytes = sizeof(double) * ((size_t)p.rows * d.rep_pad + p.rows + 64);
    char *pack = static_cast<char *>(workspace(ws_pack, pack_bytes, stream));
    KF_REQUIRE(plan && pack, KF_ERR_NOMEM, "device workspace allocation failed (%zu B)",
               pack_bytes);
    d.mol_row = reinterpret_cast<int *>(plan);
    int *src = d.mol_row + mr_n;
    d.src = src;
    int *stage = static_cast<int *>(host_staging(ws_plan, sizeof(int) * (mr_n + src_n)));
    KF_REQUIRE(stage, KF_ERR_NOMEM, "pinned staging allocation failed");
    std::copy(p.mol_row.begin(), p.mol_row.end(), stage);
    std::copy(p.src.begin(), p.src.end(), stage + mr_n);
    KF_HIP_TRY(hipMemcpyAsync(plan, stage, sizeof(int) * (mr_n + src_n), hipMemcpyHostToDevice,
                              stream));
    d.Xp = reinterpret_cast<double *>(pack);
    d.n2 = d.Xp + (size_t)p.rows * d.rep_pad;
    if (p.rows > 0) {
        hipLaunchKernelGGL(pack_rows_f64, dim3((p.rows + 3) / 4), dim3(256), 0, stream, X, src,
                           p.rows, rep, d.rep_pad, d.Xp, d.n2);
        KF_HIP_TRY(hipGetLastError());
    }
    return KF_OK;
}

// Output-block order: 8 x 8 super-blocks of blocks, super-block rows then columns,
// blocks inside a super-block row-major (upper triangle only when symmetric).
int upload_block_order(int nb1, int nb2, bool symm, hipStream_t stream, const int2 **dev) {
    constexpr int S = 8;
    std::vector<int2> order;
    order.reserve(symm ? (size_t)nb1 * (nb1 + 1) / 2 : (size_t)nb1 * nb2);
    for (int P = 0; P < nb1; P += S)
        for (int Q = symm ? P : 0; Q < nb2; Q += S)
            for (int a = P; a < std::min(P + S, nb1); ++a)
                for (int b = Q; b < std::min(Q + S, nb2); ++b)
                    if (!symm || a <= b) order.push_back(make_int2(a, b));
    const size_t bytes = order.size() * sizeof(int2);
    int2 *d = static_cast<int2 *>(workspace(WS_AUX1, bytes, stream));
    int2 *h = static_cast<int2 *>(host_staging(WS_AUX1, bytes));
    KF_REQUIRE(d && h, KF_ERR_NOMEM, "workspace allocation failed");
    std::copy(order.begin(), order.end(), h);
    KF_HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    *dev = d;
    return KF_OK;
}

int check_common(const double *X, const int *Q, const int *N, int nm, int ma, int rep,
                 double sigma) {
    KF_REQUIRE(nm >= 0 && ma >= 0 && rep >= 1, KF_ERR_ARG, "bad dimensions");
    KF_REQUIRE(std::isfinite(sigma) && sigma > 0.0, KF_ERR_SIGMA, "sigma must be > 0");
    KF_REQUIRE(nm == 0 || (X && Q && N), KF_ERR_ARG, "null pointer");
    return KF_OK;
}

template <int MODE>
int symm_impl(int precision, const double *X, const int *Q, const int *N, int nm, int ma, int rep,
              double sigma, void *out, hipStream_t stream) {
    int rc = check_common(X, Q, N, nm, ma, rep, sigma);
    if (rc) return rc;
    KF_REQUIRE(precision == KF_PREC_F64 || precision == KF_PREC_BF16X3, KF_ERR_ARG,
               "precision %d not supported here", precision);
    if (nm == 0) return KF_OK;
    KF_REQUIRE(out, KF_ERR_ARG, "null output");
    std::vector<int> hQ, hN;
    if ((rc = read_metadata(Q, N, nm, ma, hQ, hN, stream, WS_AUX0))) return rc;
    LocalPlan p;
    build_plan(hQ.data(), hN.data(), nm, ma, distinct_labels(hQ.data(), hN.data(), nm, ma), p);
    if (precision == KF_PREC_BF16X3)
        return local_mixed_symm(MODE == SYMM_DENSE, X, nm, ma, rep, sigma, p, out, stream);
    DevicePacked d;
    if ((rc = upload_and_pack(p, X, rep, WS_PACK, WS_PLAN, stream, d))) return rc;
    const int nblk = (nm + MB - 1) / MB;
    const long ntile = (long)nblk * (nblk + 1) / 2;
    const double cexp = -1.0 / (2.0 * sigma * sigma);
    if (p.nlabels == 0) {  // no atoms at all: zero output
        const size_t n = MODE == SYMM_RFP ? (size_t)nm * (nm + 1) / 2 : (size_t)nm * nm;
        KF_HIP_TRY(hipMemsetAsync(out, 0, n * sizeof(double), stream));
        return KF_OK;
    }
    const int2 *order = nullptr;
    if ((rc = upload_block_order(nblk, nblk, true, stream, &order))) return rc;
    ScopedTimer timer(TM_LOCAL_SYMM, stream);
    hipLaunchKernelGGL(local_scalar_f64_kernel<MODE>, dim3((unsigned)ntile), dim3(NT), 0, stream,
                       d.Xp, d.n2, d.mol_row, d.src, ma, d.Xp, d.n2, d.mol_row, d.src, ma,
                       p.nlabels, nm, nm, d.rep_pad, cexp, order, static_cast<double *>(out));
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_local_kernel_symm_rfp(int precision, const double *X, const int *Q,
                                        const int *N, int nm, int max_atoms, int rep,
                                        double sigma, void *arf, void *stream) {
    return kf::symm_impl<kf::SYMM_RFP>(precision, X, Q, N, nm, max_atoms, rep, sigma, arf,
                                       kf::as_stream(stream));
}

extern "C" int kf_local_kernel_symm(int precision, const double *X, const int *Q, const int *N,
                                    int nm, int max_atoms, int rep, double sigma, void *K,
                                    void *stream) {
    return kf::symm_impl<kf::SYMM_DENSE>(precision, X, Q, N, nm, max_atoms, rep, sigma, K,
                                         kf::as_stream(stream));
}

extern "C" int kf_local_kernel(int precision, const double *X1, const double *X2, const int *Q1,
                               const int *Q2, const int *N1, const int *N2, int nm1, int nm2,
                               int max_atoms1, int max_atoms2, int rep, double sigma, void *K,
                               void *stream_) {
    using namespace kf;
    hipStream_t stream = as_stream(stream_);
    int rc = check_common(X1, Q1, N1, nm1, max_atoms1, rep, sigma);
    if (rc) return rc;
    if ((rc = check_common(X2, Q2, N2, nm2, max_atoms2, rep, sigma))) return rc;
    KF_REQUIRE(precision == KF_PREC_F64 || precision == KF_PREC_BF16X3, KF_ERR_ARG,
               "precision %d not supported here", precision);
    if (nm1 == 0 || nm2 == 0) return KF_OK;
    KF_REQUIRE(K, KF_ERR_ARG, "null output");
    std::vector<int> q1, n1, q2, n2;
    if ((rc = read_metadata(Q1, N1, nm1, max_atoms1, q1, n1, stream, WS_AUX0))) return rc;
    if ((rc = read_metadata(Q2, N2, nm2, max_atoms2, q2, n2, stream, WS_AUX1))) return rc;
    std::vector<int> l1 = distinct_labels(q1.data(), n1.data(), nm1, max_atoms1);
    std::vector<int> l2 = distinct_labels(q2.data(), n2.data(), nm2, max_atoms2);
    std::vector<int> labels;
    std::set_intersection(l1.begin(), l1.end(), l2.begin(), l2.end(), std::back_inserter(labels));
    const size_t esz = precision == KF_PREC_F64 ? sizeof(double) : sizeof(float);
    if (labels.empty()) {
        KF_HIP_TRY(hipMemsetAsync(K, 0, esz * (size_t)nm1 * nm2, stream));
        return KF_OK;
    }
    LocalPlan p1, p2;
    build_plan(q1.data(), n1.data(), nm1, max_atoms1, labels, p1);
    build_plan(q2.data(), n2.data(), nm2, max_atoms2, labels, p2);
    if (precision == KF_PREC_BF16X3)
        return local_mixed_rect(X1, X2, nm1, nm2, max_atoms1, max_atoms2, rep, sigma, p1, p2,
                                static_cast<float *>(K), stream);
    DevicePacked d1, d2;
    if ((rc = upload_and_pack(p1, X1, rep, WS_PACK, WS_PLAN, stream, d1))) return rc;
    if ((rc = upload_and_pack(p2, X2, rep, WS_PACK2, WS_PLAN2, stream, d2))) return rc;
    const int nblk1 = (nm1 + MB - 1) / MB, nblk2 = (nm2 + MB - 1) / MB;
    const double cexp = -1.0 / (2.0 * sigma * sigma);
    const int2 *order = nullptr;
    if ((rc = upload_block_order(nblk1, nblk2, false, stream, &order))) return rc;
    ScopedTimer timer(TM_LOCAL_RECT, stream);
    hipLaunchKernelGGL(local_scalar_f64_kernel<RECT>, dim3((unsigned)((long)nblk1 * nblk2)),
                       dim3(NT), 0, stream, d1.Xp, d1.n2, d1.mol_row, d1.src, max_atoms1, d2.Xp,
                       d2.n2, d2.mol_row, d2.src, max_atoms2, (int)labels.size(), nm1, nm2,
                       d1.rep_pad, cexp, order, static_cast<double *>(K));
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}
