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
//    triangle of blocks for the symmetric forms), accumulated in LDS.  For every label
//    it runs the rows(A-block) x rows(B-block) Gram product as 64-row sub-tiles on FP64
//    MFMA (v_mfma_f64_16x16x4): one row tile against two column tiles at a time; wave
//    w owns columns 16w..16w+15 and skips empty 16-row/16-column fragments (95% of
//    the MFMAs issued are useful on ragged molecules, 81% with whole 64x64 tiles).  The
//    shared row tile is staged through double-buffered LDS, the wave's own column
//    operands stream from global memory straight into registers.
//  * Epilogue fused per tile, in registers: squared norms (FP64), exp, a segmented
//    scan along each wave's 16 columns (DPP row shifts) that sums the columns of one
//    molecule, then the row molecules' sums into the LDS block.  Deterministic, no
//    atomics; each K entry is written exactly once, straight into RFP (UPLO='U') or
//    dense layout.
#include <algorithm>
#include <cmath>
#include <vector>

#include "local_mixed.hpp"

namespace kf {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));

constexpr int MB = 64;       // molecules per output block
constexpr int TS = 64;       // sub-tile edge (packed rows)
constexpr int KA = 32;       // row-operand chunk staged in LDS (descriptor columns)
constexpr int KS = 16;       // column-operand sub-chunk held in registers
constexpr int LDA = KA + 2;  // LDS pitch of the staged rows (conflict-free 16-byte reads)
constexpr int PP = TS + 4;   // LDS pitch of the segment sums (<= 64 molecules + 3 window cuts)
constexpr int AP = MB + 1;   // LDS pitch of the block accumulators
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


// Gram products of one row sub-tile (64 packed rows) with NTL column sub-tiles over the
// full depth, as v_mfma_f64_16x16x4 (A[i][k] from lane i + 16k, B[k][j] from lane
// j + 16k, D row 4v + lane/16, column lane % 16).  Wave w owns columns 16w .. 16w+15 of
// every column tile and multiplies the NA non-empty 16-row fragments with its NTW
// non-empty column tiles.
//  * The row tile is shared by the four waves: it is staged global -> registers -> LDS in
//    chunks of KA columns with two LDS buffers (one barrier per chunk).  The next
//    chunk is moved in KS-column parts, each stored one sub-chunk after it was loaded.
//  * A wave's column operands are its own: they go straight from global memory into a
//    two-slot register ring, one KS sub-chunk ahead.
//  * Inside a KS sub-chunk, lane group k = lane / 16 reads the 4 consecutive columns
//    4k .. 4k+3 and k-step ks multiplies column 4k + ks, for both operands (a
//    permutation of the summation index, so one 32-byte read per operand row).
// On entry the caller's last barrier has freed both buffers; on exit a barrier has
// retired every read of them.
struct GramArgs {
    // staging row of this thread and column row of this lane, clamped into the tiles:
    // rows and columns past a tile's end repeat a valid one and their Gram entries are
    // never read, so every load is unconditional (exact wait counts, no branches)
    const double *pa, *pb0, *pb1;
    int rep_pad, lofs, r16, kq;
};

template <int NTL, int NA, int NTW>
__device__ void gram_tiles(double *__restrict__ Abuf, const GramArgs &g, v4d (&c)[2][4]) {
    constexpr int NSUB = KA / KS;
    constexpr int NB = NTW > 0 ? NTW : 1;
    static_assert(KS == 16 && NSUB == 2, "two staged row parts and two ring slots per chunk");
    double2 ra0, ra1;         // one staged row part: 4 columns of one sub-chunk
    double2 bq[NSUB][NB][2];  // column ring: sub-chunk gs in slot gs % 2
#define KF_LOAD_A(k)                                       \
    do {                                                   \
        ra0 = *(const double2 *)(g.pa + (k));              \
        ra1 = *(const double2 *)(g.pa + (k) + 2);          \
    } while (0)
#define KF_STORE_A(buf, s)                                 \
    do {                                                   \
        *(double2 *)((buf) + g.lofs + (s) * KS) = ra0;     \
        *(double2 *)((buf) + g.lofs + (s) * KS + 2) = ra1; \
    } while (0)
#define KF_LOAD_B(slot, k)                                               \
    do {                                                                 \
        bq[slot][0][0] = *(const double2 *)(g.pb0 + (k));                \
        bq[slot][0][1] = *(const double2 *)(g.pb0 + (k) + 2);            \
        if (NB == 2) {                                                   \
            bq[slot][NB - 1][0] = *(const double2 *)(g.pb1 + (k));       \
            bq[slot][NB - 1][1] = *(const double2 *)(g.pb1 + (k) + 2);   \
        }                                                                \
    } while (0)
    const int nch = g.rep_pad / KA, nsub = nch * NSUB;
#pragma unroll
    for (int s = 0; s < NSUB; ++s) {
        KF_LOAD_A(s * KS);
        KF_STORE_A(Abuf, s);
    }
    if (NTW > 0) KF_LOAD_B(0, 0);
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const double *cur = Abuf + (ch & 1) * (TS * LDA);
        double *nxt = Abuf + ((ch + 1) & 1) * (TS * LDA);
        const int cn = min(ch + 1, nch - 1);  // the last chunk re-reads itself (never used)
#pragma unroll
        for (int s = 0; s < NSUB; ++s) {
            const int gs = ch * NSUB + s;
            // Issue order is pinned (sched_barrier) and every load and store is
            // unconditional, so the wait counts are exact: each load is consumed one
            // sub-chunk after it is issued and its latency hides behind these MFMAs.
            // In the last chunk the parts land in the idle buffer and are never read.
            __builtin_amdgcn_sched_barrier(0);
            if (s > 0) KF_STORE_A(nxt, s - 1);
            KF_LOAD_A(cn * KA + s * KS);
            if (NTW > 0) KF_LOAD_B((s + 1) % NSUB, min(gs + 1, nsub - 1) * KS);
            __builtin_amdgcn_sched_barrier(0);
            if (NTW == 0) continue;
#pragma unroll
            for (int h = 0; h < 2; ++h) {  // k-steps 2h, 2h+1
                double2 a[NA];
#pragma unroll
                for (int x = 0; x < NA; ++x)
                    a[x] = *(const double2 *)(cur + (x * 16 + g.r16) * LDA + s * KS + 4 * g.kq + 2 * h);
#pragma unroll
                for (int kk = 0; kk < 2; ++kk) {
#pragma unroll
                    for (int u = 0; u < NB; ++u) {
                        const double b = kk ? bq[s][u][h].y : bq[s][u][h].x;
#pragma unroll
                        for (int x = 0; x < NA; ++x)
                            c[u][x] = __builtin_amdgcn_mfma_f64_16x16x4f64(kk ? a[x].y : a[x].x, b,
                                                                           c[u][x], 0, 0, 0);
                    }
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        KF_STORE_A(nxt, NSUB - 1);
        __syncthreads();
    }
#undef KF_LOAD_A
#undef KF_STORE_A
#undef KF_LOAD_B
}

template <int NTL, int NA>
__device__ __forceinline__ void gram_ntw(double *Abuf, const GramArgs &g, int ntw,
                                         v4d (&c)[2][4]) {
    if (ntw == 0) gram_tiles<NTL, NA, 0>(Abuf, g, c);
    else if (NTL == 1 || ntw == 1) gram_tiles<NTL, NA, 1>(Abuf, g, c);
    else gram_tiles<NTL, NA, NTL>(Abuf, g, c);
}

template <int NTL>
__device__ __forceinline__ void gram_dispatch(double *Abuf, const GramArgs &g, int na16, int ntw,
                                              v4d (&c)[2][4]) {
    switch (na16) {
        case 1: gram_ntw<NTL, 1>(Abuf, g, ntw, c); break;
        case 2: gram_ntw<NTL, 2>(Abuf, g, ntw, c); break;
        case 3: gram_ntw<NTL, 3>(Abuf, g, ntw, c); break;
        default: gram_ntw<NTL, 4>(Abuf, g, ntw, c); break;
    }
}

// v + (value of the lane D places lower in the same 16-lane row), when `take`: one step
// of a segmented inclusive scan along the 16 columns a row of MFMA output spans.
template <int D>
__device__ __forceinline__ double scan_step(double v, bool take) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int slo = __builtin_amdgcn_update_dpp(0, lo, 0x110 + D, 0xF, 0xF, true);
    const int shi = __builtin_amdgcn_update_dpp(0, hi, 0x110 + D, 0xF, 0xF, true);
    return take ? v + __hiloint2double(shi, slo) : v;
}

// Output block: 64 x 64 molecule pairs per work-group, accumulated in LDS (each entry
// owned by one thread: column molecule t % 64, row molecules 16 (t / 64) .. +15).  Per
// label, the row sub-tile sr is multiplied with two column sub-tiles at once (the row
// operand is staged once for both).  Epilogue per tile, in registers: exp of every
// Gram entry, then a segmented scan along each wave's 16 columns sums the columns of
// one molecule (a window segment: a molecule's columns inside one 16-column window);
// the segment tails go to LDS as partial sums P[row][segment], and each thread adds,
// for its column molecule's segments, the rows of its row molecules.  The molecule of a
// packed row is src[row] / max_atoms (src = a * max_atoms + i, see local_plan.hpp).
template <int MODE>
__global__ __launch_bounds__(NT, 2) void local_scalar_f64_kernel(
    const double *__restrict__ Xp1, const double *__restrict__ n2a, const int *__restrict__ mr1,
    const int *__restrict__ src1, int ma1, const double *__restrict__ Xp2,
    const double *__restrict__ n2b, const int *__restrict__ mr2, const int *__restrict__ src2,
    int ma2, int L, int nm1, int nm2, int rep_pad, double cexp, const int2 *__restrict__ order,
    double *__restrict__ out) {
    // two row-operand buffers; the segment partial sums of the epilogue reuse the same
    // memory once the Gram products are done
    constexpr int OPS = 2 * TS * LDA, EPI = TS * PP;
    __shared__ double Sh[OPS > EPI ? OPS : EPI];
    __shared__ double accL[MB * AP];
    double *const Abuf = Sh, *const P = Sh;
    __shared__ double n2r[TS], n2c[2][TS];
    __shared__ int rowmol[TS], colmol[2][TS], colseg[2][TS], colinfo[2][TS];
    __shared__ int segf[2][MB], segl[2][MB], rstart[TS + 1], nrr_s;
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
    const int lrow = tid >> 2, bcol = wave * 16 + (lane & 15);
    GramArgs ga;
    ga.rep_pad = rep_pad;
    ga.lofs = lrow * LDA + (tid & 3) * 4;
    ga.r16 = lane & 15;
    ga.kq = lane >> 4;
    for (int k = tid; k < MB * AP; k += NT) accL[k] = 0.0;

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
                if (wave == 0) {  // molecule and |x|^2 of every row; the row molecules' ranges
                    const bool v = lane < rn;
                    const int m = v ? src1[r0 + lane] / ma1 - a0 : -1;
                    rowmol[lane] = m;
                    n2r[lane] = v ? n2a[r0 + lane] : 0.0;
                    const int prev = __shfl_up(m, 1, 64);
                    const bool rh = v && (lane == 0 || prev != m);
                    const unsigned long long rheads = __ballot(rh);
                    if (rh) rstart[__popcll(rheads & ((2ull << lane) - 1)) - 1] = lane;
                    if (lane == 0) {
                        nrr_s = __popcll(rheads);
                        rstart[__popcll(rheads)] = rn;
                    }
                } else if (wave <= 2) {  // column maps of tile wave-1
                    const int u = wave - 1;
                    const bool v = lane < cn[u];
                    segf[u][lane] = -1;  // same wave: ordered before the writes below
                    const int m = v ? src2[c0[u] + lane] / ma2 - b0 : -1;
                    n2c[u][lane] = v ? n2b[c0[u] + lane] : 0.0;
                    const int prev = __shfl_up(m, 1, 64), next = __shfl_down(m, 1, 64);
                    const bool mhead = v && (lane == 0 || prev != m);
                    const bool mtail = v && (lane == 63 || next != m);
                    // window segments: a molecule's columns inside one 16-column window
                    const bool shead = v && (mhead || (lane & 15) == 0);
                    const unsigned long long heads = __ballot(shead);
                    const unsigned long long upto = heads & ((2ull << lane) - 1);
                    const int seg = __popcll(upto) - 1;
                    const int dist = v ? lane - (63 - __clzll(upto)) : 0;
                    const bool stail = v && ((lane & 15) == 15 || lane + 1 >= cn[u] ||
                                             ((heads >> (lane + 1)) & 1ull));
                    colmol[u][lane] = m;
                    colseg[u][lane] = seg;
                    colinfo[u][lane] = dist | (stail ? 256 : 0);
                    if (mhead) segf[u][m] = seg;
                    if (mtail) segl[u][m] = seg;
                }

                // ---- Gram sub-tiles on MFMA, the row operand staged once ----
                v4d c[2][4];
#pragma unroll
                for (int u = 0; u < 2; ++u)
#pragma unroll
                    for (int x = 0; x < 4; ++x) c[u][x] = (v4d){0, 0, 0, 0};
                ga.pa = Xp1 + (size_t)(r0 + min(lrow, rn - 1)) * rep_pad + (tid & 3) * 4;
                ga.pb0 = Xp2 + (size_t)(c0[0] + min(bcol, cn[0] - 1)) * rep_pad + 4 * ga.kq;
                ga.pb1 = ntile == 2 ? Xp2 + (size_t)(c0[1] + min(bcol, cn[1] - 1)) * rep_pad + 4 * ga.kq
                                    : ga.pb0;
                const int na16 = (rn + 15) >> 4;
                const int ntw = (wave * 16 < cn[1]) ? 2 : (wave * 16 < cn[0] ? 1 : 0);
                if (ntile == 2)
                    gram_dispatch<2>(Abuf, ga, na16, ntw, c);
                else
                    gram_dispatch<1>(Abuf, ga, na16, ntw, c);

                // ---- per tile: exp, segment sums in registers, row molecules ----
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && ntile == 1) break;
                    const bool dbl_self = diag && sc0 + u > sr;  // mirror sub-tile not visited
                    if (u == 1) __syncthreads();  // tile 0's row sums are done with P
                    {
                        const int col = bcol;
                        const bool colok = col < cn[u];
                        const double nc = n2c[u][col];
                        const int cm = colmol[u][col], seg = colseg[u][col];
                        const int info = colinfo[u][col], dist = info & 255;
                        const bool stail = (info >> 8) & 1;
#pragma unroll
                        for (int x = 0; x < 4; ++x) {
                            const v4d cv = c[u][x];
#pragma unroll
                            for (int v = 0; v < 4; ++v) {
                                const int row = x * 16 + mfma16_row_f64(lane, v);
                                const bool ok = colok && row < rn;
                                double e = 0.0;
                                if (ok) {
                                    e = exp((n2r[row] + nc - 2.0 * cv[v]) * cexp);
                                    if (dbl_self && rowmol[row] == cm) e *= 2.0;
                                }
                                e = scan_step<1>(e, dist >= 1);
                                e = scan_step<2>(e, dist >= 2);
                                e = scan_step<4>(e, dist >= 4);
                                e = scan_step<8>(e, dist >= 8);
                                if (ok && stail) P[row * PP + seg] = e;
                            }
                        }
                    }
                    __syncthreads();
                    // thread (wave w, lane b) adds, for column molecule b, every 4th row
                    // molecule of the tile (from w) over the rows of that molecule and the
                    // window segments of b; within a tile each (row, column) molecule pair
                    // has one owner
                    const int s0 = ob < nb ? segf[u][ob] : -1;
                    if (s0 >= 0) {
                        const int s1 = segl[u][ob], nrr = nrr_s;
                        for (int j = wave; j < nrr; j += 4) {
                            const int rs = rstart[j], re = rstart[j + 1];
                            double sum = 0.0;
                            for (int i = rs; i < re; ++i)
                                for (int sg = s0; sg <= s1; ++sg) sum += P[i * PP + sg];
                            accL[rowmol[rs] * AP + ob] += sum;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    // ---- write the block ----
#pragma unroll
    for (int k = 0; k < MB / 4; ++k) {
        const int a = oa0 + k, b = ob;
        if (a >= na || b >= nb) continue;
        const double val = accL[a * AP + b];
        if (MODE == SYMM_RFP) {
            if (!diag || a <= b) out[rfp_upper((size_t)nm1, (size_t)(a0 + a), (size_t)(b0 + b))] = val;
        } else if (MODE == RECT) {
            out[(size_t)(a0 + a) * nm2 + b0 + b] = val;
        } else if (!diag || a <= b) {
            out[(size_t)(a0 + a) * nm1 + b0 + b] = val;
            out[(size_t)(b0 + b) * nm1 + a0 + a] = val;
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
    d.rep_pad = (rep + KA - 1) / KA * KA;
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
