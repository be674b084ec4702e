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
//    triangle of blocks for the symmetric forms), accumulated in a device-memory slot of
//    the work-group (three work-groups per CU fit in LDS and registers).  For every label
//    it runs the rows(A-block) x rows(B-block) Gram product as 64-row sub-tiles on FP64
//    MFMA (v_mfma_f64_16x16x4): one row tile against two column tiles at a time; wave
//    w owns columns 16w..16w+15 and skips empty 16-row/16-column fragments (95% of
//    the MFMAs issued are useful on ragged molecules, 81% with whole 64x64 tiles).  The
//    shared row tile is staged through double-buffered LDS, the wave's own column
//    operands stream from global memory straight into registers.
//  * Epilogue fused per tile, in registers: squared norms (FP64), exp, a segmented
//    scan along each wave's 16 columns (DPP row shifts) that sums the columns of one
//    molecule, then the row molecules' sums into the block's accumulators.  Deterministic
//    (no atomics on values); each K entry is written exactly once, straight into RFP
//    (UPLO='U') or dense layout.
//  * The same kernel template also runs an FP64-accurate int8-split form (precision
//    KF_PREC_I8X6, template slice count 6): each packed row is stored as six 7-bit int8
//    slices under a power-of-two scale (pack_rows_i8), the Gram products run exactly in
//    int32 on v_mfma_i32_16x16x64_i8 (one column tile per row tile, both operands staged
//    through LDS with LDS-destination loads), and the FP64 epilogue above receives the
//    recombined Gram entries (gram_i8_tile).
#include <algorithm>
#include <mutex>
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
constexpr int AP = MB;       // row pitch of the block accumulators
constexpr int NT = 256;      // threads per workgroup

enum Mode { SYMM_RFP = 0, SYMM_DENSE = 1, RECT = 2 };

// Block accumulators in device memory.  The 64 x 64 FP64 sums of a work-group's output
// block (32 KB) live in one of the slots of a fixed device array instead of LDS, which
// brings the work-group to 40 KB of LDS so that three fit on a CU (with 168 registers):
// 207 ms instead of 222 ms at 20,000 molecules with the accumulators in LDS at two per CU.
// The slots are split by XCD (read from the hardware's XCC id): a work-group claims a free
// slot of its own XCD at its start (compare-and-swap on a busy flag) and frees it at its
// end, so a slot's cache lines only ever live in one XCD's L2 -- a slot handed between
// XCDs could be overwritten later by the other L2's write-back of the previous owner's
// lines.  Fewer work-groups than slots are resident on an XCD (the host checks), so the
// probe ends; the flags are cleared by a small kernel before every launch.  Only the owner
// thread of an entry touches it between two barriers, as in LDS; launches on one device do
// not overlap (the entry points order them).
constexpr int kXcc = 8, kSlotsPerXcc = 128, kSlots = kXcc * kSlotsPerXcc;
__device__ double g_acc_slots[(size_t)kSlots * 64 * 64];
__device__ int g_slot_busy[kSlots];

__device__ __forceinline__ int xcc_id() {
    return __builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) & 15;  // hwreg(XCC_ID, 0, 4)
}

constexpr int kXccProbe = 64;
__device__ int g_xcc_probe[kXccProbe];
__global__ __launch_bounds__(64) void xcc_probe_kernel() {
    if (threadIdx.x == 0) g_xcc_probe[blockIdx.x] = xcc_id();
}

__global__ __launch_bounds__(256) void clear_slots_kernel() {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < kSlots) g_slot_busy[i] = 0;
}

// Column windows: the descriptor columns are cut into windows of KS = 16.  A window that is
// zero in every row of a row tile, or in every row of the column tiles it meets, adds an
// exact zero to every Gram entry of that tile pair, so the kernel leaves it out.  Masks of
// the first kMaxWin windows are kept as two 64-bit words; wider descriptors run unmasked.
constexpr int kMaxWin = 128;

// Window bits of one 64-column slice from a wave ballot of "column is non-zero".
__device__ __forceinline__ unsigned long long slice_windows(unsigned long long b) {
    unsigned long long w = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if ((b >> (16 * q)) & 0xFFFFull) w |= 1ull << q;
    return w;
}

// mask[a] = windows in which some atom of molecule a is non-zero (two words per molecule).
// One wave per molecule.
__global__ __launch_bounds__(256) void mol_winmask_kernel(const double *__restrict__ X,
                                                          const int *__restrict__ N, int nm,
                                                          int ma, int rep,
                                                          unsigned long long *__restrict__ mask) {
    const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (a >= nm) return;
    const int na = min(max(N[a], 0), ma);
    const double *x = X + (size_t)a * ma * rep;
    unsigned long long w0 = 0, w1 = 0;
    for (int j = 0; j * 64 < rep && 4 * j < kMaxWin; ++j) {
        const int k = j * 64 + lane;
        bool nz = false;
        if (k < rep)
            for (int i = 0; i < na; ++i) nz |= x[(size_t)i * rep + k] != 0.0;
        const unsigned long long w = slice_windows(__ballot(nz));
        if (j < 16) w0 |= w << (4 * j);
        else w1 |= w << (4 * j - 64);
    }
    if (lane == 0) {
        mask[2 * (size_t)a] = w0;
        mask[2 * (size_t)a + 1] = w1;
    }
}

// pack: Xp[row, k] = X[xrow, k] (k < rep), 0 for rep <= k < rep_pad, where the plan's
// source s = src[row] names molecule s / ma of the (possibly reordered) set and xrow =
// perm[s / ma] * ma + s % ma its row of X (perm == nullptr: identity);
// n2[row] = |x|^2 in FP64; rowmask[row] = the row's non-zero windows.  One wave per row.
__global__ __launch_bounds__(256) void pack_rows_f64(const double *__restrict__ X,
                                                     const int *__restrict__ src,
                                                     const int *__restrict__ perm, int ma,
                                                     int rows, int rep, int rep_pad,
                                                     double *__restrict__ Xp,
                                                     double *__restrict__ n2,
                                                     unsigned long long *__restrict__ rowmask) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows) return;
    const int sr = src[row];
    const size_t xrow = perm ? (size_t)perm[sr / ma] * ma + sr % ma : (size_t)sr;
    const double *s = X + xrow * rep;
    double *d = Xp + (size_t)row * rep_pad;
    double acc = 0.0;
    unsigned long long w0 = 0, w1 = 0;
    for (int j = 0; j * 64 < rep_pad; ++j) {
        const int k = j * 64 + lane;
        const double v = k < rep ? s[k] : 0.0;
        if (k < rep_pad) d[k] = v;
        acc += v * v;
        const unsigned long long w = slice_windows(__ballot(v != 0.0));
        if (j < 16) w0 |= w << (4 * j);
        else if (j < kMaxWin / 4) w1 |= w << (4 * j - 64);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        n2[row] = acc;
        rowmask[2 * (size_t)row] = w0;
        rowmask[2 * (size_t)row + 1] = w1;
    }
}

// The int8-split form (precision KF_PREC_I8X6).  A packed row x with mu = 2^e > max|x|
// (mu = 1 for a zero row) is written as
//     x = mu * sum_{s < NS} 2^-(6 + 7 s) q_s  +  mu * r,   |r| <= 2^-(7 NS),
// each integer slice q_s in [-64, 64] the rounded remainder of the ones before it (every
// step is exact in FP64).  A Gram entry is then the exact integer dot products of slice
// pairs (s, t) with s + t < NS, summed over each anti-diagonal d = s + t in int32 on the
// MFMA and combined in FP64 as mu_x mu_y sum_d 2^-(12 + 7 d) acc_d.  Slice s is at most
// 2^-7s mu in magnitude, so a dropped pair (s + t >= NS) weighs at most 2^-7(s+t) mu_x mu_y
// per summed column (with NS = 6: five pairs on d = 6, 5.03 * 2^-42 over all of them), and
// the remainders r (half a unit of the last slice, 2^-42) add at most 2 * 2^-42 mu_x mu_y:
// a Gram entry is off by less than 8 * 2^-42 mu_x mu_y per column (in practice far less:
// kernel entries within ~1e-11 relative on FCHL19 descriptors).
// Layout: rows of nwin4 windows (16 columns each, nwin4 a multiple of 4, the columns past
// rep zero), each window its NS slices of 16 bytes (padding windows to 128-byte lines
// fetched more, not less).
typedef int v4i __attribute__((ext_vector_type(4)));
constexpr int kI8Slices = 6;
constexpr int kI8WinBytes = kI8Slices * 16;

template <int NS>
__global__ __launch_bounds__(256) void pack_rows_i8(const double *__restrict__ X,
                                                    const int *__restrict__ src,
                                                    const int *__restrict__ perm, int ma,
                                                    int rows, int rep, int nwin4,
                                                    signed char *__restrict__ Sp,
                                                    double *__restrict__ n2,
                                                    double *__restrict__ mu,
                                                    unsigned long long *__restrict__ rowmask) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    const int ncol = nwin4 * 16;
    if (row >= rows) return;
    const int sr = src[row];
    const size_t xrow = perm ? (size_t)perm[sr / ma] * ma + sr % ma : (size_t)sr;
    const double *s = X + xrow * rep;
    double acc = 0.0, mx = 0.0;
    unsigned long long w0 = 0, w1 = 0;
    for (int j = 0; j * 64 < ncol; ++j) {
        const int k = j * 64 + lane;
        const double v = k < rep ? s[k] : 0.0;
        acc += v * v;
        mx = fmax(mx, fabs(v));
        const unsigned long long w = slice_windows(__ballot(v != 0.0));
        if (j < 16) w0 |= w << (4 * j);
        else if (j < kMaxWin / 4) w1 |= w << (4 * j - 64);
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc += __shfl_xor(acc, off);
        mx = fmax(mx, __shfl_xor(mx, off));
    }
    int e = 0;
    frexp(mx, &e);  // mx = f 2^e with f in [1/2, 1), so |x| < 2^e
    const double inv = mx > 0.0 ? ldexp(1.0, -e) : 1.0;
    signed char *d = Sp + (size_t)row * nwin4 * kI8WinBytes;
    for (int j = 0; j * 64 < ncol; ++j) {
        const int k = j * 64 + lane;
        double r = (k < rep ? s[k] : 0.0) * inv;  // |r| < 1
        signed char *p = d + (k >> 4) * kI8WinBytes + (k & 15);
#pragma unroll
        for (int q = 0; q < NS; ++q) {
            const double sc = (double)(1ll << (6 + 7 * q)), isc = 1.0 / sc;
            const double t = rint(r * sc);  // in [-64, 64]: |r| <= 2^-(7 q) (q > 0: 1/2 unit)
            r = fma(-t, isc, r);            // exact: a multiple of ulp(r), below |r|
            p[q * 16] = (signed char)(int)t;
        }
    }
    if (lane == 0) {
        n2[row] = acc;
        mu[row] = mx > 0.0 ? ldexp(1.0, e) : 1.0;
        rowmask[2 * (size_t)row] = w0;
        rowmask[2 * (size_t)row + 1] = w1;
    }
}

__device__ __forceinline__ unsigned long long wave_or64(unsigned long long m) {
    unsigned lo = (unsigned)m, hi = (unsigned)(m >> 32);
    for (int off = 32; off > 0; off >>= 1) {
        lo |= (unsigned)__shfl_xor((int)lo, off);
        hi |= (unsigned)__shfl_xor((int)hi, off);
    }
    return ((unsigned long long)hi << 32) | lo;
}

// Window masks of the kernel's row tiles: for label l and molecule block A, the block's
// rows of label l are cut into 64-row tiles from its first row (the tiling the kernel
// walks); tmask[tbase[l * nblk + A] + t] = OR of the masks of tile t's rows.  One
// work-group per (label, block), a wave per tile.
__global__ __launch_bounds__(256) void tile_mask_kernel(const unsigned long long *__restrict__ rowmask,
                                                        const int *__restrict__ mol_row, int nm,
                                                        int nblk, const int *__restrict__ tbase,
                                                        ulonglong2 *__restrict__ tmask) {
    const int la = blockIdx.x, l = la / nblk, A = la - l * nblk;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int *mr = mol_row + (size_t)l * (nm + 1);
    const int r0 = mr[A * MB], r1 = mr[min(A * MB + MB, nm)];
    const int nt = (r1 - r0 + TS - 1) / TS;
    for (int t = wave; t < nt; t += 4) {
        const int r = r0 + t * TS + lane;
        unsigned long long m0 = 0, m1 = 0;
        if (r < r1) {
            m0 = rowmask[2 * (size_t)r];
            m1 = rowmask[2 * (size_t)r + 1];
        }
        m0 = wave_or64(m0);
        m1 = wave_or64(m1);
        if (lane == 0) tmask[tbase[la] + t] = make_ulonglong2(m0, m1);
    }
}

// Position of the n-th (0-based) set bit of m; n < popcount(m).
__device__ __forceinline__ int nth_set_bit(unsigned long long m, int n) {
    int pos = 0;
#pragma unroll
    for (int s = 32; s > 0; s >>= 1) {
        const int c = __popcll(m & ((1ull << s) - 1));
        if (n >= c) {
            n -= c;
            m >>= s;
            pos += s;
        }
    }
    return pos;
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
    int lofs, r16, kq;
    // the tile pair's active windows in ascending order (an even count, nsub): entry e
    // is held by lane e % 64 of w0 (e < 64) or w1 (64 <= e < kMaxWin); entries past
    // kMaxWin (unmasked wide descriptors only) are window e itself
    int nsub, w0, w1;
};

// Column offset of active window e (e uniform across the wave).
__device__ __forceinline__ int win_off(const GramArgs &g, int e) {
    const int lo = __builtin_amdgcn_readlane(g.w0, e & 63);
    const int hi = __builtin_amdgcn_readlane(g.w1, e & 63);
    return (e < 64 ? lo : (e < kMaxWin ? hi : e)) * KS;
}

template <int NTL, int NA, int NTW>
__device__ void gram_tiles(double *__restrict__ Abuf, const GramArgs &g, v4d (&c)[2][4]) {
    constexpr int NSUB = KA / KS;
    constexpr int NB = NTW > 0 ? NTW : 1;
    static_assert(KS == 16 && NSUB == 2, "two staged row parts and two ring slots per chunk");
    // the staged row chunk, 4 columns of each of its two sub-chunks (named registers: an
    // array here is left in scratch memory by the compiler)
    double2 ra0_0, ra0_1, ra1_0, ra1_1;
    double2 bq[NSUB][NB][2];  // column ring: sub-chunk gs in slot gs % 2
#define KF_LOAD_A(p, k)                                          \
    do {                                                         \
        ra##p##_0 = *(const double2 *)(g.pa + (k));              \
        ra##p##_1 = *(const double2 *)(g.pa + (k) + 2);          \
    } while (0)
#define KF_STORE_A(buf, p)                                       \
    do {                                                         \
        *(double2 *)((buf) + g.lofs + (p) * KS) = ra##p##_0;     \
        *(double2 *)((buf) + g.lofs + (p) * KS + 2) = ra##p##_1; \
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
    const int nsub = g.nsub, nch = nsub / NSUB;
    if (nch == 0) {  // no window is active: the Gram block is exactly zero
        __syncthreads();
        return;
    }
    KF_LOAD_A(0, win_off(g, 0));
    KF_LOAD_A(1, win_off(g, 1));
    KF_STORE_A(Abuf, 0);
    KF_STORE_A(Abuf, 1);
    if (NTW > 0) KF_LOAD_B(0, win_off(g, 0));
    __syncthreads();
    for (int ch = 0; ch < nch; ++ch) {
        const double *cur = Abuf + (ch & 1) * (TS * LDA);
        double *nxt = Abuf + ((ch + 1) & 1) * (TS * LDA);
        const int cn = min(ch + 1, nch - 1);  // the last chunk re-reads itself (never used)
#pragma unroll
        for (int s = 0; s < NSUB; ++s) {
            const int gs = ch * NSUB + s;
            // Issue order is pinned (sched_barrier) and every load and store is
            // unconditional, so the wait counts are exact.  The next row chunk is loaded
            // whole at the first sub-chunk and stored after the last, so its latency hides
            // behind the MFMAs of a full chunk; each column load is consumed one sub-chunk
            // after it is issued.  In the last chunk the parts land in the idle buffer and
            // are never read.
            __builtin_amdgcn_sched_barrier(0);
            if (s == 0) {
                KF_LOAD_A(0, win_off(g, cn * NSUB));
                KF_LOAD_A(1, win_off(g, cn * NSUB + 1));
            }
            if (NTW > 0) KF_LOAD_B((s + 1) % NSUB, win_off(g, min(gs + 1, nsub - 1)));
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
        KF_STORE_A(nxt, 0);
        KF_STORE_A(nxt, 1);
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

// Int8-split Gram products of one row sub-tile (64 packed rows) with one column sub-tile,
// as v_mfma_i32_16x16x64_i8 over chunks of 4 active windows (64 columns).  Lane group
// k = lane / 16 of both operands carries window 4c + k of chunk c: a permutation of the
// summation index shared by the two operands, so the integer dot products are exact
// whatever order the instruction sums its 64 products in.  acc[x][d] (d = s + t) collects
// the slice pairs of one anti-diagonal: |acc| <= NS * 64^2 * (columns), checked on the
// host against int32.  Wave w owns columns 16w .. 16w+15; D row 4 (lane / 16) + v, column
// lane % 16.
//  * Both operands go from global memory straight into LDS (16-byte LDS-destination
//    loads), one chunk ahead: the row chunk (64 rows x 4 windows x NS slices) into one of
//    two row buffers shared by the four waves, each wave's column chunk (its 16 columns)
//    into its own region, read back into registers at the start of the chunk's step.  No
//    other global loads are in flight in the loop (plain loads share the counter with
//    the LDS-destination ones and would make the compiler wait for all of them).
//  * Row buffers: slice-major, then row, then window slot, the slot of window k in row r
//    being k XOR bits 1..2 of r, so the eight lanes of a fragment read served together
//    (rows 8j .. 8j+7, 128 bytes) hit eight different 16-byte bank groups.  Each wave
//    loads 16 rows: lane i writes 16 bytes at (its wave's base) + 16 i, i.e. row
//    16 w + i / 4, slot i % 4, so it fetches window slot ^ ((i / 8) % 4).
//  * Column region of a wave: slice q of lane i at 1024 q + 16 i.
//  * A fragment's NS row slices are read from LDS one fragment ahead of its products.
// On entry the caller's last barrier has freed the LDS; on exit a barrier has retired
// every read of it.
constexpr int kI8RowBuf = kI8Slices * 4096;  // bytes of one row-chunk buffer
constexpr int kI8ColReg = kI8Slices * 1024;  // bytes of one wave's column region
constexpr int kI8Lds = 2 * kI8RowBuf + 4 * kI8ColReg;

struct GramI8Args {
    const signed char *pa;  // this lane's staging row (clamped into the tile), window 0
    const signed char *pb;  // this lane's column (clamped into the tile), window 0
    int stk, rdo;           // window slot this lane stages; LDS byte offset of its reads
    int nsub, w0, w1;       // active windows (a multiple of 4) and their list, as GramArgs
};

// s_waitcnt immediates (vmcnt in bits 3:0 and 15:14, expcnt 6:4, lgkmcnt 11:8; the others
// at their maxima): every global load done; every LDS access done.
constexpr int kVmcnt0 = 0x0F70;
constexpr int kLgkmcnt0 = 0xC07F;

// Window of active entry 4c + k (c uniform across the wave, k = 0..3 per lane): the
// chunk's four windows are read as uniform values and picked per lane (no LDS round trip).
__device__ __forceinline__ int win_i8(const GramI8Args &g, int c, int k) {
    if (c >= kMaxWin / 4) return 4 * c + k;
    const int src = c < 16 ? g.w0 : g.w1, base = (4 * c) & 63;
    const int v0 = __builtin_amdgcn_readlane(src, base), v1 = __builtin_amdgcn_readlane(src, base + 1);
    const int v2 = __builtin_amdgcn_readlane(src, base + 2), v3 = __builtin_amdgcn_readlane(src, base + 3);
    return k < 2 ? (k == 0 ? v0 : v1) : (k == 2 ? v2 : v3);
}

// Row chunk c into the row buffer buf: wave w fills bytes [q 4096 + w 1024, +1024) of
// every slice q.
template <int NS>
__device__ __forceinline__ void stage_rows_i8(signed char *buf, const GramI8Args &g, int c) {
    const signed char *src = g.pa + win_i8(g, c, g.stk) * kI8WinBytes;
    signed char *dst = buf + (threadIdx.x >> 6) * 1024;
#pragma unroll
    for (int q = 0; q < NS; ++q)
        __builtin_amdgcn_global_load_lds(
            (const void *)(src + 16 * q),
            (__attribute__((address_space(3))) void *)(dst + q * 4096), 16, 0, 0);
}

// This wave's column chunk c into its region reg.
template <int NS>
__device__ __forceinline__ void stage_cols_i8(signed char *reg, const GramI8Args &g, int c) {
    const signed char *src = g.pb + win_i8(g, c, (threadIdx.x & 63) >> 4) * kI8WinBytes;
#pragma unroll
    for (int q = 0; q < NS; ++q)
        __builtin_amdgcn_global_load_lds(
            (const void *)(src + 16 * q),
            (__attribute__((address_space(3))) void *)(reg + q * 1024), 16, 0, 0);
}

// One chunk: products from row buffer cb and the column region while the next chunk
// lands in row buffer nb and (after its reads) the column region.
template <int NS, int NA, bool HB>
__device__ __forceinline__ void gram_i8_step(const signed char *cb, signed char *nb,
                                             signed char *reg, const GramI8Args &g, int ch,
                                             int nch, v4i (&acc)[4][NS]) {
    const int lane = threadIdx.x & 63;
    const bool more = ch + 1 < nch;
    v4i b[NS], a0[NS], a1[NS];
    __builtin_amdgcn_sched_barrier(0);
    if (HB) {
#pragma unroll
        for (int q = 0; q < NS; ++q) b[q] = *(const v4i *)(reg + q * 1024 + 16 * lane);
#pragma unroll
        for (int q = 0; q < NS; ++q) a0[q] = *(const v4i *)(cb + q * 4096 + g.rdo);
        if (more) __builtin_amdgcn_s_waitcnt(kLgkmcnt0);  // the region is rewritten next
    }
    if (more) {
        stage_rows_i8<NS>(nb, g, ch + 1);
        if (HB) stage_cols_i8<NS>(reg, g, ch + 1);
    }
    __builtin_amdgcn_sched_barrier(0);
    if (HB) {
#pragma unroll
        for (int x = 0; x < NA; ++x) {
            v4i(&ac)[NS] = (x & 1) ? a1 : a0;
            v4i(&an)[NS] = (x & 1) ? a0 : a1;
            if (x + 1 < NA) {
#pragma unroll
                for (int q = 0; q < NS; ++q)
                    an[q] = *(const v4i *)(cb + q * 4096 + (x + 1) * 1024 + g.rdo);
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int q = 0; q < NS; ++q)
#pragma unroll
                for (int p = 0; p < NS - q; ++p)
                    acc[x][q + p] = __builtin_amdgcn_mfma_i32_16x16x64_i8(ac[q], b[p],
                                                                          acc[x][q + p], 0, 0, 0);
            __builtin_amdgcn_sched_barrier(0);
        }
    }
    // the next chunk has landed (this wave's loads; the barrier covers the other waves')
    __builtin_amdgcn_s_waitcnt(kVmcnt0);
    __syncthreads();
}

// lds: the kernel's shared array (kI8Lds bytes): row buffers 0 and 1, then the column
// regions of waves 0..3 (one array, constant offsets: the compiler sees that loads into
// one part do not touch another).
template <int NS, int NA, bool HB>
__device__ void gram_i8_tile(signed char *__restrict__ lds, const GramI8Args &g,
                             v4i (&acc)[4][NS]) {
    const int nch = g.nsub >> 2;
    if (nch == 0) {  // no window is active: the Gram block is exactly zero
        __syncthreads();
        return;
    }
    signed char *const buf0 = lds, *const buf1 = lds + kI8RowBuf;
    signed char *const reg = lds + 2 * kI8RowBuf + (threadIdx.x >> 6) * kI8ColReg;
    stage_rows_i8<NS>(buf0, g, 0);
    if (HB) stage_cols_i8<NS>(reg, g, 0);
    __builtin_amdgcn_s_waitcnt(kVmcnt0);
    __syncthreads();
    for (int ch = 0; ch < nch; ch += 2) {
        gram_i8_step<NS, NA, HB>(buf0, buf1, reg, g, ch, nch, acc);
        if (ch + 1 < nch) gram_i8_step<NS, NA, HB>(buf1, buf0, reg, g, ch + 1, nch, acc);
    }
}

template <int NS>
__device__ __forceinline__ void gram_i8_dispatch(signed char *lds, const GramI8Args &g, int na16,
                                                 bool hb, v4i (&acc)[4][NS]) {
#define KF_I8_CASE(n)                                     \
    case n:                                               \
        if (hb) gram_i8_tile<NS, n, true>(lds, g, acc);   \
        else gram_i8_tile<NS, n, false>(lds, g, acc);     \
        break;
    switch (na16) {
        KF_I8_CASE(1)
        KF_I8_CASE(2)
        KF_I8_CASE(3)
        default:
            if (hb) gram_i8_tile<NS, 4, true>(lds, g, acc);
            else gram_i8_tile<NS, 4, false>(lds, g, acc);
            break;
    }
#undef KF_I8_CASE
}

// v + take * (value of the lane D places lower in the same 16-lane row), take = 0 or 1:
// one step of a segmented inclusive scan along the 16 columns a row of MFMA output spans
// (exact: the product is the shifted value or zero).
template <int D>
__device__ __forceinline__ double scan_step(double v, double take) {
    const int lo = __double2loint(v), hi = __double2hiint(v);
    const int slo = __builtin_amdgcn_update_dpp(0, lo, 0x110 + D, 0xF, 0xF, true);
    const int shi = __builtin_amdgcn_update_dpp(0, hi, 0x110 + D, 0xF, 0xF, true);
    return fma(__hiloint2double(shi, slo), take, v);
}

// Output block: 64 x 64 molecule pairs per work-group, accumulated in an accumulator slot
// in device memory (each entry owned by one thread at a time).  Per
// label, the row sub-tile sr is multiplied with two column sub-tiles at once (the row
// operand is staged once for both).  Epilogue per tile, in registers: exp of every
// Gram entry, then a segmented scan along each wave's 16 columns sums the columns of
// one molecule (a window segment: a molecule's columns inside one 16-column window);
// the segment tails go to LDS as partial sums P[row][segment], and each thread adds,
// for its column molecule's segments, the rows of its row molecules.  The molecule of a
// packed row is src[row] / max_atoms (src = a * max_atoms + i, see local_plan.hpp).
// The molecules of each set may be reordered (molecules that use the same descriptor
// windows sorted together, so that blocks share their zero windows): the kernel works in
// that order and perm maps a molecule back to its row / column of the output.
// This lane's column of a Gram tile in the epilogue.
struct TileCol {
    double n2;                  // cy * |x|^2 of its row of the column set
    double t1, t2, t4, t8;      // 1 where the scan step of that distance stays in the segment
    int mol;                    // its molecule (block-relative)
    int slot;                   // its segment if it is the segment's last column, else PP - 1
};

// Epilogue of one 64 x 64 Gram tile c (this wave's 16 columns): exp of every entry, doubled
// where DS and the row and column atoms are of one molecule (a sub-tile of a diagonal block
// whose mirror is not visited), then the segmented scan along the columns; the segment tails
// go to P[row][segment].  No entry is masked or branched on: rows past the tile are never
// read back, columns past it are the last lanes of a row and never feed a segment of the
// tile (a scan only takes from lower lanes of its own segment), and lanes that hold no
// segment tail store to the spare last slot of their row, which no segment uses.  n2r holds
// cy * |x|^2 of the rows.  I8: c is in the int32 MFMA layout (row 4 (lane / 16) + v) and
// scaled: the Gram entry is c * mur[row] * (this lane's column scale, folded into m2cy).
template <bool DS, bool I8>
__device__ __forceinline__ void tile_epilogue(const v4d (&c)[4], int lane, const TileCol &tc,
                                              const double *n2r, const double *mur,
                                              const int *rowmol, const double *etab, double m2cy,
                                              double *P) {
    double *const pl = P + (I8 ? 4 * (lane >> 4) : (lane >> 4)) * PP + tc.slot;
#pragma unroll
    for (int x = 0; x < 4; ++x) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = x * 16 + (I8 ? mfma16_row_f32(lane, v) : mfma16_row_f64(lane, v));
            const double g = I8 ? c[x][v] * mur[row] : c[x][v];
            double e = exp_scaled(fma(g, m2cy, n2r[row] + tc.n2), etab);
            if (DS && rowmol[row] == tc.mol) e *= 2.0;
            e = scan_step<1>(e, tc.t1);
            e = scan_step<2>(e, tc.t2);
            e = scan_step<4>(e, tc.t4);
            e = scan_step<8>(e, tc.t8);
            pl[(x * 16 + (I8 ? v : 4 * v)) * PP] = e;  // row x*16 + (the row above)
        }
    }
}

struct WinArgs {
    const ulonglong2 *tm1, *tm2;  // tile window masks (null: every window active)
    const int *tb1, *tb2;         // first tile mask of (label, block), l * nblk + block
    const int *perm1, *perm2;     // output index of each molecule (null: identity)
    const double *mu1, *mu2;      // int8-split form: the rows' scales (null: FP64 form)
    int nblk1, nblk2, nwin;       // blocks per set, windows of the padded descriptor
};

// NS = 0: FP64 operands (rows of rep_pad doubles), two column tiles per row tile;
// NS > 0: the int8-split form with NS slices (rows of rep_pad bytes, pack_rows_i8), one
// column tile per row tile (its accumulators take the registers of the second).
template <int MODE, int NS>
__global__ __launch_bounds__(NT, NS ? 2 : 3) void local_scalar_f64_kernel(
    const void *__restrict__ Xp1, const double *__restrict__ n2a, const int *__restrict__ mr1,
    const int *__restrict__ src1, int ma1, const void *__restrict__ Xp2,
    const double *__restrict__ n2b, const int *__restrict__ mr2, const int *__restrict__ src2,
    int ma2, int L, int nm1, int nm2, int rep_pad, double cy, const int2 *__restrict__ order,
    const WinArgs wa, double *__restrict__ out) {
    // two row-operand buffers (int8-split form: and the waves' column regions); the
    // segment partial sums of the epilogue reuse the same memory once the Gram products
    // are done
    constexpr int OPS = NS ? kI8Lds / 8 : 2 * TS * LDA, EPI = TS * PP;
    constexpr int NTM = NS ? 1 : 2;  // column tiles per row tile
    __shared__ double Sh[OPS > EPI ? OPS : EPI];
    __shared__ int s_slot;
    double *const Abuf = Sh, *const P = Sh;
    __shared__ double n2r[TS], n2c[2][TS], mur[NS ? TS : 1], muc[NS ? TS : 1];
    __shared__ int rowmol[TS], colmol[2][TS], colseg[2][TS], colinfo[2][TS];
    __shared__ int segf[2][MB], segl[2][MB], rstart[TS + 1], nrr_s;
    __shared__ int mrA[MB + 1], mrB[MB + 1];
    __shared__ double etab[64];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // work-group -> output block through the host-built order: the hardware deals
    // work-groups round-robin over the 8 XCDs, so logical slot (g % 8) * (G / 8) + g / 8
    // gives each XCD a contiguous run of the order (its list of column-block groups).
    const int gi = blockIdx.x, G = gridDim.x;
    const int2 ab = order[(G % 8 == 0) ? (gi % 8) * (G / 8) + gi / 8 : gi];
    if (ab.x < 0) return;  // padding of a shorter XCD list
    if (threadIdx.x == 0) {  // claim a free accumulator slot of this XCD
        const int base = (xcc_id() & (kXcc - 1)) * kSlotsPerXcc;
        int sl = (int)((blockIdx.x / kXcc) % kSlotsPerXcc);
        for (int t = 0; t < kSlotsPerXcc; ++t) {
            if (atomicCAS(&g_slot_busy[base + sl], 0, 1) == 0) break;
            sl = sl + 1 == kSlotsPerXcc ? 0 : sl + 1;
        }
        s_slot = base + sl;
    }
    __syncthreads();
    double *const accL = g_acc_slots + (size_t)s_slot * (MB * AP);
    const int A = ab.x, B = ab.y;
    const int a0 = A * MB, na = min(MB, nm1 - a0);
    const int b0 = B * MB, nb = min(MB, nm2 - b0);
    const bool diag = (MODE != RECT) && (A == B);
    const int ob = tid & (MB - 1), oa0 = (tid / MB) * (MB / 4);
    const int lrow = tid >> 2, bcol = wave * 16 + (lane & 15);
    GramArgs ga;
    ga.lofs = lrow * LDA + (tid & 3) * 4;
    ga.r16 = lane & 15;
    ga.kq = lane >> 4;
    GramI8Args g8;
    g8.stk = (lane & 3) ^ ((lane >> 3) & 3);
    g8.rdo = (lane & 15) * 64 + 16 * ((lane >> 4) ^ (((lane & 15) >> 1) & 3));
    for (int k = tid; k < MB * AP; k += NT) accL[k] = 0.0;
    if (tid < 64) etab[tid] = kExp2Tab64[tid];  // ordered by the label loop's first barrier

    for (int l = 0; l < L; ++l) {
        __syncthreads();
        if (tid <= na) mrA[tid] = mr1[(size_t)l * (nm1 + 1) + a0 + tid];
        if (tid >= 128 && tid - 128 <= nb) mrB[tid - 128] = mr2[(size_t)l * (nm2 + 1) + b0 + tid - 128];
        __syncthreads();
        const int ra0 = mrA[0], ra1 = mrA[na], rb0 = mrB[0], rb1 = mrB[nb];
        if (ra1 == ra0 || rb1 == rb0) continue;
        const int tbA = wa.tm1 ? wa.tb1[l * wa.nblk1 + A] : 0;
        const int tbB = wa.tm1 ? wa.tb2[l * wa.nblk2 + B] : 0;
        const int nsr = (ra1 - ra0 + TS - 1) / TS, nsc = (rb1 - rb0 + TS - 1) / TS;
        for (int sr = 0; sr < nsr; ++sr) {
            for (int sc0 = diag ? sr : 0; sc0 < nsc; sc0 += NTM) {
                const int r0 = ra0 + sr * TS, rn = min(TS, ra1 - r0);
                const int ntile = NTM == 2 && sc0 + 1 < nsc ? 2 : 1;
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
                    n2r[lane] = v ? cy * n2a[r0 + lane] : 0.0;
                    if (NS) mur[lane] = v ? wa.mu1[r0 + lane] : 0.0;
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
                    n2c[u][lane] = v ? cy * n2b[c0[u] + lane] : 0.0;
                    if (NS && u == 0) muc[lane] = v ? wa.mu2[c0[0] + lane] : 0.0;
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
                // active windows: non-zero in the row tile and in one of the column tiles,
                // padded with inactive ones (they add exact zeros) to a whole number of
                // MFMA steps
                constexpr int WQ = NS ? 4 : 2;
                unsigned long long m0 = ~0ull, m1 = ~0ull;
                int nsub = wa.nwin;
                if (wa.tm1) {
                    const ulonglong2 ta = wa.tm1[tbA + sr], tc0 = wa.tm2[tbB + sc0];
                    const ulonglong2 tc1 = ntile == 2 ? wa.tm2[tbB + sc0 + 1] : make_ulonglong2(0, 0);
                    m0 = ta.x & (tc0.x | tc1.x);
                    m1 = ta.y & (tc0.y | tc1.y);
                    nsub = __popcll(m0) + __popcll(m1);
                    const int nw = wa.nwin;
                    const unsigned long long v0 = nw >= 64 ? ~0ull : (1ull << nw) - 1;
                    const unsigned long long v1 = nw >= 128 ? ~0ull : (nw > 64 ? (1ull << (nw - 64)) - 1 : 0ull);
                    for (; nsub % WQ; ++nsub) {
                        const unsigned long long z0 = ~m0 & v0, z1 = ~m1 & v1;
                        if (z0) m0 |= z0 & (~z0 + 1);
                        else m1 |= z1 & (~z1 + 1);
                    }
                }
                int w0, w1;
                {
                    const int n0 = __popcll(m0);
                    w0 = lane < n0 ? nth_set_bit(m0, lane) : 64 + nth_set_bit(m1, lane - n0);
                    w1 = 64 + nth_set_bit(m1, 64 + lane - n0);
                }
                const int na16 = (rn + 15) >> 4;
                if constexpr (NS == 0) {
                    ga.pa = static_cast<const double *>(Xp1) + (size_t)(r0 + min(lrow, rn - 1)) * rep_pad + (tid & 3) * 4;
                    ga.pb0 = static_cast<const double *>(Xp2) + (size_t)(c0[0] + min(bcol, cn[0] - 1)) * rep_pad + 4 * ga.kq;
                    ga.pb1 = ntile == 2 ? static_cast<const double *>(Xp2) + (size_t)(c0[1] + min(bcol, cn[1] - 1)) * rep_pad + 4 * ga.kq
                                        : ga.pb0;
                    ga.w0 = w0;
                    ga.w1 = w1;
                    ga.nsub = nsub;
                    const int ntw = (wave * 16 < cn[1]) ? 2 : (wave * 16 < cn[0] ? 1 : 0);
                    if (ntile == 2)
                        gram_dispatch<2>(Abuf, ga, na16, ntw, c);
                    else
                        gram_dispatch<1>(Abuf, ga, na16, ntw, c);
                } else {
                    g8.pa = static_cast<const signed char *>(Xp1) + (size_t)(r0 + min(lrow, rn - 1)) * rep_pad;
                    g8.pb = static_cast<const signed char *>(Xp2) + (size_t)(c0[0] + min(bcol, cn[0] - 1)) * rep_pad;
                    g8.w0 = w0;
                    g8.w1 = w1;
                    g8.nsub = nsub;
                    v4i acc[4][NS ? NS : 1];
#pragma unroll
                    for (int x = 0; x < 4; ++x)
#pragma unroll
                        for (int d = 0; d < NS; ++d) acc[x][d] = (v4i){0, 0, 0, 0};
                    gram_i8_dispatch<NS>(reinterpret_cast<signed char *>(Sh), g8, na16,
                                         wave * 16 < cn[0], acc);
                    // sum_d 2^-7d acc_d in FP64 (the common 2^-12 and the scales in the epilogue)
#pragma unroll
                    for (int x = 0; x < 4; ++x)
#pragma unroll
                        for (int v = 0; v < 4; ++v) {
                            double t = (double)acc[x][NS - 1][v];
#pragma unroll
                            for (int d = NS - 2; d >= 0; --d) t = fma(t, 0x1p-7, (double)acc[x][d][v]);
                            c[0][x][v] = t;
                        }
                }

                // ---- per tile: exp, segment sums in registers, row molecules ----
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    if (u == 1 && ntile == 1) break;
                    const bool dbl_self = diag && sc0 + u > sr;  // mirror sub-tile not visited
                    if (u == 1) __syncthreads();  // tile 0's row sums are done with P
                    {
                        const int col = bcol;
                        const int info = colinfo[u][col], dist = info & 255;
                        const bool tail = col < cn[u] && ((info >> 8) & 1);
                        const TileCol tc{n2c[u][col],
                                         dist >= 1 ? 1.0 : 0.0, dist >= 2 ? 1.0 : 0.0,
                                         dist >= 4 ? 1.0 : 0.0, dist >= 8 ? 1.0 : 0.0,
                                         colmol[u][col], tail ? colseg[u][col] : PP - 1};
                        // int8-split form: the column's scale and the common 2^-12 here
                        const double m2 = NS ? -2.0 * cy * 0x1p-12 * muc[col] : -2.0 * cy;
                        if (dbl_self)
                            tile_epilogue<true, NS != 0>(c[u], lane, tc, n2r, mur, rowmol, etab, m2, P);
                        else
                            tile_epilogue<false, NS != 0>(c[u], lane, tc, n2r, mur, rowmol, etab, m2, P);
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
                            // four independent partial sums over the rows, so the LDS reads
                            // of a row molecule are in flight together
                            double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
                            for (int sg = s0; sg <= s1; ++sg) {
                                int i = rs;
                                for (; i + 3 < re; i += 4) {
                                    a0 += P[i * PP + sg];
                                    a1 += P[(i + 1) * PP + sg];
                                    a2 += P[(i + 2) * PP + sg];
                                    a3 += P[(i + 3) * PP + sg];
                                }
                                for (; i < re; ++i) a0 += P[i * PP + sg];
                            }
                            accL[rowmol[rs] * AP + ob] += (a0 + a1) + (a2 + a3);
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
        const size_t ia = wa.perm1 ? wa.perm1[a0 + a] : a0 + a;
        const size_t ib = wa.perm2 ? wa.perm2[b0 + b] : b0 + b;
        if (MODE == SYMM_RFP) {
            if (!diag || a <= b) out[rfp_upper((size_t)nm1, ia < ib ? ia : ib, ia < ib ? ib : ia)] = val;
        } else if (MODE == RECT) {
            out[ia * nm2 + ib] = val;
        } else if (!diag || a <= b) {
            out[ia * nm1 + ib] = val;
            out[ib * nm1 + ia] = val;
        }
    }
    __syncthreads();  // every read of the slot is done
    if (threadIdx.x == 0) atomicExch(&g_slot_busy[s_slot], 0);
}

// Before a launch of the scalar kernel: check, once per device, that the hardware reports
// XCC ids 0..kXcc-1 (the slot partition relies on them) and that fewer work-groups than
// slots of one XCD can be resident on an XCD; then clear the slot flags.
int check_xcc_ids(hipStream_t stream) {
    static std::mutex mu;
    static int checked[64] = {};
    int dev = 0;
    KF_HIP_TRY(hipGetDevice(&dev));
    KF_REQUIRE(dev >= 0 && dev < 64, KF_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(mu);
    if (checked[dev]) return KF_OK;
    hipLaunchKernelGGL(xcc_probe_kernel, dim3(kXccProbe), dim3(64), 0, stream);
    int h[kXccProbe];
    KF_HIP_TRY(hipMemcpyFromSymbolAsync(h, HIP_SYMBOL(g_xcc_probe), sizeof(h), 0,
                                        hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    unsigned seen = 0;
    for (int i = 0; i < kXccProbe; ++i) {
        KF_REQUIRE(h[i] >= 0 && h[i] < kXcc, KF_ERR_HIP, "unexpected XCC id %d", h[i]);
        seen |= 1u << h[i];
    }
    KF_REQUIRE(seen == (1u << kXcc) - 1, KF_ERR_HIP,
               "work-groups did not spread over %d XCC ids (mask 0x%x)", kXcc, seen);
    checked[dev] = 1;
    return KF_OK;
}

template <int MODE, int NS>
int prepare_acc_slots(hipStream_t stream) {
    int dev = 0, ncu = 0, per_cu = 0, rc = 0;
    if ((rc = check_xcc_ids(stream))) return rc;
    KF_HIP_TRY(hipGetDevice(&dev));
    KF_HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    KF_HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
        &per_cu, local_scalar_f64_kernel<MODE, NS>, NT, 0));
    KF_REQUIRE((long)(ncu + kXcc - 1) / kXcc * per_cu < kSlotsPerXcc, KF_ERR_HIP,
               "%d CUs x %d resident work-groups exceed the %d accumulator slots per XCD", ncu,
               per_cu, kSlotsPerXcc);
    hipLaunchKernelGGL(clear_slots_kernel, dim3(kSlots / 256), dim3(256), 0, stream);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

// Device copy of a plan plus the packed rows, their window masks and the row tiles'.
// FP64 form (ns = 0): Xp rows of rep_pad doubles; int8-split form (ns slices): Xp rows of
// rep_pad bytes (nwin windows of kI8WinBytes, pack_rows_i8) and mu the rows' scales.
struct DevicePacked {
    void *Xp = nullptr;
    double *n2 = nullptr, *mu = nullptr;
    int *mol_row = nullptr, *src = nullptr, *perm = nullptr, *tbase = nullptr;
    unsigned long long *rowmask = nullptr;
    ulonglong2 *tmask = nullptr;  // null: windows not masked
    int rows = 0, rep_pad = 0, nblk = 0, nwin = 0;
};

inline int padded_rep(int rep) { return (rep + KA - 1) / KA * KA; }
// windows of the int8-split rows: a multiple of 4 (one MFMA step)
inline int i8_windows(int rep) { return (rep + 63) / 64 * 4; }
inline bool windows_masked(int rep, int ns) {
    return (ns ? i8_windows(rep) : padded_rep(rep) / KS) <= kMaxWin;
}

// Window masks of every molecule of a set (device X) -> host (two words per molecule).
int molecule_windows(const double *X, const int *N, int nm, int ma, int rep, int ws,
                     hipStream_t stream, std::vector<unsigned long long> &out) {
    out.assign(2 * (size_t)nm, 0ull);
    if (nm == 0) return KF_OK;
    const size_t bytes = 2 * sizeof(unsigned long long) * (size_t)nm;
    auto *dm = static_cast<unsigned long long *>(workspace(ws, bytes, stream));
    auto *hm = static_cast<unsigned long long *>(host_staging(ws, bytes));
    KF_REQUIRE(dm && hm, KF_ERR_NOMEM, "workspace allocation failed (%zu B)", bytes);
    hipLaunchKernelGGL(mol_winmask_kernel, dim3((nm + 3) / 4), dim3(256), 0, stream, X, N, nm, ma,
                       rep, dm);
    KF_HIP_TRY(hipGetLastError());
    KF_HIP_TRY(hipMemcpyAsync(hm, dm, bytes, hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    std::copy(hm, hm + 2 * (size_t)nm, out.begin());
    return KF_OK;
}

// Windows ranked rarest first by the number of molecules (of all given sets) using them.
// The counts come from a histogram of byte values at each of the 16 byte positions of a
// molecule's mask pair (16 increments a molecule instead of one per set bit).
std::vector<int> rank_windows(const std::vector<const std::vector<unsigned long long> *> &sets,
                              int nwin) {
    std::vector<long> hist(16 * 256, 0);
    for (const auto *m : sets)
        for (size_t e = 0; e < m->size(); ++e) {  // word e & 1 holds windows 64 (e & 1) + ..
            const unsigned long long v = (*m)[e];
            if (!v) continue;
            const int h = (int)(e & 1);
            for (int by = 0; by < 8; ++by) ++hist[256 * (8 * h + by) + ((v >> (8 * by)) & 255)];
        }
    std::vector<long> freq(nwin, 0);
    for (int pos = 0; pos < 16; ++pos)
        for (int v = 1; v < 256; ++v) {
            const long c = hist[256 * pos + v];
            if (!c) continue;
            for (int b = 0; b < 8; ++b)
                if ((v >> b) & 1) {
                    const int w = 8 * pos + b;
                    if (w < nwin) freq[w] += c;
                }
        }
    std::vector<int> rank(nwin);
    for (int w = 0; w < nwin; ++w) rank[w] = w;
    std::stable_sort(rank.begin(), rank.end(), [&](int x, int y) { return freq[x] < freq[y]; });
    return rank;
}

// Molecule order: ascending by the bit string of the molecule's windows read rarest
// first (the 64 rarest windows), ties in input order.  Molecules that use the same
// rare windows end up in the same blocks, so the other blocks' tiles skip them.  A key is
// 16 table lookups (the key bits of every byte value at each byte position of the mask
// pair); the order is a stable LSD radix sort of the indices by key, 8 bits a pass,
// skipping the passes whose byte is the same for every key (the same order as sorting
// (key, index) pairs).
std::vector<int> order_molecules(const std::vector<unsigned long long> &mask,
                                 const std::vector<int> &rank) {
    typedef unsigned long long u64;
    const int nm = (int)(mask.size() / 2), nk = std::min<int>(64, (int)rank.size());
    // bit of window w in the key: rarest (rank position 0) most significant
    int shift[2 * 64];
    std::fill(shift, shift + 2 * 64, -1);
    for (int t = 0; t < nk; ++t) shift[rank[t]] = nk - 1 - t;
    std::vector<u64> tab(16 * 256, 0ull);
    for (int pos = 0; pos < 16; ++pos)
        for (int v = 1; v < 256; ++v) {
            u64 k = 0;
            for (int b = 0; b < 8; ++b)
                if ((v >> b) & 1) {
                    const int sh = shift[8 * pos + b];
                    if (sh >= 0) k |= 1ull << sh;
                }
            tab[256 * pos + v] = k;
        }
    std::vector<u64> key(nm);
    for (int a = 0; a < nm; ++a) {
        u64 k = 0;
        for (int h = 0; h < 2; ++h) {
            const u64 m = mask[2 * (size_t)a + h];
            if (!m) continue;
            for (int by = 0; by < 8; ++by) k |= tab[256 * (8 * h + by) + ((m >> (8 * by)) & 255)];
        }
        key[a] = k;
    }
    std::vector<int> perm(nm), tmp(nm);
    for (int a = 0; a < nm; ++a) perm[a] = a;
    for (int pass = 0; pass < 8 && nm > 0; ++pass) {
        const int sh = 8 * pass;
        size_t cnt[257] = {0};
        for (int a = 0; a < nm; ++a) ++cnt[((key[a] >> sh) & 255) + 1];
        if (cnt[((key[0] >> sh) & 255) + 1] == (size_t)nm) continue;  // nothing moves
        for (int v = 0; v < 256; ++v) cnt[v + 1] += cnt[v];
        for (int a = 0; a < nm; ++a) {
            const int m = perm[a];
            tmp[cnt[(key[m] >> sh) & 255]++] = m;
        }
        perm.swap(tmp);
    }
    return perm;
}

// Metadata of a set taken in the order perm (new molecule t = old molecule perm[t]).
void permute_metadata(const std::vector<int> &Q, const std::vector<int> &N, int ma,
                      const std::vector<int> &perm, std::vector<int> &Qp, std::vector<int> &Np) {
    const size_t nm = perm.size();
    Qp.resize(nm * ma);
    Np.resize(nm);
    for (size_t t = 0; t < nm; ++t) {
        std::copy(Q.begin() + (size_t)perm[t] * ma, Q.begin() + (size_t)(perm[t] + 1) * ma,
                  Qp.begin() + t * ma);
        Np[t] = N[perm[t]];
    }
}

// Upload the plan (plus the molecule order, if any), pack the rows and, when the windows
// are masked, build the masks of the kernel's row tiles.
int upload_and_pack(const LocalPlan &p, const std::vector<int> &perm, const double *X, int rep,
                    int ns, int ws_pack, int ws_plan, hipStream_t stream, DevicePacked &d) {
    d.rows = p.rows;
    d.nwin = ns ? i8_windows(rep) : padded_rep(rep) / KS;
    d.rep_pad = ns ? d.nwin * kI8WinBytes : padded_rep(rep);
    d.nblk = (p.nm + MB - 1) / MB;
    const bool masked = windows_masked(rep, ns);
    const size_t mr_n = p.mol_row.size(), src_n = p.src.size(), perm_n = perm.size();
    const size_t nla = (size_t)p.nlabels * d.nblk;
    // tile-mask bases from the host plan: tiles of (label, block) in the kernel's tiling
    std::vector<int> tbase(nla + 1, 0);
    for (int l = 0; l < p.nlabels; ++l)
        for (int A = 0; A < d.nblk; ++A) {
            const int *mr = p.mol_row.data() + (size_t)l * (p.nm + 1);
            const int rows = mr[std::min(A * MB + MB, p.nm)] - mr[A * MB];
            tbase[(size_t)l * d.nblk + A + 1] = tbase[(size_t)l * d.nblk + A] + (rows + TS - 1) / TS;
        }
    const size_t ntile_masks = masked ? (size_t)tbase[nla] : 0;
    const size_t plan_n = mr_n + src_n + perm_n + nla + 1;
    char *plan = static_cast<char *>(workspace(ws_plan, sizeof(int) * plan_n, stream));
    // operand rows, then n2 (and mu) of the rows
    const size_t op_bytes = ((ns ? (size_t)p.rows * d.rep_pad
                                 : sizeof(double) * (size_t)p.rows * d.rep_pad) + 255) / 256 * 256;
    const size_t xp_bytes = op_bytes + sizeof(double) * (2 * (size_t)p.rows + 64);
    const size_t rm_off = (xp_bytes + 255) / 256 * 256;
    const size_t tm_off = rm_off + (2 * sizeof(unsigned long long) * (size_t)p.rows + 255) / 256 * 256;
    const size_t pack_bytes = tm_off + sizeof(ulonglong2) * (ntile_masks + 1);
    char *pack = static_cast<char *>(workspace(ws_pack, pack_bytes, stream));
    KF_REQUIRE(plan && pack, KF_ERR_NOMEM, "device workspace allocation failed (%zu B)",
               pack_bytes);
    d.mol_row = reinterpret_cast<int *>(plan);
    d.src = d.mol_row + mr_n;
    d.perm = perm_n ? d.src + src_n : nullptr;
    d.tbase = d.src + src_n + perm_n;
    int *stage = static_cast<int *>(host_staging(ws_plan, sizeof(int) * plan_n));
    KF_REQUIRE(stage, KF_ERR_NOMEM, "pinned staging allocation failed");
    std::copy(p.mol_row.begin(), p.mol_row.end(), stage);
    std::copy(p.src.begin(), p.src.end(), stage + mr_n);
    std::copy(perm.begin(), perm.end(), stage + mr_n + src_n);
    std::copy(tbase.begin(), tbase.end(), stage + mr_n + src_n + perm_n);
    KF_HIP_TRY(hipMemcpyAsync(plan, stage, sizeof(int) * plan_n, hipMemcpyHostToDevice, stream));
    d.Xp = pack;
    d.n2 = reinterpret_cast<double *>(pack + op_bytes);
    d.mu = ns ? d.n2 + p.rows : nullptr;
    d.rowmask = reinterpret_cast<unsigned long long *>(pack + rm_off);
    d.tmask = masked ? reinterpret_cast<ulonglong2 *>(pack + tm_off) : nullptr;
    if (p.rows > 0) {
        if (ns == kI8Slices)
            hipLaunchKernelGGL(pack_rows_i8<kI8Slices>, dim3((p.rows + 3) / 4), dim3(256), 0,
                               stream, X, d.src, d.perm, p.max_atoms, p.rows, rep, d.nwin,
                               static_cast<signed char *>(d.Xp), d.n2, d.mu, d.rowmask);
        else
            hipLaunchKernelGGL(pack_rows_f64, dim3((p.rows + 3) / 4), dim3(256), 0, stream, X,
                               d.src, d.perm, p.max_atoms, p.rows, rep, d.rep_pad,
                               static_cast<double *>(d.Xp), d.n2, d.rowmask);
        KF_HIP_TRY(hipGetLastError());
        if (masked && nla > 0) {
            hipLaunchKernelGGL(tile_mask_kernel, dim3((unsigned)nla), dim3(256), 0, stream,
                               d.rowmask, d.mol_row, p.nm, d.nblk, d.tbase, d.tmask);
            KF_HIP_TRY(hipGetLastError());
        }
    }
    return KF_OK;
}

WinArgs win_args(const DevicePacked &d1, const DevicePacked &d2) {
    WinArgs w;
    w.tm1 = d1.tmask && d2.tmask ? d1.tmask : nullptr;
    w.tm2 = w.tm1 ? d2.tmask : nullptr;
    w.tb1 = d1.tbase;
    w.tb2 = d2.tbase;
    w.perm1 = d1.perm;
    w.perm2 = d2.perm;
    w.mu1 = d1.mu;
    w.mu2 = d2.mu;
    w.nblk1 = d1.nblk;
    w.nblk2 = d2.nblk;
    w.nwin = d1.nwin;
    return w;
}

// Output-block order.  Work-group g runs on XCD g % 8 (the hardware deals work-groups
// round-robin over the 8 XCDs), and the kernel maps it to logical slot
// (g % 8) * (G / 8) + g / 8, so each XCD walks its own contiguous list.  The list of an XCD
// is a sequence of column-block groups -- every row block A paired with one column block B
// -- so the work-groups resident on an XCD at one time share their column operands (the
// larger operand stream, read by every wave straight from memory) through its L2.  Column
// blocks are dealt to the XCDs largest group first, each to the list with the fewest
// pairs; lists are padded to one length with (-1, -1) entries, which exit at once.
int upload_block_order(int nb1, int nb2, bool symm, hipStream_t stream, const int2 **dev,
                       unsigned *grid) {
    constexpr int X = 8;
    std::vector<std::vector<int2>> lists(X);
    std::vector<int> cols(nb2);
    for (int b = 0; b < nb2; ++b) cols[b] = b;
    if (symm) std::reverse(cols.begin(), cols.end());  // largest groups first
    for (int b : cols) {
        size_t x = 0;
        for (size_t y = 1; y < lists.size(); ++y)
            if (lists[y].size() < lists[x].size()) x = y;
        const int na = symm ? b + 1 : nb1;
        for (int a = 0; a < na; ++a) lists[x].push_back(make_int2(a, b));
    }
    size_t len = 0;
    for (const auto &l : lists) len = std::max(len, l.size());
    std::vector<int2> order;
    order.reserve(X * len);
    for (auto &l : lists) {
        l.resize(len, make_int2(-1, -1));
        order.insert(order.end(), l.begin(), l.end());
    }
    const size_t bytes = order.size() * sizeof(int2);
    int2 *d = static_cast<int2 *>(workspace(WS_AUX1, bytes, stream));
    int2 *h = static_cast<int2 *>(host_staging(WS_AUX1, bytes));
    KF_REQUIRE(d && h, KF_ERR_NOMEM, "workspace allocation failed");
    std::copy(order.begin(), order.end(), h);
    KF_HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream));
    *dev = d;
    *grid = (unsigned)order.size();
    return KF_OK;
}

// Launch of the scalar kernel in its FP64 form (ns = 0) or int8-split form (ns slices).
template <int MODE>
int launch_scalar(int ns, const DevicePacked &d1, int ma1, const DevicePacked &d2, int ma2, int L,
                  int nm1, int nm2, double cy, const int2 *order, unsigned grid, double *out,
                  int timer, hipStream_t stream) {
    int rc = 0;
    if (ns == kI8Slices) {
        // int32 anti-diagonal sums: ns slice pairs of |q| <= 64 over the padded columns
        KF_REQUIRE((long long)ns * d1.nwin * 16 * 64 * 64 < (1ll << 31), KF_ERR_ARG,
                   "descriptor too wide (%d windows) for the int8-split form", d1.nwin);
        if ((rc = prepare_acc_slots<MODE, kI8Slices>(stream))) return rc;
        ScopedTimer t(timer, stream);
        hipLaunchKernelGGL((local_scalar_f64_kernel<MODE, kI8Slices>), dim3(grid), dim3(NT), 0,
                           stream, d1.Xp, d1.n2, d1.mol_row, d1.src, ma1, d2.Xp, d2.n2,
                           d2.mol_row, d2.src, ma2, L, nm1, nm2, d1.rep_pad, cy, order,
                           win_args(d1, d2), out);
        KF_HIP_TRY(hipGetLastError());
        return KF_OK;
    }
    if ((rc = prepare_acc_slots<MODE, 0>(stream))) return rc;
    ScopedTimer t(timer, stream);
    hipLaunchKernelGGL((local_scalar_f64_kernel<MODE, 0>), dim3(grid), dim3(NT), 0, stream, d1.Xp,
                       d1.n2, d1.mol_row, d1.src, ma1, d2.Xp, d2.n2, d2.mol_row, d2.src, ma2, L,
                       nm1, nm2, d1.rep_pad, cy, order, win_args(d1, d2), out);
    KF_HIP_TRY(hipGetLastError());
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
    KF_REQUIRE(precision == KF_PREC_F64 || precision == KF_PREC_BF16X3 ||
                   precision == KF_PREC_I8X6,
               KF_ERR_ARG, "precision %d not supported here", precision);
    const int ns = precision == KF_PREC_I8X6 ? kI8Slices : 0;
    if (nm == 0) return KF_OK;
    KF_REQUIRE(out, KF_ERR_ARG, "null output");
    std::vector<int> hQ, hN;
    if ((rc = read_metadata(Q, N, nm, ma, hQ, hN, stream, WS_AUX0))) return rc;
    const std::vector<int> labels = distinct_labels(hQ.data(), hN.data(), nm, ma);
    LocalPlan p;
    if (precision == KF_PREC_BF16X3) {
        build_plan(hQ.data(), hN.data(), nm, ma, labels, p);
        return local_mixed_symm(MODE == SYMM_DENSE, X, nm, ma, rep, sigma, p, out, stream);
    }
    std::vector<int> perm, Qp, Np;
    if (windows_masked(rep, ns)) {
        std::vector<unsigned long long> mm;
        if ((rc = molecule_windows(X, N, nm, ma, rep, WS_AUX2, stream, mm))) return rc;
        perm = order_molecules(mm, rank_windows({&mm}, padded_rep(rep) / KS));
        permute_metadata(hQ, hN, ma, perm, Qp, Np);
        build_plan(Qp.data(), Np.data(), nm, ma, labels, p);
    } else {
        build_plan(hQ.data(), hN.data(), nm, ma, labels, p);
    }
    DevicePacked d;
    if ((rc = upload_and_pack(p, perm, X, rep, ns, WS_PACK, WS_PLAN, stream, d))) return rc;
    const int nblk = (nm + MB - 1) / MB;
    const double cy = -kExpScale / (2.0 * sigma * sigma);  // squared distance -> exp_scaled's y
    if (p.nlabels == 0) {  // no atoms at all: zero output
        const size_t n = MODE == SYMM_RFP ? (size_t)nm * (nm + 1) / 2 : (size_t)nm * nm;
        KF_HIP_TRY(hipMemsetAsync(out, 0, n * sizeof(double), stream));
        return KF_OK;
    }
    const int2 *order = nullptr;
    unsigned grid = 0;
    if ((rc = upload_block_order(nblk, nblk, true, stream, &order, &grid))) return rc;
    return launch_scalar<MODE>(ns, d, ma, d, ma, p.nlabels, nm, nm, cy, order, grid,
                               static_cast<double *>(out), TM_LOCAL_SYMM, stream);
}

}  // namespace
}  // namespace kf

extern "C" int kf_local_kernel_symm_rfp(int precision, const double *X, const int *Q,
                                        const int *N, int nm, int max_atoms, int rep,
                                        double sigma, void *arf, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return kf::symm_impl<kf::SYMM_RFP>(precision, X, Q, N, nm, max_atoms, rep, sigma, arf,
                                       kf::as_stream(stream));
}

extern "C" int kf_local_kernel_symm(int precision, const double *X, const int *Q, const int *N,
                                    int nm, int max_atoms, int rep, double sigma, void *K,
                                    void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    return kf::symm_impl<kf::SYMM_DENSE>(precision, X, Q, N, nm, max_atoms, rep, sigma, K,
                                         kf::as_stream(stream));
}

extern "C" int kf_local_kernel(int precision, const double *X1, const double *X2, const int *Q1,
                               const int *Q2, const int *N1, const int *N2, int nm1, int nm2,
                               int max_atoms1, int max_atoms2, int rep, double sigma, void *K,
                               void *stream_) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream_));
    using namespace kf;
    hipStream_t stream = as_stream(stream_);
    int rc = check_common(X1, Q1, N1, nm1, max_atoms1, rep, sigma);
    if (rc) return rc;
    if ((rc = check_common(X2, Q2, N2, nm2, max_atoms2, rep, sigma))) return rc;
    KF_REQUIRE(precision == KF_PREC_F64 || precision == KF_PREC_BF16X3 ||
                   precision == KF_PREC_I8X6,
               KF_ERR_ARG, "precision %d not supported here", precision);
    const int ns = precision == KF_PREC_I8X6 ? kI8Slices : 0;
    if (nm1 == 0 || nm2 == 0) return KF_OK;
    KF_REQUIRE(K, KF_ERR_ARG, "null output");
    std::vector<int> q1, n1, q2, n2;
    if ((rc = read_metadata(Q1, N1, nm1, max_atoms1, q1, n1, stream, WS_AUX0))) return rc;
    if ((rc = read_metadata(Q2, N2, nm2, max_atoms2, q2, n2, stream, WS_AUX1))) return rc;
    std::vector<int> l1 = distinct_labels(q1.data(), n1.data(), nm1, max_atoms1);
    std::vector<int> l2 = distinct_labels(q2.data(), n2.data(), nm2, max_atoms2);
    std::vector<int> labels;
    std::set_intersection(l1.begin(), l1.end(), l2.begin(), l2.end(), std::back_inserter(labels));
    const size_t esz = precision == KF_PREC_BF16X3 ? sizeof(float) : sizeof(double);
    if (labels.empty()) {
        KF_HIP_TRY(hipMemsetAsync(K, 0, esz * (size_t)nm1 * nm2, stream));
        return KF_OK;
    }
    LocalPlan p1, p2;
    if (precision == KF_PREC_BF16X3) {
        build_plan(q1.data(), n1.data(), nm1, max_atoms1, labels, p1);
        build_plan(q2.data(), n2.data(), nm2, max_atoms2, labels, p2);
        return local_mixed_rect(X1, X2, nm1, nm2, max_atoms1, max_atoms2, rep, sigma, p1, p2,
                                static_cast<float *>(K), stream);
    }
    std::vector<int> perm1, perm2;
    if (windows_masked(rep, ns)) {
        std::vector<unsigned long long> m1, m2;
        if ((rc = molecule_windows(X1, N1, nm1, max_atoms1, rep, WS_AUX2, stream, m1))) return rc;
        if ((rc = molecule_windows(X2, N2, nm2, max_atoms2, rep, WS_AUX3, stream, m2))) return rc;
        const std::vector<int> rank = rank_windows({&m1, &m2}, padded_rep(rep) / KS);
        perm1 = order_molecules(m1, rank);
        perm2 = order_molecules(m2, rank);
        std::vector<int> qp, np;
        permute_metadata(q1, n1, max_atoms1, perm1, qp, np);
        build_plan(qp.data(), np.data(), nm1, max_atoms1, labels, p1);
        permute_metadata(q2, n2, max_atoms2, perm2, qp, np);
        build_plan(qp.data(), np.data(), nm2, max_atoms2, labels, p2);
    } else {
        build_plan(q1.data(), n1.data(), nm1, max_atoms1, labels, p1);
        build_plan(q2.data(), n2.data(), nm2, max_atoms2, labels, p2);
    }
    DevicePacked d1, d2;
    if ((rc = upload_and_pack(p1, perm1, X1, rep, ns, WS_PACK, WS_PLAN, stream, d1))) return rc;
    if ((rc = upload_and_pack(p2, perm2, X2, rep, ns, WS_PACK2, WS_PLAN2, stream, d2))) return rc;
    const int nblk1 = (nm1 + MB - 1) / MB, nblk2 = (nm2 + MB - 1) / MB;
    const double cy = -kExpScale / (2.0 * sigma * sigma);  // squared distance -> exp_scaled's y
    const int2 *order = nullptr;
    unsigned grid = 0;
    if ((rc = upload_block_order(nblk1, nblk2, false, stream, &order, &grid))) return rc;
    return launch_scalar<RECT>(ns, d1, max_atoms1, d2, max_atoms2, (int)labels.size(), nm1, nm2,
                               cy, order, grid, static_cast<double *>(K), TM_LOCAL_RECT, stream);
}
