// rfp_chol.hip -- Cholesky factorisation and solve of an RFP-packed SPD matrix on gfx950.
//
// Reference: kf::math::solve_cholesky_rfp (src/math.cpp:134-171) = add l2 to the RFP
// diagonal, LAPACK dpftrf('N','U') then dpftrs('N','U').
//
// Design: the RFP array (TRANSR='N', UPLO='U') is addressed as the LOWER triangle of
// the full matrix (rfp_lower in kf_common.hpp): rows < n/2 form a column-major block,
// rows >= n/2 a row-major block, both with leading dimension s.  The panel partition
// is aligned at n/2, so every 64x64 tile lies in one region and is a plain strided
// matrix.  A right-looking blocked Cholesky then runs in place on the packed array:
//   per panel J:  POTRF of the 64x64 diagonal tile + its triangular inverse (one WG,
//                 LDS), panel TRSM as an MFMA product with the inverse, and the
//                 trailing SYRK/GEMM update on MFMA (v_mfma_f64_16x16x4 / f32_16x16x4).
// The factor lands exactly where LAPACK's dpftrf leaves U (= L^T of this layout).
// Solves run forward (L) and backward (L^T) over the same partition in FP64.
#include <cmath>

#include "kf_common.hpp"

namespace kf {
namespace {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

constexpr int NB = 64;        // panel width / tile edge
constexpr int TP = NB + 1;    // LDS pitch

__device__ inline v4d mfma4(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ inline v4f mfma4(float a, float b, v4f c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
template <typename T> struct Vec4;
template <> struct Vec4<double> { typedef v4d type; };
template <> struct Vec4<float> { typedef v4f type; };

// Partition of [0, n): 64-wide blocks inside [0, k) then inside [k, n).
struct Part {
    int n, k, nlo, nblk;  // nlo = number of blocks in [0, k)
    __host__ __device__ int start(int J) const {
        return J < nlo ? J * NB : k + (J - nlo) * NB;
    }
    __host__ __device__ int width(int J) const {
        const int s = start(J);
        const int end = J < nlo ? k : n;
        return (end - s) < NB ? (end - s) : NB;
    }
};
inline Part make_part(int n) {
    Part p;
    p.n = n;
    p.k = n / 2;
    p.nlo = (p.k + NB - 1) / NB;
    p.nblk = p.nlo + (n - p.k + NB - 1) / NB;
    return p;
}

// Tile view: element (r, c) of the lower triangle at base + r*rs + c*cs (tile-local).
struct Tile {
    size_t base;
    long rs, cs;
};
__device__ inline Tile tile_at(const Part &P, int r0, int c0) {
    const size_t s = (P.n & 1) ? P.n : P.n + 1;
    Tile t;
    if (r0 < P.k) {  // column-major region
        t.base = (size_t)(P.k + 1 + r0) + (size_t)c0 * s;
        t.rs = 1;
        t.cs = (long)s;
    } else {  // row-major region
        t.base = (size_t)(r0 - P.k) * s + c0;
        t.rs = (long)s;
        t.cs = 1;
    }
    return t;
}

// Load a (rn x cn) tile into LDS dst[r*TP + c] with NT threads; coalesced along the
// contiguous axis.  The trip count is a compile-time constant and the loop is
// unrolled so that the loads are in flight together (a one-at-a-time loop pays the
// memory latency NB*NB/NT times).
template <typename T, bool LOWER_ONLY, int NT = 256>
__device__ inline void load_tile(const T *__restrict__ a, const Tile &t, int rn, int cn, T *dst) {
#pragma unroll 16
    for (int e = 0; e < NB * NB / NT; ++e) {
        const int idx = threadIdx.x + NT * e;
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        // unconditional load of a clamped (valid) element, then select.  The condition is
        // formed with non-short-circuit '&': with '&&' the compiler sinks the load into
        // the branch of the first comparison, and every load of the unrolled loop is then
        // waited for before the next one is issued.
        const T v = a[t.base + min(r, rn - 1) * t.rs + min(c, cn - 1) * t.cs];
        const bool keep = (r < rn) & (c < cn) & (!LOWER_ONLY | (r >= c));
        dst[r * TP + c] = keep ? v : T(0);
    }
}
template <typename T, bool LOWER_ONLY>
__device__ inline void store_tile(T *__restrict__ a, const Tile &t, int rn, int cn, const T *src) {
    for (int idx = threadIdx.x; idx < NB * NB; idx += blockDim.x) {
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        if (r < rn && c < cn && (!LOWER_ONLY || r >= c)) a[t.base + r * t.rs + c * t.cs] = src[r * TP + c];
    }
}

template <typename T>
__global__ void add_diag(T *a, int n, double l2) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < n) a[rfp_lower((size_t)n, j, j)] += T(l2);
}

// Li = L^-1 for a lower-triangular w x w LDS tile (pitch TP), right-looking
// substitution on the identity, one barrier per row: phase q scales row q into the
// unused upper triangle of Li (and dgi[q]) and subtracts L[r][q] * row q / L[q][q] from
// every row r > q; the lower triangle is restored at the end.  Li is w x w, fully
// written (zeros above the diagonal).
template <typename T, typename U>
__device__ inline void tri_inverse(const T *L, U *Li, U *dgi, int w) {
    const int ti = threadIdx.x >> 4, tc = threadIdx.x & 15;
    for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
        const int i = idx / w, c = idx - i * w;
        Li[i * TP + c] = (i == c) ? U(1) : U(0);
    }
    __syncthreads();
    for (int q = 0; q < w; ++q) {
        const U rq = U(1) / (U)L[q * TP + q];
        if (ti == 0)
            for (int c = tc; c <= q; c += 16) {
                const U v = Li[q * TP + c] * rq;
                if (c < q) Li[c * TP + q] = v; else dgi[q] = v;
            }
        for (int r = q + 1 + ti; r < w; r += 16) {
            const U lr = (U)L[r * TP + q] * rq;
            for (int c = tc; c <= q; c += 16) Li[r * TP + c] -= lr * Li[q * TP + c];
        }
        __syncthreads();
    }
    for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
        const int i = idx / w, c = idx - i * w;
        if (i > c) {
            Li[i * TP + c] = Li[c * TP + i];
        } else if (i == c) {
            Li[i * TP + i] = dgi[i];
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < w * w; idx += blockDim.x) {
        const int i = idx / w, c = idx - i * w;
        if (i < c) Li[i * TP + c] = U(0);
    }
    __syncthreads();
}

// POTRF of diagonal tile J plus its inverse (lower) into inv[NB*NB], four waves.
// Thread (wave q, lane i) keeps columns 16q .. 16q+15 of row i in registers.  Cholesky,
// blocked right-looking over the four 16-column quarters: the wave that owns a quarter
// factors its 64 x 16 panel alone (left-looking, the in-panel updates through readlane,
// no barrier) and publishes it in LDS; after one barrier every wave right of it applies the
// rank-16 update to its own columns.  Inverse, blocked on the two 32-row halves,
//   L = [[L0, 0], [B, L1]],  L^-1 = [[L0^-1, 0], [-L1^-1 B L0^-1, L1^-1]]:
// lane i of wave 0 runs forward substitution for column i % 32 of the inverse of diagonal
// block i / 32 (both blocks at once, 32 steps) while waves 1..3 store the factor, and the
// off-diagonal block is two 32x32x32 MFMA products, one 16x16 fragment per wave.  A partial
// tile (w < 64) is padded with the identity, which keeps both factors block diagonal.
template <typename T>
__device__ __forceinline__ T lane_value(T v, int l);
template <>
__device__ __forceinline__ double lane_value<double>(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
template <>
__device__ __forceinline__ float lane_value<float>(float v, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ double recip_sqrt(double v) { return rsqrt(v); }
__device__ __forceinline__ float recip_sqrt(float v) { return rsqrtf(v); }

constexpr int HB = NB / 2;   // half block of the blocked inverse
constexpr int HP = HB + 1;   // LDS pitch of its pieces

template <typename T>
__global__ __launch_bounds__(256) void potrf_diag(T *a, Part P, int J, T *inv, int *info) {
    constexpr int PQ = 18;  // pitch of a published quarter panel (16-byte aligned rows)
    constexpr int WK = 2 * NB * PQ > (NB + HB) * HP ? 2 * NB * PQ : (NB + HB) * HP;
    __shared__ T S[NB * TP];
    __shared__ T Wk[WK];  // the quarter panels while factoring, then Di and Zs
    __shared__ T dinv[NB];
    __shared__ int fail_s[2];
    T *const Di = Wk;            // row 32h + q, column c: (L_h^-1)[q][c]
    T *const Zs = Wk + NB * HP;  // B L0^-1
    if (*info) return;
    // these waves are the whole critical path of the factorization: let them issue ahead of
    // the trailing-update waves that share their CU
    __builtin_amdgcn_s_setprio(3);
    const int tid = threadIdx.x, i = tid & 63, q = tid >> 6;
    const int s0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, s0, s0);
    load_tile<T, true, 256>(a, t, w, w, S);
    __syncthreads();
    if (tid >= w && tid < NB) S[tid * TP + tid] = T(1);
    __syncthreads();
    T r[16];
#pragma unroll
    for (int c = 0; c < 16; ++c) r[c] = 16 * q + c <= i ? S[i * TP + 16 * q + c] : T(0);
    // Quarter jb: its wave factors the 64 x 16 panel alone (left-looking, the in-panel
    // updates by readlane, no barrier), zeroing the entries above the diagonal, and
    // publishes it; after one barrier every wave right of it applies the rank-16 update to
    // its own columns.  The panel and the failure flag alternate between two buffers
    // (rewritten two quarters later, after every wave has passed the barrier that follows
    // its reads).
    int fail = 0;
    for (int jb = 0; jb < 4; ++jb) {
        T *const pq = Wk + (jb & 1) * NB * PQ;
        if (q == jb) {  // uniform
            int f = 0;
#pragma unroll
            for (int cj = 0; cj < 16; ++cj) {
                const int j = 16 * jb + cj;
                // left-looking inside the panel: column j first receives its finished
                // columns' contributions (readlanes of final values, issued back to back)
                if (cj > 0) {
                    T acc[4] = {r[cj], T(0), T(0), T(0)};
#pragma unroll
                    for (int m = 0; m < cj; ++m) acc[m & 3] -= r[m] * lane_value<T>(r[m], j);
                    r[cj] = (acc[0] + acc[1]) + (acc[2] + acc[3]);
                }
                const T piv = lane_value<T>(r[cj], j);
                if (!f && !(piv > T(0))) f = j + 1;  // also catches NaN; later values unused
                const T rd = recip_sqrt(piv);
                r[cj] = i == j ? piv * rd : (i > j ? r[cj] * rd : T(0));
                if (i == 0) dinv[j] = rd;
            }
#pragma unroll
            for (int m = 0; m < 16; ++m) pq[i * PQ + m] = r[m];
            if (i == 0) fail_s[jb & 1] = f;
        }
        __syncthreads();
        fail = fail_s[jb & 1];
        if (fail) break;
        if (q > jb) {
            T li[16];
#pragma unroll
            for (int m = 0; m < 16; ++m) li[m] = pq[i * PQ + m];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                const T *const pk = pq + (16 * q + c) * PQ;
                T acc = r[c];
#pragma unroll
                for (int m = 0; m < 16; ++m) acc -= li[m] * pk[m];
                r[c] = acc;
            }
        }
    }
    if (fail) {
        if (tid == 0 && fail <= w) atomicCAS(info, 0, s0 + fail);
        return;
    }
#pragma unroll
    for (int c = 0; c < 16; ++c) S[i * TP + 16 * q + c] = 16 * q + c <= i ? r[c] : T(0);
    __syncthreads();
    const int h = i >> 5, cc = i & (HB - 1);
    if (q == 0) {
        // column cc of L_h^-1 (h = i / 32): x[p] = (delta_p,cc - sum_{k<p} L_h[p][k] x[k]) / L_h[p][p],
        // with the reciprocals of the diagonal kept from the factorization
        T x[HB];
#pragma unroll
        for (int p = 0; p < HB; ++p) {
            const int g = h * HB + p;
            T lq[HB];
#pragma unroll
            for (int k = 0; k < p; ++k) lq[k] = S[g * TP + h * HB + k];
            const T dq = dinv[g];
            __builtin_amdgcn_sched_barrier(0);
            T acc[4] = {p == cc ? T(1) : T(0), T(0), T(0), T(0)};
#pragma unroll
            for (int k = 0; k < p; ++k) acc[k & 3] -= lq[k] * x[k];
            x[p] = ((acc[0] + acc[1]) + (acc[2] + acc[3])) * dq;
        }
#pragma unroll
        for (int p = 0; p < HB; ++p) Di[(h * HB + p) * HP + cc] = x[p];
    } else {  // waves 1..3 store the factor meanwhile (coalesced along the contiguous axis)
        for (int idx = tid - 64; idx < NB * NB; idx += 192) {
            int rr, c;
            if (t.rs == 1) { rr = idx % NB; c = idx / NB; } else { c = idx % NB; rr = idx / NB; }
            if (rr < w && c < w && rr >= c) a[t.base + rr * t.rs + c * t.cs] = S[rr * TP + c];
        }
    }
    __syncthreads();
    // Z = B L0^-1, then W = -L1^-1 Z, as 16x16x4 MFMA fragments (A[m][k] from lane
    // m + 16k, B[k][n] from lane n + 16k), fragment q on wave q
    typedef typename Vec4<T>::type V4;
    const int l16 = i & 15, kq = i >> 4, u = q >> 1, v = q & 1;
    {
        V4 z = V4{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < HB; k0 += 4)
            z = mfma4(S[(HB + u * 16 + l16) * TP + k0 + kq], Di[(k0 + kq) * HP + v * 16 + l16], z);
#pragma unroll
        for (int e = 0; e < 4; ++e) Zs[(u * 16 + mfma16_row<T>(i, e)) * HP + v * 16 + l16] = z[e];
    }
    __syncthreads();
    {
        V4 z = V4{0, 0, 0, 0};
#pragma unroll
        for (int k0 = 0; k0 < HB; k0 += 4)
            z = mfma4(-Di[(HB + u * 16 + l16) * HP + k0 + kq], Zs[(k0 + kq) * HP + v * 16 + l16], z);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            const int rr = HB + u * 16 + mfma16_row<T>(i, e), c = v * 16 + l16;
            inv[rr * NB + c] = (rr < w && c < w) ? z[e] : T(0);
        }
    }
    // inv (row-major, NB x NB): the diagonal blocks from Di, zeros above them; rows 16q ..
    // 16q+15 on wave q
#pragma unroll
    for (int p = 16 * q; p < 16 * q + 16; ++p) {
        const int hq = p / HB;
        if (hq == h) inv[p * NB + i] = (p < w && i < w) ? Di[p * HP + cc] : T(0);
        else if (hq == 0) inv[p * NB + i] = T(0);  // above the diagonal blocks
    }
}

// Panel: A_IJ <- A_IJ * inv^T for every block row I > J (one WG per I).
template <typename T>
__global__ __launch_bounds__(256) void trsm_panel(T *a, Part P, int J, const T *inv, const int *info) {
    __shared__ T Pn[NB * TP];
    __shared__ T Iv[NB * TP];
    if (*info) return;
    const int I = J + 1 + blockIdx.x;
    const int r0 = P.start(I), rn = P.width(I), c0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, r0, c0);
    load_tile<T, false>(a, t, rn, w, Pn);
#pragma unroll
    for (int e = 0; e < NB * NB / 256; ++e) {
        const int idx = threadIdx.x + 256 * e;
        Iv[(idx / NB) * TP + idx % NB] = inv[idx];
    }
    __syncthreads();
    // out[r][c] = sum_q Pn[r][q] * Iv[c][q]   (q <= c)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    typename Vec4<T>::type acc[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) acc[i][j] = typename Vec4<T>::type{0, 0, 0, 0};
    for (int k0 = 0; k0 < NB; k0 += 4) {
        const int kk = k0 + (lane >> 4);
        T av[2], bv[2];
        for (int i = 0; i < 2; ++i) av[i] = Pn[(wr * 32 + i * 16 + (lane & 15)) * TP + kk];
        for (int j = 0; j < 2; ++j) bv[j] = Iv[(wc * 32 + j * 16 + (lane & 15)) * TP + kk];
        for (int i = 0; i < 2; ++i)
            for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(av[i], bv[j], acc[i][j]);
    }
    __syncthreads();
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
            for (int v = 0; v < 4; ++v)
                Pn[(wr * 32 + i * 16 + mfma16_row<T>(lane, v)) * TP + wc * 32 + j * 16 + (lane & 15)] = acc[i][j][v];
    __syncthreads();
    store_tile<T, false>(a, t, rn, w, Pn);
}

// Trailing update over a range of panels:
//   A_IK -= sum_{J in [Jlo, Jhi)} A_IJ A_KJ^T   for K in [Kbeg, Kend), I in [K, nblk).
// One WG per (I, K) tile.  The C tile lives in MFMA accumulators (loaded from and
// stored to the packed array directly); panels stream through LDS with a register
// prefetch of the next panel.  In the column-major region (rows < n/2) the WG
// computes C^T so that consecutive lanes touch consecutive (contiguous) rows.
template <typename T>
__global__ __launch_bounds__(256) void syrk_range(T *a, Part P, int Jlo, int Jhi, int Kbeg,
                                                  int Kend, const int *info) {
    __shared__ T Ai[NB * TP];
    __shared__ T Ak[NB * TP];
    if (*info) return;
    // decode t -> (K, I): K = Kbeg + q has (nblk - K) tiles I = K..nblk-1
    const long t = blockIdx.x;
    const long r = P.nblk - Kbeg;
    long q = (long)((2.0 * r + 1.0 - sqrt((2.0 * r + 1.0) * (2.0 * r + 1.0) - 8.0 * (double)t)) * 0.5);
    if (q < 0) q = 0;
    while (q > 0 && q * r - q * (q - 1) / 2 > t) --q;
    while ((q + 1) * r - (q + 1) * q / 2 <= t) ++q;
    const int K = Kbeg + (int)q;
    if (K >= Kend) return;
    const int I = K + (int)(t - (q * r - q * (q - 1) / 2));
    if (I >= P.nblk) return;
    const int ri = P.start(I), rni = P.width(I), rk = P.start(K), rnk = P.width(K);
    const bool diagt = (I == K);
    const Tile tc = tile_at(P, ri, rk);
    const bool trans = tc.rs == 1;  // column-major region: compute C^T
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    typedef typename Vec4<T>::type V4;
    V4 acc[2][2];
    // load C (or C^T) into the accumulators
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int x = wr * 32 + i * 16 + mfma16_row<T>(lane, v);
                const int y = wc * 32 + j * 16 + (lane & 15);
                const int row = trans ? y : x, col = trans ? x : y;
                const T c = a[tc.base + (size_t)min(row, rni - 1) * tc.rs + (size_t)min(col, rnk - 1) * tc.cs];
                acc[i][j][v] = ((row < rni) & (col < rnk) & (!diagt | (row >= col))) ? c : T(0);
            }
    // panel tiles: thread owns 16 elements of each (coalesced along the contiguous axis)
    T pi[16], pk[16];
    auto gload = [&](int J) {
        const int c0 = P.start(J), w = P.width(J);
        const Tile ti = tile_at(P, ri, c0), tk = tile_at(P, rk, c0);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int idx = threadIdx.x + 256 * e;
            int rr, cc;
            if (ti.rs == 1) { rr = idx % NB; cc = idx / NB; } else { cc = idx % NB; rr = idx / NB; }
            // raw (clamped) values: the padding is zeroed at the LDS store, so the loads
            // are not waited for until then
            pi[e] = a[ti.base + (size_t)min(rr, rni - 1) * ti.rs + (size_t)min(cc, w - 1) * ti.cs];
            if (tk.rs == 1) { rr = idx % NB; cc = idx / NB; } else { cc = idx % NB; rr = idx / NB; }
            pk[e] = a[tk.base + (size_t)min(rr, rnk - 1) * tk.rs + (size_t)min(cc, w - 1) * tk.cs];
        }
    };
    auto sstore = [&](int J) {
        const int c0 = P.start(J), w = P.width(J);
        const Tile ti = tile_at(P, ri, c0), tk = tile_at(P, rk, c0);
#pragma unroll
        for (int e = 0; e < 16; ++e) {
            const int idx = threadIdx.x + 256 * e;
            int rr, cc;
            if (ti.rs == 1) { rr = idx % NB; cc = idx / NB; } else { cc = idx % NB; rr = idx / NB; }
            Ai[rr * TP + cc] = ((rr < rni) & (cc < w)) ? pi[e] : T(0);
            if (tk.rs == 1) { rr = idx % NB; cc = idx / NB; } else { cc = idx % NB; rr = idx / NB; }
            Ak[rr * TP + cc] = ((rr < rnk) & (cc < w)) ? pk[e] : T(0);
        }
    };
    gload(Jlo);
    for (int J = Jlo; J < Jhi; ++J) {
        __syncthreads();
        sstore(J);
        __syncthreads();
        if (J + 1 < Jhi) gload(J + 1);
        const T *X = trans ? Ak : Ai;  // rows of the computed tile
        const T *Y = trans ? Ai : Ak;  // columns
#pragma unroll 4
        for (int k0 = 0; k0 < NB; k0 += 4) {
            const int kk = k0 + (lane >> 4);
            T av[2], bv[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) av[i] = X[(wr * 32 + i * 16 + (lane & 15)) * TP + kk];
#pragma unroll
            for (int j = 0; j < 2; ++j) bv[j] = Y[(wc * 32 + j * 16 + (lane & 15)) * TP + kk];
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) acc[i][j] = mfma4(-av[i], bv[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int x = wr * 32 + i * 16 + mfma16_row<T>(lane, v);
                const int y = wc * 32 + j * 16 + (lane & 15);
                const int row = trans ? y : x, col = trans ? x : y;
                if (row < rni && col < rnk && (!diagt || row >= col))
                    a[tc.base + (size_t)row * tc.rs + (size_t)col * tc.cs] = acc[i][j][v];
            }
}

// Trailing update of everything right of an outer panel on 128 x 128 super-tiles:
//   A_IK -= sum_{J in [Jlo, Jhi)} A_IJ A_KJ^T   for Kbeg <= K <= I < nblk.
// Blocks are paired inside each RFP region (pairs never straddle n/2; a region's odd
// last block pairs with nothing), and a work-group owns (I-pair x K-pair): wave
// (wi, wj) computes block (I_wi, K_wj) as 4 x 4 MFMA fragments, so each staged panel
// row block feeds two waves (half the panel traffic of one-block tiles).  Panels
// stream through LDS in 16-column chunks with two buffers (one barrier per chunk).
// The region of each pair fixes its load pattern at compile time (coalesced along
// the contiguous axis); every global load is unconditional (rows clamped into the
// block, columns past a panel's width read a valid column and are zeroed), so the
// wait counts are exact.  In the column-major region the waves compute C^T so that
// consecutive lanes touch contiguous rows.
// Chunk depth of the trailing update.  8 keeps two work-groups at 36 KB of LDS each, so a
// CU running two of them still has room for a panel-factorization kernel (the critical
// path of the look-ahead) instead of holding it back until one of them retires.
constexpr int SK = 8;
constexpr int SP = SK + 1;   // LDS pitch

struct PairList {  // pairs of blocks from Kbeg: lower region first, then upper
    int kbeg, npl, u0, np;
    __device__ int first(int p) const { return p < npl ? kbeg + 2 * p : u0 + 2 * (p - npl); }
};
inline PairList make_pairs(const Part &P, int Kbeg) {
    PairList L;
    L.kbeg = Kbeg;
    L.npl = Kbeg < P.nlo ? (P.nlo - Kbeg + 1) / 2 : 0;
    L.u0 = Kbeg > P.nlo ? Kbeg : P.nlo;
    L.np = L.npl + (P.nblk - L.u0 + 1) / 2;
    return L;
}

template <typename T, bool CMI, bool CMK>
__device__ __forceinline__ void syrk2_body(T *__restrict__ a, const Part &P, int Jlo, int Jhi,
                                           const int (&bI)[2], const int (&bK)[2], T *Lb) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wi = wave >> 1, wj = wave & 1;
    const int I = bI[wi], K = bK[wj];
    const bool active = I >= 0 && K >= 0 && I >= K;
    const int rni = active ? P.width(I) : 0, rnk = active ? P.width(K) : 0;
    const bool diagt = I == K;
    constexpr bool trans = CMI;  // C rows are I rows: column-major region -> compute C^T
    const size_t s = (P.n & 1) ? P.n : P.n + 1;
    typedef typename Vec4<T>::type V4;
    V4 acc[4][4];
    if (active) {  // unconditional loads of clamped elements, then select
        const Tile tc = tile_at(P, P.start(I), P.start(K));
        const T *cb = a + tc.base;
        const int trs = (int)tc.rs, tcs = (int)tc.cs;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int x = i * 16 + mfma16_row<T>(lane, v), y = j * 16 + (lane & 15);
                    const int row = trans ? y : x, col = trans ? x : y;
                    const T c = cb[min(row, rni - 1) * trs + min(col, rnk - 1) * tcs];
                    acc[i][j][v] = ((row < rni) & (col < rnk) & (!diagt | (row >= col))) ? c : T(0);
                }
    } else {
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = V4{0, 0, 0, 0};
    }
    // staging: slot b = 0,1 (I blocks), 2,3 (K blocks), 4 elements per thread each.
    // Column-major region: element (r, c) at k + 1 + r0 + r + c * s; row-major region:
    // (r0 - k + r) * s + c (c a global column).  Thread -> (r, c): column-major
    // r = tid % 64, c = tid / 64 + 4e; row-major r = tid / 16 + 16e, c = tid % 16.
    const T *src[4];
    int rmax[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        const int blk = b < 2 ? bI[b] : bK[b - 2];
        const int bb = blk >= 0 ? blk : (b < 2 ? bI[0] : bK[0]);
        const bool cm = b < 2 ? CMI : CMK;
        const int r0 = P.start(bb);
        rmax[b] = P.width(bb) - 1;
        src[b] = cm ? a + (size_t)(P.k + 1 + r0) : a + (size_t)(r0 - P.k) * s;
    }
    // staged elements per thread and block: column-major blocks take row tid % 64 and
    // columns tid / 64 + 4e, row-major blocks column tid % SK and rows tid / SK + RS e
    constexpr int SE = NB * SK / 256, RS = 256 / SK;
    const int rc = tid & 63, cc_c = tid >> 6, rr_r = tid / SK, cc_r = tid % SK;
    T st[2][4][SE];  // two chunks in flight: chunk c is staged in slot c & 1
    // one staged block b of chunk kc of panel column block J into slot sl: the loads are unconditional
    // (clamped addresses); columns past the block's width are zeroed when the values are
    // stored to LDS, so nothing waits on a load until then
#define KF_SYRK_LOAD(sl, b, J, kc)                                                                \
    do {                                                                                          \
        const int c0_ = P.start(J), wl_ = P.width(J) - 1;                                         \
        constexpr bool cm_ = (b) < 2 ? CMI : CMK;                                                 \
        _Pragma("unroll") for (int e = 0; e < SE; ++e) {                                          \
            const int r_ = cm_ ? rc : rr_r + RS * e;                                              \
            const int c_ = (kc) * SK + (cm_ ? cc_c + 4 * e : cc_r);                               \
            const int rcl = min(r_, rmax[b]), ccl = c0_ + min(c_, wl_);                           \
            st[sl][b][e] = cm_ ? src[b][rcl + (size_t)ccl * s] : src[b][(size_t)rcl * s + ccl];   \
        }                                                                                         \
    } while (0)
#define KF_SYRK_STORE(buf, sl, b, J, kc)                                                          \
    do {                                                                                          \
        const int wl_ = P.width(J) - 1;                                                           \
        constexpr bool cm_ = (b) < 2 ? CMI : CMK;                                                 \
        _Pragma("unroll") for (int e = 0; e < SE; ++e) {                                          \
            const int r_ = cm_ ? rc : rr_r + RS * e;                                              \
            const int c_ = cm_ ? cc_c + 4 * e : cc_r;                                             \
            (buf)[((b) * NB + r_) * SP + c_] = (kc) * SK + c_ <= wl_ ? st[sl][b][e] : T(0);      \
        }                                                                                         \
    } while (0)
#define KF_SYRK_ALL(X) X(0); X(1); X(2); X(3)
    const int nch = (Jhi - Jlo) * (NB / SK);
    // prologue: chunk 0 to LDS, chunks 1 and 2 in flight (clamped to the last chunk, which
    // is then staged again into a buffer nobody reads: every load and store is unconditional)
    const int p1 = min(1, nch - 1), p2 = min(2, nch - 1);
#define KF_L0(b) KF_SYRK_LOAD(0, b, Jlo, 0)
#define KF_S0(b) KF_SYRK_STORE(Lb, 0, b, Jlo, 0)
#define KF_L1(b) KF_SYRK_LOAD(1, b, Jlo + p1 / (NB / SK), p1 % (NB / SK))
#define KF_L2(b) KF_SYRK_LOAD(0, b, Jlo + p2 / (NB / SK), p2 % (NB / SK))
    KF_SYRK_ALL(KF_L0);
    KF_SYRK_ALL(KF_S0);
    KF_SYRK_ALL(KF_L1);
    KF_SYRK_ALL(KF_L2);
#undef KF_L0
#undef KF_S0
#undef KF_L1
#undef KF_L2
    __syncthreads();
    // rows of the computed tile (X) and its columns (Y) inside a buffer
    const int xo = trans ? (2 + wj) * NB : wi * NB, yo = trans ? wi * NB : (2 + wj) * NB;
    // one chunk ch whose successor sits in register slot SL = (ch + 1) & 1 (a constant: the
    // loop below takes two chunks per trip, and nch is a multiple of NB / SK = 8)
    // the k-steps keep their branch on `active`: without it the compiler overlaps the next
    // k-step's LDS reads with the MFMAs at 246 VGPRs, and the diagonal-tile kernel of the
    // panel chain (296) then no longer fits beside a trailing-update wave on a SIMD
#define KF_SYRK_MMA(cur, ks)                                                                      \
        do {                                                                                      \
            if (active) {                                                                         \
                const int kk = 4 * (ks) + (lane >> 4);                                            \
                T av[4], bv[4];                                                                   \
                _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                    av[i] = (cur)[(xo + i * 16 + (lane & 15)) * SP + kk];                         \
                _Pragma("unroll") for (int j = 0; j < 4; ++j)                                     \
                    bv[j] = (cur)[(yo + j * 16 + (lane & 15)) * SP + kk];                         \
                _Pragma("unroll") for (int i = 0; i < 4; ++i)                                     \
                    _Pragma("unroll") for (int j = 0; j < 4; ++j)                                 \
                        acc[i][j] = mfma4(-av[i], bv[j], acc[i][j]);                              \
            }                                                                                     \
        } while (0)
#define KF_SYRK_STAGE(nxt, SL, b, J1, kc1, J3, kc3)                                               \
        do {                                                                                      \
            KF_SYRK_STORE(nxt, SL, b, J1, kc1);                                                   \
            KF_SYRK_LOAD(SL, b, J3, kc3);                                                         \
        } while (0)
    // the last chunks re-stage the last one into the idle buffer (never read): every load
    // and store is unconditional, so the wait counts stay exact across iterations.  Each
    // k-step is followed by staged blocks of the next chunk moved to LDS and the same
    // blocks of the chunk three ahead fetched into the freed slot: the staging spreads
    // over the MFMAs and each load has two chunks of MFMAs to arrive.
#define KF_SYRK_ITER(ch, SL)                                                                      \
    do {                                                                                          \
        const T *cur_ = Lb + ((ch) & 1) * (4 * NB * SP);                                          \
        T *nxt_ = Lb + (((ch) + 1) & 1) * (4 * NB * SP);                                          \
        const int c1_ = min((ch) + 1, nch - 1), c3_ = min((ch) + 3, nch - 1);                     \
        const int J1_ = Jlo + c1_ / (NB / SK), kc1_ = c1_ % (NB / SK);                            \
        const int J3_ = Jlo + c3_ / (NB / SK), kc3_ = c3_ % (NB / SK);                            \
        KF_SYRK_MMA(cur_, 0);                                                                     \
        KF_SYRK_STAGE(nxt_, SL, 0, J1_, kc1_, J3_, kc3_);                                         \
        KF_SYRK_STAGE(nxt_, SL, 1, J1_, kc1_, J3_, kc3_);                                         \
        KF_SYRK_MMA(cur_, 1);                                                                     \
        KF_SYRK_STAGE(nxt_, SL, 2, J1_, kc1_, J3_, kc3_);                                         \
        KF_SYRK_STAGE(nxt_, SL, 3, J1_, kc1_, J3_, kc3_);                                         \
        __syncthreads();                                                                          \
    } while (0)
    for (int ch = 0; ch < nch; ch += 2) {
        KF_SYRK_ITER(ch, 1);
        KF_SYRK_ITER(ch + 1, 0);
    }
#undef KF_SYRK_ITER
#undef KF_SYRK_MMA
#undef KF_SYRK_STAGE
#undef KF_SYRK_ALL
#undef KF_SYRK_LOAD
#undef KF_SYRK_STORE
    if (!active) return;
    {
        // recompute the addresses here rather than keep 64 of them live through the loop
        int ln = lane;
        asm volatile("" : "+v"(ln));
        const Tile tc = tile_at(P, P.start(I), P.start(K));
        T *cb = a + tc.base;
        const int trs = (int)tc.rs, tcs = (int)tc.cs;
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int x = i * 16 + mfma16_row<T>(ln, v), y = j * 16 + (ln & 15);
                    const int row = trans ? y : x, col = trans ? x : y;
                    if (row < rni && col < rnk && (!diagt || row >= col)) cb[row * trs + col * tcs] = acc[i][j][v];
                }
    }
}

template <typename T>
__global__ __launch_bounds__(256, 2) void syrk2_range(T *a, Part P, PairList L, int Jlo, int Jhi,
                                                      const int *info) {
    __shared__ T Lb[2 * 4 * NB * SP];  // per buffer: two I blocks then two K blocks
    if (*info) return;
    // decode t -> (q, SI): pair column q has np - q super-tiles SI = q..np-1
    const long t = blockIdx.x;
    const long r = L.np;
    long q = (long)((2.0 * r + 1.0 - sqrt((2.0 * r + 1.0) * (2.0 * r + 1.0) - 8.0 * (double)t)) * 0.5);
    if (q < 0) q = 0;
    while (q > 0 && q * r - q * (q - 1) / 2 > t) --q;
    while ((q + 1) * r - (q + 1) * q / 2 <= t) ++q;
    const int SI = (int)q + (int)(t - (q * r - q * (q - 1) / 2));
    if (q >= r || SI >= r) return;
    int bI[2], bK[2];
    bI[0] = L.first(SI);
    bI[1] = (bI[0] + 1 < P.nblk && bI[0] + 1 != P.nlo) ? bI[0] + 1 : -1;  // pairs stay in a region
    bK[0] = L.first((int)q);
    bK[1] = (bK[0] + 1 < P.nblk && bK[0] + 1 != P.nlo) ? bK[0] + 1 : -1;
    const bool cmI = bI[0] < P.nlo, cmK = bK[0] < P.nlo;
    if (cmI) {
        if (cmK) syrk2_body<T, true, true>(a, P, Jlo, Jhi, bI, bK, Lb);
        else syrk2_body<T, true, false>(a, P, Jlo, Jhi, bI, bK, Lb);
    } else {
        if (cmK) syrk2_body<T, false, true>(a, P, Jlo, Jhi, bI, bK, Lb);
        else syrk2_body<T, false, false>(a, P, Jlo, Jhi, bI, bK, Lb);
    }
}

inline long range_tiles(const Part &P, int Kbeg, int Kend) {
    long n = 0;
    for (int K = Kbeg; K < Kend; ++K) n += P.nblk - K;
    return n;
}

// ---------------------------------------------------------------------------
// Triangular solves, FP64 accumulation.  b is (nrhs, n) row-major.
//
// The 64-wide blocks are taken in groups of kGrp consecutive blocks.  Per solve, the
// inverse X_g of each group's diagonal block of L (lower triangular, kGrp x kGrp tiles of
// 64 x 64) is formed once: the diagonal tiles are the inverses of the 64 x 64 diagonal
// blocks, and column j of tiles follows by block forward substitution,
//   X(i, j) = -X(i, i) * sum_{k = j .. i-1} L(i, k) X(k, j).
// A forward step over group g is then two launches: one forms y_g = X_g b_g (a work-group
// per block of the group), the next applies it to every block row below the group, one
// work-group each, b_I -= sum_j L(I, j) y_j; the backward steps mirror this with X_g^T and
// the block columns left of the group.  The chain of dependent launches is 2 / kGrp of one
// launch per 64-wide block, and X_g is read once per step.
//
// X storage: per group kGrpTiles row-major 64 x 64 FP64 tiles, tile (i, j) (j <= i) at
// index i (i + 1) / 2 + j; entries outside the blocks' widths are zero.
// ---------------------------------------------------------------------------
constexpr int kGrp = 4;
constexpr int kGrpTiles = kGrp * (kGrp + 1) / 2;
constexpr int kTileElems = NB * NB;
static_assert(kGrp * NB == 256, "a step maps one thread per (group block, row)");

inline int n_groups(const Part &P) { return (P.nblk + kGrp - 1) / kGrp; }
__host__ __device__ inline int gtile(int i, int j) { return i * (i + 1) / 2 + j; }
__device__ inline double *xtile(double *X, int g, int i, int j) {
    return X + ((size_t)g * kGrpTiles + gtile(i, j)) * kTileElems;
}
__device__ inline const double *xtile(const double *X, int g, int i, int j) {
    return X + ((size_t)g * kGrpTiles + gtile(i, j)) * kTileElems;
}

// Diagonal tiles of every group's X: the inverses of the 64 x 64 diagonal blocks.
template <typename T>
__global__ __launch_bounds__(256) void inv_diag_all(const T *a, Part P, double *X) {
    __shared__ double L[NB * TP];
    __shared__ double Li[NB * TP];
    const int J = blockIdx.x;
    const int s0 = P.start(J), w = P.width(J);
    const Tile t = tile_at(P, s0, s0);
#pragma unroll
    for (int e = 0; e < kTileElems / 256; ++e) {  // clamped loads, masked at the LDS store
        const int idx = threadIdx.x + 256 * e;
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        const T v = a[t.base + (size_t)min(r, w - 1) * t.rs + (size_t)min(c, w - 1) * t.cs];
        L[r * TP + c] = ((r < w) & (c < w) & (r >= c)) ? (double)v : 0.0;
        Li[r * TP + c] = 0.0;
    }
    __shared__ double dgi[NB];
    __syncthreads();
    tri_inverse<double, double>(L, Li, dgi, w);
    double *dst = xtile(X, J / kGrp, J % kGrp, J % kGrp);
    for (int idx = threadIdx.x; idx < kTileElems; idx += blockDim.x) {
        const int r = idx / NB, c = idx % NB;
        dst[idx] = (r < w && c < w) ? Li[r * TP + c] : 0.0;
    }
}

// 64 x 64 FP64 product of two LDS tiles (pitch TP), 4 x 4 outputs per thread:
// acc[u][v] += sum_k A[(4 ty + u) * TP + k] * B[k * TP + 4 tx + v].
__device__ __forceinline__ void tile_mm(const double *A, const double *B, double (&acc)[4][4]) {
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
#pragma unroll 4
    for (int k = 0; k < NB; ++k) {
        double av[4], bv[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) av[u] = A[(4 * ty + u) * TP + k];
#pragma unroll
        for (int v = 0; v < 4; ++v) bv[v] = B[k * TP + 4 * tx + v];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) acc[u][v] += av[u] * bv[v];
    }
}

// A row-major 64 x 64 FP64 tile from global memory into LDS (pitch TP).
__device__ __forceinline__ void xtile_to_lds(const double *src, double *dst) {
#pragma unroll
    for (int e = 0; e < kTileElems / 256; ++e) {
        const int idx = threadIdx.x + 256 * e;
        dst[(idx / NB) * TP + idx % NB] = src[idx];
    }
}

// Off-diagonal tiles of X: one work-group per (group, tile column j < kGrp - 1), walking
// the rows i = j+1 .. m-1 of its column in order (each needs the tiles above it).
template <typename T>
__global__ __launch_bounds__(256) void inv_group_offdiag(const T *a, Part P, double *X) {
    __shared__ double As[NB * TP];
    __shared__ double Bs[NB * TP];
    const int g = blockIdx.x / (kGrp - 1), j = blockIdx.x % (kGrp - 1);
    const int J0 = g * kGrp, m = min(kGrp, P.nblk - J0);
    if (j + 1 >= m) return;
    const int ty = threadIdx.x >> 4, tx = threadIdx.x & 15;
    for (int i = j + 1; i < m; ++i) {
        double acc[4][4] = {};
        const int ri = P.start(J0 + i), wi = P.width(J0 + i);
        for (int k = j; k < i; ++k) {
            __syncthreads();  // the previous product is done with As and Bs
            const Tile t = tile_at(P, ri, P.start(J0 + k));
            const int wk = P.width(J0 + k);
#pragma unroll
            for (int e = 0; e < kTileElems / 256; ++e) {  // L(i, k), zero outside the blocks
                const int idx = threadIdx.x + 256 * e;
                int r, c;
                if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
                const T v = a[t.base + (size_t)min(r, wi - 1) * t.rs + (size_t)min(c, wk - 1) * t.cs];
                As[r * TP + c] = ((r < wi) & (c < wk)) ? (double)v : 0.0;
            }
            // X(k, j): written by inv_diag_all (k == j) or by this work-group's earlier rows,
            // whose stores the barriers order before these loads
            xtile_to_lds(xtile(X, g, k, j), Bs);
            __syncthreads();
            tile_mm(As, Bs, acc);
        }
        __syncthreads();
        xtile_to_lds(xtile(X, g, i, i), As);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) Bs[(4 * ty + u) * TP + 4 * tx + v] = acc[u][v];
        __syncthreads();
        double out[4][4] = {};
        tile_mm(As, Bs, out);
        double *dst = xtile(X, g, i, j);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) dst[(4 * ty + u) * NB + 4 * tx + v] = -out[u][v];
        __syncthreads();  // the stores of X(i, j) before the next row reads them
    }
}

// Update tiles of a step in registers: thread t holds elements t + 256 e of each tile,
// coalesced along its contiguous axis: in the column-major region (rs == 1) row t % 64 of
// columns t / 64 + 4 e; in the row-major region column t % 64 of rows t / 64 + 4 e.  The
// loads are of clamped (valid) elements and unconditional; the padding is applied where
// the values are used (a select right after a load would wait for it there).
constexpr int SE_T = kTileElems / 256;
template <typename T>
__device__ __forceinline__ void step_tile_load(const T *a, const Tile &t, int rn, int cn,
                                               T (&v)[SE_T]) {
#pragma unroll
    for (int e = 0; e < SE_T; ++e) {
        const int idx = threadIdx.x + 256 * e;
        int r, c;
        if (t.rs == 1) { r = idx % NB; c = idx / NB; } else { c = idx % NB; r = idx / NB; }
        v[e] = a[t.base + (size_t)min(r, rn - 1) * t.rs + (size_t)min(c, cn - 1) * t.cs];
    }
}

// Sum over the 64 lanes of a wave of 16 values per lane: on return, the lanes with
// (lane & 3) == 0 hold the total of value index e = lane >> 2 (the others partial sums).
// Four halving exchanges (xor 32, 16, 8, 4) each keep half of the remaining indices, then
// two plain exchanges finish the sums: 17 exchanges instead of 16 x 6.
__device__ __forceinline__ double wave_sum16(double (&v)[16]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 8, bit = 32; h >= 1; h >>= 1, bit >>= 1) {
        const bool up = (lane & bit) != 0;  // keep indices [h, 2h) of the current half
#pragma unroll
        for (int e = 0; e < h; ++e) {
            const double keep = up ? v[e + h] : v[e], send = up ? v[e] : v[e + h];
            v[e] = keep + __shfl_xor(send, bit);
        }
    }
    double s = v[0];
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    return s;
}

// Forward step over group g, right-hand side blockIdx.y, as two launches.
// fwd_group_y: work-group i forms block i of y_g = X_g b_g, y_i = sum_{j <= i} X(i, j) b_j
// (4 threads per row), and stores it to y.
__global__ __launch_bounds__(256) void fwd_group_y(Part P, int g, const double *X,
                                                   const double *b, double *y, int n) {
    __shared__ double bg[kGrp][NB];
    const int J0 = g * kGrp, i = blockIdx.x, tid = threadIdx.x;
    const double *bq = b + (size_t)blockIdx.y * n;
    {
        const int j = tid >> 6, c = tid & 63;
        const int jc = min(j, i), wj = P.width(J0 + jc);  // blocks j <= i < m exist
        const double v = bq[P.start(J0 + jc) + min(c, wj - 1)];
        bg[j][c] = ((j <= i) & (c < wj)) ? v : 0.0;
    }
    __syncthreads();
    const int r = tid >> 2, qq = tid & 3;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < kGrp; ++j) {  // blocks j > i of bg are zero: tile (i, i) stands in
        const double *xt = xtile(X, g, i, min(j, i)) + r * NB + qq;
#pragma unroll
        for (int k = 0; k < NB / 4; ++k) s += xt[4 * k] * bg[j][qq + 4 * k];
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (qq == 0 && r < P.width(J0 + i)) y[(size_t)blockIdx.y * n + P.start(J0 + i) + r] = s;
}

// fwd_group_upd: b_I -= sum_j L(I, J0 + j) y_j for the block row I = J0 + m + blockIdx.x
// below the group, from the update tiles in registers (loaded while y_g is read).
template <typename T>
__global__ __launch_bounds__(256) void fwd_group_upd(const T *a, Part P, int g, const double *y,
                                                     double *b, int n) {
    __shared__ double yg[kGrp][NB];
    __shared__ double red[4][NB];
    const int J0 = g * kGrp, m = min(kGrp, P.nblk - J0);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int I = J0 + m + (int)blockIdx.x;
    const int r0 = P.start(I), rn = P.width(I);
    const bool cm = tile_at(P, r0, P.start(J0)).rs == 1;  // block row I lies in one region
    T lt[kGrp][SE_T];
#pragma unroll
    for (int j = 0; j < kGrp; ++j) {  // clamped to the group's last block: valid tiles
        const int jc = min(j, m - 1);
        step_tile_load(a, tile_at(P, r0, P.start(J0 + jc)), rn, P.width(J0 + jc), lt[j]);
    }
    double *bq = b + (size_t)blockIdx.y * n;
    const double bI = bq[r0 + min(lane, rn - 1)];
    {
        const int j = tid >> 6, c = tid & 63;
        const int jc = min(j, m - 1), wj = P.width(J0 + jc);
        const double v = y[(size_t)blockIdx.y * n + P.start(J0 + jc) + min(c, wj - 1)];
        yg[j][c] = ((j < m) & (c < wj)) ? v : 0.0;  // zero past the group and the widths
    }
    __syncthreads();
    if (cm) {  // thread: row lane, columns wv + 4e -> partial row sums, then across waves
        double s = 0.0;
#pragma unroll
        for (int j = 0; j < kGrp; ++j) {
            const int wj = P.width(J0 + min(j, m - 1));
#pragma unroll
            for (int e = 0; e < SE_T; ++e) {
                const int c = wv + 4 * e;
                s += (c < wj ? (double)lt[j][e] : 0.0) * yg[j][c];
            }
        }
        red[wv][lane] = s;
    } else {  // thread: column lane, rows wv + 4e -> per-row partials, summed over lanes
        double v[16];
#pragma unroll
        for (int e = 0; e < 16; ++e) v[e] = 0.0;
#pragma unroll
        for (int j = 0; j < kGrp; ++j) {
            const double yv = yg[j][lane];  // zero past the block width
#pragma unroll
            for (int e = 0; e < SE_T; ++e) v[e] += (double)lt[j][e] * yv;
        }
        const double tot = wave_sum16(v);  // row wv + 4 (lane >> 2) on lanes 4k
        if ((lane & 3) == 0) red[0][wv + 4 * (lane >> 2)] = tot;
    }
    __syncthreads();
    if (tid < NB && tid < rn) {
        const double tot = cm ? red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid] : red[0][tid];
        bq[r0 + tid] = bI - tot;
    }
}

// Backward step over group g, right-hand side blockIdx.y, as two launches.
// bwd_group_x: work-group j forms block j of x_g = X_g^T y_g,
// x_j = sum_{i >= j} X(i, j)^T y_i (4 threads per column), and stores it to x.
__global__ __launch_bounds__(256) void bwd_group_x(Part P, int g, const double *X,
                                                   const double *y, double *x, int n) {
    __shared__ double yg[kGrp][NB];
    const int J0 = g * kGrp, m = min(kGrp, P.nblk - J0), j = blockIdx.x, tid = threadIdx.x;
    const double *yq = y + (size_t)blockIdx.y * n;
    {
        const int i = tid >> 6, r = tid & 63;
        const int ic = min(max(i, j), m - 1), wi = P.width(J0 + ic);  // blocks j <= i < m
        const double v = yq[P.start(J0 + ic) + min(r, wi - 1)];
        yg[i][r] = ((i >= j) & (i < m) & (r < wi)) ? v : 0.0;
    }
    __syncthreads();
    const int c = tid >> 2, qq = tid & 3;
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < kGrp; ++i) {  // blocks i < j or i >= m of yg are zero: a valid tile
        const double *xt = xtile(X, g, min(max(i, j), m - 1), j) + qq * NB + c;  // stands in
#pragma unroll
        for (int k = 0; k < NB / 4; ++k) s += xt[4 * k * NB] * yg[i][qq + 4 * k];
    }
    s += __shfl_xor(s, 1);
    s += __shfl_xor(s, 2);
    if (qq == 0 && c < P.width(J0 + j)) x[(size_t)blockIdx.y * n + P.start(J0 + j) + c] = s;
}

// bwd_group_upd: y_I -= sum_j L(J0 + j, I)^T x_j for the block column I = blockIdx.x left of
// the group, from the update tiles in registers (loaded while x_g is read).
template <typename T>
__global__ __launch_bounds__(256) void bwd_group_upd(const T *a, Part P, int g, const double *x,
                                                     double *y, int n) {
    __shared__ double xg[kGrp][NB];
    __shared__ double red[4][NB];
    __shared__ double colA[NB];
    const int J0 = g * kGrp, m = min(kGrp, P.nblk - J0);
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int I = blockIdx.x;
    const int c0 = P.start(I), cn = P.width(I);
    T lt[kGrp][SE_T];
    bool cmj[kGrp];  // tile L(J0 + j, I) in the column-major region
#pragma unroll
    for (int j = 0; j < kGrp; ++j) {
        const int jc = min(j, m - 1);
        const Tile t = tile_at(P, P.start(J0 + jc), c0);
        cmj[j] = t.rs == 1;
        step_tile_load(a, t, P.width(J0 + jc), cn, lt[j]);
    }
    double *yq = y + (size_t)blockIdx.y * n;
    const double yI = yq[c0 + min(lane, cn - 1)];
    {
        const int j = tid >> 6, r = tid & 63;
        const int jc = min(j, m - 1), wj = P.width(J0 + jc);
        const double v = x[(size_t)blockIdx.y * n + P.start(J0 + jc) + min(r, wj - 1)];
        xg[j][r] = ((j < m) & (r < wj)) ? v : 0.0;  // zero past the group and the widths
    }
    __syncthreads();
    double pB = 0.0, v[16];
#pragma unroll
    for (int e = 0; e < 16; ++e) v[e] = 0.0;
#pragma unroll
    for (int j = 0; j < kGrp; ++j) {
        const int wj = P.width(J0 + min(j, m - 1));
        if (cmj[j]) {  // thread: row lane, columns wv + 4e -> per-column partials
            const double xv = xg[j][lane];  // zero past the block width
#pragma unroll
            for (int e = 0; e < SE_T; ++e) v[e] += (double)lt[j][e] * xv;
        } else {  // thread: column lane, rows wv + 4e -> partial of column lane
#pragma unroll
            for (int e = 0; e < SE_T; ++e) {
                const int r = wv + 4 * e;
                pB += (r < wj ? (double)lt[j][e] : 0.0) * xg[j][r];
            }
        }
    }
    const double totA = wave_sum16(v);  // column wv + 4 (lane >> 2) on lanes 4k
    if ((lane & 3) == 0) colA[wv + 4 * (lane >> 2)] = totA;
    red[wv][lane] = pB;
    __syncthreads();
    if (tid < NB && tid < cn)
        yq[c0 + tid] = yI - (colA[tid] + red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
}

// Outer panels of 64-wide blocks: 8 (a 512-deep trailing update) while more than
// kWideFrom blocks remain, 4 (256-deep) after that.  A deeper update reads and writes the
// trailing matrix half as often; near the end the panel chain is the critical path and
// narrow panels keep it short.  Measured at orders 20,000 / 32,000 / 64,000: 75.4 /
// 237.8 / 1680 ms against 74.9 / 248.0 / 1828 ms with 4 throughout; 16-wide panels at the
// start are slower (246.8 ms at 32,000).
constexpr int kWideFrom = 200;

template <typename T>
int potrf_impl(T *a, int n, double l2, hipStream_t stream) {
    KF_REQUIRE(a && n > 0, KF_ERR_ARG, "n must be > 0");
    const Part P = make_part(n);
    char *ws = static_cast<char *>(workspace(WS_AUX2, sizeof(T) * NB * NB + 256, stream));
    KF_REQUIRE(ws, KF_ERR_NOMEM, "workspace allocation failed");
    int *info = reinterpret_cast<int *>(ws);
    T *inv = reinterpret_cast<T *>(ws + 256);
    KF_HIP_TRY(hipMemsetAsync(info, 0, sizeof(int), stream));
    ScopedTimer timer(TM_POTRF, stream);
    if (l2 != 0.0) {
        hipLaunchKernelGGL(add_diag<T>, dim3((n + 255) / 256), dim3(256), 0, stream, a, n, l2);
    }
    // Look-ahead over outer panels (widths from next()).  factor(p) = diagonal POTRF + inverse,
    // panel TRSM and the in-panel updates of panel p; it runs on the aux stream while
    // the caller's stream applies panel p-1 to everything right of panel p.  Per p:
    //   caller: wait factor(p); update panel p+1's columns from p (64-wide tiles); mark N
    //   aux:    wait N; factor(p+1); mark F
    //   caller: update the columns right of panel p+1 from p (128-wide super-tiles)
    // factor(p+1) touches only panel p+1's columns, disjoint from the concurrent update.
    hipStream_t s2 = aux_stream();
    hipEvent_t ev_start = aux_event(0), ev_f = aux_event(1), ev_n = aux_event(2);
    KF_REQUIRE(s2 && ev_start && ev_f && ev_n, KF_ERR_HIP, "aux stream/event creation failed");
    // Inside an outer panel the updates are left-looking: block column J receives the
    // contributions of blocks J0 .. J-1 in one deep update just before it is factored, so
    // the panel chain runs one update launch per block with the C tiles read and written
    // once, instead of a rank-64 update of every remaining panel column after each block.
    auto factor = [&](int J0, int J1) {
        for (int J = J0; J < J1; ++J) {
            if (J > J0)
                hipLaunchKernelGGL(syrk_range<T>, dim3((unsigned)range_tiles(P, J, J + 1)),
                                   dim3(256), 0, s2, a, P, J0, J, J, J + 1, info);
            hipLaunchKernelGGL(potrf_diag<T>, dim3(1), dim3(256), 0, s2, a, P, J, inv, info);
            const int m = P.nblk - J - 1;
            if (m > 0)
                hipLaunchKernelGGL(trsm_panel<T>, dim3(m), dim3(256), 0, s2, a, P, J, inv, info);
        }
    };
    auto next = [&](int J) {  // end of the outer panel that starts at block J
        const int w = P.nblk - J > kWideFrom ? 8 : 4;
        return J + w < P.nblk ? J + w : P.nblk;
    };
    KF_HIP_TRY(hipEventRecord(ev_start, stream));
    KF_HIP_TRY(hipStreamWaitEvent(s2, ev_start, 0));
    factor(0, next(0));
    KF_HIP_TRY(hipEventRecord(ev_f, s2));
    for (int J0 = 0; J0 < P.nblk; J0 = next(J0)) {
        const int J1 = next(J0);
        const int J2 = J1 < P.nblk ? next(J1) : P.nblk;
        KF_HIP_TRY(hipStreamWaitEvent(stream, ev_f, 0));
        if (J1 < P.nblk) {
            const long nt = range_tiles(P, J1, J2);
            hipLaunchKernelGGL(syrk_range<T>, dim3((unsigned)nt), dim3(256), 0, stream, a, P, J0,
                               J1, J1, J2, info);
            KF_HIP_TRY(hipEventRecord(ev_n, stream));
            KF_HIP_TRY(hipStreamWaitEvent(s2, ev_n, 0));
            factor(J1, J2);
            KF_HIP_TRY(hipEventRecord(ev_f, s2));
        }
        if (J2 < P.nblk) {  // everything right of the next outer panel, panel p deep
            const PairList L = make_pairs(P, J2);
            const long nt = (long)L.np * (L.np + 1) / 2;
            hipLaunchKernelGGL(syrk2_range<T>, dim3((unsigned)nt), dim3(256), 0, stream, a, P, L,
                               J0, J1, info);
        }
    }
    KF_HIP_TRY(hipGetLastError());
    timer.stop();
    int h = 0;
    KF_HIP_TRY(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    return h;
}

// Scratch of a solve: the group inverses X, then the intermediate y = L^-1 b.  Every byte
// a solve reads from it is first written by the same solve (inv_diag_all and
// inv_group_offdiag write the whole tiles a group uses, the forward steps every row of y),
// so it needs no clearing.
inline size_t potrs_x_bytes(const Part &P) {
    return sizeof(double) * (size_t)n_groups(P) * kGrpTiles * kTileElems;
}
inline size_t potrs_scratch_bytes(const Part &P, int n, int nrhs) {
    return potrs_x_bytes(P) + sizeof(double) * (size_t)n * nrhs;
}

template <typename T>
int potrs_run(const T *a, const Part &P, double *b, int nrhs, char *scratch, hipStream_t stream,
              bool timed) {
    const int n = P.n, ng = n_groups(P);
    double *X = reinterpret_cast<double *>(scratch);
    double *y = reinterpret_cast<double *>(scratch + potrs_x_bytes(P));
    // the kernel timer creates HIP events when enabled, so the allocation-free entry
    // point runs untimed
    ScopedTimer timer(timed ? TM_POTRS : -1, stream);
    hipLaunchKernelGGL(inv_diag_all<T>, dim3(P.nblk), dim3(256), 0, stream, a, P, X);
    hipLaunchKernelGGL(inv_group_offdiag<T>, dim3(ng * (kGrp - 1)), dim3(256), 0, stream, a, P, X);
    // right-hand sides along the grid's y dimension, at most 65,535 per launch
    for (int q0 = 0; q0 < nrhs; q0 += 65535) {
        const int nq = min(nrhs - q0, 65535);
        double *bq = b + (size_t)q0 * n, *yq = y + (size_t)q0 * n;
        for (int g = 0; g < ng; ++g) {
            const int m = min(kGrp, P.nblk - g * kGrp), below = P.nblk - g * kGrp - m;
            hipLaunchKernelGGL(fwd_group_y, dim3(m, nq), dim3(256), 0, stream, P, g, X, bq, yq, n);
            if (below > 0)
                hipLaunchKernelGGL(fwd_group_upd<T>, dim3(below, nq), dim3(256), 0, stream, a, P,
                                   g, yq, bq, n);
        }
        for (int g = ng - 1; g >= 0; --g) {
            const int m = min(kGrp, P.nblk - g * kGrp);
            hipLaunchKernelGGL(bwd_group_x, dim3(m, nq), dim3(256), 0, stream, P, g, X, yq, bq, n);
            if (g > 0)
                hipLaunchKernelGGL(bwd_group_upd<T>, dim3(g * kGrp, nq), dim3(256), 0, stream, a,
                                   P, g, bq, yq, n);
        }
    }
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

template <typename T>
int potrs_impl(const T *a, int n, double *b, int nrhs, hipStream_t stream) {
    KF_REQUIRE(a && b && n > 0 && nrhs >= 1, KF_ERR_ARG, "bad potrs arguments");
    const Part P = make_part(n);
    char *ws = static_cast<char *>(workspace(WS_AUX3, potrs_scratch_bytes(P, n, nrhs) + 256, stream));
    KF_REQUIRE(ws, KF_ERR_NOMEM, "workspace allocation failed");
    return potrs_run(a, P, b, nrhs, ws, stream, true);
}

// ---------------------------------------------------------------------------
// Allocation-free solve: the same kernels with their scratch in a fixed device array.
// ---------------------------------------------------------------------------
constexpr size_t kNoAllocScratchBytes = size_t(48) << 20;  // 48 MiB per device
static __device__ __align__(256) char scratch[kNoAllocScratchBytes];

constexpr int kMaxDev = 16;
std::mutex g_scratch_mutex;
char *g_scratch_ptr[kMaxDev] = {};

// Device address of `scratch` on the current device, looked up once per device (the
// host-side name of a __device__ array is only a stand-in; kernels must be given this).
int scratch_address(char **out) {
    int dev = 0;
    KF_HIP_TRY(hipGetDevice(&dev));
    KF_REQUIRE(dev >= 0 && dev < kMaxDev, KF_ERR_ARG, "device %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        KF_HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        KF_REQUIRE(p, KF_ERR_HIP, "scratch symbol has no device address");
        g_scratch_ptr[dev] = static_cast<char *>(p);
    }
    *out = g_scratch_ptr[dev];
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_rfp_potrf(int precision, void *arf, int n, double l2, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    if (precision == KF_PREC_F64) return potrf_impl(static_cast<double *>(arf), n, l2, as_stream(stream));
    if (precision == KF_PREC_F32) return potrf_impl(static_cast<float *>(arf), n, l2, as_stream(stream));
    set_last_error("precision %d not supported by kf_rfp_potrf", precision);
    return KF_ERR_ARG;
}

extern "C" int kf_rfp_potrs_noalloc(int precision, const void *arf, int n, double *b, int nrhs,
                                    void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    KF_REQUIRE(precision == KF_PREC_F64 || precision == KF_PREC_F32, KF_ERR_ARG,
               "precision %d not supported by kf_rfp_potrs_noalloc", precision);
    KF_REQUIRE(arf && b && n > 0 && nrhs >= 1, KF_ERR_ARG, "bad potrs arguments");
    const Part P = make_part(n);
    const size_t need = potrs_scratch_bytes(P, n, nrhs);
    KF_REQUIRE(need <= kNoAllocScratchBytes, KF_ERR_ARG,
               "kf_rfp_potrs_noalloc: order %d with %d right-hand sides needs %zu bytes of "
               "scratch, more than its fixed %zu (use kf_rfp_potrs)",
               n, nrhs, need, kNoAllocScratchBytes);
    char *scr = nullptr;
    int rc = scratch_address(&scr);
    if (rc) return rc;
    hipStream_t st = as_stream(stream);
    if (precision == KF_PREC_F64)
        return potrs_run(static_cast<const double *>(arf), P, b, nrhs, scr, st, false);
    return potrs_run(static_cast<const float *>(arf), P, b, nrhs, scr, st, false);
}

extern "C" int kf_rfp_potrs(int precision, const void *arf, int n, double *b, int nrhs, void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    if (precision == KF_PREC_F64)
        return potrs_impl(static_cast<const double *>(arf), n, b, nrhs, as_stream(stream));
    if (precision == KF_PREC_F32)
        return potrs_impl(static_cast<const float *>(arf), n, b, nrhs, as_stream(stream));
    set_last_error("precision %d not supported by kf_rfp_potrs", precision);
    return KF_ERR_ARG;
}
