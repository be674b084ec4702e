// local_mixed.hip -- BF16x3 split-precision local Gaussian kernel on gfx950 (MFMA).
//
//   K[a,b] = sum_{i in a, j in b, Q_i == Q_j} exp(-(|x_i|^2 + |x_j|^2 - 2 x_i.x_j) / 2 sigma^2)
//
// Same mathematics as the reference's kernel_gaussian_symm_rfp / kernel_gaussian
// (src/local_kernels.cpp:155-298, 403-577); different arithmetic:
//  * each label's rows are centred on a per-label reference point (d^2 is
//    translation invariant; centring shrinks |x| and with it the absolute error of
//    the split product), split x = hi + lo with hi = bf16(x), lo = bf16(x - hi), and
//    x_i . x_j ~ hi.hi + hi.lo + lo.hi accumulated in FP32 on
//    v_mfma_f32_16x16x32_bf16; squared norms stay FP64; exp in FP32.
//  * one workgroup = one 128x128 tile of a label's packed Gram matrix (upper
//    triangle of tiles for the symmetric form, XCD-grouped so tiles sharing a row
//    block run on the same L2).  Operands stream through LDS with register
//    prefetch; 4 waves each own a 64x64 quadrant (4x4 MFMA blocks x 3 products).
//  * fused epilogue: weights (x2 for same-molecule pairs i<j), exp, then a
//    segmented reduction in LDS (row-wise runs over column molecules, in place,
//    then column-run sums over row molecules) and one FP32 atomic add per
//    (molecule, molecule) pair present in the tile into the RFP / dense output.
#include <cmath>

#include "local_mixed.hpp"

namespace kf {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int TM = 128;       // tile edge (packed rows)
constexpr int TK = 32;        // k per MFMA step
constexpr int LP = 40;        // LDS row pitch of operand tiles, in bf16 (80 B)
constexpr int EPITCH = TM + 1;
constexpr int NT = 256;
constexpr int CENTER_ROWS = 4096;  // rows averaged for a label's centring point

enum MMode { M_SYMM_RFP = 0, M_SYMM_DENSE = 1, M_RECT = 2 };

// ---------------------------------------------------------------------------
// centring point: mean of the first min(R_l, 4096) rows of each label (FP64)
// grid (L, ceil(rep/64)), 256 threads: thread (t%64 -> k, t/64 -> row phase)
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void label_center(const double *__restrict__ X,
                                                    const int *__restrict__ src,
                                                    const int *__restrict__ lab_row0,
                                                    const int *__restrict__ lab_rows, int rep,
                                                    double *__restrict__ center) {
    __shared__ double part[4][64];
    const int l = blockIdx.x, k = blockIdx.y * 64 + (threadIdx.x & 63), ph = threadIdx.x >> 6;
    const int r0 = lab_row0[l], nr = min(lab_rows[l], CENTER_ROWS);
    double s = 0.0;
    if (k < rep)
        for (int r = ph; r < nr; r += 4) s += X[(size_t)src[r0 + r] * rep + k];
    part[ph][threadIdx.x & 63] = s;
    __syncthreads();
    if (ph == 0 && k < rep) {
        const double t = part[0][threadIdx.x] + part[1][threadIdx.x] + part[2][threadIdx.x] +
                         part[3][threadIdx.x];
        center[(size_t)l * rep + k] = nr > 0 ? t / nr : 0.0;
    }
}

// pack: one wave per padded row.  rows of label l live at [pad0[l], pad0[l] + R_l),
// padding up to the next multiple of 128 is zero with rmol = -1.
__global__ __launch_bounds__(256) void pack_split_bf16(
    const double *__restrict__ X, const int *__restrict__ src, const int *__restrict__ srow_of,
    const int *__restrict__ lab_of, const double *__restrict__ center, int rows_pad, int rep,
    int kp, int ma, __bf16 *__restrict__ Xh, __bf16 *__restrict__ Xl, double *__restrict__ n2,
    int *__restrict__ rmol) {
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= rows_pad) return;
    const int sr = srow_of[row];  // packed source row, or -1 for padding
    __bf16 *dh = Xh + (size_t)row * kp;
    __bf16 *dl = Xl + (size_t)row * kp;
    if (sr < 0) {
        for (int k = lane; k < kp; k += 64) {
            dh[k] = __bf16(0.0f);
            dl[k] = __bf16(0.0f);
        }
        if (lane == 0) {
            n2[row] = 0.0;
            rmol[row] = -1;
        }
        return;
    }
    const int atom = src[sr];
    const double *s = X + (size_t)atom * rep;
    const double *c = center + (size_t)lab_of[row] * rep;
    double acc = 0.0;
    for (int k = lane; k < kp; k += 64) {
        float hi = 0.0f, lo = 0.0f;
        if (k < rep) {
            const double v = s[k] - c[k];
            acc += v * v;
            const float vf = (float)v;
            const __bf16 h = __bf16(vf);
            hi = (float)h;
            lo = (float)__bf16((float)(v - (double)hi));
        }
        dh[k] = __bf16(hi);
        dl[k] = __bf16(lo);
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off);
    if (lane == 0) {
        n2[row] = acc;
        rmol[row] = atom / ma;
    }
}

struct MixedSide {
    const __bf16 *Xh, *Xl;
    const double *n2;
    const int *rmol;
    const int *pad0;   // L   first padded row of each label
    const int *rows;   // L   real rows of each label
};

// ---------------------------------------------------------------------------
// The tile kernel
// ---------------------------------------------------------------------------
template <int MODE>
__global__ __launch_bounds__(NT, 2) void local_bf16x3_kernel(MixedSide A, MixedSide B,
                                                             const long *__restrict__ tile_off,
                                                             int L, long ntiles, int kp,
                                                             float cexp, int nm_out,
                                                             float *__restrict__ out) {
    extern __shared__ __align__(16) char smem[];
    __shared__ int rowm[TM], colm[TM];
    __shared__ int rrun_s[TM + 1], crun_s[TM + 1];
    __shared__ int nrr, ncr;

    // XCD-grouped tile order: hardware dispatches consecutive workgroups round-robin
    // over the 8 XCDs; give each XCD a contiguous range of tiles instead.
    long t;
    {
        const long g = blockIdx.x, per = (ntiles + 7) / 8;  // grid = 8 * per
        t = (g % 8) * per + g / 8;                          // a permutation of [0, 8 per)
    }
    if (t >= ntiles) return;
    int l = 0;
    while (l + 1 < L && tile_off[l + 1] <= t) ++l;
    const long tl = t - tile_off[l];
    int I, J;
    if (MODE == M_RECT) {
        const int tb = (B.rows[l] + TM - 1) / TM;
        I = (int)(tl / tb);
        J = (int)(tl % tb);
    } else {
        const int nb = (A.rows[l] + TM - 1) / TM;
        const double dn = 2.0 * nb + 1.0;
        int a = (int)floor((dn - sqrt(dn * dn - 8.0 * (double)tl)) * 0.5);
        if (a < 0) a = 0;
        while (a > 0 && (long)a * nb - (long)a * (a - 1) / 2 > tl) --a;
        while ((long)(a + 1) * nb - (long)(a + 1) * a / 2 <= tl) ++a;
        I = a;
        J = a + (int)(tl - ((long)a * nb - (long)a * (a - 1) / 2));
    }
    const int r0 = A.pad0[l] + I * TM, c0 = B.pad0[l] + J * TM;
    const bool diag = (MODE != M_RECT) && (I == J);

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    __bf16 *sAh = reinterpret_cast<__bf16 *>(smem);
    __bf16 *sAl = sAh + TM * LP;
    __bf16 *sBh = sAl + TM * LP;
    __bf16 *sBl = sBh + TM * LP;

    // global -> register prefetch: segment s = tid + 256 q (q = 0,1): row s/4, part s%4
    bf16x8 ra_h[2], ra_l[2], rb_h[2], rb_l[2];
    auto gload = [&](int k0) {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s = tid + NT * q, row = s >> 2, part = s & 3;
            const size_t oa = (size_t)(r0 + row) * kp + k0 + part * 8;
            const size_t ob = (size_t)(c0 + row) * kp + k0 + part * 8;
            ra_h[q] = *reinterpret_cast<const bf16x8 *>(A.Xh + oa);
            ra_l[q] = *reinterpret_cast<const bf16x8 *>(A.Xl + oa);
            rb_h[q] = *reinterpret_cast<const bf16x8 *>(B.Xh + ob);
            rb_l[q] = *reinterpret_cast<const bf16x8 *>(B.Xl + ob);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int s = tid + NT * q, row = s >> 2, part = s & 3;
            const int o = row * LP + part * 8;
            *reinterpret_cast<bf16x8 *>(sAh + o) = ra_h[q];
            *reinterpret_cast<bf16x8 *>(sAl + o) = ra_l[q];
            *reinterpret_cast<bf16x8 *>(sBh + o) = rb_h[q];
            *reinterpret_cast<bf16x8 *>(sBl + o) = rb_l[q];
        }
    };

    f32x4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};

    gload(0);
    for (int k0 = 0; k0 < kp; k0 += TK) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + TK < kp) gload(k0 + TK);
        const int g8 = (lane >> 4) * 8;
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int o = (wr * 64 + i * 16 + (lane & 15)) * LP + g8;
            ah[i] = *reinterpret_cast<const bf16x8 *>(sAh + o);
            al[i] = *reinterpret_cast<const bf16x8 *>(sAl + o);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int o = (wc * 64 + j * 16 + (lane & 15)) * LP + g8;
            bh[j] = *reinterpret_cast<const bf16x8 *>(sBh + o);
            bl[j] = *reinterpret_cast<const bf16x8 *>(sBl + o);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
                acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            }
    }

    // ---- epilogue ----
    if (tid < TM) rowm[tid] = A.rmol[r0 + tid];
    else colm[tid - TM] = B.rmol[c0 + tid - TM];
    __syncthreads();  // also: every wave is done with the operand tiles
    float *E = reinterpret_cast<float *>(smem);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
            const int row = wr * 64 + i * 16 + mfma16_row_f32(lane, v);
            const int rm = rowm[row];
            const double n2r = A.n2[r0 + row];
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                const int col = wc * 64 + j * 16 + (lane & 15);
                const int cm = colm[col];
                float e = 0.0f;
                const bool lower = diag && row > col;
                if (rm >= 0 && cm >= 0 && !lower) {
                    const double d2 = n2r + B.n2[c0 + col] - 2.0 * (double)acc[i][j][v];
                    e = expf((float)d2 * cexp);
                    if (MODE != M_RECT && rm == cm && !(diag && row == col)) e *= 2.0f;
                }
                E[row * EPITCH + col] = e;
            }
        }
    }
    // runs of equal molecule along rows and columns
    if (tid == 0) {
        int n = 0;
        for (int r = 0; r < TM; ++r)
            if (r == 0 || rowm[r] != rowm[r - 1]) rrun_s[n++] = r;
        rrun_s[n] = TM;
        nrr = n;
    } else if (tid == 64) {
        int n = 0;
        for (int c = 0; c < TM; ++c)
            if (c == 0 || colm[c] != colm[c - 1]) crun_s[n++] = c;
        crun_s[n] = TM;
        ncr = n;
    }
    __syncthreads();
    // stage 1 (in place): E[row][q] = sum over column run q of E[row][..]
    if (tid < TM) {
        float *er = E + tid * EPITCH;
        const int nc = ncr;
        for (int q = 0; q < nc; ++q) {
            float s = 0.0f;
            for (int c = crun_s[q]; c < crun_s[q + 1]; ++c) s += er[c];
            er[q] = s;
        }
    }
    __syncthreads();
    // stage 2: (row run, column run) -> one atomic add
    const int nr_ = nrr, nc_ = ncr;
    for (int item = tid; item < nr_ * nc_; item += NT) {
        const int p = item / nc_, q = item % nc_;
        const int rm = rowm[rrun_s[p]], cm = colm[crun_s[q]];
        if (rm < 0 || cm < 0) continue;
        if (MODE != M_RECT && rm > cm) continue;  // only reached by masked (zero) entries
        float s = 0.0f;
        for (int r = rrun_s[p]; r < rrun_s[p + 1]; ++r) s += E[r * EPITCH + q];
        if (s == 0.0f) continue;
        if (MODE == M_SYMM_RFP) {
            atomicAdd(out + rfp_upper((size_t)nm_out, (size_t)rm, (size_t)cm), s);
        } else if (MODE == M_SYMM_DENSE) {
            atomicAdd(out + (size_t)rm * nm_out + cm, s);
            if (rm != cm) atomicAdd(out + (size_t)cm * nm_out + rm, s);
        } else {
            atomicAdd(out + (size_t)rm * nm_out + cm, s);
        }
    }
}

// Host side ------------------------------------------------------------------

struct MixedPacked {
    MixedSide side{};
    const double *center = nullptr;
    std::vector<int> pad0, rows;
    int rows_pad = 0;
};

// Lay out labels in 128-row padded blocks and pack centred BF16 hi/lo operands.
// `center_in` (device, L x rep) overrides the per-label centring point; the
// rectangular form packs set 2 with set 1's points so both sides share the shift.
int mixed_pack(const LocalPlan &p, const double *X, int rep, int kp, int ma, int ws_pack,
               int ws_plan, hipStream_t stream, MixedPacked &m,
               const double *center_in = nullptr) {
    const int L = p.nlabels;
    m.pad0.assign(L, 0);
    m.rows.assign(L, 0);
    int acc = 0;
    for (int l = 0; l < L; ++l) {
        m.pad0[l] = acc;
        m.rows[l] = p.mol_row[(size_t)l * (p.nm + 1) + p.nm] - p.mol_row[(size_t)l * (p.nm + 1)];
        acc += (m.rows[l] + TM - 1) / TM * TM;
    }
    m.rows_pad = acc;
    // host tables: srow_of[padded row] (-1 for padding), lab_of[padded row], lab_row0[l]
    std::vector<int> srow_of(acc > 0 ? acc : 1, -1), lab_of(acc > 0 ? acc : 1, 0), row0(L);
    for (int l = 0; l < L; ++l) {
        row0[l] = p.mol_row[(size_t)l * (p.nm + 1)];
        for (int r = 0; r < (m.rows[l] + TM - 1) / TM * TM; ++r) {
            lab_of[m.pad0[l] + r] = l;
            if (r < m.rows[l]) srow_of[m.pad0[l] + r] = row0[l] + r;
        }
    }
    const size_t n_int = p.src.size() + 2 * (size_t)srow_of.size() + 3 * (size_t)L + 8;
    int *hst = static_cast<int *>(host_staging(ws_plan, sizeof(int) * n_int));
    KF_REQUIRE(hst, KF_ERR_NOMEM, "pinned staging allocation failed");
    size_t o = 0;
    const size_t o_src = o;
    std::copy(p.src.begin(), p.src.end(), hst + o);
    o += p.src.size();
    const size_t o_srow = o;
    std::copy(srow_of.begin(), srow_of.end(), hst + o);
    o += srow_of.size();
    const size_t o_lab = o;
    std::copy(lab_of.begin(), lab_of.end(), hst + o);
    o += lab_of.size();
    const size_t o_row0 = o;
    std::copy(row0.begin(), row0.end(), hst + o);
    o += L;
    const size_t o_rows = o;
    std::copy(m.rows.begin(), m.rows.end(), hst + o);
    o += L;
    const size_t o_pad0 = o;
    std::copy(m.pad0.begin(), m.pad0.end(), hst + o);
    o += L;
    int *dplan = static_cast<int *>(workspace(ws_plan, sizeof(int) * (o + 8), stream));
    const size_t bytes_ops = (size_t)acc * kp * sizeof(__bf16);
    const size_t pack_bytes = 2 * bytes_ops + sizeof(double) * (size_t)acc + sizeof(int) * (size_t)acc +
                              sizeof(double) * (size_t)L * rep + 1024;
    char *pk = static_cast<char *>(workspace(ws_pack, pack_bytes, stream));
    KF_REQUIRE(dplan && pk, KF_ERR_NOMEM, "device workspace allocation failed (%zu B)", pack_bytes);
    KF_HIP_TRY(hipMemcpyAsync(dplan, hst, sizeof(int) * o, hipMemcpyHostToDevice, stream));
    size_t off = 0;
    __bf16 *Xh = reinterpret_cast<__bf16 *>(pk + off);
    off += bytes_ops;
    __bf16 *Xl = reinterpret_cast<__bf16 *>(pk + off);
    off += bytes_ops;
    off = (off + 255) & ~size_t(255);
    double *n2 = reinterpret_cast<double *>(pk + off);
    off += sizeof(double) * acc;
    double *center = reinterpret_cast<double *>(pk + off);
    off += sizeof(double) * (size_t)L * rep;
    int *rmol = reinterpret_cast<int *>(pk + off);
    if (center_in) {
        center = const_cast<double *>(center_in);
    } else if (L > 0) {
        hipLaunchKernelGGL(label_center, dim3(L, (rep + 63) / 64), dim3(256), 0, stream, X,
                           dplan + o_src, dplan + o_row0, dplan + o_rows, rep, center);
        KF_HIP_TRY(hipGetLastError());
    }
    if (acc > 0) {
        hipLaunchKernelGGL(pack_split_bf16, dim3((acc + 3) / 4), dim3(256), 0, stream, X,
                           dplan + o_src, dplan + o_srow, dplan + o_lab, center, acc, rep, kp,
                           ma, Xh, Xl, n2, rmol);
        KF_HIP_TRY(hipGetLastError());
    }
    m.side = MixedSide{Xh, Xl, n2, rmol, dplan + o_pad0, dplan + o_rows};
    m.center = center;
    return KF_OK;
}

constexpr size_t kMixedLds = 4 * TM * LP * sizeof(__bf16) > sizeof(float) * TM * EPITCH
                                 ? 4 * TM * LP * sizeof(__bf16)
                                 : sizeof(float) * TM * EPITCH;

}  // namespace

// Shared with local_scalar.hip's entry points (precision == KF_PREC_BF16X3).
int local_mixed_symm(int dense, const double *X, int nm, int ma, int rep, double sigma,
                     const LocalPlan &p, void *out, hipStream_t stream) {
    const int kp = (rep + TK - 1) / TK * TK;
    MixedPacked m;
    int rc = mixed_pack(p, X, rep, kp, ma, WS_PACK, WS_PLAN, stream, m);
    if (rc) return rc;
    const int L = p.nlabels;
    std::vector<long> toff(L + 1, 0);
    for (int l = 0; l < L; ++l) {
        const long nb = (m.rows[l] + TM - 1) / TM;
        toff[l + 1] = toff[l] + nb * (nb + 1) / 2;
    }
    const size_t nout = dense ? (size_t)nm * nm : (size_t)nm * (nm + 1) / 2;
    KF_HIP_TRY(hipMemsetAsync(out, 0, nout * sizeof(float), stream));
    const long ntiles = toff[L];
    if (ntiles == 0) return KF_OK;
    long *dtoff = static_cast<long *>(workspace(WS_AUX1, sizeof(long) * (L + 1), stream));
    KF_REQUIRE(dtoff, KF_ERR_NOMEM, "workspace allocation failed");
    long *htoff = static_cast<long *>(host_staging(WS_AUX1, sizeof(long) * (L + 1)));
    KF_REQUIRE(htoff, KF_ERR_NOMEM, "pinned staging allocation failed");
    std::copy(toff.begin(), toff.end(), htoff);
    KF_HIP_TRY(hipMemcpyAsync(dtoff, htoff, sizeof(long) * (L + 1), hipMemcpyHostToDevice, stream));
    const float cexp = (float)(-1.0 / (2.0 * sigma * sigma));
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
    ScopedTimer timer(TM_LOCAL_SYMM, stream);
    if (dense)
        hipLaunchKernelGGL(local_bf16x3_kernel<M_SYMM_DENSE>, dim3(grid), dim3(NT), kMixedLds,
                           stream, m.side, m.side, dtoff, L, ntiles, kp, cexp, nm,
                           static_cast<float *>(out));
    else
        hipLaunchKernelGGL(local_bf16x3_kernel<M_SYMM_RFP>, dim3(grid), dim3(NT), kMixedLds,
                           stream, m.side, m.side, dtoff, L, ntiles, kp, cexp, nm,
                           static_cast<float *>(out));
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

int local_mixed_rect(const double *X1, const double *X2, int nm1, int nm2, int ma1, int ma2,
                     int rep, double sigma, const LocalPlan &p1, const LocalPlan &p2, float *out,
                     hipStream_t stream) {
    const int kp = (rep + TK - 1) / TK * TK;
    MixedPacked m1, m2;
    int rc = mixed_pack(p1, X1, rep, kp, ma1, WS_PACK, WS_PLAN, stream, m1);
    if (rc) return rc;
    if ((rc = mixed_pack(p2, X2, rep, kp, ma2, WS_PACK2, WS_PLAN2, stream, m2, m1.center)))
        return rc;
    const int L = p1.nlabels;
    std::vector<long> toff(L + 1, 0);
    for (int l = 0; l < L; ++l) {
        const long a = (m1.rows[l] + TM - 1) / TM, b = (m2.rows[l] + TM - 1) / TM;
        toff[l + 1] = toff[l] + a * b;
    }
    KF_HIP_TRY(hipMemsetAsync(out, 0, sizeof(float) * (size_t)nm1 * nm2, stream));
    const long ntiles = toff[L];
    if (ntiles == 0) return KF_OK;
    long *dtoff = static_cast<long *>(workspace(WS_AUX1, sizeof(long) * (L + 1), stream));
    long *htoff = static_cast<long *>(host_staging(WS_AUX1, sizeof(long) * (L + 1)));
    KF_REQUIRE(dtoff && htoff, KF_ERR_NOMEM, "workspace allocation failed");
    std::copy(toff.begin(), toff.end(), htoff);
    KF_HIP_TRY(hipMemcpyAsync(dtoff, htoff, sizeof(long) * (L + 1), hipMemcpyHostToDevice, stream));
    const float cexp = (float)(-1.0 / (2.0 * sigma * sigma));
    const unsigned grid = (unsigned)((ntiles + 7) / 8 * 8);
    ScopedTimer timer(TM_LOCAL_RECT, stream);
    hipLaunchKernelGGL(local_bf16x3_kernel<M_RECT>, dim3(grid), dim3(NT), kMixedLds, stream,
                       m1.side, m2.side, dtoff, L, ntiles, kp, cexp, nm2, out);
    KF_HIP_TRY(hipGetLastError());
    return KF_OK;
}

}  // namespace kf
