// fchl19.hip -- FCHL19 atom-centred representation and its coordinate Jacobian
// on gfx950, FP64.
//
// Maths: reference src/fchl19_repr.cpp:89-299 (values) and :308-778 (values +
// Jacobian).  The Jacobian here is derived from the closed form (product rule
// over angular x radial x cutoff x Axilrod-Teller-Muto factors) rather than
// transcribed; it agrees with the reference to FP64 rounding.
//
// Work decomposition
//  * representation: one workgroup (4 waves) per molecule.  Two-body: one
//    thread per (atom, 8-wide radial chunk), pair quantities recomputed per
//    chunk, every output written once (no atomics).  Three-body: one wave per
//    centre; lanes first evaluate up to 64 triplets' geometry into LDS records,
//    then lanes switch to owning output slots (l, t) and stream the records, so
//    each rep entry has exactly one writer and a fixed summation order.
//  * Jacobian: one workgroup per (molecule, centre).  Triplet records carry the
//    value and the 9-component gradients of the angular factor, of
//    W = cutoff x ATM and of rbar; threads own (slot, xyz) columns of an LDS
//    block [40 slots][3*max_atoms] per element-pair channel, which is flushed
//    to HBM as one contiguous run (dX rows are contiguous per (centre, feature)).
#include <cmath>

#include "kf_common.hpp"

namespace kf {
namespace {

constexpr int kMaxElem = 16;
constexpr int kMaxRs = 64;
constexpr int kMaxHarm = 8;
constexpr int kMaxAtoms = 128;
constexpr double kEps = 1e-12;
constexpr double kSqrt2Pi = 2.5066282746310005;

struct FchlParams {
    int nelem, nb2, nb3, na, nharm, nchan, rs, three_off, max_atoms, grad_path;
    int elements[kMaxElem];
    double Rs2[kMaxRs], logRs2[kMaxRs], Rs3[kMaxRs];
    double angw[kMaxHarm];
    double eta2, eta3, rcut, acut, d2, d3, w3;
};

__device__ inline int elem_index(const FchlParams &P, int z) {
    for (int e = 0; e < P.nelem; ++e)
        if (P.elements[e] == z) return e;
    return -1;
}

__device__ inline int pair_channel(int nelem, int p, int q) {
    if (p > q) { int t = p; p = q; q = t; }
    return p * nelem - p * (p - 1) / 2 + (q - p);
}

__device__ inline double cut_fn(double r, double cut) {
    const double inv = cut > 0.0 ? 1.0 / cut : 0.0;
    return 0.5 * (cos(M_PI * r * inv) + 1.0);
}
__device__ inline double cut_dfn(double r, double cut) {
    const double inv = cut > 0.0 ? 1.0 / cut : 0.0;
    return -0.5 * M_PI * inv * sin(M_PI * r * inv);
}

__device__ inline double dist(const double *R, int i, int j) {
    const double dx = R[3 * j] - R[3 * i], dy = R[3 * j + 1] - R[3 * i + 1],
                 dz = R[3 * j + 2] - R[3 * i + 2];
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Two-body pair quantities.  value_k = f * pref0 * exp(-L^2 / 2 s2) / Rs2_k with
// L = ln Rs2_k - mu; d value_k / dr = pref0 exp(..)/Rs2_k * (f * dlog0_k + df),
// dlog0_k = -d2/r - s'/s + L mu'/s2 + L^2 s'/s^3.  Returns false when the
// log-normal width vanishes.
struct TwoBody {
    double s2, s, mu, pref0, f, df, ds, dmu;
};
__device__ inline bool two_body(const FchlParams &P, double r, TwoBody &t, bool want_grad) {
    const double r2 = r * r;
    t.s2 = log1p(P.eta2 / (r2 > kEps ? r2 : kEps));
    t.s = sqrt(t.s2 > 0.0 ? t.s2 : 0.0);
    if (t.s < kEps) return false;
    t.mu = log(r) - 0.5 * t.s2;
    t.f = cut_fn(r, P.rcut);
    t.pref0 = 1.0 / (t.s * kSqrt2Pi * pow(r, P.d2));
    if (want_grad) {
        const double ds2 = -2.0 * P.eta2 / (r * (r2 + P.eta2));
        t.ds = ds2 / (2.0 * t.s);
        t.dmu = 1.0 / r - 0.5 * ds2;
        t.df = cut_dfn(r, P.rcut);
    }
    return true;
}

__device__ inline void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ---------------------------------------------------------------------------
// Representation kernel: one workgroup per molecule.
// dynamic LDS: coords (max_atoms*3) | elem (max_atoms ints) |
//              per wave: records (64 * (3 + na)) + acc (nchan*nb3*na)
// ---------------------------------------------------------------------------
constexpr int kRepThreads = 256;
constexpr int kRepWaves = kRepThreads / 64;
constexpr int kK2Chunk = 8;

__global__ __launch_bounds__(kRepThreads) void fchl19_rep_kernel(
    const double *__restrict__ coords, const int *__restrict__ Q, const int *__restrict__ N,
    FchlParams P, double *__restrict__ X, int *__restrict__ err) {
    extern __shared__ double smem[];
    const int m = blockIdx.x;
    const int ma = P.max_atoms;
    int n = N[m];
    n = n < 0 ? 0 : (n > ma ? ma : n);
    double *R = smem;
    int *el = reinterpret_cast<int *>(R + 3 * ma);
    double *wave_base = reinterpret_cast<double *>(el + ((ma + 1) & ~1));
    const int rec_w = 3 + P.na;
    const int acc_n = P.nchan * P.nb3 * P.na;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    double *recs = wave_base + (size_t)wave * (64 * rec_w + acc_n);
    double *acc = recs + 64 * rec_w;

    const double *Cm = coords + (size_t)m * ma * 3;
    for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) R[t] = Cm[t];
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int e = elem_index(P, Q[(size_t)m * ma + t]);
        el[t] = e;
        if (e < 0) atomicOr(err, 1);
    }
    __syncthreads();
    // a molecule with an unknown element is reported through *err; its rows are zeroed
    bool bad = false;
    for (int t = 0; t < n; ++t) bad |= el[t] < 0;

    double *Xm = X + (size_t)m * ma * P.rs;
    // zero the padding rows
    for (size_t t = (size_t)n * P.rs + threadIdx.x; t < (size_t)ma * P.rs; t += blockDim.x)
        Xm[t] = 0.0;
    if (bad) {
        for (size_t t = threadIdx.x; t < (size_t)n * P.rs; t += blockDim.x) Xm[t] = 0.0;
        return;
    }

    // ---- two-body: thread per (atom i, radial chunk) ----
    const int nchunk = (P.nb2 + kK2Chunk - 1) / kK2Chunk;
    for (int item = threadIdx.x; item < n * nchunk; item += blockDim.x) {
        const int i = item / nchunk, k0 = (item % nchunk) * kK2Chunk;
        const int kn = (P.nb2 - k0) < kK2Chunk ? (P.nb2 - k0) : kK2Chunk;
        for (int e = 0; e < P.nelem; ++e) {
            double v[kK2Chunk];
#pragma unroll
            for (int kk = 0; kk < kK2Chunk; ++kk) v[kk] = 0.0;
            for (int j = 0; j < n; ++j) {
                if (j == i || el[j] != e) continue;
                const double r = dist(R, i, j);
                if (r > P.rcut) continue;
                TwoBody tb;
                if (!two_body(P, r, tb, false)) continue;
#pragma unroll
                for (int kk = 0; kk < kK2Chunk; ++kk) {
                    if (kk < kn) {
                        const double L = P.logRs2[k0 + kk] - tb.mu;
                        v[kk] += tb.f * tb.pref0 * exp(-0.5 * L * L / tb.s2) / P.Rs2[k0 + kk];
                    }
                }
            }
            double *dst = Xm + (size_t)i * P.rs + e * P.nb2 + k0;
            for (int kk = 0; kk < kn; ++kk) dst[kk] = v[kk];
        }
    }

    // ---- three-body: one wave per centre ----
    const int nslot = P.nb3 * P.na;
    for (int i = wave; i < n; i += kRepWaves) {
        for (int t = lane; t < acc_n; t += 64) acc[t] = 0.0;
        const double *Ri = R + 3 * i;
        const int npairs = n * (n - 1) / 2;
        for (int p0 = 0; p0 < npairs; p0 += 64) {
            // phase A: lane -> pair index p0 + lane -> (j, k), j < k
            int valid = 0;
            double rec_chan = 0, rec_rbar = 0, rec_fa = 0;
            double ang[2 * kMaxHarm];
            const int p = p0 + lane;
            if (p < npairs) {
                // invert p = j*n - j(j+1)/2 + (k - j - 1)
                int j = 0, rem = p;
                while (rem >= n - 1 - j) { rem -= n - 1 - j; ++j; }
                const int k = j + 1 + rem;
                if (j != i && k != i) {
                    const double rij = dist(R, i, j), rik = dist(R, i, k);
                    if (rij <= P.acut && rik <= P.acut) {
                        const double rjk = dist(R, j, k);
                        const double *Rj = R + 3 * j, *Rk = R + 3 * k;
                        const double iij = 1.0 / (rij > kEps ? rij : kEps);
                        const double iik = 1.0 / (rik > kEps ? rik : kEps);
                        const double ijk = 1.0 / (rjk > kEps ? rjk : kEps);
                        double eij[3], eik[3], ejk[3];
                        for (int c = 0; c < 3; ++c) {
                            eij[c] = (Rj[c] - Ri[c]) * iij;
                            eik[c] = (Rk[c] - Ri[c]) * iik;
                            ejk[c] = (Rk[c] - Rj[c]) * ijk;
                        }
                        double ci = eij[0] * eik[0] + eij[1] * eik[1] + eij[2] * eik[2];
                        ci = ci > 1.0 ? 1.0 : (ci < -1.0 ? -1.0 : ci);
                        const double cj = -(eij[0] * ejk[0] + eij[1] * ejk[1] + eij[2] * ejk[2]);
                        const double ck = eik[0] * ejk[0] + eik[1] * ejk[1] + eik[2] * ejk[2];
                        const double prod = rik * rij * rjk;
                        const double ksi3 = (1.0 + 3.0 * ci * cj * ck) * P.w3 /
                                            pow(prod > kEps ? prod : kEps, P.d3);
                        const double si = sqrt(fmax(0.0, 1.0 - ci * ci));
                        double cp = 1.0, sp = 0.0, cc = ci, sc = si;
                        int o = 1;
                        for (int h = 0; h < P.nharm; ++h) {
                            while (o < 2 * h + 1) {
                                const double cn = 2.0 * ci * cc - cp, sn = 2.0 * ci * sc - sp;
                                cp = cc; sp = sc; cc = cn; sc = sn; ++o;
                            }
                            ang[2 * h] = P.angw[h] * cc;
                            ang[2 * h + 1] = P.angw[h] * sc;
                        }
                        for (int t = 2 * P.nharm; t < P.na; ++t) ang[t] = 0.0;
                        rec_chan = pair_channel(P.nelem, el[j], el[k]);
                        rec_rbar = 0.5 * (rij + rik);
                        rec_fa = cut_fn(rij, P.acut) * cut_fn(rik, P.acut) * ksi3;
                        valid = 1;
                    }
                }
            }
            // compact valid records (wave ballot prefix)
            const unsigned long long mask = __ballot(valid);
            const int cnt = __popcll(mask);
            if (valid) {
                const int slot = __popcll(mask & ((1ull << lane) - 1ull));
                double *rr = recs + slot * rec_w;
                rr[0] = rec_chan;
                rr[1] = rec_rbar;
                rr[2] = rec_fa;
                for (int t = 0; t < P.na; ++t) rr[3 + t] = ang[t];
            }
            wave_sync();
            // phase B: lane owns slots s = (l, t)
            for (int s = lane; s < nslot; s += 64) {
                const int l = s / P.na, t = s % P.na;
                const double rs3 = P.Rs3[l];
                for (int q = 0; q < cnt; ++q) {
                    const double *rr = recs + q * rec_w;
                    const double dr = rr[1] - rs3;
                    const double v = rr[3 + t] * exp(-P.eta3 * dr * dr) * rr[2];
                    acc[(int)rr[0] * nslot + s] += v;
                }
            }
            wave_sync();
        }
        double *dst = Xm + (size_t)i * P.rs + P.three_off;
        for (int t = lane; t < acc_n; t += 64) dst[t] = acc[t];
        wave_sync();
    }
}

// ---------------------------------------------------------------------------
// Representation + Jacobian kernel: one workgroup per (molecule, centre).
// ---------------------------------------------------------------------------
constexpr int kGradThreads = 256;
constexpr int kRecW = 42;  // a0,a1 | ga0[9] | ga1[9] | W | gW[9] | rbar | grbar[9] | chan | pad
constexpr int kRecChan = 40;

__global__ __launch_bounds__(kGradThreads) void fchl19_grad_kernel(
    const double *__restrict__ coords, const int *__restrict__ Q, const int *__restrict__ N,
    FchlParams P, double *__restrict__ X, double *__restrict__ dX, int *__restrict__ err) {
    extern __shared__ double smem[];
    const int m = blockIdx.y, i = blockIdx.x;
    const int ma = P.max_atoms;
    const int ld = 3 * ma;
    int n = N[m];
    n = n < 0 ? 0 : (n > ma ? ma : n);
    double *Xi = X + ((size_t)m * ma + i) * P.rs;
    double *dXi = dX + ((size_t)m * ma + i) * P.rs * ld;
    if (i >= n) {  // padding centre: zero row and Jacobian rows
        for (int t = threadIdx.x; t < P.rs; t += blockDim.x) Xi[t] = 0.0;
        for (size_t t = threadIdx.x; t < (size_t)P.rs * ld; t += blockDim.x) dXi[t] = 0.0;
        return;
    }
    double *R = smem;                                             // 3*ma
    int *el = reinterpret_cast<int *>(R + 3 * ma);                // ma (padded to even)
    int *tri_jk = el + ((ma + 1) & ~1);                           // 2 * ntri_cap ints
    const int ntri_cap = (ma * (ma - 1)) / 2;
    double *recs = reinterpret_cast<double *>(tri_jk + 2 * ((ntri_cap + 1) & ~1));
    double *blk = recs + (size_t)ntri_cap * kRecW;                // [slots][ld]
    int *chan_cnt = reinterpret_cast<int *>(blk + (size_t)64 * ld);   // nchan + 1
    int *chan_list = chan_cnt + ((P.nchan + 2) & ~1);                // ntri_cap

    const double *Cm = coords + (size_t)m * ma * 3;
    for (int t = threadIdx.x; t < 3 * n; t += blockDim.x) R[t] = Cm[t];
    for (int t = threadIdx.x; t < n; t += blockDim.x) {
        const int e = elem_index(P, Q[(size_t)m * ma + t]);
        el[t] = e;
        if (e < 0) atomicOr(err, 1);
    }
    __syncthreads();
    bool bad = false;
    for (int t = 0; t < n; ++t) bad |= el[t] < 0;
    if (bad) {
        for (int t = threadIdx.x; t < P.rs; t += blockDim.x) Xi[t] = 0.0;
        for (size_t t = threadIdx.x; t < (size_t)P.rs * ld; t += blockDim.x) dXi[t] = 0.0;
        return;
    }
    const double *Ri = R + 3 * i;

    // ---------------- two-body block: features [0, nelem*nb2) ----------------
    // thread per (feature f = e*nb2 + k): value + gradient columns of atoms j (e_j = e) and i
    for (int f0 = 0; f0 < P.nelem * P.nb2; f0 += 64) {
        const int nf = (P.nelem * P.nb2 - f0) < 64 ? (P.nelem * P.nb2 - f0) : 64;
        for (int t = threadIdx.x; t < nf * ld; t += blockDim.x) blk[t] = 0.0;
        __syncthreads();
        for (int item = threadIdx.x; item < nf * 3; item += blockDim.x) {
            const int fl = item / 3, c = item % 3, f = f0 + fl;
            const int e = f / P.nb2, k = f % P.nb2;
            double val = 0.0, gi = 0.0;
            double *row = blk + (size_t)fl * ld;
            for (int j = 0; j < n; ++j) {
                if (j == i || el[j] != e) continue;
                const double r = dist(R, i, j);
                if (r > P.rcut) continue;
                TwoBody tb;
                if (!two_body(P, r, tb, true)) continue;
                const double L = P.logRs2[k] - tb.mu;
                const double v0 = tb.pref0 * exp(-0.5 * L * L / tb.s2) / P.Rs2[k];
                const double dlog0 = -P.d2 / r - tb.ds / tb.s + (L / tb.s2) * tb.dmu +
                                     (L * L / (tb.s2 * tb.s)) * tb.ds;
                const double dv = v0 * (tb.f * dlog0 + tb.df);
                const double g = dv * (Ri[c] - R[3 * j + c]) / r;  // d v / d R_i,c
                val += tb.f * v0;
                gi += g;
                row[3 * j + c] = -g;
            }
            row[3 * i + c] = gi;
            if (c == 0) Xi[f] = val;
        }
        __syncthreads();
        double *dst = dXi + (size_t)f0 * ld;
        for (int t = threadIdx.x; t < nf * ld; t += blockDim.x) dst[t] = blk[t];
        __syncthreads();
    }

    // ---------------- three-body triplet records ----------------
    // enumerate pairs (j<k), j,k != i, within acut
    if (threadIdx.x == 0) {
        int cnt = 0;
        for (int j = 0; j + 1 < n; ++j) {
            if (j == i) continue;
            for (int k = j + 1; k < n; ++k) {
                if (k == i) continue;
                tri_jk[2 * cnt] = j;
                tri_jk[2 * cnt + 1] = k;
                ++cnt;
            }
        }
        chan_cnt[P.nchan] = cnt;
    }
    __syncthreads();
    const int ntri = chan_cnt[P.nchan];
    const double inv_ac = P.acut > 0.0 ? 1.0 / P.acut : 0.0;
    for (int tr = threadIdx.x; tr < ntri; tr += blockDim.x) {
        const int j = tri_jk[2 * tr], k = tri_jk[2 * tr + 1];
        double *rr = recs + (size_t)tr * kRecW;
        const double rij = dist(R, i, j), rik = dist(R, i, k);
        if (rij > P.acut || rik > P.acut) {
            rr[kRecChan] = -1.0;  // channel marker: skip
            continue;
        }
        const double rjk = dist(R, j, k);
        const double *Rj = R + 3 * j, *Rk = R + 3 * k;
        double u[3], v[3], w[3];
        for (int c = 0; c < 3; ++c) {
            u[c] = Rj[c] - Ri[c];
            v[c] = Rk[c] - Ri[c];
            w[c] = Rk[c] - Rj[c];
        }
        const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
        const double uw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
        const double vw = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
        double ci = uv / (rij * rik);
        ci = ci > 1.0 ? 1.0 : (ci < -1.0 ? -1.0 : ci);
        const double cj = -uw / (rij * rjk);
        const double ck = vw / (rik * rjk);
        const double prod = rij * rik * rjk;
        const double P3 = pow(prod > kEps ? prod : kEps, -P.d3);
        const double T = (1.0 + 3.0 * ci * cj * ck) * P.w3 * P3;
        const double fij = cut_fn(rij, P.acut), fik = cut_fn(rik, P.acut);
        const double dfij = -0.5 * M_PI * inv_ac * sin(M_PI * rij * inv_ac);
        const double dfik = -0.5 * M_PI * inv_ac * sin(M_PI * rik * inv_ac);
        const double F = fij * fik;
        const double sden = sqrt(fmax(1e-10, rij * rij * rik * rik - uv * uv)) / (rij * rik);
        const double aw = P.angw[0];
        rr[0] = aw * ci;
        rr[1] = aw * sqrt(fmax(0.0, 1.0 - ci * ci));
        rr[20] = F * T;
        rr[30] = 0.5 * (rij + rik);
        for (int c = 0; c < 9; ++c) {
            const int a = c / 3, d = c % 3;  // atom role (i, j, k), xyz
            double gci, gcj, gck, grij, grik, grjk;
            const double dqj = v[d] / (rij * rik) - ci * u[d] / (rij * rij);
            const double dqk = u[d] / (rij * rik) - ci * v[d] / (rik * rik);
            const double dji = w[d] / (rij * rjk) + cj * u[d] / (rij * rij);
            const double djk = -u[d] / (rij * rjk) - cj * w[d] / (rjk * rjk);
            const double dki = -w[d] / (rik * rjk) + ck * v[d] / (rik * rik);
            const double dkj = -v[d] / (rik * rjk) + ck * w[d] / (rjk * rjk);
            if (a == 0) {
                gci = -(dqj + dqk); gcj = dji; gck = dki;
                grij = -u[d] / rij; grik = -v[d] / rik; grjk = 0.0;
            } else if (a == 1) {
                gci = dqj; gcj = -(dji + djk); gck = dkj;
                grij = u[d] / rij; grik = 0.0; grjk = -w[d] / rjk;
            } else {
                gci = dqk; gcj = djk; gck = -(dki + dkj);
                grij = 0.0; grik = v[d] / rik; grjk = w[d] / rjk;
            }
            const double gF = dfij * fik * grij + fij * dfik * grik;
            const double gT = P.w3 * P3 *
                              (3.0 * (gci * cj * ck + ci * gcj * ck + ci * cj * gck) -
                               (1.0 + 3.0 * ci * cj * ck) * P.d3 *
                                   (grij / rij + grik / rik + grjk / rjk));
            rr[2 + c] = aw * gci;                    // grad a0
            rr[11 + c] = -aw * ci * gci / sden;      // grad a1
            rr[21 + c] = gF * T + F * gT;            // grad W
            rr[31 + c] = 0.5 * (grij + grik);        // grad rbar
        }
        rr[kRecChan] = (double)pair_channel(P.nelem, el[j], el[k]);
    }
    __syncthreads();
    // per-channel triplet lists (serial counting sort: ntri <= ma^2/2)
    if (threadIdx.x == 0) {
        for (int c = 0; c < P.nchan; ++c) chan_cnt[c] = 0;
        for (int tr = 0; tr < ntri; ++tr) {
            const int c = (int)recs[(size_t)tr * kRecW + kRecChan];
            if (c >= 0) chan_cnt[c]++;
        }
        int off = 0;
        for (int c = 0; c < P.nchan; ++c) {
            const int t = chan_cnt[c];
            chan_cnt[c] = off;
            off += t;
        }
        for (int tr = 0; tr < ntri; ++tr) {
            const int c = (int)recs[(size_t)tr * kRecW + kRecChan];
            if (c >= 0) chan_list[chan_cnt[c]++] = tr;
        }
        // chan_cnt[c] now holds the END of channel c's list
    }
    __syncthreads();

    // ---------------- three-body blocks, one channel at a time ----------------
    const int nslot = P.nb3 * P.na;
    for (int ch = 0; ch < P.nchan; ++ch) {
        const int lo = ch == 0 ? 0 : chan_cnt[ch - 1], hi = chan_cnt[ch];
        for (int s0 = 0; s0 < nslot; s0 += 64) {  // slot chunks of the LDS block
            const int ns = (nslot - s0) < 64 ? (nslot - s0) : 64;
            for (int t = threadIdx.x; t < ns * ld; t += blockDim.x) blk[t] = 0.0;
            __syncthreads();
            for (int item = threadIdx.x; item < ns * 3; item += blockDim.x) {
                const int sl = item / 3, d = item % 3, s = s0 + sl;
                const int l = s / P.na, ta = s % P.na;
                const double rs3 = P.Rs3[l];
                double val = 0.0;
                double *row = blk + (size_t)sl * ld;
                for (int q = lo; q < hi; ++q) {
                    const int tr = chan_list[q];
                    const double *rr = recs + (size_t)tr * kRecW;
                    const int j = tri_jk[2 * tr], k = tri_jk[2 * tr + 1];
                    const double dr = rr[30] - rs3;
                    const double rad = exp(-P.eta3 * dr * dr);
                    const double drad = -2.0 * P.eta3 * dr * rad;
                    const double a = ta < 2 ? rr[ta] : 0.0;
                    const double W = rr[20];
                    if (d == 0) val += a * rad * W;
                    if (ta < 2) {
                        const double *ga = rr + (ta == 0 ? 2 : 11);
                        const int atom[3] = {i, j, k};
#pragma unroll
                        for (int role = 0; role < 3; ++role) {
                            const int c = 3 * role + d;
                            row[3 * atom[role] + d] +=
                                rad * (ga[c] * W + a * rr[21 + c]) + a * W * drad * rr[31 + c];
                        }
                    }
                }
                if (d == 0) Xi[P.three_off + ch * nslot + s] = val;
            }
            __syncthreads();
            double *dst = dXi + (size_t)(P.three_off + ch * nslot + s0) * ld;
            for (int t = threadIdx.x; t < ns * ld; t += blockDim.x) dst[t] = blk[t];
            __syncthreads();
        }
    }
}

int build_params(FchlParams &P, int max_atoms, const int *elements, int nelem, const double *Rs2,
                 int nRs2, const double *Rs3, int nRs3, int nabasis, double eta2, double eta3,
                 double zeta, double rcut, double acut, double d2, double d3, double w3,
                 bool grad) {
    KF_REQUIRE(nelem >= 1 && nelem <= kMaxElem, KF_ERR_BASIS, "nelem must be in [1, %d]",
               kMaxElem);
    KF_REQUIRE(nRs2 >= 1 && nRs2 <= kMaxRs && nRs3 >= 1 && nRs3 <= kMaxRs, KF_ERR_BASIS,
               "nRs2/nRs3 must be in [1, %d]", kMaxRs);
    KF_REQUIRE(nabasis >= 0 && nabasis % 2 == 0 && nabasis <= 2 * kMaxHarm, KF_ERR_BASIS,
               "Ts.size() (nabasis) must be even and <= %d", 2 * kMaxHarm);
    KF_REQUIRE(max_atoms >= 1, KF_ERR_ARG, "max_atoms must be >= 1");
    for (int k = 0; k < nRs2; ++k)
        KF_REQUIRE(Rs2[k] > 0.0, KF_ERR_BASIS, "All Rs2 must be > 0");
    P.nelem = nelem;
    P.nb2 = nRs2;
    P.nb3 = nRs3;
    P.na = nabasis;
    P.nharm = grad ? (nabasis >= 2 ? 1 : 0) : nabasis / 2;
    P.nchan = nelem * (nelem + 1) / 2;
    P.three_off = nelem * nRs2;
    P.rs = P.three_off + P.nchan * nRs3 * nabasis;
    P.max_atoms = max_atoms;
    P.grad_path = grad ? 1 : 0;
    for (int e = 0; e < kMaxElem; ++e) P.elements[e] = e < nelem ? elements[e] : 0;
    for (int k = 0; k < kMaxRs; ++k) {
        P.Rs2[k] = k < nRs2 ? Rs2[k] : 1.0;
        P.logRs2[k] = log(P.Rs2[k]);
        P.Rs3[k] = k < nRs3 ? Rs3[k] : 0.0;
    }
    for (int h = 0; h < kMaxHarm; ++h) {
        const double t = zeta * (2 * h + 1);
        P.angw[h] = 2.0 * exp(-0.5 * t * t);
    }
    P.eta2 = eta2;
    P.eta3 = eta3;
    P.rcut = rcut;
    P.acut = acut;
    P.d2 = d2;
    P.d3 = d3;
    P.w3 = w3;
    return KF_OK;
}

}  // namespace
}  // namespace kf

extern "C" int kf_fchl19_f64(const double *coords, const int *Q, const int *N, int nm,
                             int max_atoms, const int *elements, int nelem, const double *Rs2,
                             int nRs2, const double *Rs3, int nRs3, int nabasis, double eta2,
                             double eta3, double zeta, double rcut, double acut,
                             double two_body_decay, double three_body_decay, double w3,
                             double *X, double *dX, void *stream_) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream_));
    using namespace kf;
    hipStream_t stream = as_stream(stream_);
    KF_REQUIRE(nm >= 0, KF_ERR_ARG, "nm must be >= 0");
    if (nm == 0) return KF_OK;
    KF_REQUIRE(coords && Q && N && X && elements && Rs2 && Rs3, KF_ERR_ARG, "null pointer");
    FchlParams P;
    int rc = build_params(P, max_atoms, elements, nelem, Rs2, nRs2, Rs3, nRs3, nabasis, eta2,
                          eta3, zeta, rcut, acut, two_body_decay, three_body_decay, w3,
                          dX != nullptr);
    if (rc) return rc;
    // LDS of the kernels below (whole molecules staged per workgroup)
    const int ma = max_atoms, ld = 3 * ma;
    const int ntri_cap = ma * (ma - 1) / 2;
    const size_t lds_rep = sizeof(double) * (3 * ma) + sizeof(int) * ((ma + 1) & ~1) +
                           sizeof(double) * kRepWaves * (64 * (3 + P.na) + P.nchan * P.nb3 * P.na);
    const size_t lds_grad = sizeof(double) * 3 * ma + sizeof(int) * ((ma + 1) & ~1) +
                            sizeof(int) * 2 * ((ntri_cap + 1) & ~1) +
                            sizeof(double) * (size_t)ntri_cap * kRecW +
                            sizeof(double) * (size_t)64 * ld +
                            sizeof(int) * (((P.nchan + 2) & ~1) + ntri_cap);
    const size_t lds = dX ? lds_grad : lds_rep;
    if (ma > kMaxAtoms || lds > 160 * 1024) {
        // Large molecules: the FCHL19v2 kernel with the log-normal two-body and odd-Fourier
        // r_bar three-body bases is this representation (same values and Jacobian, the
        // gradient path's first-harmonic convention included) and stages no molecule in LDS.
        return kf_fchl19v2_f64(coords, Q, N, nm, max_atoms, elements, nelem, Rs2, nRs2, Rs3,
                               nRs3, nullptr, 0, nabasis, 0, 0, eta2, eta3, eta3, zeta, rcut,
                               acut, two_body_decay, three_body_decay, w3, X, dX, stream_);
    }
    int *err = static_cast<int *>(workspace(WS_AUX3, 64, stream));
    KF_REQUIRE(err, KF_ERR_NOMEM, "workspace allocation failed");
    KF_HIP_TRY(hipMemsetAsync(err, 0, sizeof(int), stream));
    ScopedTimer timer(TM_FCHL19, stream);
    if (!dX) {
        hipLaunchKernelGGL(fchl19_rep_kernel, dim3(nm), dim3(kRepThreads), lds, stream, coords, Q,
                           N, P, X, err);
    } else {
        KF_REQUIRE(P.nelem * P.nb2 > 0, KF_ERR_BASIS, "empty two-body basis");
        hipLaunchKernelGGL(fchl19_grad_kernel, dim3(ma, nm), dim3(kGradThreads), lds, stream,
                           coords, Q, N, P, X, dX, err);
    }
    KF_HIP_TRY(hipGetLastError());
    int herr = 0;
    KF_HIP_TRY(hipMemcpyAsync(&herr, err, sizeof(int), hipMemcpyDeviceToHost, stream));
    KF_HIP_TRY(hipStreamSynchronize(stream));
    KF_REQUIRE(herr == 0, KF_ERR_ELEMENT, "nuclear_z contains an element not present in elements");
    return KF_OK;
}
