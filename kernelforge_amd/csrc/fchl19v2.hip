// fchl19v2.hip -- FCHL19v2 representations (selectable two-body and three-body bases)
// and their coordinate Jacobians on gfx950, FP64.
//
// Maths: reference src/fchl19v2_repr.cpp (values :145-1329, gradient path :436-2870):
//   two-body  T1 log-normal, T2 Gaussian in r, T3 Gaussian in ln r, T4 Gaussian in r
//             without r^-decay, T5 radial Bessel sin(k pi r/rcut)/r;
//   three-body A1 odd Fourier, r_bar radial, ATM; A2 cosine series, r_bar, ATM;
//             A3 odd Fourier, split radial (r_ij + r_ik, |r_ij - r_ik|), ATM;
//             A4 cosine series, split, ATM; A5 cosine series, split, no ATM.
// Every three-body factor is written as a function of the triplet's three distances
// (the centre angle by the law of cosines), so a Jacobian entry is three partial
// derivatives times the pair unit vectors; the derivatives are derived here, not
// transcribed.  Two conventions of the reference's gradient path are kept because they
// define its output: A1 keeps only the first harmonic there (values and Jacobian), and
// the split bases differentiate |r_ij - r_ik| as sqrt((r_ij - r_ik)^2 + 1e-8).
//
// Work decomposition: one workgroup (256 threads) per (molecule, centre atom i).  The
// triplets (j < k, both != i) of the centre are taken in chunks of kChunk; a chunk's
// triplets within both cutoffs are evaluated once into LDS records (distances, cos of
// the centre angle, the angular-independent weight W and its three partials), ordered
// by element-pair channel with a stable rank, so that a thread that owns a feature
// (and, for the Jacobian, one atom's three coordinates) streams only its channel's
// records.  Every output entry has one writer; later chunks add to what earlier
// chunks stored, in a fixed order (deterministic, no atomics).  There is no limit on
// the molecule size.
#include <cmath>

#include "kf_common.hpp"

namespace kf {
namespace {

constexpr int kV2MaxElem = 16;
constexpr int kV2MaxRs = 64;
constexpr int kV2MaxHarm = 16;
constexpr int kV2Threads = 256;
constexpr int kChunk = 512;
constexpr double kV2Eps = 1e-12;
constexpr double kV2Sqrt2Pi = 2.5066282746310005;
constexpr double kSplitSmooth = 1e-8;

enum TwoBody { TB_LOGNORMAL = 0, TB_GAUSS_R, TB_GAUSS_LOGR, TB_GAUSS_R_NOPOW, TB_BESSEL };
enum ThreeBody { AB_ODD_RBAR = 0, AB_COS_RBAR, AB_ODD_SPLIT, AB_COS_SPLIT, AB_COS_SPLIT_NOATM };

struct V2Params {
    int nelem, nb2, nb3, nb3m, nab, tb, ab, grad_path;
    int rep, three_off, nrad, max_atoms;
    int elements[kV2MaxElem];
    double Rs2[kV2MaxRs], Rs3[kV2MaxRs], Rs3m[kV2MaxRs];
    double ow[kV2MaxHarm];  // 2 exp(-(zeta o)^2 / 2), o = 2l + 1
    double eta2, eta3, eta3m, rcut, acut, d2, d3, w3;
};

__device__ inline bool is_split(int ab) { return ab >= AB_ODD_SPLIT; }
__device__ inline bool is_odd(int ab) { return ab == AB_ODD_RBAR || ab == AB_ODD_SPLIT; }

__device__ inline int v2_elem(const V2Params &P, int z) {
    for (int e = 0; e < P.nelem; ++e)
        if (P.elements[e] == z) return e;
    return -1;
}

__device__ inline int v2_pair(int nelem, int p, int q) {
    if (p > q) { const int t = p; p = q; q = t; }
    return -p * (p + 1) / 2 + q + nelem * p;
}

// cutoff decay 1/2 (cos(pi r / c) + 1) and its derivative (c <= 0: no decay)
__device__ inline double decay(double r, double c) {
    return 0.5 * (cos(M_PI * r * (c > 0.0 ? 1.0 / c : 0.0)) + 1.0);
}
__device__ inline double ddecay(double r, double c) {
    const double inv = c > 0.0 ? 1.0 / c : 0.0;
    return -0.5 * M_PI * inv * sin(M_PI * r * inv);
}

__device__ inline void coord(const double *R, int a, double &x, double &y, double &z) {
    x = R[3 * a];
    y = R[3 * a + 1];
    z = R[3 * a + 2];
}
__device__ inline double distance(const double *R, int a, int b) {
    double xa, ya, za, xb, yb, zb;
    coord(R, a, xa, ya, za);
    coord(R, b, xb, yb, zb);
    const double dx = xb - xa, dy = yb - ya, dz = zb - za;
    return sqrt(dx * dx + dy * dy + dz * dz);
}

// Two-body basis function k at distance r and its derivative; false when the pair is
// skipped (outside the cutoff, vanishing log-normal width, r ~ 0 for the Bessel basis).
__device__ inline bool two_body_fn(const V2Params &P, int k, double r, double &v, double &dv) {
    if (r > P.rcut) return false;
    const double dec = decay(r, P.rcut), ddec = ddecay(r, P.rcut);
    switch (P.tb) {
        case TB_LOGNORMAL: {
            const double r2 = r * r;
            const double s2 = log1p(P.eta2 / (r2 > kV2Eps ? r2 : kV2Eps));
            const double s = sqrt(s2 > 0.0 ? s2 : 0.0);
            if (s < kV2Eps) return false;
            const double mu = log(r) - 0.5 * s2;
            const double L = log(P.Rs2[k]) - mu;
            const double g = exp(-0.5 * L * L / (s * s)) / (s * kV2Sqrt2Pi * pow(r, P.d2) * P.Rs2[k]);
            v = dec * g;
            // d ln g / dr with ds2 = d s^2/dr, mu' = 1/r - ds2/2
            const double ds2 = -2.0 * P.eta2 / (r * (r2 + P.eta2));
            const double dmu = 1.0 / r - 0.5 * ds2;
            const double dlng = L * dmu / s2 + 0.5 * L * L * ds2 / (s2 * s2) - 0.5 * ds2 / s2 - P.d2 / r;
            dv = g * (ddec + dec * dlng);
            return true;
        }
        case TB_GAUSS_R: {
            const double t = r - P.Rs2[k];
            const double g = pow(r, -P.d2) * exp(-P.eta2 * t * t);
            v = dec * g;
            dv = g * (ddec + dec * (-P.d2 / r - 2.0 * P.eta2 * t));
            return true;
        }
        case TB_GAUSS_LOGR: {
            const double t = log(r) - log(P.Rs2[k]);
            const double g = pow(r, -P.d2) * exp(-P.eta2 * t * t);
            v = dec * g;
            dv = g * (ddec + dec * (-P.d2 / r - 2.0 * P.eta2 * t / r));
            return true;
        }
        case TB_GAUSS_R_NOPOW: {
            const double t = r - P.Rs2[k];
            const double g = exp(-P.eta2 * t * t);
            v = dec * g;
            dv = g * (ddec - 2.0 * P.eta2 * t * dec);
            return true;
        }
        default: {  // TB_BESSEL
            if (r < kV2Eps) return false;
            const double w = (k + 1) * M_PI / P.rcut, nrm = sqrt(2.0 / P.rcut);
            const double sn = sin(w * r), cs = cos(w * r);
            v = dec * nrm * sn / r;
            dv = nrm * (ddec * sn / r + dec * (w * cs * r - sn) / (r * r));
            return true;
        }
    }
}

// Angular basis entry t at cos(theta) = c, and dA/dc.  Odd Fourier: entries 2l, 2l+1 =
// w_l cos(o theta), w_l sin(o theta), o = 2l+1; cosine series: entry m = cos(m theta) =
// T_m(c).  isin = 1 / sin(theta) (with the reference's clamp of r_ij r_ik sin(theta)).
__device__ inline void angular_fn(const V2Params &P, int t, double c, double s, double isin,
                                  double &A, double &dA) {
    if (is_odd(P.ab)) {
        const int l = t >> 1, o = 2 * l + 1;
        // cos(o theta), sin(o theta) by the Chebyshev recurrences in c
        double c0 = 1.0, s0 = 0.0, c1 = c, s1 = s;
        for (int q = 2; q <= o; ++q) {
            const double cn = 2.0 * c * c1 - c0, sn = 2.0 * c * s1 - s0;
            c0 = c1; s0 = s1; c1 = cn; s1 = sn;
        }
        // d cos(o theta)/dc = o sin(o theta)/sin(theta), d sin(o theta)/dc = -o cos(o theta)/sin(theta)
        if (t & 1) {
            A = P.ow[l] * s1;
            dA = -P.ow[l] * o * c1 * isin;
        } else {
            A = P.ow[l] * c1;
            dA = P.ow[l] * o * s1 * isin;
        }
        return;
    }
    // T_t(c) and T_t'(c) = t U_{t-1}(c)
    double T0 = 1.0, T1 = c, U0 = 1.0, U1 = 2.0 * c;  // U_{m}: U0 = 1, U1 = 2c
    if (t == 0) { A = 1.0; dA = 0.0; return; }
    if (t == 1) { A = c; dA = 1.0; return; }
    for (int m = 2; m <= t; ++m) {
        const double Tn = 2.0 * c * T1 - T0;
        T0 = T1; T1 = Tn;
    }
    // U_{t-1}
    for (int m = 2; m <= t - 1; ++m) {
        const double Un = 2.0 * c * U1 - U0;
        U0 = U1; U1 = Un;
    }
    const double Ut1 = (t - 1 == 0) ? 1.0 : U1;
    A = T1;
    dA = t * Ut1;
}

// Radial factor of radial index lr at (a = r_ij, b = r_ik), with dR/da, dR/db.
__device__ inline void radial_fn(const V2Params &P, int lr, double a, double b, double &R,
                                 double &dRa, double &dRb) {
    if (!is_split(P.ab)) {
        const double t = 0.5 * (a + b) - P.Rs3[lr];
        R = exp(-P.eta3 * t * t);
        dRa = dRb = -P.eta3 * t * R;  // d/da of (a+b)/2 is 1/2
        return;
    }
    const int l1 = lr / P.nb3m, l2 = lr - l1 * P.nb3m;
    const double tp = a + b - P.Rs3[l1];
    const double php = exp(-P.eta3 * tp * tp), dphp = -2.0 * P.eta3 * tp * php;
    const double dm = a - b;
    const double tm = fabs(dm) - P.Rs3m[l2];
    const double phm = exp(-P.eta3m * tm * tm);
    // derivative through the smoothed |a - b| (reference gradient-path convention)
    const double sm = sqrt(dm * dm + kSplitSmooth);
    const double dphm = phm * (-2.0 * P.eta3m * (sm - P.Rs3m[l2])) * (dm / sm);
    R = php * phm;
    dRa = dphp * phm + php * dphm;
    dRb = dphp * phm - php * dphm;
}

// LDS triplet records of one chunk, sorted by channel.
struct Records {
    int j[kChunk], k[kChunk];
    double a[kChunk], b[kChunk], d[kChunk], c[kChunk], W[kChunk];
    double Wa[kChunk], Wb[kChunk], Wd[kChunk];
    int chan[kChunk];    // per candidate slot (unsorted), -1 outside the cutoffs
    int start[kV2MaxElem * (kV2MaxElem + 1) / 2 + 1];
    int nvalid;
};

// (j, k) of candidate pair t of the M = n - 1 atoms other than i (pairs o_j < o_k of
// [0, M) in row order; S(r) = r M - r (r + 1)/2 pairs precede row r).
__device__ inline void pair_of(long t, int M, int i, int &j, int &k) {
    const double m2 = 2.0 * M - 1.0;
    long r = (long)floor((m2 - sqrt(fmax(m2 * m2 - 8.0 * (double)t, 0.0))) * 0.5);
    if (r < 0) r = 0;
    auto S = [&](long q) { return q * M - q * (q + 1) / 2; };
    while (r + 1 < M && S(r + 1) <= t) ++r;
    while (r > 0 && S(r) > t) --r;
    const long c = t - S(r) + r + 1;
    j = (int)r + ((int)r >= i ? 1 : 0);
    k = (int)c + ((int)c >= i ? 1 : 0);
}

template <bool GRAD>
__global__ __launch_bounds__(kV2Threads) void fchl19v2_kernel(
    const double *__restrict__ coords, const int *__restrict__ Q, const int *__restrict__ N,
    V2Params P, double *__restrict__ X, double *__restrict__ dX, int *__restrict__ err) {
    __shared__ Records S;
    const int m = blockIdx.x, i = blockIdx.y;
    const int ma = P.max_atoms, rep = P.rep, ncol = 3 * ma;
    const int n = min(max(N[m], 0), ma);
    double *xrow = X + ((size_t)m * ma + i) * rep;
    double *jrow = GRAD ? dX + ((size_t)m * ma + i) * rep * (size_t)ncol : nullptr;
    const double *R = coords + (size_t)m * ma * 3;
    const int *Qm = Q + (size_t)m * ma;
    if (i >= n) {  // padded centre: zero row (and Jacobian block)
        for (int f = threadIdx.x; f < rep; f += kV2Threads) xrow[f] = 0.0;
        if (GRAD)
            for (size_t e = threadIdx.x; e < (size_t)rep * ncol; e += kV2Threads) jrow[e] = 0.0;
        return;
    }
    for (int a = threadIdx.x; a < n; a += kV2Threads)
        if (v2_elem(P, Qm[a]) < 0) atomicOr(err, 1);
    // ---- two-body features f = e * nb2 + k (every pair (i, j), j != i, of element e)
    const int n2 = P.nelem * P.nb2;
    const int per_f = GRAD ? ma : 1;
    for (long idx = threadIdx.x; idx < (long)n2 * per_f; idx += kV2Threads) {
        const int f = (int)(idx / per_f), a = GRAD ? (int)(idx - (long)f * per_f) : i;
        const int e = f / P.nb2, k = f - e * P.nb2;
        double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
        if (a < n) {
            double xi, yi, zi;
            coord(R, i, xi, yi, zi);
            for (int j = 0; j < n; ++j) {
                if (j == i || v2_elem(P, Qm[j]) != e) continue;
                if (GRAD && a != i && a != j) continue;
                const double r = distance(R, i, j);
                double v, dv;
                if (!two_body_fn(P, k, r, v, dv)) continue;
                if (a == i) val += v;
                if (GRAD) {
                    double xj, yj, zj;
                    coord(R, j, xj, yj, zj);
                    const double s = (a == i ? -dv : dv) / r;  // d r/d R_i = -(R_j - R_i)/r
                    gx += s * (xj - xi);
                    gy += s * (yj - yi);
                    gz += s * (zj - zi);
                }
            }
        }
        if (a == i) xrow[f] = val;
        if (GRAD) {
            double *o = jrow + (size_t)f * ncol + 3 * a;
            o[0] = gx;
            o[1] = gy;
            o[2] = gz;
        }
    }
    // ---- three-body features, chunks of candidate triplets
    const int nthree = rep - P.three_off;
    const int npair = P.nelem * (P.nelem + 1) / 2;
    const int per_chan = P.nrad * P.nab;
    const int M = n - 1;
    const long ntrip = M >= 2 ? (long)M * (M - 1) / 2 : 0;
    const bool keep_first_only = GRAD && P.grad_path && P.ab == AB_ODD_RBAR;
    for (long t0 = 0, chunk = 0; chunk == 0 || t0 < ntrip; t0 += kChunk, ++chunk) {
        const int cnt = (int)min((long)kChunk, max(ntrip - t0, 0L));
        __syncthreads();  // the previous chunk's records are no longer read
        // channel of each candidate (-1 outside the cutoffs)
        for (int s = threadIdx.x; s < cnt; s += kV2Threads) {
            int j, k;
            pair_of(t0 + s, M, i, j, k);
            const double rij = distance(R, i, j), rik = distance(R, i, k);
            S.chan[s] = (rij > P.acut || rik > P.acut)
                            ? -1
                            : v2_pair(P.nelem, v2_elem(P, Qm[j]), v2_elem(P, Qm[k]));
        }
        __syncthreads();
        if ((int)threadIdx.x <= npair) {  // counts per channel, then exclusive offsets
            int c = 0;
            if ((int)threadIdx.x < npair)
                for (int s = 0; s < cnt; ++s) c += S.chan[s] == (int)threadIdx.x;
            S.start[threadIdx.x] = c;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            int acc = 0;
            for (int p = 0; p <= npair; ++p) {
                const int c = S.start[p];
                S.start[p] = acc;
                acc += c;
            }
            S.nvalid = acc;
        }
        __syncthreads();
        // records at their sorted positions (stable rank inside the channel)
        for (int s = threadIdx.x; s < cnt; s += kV2Threads) {
            const int ch = S.chan[s];
            if (ch < 0) continue;
            int rank = 0;
            for (int u = 0; u < s; ++u) rank += S.chan[u] == ch;
            const int pos = S.start[ch] + rank;
            int j, k;
            pair_of(t0 + s, M, i, j, k);
            const double a = distance(R, i, j), b = distance(R, i, k), d = distance(R, j, k);
            double c = (a * a + b * b - d * d) / (2.0 * a * b);
            c = fmin(1.0, fmax(-1.0, c));
            const double deca = decay(a, P.acut), decb = decay(b, P.acut);
            const double ddeca = ddecay(a, P.acut), ddecb = ddecay(b, P.acut);
            double W, Wa, Wb, Wd;
            if (P.ab == AB_COS_SPLIT_NOATM) {
                W = P.w3 * deca * decb;
                Wa = P.w3 * ddeca * decb;
                Wb = P.w3 * deca * ddecb;
                Wd = 0.0;
            } else {
                // ATM = (1 + 3 c_i c_j c_k) w3 (a b d)^-d3, angles by the law of cosines
                const double cj = (a * a + d * d - b * b) / (2.0 * a * d);
                const double ck = (b * b + d * d - a * a) / (2.0 * b * d);
                const double dci_a = 1.0 / b - c / a, dci_b = 1.0 / a - c / b, dci_d = -d / (a * b);
                const double dcj_a = 1.0 / d - cj / a, dcj_d = 1.0 / a - cj / d, dcj_b = -b / (a * d);
                const double dck_b = 1.0 / d - ck / b, dck_d = 1.0 / b - ck / d, dck_a = -a / (b * d);
                const double poly = 1.0 + 3.0 * c * cj * ck;
                const double prod = a * b * d;
                const double pw = P.w3 * pow(prod > kV2Eps ? prod : kV2Eps, -P.d3);
                const double atm = poly * pw;
                const double dpa = 3.0 * (cj * ck * dci_a + c * ck * dcj_a + c * cj * dck_a);
                const double dpb = 3.0 * (cj * ck * dci_b + c * ck * dcj_b + c * cj * dck_b);
                const double dpd = 3.0 * (cj * ck * dci_d + c * ck * dcj_d + c * cj * dck_d);
                const double atm_a = dpa * pw - P.d3 * atm / a;
                const double atm_b = dpb * pw - P.d3 * atm / b;
                const double atm_d = dpd * pw - P.d3 * atm / d;
                W = atm * deca * decb;
                Wa = atm_a * deca * decb + atm * ddeca * decb;
                Wb = atm_b * deca * decb + atm * deca * ddecb;
                Wd = atm_d * deca * decb;
            }
            S.j[pos] = j;
            S.k[pos] = k;
            S.a[pos] = a;
            S.b[pos] = b;
            S.d[pos] = d;
            S.c[pos] = c;
            S.W[pos] = W;
            S.Wa[pos] = Wa;
            S.Wb[pos] = Wb;
            S.Wd[pos] = Wd;
        }
        __syncthreads();
        for (long idx = threadIdx.x; idx < (long)nthree * per_f; idx += kV2Threads) {
            const int f3 = (int)(idx / per_f), a = GRAD ? (int)(idx - (long)f3 * per_f) : i;
            const int p = f3 / per_chan, rem = f3 - p * per_chan;
            const int lr = rem / P.nab, t = rem - lr * P.nab;
            double val = 0.0, gx = 0.0, gy = 0.0, gz = 0.0;
            if (a < n && !(keep_first_only && t >= 2)) {
                double xi, yi, zi;
                coord(R, i, xi, yi, zi);
                for (int q = S.start[p]; q < S.start[p + 1]; ++q) {
                    const int j = S.j[q], k = S.k[q];
                    if (GRAD && a != i && a != j && a != k) continue;
                    const double ra = S.a[q], rb = S.b[q], rd = S.d[q], c = S.c[q], W = S.W[q];
                    double Rv, dRa, dRb;
                    radial_fn(P, lr, ra, rb, Rv, dRa, dRb);
                    const double s = sqrt(fmax(0.0, 1.0 - c * c));
                    const double den = sqrt(fmax(1e-10, ra * ra * rb * rb * (1.0 - c * c)));
                    const double isin = ra * rb / den;
                    double A, dA;
                    angular_fn(P, t, c, s, isin, A, dA);
                    if (a == i) val += Rv * A * W;
                    if (!GRAD) continue;
                    const double dca = 1.0 / rb - c / ra, dcb = 1.0 / ra - c / rb, dcd = -rd / (ra * rb);
                    // dV/dr for r in (r_ij, r_ik, r_jk)
                    const double Va = dRa * A * W + Rv * dA * dca * W + Rv * A * S.Wa[q];
                    const double Vb = dRb * A * W + Rv * dA * dcb * W + Rv * A * S.Wb[q];
                    const double Vd = Rv * dA * dcd * W + Rv * A * S.Wd[q];
                    double xj, yj, zj, xk, yk, zk;
                    coord(R, j, xj, yj, zj);
                    coord(R, k, xk, yk, zk);
                    // r_ij: d/dR_i = -(R_j - R_i)/r_ij, d/dR_j = +(R_j - R_i)/r_ij; r_jk likewise
                    double cx = 0.0, cy = 0.0, cz = 0.0;
                    if (a == i) {
                        cx = -(Va * (xj - xi) / ra + Vb * (xk - xi) / rb);
                        cy = -(Va * (yj - yi) / ra + Vb * (yk - yi) / rb);
                        cz = -(Va * (zj - zi) / ra + Vb * (zk - zi) / rb);
                    } else if (a == j) {
                        cx = Va * (xj - xi) / ra - Vd * (xk - xj) / rd;
                        cy = Va * (yj - yi) / ra - Vd * (yk - yj) / rd;
                        cz = Va * (zj - zi) / ra - Vd * (zk - zj) / rd;
                    } else {
                        cx = Vb * (xk - xi) / rb + Vd * (xk - xj) / rd;
                        cy = Vb * (yk - yi) / rb + Vd * (yk - yj) / rd;
                        cz = Vb * (zk - zi) / rb + Vd * (zk - zj) / rd;
                    }
                    gx += cx;
                    gy += cy;
                    gz += cz;
                }
            }
            const int f = P.three_off + f3;
            if (a == i) xrow[f] = chunk == 0 ? val : xrow[f] + val;
            if (GRAD) {
                double *o = jrow + (size_t)f * ncol + 3 * a;
                if (chunk == 0) {
                    o[0] = gx;
                    o[1] = gy;
                    o[2] = gz;
                } else {
                    o[0] += gx;
                    o[1] += gy;
                    o[2] += gz;
                }
            }
        }
    }
    // Jacobian columns of padded atoms a in [n, ma) are written by the loops above (zero)
}

}  // namespace
}  // namespace kf

// Representation (and optionally its Jacobian) of FCHL19v2 for a padded batch.
extern "C" int kf_fchl19v2_f64(const double *coords, const int *Q, const int *N, int nm,
                               int max_atoms, const int *elements, int nelem, const double *Rs2,
                               int nRs2, const double *Rs3, int nRs3, const double *Rs3_minus,
                               int nRs3_minus, int nabasis, int two_body_type,
                               int three_body_type, double eta2, double eta3, double eta3_minus,
                               double zeta, double rcut, double acut, double two_body_decay,
                               double three_body_decay, double w3, double *X, double *dX,
                               void *stream) {
    kf::EntryGuard kf_entry_guard_(kf::as_stream(stream));
    using namespace kf;
    KF_REQUIRE(nm >= 0 && max_atoms >= 0, KF_ERR_ARG, "bad sizes");
    KF_REQUIRE(nelem >= 1 && nelem <= kV2MaxElem, KF_ERR_BASIS, "1..%d elements supported",
               kV2MaxElem);
    KF_REQUIRE(nRs2 >= 0 && nRs2 <= kV2MaxRs && nRs3 >= 0 && nRs3 <= kV2MaxRs &&
                   nRs3_minus >= 0 && nRs3_minus <= kV2MaxRs,
               KF_ERR_BASIS, "at most %d radial functions per grid", kV2MaxRs);
    KF_REQUIRE(two_body_type >= 0 && two_body_type <= 4, KF_ERR_ARG, "unknown two_body_type");
    KF_REQUIRE(three_body_type >= 0 && three_body_type <= 4, KF_ERR_ARG,
               "unknown three_body_type");
    const bool split = three_body_type >= AB_ODD_SPLIT;
    const bool odd = three_body_type == AB_ODD_RBAR || three_body_type == AB_ODD_SPLIT;
    KF_REQUIRE(!split || nRs3_minus > 0, KF_ERR_BASIS,
               "nbasis3_minus must be > 0 for SplitR variants");
    KF_REQUIRE(nabasis >= 0 && (!odd || (nabasis % 2 == 0 && nabasis / 2 <= kV2MaxHarm)) &&
                   nabasis <= 2 * kV2MaxHarm,
               KF_ERR_BASIS, "unsupported nabasis %d", nabasis);
    V2Params P{};
    P.nelem = nelem;
    P.nb2 = nRs2;
    P.nb3 = nRs3;
    P.nb3m = split ? nRs3_minus : 1;
    P.nab = nabasis;
    P.tb = two_body_type;
    P.ab = three_body_type;
    P.grad_path = dX != nullptr;
    P.nrad = split ? nRs3 * nRs3_minus : nRs3;
    P.three_off = nelem * nRs2;
    P.rep = P.three_off + nelem * (nelem + 1) / 2 * P.nrad * nabasis;
    P.max_atoms = max_atoms;
    for (int e = 0; e < nelem; ++e) P.elements[e] = elements[e];
    const bool log_grid = two_body_type == TB_LOGNORMAL || two_body_type == TB_GAUSS_LOGR;
    for (int k = 0; k < nRs2; ++k) {
        KF_REQUIRE(!log_grid || Rs2[k] > 0.0, KF_ERR_BASIS, "All Rs2 must be > 0");
        P.Rs2[k] = Rs2[k];
    }
    for (int k = 0; k < nRs3; ++k) P.Rs3[k] = Rs3[k];
    for (int k = 0; k < nRs3_minus; ++k) P.Rs3m[k] = Rs3_minus[k];
    for (int l = 0; l < nabasis / 2 && l < kV2MaxHarm; ++l) {
        const double t = zeta * (2 * l + 1);
        P.ow[l] = 2.0 * std::exp(-0.5 * t * t);
    }
    P.eta2 = eta2;
    P.eta3 = eta3;
    P.eta3m = eta3_minus;
    P.rcut = rcut;
    P.acut = acut;
    P.d2 = two_body_decay;
    P.d3 = three_body_decay;
    P.w3 = w3;
    if (nm == 0 || max_atoms == 0) return KF_OK;
    KF_REQUIRE(coords && Q && N && X, KF_ERR_ARG, "null pointer");
    hipStream_t st = as_stream(stream);
    int *err = static_cast<int *>(workspace(WS_AUX3, 64, st));
    KF_REQUIRE(err, KF_ERR_NOMEM, "workspace allocation failed");
    KF_HIP_TRY(hipMemsetAsync(err, 0, sizeof(int), st));
    ScopedTimer timer(TM_FCHL19, st);
    const dim3 grid((unsigned)nm, (unsigned)max_atoms);
    if (dX)
        hipLaunchKernelGGL(fchl19v2_kernel<true>, grid, dim3(kV2Threads), 0, st, coords, Q, N, P,
                           X, dX, err);
    else
        hipLaunchKernelGGL(fchl19v2_kernel<false>, grid, dim3(kV2Threads), 0, st, coords, Q, N, P,
                           X, (double *)nullptr, err);
    KF_HIP_TRY(hipGetLastError());
    timer.stop();
    int h = 0;
    KF_HIP_TRY(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, st));
    KF_HIP_TRY(hipStreamSynchronize(st));
    KF_REQUIRE(h == 0, KF_ERR_ELEMENT, "nuclear_z contains an element not present in elements");
    return KF_OK;
}
