/*
 * kf_oracle.c -- CPU FP64 restatement of kernelforge's FCHL19 KRR hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle and the CPU
 * timing baseline.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product (libkf_hip.so + the pybind11
 * modules) never links, calls or falls back to it.
 *
 * Every function restates the algorithm of the reference file:line cited in its
 * comment.  BLAS/LAPACK come from the caller through kfo_set_blas(): the Python
 * loader hands over the Fortran-ABI entry points that scipy.linalg.cython_blas /
 * cython_lapack export (OpenBLAS bundled with scipy), so no header is needed.
 *
 * Layouts (all row-major, as the reference's pybind layer hands them over):
 *   X   (nm, max_atoms, rep)            Q (nm, max_atoms) int32   N (nm,) int32
 *   dX  (nm, max_atoms, rep, 3*max_atoms)
 *   RFP: TRANSR='N', UPLO='U', length n(n+1)/2 (src/rfp_utils.hpp:14-22)
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define KFO_EPS 1e-12
#define KFO_SQRT_2PI 2.5066282746310005

/* ------------------------------------------------------------------------- */
/* BLAS / LAPACK entry points (Fortran ABI, column-major)                     */
/* ------------------------------------------------------------------------- */
typedef void (*dgemm_fn)(const char *, const char *, const int *, const int *, const int *,
                         const double *, const double *, const int *, const double *,
                         const int *, const double *, double *, const int *);
typedef void (*dsyrk_fn)(const char *, const char *, const int *, const int *, const double *,
                         const double *, const int *, const double *, double *, const int *);
typedef void (*dgemv_fn)(const char *, const int *, const int *, const double *,
                         const double *, const int *, const double *, const int *,
                         const double *, double *, const int *);
typedef void (*dger_fn)(const int *, const int *, const double *, const double *, const int *,
                        const double *, const int *, double *, const int *);
typedef void (*dpftrf_fn)(const char *, const char *, const int *, double *, int *);
typedef void (*dpftrs_fn)(const char *, const char *, const int *, const int *, const double *,
                          double *, const int *, int *);

static dgemm_fn p_dgemm;
static dsyrk_fn p_dsyrk;
static dgemv_fn p_dgemv;
static dger_fn p_dger;
static dpftrf_fn p_dpftrf;
static dpftrs_fn p_dpftrs;

int kfo_set_blas(void *dgemm, void *dsyrk, void *dgemv, void *dger, void *dpftrf, void *dpftrs) {
    p_dgemm = (dgemm_fn)dgemm;
    p_dsyrk = (dsyrk_fn)dsyrk;
    p_dgemv = (dgemv_fn)dgemv;
    p_dger = (dger_fn)dger;
    p_dpftrf = (dpftrf_fn)dpftrf;
    p_dpftrs = (dpftrs_fn)dpftrs;
    return (p_dgemm && p_dsyrk && p_dgemv && p_dger && p_dpftrf && p_dpftrs) ? 0 : -1;
}

int kfo_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* Row-major C(m,n) = alpha * A(m,k) * B(n,k)^T + beta * C. */
static void rm_gemm_nt(int m, int n, int k, double alpha, const double *A, int lda,
                       const double *B, int ldb, double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    p_dgemm("T", "N", &n, &m, &k, &alpha, B, &ldb, A, &lda, &beta, C, &ldc);
}
/* Row-major C(m,n) = alpha * A(k,m)^T * B(k,n) + beta * C. */
static void rm_gemm_tn(int m, int n, int k, double alpha, const double *A, int lda,
                       const double *B, int ldb, double beta, double *C, int ldc) {
    if (m <= 0 || n <= 0) return;
    p_dgemm("N", "T", &n, &m, &k, &alpha, B, &ldb, A, &lda, &beta, C, &ldc);
}
/* Row-major upper triangle of C(n,n) = alpha * A(n,k) A^T. */
static void rm_syrk_upper(int n, int k, double alpha, const double *A, int lda, double *C,
                          int ldc) {
    const double beta = 0.0;
    if (n <= 0) return;
    p_dsyrk("L", "T", &n, &k, &alpha, A, &lda, &beta, C, &ldc);
}
/* y(n) = alpha * A(m,n)^T x(m) + beta y, A row-major with leading dim lda. */
static void rm_gemv_t(int m, int n, double alpha, const double *A, int lda, const double *x,
                      double beta, double *y) {
    const int one = 1;
    if (n <= 0) return;
    p_dgemv("N", &n, &m, &alpha, A, &lda, x, &one, &beta, y, &one);
}
/* y(m) = alpha * A(m,n) x(n) + beta y. */
static void rm_gemv_n(int m, int n, double alpha, const double *A, int lda, const double *x,
                      double beta, double *y) {
    const int one = 1;
    if (m <= 0) return;
    p_dgemv("T", &n, &m, &alpha, A, &lda, x, &one, &beta, y, &one);
}
/* A(m,n) += alpha * x(m) y(n)^T, row-major. */
static void rm_ger(int m, int n, double alpha, const double *x, const double *y, double *A,
                   int lda) {
    const int one = 1;
    if (m <= 0 || n <= 0) return;
    p_dger(&n, &m, &alpha, y, &one, x, &one, A, &lda);
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* RFP position of (i,j), i <= j < n, TRANSR='N', UPLO='U' (src/rfp_utils.hpp:14-22). */
static inline size_t rfp_idx(size_t n, size_t i, size_t j) {
    const size_t k = n / 2;
    const size_t s = (n % 2 == 0) ? n + 1 : n;
    return (j >= k) ? (j - k) * s + i : i * s + (j + k + 1);
}
size_t kfo_rfp_index(size_t n, size_t i, size_t j) {
    if (i > j) { size_t t = i; i = j; j = t; }
    return rfp_idx(n, i, j);
}

/* ========================================================================= */
/* FCHL19 descriptor                                                          */
/* ========================================================================= */

/* compute_rep_size (src/fchl19_repr.cpp:24-31). */
long kfo_rep_size(long nelem, long nbasis2, long nbasis3, long nabasis) {
    return nelem * nbasis2 + nelem * (nelem + 1) / 2 * nbasis3 * nabasis;
}

/* Symmetric element-pair channel (src/fchl19_repr.cpp:47-52): for p <= q,
   the pairs (0,0),(0,1)..(0,N-1),(1,1).. are numbered consecutively. */
static inline int pair_channel(int nelem, int p, int q) {
    if (p > q) { int t = p; p = q; q = t; }
    return p * nelem - p * (p - 1) / 2 + (q - p);
}

typedef struct {
    int nelem;
    const int *elements;
    int nRs2, nRs3, nabasis;
    const double *Rs2, *Rs3;
    double eta2, eta3, zeta, rcut, acut, d2, d3, w3;
} fchl_params;

static int map_elements(const fchl_params *P, const int *z, int n, int *eidx) {
    for (int i = 0; i < n; ++i) {
        eidx[i] = -1;
        for (int e = 0; e < P->nelem; ++e)
            if (P->elements[e] == z[i]) { eidx[i] = e; break; }
        if (eidx[i] < 0) return -1;
    }
    return 0;
}

static inline double half_cos_cut(double r, double cut) {
    const double inv = cut > 0 ? 1.0 / cut : 0.0;
    return 0.5 * (cos(M_PI * r * inv) + 1.0);
}

/* Two-body log-normal radial term and its r-derivative, for one pair
   (src/fchl19_repr.cpp:149-195 values; :372-443 derivatives, restated from the
   closed form: g_k = f(r) r^-d2 exp(-(ln Rs_k - mu)^2 / 2s^2) / (s sqrt(2pi) Rs_k),
   s^2 = ln(1 + eta2/r^2), mu = ln r - s^2/2, f = (cos(pi r/rcut)+1)/2). */
static int two_body_terms(const fchl_params *P, double r, double *val, double *dval) {
    const double r2 = r * r;
    const double s2 = log1p(P->eta2 / (r2 > KFO_EPS ? r2 : KFO_EPS));
    const double s = sqrt(s2 > 0 ? s2 : 0.0);
    if (s < KFO_EPS) return 0;
    const double mu = log(r) - 0.5 * s2;
    const double f = half_cos_cut(r, P->rcut);
    const double inv_rc = P->rcut > 0 ? 1.0 / P->rcut : 0.0;
    const double df = -0.5 * M_PI * inv_rc * sin(M_PI * r * inv_rc);
    const double rp = pow(r, -P->d2);
    const double pref0 = rp / (s * KFO_SQRT_2PI);
    /* d(s^2)/dr, ds/dr, dmu/dr */
    const double ds2 = -2.0 * P->eta2 / (r * (r2 + P->eta2));
    const double ds = ds2 / (2.0 * s);
    const double dmu = 1.0 / r - 0.5 * ds2;
    for (int k = 0; k < P->nRs2; ++k) {
        const double L = log(P->Rs2[k]) - mu;
        const double v0 = pref0 * exp(-0.5 * L * L / s2) / P->Rs2[k];
        val[k] = f * v0;
        if (dval) {
            /* d/dr [f v0] = v0 (f dlog(v0)/dr + f'), no division by f */
            const double dlog0 = -P->d2 / r - ds / s + (L / s2) * dmu + (L * L / (s2 * s)) * ds;
            dval[k] = v0 * (f * dlog0 + df);
        }
    }
    return 1;
}

/* One molecule: representation (and optionally the Jacobian).
   rep  (n, rep_size);  grad (n, rep_size, n, 3) or NULL.
   Values follow generate_fchl_acsf (src/fchl19_repr.cpp:89-299) when grad is NULL
   and generate_fchl_acsf_and_gradients (:308-778) otherwise; the latter keeps
   only the first Fourier harmonic in the three-body angular term (:449,523-528). */
static int fchl19_one(const fchl_params *P, const double *R, const int *z, int n, double *rep,
                      double *grad) {
    const int nb2 = P->nRs2, nb3 = P->nRs3, na = P->nabasis;
    const long rs = kfo_rep_size(P->nelem, nb2, nb3, na);
    const int three_off = P->nelem * nb2;
    if (na % 2 != 0) return -2;
    for (int k = 0; k < nb2; ++k)
        if (P->Rs2[k] <= 0.0) return -3;
    int *e = (int *)malloc(sizeof(int) * (n > 0 ? n : 1));
    if (map_elements(P, z, n, e)) { free(e); return -1; }
    memset(rep, 0, sizeof(double) * (size_t)n * rs);
    if (grad) memset(grad, 0, sizeof(double) * (size_t)n * rs * n * 3);

    double *D = (double *)malloc(sizeof(double) * (size_t)n * n);
    for (int i = 0; i < n; ++i) {
        D[i * n + i] = 0.0;
        for (int j = i + 1; j < n; ++j) {
            const double dx = R[3 * j] - R[3 * i], dy = R[3 * j + 1] - R[3 * i + 1],
                         dz = R[3 * j + 2] - R[3 * i + 2];
            D[i * n + j] = D[j * n + i] = sqrt(dx * dx + dy * dy + dz * dz);
        }
    }
    double *v2 = (double *)malloc(sizeof(double) * nb2);
    double *dv2 = (double *)malloc(sizeof(double) * nb2);

    /* ---- two-body ---- */
    for (int i = 0; i < n; ++i) {
        for (int j = i + 1; j < n; ++j) {
            const double r = D[i * n + j];
            if (r > P->rcut) continue;
            if (!two_body_terms(P, r, v2, grad ? dv2 : NULL)) continue;
            const int ci = e[j] * nb2, cj = e[i] * nb2;
            for (int k = 0; k < nb2; ++k) {
                rep[(size_t)i * rs + ci + k] += v2[k];
                rep[(size_t)j * rs + cj + k] += v2[k];
            }
            if (!grad) continue;
            for (int t = 0; t < 3; ++t) {
                const double u = (R[3 * i + t] - R[3 * j + t]) / r; /* dr/dR_i */
                for (int k = 0; k < nb2; ++k) {
                    const double g = dv2[k] * u;
                    grad[(((size_t)i * rs + ci + k) * n + i) * 3 + t] += g;
                    grad[(((size_t)i * rs + ci + k) * n + j) * 3 + t] -= g;
                    grad[(((size_t)j * rs + cj + k) * n + i) * 3 + t] += g;
                    grad[(((size_t)j * rs + cj + k) * n + j) * 3 + t] -= g;
                }
            }
        }
    }

    /* ---- three-body ---- */
    const int nharm = grad ? 1 : na / 2;
    double *angw = (double *)malloc(sizeof(double) * (nharm > 0 ? nharm : 1));
    for (int h = 0; h < nharm; ++h) {
        const double t = P->zeta * (2 * h + 1);
        angw[h] = 2.0 * exp(-0.5 * t * t);
    }
    double *ang = (double *)calloc(na > 0 ? na : 1, sizeof(double));
    double *rad = (double *)malloc(sizeof(double) * nb3);
    const double inv_ac = P->acut > 0 ? 1.0 / P->acut : 0.0;

    for (int i = 0; i < n; ++i) {
        const double *Ri = R + 3 * i;
        for (int j = 0; j + 1 < n; ++j) {
            if (j == i) continue;
            const double rij = D[i * n + j];
            if (rij > P->acut) continue;
            const double *Rj = R + 3 * j;
            for (int k = j + 1; k < n; ++k) {
                if (k == i) continue;
                const double rik = D[i * n + k];
                if (rik > P->acut) continue;
                const double *Rk = R + 3 * k;
                const double rjk = D[j * n + k];
                double u[3], v[3], w[3]; /* u = Rj-Ri, v = Rk-Ri, w = Rk-Rj */
                for (int t = 0; t < 3; ++t) {
                    u[t] = Rj[t] - Ri[t];
                    v[t] = Rk[t] - Ri[t];
                    w[t] = Rk[t] - Rj[t];
                }
                const double uv = u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
                const double uw = u[0] * w[0] + u[1] * w[1] + u[2] * w[2];
                const double vw = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
                double ci = uv / (rij * rik);
                ci = ci > 1.0 ? 1.0 : (ci < -1.0 ? -1.0 : ci);
                const double cj = -uw / (rij * rjk); /* angle at j between (Ri-Rj),(Rk-Rj) */
                const double ck = vw / (rik * rjk);  /* angle at k between (Ri-Rk),(Rj-Rk) */
                double prod = rij * rik * rjk;
                const double P3 = pow(prod > KFO_EPS ? prod : KFO_EPS, -P->d3);
                const double atm = (1.0 + 3.0 * ci * cj * ck) * P->w3 * P3;
                const double si = sqrt(1.0 - ci * ci > 0 ? 1.0 - ci * ci : 0.0);
                /* angular basis: cos/sin of odd multiples of theta_i (Chebyshev) */
                {
                    double c_prev = 1.0, s_prev = 0.0, c_cur = ci, s_cur = si;
                    for (int h = 0, o = 1; h < nharm; ++h) {
                        const int target = 2 * h + 1;
                        while (o < target) {
                            const double cn = 2.0 * ci * c_cur - c_prev;
                            const double sn = 2.0 * ci * s_cur - s_prev;
                            c_prev = c_cur; s_prev = s_cur; c_cur = cn; s_cur = sn; ++o;
                        }
                        ang[2 * h] = angw[h] * c_cur;
                        ang[2 * h + 1] = angw[h] * s_cur;
                    }
                }
                const double fij = half_cos_cut(rij, P->acut), fik = half_cos_cut(rik, P->acut);
                const double F = fij * fik;
                const double rbar = 0.5 * (rij + rik);
                for (int l = 0; l < nb3; ++l) {
                    const double dr = rbar - P->Rs3[l];
                    rad[l] = exp(-P->eta3 * dr * dr);
                }
                const size_t base = (size_t)i * rs + three_off +
                                    (size_t)pair_channel(P->nelem, e[j], e[k]) * nb3 * na;
                for (int l = 0; l < nb3; ++l) {
                    const double sc = rad[l] * F * atm;
                    for (int t = 0; t < na; ++t) rep[base + (size_t)l * na + t] += ang[t] * sc;
                }
                if (!grad) continue;

                /* gradients of the four factors w.r.t. (Ri, Rj, Rk), 9 components */
                double g_ci[9], g_cj[9], g_ck[9], g_rij[9], g_rik[9], g_rjk[9];
                for (int t = 0; t < 3; ++t) {
                    /* cos at i: a = u, b = v */
                    const double dqj = v[t] / (rij * rik) - ci * u[t] / (rij * rij);
                    const double dqk = u[t] / (rij * rik) - ci * v[t] / (rik * rik);
                    g_ci[3 + t] = dqj; g_ci[6 + t] = dqk; g_ci[t] = -(dqj + dqk);
                    /* cos at j: a = Ri-Rj = -u, b = Rk-Rj = w */
                    const double dji = w[t] / (rij * rjk) + cj * u[t] / (rij * rij);
                    const double djk = -u[t] / (rij * rjk) - cj * w[t] / (rjk * rjk);
                    g_cj[t] = dji; g_cj[6 + t] = djk; g_cj[3 + t] = -(dji + djk);
                    /* cos at k: a = Ri-Rk = -v, b = Rj-Rk = -w */
                    const double dki = -w[t] / (rik * rjk) + ck * v[t] / (rik * rik);
                    const double dkj = -v[t] / (rik * rjk) + ck * w[t] / (rjk * rjk);
                    g_ck[t] = dki; g_ck[3 + t] = dkj; g_ck[6 + t] = -(dki + dkj);
                    /* distances */
                    g_rij[t] = -u[t] / rij; g_rij[3 + t] = u[t] / rij; g_rij[6 + t] = 0.0;
                    g_rik[t] = -v[t] / rik; g_rik[3 + t] = 0.0; g_rik[6 + t] = v[t] / rik;
                    g_rjk[t] = 0.0; g_rjk[3 + t] = -w[t] / rjk; g_rjk[6 + t] = w[t] / rjk;
                }
                const double dfij = -0.5 * M_PI * inv_ac * sin(M_PI * rij * inv_ac);
                const double dfik = -0.5 * M_PI * inv_ac * sin(M_PI * rik * inv_ac);
                double gF[9], gA[9], gang0[9], gang1[9], grbar[9];
                const double sden = sqrt(fmax(1e-10, rij * rij * rik * rik - uv * uv)) /
                                    (rij * rik); /* = sin(theta_i), floored */
                for (int c = 0; c < 9; ++c) {
                    gF[c] = dfij * fik * g_rij[c] + fij * dfik * g_rik[c];
                    gA[c] = P->w3 * P3 *
                            (3.0 * (g_ci[c] * cj * ck + ci * g_cj[c] * ck + ci * cj * g_ck[c]) -
                             (1.0 + 3.0 * ci * cj * ck) * P->d3 *
                                 (g_rij[c] / rij + g_rik[c] / rik + g_rjk[c] / rjk));
                    gang0[c] = angw[0] * g_ci[c];
                    gang1[c] = -angw[0] * ci * g_ci[c] / sden;
                    grbar[c] = 0.5 * (g_rij[c] + g_rik[c]);
                }
                const int atoms[3] = {i, j, k};
                for (int l = 0; l < nb3; ++l) {
                    const double dr = rbar - P->Rs3[l];
                    const double drad = -2.0 * P->eta3 * dr * rad[l];
                    for (int t = 0; t < na; ++t) {
                        const double a_t = ang[t];
                        const double *ga = (t == 0) ? gang0 : (t == 1 ? gang1 : NULL);
                        const size_t frow = (base + (size_t)l * na + t) * n;
                        for (int c = 0; c < 9; ++c) {
                            double g = a_t * (drad * grbar[c] * F * atm + rad[l] * gF[c] * atm +
                                              rad[l] * F * gA[c]);
                            if (ga) g += ga[c] * rad[l] * F * atm;
                            grad[(frow + atoms[c / 3]) * 3 + (c % 3)] += g;
                        }
                    }
                }
            }
        }
    }
    free(ang); free(rad); free(angw); free(v2); free(dv2); free(D); free(e);
    return 0;
}

/* Single molecule entry points.  Rs2/Rs3 grids and w3 = sqrt(eta3/pi)*weight are
   built by the caller exactly as src/fchl19_repr_bindings.cpp:72-84,125-144 do. */
int kfo_fchl19(const double *coords, const int *z, int n, const int *elements, int nelem,
               const double *Rs2, int nRs2, const double *Rs3, int nRs3, int nabasis,
               double eta2, double eta3, double zeta, double rcut, double acut, double d2,
               double d3, double w3, double *rep, double *grad) {
    fchl_params P = {nelem, elements, nRs2, nRs3, nabasis, Rs2, Rs3,
                     eta2,  eta3,     zeta, rcut, acut,    d2,  d3,  w3};
    return fchl19_one(&P, coords, z, n, rep, grad);
}

/* Batched over a padded set: coords (nm, max_atoms, 3), Q (nm, max_atoms), N (nm).
   X (nm, max_atoms, rep) and dX (nm, max_atoms, rep, 3*max_atoms) or NULL, in the
   padded layout models/representations.py:126-145 builds.  OpenMP over molecules. */
int kfo_fchl19_batch(const double *coords, const int *Q, const int *N, int nm, int max_atoms,
                     const int *elements, int nelem, const double *Rs2, int nRs2,
                     const double *Rs3, int nRs3, int nabasis, double eta2, double eta3,
                     double zeta, double rcut, double acut, double d2, double d3, double w3,
                     double *X, double *dX) {
    fchl_params P = {nelem, elements, nRs2, nRs3, nabasis, Rs2, Rs3,
                     eta2,  eta3,     zeta, rcut, acut,    d2,  d3,  w3};
    const long rs = kfo_rep_size(nelem, nRs2, nRs3, nabasis);
    const size_t ldx = (size_t)3 * max_atoms;
    memset(X, 0, sizeof(double) * (size_t)nm * max_atoms * rs);
    if (dX) memset(dX, 0, sizeof(double) * (size_t)nm * max_atoms * rs * ldx);
    int err = 0;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < nm; ++m) {
        const int n = clampi(N[m], 0, max_atoms);
        if (n == 0) continue;
        double *rep = (double *)malloc(sizeof(double) * (size_t)n * rs);
        double *g = dX ? (double *)malloc(sizeof(double) * (size_t)n * rs * n * 3) : NULL;
        int rc = fchl19_one(&P, coords + (size_t)m * max_atoms * 3, Q + (size_t)m * max_atoms,
                            n, rep, g);
        if (rc) {
#pragma omp atomic write
            err = rc;
        } else {
            memcpy(X + (size_t)m * max_atoms * rs, rep, sizeof(double) * (size_t)n * rs);
            if (dX)
                for (int i = 0; i < n; ++i)
                    for (long f = 0; f < rs; ++f)
                        memcpy(dX + (((size_t)m * max_atoms + i) * rs + f) * ldx,
                               g + ((size_t)i * rs + f) * n * 3, sizeof(double) * n * 3);
        }
        free(rep);
        free(g);
    }
    return err;
}

/* ========================================================================= */
/* Local Gaussian kernels                                                     */
/* ========================================================================= */

/* Rows of one label, packed in (molecule, atom) order, with per-molecule ranges
   (src/local_kernels.cpp:59-149,325-378). */
typedef struct {
    int R;
    double *A;     /* (R, rep) */
    double *n2;    /* (R) squared norms */
    int *mol_lo;   /* (nm) first packed row of molecule */
    int *mol_hi;   /* (nm) one past last */
} packed_label;

static void pack_label(int label, const double *x, const int *q, const int *nn, int nm,
                       int max_atoms, int rep, packed_label *pk) {
    int R = 0;
    for (int a = 0; a < nm; ++a) {
        const int na = clampi(nn[a], 0, max_atoms);
        for (int j = 0; j < na; ++j) R += q[(size_t)a * max_atoms + j] == label;
    }
    pk->R = R;
    pk->A = (double *)malloc(sizeof(double) * ((size_t)R * rep + 1));
    pk->n2 = (double *)malloc(sizeof(double) * (R + 1));
    pk->mol_lo = (int *)malloc(sizeof(int) * (nm + 1));
    pk->mol_hi = (int *)malloc(sizeof(int) * (nm + 1));
    int r = 0;
    for (int a = 0; a < nm; ++a) {
        const int na = clampi(nn[a], 0, max_atoms);
        pk->mol_lo[a] = r;
        for (int j = 0; j < na; ++j) {
            if (q[(size_t)a * max_atoms + j] != label) continue;
            const double *src = x + ((size_t)a * max_atoms + j) * rep;
            double s = 0.0;
            for (int k = 0; k < rep; ++k) {
                pk->A[(size_t)r * rep + k] = src[k];
                s += src[k] * src[k];
            }
            pk->n2[r] = s;
            ++r;
        }
        pk->mol_hi[a] = r;
    }
}
static void free_label(packed_label *pk) {
    free(pk->A); free(pk->n2); free(pk->mol_lo); free(pk->mol_hi);
}

static int collect_labels(const int *q1, const int *n1, int nm1, int ma1, const int *q2,
                          const int *n2, int nm2, int ma2, int **out) {
    int cap = 16, cnt = 0;
    int *L = (int *)malloc(sizeof(int) * cap);
    for (int s = 0; s < 2; ++s) {
        const int *q = s ? q2 : q1, *nn = s ? n2 : n1;
        const int nm = s ? nm2 : nm1, ma = s ? ma2 : ma1;
        if (!q) continue;
        for (int a = 0; a < nm; ++a) {
            const int na = clampi(nn[a], 0, ma);
            for (int j = 0; j < na; ++j) {
                const int lb = q[(size_t)a * ma + j];
                int found = 0;
                for (int t = 0; t < cnt; ++t)
                    if (L[t] == lb) { found = 1; break; }
                if (!found) {
                    if (cnt == cap) { cap *= 2; L = (int *)realloc(L, sizeof(int) * cap); }
                    L[cnt++] = lb;
                }
            }
        }
    }
    for (int i = 1; i < cnt; ++i) /* insertion sort */
        for (int j = i; j > 0 && L[j - 1] > L[j]; --j) { int t = L[j]; L[j] = L[j - 1]; L[j - 1] = t; }
    *out = L;
    return cnt;
}

/* Row-tile edge of the BLAS tiles (8192 in the reference); settable so tests can
   exercise the multi-tile bookkeeping on small inputs. */
static int kfo_tile = 8192;
void kfo_set_tile(int t) { kfo_tile = t > 0 ? t : 8192; }

/* Symmetric scalar kernel, K[a,b] for a <= b, written through `put`
   (src/local_kernels.cpp:403-577): per label, pack rows; for each 8192-row tile
   pair run DSYRK (diagonal) or DGEMM with alpha=-2 (off-diagonal) into a scratch
   tile, then an OpenMP exp-sum over molecule pairs; cross-tile self pairs count
   twice.  Molecule rows are contiguous, so each molecule's bucket for a tile is a
   row range [max(lo,i0), min(hi,i0+B)). */
static int symm_kernel_core(const double *x, const int *q, const int *nn, int nm, int max_atoms,
                            int rep, double sigma, double *out, int rfp) {
    if (!(sigma > 0.0) || !isfinite(sigma)) return -4;
    const double c = -1.0 / (2.0 * sigma * sigma);
    int *labels;
    const int nl = collect_labels(q, nn, nm, max_atoms, NULL, NULL, 0, 0, &labels);
    const int B = kfo_tile;
    for (int il = 0; il < nl; ++il) {
        packed_label pk;
        pack_label(labels[il], x, q, nn, nm, max_atoms, rep, &pk);
        const int R = pk.R;
        if (R == 0) { free_label(&pk); continue; }
        const int tb = R < B ? R : B;
        double *C = (double *)malloc(sizeof(double) * (size_t)tb * tb);
        for (int i0 = 0; i0 < R; i0 += B) {
            const int ib = (R - i0) < B ? (R - i0) : B;
            for (int j0 = i0; j0 < R; j0 += B) {
                const int jb = (R - j0) < B ? (R - j0) : B;
                const int diag = (j0 == i0);
                if (diag)
                    rm_syrk_upper(ib, rep, -2.0, pk.A + (size_t)i0 * rep, rep, C, ib);
                else
                    rm_gemm_nt(ib, jb, rep, -2.0, pk.A + (size_t)i0 * rep, rep,
                               pk.A + (size_t)j0 * rep, rep, 0.0, C, jb);
                const int ldc = diag ? ib : jb;
#pragma omp parallel for schedule(guided)
                for (int a = 0; a < nm; ++a) {
                    const int alo = pk.mol_lo[a] > i0 ? pk.mol_lo[a] : i0;
                    const int ahi = pk.mol_hi[a] < i0 + ib ? pk.mol_hi[a] : i0 + ib;
                    if (alo >= ahi) continue;
                    for (int b = a; b < nm; ++b) {
                        const int blo = pk.mol_lo[b] > j0 ? pk.mol_lo[b] : j0;
                        const int bhi = pk.mol_hi[b] < j0 + jb ? pk.mol_hi[b] : j0 + jb;
                        if (blo >= bhi) {
                            if (pk.mol_lo[b] >= j0 + jb) break;
                            continue;
                        }
                        double kab = 0.0;
                        for (int gi = alo; gi < ahi; ++gi) {
                            const int li = gi - i0;
                            for (int gj = blo; gj < bhi; ++gj) {
                                const int lj = gj - j0;
                                double g;
                                if (diag) {
                                    const int r0 = li <= lj ? li : lj, c0 = li <= lj ? lj : li;
                                    g = C[(size_t)r0 * ldc + c0];
                                } else {
                                    g = C[(size_t)li * ldc + lj];
                                }
                                kab += exp((pk.n2[gi] + pk.n2[gj] + g) * c);
                            }
                        }
                        if (!diag && a == b) kab *= 2.0;
                        if (rfp) {
                            out[rfp_idx(nm, a, b)] += kab;
                        } else {
                            out[(size_t)a * nm + b] += kab;
                        }
                    }
                }
            }
        }
        free(C);
        free_label(&pk);
    }
    free(labels);
    if (!rfp)
        for (int a = 0; a < nm; ++a)
            for (int b = a + 1; b < nm; ++b) out[(size_t)b * nm + a] = out[(size_t)a * nm + b];
    return 0;
}

/* kernel_gaussian_symm_rfp (src/local_kernels.cpp:403-577). */
int kfo_kernel_symm_rfp(const double *x, const int *q, const int *nn, int nm, int max_atoms,
                        int rep, double sigma, double *arf) {
    memset(arf, 0, sizeof(double) * (size_t)nm * (nm + 1) / 2);
    return symm_kernel_core(x, q, nn, nm, max_atoms, rep, sigma, arf, 1);
}

/* kernel_gaussian_symm (src/local_kernels.cpp:584-758): dense, both triangles. */
int kfo_kernel_symm(const double *x, const int *q, const int *nn, int nm, int max_atoms, int rep,
                    double sigma, double *K) {
    memset(K, 0, sizeof(double) * (size_t)nm * nm);
    return symm_kernel_core(x, q, nn, nm, max_atoms, rep, sigma, K, 0);
}

/* kernel_gaussian (rectangular, src/local_kernels.cpp:155-298): per label,
   DGEMM alpha=-2 on 8192-row tiles, then exp-sum per molecule pair. */
int kfo_kernel_rect(const double *x1, const double *x2, const int *q1, const int *q2,
                    const int *n1, const int *n2, int nm1, int nm2, int ma1, int ma2, int rep,
                    double sigma, double *K) {
    if (!(sigma > 0.0) || !isfinite(sigma)) return -4;
    memset(K, 0, sizeof(double) * (size_t)nm1 * nm2);
    const double c = -1.0 / (2.0 * sigma * sigma);
    int *labels;
    const int nl = collect_labels(q1, n1, nm1, ma1, q2, n2, nm2, ma2, &labels);
    const int B = kfo_tile;
    for (int il = 0; il < nl; ++il) {
        packed_label p1, p2;
        pack_label(labels[il], x1, q1, n1, nm1, ma1, rep, &p1);
        pack_label(labels[il], x2, q2, n2, nm2, ma2, rep, &p2);
        if (p1.R && p2.R) {
            const int tb1 = p1.R < B ? p1.R : B, tb2 = p2.R < B ? p2.R : B;
            double *C = (double *)malloc(sizeof(double) * (size_t)tb1 * tb2);
            for (int i0 = 0; i0 < p1.R; i0 += B) {
                const int ib = (p1.R - i0) < B ? (p1.R - i0) : B;
                for (int j0 = 0; j0 < p2.R; j0 += B) {
                    const int jb = (p2.R - j0) < B ? (p2.R - j0) : B;
                    rm_gemm_nt(ib, jb, rep, -2.0, p1.A + (size_t)i0 * rep, rep,
                               p2.A + (size_t)j0 * rep, rep, 0.0, C, jb);
#pragma omp parallel for schedule(guided)
                    for (int a = 0; a < nm1; ++a) {
                        const int alo = p1.mol_lo[a] > i0 ? p1.mol_lo[a] : i0;
                        const int ahi = p1.mol_hi[a] < i0 + ib ? p1.mol_hi[a] : i0 + ib;
                        if (alo >= ahi) continue;
                        for (int b = 0; b < nm2; ++b) {
                            const int blo = p2.mol_lo[b] > j0 ? p2.mol_lo[b] : j0;
                            const int bhi = p2.mol_hi[b] < j0 + jb ? p2.mol_hi[b] : j0 + jb;
                            if (blo >= bhi) continue;
                            double kab = 0.0;
                            for (int gi = alo; gi < ahi; ++gi)
                                for (int gj = blo; gj < bhi; ++gj)
                                    kab += exp((p1.n2[gi] + p2.n2[gj] +
                                                C[(size_t)(gi - i0) * jb + (gj - j0)]) * c);
                            K[(size_t)a * nm2 + b] += kab;
                        }
                    }
                }
            }
            free(C);
        }
        free_label(&p1);
        free_label(&p2);
    }
    free(labels);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Derivative kernels.  For one molecule pair (a from set 1, b from set 2) the
   E+F block is, summed over label-matched atom pairs (i,j), d = x_ai - x_bj,
   e = exp(-|d|^2 / 2 sigma^2):
     EE  = e
     EF  = +(e/s^2) J_bj^T d          (dK/dR_b, 1 x 3nb)
     FE  = -(e/s^2) J_ai^T d          (dK/dR_a, 3na x 1)
     FF  = J_ai^T [(e/s^2) I - (e/s^4) d d^T] J_bj   (3na x 3nb)
   following the reference's per-pair DGEMV/DGER + static S_sum DGEMM scheme
   (src/local_kernels.cpp:1171-1419, 2511-2815).  Any of the outputs may be NULL. */
/* ------------------------------------------------------------------------- */
typedef struct {
    const double *x, *dx;
    const int *q, *n;
    int nm, ma, rep;
    int *offs; /* 3*sum of clamped n over preceding molecules */
} side_t;

static void side_init(side_t *s, const double *x, const double *dx, const int *q, const int *n,
                      int nm, int ma, int rep) {
    s->x = x; s->dx = dx; s->q = q; s->n = n; s->nm = nm; s->ma = ma; s->rep = rep;
    s->offs = (int *)malloc(sizeof(int) * (nm + 1));
    int acc = 0;
    for (int m = 0; m < nm; ++m) { s->offs[m] = acc; acc += 3 * clampi(n[m], 0, ma); }
    s->offs[nm] = acc;
}

typedef struct {
    double *d, *Dw, *W, *V, *S, *tmp;
} pair_ws;

static void ws_alloc(pair_ws *w, int rep, int ncmax) {
    w->d = (double *)malloc(sizeof(double) * rep);
    w->Dw = (double *)malloc(sizeof(double) * rep);
    w->W = (double *)malloc(sizeof(double) * (ncmax + 1));
    w->V = (double *)malloc(sizeof(double) * (ncmax + 1));
    w->S = (double *)malloc(sizeof(double) * (size_t)rep * (ncmax + 1));
    w->tmp = (double *)malloc(sizeof(double) * (ncmax + 1));
}
static void ws_free(pair_ws *w) {
    free(w->d); free(w->Dw); free(w->W); free(w->V); free(w->S); free(w->tmp);
}

/* FF has row stride ldff; EF/FE are contiguous vectors. */
static void pair_block(const side_t *A, int a, const side_t *B, int b, double sigma,
                       double *EE, double *EF, double *FE, double *FF, int ldff, pair_ws *w) {
    const int rep = A->rep;
    const int na = clampi(A->n[a], 0, A->ma), nb = clampi(B->n[b], 0, B->ma);
    const int nca = 3 * na, ncb = 3 * nb, lda = 3 * A->ma, ldb = 3 * B->ma;
    const double s2 = sigma * sigma, s4 = s2 * s2, c = -1.0 / (2.0 * s2);
    for (int j = 0; j < nb; ++j) {
        const int lb = B->q[(size_t)b * B->ma + j];
        const double *xb = B->x + ((size_t)b * B->ma + j) * rep;
        const double *Jb = B->dx ? B->dx + ((size_t)b * B->ma + j) * rep * ldb : NULL;
        int any = 0;
        for (int i = 0; i < na; ++i) any |= (A->q[(size_t)a * A->ma + i] == lb);
        if (!any) continue;
        if (FF) memset(w->S, 0, sizeof(double) * (size_t)rep * nca);
        if (EF) memset(w->Dw, 0, sizeof(double) * rep);
        for (int i = 0; i < na; ++i) {
            if (A->q[(size_t)a * A->ma + i] != lb) continue;
            const double *xa = A->x + ((size_t)a * A->ma + i) * rep;
            const double *Ja = A->dx ? A->dx + ((size_t)a * A->ma + i) * rep * lda : NULL;
            double l2 = 0.0;
            for (int k = 0; k < rep; ++k) { w->d[k] = xa[k] - xb[k]; l2 += w->d[k] * w->d[k]; }
            const double e = exp(l2 * c);
            if (EE) *EE += e;
            if (EF) for (int k = 0; k < rep; ++k) w->Dw[k] += (e / s2) * w->d[k];
            if (FE) rm_gemv_t(rep, nca, -e / s2, Ja, lda, w->d, 1.0, FE);
            if (FF) {
                for (int k = 0; k < rep; ++k) {
                    double *srow = w->S + (size_t)k * nca;
                    const double *jrow = Ja + (size_t)k * lda;
                    for (int t = 0; t < nca; ++t) srow[t] += (e / s2) * jrow[t];
                }
                const double sc = sqrt(e / s4);
                for (int k = 0; k < rep; ++k) w->d[k] *= sc;
                rm_gemv_t(rep, ncb, 1.0, Jb, ldb, w->d, 0.0, w->W);
                rm_gemv_t(rep, nca, -1.0, Ja, lda, w->d, 0.0, w->V);
                rm_ger(nca, ncb, 1.0, w->V, w->W, FF, ldff);
            }
        }
        if (EF) rm_gemv_t(rep, ncb, 1.0, Jb, ldb, w->Dw, 1.0, EF);
        if (FF) rm_gemm_tn(nca, ncb, rep, 1.0, w->S, nca, Jb, ldb, 1.0, FF, ldff);
    }
}

static int check_sigma(double sigma) { return (sigma > 0.0 && isfinite(sigma)) ? 0 : -4; }

/* kernel_gaussian_jacobian (src/local_kernels.cpp:791-974): (nm1, naq2), +dK/dR2. */
int kfo_jacobian(const double *x1, const double *x2, const double *dx2, const int *q1,
                 const int *q2, const int *n1, const int *n2, int nm1, int nm2, int ma1, int ma2,
                 int rep, double sigma, double *K) {
    if (check_sigma(sigma)) return -4;
    side_t A, B;
    side_init(&A, x1, NULL, q1, n1, nm1, ma1, rep);
    side_init(&B, x2, dx2, q2, n2, nm2, ma2, rep);
    const int naq2 = B.offs[nm2];
    memset(K, 0, sizeof(double) * (size_t)nm1 * naq2);
#pragma omp parallel
    {
        pair_ws w;
        ws_alloc(&w, rep, 3 * (ma1 > ma2 ? ma1 : ma2));
#pragma omp for schedule(dynamic)
        for (int a = 0; a < nm1; ++a)
            for (int b = 0; b < nm2; ++b)
                pair_block(&A, a, &B, b, sigma, NULL, K + (size_t)a * naq2 + B.offs[b], NULL, NULL,
                           0, &w);
        ws_free(&w);
    }
    free(A.offs); free(B.offs);
    return 0;
}

/* kernel_gaussian_jacobian_t (src/local_kernels.cpp:991-1166): (naq1, nm2),
   = +(e/s^2) J1^T d = -dK/dR1. */
int kfo_jacobian_t(const double *x1, const double *x2, const double *dx1, const int *q1,
                   const int *q2, const int *n1, const int *n2, int nm1, int nm2, int ma1,
                   int ma2, int rep, double sigma, double *K) {
    if (check_sigma(sigma)) return -4;
    side_t A, B;
    side_init(&A, x1, dx1, q1, n1, nm1, ma1, rep);
    side_init(&B, x2, NULL, q2, n2, nm2, ma2, rep);
    const int naq1 = A.offs[nm1];
    memset(K, 0, sizeof(double) * (size_t)naq1 * nm2);
#pragma omp parallel
    {
        pair_ws w;
        ws_alloc(&w, rep, 3 * (ma1 > ma2 ? ma1 : ma2));
        double *fe = (double *)malloc(sizeof(double) * (3 * ma1 + 1));
#pragma omp for schedule(dynamic)
        for (int a = 0; a < nm1; ++a) {
            const int nca = A.offs[a + 1] - A.offs[a];
            for (int b = 0; b < nm2; ++b) {
                memset(fe, 0, sizeof(double) * (nca + 1));
                pair_block(&A, a, &B, b, sigma, NULL, NULL, fe, NULL, 0, &w);
                for (int r = 0; r < nca; ++r) K[(size_t)(A.offs[a] + r) * nm2 + b] = -fe[r];
            }
        }
        free(fe);
        ws_free(&w);
    }
    free(A.offs); free(B.offs);
    return 0;
}

/* Generic E+F assembly into a dense row-major (R1 x C2) matrix:
   rows: [energies of set 1 if with_e1][forces of set 1]
   cols: [energies of set 2 if with_e2][forces of set 2]. */
static int full_dense(const double *x1, const double *dx1, const double *x2, const double *dx2,
                      const int *q1, const int *q2, const int *n1, const int *n2, int nm1,
                      int nm2, int ma1, int ma2, int rep, double sigma, int with_e, double *K) {
    if (check_sigma(sigma)) return -4;
    side_t A, B;
    side_init(&A, x1, dx1, q1, n1, nm1, ma1, rep);
    side_init(&B, x2, dx2, q2, n2, nm2, ma2, rep);
    const int e1 = with_e ? nm1 : 0, e2 = with_e ? nm2 : 0;
    const size_t ld = (size_t)e2 + B.offs[nm2];
    const size_t rows = (size_t)e1 + A.offs[nm1];
    memset(K, 0, sizeof(double) * rows * ld);
#pragma omp parallel
    {
        pair_ws w;
        ws_alloc(&w, rep, 3 * (ma1 > ma2 ? ma1 : ma2));
#pragma omp for schedule(dynamic)
        for (int a = 0; a < nm1; ++a)
            for (int b = 0; b < nm2; ++b) {
                const int nca = A.offs[a + 1] - A.offs[a];
                double ee = 0.0;
                double *fe = (double *)calloc(nca + 1, sizeof(double));
                double *ef = with_e ? K + (size_t)a * ld + e2 + B.offs[b] : NULL;
                double *ff = K + ((size_t)e1 + A.offs[a]) * ld + e2 + B.offs[b];
                pair_block(&A, a, &B, b, sigma, with_e ? &ee : NULL, ef, with_e ? fe : NULL, ff,
                           (int)ld, &w);
                if (with_e) {
                    K[(size_t)a * ld + b] = ee;
                    for (int r = 0; r < nca; ++r) K[((size_t)e1 + A.offs[a] + r) * ld + b] = fe[r];
                }
                free(fe);
            }
        ws_free(&w);
    }
    free(A.offs); free(B.offs);
    return 0;
}

/* kernel_gaussian_hessian (src/local_kernels.cpp:1171-1419): (naq1, naq2). */
int kfo_hessian(const double *x1, const double *dx1, const double *x2, const double *dx2,
                const int *q1, const int *q2, const int *n1, const int *n2, int nm1, int nm2,
                int ma1, int ma2, int rep, double sigma, double *K) {
    return full_dense(x1, dx1, x2, dx2, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma, 0, K);
}

/* kernel_gaussian_full (src/local_kernels.cpp:1918-2195): (nm1+naq1, nm2+naq2). */
int kfo_full(const double *x1, const double *dx1, const double *x2, const double *dx2,
             const int *q1, const int *q2, const int *n1, const int *n2, int nm1, int nm2,
             int ma1, int ma2, int rep, double sigma, double *K) {
    return full_dense(x1, dx1, x2, dx2, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma, 1, K);
}

/* Dense symmetric (n x n) -> RFP (upper).  Used by the *_symm_rfp oracles. */
static void dense_to_rfp_upper(const double *K, size_t n, double *arf) {
    for (size_t i = 0; i < n; ++i)
        for (size_t j = i; j < n; ++j) arf[rfp_idx(n, i, j)] = K[i * n + j];
}

/* kernel_gaussian_hessian_symm / _rfp (src/local_kernels.cpp:1421-1900). */
int kfo_hessian_symm(const double *x, const double *dx, const int *q, const int *n, int nm,
                     int ma, int rep, double sigma, double *K) {
    return full_dense(x, dx, x, dx, q, q, n, n, nm, nm, ma, ma, rep, sigma, 0, K);
}
int kfo_hessian_symm_rfp(const double *x, const double *dx, const int *q, const int *n, int nm,
                         int ma, int rep, double sigma, double *arf) {
    int naq = 0;
    for (int m = 0; m < nm; ++m) naq += 3 * clampi(n[m], 0, ma);
    double *K = (double *)malloc(sizeof(double) * (size_t)naq * naq + 8);
    int rc = kfo_hessian_symm(x, dx, q, n, nm, ma, rep, sigma, K);
    if (!rc) dense_to_rfp_upper(K, naq, arf);
    free(K);
    return rc;
}

/* kernel_gaussian_full_symm / _rfp (src/local_kernels.cpp:2201-2815). */
int kfo_full_symm(const double *x, const double *dx, const int *q, const int *n, int nm, int ma,
                  int rep, double sigma, double *K) {
    return full_dense(x, dx, x, dx, q, q, n, n, nm, nm, ma, ma, rep, sigma, 1, K);
}

/* Upper-triangle RFP assembly pair by pair (a <= b) as the reference's
   kernel_gaussian_full_symm_rfp does (src/local_kernels.cpp:2511-2815). */
int kfo_full_symm_rfp(const double *x, const double *dx, const int *q, const int *n, int nm,
                      int ma, int rep, double sigma, double *arf) {
    if (check_sigma(sigma)) return -4;
    side_t S;
    side_init(&S, x, dx, q, n, nm, ma, rep);
    const size_t big = (size_t)nm + S.offs[nm];
    memset(arf, 0, sizeof(double) * big * (big + 1) / 2);
#pragma omp parallel
    {
        pair_ws w;
        ws_alloc(&w, rep, 3 * ma);
        const int ncm = 3 * ma;
        double *ff = (double *)malloc(sizeof(double) * (size_t)ncm * ncm + 8);
        double *ef = (double *)malloc(sizeof(double) * (ncm + 1));
        double *fe = (double *)malloc(sizeof(double) * (ncm + 1));
#pragma omp for schedule(dynamic)
        for (int b = 0; b < nm; ++b) {
            const int ncb = S.offs[b + 1] - S.offs[b];
            for (int a = 0; a <= b; ++a) {
                const int nca = S.offs[a + 1] - S.offs[a];
                double ee = 0.0;
                memset(ff, 0, sizeof(double) * (size_t)nca * ncb);
                memset(ef, 0, sizeof(double) * (ncb + 1));
                memset(fe, 0, sizeof(double) * (nca + 1));
                pair_block(&S, a, &S, b, sigma, &ee, ef, fe, ff, ncb, &w);
                /* rows/cols owned by this (a,b) are written by exactly one thread:
                   (a,b) EE, (a, F_b), (b, F_a) [a != b], (F_a, F_b). */
#pragma omp atomic
                arf[rfp_idx(big, a, b)] += ee;
                for (int r = 0; r < ncb; ++r) {
#pragma omp atomic
                    arf[rfp_idx(big, a, nm + S.offs[b] + r)] += ef[r];
                }
                if (a != b)
                    for (int r = 0; r < nca; ++r) {
#pragma omp atomic
                        arf[rfp_idx(big, b, nm + S.offs[a] + r)] += fe[r];
                    }
                for (int r = 0; r < nca; ++r)
                    for (int t = (a == b ? r : 0); t < ncb; ++t) {
                        const size_t i = nm + S.offs[a] + r, j = nm + S.offs[b] + t;
#pragma omp atomic
                        arf[rfp_idx(big, i, j)] += ff[(size_t)r * ncb + t];
                    }
            }
        }
        free(ff); free(ef); free(fe);
        ws_free(&w);
    }
    free(S.offs);
    return 0;
}

/* kernel_gaussian_local_compute_alpha_desc (src/local_kernels.cpp:2821-2893):
   alpha_desc[b,i,:] = dX[b,i,:,:3n_b] @ alpha[offs_b : offs_b+3n_b]. */
int kfo_alpha_desc(const double *dx, const int *q, const int *n, int nm, int ma, int rep,
                   const double *alpha, double *out) {
    (void)q;
    const int ld = 3 * ma;
    int *offs = (int *)malloc(sizeof(int) * (nm + 1));
    int acc = 0;
    for (int m = 0; m < nm; ++m) { offs[m] = acc; acc += 3 * clampi(n[m], 0, ma); }
    memset(out, 0, sizeof(double) * (size_t)nm * ma * rep);
#pragma omp parallel for schedule(static)
    for (int b = 0; b < nm; ++b) {
        const int nb = clampi(n[b], 0, ma);
        if (!nb) continue;
        rm_gemv_n(nb * rep, 3 * nb, 1.0, dx + (size_t)b * ma * rep * ld, ld, alpha + offs[b], 0.0,
                  out + (size_t)b * ma * rep);
    }
    free(offs);
    return 0;
}

/* kernel_gaussian_local_full_matvec (src/local_kernels.cpp:3263-3425) and its
   pieces: hessian_matvec (:2895-3145, F only, alpha_E = 0) and jacobian_t_matvec
   (:3155-3253, E from alpha_desc only).  Per query atom (a,i):
     E[a] += e * aE[b] + (e/s^2) (d . ad_bj)
     G    += (e/s^2) ad_bj - (e/s^4)(d . ad_bj) d + (e/s^2) aE[b] x_bj
     G    -= (sum (e/s^2) aE[b]) x_ai
     F[offs_a:] += J_ai^T G.
   mode bit 1: compute E, bit 2: compute F, bit 4: include the alpha_E terms. */
static int matvec_core(const double *x1, const double *dx1, const double *x2, const double *ad,
                       const double *aE, const int *q1, const int *q2, const int *n1,
                       const int *n2, int nm1, int nm2, int ma1, int ma2, int rep, double sigma,
                       double *E, double *F, int mode) {
    if (check_sigma(sigma)) return -4;
    const double s2 = sigma * sigma, c = -1.0 / (2.0 * s2);
    const int ld1 = 3 * ma1;
    int *offs = (int *)malloc(sizeof(int) * (nm1 + 1));
    int acc = 0;
    for (int m = 0; m < nm1; ++m) { offs[m] = acc; acc += 3 * clampi(n1[m], 0, ma1); }
    if (E) memset(E, 0, sizeof(double) * nm1);
    if (F) memset(F, 0, sizeof(double) * acc);
#pragma omp parallel
    {
        double *G = (double *)malloc(sizeof(double) * rep);
        double *d = (double *)malloc(sizeof(double) * rep);
#pragma omp for schedule(dynamic)
        for (int a = 0; a < nm1; ++a) {
            const int na = clampi(n1[a], 0, ma1);
            double Ea = 0.0;
            for (int i = 0; i < na; ++i) {
                const int lb = q1[(size_t)a * ma1 + i];
                const double *xa = x1 + ((size_t)a * ma1 + i) * rep;
                memset(G, 0, sizeof(double) * rep);
                double wE = 0.0;
                for (int b = 0; b < nm2; ++b) {
                    const int nb = clampi(n2[b], 0, ma2);
                    for (int j = 0; j < nb; ++j) {
                        if (q2[(size_t)b * ma2 + j] != lb) continue;
                        const double *xb = x2 + ((size_t)b * ma2 + j) * rep;
                        const double *adb = ad + ((size_t)b * ma2 + j) * rep;
                        double l2 = 0.0, inner = 0.0;
                        for (int k = 0; k < rep; ++k) {
                            d[k] = xa[k] - xb[k];
                            l2 += d[k] * d[k];
                            inner += d[k] * adb[k];
                        }
                        const double e = exp(l2 * c);
                        const double ed = e / s2, edd = -e / (s2 * s2);
                        const double ae = (mode & 4) ? aE[b] : 0.0;
                        Ea += e * ae + ed * inner;
                        if (mode & 2) {
                            wE += ed * ae;
                            for (int k = 0; k < rep; ++k)
                                G[k] += ed * adb[k] + edd * inner * d[k] + ed * ae * xb[k];
                        }
                    }
                }
                if (mode & 2) {
                    for (int k = 0; k < rep; ++k) G[k] -= wE * xa[k];
                    rm_gemv_t(rep, 3 * na, 1.0, dx1 + ((size_t)a * ma1 + i) * rep * ld1, ld1, G,
                              1.0, F + offs[a]);
                }
            }
            if (E) E[a] = Ea;
        }
        free(G);
        free(d);
    }
    free(offs);
    return 0;
}

int kfo_full_matvec(const double *x1, const double *dx1, const double *x2, const double *ad,
                    const double *aE, const int *q1, const int *q2, const int *n1, const int *n2,
                    int nm1, int nm2, int ma1, int ma2, int rep, double sigma, double *E,
                    double *F) {
    return matvec_core(x1, dx1, x2, ad, aE, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma, E, F,
                       1 | 2 | 4);
}
int kfo_hessian_matvec(const double *x1, const double *dx1, const double *x2, const double *ad,
                       const int *q1, const int *q2, const int *n1, const int *n2, int nm1,
                       int nm2, int ma1, int ma2, int rep, double sigma, double *F) {
    return matvec_core(x1, dx1, x2, ad, NULL, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma,
                       NULL, F, 2);
}
int kfo_jacobian_t_matvec(const double *x1, const double *x2, const double *ad, const int *q1,
                          const int *q2, const int *n1, const int *n2, int nm1, int nm2, int ma1,
                          int ma2, int rep, double sigma, double *E) {
    return matvec_core(x1, NULL, x2, ad, NULL, q1, q2, n1, n2, nm1, nm2, ma1, ma2, rep, sigma, E,
                       NULL, 1);
}

/* ========================================================================= */
/* kernelmath: RFP Cholesky solve (src/math.cpp:118-171)                      */
/* ========================================================================= */

/* Adds l2 on the RFP diagonal, then LAPACK dpftrf/dpftrs ('N','U').
   Overwrites arf with the factor.  Returns LAPACK info (0 ok). */
int kfo_cho_solve_rfp(double *arf, const double *y, int n, double l2, double *alpha) {
    memcpy(alpha, y, sizeof(double) * n);
    if (l2 != 0.0)
        for (int j = 0; j < n; ++j) arf[rfp_idx(n, j, j)] += l2;
    int info = 0;
    p_dpftrf("N", "U", &n, arf, &info);
    if (info) return info;
    const int one = 1;
    p_dpftrs("N", "U", &n, &one, arf, alpha, &n, &info);
    return info ? -1000 - info : 0;
}
