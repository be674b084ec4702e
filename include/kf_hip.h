/*
 * kf_hip.h -- C ABI of libkf_hip.so, the MI355X (gfx950) implementation of
 * kernelforge's FCHL19 local-kernel KRR hot path.
 *
 * Conventions
 *  - Every array argument is a DEVICE pointer (HBM), except the small parameter
 *    arrays marked "host" (element lists, basis grids).  Layouts are the
 *    reference's row-major ones:
 *        X   (nm, max_atoms, rep)                  float64
 *        dX  (nm, max_atoms, rep, 3*max_atoms)     float64
 *        Q   (nm, max_atoms) int32 labels, N (nm,) int32 atom counts
 *        RFP arrays: LAPACK Rectangular Full Packed, TRANSR='N', UPLO='U',
 *        length n(n+1)/2 (reference src/rfp_utils.hpp:14-22).
 *    Atom counts are clamped to [0, max_atoms] as the reference does
 *    (src/local_kernels.cpp:40).  naq = 3 * sum(clamped N).
 *  - `stream` is a hipStream_t (NULL = the null stream).  Calls are
 *    stream-ordered; entry points that need host-side metadata (labels, atom
 *    counts) copy it back and synchronise the stream.  Calls on one device share
 *    the library's device workspace, so the library serialises them: each entry
 *    point holds a per-device lock for its host-side body, and a call on a stream
 *    other than the previous call's first makes its stream wait for the previous
 *    call's work (an event).  Concurrent calls from several threads are therefore
 *    safe but do not run concurrently on the GPU.
 *  - Return value: KF_OK (0) on success, a negative KF_ERR_* code for invalid
 *    arguments or runtime failures (kf_last_error() has the message), or a
 *    positive factorisation `info` (leading minor that is not positive
 *    definite, 1-based) from the Cholesky entry points.
 *
 * Each entry point names the reference interface it replaces (file:line in
 * andersx/kernelforge).
 */
#ifndef KF_HIP_H
#define KF_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define KF_OK 0
#define KF_ERR_ARG (-1)     /* bad shape / size / pointer                 -> ValueError   */
#define KF_ERR_SIGMA (-2)   /* sigma <= 0 or not finite                    -> ValueError   */
#define KF_ERR_ELEMENT (-3) /* nuclear charge not in `elements`            -> RuntimeError */
#define KF_ERR_BASIS (-4)   /* Rs2 <= 0, odd nabasis, unsupported sizes    -> ValueError   */
#define KF_ERR_HIP (-5)     /* HIP runtime failure                         -> RuntimeError */
#define KF_ERR_NOMEM (-6)   /* device allocation failed                    -> MemoryError  */
#define KF_ERR_NOCONV (-7)  /* iterative solver did not converge           -> RuntimeError */

/* Arithmetic of a kernel / factorisation. */
#define KF_PREC_F64 0    /* FP64 throughout (the reference's precision)                   */
#define KF_PREC_BF16X3 1 /* BF16x3 split MFMA contraction, FP32 accumulate/exp, FP32 out  */
#define KF_PREC_F32 2    /* FP32 (factorisations)                                         */
#define KF_PREC_I8X6 3   /* FP64-accurate int8 split: each descriptor row as 6 int8 slices */
                         /* of 7 bits (per-row power-of-two scale), exact int32 MFMA slice */
                         /* products, FP64 norms/exp/sums/out; entries ~1e-11 relative     */

/* Library / device information. */
int kf_version(void);
const char *kf_last_error(void);
int kf_device_name(char *buf, int len);

/* Device-memory plumbing (no reference counterpart): for FFI callers that hold no
 * GPU allocator of their own (ctypes, cgo, JNI).  kf_memcpy kind: 0 host->device,
 * 1 device->host (both complete before returning), 2 device->device (stream-ordered). */
int kf_malloc(void **ptr, size_t bytes);
int kf_free(void *ptr);
int kf_memcpy(void *dst, const void *src, size_t bytes, int kind, void *stream);
int kf_memset_zero(void *dst, size_t bytes, void *stream);
int kf_stream_synchronize(void *stream);

/* Kernel timers (instrumentation; no reference counterpart -- the reference prints
 * ad-hoc phase timers, src/local_kernels.cpp:2564-2568).  When enabled, entry
 * points bracket their main kernel launch(es) with HIP events on the launch
 * stream; kf_timer_read synchronises on them and returns the summed milliseconds
 * and the number of bracketed calls for one timer id:
 *   0 fchl19, 1 local symmetric kernel, 2 local rectangular kernel,
 *   3 RFP potrf, 4 RFP potrs, 5 E+F kernels, 6 matvecs. */
int kf_timer_enable(int on);
int kf_timer_reset(void);
int kf_timer_read(int id, double *total_ms, int *count);

/* ------------------------------------------------------------------------ */
/* FCHL19 representation                                                    */
/* ------------------------------------------------------------------------ */

/* Representation (and optionally its coordinate Jacobian) of a padded batch of
 * molecules.  Replaces kf::fchl19::generate_fchl_acsf (src/fchl19_repr.cpp:89)
 * and generate_fchl_acsf_and_gradients (src/fchl19_repr.cpp:308), called per
 * molecule by models/representations.py:107-145.
 *   coords (nm, max_atoms, 3), Q (nm, max_atoms) nuclear charges, N (nm,)
 *   elements/Rs2/Rs3 are HOST arrays; w3 = sqrt(eta3/pi) * three_body_weight
 *   (src/fchl19_repr_bindings.cpp:87).
 *   X  (nm, max_atoms, rep) out, rows >= N[m] zeroed.
 *   dX (nm, max_atoms, rep, 3*max_atoms) out, or NULL for representation only.
 *   With dX != NULL the three-body angular term keeps only the first harmonic,
 *   as the reference's gradient path does (src/fchl19_repr.cpp:449,523-528). */
int kf_fchl19_f64(const double *coords, const int *Q, const int *N, int nm, int max_atoms,
                  const int *elements, int nelem, const double *Rs2, int nRs2, const double *Rs3,
                  int nRs3, int nabasis, double eta2, double eta3, double zeta, double rcut,
                  double acut, double two_body_decay, double three_body_decay, double w3,
                  double *X, double *dX, void *stream);

/* FCHL19v2 representation (and optionally its Jacobian) of a padded batch, with
 * selectable bases.  Replaces kf::fchl19v2::generate (src/fchl19v2_repr.cpp:2874) and
 * generate_and_gradients (:2919).  two_body_type / three_body_type index the reference's
 * TwoBodyType {log_normal, gaussian_r, gaussian_log_r, gaussian_r_no_pow, bessel} and
 * ThreeBodyType {odd_fourier_rbar, cosine_rbar, odd_fourier_split_r, cosine_split_r,
 * cosine_split_r_no_atm} (src/fchl19v2_repr.hpp:11-27).  elements / Rs2 / Rs3 /
 * Rs3_minus are HOST arrays; w3 = sqrt(eta3/pi) * three_body_weight; Rs3_minus is used
 * by the split bases only (nRs3_minus > 0 there).  Layouts as kf_fchl19_f64, with
 * rep = nelem nRs2 + nelem(nelem+1)/2 nRs3 [nRs3_minus] nabasis.  No limit on
 * max_atoms; up to 16 elements and 64 functions per radial grid. */
int kf_fchl19v2_f64(const double *coords, const int *Q, const int *N, int nm, int max_atoms,
                    const int *elements, int nelem, const double *Rs2, int nRs2,
                    const double *Rs3, int nRs3, const double *Rs3_minus, int nRs3_minus,
                    int nabasis, int two_body_type, int three_body_type, double eta2,
                    double eta3, double eta3_minus, double zeta, double rcut, double acut,
                    double two_body_decay, double three_body_decay, double w3, double *X,
                    double *dX, void *stream);

/* ------------------------------------------------------------------------ */
/* Local (element-matched) Gaussian kernels                                 */
/*   K[a,b] = sum_{i in a, j in b, Q_i == Q_j} exp(-|x_i - x_j|^2 / 2 sigma^2) */
/* ------------------------------------------------------------------------ */

/* Symmetric kernel written straight into RFP (upper).  Replaces
 * kf::fchl19::kernel_gaussian_symm_rfp (src/local_kernels.cpp:403).
 * precision KF_PREC_F64 or KF_PREC_I8X6: arf is double*; KF_PREC_BF16X3: arf is float*. */
int kf_local_kernel_symm_rfp(int precision, const double *X, const int *Q, const int *N, int nm,
                             int max_atoms, int rep, double sigma, void *arf, void *stream);

/* Symmetric kernel, dense (nm, nm) row-major, both triangles.  Replaces
 * kf::fchl19::kernel_gaussian_symm (src/local_kernels.cpp:584).  K is double*
 * (KF_PREC_F64, KF_PREC_I8X6) or float* (KF_PREC_BF16X3). */
int kf_local_kernel_symm(int precision, const double *X, const int *Q, const int *N, int nm,
                         int max_atoms, int rep, double sigma, void *K, void *stream);

/* Rectangular kernel (nm1, nm2) row-major.  Replaces kf::fchl19::kernel_gaussian
 * (src/local_kernels.cpp:155).  K is double* (F64, I8X6) or float* (BF16X3). */
int kf_local_kernel(int precision, const double *X1, const double *X2, const int *Q1,
                    const int *Q2, const int *N1, const int *N2, int nm1, int nm2,
                    int max_atoms1, int max_atoms2, int rep, double sigma, void *K, void *stream);

/* ------------------------------------------------------------------------ */
/* Energy+force forms.  dX arrays are (nm, max_atoms, rep, 3*max_atoms): the  */
/* Jacobian of each atom's descriptor with respect to its molecule's         */
/* coordinates; naq = 3 * sum(clamped N).  Any max_atoms: up to 32 atoms per */
/* molecule run the MFMA kernels, larger molecules a size-general kernel.    */
/* All arrays are device pointers.                                           */
/* ------------------------------------------------------------------------ */

/* (nm1, naq2) row-major, +dK/dR2.  Replaces kf::fchl19::kernel_gaussian_jacobian
 * (src/local_kernels.hpp:26-42, src/local_kernels.cpp:791-974). */
int kf_local_kernel_jacobian(const double *X1, const double *X2, const double *dX2,
                             const int *Q1, const int *Q2, const int *N1, const int *N2, int nm1,
                             int nm2, int max_atoms1, int max_atoms2, int rep, double sigma,
                             double *K, void *stream);

/* (naq1, nm2) row-major, -dK/dR1.  Replaces kernel_gaussian_jacobian_t
 * (src/local_kernels.hpp:44-56, src/local_kernels.cpp:991-1166). */
int kf_local_kernel_jacobian_t(const double *X1, const double *X2, const double *dX1,
                               const int *Q1, const int *Q2, const int *N1, const int *N2,
                               int nm1, int nm2, int max_atoms1, int max_atoms2, int rep,
                               double sigma, double *K, void *stream);

/* (naq1, naq2) row-major Hessian kernel.  Replaces kernel_gaussian_hessian
 * (src/local_kernels.hpp:58-72, src/local_kernels.cpp:1171-1419). */
int kf_local_kernel_hessian(const double *X1, const double *X2, const double *dX1,
                            const double *dX2, const int *Q1, const int *Q2, const int *N1,
                            const int *N2, int nm1, int nm2, int max_atoms1, int max_atoms2,
                            int rep, double sigma, double *K, void *stream);

/* (nm1+naq1, nm2+naq2) row-major [[EE, EF], [FE, FF]].  Replaces
 * kernel_gaussian_full (src/local_kernels.hpp:99-107, src/local_kernels.cpp:1918-2195). */
int kf_local_kernel_full(const double *X1, const double *X2, const double *dX1,
                         const double *dX2, const int *Q1, const int *Q2, const int *N1,
                         const int *N2, int nm1, int nm2, int max_atoms1, int max_atoms2,
                         int rep, double sigma, double *K, void *stream);

/* (naq, naq) dense symmetric Hessian kernel.  Replaces kernel_gaussian_hessian_symm
 * (src/local_kernels.hpp:74-83, src/local_kernels.cpp:1421-1662). */
int kf_local_kernel_hessian_symm(const double *X, const double *dX, const int *Q, const int *N,
                                 int nm, int max_atoms, int rep, double sigma, double *K,
                                 void *stream);

/* Hessian kernel in RFP (TRANSR='N', UPLO='U'), naq(naq+1)/2 doubles.  Replaces
 * kernel_gaussian_hessian_symm_rfp (src/local_kernels.hpp:85-97, :1664-1900). */
int kf_local_kernel_hessian_symm_rfp(const double *X, const double *dX, const int *Q,
                                     const int *N, int nm, int max_atoms, int rep,
                                     double sigma, double *arf, void *stream);

/* (nm+naq)^2 dense symmetric full kernel.  Replaces kernel_gaussian_full_symm
 * (src/local_kernels.hpp:109-115, src/local_kernels.cpp:2201-2505). */
int kf_local_kernel_full_symm(const double *X, const double *dX, const int *Q, const int *N,
                              int nm, int max_atoms, int rep, double sigma, double *K,
                              void *stream);

/* Full kernel in RFP (TRANSR='N', UPLO='U') of order BIG = nm + naq.  Replaces
 * kernel_gaussian_full_symm_rfp (src/local_kernels.hpp:117-121, :2511-2815). */
int kf_local_kernel_full_symm_rfp(const double *X, const double *dX, const int *Q,
                                  const int *N, int nm, int max_atoms, int rep, double sigma,
                                  double *arf, void *stream);

/* ------------------------------------------------------------------------ */
/* Matrix-free products with the energy+force kernel (the PCG operator).     */
/* ------------------------------------------------------------------------ */

/* alpha_desc[b,j,:] = dX[b,j,:,:3N_b] @ alpha[offs_b : offs_b + 3N_b], zero rows for
 * j >= N_b; out is (nm, max_atoms, rep).  Replaces
 * kernel_gaussian_local_compute_alpha_desc (src/local_kernels.hpp:141-169,
 * src/local_kernels.cpp:2821-2893).  Q is accepted for signature parity. */
int kf_local_alpha_desc(const double *dX, const int *Q, const int *N, int nm, int max_atoms,
                        int rep, const double *alpha, double *out, void *stream);

/* (E (nm1,), F (naq1,)) = K_full(set1, set2) @ [alpha_E; alpha_F] with alpha_desc from
 * kf_local_alpha_desc(dX2, .., alpha_F).  Replaces kernel_gaussian_local_full_matvec
 * (src/local_kernels.hpp:214-229, src/local_kernels.cpp:3263-3425). */
int kf_local_full_matvec(const double *X1, const double *dX1, const double *X2,
                         const double *alpha_desc, const double *alpha_E, const int *Q1,
                         const int *Q2, const int *N1, const int *N2, int nm1, int nm2,
                         int max_atoms1, int max_atoms2, int rep, double sigma, double *E,
                         double *F, void *stream);

/* F (naq1,) = Hessian kernel @ alpha_F.  Replaces kernel_gaussian_local_hessian_matvec
 * (src/local_kernels.hpp:171-191, src/local_kernels.cpp:2895-3145). */
int kf_local_hessian_matvec(const double *X1, const double *dX1, const double *X2,
                            const double *alpha_desc, const int *Q1, const int *Q2,
                            const int *N1, const int *N2, int nm1, int nm2, int max_atoms1,
                            int max_atoms2, int rep, double sigma, double *F, void *stream);

/* E (nm1,) = Jacobian kernel @ alpha_F.  Replaces kernel_gaussian_local_jacobian_t_matvec
 * (src/local_kernels.hpp:193-212, src/local_kernels.cpp:3155-3253). */
int kf_local_jacobian_t_matvec(const double *X1, const double *X2, const double *alpha_desc,
                               const int *Q1, const int *Q2, const int *N1, const int *N2,
                               int nm1, int nm2, int max_atoms1, int max_atoms2, int rep,
                               double sigma, double *E, void *stream);

/* Planned full matvec for repeated products with a fixed training side and fixed query
 * labels (molecular dynamics: the query geometry changes, its atoms and labels do not).
 * kf_matvec_plan_create reads the labels once, packs the training rows of every label
 * with their norms, alpha_desc dot products and alpha_E (the quantities of the
 * reference's precompute_train, src/cuda_local_kernels.hpp:178-230) and allocates the
 * plan's own device workspace; kf_matvec_plan_execute then computes
 * (E (nm1,), F (naq1,)) = kf_local_full_matvec(X1, dX1, ...) for new query descriptors
 * and Jacobians with no host synchronisation and no allocation, so it can be captured
 * into a graph once the plan exists.  Replaces kernel_gaussian_full_matvec_cached and
 * _cached_graph (src/cuda_local_kernels.hpp:232-289).  A plan belongs to the device that
 * was current at creation; executions on one plan must not overlap (they share its
 * workspace).  kf_matvec_plan_destroy waits for the device before freeing. */
typedef struct kf_matvec_plan kf_matvec_plan;
int kf_matvec_plan_create(const int *Q1, const int *N1, int nm1, int max_atoms1, const double *X2,
                          const double *alpha_desc, const double *alpha_E, const int *Q2,
                          const int *N2, int nm2, int max_atoms2, int rep, double sigma,
                          kf_matvec_plan **plan, void *stream);
int kf_matvec_plan_execute(const kf_matvec_plan *plan, const double *X1, const double *dX1,
                           double *E, double *F, void *stream);
int kf_matvec_plan_destroy(kf_matvec_plan *plan);

/* ------------------------------------------------------------------------ */
/* Energy+force solve by PCG with a (learned) block-Jacobi preconditioner.    */
/* The E+F vector order is [E_0..E_{nm-1}, F_0 (3 N_0), F_1, ...]; block b    */
/* holds E_b then F_b, bs = 1 + 3 max_atoms.  These have no counterpart in    */
/* the reference (which solves E+F systems only by RFP Cholesky,             */
/* src/math.cpp:134-171); they serve the matrix-free solve of that system.   */
/* ------------------------------------------------------------------------ */

/* blocks (nm, bs, bs): the diagonal (E_b, F_b) blocks of the symmetric full kernel
 * (the entries kf_local_kernel_full_symm places at those rows / columns); entries
 * beyond 1 + 3 N_b are left untouched. */
int kf_local_full_diag_blocks(const double *X, const double *dX, const int *Q, const int *N,
                              int nm, int max_atoms, int rep, double sigma, double *blocks,
                              void *stream);

/* inv_b = (blocks_b + lambda I + diag(shift_b))^-1 on the live 1 + 3 N_b rows (identity on
 * padded rows); shift (nb, bs) may be null.  info (nb ints, may be null) gets 0 or the
 * 1-based pivot at which the block was not positive definite. */
int kf_block_factor(const double *blocks, int nb, int bs, const int *N, int max_atoms,
                    double lambda, const double *shift, double *inv, int *info, void *stream);

/* z = blockdiag(inv) r in the E+F vector order. */
int kf_block_apply(const double *inv, int nb, int bs, const int *N, int max_atoms,
                   const double *r, double *z, void *stream);

/* PCG on (A + lambda I) x = y, A an explicit RFP (TRANSR='N', UPLO='U') E+F kernel of
 * order n = nb + 3 sum N; inv from kf_block_factor, or null for plain CG.  Stops at
 * ||r||/||y|| <= tol or max_iter; *rel is the true final residual; history (max_iter+1
 * doubles, may be null) receives the recurrence residual per iteration. */
int kf_ef_pcg_rfp(const double *arf, int n, double lambda, const double *inv, int nb, int bs,
                  const int *N, int max_atoms, const double *y, double *x, double tol,
                  int max_iter, int *iters, double *rel, double *history, void *stream);

/* The same with the matrix-free operator K_full x = full_matvec(alpha_desc(x_F), x_E)
 * of the training set (X, dX). */
int kf_ef_pcg_matfree(const double *X, const double *dX, const int *Q, const int *N, int nm,
                      int max_atoms, int rep, double sigma, double lambda, const double *inv,
                      int bs, const double *y, double *x, double tol, int max_iter, int *iters,
                      double *rel, double *history, void *stream);

/* loss = log(||r_steps|| / ||y||) after `steps` PCG iterations from x = 0 with the RFP
 * operator and inv = (K_bb + lambda + diag(s))^-1; adds d loss / d s to grad_shift
 * (nb, bs) by reverse-mode differentiation of the recurrences. */
int kf_pcg_unrolled_grad_rfp(const double *arf, int n, double lambda, const double *inv, int nb,
                             int bs, const int *N, int max_atoms, const double *y, int steps,
                             double *loss, double *grad_shift, void *stream);

/* Preconditioner MLP (BF16 weights W (fan_in x fan_out) as uint16, FP32 everything
 * else): Y = relu(relu(X W1 + b1) W2 + b2) W3 + b3, keeping Z1, Z2 (B x H) for the
 * backward.  No counterpart in the reference. */
int kf_mlp_forward(const float *X, int B, int D0, int H, int D3, const uint16_t *W1,
                   const float *b1, const uint16_t *W2, const float *b2, const uint16_t *W3,
                   const float *b3, float *Z1, float *Z2, float *Y, void *stream);
int kf_mlp_backward(const float *X, int B, int D0, int H, int D3, const uint16_t *W1,
                    const uint16_t *W2, const uint16_t *W3, const float *Z1, const float *Z2,
                    const float *dY, float *dW1, float *db1, float *dW2, float *db2,
                    float *dW3, float *db3, void *stream);
/* Adam step on FP32 master weights p (bias-corrected, step t >= 1); refreshes the BF16
 * copy p_bf16 (round to nearest even) when given. */
int kf_mlp_adam(float *p, const float *g, float *m, float *v, int n, float lr, float beta1,
                float beta2, float eps, int t, uint16_t *p_bf16, void *stream);
int kf_cast_bf16(const float *p, int n, uint16_t *out, void *stream);

/* ------------------------------------------------------------------------ */
/* kernelmath: RFP Cholesky                                                  */
/* ------------------------------------------------------------------------ */

/* In-place Cholesky of an RFP (TRANSR='N', UPLO='U') matrix after adding l2 to
 * its diagonal; on return arf holds the factor exactly where LAPACK dpftrf
 * leaves it.  Replaces the dpftrf_ call of kf::math::solve_cholesky_rfp
 * (src/math.cpp:156-165).  precision KF_PREC_F64 (double*) or KF_PREC_F32
 * (float*).  Returns 0 or the positive LAPACK-style info. */
int kf_rfp_potrf(int precision, void *arf, int n, double l2, void *stream);

/* Solve (L L^T) x = b in place for nrhs right-hand sides stored as columns of b
 * (b is (nrhs, n) row-major, i.e. each RHS contiguous) with the factor from
 * kf_rfp_potrf.  Replaces the dpftrs_ call (src/math.cpp:167-170).
 * precision KF_PREC_F64 (double arf, double b) or KF_PREC_F32 (float arf,
 * double b: the solve accumulates in FP64). */
int kf_rfp_potrs(int precision, const void *arf, int n, double *b, int nrhs, void *stream);

/* kf_rfp_potrs without allocation: the same solve, with its scratch (the inverses of the
 * diagonal blocks of groups of four 64-wide blocks, and the intermediate L^-1 b) in a fixed
 * 48 MiB device array of the library instead of allocated workspace.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code under graph
 * capture, tight serving loops that solve many right-hand sides with one factor).  Orders
 * up to 38,912 with one right-hand side fit (the scratch needs 327,680 * ceil(nblk / 4) +
 * 8 * n * nrhs bytes, nblk = ceil((n/2)/64) + ceil((n - n/2)/64)); larger requests return
 * KF_ERR_ARG.  The first call on a device looks up the array's device address; no call
 * allocates.  Calls on one device share the array: calls that overlap on one device race
 * (the library's per-device serialisation of entry points orders calls on different
 * streams). */
int kf_rfp_potrs_noalloc(int precision, const void *arf, int n, double *b, int nrhs,
                         void *stream);

/* y = (A + lambda I) x for a symmetric RFP matrix A (precision KF_PREC_F64: double*,
 * KF_PREC_F32: float*), FP64 vectors and accumulation, deterministic.  The operator
 * of the refinement residual r = y - (K + lambda I) alpha (no reference counterpart:
 * the reference only factorises, src/math.cpp:134-171). */
int kf_rfp_symv(int precision, const void *arf, int n, double lambda, const double *x, double *y,
                void *stream);

/* Preconditioned conjugate gradients on (A + lambda I) alpha = y, A an RFP matrix
 * (a_precision F64/F32), preconditioner M = L L^T from an RFP Cholesky factor
 * (m_precision F64/F32) of a shifted A.  alpha holds the initial guess on entry.
 * Stops at ||r|| / ||y|| <= tol (true FP64 residual); returns KF_ERR_NOCONV when
 * max_iter is reached first.  No reference counterpart (the reference's GPU model
 * runs a Jacobi-PCG on dense FP32 K, models/cuda_local_krr.py:214-287). */
int kf_rfp_pcg(int a_precision, const void *arf, int m_precision, const void *lfac, int n,
               double lambda, const double *y, double *alpha, double tol, int max_iter, int *iters,
               double *rel_residual, void *stream);

/* Mixed-precision replacement of cho_solve_rfp (src/math.cpp:134-171) for an FP32
 * RFP kernel K32 (e.g. from kf_local_kernel_symm_rfp with KF_PREC_BF16X3): FP32 RFP
 * Cholesky of K32 + mu I into lwork (n(n+1)/2 floats; mu = lambda, raised only if
 * the FP32 factorisation breaks down) used as the preconditioner of PCG on
 * (K32 + lambda I) with FP64 residuals, to ||r||/||y|| <= tol.  The system solved is
 * always K32 + lambda I: if PCG stalls or breaks down (K32 + lambda I not positive
 * definite in FP32 storage) the call returns KF_ERR_NOCONV and the caller falls back to
 * the FP64 kernel and kf_rfp_potrf/potrs.  *lambda_used (may be null) returns mu, the
 * shift of the preconditioner's factor.  K32 is not modified. */
int kf_krr_solve_mixed(const float *K32, float *lwork, int n, double lambda, const double *y,
                       double *alpha, double tol, int max_iter, int *iters, double *rel_residual,
                       double *lambda_used, void *stream);

/* Full (column-major, leading dimension lda) -> RFP, LAPACK dtrttf semantics for
 * all TRANSR in {'N','T'} x UPLO in {'U','L'}.  Replaces the dtrttf_ call of
 * kf::math::full_to_rfp (src/math.cpp:173-179); kernelmath.full_to_rfp passes the
 * C-order buffer with UPLO swapped exactly as src/math_bindings.cpp:89-120 does. */
int kf_trttf(char transr, char uplo, int n, const double *A, int lda, double *arf, void *stream);

/* RFP -> full (column-major), LAPACK dtfttr semantics; with symmetrize != 0 the
 * other triangle is mirrored as kf::math::rfp_to_full does (src/math.cpp:181-211). */
int kf_tfttr(char transr, char uplo, int n, const double *arf, double *A, int lda, int symmetrize,
             void *stream);

/* Symmetric RFP re-layout between the four TRANSR x UPLO variants (column-major LAPACK
 * semantics), FP64 in, FP64 or FP32 (out_precision) out.  No reference counterpart as a
 * function: it converts this library's native packing (TRANSR='N', UPLO='U', that of
 * src/local_kernels.cpp:386-400) to the reference GPU kernels' TRANSR='N', UPLO='L'
 * (src/cuda_local_kernels.hpp:52) for the torch-tensor interface. */
int kf_rfp_relayout(char transr_in, char uplo_in, char transr_out, char uplo_out, int n,
                    const double *in, int out_precision, void *out, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* KF_HIP_H */
