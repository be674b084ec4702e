"""Torch-tensor interface of the local Gaussian kernels, with the function names and
arguments of the reference's GPU module (python/kernelforge/cuda_local_kernels.pyi:9-349,
src/cuda_local_kernels.hpp), on the HIP kernels of libkf_hip.so.

Every input is a GPU tensor; results are GPU tensors produced on torch's current stream.
The kernels compute in FP64.  Outputs the reference returns as float32 come back as
float32 when the inputs are float32 and as float64 when any input is float64; the
rectangular and symmetric dense kernels are float64 as in the reference.  RFP outputs use
the reference GPU kernels' packing, TRANSR='N', UPLO='L' (unpack with
kernelmath.rfp_to_full(K_rfp, n, uplo='U', transr='N')).

Import torch before any other kernelforge_amd module (see _torch_bridge).
"""

from __future__ import annotations

import ctypes

import torch

from . import _torch_bridge as tb
from . import capi

__all__ = [
    "kernel_gaussian_rect", "kernel_gaussian_symm", "kernel_gaussian_symm_rfp",
    "kernel_gaussian_full_symm", "kernel_gaussian_full_symm_rfp", "compute_alpha_desc",
    "kernel_gaussian_full_matvec", "precompute_train", "kernel_gaussian_full_matvec_cached",
    "kernel_gaussian_full_matvec_cached_graph",
    "kernel_gaussian_full_matvec_cached_graph_with_offsets",
]


def _xqn(X, Q, N, tag=""):
    X = tb.f64(X, "X" + tag)
    if X.dim() != 3:
        raise ValueError(f"X{tag} must have shape (nm, max_atoms, rep)")
    nm, ma, rep = X.shape
    Q, N = tb.i32(Q, "Q" + tag), tb.i32(N, "N" + tag)
    if tuple(Q.shape) != (nm, ma) or tuple(N.shape) != (nm,):
        raise ValueError(f"Q{tag} must be (nm, max_atoms) and N{tag} (nm,)")
    return X, Q, N, nm, ma, rep


def _dx(dX, nm, ma, rep, tag=""):
    dX = tb.f64(dX, "dX" + tag)
    if dX.numel() != nm * ma * rep * 3 * ma:
        raise ValueError(f"dX{tag} must have shape (nm, max_atoms, rep, 3*max_atoms)")
    return dX.reshape(nm, ma, rep, 3 * ma)


def _check_train(ad, aE, nm, ma, rep):
    if ad.numel() != nm * ma * rep:
        raise ValueError("alpha_desc must have shape (nm_t, max_atoms_t, rep)")
    if aE.numel() != nm:
        raise ValueError("alpha_E must have nm_t entries")


def _sigma(sigma):
    sigma = float(sigma)
    if not sigma > 0.0:
        raise ValueError("sigma must be > 0")
    return sigma


def _to_lower_rfp(native: torch.Tensor, n: int, dtype: torch.dtype) -> torch.Tensor:
    out = torch.empty(native.shape, dtype=dtype, device=native.device)
    prec = capi.KF_PREC_F64 if dtype == torch.float64 else capi.KF_PREC_F32
    tb.check(tb.lib().kf_rfp_relayout(b"N", b"U", b"N", b"L", n, native.data_ptr(), prec,
                                      out.data_ptr(), tb.stream_of(native)), "kf_rfp_relayout")
    return out


def kernel_gaussian_rect(X_q, Q_q, N_q, X_t, Q_t, N_t, sigma):
    """K[a, b] over label-matched atom pairs, (nm_q, nm_t) float64."""
    dt = tb.same_device(X_q, Q_q, N_q, X_t, Q_t, N_t)
    X1, Q1, N1, nm1, ma1, rep = _xqn(X_q, Q_q, N_q, "_q")
    X2, Q2, N2, nm2, ma2, rep2 = _xqn(X_t, Q_t, N_t, "_t")
    if rep2 != rep:
        raise ValueError("X_q and X_t must have the same representation size")
    K = torch.empty((nm1, nm2), dtype=torch.float64, device=dt)
    tb.check(tb.lib().kf_local_kernel(capi.KF_PREC_F64, X1.data_ptr(), X2.data_ptr(),
                                      Q1.data_ptr(), Q2.data_ptr(), N1.data_ptr(), N2.data_ptr(),
                                      nm1, nm2, ma1, ma2, rep, _sigma(sigma), K.data_ptr(),
                                      tb.stream_of(K)), "kernel_gaussian_rect")
    return K


def kernel_gaussian_symm(X, Q, N, sigma):
    """Symmetric (nm, nm) float64 kernel, both triangles."""
    dt = tb.same_device(X, Q, N)
    X, Q, N, nm, ma, rep = _xqn(X, Q, N)
    K = torch.empty((nm, nm), dtype=torch.float64, device=dt)
    tb.check(tb.lib().kf_local_kernel_symm(capi.KF_PREC_F64, X.data_ptr(), Q.data_ptr(),
                                           N.data_ptr(), nm, ma, rep, _sigma(sigma), K.data_ptr(),
                                           tb.stream_of(K)), "kernel_gaussian_symm")
    return K


def kernel_gaussian_symm_rfp(X, Q, N, sigma):
    """Symmetric kernel in RFP (TRANSR='N', UPLO='L'), nm(nm+1)/2 entries."""
    tb.same_device(X, Q, N)
    dtype = tb.out_dtype(X)
    X, Q, N, nm, ma, rep = _xqn(X, Q, N)
    native = torch.empty(nm * (nm + 1) // 2, dtype=torch.float64, device=X.device)
    tb.check(tb.lib().kf_local_kernel_symm_rfp(capi.KF_PREC_F64, X.data_ptr(), Q.data_ptr(),
                                               N.data_ptr(), nm, ma, rep, _sigma(sigma),
                                               native.data_ptr(), tb.stream_of(X)),
             "kernel_gaussian_symm_rfp")
    return _to_lower_rfp(native, nm, dtype)


def kernel_gaussian_full_symm(X, dX, Q, N, sigma):
    """Symmetric (nm + naq)^2 energy+force kernel [[EE, EF], [FE, FF]], naq = 3 sum(N)."""
    tb.same_device(X, dX, Q, N)
    dtype = tb.out_dtype(X, dX)
    X, Q, N, nm, ma, rep = _xqn(X, Q, N)
    dXf = _dx(dX, nm, ma, rep)
    big = nm + tb.naq(N, ma)
    K = torch.empty((big, big), dtype=torch.float64, device=X.device)
    tb.check(tb.lib().kf_local_kernel_full_symm(X.data_ptr(), dXf.data_ptr(), Q.data_ptr(),
                                                N.data_ptr(), nm, ma, rep, _sigma(sigma),
                                                K.data_ptr(), tb.stream_of(X)),
             "kernel_gaussian_full_symm")
    return K if dtype == torch.float64 else K.to(dtype)


def kernel_gaussian_full_symm_rfp(X, dX, Q, N, sigma):
    """The energy+force kernel in RFP (TRANSR='N', UPLO='L'), BIG(BIG+1)/2 entries."""
    tb.same_device(X, dX, Q, N)
    dtype = tb.out_dtype(X, dX)
    X, Q, N, nm, ma, rep = _xqn(X, Q, N)
    dXf = _dx(dX, nm, ma, rep)
    big = nm + tb.naq(N, ma)
    native = torch.empty(big * (big + 1) // 2, dtype=torch.float64, device=X.device)
    tb.check(tb.lib().kf_local_kernel_full_symm_rfp(X.data_ptr(), dXf.data_ptr(), Q.data_ptr(),
                                                    N.data_ptr(), nm, ma, rep, _sigma(sigma),
                                                    native.data_ptr(), tb.stream_of(X)),
             "kernel_gaussian_full_symm_rfp")
    return _to_lower_rfp(native, big, dtype)


def compute_alpha_desc(dX, N, alpha_F):
    """alpha_desc[b, i, :] = dX[b, i, :, :3 N_b] @ alpha_F[offs_b : offs_b + 3 N_b]."""
    tb.same_device(dX, N, alpha_F)
    dtype = tb.out_dtype(dX, alpha_F)
    N = tb.i32(N, "N")
    if dX.dim() not in (4, 5):
        raise ValueError("dX must have shape (nm, max_atoms, rep, 3*max_atoms)")
    nm, ma, rep = dX.shape[0], dX.shape[1], dX.shape[2]
    dXf = _dx(dX, nm, ma, rep)
    a = tb.f64(alpha_F, "alpha_F").reshape(-1)
    if a.numel() != tb.naq(N, ma):
        raise ValueError("alpha_F must have naq = 3 * sum(N) entries")
    out = torch.empty((nm, ma, rep), dtype=torch.float64, device=dXf.device)
    tb.check(tb.lib().kf_local_alpha_desc(dXf.data_ptr(), None, N.data_ptr(), nm, ma, rep,
                                          a.data_ptr(), out.data_ptr(), tb.stream_of(dXf)),
             "compute_alpha_desc")
    return out if dtype == torch.float64 else out.to(dtype)


def kernel_gaussian_full_matvec(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc_F,
                                sigma):
    """(E_pred (nm_q,), F_pred (naq_q,)) = K_full(query, train) @ [alpha_E; alpha_F] with
    alpha_desc_F from compute_alpha_desc (the full kernel is never formed)."""
    tb.same_device(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc_F)
    dtype = tb.out_dtype(X_q, dX_q, X_t, alpha_E, alpha_desc_F)
    X1, Q1, N1, nm1, ma1, rep = _xqn(X_q, Q_q, N_q, "_q")
    X2, Q2, N2, nm2, ma2, _ = _xqn(X_t, Q_t, N_t, "_t")
    dX1 = _dx(dX_q, nm1, ma1, rep, "_q")
    ad = tb.f64(alpha_desc_F, "alpha_desc_F")
    aE = tb.f64(alpha_E, "alpha_E").reshape(-1)
    _check_train(ad, aE, nm2, ma2, rep)
    E = torch.empty(nm1, dtype=torch.float64, device=X1.device)
    F = torch.empty(tb.naq(N1, ma1), dtype=torch.float64, device=X1.device)
    tb.check(tb.lib().kf_local_full_matvec(X1.data_ptr(), dX1.data_ptr(), X2.data_ptr(),
                                           ad.data_ptr(), aE.data_ptr(), Q1.data_ptr(),
                                           Q2.data_ptr(), N1.data_ptr(), N2.data_ptr(), nm1, nm2,
                                           ma1, ma2, rep, _sigma(sigma), E.data_ptr(),
                                           F.data_ptr() if F.numel() else None, tb.stream_of(E)),
             "kernel_gaussian_full_matvec")
    return (E, F) if dtype == torch.float64 else (E.to(dtype), F.to(dtype))


def precompute_train(X_t, Q_t, N_t, alpha_E, alpha_desc):
    """Training-side constants of the cached matvec, as the reference defines them
    (src/cuda_local_kernels.cu:1564-1640, 2620-2650): norms_t = ||X_t[t]||^2,
    S_adF = X_t[t] . alpha_desc[t], alpha_E_t = alpha_E[mol(t)] and
    combined_t = alpha_desc + alpha_E_t[t] X_t, zero on padded atoms, each flattened
    over (nm_t * max_atoms_t).  The cached matvec below keeps its own packed copies of
    these (a plan); the tensors are returned for callers that use them."""
    tb.same_device(X_t, Q_t, N_t, alpha_E, alpha_desc)
    dtype = tb.out_dtype(X_t, alpha_E, alpha_desc)
    X, _, N, nm, ma, rep = _xqn(X_t, Q_t, N_t, "_t")
    ad = tb.f64(alpha_desc, "alpha_desc").reshape(nm, ma, rep)
    aE = tb.f64(alpha_E, "alpha_E").reshape(nm)
    live = (torch.arange(ma, device=X.device)[None, :] < N.clamp(0, ma)[:, None]).to(torch.float64)
    norms = (X * X).sum(-1).reshape(-1)
    s_adf = ((X * ad).sum(-1) * live).reshape(-1)
    aE_t = (aE[:, None] * live).reshape(-1)
    comb = ((ad + aE_t.reshape(nm, ma, 1) * X) * live[..., None]).reshape(nm * ma, rep)
    return tuple(t.to(dtype) for t in (norms, s_adf, aE_t, comb))


class _Plan:
    """One planned matvec (kf_matvec_plan) with the key it was made for."""

    def __init__(self, key, handle, naq, nm, ma, rep, device):
        self.key, self.handle, self.naq, self.nm, self.ma, self.rep = key, handle, naq, nm, ma, rep
        self.E = torch.empty(nm, dtype=torch.float64, device=device)
        self.F = torch.empty(max(naq, 1), dtype=torch.float64, device=device)

    def __del__(self):
        try:
            if self.handle:
                capi.lib().kf_matvec_plan_destroy(self.handle)
        except Exception:
            pass


_PLAN: _Plan | None = None


def _sig(t: torch.Tensor):
    return (t.data_ptr(), tuple(t.shape), t.dtype, t._version)


def _cached_plan(X_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc, sigma) -> _Plan:
    """The plan for these query labels and this training side; made on first use and kept
    while none of the tensors it was made from changes (storage, shape, dtype or an
    in-place write, by torch's version counter)."""
    global _PLAN
    sigma = _sigma(sigma)
    key = (X_q.device, sigma, X_q.shape[2],
           *(_sig(t) for t in (Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc)))
    if _PLAN is not None and _PLAN.key == key:
        return _PLAN
    if torch.cuda.is_current_stream_capturing():
        raise RuntimeError("the cached matvec plans on its first call, which synchronises: "
                           "call it once before graph capture")
    _, Q1, N1, nm1, ma1, rep = _xqn(X_q, Q_q, N_q, "_q")
    X2, Q2, N2, nm2, ma2, rep2 = _xqn(X_t, Q_t, N_t, "_t")
    if rep2 != rep:
        raise ValueError("X_q and X_t must have the same representation size")
    ad = tb.f64(alpha_desc, "alpha_desc")
    aE = tb.f64(alpha_E, "alpha_E").reshape(-1)
    _check_train(ad, aE, nm2, ma2, rep)
    h = ctypes.c_void_p()
    tb.check(tb.lib().kf_matvec_plan_create(Q1.data_ptr(), N1.data_ptr(), nm1, ma1,
                                            X2.data_ptr(), ad.data_ptr(), aE.data_ptr(),
                                            Q2.data_ptr(), N2.data_ptr(), nm2, ma2, rep, sigma,
                                            ctypes.byref(h), tb.stream_of(X2)),
             "kf_matvec_plan_create")
    _PLAN = None  # destroy the previous plan before keeping the new one
    _PLAN = _Plan(key, h.value, tb.naq(N1, ma1), nm1, ma1, rep, X_q.device)
    return _PLAN


def _run_plan(P: _Plan, X_q, dX_q):
    # the plan's kernels index X_q and dX_q by its own sizes: check them on the host
    if X_q.dim() != 3 or tuple(X_q.shape) != (P.nm, P.ma, P.rep):
        raise ValueError(f"X_q must have shape ({P.nm}, {P.ma}, {P.rep}) like Q_q")
    nm, ma, rep = X_q.shape
    X1 = tb.f64(X_q, "X_q")
    dX1 = _dx(dX_q, nm, ma, rep, "_q")
    tb.check(tb.lib().kf_matvec_plan_execute(P.handle, X1.data_ptr(), dX1.data_ptr(),
                                             P.E.data_ptr(), P.F.data_ptr(), tb.stream_of(X1)),
             "kf_matvec_plan_execute")


def kernel_gaussian_full_matvec_cached(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc,
                                       norms_t, S_adF, alpha_E_t, combined_t, sigma):
    """kernel_gaussian_full_matvec for repeated calls with fixed training data and query
    labels (molecular dynamics): after the first call, each call runs the planned matvec
    with no host synchronisation and no allocation in the library."""
    tb.same_device(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc, norms_t, S_adF,
                   alpha_E_t, combined_t)
    nt = X_t.shape[0] * X_t.shape[1]
    if norms_t.numel() != nt or S_adF.numel() != nt or alpha_E_t.numel() != nt \
            or combined_t.numel() != nt * X_t.shape[2]:
        raise ValueError("precomputed training tensors do not match X_t (use precompute_train)")
    dtype = tb.out_dtype(X_q, dX_q, X_t, alpha_E, alpha_desc)
    P = _cached_plan(X_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc, sigma)
    _run_plan(P, X_q, dX_q)
    E, F = P.E.clone(), P.F[: P.naq].clone()
    return (E, F) if dtype == torch.float64 else (E.to(dtype), F.to(dtype))


def kernel_gaussian_full_matvec_cached_graph(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E,
                                             alpha_desc, norms_t, S_adF, alpha_E_t, combined_t,
                                             query_indices_by_group, query_counts_by_group,
                                             E_pred, F_pred, sigma):
    """The cached matvec into caller-provided E_pred (nm_q,) and F_pred (naq_q,), with no
    synchronisation once planned, so it can be captured into a graph after one warm-up
    call.  The per-label query groups are part of the plan here; the arguments are
    accepted for the reference's signature."""
    del query_indices_by_group, query_counts_by_group
    P = _cached_plan(X_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E, alpha_desc, sigma)
    if E_pred.numel() != P.nm or F_pred.numel() != P.naq:
        raise ValueError("E_pred must have nm_q and F_pred naq_q entries")
    _run_plan(P, X_q, dX_q)
    E_pred.copy_(P.E.reshape(E_pred.shape))
    F_pred.copy_(P.F[: P.naq].reshape(F_pred.shape))


def kernel_gaussian_full_matvec_cached_graph_with_offsets(X_q, dX_q, Q_q, N_q, offs_q, X_t, Q_t,
                                                          N_t, alpha_E, alpha_desc, norms_t,
                                                          S_adF, alpha_E_t, combined_t,
                                                          query_indices_by_group,
                                                          query_counts_by_group, E_pred, F_pred,
                                                          sigma):
    """kernel_gaussian_full_matvec_cached_graph with the query force offsets given (they
    equal the cumulative 3 N_q, which the plan holds)."""
    del offs_q
    kernel_gaussian_full_matvec_cached_graph(X_q, dX_q, Q_q, N_q, X_t, Q_t, N_t, alpha_E,
                                             alpha_desc, norms_t, S_adF, alpha_E_t, combined_t,
                                             query_indices_by_group, query_counts_by_group,
                                             E_pred, F_pred, sigma)
