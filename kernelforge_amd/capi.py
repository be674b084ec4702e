"""ctypes binding of the C ABI in ``include/kf_hip.h`` (``libkf_hip.so``).

This is the binding a Python caller holding device memory uses: every array argument
is a device pointer passed as an integer (``DeviceArray.ptr`` below, or any other
allocation on the same HIP runtime), ``stream`` is a ``hipStream_t`` or 0.  The
library links the system ROCm runtime; a process must not also load a second,
differently-versioned HIP runtime (PyTorch wheels bundle their own), so device
memory here comes from the library's own kf_malloc.  Status codes are turned into the exceptions the reference
raises (ValueError / RuntimeError / MemoryError).
"""

from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

KF_OK = 0
KF_ERR_ARG, KF_ERR_SIGMA, KF_ERR_ELEMENT, KF_ERR_BASIS = -1, -2, -3, -4
KF_ERR_HIP, KF_ERR_NOMEM, KF_ERR_NOCONV = -5, -6, -7
KF_PREC_F64, KF_PREC_BF16X3, KF_PREC_F32, KF_PREC_I8X6 = 0, 1, 2, 3

_c_int, _c_double, _c_char, _vp = ctypes.c_int, ctypes.c_double, ctypes.c_char, ctypes.c_void_p

# name -> argtypes (restype int unless listed in _RESTYPE)
SIGNATURES = {
    "kf_version": [],
    "kf_last_error": [],
    "kf_device_name": [ctypes.c_char_p, _c_int],
    "kf_malloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
    "kf_free": [_vp],
    "kf_memcpy": [_vp, _vp, ctypes.c_size_t, _c_int, _vp],
    "kf_memset_zero": [_vp, ctypes.c_size_t, _vp],
    "kf_stream_synchronize": [_vp],
    "kf_timer_enable": [_c_int],
    "kf_timer_reset": [],
    "kf_timer_read": [_c_int, ctypes.POINTER(_c_double), ctypes.POINTER(_c_int)],
    "kf_fchl19_f64": [_vp, _vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int,
                      _c_int, _c_double, _c_double, _c_double, _c_double, _c_double, _c_double,
                      _c_double, _c_double, _vp, _vp, _vp],
    "kf_fchl19v2_f64": [_vp, _vp, _vp, _c_int, _c_int, _vp, _c_int, _vp, _c_int, _vp, _c_int,
                        _vp, _c_int, _c_int, _c_int, _c_int] + [_c_double] * 11 + [_vp, _vp, _vp],
    "kf_local_kernel_symm_rfp": [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_double, _vp,
                                 _vp],
    "kf_local_kernel_symm": [_c_int, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_double, _vp, _vp],
    "kf_local_kernel": [_c_int, _vp, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int, _c_int, _c_int,
                        _c_int, _c_double, _vp, _vp],
    "kf_rfp_potrf": [_c_int, _vp, _c_int, _c_double, _vp],
    "kf_rfp_potrs": [_c_int, _vp, _c_int, _vp, _c_int, _vp],
    "kf_rfp_potrs_noalloc": [_c_int, _vp, _c_int, _vp, _c_int, _vp],
    "kf_rfp_symv": [_c_int, _vp, _c_int, _c_double, _vp, _vp, _vp],
    "kf_rfp_pcg": [_c_int, _vp, _c_int, _vp, _c_int, _c_double, _vp, _vp, _c_double, _c_int,
                   ctypes.POINTER(_c_int), ctypes.POINTER(_c_double), _vp],
    "kf_krr_solve_mixed": [_vp, _vp, _c_int, _c_double, _vp, _vp, _c_double, _c_int,
                           ctypes.POINTER(_c_int), ctypes.POINTER(_c_double),
                           ctypes.POINTER(_c_double), _vp],
    "kf_local_kernel_jacobian": [_vp] * 7 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_local_kernel_jacobian_t": [_vp] * 7 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_local_kernel_hessian": [_vp] * 8 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_local_kernel_full": [_vp] * 8 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_local_kernel_hessian_symm": [_vp] * 4 + [_c_int] * 3 + [_c_double, _vp, _vp],
    "kf_local_kernel_hessian_symm_rfp": [_vp] * 4 + [_c_int] * 3 + [_c_double, _vp, _vp],
    "kf_local_kernel_full_symm": [_vp] * 4 + [_c_int] * 3 + [_c_double, _vp, _vp],
    "kf_local_kernel_full_symm_rfp": [_vp] * 4 + [_c_int] * 3 + [_c_double, _vp, _vp],
    "kf_local_alpha_desc": [_vp, _vp, _vp, _c_int, _c_int, _c_int, _vp, _vp, _vp],
    "kf_local_full_matvec": [_vp] * 9 + [_c_int] * 5 + [_c_double, _vp, _vp, _vp],
    "kf_local_hessian_matvec": [_vp] * 8 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_local_jacobian_t_matvec": [_vp] * 7 + [_c_int] * 5 + [_c_double, _vp, _vp],
    "kf_matvec_plan_create": [_vp, _vp, _c_int, _c_int, _vp, _vp, _vp, _vp, _vp, _c_int, _c_int,
                              _c_int, _c_double, ctypes.POINTER(_vp), _vp],
    "kf_matvec_plan_execute": [_vp, _vp, _vp, _vp, _vp, _vp],
    "kf_matvec_plan_destroy": [_vp],
    "kf_local_full_diag_blocks": [_vp] * 4 + [_c_int] * 3 + [_c_double, _vp, _vp],
    "kf_block_factor": [_vp, _c_int, _c_int, _vp, _c_int, _c_double, _vp, _vp, _vp, _vp],
    "kf_block_apply": [_vp, _c_int, _c_int, _vp, _c_int, _vp, _vp, _vp],
    "kf_ef_pcg_rfp": [_vp, _c_int, _c_double, _vp, _c_int, _c_int, _vp, _c_int, _vp, _vp,
                      _c_double, _c_int, ctypes.POINTER(_c_int), ctypes.POINTER(_c_double), _vp,
                      _vp],
    "kf_ef_pcg_matfree": [_vp] * 4 + [_c_int] * 3 + [_c_double, _c_double, _vp, _c_int, _vp, _vp,
                                                      _c_double, _c_int, ctypes.POINTER(_c_int),
                                                      ctypes.POINTER(_c_double), _vp, _vp],
    "kf_pcg_unrolled_grad_rfp": [_vp, _c_int, _c_double, _vp, _c_int, _c_int, _vp, _c_int, _vp,
                                 _c_int, ctypes.POINTER(_c_double), _vp, _vp],
    "kf_mlp_forward": [_vp, _c_int, _c_int, _c_int, _c_int] + [_vp] * 10,
    "kf_mlp_backward": [_vp, _c_int, _c_int, _c_int, _c_int] + [_vp] * 13,
    "kf_mlp_adam": [_vp, _vp, _vp, _vp, _c_int, ctypes.c_float, ctypes.c_float, ctypes.c_float,
                    ctypes.c_float, _c_int, _vp, _vp],
    "kf_cast_bf16": [_vp, _c_int, _vp, _vp],
    "kf_trttf": [_c_char, _c_char, _c_int, _vp, _c_int, _vp, _vp],
    "kf_tfttr": [_c_char, _c_char, _c_int, _vp, _vp, _c_int, _c_int, _vp],
    "kf_rfp_relayout": [_c_char, _c_char, _c_char, _c_char, _c_int, _vp, _c_int, _vp, _vp],
}
_RESTYPE = {"kf_last_error": ctypes.c_char_p}


def lib() -> ctypes.CDLL:
    """Load libkf_hip.so (raises OSError when it has not been built)."""
    global _LIB
    if _LIB is None:
        L = ctypes.CDLL(os.path.join(_HERE, "libkf_hip.so"))
        for name, args in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = _RESTYPE.get(name, _c_int)
        _LIB = L
    return _LIB


TM_FCHL19, TM_LOCAL_SYMM, TM_LOCAL_RECT, TM_POTRF, TM_POTRS, TM_FULL, TM_MATVEC = range(7)


def timer_read(tid: int):
    """(total milliseconds, number of bracketed calls) of one kernel timer."""
    ms, cnt = _c_double(0.0), _c_int(0)
    check(lib().kf_timer_read(tid, ctypes.byref(ms), ctypes.byref(cnt)), "kf_timer_read")
    return ms.value, cnt.value


class NoConvergence(RuntimeError):
    """An iterative solve stopped above its tolerance (KF_ERR_NOCONV)."""


def check(rc: int, what: str) -> int:
    if rc == KF_OK:
        return rc
    msg = f"{what}: {lib().kf_last_error().decode(errors='replace')}"
    if rc in (KF_ERR_ARG, KF_ERR_SIGMA, KF_ERR_BASIS):
        raise ValueError(msg)
    if rc == KF_ERR_NOMEM:
        raise MemoryError(msg)
    if rc == KF_ERR_NOCONV:
        raise NoConvergence(msg)
    if rc > 0:
        raise RuntimeError(f"{what} failed, info={rc}")
    raise RuntimeError(msg)


class DeviceArray:
    """A device allocation made through the C ABI (kf_malloc), with NumPy transfer.

    ``ptr`` is the device address to pass to the kf_* entry points."""

    def __init__(self, shape, dtype=np.float64):
        self.shape = tuple(int(s) for s in np.atleast_1d(shape)) if np.ndim(shape) else (int(shape),)
        self.dtype = np.dtype(dtype)
        self.nbytes = int(np.prod(self.shape, dtype=np.int64)) * self.dtype.itemsize
        p = _vp()
        check(lib().kf_malloc(ctypes.byref(p), self.nbytes), "kf_malloc")
        self.ptr = p.value or 0

    @classmethod
    def from_numpy(cls, a):
        a = np.ascontiguousarray(a)
        d = cls(a.shape, a.dtype)
        if d.nbytes:
            check(lib().kf_memcpy(d.ptr, a.ctypes.data, d.nbytes, 0, None), "kf_memcpy")
        return d

    def numpy(self):
        out = np.empty(self.shape, self.dtype)
        if self.nbytes:
            check(lib().kf_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1, None), "kf_memcpy")
        return out

    def copy_from(self, other, stream=0):
        assert other.nbytes == self.nbytes
        check(lib().kf_memcpy(self.ptr, other.ptr, self.nbytes, 2, stream or None), "kf_memcpy")

    def zero(self, stream=0):
        check(lib().kf_memset_zero(self.ptr, self.nbytes, stream or None), "kf_memset_zero")

    def free(self):
        if self.ptr and _LIB is not None:
            _LIB.kf_free(self.ptr)
        self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def synchronize(stream=0):
    check(lib().kf_stream_synchronize(stream or None), "kf_stream_synchronize")


def fchl_basis(elements=(1, 6, 7, 8, 16), nRs2=24, nRs3=20, nFourier=1, rcut=8.0, acut=8.0,
               eta3=2.7, three_body_weight=13.4, grad=False):
    """Grids as the reference bindings build them (src/fchl19_repr_bindings.cpp:72-87)."""
    Rs2 = [rcut * i / nRs2 for i in range(1, nRs2 + 1)]
    Rs3 = [acut * i / nRs3 for i in range(1, nRs3 + 1)]
    nab = 2 * nFourier
    if grad:
        nab = max(nab, 2)
    w3 = math.sqrt(eta3 / math.pi) * three_body_weight
    return list(elements), Rs2, Rs3, nab, w3


def fchl19(coords, Q, N, nm, max_atoms, X, dX=0, stream=0, elements=(1, 6, 7, 8, 16), nRs2=24,
           nRs3=20, nFourier=1, eta2=0.32, eta3=2.7, zeta=math.pi, rcut=8.0, acut=8.0,
           two_body_decay=1.8, three_body_decay=0.57, three_body_weight=13.4):
    """kf_fchl19_f64 on device pointers (dX=0 for the representation only)."""
    el, Rs2, Rs3, nab, w3 = fchl_basis(elements, nRs2, nRs3, nFourier, rcut, acut, eta3,
                                       three_body_weight, grad=bool(dX))
    el_a = (ctypes.c_int * len(el))(*el)
    r2 = (ctypes.c_double * len(Rs2))(*Rs2)
    r3 = (ctypes.c_double * len(Rs3))(*Rs3)
    rc = lib().kf_fchl19_f64(coords, Q, N, nm, max_atoms, ctypes.cast(el_a, _vp), len(el),
                             ctypes.cast(r2, _vp), len(Rs2), ctypes.cast(r3, _vp), len(Rs3), nab,
                             eta2, eta3, zeta, rcut, acut, two_body_decay, three_body_decay, w3,
                             X, dX or None, stream or None)
    return check(rc, "kf_fchl19_f64")


def rep_size(nelem=5, nRs2=24, nRs3=20, nabasis=2):
    return nelem * nRs2 + nelem * (nelem + 1) // 2 * nRs3 * nabasis


def local_kernel_symm_rfp(precision, X, Q, N, nm, max_atoms, rep, sigma, arf, stream=0):
    return check(lib().kf_local_kernel_symm_rfp(precision, X, Q, N, nm, max_atoms, rep, sigma,
                                                arf, stream or None), "kf_local_kernel_symm_rfp")


def local_kernel(precision, X1, X2, Q1, Q2, N1, N2, nm1, nm2, ma1, ma2, rep, sigma, K, stream=0):
    return check(lib().kf_local_kernel(precision, X1, X2, Q1, Q2, N1, N2, nm1, nm2, ma1, ma2, rep,
                                       sigma, K, stream or None), "kf_local_kernel")


def local_kernel_full_symm_rfp(X, dX, Q, N, nm, max_atoms, rep, sigma, arf, stream=0):
    return check(lib().kf_local_kernel_full_symm_rfp(X, dX, Q, N, nm, max_atoms, rep, sigma, arf,
                                                     stream or None),
                 "kf_local_kernel_full_symm_rfp")


def local_alpha_desc(dX, Q, N, nm, max_atoms, rep, alpha, out, stream=0):
    return check(lib().kf_local_alpha_desc(dX, Q, N, nm, max_atoms, rep, alpha, out,
                                           stream or None), "kf_local_alpha_desc")


def local_full_matvec(X1, dX1, X2, ad, aE, Q1, Q2, N1, N2, nm1, nm2, ma1, ma2, rep, sigma, E, F,
                      stream=0):
    return check(lib().kf_local_full_matvec(X1, dX1, X2, ad, aE, Q1, Q2, N1, N2, nm1, nm2, ma1,
                                            ma2, rep, sigma, E, F, stream or None),
                 "kf_local_full_matvec")


def rfp_potrf(precision, arf, n, l2, stream=0):
    return check(lib().kf_rfp_potrf(precision, arf, n, l2, stream or None), "DPFTRF")


def rfp_potrs(precision, arf, n, b, nrhs=1, stream=0):
    return check(lib().kf_rfp_potrs(precision, arf, n, b, nrhs, stream or None), "DPFTRS")


def rfp_symv(precision, arf, n, lam, x, y, stream=0):
    return check(lib().kf_rfp_symv(precision, arf, n, lam, x, y, stream or None), "kf_rfp_symv")


def krr_solve_mixed(K32, lwork, n, lam, y, alpha, tol=1e-10, max_iter=200, stream=0):
    """Returns (iterations, relative residual, shift of the FP32 preconditioner factor).

    Raises NoConvergence when PCG on (K32 + lam I) stops above ``tol``."""
    it, rel, mu = _c_int(0), _c_double(0.0), _c_double(0.0)
    check(lib().kf_krr_solve_mixed(K32, lwork, n, lam, y, alpha, tol, max_iter, ctypes.byref(it),
                                   ctypes.byref(rel), ctypes.byref(mu), stream or None),
          "kf_krr_solve_mixed")
    return it.value, rel.value, mu.value
