"""CPU FP64 oracle for the FCHL19 KRR hot path.

TEST INFRASTRUCTURE ONLY.  Loaded by tests/, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg as the checker; the product package
``kernelforge_amd`` never imports it.

``libkf_oracle.so`` (built from ``oracle/kf_oracle.c`` by ``oracle/Makefile``)
restates the reference's algorithms (file:line cited per function there).  BLAS
and LAPACK are the OpenBLAS that scipy bundles, handed to the C code as function
pointers taken from ``scipy.linalg.cython_blas`` / ``cython_lapack`` capsules.

``ref()`` loads ``oracle/_ref/libkf_ref.so`` -- the reference's own
``src/fchl19_repr.cpp`` and ``src/math.cpp`` compiled from their sources -- when
that was built (only where the reference tree is present); ``None`` otherwise.

Parity status of each function (what pins it) is listed in the test modules
``tests/test_oracle_*.py``.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(ref: bool | None = None) -> None:
    """Compile the oracle (and the reference shim when the reference is present)."""
    subprocess.run(["make", "-s", "-C", _HERE, "libkf_oracle.so"], check=True)
    if ref is None:
        ref = os.path.isdir("/root/reference/src")
    if ref:
        subprocess.run(["make", "-s", "-C", _HERE, "ref"], check=True)


def _capsule_ptr(capi: dict, name: str) -> int:
    cap = capi[name]
    get_name = ctypes.pythonapi.PyCapsule_GetName
    get_name.restype = ctypes.c_char_p
    get_name.argtypes = [ctypes.py_object]
    get_ptr = ctypes.pythonapi.PyCapsule_GetPointer
    get_ptr.restype = ctypes.c_void_p
    get_ptr.argtypes = [ctypes.py_object, ctypes.c_char_p]
    return get_ptr(cap, get_name(cap))


def lib() -> ctypes.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(_HERE, "libkf_oracle.so")
    if not os.path.exists(path):
        build(ref=False)
    L = ctypes.CDLL(path)
    from scipy.linalg import cython_blas, cython_lapack

    b, lp = cython_blas.__pyx_capi__, cython_lapack.__pyx_capi__
    ptrs = [_capsule_ptr(b, n) for n in ("dgemm", "dsyrk", "dgemv", "dger")]
    ptrs += [_capsule_ptr(lp, n) for n in ("dpftrf", "dpftrs")]
    L.kfo_set_blas.argtypes = [ctypes.c_void_p] * 6
    if L.kfo_set_blas(*ptrs) != 0:
        raise RuntimeError("oracle: BLAS entry points missing")
    L.kfo_rfp_index.restype = ctypes.c_size_t
    L.kfo_rfp_index.argtypes = [ctypes.c_size_t] * 3
    L.kfo_rep_size.restype = ctypes.c_long
    L.kfo_rep_size.argtypes = [ctypes.c_long] * 4
    _LIB = L
    return L


def ref() -> ctypes.CDLL | None:
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libkf_ref.so")
        if not os.path.exists(path):
            return None
        _REF = ctypes.CDLL(path)
    return _REF


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _i(a):
    return a.ctypes.data_as(_ip)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


DEFAULTS = dict(elements=(1, 6, 7, 8, 16), nRs2=24, nRs3=20, nFourier=1, eta2=0.32, eta3=2.7,
                zeta=np.pi, rcut=8.0, acut=8.0, two_body_decay=1.8, three_body_decay=0.57,
                three_body_weight=13.4)


def _grids(p, grad):
    """Basis grids and w3 as src/fchl19_repr_bindings.cpp:72-84,125-144 build them."""
    nRs2, nRs3, nF = p["nRs2"], p["nRs3"], p["nFourier"]
    Rs2 = np.array([p["rcut"] * i / nRs2 for i in range(1, nRs2 + 1)], dtype=np.float64)
    Rs3 = np.array([p["acut"] * i / nRs3 for i in range(1, nRs3 + 1)], dtype=np.float64)
    nT = max(2 * nF, 2) if grad else 2 * nF
    w3 = np.sqrt(p["eta3"] / np.pi) * p["three_body_weight"]
    return Rs2, Rs3, nT, w3


def _params(kw):
    p = dict(DEFAULTS)
    p.update(kw)
    if not p["elements"]:
        p["elements"] = DEFAULTS["elements"]
    return p


def rep_size(nelem, nRs2=24, nRs3=20, nabasis=2):
    return int(lib().kfo_rep_size(nelem, nRs2, nRs3, nabasis))


def fchl19(coords, z, grad=False, **kw):
    """Oracle generate_fchl_acsf[_and_gradients] for one molecule.
    Returns rep (n, rs) or (rep, grad (n, rs, 3n))."""
    p = _params(kw)
    c = _f64(coords).reshape(-1)
    z = _i32(z)
    n = z.shape[0]
    el = _i32(p["elements"])
    Rs2, Rs3, nT, w3 = _grids(p, grad)
    rs = rep_size(len(el), len(Rs2), len(Rs3), nT)
    rep = np.zeros((n, rs))
    g = np.zeros((n, rs, 3 * n)) if grad else None
    L = lib()
    rc = L.kfo_fchl19(_d(c), _i(z), n, _i(el), len(el), _d(Rs2), len(Rs2), _d(Rs3), len(Rs3), nT,
                      ctypes.c_double(p["eta2"]), ctypes.c_double(p["eta3"]),
                      ctypes.c_double(p["zeta"]), ctypes.c_double(p["rcut"]),
                      ctypes.c_double(p["acut"]), ctypes.c_double(p["two_body_decay"]),
                      ctypes.c_double(p["three_body_decay"]), ctypes.c_double(w3), _d(rep), _d(g))
    if rc:
        raise ValueError(f"oracle fchl19 failed ({rc})")
    return (rep, g) if grad else rep


def fchl19_batch(C, Q, N, grad=False, **kw):
    """Padded batch: C (nm, ma, 3), Q (nm, ma), N (nm,) -> X (nm, ma, rs)[, dX (nm, ma, rs, 3ma)]."""
    p = _params(kw)
    C, Q, N = _f64(C), _i32(Q), _i32(N)
    nm, ma = Q.shape
    el = _i32(p["elements"])
    Rs2, Rs3, nT, w3 = _grids(p, grad)
    rs = rep_size(len(el), len(Rs2), len(Rs3), nT)
    X = np.empty((nm, ma, rs))
    dX = np.empty((nm, ma, rs, 3 * ma)) if grad else None
    rc = lib().kfo_fchl19_batch(
        _d(C), _i(Q), _i(N), nm, ma, _i(el), len(el), _d(Rs2), len(Rs2), _d(Rs3), len(Rs3), nT,
        ctypes.c_double(p["eta2"]), ctypes.c_double(p["eta3"]), ctypes.c_double(p["zeta"]),
        ctypes.c_double(p["rcut"]), ctypes.c_double(p["acut"]),
        ctypes.c_double(p["two_body_decay"]), ctypes.c_double(p["three_body_decay"]),
        ctypes.c_double(w3), _d(X), _d(dX))
    if rc:
        raise ValueError(f"oracle fchl19_batch failed ({rc})")
    return (X, dX) if grad else X


def ref_fchl19(coords, z, grad=False, **kw):
    """The reference's own generate_fchl_acsf[_and_gradients] (oracle/_ref), or None."""
    R = ref()
    if R is None:
        return None
    p = _params(kw)
    c = _f64(coords).reshape(-1)
    z = _i32(z)
    n = z.shape[0]
    el = _i32(p["elements"])
    _, _, nT, _ = _grids(p, grad)
    rs = rep_size(len(el), p["nRs2"], p["nRs3"], nT)
    rep = np.zeros((n, rs))
    g = np.zeros((n, rs, 3 * n)) if grad else None
    rc = R.kfr_fchl19(_d(c), _i(z), n, _i(el), len(el), p["nRs2"], p["nRs3"], p["nFourier"],
                      ctypes.c_double(p["eta2"]), ctypes.c_double(p["eta3"]),
                      ctypes.c_double(p["zeta"]), ctypes.c_double(p["rcut"]),
                      ctypes.c_double(p["acut"]), ctypes.c_double(p["two_body_decay"]),
                      ctypes.c_double(p["three_body_decay"]),
                      ctypes.c_double(p["three_body_weight"]), _d(rep), _d(g))
    if rc:
        raise ValueError("reference fchl19 raised")
    return (rep, g) if grad else rep


def ref_fchl19v2(coords, z, grad=False, elements=(1, 6, 7, 8, 16), **kw):
    """The reference's own fchl19v2_repr.generate[_and_gradients] (oracle/_ref), or None.
    Keywords as oracle.fchl19v2.params (two_body_type / three_body_type by name)."""
    from . import fchl19v2 as v2

    R = ref()
    if R is None:
        return None
    p = v2.params(**kw)
    c = _f64(coords).reshape(-1)
    z = _i32(z)
    n = z.shape[0]
    el = _i32(elements)
    nab = v2.nabasis(p["three_body"], p["nFourier"], p["nCosine"])
    rs = v2.rep_size(len(el), p["nRs2"], p["nRs3"], nab, p["three_body"], p["nRs3_minus"])
    rep = np.zeros((n, rs))
    g = np.zeros((n, rs, 3 * n)) if grad else None
    dbl = ctypes.c_double
    rc = R.kfr_fchl19v2(_d(c), _i(z), n, _i(el), len(el), p["nRs2"], p["nRs3"], p["nRs3_minus"],
                        p["nFourier"], p["nCosine"], dbl(p["eta2"]), dbl(p["eta3"]),
                        dbl(p["eta3_minus"]), dbl(p["zeta"]), dbl(p["rcut"]), dbl(p["acut"]),
                        dbl(p["two_body_decay"]), dbl(p["three_body_decay"]),
                        dbl(p["three_body_weight"]), v2.TWO_BODY.index(p["two_body"]),
                        v2.THREE_BODY.index(p["three_body"]), _d(rep), _d(g))
    if rc:
        raise ValueError("reference fchl19v2 raised")
    return (rep, g) if grad else rep


def ref_cho_solve_rfp(arf, y, l2=0.0):
    R = ref()
    if R is None:
        return None
    a = _f64(arf).copy()
    y = _f64(y)
    out = np.empty_like(y)
    if R.kfr_cho_solve_rfp(_d(a), _d(y), y.shape[0], ctypes.c_double(l2), _d(out)):
        raise RuntimeError("reference cho_solve_rfp raised")
    return out


def _naq(N, ma):
    return int(3 * np.sum(np.clip(N, 0, ma)))


def _check(rc):
    if rc == -4:
        raise ValueError("sigma must be > 0")
    if rc:
        raise RuntimeError(f"oracle error {rc}")


def kernel_symm_rfp(X, Q, N, sigma):
    X, Q, N = _f64(X), _i32(Q), _i32(N)
    nm, ma, rep = X.shape
    out = np.empty(nm * (nm + 1) // 2)
    _check(lib().kfo_kernel_symm_rfp(_d(X), _i(Q), _i(N), nm, ma, rep, ctypes.c_double(sigma),
                                     _d(out)))
    return out


def kernel_symm(X, Q, N, sigma):
    X, Q, N = _f64(X), _i32(Q), _i32(N)
    nm, ma, rep = X.shape
    out = np.empty((nm, nm))
    _check(lib().kfo_kernel_symm(_d(X), _i(Q), _i(N), nm, ma, rep, ctypes.c_double(sigma),
                                 _d(out)))
    return out


def kernel_rect(X1, X2, Q1, Q2, N1, N2, sigma):
    X1, X2, Q1, Q2, N1, N2 = _f64(X1), _f64(X2), _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    out = np.empty((nm1, nm2))
    _check(lib().kfo_kernel_rect(_d(X1), _d(X2), _i(Q1), _i(Q2), _i(N1), _i(N2), nm1, nm2, ma1,
                                 ma2, rep, ctypes.c_double(sigma), _d(out)))
    return out


def jacobian(X1, X2, dX2, Q1, Q2, N1, N2, sigma):
    X1, X2, dX2 = _f64(X1), _f64(X2), _f64(dX2)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    out = np.empty((nm1, _naq(N2, ma2)))
    _check(lib().kfo_jacobian(_d(X1), _d(X2), _d(dX2), _i(Q1), _i(Q2), _i(N1), _i(N2), nm1, nm2,
                              ma1, ma2, rep, ctypes.c_double(sigma), _d(out)))
    return out


def jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, sigma):
    X1, X2, dX1 = _f64(X1), _f64(X2), _f64(dX1)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    out = np.empty((_naq(N1, ma1), nm2))
    _check(lib().kfo_jacobian_t(_d(X1), _d(X2), _d(dX1), _i(Q1), _i(Q2), _i(N1), _i(N2), nm1,
                                nm2, ma1, ma2, rep, ctypes.c_double(sigma), _d(out)))
    return out


def _four(fn, X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma, with_e):
    X1, X2, dX1, dX2 = _f64(X1), _f64(X2), _f64(dX1), _f64(dX2)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    r = _naq(N1, ma1) + (nm1 if with_e else 0)
    c = _naq(N2, ma2) + (nm2 if with_e else 0)
    out = np.empty((r, c))
    _check(fn(_d(X1), _d(dX1), _d(X2), _d(dX2), _i(Q1), _i(Q2), _i(N1), _i(N2), nm1, nm2, ma1,
              ma2, rep, ctypes.c_double(sigma), _d(out)))
    return out


def hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma):
    return _four(lib().kfo_hessian, X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma, False)


def full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma):
    return _four(lib().kfo_full, X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma, True)


def _symm2(fn, X, dX, Q, N, sigma, size):
    X, dX, Q, N = _f64(X), _f64(dX), _i32(Q), _i32(N)
    nm, ma, rep = X.shape
    out = np.empty(size(nm, _naq(N, ma)))
    _check(fn(_d(X), _d(dX), _i(Q), _i(N), nm, ma, rep, ctypes.c_double(sigma), _d(out)))
    return out


def hessian_symm(X, dX, Q, N, sigma):
    return _symm2(lib().kfo_hessian_symm, X, dX, Q, N, sigma, lambda nm, naq: (naq, naq))


def hessian_symm_rfp(X, dX, Q, N, sigma):
    return _symm2(lib().kfo_hessian_symm_rfp, X, dX, Q, N, sigma,
                  lambda nm, naq: naq * (naq + 1) // 2)


def full_symm(X, dX, Q, N, sigma):
    return _symm2(lib().kfo_full_symm, X, dX, Q, N, sigma,
                  lambda nm, naq: (nm + naq, nm + naq))


def full_symm_rfp(X, dX, Q, N, sigma):
    return _symm2(lib().kfo_full_symm_rfp, X, dX, Q, N, sigma,
                  lambda nm, naq: (nm + naq) * (nm + naq + 1) // 2)


def alpha_desc(dX, Q, N, alpha):
    dX, Q, N, alpha = _f64(dX), _i32(Q), _i32(N), _f64(alpha)
    nm, ma, rep, _ = dX.shape
    out = np.empty((nm, ma, rep))
    _check(lib().kfo_alpha_desc(_d(dX), _i(Q), _i(N), nm, ma, rep, _d(alpha), _d(out)))
    return out


def full_matvec(X1, dX1, X2, ad, aE, Q1, Q2, N1, N2, sigma):
    X1, dX1, X2, ad, aE = _f64(X1), _f64(dX1), _f64(X2), _f64(ad), _f64(aE)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    E = np.empty(nm1)
    F = np.empty(_naq(N1, ma1))
    _check(lib().kfo_full_matvec(_d(X1), _d(dX1), _d(X2), _d(ad), _d(aE), _i(Q1), _i(Q2), _i(N1),
                                 _i(N2), nm1, nm2, ma1, ma2, rep, ctypes.c_double(sigma), _d(E),
                                 _d(F)))
    return E, F


def hessian_matvec(X1, dX1, X2, ad, Q1, Q2, N1, N2, sigma):
    X1, dX1, X2, ad = _f64(X1), _f64(dX1), _f64(X2), _f64(ad)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    F = np.empty(_naq(N1, ma1))
    _check(lib().kfo_hessian_matvec(_d(X1), _d(dX1), _d(X2), _d(ad), _i(Q1), _i(Q2), _i(N1),
                                    _i(N2), nm1, nm2, ma1, ma2, rep, ctypes.c_double(sigma),
                                    _d(F)))
    return F


def jacobian_t_matvec(X1, X2, ad, Q1, Q2, N1, N2, sigma):
    X1, X2, ad = _f64(X1), _f64(X2), _f64(ad)
    Q1, Q2, N1, N2 = _i32(Q1), _i32(Q2), _i32(N1), _i32(N2)
    nm1, ma1, rep = X1.shape
    nm2, ma2, _ = X2.shape
    E = np.empty(nm1)
    _check(lib().kfo_jacobian_t_matvec(_d(X1), _d(X2), _d(ad), _i(Q1), _i(Q2), _i(N1), _i(N2),
                                       nm1, nm2, ma1, ma2, rep, ctypes.c_double(sigma), _d(E)))
    return E


def cho_solve_rfp(arf, y, l2=0.0):
    """Oracle cho_solve_rfp on a copy of arf (the product API overwrites its input)."""
    a = _f64(arf).copy()
    y = _f64(y)
    out = np.empty_like(y)
    info = lib().kfo_cho_solve_rfp(_d(a), _d(y), y.shape[0], ctypes.c_double(l2), _d(out))
    if info:
        raise RuntimeError(f"DPFTRF failed, info={info}")
    return out


def rfp_index(n, i, j):
    return int(lib().kfo_rfp_index(n, i, j))
