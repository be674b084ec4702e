"""GPU parity: RFP Cholesky (kf_rfp_potrf / kf_rfp_potrs) and RFP conversions vs LAPACK
(scipy), numpy.linalg.solve and the reference math.cpp golden."""

import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def km():
    from kernelforge_amd import kernelmath

    return kernelmath


def spd(n, seed, shift=1.0):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    return A @ A.T / n + shift * np.eye(n)


def to_rfp(K):
    trttf = sl.get_lapack_funcs("trttf", (K,))
    arf, info = trttf(np.asfortranarray(K), transr="N", uplo="U")
    assert info == 0
    return arf


def test_golden(km):
    g = np.load(os.path.join(ROOT, "tests", "golden", "rfp_ref.npz"))
    a = km.cho_solve_rfp(g["arf"].copy(), g["y"], l2=float(g["l2"]))
    assert rel_err(a, g["alpha"]) < 1e-12


@pytest.mark.parametrize("n", [1, 2, 3, 63, 64, 65, 128, 129, 200, 511, 1001, 1500, 2311])
def test_cho_solve_rfp_and_factor(km, n):
    K = spd(n, n)
    y = np.random.default_rng(1).standard_normal(n)
    arf = to_rfp(K)
    work = arf.copy()
    a = km.cho_solve_rfp(work, y, l2=1e-3)
    ref = np.linalg.solve(K + 1e-3 * np.eye(n), y)
    assert rel_err(a, ref) < 1e-10
    # factor overwrites K_rfp exactly where LAPACK dpftrf leaves it
    pftrf = sl.get_lapack_funcs("pftrf", (K,))
    arf_l = arf.copy()
    for j in range(n):
        arf_l[oracle.rfp_index(n, j, j)] += 1e-3
    fac, info = pftrf(n, arf_l, transr="N", uplo="U")
    assert info == 0
    assert rel_err(work, fac) < 1e-11


def test_cho_solve_rfp_wide_panels(km):
    """Order 16,000 (250 blocks of 64): the first outer panels are 512 columns wide (while
    more than 200 blocks remain), the rest 256.  A symmetric, strictly diagonally dominant
    matrix (cheap to build at this size); the solution is checked by its residual."""
    n = 16000
    rng = np.random.default_rng(16)
    K = rng.random((n, n)) - 0.5
    K = K + K.T
    K[np.diag_indices(n)] += n
    y = rng.standard_normal(n)
    a = km.cho_solve_rfp(to_rfp(K), y, l2=0.0)
    r = K @ a - y
    assert np.linalg.norm(r) / np.linalg.norm(y) < 1e-12


@pytest.mark.parametrize("n,col", [(70, 40), (70, 0), (70, 5), (70, 16), (70, 63), (70, 64),
                                   (70, 69), (301, 149), (301, 150), (301, 200), (1001, 700)])
def test_not_spd_info_matches_lapack(km, n, col):
    """A negative diagonal entry at column col: the factorization stops there with LAPACK's
    info (1-based column), whichever 16-column quarter of a 64-wide diagonal tile, partial
    tile (70 and 301 end in one) or RFP region (the upper region starts at n / 2) it is in."""
    K = spd(n, 3)
    K[col, col] = -5.0
    arf = to_rfp(K)
    pftrf = sl.get_lapack_funcs("pftrf", (K,))
    _, info = pftrf(n, arf.copy(), transr="N", uplo="U")
    assert info == col + 1
    with pytest.raises(RuntimeError, match=f"DPFTRF failed, info={info}"):
        km.cho_solve_rfp(arf.copy(), np.ones(n), 0.0)


def test_solve_cholesky(km):
    n = 150
    K = spd(n, 9)
    y = np.arange(n, dtype=np.float64)
    K0 = K.copy()
    a = km.solve_cholesky(K, y, regularize=0.1)
    assert rel_err(a, np.linalg.solve(K0 + 0.1 * np.eye(n), y)) < 1e-10
    np.testing.assert_array_equal(K, K0)
    with pytest.raises(RuntimeError):
        km.solve_cholesky(np.ones((3, 4)), np.ones(3))


@pytest.mark.parametrize("n", [1, 6, 7, 64, 101])
@pytest.mark.parametrize("uplo", ["U", "L"])
@pytest.mark.parametrize("transr", ["N", "T"])
def test_rfp_conversions(km, n, uplo, transr):
    K = spd(n, n + 5)
    arf = km.full_to_rfp(K, uplo=uplo, transr=transr)
    # reference semantics: dtrttf on the C-order buffer with UPLO swapped (math_bindings.cpp:89-120)
    trttf = sl.get_lapack_funcs("trttf", (K,))
    ref, info = trttf(np.asfortranarray(K.T), transr=transr, uplo="L" if uplo == "U" else "U")
    np.testing.assert_array_equal(arf, ref)
    back = km.rfp_to_full(arf, n, uplo=uplo, transr=transr)
    np.testing.assert_array_equal(back, K)


def test_kernel_layout_is_full_to_rfp_lower(km):
    K = spd(9, 2)
    np.testing.assert_array_equal(km.full_to_rfp(K, uplo="L"), to_rfp(K))


@pytest.mark.parametrize("n", [300, 1300])
def test_fp32_potrf_potrs_c_abi(n):
    from kernelforge_amd import capi

    K = spd(n, 4, shift=2.0)
    arf = to_rfp(K)
    y = np.random.default_rng(2).standard_normal(n)
    a32 = capi.DeviceArray.from_numpy(arf.astype(np.float32))
    b = capi.DeviceArray.from_numpy(y)
    capi.rfp_potrf(capi.KF_PREC_F32, a32.ptr, n, 0.0)
    capi.rfp_potrs(capi.KF_PREC_F32, a32.ptr, n, b.ptr, 1)
    capi.synchronize()
    assert rel_err(b.numpy(), np.linalg.solve(K, y)) < 1e-4


@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_potrs_repeated_on_same_buffers(prec):
    """Repeated solves on the same factor and right-hand-side buffers with new right-hand
    sides, with other solves and host synchronisations in between, each match numpy (the
    inverse diagonal blocks and the intermediate vector live in a reused workspace)."""
    from kernelforge_amd import capi

    n, n2 = 700, 333
    p = capi.KF_PREC_F64 if prec == "f64" else capi.KF_PREC_F32
    dt = np.float64 if prec == "f64" else np.float32
    tol = 1e-10 if prec == "f64" else 1e-4
    K, K2 = spd(n, 7, shift=2.0), spd(n2, 8, shift=2.0)
    A = capi.DeviceArray.from_numpy(to_rfp(K).astype(dt))
    A2 = capi.DeviceArray.from_numpy(to_rfp(K2).astype(dt))
    capi.rfp_potrf(p, A.ptr, n, 0.0)
    capi.rfp_potrf(p, A2.ptr, n2, 0.0)
    B = capi.DeviceArray((n,))
    rng = np.random.default_rng(3)
    for rep in range(4):
        y = rng.standard_normal(n)
        B.copy_from(capi.DeviceArray.from_numpy(y))
        capi.rfp_potrs(p, A.ptr, n, B.ptr, 1)
        capi.synchronize()
        assert rel_err(B.numpy(), np.linalg.solve(K, y)) < tol, rep
        y2 = rng.standard_normal(n2)  # another solve in between, on fresh buffers
        B2 = capi.DeviceArray.from_numpy(y2)
        capi.rfp_potrs(p, A2.ptr, n2, B2.ptr, 1)
        capi.synchronize()
        assert rel_err(B2.numpy(), np.linalg.solve(K2, y2)) < tol, rep


@pytest.mark.parametrize("n", [63, 300, 1027])
def test_potrs_several_right_hand_sides(n):
    """kf_rfp_potrs with nrhs = 3 (right-hand sides stored one after another)."""
    from kernelforge_amd import capi

    K = spd(n, 9, shift=1.5)
    Y = np.random.default_rng(4).standard_normal((3, n))
    A = capi.DeviceArray.from_numpy(to_rfp(K))
    B = capi.DeviceArray.from_numpy(Y.copy())
    capi.rfp_potrf(capi.KF_PREC_F64, A.ptr, n, 0.0)
    capi.rfp_potrs(capi.KF_PREC_F64, A.ptr, n, B.ptr, 3)
    capi.synchronize()
    X = B.numpy().reshape(3, n)
    for q in range(3):
        assert rel_err(X[q], np.linalg.solve(K, Y[q])) < 1e-10, q
