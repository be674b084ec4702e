"""Pins the oracle's local-kernel and RFP-solve restatements (oracle/kf_oracle.c).

What fixes each function independently of the oracle's own formulas:
  scalar kernels  -- brute-force triple loop over atom pairs (the reference test
                     tests/test_local_scalar_kernels.py:19-85), multi-tile bookkeeping
                     (tile edge forced to 3/7 rows), known answers (1-atom diagonal = 1,
                     disjoint labels = 0), RFP layout via LAPACK dtrttf;
  jacobian(_t)    -- central finite differences of K(X(R)) through the descriptor
                     (dK/dR2 and -dK/dR1), transpose identity (local_kernels.hpp:40-43);
  hessian         -- finite differences of the jacobian; symmetry of the symm form;
  full(_symm/_rfp)-- block identities (tests/test_fchl19_full_kernel.py:36-342), RFP == dense;
  alpha_desc      -- einsum over the Jacobian;
  matvecs         -- dense full kernel @ alpha (tests/test_local_jacobian_t_matvec.py:59-79);
  cho_solve_rfp   -- numpy.linalg.solve, scipy dpftrf/dpftrs, and the reference math.cpp
                     golden (tests/golden/rfp_ref.npz).
"""

import os

import numpy as np
import pytest
import scipy.linalg as sl

import oracle
from kernelforge_amd import synthetic

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def brute_kernel(x1, x2, q1, q2, n1, n2, sigma):
    K = np.zeros((x1.shape[0], x2.shape[0]))
    for a in range(x1.shape[0]):
        for b in range(x2.shape[0]):
            for i in range(int(np.clip(n1[a], 0, x1.shape[1]))):
                for j in range(int(np.clip(n2[b], 0, x2.shape[1]))):
                    if q1[a, i] == q2[b, j]:
                        d = x1[a, i] - x2[b, j]
                        K[a, b] += np.exp(-d @ d / (2 * sigma * sigma))
    return K


def random_set(nm, ma, rep, nlab, seed, scale=0.4):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nm, ma, rep)) * scale
    q = rng.integers(0, nlab, (nm, ma)).astype(np.int32)
    n = rng.integers(1, ma + 1, nm).astype(np.int32)
    return x, q, n


def to_rfp(K):
    n = K.shape[0]
    arf = np.zeros(n * (n + 1) // 2)
    for i in range(n):
        for j in range(i, n):
            arf[oracle.rfp_index(n, i, j)] = K[i, j]
    return arf


@pytest.mark.parametrize("tile", [8192, 3, 7])
def test_scalar_kernels_match_brute_force(tile):
    x, q, n = random_set(13, 5, 9, 3, 1)
    x2, q2, n2 = random_set(7, 4, 9, 3, 2)
    oracle.lib().kfo_set_tile(tile)
    try:
        K = brute_kernel(x, x, q, q, n, n, 1.3)
        np.testing.assert_allclose(oracle.kernel_symm(x, q, n, 1.3), K, rtol=1e-12, atol=1e-13)
        np.testing.assert_allclose(oracle.kernel_symm_rfp(x, q, n, 1.3), to_rfp(K), rtol=1e-12,
                                   atol=1e-13)
        np.testing.assert_allclose(oracle.kernel_rect(x, x2, q, q2, n, n2, 1.3),
                                   brute_kernel(x, x2, q, q2, n, n2, 1.3), rtol=1e-12, atol=1e-13)
    finally:
        oracle.lib().kfo_set_tile(8192)


def test_rfp_index_is_lapack_layout():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 10):
        A = rng.standard_normal((n, n))
        A = A + A.T
        trttf = sl.get_lapack_funcs("trttf", (A,))
        arf, info = trttf(np.asfortranarray(A), transr="N", uplo="U")
        np.testing.assert_array_equal(to_rfp(A), arf)


def test_known_answers():
    x = np.zeros((2, 1, 4))
    x[1, 0, 0] = 1.0
    q = np.array([[1], [1]], np.int32)
    n = np.array([1, 1], np.int32)
    K = oracle.kernel_symm(x, q, n, 1 / np.sqrt(2))
    np.testing.assert_allclose(K, [[1, np.exp(-1)], [np.exp(-1), 1]], rtol=1e-15)
    q2 = np.array([[1], [2]], np.int32)
    assert oracle.kernel_symm(x, q2, n, 1.0)[0, 1] == 0.0


def _molset(nm, seed, n_atoms=None):
    coords, zs, _ = synthetic.random_molecules(nm, seed)
    C, Q, N = synthetic.pad_batch(coords, zs)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    return C, Q, N, X, dX


def _k_of_coords(C1, Q1, N1, C2, Q2, N2, sigma):
    X1 = oracle.fchl19_batch(C1, Q1, N1)
    X2 = oracle.fchl19_batch(C2, Q2, N2)
    return oracle.kernel_rect(X1, X2, Q1, Q2, N1, N2, sigma)


def test_jacobians_are_coordinate_derivatives():
    sigma, h = 2.0, 1e-5
    C1, Q1, N1, X1, dX1 = _molset(2, 21)
    C2, Q2, N2, X2, dX2 = _molset(3, 22)
    J = oracle.jacobian(X1, X2, dX2, Q1, Q2, N1, N2, sigma)        # dK/dR2
    Jt = oracle.jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, sigma)     # -dK/dR1
    offs2 = np.concatenate([[0], np.cumsum(3 * N2)])
    for b in range(len(N2)):
        for a_ in range(N2[b]):
            for d in range(3):
                Cp, Cm = C2.copy(), C2.copy()
                Cp[b, a_, d] += h
                Cm[b, a_, d] -= h
                fd = (_k_of_coords(C1, Q1, N1, Cp, Q2, N2, sigma)
                      - _k_of_coords(C1, Q1, N1, Cm, Q2, N2, sigma)) / (2 * h)
                np.testing.assert_allclose(J[:, offs2[b] + 3 * a_ + d], fd[:, b], rtol=2e-6,
                                           atol=1e-8)
    offs1 = np.concatenate([[0], np.cumsum(3 * N1)])
    for a in range(len(N1)):
        for a_ in range(N1[a]):
            d = a_ % 3
            Cp, Cm = C1.copy(), C1.copy()
            Cp[a, a_, d] += h
            Cm[a, a_, d] -= h
            fd = (_k_of_coords(Cp, Q1, N1, C2, Q2, N2, sigma)
                  - _k_of_coords(Cm, Q1, N1, C2, Q2, N2, sigma)) / (2 * h)
            np.testing.assert_allclose(-Jt[offs1[a] + 3 * a_ + d], fd[a], rtol=2e-6, atol=1e-8)
    # transpose identity
    np.testing.assert_allclose(Jt, -oracle.jacobian(X2, X1, dX1, Q2, Q1, N2, N1, sigma).T,
                               rtol=1e-12, atol=1e-14)


def test_hessian_is_derivative_of_jacobian():
    sigma, h = 2.0, 1e-5
    C1, Q1, N1, X1, dX1 = _molset(2, 31)
    C2, Q2, N2, X2, dX2 = _molset(2, 32)
    H = oracle.hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, sigma)
    offs1 = np.concatenate([[0], np.cumsum(3 * N1)])
    for a in range(len(N1)):
        for a_ in (0, N1[a] - 1):
            for d in range(3):
                Cp, Cm = C1.copy(), C1.copy()
                Cp[a, a_, d] += h
                Cm[a, a_, d] -= h
                Jp = oracle.jacobian(oracle.fchl19_batch(Cp, Q1, N1), X2, dX2, Q1, Q2, N1, N2,
                                     sigma)
                Jm = oracle.jacobian(oracle.fchl19_batch(Cm, Q1, N1), X2, dX2, Q1, Q2, N1, N2,
                                     sigma)
                fd = (Jp[a] - Jm[a]) / (2 * h)
                np.testing.assert_allclose(H[offs1[a] + 3 * a_ + d], fd, rtol=5e-6, atol=1e-8)


@pytest.mark.parametrize("seed", [0, 42])
def test_full_kernel_blocks(seed):
    rng = np.random.default_rng(seed)
    nm, ma, rep, sigma = 4, 4, 7, 1.5
    x = rng.normal(size=(nm, ma, rep))
    dx = rng.normal(size=(nm, ma, rep, 3 * ma))
    q = rng.integers(0, 3, size=(nm, ma)).astype(np.int32)
    n = rng.integers(1, ma + 1, size=nm).astype(np.int32)
    naq = int(3 * n.sum())
    F = oracle.full(x, dx, x, dx, q, q, n, n, sigma)
    np.testing.assert_allclose(F[:nm, :nm], oracle.kernel_rect(x, x, q, q, n, n, sigma),
                               rtol=1e-12, atol=1e-14)
    Jt = oracle.jacobian_t(x, dx, x, q, q, n, n, sigma)
    np.testing.assert_allclose(F[nm:, :nm], -Jt, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(F[:nm, nm:], -Jt.T, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(F[nm:, nm:], oracle.hessian(x, dx, x, dx, q, q, n, n, sigma),
                               rtol=1e-12, atol=1e-14)
    Fs = oracle.full_symm(x, dx, q, n, sigma)
    np.testing.assert_allclose(Fs, F, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(Fs, Fs.T, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(oracle.full_symm_rfp(x, dx, q, n, sigma), to_rfp(Fs), rtol=1e-12,
                               atol=1e-14)
    Hs = oracle.hessian_symm(x, dx, q, n, sigma)
    np.testing.assert_allclose(Hs, F[nm:, nm:], rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(oracle.hessian_symm_rfp(x, dx, q, n, sigma), to_rfp(Hs),
                               rtol=1e-12, atol=1e-14)
    assert F.shape == (nm + naq, nm + naq)


def test_alpha_desc_and_matvecs():
    rng = np.random.default_rng(5)
    nm, ma, rep, sigma = 5, 4, 6, 1.2
    x = rng.normal(size=(nm, ma, rep)) * 0.5
    dx = rng.normal(size=(nm, ma, rep, 3 * ma))
    q = rng.integers(0, 2, size=(nm, ma)).astype(np.int32)
    n = rng.integers(1, ma + 1, size=nm).astype(np.int32)
    naq = int(3 * n.sum())
    aE = rng.normal(size=nm)
    aF = rng.normal(size=naq)
    ad = oracle.alpha_desc(dx, q, n, aF)
    offs = np.concatenate([[0], np.cumsum(3 * n)])
    for b in range(nm):
        ref = np.einsum("irc,c->ir", dx[b, : n[b], :, : 3 * n[b]], aF[offs[b]: offs[b + 1]])
        np.testing.assert_allclose(ad[b, : n[b]], ref, rtol=1e-12, atol=1e-14)
        assert np.all(ad[b, n[b]:] == 0)
    F = oracle.full_symm(x, dx, q, n, sigma)
    E, Fv = oracle.full_matvec(x, dx, x, ad, aE, q, q, n, n, sigma)
    full = F @ np.concatenate([aE, aF])
    np.testing.assert_allclose(E, full[:nm], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(Fv, full[nm:], rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.hessian_matvec(x, dx, x, ad, q, q, n, n, sigma),
                               F[nm:, nm:] @ aF, rtol=1e-11, atol=1e-12)
    np.testing.assert_allclose(oracle.jacobian_t_matvec(x, x, ad, q, q, n, n, sigma),
                               F[:nm, nm:] @ aF, rtol=1e-11, atol=1e-12)


def test_cho_solve_rfp():
    g = np.load(os.path.join(ROOT, "tests", "golden", "rfp_ref.npz"))
    a = oracle.cho_solve_rfp(g["arf"], g["y"], float(g["l2"]))
    np.testing.assert_allclose(a, g["alpha"], rtol=1e-12, atol=1e-14)
    n = g["y"].shape[0]
    K = sl.get_lapack_funcs("tfttr", (g["arf"],))(n, g["arf"], transr="N", uplo="U")[0]
    K = np.triu(K) + np.triu(K, 1).T
    np.testing.assert_allclose(a, np.linalg.solve(K + float(g["l2"]) * np.eye(n), g["y"]),
                               rtol=1e-10)


def test_cho_solve_rfp_not_spd_raises():
    K = -np.eye(4)
    with pytest.raises(RuntimeError, match="DPFTRF failed"):
        oracle.cho_solve_rfp(to_rfp(K), np.ones(4), 0.0)
