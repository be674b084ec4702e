"""GPU parity: energy+force forms of the local kernel (jacobian, jacobian_t, hessian,
full, and the symmetric dense / RFP forms) vs the FP64 oracle.

Bar: every entry within 1e-10 relative to the largest |entry| of the oracle's block
(FP64 throughout; the GPU sums in a different order, so bit equality is not expected).
Random inputs use random (not physical) Jacobians, which exercises the algebra of every
block independently; the conformer case uses real FCHL19 descriptors and Jacobians."""

import numpy as np
import pytest

import oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10


@pytest.fixture(scope="module")
def lk():
    from kernelforge_amd import local_kernels

    return local_kernels


def rand_set(nm, ma, rep, nlab, seed, nmin=0):
    rng = np.random.default_rng(seed)
    X = rng.standard_normal((nm, ma, rep)) * 0.3
    dX = rng.standard_normal((nm, ma, rep, 3 * ma)) * 0.2
    Q = rng.integers(1, nlab + 1, (nm, ma)).astype(np.int32)
    N = rng.integers(nmin, ma + 1, nm).astype(np.int32)
    return X, dX, Q, N


# (nm1, ma1, nm2, ma2, rep, nlab): single-atom molecules (one 16-column block), 2 and 4
# force-column blocks, more than 64 columns (a second round of column blocks), the
# 32-atom maximum, and more partners per molecule than one projection group holds (8).
CASES = [(3, 1, 4, 2, 9, 1), (5, 3, 6, 4, 33, 2), (4, 7, 5, 9, 40, 3), (3, 21, 4, 21, 72, 3),
         (2, 23, 3, 25, 50, 2), (2, 32, 2, 32, 24, 2), (3, 5, 20, 7, 16, 2)]


@pytest.mark.parametrize("case", CASES)
def test_rect_forms(lk, case):
    nm1, ma1, nm2, ma2, rep, nlab = case
    X1, dX1, Q1, N1 = rand_set(nm1, ma1, rep, nlab, 1 + ma1)
    X2, dX2, Q2, N2 = rand_set(nm2, ma2, rep, nlab, 100 + ma2)
    s = 1.7
    got = lk.kernel_gaussian_jacobian(X1, X2, dX2, Q1, Q2, N1, N2, s)
    assert got.shape == (nm1, 3 * int(np.clip(N2, 0, ma2).sum()))
    assert rel_err(got, oracle.jacobian(X1, X2, dX2, Q1, Q2, N1, N2, s)) < TOL
    got = lk.kernel_gaussian_jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, s)
    assert rel_err(got, oracle.jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, s)) < TOL
    got = lk.kernel_gaussian_hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)
    assert rel_err(got, oracle.hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)) < TOL
    got = lk.kernel_gaussian_full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)
    ref = oracle.full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)
    assert got.shape == ref.shape
    assert rel_err(got, ref) < TOL


@pytest.mark.parametrize("case", [(4, 2, 10, 2), (7, 6, 33, 3), (5, 21, 64, 3), (3, 24, 20, 2),
                                  (2, 32, 12, 1), (19, 6, 20, 3)])
def test_symm_forms(lk, case):
    nm, ma, rep, nlab = case
    X, dX, Q, N = rand_set(nm, ma, rep, nlab, 7 * ma + nm)
    s = 1.3
    for name, fn in (("kernel_gaussian_hessian_symm", oracle.hessian_symm),
                     ("kernel_gaussian_hessian_symm_rfp", oracle.hessian_symm_rfp),
                     ("kernel_gaussian_full_symm", oracle.full_symm),
                     ("kernel_gaussian_full_symm_rfp", oracle.full_symm_rfp)):
        got = getattr(lk, name)(X, dX, Q, N, s)
        ref = fn(X, dX, Q, N, s)
        assert got.shape == ref.shape, name
        assert rel_err(got, ref) < TOL, name


def test_zero_columns(lk):
    """Columns that are zero in X and in the Jacobian for every atom of both sets are
    dropped before the kernels run; columns zero in X but not in the Jacobian stay (their
    Jacobian rows still enter the force-force blocks)."""
    X1, dX1, Q1, N1 = rand_set(4, 6, 50, 2, 31)
    X2, dX2, Q2, N2 = rand_set(5, 5, 50, 2, 32)
    x_only = [3, 4, 17, 30, 31, 32]     # zero in X, Jacobian kept
    both = [0, 1, 2, 10, 11, 40, 49]    # zero in X and in the Jacobian
    for X, dX in ((X1, dX1), (X2, dX2)):
        X[..., x_only + both] = 0.0
        dX[..., both, :] = 0.0
    dX2[..., 5, :] = 0.0                # zero in one set's Jacobian only: kept
    s = 1.3
    assert rel_err(lk.kernel_gaussian_full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s),
                   oracle.full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)) < TOL
    assert rel_err(lk.kernel_gaussian_jacobian(X1, X2, dX2, Q1, Q2, N1, N2, s),
                   oracle.jacobian(X1, X2, dX2, Q1, Q2, N1, N2, s)) < TOL
    got = lk.kernel_gaussian_full_symm_rfp(X1, dX1, Q1, N1, s)
    assert rel_err(got, oracle.full_symm_rfp(X1, dX1, Q1, N1, s)) < TOL
    got = lk.kernel_gaussian_hessian_symm(X2, dX2, Q2, N2, s)
    assert rel_err(got, oracle.hessian_symm(X2, dX2, Q2, N2, s)) < TOL
    # every column zero: the kernel of the constant descriptor
    Z, dZ = np.zeros_like(X1), np.zeros_like(dX1)
    assert rel_err(lk.kernel_gaussian_full_symm(Z, dZ, Q1, N1, s),
                   oracle.full_symm(Z, dZ, Q1, N1, s)) < TOL


def test_full_symm_conformers(lk):
    """Real FCHL19 descriptors + Jacobians of 10 C9H8O4 conformers: the RFP E+F
    kernel, its dense form, and the block identities of the full kernel."""
    from kernelforge_amd import synthetic

    base, z = synthetic.aspirin_base()
    C, Z, _, _ = synthetic.conformers(base, z, 10, 5)
    Q = np.tile(Z, (10, 1)).astype(np.int32)
    N = np.full(10, 21, np.int32)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    arf = lk.kernel_gaussian_full_symm_rfp(X, dX, Q, N, 2.0)
    assert rel_err(arf, oracle.full_symm_rfp(X, dX, Q, N, 2.0)) < TOL
    K = lk.kernel_gaussian_full_symm(X, dX, Q, N, 2.0)
    nm = 10
    # blocks of the symmetric full kernel against the standalone forms
    assert rel_err(K[:nm, :nm], lk.kernel_gaussian_symm(X, Q, N, 2.0)) < TOL
    jt = lk.kernel_gaussian_jacobian_t(X, dX, X, Q, Q, N, N, 2.0)
    assert rel_err(K[nm:, :nm], -jt) < TOL            # FE = dK/dR1 = -jacobian_t
    assert rel_err(K[:nm, nm:], -jt.T) < TOL          # EF = dK/dR2
    assert rel_err(K[nm:, nm:], lk.kernel_gaussian_hessian_symm(X, dX, Q, N, 2.0)) < TOL
    assert rel_err(K, K.T) < 1e-13


def test_zero_atoms_and_no_shared_labels(lk):
    X, dX, Q, N = rand_set(4, 3, 8, 2, 3, nmin=0)
    N[:] = [0, 2, 0, 3]
    got = lk.kernel_gaussian_full_symm_rfp(X, dX, Q, N, 1.1)
    assert rel_err(got, oracle.full_symm_rfp(X, dX, Q, N, 1.1)) < TOL
    K = lk.kernel_gaussian_full(X, dX, X, dX, Q, Q + 10, N, N, 1.1)
    assert np.all(K == 0.0)
    N0 = np.zeros(4, np.int32)
    assert lk.kernel_gaussian_hessian(X, dX, X, dX, Q, Q, N0, N, 1.0).shape == (0, 15)
    assert lk.kernel_gaussian_jacobian_t(X, dX, X, Q, Q, N0, N, 1.0).shape == (0, 4)


def test_errors(lk):
    X, dX, Q, N = rand_set(3, 2, 6, 2, 4)
    with pytest.raises(ValueError):
        lk.kernel_gaussian_full_symm(X, dX, Q, N, 0.0)
    with pytest.raises(ValueError):
        lk.kernel_gaussian_full_symm(X, dX[..., :5], Q, N, 1.0)
    with pytest.raises(ValueError):
        lk.kernel_gaussian_hessian(X, dX, X[..., :4], dX, Q, Q, N, N, 1.0)
    # more than 32 atoms per molecule is no error: the size-general kernel takes it
    Xb, dXb, Qb, Nb = rand_set(2, 33, 4, 1, 5)
    assert rel_err(lk.kernel_gaussian_hessian_symm(Xb, dXb, Qb, Nb, 1.0),
                   oracle.hessian_symm(Xb, dXb, Qb, Nb, 1.0)) < 1e-12


def test_keywords(lk):
    X, dX, Q, N = rand_set(3, 2, 6, 2, 6)
    a = lk.kernel_gaussian_full(X1=X, dX1=dX, X2=X, dX2=dX, Q1=Q, Q2=Q, N1=N, N2=N, sigma=1.2)
    np.testing.assert_array_equal(a, lk.kernel_gaussian_full(X, dX, X, dX, Q, Q, N, N, 1.2))
    b = lk.kernel_gaussian_jacobian(X1=X, X2=X, dX2=dX, Q1=Q, Q2=Q, N1=N, N2=N, sigma=1.2)
    np.testing.assert_array_equal(b, lk.kernel_gaussian_jacobian(X, X, dX, Q, Q, N, N, 1.2))


# ---------------------------------------------------------------------------
# alpha_desc and the matrix-free matvecs (the PCG operator) vs the oracle and
# vs the explicit full kernel: K_full @ [aE; aF] == full_matvec(aE, alpha_desc(aF)).
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("case", [(4, 3, 5, 4, 17, 2), (6, 9, 7, 8, 40, 3), (3, 21, 4, 21, 72, 3),
                                  (2, 30, 3, 26, 20, 2)])
def test_alpha_desc_and_matvecs(lk, case):
    nm1, ma1, nm2, ma2, rep, nlab = case
    X1, dX1, Q1, N1 = rand_set(nm1, ma1, rep, nlab, 11 + ma1)
    X2, dX2, Q2, N2 = rand_set(nm2, ma2, rep, nlab, 111 + ma2)
    rng = np.random.default_rng(5)
    naq2 = 3 * int(np.clip(N2, 0, ma2).sum())
    aF, aE = rng.standard_normal(naq2), rng.standard_normal(nm2)
    ad = lk.kernel_gaussian_local_compute_alpha_desc(dX2, Q2, N2, aF)
    ad_o = oracle.alpha_desc(dX2, Q2, N2, aF)
    assert ad.shape == ad_o.shape and rel_err(ad, ad_o) < 1e-13
    s = 1.6
    E, F = lk.kernel_gaussian_local_full_matvec(X1, dX1, X2, ad_o, aE, Q1, Q2, N1, N2, s)
    Eo, Fo = oracle.full_matvec(X1, dX1, X2, ad_o, aE, Q1, Q2, N1, N2, s)
    assert rel_err(E, Eo) < TOL and rel_err(F, Fo) < TOL
    Fh = lk.kernel_gaussian_local_hessian_matvec(X1, dX1, X2, ad_o, Q1, Q2, N1, N2, s)
    assert rel_err(Fh, oracle.hessian_matvec(X1, dX1, X2, ad_o, Q1, Q2, N1, N2, s)) < TOL
    Ej = lk.kernel_gaussian_local_jacobian_t_matvec(X1, X2, ad_o, Q1, Q2, N1, N2, s)
    assert rel_err(Ej, oracle.jacobian_t_matvec(X1, X2, ad_o, Q1, Q2, N1, N2, s)) < TOL
    # the matvec is the explicit full kernel applied to [aE; aF]
    K = lk.kernel_gaussian_full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, s)
    y = K @ np.concatenate([aE, aF])
    assert rel_err(np.concatenate([E, F]), y) < 1e-9


def test_matvec_conformers_symmetric(lk):
    """Training-set operator on C9H8O4 conformers equals the RFP E+F kernel times alpha."""
    from kernelforge_amd import synthetic

    base, z = synthetic.aspirin_base()
    C, Z, _, _ = synthetic.conformers(base, z, 8, 9)
    Q = np.tile(Z, (8, 1)).astype(np.int32)
    N = np.full(8, 21, np.int32)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    rng = np.random.default_rng(2)
    a = rng.standard_normal(8 + 8 * 63)
    ad = lk.kernel_gaussian_local_compute_alpha_desc(dX, Q, N, a[8:])
    E, F = lk.kernel_gaussian_local_full_matvec(X, dX, X, ad, a[:8], Q, Q, N, N, 2.0)
    K = oracle.full_symm(X, dX, Q, N, 2.0)
    assert rel_err(np.concatenate([E, F]), K @ a) < 1e-9
