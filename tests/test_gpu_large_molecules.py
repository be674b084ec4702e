"""GPU parity for molecules beyond the fixed-size fast kernels: 33 to 48 atoms in the
derivative forms and matvecs (the MFMA kernels take up to 32 atoms per molecule, larger
ones run the size-general kernel), Jacobians whose LDS staging does not fit one
workgroup, and a 130-atom representation (more than the 128 atoms the FCHL19 kernels
stage; both route to the FCHL19v2 kernel with the FCHL19 bases).  Against the oracle,
which has no size limits.  Bars as in the fixed-size tests."""

import numpy as np
import pytest

import oracle
from tests.helpers import rel_err, row_rel_err

pytestmark = pytest.mark.gpu
SIGMA = 2.0


def molecule(n, seed, elements=(1, 6, 7, 8)):
    rng = np.random.default_rng(seed)
    side = 1.6 * n ** (1 / 3)
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, side, 3)
        if all(np.linalg.norm(p - q) > 0.9 for q in pts):
            pts.append(p)
    return np.array(pts), rng.choice(elements, n).astype(np.int32)


def padded(mols):
    from kernelforge_amd import synthetic

    return synthetic.pad_batch([c for c, _ in mols], [z for _, z in mols])


@pytest.fixture(scope="module")
def sets():
    """Set 1: 33- and 48-atom molecules; set 2: 40 atoms and a 12-atom one (so both the
    pair sizes and the padding vary)."""
    from kernelforge_amd import fchl19_repr

    s1 = padded([molecule(33, 1), molecule(48, 2)])
    s2 = padded([molecule(40, 3), molecule(12, 4)])
    out = []
    for C, Q, N in (s1, s2):
        X, dX = fchl19_repr.generate_fchl_acsf_batch(C, Q, N, with_gradients=True)
        Xo, dXo = oracle.fchl19_batch(C, Q, N, grad=True)
        out.append((C, Q, N, X, dX, Xo, dXo))
    return out


def test_descriptors_and_jacobians(sets):
    for C, Q, N, X, dX, Xo, dXo in sets:
        assert row_rel_err(X, Xo, lead=2) < 1e-12
        assert row_rel_err(dX, dXo, lead=2) < 1e-12


def test_rectangular_derivative_forms(sets):
    from kernelforge_amd import local_kernels as lk

    (_, Q1, N1, _, _, X1, dX1), (_, Q2, N2, _, _, X2, dX2) = sets
    got = lk.kernel_gaussian_jacobian(X1, X2, dX2, Q1, Q2, N1, N2, SIGMA)
    assert rel_err(got, oracle.jacobian(X1, X2, dX2, Q1, Q2, N1, N2, SIGMA)) < 1e-10
    got = lk.kernel_gaussian_jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, SIGMA)
    assert rel_err(got, oracle.jacobian_t(X1, dX1, X2, Q1, Q2, N1, N2, SIGMA)) < 1e-10
    got = lk.kernel_gaussian_hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, SIGMA)
    assert rel_err(got, oracle.hessian(X1, dX1, X2, dX2, Q1, Q2, N1, N2, SIGMA)) < 1e-10
    got = lk.kernel_gaussian_full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, SIGMA)
    assert rel_err(got, oracle.full(X1, dX1, X2, dX2, Q1, Q2, N1, N2, SIGMA)) < 1e-10


def test_symmetric_derivative_forms(sets):
    from kernelforge_amd import local_kernels as lk

    _, Q, N, _, _, X, dX = sets[0]
    assert rel_err(lk.kernel_gaussian_full_symm(X, dX, Q, N, SIGMA),
                   oracle.full_symm(X, dX, Q, N, SIGMA)) < 1e-10
    assert rel_err(lk.kernel_gaussian_full_symm_rfp(X, dX, Q, N, SIGMA),
                   oracle.full_symm_rfp(X, dX, Q, N, SIGMA)) < 1e-10
    assert rel_err(lk.kernel_gaussian_hessian_symm_rfp(X, dX, Q, N, SIGMA),
                   oracle.hessian_symm_rfp(X, dX, Q, N, SIGMA)) < 1e-10


def test_alpha_desc_and_full_matvec(sets):
    from kernelforge_amd import local_kernels as lk

    (_, Q1, N1, _, _, X1, dX1), (_, Q2, N2, _, _, X2, dX2) = sets
    naq2 = 3 * int(N2.sum())
    rng = np.random.default_rng(5)
    aF, aE = rng.standard_normal(naq2), rng.standard_normal(len(N2))
    ad = lk.kernel_gaussian_local_compute_alpha_desc(dX2, Q2, N2, aF)
    ado = oracle.alpha_desc(dX2, Q2, N2, aF)
    assert rel_err(ad, ado) < 1e-12
    E, F = lk.kernel_gaussian_local_full_matvec(X1, dX1, X2, ado, aE, Q1, Q2, N1, N2, SIGMA)
    Eo, Fo = oracle.full_matvec(X1, dX1, X2, ado, aE, Q1, Q2, N1, N2, SIGMA)
    assert rel_err(E, Eo) < 1e-10 and rel_err(F, Fo) < 1e-10


def test_representation_of_130_atoms():
    from kernelforge_amd import fchl19_repr

    c, z = molecule(130, 9, elements=(1, 6, 8))
    assert row_rel_err(fchl19_repr.generate_fchl_acsf(c, z), oracle.fchl19(c, z)) < 1e-12
