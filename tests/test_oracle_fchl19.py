"""Pins the oracle's FCHL19 restatement (oracle/kf_oracle.c).

* against the reference's own code: golden vectors written by
  tests/golden/make_golden.py from src/fchl19_repr.cpp compiled as-is, and the live
  reference build oracle/_ref when present;
* against mathematics the formula fixes: central finite differences of the
  representation (the Jacobian), translation/rotation invariance, weight-0
  zeroing of the three-body block (tests/test_fchl19.py:146-193 in the reference),
  rep of the gradient path == rep path for nFourier=1 (:220-251), and the basis
  grids of the bindings (:196-217).
"""

import numpy as np
import pytest

import oracle
from tests.helpers import row_rel_err
from kernelforge_amd import synthetic


def test_rep_matches_reference_golden(golden_fchl19):
    g = golden_fchl19
    for t in range(10):
        rep = oracle.fchl19(g[f"coords_{t}"], g[f"z_{t}"])
        assert rep.shape == g[f"rep_{t}"].shape
        assert row_rel_err(rep, g[f"rep_{t}"]) < 1e-13


def test_grad_matches_reference_golden(golden_fchl19):
    g = golden_fchl19
    for t in range(3):
        rep, grad = oracle.fchl19(g[f"gcoords_{t}"], g[f"gz_{t}"], grad=True)
        assert row_rel_err(rep, g[f"grep_{t}"]) < 1e-13
        assert row_rel_err(grad, g[f"ggrad_{t}"]) < 1e-12


def test_nondefault_params_match_reference_golden(golden_fchl19):
    from tests.golden.make_golden import NONDEFAULT

    g = golden_fchl19
    rep = oracle.fchl19(g["nd_coords"], g["nd_z"], **NONDEFAULT)
    assert row_rel_err(rep, g["nd_rep"]) < 1e-13
    rep, grad = oracle.fchl19(g["nd_coords"], g["nd_z"], grad=True, **NONDEFAULT)
    assert row_rel_err(rep, g["nd_grep"]) < 1e-13
    assert row_rel_err(grad, g["nd_ggrad"]) < 1e-12


@pytest.mark.skipif(oracle.ref() is None, reason="reference build oracle/_ref not present")
@pytest.mark.parametrize("kw", [{}, dict(nFourier=3), dict(three_body_weight=0.0),
                                dict(rcut=2.8, acut=2.3)])
def test_oracle_matches_live_reference(kw):
    coords, zs, _ = synthetic.random_molecules(6, 11)
    for c, z in zip(coords, zs):
        assert row_rel_err(oracle.fchl19(c, z, **kw), oracle.ref_fchl19(c, z, **kw)) < 1e-13
        r1, g1 = oracle.fchl19(c, z, grad=True, **kw)
        r2, g2 = oracle.ref_fchl19(c, z, grad=True, **kw)
        assert row_rel_err(g1, g2) < 1e-12


def test_jacobian_is_derivative_of_rep():
    coords, zs, _ = synthetic.random_molecules(3, 4)
    c, z = coords[0], zs[0]
    rep, grad = oracle.fchl19(c, z, grad=True)
    h = 1e-5
    n = len(z)
    for a in range(n):
        for d in range(3):
            cp, cm = c.copy(), c.copy()
            cp[a, d] += h
            cm[a, d] -= h
            fd = (oracle.fchl19(cp, z) - oracle.fchl19(cm, z)) / (2 * h)
            an = grad[:, :, 3 * a + d]
            assert np.max(np.abs(fd - an)) < 2e-6 * max(1.0, np.max(np.abs(an)))


def test_translation_rotation_invariance():
    coords, zs, _ = synthetic.random_molecules(2, 9)
    c, z = coords[1], zs[1]
    rep = oracle.fchl19(c, z)
    th = 0.7
    Rz = np.array([[np.cos(th), -np.sin(th), 0], [np.sin(th), np.cos(th), 0], [0, 0, 1]])
    rep2 = oracle.fchl19(c @ Rz.T + np.array([1.3, -2.0, 0.4]), z)
    assert row_rel_err(rep2, rep) < 1e-11
    # Jacobian rows sum to zero over atoms for a uniform shift
    _, grad = oracle.fchl19(c, z, grad=True)
    s = grad.reshape(grad.shape[0], grad.shape[1], -1, 3).sum(axis=2)
    assert np.max(np.abs(s)) < 1e-10 * np.max(np.abs(grad))


def test_three_body_block_zero_when_weight_zero():
    coords, zs, _ = synthetic.random_molecules(2, 3)
    rep = oracle.fchl19(coords[0], zs[0], three_body_weight=0.0)
    assert np.all(rep[:, 5 * 24:] == 0.0)
    assert np.any(rep[:, : 5 * 24] != 0.0)


def test_gradient_path_rep_equals_rep_path_for_one_harmonic():
    coords, zs, _ = synthetic.random_molecules(3, 2)
    for c, z in zip(coords, zs):
        r1 = oracle.fchl19(c, z)
        r2, _ = oracle.fchl19(c, z, grad=True)
        assert row_rel_err(r2, r1) < 1e-13


def test_rep_size_closed_form():
    assert oracle.rep_size(5, 24, 20, 2) == 720
    assert oracle.rep_size(4, 10, 7, 6) == 4 * 10 + 10 * 7 * 6


def test_batch_equals_per_molecule(config0):
    coords, zs, _, C, Q, N = config0
    X = oracle.fchl19_batch(C[:7], Q[:7], N[:7])
    for m in range(7):
        n = len(zs[m])
        np.testing.assert_array_equal(X[m, :n], oracle.fchl19(coords[m], zs[m]))
        assert np.all(X[m, n:] == 0.0)
