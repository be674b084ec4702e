"""CPU: the FCHL19v2 oracle (oracle/fchl19v2.py) pinned by the reference's own
src/fchl19v2_repr.cpp -- through the committed golden vectors, and, where oracle/_ref is
built, directly on config-0 molecules for every basis pair -- plus the host-side
argument handling of the fchl19v2_repr module (no GPU calls)."""

import itertools
import os

import numpy as np
import pytest

import oracle
from oracle import fchl19v2 as v2
from tests.golden.make_golden import V2_ELEMENTS, v2_golden_kw
from tests.helpers import row_rel_err

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PAIRS = list(itertools.product(v2.TWO_BODY, v2.THREE_BODY))


@pytest.fixture(scope="module")
def golden():
    return np.load(os.path.join(ROOT, "tests", "golden", "fchl19v2_ref.npz"))


@pytest.mark.parametrize("tb,ab", PAIRS)
def test_oracle_matches_reference_golden(golden, tb, ab):
    c, z = golden["coords"], golden["z"]
    kw = v2_golden_kw(tb, ab)
    rep = v2.generate(c, z, elements=V2_ELEMENTS, **kw)
    assert row_rel_err(rep, golden[f"rep_{tb}_{ab}"]) < 1e-13
    if f"ggrad_{tb}_{ab}" in golden.files:
        r, g = v2.generate_and_gradients(c, z, elements=V2_ELEMENTS, **kw)
        assert row_rel_err(r, golden[f"grep_{tb}_{ab}"]) < 1e-13
        assert row_rel_err(g, golden[f"ggrad_{tb}_{ab}"]) < 1e-13


def default_kw(tb, ab):
    return dict(two_body_type=tb, three_body_type=ab, nRs3_minus=4 if ab in v2.SPLIT else 0,
                nFourier=2 if ab in v2.ODD else 1, nCosine=3 if "cosine" in ab else 0)


@pytest.mark.parametrize("tb,ab", PAIRS)
def test_oracle_matches_reference_build(tb, ab):
    if oracle.ref() is None:
        pytest.skip("oracle/_ref (the reference build) is not available")
    from kernelforge_amd import synthetic

    coords, zs, _ = synthetic.random_molecules(8, 0)
    m = int(np.argmin([len(q) for q in zs]))
    kw = default_kw(tb, ab)
    ref = oracle.ref_fchl19v2(coords[m], zs[m], **kw)
    assert row_rel_err(v2.generate(coords[m], zs[m], **kw), ref) < 1e-13
    rr, gr = oracle.ref_fchl19v2(coords[m], zs[m], grad=True, **kw)
    r, g = v2.generate_and_gradients(coords[m], zs[m], **kw)
    assert row_rel_err(r, rr) < 1e-13 and row_rel_err(g, gr) < 1e-12


def test_gradient_path_conventions():
    """The two conventions of the reference's gradient path that the oracle keeps: A1
    drops the higher harmonics (values too), and the split bases' Jacobian differs from
    the exact derivative of |r_ij - r_ik| by the 1e-8 smoothing (small but nonzero)."""
    from kernelforge_amd import synthetic

    coords, zs, _ = synthetic.random_molecules(8, 0)
    m = int(np.argmin([len(q) for q in zs]))
    c, z = coords[m], zs[m]
    kw = dict(three_body_type="odd_fourier_rbar", nFourier=2)
    full = v2.generate(c, z, **kw)
    gpath, _ = v2.generate_and_gradients(c, z, **kw)
    off = 5 * 24
    blk = gpath[:, off:].reshape(len(z), 15, 20, 4)
    assert np.all(blk[..., 2:] == 0.0) and np.any(full[:, off:].reshape(len(z), 15, 20, 4)[..., 2:])
    assert np.allclose(blk[..., :2], full[:, off:].reshape(len(z), 15, 20, 4)[..., :2],
                       rtol=1e-14, atol=0)


def test_module_rejects_bad_types_and_sizes():
    from kernelforge_amd import fchl19v2_repr as m

    assert m.compute_rep_size(5, 24, 20, 2) == 720
    assert m.compute_rep_size(5, 24, 20, 2, "odd_fourier_split_r", 4) == 5 * 24 + 15 * 20 * 4 * 2
    assert m.compute_rep_size(5, 24, 20, 2) == v2.rep_size(5, 24, 20, 2)
    with pytest.raises(ValueError):
        m.compute_rep_size(5, 24, 20, 2, "cosine_split_r", 0)
    with pytest.raises(ValueError):
        m.compute_rep_size(5, 24, 20, 2, "no_such_basis")
    c = np.zeros((3, 3))
    z = np.array([1, 1, 1], dtype=np.int32)
    with pytest.raises(ValueError):
        m.generate(c, z, two_body_type="no_such_basis")
    with pytest.raises(ValueError):
        m.generate(c, z, three_body_type="cosine_split_r")  # nRs3_minus = 0


def test_fchl19_is_the_v2_lognormal_odd_fourier_basis():
    """FCHL19 equals FCHL19v2 with the log-normal two-body and odd-Fourier r_bar three-body
    bases, values and Jacobians (the gradient path's first-harmonic convention included),
    which lets the library run large molecules of FCHL19 on the FCHL19v2 kernel."""
    from kernelforge_amd import synthetic

    coords, zs, _ = synthetic.random_molecules(6, 3)
    nd = dict(nFourier=2, nRs2=12, nRs3=9, eta2=0.41, eta3=2.1, rcut=5.5, acut=4.5,
              two_body_decay=1.5, three_body_decay=0.7, three_body_weight=9.0)
    for m in range(3):
        c, z = coords[m], zs[m]
        for kw in ({}, nd):
            assert row_rel_err(v2.generate(c, z, **kw), oracle.fchl19(c, z, **kw)) < 1e-14
            r1, g1 = oracle.fchl19(c, z, grad=True, **kw)
            r2, g2 = v2.generate_and_gradients(c, z, **kw)
            assert row_rel_err(r2, r1) < 1e-14 and row_rel_err(g2, g1) < 1e-13
