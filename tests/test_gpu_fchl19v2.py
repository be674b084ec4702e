"""GPU parity of kernelforge_amd.fchl19v2_repr (kf_fchl19v2_f64) against the FCHL19v2
oracle (oracle/fchl19v2.py, pinned by the reference build in tests/test_oracle_fchl19v2.py).
Bar: 1e-12 per atom (max |diff| over an atom's row / Jacobian block over its max)."""

import itertools

import numpy as np
import pytest

from oracle import fchl19v2 as v2
from tests.helpers import row_rel_err

pytestmark = pytest.mark.gpu
PAIRS = list(itertools.product(v2.TWO_BODY, v2.THREE_BODY))


def kw_for(tb, ab):
    return dict(two_body_type=tb, three_body_type=ab, nRs3_minus=4 if ab in v2.SPLIT else 0,
                nFourier=2 if ab in v2.ODD else 1, nCosine=3 if "cosine" in ab else 0)


@pytest.fixture(scope="module")
def mols():
    from kernelforge_amd import synthetic

    coords, zs, _ = synthetic.random_molecules(12, 0)
    order = np.argsort([len(q) for q in zs])
    return [(coords[m], zs[m]) for m in order[:2]]


@pytest.mark.parametrize("tb,ab", PAIRS)
def test_generate_and_gradients_vs_oracle(mols, tb, ab):
    from kernelforge_amd import fchl19v2_repr as m

    kw = kw_for(tb, ab)
    for c, z in mols:
        assert row_rel_err(m.generate(c, z, **kw), v2.generate(c, z, **kw)) < 1e-12
        r, g = m.generate_and_gradients(c, z, **kw)
        ro, go = v2.generate_and_gradients(c, z, **kw)
        assert row_rel_err(r, ro) < 1e-12
        assert row_rel_err(g, go) < 1e-12


def test_batch_padding_matches_single(mols):
    from kernelforge_amd import fchl19v2_repr as m
    from kernelforge_amd import synthetic

    coords, zs, _ = synthetic.random_molecules(9, 4)
    C, Q, N = synthetic.pad_batch(coords, zs)
    kw = kw_for("bessel", "cosine_split_r")
    X, dX = m.generate_batch(C, Q, N, with_gradients=True, **kw)
    for a in range(len(zs)):
        n = len(zs[a])
        r, g = m.generate_and_gradients(coords[a], zs[a], **kw)
        assert row_rel_err(X[a, :n], r) < 1e-13
        assert np.all(X[a, n:] == 0) and np.all(dX[a, n:] == 0)
        assert np.all(dX[a, :n, :, 3 * n:] == 0)
        assert row_rel_err(dX[a, :n, :, :3 * n], g) < 1e-13


def test_large_molecule_several_triplet_chunks():
    """42 atoms: 820 triplets per centre, more than one chunk of 512 per workgroup."""
    from kernelforge_amd import fchl19v2_repr as m

    rng = np.random.default_rng(11)
    n = 42
    pts = []
    while len(pts) < n:
        p = rng.uniform(0, 1.6 * n ** (1 / 3), 3)
        if all(np.linalg.norm(p - q) > 0.9 for q in pts):
            pts.append(p)
    c = np.array(pts)
    z = rng.choice([1, 6, 7, 8], n).astype(np.int32)
    kw = kw_for("log_normal", "odd_fourier_split_r")
    kw.update(nRs2=8, nRs3=6, nRs3_minus=3)
    assert row_rel_err(m.generate(c, z, **kw), v2.generate(c, z, **kw)) < 1e-12
    cols = [0, 1, 2, 61, 125]
    r, g = m.generate_and_gradients(c, z, **kw)
    ro, go = v2.generate_and_gradients(c, z, columns=cols, **kw)
    assert row_rel_err(r, ro) < 1e-12
    assert row_rel_err(g[:, :, cols], go[:, :, cols]) < 1e-12


def test_unknown_element_raises():
    from kernelforge_amd import fchl19v2_repr as m

    with pytest.raises(RuntimeError):
        m.generate(np.eye(3), np.array([1, 6, 9], dtype=np.int32))
