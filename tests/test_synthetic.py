"""The seeded workload generator: frozen draw order, geometry constraints, analytic forces."""

import numpy as np

from kernelforge_amd import synthetic


def test_deterministic_and_constraints():
    c1, z1, e1 = synthetic.random_molecules(50, 0)
    c2, z2, e2 = synthetic.random_molecules(50, 0)
    for a, b in zip(c1, c2):
        np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(e1, e2)
    for c, z in zip(c1, z1):
        n = len(z)
        assert 5 <= n <= 23
        assert set(z.tolist()) <= {1, 6, 7, 8, 16}
        d = np.sqrt(((c[:, None] - c[None]) ** 2).sum(-1))
        assert np.min(d[np.triu_indices(n, 1)]) >= 0.9
        assert np.all((c >= 0) & (c <= 1.6 * n ** (1 / 3)))


def test_frozen_first_molecule():
    c, z, e = synthetic.random_molecules(1, 0)
    # freezes the draw order (n, Z, then placement) of seed 0
    assert len(z[0]) == 21
    assert abs(e[0] - (-547.97281148)) < 1e-6


def test_conformer_forces_are_minus_gradient():
    base, z = synthetic.aspirin_base()
    assert sorted(z.tolist()) == sorted([6] * 9 + [1] * 8 + [8] * 4)
    C, Z, E, F = synthetic.conformers(base, z, 3, 2)
    h = 1e-4  # energies are ~ -630 Ha: a larger step keeps cancellation below 1e-9
    for a in (0, 7, 20):
        for d in range(3):
            Cp, Cm = C.copy(), C.copy()
            Cp[:, a, d] += h
            Cm[:, a, d] -= h
            fd = -(synthetic.harmonic_morse(Cp, base, z)[0] - synthetic.harmonic_morse(Cm, base, z)[0]) / (2 * h)
            np.testing.assert_allclose(F[:, a, d], fd, rtol=1e-6, atol=1e-7)
