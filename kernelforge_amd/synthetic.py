"""Seeded synthetic molecule generators for the FCHL19 KRR workloads.

The reference ships no data generator; these are the build's own, and the RNG
draw order is frozen here (changing it changes every fixture and benchmark):

``random_molecules(n_mols, seed)`` — one ``numpy.random.default_rng(seed)``
stream, per molecule in order:
  1. ``n = rng.integers(5, 24)``                         (atom count, 5..23)
  2. ``Z = rng.choice([1,6,7,8,16], n, p=[.5,.3,.1,.08,.02])``
  3. atoms placed one at a time: ``p = rng.uniform(0, L, 3)`` with
     ``L = 1.6 * n**(1/3)`` Angstrom, redrawn while any placed atom is
     closer than 0.9 Angstrom.
Energy (Hartree) = sum of per-element offsets + sum over all pairs i<j of the
Morse term ``De * (1 - exp(-a (r - r0)))**2`` (De=0.1, a=1.5/A, r0=1.4 A; zero
at the minimum).

``conformers(base_coords, base_z, n_conf, seed)`` — Gaussian displacement
(sigma 0.05 A) of every atom, drawn as ``rng.normal(0, 0.05, (n_conf, n, 3))``
in one call.  Energy = harmonic bonds ``0.5 k (r - r0)^2`` (k=0.5 Ha/A^2, bonds
= base pairs closer than 1.7 A, r0 = their base distance) + Morse on the
non-bonded pairs.  Forces are the analytic ``-dE/dR``.

The aspirin-composition base molecule (C9H8O4, element order
``[6]*9 + [1]*8 + [8]*4``) takes its coordinates from the same placement rule
with ``default_rng(0)`` and n = 21.
"""

from __future__ import annotations

import numpy as np

ELEMENTS = (1, 6, 7, 8, 16)
ELEMENT_P = (0.5, 0.3, 0.1, 0.08, 0.02)
ELEMENT_OFFSET = {1: -0.5, 6: -38.0, 7: -54.6, 8: -75.0, 16: -398.0}
MORSE_DE, MORSE_A, MORSE_R0 = 0.1, 1.5, 1.4
BOND_K, BOND_CUT = 0.5, 1.7
MIN_DIST = 0.9
ASPIRIN_Z = np.array([6] * 9 + [1] * 8 + [8] * 4, dtype=np.int32)


def _place_atoms(rng: np.random.Generator, n: int) -> np.ndarray:
    side = 1.6 * n ** (1.0 / 3.0)
    out = np.empty((n, 3), dtype=np.float64)
    placed = 0
    while placed < n:
        p = rng.uniform(0.0, side, size=3)
        if placed:
            d2 = np.sum((out[:placed] - p) ** 2, axis=1)
            if np.min(d2) < MIN_DIST * MIN_DIST:
                continue
        out[placed] = p
        placed += 1
    return out


def _pair_dist(coords: np.ndarray):
    n = coords.shape[-2]
    iu, ju = np.triu_indices(n, 1)
    diff = coords[..., iu, :] - coords[..., ju, :]
    r = np.sqrt(np.sum(diff * diff, axis=-1))
    return iu, ju, diff, r


def morse(r: np.ndarray) -> np.ndarray:
    t = 1.0 - np.exp(-MORSE_A * (r - MORSE_R0))
    return MORSE_DE * t * t


def morse_dr(r: np.ndarray) -> np.ndarray:
    e = np.exp(-MORSE_A * (r - MORSE_R0))
    return 2.0 * MORSE_DE * MORSE_A * (1.0 - e) * e


def molecule_energy(coords: np.ndarray, z: np.ndarray) -> float:
    _, _, _, r = _pair_dist(coords)
    e0 = sum(ELEMENT_OFFSET[int(zi)] for zi in z)
    return float(e0 + np.sum(morse(r)))


def random_molecules(n_mols: int, seed: int):
    """Return (coords_list, z_list, energies) for ``n_mols`` seeded molecules."""
    rng = np.random.default_rng(seed)
    coords_list, z_list = [], []
    energies = np.empty(n_mols, dtype=np.float64)
    for m in range(n_mols):
        n = int(rng.integers(5, 24))
        z = rng.choice(np.array(ELEMENTS), size=n, p=ELEMENT_P).astype(np.int32)
        c = _place_atoms(rng, n)
        coords_list.append(c)
        z_list.append(z)
        energies[m] = molecule_energy(c, z)
    return coords_list, z_list, energies


def aspirin_base():
    """Fixed 21-atom C9H8O4 base geometry (placement rule, default_rng(0))."""
    rng = np.random.default_rng(0)
    return _place_atoms(rng, 21), ASPIRIN_Z.copy()


def harmonic_morse(coords: np.ndarray, base_coords: np.ndarray, z: np.ndarray):
    """Energy (n_conf,) and forces (n_conf,n,3) of conformers ``coords`` (n_conf,n,3):
    harmonic bonds on base pairs < 1.7 A, Morse on the other pairs."""
    iu, ju, _, r_base = _pair_dist(base_coords)
    bonded = r_base < BOND_CUT
    _, _, diff, r = _pair_dist(coords)          # diff = R_i - R_j, (n_conf, npairs, 3)
    e_offset = sum(ELEMENT_OFFSET[int(zi)] for zi in z)
    e_bond = 0.5 * BOND_K * (r - r_base[None, :]) ** 2
    pair_e = np.where(bonded[None, :], e_bond, morse(r))
    energies = e_offset + np.sum(pair_e, axis=1)
    dv = np.where(bonded[None, :], BOND_K * (r - r_base[None, :]), morse_dr(r))
    g = (dv / r)[..., None] * diff                 # dE/dR_i for each pair
    grad = np.zeros_like(coords)
    np.add.at(grad, (slice(None), iu), g)
    np.add.at(grad, (slice(None), ju), -g)
    return energies, -grad


def conformers(base_coords: np.ndarray, base_z: np.ndarray, n_conf: int, seed: int):
    """Return (coords (n_conf,n,3), z (n,), energies (n_conf,), forces (n_conf,n,3))."""
    rng = np.random.default_rng(seed)
    n = base_coords.shape[0]
    coords = base_coords[None, :, :] + rng.normal(0.0, 0.05, size=(n_conf, n, 3))
    energies, forces = harmonic_morse(coords, base_coords, base_z)
    return coords, base_z.copy(), energies, forces


def pad_batch(coords_list, z_list):
    """Zero-padded (coords (nm,max,3), Q (nm,max) int32, N (nm,) int32)."""
    nm = len(z_list)
    N = np.array([len(z) for z in z_list], dtype=np.int32)
    max_atoms = int(N.max()) if nm else 0
    C = np.zeros((nm, max_atoms, 3), dtype=np.float64)
    Q = np.zeros((nm, max_atoms), dtype=np.int32)
    for i, (c, z) in enumerate(zip(coords_list, z_list)):
        C[i, : len(z)] = c
        Q[i, : len(z)] = z
    return C, Q, N


def config_molecules(config: int):
    """Workload inputs by config number (0: 1k mols seed 0; 1/2: 20k+2k mols seed 1;
    3: 500 C9H8O4 conformers; 4: 2000 conformers, seed 3)."""
    if config == 0:
        return random_molecules(1000, 0)
    if config in (1, 2):
        return random_molecules(22000, 1)
    base, z = aspirin_base()
    if config == 3:
        return conformers(base, z, 500, 2)
    if config == 4:
        return conformers(base, z, 2000, 3)
    raise ValueError(f"unknown config {config}")
