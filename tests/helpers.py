"""Numeric comparison helpers shared by the tests."""

import numpy as np


def rel_err(a, b, floor=0.0):
    """max|a-b| / max|b| (with an optional floor on the denominator)."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), floor, 1e-300)) if b.size else 0.0


def row_rel_err(a, b, lead=1):
    """Descriptor parity norm: the arrays are split into per-atom blocks along the
    first `lead` axes (rep (n, rs): one row per atom; grad (n, rs, 3n): one
    (rs, 3n) block per centre atom; padded batches (nm, ma, ...): lead=2), and the
    result is max over atoms of max|a-b| / max|b| within the atom's block.  Atom
    blocks that are identically zero in b must be identically zero in a."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    na = int(np.prod(a.shape[:lead]))
    a = a.reshape(na, -1)
    b = b.reshape(na, -1)
    num = np.max(np.abs(a - b), axis=1)
    den = np.max(np.abs(b), axis=1)
    zero = den == 0
    if np.any(num[zero] != 0):
        return float("inf")
    return float(np.max(num[~zero] / den[~zero])) if (~zero).any() else 0.0
