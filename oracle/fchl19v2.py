"""FCHL19v2 representation and coordinate Jacobian: NumPy restatement (oracle).

TEST INFRASTRUCTURE ONLY (tests/, the smoke check and the CPU baseline use it as the
checker; the product package never imports it).  Pure NumPy, per molecule, for the
small molecules of the tests.

Restates src/fchl19v2_repr.cpp of the reference:
  * grids and nabasis as the bindings build them (src/fchl19v2_repr_bindings.cpp:56-88);
    w3 = sqrt(eta3/pi) * three_body_weight (:120);
  * rep size (src/fchl19v2_repr.cpp:118-141), pair index (:57-62), cutoff decay (:65-73);
  * two-body bases T1-T5 (:145-391): log-normal, Gaussian in r, Gaussian in ln r,
    Gaussian in r without the r^-decay factor, radial Bessel;
  * three-body bases A1-A5 (:837-1329): odd Fourier or cosine series in the centre angle,
    r_bar = (r_ij + r_ik)/2 or split (r_ij + r_ik, |r_ij - r_ik|) radial, with or without
    the Axilrod-Teller-Muto factor;
  * the gradient path (:2919-2986): for A1 it keeps only the first harmonic in both the
    values and the Jacobian (:1458-1462), as the original FCHL19 gradient path does; for
    the split radial bases it differentiates |r_ij - r_ik| as sqrt((r_ij - r_ik)^2 + 1e-8)
    (:2072-2096) while the values keep the exact |r_ij - r_ik|.

The Jacobian is the exact derivative of this restatement, obtained by complex-step
differentiation (Squire & Trapp, SIAM Review 40(1), 1998): every operation here is
analytic in the coordinates away from the cutoffs, so Im f(x + ih)/h with h = 1e-30 is
the derivative to rounding, with no cancellation.  |.|, max and the comparisons act on
real parts.
"""

from __future__ import annotations

import math

import numpy as np

EPS = 1e-12
SQRT_2PI = 2.5066282746310005
TWO_BODY = ("log_normal", "gaussian_r", "gaussian_log_r", "gaussian_r_no_pow", "bessel")
THREE_BODY = ("odd_fourier_rbar", "cosine_rbar", "odd_fourier_split_r", "cosine_split_r",
              "cosine_split_r_no_atm")
SPLIT = ("odd_fourier_split_r", "cosine_split_r", "cosine_split_r_no_atm")
ODD = ("odd_fourier_rbar", "odd_fourier_split_r")


def nabasis(three_body, nFourier, nCosine):
    """src/fchl19v2_repr_bindings.cpp:76-86."""
    if three_body in ODD:
        return 2 * nFourier
    return nCosine if nCosine > 0 else 2 * nFourier


def rep_size(nelem, nRs2, nRs3, nab, three_body="odd_fourier_rbar", nRs3_minus=0):
    """src/fchl19v2_repr.cpp:118-141."""
    npair = nelem * (nelem + 1) // 2
    if three_body in SPLIT:
        if nRs3_minus <= 0:
            raise ValueError("nbasis3_minus must be > 0 for SplitR variants")
        return nelem * nRs2 + npair * nRs3 * nRs3_minus * nab
    return nelem * nRs2 + npair * nRs3 * nab


def pair_index(nelem, p, q):
    p, q = min(p, q), max(p, q)
    return -p * (p + 1) // 2 + q + nelem * p


def _rmax(x, lo):
    """max(x, lo) on the real part, keeping x's imaginary part where x wins."""
    return np.where(np.real(x) > lo, x, lo)


def _rabs(x):
    return np.where(np.real(x) < 0, -x, x)


def _clip1(x):
    r = np.real(x)
    return np.where(r > 1.0, 1.0, np.where(r < -1.0, -1.0, x))


def _forward(R, elem_of, nelem, P, grad_path):
    """rep (n, rep_size) of coordinates R (n, 3), real or complex."""
    n = R.shape[0]
    dt = R.dtype
    Rs2 = np.array([P["rcut"] * i / P["nRs2"] for i in range(1, P["nRs2"] + 1)])
    Rs3 = np.array([P["acut"] * i / P["nRs3"] for i in range(1, P["nRs3"] + 1)])
    nb3m = P["nRs3_minus"]
    Rs3m = np.array([P["acut"] * i / nb3m for i in range(1, nb3m + 1)])
    tb, ab = P["two_body"], P["three_body"]
    nab = nabasis(ab, P["nFourier"], P["nCosine"])
    nb2, nb3 = len(Rs2), len(Rs3)
    size = rep_size(nelem, nb2, nb3, nab, ab, nb3m)
    rep = np.zeros((n, size), dtype=dt)
    diff = R[None, :, :] - R[:, None, :]  # diff[i, j] = R_j - R_i
    D = np.sqrt(np.sum(diff * diff, axis=-1) + (np.eye(n) if n else 0.0)) * (1 - np.eye(n))
    rcut, acut = P["rcut"], P["acut"]
    dec2 = 0.5 * (np.cos(math.pi * D * (1.0 / rcut if rcut > 0 else 0.0)) + 1.0)
    dec3 = 0.5 * (np.cos(math.pi * D * (1.0 / acut if acut > 0 else 0.0)) + 1.0)
    eta2, d2 = P["eta2"], P["two_body_decay"]
    # ---- two-body (src/fchl19v2_repr.cpp:145-391)
    for i in range(n):
        for j in range(i + 1, n):
            r = D[i, j]
            if np.real(r) > rcut:
                continue
            dec = dec2[i, j]
            if tb == "log_normal":
                if np.any(Rs2 <= 0):
                    raise ValueError("All Rs2 must be > 0")
                lp = np.log1p(eta2 / _rmax(r * r, EPS))
                sig = np.sqrt(_rmax(lp, 0.0))
                if np.real(sig) < EPS:
                    continue
                mu = np.log(r) - 0.5 * lp
                pref = dec / (sig * SQRT_2PI * r ** d2)
                dl = np.log(Rs2) - mu
                val = pref * np.exp(-0.5 * dl * dl / (sig * sig)) / Rs2
            elif tb == "gaussian_r":
                val = dec / r ** d2 * np.exp(-eta2 * (r - Rs2) ** 2)
            elif tb == "gaussian_log_r":
                if np.any(Rs2 <= 0):
                    raise ValueError("All Rs2 must be > 0")
                val = dec / r ** d2 * np.exp(-eta2 * (np.log(r) - np.log(Rs2)) ** 2)
            elif tb == "gaussian_r_no_pow":
                val = dec * np.exp(-eta2 * (r - Rs2) ** 2)
            elif tb == "bessel":
                if np.real(r) < EPS:
                    continue
                k = np.arange(1, nb2 + 1)
                val = dec * math.sqrt(2.0 / rcut) * np.sin(k * math.pi / rcut * r) / r
            else:
                raise ValueError(f"Unknown two_body_type: {tb}")
            rep[i, elem_of[j] * nb2:(elem_of[j] + 1) * nb2] += val
            rep[j, elem_of[i] * nb2:(elem_of[i] + 1) * nb2] += val
    # ---- three-body (src/fchl19v2_repr.cpp:837-1329)
    w3 = math.sqrt(P["eta3"] / math.pi) * P["three_body_weight"]
    if w3 == 0.0:
        return rep
    eta3, eta3m, d3, zeta = P["eta3"], P["eta3_minus"], P["three_body_decay"], P["zeta"]
    off = nelem * nb2
    nh = nab // 2
    ow = [2.0 * math.exp(-0.5 * (zeta * (2 * l + 1)) ** 2) for l in range(nh)]
    for i in range(n):
        for j in range(n - 1):
            if j == i or np.real(D[i, j]) > acut:
                continue
            for k in range(j + 1, n):
                if k == i or np.real(D[i, k]) > acut:
                    continue
                rij, rik, rjk = D[i, j], D[i, k], D[j, k]
                eij = diff[i, j] / _rmax(rij, EPS)
                eik = diff[i, k] / _rmax(rik, EPS)
                ci = _clip1(np.sum(eij * eik))
                if ab == "cosine_split_r_no_atm":
                    ksi = w3
                else:
                    ejk = diff[j, k] / _rmax(rjk, EPS)
                    cj = -np.sum(eij * ejk)
                    ck = np.sum(eik * ejk)
                    ksi = (1.0 + 3.0 * ci * cj * ck) * (w3 / _rmax(rik * rij * rjk, EPS) ** d3)
                if ab in ODD:
                    si = np.sqrt(_rmax(1.0 - ci * ci, 0.0))
                    ang = np.zeros(nab, dtype=dt)
                    # cos(o theta), sin(o theta) for o = 1, 3, ... by the Chebyshev recurrence
                    c2, s2, c1, s1 = 1.0, 0.0, ci, si
                    cos_o, sin_o = {1: ci}, {1: si}
                    for o in range(2, 2 * nh):
                        c, s = 2.0 * ci * c1 - c2, 2.0 * ci * s1 - s2
                        c2, s2, c1, s1 = c1, s1, c, s
                        cos_o[o], sin_o[o] = c, s
                    keep = 1 if (grad_path and ab == "odd_fourier_rbar") else nh
                    for l in range(keep):
                        o = 2 * l + 1
                        ang[2 * l] = ow[l] * cos_o[o]
                        if 2 * l + 1 < nab:
                            ang[2 * l + 1] = ow[l] * sin_o[o]
                else:
                    ang = np.zeros(nab, dtype=dt)
                    if nab > 0:
                        ang[0] = 1.0
                    if nab > 1:
                        ang[1] = ci
                    for m in range(2, nab):
                        ang[m] = 2.0 * ci * ang[m - 1] - ang[m - 2]
                p = pair_index(nelem, elem_of[j], elem_of[k])
                dd = dec3[i, j] * dec3[i, k]
                if ab in SPLIT:
                    rp, dm = rij + rik, rij - rik
                    php = np.exp(-eta3 * (rp - Rs3) ** 2)
                    phm = np.exp(-eta3m * (_rabs(dm) - Rs3m) ** 2)
                    if grad_path and np.iscomplexobj(dm):
                        # the gradient path differentiates |r_ij - r_ik| as the smoothed
                        # sqrt((r_ij - r_ik)^2 + 1e-8) (src/fchl19v2_repr.cpp:2072-2096),
                        # while its values keep the exact |r_ij - r_ik|
                        dr, sm = dm.real, math.sqrt(dm.real ** 2 + 1e-8)
                        val = phm.real
                        phm = val + 1j * dm.imag * (val * (-2.0 * eta3m * (sm - Rs3m)) * dr / sm)
                    rad = (php[:, None] * phm[None, :] * dd).reshape(-1)
                    base = off + p * nb3 * nb3m * nab
                else:
                    rad = np.exp(-eta3 * (0.5 * (rij + rik) - Rs3) ** 2) * dd
                    base = off + p * nb3 * nab
                rep[i, base:base + rad.size * nab] += (rad[:, None] * ang[None, :] * ksi).reshape(-1)
    return rep


def params(nRs2=24, nRs3=20, nRs3_minus=0, nFourier=1, nCosine=0, eta2=0.32, eta3=2.7,
           eta3_minus=2.7, zeta=math.pi, rcut=8.0, acut=8.0, two_body_decay=1.8,
           three_body_decay=0.57, three_body_weight=13.4, two_body_type="log_normal",
           three_body_type="odd_fourier_rbar"):
    if two_body_type not in TWO_BODY:
        raise ValueError(f"Unknown two_body_type: {two_body_type}")
    if three_body_type not in THREE_BODY:
        raise ValueError(f"Unknown three_body_type: {three_body_type}")
    return dict(nRs2=nRs2, nRs3=nRs3, nRs3_minus=nRs3_minus, nFourier=nFourier,
                nCosine=nCosine, eta2=eta2, eta3=eta3, eta3_minus=eta3_minus, zeta=zeta,
                rcut=rcut, acut=acut, two_body_decay=two_body_decay,
                three_body_decay=three_body_decay, three_body_weight=three_body_weight,
                two_body=two_body_type, three_body=three_body_type)


def _elem_of(z, elements):
    lut = {int(e): i for i, e in enumerate(elements)}
    try:
        return [lut[int(q)] for q in z]
    except KeyError:
        raise RuntimeError("nuclear_z contains an element not present in elements") from None


def generate(coords, z, elements=(1, 6, 7, 8, 16), **kw):
    """rep (n, rep_size), as kernelforge.fchl19v2_repr.generate."""
    R = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    return _forward(R, _elem_of(z, elements), len(elements), params(**kw), False).real


def generate_and_gradients(coords, z, elements=(1, 6, 7, 8, 16), columns=None, **kw):
    """(rep (n, rep_size), grad (n, rep_size, 3n)) as generate_and_gradients.  With
    `columns` (indices into the 3n coordinates) only those Jacobian columns are formed
    (the others stay zero)."""
    R = np.asarray(coords, dtype=np.float64).reshape(-1, 3)
    n = R.shape[0]
    P = params(**kw)
    el = _elem_of(z, elements)
    rep = _forward(R, el, len(elements), P, True).real
    h = 1e-30
    grad = np.zeros(rep.shape + (3 * n,))
    for col in (range(3 * n) if columns is None else columns):
        Rc = R.astype(np.complex128)
        Rc[col // 3, col % 3] += 1j * h
        grad[:, :, col] = _forward(Rc, el, len(elements), P, True).imag / h
    return rep, grad
