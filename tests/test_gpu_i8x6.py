"""GPU parity: the int8-split (precision='i8x6') form of the scalar local kernels vs the
oracle.  Each descriptor row is split into six 7-bit integer slices under a power-of-two
row scale; the slice products run exactly in int32 on the int8 MFMA, the pairs of slices
whose weights sum past 2^-42 are dropped, and norms, exp and sums are FP64.

Bars:
* rows whose bits fit the kept slice pairs: Gram entries exact, so the kernel must match
  the oracle to the exact-argument exp tolerance of the FP64 form;
* random and windowed sets: entry by entry within the method's bound.  Slice s is at most
  2^-7s mu, so a dropped pair (s, t), s + t >= 6, weighs at most 2^-(7 (s+t)) mu_i mu_j
  per column (5.03 * 2^-42 over all of them), and the remainder of each row after six
  slices is at most 2^-42 mu (half a unit of the last slice), adding 2 * 2^-42 mu_i mu_j:
  |dG_ij| < 8 * 2^-42 mu_i mu_j per summed column, and each exp term moves by at most its
  value times expm1(2 |dG_ij| / 2 sigma^2) (plus the FP64 form's rounding).  A dropped anti-diagonal, a wrong slice weight or a
  wrong row scale breaks the bound by orders of magnitude;
* FCHL19 descriptors (config-0 molecules, sigma 2): entries within 1e-10 relative;
* the energy fit of the 800/200 config-0 split at lambda = 1e-8: within 1e-6 relative of
  the oracle's FP64 fit."""

import numpy as np
import pytest

import oracle
from tests.helpers import rel_err
from tests.test_gpu_local_kernels import random_set, windowed_set

pytestmark = pytest.mark.gpu

P = "i8x6"


@pytest.fixture(scope="module")
def lk():
    from kernelforge_amd import local_kernels

    return local_kernels


def _rows(x, q, n):
    """Packed atom rows, labels and molecule index of a padded set."""
    R, L, M = [], [], []
    for a in range(len(n)):
        for i in range(min(max(int(n[a]), 0), x.shape[1])):
            R.append(x[a, i])
            L.append(q[a, i])
            M.append(a)
    return np.array(R).reshape(len(R), x.shape[2]), np.array(L), np.array(M)


def i8_bound(x1, q1, n1, x2, q2, n2, sigma):
    """Entry-wise bound on |K_i8x6 - K| (see the module docstring)."""
    R1, L1, M1 = _rows(x1, q1, n1)
    R2, L2, M2 = _rows(x2, q2, n2)
    def mu(R):
        m = np.abs(R).max(1)
        return np.where(m > 0, 2.0 ** np.frexp(m)[1], 1.0)
    d2 = (R1 * R1).sum(1)[:, None] + (R2 * R2).sum(1)[None, :] - 2.0 * R1 @ R2.T
    c = 1.0 / (2.0 * sigma * sigma)
    term = np.exp(-np.maximum(d2, 0.0) * c) * (L1[:, None] == L2[None, :])
    dg = 8.0 * 2.0**-42 * mu(R1)[:, None] * mu(R2)[None, :] * x1.shape[2]
    b = term * np.expm1(2.0 * dg * c)
    out = np.zeros((len(n1), len(n2)))
    np.add.at(out, (np.repeat(M1, len(M2)), np.tile(M2, len(M1))), b.ravel())
    return out


def to_rfp(K):
    """Dense symmetric K in the RFP order of the oracle's index map."""
    n = K.shape[0]
    out = np.empty(n * (n + 1) // 2)
    for i in range(n):
        for j in range(i, n):
            out[oracle.rfp_index(n, i, j)] = K[i, j]
    return out


def assert_within_bound(K8, Ko, bound):
    slack = 1e-12 * np.max(np.abs(Ko))  # the FP64 form's own rounding
    assert np.all(np.abs(K8 - Ko) <= bound + slack), float(np.max(np.abs(K8 - Ko) - bound))


@pytest.mark.parametrize("nm,ma,rep,nlab", [(1, 3, 7, 2), (13, 5, 9, 3), (33, 4, 16, 1),
                                            (70, 6, 33, 4), (130, 9, 720, 5), (67, 5, 64, 2)])
def test_symm_and_rfp_random(lk, nm, ma, rep, nlab):
    x, q, n = random_set(nm, ma, rep, nlab, nm)
    K = lk.kernel_gaussian_symm(x, q, n, 1.3, precision=P)
    assert K.dtype == np.float64
    Ko = oracle.kernel_symm(x, q, n, 1.3)
    bound = i8_bound(x, q, n, x, q, n, 1.3)
    assert_within_bound(K, Ko, bound)
    assert np.array_equal(K, K.T)
    # the RFP form holds the dense form's entries (one kernel, two output layouts)
    arf = lk.kernel_gaussian_symm_rfp(x, q, n, 1.3, precision=P)
    assert rel_err(arf, to_rfp(K)) < 1e-14


@pytest.mark.parametrize("nm1,nm2", [(1, 1), (5, 40), (33, 65), (100, 7)])
def test_rect_random(lk, nm1, nm2):
    x1, q1, n1 = random_set(nm1, 5, 24, 3, nm1)
    x2, q2, n2 = random_set(nm2, 7, 24, 3, nm2 + 100)
    K = lk.kernel_gaussian(x1, x2, q1, q2, n1, n2, 0.9, precision=P)
    assert K.shape == (nm1, nm2) and K.dtype == np.float64
    assert_within_bound(K, oracle.kernel_rect(x1, x2, q1, q2, n1, n2, 0.9),
                        i8_bound(x1, q1, n1, x2, q2, n2, 0.9))


@pytest.mark.parametrize("nm,ma,rep,nlab", [(300, 7, 720, 3), (200, 5, 1500, 2), (150, 4, 2100, 2),
                                            (129, 6, 40, 2)])
def test_symm_zero_windows(lk, nm, ma, rep, nlab):
    """Zero windows skipped per tile pair; 2100 columns (132 windows) run unmasked."""
    x, q, n = windowed_set(nm, ma, rep, nlab, rep + nm)
    sigma = 0.35 * np.sqrt(rep)
    K = lk.kernel_gaussian_symm(x, q, n, sigma, precision=P)
    assert_within_bound(K, oracle.kernel_symm(x, q, n, sigma), i8_bound(x, q, n, x, q, n, sigma))
    arf = lk.kernel_gaussian_symm_rfp(x, q, n, sigma, precision=P)
    assert rel_err(arf, to_rfp(K)) < 1e-14


def test_rect_zero_windows(lk):
    x1, q1, n1 = windowed_set(250, 6, 720, 3, 11)
    x2, q2, n2 = windowed_set(190, 5, 720, 3, 12)
    sigma = 0.35 * np.sqrt(720)
    K = lk.kernel_gaussian(x1, x2, q1, q2, n1, n2, sigma, precision=P)
    assert_within_bound(K, oracle.kernel_rect(x1, x2, q1, q2, n1, n2, sigma),
                        i8_bound(x1, q1, n1, x2, q2, n2, sigma))


@pytest.mark.parametrize("sigma", [4.0, 1.0, 0.5])
def test_exact_grid(lk, sigma):
    """One atom per molecule at t = 0..24 (one slice: t / 32 is a multiple of 2^-6): the
    Gram entries are exact, so every entry is exp of an exact argument, as in the FP64
    form's test; also pins the row scale and the row/column order of the int32 MFMA
    output."""
    t = np.random.default_rng(5).permutation(25).astype(np.float64)
    x = t.reshape(-1, 1, 1)
    q = np.ones((25, 1), np.int32)
    n = np.ones(25, np.int32)
    c = 1.0 / (2.0 * sigma * sigma)
    d2 = (t[:, None] - t[None, :]) ** 2
    ref = np.exp(-d2 * c)
    tol = 2.3e-16 * (4.0 + 4.0 * d2 * c + 4.0 * (t[:, None] ** 2 + t[None, :] ** 2) * c)
    for K in (lk.kernel_gaussian_symm(x, q, n, sigma, precision=P),
              lk.kernel_gaussian(x[:11], x, q[:11], q, n[:11], n, sigma, precision=P)):
        r = ref[: K.shape[0]]
        normal = r > 1e-300
        assert np.all(np.abs(K - r)[normal] <= (tol[: K.shape[0]] * r)[normal])


def test_exact_three_slices(lk):
    """Rows whose values are multiples of 2^-20 of their power-of-two scale need three
    slices; every slice pair then has weight above 2^-42, none is dropped, and the Gram
    entries, norms and distances are exact in FP64 (integers below 2^53 after scaling), so
    the result must equal the FP64 form's to the exp's rounding.  A wrong slice weight,
    a dropped anti-diagonal or a mis-scaled row breaks this by far more."""
    rng = np.random.default_rng(9)
    nm, ma, rep = 90, 4, 40
    k = rng.integers(-(2**19), 2**19, (nm, ma, rep)).astype(np.float64)
    scale = 2.0 ** rng.integers(-1, 2, (nm, ma, 1)).astype(np.float64)
    x = k * 2.0**-20 * scale
    x[rng.random((nm, ma, rep)) < 0.3] = 0.0
    q = rng.integers(0, 2, (nm, ma)).astype(np.int32)
    n = rng.integers(1, ma + 1, nm).astype(np.int32)
    sigma = 3.0
    K8 = lk.kernel_gaussian_symm(x, q, n, sigma, precision=P)
    K64 = lk.kernel_gaussian_symm(x, q, n, sigma)
    assert np.max(np.abs(K8 - K64) / np.maximum(np.abs(K64), 1e-300)) < 1e-14


def test_config0_descriptors(lk, config0):
    """FCHL19 descriptors of the 1,000 config-0 molecules: entries within 1e-10 relative."""
    _, _, _, C, Q, N = config0
    X = oracle.fchl19_batch(C, Q, N)
    arf = lk.kernel_gaussian_symm_rfp(X, Q, N, 2.0, precision=P)
    arf_o = oracle.kernel_symm_rfp(X, Q, N, 2.0)
    big = np.abs(arf_o) > 1e-200
    assert np.all(arf[~big] == pytest.approx(arf_o[~big], abs=1e-200))
    assert np.max(np.abs(arf - arf_o)[big] / np.abs(arf_o)[big]) < 1e-10
    Kr = lk.kernel_gaussian(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0, precision=P)
    Ko = oracle.kernel_rect(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0)
    assert rel_err(Kr, Ko) < 1e-11


def test_config0_energy_fit(lk, config0):
    """800/200 split of config 0, lambda 1e-8, per-element baseline: the test energies of
    the i8x6 kernel with the FP64 RFP Cholesky within 1e-6 relative of the oracle's FP64
    fit."""
    from kernelforge_amd import kernelmath

    _, zs, E, C, Q, N = config0
    X = oracle.fchl19_batch(C, Q, N)
    els = np.array([1, 6, 7, 8, 16])
    cnt = np.array([[np.sum(z == e) for e in els] for z in zs], dtype=np.float64)
    coef = np.linalg.lstsq(cnt[:800], E[:800], rcond=None)[0]
    y = E[:800] - cnt[:800] @ coef

    a8 = kernelmath.cho_solve_rfp(lk.kernel_gaussian_symm_rfp(X[:800], Q[:800], N[:800], 2.0,
                                                                precision=P), y.copy(), l2=1e-8)
    ao = oracle.cho_solve_rfp(oracle.kernel_symm_rfp(X[:800], Q[:800], N[:800], 2.0), y.copy(), 1e-8)
    Ks8 = lk.kernel_gaussian(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0, precision=P)
    Kso = oracle.kernel_rect(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0)
    e8 = Ks8 @ a8 + cnt[800:] @ coef
    eo = Kso @ ao + cnt[800:] @ coef
    assert np.max(np.abs(e8 - eo)) / np.max(np.abs(eo)) < 1e-6


def test_too_wide_for_int32(lk):
    """87,392 or more columns could overflow the int32 anti-diagonal sums: an error."""
    x = np.ones((1, 1, 90000))
    with pytest.raises(ValueError):
        lk.kernel_gaussian_symm(x, np.ones((1, 1), np.int32), np.ones(1, np.int32), 1.0, precision=P)
