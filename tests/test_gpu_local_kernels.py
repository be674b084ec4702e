"""GPU parity: scalar local Gaussian kernels (kf_local_kernel_symm_rfp / _symm / kf_local_kernel)
vs the oracle.  Bar: FP64 kernel entries within 1e-10 relative (to max |K|)."""

import numpy as np
import pytest

import oracle
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def lk():
    from kernelforge_amd import local_kernels

    return local_kernels


def random_set(nm, ma, rep, nlab, seed, scale=0.4, nmin=1):
    rng = np.random.default_rng(seed)
    x = rng.standard_normal((nm, ma, rep)) * scale
    q = rng.integers(0, nlab, (nm, ma)).astype(np.int32)
    n = rng.integers(nmin, ma + 1, nm).astype(np.int32)
    return x, q, n


@pytest.mark.parametrize("nm,ma,rep,nlab", [(1, 3, 7, 2), (13, 5, 9, 3), (33, 4, 16, 1),
                                            (70, 6, 33, 4), (130, 9, 720, 5)])
def test_symm_and_rfp_random(lk, nm, ma, rep, nlab):
    x, q, n = random_set(nm, ma, rep, nlab, nm)
    Ko = oracle.kernel_symm(x, q, n, 1.3)
    assert rel_err(lk.kernel_gaussian_symm(x, q, n, 1.3), Ko) < 1e-12
    assert rel_err(lk.kernel_gaussian_symm_rfp(x, q, n, 1.3), oracle.kernel_symm_rfp(x, q, n, 1.3)) < 1e-12


@pytest.mark.parametrize("nm1,nm2", [(1, 1), (5, 40), (33, 65), (100, 7)])
def test_rect_random(lk, nm1, nm2):
    x1, q1, n1 = random_set(nm1, 5, 24, 3, nm1)
    x2, q2, n2 = random_set(nm2, 7, 24, 3, nm2 + 100)
    K = lk.kernel_gaussian(x1, x2, q1, q2, n1, n2, 0.9)
    assert K.shape == (nm1, nm2)
    assert rel_err(K, oracle.kernel_rect(x1, x2, q1, q2, n1, n2, 0.9)) < 1e-12


def windowed_set(nm, ma, rep, nlab, seed, ngroups=7):
    """Random rows whose columns fall into groups (spans not aligned to the kernel's
    16-column windows) that each molecule uses or leaves at zero, single atoms dropping
    further groups, and a few all-zero molecules: the pattern of FCHL19 channels of
    absent neighbour elements, at random."""
    rng = np.random.default_rng(seed)
    x, q, n = random_set(nm, ma, rep, nlab, seed)
    cuts = np.sort(rng.choice(np.arange(1, rep), ngroups - 1, replace=False))
    spans = np.split(np.arange(rep), cuts)
    use = rng.random((nm, ngroups)) < 0.45
    use[:, 0] = True
    for a in range(nm):
        for g, cols in enumerate(spans):
            if not use[a, g]:
                x[a][:, cols] = 0.0
        for i in range(ma):
            if rng.random() < 0.2:
                x[a, i, spans[rng.integers(ngroups)]] = 0.0
    x[rng.random(nm) < 0.05] = 0.0
    return x, q, n


@pytest.mark.parametrize("nm,ma,rep,nlab", [(300, 7, 720, 3), (200, 5, 1500, 2), (150, 4, 2100, 2),
                                            (129, 6, 40, 2)])
def test_symm_zero_windows(lk, nm, ma, rep, nlab):
    """Molecules are reordered by the windows they use and zero windows are skipped per
    tile pair (descriptors up to 128 windows of 16 columns; 2100 columns run unmasked)."""
    x, q, n = windowed_set(nm, ma, rep, nlab, rep + nm)
    sigma = 0.35 * np.sqrt(rep)
    assert rel_err(lk.kernel_gaussian_symm(x, q, n, sigma), oracle.kernel_symm(x, q, n, sigma)) < 1e-12
    assert rel_err(lk.kernel_gaussian_symm_rfp(x, q, n, sigma),
                   oracle.kernel_symm_rfp(x, q, n, sigma)) < 1e-12


@pytest.mark.parametrize("nm1,nm2,rep", [(250, 190, 720), (70, 130, 1100)])
def test_rect_zero_windows(lk, nm1, nm2, rep):
    x1, q1, n1 = windowed_set(nm1, 6, rep, 3, 11)
    x2, q2, n2 = windowed_set(nm2, 5, rep, 3, 12)
    sigma = 0.35 * np.sqrt(rep)
    K = lk.kernel_gaussian(x1, x2, q1, q2, n1, n2, sigma)
    assert rel_err(K, oracle.kernel_rect(x1, x2, q1, q2, n1, n2, sigma)) < 1e-12


def test_config0_descriptors(lk, config0):
    _, _, _, C, Q, N = config0
    X = oracle.fchl19_batch(C, Q, N)
    arf = lk.kernel_gaussian_symm_rfp(X, Q, N, 2.0)
    arf_o = oracle.kernel_symm_rfp(X, Q, N, 2.0)
    assert rel_err(arf, arf_o) < 1e-10
    Kr = lk.kernel_gaussian(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0)
    assert rel_err(Kr, oracle.kernel_rect(X[800:], X[:800], Q[800:], Q[:800], N[800:], N[:800], 2.0)) < 1e-10


@pytest.mark.parametrize("sigma", [4.0, 1.0, 0.5, 0.25])
def test_exp_range_exact_distances(lk, sigma):
    """One atom per molecule on an integer grid t = 0..24: every squared distance
    (t_a - t_b)^2 and 1/(2 sigma^2) is exact, so each entry is exp of an exact argument,
    from 0 down to -4608 (through the subnormal range and past underflow).  Checked entry
    by entry against np.exp; the tolerance covers the Gram form's cancellation
    (|x_a|^2 + |x_b|^2 - 2 x_a.x_b, rounded in units of t_a^2 + t_b^2) and the argument's
    rounding, a few ulp per unit of |argument|, not the exp: a wrong table entry or a wrong
    polynomial coefficient is far outside it."""
    t = np.random.default_rng(5).permutation(25).astype(np.float64)
    x = t.reshape(-1, 1, 1)
    q = np.ones((25, 1), np.int32)
    n = np.ones(25, np.int32)
    c = 1.0 / (2.0 * sigma * sigma)
    d2 = (t[:, None] - t[None, :]) ** 2
    ref = np.exp(-d2 * c)
    tol = 2.3e-16 * (4.0 + 4.0 * d2 * c + 4.0 * (t[:, None] ** 2 + t[None, :] ** 2) * c)
    for K in (lk.kernel_gaussian_symm(x, q, n, sigma),
              lk.kernel_gaussian(x[:11], x, q[:11], q, n[:11], n, sigma)):
        r = ref[: K.shape[0]]
        normal = r > 1e-300
        assert np.all(np.abs(K - r)[normal] <= (tol[: K.shape[0]] * r)[normal])
        assert np.all(np.abs(K - r)[~normal] <= 1e-300)


def test_no_shared_labels_and_clamping(lk):
    x, q, n = random_set(9, 4, 8, 2, 3)
    q2 = q + 10
    K = lk.kernel_gaussian(x, x, q, q2, n, n, 1.0)
    assert np.all(K == 0.0)
    n_bad = n.copy()
    n_bad[0] = 99
    n_bad[1] = -3
    assert rel_err(lk.kernel_gaussian_symm(x, q, n_bad, 1.0), oracle.kernel_symm(x, q, n_bad, 1.0)) < 1e-12


def test_errors(lk):
    x, q, n = random_set(4, 3, 5, 2, 1)
    with pytest.raises(ValueError):
        lk.kernel_gaussian_symm_rfp(x, q, n, 0.0)
    with pytest.raises(ValueError):
        lk.kernel_gaussian_symm_rfp(x, q, n, float("nan"))
    with pytest.raises(ValueError):
        lk.kernel_gaussian_symm_rfp(x, q[:, :2], n, 1.0)
    with pytest.raises(ValueError):
        lk.kernel_gaussian(x, x[..., :4], q, q, n, n, 1.0)


def test_keywords(lk):
    x, q, n = random_set(6, 3, 5, 2, 8)
    a = lk.kernel_gaussian_symm_rfp(X=x, Q=q, N=n, sigma=1.1)
    b = lk.kernel_gaussian_symm_rfp(x, q, n, 1.1)
    np.testing.assert_array_equal(a, b)
    c = lk.kernel_gaussian(X1=x, X2=x, Q1=q, Q2=q, N1=n, N2=n, sigma=1.1)
    np.testing.assert_array_equal(c, lk.kernel_gaussian(x, x, q, q, n, n, 1.1))


# ---------------------------------------------------------------------------
# BF16x3 split-precision path (precision='bf16x3'): accuracy bar vs the FP64 oracle.
# Tolerances: max |K_mixed - K_fp64| <= 2e-5 * max |K| (measured ~3e-6 in emulation).
# ---------------------------------------------------------------------------
def test_bf16x3_symm_config0(lk, config0):
    _, _, _, C, Q, N = config0
    X = oracle.fchl19_batch(C[:400], Q[:400], N[:400])
    K64 = oracle.kernel_symm(X, Q[:400], N[:400], 2.0)
    K = lk.kernel_gaussian_symm(X, Q[:400], N[:400], 2.0, precision="bf16x3")
    assert K.dtype == np.float32
    assert rel_err(K, K64) < 2e-5
    arf = lk.kernel_gaussian_symm_rfp(X, Q[:400], N[:400], 2.0, precision="bf16x3")
    assert rel_err(arf, oracle.kernel_symm_rfp(X, Q[:400], N[:400], 2.0)) < 2e-5


def test_bf16x3_rect_and_random(lk, config0):
    _, _, _, C, Q, N = config0
    X = oracle.fchl19_batch(C[:300], Q[:300], N[:300])
    Q, N = Q[:300], N[:300]
    Kr = lk.kernel_gaussian(X[:70], X[70:], Q[:70], Q[70:], N[:70], N[70:], 2.0, precision="bf16x3")
    assert rel_err(Kr, oracle.kernel_rect(X[:70], X[70:], Q[:70], Q[70:], N[:70], N[70:], 2.0)) < 2e-5
    x, q, n = random_set(150, 6, 33, 3, 5)
    assert rel_err(lk.kernel_gaussian_symm(x, q, n, 1.3, precision="bf16x3"),
                   oracle.kernel_symm(x, q, n, 1.3)) < 2e-5


def test_two_threads_share_the_device_workspace(lk):
    """Two Python threads call the kernel at once (the modules release the GIL); the
    library serialises entry points on one device, so each thread gets its own kernel."""
    import threading

    sets = [random_set(90, 8, 720, 5, 501 + t) for t in range(2)]
    want = [oracle.kernel_symm_rfp(x, q, n, 2.0) for x, q, n in sets]
    got = [[None] * 6 for _ in sets]
    errors = []

    def work(t):
        try:
            x, q, n = sets[t]
            for k in range(6):
                got[t][k] = lk.kernel_gaussian_symm_rfp(x, q, n, 2.0)
        except Exception as exc:  # reported by the assertion below
            errors.append(exc)

    th = [threading.Thread(target=work, args=(t,)) for t in range(2)]
    for h in th:
        h.start()
    for h in th:
        h.join()
    assert not errors, errors
    for t in range(2):
        for k in range(6):
            assert rel_err(got[t][k], want[t]) < 1e-12
