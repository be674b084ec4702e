"""GPU: kf_rfp_potrs_noalloc (RFP triangular solves with the scratch in a fixed device
array) against kf_rfp_potrs on the same factor, twice in a row with the second order
smaller, and its size limit.  Both entry points run the same kernels, so the solutions
must be bit-identical; each is also checked against numpy.linalg.solve."""

import numpy as np
import pytest
import scipy.linalg as sl

from tests.helpers import rel_err



def _spd_rfp(n, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((n, n))
    K = A @ A.T / n + np.eye(n)
    arf = sl.get_lapack_funcs("trttf", (K,))(np.asfortranarray(K), transr="N", uplo="U")[0]
    return K, arf.astype(dtype)


def _noalloc(L):
    return L.kf_rfp_potrs_noalloc  # argtypes bound by capi.lib() from the header table


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["f64", "f32"])
def test_noalloc_matches_potrs_two_calls(prec):
    from kernelforge_amd import capi

    L = capi.lib()
    noalloc = _noalloc(L)
    P = capi.KF_PREC_F64 if prec == "f64" else capi.KF_PREC_F32
    dt = np.float64 if prec == "f64" else np.float32
    for n, nrhs, seed in ((1000, 2, 1), (300, 1, 2)):  # second call smaller
        K, arf = _spd_rfp(n, seed, dt)
        fac = capi.DeviceArray.from_numpy(arf)
        capi.rfp_potrf(P, fac.ptr, n, 0.0)
        B = np.random.default_rng(seed + 10).standard_normal((nrhs, n))
        b_ref, b_new = capi.DeviceArray.from_numpy(B), capi.DeviceArray.from_numpy(B)
        capi.rfp_potrs(P, fac.ptr, n, b_ref.ptr, nrhs)
        rc = noalloc(P, fac.ptr, n, b_new.ptr, nrhs, None)
        capi.check(rc, "kf_rfp_potrs_noalloc")
        capi.synchronize()
        x_ref, x_new = b_ref.numpy(), b_new.numpy()
        print(f"noalloc {prec} n={n} nrhs={nrhs}: max |diff| vs kf_rfp_potrs "
              f"{np.max(np.abs(x_new - x_ref)):.1e}", flush=True)
        assert np.array_equal(x_new, x_ref)
        Kd = K.astype(dt).astype(np.float64) if prec == "f32" else K
        want = np.linalg.solve(Kd, B.T).T
        assert rel_err(x_new, want) < (1e-10 if prec == "f64" else 1e-3)


def _scratch_need(n, nrhs):
    """Bytes of scratch a solve of order n needs: the inverses of the diagonal blocks of
    groups of four 64-wide blocks (ten 64 x 64 FP64 tiles per group; blocks of 64 inside
    [0, n/2) and inside [n/2, n)) plus the n x nrhs intermediate."""
    k = n // 2
    nblk = -(-k // 64) + -(-(n - k) // 64)
    return 8 * -(-nblk // 4) * 10 * 64 * 64 + 8 * n * nrhs


@pytest.mark.cpu_only
def test_noalloc_rejects_orders_beyond_its_scratch():
    """Runs only where no GPU is visible: the placeholder pointers below must never reach
    a kernel, and the argument check that rejects the size comes before any device call."""
    from kernelforge_amd import capi

    L = capi.lib()
    if L.kf_device_name(None, 0) == capi.KF_OK:
        pytest.skip("placeholder device pointers: CPU only")
    noalloc = _noalloc(L)
    limit = 48 << 20
    for n, nrhs in ((60000, 1), (20000, 400)):
        assert _scratch_need(n, nrhs) > 1.4 * limit
        assert noalloc(capi.KF_PREC_F64, 4096, n, 4096, nrhs, None) == capi.KF_ERR_ARG
        assert b"scratch" in L.kf_last_error()
