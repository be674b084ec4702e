"""GPU parity at the workloads' own sizes and regularisation (the other GPU tests use
random or small inputs):

* 1,000 molecules (seed 0), 800 training: alpha of the RFP Cholesky solve at
  lambda = 1e-8 against the oracle's cho_solve_rfp, on the same kernel and end to end.
* 20,000 molecules (seed 1): descriptors of a 64-molecule sample, a 64 x 20,000 strip of
  the FP64 RFP kernel, of the int8-split (i8x6) kernel and of the BF16x3 kernel against
  the oracle, and the FP64 RFP
  Cholesky at lambda = 1e-8 judged by its backward error, with the residual taken
  through kf_rfp_symv on a kept copy of the kernel.
* the learned-PCG pieces at the 2,000-conformer workload's lambda = 1e-10 (8 conformers
  of its seed-3 generator): block inverses and the unrolled-PCG loss and gradient.
"""

import numpy as np
import pytest

import oracle
from oracle import learned_pcg as lp
from tests.helpers import rel_err, row_rel_err

pytestmark = pytest.mark.gpu

LAMBDA = 1e-8


def rfp_upper_index(n, i, j):
    """Vectorised RFP (TRANSR='N', UPLO='U') position of upper element (i, j), i <= j."""
    i, j = np.minimum(i, j), np.maximum(i, j)
    k = n // 2
    s = n + 1 if n % 2 == 0 else n
    return np.where(j >= k, (j - k) * s + i, i * s + j + k + 1)


def test_rfp_upper_index_matches_oracle():
    rng = np.random.default_rng(0)
    for n in (1, 2, 7, 64, 1001, 20000):
        i, j = rng.integers(0, n, 50), rng.integers(0, n, 50)
        want = [oracle.rfp_index(n, min(a, b), max(a, b)) for a, b in zip(i, j)]
        np.testing.assert_array_equal(rfp_upper_index(n, i, j), want)


def test_config0_alpha_at_workload_lambda(config0):
    from kernelforge_amd import fchl19_repr, kernelmath, local_kernels

    _, _, energies, C, Q, N = config0
    tr = slice(0, 800)
    y = energies[tr] - energies[tr].mean()
    Xo = oracle.fchl19_batch(C[tr], Q[tr], N[tr])
    arf_o = oracle.kernel_symm_rfp(Xo, Q[tr], N[tr], 2.0)
    a_o = oracle.cho_solve_rfp(arf_o.copy(), y, LAMBDA)
    # the solve alone, on the oracle's kernel
    a_g = kernelmath.cho_solve_rfp(arf_o.copy(), y, l2=LAMBDA)
    # end to end: GPU descriptors -> GPU kernel -> GPU solve
    Xg = fchl19_repr.generate_fchl_acsf_batch(C[tr], Q[tr], N[tr])
    arf_g = local_kernels.kernel_gaussian_symm_rfp(Xg, Q[tr], N[tr], 2.0)
    a_e = kernelmath.cho_solve_rfp(arf_g, y, l2=LAMBDA)
    e_solve, e_e2e = rel_err(a_g, a_o), rel_err(a_e, a_o)
    print(f"config 0 alpha at lambda 1e-8: solve {e_solve:.2e}, end to end {e_e2e:.2e}")
    assert e_solve < 1e-10 and e_e2e < 1e-10


@pytest.fixture(scope="module")
def config1():
    """20,000 molecules of seed 1 on the device: descriptors, the FP64 RFP kernel and a
    64-molecule sample with the oracle's kernel strip against all 20,000."""
    from kernelforge_amd import capi, synthetic

    coords, zs, energies = synthetic.config_molecules(1)
    nm = 20000
    C, Q, N = synthetic.pad_batch(coords[:nm], zs[:nm])
    ma, rs = C.shape[1], capi.rep_size()
    DA = capi.DeviceArray
    dC, dQ, dN = DA.from_numpy(C), DA.from_numpy(Q), DA.from_numpy(N)
    X = DA((nm, ma, rs))
    capi.fchl19(dC.ptr, dQ.ptr, dN.ptr, nm, ma, X.ptr)
    arf = DA(nm * (nm + 1) // 2)
    capi.local_kernel_symm_rfp(capi.KF_PREC_F64, X.ptr, dQ.ptr, dN.ptr, nm, ma, rs, 2.0, arf.ptr)
    capi.synchronize()
    sample = np.sort(np.random.default_rng(20).choice(nm, 64, replace=False))
    Xh = X.numpy()
    strip = oracle.kernel_rect(Xh[sample], Xh, Q[sample], Q, N[sample], N, 2.0)
    y = energies[:nm] - energies[:nm].mean()
    return dict(nm=nm, ma=ma, rs=rs, C=C, Q=Q, N=N, dQ=dQ, dN=dN, X=X, Xh=Xh, arf=arf,
                sample=sample, strip=strip, y=y)


def _strip_from_rfp(arf_h, n, sample):
    cols = np.arange(n)
    return np.stack([arf_h[rfp_upper_index(n, np.full(n, i), cols)] for i in sample])


def test_config1_descriptors_sample(config1):
    c = config1
    s = c["sample"]
    Xo = oracle.fchl19_batch(c["C"][s], c["Q"][s], c["N"][s])
    assert row_rel_err(c["Xh"][s], Xo, lead=2) < 1e-12


def test_config1_fp64_kernel_strip(config1):
    c = config1
    got = _strip_from_rfp(c["arf"].numpy(), c["nm"], c["sample"])
    err = rel_err(got, c["strip"])
    print(f"20k FP64 kernel strip (64 x 20,000) vs oracle: {err:.2e}")
    assert err < 1e-12


def test_config1_bf16x3_kernel_strip(config1):
    from kernelforge_amd import capi

    c = config1
    nm = c["nm"]
    arf32 = capi.DeviceArray(nm * (nm + 1) // 2, np.float32)
    capi.local_kernel_symm_rfp(capi.KF_PREC_BF16X3, c["X"].ptr, c["dQ"].ptr, c["dN"].ptr, nm,
                               c["ma"], c["rs"], 2.0, arf32.ptr)
    capi.synchronize()
    got = _strip_from_rfp(arf32.numpy().astype(np.float64), nm, c["sample"])
    err = rel_err(got, c["strip"])
    print(f"20k BF16x3 kernel strip (64 x 20,000) vs oracle: {err:.2e}")
    assert err < 2e-5


def test_config1_i8x6_kernel_strip(config1):
    """The FP64-accurate int8-split form at the workload's size: every entry of the strip
    within 1e-10 relative of the oracle (the form's error on FCHL19 descriptors is ~1e-11)."""
    from kernelforge_amd import capi

    c = config1
    nm = c["nm"]
    arf8 = capi.DeviceArray(nm * (nm + 1) // 2)
    capi.local_kernel_symm_rfp(capi.KF_PREC_I8X6, c["X"].ptr, c["dQ"].ptr, c["dN"].ptr, nm,
                               c["ma"], c["rs"], 2.0, arf8.ptr)
    capi.synchronize()
    got = _strip_from_rfp(arf8.numpy(), nm, c["sample"])
    ref = c["strip"]
    big = np.abs(ref) > 1e-200
    entry = float(np.max(np.abs(got - ref)[big] / np.abs(ref)[big]))
    print(f"20k i8x6 kernel strip (64 x 20,000) vs oracle: max entry rel {entry:.2e}, "
          f"normwise {rel_err(got, ref):.2e}")
    assert entry < 1e-10
    assert np.all(np.abs(got - ref)[~big] <= 1e-200)


def test_config1_fp64_cholesky_backward_error(config1):
    """(K + lambda I) alpha = y at n = 20,000, lambda = 1e-8.  The normwise backward error
    ||r|| / (||K|| ||alpha|| + ||y||) of a backward-stable Cholesky solve is O(n u); it is
    computed here with max_i K_ii <= ||K||_2 in the denominator (an overestimate), and a
    wrong factor or solve gives O(1)."""
    from kernelforge_amd import capi

    c = config1
    nm = c["nm"]
    DA = capi.DeviceArray
    keep = DA(nm * (nm + 1) // 2)
    keep.copy_from(c["arf"])
    fac = c["arf"]
    capi.rfp_potrf(capi.KF_PREC_F64, fac.ptr, nm, LAMBDA)
    alpha = DA.from_numpy(c["y"])
    capi.rfp_potrs(capi.KF_PREC_F64, fac.ptr, nm, alpha.ptr, 1)
    Ka = DA(nm)
    capi.rfp_symv(capi.KF_PREC_F64, keep.ptr, nm, LAMBDA, alpha.ptr, Ka.ptr)
    capi.synchronize()
    a, y = alpha.numpy(), c["y"]
    r = Ka.numpy() - y
    kh = keep.numpy()
    dmax = float(np.max(kh[rfp_upper_index(nm, np.arange(nm), np.arange(nm))]))
    eta = np.linalg.norm(r) / (dmax * np.linalg.norm(a) + np.linalg.norm(y))
    print(f"20k FP64 RFP Cholesky at lambda 1e-8: relative residual "
          f"{np.linalg.norm(r) / np.linalg.norm(y):.2e}, backward error (max-diagonal norm) "
          f"{eta:.2e}, max|alpha| {np.max(np.abs(a)):.2e}")
    assert np.all(np.isfinite(a))
    assert eta < 1e-10


def test_config4_regime_blocks_and_unrolled_gradient():
    """Block inverses and the unrolled-PCG loss/gradient at lambda = 1e-10, with the
    network's initial relative shift (1e-6 |diag K_bb|), on 8 conformers of the seed-3
    generator, against the NumPy oracle on the dense full kernel."""
    from kernelforge_amd import capi, synthetic
    from kernelforge_amd import learned_pcg as L

    lam = 1e-10
    base, z = synthetic.aspirin_base()
    C, Z, E, F = synthetic.conformers(base, z, 8, 3)
    s = L.DeviceSystem(C, Z, E, F)
    Q = np.tile(Z, (8, 1)).astype(np.int32)
    N = np.full(8, 21, np.int32)
    X, dX = oracle.fchl19_batch(C, Q, N, grad=True)
    K = oracle.full_symm(X, dX, Q, N, 2.0)
    y = np.concatenate([E, -F.reshape(-1)])
    shift = L.INIT_REL_SHIFT * s.dscale
    inv = L.factor(s, lam, shift)
    ref = lp.block_inverses(lp.diag_blocks(K, N, 21), N, 21, lam, shift)
    r = np.random.default_rng(5).standard_normal(s.n)
    dr, dz = capi.DeviceArray.from_numpy(r), capi.DeviceArray(s.n)
    capi.check(capi.lib().kf_block_apply(inv.ptr, s.nm, s.bs, s.dN.ptr, s.ma, dr.ptr, dz.ptr,
                                         None), "kf_block_apply")
    e_apply = rel_err(dz.numpy(), lp.apply_blocks(ref, N, 21, r))
    s.build_rfp()
    loss, g = L.unrolled_loss_grad(s, lam, inv, steps=10)
    loss_o, g_o = lp.unrolled_loss_grad(K + lam * np.eye(len(y)), y, ref, N, 21, 10)
    e_grad = rel_err(g, g_o)
    print(f"lambda 1e-10 (8 conformers): block apply {e_apply:.2e}, unrolled loss "
          f"{loss:.6f} vs {loss_o:.6f}, gradient {e_grad:.2e}")
    assert e_apply < 1e-6
    assert abs(loss - loss_o) < 1e-6
    assert e_grad < 1e-4
