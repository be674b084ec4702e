"""GPU parity: FCHL19 descriptors and Jacobians (kf_fchl19_f64) vs the oracle and the
reference golden vectors.  Bar: per-atom relative error <= 1e-12 (tests/helpers.py)."""

import numpy as np
import pytest

import oracle
from tests.helpers import row_rel_err

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def fr():
    from kernelforge_amd import fchl19_repr

    return fchl19_repr


def test_golden_rep_and_grad(fr, golden_fchl19):
    g = golden_fchl19
    for t in range(10):
        rep = fr.generate_fchl_acsf(g[f"coords_{t}"], g[f"z_{t}"])
        assert rep.shape == g[f"rep_{t}"].shape
        assert row_rel_err(rep, g[f"rep_{t}"]) < 1e-12
    for t in range(3):
        rep, grad = fr.generate_fchl_acsf_and_gradients(g[f"gcoords_{t}"], g[f"gz_{t}"])
        assert row_rel_err(rep, g[f"grep_{t}"]) < 1e-12
        assert grad.shape == g[f"ggrad_{t}"].shape
        assert row_rel_err(grad, g[f"ggrad_{t}"]) < 1e-12


def test_golden_nondefault(fr, golden_fchl19):
    from tests.golden.make_golden import NONDEFAULT

    g = golden_fchl19
    assert row_rel_err(fr.generate_fchl_acsf(g["nd_coords"], g["nd_z"], **NONDEFAULT),
                       g["nd_rep"]) < 1e-12
    rep, grad = fr.generate_fchl_acsf_and_gradients(g["nd_coords"], g["nd_z"], **NONDEFAULT)
    assert row_rel_err(rep, g["nd_grep"]) < 1e-12
    assert row_rel_err(grad, g["nd_ggrad"]) < 1e-12


def test_batch_config0_vs_oracle(fr, config0):
    _, _, _, C, Q, N = config0
    X = fr.generate_fchl_acsf_batch(C, Q, N)
    Xo = oracle.fchl19_batch(C, Q, N)
    assert X.shape == Xo.shape == (1000, 23, 720)
    assert row_rel_err(X, Xo, lead=2) < 1e-12


def test_batch_gradients_vs_oracle(fr, config0):
    _, _, _, C, Q, N = config0
    sl = slice(0, 40)
    X, dX = fr.generate_fchl_acsf_batch(C[sl], Q[sl], N[sl], with_gradients=True)
    Xo, dXo = oracle.fchl19_batch(C[sl], Q[sl], N[sl], grad=True)
    assert dX.shape == dXo.shape
    assert row_rel_err(X, Xo, lead=2) < 1e-12
    assert row_rel_err(dX, dXo, lead=2) < 1e-12


@pytest.mark.parametrize("kw", [dict(nFourier=2), dict(nRs2=10, nRs3=7, rcut=3.0, acut=2.6),
                                dict(three_body_weight=0.0), dict(elements=[1, 6, 7, 8, 16, 9])])
def test_parameter_variants(fr, config0, kw):
    coords, zs, *_ = config0
    for m in range(4):
        assert row_rel_err(fr.generate_fchl_acsf(coords[m], zs[m], **kw),
                           oracle.fchl19(coords[m], zs[m], **kw)) < 1e-12
        r, g = fr.generate_fchl_acsf_and_gradients(coords[m], zs[m], **kw)
        ro, go = oracle.fchl19(coords[m], zs[m], grad=True, **kw)
        assert row_rel_err(r, ro) < 1e-12
        assert row_rel_err(g, go) < 1e-12


def test_flat_coords_and_errors(fr, config0):
    coords, zs, *_ = config0
    a = fr.generate_fchl_acsf(coords[0], zs[0])
    b = fr.generate_fchl_acsf(coords[0].reshape(-1), zs[0])
    np.testing.assert_array_equal(a, b)
    with pytest.raises(RuntimeError):
        fr.generate_fchl_acsf(coords[0], np.full(len(zs[0]), 9, np.int32))
    with pytest.raises(ValueError):
        fr.generate_fchl_acsf(np.zeros((0, 3)), np.zeros(0, np.int32))
    with pytest.raises(ValueError):
        fr.generate_fchl_acsf(np.zeros((4, 2)), np.ones(4, np.int32))
    assert fr.compute_rep_size(5, 24, 20, 2) == 720


def test_c_abi_device_resident(config0):
    from kernelforge_amd import capi

    _, _, _, C, Q, N = config0
    ma = Q.shape[1]
    dC = capi.DeviceArray.from_numpy(C[:64])
    dQ = capi.DeviceArray.from_numpy(Q[:64])
    dN = capi.DeviceArray.from_numpy(N[:64])
    X = capi.DeviceArray((64, ma, 720))
    dX = capi.DeviceArray((64, ma, 720, 3 * ma))
    capi.fchl19(dC.ptr, dQ.ptr, dN.ptr, 64, ma, X.ptr, dX.ptr)
    capi.synchronize()
    Xo, dXo = oracle.fchl19_batch(C[:64], Q[:64], N[:64], grad=True)
    assert row_rel_err(X.numpy(), Xo, lead=2) < 1e-12
    assert row_rel_err(dX.numpy(), dXo, lead=2) < 1e-12
