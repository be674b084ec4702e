"""CPU checks of the drop-in boundary: libkf_hip.so loads and exports every symbol that
include/kf_hip.h declares, the ctypes binding covers them, and the pybind11 modules
expose the reference's function names and keyword names.  No GPU calls."""

import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "kf_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(kf_\w+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("kf_fchl19_f64", "kf_local_kernel_symm_rfp", "kf_rfp_potrf", "kf_rfp_potrs"):
        assert must in names


def test_library_exports_every_declared_symbol():
    import ctypes

    lib = ctypes.CDLL(os.path.join(ROOT, "kernelforge_amd", "libkf_hip.so"))
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_ctypes_binding_covers_header():
    from kernelforge_amd import capi

    assert sorted(capi.SIGNATURES) == declared()
    capi.lib()  # argtypes bind without error


def test_pybind_modules_mirror_reference_names():
    from kernelforge_amd import fchl19_repr, kernelmath, local_kernels

    for f in ("compute_rep_size", "generate_fchl_acsf", "generate_fchl_acsf_and_gradients"):
        assert hasattr(fchl19_repr, f)
    for f in ("cho_solve_rfp", "solve_cholesky", "full_to_rfp", "rfp_to_full"):
        assert hasattr(kernelmath, f)
    for f in ("kernel_gaussian", "kernel_gaussian_symm", "kernel_gaussian_symm_rfp"):
        assert hasattr(local_kernels, f)
    doc = fchl19_repr.generate_fchl_acsf.__doc__
    for kw in ("coords", "nuclear_z", "elements", "nRs2", "nRs3", "nFourier", "eta2", "eta3",
               "zeta", "rcut", "acut", "two_body_decay", "three_body_decay", "three_body_weight"):
        assert kw in doc
    assert "K_rfp" in kernelmath.cho_solve_rfp.__doc__ and "l2" in kernelmath.cho_solve_rfp.__doc__


def test_compute_rep_size_is_host_only():
    from kernelforge_amd import fchl19_repr

    assert fchl19_repr.compute_rep_size(5, 24, 20, 2) == 720
    assert fchl19_repr.compute_rep_size(nelements=4, nbasis2=10, nbasis3=7, nabasis=6) == 460


def test_pcg_entry_points_reject_block_size_before_any_launch():
    """kf_ef_pcg_rfp / kf_ef_pcg_matfree / kf_pcg_unrolled_grad_rfp check the block size
    (bs == 1 + 3 max_atoms <= 97, the LDS block of the preconditioner kernels) before
    touching any pointer or launching a kernel, so this runs without a GPU."""
    import ctypes

    from kernelforge_amd import capi

    L = capi.lib()
    if L.kf_device_name(None, 0) == capi.KF_OK:
        pytest.skip("placeholder device pointers: CPU only")
    fake = 4096  # never dereferenced: the argument check comes first
    it, rel = ctypes.c_int(0), ctypes.c_double(0.0)
    loss = ctypes.c_double(0.0)
    for ma, bs in ((32, 98), (33, 100), (4, 12)):
        assert L.kf_ef_pcg_rfp(fake, 10, 1e-3, fake, 2, bs, fake, ma, fake, fake, 1e-8, 5,
                               ctypes.byref(it), ctypes.byref(rel), None, None) == capi.KF_ERR_ARG
        assert L.kf_ef_pcg_matfree(fake, fake, fake, fake, 2, ma, 720, 2.0, 1e-3, fake, bs, fake,
                                   fake, 1e-8, 5, ctypes.byref(it), ctypes.byref(rel), None,
                                   None) == capi.KF_ERR_ARG
        assert L.kf_pcg_unrolled_grad_rfp(fake, 10, 1e-3, fake, 2, bs, fake, ma, fake, 3,
                                          ctypes.byref(loss), fake, None) == capi.KF_ERR_ARG
        assert b"block size" in L.kf_last_error()
