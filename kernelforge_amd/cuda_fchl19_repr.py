"""Torch-tensor interface of the FCHL19 representation, with the function names and
arguments of the reference's GPU module (python/kernelforge/cuda_fchl19_repr.pyi:5-81),
on the FP64 HIP kernel of libkf_hip.so (kf_fchl19_f64).

Q holds element indices in [0, nelements) as in the reference's GPU module (the host
module fchl19_repr takes nuclear charges and an element list instead).  The results
come back in the coordinates' dtype: float32 for the reference's float32 inputs,
float64 for float64 inputs.  Jacobians have the reference's shape
(nm, max_atoms, rep, max_atoms, 3), the same memory as (nm, max_atoms, rep, 3*max_atoms).
With Jacobians the three-body angular term keeps its first harmonic only, as the
reference's gradient path does (src/fchl19_repr.cpp:449,523-528).  ``deterministic`` is
accepted for the reference's signature: these kernels always are.

Import torch before any other kernelforge_amd module (see _torch_bridge).
"""

from __future__ import annotations

import math

import torch

from . import _torch_bridge as tb
from . import capi

__all__ = ["generate_fchl_acsf", "generate_fchl_acsf_and_gradients",
           "generate_fchl_acsf_and_gradients_md_single"]


def _run(coords, Q, N, nelements, grad, nRs2, nRs3, nFourier, eta2, eta3, zeta, rcut, acut,
         two_body_decay, three_body_decay, three_body_weight):
    tb.same_device(coords, Q, N)
    dtype = tb.out_dtype(coords)
    C = tb.f64(coords, "coords")
    Qi, Ni = tb.i32(Q, "Q"), tb.i32(N, "N")
    if Qi.dim() != 2:
        raise ValueError("Q must have shape (nm, max_atoms)")
    nm, ma = Qi.shape
    if C.numel() != nm * ma * 3 or tuple(Ni.shape) != (nm,):
        raise ValueError("coords must be (nm, max_atoms, 3) and N (nm,)")
    nelements = int(nelements)
    if nelements < 1:
        raise ValueError("nelements must be >= 1")
    nab = max(2 * int(nFourier), 2) if grad else 2 * int(nFourier)
    rep = capi.rep_size(nelements, nRs2, nRs3, nab)
    X = torch.empty((nm, ma, rep), dtype=torch.float64, device=C.device)
    dX = torch.empty((nm, ma, rep, ma, 3), dtype=torch.float64, device=C.device) if grad else None
    tb.lib()  # binds the library to torch's runtime (checked once)
    capi.fchl19(C.data_ptr(), Qi.data_ptr(), Ni.data_ptr(), nm, ma, X.data_ptr(),
                dX.data_ptr() if grad else 0, tb.stream_of(C), elements=tuple(range(nelements)),
                nRs2=nRs2, nRs3=nRs3, nFourier=nFourier, eta2=eta2, eta3=eta3, zeta=zeta,
                rcut=rcut, acut=acut, two_body_decay=two_body_decay,
                three_body_decay=three_body_decay, three_body_weight=three_body_weight)
    if dtype != torch.float64:
        X = X.to(dtype)
        dX = dX.to(dtype) if grad else None
    return (X, dX) if grad else X


def generate_fchl_acsf(coords, Q, N, nelements, nRs2=24, nRs3=20, nFourier=1, eta2=0.32,
                       eta3=2.7, zeta=math.pi, rcut=8.0, acut=8.0, two_body_decay=1.8,
                       three_body_decay=0.57, three_body_weight=13.4, deterministic=False):
    """Representations (nm, max_atoms, rep) of a padded batch; rows >= N[m] are zero."""
    del deterministic
    return _run(coords, Q, N, nelements, False, nRs2, nRs3, nFourier, eta2, eta3, zeta, rcut,
                acut, two_body_decay, three_body_decay, three_body_weight)


def generate_fchl_acsf_and_gradients(coords, Q, N, nelements, nRs2=24, nRs3=20, nFourier=1,
                                     eta2=0.32, eta3=2.7, zeta=math.pi, rcut=8.0, acut=8.0,
                                     two_body_decay=1.8, three_body_decay=0.57,
                                     three_body_weight=13.4, deterministic=False):
    """(rep (nm, max_atoms, rep), grad (nm, max_atoms, rep, max_atoms, 3))."""
    del deterministic
    return _run(coords, Q, N, nelements, True, nRs2, nRs3, nFourier, eta2, eta3, zeta, rcut,
                acut, two_body_decay, three_body_decay, three_body_weight)


def generate_fchl_acsf_and_gradients_md_single(coords, Q, N, nelements, nRs2=24, nRs3=20,
                                               nFourier=1, eta2=0.32, eta3=2.7, zeta=math.pi,
                                               rcut=8.0, acut=8.0, two_body_decay=1.8,
                                               three_body_decay=0.57, three_body_weight=13.4,
                                               deterministic=False):
    """generate_fchl_acsf_and_gradients for one molecule (nm == 1), the MD step's call."""
    if Q.dim() != 2 or Q.shape[0] != 1:
        raise ValueError("the md_single path takes one molecule (nm == 1)")
    del deterministic
    return _run(coords, Q, N, nelements, True, nRs2, nRs3, nFourier, eta2, eta3, zeta, rcut,
                acut, two_body_decay, three_body_decay, three_body_weight)
