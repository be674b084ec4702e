"""kernelforge_amd -- MI355X (gfx950) implementation of kernelforge's FCHL19
local-kernel KRR hot path.

Drop-in modules (same function names, keyword names, defaults and error types as
the reference's pybind11 modules):

    from kernelforge_amd import fchl19_repr, fchl19v2_repr, local_kernels, kernelmath

and the torch-tensor interface of the reference's GPU modules (import torch first):

    import torch
    from kernelforge_amd import cuda_local_kernels, cuda_fchl19_repr

Each is a compiled pybind11 extension that calls hand-written HIP kernels through
the C ABI of ``libkf_hip.so`` (``include/kf_hip.h``).  There is no host fallback:
importing a module whose extension is not built raises ImportError.  Build with
``python -m kernelforge_amd.build``.

Submodules are imported lazily, so ``import kernelforge_amd`` costs nothing.
"""

from __future__ import annotations

import importlib

__all__ = ["fchl19_repr", "fchl19v2_repr", "local_kernels", "kernelmath", "capi", "synthetic",
           "cuda_local_kernels", "cuda_fchl19_repr"]
__version__ = "0.1.0"


def __getattr__(name):
    if name in __all__:
        mod = importlib.import_module(f"{__name__}.{name}")
        globals()[name] = mod
        return mod
    raise AttributeError(name)
