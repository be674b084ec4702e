"""Bridge between torch device tensors and the C ABI of libkf_hip.so.

The torch-tensor modules (``cuda_local_kernels``, ``cuda_fchl19_repr``) pass tensor
storage and torch's current stream straight to the kf_* entry points, so their results
stay in device memory and their work is ordered on torch's stream.

One HIP runtime per process.  PyTorch wheels bundle their own ``libamdhip64.so`` with
the same soname (``libamdhip64.so.7``) as the system ROCm runtime the library links.
When torch is imported first, the dynamic loader resolves the library's dependency to
the runtime torch already loaded, and both share devices, streams and allocations.
Loaded the other way round, the process ends up with two runtimes that do not see
each other's devices, so this module imports torch before it loads the library and
refuses to work when a second runtime is mapped.
"""

from __future__ import annotations

import torch  # noqa: I001 -- must precede loading libkf_hip.so (see the module docstring)

from . import capi

_CHECKED = False


def _hip_runtimes() -> set[str]:
    paths = set()
    with open("/proc/self/maps") as f:
        for line in f:
            if "libamdhip64" in line:
                paths.add(line.split()[-1])
    return paths


def lib():
    """libkf_hip.so bound to torch's HIP runtime (ImportError when two runtimes are mapped)."""
    global _CHECKED
    L = capi.lib()
    if not _CHECKED:
        if not torch.cuda.is_available():
            raise RuntimeError("the torch-tensor interface needs a visible GPU")
        rts = _hip_runtimes()
        if len(rts) > 1:
            raise ImportError(
                "two HIP runtimes are loaded (" + ", ".join(sorted(rts)) + "): import torch "
                "before any kernelforge_amd module that loads libkf_hip.so")
        _CHECKED = True
    return L


def check(rc: int, what: str) -> None:
    capi.check(rc, what)


def stream_of(t: torch.Tensor) -> int:
    return torch.cuda.current_stream(t.device).cuda_stream


def device_tensor(t, name: str) -> torch.Tensor:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.device.type != "cuda":
        raise ValueError(f"{name} must be a GPU tensor")
    return t


def f64(t: torch.Tensor, name: str) -> torch.Tensor:
    """FP64 contiguous view or copy (the kernels compute in FP64)."""
    device_tensor(t, name)
    if not (t.is_floating_point()):
        raise TypeError(f"{name} must be a floating-point tensor")
    return t.to(torch.float64).contiguous()


def i32(t: torch.Tensor, name: str) -> torch.Tensor:
    device_tensor(t, name)
    if t.is_floating_point() or t.is_complex():
        raise TypeError(f"{name} must be an integer tensor")
    return t.to(torch.int32).contiguous()


def same_device(*ts: torch.Tensor) -> torch.device:
    dev = ts[0].device
    for t in ts[1:]:
        if t.device != dev:
            raise ValueError("all tensors must be on the same device")
    return dev


def out_dtype(*ts: torch.Tensor) -> torch.dtype:
    """float64 when any floating input is float64, else float32 (the reference's dtype)."""
    return torch.float64 if any(t.dtype == torch.float64 for t in ts) else torch.float32


def naq(N: torch.Tensor, max_atoms: int) -> int:
    """3 * sum(clamp(N, 0, max_atoms)) (reads N back to the host)."""
    return 3 * int(N.clamp(0, max_atoms).sum().item())
