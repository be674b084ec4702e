"""Build the gfx950 library and the pybind11 modules in-tree.

    python -m kernelforge_amd.build            # everything
    python -m kernelforge_amd.build --jobs 8

Outputs (next to this file):
  libkf_hip.so                          hand-written HIP kernels + the C ABI (include/kf_hip.h)
  fchl19_repr / local_kernels / kernelmath .cpython-*.so   pybind11 host layer over the C ABI

hipcc compiles device code for gfx950 only; packed-FP32 instruction selection is
disabled (-target-feature -packed-fp32-ops).  No CMake, no JIT cache.
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
ROOT = os.path.dirname(HERE)
BUILD = os.path.join(HERE, "_build")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
HIPCC = os.path.join(ROCM, "bin", "hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kf_runtime.hip", "fchl19.hip", "local_scalar.hip", "local_mixed.hip", "rfp_chol.hip",
               "rfp_convert.hip", "rfp_solve.hip", "local_deriv.hip",
               "local_matvec.hip", "ef_pcg.hip", "mlp.hip", "feature_compact.hip",
               "fchl19v2.hip"]
PY_MODULES = {"fchl19_repr": "py_fchl19_repr.cpp", "local_kernels": "py_local_kernels.cpp",
              "kernelmath": "py_kernelmath.cpp", "fchl19v2_repr": "py_fchl19v2_repr.cpp"}

HIP_FLAGS = ["--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC", "-Wall",
             "-Wno-unused-function", "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if r.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" + out)
    return out


def _hip_obj(src):
    obj = os.path.join(BUILD, os.path.basename(src) + ".o")
    _run([HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, src), "-o", obj])
    return obj


def _py_module(name, src, ext, pyinc, pybinc):
    out = os.path.join(HERE, name + ext)
    _run(["g++", "-O2", "-shared", "-fPIC", "-std=c++17", "-fvisibility=hidden",
          "-D__HIP_PLATFORM_AMD__", "-I" + pybinc,
          "-I" + pyinc, "-I" + os.path.join(ROCM, "include"),
          "-I" + os.path.join(ROOT, "include"), os.path.join(CSRC, src), "-o", out,
          "-L" + HERE, "-lkf_hip", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
          "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    return out


def build(jobs: int | None = None, verbose: bool = False) -> list[str]:
    os.makedirs(BUILD, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_hip_obj, HIP_SOURCES))
    lib = os.path.join(HERE, "libkf_hip.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, lib)
    import pybind11

    # sysconfig is not safe to initialise from several threads at once
    ext = sysconfig.get_config_var("EXT_SUFFIX")
    pyinc = sysconfig.get_paths()["include"]
    pybinc = pybind11.get_include()
    with cf.ThreadPoolExecutor(jobs) as ex:
        mods = list(ex.map(lambda kv: _py_module(kv[0], kv[1], ext, pyinc, pybinc),
                           PY_MODULES.items()))
    if verbose:
        print("built", lib, *mods, sep="\n  ")
    return [lib, *mods]


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=None)
    args = ap.parse_args(argv)
    build(args.jobs, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
