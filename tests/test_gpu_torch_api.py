"""GPU: the torch-tensor interface (cuda_local_kernels / cuda_fchl19_repr names) against
the oracle, including 100 MD-style steps of the cached matvec and its replay from a
captured graph.  The checks run in one child process (tests/torch_api_checks.py):
torch must load its HIP runtime before libkf_hip.so, which this process has already
loaded without torch."""

import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def test_torch_interface_vs_oracle():
    pytest.importorskip("torch")
    proc = subprocess.run([sys.executable, "-u", os.path.join(HERE, "torch_api_checks.py")],
                          stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True,
                          timeout=300)
    print(proc.stdout)
    assert proc.returncode == 0 and "TORCH_API_OK" in proc.stdout, proc.stdout[-3000:]
