"""Shared test setup.

CPU tests (`-m "not gpu"`): the oracle against golden vectors and the
reference's own CPU join, the host logic of the library, and its C ABI
surface (symbols, error paths, output.bin) without any device call.
GPU tests (`-m gpu`): parity of the HIP path against the oracle, through the
C ABI.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
sys.path.insert(0, os.path.join(ROOT, "brig-join_amd"))

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a HIP GPU (MI355X)")
    # the oracle is CPU test infrastructure: build it if it is missing
    lib = os.path.join(ROOT, "oracle", "lib", "libbjoracle.so")
    if not os.path.exists(lib):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "all"], check=True)


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    import brigjoin as bj
    assert torch.cuda.is_available(), "GPU tests need a HIP device"
    torch.cuda.init()
    ctx = bj.Context(0)
    yield ctx
    ctx.close()
