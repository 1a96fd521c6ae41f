import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X GPU")
    config.addinivalue_line("markers", "slow: longer CPU test")


@pytest.fixture(scope="session")
def emu_lib(tmp_path_factory):
    """g++ build of tests/emu/fft_emu.cpp (CPU emulation of the FFT kernels)."""
    import ctypes
    d = tmp_path_factory.mktemp("emu")
    so = str(d / "libemu.so")
    src = os.path.join(ROOT, "tests", "emu", "fft_emu.cpp")
    inc = os.path.join(ROOT, "als_speckle_amd", "csrc")
    subprocess.run(["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-I", inc, src, "-o", so],
                   check=True)
    lib = ctypes.CDLL(so)
    lib.emu_fft2.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                             ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    lib.emu_fft2.restype = ctypes.c_int
    lib.emu_stockham_out.argtypes = [ctypes.c_int] * 4
    lib.emu_stockham_out.restype = ctypes.c_int
    return lib


@pytest.fixture(scope="session")
def gpu_ctx():
    import torch
    if not torch.cuda.is_available():
        pytest.fail("GPU test run without a visible GPU")
    from als_speckle_amd.device import Context
    return Context(0)
