"""CPU emulation of the FFT kernels' per-thread code (tests/emu/fft_emu.cpp).

The emulator compiles spk_fft_core.hpp -- the radix butterflies, the Stockham
stage code and the tiled-layout addressing that the kernels execute -- with
g++, and runs every workgroup thread in turn between barriers.
"""
import numpy as np
import pytest

SIZES = [16, 32, 64, 128, 256, 512, 1024, 2048, 4096]


@pytest.mark.parametrize("n", SIZES)
@pytest.mark.parametrize("tile", [2, 4, 8])
@pytest.mark.parametrize("inverse", [0, 1])
@pytest.mark.parametrize("tw_mode", [0, 1, 2])
@pytest.mark.parametrize("ept", [8, 16])
def test_emulated_fft2_matches_numpy(emu_lib, n, tile, inverse, tw_mode, ept):
    two_level = tw_mode == 1
    if n == 4096 and tile == 8:
        pytest.skip("tile 8 is not built for 4096")
    if (n // ept) * tile > 1024:
        pytest.skip("workgroup would exceed 1024 threads")
    if ept == 8 and (two_level or (n >= 2048 and tile != 2)):
        pytest.skip("8 values per thread: small and medium frames, full twiddle table")
    if n == 4096 and inverse:
        pytest.skip("one direction is enough at 4096 (2 s each)")
    if two_level and (n < 64 or (n < 1024 and tile == 8)):
        pytest.skip("two-level twiddles are used for large frames only")
    rng = np.random.default_rng(n + tile)
    x = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))).astype(np.complex64)
    out = np.zeros_like(x)
    if tw_mode == 0 and n > 256:
        pytest.skip("the exponent-indexed table is checked on small frames")
    assert emu_lib.emu_fft2(n, tile, inverse, x.ctypes.data, out.ctypes.data, tw_mode, ept) == 0
    xd = x.astype(np.complex128)
    ref = np.fft.ifft2(xd) * n * n if inverse else np.fft.fft2(xd)
    rel = np.linalg.norm(out - ref) / np.linalg.norm(ref)
    # FP32 Stockham: error grows ~ log2(n) * eps
    assert rel < 4e-8 * np.log2(n) + 1e-7, rel


@pytest.mark.parametrize("n,radices", [(1024, (16, 16, 4)), (512, (16, 16, 2)), (4096, (16, 16, 16))])
def test_stockham_output_is_a_permutation(emu_lib, n, radices):
    ns = 1
    for R in radices:
        seen = np.zeros(n, bool)
        for j in range(n // R):
            for r in range(R):
                seen[emu_lib.emu_stockham_out(j, r, ns, R)] = True
        assert seen.all()
        ns *= R


def test_emulator_under_address_and_undefined_sanitizers(tmp_path):
    """The kernels' index arithmetic (LDS slots, per-stage twiddle indices,
    layout positions) run under g++ ASan/UBSan with exactly sized buffers;
    every built shape up to 1024 round-trips."""
    import os
    import shutil
    import subprocess
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "emu_check")
    flags = ["-O0", "-g", "-std=c++17",
             "-fsanitize=address,undefined", "-fno-sanitize-recover=all"]
    subprocess.run(["g++", *flags, "-I", os.path.join(root, "als_speckle_amd", "csrc"),
                    os.path.join(root, "tests", "emu", "fft_emu.cpp"),
                    os.path.join(root, "tests", "emu", "emu_check_main.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    assert "0 bad" in r.stdout
