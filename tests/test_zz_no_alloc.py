"""The allocation-free alignment entry point (spk_align_global_phase_noalloc)
against spk_align_global_phase, through the C ABI.

The two run the same kernels on the same partition of the sums, so their
results agree bitwise.  Two calls in a row, the second on a smaller stack
(the static workspace keeps the first call's partial sums; every entry the
second call reads is rewritten by it), and a stack larger than the
workspace is refused with SPK_EINVAL before any launch.
"""
import ctypes

import numpy as np
import pytest
import torch

from als_speckle_amd import _capi

pytestmark = pytest.mark.gpu


def stack(frames, rows, cols, seed):
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((frames, rows, cols)) + 1j * rng.standard_normal((frames, rows, cols))
    a *= np.exp(1j * rng.uniform(0, 2 * np.pi, (frames, 1, 1)))
    return torch.from_numpy(a.astype(np.complex64)).to("cuda")


def test_noalloc_alignment_equals_the_allocating_entry_point(gpu_ctx):
    lib = _capi.load()
    h = gpu_ctx.handle
    for frames, rows, cols, seed in ((64, 100, 120, 1), (5, 37, 29, 2), (819, 128, 128, 3)):
        print(f"[noalloc align {frames}x{rows}x{cols}]", flush=True)
        x = stack(frames, rows, cols, seed)
        a, b = x.clone(), x.clone()
        _capi.check(lib.spk_align_global_phase(h, ctypes.c_void_p(a.data_ptr()), frames, rows, cols))
        _capi.check(lib.spk_align_global_phase_noalloc(h, ctypes.c_void_p(b.data_ptr()), frames,
                                                       rows, cols))
        torch.cuda.synchronize()
        assert torch.equal(a, b)
        assert not torch.equal(a, x)           # the stack was rotated
    big = stack(820, 128, 128, 4)
    before = big.clone()
    rc = lib.spk_align_global_phase_noalloc(h, ctypes.c_void_p(big.data_ptr()), 820, 128, 128)
    assert rc == -1 and "static workspace" in _capi.last_error()
    torch.cuda.synchronize()
    assert torch.equal(big, before)
