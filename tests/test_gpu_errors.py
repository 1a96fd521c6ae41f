"""Error behaviour at the boundary.

- The C ABI returns a negative SPK_E* code with a message for a bad size,
  frame, state or step, and the plan stays usable afterwards.
- A step table with one bad entry changes nothing: every entry is checked
  before the first launch, so the iterate is bitwise what it was.
- The drop-in class and module functions raise the reference's
  AssertionErrors for the same misuse (speckle/phasing.py:241, 331-353,
  384-386, 417-418, 432-434, 556, 999, 1333, 1374-1375, 1429, 1556).
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan
from als_speckle_amd.phasing import align_global_phase, phasing, prtf, rftf

pytestmark = pytest.mark.gpu

EINVAL, ESTATE, EUNSUP = -1, -4, -5
HIO = (_capi.HIO, _capi.SEM_NUMPY, 0.8, 0.0, 0.0)


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def loaded(ctx, n=128):
    f = synth.make_frame(n, 5, shape="disk", radius=n // 6, photons=1e6)
    rng = np.random.default_rng(1)
    g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
    p = Plan(ctx, n)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    p.set_psi(dev(g, torch.complex64))
    return p


def raises_code(code, fn, *a, **kw):
    with pytest.raises(_capi.SpkError) as e:
        fn(*a, **kw)
    assert e.value.code == code, str(e.value)
    assert str(e.value).split(": ", 1)[1], "no message"
    return str(e.value)


@pytest.mark.parametrize("n,kw,code", [
    (1000, {}, EINVAL),              # not a power of two
    (64, dict(batch=0), EINVAL),
    (8, {}, EUNSUP),                 # below the smallest built size
    (8192, {}, EUNSUP),              # above the largest
    (4096, dict(tile=8), EUNSUP),    # no tile 8 at 4096
    (64, dict(ept=4), EUNSUP),       # 8 or 16 values per thread
    (2048, dict(tile=8, ept=8), EUNSUP),  # 1024 threads per workgroup at most
])
def test_plan_shapes_rejected(gpu_ctx, n, kw, code):
    raises_code(code, Plan, gpu_ctx, n, **kw)


def test_calls_out_of_order_and_bad_frames(gpu_ctx):
    n = 64
    p = Plan(gpu_ctx, n)
    msg = raises_code(ESTATE, p.iterate, [HIO])
    assert "modulus" in msg
    raises_code(ESTATE, p.set_psi, dev(np.ones((n, n), np.complex64), torch.complex64),
                mult_support=True)
    raises_code(ESTATE, p.shrinkwrap, 2.0, 0.2)
    p.close()
    m = np.ones((n, n), np.float32)
    shared = Plan(gpu_ctx, n, batch=2, shared=True)
    raises_code(EINVAL, shared.set_modulus, dev(m, torch.float32), 1)    # one modulus
    raises_code(EINVAL, shared.set_psi, dev(m, torch.complex64), 2, False)   # frames 0, 1
    shared.close()
    own = Plan(gpu_ctx, n, batch=2, shared=False)
    raises_code(EINVAL, own.set_support, dev(m, torch.float32), 2)
    raises_code(EINVAL, own.set_support, dev(m, torch.float32), -2)
    own.close()


@pytest.mark.parametrize("bad,code", [
    ((99, _capi.SEM_NUMPY, 0.8, 0.0, 0.0), EINVAL),          # unknown algorithm
    ((_capi.HIO, 7, 0.8, 0.0, 0.0), EINVAL),                 # unknown semantics
    ((_capi.DM, _capi.SEM_NUMPY, 0.0, 0.0, 0.0), EINVAL),    # DM needs beta != 0
    ((_capi.HIO, _capi.SEM_NUMPY, 0.8, 2.0, 1.5), EINVAL),   # tau outside [0, 1]
    ((_capi.HIO, _capi.SEM_NUMPY, 0.8, 20.0, 0.2), EUNSUP),  # blur radius beyond 47 px
])
def test_bad_step_changes_nothing(gpu_ctx, bad, code):
    p = loaded(gpu_ctx)
    before = p.get_psi().cpu().numpy()
    sup = p.get_support().cpu().numpy()
    sched = [HIO, (_capi.HIO, _capi.SEM_NUMPY, 0.8, 2.0, 0.2), HIO, bad, HIO]
    raises_code(code, p.iterate, sched)
    torch.cuda.synchronize()
    assert np.array_equal(p.get_psi().cpu().numpy(), before)
    assert np.array_equal(p.get_support().cpu().numpy(), sup)
    tr = p.iterate([HIO] * 3).cpu().numpy()         # still usable
    assert tr.shape == (3, 1) and np.all(np.isfinite(tr))
    assert not np.array_equal(p.get_psi().cpu().numpy(), before)
    p.close()


def test_dropin_load_assertions():
    r = phasing()
    with pytest.raises(AssertionError):
        r.seed()                                          # phasing.py:999: N unset
    with pytest.raises(AssertionError):
        r.load(modulus=[[1.0, 2.0], [3.0, 4.0]])          # :331 ndarray only
    with pytest.raises(AssertionError):
        r.load(modulus=np.ones((2, 64, 64)))              # :347 2-D
    with pytest.raises(AssertionError):
        r.load(modulus=np.ones((64, 32)))                 # :353 square
    f = synth.make_frame(64, 1, shape="disk", radius=10, photons=1e6)
    r.load(modulus=f.modulus)
    with pytest.raises(AssertionError):
        r.load(modulus=np.ones((128, 128)))               # :351 same shape on reload
    with pytest.raises(AssertionError):
        r.load(support=np.ones((2, 32, 32)))              # :384 2-D
    with pytest.raises(AssertionError):
        r.load(support=np.ones((32, 16)))                 # :386 square
    with pytest.raises(AssertionError):
        r.load(ipsf=np.ones((64, 32)))                    # :418 square
    with pytest.raises(AssertionError):
        r.load(spectrum=np.ones((3, 10)))                 # :434 two rows
    r.load(support=f.support, numtrials=1)
    r.seed(0)
    with pytest.raises(AssertionError):
        r.iterate(5, order=(("hio", 4), ("xyz", 1)))      # :1333 unknown algorithm
    with pytest.raises(AssertionError):
        r.iterate(5, order="hio")                         # :241 order type
    with pytest.raises(AssertionError):
        r.optimize_coherence(np.ones((8, 8), np.complex64), method="lbfgs")   # :556
    r.iterate(5, order=(("hio", 4), ("er", 1)))           # still usable
    assert np.all(np.isfinite(r.get_error_trace()))


def test_module_function_assertions():
    with pytest.raises(AssertionError):
        align_global_phase(np.ones(16, np.complex64))     # :1374 2-D or 3-D
    with pytest.raises(AssertionError):
        align_global_phase(np.ones((4, 4)))               # :1375 complex
    with pytest.raises(AssertionError):
        prtf(np.ones((16, 16), np.complex64))             # :1429 a stack
    with pytest.raises(AssertionError):
        rftf(np.ones((32, 32), np.complex64), np.ones((16, 16), np.float32))   # :1556
