"""Host-side logic of the product, on the CPU: helpers of the drop-in API,
the synthetic generators, and the C-ABI library's exports (no compute calls).
"""
import ctypes
import os
import re

import numpy as np
import pytest

from als_speckle_amd import masking, synth
from als_speckle_amd import phasing as P
from als_speckle_amd import _capi
from oracle import phasing_c128 as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bounding_box_reference_vectors():
    box = np.zeros((256, 256))
    box[18:90, 45:63] = 1
    assert tuple(masking.bounding_box(box)) == (18, 90, 45, 63)
    assert tuple(masking.bounding_box(box, pad=14)) == (4, 104, 31, 77)


def test_bounding_box_matches_oracle():
    rng = np.random.default_rng(3)
    for _ in range(200):
        n = int(rng.integers(16, 64))
        m = np.zeros((n, n))
        r0, c0 = rng.integers(0, n - 2, 2)
        m[r0:r0 + rng.integers(1, n - r0), c0:c0 + rng.integers(1, n - c0)] = 1
        kw = dict(pad=int(rng.integers(0, 12)), force_to_square=bool(rng.integers(0, 2)),
                  power_of_2=bool(rng.integers(0, 2)))
        assert tuple(masking.bounding_box(m, **kw)) == tuple(O.bounding_box(m, **kw)), kw


def test_find_all_objects_matches_oracle():
    rng = np.random.default_rng(4)
    for _ in range(20):
        m = (rng.random((20, 20)) > 0.75).astype(float)
        a = masking.find_all_objects(m)
        b = O.find_all_objects(m)
        assert a.shape == b.shape
        assert np.array_equal(a, b)


def test_expand_order_matches_oracle():
    for order, it in [(None, 250), ((("hio", 5), ("er", 5), ("sf", 1)), 13),
                      ((("raar", 3), ("dm", 2)), 17)]:
        assert P.expand_order(order, it) == O.expand_order(order, it)
    with pytest.raises(AssertionError):
        P.expand_order((("xyz", 2),), 4)
    with pytest.raises(AssertionError):
        P.expand_order((("hio",),), 4)


def test_shrinkwrap_schedule_matches_oracle():
    sw = P.shrinkwrap_schedule(1000)
    assert sorted(sw) == [20 * k - 1 for k in range(1, 51)]
    sig = O.sigma_schedule(50)
    assert [sw[k][0] for k in sorted(sw)] == pytest.approx(sig)


def test_synth_frame_properties():
    f = synth.make_frame(256, 0, fill=0.15, photons=1e6)
    fill = f.true_support.mean()
    assert 0.12 < fill < 0.18
    assert f.modulus.dtype == np.float32 and f.modulus.shape == (256, 256)
    assert abs(np.sum(f.modulus.astype(np.float64) ** 2) - 1e6) < 5 * np.sqrt(1e6)
    assert np.all(f.support >= f.true_support)
    d = synth.make_frame(256, 1, shape="disk", radius=40, photons=1e6)
    assert abs(d.true_support.sum() - np.pi * 40 ** 2) < 0.01 * np.pi * 40 ** 2
    # deterministic
    g = synth.make_frame(256, 0, fill=0.15, photons=1e6)
    assert np.array_equal(g.modulus, f.modulus)


def _declared_symbols():
    src = open(os.path.join(ROOT, "include", "speckle_hip.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(spk_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    if not os.path.exists(_capi.LIB_PATH):
        pytest.fail("libspeckle_hip.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(_capi.LIB_PATH)
    decl = _declared_symbols()
    assert len(decl) >= 20
    for name in decl:
        assert hasattr(lib, name), name
    assert set(decl) == set(_capi.SIGNATURES), "ctypes table and header disagree"
    # host-only calls that need no GPU
    lib.spk_version.restype = ctypes.c_char_p
    assert b"gfx950" in lib.spk_version()
    lib.spk_supported_sizes.restype = ctypes.c_uint32
    assert lib.spk_supported_sizes() == sum(1 << k for k in range(4, 13))


def test_library_rejects_bad_arguments_without_gpu():
    lib = _capi.load()
    out = ctypes.c_size_t()
    assert lib.spk_plan_bytes(1000, 1, 1, 0, ctypes.byref(out)) == -1   # not a power of 2
    assert "bad plan size" in _capi.last_error()
    assert lib.spk_plan_bytes(1024, 1, 1, 0, ctypes.byref(out)) == 0
    assert out.value >= 1024 * 1024 * (16 + 8)


def test_phasing_has_no_cpu_path():
    with pytest.raises(RuntimeError):
        P.phasing(force_cpu=True)
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="GPU"):
            P.phasing()


def test_server_budget_sizing():
    from als_speckle_amd.device import plan_bytes
    from als_speckle_amd.serve import Server
    mix = {256: 0.4, 512: 0.3, 1024: 0.2, 2048: 0.1}
    slots = Server.size_slots(8 << 30, mix)
    used = sum(plan_bytes(n, k, False) for n, k in slots.items())
    assert used <= (8 << 30) * 1.01
    assert slots[2048] >= 1 and slots[256] >= 1
    # larger frames get more of the budget, smaller frames more slots
    assert slots[256] >= slots[2048]


def test_autotune_candidates_are_built_shapes():
    from als_speckle_amd.tune import candidates
    for n in (16, 64, 256, 1024, 2048, 4096):
        cands = candidates(n)
        assert cands and all(p in (False, True) for _, _, p in cands)
        for tile, ept, _ in cands:
            # every candidate is a shape the library builds
            assert (n // ept) * tile <= 1024 and not (n == 4096 and tile == 8)
        assert (4, 16, False) in cands
    assert (2, 8, False) in candidates(4096) and (4, 16, False) in candidates(4096)
    assert all(t != 8 for t, _, _ in candidates(4096))


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and not r.stdout.strip()


def test_bench_headline_schedule():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    bench = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(bench)
    steps, nsw = bench.headline_steps(1000)
    assert len(steps) == 1000 and nsw == 50
    algs = [s[0] for s in steps]
    assert algs.count(_capi.HIO) == 900 and algs.count(_capi.ER) == 100
    assert all(s[2] == 0.8 and s[1] == _capi.SEM_NUMPY for s in steps)
    sw = [(k, s[3], s[4]) for k, s in enumerate(steps) if s[3] > 0]
    assert [k for k, _, _ in sw] == list(range(19, 1000, 20))
    assert sw[0][1] == 3.0 and sw[-1][1] == 1.5 and all(t == 0.2 for _, _, t in sw)


def _req(rid, n, arrival, iterations=30, shrinkwrap=None):
    from als_speckle_amd.serve import Request
    z = np.zeros((n, n), np.float32)
    return Request(rid, z, z, z.astype(np.complex64), iterations=iterations,
                   shrinkwrap=shrinkwrap, arrival=arrival)


def test_serving_schedule_of_a_request():
    from als_speckle_amd.serve import _steps_for
    r = _req(0, 64, 0.0, iterations=45, shrinkwrap=True)
    r.order = (("hio", 4), ("er", 1))
    t = _steps_for(r)
    assert t.dtype == _capi.STEP_DTYPE and t.size == 45
    assert list(t["alg"][:6]) == [_capi.HIO] * 4 + [_capi.ER, _capi.HIO]
    sw = np.nonzero(t["sw_sigma"] > 0)[0]
    assert list(sw) == [19, 39] and t["sw_sigma"][19] == 3.0 and t["sw_sigma"][39] == 1.5
    assert np.all(t["beta"] == np.float32(0.8))


def test_serving_block_table_advances_each_slot_on_its_own_schedule():
    from als_speckle_amd.serve import _steps_for, block_table
    a = _steps_for(_req(0, 64, 0.0, iterations=30))
    b = _steps_for(_req(1, 64, 0.0, iterations=7, shrinkwrap=dict(every=3)))
    # slot 0 is 25 steps into its schedule, slot 2 just admitted, slot 1 idle
    table, takes = block_table({0: (a, 25), 2: (b, 0)}, 21, 3)
    assert table.shape == (21, 3) and takes == {0: 5, 2: 7}
    assert np.array_equal(table[:5, 0], a[25:30]) and np.all(table["alg"][5:, 0] == _capi.NOOP)
    assert np.all(table["alg"][:, 1] == _capi.NOOP)
    assert np.array_equal(table[:7, 2], b) and np.all(table["alg"][7:, 2] == _capi.NOOP)


def test_serving_admission_order():
    from als_speckle_amd.serve import admission_order
    q = [_req(0, 512, 0.0), _req(1, 512, 0.1), _req(2, 256, 0.2), _req(3, 512, 0.3),
         _req(4, 256, 5.0)]
    free = {512: 1, 256: 2}
    admitted, waiting = admission_order(q, free, now=1.0)
    # first come first served within a size; a full class does not block another;
    # requests that have not arrived wait
    assert [r.rid for r in admitted] == [0, 2]
    assert [r.rid for r in waiting] == [1, 3, 4]
    assert free == {512: 0, 256: 1}
    admitted, waiting = admission_order(waiting, {512: 2, 256: 1}, now=10.0)
    assert [r.rid for r in admitted] == [1, 3, 4] and not waiting
