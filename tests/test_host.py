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
    assert lib.spk_plan_bytes(1024, 1, 1, 64, ctypes.byref(out)) == -1   # unknown option bit


def test_mixed_size_shapes_give_one_workgroup_shape():
    """spk_group_shape: every member size runs 256-thread workgroups of eight
    values per thread ((n / 8) * tile == 256); other sizes are refused."""
    from als_speckle_amd.device import group_shape
    for n, tile in ((256, 8), (512, 4), (1024, 2)):
        assert group_shape(n) == (tile, 8)
        assert (n // 8) * tile == 256
    for n in (64, 128, 2048, 4096, 1000):
        with pytest.raises(_capi.SpkError, match="256, 512 and 1024"):
            group_shape(n)
    lib = _capi.load()
    assert lib.spk_group_create(None, None) == -1            # SPK_EINVAL, no GPU call
    assert lib.spk_group_iterate(None, None, 1, None, 1, None, None) == -1


def test_plan_bytes_counts_the_buffers_of_each_option():
    """spk_plan_bytes grows by each option's buffers (per frame: phib 8 B/px,
    shrinkwrap maps 8 B/px, ipsf work 8 B/px, spectrum maps 8 B/px; two Phi
    layouts at 4096)."""
    from als_speckle_amd.device import plan_bytes
    n, B = 1024, 3
    nn = n * n
    base = plan_bytes(n, B, False)
    assert base >= B * nn * (8 + 8 + 4 + 4)
    assert plan_bytes(n, B, False, _capi.PB_DM) - base == B * nn * 8
    assert plan_bytes(n, B, False, _capi.PB_SHRINKWRAP) - base >= B * nn * 8
    assert plan_bytes(n, B, False, _capi.PB_IPSF) - base >= (B + 1) * nn * 8
    assert plan_bytes(n, B, False, _capi.PB_SPECTRUM) - base >= B * nn * 8
    assert plan_bytes(n, B, False, _capi.PB_TRACE) > base
    every = sum((_capi.PB_DM, _capi.PB_SHRINKWRAP, _capi.PB_IPSF, _capi.PB_SPECTRUM,
                 _capi.PB_TRACE))
    # DM with an ipsf or a spectrum adds the work buffers of its second projection
    cross = B * nn * 8 + B * nn * 8
    assert plan_bytes(n, B, False, every) - base == cross + sum(
        plan_bytes(n, B, False, o) - base for o in (1, 2, 4, 8, 16))
    # the dual Phi layout at 4096: psi + two Phi layouts, and two phib layouts
    big = 4096 * 4096
    assert plan_bytes(4096, 1, True) >= big * (8 * 3 + 8)
    assert plan_bytes(4096, 1, True, _capi.PB_DM) - plan_bytes(4096, 1, True) == big * 16


def test_phasing_has_no_cpu_path():
    with pytest.raises(RuntimeError):
        P.phasing(force_cpu=True)
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(RuntimeError, match="GPU"):
            P.phasing()


def test_server_budget_sizing():
    from als_speckle_amd.device import plan_bytes
    from als_speckle_amd.serve import CHUNK_OPTIONS, Server
    mix = {256: 0.4, 512: 0.3, 1024: 0.2, 2048: 0.1}
    slots = Server.size_slots(8 << 30, mix)
    used = sum(plan_bytes(n, k, False, CHUNK_OPTIONS) for n, k in slots.items())
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
        # the persistent launch mode is searched only on request
        assert not any(p for _, _, p in cands)
        assert {c[:2] for c in candidates(n, persist=True)} == {c[:2] for c in cands}
        assert len(candidates(n, persist=True)) == 2 * len(cands)
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


def _admit(q, now, free, idle=(), budget=10**9, in_use=0, growable=True):
    from als_speckle_amd.serve import plan_admission
    slots = {256: 4, 512: 2, 2048: 1}
    byts = {256: 100, 512: 200, 2048: 600}
    return plan_admission(q, now, free, list(idle), slots, byts, budget, in_use, growable)


def test_serving_admission_static_slots():
    """Fixed slots (no growth): first come first served within a size, a
    full size does not block another, requests that have not arrived wait."""
    q = [_req(0, 512, 0.0), _req(1, 512, 0.1), _req(2, 256, 0.2), _req(3, 512, 0.3),
         _req(4, 256, 5.0)]
    a = _admit(q, 1.0, {512: 1, 256: 2}, growable=False)
    assert [r.rid for r in a.admitted] == [0, 2]
    assert [r.rid for r in a.waiting] == [1, 3, 4]
    assert not a.allocate and not a.release
    a = _admit(a.waiting, 10.0, {512: 2, 256: 1}, growable=False)
    assert [r.rid for r in a.admitted] == [1, 3, 4] and not a.waiting


def test_serving_admission_grows_chunks_within_the_budget():
    # budget 1000, nothing allocated: 5 x 256 need two 4-slot chunks (200),
    # 3 x 512 two 2-slot chunks (400); the 2048 chunk (600) no longer fits
    q = [_req(i, 256, 0.0) for i in range(5)] + [_req(5 + i, 512, 0.0) for i in range(3)] + \
        [_req(8, 2048, 0.0), _req(9, 256, 0.0)]
    a = _admit(q, 0.0, {}, budget=1000)
    assert a.allocate == {256: 2, 512: 2}
    assert [r.rid for r in a.admitted] == [0, 1, 2, 3, 4, 5, 6, 7, 9]
    assert [r.rid for r in a.waiting] == [8] and not a.release


def test_serving_admission_shares_the_budget_across_sizes():
    """A size with no free slot takes memory from idle chunks of other sizes
    (oldest idle first), never from chunks with running requests."""
    # 800 of 1000 in use: two idle 256 chunks (oldest first), one 512 chunk
    # with a running request, one idle 512 chunk
    idle = [(256, 0), (512, 1), (256, 1)]
    a = _admit([_req(0, 2048, 0.0)], 0.0, {256: 8, 512: 3}, idle, budget=1000, in_use=800)
    # needs 600: 200 free, + (256, 0) 100 + (512, 1) 200 + (256, 1) 100 = 600
    assert a.release == [(256, 0), (512, 1), (256, 1)]
    assert a.allocate == {2048: 1} and [r.rid for r in a.admitted] == [0]
    # not enough reclaimable memory: the request waits and nothing is released,
    # while a request of another size that fits still goes ahead
    a = _admit([_req(0, 2048, 0.0), _req(1, 256, 0.0)], 0.0, {256: 1}, [(256, 0)],
               budget=1000, in_use=900)
    assert not a.release and not a.allocate
    assert [r.rid for r in a.admitted] == [1] and [r.rid for r in a.waiting] == [0]
    # a later request of the same size waits behind the first (FCFS within a size)
    a = _admit([_req(0, 2048, 0.0), _req(1, 2048, 0.0)], 0.0, {}, [], budget=500)
    assert [r.rid for r in a.waiting] == [0, 1] and not a.allocate


def test_serving_chunk_shapes_fit_the_budget():
    from als_speckle_amd.device import plan_bytes
    from als_speckle_amd.serve import CHUNK_OPTIONS, chunk_shapes
    slots, byts = chunk_shapes((256, 512, 1024, 2048), 48 << 30)
    for n in slots:
        assert byts[n] == plan_bytes(n, slots[n], False, CHUNK_OPTIONS)
        assert 1 <= slots[n] <= 64 and (byts[n] <= (1 << 30) or slots[n] == 1)
    assert slots[256] >= slots[1024] >= slots[2048]
    small, _ = chunk_shapes((2048,), 1 << 30)       # an eighth of a small budget
    assert small[2048] == 1


def test_bench_finds_the_traffic_summary_of_either_dominant_pass():
    """roofline.traffic comes from the committed PMC summary of whichever
    pass dominates the headline (column or row pass)."""
    import bench
    for dom in ("column", "row"):
        path = bench.traffic_file(dom)
        assert os.path.exists(path), path
        t = bench.measured_traffic(dom)
        assert isinstance(t, int) and t > 0
