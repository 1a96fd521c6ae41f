"""Batched trials and the continuous-batching server on the GPU.

The identity checks are bitwise: a frame's kernels run the same code on the
same inputs whatever else is resident, and no reduction spans frames.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import synth
from als_speckle_amd.batch import phasing_batch
from als_speckle_amd.phasing import phasing
from als_speckle_amd.serve import Request, Server
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu


def test_batch_trials_equal_sequential_trials():
    f = synth.make_frame(128, 2, shape="polygon", fill=0.1, photons=1e7)
    seeds = [11, 12, 13, 14]
    b = phasing_batch(4)
    b.load(f.modulus, f.support)
    b.seed(seeds)
    b.iterate(30, order=(("raar", 1),), beta=0.87)
    tr = b.get_error_trace()
    stack = b.get(b.roi_stack())
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=4)
    for t, sd in enumerate(seeds):
        r.seed(sd)
        r.iterate(30, order=(("raar", 1),), beta=0.87)
        assert np.array_equal(r.get_error_trace(), tr[:, t])
    buf = r.get(r.savebuffer)
    assert np.array_equal(buf[:4], stack)
    aligned, avg = b.aligned_average()
    want = O.align_global_phase(stack.astype(np.complex128))
    got = b.get(aligned)
    assert np.linalg.norm(got - want) / np.linalg.norm(want) < 1e-5
    assert np.allclose(b.get(avg), want.mean(0), atol=1e-5 * np.abs(want).max())


def test_batch_dm_runs_and_matches_single():
    f = synth.make_frame(64, 3, shape="disk", radius=10, photons=1e6)
    b = phasing_batch(3)
    b.load(f.modulus, f.support)
    b.seed([1, 2, 3])
    b.iterate(12, order=(("dm", 1),), beta=0.7)
    tr = b.get_error_trace()
    r = phasing()
    r.load(modulus=f.modulus, support=f.support, numtrials=1)
    r.seed(2)
    r.iterate(12, order=(("dm", 1),), beta=0.7)
    assert np.array_equal(r.get_error_trace(), tr[:, 1])


def _requests(specs):
    reqs = []
    for rid, (n, seed, arrival) in enumerate(specs):
        f = synth.make_frame(n, seed, shape="polygon", fill=0.12, photons=1e7)
        rng = np.random.default_rng(100 + rid)
        g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
        reqs.append(Request(rid, f.modulus, f.support, g, iterations=60,
                            shrinkwrap=True, arrival=arrival))
    return reqs


@pytest.mark.parametrize("block", [20, 21])
def test_server_result_same_alone_and_in_mix(block):
    """block 20: every shrinkwrap ends a launch (stand-alone kernels); 21:
    most run fused into the passes, and whether they do depends on the other
    requests in the pool, which must not change any result."""
    specs = [(64, 1, 0.0), (128, 2, 0.0), (64, 3, 0.0), (64, 4, 0.0), (128, 5, 0.0),
             (64, 6, 0.0)]
    mix = Server(slots={64: 2, 128: 1}, block=block)
    res = {r.rid: r for r in mix.run(_requests(specs))}
    assert sorted(res) == list(range(len(specs)))
    # queueing happened: more 64^2 requests than slots
    assert max(r.admitted for r in res.values()) > 0
    for rid in (2, 4):
        alone = Server(slots={64: 2, 128: 1}, block=block)
        one = alone.run([_requests(specs)[rid]])[0]
        assert np.array_equal(one.psi.cpu().numpy(), res[rid].psi.cpu().numpy())
        assert np.array_equal(one.support.cpu().numpy(), res[rid].support.cpu().numpy())
        assert np.array_equal(one.trace, res[rid].trace)


@pytest.mark.parametrize("block", [20, 21])
def test_server_matches_dropin_object(block):
    specs = [(64, 7, 0.0)]
    req = _requests(specs)[0]
    res = Server(slots={64: 3}, block=block).run([req])[0]
    r = phasing()
    r.load(modulus=req.modulus, support=req.support, numtrials=1)
    # the drop-in seeds from numpy; load the same start through the C ABI plan
    r._plan.set_psi(torch.from_numpy(req.guess).cuda(), -1, mult_support=True)
    r._seeded = True
    r.iterate(60, order=req.order, beta=req.beta, shrinkwrap=True)
    assert np.array_equal(r.get(r.psi_in), res.psi.cpu().numpy())
    assert np.array_equal(r.get_error_trace(), res.trace)


def test_autotune_picks_a_built_shape_and_caches(tmp_path):
    from als_speckle_amd.device import Context
    from als_speckle_amd.tune import autotune, candidates
    ctx = Context(0)
    path = str(tmp_path / "tune.json")
    r = autotune(ctx, 64, batch=3, shared=True, iters=4, cache_path=path)
    assert (r.tile, r.ept, r.persist) in candidates(64)
    assert len(r.table) == len(candidates(64))
    assert r.us_per_iter == min(row["us_per_iter"] for row in r.table)
    assert autotune(ctx, 64, batch=3, shared=True) is r          # memory cache
    import json
    disk = json.load(open(path))
    assert any(k.endswith("n=64|batch=3|shared=1") for k in disk)
    # a tuned batch gives the same trials as a default one to FP32 rounding
    # (different radix schedules round differently)
    f = synth.make_frame(64, 4, shape="disk", radius=10, photons=1e6)
    runs = []
    for auto in (False, True):
        b = phasing_batch(3, gpu_info=ctx, autotune=auto)
        b.load(f.modulus, f.support)
        b.seed([5, 6, 7])
        b.iterate(5, order=(("hio", 1),), beta=0.9)
        runs.append(b.get_error_trace())
    assert np.allclose(runs[0], runs[1], rtol=1e-5)


def test_server_cancel_and_a_bad_request_do_not_touch_the_others():
    """Cancelling a waiting and a running request, and a request whose data is
    NaN, leave the other results bitwise as they are alone."""
    specs = [(64, 11, 0.0), (64, 12, 0.0), (64, 13, 0.0), (64, 14, 0.0)]
    reqs = _requests(specs)
    bad = _requests([(64, 15, 0.0)])[0]
    bad.rid = 99
    bad.modulus = np.full_like(bad.modulus, np.nan)
    srv = Server(slots={64: 2}, block=21)
    for r in reqs[:3] + [bad]:
        srv.submit(r)
    srv.tick()                       # requests 0 and 1 run, 2 and 99 wait
    assert srv.cancel(1) and srv.cancel(2) and not srv.cancel(7)
    while srv.busy():
        srv.tick()
    res = {r.rid: r for r in srv.results}
    assert sorted(res) == [0, 99]
    assert np.all(np.isnan(res[99].trace))
    alone = Server(slots={64: 2}, block=21).run([_requests(specs)[0]])[0]
    assert np.array_equal(alone.psi.cpu().numpy(), res[0].psi.cpu().numpy())
    assert np.array_equal(alone.trace, res[0].trace)


# --- the serving workload's sizes: 256^2, 512^2, 1024^2 and 2048^2 ---------

MIX_SIZES = (256, 512, 1024, 2048)
MIX_CHUNKS = {256: 2, 512: 2, 1024: 1, 2048: 1}


def _mix_requests():
    """Ten requests of the serving mix's sizes (polygons 10-15 % fill,
    Poisson 1e7), 25 iterations of 90 HIO / 10 ER with a shrinkwrap after
    iteration 20, all arriving at once."""
    specs = [(256, 1), (256, 2), (256, 3), (512, 4), (512, 5), (512, 6), (1024, 7), (1024, 8),
             (2048, 9), (2048, 10)]
    reqs = []
    for rid, (n, seed) in enumerate(specs):
        rng = np.random.default_rng(200 + rid)
        f = synth.make_frame(n, 300 + seed, shape="polygon", fill=float(rng.uniform(0.10, 0.15)),
                             photons=1e7)
        g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
        reqs.append(Request(rid, f.modulus, f.support, g, iterations=25,
                            shrinkwrap=dict(every=20), arrival=0.0))
    return reqs


def _chunk_bytes(n, k):
    from als_speckle_amd.device import plan_bytes
    from als_speckle_amd.serve import CHUNK_OPTIONS
    return plan_bytes(n, k, False, CHUNK_OPTIONS)


def test_serving_mix_sizes_budget_sharing_and_alone_identity():
    """A budget of one chunk per size: the 2048^2 requests wait until the
    chunks of the smaller sizes go idle and are released to them; every
    result is bitwise the same as the request served alone."""
    print("[serving mix 256..2048, shared budget]", flush=True)
    budget = sum(_chunk_bytes(n, k) for n, k in MIX_CHUNKS.items())
    srv = Server(budget_bytes=budget, sizes=MIX_SIZES, chunk_slots=MIX_CHUNKS)
    res = {r.rid: r for r in srv.run(_mix_requests(), lockstep=True)}
    assert sorted(res) == list(range(10))
    assert srv.peak_bytes <= budget
    assert srv.chunks_released >= 1, "the budget never had to move between sizes"
    assert max(r.admitted for r in res.values() if r.n == 2048) > 0     # they queued
    srv.close()
    for rid in range(10):
        alone = Server(budget_bytes=64 << 30, sizes=MIX_SIZES, chunk_slots=MIX_CHUNKS)
        one = alone.run([_mix_requests()[rid]])[0]
        assert np.array_equal(one.psi.cpu().numpy(), res[rid].psi.cpu().numpy()), rid
        assert np.array_equal(one.support.cpu().numpy(), res[rid].support.cpu().numpy()), rid
        assert np.array_equal(one.trace, res[rid].trace), rid
        assert len(one.trace) == 25 and np.all(np.isfinite(one.trace))
        alone.close()


def test_serving_mix_sizes_lockstep_against_the_oracle():
    """The first iterations of one request per size, read back from the
    server after each one-iteration launch round, against the oracle's
    iteration from the same state (1e-5 relative)."""
    reqs = [r for r in _mix_requests() if r.rid in (0, 3, 6, 8)]
    srv = Server(budget_bytes=64 << 30, sizes=MIX_SIZES, chunk_slots=MIX_CHUNKS, block=1)
    for r in reqs:
        srv.submit(r)
    state = {r.rid: (r.guess.astype(np.complex128) * r.support) for r in reqs}
    for k in range(3):
        srv.tick()
        for r in reqs:
            print(f"[serving lockstep n={r.n} step {k}]", flush=True)
            done, psi = srv.peek(r.rid)
            assert done == k + 1
            got = psi.cpu().numpy()
            want, _, _ = O.iteration(state[r.rid], r.modulus, r.support, "hio", r.beta, "numpy")
            rel = np.linalg.norm(got - want) / np.linalg.norm(want)
            assert rel < 1e-5, (r.n, k, rel)
            state[r.rid] = got.astype(np.complex128)
    for r in reqs:
        assert srv.cancel(r.rid)
    srv.close()


def test_merged_size_classes_same_alone_and_in_mix():
    """merge_sizes: the 256^2, 512^2 and 1024^2 classes run every block as
    merged launches (spk_group_iterate); a request's result does not depend
    on what else is resident, and the merged launches are used."""
    print("[serve merged sizes]", flush=True)
    specs = [(256, 21, 0.0), (512, 22, 0.0), (256, 23, 0.0), (1024, 24, 0.0), (512, 25, 0.0),
             (256, 26, 0.0)]
    slots = {256: 2, 512: 1, 1024: 1}
    mix = Server(slots=slots, block=21, merge_sizes=True)
    assert mix.group is not None
    res = {r.rid: r for r in mix.run(_requests(specs))}
    assert sorted(res) == list(range(len(specs)))
    assert mix.group.counts()[0] > 0 and mix.group.counts()[1] > mix.group.counts()[0]
    mix.close()
    for rid in (1, 3, 5):
        alone = Server(slots=slots, block=21, merge_sizes=True)
        one = alone.run([_requests(specs)[rid]])[0]
        assert np.array_equal(one.psi.cpu().numpy(), res[rid].psi.cpu().numpy())
        assert np.array_equal(one.support.cpu().numpy(), res[rid].support.cpu().numpy())
        assert np.array_equal(one.trace, res[rid].trace)
        alone.close()
