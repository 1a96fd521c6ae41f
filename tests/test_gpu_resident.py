"""Resident single-frame launches (k_resident: the frame's modulus, support
and iterate held in registers across a block of steps, Phi exchanged through
a grid barrier) against the one-launch-per-pass path.

The resident kernel runs the per-pass arithmetic (same transforms, projection
and update code; the library is built without implicit FMA contraction), so
iterates, error traces and supports agree bitwise, for every exchange mode
of Phi across the barrier.  Schedules cover HIO, RAAR, SF and ER in both
semantics (ER with numpy semantics skips the projected Phi), idle steps,
shrinkwraps (which end a block and run as stand-alone kernels), DM steps
(which run as per-pass launches inside the same call), calls with and
without an error trace, and the 1024^2 headline frame and schedule.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan
from als_speckle_amd.phasing import expand_order, shrinkwrap_schedule

pytestmark = pytest.mark.gpu


def step(alg, beta=0.8, sem="numpy", sigma=0.0, tau=0.0):
    if alg == "noop":
        return (_capi.NOOP, 0, 0.0, 0.0, 0.0)
    return (_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, sigma, tau)


def run(ctx, frame, guess, calls, resident, exchange=0, trace=True, tile=0):
    n = frame.modulus.shape[0]
    p = Plan(ctx, n, resident=1 if resident else 2, exchange=exchange, tile=tile)
    p.set_modulus(torch.from_numpy(frame.modulus).to(ctx.device))
    p.set_support(torch.from_numpy(frame.support).to(ctx.device))
    p.set_psi(torch.from_numpy(guess).to(ctx.device))
    errs = []
    for steps in calls:
        e = p.iterate(steps, error_trace=trace)
        errs.append(e.cpu().numpy() if e is not None else None)
    p.check()
    out = (p.get_psi().cpu().numpy(), errs, p.get_support().cpu().numpy(), p.resident_counts())
    p.close()
    return out


def same(a, b):
    assert np.array_equal(a[0], b[0]), "iterate"
    for x, y in zip(a[1], b[1]):
        if x is None:
            assert y is None
        else:
            assert np.array_equal(x, y, equal_nan=True), "error trace"
    assert np.array_equal(a[2], b[2]), "support"


def data(n, seed):
    f = synth.make_frame(n, seed, shape="polygon", fill=0.15, photons=1e7)
    rng = np.random.default_rng(seed + 7)
    g = (rng.random((n, n)) + 1j * rng.random((n, n))).astype(np.complex64)
    return f, g


MIXED = ([step("hio", 0.9)] * 6 + [step("er")] * 2 + [step("raar", 0.87)] * 3 +
         [step("noop")] + [step("sf", sem="opencl")] * 2 + [step("er", sem="opencl")] +
         [step("sf")] + [step("hio", sigma=2.0, tau=0.2)] + [step("hio")] * 4 +
         [step("dm", 0.7)] * 2 + [step("er")] + [step("hio")] * 3)


@pytest.mark.parametrize("n", [256, 512, 1024])
@pytest.mark.parametrize("exchange", [1, 2, 3])
def test_resident_equals_per_pass(gpu_ctx, n, exchange):
    print(f"[resident n={n} exchange={exchange}]", flush=True)
    f, g = data(n, 11)
    calls = [MIXED, [step("hio")] * 5 + [step("er")] * 3]
    ref = run(gpu_ctx, f, g, calls, resident=False)
    got = run(gpu_ctx, f, g, calls, resident=True, exchange=exchange)
    same(got, ref)
    launches, steps = got[3]
    assert launches >= 3 and steps >= 20       # the blocks ran resident
    assert ref[3] == (0, 0)


@pytest.mark.parametrize("n", [256, 512])
def test_resident_in_tile_2_equals_per_pass(gpu_ctx, n):
    """The resident launch is also built in tile 2 for 256^2 and 512^2 frames
    (twice the workgroups, half the lines each): the same bitwise agreement
    with the per-pass launches of a tile-2 plan, fused shrinkwraps included."""
    print(f"[resident tile 2 n={n}]", flush=True)
    f, g = data(n, 12)
    sw = [step("hio", sigma=2.0, tau=0.2) if k % 5 == 4 else step("hio") for k in range(20)]
    calls = [MIXED, sw + [step("er")] * 3]
    ref = run(gpu_ctx, f, g, calls, resident=False, tile=2)
    got = run(gpu_ctx, f, g, calls, resident=True, tile=2)
    same(got, ref)
    assert got[3][0] >= 3 and ref[3] == (0, 0)


def test_resident_without_error_trace_and_single_step_calls(gpu_ctx):
    print("[resident no trace]", flush=True)
    f, g = data(1024, 12)
    calls = [[step("hio")] * 9 + [step("er")], [step("hio")], [step("er")] * 4]
    ref = run(gpu_ctx, f, g, calls, resident=False, trace=False)
    got = run(gpu_ctx, f, g, calls, resident=True, trace=False)
    same(got, ref)
    assert got[3][0] == 2            # a one-step call runs per pass


def test_resident_headline_schedule(gpu_ctx):
    """The first 200 iterations of the 1024^2 headline schedule (90 HIO /
    10 ER, beta 0.8, shrinkwrap every 20 with sigma 3 -> 1.5, tau 0.2)."""
    print("[resident headline 200]", flush=True)
    c = synth.CONFIGS["single_1024"]
    f = synth.config_frame("single_1024", 0)
    np.random.seed(1)
    g = (np.random.rand(1024, 1024) + 1j * np.random.rand(1024, 1024)).astype(np.complex64)
    sw = shrinkwrap_schedule(c["iterations"], **c["shrinkwrap"])
    steps = [step(a, c["beta"], "numpy", *sw.get(k, (0.0, 0.0)))
             for k, a in enumerate(expand_order(c["order"], c["iterations"]))][:200]
    ref = run(gpu_ctx, f, g, [steps], resident=False)
    got = run(gpu_ctx, f, g, [steps], resident=True)
    same(got, ref)
    # one launch: the shrinkwraps after steps 19, 39, ..., 179 run inside it,
    # the one after the call's last step as stand-alone kernels
    assert got[3] == (1, 200)


def test_resident_shrinkwrap_cases(gpu_ctx):
    """Shrinkwraps inside a resident block: back to back, followed by an
    idle step, with a blur radius beyond the fused window (sigma 5: the
    block ends and the stand-alone kernels run), with ER steps whose update
    skips the projected Phi, and as the last step of a call."""
    print("[resident shrinkwrap cases]", flush=True)
    f, g = data(512, 13)
    sw = lambda alg, sigma: step(alg, 0.9, "numpy", sigma, 0.25)
    calls = [[step("hio", 0.9)] * 3 + [sw("hio", 2.0), sw("hio", 3.5), step("hio", 0.9),
                                       sw("er", 1.5), step("noop"), step("hio", 0.9),
                                       sw("hio", 5.0), step("hio", 0.9), step("hio", 0.9),
                                       sw("er", 2.5), step("er"), step("er"), sw("hio", 1.0)],
             [step("hio", 0.9), sw("hio", 2.0)]]
    for trace in (True, False):
        ref = run(gpu_ctx, f, g, calls, resident=False, trace=trace)
        got = run(gpu_ctx, f, g, calls, resident=True, trace=trace)
        same(got, ref)
        assert got[3][0] == 3          # sigma 5 splits the first call; the second is one block


def test_resident_error_only_steps_and_calls_between_blocks(gpu_ctx):
    """ER steps (numpy semantics) with an error trace, whose column phase only
    forms the error partials: odd and even runs of them inside a block and at
    its end, and a one-step call (per-pass launches reading the Phi the block
    left) after each block: bitwise the per-pass results."""
    print("[resident error-only steps]", flush=True)
    f, g = data(512, 14)
    calls = [[step("hio")] * 3 + [step("er")] * 3 + [step("hio")] * 2,
             [step("er")],
             [step("hio")] * 2 + [step("er")] * 2 + [step("raar", 0.87)] + [step("er")],
             [step("hio")],
             [step("er")] * 5,
             [step("sf")]]
    ref = run(gpu_ctx, f, g, calls, resident=False)
    got = run(gpu_ctx, f, g, calls, resident=True)
    same(got, ref)
    assert got[3][0] == 3          # three blocks ran resident


@pytest.mark.parametrize("kind", ["resident", "persist"])
def test_barrier_give_up_is_reported_and_cleared(gpu_ctx, kind):
    """A grid-barrier wait that gives up, forced from the host with a poll
    limit of 1 (the first workgroups to arrive stop waiting), lets the launch
    drain: the call returns, spk_plan_check reports SPK_EHIP once and clears
    the flag, and the next launch of a plan with the default limit gives the
    per-pass results.  Every access of the kernels stays in bounds whatever
    the values, so a launch past a given-up barrier computes garbage, never
    out of range."""
    print(f"[barrier give-up {kind}]", flush=True)
    from als_speckle_amd._capi import SpkError
    n = 1024
    f, g = data(n, 16)
    kw = dict(resident=1) if kind == "resident" else dict(resident=2, persist=True)
    p = Plan(gpu_ctx, n, barrier_polls=1, **kw)
    p.set_modulus(torch.from_numpy(f.modulus).to(gpu_ctx.device))
    p.set_support(torch.from_numpy(f.support).to(gpu_ctx.device))
    p.set_psi(torch.from_numpy(g).to(gpu_ctx.device))
    p.iterate([step("hio")] * 12, error_trace=True)
    with pytest.raises(SpkError) as e:
        p.check()
    assert e.value.code == -2
    p.check()                                   # the flag was cleared
    if kind == "resident":
        assert p.resident_counts()[0] >= 1
    p.close()
    calls = [[step("hio")] * 6 + [step("er")] * 2]
    ref = run(gpu_ctx, f, g, calls, resident=False)
    got = run(gpu_ctx, f, g, calls, resident=True)
    same(got, ref)
