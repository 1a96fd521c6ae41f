"""Per-frame schedules with DM and idle steps, each frame against its own
single-frame run (bitwise: a frame's kernels run the same code on the same
inputs whatever the other frames do).

- a frame idle at step k and DM at step k+1, while another frame's row pass
  at step k emits its DM estimate for step k+1: the idle frame's estimate must
  be recomputed, not taken as fresh;
- a step where one frame is DM and another is not: the non-DM frame must not
  read its (stale, here NaN) DM estimate buffer;
- whole-step NOOP entries with one step table for every frame: the error
  trace is NaN there and the iterate is untouched.
"""
import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan

pytestmark = pytest.mark.gpu

N = 256


def step(alg, beta=0.7):
    if alg == "noop":
        return (_capi.NOOP, 0, 0.0, 0.0, 0.0)
    return (_capi.ALG_CODES[alg], _capi.SEM_NUMPY, beta, 0.0, 0.0)


def frames(count):
    out = []
    for k in range(count):
        f = synth.make_frame(N, 40 + k, shape="polygon", fill=0.12, photons=1e7)
        rng = np.random.default_rng(70 + k)
        g = (rng.random((N, N)) + 1j * rng.random((N, N))).astype(np.complex64)
        out.append((f.modulus, f.support, g))
    return out


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def load(ctx, data):
    p = Plan(ctx, N, batch=len(data), shared=False)
    for k, (m, s, g) in enumerate(data):
        p.set_modulus(dev(m, torch.float32), k)
        p.set_support(dev(s, torch.float32), k)
        p.set_psi(dev(g, torch.complex64), k)
    return p


def single(ctx, frame, algs):
    """One frame alone, running its non-idle steps."""
    p = load(ctx, [frame])
    tr = p.iterate([step(a) for a in algs if a != "noop"]).cpu().numpy()[:, 0]
    psi = p.get_psi().cpu().numpy()
    p.close()
    return psi, tr


def table(scheds):
    """Per-frame step table (step-major, frames fastest) from per-frame lists."""
    return [step(scheds[f][k]) for k in range(len(scheds[0])) for f in range(len(scheds))]


def check_frames(p, data, scheds, tr):
    for f, algs in enumerate(scheds):
        psi, want_tr = single(p.ctx, data[f], algs)
        got = p.get_psi(f).cpu().numpy()
        assert np.array_equal(got, psi), f"frame {f}: iterate differs from its run alone"
        active = [k for k, a in enumerate(algs) if a != "noop"]
        assert np.array_equal(tr[active, f], want_tr), f"frame {f}: trace"
        assert np.all(np.isnan(tr[[k for k in range(len(algs)) if k not in active], f]))


def test_dm_after_an_idle_step_recomputes_its_estimate(gpu_ctx):
    print("[per-frame DM after NOOP]", flush=True)
    data = frames(3)
    scheds = [["hio", "dm", "dm", "er"],
              ["noop", "dm", "dm", "er"],
              ["hio", "noop", "dm", "dm"]]
    p = load(gpu_ctx, data)
    tr = p.iterate(table(scheds), per_frame=True).cpu().numpy()
    check_frames(p, data, scheds, tr)
    # and across calls: frame 1 idle in the first call and DM in the second
    # (a row pass emits only for a DM step of the same call)
    p2 = load(gpu_ctx, data[:2])
    t1 = p2.iterate(table([["hio"], ["noop"]]), per_frame=True).cpu().numpy()
    t2 = p2.iterate(table([["dm"], ["dm"]]), per_frame=True).cpu().numpy()
    check_frames(p2, data[:2], [["hio", "dm"], ["noop", "dm"]], np.concatenate([t1, t2]))
    p.close()
    p2.close()


def test_non_dm_frame_of_a_dm_step_ignores_its_estimate_buffer(gpu_ctx):
    print("[mixed DM step, NaN estimate buffer]", flush=True)
    data = frames(2)
    p = load(gpu_ctx, data)
    # fill frame 0's DM estimate buffer with NaN: a DM step on a NaN iterate
    nan = np.full((N, N), np.nan, np.complex64)
    p.set_psi(dev(nan, torch.complex64), 0, mult_support=False)
    p.iterate(table([["dm"], ["noop"]]), per_frame=True)
    p.set_psi(dev(data[0][2], torch.complex64), 0)
    scheds = [["hio", "raar", "er"], ["dm", "dm", "er"]]
    tr = p.iterate(table(scheds), per_frame=True).cpu().numpy()
    assert np.all(np.isfinite(p.get_psi(0).cpu().numpy()))
    check_frames(p, data, scheds, tr)
    p.close()


def test_idle_steps_of_a_shared_step_table(gpu_ctx):
    print("[NOOP steps without per-frame tables]", flush=True)
    data = frames(1)
    p = load(gpu_ctx, data)
    tr = p.iterate([step("hio"), step("noop"), step("noop"), step("hio"), step("er")]).cpu().numpy()[:, 0]
    psi = p.get_psi().cpu().numpy()
    want, want_tr = single(gpu_ctx, data[0], ["hio", "hio", "er"])
    assert np.array_equal(psi, want)
    assert np.all(np.isnan(tr[1:3]))
    assert np.array_equal(tr[[0, 3, 4]], want_tr)
    p.close()
