"""Parity of the HIP path with the complex128 oracle, through the C ABI.

Tolerances (the bar stated for this path): per-iteration modulus error within
1e-5 relative; iterates within 1e-5 relative L2 per step (FP32 FFT error is
~2e-7); free-running final reconstructions within 1e-3 relative L2 after
global-phase alignment where the iteration is contractive (ER).  HIO is
chaotic: any FP32 implementation leaves the FP64 trajectory after ~15
iterations, so HIO parity is checked step by step along the GPU trajectory
(lockstep), which pins the iteration map itself.
"""
import math

import numpy as np
import pytest
import torch

from als_speckle_amd import _capi, synth
from als_speckle_amd.device import Plan
from oracle import phasing_c128 as O

pytestmark = pytest.mark.gpu

TOL_STEP = 1e-5
TOL_TRACE = 1e-5


def dev(a, dtype):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def frame(n, seed=0, noise=True):
    return synth.make_frame(n, seed, shape="disk", radius=max(3, n * 40 // 256),
                            photons=1e6, noise=noise)


def start_psi(n, s, seed=5):
    np.random.seed(seed)
    g = (np.random.rand(n, n) + 1j * np.random.rand(n, n)).astype(np.complex64)
    return (g * s).astype(np.complex64)


def load_plan(ctx, f, psi, batch=1, tile=0, ept=0):
    p = Plan(ctx, f.modulus.shape[0], batch=batch, shared=True, tile=tile, ept=ept)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    if batch == 1:
        p.set_psi(dev(psi, torch.complex64), mult_support=False)
    else:
        p.set_psi(dev(np.stack([psi] * batch), torch.complex64), mult_support=False)
    return p


def step(alg, beta=0.9, sem="numpy", sigma=0.0, tau=0.0):
    return (_capi.ALG_CODES[alg], _capi.SEM_CODES[sem], beta, sigma, tau)


# ------------------------------------------------------------------ FFT
@pytest.mark.parametrize("n", [16, 32, 64, 128, 256, 512, 1024, 2048, 4096])
@pytest.mark.parametrize("tile", [2, 4, 8])
@pytest.mark.parametrize("ept", [16, 8])
def test_fft2_matches_numpy(gpu_ctx, n, tile, ept):
    if n == 4096 and tile == 8:
        pytest.skip("tile 8 not built at 4096")
    if (n // ept) * tile > 1024:
        pytest.skip("shape not built (over 1024 threads)")
    print(f"[fft2 n={n} tile={tile} ept={ept}]", flush=True)
    rng = np.random.default_rng(n)
    cnt = 2 if n <= 1024 else 1
    x = (rng.standard_normal((cnt, n, n)) + 1j * rng.standard_normal((cnt, n, n))).astype(np.complex64)
    p = Plan(gpu_ctx, n, batch=1, tile=tile, ept=ept)
    fx = p.fft2(dev(x, torch.complex64)).cpu().numpy()
    ix = p.fft2(dev(x, torch.complex64), inverse=True).cpu().numpy()
    xd = x.astype(np.complex128)
    assert rel(fx, np.fft.fft2(xd)) < 1e-6
    assert rel(ix, np.fft.ifft2(xd)) < 1e-6


# ------------------------------------------------------- one-step parity
@pytest.mark.parametrize("n", [64, 256])
@pytest.mark.parametrize("alg,sem", [("hio", "numpy"), ("er", "numpy"), ("er", "opencl"),
                                     ("raar", "numpy"), ("sf", "numpy"), ("sf", "opencl"),
                                     ("dm", "numpy")])
@pytest.mark.parametrize("tile,ept", [(4, 16), (8, 16), (4, 8), (2, 8), (2, 16)])
def test_one_step_matches_oracle(gpu_ctx, n, alg, sem, tile, ept):
    print(f"[one-step n={n} {alg}/{sem} tile={tile} ept={ept}]", flush=True)
    f = frame(n)
    psi = start_psi(n, f.support)
    beta = 0.7 if alg == "dm" else 0.9
    p = load_plan(gpu_ctx, f, psi, tile=tile, ept=ept)
    err = p.iterate([step(alg, beta, sem)]).cpu().numpy()[0, 0]
    got = p.get_psi().cpu().numpy()
    want, _, e = O.iteration(psi.astype(np.complex128), f.modulus, f.support, alg, beta, sem)
    assert rel(got, want) < TOL_STEP
    assert abs(err - e) / e < TOL_TRACE


# ------------------------------------------------ lockstep HIO trajectory
def test_lockstep_hio_trace_first_50(gpu_ctx):
    """cfg-1 style disk frame: 50 HIO (beta 0.9) iterations; at every
    iteration the oracle is restarted from the GPU iterate."""
    n = 256
    print("[lockstep hio 256 x 50]", flush=True)
    f = frame(n, seed=1)
    psi = start_psi(n, f.support, seed=7)
    p = load_plan(gpu_ctx, f, psi)
    cur = psi.astype(np.complex128)
    worst_t, worst_s = 0.0, 0.0
    traces = []
    for k in range(50):
        e_gpu = p.iterate([step("hio", 0.9)]).cpu().numpy()[0, 0]
        nxt = p.get_psi().cpu().numpy()
        want, _, e = O.iteration(cur, f.modulus, f.support, "hio", 0.9)
        worst_t = max(worst_t, abs(e_gpu - e) / e)
        worst_s = max(worst_s, rel(nxt, want))
        traces.append(e_gpu)
        cur = nxt.astype(np.complex128)
    print(f"worst trace rel {worst_t:.2e}, worst step rel {worst_s:.2e}")
    assert worst_t < TOL_TRACE
    assert worst_s < TOL_STEP
    # the same 50 iterations in one call give the identical trace
    q = load_plan(gpu_ctx, f, psi)
    one = q.iterate([step("hio", 0.9)] * 50).cpu().numpy()[:, 0]
    assert np.array_equal(one, np.array(traces))


def test_free_running_er_matches_oracle(gpu_ctx):
    """ER (opencl form, s*psi_out) is contractive: free-running parity."""
    n = 256
    print("[free-run er 256 x 60]", flush=True)
    f = frame(n, seed=2)
    psi = start_psi(n, f.support, seed=3)
    p = load_plan(gpu_ctx, f, psi)
    tr = p.iterate([step("er", 0.9, "opencl")] * 60).cpu().numpy()[:, 0]
    got = p.get_psi().cpu().numpy()
    want, otr, _ = O.iterate(psi, f.modulus, f.support, 60, (("er", 1),), 0.9, "opencl")
    assert np.max(np.abs(tr - otr) / otr) < TOL_TRACE
    a = O.align_global_phase_closed(got)
    b = O.align_global_phase_closed(want)
    assert rel(a, b) < 1e-3


@pytest.mark.parametrize("alg,sem", [("hio", "numpy"), ("er", "numpy"), ("er", "opencl"),
                                     ("raar", "numpy"), ("sf", "numpy"), ("sf", "opencl")])
def test_known_object_fixed_point(gpu_ctx, alg, sem):
    n = 128
    f = frame(n, seed=4, noise=False)
    p = load_plan(gpu_ctx, f, f.obj.astype(np.complex64))
    e = p.iterate([step(alg, 0.9, sem)] * 3).cpu().numpy()[:, 0]
    got = p.get_psi().cpu().numpy()
    assert rel(got, f.obj) < 2e-6
    assert np.all(e < 1e-6)


def test_dm_doubles_solution(gpu_ctx):
    n = 128
    f = frame(n, seed=4, noise=False)
    p = load_plan(gpu_ctx, f, f.obj.astype(np.complex64))
    p.iterate([step("dm", 0.7)])
    assert rel(p.get_psi().cpu().numpy(), 2 * f.obj) < 2e-6


@pytest.mark.parametrize("tile,ept", [(4, 16), (8, 8), (2, 8)])
def test_dm_multi_step_lockstep(gpu_ctx, tile, ept):
    """DM over several iterations uses the row pass's phib emission; the
    result equals step-by-step calls (phib rebuilt each call) and the oracle
    per step."""
    n = 128
    f = frame(n, seed=6)
    psi = start_psi(n, f.support, seed=8)
    p = load_plan(gpu_ctx, f, psi, tile=tile, ept=ept)
    tr_all = p.iterate([step("dm", 0.7)] * 8).cpu().numpy()[:, 0]
    end_all = p.get_psi().cpu().numpy()
    q = load_plan(gpu_ctx, f, psi, tile=tile, ept=ept)
    cur = psi.astype(np.complex128)
    for k in range(8):
        e = q.iterate([step("dm", 0.7)]).cpu().numpy()[0, 0]
        nxt = q.get_psi().cpu().numpy()
        want, _, eo = O.iteration(cur, f.modulus, f.support, "dm", 0.7)
        assert rel(nxt, want) < TOL_STEP
        assert abs(e - eo) / eo < TOL_TRACE
        assert e == tr_all[k]
        cur = nxt.astype(np.complex128)
    assert np.array_equal(nxt, end_all)


def test_mixed_schedule_with_dm(gpu_ctx):
    n = 64
    f = frame(n, seed=9)
    psi = start_psi(n, f.support, seed=1)
    algs = ["hio", "dm", "dm", "raar", "er", "dm", "sf"]
    p = load_plan(gpu_ctx, f, psi)
    tr = p.iterate([step(a, 0.75) for a in algs]).cpu().numpy()[:, 0]
    cur = psi.astype(np.complex128)
    q = load_plan(gpu_ctx, f, psi)
    for k, a in enumerate(algs):
        q.iterate([step(a, 0.75)])
        nxt = q.get_psi().cpu().numpy()
        want, _, eo = O.iteration(cur, f.modulus, f.support, a, 0.75)
        assert rel(nxt, want) < TOL_STEP
        assert abs(tr[k] - eo) / eo < TOL_TRACE
        cur = nxt.astype(np.complex128)


def test_zero_guard(gpu_ctx):
    n = 64
    f = frame(n)
    p = load_plan(gpu_ctx, f, np.zeros((n, n), np.complex64))
    e = p.iterate([step("hio", 0.9)]).cpu().numpy()[0, 0]
    got = p.get_psi().cpu().numpy()
    assert np.all(np.isfinite(got)) and np.all(got == 0)
    assert math.isclose(e, 1.0, rel_tol=1e-12)


# ------------------------------------------------------------ shrinkwrap
@pytest.mark.parametrize("n,sigma,tile", [(64, 3.0, 0), (256, 2.0, 0), (512, 1.5, 0),
                                          (256, 2.0, 2), (256, 3.0, 8), (2048, 2.5, 0)])
def test_shrinkwrap_matches_oracle(gpu_ctx, n, sigma, tile):
    print(f"[shrinkwrap n={n} sigma={sigma} tile={tile}]", flush=True)
    f = frame(n, seed=3)
    rng = np.random.default_rng(n)
    psi = (f.obj + 0.05 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))).astype(np.complex64)
    p = load_plan(gpu_ctx, f, psi, tile=tile)
    p.shrinkwrap(sigma, 0.2)
    got = p.get_support().cpu().numpy()
    want, b = O.shrinkwrap_support(psi.astype(np.complex128), f.support, sigma, 0.2)
    diff = got != want
    near = np.abs(b - 0.2 * b.max()) < 1e-5 * b.max()
    assert np.all(~diff | near), int(diff.sum())
    assert np.all(got <= f.support)


def test_shrinkwrap_in_loop_lockstep(gpu_ctx):
    n = 256
    f = frame(n, seed=5)
    psi = start_psi(n, f.support, seed=2)
    p = load_plan(gpu_ctx, f, psi)
    sw = {19: (3.0, 0.2), 39: (1.5, 0.2)}
    s = f.support.astype(np.float64)
    cur = psi.astype(np.complex128)
    for k in range(40):
        sig, tau = sw.get(k, (0.0, 0.0))
        p.iterate([step("hio", 0.9, "numpy", sig, tau)])
        nxt = p.get_psi().cpu().numpy()
        want, _, _ = O.iteration(cur, f.modulus, s, "hio", 0.9)
        assert rel(nxt, want) < TOL_STEP
        cur = nxt.astype(np.complex128)
        if k in sw:
            sg = p.get_support().cpu().numpy()
            so, b = O.shrinkwrap_support(cur, s, sig, tau)
            near = np.abs(b - tau * b.max()) < 1e-5 * b.max()
            assert np.all((sg == so) | near)
            s = sg.astype(np.float64)


# ------------------------------------------------------- batching identity
def test_batch_frames_equal_single_runs(gpu_ctx):
    n = 128
    f = frame(n, seed=7)
    psis = [start_psi(n, f.support, seed=s) for s in (1, 2, 3)]
    p = Plan(gpu_ctx, n, batch=3, shared=True)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.set_support(dev(f.support, torch.float32))
    p.set_psi(dev(np.stack(psis), torch.complex64), mult_support=False)
    sched = [step("raar", 0.87)] * 5 + [step("dm", 0.7)] * 3 + [step("er", 0.9)] * 2
    tr = p.iterate(sched).cpu().numpy()
    out = p.get_psi().cpu().numpy()
    for i, psi in enumerate(psis):
        q = load_plan(gpu_ctx, f, psi)
        t1 = q.iterate(sched).cpu().numpy()[:, 0]
        assert np.array_equal(t1, tr[:, i])
        assert np.array_equal(q.get_psi().cpu().numpy(), out[i])


def test_per_frame_steps_and_idle_frames(gpu_ctx):
    n = 64
    fr = [frame(n, seed=s) for s in (11, 12)]
    psis = [start_psi(n, x.support, seed=s) for x, s in zip(fr, (1, 2))]
    p = Plan(gpu_ctx, n, batch=2, shared=False)
    p.set_modulus(dev(np.stack([x.modulus for x in fr]), torch.float32))
    p.set_support(dev(np.stack([x.support for x in fr]), torch.float32))
    p.set_psi(dev(np.stack(psis), torch.complex64), mult_support=False)
    noop = (_capi.NOOP, 0, 0.0, 0.0, 0.0)
    steps = []
    for k in range(4):
        steps += [step("hio", 0.9), noop if k < 2 else step("er", 0.9)]
    tr = p.iterate(steps, per_frame=True).cpu().numpy()
    assert np.all(np.isnan(tr[:2, 1])) and np.all(np.isfinite(tr[2:, 1]))
    out = p.get_psi().cpu().numpy()
    a = load_plan(gpu_ctx, fr[0], psis[0])
    ta = a.iterate([step("hio", 0.9)] * 4).cpu().numpy()[:, 0]
    assert np.array_equal(ta, tr[:, 0]) and np.array_equal(a.get_psi().cpu().numpy(), out[0])
    b = load_plan(gpu_ctx, fr[1], psis[1])
    tb = b.iterate([step("er", 0.9)] * 2).cpu().numpy()[:, 0]
    assert np.array_equal(tb, tr[2:, 1]) and np.array_equal(b.get_psi().cpu().numpy(), out[1])


# ------------------------------------------------------------- align phase
def test_align_global_phase_matches_fminbound(gpu_ctx):
    rng = np.random.default_rng(21)
    data = (rng.standard_normal((5, 40, 48)) + 1j * rng.standard_normal((5, 40, 48)) + 2.0)
    data = (data * np.exp(1j * rng.uniform(0, 2 * np.pi, (5, 1, 1)))).astype(np.complex64)
    t = dev(data, torch.complex64)
    gpu_ctx.align_global_phase(t)
    want = O.align_global_phase(data.astype(np.complex128))
    assert rel(t.cpu().numpy(), want) < 1e-5


def test_context_outlives_plans():
    """Destroying a context before its plans is safe (plans hold a reference)."""
    from als_speckle_amd.device import Context
    c = Context(0)
    p = Plan(c, 64)
    q = Plan(c, 32)
    c.close()
    f = frame(64)
    p.set_modulus(dev(f.modulus, torch.float32))
    p.close()
    q.close()


def test_graph_replay_matches_eager():
    """A captured reconstruction block replays to the eager result, every time,
    with eager work and synchronisations between the replays."""
    from als_speckle_amd.device import Context
    n = 128
    f = frame(n, seed=12)
    psi = start_psi(n, f.support, seed=4)
    stream = torch.cuda.Stream()
    ctx = Context(0, stream=stream)
    p = Plan(ctx, n)
    m, s, g = dev(f.modulus, torch.float32), dev(f.support, torch.float32), dev(psi, torch.complex64)
    err = torch.empty((30, 1), dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    sched = [step("hio", 0.9, "numpy", 2.0 if k == 19 else 0.0, 0.2) for k in range(25)] + \
            [step("er", 0.9)] * 5

    def block():
        p.set_support(s)
        p.set_psi(g, mult_support=False)
        p.iterate(sched, err_out=err)

    with torch.cuda.stream(stream):
        p.set_modulus(m)
        block()
        ctx.sync()
        want_psi = p.get_psi().cpu().numpy()
        want_err = err.cpu().numpy().copy()
        graph = p.capture(block)
        for rep in range(3):
            err.zero_()
            graph.launch()
            ctx.sync()
            assert np.array_equal(p.get_psi().cpu().numpy(), want_psi), rep
            assert np.array_equal(err.cpu().numpy(), want_err), rep
            q = load_plan(ctx, f, psi)   # unrelated eager work in between
            q.iterate([step("raar", 0.8)] * 3)
            # eager work on the captured plan itself, with other step tables
            p.iterate([step("hio", 0.5, "numpy", 1.0 + rep, 0.3)] * 4 + [step("er")])
            ctx.sync()
    graph.close()


def test_two_graphs_of_one_plan_replay_their_own_tables():
    """Two blocks with different step tables captured from one plan: each
    graph replays its own block (the tables live in the graph), also after
    the other graph has been destroyed."""
    from als_speckle_amd.device import Context
    n = 64
    f = frame(n, seed=13)
    psi = start_psi(n, f.support, seed=5)
    stream = torch.cuda.Stream()
    ctx = Context(0, stream=stream)
    p = Plan(ctx, n)
    m, s, g = dev(f.modulus, torch.float32), dev(f.support, torch.float32), dev(psi, torch.complex64)
    torch.cuda.synchronize()
    scheds = [[step("hio", 0.9, "numpy", 1.5 if k == 9 else 0.0, 0.2) for k in range(12)],
              [step("raar", 0.8)] * 6 + [step("er", 0.9, "numpy", 2.0, 0.25)] + [step("er")] * 3]

    def block(sched):
        p.set_support(s)
        p.set_psi(g, mult_support=False)
        p.iterate(sched, error_trace=False)

    with torch.cuda.stream(stream):
        p.set_modulus(m)
        want = []
        for sc in scheds:
            block(sc)
            ctx.sync()
            want.append(p.get_psi().cpu().numpy())
        graphs = [p.capture(lambda sc=sc: block(sc)) for sc in scheds]
        for rep in range(2):
            for gr, w in zip(graphs, want):
                gr.launch()
                ctx.sync()
                assert np.array_equal(p.get_psi().cpu().numpy(), w), rep
        graphs[0].close()
        graphs[1].launch()
        ctx.sync()
        assert np.array_equal(p.get_psi().cpu().numpy(), want[1])
        graphs[1].close()


def test_er_numpy_steps_without_trace(gpu_ctx):
    """ER with numpy semantics (s * psi_in) does not need the projected Phi:
    without an error trace its column pass is skipped.  The iterates must
    not depend on whether the trace is formed, and one ER step must match
    the oracle."""
    n = 128
    f = frame(n, seed=21)
    psi = start_psi(n, f.support, seed=8)
    sched = [step("hio", 0.9)] * 4 + [step("er", 0.9)] * 3 + [step("hio", 0.9)] * 2 + \
            [step("er", 0.9)]
    a = load_plan(gpu_ctx, f, psi)
    ta = a.iterate(sched).cpu().numpy()[:, 0]
    b = load_plan(gpu_ctx, f, psi)
    assert b.iterate(sched, error_trace=False) is None
    assert np.array_equal(a.get_psi().cpu().numpy(), b.get_psi().cpu().numpy())
    assert np.all(np.isfinite(ta))
    c = load_plan(gpu_ctx, f, psi)
    c.iterate([step("er", 0.9)], error_trace=False)
    want, _, _ = O.iteration(psi.astype(np.complex128), f.modulus, f.support, "er", 0.9)
    assert rel(c.get_psi().cpu().numpy(), want) < TOL_STEP


def test_single_1024_frame_equals_batched_copy(gpu_ctx):
    """One 1024^2 frame alone (row-pass stores non-temporal) and the same
    frame in a batch of two (plain stores) give identical iterates and
    traces: the store form changes no value."""
    n = 1024
    f = synth.config_frame("single_1024", seed=0)
    psi = start_psi(n, f.support, seed=9)
    sched = [step("hio", 0.8)] * 6 + [step("er", 0.8)] * 2 + [step("hio", 0.8)] * 2
    one = load_plan(gpu_ctx, f, psi)
    t1 = one.iterate(sched).cpu().numpy()[:, 0]
    two = load_plan(gpu_ctx, f, psi, batch=2)
    t2 = two.iterate(sched).cpu().numpy()
    assert np.array_equal(t1, t2[:, 0]) and np.array_equal(t2[:, 0], t2[:, 1])
    assert np.array_equal(one.get_psi().cpu().numpy(), two.get_psi().cpu().numpy()[0])
